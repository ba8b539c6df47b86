// Mirrors the reference's tests/test_fsp_mat.cu: the 6-reaction toggle on the
// (31, 31) box in CUDA_CSR, HYB and KRONECKER form; A.0 = 0 for each, CSR ==
// HYB (L2 <= 1e-14) and KRON == HYB (L2 <= 1e-12) on x = 1.
#include <cassert>
#include <cmath>
#include <cstdio>
#include <vector>

#include "FSPMat.h"

const double ayx{2.6e-3}, axy{6.1e-3}, nyx{3e0}, nxy{2.1e0}, kx0{2.2e-3}, kx{1.7e-2},
    dx{3.8e-4}, ky0{6.8e-5}, ky{1.6e-2}, dy{3.8e-4};

double toggle_propensity(int *x, int reaction)
{
    switch (reaction) {
    case 0: return 1.0;
    case 1: return 1.0 / (1.0 + ayx * std::pow(1.0 * x[1], nyx));
    case 2: return 1.0 * x[0];
    case 3: return 1.0;
    case 4: return 1.0 / (1.0 + axy * std::pow(1.0 * x[0], nxy));
    case 5: return 1.0 * x[1];
    default: return 0.0;
    }
}

double toggle_propensity_factor(int x, int species, int reaction)
{
    double v = 1.0;
    switch (reaction) {
    case 1: if (species == 1) v = 1.0 / (1.0 + ayx * std::pow(1.0 * x, nyx)); break;
    case 2: if (species == 0) v = 1.0 * x; break;
    case 4: if (species == 0) v = 1.0 / (1.0 + axy * std::pow(1.0 * x, nxy)); break;
    case 5: if (species == 1) v = 1.0 * x; break;
    default: break;
    }
    return v;
}

void t_func(double, double *out)
{
    out[0] = kx0; out[1] = kx; out[2] = dx; out[3] = ky0; out[4] = ky; out[5] = dy;
}

static double l2_diff(cuFSP::dvec &a, cuFSP::dvec &b)
{
    std::vector<double> x = a.to_host(), y = b.to_host();
    double s = 0.0;
    for (size_t i = 0; i < x.size(); ++i) s += (x[i] - y[i]) * (x[i] - y[i]);
    return std::sqrt(s);
}

int main()
{
    int n_species = 2, n_reactions = 6;
    int stoich_vals[] = {1, 1, -1, 1, 1, -1};
    int stoich_colidxs[] = {0, 0, 0, 1, 1, 1};
    int stoich_rowptrs[] = {0, 1, 2, 3, 4, 5, 6};
    cuFSP::CSRMatInt stoich;
    stoich.vals = stoich_vals;
    stoich.col_idxs = stoich_colidxs;
    stoich.row_ptrs = stoich_rowptrs;
    stoich.n_rows = 6;
    stoich.n_cols = 2;
    stoich.nnz = 6;
    int n_bounds[2] = {(1 << 5) - 1, (1 << 5) - 1};
    int n_states = cuFSP::rect_fsp_num_states(n_species, n_bounds);

    cuFSP::FSPMat A(n_reactions, n_species, n_bounds, stoich, &t_func, &toggle_propensity);
    cuFSP::FSPMat A2(n_reactions, n_species, n_bounds, stoich, &t_func, &toggle_propensity,
                     cuFSP::HYB);
    cuFSP::FSPMat A3(n_reactions, n_species, n_bounds, stoich, &t_func,
                     &toggle_propensity_factor, cuFSP::KRONECKER);

    cuFSP::thrust_dvec v(n_states), w(n_states), w2(n_states), w3(n_states);
    A.action(1.0, v.data(), w.data());
    assert(std::abs(w.sum()) <= 1.0e-14);
    A2.action(1.0, v.data(), w2.data());
    assert(std::abs(w2.sum()) <= 1.0e-14);
    A3.action(1.0, v.data(), w3.data());
    assert(std::abs(w3.sum()) <= 1.0e-14);

    v.fill(1.0);
    A.action(1.0, v.data(), w.data());
    A2.action(1.0, v.data(), w2.data());
    A3.action(1.0, v.data(), w3.data());
    double e1 = l2_diff(w, w2), e2 = l2_diff(w3, w2);
    std::printf("error_l2 = %.3e\nerror_l2 = %.3e\n", e1, e2);
    assert(e1 <= 1.0e-14);
    assert(e2 <= 1.0e-12);
    std::printf("Matvec test successful.\n");
    return 0;
}
