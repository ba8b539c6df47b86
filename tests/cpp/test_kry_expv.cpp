// Mirrors the reference's tests/test_kry_expv.cu: toggle on (256, 256), HYB,
// p0 = e_0, t = 8 h, m = 30, tol 1e-8, IOP; mass kept to 1e-10.  Runs both
// the reference's std::function matvec route and the fused FSPMat route.
#include <cassert>
#include <cmath>
#include <cstdio>

#include "KryExpvFSP.h"

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

void t_func(double, double *out)
{
    out[0] = kx0; out[1] = kx; out[2] = dx; out[3] = ky0; out[4] = ky; out[5] = dy;
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
    int n_bounds[] = {1 << 8, 1 << 8};
    int n_states = cuFSP::rect_fsp_num_states(n_species, n_bounds);
    cuFSP::FSPMat A(n_reactions, n_species, n_bounds, stoich, &t_func, &toggle_propensity,
                    cuFSP::HYB);

    double t_final = 8 * 3600, tol = 1.0e-8;
    int m = 30;
    cuFSP::thrust_dvec v(n_states);
    v[0] = 1.0;
    std::function<void(double *, double *)> matvec = [&](double *x, double *y) {
        A.action(1.0, x, y);
    };
    cuFSP::KryExpvFSP expv(t_final, matvec, v, m, tol, true);
    expv.solve();
    double vsum = v.sum();
    std::printf("callback route: 1 - sum = %.3e, steps %lld\n", 1.0 - vsum,
                (long long)expv.stats().i_step);
    assert(std::abs(1.0 - vsum) <= 1.0e-10);

    cuFSP::thrust_dvec v2(n_states);
    v2[0] = 1.0;
    cuFSP::KryExpvFSP expv2(t_final, A, 1.0, v2, m, tol, true);
    expv2.solve();
    std::vector<double> a = v.to_host(), b = v2.to_host();
    double l1 = 0.0;
    for (size_t i = 0; i < a.size(); ++i) l1 += std::abs(a[i] - b[i]);
    std::printf("fused route: 1 - sum = %.3e, L1 vs callback %.3e\n", 1.0 - v2.sum(), l1);
    assert(std::abs(1.0 - v2.sum()) <= 1.0e-10);
    assert(l1 <= 1.0e-12);
    return 0;
}
