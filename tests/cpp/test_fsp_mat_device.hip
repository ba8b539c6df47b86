// The reference's tests/test_fsp_mat.cu with its device PropFun route: the
// 6-reaction toggle on the (31, 31) box, the propensity a __device__
// function (toggle_device.hip) whose address comes from hipMemcpyFromSymbol
// (tests/test_fsp_mat.cu:123-124) and goes to the FSPMat constructor
// unchanged.  Checks the reference's own (A 0 sums to 0 for CUDA_CSR and
// HYB), CSR == HYB on a random vector (1e-14 relative), and that the
// device-tabulated generator applies as the one built from the same
// propensity evaluated on the host (the two pow implementations may differ
// in the last bit: 1e-13 relative).
#include <cassert>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "FSPMat.h"

extern __device__ cuFSP::PropFun toggle_prop_pointer;

static double toggle_propensity_host(int *x, int reaction)
{
    const double ayx = 2.6e-3, axy = 6.1e-3, nyx = 3e0, nxy = 2.1e0;
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

static void t_func(double, double *out)
{
    const double kx0 = 2.2e-3, kx = 1.7e-2, dx = 3.8e-4, ky0 = 6.8e-5, ky = 1.6e-2, dy = 3.8e-4;
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
    int n_bounds[2] = {(1 << 5) - 1, (1 << 5) - 1};
    const int n_states = cuFSP::rect_fsp_num_states(n_species, n_bounds);

    cuFSP::PropFun host_prop_ptr = nullptr;
    if (hipMemcpyFromSymbol(&host_prop_ptr, HIP_SYMBOL(toggle_prop_pointer),
                            sizeof(cuFSP::PropFun)) != hipSuccess) {
        std::printf("hipMemcpyFromSymbol failed\n");
        return 1;
    }
    cuFSP::FSPMat A(n_reactions, n_species, n_bounds, stoich, &t_func, host_prop_ptr);
    cuFSP::FSPMat A2(n_reactions, n_species, n_bounds, stoich, &t_func, host_prop_ptr,
                     cuFSP::HYB);
    cuFSP::FSPMat Ah(n_reactions, n_species, n_bounds, stoich, &t_func,
                     &toggle_propensity_host);
    std::printf("device PropFun: CUDA_CSR and HYB generators built\n");
    std::fflush(stdout);

    // the reference's checks (tests/test_fsp_mat.cu:131-150): A v with v = 0
    cuFSP::thrust_dvec v(n_states), w(n_states), w2(n_states);
    A.action(1.0, v.data(), w.data());
    A2.action(1.0, v.data(), w2.data());
    std::printf("sum(A 0) = %.3e (CSR), %.3e (HYB)\n", w.sum(), w2.sum());
    assert(std::abs(w.sum()) <= 1.0e-14);
    assert(std::abs(w2.sum()) <= 1.0e-14);
    std::mt19937_64 rng(3);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<double> xh(n_states);
    for (double &e : xh) e = U(rng);
    cuFSP::thrust_dvec x(n_states), yd(n_states), yd2(n_states), yh(n_states);
    x.from_host(xh);
    A.action(0.0, x.data(), yd.data());
    A2.action(0.0, x.data(), yd2.data());
    Ah.action(0.0, x.data(), yh.data());
    std::vector<double> p = yd.to_host(), p2 = yd2.to_host(), q = yh.to_host();
    double dmax = 0.0, d2max = 0.0, ymax = 0.0;
    for (int i = 0; i < n_states; ++i) {
        dmax = std::max(dmax, std::abs(p[i] - q[i]));
        d2max = std::max(d2max, std::abs(p[i] - p2[i]));
        ymax = std::max(ymax, std::abs(q[i]));
    }
    std::printf("device propensity, CSR vs HYB: max |dy| = %.3e; device vs host propensity: "
                "max |dy| = %.3e (max |y| = %.3e)\n", d2max, dmax, ymax);
    std::fflush(stdout);
    assert(d2max <= 1e-14 * ymax);
    assert(dmax <= 1e-13 * ymax);
    std::printf("device PropFun test done\n");
    return 0;
}
