// FSPMat.h -- drop-in for the reference's src/FSPMat.h (namespace cuFSP).
//
// The generator A(t) = sum_r c_r(t) A_r of the CME on a rectangular FSP box,
// held on the GPU by the engine (fsp_mat).  Same constructors, action()
// semantics (y += A(t) x, t_func called on every action) and getters as
// src/FSPMat.h:38-55; in addition every format accepts both propensity
// forms, a precision and an optional sink state.
#pragma once

#include <functional>

#include "cme_util.h"

namespace cuFSP {

// src/FSPMat.h:16
typedef std::function<void(double, double *)> TcoefFun;
// src/FSPMat.h:17-19
enum MatrixFormat { CUDA_CSR, HYB, KRONECKER };
// FP64 throughout, or FP32 values/basis with FP64 sums (no reference counterpart)
enum Precision { FP64, MIXED };

class FSPMat {
public:
    // src/FSPMat.cu:10-45.  prop_func is either a host function, evaluated
    // on the host over the box a chunk at a time, or (as in the reference)
    // the address of a __device__ function fetched with hipMemcpyFromSymbol,
    // evaluated on the GPU by the tabulator at the end of this header.  The
    // table [R][n] lives on the GPU.
    explicit FSPMat(int n_reactions, int n_species, int *fsp_dim, CSRMatInt stoich,
                    TcoefFun t_func, PropFun prop_func, MatrixFormat format = CUDA_CSR,
                    Precision precision = FP64, bool with_sink = false);
    // src/FSPMat.cu:49-76 (separable propensity factors, tabulated on the host)
    explicit FSPMat(int n_reactions, int n_species, int *fsp_dim, CSRMatInt stoich,
                    TcoefFun t_func, PropFactorFun pffunc, MatrixFormat format,
                    Precision precision = FP64, bool with_sink = false);
    ~FSPMat();
    FSPMat(const FSPMat &) = delete;
    FSPMat &operator=(const FSPMat &) = delete;

    // src/FSPMat.cu:102-105: y += A(t) x on device vectors of get_n_rows()
    void action(double t, double *x, double *y);

    int get_n_rows();       // box states (+1 with the sink)
    int get_n_species();
    int get_n_reactions();

    // engine handle (for the fused Krylov path) and the coefficients at t
    fsp_mat handle() const { return h_; }
    void coefficients(double t, std::vector<double> &c);
    void set_coefficients(const std::vector<double> &c);

private:
    void init(int n_reactions, int n_species, int *fsp_dim, CSRMatInt stoich,
              const double *factors, const double *table_dev, MatrixFormat format,
              Precision precision, bool with_sink);
    fsp_mat h_ = nullptr;
    int nst_ = 0, ns_ = 0, nr_ = 0, rows_ = 0;
    TcoefFun tcoeffunc_;
    std::vector<double> tcoef_;
    double *table_dev_ = nullptr;
};

}  // namespace cuFSP

#if defined(__HIPCC__)
// ------------------------------------------------------ device propensities
// The reference's PropFun is a __device__ function pointer, fetched with
// cudaMemcpyFromSymbol and called inside its assembly kernels
// (src/cme_util.h:25, src/fspmat_csr_kernels.cu:79,83,
// tests/test_fsp_mat.cu:86,123-124).  The same code works here with
// hipMemcpyFromSymbol: the FSPMat constructor sees an address that is no
// host code and calls cufsp_tabulate_device_propensity below, which
// evaluates a_r(x) on the GPU over the box into the device table
// [R][n] (one thread per state, the R reactions in order; each call gets
// its own copy of the state).
//
// The tabulator is compiled into every hipcc translation unit that
// includes this header (a weak definition: the program keeps one).  An
// indirect call needs the propensity and the calling kernel in one code
// object, so a program whose device propensity lives in another
// translation unit is built with -fgpu-rdc, as the reference builds with
// separable compilation.
#include <hip/hip_runtime.h>

namespace cuFSP {
namespace device_propensity {
struct Box {
    int n_species, n_reactions, n_states;
    int bounds[kMaxSpeciesFacade];
};

static __global__ void __launch_bounds__(256) k_tabulate(PropFun f, Box box, double *table)
{
    const long long n = box.n_states;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        // index -> state, species 0 fastest (src/cme_util.cu:5-11)
        int x[kMaxSpeciesFacade];
        long long rem = i;
        for (int s = 0; s < box.n_species; ++s) {
            x[s] = (int)(rem % (box.bounds[s] + 1));
            rem /= box.bounds[s] + 1;
        }
        for (int r = 0; r < box.n_reactions; ++r) {
            int xc[kMaxSpeciesFacade];
            for (int s = 0; s < box.n_species; ++s) xc[s] = x[s];
            table[(size_t)r * (size_t)n + (size_t)i] = f(xc, r);
        }
    }
}
}  // namespace device_propensity
}  // namespace cuFSP

// Fills table_dev[r * n_states + i] = f(x_i, r) on the GPU; 0 on success.
// The caller has checked 1 <= n_species <= kMaxSpeciesFacade.
extern "C" __attribute__((weak)) int cufsp_tabulate_device_propensity(
    cuFSP::PropFun f, int n_species, int n_reactions, const int *bounds, int n_states,
    double *table_dev)
{
    cuFSP::device_propensity::Box box{};
    box.n_species = n_species;
    box.n_reactions = n_reactions;
    box.n_states = n_states;
    for (int s = 0; s < n_species; ++s) box.bounds[s] = bounds[s];
    const long long blocks = ((long long)n_states + 255) / 256;
    const unsigned grid = (unsigned)(blocks < 16384 ? (blocks > 0 ? blocks : 1) : 16384);
    hipLaunchKernelGGL(cuFSP::device_propensity::k_tabulate, dim3(grid), dim3(256), 0, nullptr,
                       f, box, table_dev);
    if (hipGetLastError() != hipSuccess) return 1;
    return hipDeviceSynchronize() == hipSuccess ? 0 : 2;
}
#endif  // __HIPCC__
