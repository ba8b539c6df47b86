// Error plumbing for the CPU (SIMT stand-in) build of the batched solver:
// the same thread-local last-error string and fail() as the HIP library's
// C ABI, so fsp_batch_* report errors through fsp_last_error().
#include <string>

#include "fsp_batch_impl.h"
#include "fsp_expm.h"

namespace fsp {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

fsp_status fail(fsp_status st, const std::string &msg)
{
    set_error(msg);
    return st;
}

// The stream scheduler drives the HIP runtime's streams and the
// single-instance engine, which this CPU build does not contain; the
// simulated device always takes the persistent-kernel path.
fsp_status batch_solve_streams(BatchImpl &)
{
    return fail(FSP_ERR_UNSUPPORTED, "no stream scheduler in the CPU build");
}
fsp_status batch_autotune_streams(BatchImpl &, double *, int *)
{
    return fail(FSP_ERR_UNSUPPORTED, "no stream scheduler in the CPU build");
}
void batch_streams_free(BatchImpl &) {}

}  // namespace fsp

extern "C" const char *fsp_last_error(void) { return fsp::g_last_error.c_str(); }

// The device Pade exponential (fsp::expm_block) of an N x N column-major
// matrix by one workgroup of `block` lanes, as the library's k_expm_dense
// runs it: returns 0, or 2 when the LDS is not 16-byte aligned (the
// 32 x 32 Hessenberg path would not be taken).
namespace {
__global__ void k_simt_expm(int N, const double *Ag, double *Fg, int *aligned)
{
    HIP_DYNAMIC_SHARED(double, smem)
    const int NN = N * N;
    double *A = smem, *X = A + NN, *T = X + NN, *Nm = T + NN, *Dm = Nm + NN;
    __shared__ double red[fsp::kBlock / fsp::kWave];
    if (threadIdx.x == 0) *aligned = (reinterpret_cast<uintptr_t>(smem) & 15) == 0;
    for (int e = threadIdx.x; e < NN; e += blockDim.x) A[e] = Ag[e];
    __syncthreads();
    fsp::expm_block(N, A, X, T, Nm, Dm, red);
    for (int e = threadIdx.x; e < NN; e += blockDim.x) Fg[e] = Nm[e];
}
}  // namespace

extern "C" int simt_expm_dense(int N, const double *A, double *F, int block)
{
    int aligned = 0;
    hipLaunchKernelGGL(k_simt_expm, dim3(1), dim3(block), sizeof(double) * 5 * (size_t)N * N,
                       nullptr, N, A, F, &aligned);
    if (hipGetLastError() != hipSuccess) return 1;
    return aligned ? 0 : 2;
}
