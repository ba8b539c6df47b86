// fsp_internal.h -- shared host/device declarations of the HIP FSP engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fsp_capi.h"

namespace fsp {

constexpr int kMaxSpecies = 8;
constexpr int kMaxReactions = 64;
constexpr int kMaxKrylov = 50;      // m <= 50, so (m+2)^2 doubles fit in LDS
constexpr int kBlock = 256;         // rows per workgroup in the row kernels
constexpr int kMaxRowsPerApplyBlock = kBlock;

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
fsp_status fail(fsp_status st, const std::string &msg);

#define FSP_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e__ = (call);                                               \
        if (e__ != hipSuccess)                                                 \
            return ::fsp::fail(FSP_ERR_HIP, std::string(#call) + ": " +        \
                                                hipGetErrorString(e__));       \
    } while (0)

#define FSP_TRY(expr)                                                          \
    do {                                                                       \
        fsp_status s__ = (expr);                                               \
        if (s__ != FSP_OK) return s__;                                         \
    } while (0)

// ------------------------------------------------ division by invariants
// q / div for 0 <= q < 2^31 as (q * mul) >> shift, with l = ceil(log2 div),
// shift = 31 + l, mul = ceil(2^shift / div) (Granlund & Montgomery, PLDI
// 1994, Thm 4.2 with N = 31): exact for every such q, and mul < 2^32.
struct FastDiv {
    uint64_t mul;
    uint32_t shift;
    uint32_t div;
};

inline FastDiv make_fastdiv(uint32_t div)
{
    FastDiv f;
    uint32_t l = 0;
    while ((1ull << l) < div) ++l;
    f.shift = 31 + l;
    f.mul = ((1ull << f.shift) + div - 1) / div;
    f.div = div;
    return f;
}

__host__ __device__ inline uint32_t fast_div(uint32_t q, const FastDiv &f)
{
    return (uint32_t)(((uint64_t)q * f.mul) >> f.shift);
}

// ------------------------------------------------------ model on device
// Per-reaction description.  "Involved" species are those with a nonzero
// net change or a non-constant propensity factor, ascending, so products of
// factors are formed in species order (a factor equal to 1.0 changes no bit).
struct RxnDev {
    int64_t delta;            // source of the inflow into row i is i - delta
    int zero_nu;              // net change 0: contributes exactly 0, skipped
    int slot;                 // unique-net-change slot, -1 if zero_nu
    int nsp;                  // number of involved species
    int sp[kMaxSpecies];
    int nu[kMaxSpecies];
    int foff[kMaxSpecies];    // factor table offset in the compact table, -1 = 1.0
};

struct DevModel {
    int d, R, U;
    int use_table;            // propensities from prop_table (R x n) instead of factors
    int64_t n;
    int radix[kMaxSpecies];
    int64_t stride[kMaxSpecies];
    FastDiv fd[kMaxSpecies];
    RxnDev rx[kMaxReactions];
    int64_t slot_delta[kMaxReactions];
    int slot_ptr[kMaxReactions + 1];   // members of slot u: slot_rxn[slot_ptr[u]..slot_ptr[u+1])
    int slot_rxn[kMaxReactions];       // in reaction order
    // merged-CSR row layout: slots (and the diagonal) sorted by column
    int csr_order[kMaxReactions + 1];  // entry position k -> slot (-1 = diagonal)
    int compact_size;         // doubles in the compact factor table
};

// Host-side model: validated copy of an fsp_model plus derived geometry.
struct HostModel {
    DevModel dm;
    std::vector<double> compact;     // compact factor table (non-constant tables)
    std::vector<int> nu_dense;       // R x d
    const double *prop_table_dev = nullptr;
    bool has_factors = false;
};

fsp_status build_host_model(const fsp_model *m, HostModel &hm);

// Blocking copy and fill ordered on stream s, used in place of hipMemcpy /
// hipMemset: those run on the legacy default stream, and HIP refuses a
// legacy-stream operation while another thread of the process is capturing
// a graph.  s = NULL: the calling thread's own default stream, which still
// orders after earlier work on the legacy stream.
inline hipError_t copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind,
                            hipStream_t s)
{
    if (!s) s = hipStreamPerThread;
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

inline hipError_t fill_sync(void *dst, int value, size_t bytes, hipStream_t s)
{
    if (!s) s = hipStreamPerThread;
    const hipError_t e = hipMemsetAsync(dst, value, bytes, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

// ------------------------------------------------------------ matrices
struct MatImpl {
    HostModel hm;
    fsp_format format;
    fsp_precision prec;
    int with_sink;
    // HYB F64 without a sink only: the arrays hold the transpose A^T (row j:
    // outflow rates of state j at the columns of its targets)
    int transposed = 0;
    int64_t n, n_rows;
    std::vector<double> coefs;       // coefficients the stored values hold
    // device copies
    DevModel *d_model = nullptr;
    double *d_compact = nullptr;     // compact factor table (FP64)
    float *d_compact_f = nullptr;    // same, rounded to FP32 (mixed KRON)
    // KRON slot form (one coefficient-weighted factor table per unique net
    // change), when the model allows it; rebuilt with the coefficients
    int kron_slots = 0;
    double *d_slot_tab = nullptr;
    float *d_slot_tab_f = nullptr;   // same, rounded to FP32 (mixed)
    double *d_coefs = nullptr;       // R coefficients (device copy)
    // CSR (rows 0..n-1) + sink row stored separately in block order
    int *rp = nullptr;
    int *cols = nullptr;
    void *vals = nullptr;            // double or float
    int64_t nnz_rows = 0;            // entries in rows 0..n-1
    int max_row = 0;
    int *sink_cols = nullptr;        // sink row (state indices, ascending)
    void *sink_vals = nullptr;
    int64_t nnz_sink = 0;
    int *sink_block_ptr = nullptr;   // first sink entry of rows >= i (n + 1 entries)
    // HYB
    void *diag = nullptr;
    void *slot_vals = nullptr;
    int *slot_off = nullptr;         // U column offsets (device): col = row + off[u]
    void *sink_out = nullptr;        // dense outflow rates (HYB, n)
    // reductions (sink row, deterministic last-block finalize)
    double *partials = nullptr;
    unsigned int *ticket = nullptr;
    int64_t n_partials = 0;
};

fsp_status mat_assemble(MatImpl &M, hipStream_t s);
fsp_status mat_revalue(MatImpl &M, const double *coefs, hipStream_t s);
fsp_status mat_apply(MatImpl &M, const double *d_x, double *d_y, int accumulate,
                     hipStream_t s, const int *d_skip = nullptr);
fsp_status mat_apply_f32x(MatImpl &M, const float *d_x, double *d_y, hipStream_t s,
                          const int *d_skip);
fsp_status mat_apply_transpose(MatImpl &M, const double *d_x, double *d_y,
                               int accumulate, hipStream_t s);
fsp_status mat_apply_term(MatImpl &M, int r, double c, const double *d_x, double *d_y,
                          int accumulate, hipStream_t s);
// Generator diagonal -sum_r c_r a_r(x_i), i < n (matrix free, any format).
fsp_status mat_diagonal(MatImpl &M, double *d_diag, hipStream_t s);
void mat_free(MatImpl &M);
// Explicit HYB slot columns [U][n] of a (non-transposed) HYB matrix.
fsp_status mat_hyb_cols(const MatImpl &M, int *d_cols, hipStream_t s);
double mat_bytes_per_apply(const MatImpl &M);

fsp_status term_csr(const HostModel &hm, int r, int *d_rp, int *d_cols, double *d_vals,
                    int64_t *h_nnz, hipStream_t s);

// -------------------------------------------------------------- scans
// Exclusive scan of int32 counts into int32 offsets (n+1 entries written:
// out[0] = 0, out[n] = total); FSP_ERR_OVERFLOW past 2^31-1.
fsp_status scan_counts(const int *d_counts, int *d_out, int64_t n, int64_t *h_total,
                       hipStream_t s);

// -------------------------------------------------------------- vectors
fsp_status vec_sum(const double *d_x, int64_t n, double *h_sum, hipStream_t s);

}  // namespace fsp

// Internal view of a matrix handle, shared by the library's sources and not
// exported from the shared object.
extern "C" __attribute__((visibility("hidden"))) fsp_status fsp_mat_internal(fsp_mat A,
                                                                          fsp::MatImpl **impl);
// Krylov solve to the final time that also writes v at the output times
// t_out[0..n_out) (increasing, within (t_now, t_final]) from inside the
// accepted steps; fused matrix operators only.  Not exported.
extern "C" __attribute__((visibility("hidden"))) fsp_status fsp_expv_solve_dense(
    fsp_expv e, double *d_v, int n_out, const double *t_out, double *const *d_out);

// Transposed generator A^T of a model (HYB, F64, no sink), for solves with
// A^T such as the adjoint pass of the likelihood gradient; not exported.
extern "C" __attribute__((visibility("hidden"))) fsp_status fsp_mat_create_transposed(
    const fsp_model *model, const double *coefs, void *stream, fsp_mat *out);
