// fsp_capi.hip -- C ABI entry points (include/fsp_capi.h): errors, device
// memory helpers, state enumeration, vector reductions and fsp_mat.
#include <cstring>
#include <mutex>
#include <string>

#include "fsp_formats.h"

namespace fsp {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

fsp_status fail(fsp_status st, const std::string &msg)
{
    set_error(msg);
    return st;
}

namespace {

// fsp_get_states (src/cme_util.cu:24-41): AoS int[n][d], species 0 fastest.
template <int D>
__global__ void k_enumerate(const DevModel *__restrict__ dm, uint32_t n, int *__restrict__ out)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    int xs[D];
    decode<D>(i, dm, xs);
#pragma unroll
    for (int s = 0; s < D; ++s) out[(size_t)i * D + s] = xs[s];
}

// state2indx (src/cme_util.cu:13-22) over a batch; outside the box -> -1.
__global__ void k_to_index(const DevModel *__restrict__ dm, const int *__restrict__ states,
                           int64_t count, int *__restrict__ idx)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    const int d = dm->d;
    int64_t v = 0;
    bool ok = true;
    for (int s = 0; s < d; ++s) {
        const int x = states[k * d + s];
        ok = ok && x >= 0 && x < dm->radix[s];
        v += (int64_t)x * dm->stride[s];
    }
    idx[k] = ok ? (int)v : -1;
}

__device__ inline void put_partial(double *partials, dd v, double *red)
{
    v = block_sum_dd(v, red);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = v.s;
        partials[2 * blockIdx.x + 1] = v.c;
    }
}

__global__ void k_sum_partials(const double *__restrict__ x, int64_t n, double *partials)
{
    __shared__ double red[2 * kBlock / kWave];
    dd s = dd_zero();
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        s = dd_add(s, x[i]);
    put_partial(partials, s, red);
}

// out = fl(sum of G dd partials)
__global__ void k_sum_final(const double *__restrict__ partials, int G, double *out)
{
    __shared__ double red[2 * kBlock / kWave];
    const double s = dd_value(sum_dd_partials(partials, G, red));
    if (threadIdx.x == 0) *out = s;
}

__global__ void k_axpy(double a, const double *__restrict__ x, double *__restrict__ y, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        y[i] = a * x[i] + y[i];
}

__global__ void k_fill(double *__restrict__ x, double v, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        x[i] = v;
}

// log-likelihood partials: sum_k log(max(floor, p[idx[k]])) (idx < 0 -> 0)
__global__ void k_loglik_idx(const double *__restrict__ p, const int *__restrict__ idx,
                             int64_t count, double floor_val, double *partials)
{
    __shared__ double red[2 * kBlock / kWave];
    dd s = dd_zero();
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < count;
         k += (int64_t)gridDim.x * kBlock) {
        const int i = idx[k];
        const double v = i >= 0 ? p[i] : 0.0;
        s = dd_add(s, log(v > floor_val ? v : floor_val));
    }
    put_partial(partials, s, red);
}

__global__ void k_loglik_w(const double *__restrict__ q, const double *__restrict__ w,
                           int64_t count, double floor_val, double *partials)
{
    __shared__ double red[2 * kBlock / kWave];
    dd s = dd_zero();
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < count;
         k += (int64_t)gridDim.x * kBlock) {
        const double v = q[k];
        if (w[k] != 0.0) s = dd_add_prod(s, w[k], log(v > floor_val ? v : floor_val));
    }
    put_partial(partials, s, red);
}

struct MargGeom {
    int nk, nd;
    int krad[kMaxSpecies], drad[kMaxSpecies];
    int64_t kstride[kMaxSpecies], dstride[kMaxSpecies];
    int64_t n_out, n_drop;
};

// One output cell per thread; the dropped coordinates are summed in
// ascending index order (deterministic, no atomics).
__global__ void k_marginal(const double *__restrict__ p, MargGeom g, double *__restrict__ out)
{
    const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (o >= g.n_out) return;
    int64_t base = 0, q = o;
    for (int k = 0; k < g.nk; ++k) {
        base += (q % g.krad[k]) * g.kstride[k];
        q /= g.krad[k];
    }
    int y[kMaxSpecies] = {0};
    int64_t off = 0;
    double s = 0.0;
    for (int64_t c = 0; c < g.n_drop; ++c) {
        s += p[base + off];
        for (int k = 0; k < g.nd; ++k) {  // odometer over the dropped species
            if (++y[k] < g.drad[k]) {
                off += g.dstride[k];
                break;
            }
            off -= (int64_t)(g.drad[k] - 1) * g.dstride[k];
            y[k] = 0;
        }
    }
    out[o] = s;
}

// tmp holds G dd partials (2G doubles) and one more double for the result.
fsp_status reduce_partials(const double *tmp, int G, double *h_out, hipStream_t s)
{
    k_sum_final<<<1, kBlock, 0, s>>>(tmp, G, const_cast<double *>(tmp) + 2 * G);
    FSP_HIP(hipGetLastError());
    FSP_HIP(hipMemcpyAsync(h_out, tmp + 2 * G, sizeof(double), hipMemcpyDeviceToHost, s));
    FSP_HIP(hipStreamSynchronize(s));
    return FSP_OK;
}

// Geometry-only model (bounds) for the state-space calls.
fsp_status geom_model(int d, const int *bounds, HostModel &hm)
{
    if (d < 1 || d > kMaxSpecies) return fail(FSP_ERR_UNSUPPORTED, "n_species must be in 1..8");
    if (!bounds) return fail(FSP_ERR_INVALID, "bounds is NULL");
    std::vector<int> sv(1, 1), sc(1, 0), sr = {0, 0};
    std::vector<double> f;
    int stride = 0;
    for (int s = 0; s < d; ++s) {
        if (bounds[s] < 0) return fail(FSP_ERR_INVALID, "negative bound");
        stride += bounds[s] + 1;
    }
    f.assign((size_t)stride, 1.0);
    fsp_model m{d, 1, bounds, sv.data(), sc.data(), sr.data(), f.data(), nullptr};
    return build_host_model(&m, hm);
}

}  // namespace

fsp_status vec_sum(const double *d_x, int64_t n, double *h_sum, hipStream_t s)
{
    const int G = grid_for((n + kBlock - 1) / kBlock, 1024);
    double *tmp = nullptr;
    FSP_HIP(hipMalloc(&tmp, sizeof(double) * (size_t)(2 * G + 1)));
    k_sum_partials<<<G, kBlock, 0, s>>>(d_x, n, tmp);
    FSP_HIP(hipGetLastError());
    fsp_status st = reduce_partials(tmp, G, h_sum, s);
    (void)hipFree(tmp);
    return st;
}

}  // namespace fsp

using namespace fsp;

struct fsp_mat_s {
    MatImpl impl;
};

extern "C" {

const char *fsp_last_error(void) { return g_last_error.c_str(); }

int fsp_version(void) { return 1; }

fsp_status fsp_device_count(int *count)
{
    if (!count) return fail(FSP_ERR_INVALID, "count is NULL");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    *count = c;
    return FSP_OK;
}

fsp_status fsp_malloc(void **d_ptr, size_t bytes)
{
    if (!d_ptr) return fail(FSP_ERR_INVALID, "d_ptr is NULL");
    hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(FSP_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return FSP_OK;
}

fsp_status fsp_free(void *d_ptr)
{
    if (d_ptr) FSP_HIP(hipFree(d_ptr));
    return FSP_OK;
}

fsp_status fsp_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes)
{
    FSP_HIP(copy_sync(d_dst, h_src, bytes, hipMemcpyHostToDevice, nullptr));
    return FSP_OK;
}

fsp_status fsp_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes)
{
    FSP_HIP(copy_sync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, nullptr));
    return FSP_OK;
}

fsp_status fsp_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, void *stream)
{
    FSP_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return FSP_OK;
}

fsp_status fsp_memset_zero(void *d_ptr, size_t bytes, void *stream)
{
    FSP_HIP(hipMemsetAsync(d_ptr, 0, bytes, (hipStream_t)stream));
    return FSP_OK;
}

fsp_status fsp_stream_sync(void *stream)
{
    FSP_HIP(hipStreamSynchronize((hipStream_t)stream));
    return FSP_OK;
}

fsp_status fsp_vec_sum(const double *d_x, int64_t n, double *h_sum, void *stream)
{
    if (!d_x || !h_sum || n < 0) return fail(FSP_ERR_INVALID, "bad argument");
    if (n == 0) {
        *h_sum = 0.0;
        return FSP_OK;
    }
    return vec_sum(d_x, n, h_sum, (hipStream_t)stream);
}

fsp_status fsp_vec_axpy(double a, const double *d_x, double *d_y, int64_t n, void *stream)
{
    if (n <= 0) return FSP_OK;
    const int G = grid_for((n + kBlock - 1) / kBlock, 4096);
    k_axpy<<<G, kBlock, 0, (hipStream_t)stream>>>(a, d_x, d_y, n);
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}

fsp_status fsp_vec_fill(double *d_x, double value, int64_t n, void *stream)
{
    if (n <= 0) return FSP_OK;
    const int G = grid_for((n + kBlock - 1) / kBlock, 4096);
    k_fill<<<G, kBlock, 0, (hipStream_t)stream>>>(d_x, value, n);
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}

fsp_status fsp_num_states(int n_species, const int *bounds, int64_t *n_states)
{
    if (!n_states) return fail(FSP_ERR_INVALID, "n_states is NULL");
    if (!bounds || n_species < 1) return fail(FSP_ERR_INVALID, "bad bounds");
    int64_t n = 1;
    for (int s = 0; s < n_species; ++s) {
        if (bounds[s] < 0) return fail(FSP_ERR_INVALID, "negative bound");
        n *= (int64_t)bounds[s] + 1;
        if (n > 2147483647LL) return fail(FSP_ERR_OVERFLOW, "more than 2^31-1 states");
    }
    *n_states = n;
    return FSP_OK;
}

fsp_status fsp_states_enumerate(int n_species, const int *bounds, int *d_states, void *stream)
{
    HostModel hm;
    FSP_TRY(geom_model(n_species, bounds, hm));
    if (!d_states) return fail(FSP_ERR_INVALID, "d_states is NULL");
    DevModel *dm = nullptr;
    FSP_HIP(hipMalloc(&dm, sizeof(DevModel)));
    FSP_HIP(copy_sync(dm, &hm.dm, sizeof(DevModel), hipMemcpyHostToDevice, (hipStream_t)stream));
    const uint32_t n = (uint32_t)hm.dm.n;
    const unsigned blocks = (n + kBlock - 1) / kBlock;
    hipStream_t s = (hipStream_t)stream;
    FSP_SWITCH_D(n_species, (k_enumerate<D_><<<blocks, kBlock, 0, s>>>(dm, n, d_states)));
    FSP_HIP(hipGetLastError());
    FSP_HIP(hipStreamSynchronize(s));
    (void)hipFree(dm);
    return FSP_OK;
}

fsp_status fsp_states_to_index(int n_species, const int *bounds, const int *d_states,
                               int64_t count, int *d_index, void *stream)
{
    HostModel hm;
    FSP_TRY(geom_model(n_species, bounds, hm));
    if (count < 0 || (count > 0 && (!d_states || !d_index)))
        return fail(FSP_ERR_INVALID, "bad argument");
    if (count == 0) return FSP_OK;
    DevModel *dm = nullptr;
    FSP_HIP(hipMalloc(&dm, sizeof(DevModel)));
    FSP_HIP(copy_sync(dm, &hm.dm, sizeof(DevModel), hipMemcpyHostToDevice, (hipStream_t)stream));
    hipStream_t s = (hipStream_t)stream;
    k_to_index<<<(unsigned)((count + kBlock - 1) / kBlock), kBlock, 0, s>>>(dm, d_states, count,
                                                                           d_index);
    FSP_HIP(hipGetLastError());
    FSP_HIP(hipStreamSynchronize(s));
    (void)hipFree(dm);
    return FSP_OK;
}

__attribute__((visibility("hidden"))) fsp_status fsp_mat_internal(fsp_mat A, MatImpl **impl)
{
    if (!A || !impl) return fail(FSP_ERR_INVALID, "NULL matrix");
    *impl = &A->impl;
    return FSP_OK;
}

static fsp_status mat_create(const fsp_model *model, fsp_format format, fsp_precision precision,
                             int with_sink, const double *coefs, void *stream, int transposed,
                             fsp_mat *out)
{
    if (!out) return fail(FSP_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (format != FSP_CSR && format != FSP_HYB && format != FSP_KRON)
        return fail(FSP_ERR_UNSUPPORTED, "requested format is not supported");
    if (precision != FSP_F64 && precision != FSP_MIXED)
        return fail(FSP_ERR_UNSUPPORTED, "requested precision is not supported");
    if (!coefs) return fail(FSP_ERR_INVALID, "coefs is NULL");
    auto *h = new fsp_mat_s();
    MatImpl &M = h->impl;
    fsp_status st = build_host_model(model, M.hm);
    if (st != FSP_OK) {
        delete h;
        return st;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        delete h;
        return fail(FSP_ERR_NO_DEVICE, "no HIP device visible");
    }
    M.format = format;
    M.prec = precision;
    M.with_sink = with_sink ? 1 : 0;
    M.transposed = transposed;
    M.coefs.assign(coefs, coefs + M.hm.dm.R);
    st = mat_assemble(M, (hipStream_t)stream);
    if (st != FSP_OK) {
        mat_free(M);
        delete h;
        return st;
    }
    *out = h;
    return FSP_OK;
}

fsp_status fsp_mat_create(const fsp_model *model, fsp_format format, fsp_precision precision,
                          int with_sink, const double *coefs, void *stream, fsp_mat *out)
{
    return mat_create(model, format, precision, with_sink, coefs, stream, 0, out);
}

__attribute__((visibility("hidden"))) fsp_status fsp_mat_create_transposed(
    const fsp_model *model, const double *coefs, void *stream, fsp_mat *out)
{
    return mat_create(model, FSP_HYB, FSP_F64, 0, coefs, stream, 1, out);
}

fsp_status fsp_mat_destroy(fsp_mat A)
{
    if (!A) return FSP_OK;
    mat_free(A->impl);
    delete A;
    return FSP_OK;
}

fsp_status fsp_mat_get_info(fsp_mat A, fsp_mat_info *info)
{
    if (!A || !info) return fail(FSP_ERR_INVALID, "NULL argument");
    const MatImpl &M = A->impl;
    info->n_states = M.n;
    info->n_rows = M.n_rows;
    info->nnz = M.nnz_rows + M.nnz_sink;
    info->n_slots = M.hm.dm.U;
    info->format = M.format;
    info->precision = M.prec;
    info->with_sink = M.with_sink;
    info->max_row = M.max_row;
    info->bytes_per_apply = mat_bytes_per_apply(M);
    return FSP_OK;
}

fsp_status fsp_mat_set_coefs(fsp_mat A, const double *h_coefs, void *stream)
{
    if (!A || !h_coefs) return fail(FSP_ERR_INVALID, "NULL argument");
    return mat_revalue(A->impl, h_coefs, (hipStream_t)stream);
}

fsp_status fsp_mat_apply(fsp_mat A, const double *h_coefs, const double *d_x, double *d_y,
                         int accumulate, void *stream)
{
    if (!A || !d_x || !d_y) return fail(FSP_ERR_INVALID, "NULL argument");
    if (h_coefs) FSP_TRY(mat_revalue(A->impl, h_coefs, (hipStream_t)stream));
    return mat_apply(A->impl, d_x, d_y, accumulate, (hipStream_t)stream);
}

fsp_status fsp_mat_apply_transpose(fsp_mat A, const double *h_coefs, const double *d_x,
                                   double *d_y, int accumulate, void *stream)
{
    if (!A || !d_x || !d_y) return fail(FSP_ERR_INVALID, "NULL argument");
    if (h_coefs) FSP_TRY(mat_revalue(A->impl, h_coefs, (hipStream_t)stream));
    return mat_apply_transpose(A->impl, d_x, d_y, accumulate, (hipStream_t)stream);
}

fsp_status fsp_mat_apply_term(fsp_mat A, int reaction, double c, const double *d_x, double *d_y,
                              int accumulate, void *stream)
{
    if (!A || !d_x || !d_y) return fail(FSP_ERR_INVALID, "NULL argument");
    return mat_apply_term(A->impl, reaction, c, d_x, d_y, accumulate, (hipStream_t)stream);
}

fsp_status fsp_mat_export_csr(fsp_mat A, int *d_row_ptrs, int *d_cols, double *d_vals,
                              void *stream)
{
    if (!A) return fail(FSP_ERR_INVALID, "NULL matrix");
    const MatImpl &M = A->impl;
    if (M.format != FSP_CSR || M.prec != FSP_F64)
        return fail(FSP_ERR_UNSUPPORTED, "export_csr needs an FP64 CSR matrix");
    hipStream_t s = (hipStream_t)stream;
    // rows 0..n-1 as stored; the sink row (if any) appended as row n
    if (d_row_ptrs) {
        FSP_HIP(hipMemcpyAsync(d_row_ptrs, M.rp, sizeof(int) * (size_t)(M.n + 1),
                               hipMemcpyDeviceToDevice, s));
        if (M.with_sink) {
            const int end = (int)(M.nnz_rows + M.nnz_sink);
            FSP_HIP(hipStreamSynchronize(s));
            FSP_HIP(copy_sync(d_row_ptrs + M.n + 1, &end, sizeof(int), hipMemcpyHostToDevice, s));
        }
    }
    if (d_cols) {
        FSP_HIP(hipMemcpyAsync(d_cols, M.cols, sizeof(int) * (size_t)M.nnz_rows,
                               hipMemcpyDeviceToDevice, s));
        if (M.with_sink && M.nnz_sink)
            FSP_HIP(hipMemcpyAsync(d_cols + M.nnz_rows, M.sink_cols, sizeof(int) * (size_t)M.nnz_sink,
                                   hipMemcpyDeviceToDevice, s));
    }
    if (d_vals) {
        FSP_HIP(hipMemcpyAsync(d_vals, M.vals, sizeof(double) * (size_t)M.nnz_rows,
                               hipMemcpyDeviceToDevice, s));
        if (M.with_sink && M.nnz_sink)
            FSP_HIP(hipMemcpyAsync(d_vals + M.nnz_rows, M.sink_vals,
                                   sizeof(double) * (size_t)M.nnz_sink, hipMemcpyDeviceToDevice, s));
    }
    FSP_HIP(hipStreamSynchronize(s));
    return FSP_OK;
}

fsp_status fsp_mat_export_hyb(fsp_mat A, double *d_diag, double *d_slot_vals, int *d_slot_cols,
                              double *d_sink_out, void *stream)
{
    if (!A) return fail(FSP_ERR_INVALID, "NULL matrix");
    const MatImpl &M = A->impl;
    if (M.format != FSP_HYB || M.prec != FSP_F64)
        return fail(FSP_ERR_UNSUPPORTED, "export_hyb needs an FP64 HYB matrix");
    hipStream_t s = (hipStream_t)stream;
    const size_t slots = (size_t)M.hm.dm.U * (size_t)M.n;
    if (d_diag)
        FSP_HIP(hipMemcpyAsync(d_diag, M.diag, sizeof(double) * (size_t)M.n, hipMemcpyDeviceToDevice, s));
    if (d_slot_vals && slots)
        FSP_HIP(hipMemcpyAsync(d_slot_vals, M.slot_vals, sizeof(double) * slots, hipMemcpyDeviceToDevice, s));
    if (d_slot_cols && slots) FSP_TRY(mat_hyb_cols(M, d_slot_cols, s));
    if (d_sink_out && M.with_sink)
        FSP_HIP(hipMemcpyAsync(d_sink_out, M.sink_out, sizeof(double) * (size_t)M.n, hipMemcpyDeviceToDevice, s));
    FSP_HIP(hipStreamSynchronize(s));
    return FSP_OK;
}

fsp_status fsp_loglik_indices(const double *d_p, const int *d_idx, int64_t count,
                              double floor_val, double *h_ll, void *stream)
{
    if (!h_ll || count < 0 || (count > 0 && (!d_p || !d_idx)))
        return fail(FSP_ERR_INVALID, "bad argument");
    if (count == 0) {
        *h_ll = 0.0;
        return FSP_OK;
    }
    hipStream_t s = (hipStream_t)stream;
    const int G = grid_for((count + kBlock - 1) / kBlock, 1024);
    double *tmp = nullptr;
    FSP_HIP(hipMalloc(&tmp, sizeof(double) * (size_t)(2 * G + 1)));
    k_loglik_idx<<<G, kBlock, 0, s>>>(d_p, d_idx, count, floor_val, tmp);
    FSP_HIP(hipGetLastError());
    fsp_status st = reduce_partials(tmp, G, h_ll, s);
    (void)hipFree(tmp);
    return st;
}

}  // extern "C"

// ------------------------------------------------ allocation-free variant
// Scratch of fsp_loglik_indices_noalloc: the G <= kNoAllocGrid dd partials
// (2 G doubles) and the result slot.  File-scope device memory: each GPU
// the library runs on holds its own copy (16 KiB) until the process ends.
// Its address on each device comes from hipGetSymbolAddress: on the host
// the array's name is not a device address.
constexpr int kNoAllocGrid = 1024;
constexpr int kNoAllocScratch = 2 * kNoAllocGrid + 1;
static __device__ __align__(256) double scratch[kNoAllocScratch];

namespace {
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the
// first call there.
fsp_status scratch_on_current_device(double **out)
{
    int dev = 0;
    FSP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(FSP_ERR_UNSUPPORTED, "device index >= 64");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(FSP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return FSP_OK;
}
}  // namespace

extern "C" {

fsp_status fsp_loglik_indices_noalloc(const double *d_p, const int *d_idx, int64_t count,
                                      double floor_val, double *h_ll, void *stream)
{
    if (!h_ll || count < 0 || (count > 0 && (!d_p || !d_idx)))
        return fail(FSP_ERR_INVALID, "bad argument");
    if (count == 0) {
        *h_ll = 0.0;
        return FSP_OK;
    }
    hipStream_t s = (hipStream_t)stream;
    // the same grid rule as fsp_loglik_indices, so both reduce their
    // partials in the same order and agree bit for bit
    const int G = grid_for((count + kBlock - 1) / kBlock, 1024);
    if (2 * (int64_t)G + 1 > kNoAllocScratch)
        return fail(FSP_ERR_UNSUPPORTED, "fsp_loglik_indices_noalloc: more than 1024 partial "
                                         "sums do not fit the 16 KiB scratch");
    double *tmp = nullptr;
    FSP_TRY(scratch_on_current_device(&tmp));
    // every partial slot 0..2G-1 is written by its block before k_sum_final
    // reads it, and k_sum_final writes slot 2G: nothing needs clearing
    k_loglik_idx<<<G, kBlock, 0, s>>>(d_p, d_idx, count, floor_val, tmp);
    FSP_HIP(hipGetLastError());
    return reduce_partials(tmp, G, h_ll, s);
}

fsp_status fsp_loglik_weighted(const double *d_q, const double *d_w, int64_t count,
                               double floor_val, double *h_ll, void *stream)
{
    if (!h_ll || count < 0 || (count > 0 && (!d_q || !d_w)))
        return fail(FSP_ERR_INVALID, "bad argument");
    if (count == 0) {
        *h_ll = 0.0;
        return FSP_OK;
    }
    hipStream_t s = (hipStream_t)stream;
    const int G = grid_for((count + kBlock - 1) / kBlock, 1024);
    double *tmp = nullptr;
    FSP_HIP(hipMalloc(&tmp, sizeof(double) * (size_t)(2 * G + 1)));
    k_loglik_w<<<G, kBlock, 0, s>>>(d_q, d_w, count, floor_val, tmp);
    FSP_HIP(hipGetLastError());
    fsp_status st = reduce_partials(tmp, G, h_ll, s);
    (void)hipFree(tmp);
    return st;
}

fsp_status fsp_marginal(const double *d_p, int n_species, const int *bounds, const int *keep_mask,
                        double *d_out, void *stream)
{
    if (!d_p || !d_out || !bounds || !keep_mask) return fail(FSP_ERR_INVALID, "bad argument");
    if (n_species < 1 || n_species > kMaxSpecies) return fail(FSP_ERR_UNSUPPORTED, "n_species");
    MargGeom g;
    std::memset(&g, 0, sizeof(g));
    int64_t stride = 1;
    g.n_out = 1;
    g.n_drop = 1;
    for (int s = 0; s < n_species; ++s) {
        if (bounds[s] < 0) return fail(FSP_ERR_INVALID, "negative bound");
        const int rad = bounds[s] + 1;
        if (keep_mask[s]) {
            g.krad[g.nk] = rad;
            g.kstride[g.nk++] = stride;
            g.n_out *= rad;
        } else {
            g.drad[g.nd] = rad;
            g.dstride[g.nd++] = stride;
            g.n_drop *= rad;
        }
        stride *= rad;
        if (stride > 2147483647LL) return fail(FSP_ERR_OVERFLOW, "more than 2^31-1 states");
    }
    hipStream_t s = (hipStream_t)stream;
    k_marginal<<<(unsigned)((g.n_out + kBlock - 1) / kBlock), kBlock, 0, s>>>(d_p, g, d_out);
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}

fsp_status fsp_term_csr(const fsp_model *model, int reaction, int *d_row_ptrs, int *d_cols,
                        double *d_vals, int64_t *h_nnz, void *stream)
{
    if (!h_nnz) return fail(FSP_ERR_INVALID, "h_nnz is NULL");
    HostModel hm;
    FSP_TRY(build_host_model(model, hm));
    return term_csr(hm, reaction, d_row_ptrs, d_cols, d_vals, h_nnz, (hipStream_t)stream);
}

}  // extern "C"
