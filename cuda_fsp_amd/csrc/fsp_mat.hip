// fsp_mat.hip -- generator assembly and apply (CSR, HYB, KRON) for gfx950.
//
// Replaces generate_fsp_mats_cuda_csr / fspmat_csr_* (src/fspmat_csr_kernels.cu),
// generate_fsp_mats_hyb / fspmat_hyb_* (src/fspmat_hyb_kernels.cu) and
// generate_fsp_mats_sdkron / fspmat_sdkron_mv (src/fspmat_kron_kernels.cu).
// All assembly runs on the device: one fused count kernel, an own int scan,
// and one fill kernel per format, instead of the reference's per-reaction
// launch + Thrust scan + host copy loop.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <type_traits>

#include "fsp_assemble.h"
#include "fsp_formats.h"

namespace fsp {

// ====================================================================== scan
namespace {
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__global__ void k_scan_tiles(const int *__restrict__ cnt, int64_t n, long long *tile_sums)
{
    __shared__ double red[kBlock / kWave];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    long long s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + (int64_t)k * kBlock + threadIdx.x;
        if (i < n) s += cnt[i];
    }
    // exact in double for counts < 2^53
    const double t = block_sum((double)s, red);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = (long long)t;
}

__global__ void k_scan_top(long long *tile_sums, int64_t tiles, long long *total)
{
    // single workgroup: exclusive scan in place, chunks of kBlock
    __shared__ long long buf[kBlock];
    long long carry = 0;
    for (int64_t c0 = 0; c0 < tiles; c0 += kBlock) {
        const int64_t i = c0 + threadIdx.x;
        long long v = (i < tiles) ? tile_sums[i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < kBlock; off <<= 1) {
            long long a = (threadIdx.x >= (unsigned)off) ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += a;
            __syncthreads();
        }
        if (i < tiles) tile_sums[i] = carry + buf[threadIdx.x] - v;
        carry += buf[kBlock - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_apply(const int *__restrict__ cnt, int64_t n,
                             const long long *__restrict__ tile_off, int *out)
{
    __shared__ long long buf[kBlock];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    long long loc[kScanItems];
    long long s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k;
        s += (i < n) ? cnt[i] : 0;
        loc[k] = s;
    }
    buf[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        long long a = (threadIdx.x >= (unsigned)off) ? buf[threadIdx.x - off] : 0;
        __syncthreads();
        buf[threadIdx.x] += a;
        __syncthreads();
    }
    const long long pre = tile_off[blockIdx.x] + buf[threadIdx.x] - s;
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k;
        if (i < n) out[i + 1] = (int)(pre + loc[k]);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
}
}  // namespace

fsp_status scan_counts(const int *d_counts, int *d_out, int64_t n, int64_t *h_total,
                       hipStream_t s)
{
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    long long *tmp = nullptr;
    FSP_HIP(hipMalloc(&tmp, sizeof(long long) * (size_t)(tiles + 1)));
    if (tiles > 0) {
        k_scan_tiles<<<(unsigned)tiles, kBlock, 0, s>>>(d_counts, n, tmp);
        k_scan_top<<<1, kBlock, 0, s>>>(tmp, tiles, tmp + tiles);
        k_scan_apply<<<(unsigned)tiles, kBlock, 0, s>>>(d_counts, n, tmp, d_out);
    } else {
        FSP_HIP(hipMemsetAsync(d_out, 0, sizeof(int), s));
        FSP_HIP(hipMemsetAsync(tmp, 0, sizeof(long long), s));
    }
    long long total = 0;
    FSP_HIP(hipMemcpyAsync(&total, tmp + tiles, sizeof(long long), hipMemcpyDeviceToHost, s));
    FSP_HIP(hipStreamSynchronize(s));
    FSP_HIP(hipGetLastError());
    (void)hipFree(tmp);
    *h_total = total;
    if (total > 2147483647LL) return fail(FSP_ERR_OVERFLOW, "more than 2^31-1 nonzeros");
    return FSP_OK;
}

// ================================================================ assembly
namespace {

// Stage the compact FP64 factor table in LDS when it fits, else read global.
__device__ inline const double *table_ptr(const DevModel *__restrict__ dm,
                                          const double *__restrict__ g, double *lds, int use_lds)
{
    if (!g) return nullptr;
    if (use_lds) {
        stage_table(lds, g, dm->compact_size);
        return lds;
    }
    return g;
}

// Merged CSR: per-row entry count and sink flag.
template <int D>
__global__ void k_count_merged(const DevModel *__restrict__ dm, uint32_t n, int diag_in_row,
                               int *__restrict__ counts, int *__restrict__ sink_cnt)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    int xs[D];
    decode<D>(i, dm, xs);
    int c = diag_in_row;
    for (int u = 0; u < dm->U; ++u)
        c += source_in_box<D>(dm->rx[dm->slot_rxn[dm->slot_ptr[u]]], xs, dm) ? 1 : 0;
    counts[i] = c;
    if (sink_cnt) {
        int has = 0;
        for (int r = 0; r < dm->R; ++r) has |= target_leaves<D>(dm->rx[r], xs, dm) ? 1 : 0;
        sink_cnt[i] = has;
    }
}

template <int D, typename V>
__global__ void k_fill_merged(const DevModel *__restrict__ dm, const double *__restrict__ tab_g,
                              int use_lds, const double *__restrict__ table,
                              const double *__restrict__ coefs, uint32_t n, int mixed,
                              const int *__restrict__ rp, int *__restrict__ cols,
                              V *__restrict__ vals, double *__restrict__ diag_out,
                              const int *__restrict__ sink_off, int *__restrict__ sink_cols,
                              double *__restrict__ sink_vals)
{
    extern __shared__ double smem[];
    const double *tab = table_ptr(dm, tab_g, smem, use_lds);
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    int xs[D];
    decode<D>(i, dm, xs);
    double dg, out;
    int has;
    diag_and_out<D>(dm, tab, table, coefs, mixed, i, xs, dg, out, has);
    int pos = rp[i];
    for (int k = 0; k <= dm->U; ++k) {
        const int u = dm->csr_order[k];
        if (u < 0) {
            if (!mixed) {
                cols[pos] = (int)i;
                vals[pos] = (V)dg;
                ++pos;
            } else {
                diag_out[i] = dg;
            }
            continue;
        }
        if (!source_in_box<D>(dm->rx[dm->slot_rxn[dm->slot_ptr[u]]], xs, dm)) continue;
        const double v = slot_value<D>(dm, tab, table, coefs, u, i, xs, 1);
        cols[pos] = (int)((int64_t)i - dm->slot_delta[u]);
        vals[pos] = (V)v;
        ++pos;
    }
    if (sink_off && has) {
        const int p = sink_off[i];
        sink_cols[p] = (int)i;
        sink_vals[p] = out;
    }
}

template <int D, typename V>
__global__ void k_fill_hyb(const DevModel *__restrict__ dm, const double *__restrict__ tab_g,
                           int use_lds, const double *__restrict__ table,
                           const double *__restrict__ coefs, uint32_t n, int mixed,
                           double *__restrict__ diag, V *__restrict__ slot_vals,
                           double *__restrict__ sink_out, int transposed)
{
    extern __shared__ double smem[];
    const double *tab = table_ptr(dm, tab_g, smem, use_lds);
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    int xs[D];
    decode<D>(i, dm, xs);
    double dg, out;
    int has;
    diag_and_out<D>(dm, tab, table, coefs, mixed, i, xs, dg, out, has);
    diag[i] = dg;
    if (sink_out) sink_out[i] = out;
    for (int u = 0; u < dm->U; ++u) {
        const size_t o = (size_t)u * n + i;
        const RxnDev &rx = dm->rx[dm->slot_rxn[dm->slot_ptr[u]]];
        if (transposed) {
            // row i of A^T: the outflow rate of state i towards its target
            const bool in = !target_leaves<D>(rx, xs, dm);
            slot_vals[o] = (V)(in ? slot_value<D>(dm, tab, table, coefs, u, i, xs, 0) : 0.0);
        } else {
            const bool in = source_in_box<D>(rx, xs, dm);
            slot_vals[o] = (V)(in ? slot_value<D>(dm, tab, table, coefs, u, i, xs, 1) : 0.0);
        }
    }
}

// Explicit HYB columns (for export): the inflow source i - delta_u when it
// lies in the box, else i.
template <int D>
__global__ void k_hyb_cols(const DevModel *__restrict__ dm, uint32_t n, int *__restrict__ cols)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    int xs[D];
    decode<D>(i, dm, xs);
    for (int u = 0; u < dm->U; ++u) {
        const RxnDev &rx = dm->rx[dm->slot_rxn[dm->slot_ptr[u]]];
        cols[(size_t)u * n + i] =
            source_in_box<D>(rx, xs, dm) ? (int)((int64_t)i - dm->slot_delta[u]) : (int)i;
    }
}

// Reference-layout per-reaction CSR (src/fspmat_csr_kernels.cu:7-94).
template <int D>
__global__ void k_term_count(const DevModel *__restrict__ dm, int r, uint32_t n, int *counts)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    int xs[D];
    decode<D>(i, dm, xs);
    counts[i] = 1 + (source_in_box<D>(dm->rx[r], xs, dm) ? 1 : 0);
}

template <int D>
__global__ void k_term_fill(const DevModel *__restrict__ dm, const double *__restrict__ tab_g,
                            int use_lds, const double *__restrict__ table, int r, uint32_t n,
                            const int *__restrict__ rp, int *__restrict__ cols,
                            double *__restrict__ vals)
{
    extern __shared__ double smem[];
    const double *tab = table_ptr(dm, tab_g, smem, use_lds);
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    int xs[D];
    decode<D>(i, dm, xs);
    const int p = rp[i];
    const double a_here = prop_at<D>(dm, tab, table, r, i, xs, 0);
    if (source_in_box<D>(dm->rx[r], xs, dm)) {
        const int64_t j = (int64_t)i - dm->rx[r].delta;
        const int idg = (j > (int64_t)i) ? p : p + 1;
        const int iof = (j > (int64_t)i) ? p + 1 : p;
        cols[idg] = (int)i;
        cols[iof] = (int)j;
        if (vals) {
            vals[idg] = -1.0 * a_here;
            vals[iof] = prop_at<D>(dm, tab, table, r, i, xs, 1);
        }
    } else {
        cols[p] = (int)i;
        if (vals) vals[p] = -1.0 * a_here;
    }
}

}  // namespace

static fsp_status upload_model(MatImpl &M, hipStream_t s)
{
    FSP_HIP(hipMalloc(&M.d_model, sizeof(DevModel)));
    FSP_HIP(copy_sync(M.d_model, &M.hm.dm, sizeof(DevModel), hipMemcpyHostToDevice, s));
    if (!M.hm.compact.empty()) {
        FSP_HIP(hipMalloc(&M.d_compact, sizeof(double) * M.hm.compact.size()));
        FSP_HIP(copy_sync(M.d_compact, M.hm.compact.data(), sizeof(double) * M.hm.compact.size(),
                          hipMemcpyHostToDevice, s));
    }
    FSP_HIP(hipMalloc(&M.d_coefs, sizeof(double) * (size_t)M.hm.dm.R));
    FSP_HIP(copy_sync(M.d_coefs, M.coefs.data(), sizeof(double) * (size_t)M.hm.dm.R,
                      hipMemcpyHostToDevice, s));
    return FSP_OK;
}

static size_t table_lds_bytes(const HostModel &hm, int &use_lds)
{
    const size_t b = sizeof(double) * hm.compact.size();
    use_lds = (!hm.compact.empty() && b <= kLdsTableMax) ? 1 : 0;
    return use_lds ? b : 0;
}

static fsp_status fill_values(MatImpl &M, hipStream_t s)
{
    const DevModel &dm = M.hm.dm;
    const uint32_t n = (uint32_t)M.n;
    const unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
    int use_lds = 0;
    const size_t lds = table_lds_bytes(M.hm, use_lds);
    const double *tab = M.d_compact;
    const double *table = M.hm.prop_table_dev;
    const int mixed = M.prec == FSP_MIXED;
    FSP_HIP(hipMemcpyAsync(M.d_coefs, M.coefs.data(), sizeof(double) * (size_t)dm.R,
                           hipMemcpyHostToDevice, s));
    if (M.format == FSP_CSR) {
        if (mixed) {
            FSP_SWITCH_D(dm.d, (k_fill_merged<D_, float><<<blocks, kBlock, lds, s>>>(
                                   M.d_model, tab, use_lds, table, M.d_coefs, n, 1, M.rp,
                                   M.cols, (float *)M.vals, (double *)M.diag,
                                   M.with_sink ? M.sink_block_ptr : nullptr, M.sink_cols,
                                   (double *)M.sink_vals)));
        } else {
            FSP_SWITCH_D(dm.d, (k_fill_merged<D_, double><<<blocks, kBlock, lds, s>>>(
                                   M.d_model, tab, use_lds, table, M.d_coefs, n, 0, M.rp,
                                   M.cols, (double *)M.vals, nullptr,
                                   M.with_sink ? M.sink_block_ptr : nullptr, M.sink_cols,
                                   (double *)M.sink_vals)));
        }
    } else if (M.format == FSP_HYB) {
        if (mixed) {
            FSP_SWITCH_D(dm.d, (k_fill_hyb<D_, float><<<blocks, kBlock, lds, s>>>(
                                   M.d_model, tab, use_lds, table, M.d_coefs, n, 1,
                                   (double *)M.diag, (float *)M.slot_vals,
                                   (double *)M.sink_out, 0)));
        } else {
            FSP_SWITCH_D(dm.d, (k_fill_hyb<D_, double><<<blocks, kBlock, lds, s>>>(
                                   M.d_model, tab, use_lds, table, M.d_coefs, n, 0,
                                   (double *)M.diag, (double *)M.slot_vals,
                                   (double *)M.sink_out, M.transposed)));
        }
    }
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}

// KRON slot tables for the current coefficients (slot form only).
static fsp_status upload_slot_tables(MatImpl &M, hipStream_t s)
{
    KronSlotPlan P;
    if (!M.kron_slots || !kron_slot_plan(M.hm, P)) return FSP_OK;
    std::vector<double> tab;
    kron_slot_tables(M.hm, P, M.coefs.data(), tab);
    FSP_HIP(copy_sync(M.d_slot_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, s));
    if (M.d_slot_tab_f) {
        std::vector<float> f(tab.begin(), tab.end());
        FSP_HIP(copy_sync(M.d_slot_tab_f, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice, s));
    }
    return FSP_OK;
}

fsp_status mat_assemble(MatImpl &M, hipStream_t s)
{
    const DevModel &dm = M.hm.dm;
    M.n = dm.n;
    M.n_rows = dm.n + (M.with_sink ? 1 : 0);
    const uint32_t n = (uint32_t)M.n;
    const unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
    const int mixed = M.prec == FSP_MIXED;
    FSP_TRY(upload_model(M, s));
    const int64_t tiles = (M.n + kBlock - 1) / kBlock;
    M.n_partials = std::max<int64_t>(tiles, 1) * 64;
    FSP_HIP(hipMalloc(&M.partials, sizeof(double) * (size_t)M.n_partials));
    FSP_HIP(hipMalloc(&M.ticket, sizeof(unsigned int) * 4));
    FSP_HIP(fill_sync(M.ticket, 0, sizeof(unsigned int) * 4, s));
    if (M.format == FSP_CSR) {
        int *cnt = nullptr, *scnt = nullptr;
        FSP_HIP(hipMalloc(&cnt, sizeof(int) * (size_t)(M.n + 1)));
        FSP_HIP(hipMalloc(&M.rp, sizeof(int) * (size_t)(M.n + 1)));
        if (M.with_sink) {
            FSP_HIP(hipMalloc(&scnt, sizeof(int) * (size_t)(M.n + 1)));
            FSP_HIP(hipMalloc(&M.sink_block_ptr, sizeof(int) * (size_t)(M.n + 1)));
        }
        FSP_SWITCH_D(dm.d, (k_count_merged<D_><<<blocks, kBlock, 0, s>>>(M.d_model, n, mixed ? 0 : 1,
                                                                         cnt, scnt)));
        FSP_HIP(hipGetLastError());
        int64_t total = 0;
        FSP_TRY(scan_counts(cnt, M.rp, M.n, &total, s));
        M.nnz_rows = total;
        if (M.with_sink) {
            int64_t st = 0;
            FSP_TRY(scan_counts(scnt, M.sink_block_ptr, M.n, &st, s));
            M.nnz_sink = st;
            FSP_HIP(hipMalloc(&M.sink_cols, sizeof(int) * (size_t)std::max<int64_t>(st, 1)));
            FSP_HIP(hipMalloc(&M.sink_vals, sizeof(double) * (size_t)std::max<int64_t>(st, 1)));
        }
        // longest row (for the LDS product buffer); counts are still in cnt
        {
            std::vector<int> h((size_t)M.n);
            FSP_HIP(copy_sync(h.data(), cnt, sizeof(int) * (size_t)M.n, hipMemcpyDeviceToHost, s));
            M.max_row = M.n ? *std::max_element(h.begin(), h.end()) : 0;
        }
        (void)hipFree(cnt);
        if (scnt) (void)hipFree(scnt);
        const size_t vb = mixed ? sizeof(float) : sizeof(double);
        FSP_HIP(hipMalloc(&M.cols, sizeof(int) * (size_t)std::max<int64_t>(M.nnz_rows, 1)));
        FSP_HIP(hipMalloc(&M.vals, vb * (size_t)std::max<int64_t>(M.nnz_rows, 1)));
        if (mixed) FSP_HIP(hipMalloc(&M.diag, sizeof(double) * (size_t)M.n));
    } else if (M.format == FSP_HYB) {
        const size_t vb = mixed ? sizeof(float) : sizeof(double);
        const size_t slots = (size_t)std::max(dm.U, 1) * (size_t)M.n;
        FSP_HIP(hipMalloc(&M.diag, sizeof(double) * (size_t)M.n));
        FSP_HIP(hipMalloc(&M.slot_vals, vb * slots));
        std::vector<int> off((size_t)std::max(dm.U, 1), 0);
        for (int u = 0; u < dm.U; ++u) {
            const int64_t d = M.transposed ? dm.slot_delta[u] : -dm.slot_delta[u];
            off[u] = (int)std::max<int64_t>(-(int64_t)M.n, std::min<int64_t>((int64_t)M.n, d));
        }
        FSP_HIP(hipMalloc(&M.slot_off, sizeof(int) * off.size()));
        FSP_HIP(copy_sync(M.slot_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, s));
        if (M.with_sink) FSP_HIP(hipMalloc(&M.sink_out, sizeof(double) * (size_t)M.n));
        M.nnz_rows = (int64_t)M.n * (dm.U + 1);
    } else {
        if (!M.hm.has_factors && !M.hm.prop_table_dev)
            return fail(FSP_ERR_INVALID, "KRON needs factor tables or a propensity table");
        if (mixed && !M.hm.compact.empty()) {
            std::vector<float> f(M.hm.compact.begin(), M.hm.compact.end());
            FSP_HIP(hipMalloc(&M.d_compact_f, sizeof(float) * f.size()));
            FSP_HIP(copy_sync(M.d_compact_f, f.data(), sizeof(float) * f.size(),
                              hipMemcpyHostToDevice, s));
        }
        KronSlotPlan P;
        if (!M.hm.prop_table_dev && kron_slot_plan(M.hm, P) &&
            sizeof(double) * (size_t)P.tab_size <= kLdsTableMax) {
            M.kron_slots = 1;
            FSP_HIP(hipMalloc(&M.d_slot_tab, sizeof(double) * (size_t)P.tab_size));
            if (mixed) FSP_HIP(hipMalloc(&M.d_slot_tab_f, sizeof(float) * (size_t)P.tab_size));
            FSP_TRY(upload_slot_tables(M, s));
        }
        M.nnz_rows = 0;
    }
    FSP_TRY(fill_values(M, s));
    FSP_HIP(hipStreamSynchronize(s));
    return FSP_OK;
}

fsp_status mat_revalue(MatImpl &M, const double *coefs, hipStream_t s)
{
    bool same = true;
    for (int r = 0; r < M.hm.dm.R; ++r) same = same && (M.coefs[r] == coefs[r]);
    if (same) return FSP_OK;
    M.coefs.assign(coefs, coefs + M.hm.dm.R);
    if (M.format == FSP_KRON) {
        FSP_HIP(hipMemcpyAsync(M.d_coefs, M.coefs.data(), sizeof(double) * (size_t)M.hm.dm.R,
                               hipMemcpyHostToDevice, s));
        // the host copy must stay valid until the copy ran
        FSP_HIP(hipStreamSynchronize(s));
        return upload_slot_tables(M, s);
    }
    FSP_TRY(fill_values(M, s));
    FSP_HIP(hipStreamSynchronize(s));
    return FSP_OK;
}

// ==================================================================== apply
namespace {

struct PlainX {
    const double *__restrict__ x;
    __device__ double operator()(uint32_t c) const { return x[c]; }
};

// Sink element y[n] from per-block dd partials (order-independent sum).
__global__ void k_finalize_sink(const double *__restrict__ partials, int count, double *y,
                                int accumulate)
{
    __shared__ double red[2 * kBlock / kWave];
    const double s = dd_value(sum_dd_partials(partials, count, red));
    if (threadIdx.x == 0) *y = accumulate ? *y + s : s;
}

__device__ inline void put_dd_partial(double *partials, dd sink, double *red)
{
    const dd s = block_sum_dd(sink, red);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = s.s;
        partials[2 * blockIdx.x + 1] = s.c;
    }
}

// Merged CSR apply: one tile of kCsrApplyRows rows per workgroup (larger
// tiles amortise the LDS round trip of csr_tile), at most kCsrApplyPer
// products per thread per chunk, tiles renumbered per XCD (xcd_block).
// Sink partial of tile t at 2 t, so the finalizer sums them in tile order.
constexpr int kCsrApplyRows = 512;
constexpr int kCsrApplyPer = 12;

template <typename V>
__global__ void __launch_bounds__(kCsrApplyRows) k_apply_csr(CsrF<V> A, PlainX X,
                                                             double *__restrict__ y, int accumulate,
                                                             double *__restrict__ partials)
{
    constexpr int NW = kCsrApplyRows / kWave;
    extern __shared__ double smem[];
    double *prod = smem;
    double *red = smem + A.cap;
    dd sink = dd_zero();
    const uint32_t t = xcd_block(blockIdx.x, gridDim.x);
    const double v = csr_tile<V, PlainX, kCsrApplyRows, kCsrApplyPer>(A, t, X, prod, sink);
    const uint32_t i = t * kCsrApplyRows + threadIdx.x;
    if (i < A.n) y[i] = accumulate ? y[i] + v : v;
    if (A.sink_off) {
        // block sum of the sink dd over NW waves
        const dd ws = wave_sum_dd(sink);
        __syncthreads();
        if (threadIdx.x % kWave == 0) {
            red[2 * (threadIdx.x / kWave)] = ws.s;
            red[2 * (threadIdx.x / kWave) + 1] = ws.c;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            dd s = dd_zero();
            for (int q = 0; q < NW; ++q) s = dd_merge(s, dd{red[2 * q], red[2 * q + 1]});
            partials[2 * t] = s.s;
            partials[2 * t + 1] = s.c;
        }
    }
}

// NU > 0: the slot count as a compile-time constant (registers sized for
// it); NU = 0: any slot count.
template <int NU, typename V>
__global__ void __launch_bounds__(kBlock) k_apply_hyb(HybF<V> A, PlainX X, double *__restrict__ y,
                                                      int accumulate, uint32_t tiles,
                                                      double *__restrict__ partials)
{
    __shared__ double red[2 * kBlock / kWave];
    dd sink = dd_zero();
    for (TileWalk w = tile_walk(tiles); w.t < w.end; w.t += w.step) {
        const uint32_t t = w.t;
        const uint32_t i = t * kBlock + threadIdx.x;
        double v = 0.0;
        if (NU > 0) {
            if (i < A.n) v = hyb_row_n<(NU > 0 ? NU : 1)>(A, i, X, sink);
        } else {
            v = hyb_row<V>(A, i, X, sink);
        }
        if (i < A.n) y[i] = accumulate ? y[i] + v : v;
    }
    if (A.sink_out) put_dd_partial(partials, sink, red);
}

template <typename V>
static void launch_hyb(const HybF<V> &A, const double *d_x, double *d_y, int accumulate,
                       uint32_t tiles, double *partials, int grid, hipStream_t s)
{
#define FSP_HYB_CASE(K)                                                                   \
    case K:                                                                               \
        k_apply_hyb<K, V><<<grid, kBlock, 0, s>>>(A, PlainX{d_x}, d_y, accumulate, tiles, \
                                                  partials);                              \
        return;
    switch (A.U) {
        FSP_HYB_CASE(1) FSP_HYB_CASE(2) FSP_HYB_CASE(3) FSP_HYB_CASE(4)
        FSP_HYB_CASE(5) FSP_HYB_CASE(6) FSP_HYB_CASE(7) FSP_HYB_CASE(8)
        FSP_HYB_CASE(9) FSP_HYB_CASE(10) FSP_HYB_CASE(11) FSP_HYB_CASE(12)
    default:
        k_apply_hyb<0, V><<<grid, kBlock, 0, s>>>(A, PlainX{d_x}, d_y, accumulate, tiles, partials);
    }
#undef FSP_HYB_CASE
}

template <int D, typename V>
__global__ void __launch_bounds__(kBlock) k_apply_kron(KronF<V> A, PlainX X, double *__restrict__ y,
                                                       int accumulate, uint32_t tiles,
                                                       double *__restrict__ partials, int use_lds,
                                                       int transpose)
{
    extern __shared__ double smem[];
    __shared__ double red[2 * kBlock / kWave];
    const V *tab = A.tab_g;
    if (use_lds && A.tab_g) {
        stage_table((V *)smem, A.tab_g, A.tab_size);
        tab = (const V *)smem;
    }
    dd sink = dd_zero();
    for (TileWalk w = tile_walk(tiles); w.t < w.end; w.t += w.step) {
        const uint32_t t = w.t;
        const uint32_t i = t * kBlock + threadIdx.x;
        const double v = transpose ? kron_row_t<D, V>(A, i, X, tab)
                                   : kron_row<D, V>(A, i, X, tab, sink);
        if (i < A.n) y[i] = accumulate ? y[i] + v : v;
    }
    if (A.with_sink && !transpose) put_dd_partial(partials, sink, red);
}

// The same product with the reaction geometry in a KronPlan (models with
// at most kKronPlanR reactions whose factor tables fit in LDS).  The tables
// are read from LDS only, so their loads are LDS loads.
template <int D, typename V>
__global__ void __launch_bounds__(kBlock) k_apply_kron_plan(KronF<V> A, KronPlan P, PlainX X,
                                                            double *__restrict__ y, int accumulate,
                                                            uint32_t tiles,
                                                            double *__restrict__ partials)
{
    extern __shared__ double smem[];
    __shared__ double red[2 * kBlock / kWave];
    stage_table((V *)smem, A.tab_g, A.tab_size);
    const V *tab = (const V *)smem;
    dd sink = dd_zero();
    for (TileWalk w = tile_walk(tiles); w.t < w.end; w.t += w.step) {
        const uint32_t i = w.t * kBlock + threadIdx.x;
        const double v = kron_row_plan<D, V>(A, P, i, X, tab, sink);
        if (i < A.n) y[i] = accumulate ? y[i] + v : v;
    }
    if (A.with_sink) put_dd_partial(partials, sink, red);
}

// Slot form of the Kronecker apply (kron_row_slot); the slot tables are
// read from LDS.
template <int D, typename V>
__global__ void __launch_bounds__(kBlock) k_apply_kron_slot(KronF<V> A, KronSlotPlan P,
                                                            const V *__restrict__ tab_g, PlainX X,
                                                            double *__restrict__ y, int accumulate,
                                                            uint32_t tiles,
                                                            double *__restrict__ partials)
{
    extern __shared__ double smem[];
    __shared__ double red[2 * kBlock / kWave];
    stage_table((V *)smem, tab_g, P.tab_size);
    const V *tab = (const V *)smem;
    dd sink = dd_zero();
    for (TileWalk w = tile_walk(tiles); w.t < w.end; w.t += w.step) {
        const uint32_t i = w.t * kBlock + threadIdx.x;
        const double v = kron_row_slot<D, V>(A, P, i, X, tab, sink);
        if (i < A.n) y[i] = accumulate ? y[i] + v : v;
    }
    if (A.with_sink) put_dd_partial(partials, sink, red);
}

// Diagonal of the generator: -sum_r c_r a_r(x_i) over the reactions that
// change the state (the whole outflow, including outflow leaving the box).
template <int D, typename V>
__global__ void __launch_bounds__(kBlock) k_kron_diag(KronF<V> A, double *__restrict__ out,
                                                      uint32_t tiles, int use_lds)
{
    extern __shared__ double smem[];
    const V *tab = A.tab_g;
    if (use_lds && A.tab_g) {
        stage_table((V *)smem, A.tab_g, A.tab_size);
        tab = (const V *)smem;
    }
    for (TileWalk w = tile_walk(tiles); w.t < w.end; w.t += w.step) {
        const uint32_t i = w.t * kBlock + threadIdx.x;
        if (i >= A.n) continue;
        int xs[D];
        decode<D>(i, A.dm, xs);
        double s = 0.0;
        for (int r = 0; r < A.R; ++r) {
            const RxnDev &rx = A.dm->rx[r];
            if (rx.zero_nu) continue;
            const double a = A.prop_table ? A.prop_table[(size_t)r * A.n + i]
                                          : (double)prop_tab<D>(rx, xs, tab, 0);
            s += A.coefs[r] * a;
        }
        out[i] = -s;
    }
}

}  // namespace

static int apply_grid_cap() { return 4096; }

fsp_status mat_diagonal(MatImpl &M, double *d_diag, hipStream_t s)
{
    KronF<double> A = kron_view<double>(M, -1, 0.0);
    const uint32_t tiles = (uint32_t)((M.n + kBlock - 1) / kBlock);
    const int grid = grid_for(tiles, apply_grid_cap());
    const size_t tb = sizeof(double) * (size_t)A.tab_size;
    const int use_lds = (A.tab_g && tb <= kLdsTableMax) ? 1 : 0;
    FSP_SWITCH_D(M.hm.dm.d, (k_kron_diag<D_, double><<<grid, kBlock, use_lds ? tb : 0, s>>>(
                                A, d_diag, tiles, use_lds)));
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}

template <typename V>
static fsp_status launch_kron(const MatImpl &M, const double *x, double *y, int accumulate,
                              hipStream_t s, int transpose, int single, double single_coef)
{
    KronF<V> A = kron_view<V>(M, single, single_coef);
    const uint32_t tiles = (uint32_t)((M.n + kBlock - 1) / kBlock);
    const int grid = grid_for(tiles, apply_grid_cap());
    const size_t tb = sizeof(V) * (size_t)A.tab_size;
    const int use_lds = (A.tab_g && tb <= kLdsTableMax) ? 1 : 0;
    const int with_sink = M.with_sink && !transpose && single < 0;
    A.with_sink = with_sink;
    KronPlan P;
    KronSlotPlan SP;
    if (!transpose && single < 0 && M.kron_slots && kron_slot_plan(M.hm, SP)) {
        const V *st = std::is_same<V, float>::value ? (const V *)(const void *)M.d_slot_tab_f
                                                    : (const V *)(const void *)M.d_slot_tab;
        FSP_SWITCH_D(M.hm.dm.d, (k_apply_kron_slot<D_, V><<<grid, kBlock, sizeof(V) * SP.tab_size, s>>>(
                                    A, SP, st, PlainX{x}, y, accumulate, tiles, M.partials)));
    } else if (!transpose && single < 0 && !A.prop_table && use_lds && kron_plan(M.hm, P)) {
        FSP_SWITCH_D(M.hm.dm.d, (k_apply_kron_plan<D_, V><<<grid, kBlock, tb, s>>>(
                                    A, P, PlainX{x}, y, accumulate, tiles, M.partials)));
    } else {
        FSP_SWITCH_D(M.hm.dm.d, (k_apply_kron<D_, V><<<grid, kBlock, use_lds ? tb : 0, s>>>(
                                    A, PlainX{x}, y, accumulate, tiles, M.partials, use_lds,
                                    transpose)));
    }
    FSP_HIP(hipGetLastError());
    if (with_sink) {
        k_finalize_sink<<<1, kBlock, 0, s>>>(M.partials, grid, y + M.n, accumulate);
    } else if (M.with_sink && !accumulate) {
        FSP_HIP(hipMemsetAsync(y + M.n, 0, sizeof(double), s));
    }
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}

fsp_status mat_apply(MatImpl &M, const double *d_x, double *d_y, int accumulate, hipStream_t s,
                     const int *)
{
    const uint32_t tiles = (uint32_t)((M.n + kBlock - 1) / kBlock);
    const int grid = grid_for(tiles, apply_grid_cap());
    const bool mixed = M.prec == FSP_MIXED;
    if (M.format == FSP_CSR) {
        // one 512-row tile per workgroup (the partials hold 64 slots per
        // 256-row tile, so one slot per 512-row tile fits)
        const uint32_t ct = (uint32_t)((M.n + kCsrApplyRows - 1) / kCsrApplyRows);
        const int cgrid = grid_for(ct, 1 << 30);
        const size_t lds_red = 2 * kCsrApplyRows / kWave;
        if (mixed) {
            CsrF<float> A = csr_view<float>(M, kCsrApplyRows, kCsrApplyPer);
            const size_t lds = sizeof(double) * ((size_t)A.cap + lds_red);
            k_apply_csr<float><<<cgrid, kCsrApplyRows, lds, s>>>(A, PlainX{d_x}, d_y, accumulate,
                                                                 M.partials);
        } else {
            CsrF<double> A = csr_view<double>(M, kCsrApplyRows, kCsrApplyPer);
            const size_t lds = sizeof(double) * ((size_t)A.cap + lds_red);
            k_apply_csr<double><<<cgrid, kCsrApplyRows, lds, s>>>(A, PlainX{d_x}, d_y, accumulate,
                                                                  M.partials);
        }
        FSP_HIP(hipGetLastError());
        if (M.with_sink) {
            k_finalize_sink<<<1, kBlock, 0, s>>>(M.partials, cgrid, d_y + M.n, accumulate);
            FSP_HIP(hipGetLastError());
        }
        return FSP_OK;
    } else if (M.format == FSP_HYB) {
        // one tile per workgroup
        const int hgrid = grid_for(tiles, 1 << 30);
        if (mixed)
            launch_hyb<float>(hyb_view<float>(M), d_x, d_y, accumulate, tiles, M.partials, hgrid, s);
        else
            launch_hyb<double>(hyb_view<double>(M), d_x, d_y, accumulate, tiles, M.partials, hgrid, s);
        FSP_HIP(hipGetLastError());
        if (M.with_sink) k_finalize_sink<<<1, kBlock, 0, s>>>(M.partials, hgrid, d_y + M.n, accumulate);
        FSP_HIP(hipGetLastError());
        return FSP_OK;
    } else {
        return (mixed && M.hm.has_factors)
                   ? launch_kron<float>(M, d_x, d_y, accumulate, s, 0, -1, 0.0)
                     : launch_kron<double>(M, d_x, d_y, accumulate, s, 0, -1, 0.0);
    }
    FSP_HIP(hipGetLastError());
    if (M.with_sink) k_finalize_sink<<<1, kBlock, 0, s>>>(M.partials, grid, d_y + M.n, accumulate);
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}

fsp_status mat_apply_transpose(MatImpl &M, const double *d_x, double *d_y, int accumulate,
                               hipStream_t s)
{
    if (M.with_sink) return fail(FSP_ERR_UNSUPPORTED, "transpose apply with a sink state");
    if (M.transposed) return fail(FSP_ERR_UNSUPPORTED, "transpose apply of a transposed generator");
    // The transpose is matrix-free from the model for every format: the
    // generator entries are the model's propensities, so no transposed copy
    // of the CSR / HYB arrays is kept.
    return launch_kron<double>(M, d_x, d_y, accumulate, s, 1, -1, 0.0);
}

fsp_status mat_apply_term(MatImpl &M, int r, double c, const double *d_x, double *d_y,
                          int accumulate, hipStream_t s)
{
    if (r < 0 || r >= M.hm.dm.R) return fail(FSP_ERR_INVALID, "reaction index out of range");
    if (M.with_sink) return fail(FSP_ERR_UNSUPPORTED, "term apply with a sink state");
    if (M.transposed) return fail(FSP_ERR_UNSUPPORTED, "term apply of a transposed generator");
    return launch_kron<double>(M, d_x, d_y, accumulate, s, 0, r, c);
}

void mat_free(MatImpl &M)
{
    void *ptrs[] = {M.d_model, M.d_compact, M.d_compact_f, M.d_coefs, M.rp, M.cols, M.vals,
                    M.sink_cols, M.sink_vals, M.sink_block_ptr, M.diag, M.slot_vals,
                    M.slot_off, M.sink_out, M.partials, M.ticket, M.d_slot_tab,
                    M.d_slot_tab_f};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
}

double mat_bytes_per_apply(const MatImpl &M)
{
    // Algorithmic bytes of one y = A x (x read once, y written once).
    const double n = (double)M.n;
    const double vb = M.prec == FSP_MIXED ? 4.0 : 8.0;
    double b = 0.0;
    if (M.format == FSP_CSR) {
        b = 4.0 * (n + 1) + (vb + 4.0) * (double)M.nnz_rows + 16.0 * n;
        if (M.prec == FSP_MIXED) b += 8.0 * n;  // separate FP64 diagonal
        if (M.with_sink) b += 12.0 * (double)M.nnz_sink;
    } else if (M.format == FSP_HYB) {
        b = 8.0 * n + vb * (double)M.hm.dm.U * n + 16.0 * n;  // columns implicit
        if (M.with_sink) b += 8.0 * n;
    } else {
        b = 16.0 * n;
    }
    return b;
}

fsp_status mat_hyb_cols(const MatImpl &M, int *d_cols, hipStream_t s)
{
    const uint32_t n = (uint32_t)M.n;
    const unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
    FSP_SWITCH_D(M.hm.dm.d, (k_hyb_cols<D_><<<blocks, kBlock, 0, s>>>(M.d_model, n, d_cols)));
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}

// ============================================================= term CSR
fsp_status term_csr(const HostModel &hm, int r, int *d_rp, int *d_cols, double *d_vals,
                    int64_t *h_nnz, hipStream_t s)
{
    const DevModel &dm = hm.dm;
    if (r < 0 || r >= dm.R) return fail(FSP_ERR_INVALID, "reaction index out of range");
    const uint32_t n = (uint32_t)dm.n;
    const unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
    DevModel *d_model = nullptr;
    double *d_tab = nullptr;
    int *cnt = nullptr, *rp = d_rp;
    FSP_HIP(hipMalloc(&d_model, sizeof(DevModel)));
    FSP_HIP(copy_sync(d_model, &dm, sizeof(DevModel), hipMemcpyHostToDevice, s));
    if (!hm.compact.empty()) {
        FSP_HIP(hipMalloc(&d_tab, sizeof(double) * hm.compact.size()));
        FSP_HIP(copy_sync(d_tab, hm.compact.data(), sizeof(double) * hm.compact.size(),
                          hipMemcpyHostToDevice, s));
    }
    FSP_HIP(hipMalloc(&cnt, sizeof(int) * (size_t)(n + 1)));
    if (!rp) FSP_HIP(hipMalloc(&rp, sizeof(int) * (size_t)(n + 1)));
    FSP_SWITCH_D(dm.d, (k_term_count<D_><<<blocks, kBlock, 0, s>>>(d_model, r, n, cnt)));
    FSP_HIP(hipGetLastError());
    int64_t total = 0;
    fsp_status st = scan_counts(cnt, rp, n, &total, s);
    if (st == FSP_OK && d_cols) {
        int use_lds = 0;
        const size_t lds = table_lds_bytes(hm, use_lds);
        FSP_SWITCH_D(dm.d, (k_term_fill<D_><<<blocks, kBlock, lds, s>>>(
                               d_model, d_tab, use_lds, hm.prop_table_dev, r, n, rp, d_cols,
                               d_vals)));
        FSP_HIP(hipGetLastError());
        FSP_HIP(hipStreamSynchronize(s));
    }
    *h_nnz = total;
    (void)hipFree(cnt);
    if (rp != d_rp) (void)hipFree(rp);
    (void)hipFree(d_model);
    if (d_tab) (void)hipFree(d_tab);
    return st;
}

}  // namespace fsp
