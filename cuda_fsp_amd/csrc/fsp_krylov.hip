// fsp_krylov.hip -- device-resident Krylov expv (Expokit scheme) for gfx950.
//
// Restates KryExpvFSP::step (src/KryExpvFSP.cu:38-182) with the whole step on
// the device: no host round trip inside a step (the reference syncs twice per
// Arnoldi iteration and runs arma::expmat on the host).
//
// Per Arnoldi iteration j (operator held by an fsp_mat):
//   k_arn_spmv   : V_j = w_j * sigma_j (the reference's V_{j+1} /= s, fused
//                  into the read), w_{j+1} = A V_j, partial dots V_k . w_{j+1}
//   k_arn_dots   : H(k, j) = sum of partials (fixed order), sink element
//   k_arn_update : w_{j+1} -= V h (gemv N of :79-81), partial |w|^2
//   (next k_arn_spmv, every workgroup): s = |w|, happy breakdown test,
//                  H(j+1, j) = s, sigma = 1/s (k_arn_norm with a host
//                  callback operator)
// With an FP32 basis (mixed precision) one classical Gram-Schmidt pass, as
// the reference does in FP64, lets the loss of orthogonality compound by
// |w| / h(j+1, j) per step until the basis is degenerate.  The mixed path
// therefore projects with the Gram matrix G = V^T V of the stored basis:
// k_arn_* also returns the dots V_k . V_j (column j of G); every workgroup
// of k_arn_update turns the Hessenberg column h into c = 2h - G h (the
// second Gram-Schmidt pass h2 = V^T (w - V h) = h - G h folded into the
// first) and subtracts V c in one sweep over the basis.
// then k_expm_control (one workgroup: Pade expm of tau*H, error estimate,
// reject loop, step-size update) and k_solution (v = beta V F0).
// Reductions are deterministic: per-wave butterflies, per-block partials in
// fixed order, one finalizing kernel.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fsp_krylov_dev.h"

namespace fsp {


constexpr long long kTauLogCap = 1 << 16;


struct ExpvImpl {
    MatImpl *A = nullptr;
    fsp_matvec_cb cb = nullptr;
    void *ctx = nullptr;
    int64_t n = 0;       // vector length (n_rows of the operator)
    int m = 30;
    double tol = 1e-8;
    int iop = 0, q_iop = 2;
    double anorm = 1.0;
    int max_reject = 1000;
    int64_t max_nstep = 10000;
    hipStream_t s = nullptr;
    bool mixed = false;
    KState *d_st = nullptr;
    KState *h_st = nullptr;      // pinned mirror
    double *V64 = nullptr;       // FP64 basis (all vectors, or V_0 when mixed)
    float *V32 = nullptr;        // mixed: V_1..V_m
    double *W[2] = {nullptr, nullptr};
    double *partials = nullptr;
    int G = 1;                   // blocks of the row kernels
    int64_t tiles = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double *graph_v = nullptr;
    bool use_graph = true;
    hipStream_t cap = nullptr;   // private stream used only for graph capture
    double *tau_log = nullptr;
    double *dcoef = nullptr;     // dense-output coefficients (m + 2)
    // per-kernel timing (fsp_expv_set_profiling): while enabled, steps run
    // without the captured graph, one per host check, and each launch is
    // bracketed by HIP events on the handle's stream
    bool profile = false;
    std::vector<hipEvent_t> pev;      // event pool, two per launch of a step
    std::vector<int> pcls;            // kernel class of each launch
    std::vector<double> pbytes;       // algorithmic bytes of each launch
    int pcount = 0;                   // launches recorded in the current step
    int64_t prof_n[FSP_PROF_CLASSES] = {};
    double prof_ms[FSP_PROF_CLASSES] = {};
    double prof_bytes[FSP_PROF_CLASSES] = {};
};

namespace {

// ------------------------------------------------------------- accessors
struct ArnArgs {
    KState *st;
    const double *in;     // w_j (or v for j = 0), unscaled
    double *out;          // w_{j+1}
    double *V64;
    void *Vb;
    double *partials;     // [(nk + 1) * G]
    uint32_t n;           // box states (rows computed by the format)
    uint32_t nv;          // vector length (n or n+1)
    uint32_t tiles;
    int j, istart, mode;  // mode 0: dots, 1: norm only (avnorm)
    int G;
    double *npart;        // [G] dd partials of |w_j|^2 from the update kernel
    int M2;
};

constexpr int kRedLen = 2 * kBlock / kWave;
// LDS of the Arnoldi kernels' end reduction: one dd per (dot, wave), for up
// to m + 1 dots plus the sink sum.
// Partial slots per Arnoldi kernel: dots 0..nk-1, the sink sum at nk and,
// with an FP32 basis, the Gram column at nk+1..2nk.
constexpr int kArnSlots = 2 * kMaxKrylov + 4;
constexpr int kArnRedLen = 2 * kArnSlots * (kBlock / kWave);
constexpr int kDotChunk = 8;
constexpr int kTileGroup = 3;  // tiles whose dot loads are issued together
// LDS of the FP64 dot phase's transposed block reduction: kDotChunk dd per
// thread.  It follows the kArnRedLen doubles of red and may share its bytes
// with a kernel's phase-1 LDS (a CSR product tile, a Kronecker table): the
// dot phase starts with a barrier after the last phase-1 access.
constexpr int kArnChunkLen = 2 * kDotChunk * kBlock;

// Dynamic LDS bytes of an Arnoldi row kernel with `extra` bytes of phase-1
// LDS after red.
inline size_t arn_lds_bytes(size_t extra)
{
    return sizeof(double) * kArnRedLen + std::max(extra, sizeof(double) * kArnChunkLen);
}

template <typename B>
__device__ inline bool arn_skip(const ArnArgs &a)
{
    const KState *st = a.st;
    if (!st->active || st->broke) return true;
    if (a.mode == 1 && st->k1 == 0) return true;
    return false;
}

// Phase 1 row epilogue: store V_j[i] and w_{j+1}[i] = y (mode 0), or
// accumulate |y|^2 (mode 1: the avnorm product A V_m).
template <typename B>
__device__ inline void arn_epi(const ArnArgs &a, uint32_t i, double y, double xi, dd &nrm)
{
    if (i >= a.n) return;
    store_basis<B>(a.V64, (B *)a.Vb, a.nv, a.j, i, xi);
    if (a.mode == 0)
        a.out[i] = y;
    else
        nrm = dd_add_prod(nrm, y, y);
}

// Merges the per-wave dd sums stored in red (slot k of wave w at
// 2 (k NW + w)) in wave order and writes the block's partials of slots
// first..last: slot k of block b at 2 (k G + b).
__device__ inline void arn_store_partials(const ArnArgs &a, int last, const double *red, int first = 0)
{
    constexpr int NW = kBlock / kWave;
    __syncthreads();
    if ((int)threadIdx.x <= last - first) {
        const int kk = first + (int)threadIdx.x;
        dd s = dd_zero();
#pragma unroll
        for (int q = 0; q < NW; ++q) s = dd_merge(s, dd{red[2 * (kk * NW + q)], red[2 * (kk * NW + q) + 1]});
        a.partials[2 * ((size_t)kk * a.G + blockIdx.x)] = s.s;
        a.partials[2 * ((size_t)kk * a.G + blockIdx.x) + 1] = s.c;
    }
}

__device__ inline void arn_put_wave(double *red, int k, dd v)
{
    constexpr int NW = kBlock / kWave;
    const dd t = wave_sum_dd(v);
    if (threadIdx.x % kWave == 0) {
        const int w = threadIdx.x / kWave;
        red[2 * (k * NW + w)] = t.s;
        red[2 * (k * NW + w) + 1] = t.c;
    }
}

// Wave sums of slots base + q, q = 0 .. kDotChunk - 1 (those with
// c0 + q < nk) from kDotChunk per-lane dd sums, by the transposed wave
// reduction (wave_sums_dd_t, through the wave's own kDotChunk kWave dd of
// LDS after red), into red in the arn_put_wave layout; arn_store_partials
// merges the waves once after the last slot.
__device__ inline void arn_chunk_waves(int c0, int nk, int base, const dd (&acc)[kDotChunk], double *red)
{
    constexpr int NW = kBlock / kWave;
    const int w = threadIdx.x / kWave;
    double *wbuf = red + kArnRedLen + (size_t)w * (2 * kDotChunk * kWave);
    int q;
    dd v;
    if (wave_sums_dd_t<kDotChunk>(acc, wbuf, q, v) && c0 + q < nk) {
        red[2 * ((base + q) * NW + w)] = v.s;
        red[2 * ((base + q) * NW + w) + 1] = v.c;
    }
}

// Phase 2 and the block reduction.  Mode 0: the dots V_k . w_{j+1}
// (k = istart..j) over this block's rows, kDotChunk dots at a time: each
// thread re-reads its own w_{j+1}[i] (written in phase 1, so cached) and a
// chunk of independent basis loads (column index clamped to j, always
// valid), so few registers are live and a chunk costs one memory round trip.
// Each chunk's wave sums come from the transposed wave reduction
// (arn_chunk_waves, kArnChunkLen doubles of LDS after red, a slice per
// wave) and the sink sum (and mode 1's norm) from wave butterflies, all into
// red (kArnRedLen doubles); then one barrier and the block partials.
// With an FP32 basis the same pass also forms the Gram dots V_k . V_j
// (slots nk+1..2nk) from the same basis loads; each thread sums both in
// plain FP64 (a fixed order, so still deterministic; no bit-exact oracle
// match is asked of the FP32 path).

// c_begin: the first chunk of FP64 dots still to form (a kernel that formed
// the first chunk with its rows passes kDotChunk).
template <typename B>
__device__ inline void arn_finish(const ArnArgs &a, dd nrm, dd sink, double *red, int c_begin = 0)
{
    constexpr bool GRAM = gram_basis<B>();
    const int nk = a.mode == 0 ? a.j - a.istart + 1 : 1;
    // the chunk slices may share bytes with a kernel's phase-1 LDS (a CSR
    // product tile, a Kronecker table) that other waves may still read
    if (a.mode == 0 && c_begin < nk) __syncthreads();
    if (a.mode == 0 && !GRAM) {
        for (int c0 = c_begin; c0 < nk; c0 += kDotChunk) {
            dd acc[kDotChunk];
#pragma unroll
            for (int q = 0; q < kDotChunk; ++q) acc[q] = dd_zero();
            // kTileGroup of the workgroup's tiles at a time: all their loads
            // are issued before the sums, which still run in tile order
            for (TileWalk t = tile_walk(a.tiles); t.t < t.end; t.t += kTileGroup * t.step) {
                double y[kTileGroup], vk[kTileGroup][kDotChunk];
                bool ok[kTileGroup];
#pragma unroll
                for (int g = 0; g < kTileGroup; ++g) {
                    const uint32_t tt = t.t + (uint32_t)g * t.step;
                    const uint32_t i = tt * kBlock + threadIdx.x;
                    ok[g] = tt < t.end && i < a.n;
                    const uint32_t r = ok[g] ? i : 0u;
                    y[g] = a.out[r];
                    // only the chunk's vectors (a uniform test, so the loads
                    // still go out together): padding loads of a short last
                    // chunk cost L1/L2 bandwidth
#pragma unroll
                    for (int q = 0; q < kDotChunk; ++q)
                        vk[g][q] = c0 + q < nk ? basis<B>(a.V64, (const B *)a.Vb, a.nv, a.istart + c0 + q, r) : 0.0;
                }
#pragma unroll
                for (int g = 0; g < kTileGroup; ++g) {
                    if (!ok[g]) continue;
#pragma unroll
                    for (int q = 0; q < kDotChunk; ++q)
                        if (c0 + q < nk) acc[q] = dd_add_prod(acc[q], y[g], vk[g][q]);
                }
            }
            arn_chunk_waves(c0, nk, c0, acc, red);
        }
    } else if (a.mode == 0) {
        for (int c0 = 0; c0 < nk; c0 += kDotChunk) {
            double acc[kDotChunk], accg[kDotChunk];
#pragma unroll
            for (int q = 0; q < kDotChunk; ++q) acc[q] = accg[q] = 0.0;
            for (TileWalk t = tile_walk(a.tiles); t.t < t.end; t.t += kTileGroup * t.step) {
                double y[kTileGroup], xj[kTileGroup], vk[kTileGroup][kDotChunk];
                bool ok[kTileGroup];
#pragma unroll
                for (int g = 0; g < kTileGroup; ++g) {
                    const uint32_t tt = t.t + (uint32_t)g * t.step;
                    const uint32_t i = tt * kBlock + threadIdx.x;
                    ok[g] = tt < t.end && i < a.n;
                    const uint32_t r = ok[g] ? i : 0u;
                    y[g] = a.out[r];
                    xj[g] = basis<B>(a.V64, (const B *)a.Vb, a.nv, a.j, r);
                    // padding loads past the chunk re-read V_j (a guard here
                    // made the FP32-basis solve 15% slower)
#pragma unroll
                    for (int q = 0; q < kDotChunk; ++q)
                        vk[g][q] = basis<B>(a.V64, (const B *)a.Vb, a.nv, min(a.istart + c0 + q, a.j), r);
                }
#pragma unroll
                for (int g = 0; g < kTileGroup; ++g) {
                    if (!ok[g]) continue;
#pragma unroll
                    for (int q = 0; q < kDotChunk; ++q) {
                        acc[q] = fma(y[g], vk[g][q], acc[q]);
                        accg[q] = fma(xj[g], vk[g][q], accg[q]);
                    }
                }
            }
            dd da[kDotChunk], dg[kDotChunk];
#pragma unroll
            for (int q = 0; q < kDotChunk; ++q) {
                da[q] = dd{acc[q], 0.0};
                dg[q] = dd{accg[q], 0.0};
            }
            arn_chunk_waves(c0, nk, c0, da, red);
            arn_chunk_waves(c0, nk, nk + 1 + c0, dg, red);
        }
    }
    // the sink sum at slot nk (mode 1: its norm at slot 0, the sink at 1);
    // without a sink entry (nv == n) the finalizer reads no sink slot
    const bool has_sink = a.nv > a.n;
    if (a.mode == 0) {
        if (has_sink || GRAM) arn_put_wave(red, nk, sink);
        arn_store_partials(a, GRAM ? 2 * nk : (has_sink ? nk : nk - 1), red);
    } else {
        arn_put_wave(red, 0, nrm);
        if (has_sink) arn_put_wave(red, 1, sink);
        arn_store_partials(a, has_sink ? 1 : 0, red);
    }
}

// Register-accumulator variant for the callback path (k_cb_dots), which
// reads rows it did not write: NKB >= number of dots.
template <int NKB>
__device__ inline void arn_end(const ArnArgs &a, dd (&acc)[NKB], dd sink, double *red)
{
    const int nk = a.mode == 0 ? a.j - a.istart + 1 : 1;
#pragma unroll
    for (int kk = 0; kk < NKB; ++kk)
        if (kk < nk) arn_put_wave(red, kk, acc[kk]);
    arn_put_wave(red, nk, sink);
    arn_store_partials(a, nk, red);
}

template <int NKB>
__device__ inline void acc_zero(dd (&acc)[NKB])
{
#pragma unroll
    for (int kk = 0; kk < NKB; ++kk) acc[kk] = dd_zero();
}

// Scale of the input vector: sigma_0 for j = 0; for j >= 1 every workgroup
// sums the update kernel's |w_j|^2 partials itself (the fixed-order sum of
// a finalizing kernel, so the same bits in every workgroup) and workgroup 0
// records H(j, j-1) = |w_j| and sigma_j, or the happy breakdown
// (src/KryExpvFSP.cu:90-98), after which every workgroup returns.  This
// replaces a one-workgroup norm kernel between the update and this kernel.
// Returns false when the kernel has nothing to do.
__device__ inline bool arn_sigma(const ArnArgs &a, double *red, double &sig)
{
    KState *st = a.st;
    if (a.j == 0) {
        sig = st->sigma[0];
        return true;
    }
    const double s = sqrt(dd_value(sum_dd_partials(a.npart, a.G, red)));
    // every wave has read the block sum out of red before any wave writes
    // its dot or norm sums there
    __syncthreads();
    const bool broke = s < st->btol;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (broke) {
            st->k1 = 0;
            st->mb = a.j;
            st->tau = st->t_final - st->t_now;
            st->broke = 1;
            st->n_breakdown += 1;
        } else {
            st->H[a.j + (a.j - 1) * a.M2] = s;
            st->sigma[a.j] = 1.0 / s;
        }
    }
    sig = 1.0 / s;
    return !broke;
}

template <typename V, typename B>
__global__ void __launch_bounds__(kBlock) k_arn_csr(CsrF<V> A, ArnArgs a)
{
    if (arn_skip<B>(a)) return;
    FSP_DYNAMIC_LDS(smem);
    double *red = smem;
    double *prod = red + kArnRedLen;
    double sig;
    if (!arn_sigma(a, red, sig)) return;
    dd nrm = dd_zero();
    const KryX<B> X{a.in, sig, a.j};
    dd sink = dd_zero();
    for (TileWalk w = tile_walk(a.tiles); w.t < w.end; w.t += w.step) {
        const double y = csr_tile<V>(A, w.t, X, prod, sink);
        const uint32_t i = w.t * kBlock + threadIdx.x;
        arn_epi<B>(a, i, y, i < a.n ? X(i) : 0.0, nrm);
    }
    arn_finish<B>(a, nrm, sink, red);
}

// NU > 0: the slot count as a compile-time constant and, up to 8 slots, the
// workgroup's tiles in a two-deep pipeline: the loads of tile t + 1 are
// issued before tile t is summed and stored (stores overlap loads; a burst
// of all tiles' loads, then all their stores, measured 2.4 us slower per
// launch at 3 tiles of 786K rows).  Stores, sink terms and norms stay in
// tile order.  NU = 0: any slot count, one tile at a time.
template <int NU, typename V, typename B>
__global__ void __launch_bounds__(kBlock) k_arn_hyb(HybF<V> A, ArnArgs a)
{
    if (arn_skip<B>(a)) return;
    FSP_DYNAMIC_LDS(smem);
    double *red = smem;
    constexpr int K = NU > 0 ? NU : 1;
    // the pipeline holds two rows' loads: above 8 slots that would cost
    // occupancy (registers), so one tile at a time there
    constexpr bool kPipe = NU > 0 && K <= 8;
    const TileWalk w = tile_walk(a.tiles);
    // rows past the box read row 0 and are discarded
    auto row = [&](uint32_t t) {
        const uint32_t i = t * kBlock + threadIdx.x;
        return i < A.n ? i : 0u;
    };
    // the first tile's loads go out before the scale sigma_j is known (its
    // x entries raw, scaled at the sum by the same kry_scale), so that they
    // overlap the norm prologue's loads of the update's partials
    HybLoads<K, V> cur, nxt;
    if (kPipe && w.t < w.end) hyb_load<K>(A, row(w.t), RawX{a.in}, cur);
    double sig;
    if (!arn_sigma(a, red, sig)) return;
    dd nrm = dd_zero();
    const KryX<B> X{a.in, sig, a.j};
    const int jj = a.j;
    auto xs = [sig, jj](double v) { return kry_scale<B>(v, sig, jj); };
    dd sink = dd_zero();
    // FP64 basis, dots mode: the first chunk of dots is formed tile by tile
    // with the rows (w_{j+1} and V_j from registers, the basis loads of the
    // vectors below j issued with the next tile's generator loads), in the
    // same per-thread order as arn_finish's pass
    const bool fuse0 = kPipe && !gram_basis<B>() && a.mode == 0;
    const int nk = a.j - a.istart + 1;
    dd acc0[kDotChunk];
#pragma unroll
    for (int q = 0; q < kDotChunk; ++q) acc0[q] = dd_zero();
    if (NU > 0) {
        for (uint32_t t = w.t; t < w.end; t += w.step) {
            const uint32_t tn = t + w.step;
            if (!kPipe)
                hyb_load<K>(A, row(t), RawX{a.in}, cur);
            else if (tn < w.end)
                hyb_load<K>(A, row(tn), RawX{a.in}, nxt);
            const uint32_t i = t * kBlock + threadIdx.x;
            const bool in = i < A.n;
            // vectors istart + q below j (any valid index otherwise: such a
            // load is not used; a guard here would cost a register and
            // occupancy)
            double vk[kDotChunk];
            if (fuse0) {
#pragma unroll
                for (int q = 0; q < kDotChunk; ++q)
                    vk[q] = basis<B>(a.V64, (const B *)a.Vb, a.nv, min(a.istart + q, max(a.j - 1, 0)),
                                     row(t));
            }
            const double y = hyb_sum_x<K>(cur, xs);
            if (A.sink_out && in) sink = dd_add_prod(sink, A.sink_out[i], X(i));
            arn_epi<B>(a, i, in ? y : 0.0, i < a.n ? X(i) : 0.0, nrm);
            if (fuse0 && i < a.n) {
                const double xj = X(i);
#pragma unroll
                for (int q = 0; q < kDotChunk; ++q)
                    if (q < nk) acc0[q] = dd_add_prod(acc0[q], y, a.istart + q < a.j ? vk[q] : xj);
            }
            if (kPipe) cur = nxt;
        }
    } else {
        for (TileWalk v = w; v.t < v.end; v.t += v.step) {
            const uint32_t i = v.t * kBlock + threadIdx.x;
            const double y = hyb_row<V>(A, i, X, sink);
            arn_epi<B>(a, i, y, i < a.n ? X(i) : 0.0, nrm);
        }
    }
    if (fuse0) arn_chunk_waves(0, nk, 0, acc0, red);
    arn_finish<B>(a, nrm, sink, red, fuse0 ? kDotChunk : 0);
}

template <int D, typename V, typename B>
__global__ void __launch_bounds__(kBlock) k_arn_kron(KronF<V> A, ArnArgs a, int use_lds)
{
    if (arn_skip<B>(a)) return;
    FSP_DYNAMIC_LDS(smem);
    double *red = smem;
    V *tabl = (V *)(red + kArnRedLen);
    const V *tab = A.tab_g;
    if (use_lds && A.tab_g) {
        stage_table(tabl, A.tab_g, A.tab_size);
        tab = tabl;
    }
    double sig;
    if (!arn_sigma(a, red, sig)) return;
    dd nrm = dd_zero();
    const KryX<B> X{a.in, sig, a.j};
    dd sink = dd_zero();
    for (TileWalk w = tile_walk(a.tiles); w.t < w.end; w.t += w.step) {
        const uint32_t i = w.t * kBlock + threadIdx.x;
        const double y = kron_row<D, V>(A, i, X, tab, sink);
        arn_epi<B>(a, i, y, i < a.n ? X(i) : 0.0, nrm);
    }
    arn_finish<B>(a, nrm, sink, red);
}

// k_arn_kron with the reaction geometry in a KronPlan; the factor tables are
// read from LDS only (LDS loads).
template <int D, typename V, typename B>
__global__ void __launch_bounds__(kBlock) k_arn_kron_plan(KronF<V> A, KronPlan P, ArnArgs a)
{
    if (arn_skip<B>(a)) return;
    FSP_DYNAMIC_LDS(smem);
    double *red = smem;
    V *tabl = (V *)(red + kArnRedLen);
    stage_table(tabl, A.tab_g, A.tab_size);
    const V *tab = tabl;
    double sig;
    if (!arn_sigma(a, red, sig)) return;
    dd nrm = dd_zero();
    const KryX<B> X{a.in, sig, a.j};
    dd sink = dd_zero();
    for (TileWalk w = tile_walk(a.tiles); w.t < w.end; w.t += w.step) {
        const uint32_t i = w.t * kBlock + threadIdx.x;
        const double y = kron_row_plan<D, V>(A, P, i, X, tab, sink);
        arn_epi<B>(a, i, y, i < a.n ? X(i) : 0.0, nrm);
    }
    arn_finish<B>(a, nrm, sink, red);
}

// k_arn_kron with the slot form (kron_row_slot); slot tables in LDS.
template <int D, typename V, typename B>
__global__ void __launch_bounds__(kBlock) k_arn_kron_slot(KronF<V> A, KronSlotPlan P,
                                                          const V *__restrict__ tab_g, ArnArgs a)
{
    if (arn_skip<B>(a)) return;
    FSP_DYNAMIC_LDS(smem);
    double *red = smem;
    V *tabl = (V *)(red + kArnRedLen);
    stage_table(tabl, tab_g, P.tab_size);
    const V *tab = tabl;
    double sig;
    if (!arn_sigma(a, red, sig)) return;
    dd nrm = dd_zero();
    const KryX<B> X{a.in, sig, a.j};
    dd sink = dd_zero();
    for (TileWalk w = tile_walk(a.tiles); w.t < w.end; w.t += w.step) {
        const uint32_t i = w.t * kBlock + threadIdx.x;
        const double y = kron_row_slot<D, V>(A, P, i, X, tab, sink);
        arn_epi<B>(a, i, y, i < a.n ? X(i) : 0.0, nrm);
    }
    arn_finish<B>(a, nrm, sink, red);
}

#define FSP_SWITCH_NK(nk, CALL)                                                \
    if ((nk) <= 8) {                                                           \
        constexpr int NK_ = 8;                                                 \
        CALL;                                                                  \
    } else if ((nk) <= 16) {                                                   \
        constexpr int NK_ = 16;                                                \
        CALL;                                                                  \
    } else if ((nk) <= 32) {                                                   \
        constexpr int NK_ = 32;                                                \
        CALL;                                                                  \
    } else {                                                                   \
        constexpr int NK_ = kMaxKrylov + 1;                                    \
        CALL;                                                                  \
    }

// Finalize dots (mode 0) or the avnorm (mode 1).  Block k sums partials of
// dot k in fixed order; the sink element (index n) is handled here.
template <typename B>
__global__ void __launch_bounds__(kBlock) k_arn_dots(ArnArgs a, int with_sink, int M2)
{
    if (arn_skip<B>(a)) return;
    __shared__ double red[kRedLen];
    KState *st = a.st;
    const int nk = a.mode == 0 ? a.j - a.istart + 1 : 1;
    double ys = 0.0, xs = 0.0;
    if (with_sink) {
        ys = dd_value(sum_dd_partials(a.partials + 2 * (size_t)nk * a.G, a.G, red));
        xs = kry_scale<B>(a.in[a.n], st->sigma[a.j], a.j);
    }
    const int k = blockIdx.x;
    dd sd = sum_dd_partials(a.partials + 2 * (size_t)k * a.G, a.G, red);
    dd sg = dd_zero();
    if (gram_basis<B>() && a.mode == 0)
        sg = sum_dd_partials(a.partials + 2 * (size_t)(nk + 1 + k) * a.G, a.G, red);
    if (threadIdx.x == 0) {
        if (a.mode == 0) {
            const int kk = a.istart + k;
            if (with_sink) {
                const double vk = (kk == a.j) ? xs : basis<B>(a.V64, (const B *)a.Vb, a.nv, kk, a.n);
                sd = dd_add_prod(sd, ys, vk);
                if (gram_basis<B>()) sg = dd_add_prod(sg, xs, vk);
            }
            const double s = dd_value(sd);
            st->H[kk + a.j * M2] = s;
            if (gram_basis<B>()) {
                st->hs[kk] = s;
                const double g = dd_value(sg);
                st->G[kk + a.j * kGramLd] = g;
                st->G[a.j + kk * kGramLd] = g;
            }
            if (k == 0) {
                st->n_matvec += 1;
                if (with_sink) {
                    a.out[a.n] = ys;
                    store_basis<B>(a.V64, (B *)a.Vb, a.nv, a.j, a.n, xs);
                }
            }
        } else {
            st->avnorm = sqrt(dd_value(dd_add_prod(sd, ys, ys)));
            st->H[(a.j + 1) + a.j * M2] = 1.0;  // H(m+1, m) = 1 (src/KryExpvFSP.cu:115)
            st->n_matvec += 1;
            if (with_sink) store_basis<B>(a.V64, (B *)a.Vb, a.nv, a.j, a.n, xs);
        }
    }
}


// w_{j+1} -= V[istart..j] H(istart..j, j); partial |w|^2.
template <typename B>
__global__ void __launch_bounds__(kBlock) k_arn_update(ArnArgs a, int M2)
{
    if (arn_skip<B>(a)) return;
    __shared__ double red[kRedLen];
    KState *st = a.st;
    const double *h = st->H + (size_t)a.j * M2;
    if (gram_basis<B>()) {
        // FP32 basis: every workgroup forms the Gram-corrected coefficients
        // 2h - G h from the raw dots in LDS; workgroup 0 also stores them as
        // the Hessenberg column (no workgroup reads that column here)
        FSP_DYNAMIC_LDS(lds);
        double *cw = lds + kGramLdsBytes / sizeof(double);
        gram_coefs_lds(st, a.istart, a.j, lds, cw);
        if (blockIdx.x == 0)
            for (int k = threadIdx.x; k <= a.j - a.istart; k += blockDim.x)
                st->H[(a.istart + k) + (size_t)a.j * M2] = cw[k];
        h = cw - a.istart;
    }
    // the rows of kUpdRows tiles per thread at a time, walked as the Arnoldi
    // row kernel walks them (so each XCD updates the rows whose w_{j+1} and
    // V_j it just wrote): the loads of all of them are issued before their
    // sums, so a group of rows costs one memory round trip per 8 basis
    // vectors instead of one per row; each row's sum still runs over k
    // ascending
    constexpr int kUpdRows = kTileGroup;
    dd nrm = dd_zero();
    for (TileWalk tw = tile_walk(a.tiles); tw.t < tw.end; tw.t += kUpdRows * tw.step) {
        uint32_t r[kUpdRows];
        bool ok[kUpdRows];
        double acc[kUpdRows], o[kUpdRows];
#pragma unroll
        for (int u = 0; u < kUpdRows; ++u) {
            const uint32_t tt = tw.t + (uint32_t)u * tw.step;
            const uint32_t i = tt * kBlock + threadIdx.x;
            ok[u] = tt < tw.end && i < a.nv;
            r[u] = ok[u] ? i : 0u;
            o[u] = a.out[r[u]];
            acc[u] = 0.0;
        }
        int k = a.istart;
        for (; k + kDotChunk <= a.j + 1; k += kDotChunk) {
            double v[kUpdRows][kDotChunk];
#pragma unroll
            for (int u = 0; u < kUpdRows; ++u)
#pragma unroll
                for (int q = 0; q < kDotChunk; ++q) v[u][q] = basis<B>(a.V64, (const B *)a.Vb, a.nv, k + q, r[u]);
#pragma unroll
            for (int u = 0; u < kUpdRows; ++u)
#pragma unroll
                for (int q = 0; q < kDotChunk; ++q) acc[u] += v[u][q] * h[k + q];
        }
        for (; k <= a.j; ++k) {
            double v[kUpdRows];
#pragma unroll
            for (int u = 0; u < kUpdRows; ++u) v[u] = basis<B>(a.V64, (const B *)a.Vb, a.nv, k, r[u]);
#pragma unroll
            for (int u = 0; u < kUpdRows; ++u) acc[u] += v[u] * h[k];
        }
#pragma unroll
        for (int u = 0; u < kUpdRows; ++u) {
            if (!ok[u]) continue;
            const double w = -1.0 * acc[u] + o[u];
            a.out[r[u]] = w;
            nrm = dd_add_prod(nrm, w, w);
        }
    }
    const dd t = block_sum_dd(nrm, red);
    if (threadIdx.x == 0) {
        a.npart[2 * blockIdx.x] = t.s;
        a.npart[2 * blockIdx.x + 1] = t.c;
    }
}

__global__ void __launch_bounds__(kBlock) k_arn_norm(ArnArgs a, int M2, int G)
{
    KState *st = a.st;
    if (!st->active || st->broke) return;
    __shared__ double red[kRedLen];
    const double s = sqrt(dd_value(sum_dd_partials(a.npart, G, red)));
    if (threadIdx.x == 0) {
        if (s < st->btol) {  // happy breakdown (src/KryExpvFSP.cu:90-98)
            st->k1 = 0;
            st->mb = a.j + 1;
            st->tau = st->t_final - st->t_now;
            st->broke = 1;
            st->n_breakdown += 1;
        } else {
            st->H[(a.j + 1) + a.j * M2] = s;
            st->sigma[a.j + 1] = 1.0 / s;
        }
    }
}

// ||v||^2 partials over the vector.
__global__ void __launch_bounds__(kBlock) k_norm_partials(const double *__restrict__ v, uint32_t nv,
                                                         double *partials, const KState *st)
{
    if (st && st->t_now >= st->t_final) return;
    __shared__ double red[kRedLen];
    dd s = dd_zero();
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nv; i += gridDim.x * kBlock)
        s = dd_add_prod(s, v[i], v[i]);
    s = block_sum_dd(s, red);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = s.s;
        partials[2 * blockIdx.x + 1] = s.c;
    }
}

// Step prologue (src/KryExpvFSP.cu:41-65) for one state: kry_step_begin.
__global__ void __launch_bounds__(kBlock) k_step_begin(KState *st, const double *partials, int G,
                                                      int m, int M2)
{
    __shared__ double red[kRedLen];
    // the host enforces max_nstep with its own message
    kry_step_begin(st, partials, G, m, M2, 0x7fffffffffffffffLL, red);
}

// Error estimate, reject loop and step update for one state: kry_control.
__global__ void __launch_bounds__(kBlock) k_expm_control(KState *st, int m, int max_reject)
{
    FSP_DYNAMIC_LDS(smem);
    kry_control(st, m, max_reject, smem);
}

// v = beta * V[:, 0:mx] F0 (src/KryExpvFSP.cu:160-166).
template <typename B>
__global__ void __launch_bounds__(kBlock) k_solution(const KState *st, const double *__restrict__ V64,
                                                    const B *__restrict__ Vb, uint32_t nv,
                                                    double *__restrict__ v)
{
    if (!st->active) return;
    const int mx = st->mx;
    const double beta = st->beta;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nv; i += gridDim.x * kBlock) {
        double acc = 0.0;
        int k = 0;
        for (; k + kDotChunk <= mx; k += kDotChunk) {
            double b[kDotChunk];
#pragma unroll
            for (int q = 0; q < kDotChunk; ++q) b[q] = basis<B>(V64, Vb, nv, k + q, i);
#pragma unroll
            for (int q = 0; q < kDotChunk; ++q) acc += b[q] * st->F0[k + q];
        }
        for (; k < mx; ++k) acc += basis<B>(V64, Vb, nv, k, i) * st->F0[k];
        v[i] = beta * acc;
    }
}

// Dense output inside the last accepted step (t_old, t_old + tau]: c = the
// first column of exp(delta H) (delta <= tau) over the step's own mx basis
// vectors, the step's Pade in place of exp(tau H).
__global__ void __launch_bounds__(kBlock) k_dense_coef(const KState *st, int m, double delta,
                                                      double *coef)
{
    FSP_DYNAMIC_LDS(smem);
    const int M2 = m + 2, NN = M2 * M2;
    double *A = smem, *X = A + NN, *T = X + NN, *Nm = T + NN, *Dm = Nm + NN;
    __shared__ double red[kBlock / kWave];
    for (int e = threadIdx.x; e < NN; e += blockDim.x) A[e] = delta * st->H[e];
    __syncthreads();
    expm_block(M2, A, X, T, Nm, Dm, red);
    const int mx = st->mx;
    for (int k = threadIdx.x; k < M2; k += blockDim.x) coef[k] = k < mx ? Nm[k] : 0.0;
}

// out = beta V[:, 0:mx] c (the solution formula with the dense coefficients).
template <typename B>
__global__ void __launch_bounds__(kBlock) k_dense_out(const KState *st, const double *__restrict__ V64,
                                                     const B *__restrict__ Vb, uint32_t nv,
                                                     const double *__restrict__ coef,
                                                     double *__restrict__ out)
{
    const int mx = st->mx;
    const double beta = st->beta;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nv; i += gridDim.x * kBlock) {
        double acc = 0.0;
        int k = 0;
        for (; k + kDotChunk <= mx; k += kDotChunk) {
            double b[kDotChunk];
#pragma unroll
            for (int q = 0; q < kDotChunk; ++q) b[q] = basis<B>(V64, Vb, nv, k + q, i);
#pragma unroll
            for (int q = 0; q < kDotChunk; ++q) acc += b[q] * coef[k + q];
        }
        for (; k < mx; ++k) acc += basis<B>(V64, Vb, nv, k, i) * coef[k];
        out[i] = beta * acc;
    }
}

// ---- callback operator path: explicit scale / zero / dots kernels ----
template <typename B>
__global__ void __launch_bounds__(kBlock) k_cb_scale(ArnArgs a)
{
    if (!a.st->active || a.st->broke) return;
    const double sig = a.st->sigma[a.j];
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < a.nv; i += gridDim.x * kBlock) {
        store_basis<B>(a.V64, (B *)a.Vb, a.nv, a.j, i, kry_scale<B>(a.in[i], sig, a.j));
        a.out[i] = 0.0;
    }
}

template <typename B, int NKB>
__global__ void __launch_bounds__(kBlock) k_cb_dots(ArnArgs a)
{
    if (arn_skip<B>(a)) return;
    __shared__ double red[kArnRedLen];
    dd acc[NKB];
    acc_zero(acc);
    const int nk = a.mode == 0 ? a.j - a.istart + 1 : 1;
    for (TileWalk w = tile_walk(a.tiles); w.t < w.end; w.t += w.step) {
        const uint32_t i = w.t * kBlock + threadIdx.x;
        if (i >= a.nv) continue;
        const double y = a.out[i];
#pragma unroll
        for (int kk = 0; kk < NKB; ++kk) {
            if (kk < nk) {
                const double vk = a.mode == 0 ? basis<B>(a.V64, (const B *)a.Vb, a.nv, a.istart + kk, i) : y;
                acc[kk] = dd_add_prod(acc[kk], y, vk);
            }
        }
    }
    arn_end<NKB>(a, acc, dd_zero(), red);
}

}  // namespace

// ======================================================================= host
// ------------------------------------------------------------ profiling
// Event pair k brackets the k-th launch of the step being enqueued.
static fsp_status prof_mark(ExpvImpl &E, int cls, double bytes)
{
    if (!E.profile) return FSP_OK;
    const size_t k = (size_t)E.pcount;
    while (E.pev.size() < 2 * (k + 1)) {
        // timing only (read after the step's stream synchronisation): no
        // system-scope fence at record, whose cache writeback would be
        // counted in the bracketed launch
        hipEvent_t ev = nullptr;
        FSP_HIP(hipEventCreateWithFlags(&ev, hipEventDisableSystemFence));
        E.pev.push_back(ev);
    }
    if (E.pcls.size() < k + 1) {
        E.pcls.resize(k + 1);
        E.pbytes.resize(k + 1);
    }
    E.pcls[k] = cls;
    E.pbytes[k] = bytes;
    FSP_HIP(hipEventRecord(E.pev[2 * k], E.s));
    return FSP_OK;
}

static fsp_status prof_done(ExpvImpl &E)
{
    if (!E.profile) return FSP_OK;
    FSP_HIP(hipEventRecord(E.pev[2 * (size_t)E.pcount + 1], E.s));
    ++E.pcount;
    return FSP_OK;
}

// Adds the recorded launches of the last step to the totals.  Launches of
// an inactive step (the final time reached, or a failed solve) return at
// once and are not counted; after a happy breakdown the remaining Arnoldi
// launches do nothing and are not counted either.
static fsp_status prof_collect(ExpvImpl &E, bool active, int mb)
{
    if (!E.profile || E.pcount == 0) return FSP_OK;
    FSP_HIP(hipEventSynchronize(E.pev[2 * (size_t)E.pcount - 1]));
    int arn = 0;
    for (int k = 0; k < E.pcount; ++k) {
        const int c = E.pcls[k];
        bool count = active;
        if (c == FSP_PROF_ARNOLDI || c == FSP_PROF_DOTS || c == FSP_PROF_UPDATE) {
            // launch j of its class; the avnorm launch (j = m) runs unless
            // the step broke down
            const int j = c == FSP_PROF_ARNOLDI ? arn++ : (c == FSP_PROF_DOTS ? arn - 1 : arn - 1);
            count = count && (mb == E.m ? true : j < mb);
        }
        if (!count) continue;
        float ms = 0.0f;
        FSP_HIP(hipEventElapsedTime(&ms, E.pev[2 * (size_t)k], E.pev[2 * (size_t)k + 1]));
        E.prof_n[c] += 1;
        E.prof_ms[c] += ms;
        E.prof_bytes[c] += E.pbytes[k];
    }
    E.pcount = 0;
    return FSP_OK;
}

// Algorithmic bytes of the step's kernels: every vector element and stored
// matrix value the operation needs once (cached re-reads not counted).
// Basis vector k has elem_bytes(k) bytes per entry.
static double basis_elem(const ExpvImpl &E, int k) { return (!E.mixed || k == 0) ? 8.0 : 4.0; }

static double arn_bytes(const ExpvImpl &E, int j, int istart, int mode)
{
    const double n = (double)E.n;
    const double apply = E.A ? mat_bytes_per_apply(*E.A) : 16.0 * n;  // matrix + x + y
    double b = apply + basis_elem(E, j) * n;                            // + V_j store
    if (mode == 0) {
        for (int k = istart; k <= j; ++k) b += basis_elem(E, k) * n;    // dots V_k . w
    } else {
        b -= 8.0 * n;                                                   // norm only, no w store
    }
    return b;
}

static double update_bytes(const ExpvImpl &E, int j, int istart)
{
    double b = 16.0 * (double)E.n;                                      // w read and written
    for (int k = istart; k <= j; ++k) b += basis_elem(E, k) * (double)E.n;
    return b;
}

static fsp_status expv_alloc(ExpvImpl &E)
{
    const size_t nv = (size_t)E.n;
    FSP_HIP(hipMalloc(&E.d_st, sizeof(KState)));
    FSP_HIP(hipHostMalloc(&E.h_st, sizeof(KState), hipHostMallocDefault));
    std::memset(E.h_st, 0, sizeof(KState));
    if (E.mixed) {
        FSP_HIP(hipMalloc(&E.V64, sizeof(double) * nv));
        FSP_HIP(hipMalloc(&E.V32, sizeof(float) * nv * (size_t)(E.m + 1)));
    } else {
        FSP_HIP(hipMalloc(&E.V64, sizeof(double) * nv * (size_t)(E.m + 2)));
    }
    FSP_HIP(hipMalloc(&E.W[0], sizeof(double) * nv));
    FSP_HIP(hipMalloc(&E.W[1], sizeof(double) * nv));
    E.tiles = ((int64_t)nv + kBlock - 1) / kBlock;
    // row-kernel grid: 4 workgroups per CU measured fastest on the 786K-state
    // solve (1.34 s against 1.43 s at 768, 1.46 s at 1280 and 1.77 s at
    // 2048; the finalizers and the norm prologue sum one partial per
    // workgroup); FSP_ARN_GRID overrides the cap for timing sweeps
    int cap = 1024;
    if (const char *g = std::getenv("FSP_ARN_GRID")) cap = std::max(8, std::atoi(g));
    E.G = grid_for(E.tiles, cap);
    // row-kernel slots, then the update kernel's |w|^2 partials
    FSP_HIP(hipMalloc(&E.partials, 2 * sizeof(double) * (size_t)E.G * (kArnSlots + 1)));
    KState &h = *E.h_st;
    h.t_final = 0.0;
    h.t_now = 0.0;
    h.k1 = 2;
    h.mb = E.m;
    h.tol = E.tol;
    h.anorm = E.anorm;
    // fact = ((m+1)/e)^(m+1) sqrt(2 pi (m+1)) with pi written 3.1416 as in
    // src/KryExpvFSP.cu:48; computed once on the host.
    h.fact = std::pow((E.m + 1) / std::exp(1.0), E.m + 1) * std::sqrt(2 * (3.1416) * (E.m + 1));
    for (int k = kP10Lo; k <= kP10Hi; ++k) {
        char buf[16];
        std::snprintf(buf, sizeof buf, "1e%d", k);
        h.p10[k - kP10Lo] = std::strtod(buf, nullptr);
    }
    FSP_HIP(hipMalloc(&E.tau_log, sizeof(double) * (size_t)kTauLogCap));
    FSP_HIP(hipMalloc(&E.dcoef, sizeof(double) * (size_t)(E.m + 2)));
    h.tau_log = E.tau_log;
    h.tau_cap = kTauLogCap;
    FSP_HIP(copy_sync(E.d_st, &h, sizeof(KState), hipMemcpyHostToDevice, E.s));
    return FSP_OK;
}

static void expv_free(ExpvImpl &E)
{
    if (E.gexec) (void)hipGraphExecDestroy(E.gexec);
    if (E.graph) (void)hipGraphDestroy(E.graph);
    if (E.cap) (void)hipStreamDestroy(E.cap);
    void *ptrs[] = {E.d_st, E.V64, E.V32, E.W[0], E.W[1], E.partials, E.tau_log, E.dcoef};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (E.h_st) (void)hipHostFree(E.h_st);
    for (hipEvent_t ev : E.pev) (void)hipEventDestroy(ev);
}

template <typename V, typename B>
static void launch_arn_hyb(const HybF<V> &A, const ArnArgs &a, int G, size_t red_b, hipStream_t s)
{
#define FSP_ARN_HYB_CASE(K)                                          \
    case K:                                                          \
        k_arn_hyb<K, V, B><<<G, kBlock, red_b, s>>>(A, a);           \
        return;
    switch (A.U) {
        FSP_ARN_HYB_CASE(1) FSP_ARN_HYB_CASE(2) FSP_ARN_HYB_CASE(3) FSP_ARN_HYB_CASE(4)
        FSP_ARN_HYB_CASE(5) FSP_ARN_HYB_CASE(6) FSP_ARN_HYB_CASE(7) FSP_ARN_HYB_CASE(8)
        FSP_ARN_HYB_CASE(9) FSP_ARN_HYB_CASE(10) FSP_ARN_HYB_CASE(11) FSP_ARN_HYB_CASE(12)
    default:
        k_arn_hyb<0, V, B><<<G, kBlock, red_b, s>>>(A, a);
    }
#undef FSP_ARN_HYB_CASE
}

template <typename B>
static fsp_status launch_arn_spmv(ExpvImpl &E, ArnArgs &a)
{
    MatImpl &M = *E.A;
    const size_t red_b = arn_lds_bytes(0);
    const bool mixed = M.prec == FSP_MIXED;
    if (M.format == FSP_CSR) {
        if (mixed) {
            CsrF<float> A = csr_view<float>(M);
            k_arn_csr<float, B><<<E.G, kBlock, arn_lds_bytes(sizeof(double) * A.cap), E.s>>>(A, a);
        } else {
            CsrF<double> A = csr_view<double>(M);
            k_arn_csr<double, B><<<E.G, kBlock, arn_lds_bytes(sizeof(double) * A.cap), E.s>>>(A, a);
        }
    } else if (M.format == FSP_HYB) {
        if (mixed)
            launch_arn_hyb<float, B>(hyb_view<float>(M), a, E.G, red_b, E.s);
        else
            launch_arn_hyb<double, B>(hyb_view<double>(M), a, E.G, red_b, E.s);
    } else {
        KronPlan P;
        KronSlotPlan SP;
        const bool plan0 = !M.hm.prop_table_dev && kron_plan(M.hm, P);
        const bool slots = M.kron_slots && kron_slot_plan(M.hm, SP);
        if (slots && mixed) {
            KronF<float> A = kron_view<float>(M, -1, 0.0);
            const size_t lds = arn_lds_bytes(sizeof(float) * (size_t)SP.tab_size);
            FSP_SWITCH_D(M.hm.dm.d, (k_arn_kron_slot<D_, float, B><<<E.G, kBlock, lds, E.s>>>(
                                        A, SP, M.d_slot_tab_f, a)));
        } else if (slots) {
            KronF<double> A = kron_view<double>(M, -1, 0.0);
            const size_t lds = arn_lds_bytes(sizeof(double) * (size_t)SP.tab_size);
            FSP_SWITCH_D(M.hm.dm.d, (k_arn_kron_slot<D_, double, B><<<E.G, kBlock, lds, E.s>>>(
                                        A, SP, M.d_slot_tab, a)));
        } else if (mixed && M.hm.has_factors) {
            KronF<float> A = kron_view<float>(M, -1, 0.0);
            const size_t tb = sizeof(float) * (size_t)A.tab_size;
            const int use_lds = (A.tab_g && tb <= kLdsTableMax) ? 1 : 0;
            const size_t lds = arn_lds_bytes(use_lds ? tb : 0);
            if (plan0 && use_lds) {
                FSP_SWITCH_D(M.hm.dm.d, (k_arn_kron_plan<D_, float, B><<<E.G, kBlock, lds, E.s>>>(A, P, a)));
            } else {
                FSP_SWITCH_D(M.hm.dm.d, (k_arn_kron<D_, float, B><<<E.G, kBlock, lds, E.s>>>(A, a, use_lds)));
            }
        } else {
            KronF<double> A = kron_view<double>(M, -1, 0.0);
            const size_t tb = sizeof(double) * (size_t)A.tab_size;
            const int use_lds = (A.tab_g && tb <= kLdsTableMax) ? 1 : 0;
            const size_t lds = arn_lds_bytes(use_lds ? tb : 0);
            if (plan0 && use_lds) {
                FSP_SWITCH_D(M.hm.dm.d, (k_arn_kron_plan<D_, double, B><<<E.G, kBlock, lds, E.s>>>(A, P, a)));
            } else {
                FSP_SWITCH_D(M.hm.dm.d, (k_arn_kron<D_, double, B><<<E.G, kBlock, lds, E.s>>>(A, a, use_lds)));
            }
        }
    }
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}


template <typename B>
static void launch_cb_dots(ExpvImpl &E, ArnArgs &a)
{
    const int nk = a.mode == 0 ? a.j - a.istart + 1 : 1;
    FSP_SWITCH_NK(nk, (k_cb_dots<B, NK_><<<E.G, kBlock, 0, E.s>>>(a)));
}

// Enqueue one full step (no host synchronisation).
template <typename B>
static fsp_status enqueue_step(ExpvImpl &E, double *d_v)
{
    const int m = E.m, M2 = m + 2;
    const uint32_t nv = (uint32_t)E.n;
    const uint32_t nbox = E.A ? (uint32_t)E.A->n : nv;
    const int with_sink = (E.A && E.A->with_sink) ? 1 : 0;
    FSP_TRY(prof_mark(E, FSP_PROF_PROLOGUE, 8.0 * nv));
    k_norm_partials<<<E.G, kBlock, 0, E.s>>>(d_v, nv, E.partials, E.d_st);
    k_step_begin<<<1, kBlock, 0, E.s>>>(E.d_st, E.partials, E.G, m, M2);
    FSP_TRY(prof_done(E));
    ArnArgs a;
    a.st = E.d_st;
    a.V64 = E.V64;
    a.Vb = E.mixed ? (void *)E.V32 : (void *)E.V64;
    a.partials = E.partials;
    a.n = nbox;
    a.nv = nv;
    a.tiles = (uint32_t)E.tiles;
    a.G = E.G;
    a.npart = E.partials + 2 * (size_t)E.G * kArnSlots;
    a.M2 = M2;
    int istart = 0;
    for (int j = 0; j <= m; ++j) {
        const bool last = (j == m);
        if (!last && E.iop) istart = (j >= E.q_iop - 1) ? j - E.q_iop + 1 : 0;
        a.in = (j == 0) ? d_v : E.W[j & 1];
        a.out = E.W[(j + 1) & 1];
        a.j = j;
        a.istart = last ? 0 : istart;
        a.mode = last ? 1 : 0;
        const int nk = last ? 1 : j - istart + 1;
        if (E.A) {
            FSP_TRY(prof_mark(E, FSP_PROF_ARNOLDI, arn_bytes(E, j, a.istart, a.mode)));
            FSP_TRY(launch_arn_spmv<B>(E, a));
            FSP_TRY(prof_done(E));
            FSP_TRY(prof_mark(E, FSP_PROF_DOTS, 16.0 * E.G * nk));
            k_arn_dots<B><<<nk, kBlock, 0, E.s>>>(a, with_sink, M2);
            FSP_TRY(prof_done(E));
        } else {
            k_cb_scale<B><<<E.G, kBlock, 0, E.s>>>(a);
            FSP_HIP(hipGetLastError());
            E.cb(E.ctx, a.V64 + (size_t)j * nv, a.out);
            launch_cb_dots<B>(E, a);
            k_arn_dots<B><<<nk, kBlock, 0, E.s>>>(a, 0, M2);
        }
        FSP_HIP(hipGetLastError());
        if (last) break;
        // with an FP32 basis the update forms the Gram-corrected coefficients
        const size_t upd_lds = E.mixed ? kGramLdsBytes + sizeof(double) * (kMaxKrylov + 2) : 0;
        FSP_TRY(prof_mark(E, FSP_PROF_UPDATE, update_bytes(E, j, istart)));
        k_arn_update<B><<<E.G, kBlock, upd_lds, E.s>>>(a, M2);
        FSP_TRY(prof_done(E));
        // a fused operator's next row kernel finishes |w| itself
        if (!E.A) k_arn_norm<<<1, kBlock, 0, E.s>>>(a, M2, E.G);
        FSP_HIP(hipGetLastError());
    }
    const size_t lds = sizeof(double) * 5 * (size_t)M2 * M2;
    FSP_TRY(prof_mark(E, FSP_PROF_CONTROL, 0.0));
    k_expm_control<<<1, kBlock, lds, E.s>>>(E.d_st, m, E.max_reject);
    FSP_TRY(prof_done(E));
    // v = beta V F0 over the basis vectors the step used (mx <= m + 1; the
    // host counts m + 1)
    double sol_bytes = 8.0 * nv;
    for (int k = 0; k <= m; ++k) sol_bytes += basis_elem(E, k) * nv;
    FSP_TRY(prof_mark(E, FSP_PROF_SOLUTION, sol_bytes));
    k_solution<B><<<E.G, kBlock, 0, E.s>>>(E.d_st, E.V64, (const B *)(E.mixed ? (void *)E.V32 : (void *)E.V64), nv, d_v);
    FSP_TRY(prof_done(E));
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}

static fsp_status enqueue_any(ExpvImpl &E, double *d_v)
{
    return E.mixed ? enqueue_step<float>(E, d_v) : enqueue_step<double>(E, d_v);
}

static fsp_status launch_step(ExpvImpl &E, double *d_v)
{
    if (!E.A || !E.use_graph || E.profile) return enqueue_any(E, d_v);
    if (!E.gexec || E.graph_v != d_v) {
        if (E.gexec) (void)hipGraphExecDestroy(E.gexec);
        if (E.graph) (void)hipGraphDestroy(E.graph);
        E.gexec = nullptr;
        E.graph = nullptr;
        if (!E.cap) FSP_HIP(hipStreamCreateWithFlags(&E.cap, hipStreamNonBlocking));
        // the step's kernels are recorded on the private stream, then the
        // graph is launched on the caller's stream
        FSP_HIP(hipStreamSynchronize(E.s));
        hipStream_t user = E.s;
        E.s = E.cap;
        FSP_HIP(hipStreamBeginCapture(E.cap, hipStreamCaptureModeThreadLocal));
        fsp_status st = enqueue_any(E, d_v);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(E.cap, &g);
        E.s = user;
        FSP_TRY(st);
        if (ce != hipSuccess) return fail(FSP_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
        E.graph = g;
        FSP_HIP(hipGraphInstantiate(&E.gexec, g, nullptr, nullptr, 0));
        E.graph_v = d_v;
    }
    FSP_HIP(hipGraphLaunch(E.gexec, E.s));
    return FSP_OK;
}

// Device-to-host copy of the step state, whatever its status: the setters
// and getters use it, so a handle whose last solve failed can still report
// its statistics and be restarted.
static fsp_status read_state(ExpvImpl &E)
{
    FSP_HIP(hipMemcpyAsync(E.h_st, E.d_st, sizeof(KState), hipMemcpyDeviceToHost, E.s));
    FSP_HIP(hipStreamSynchronize(E.s));
    return FSP_OK;
}

// read_state plus the status check of a step or solve.
static fsp_status pull_state(ExpvImpl &E)
{
    FSP_TRY(read_state(E));
    if (E.h_st->status != FSP_OK)
        return fail((fsp_status)E.h_st->status, "Krylov reject loop exhausted max_reject");
    return FSP_OK;
}

static fsp_status push_state(ExpvImpl &E)
{
    FSP_HIP(hipMemcpyAsync(E.d_st, E.h_st, sizeof(KState), hipMemcpyHostToDevice, E.s));
    FSP_HIP(hipStreamSynchronize(E.s));
    return FSP_OK;
}

}  // namespace fsp

// ===================================================================== C ABI
using fsp::ExpvImpl;

struct fsp_expv_s {
    ExpvImpl impl;
};

extern "C" fsp_status fsp_expv_create(fsp_mat A, fsp_matvec_cb cb, void *cb_ctx, int64_t n,
                                      double t_final, int m, double tol, int iop, int q_iop,
                                      double anorm, void *stream, fsp_expv *out);

extern "C" fsp_status fsp_expv_create(fsp_mat A, fsp_matvec_cb cb, void *cb_ctx, int64_t n,
                                      double t_final, int m, double tol, int iop, int q_iop,
                                      double anorm, void *stream, fsp_expv *out)
{
    using namespace fsp;
    if (!out) return fail(FSP_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (m < 1 || m > kMaxKrylov) return fail(FSP_ERR_UNSUPPORTED, "m must be in 1..50");
    if (!A && !cb) return fail(FSP_ERR_INVALID, "need an fsp_mat or a matvec callback");
    if (n < 1) return fail(FSP_ERR_INVALID, "n < 1");
    if (n >= 2147483647LL) return fail(FSP_ERR_OVERFLOW, "vector too long");
    if (!(tol > 0.0) || !(anorm > 0.0)) return fail(FSP_ERR_INVALID, "tol and anorm must be > 0");
    if (iop && q_iop < 1) return fail(FSP_ERR_INVALID, "q_iop < 1");
    auto *h = new fsp_expv_s();
    ExpvImpl &E = h->impl;
    if (A) {
        FSP_TRY(fsp_mat_internal(A, &E.A));
        if (n != E.A->n_rows) {
            delete h;
            return fail(FSP_ERR_INVALID, "n differs from the matrix rows");
        }
        E.mixed = E.A->prec == FSP_MIXED;
    }
    E.cb = cb;
    E.ctx = cb_ctx;
    E.n = n;
    E.m = m;
    E.tol = tol;
    E.iop = iop;
    E.q_iop = q_iop;
    E.anorm = anorm;
    E.s = (hipStream_t)stream;
    // FSP_EXPV_GRAPH=0 launches each step's kernels directly instead of
    // replaying a captured HIP graph (for tools that do not follow graphs)
    if (const char *g = std::getenv("FSP_EXPV_GRAPH")) E.use_graph = std::atoi(g) != 0;
    fsp_status st = expv_alloc(E);
    if (st != FSP_OK) {
        expv_free(E);
        delete h;
        return st;
    }
    E.h_st->t_final = t_final;
    st = push_state(E);
    if (st != FSP_OK) {
        expv_free(E);
        delete h;
        return st;
    }
    *out = h;
    return FSP_OK;
}

extern "C" fsp_status fsp_expv_destroy(fsp_expv e)
{
    if (!e) return FSP_OK;
    fsp::expv_free(e->impl);
    delete e;
    return FSP_OK;
}

extern "C" fsp_status fsp_expv_step(fsp_expv e, double *d_v)
{
    using namespace fsp;
    if (!e || !d_v) return fail(FSP_ERR_INVALID, "NULL argument");
    ExpvImpl &E = e->impl;
    const int64_t steps0 = E.h_st->i_step;
    FSP_TRY(launch_step(E, d_v));
    const fsp_status st = pull_state(E);
    FSP_TRY(prof_collect(E, E.h_st->i_step > steps0, E.h_st->mb));
    return st;
}

extern "C" fsp_status fsp_expv_solve(fsp_expv e, double *d_v)
{
    using namespace fsp;
    if (!e || !d_v) return fail(FSP_ERR_INVALID, "NULL argument");
    ExpvImpl &E = e->impl;
    FSP_TRY(pull_state(E));
    // Steps are launched ahead of the host check (inactive steps do nothing),
    // up to 4, but no more than the remaining interval needs at the current
    // step size: a solve to a nearby output time then launches one step, not
    // three more that would only return.
    const int ahead = (E.A && !E.profile) ? 4 : 1;
    while (E.h_st->t_now < E.h_st->t_final) {
        const KState &h = *E.h_st;
        int launch = ahead;
        if (ahead > 1 && h.i_step > 0 && h.t_new > 0.0) {
            const double est = std::ceil((h.t_final - h.t_now) / h.t_new);
            if (std::isfinite(est)) launch = (int)std::max(1.0, std::min((double)ahead, est));
        }
        const int64_t steps0 = h.i_step;
        for (int k = 0; k < launch; ++k) FSP_TRY(launch_step(E, d_v));
        const fsp_status pst = pull_state(E);
        FSP_TRY(prof_collect(E, E.h_st->i_step > steps0, E.h_st->mb));
        FSP_TRY(pst);
        if (E.h_st->i_step > E.max_nstep)
            return fail(FSP_ERR_NOT_CONVERGED, "Krylov step count exceeded max_nstep");
    }
    return FSP_OK;
}

// Solve to the final time and write v(t_out[q]) to d_out[q] for output
// times t_now < t_out[0] < ... <= t_final, from the steps the plain solve
// takes (one step per launch, so that each step's basis and Hessenberg are
// still in place for the outputs inside it).  Internal (adjoint).
__attribute__((visibility("hidden"))) fsp_status fsp_expv_solve_dense(fsp_expv e, double *d_v,
                                                                     int n_out,
                                                                     const double *t_out,
                                                                     double *const *d_out)
{
    using namespace fsp;
    if (!e || !d_v || (n_out > 0 && (!t_out || !d_out))) return fail(FSP_ERR_INVALID, "NULL argument");
    ExpvImpl &E = e->impl;
    if (!E.A) return fail(FSP_ERR_UNSUPPORTED, "dense output needs a matrix operator");
    FSP_TRY(pull_state(E));
    for (int q = 0; q < n_out; ++q)
        if (!(t_out[q] > E.h_st->t_now) || t_out[q] > E.h_st->t_final || (q && !(t_out[q] > t_out[q - 1])))
            return fail(FSP_ERR_INVALID, "output times must increase within (t_now, t_final]");
    const uint32_t nv = (uint32_t)E.n;
    const size_t lds = sizeof(double) * 5 * (size_t)(E.m + 2) * (E.m + 2);
    int q = 0;
    while (E.h_st->t_now < E.h_st->t_final) {
        const double t0 = E.h_st->t_now;
        FSP_TRY(launch_step(E, d_v));
        FSP_TRY(pull_state(E));
        if (E.h_st->i_step > E.max_nstep)
            return fail(FSP_ERR_NOT_CONVERGED, "Krylov step count exceeded max_nstep");
        const double t1 = E.h_st->t_now;
        for (; q < n_out && t_out[q] <= t1; ++q) {
            if (t_out[q] == t1) {
                FSP_HIP(hipMemcpyAsync(d_out[q], d_v, sizeof(double) * nv, hipMemcpyDeviceToDevice, E.s));
                continue;
            }
            k_dense_coef<<<1, kBlock, lds, E.s>>>(E.d_st, E.m, t_out[q] - t0, E.dcoef);
            if (E.mixed)
                k_dense_out<float><<<E.G, kBlock, 0, E.s>>>(E.d_st, E.V64, E.V32, nv, E.dcoef, d_out[q]);
            else
                k_dense_out<double><<<E.G, kBlock, 0, E.s>>>(E.d_st, E.V64, E.V64, nv, E.dcoef, d_out[q]);
            FSP_HIP(hipGetLastError());
        }
    }
    FSP_HIP(hipStreamSynchronize(E.s));
    return FSP_OK;
}

extern "C" fsp_status fsp_expv_reset_time(fsp_expv e)
{
    using namespace fsp;
    if (!e) return fail(FSP_ERR_INVALID, "NULL argument");
    FSP_TRY(read_state(e->impl));
    e->impl.h_st->t_now = 0.0;
    return push_state(e->impl);
}

extern "C" fsp_status fsp_expv_restart(fsp_expv e, double t_final)
{
    using namespace fsp;
    if (!e) return fail(FSP_ERR_INVALID, "NULL argument");
    FSP_TRY(read_state(e->impl));
    KState &h = *e->impl.h_st;
    h.t_now = 0.0;
    h.t_final = t_final;
    h.t_new = 0.0;
    h.i_step = 0;
    h.k1 = 2;
    h.status = 0;
    h.n_reject = h.n_breakdown = h.n_matvec = 0;
    return push_state(e->impl);
}

extern "C" fsp_status fsp_expv_set_final_time(fsp_expv e, double t_final)
{
    using namespace fsp;
    if (!e) return fail(FSP_ERR_INVALID, "NULL argument");
    FSP_TRY(read_state(e->impl));
    e->impl.h_st->t_final = t_final;
    e->impl.h_st->k1 = 2;
    return push_state(e->impl);
}

extern "C" fsp_status fsp_expv_final_time_reached(fsp_expv e, int *reached)
{
    using namespace fsp;
    if (!e || !reached) return fail(FSP_ERR_INVALID, "NULL argument");
    FSP_TRY(read_state(e->impl));
    *reached = e->impl.h_st->t_now >= e->impl.h_st->t_final;
    return FSP_OK;
}

extern "C" fsp_status fsp_expv_get_stats(fsp_expv e, fsp_expv_stats *s)
{
    using namespace fsp;
    if (!e || !s) return fail(FSP_ERR_INVALID, "NULL argument");
    FSP_TRY(read_state(e->impl));
    const KState &h = *e->impl.h_st;
    s->t_now = h.t_now;
    s->t_new = h.t_new;
    s->tau_last = h.tau_last;
    s->err_last = h.err_loc;
    s->beta_last = h.beta;
    s->i_step = h.i_step;
    s->n_matvec = h.n_matvec;
    s->n_reject = h.n_reject;
    s->n_breakdown = h.n_breakdown;
    return FSP_OK;
}

extern "C" fsp_status fsp_expv_get_tau_history(fsp_expv e, double *h_taus, int64_t cap,
                                               int64_t *count)
{
    using namespace fsp;
    if (!e || !count) return fail(FSP_ERR_INVALID, "NULL argument");
    FSP_TRY(read_state(e->impl));
    const int64_t n = std::min<int64_t>(e->impl.h_st->i_step, kTauLogCap);
    *count = n;
    if (h_taus && cap > 0 && n > 0)
        FSP_HIP(copy_sync(h_taus, e->impl.tau_log, sizeof(double) * (size_t)std::min(n, cap),
                          hipMemcpyDeviceToHost, e->impl.s));
    return FSP_OK;
}

extern "C" fsp_status fsp_expv_set_basis_precision(fsp_expv e, int precision)
{
    using namespace fsp;
    if (!e) return fail(FSP_ERR_INVALID, "NULL argument");
    if (precision != FSP_F64 && precision != FSP_MIXED) return fail(FSP_ERR_INVALID, "unknown precision");
    ExpvImpl &E = e->impl;
    if (!E.A) return fail(FSP_ERR_UNSUPPORTED, "the basis precision needs a matrix operator");
    const bool mixed = precision == FSP_MIXED;
    if (mixed && E.A->prec != FSP_MIXED)
        return fail(FSP_ERR_UNSUPPORTED, "an FP32 basis needs a FSP_MIXED matrix");
    if (mixed == E.mixed) return FSP_OK;
    FSP_HIP(hipStreamSynchronize(E.s));
    const double t_final = E.h_st->t_final;
    expv_free(E);
    E.d_st = nullptr;
    E.h_st = nullptr;
    E.V64 = nullptr;
    E.V32 = nullptr;
    E.W[0] = E.W[1] = nullptr;
    E.partials = nullptr;
    E.tau_log = nullptr;
    E.dcoef = nullptr;
    E.graph = nullptr;
    E.gexec = nullptr;
    E.graph_v = nullptr;
    E.cap = nullptr;
    E.pev.clear();
    E.mixed = mixed;
    FSP_TRY(expv_alloc(E));
    E.h_st->t_final = t_final;
    return push_state(E);
}

extern "C" fsp_status fsp_expv_set_profiling(fsp_expv e, int enable)
{
    using namespace fsp;
    if (!e) return fail(FSP_ERR_INVALID, "NULL argument");
    ExpvImpl &E = e->impl;
    if (E.profile && !enable) FSP_HIP(hipStreamSynchronize(E.s));
    E.profile = enable != 0 && E.A != nullptr;
    E.pcount = 0;
    for (int c = 0; c < FSP_PROF_CLASSES; ++c) {
        E.prof_n[c] = 0;
        E.prof_ms[c] = 0.0;
        E.prof_bytes[c] = 0.0;
    }
    if (enable && !E.A) return fail(FSP_ERR_UNSUPPORTED, "profiling needs a matrix operator");
    return FSP_OK;
}

extern "C" fsp_status fsp_expv_get_profile(fsp_expv e, fsp_expv_profile *p)
{
    using namespace fsp;
    if (!e || !p) return fail(FSP_ERR_INVALID, "NULL argument");
    const ExpvImpl &E = e->impl;
    for (int c = 0; c < FSP_PROF_CLASSES; ++c) {
        p->launches[c] = E.prof_n[c];
        p->ms[c] = E.prof_ms[c];
        p->bytes[c] = E.prof_bytes[c];
    }
    return FSP_OK;
}

extern "C" fsp_status fsp_expv_set_iop(fsp_expv e, int iop, int q_iop)
{
    using namespace fsp;
    if (!e || (iop && q_iop < 1)) return fail(FSP_ERR_INVALID, "bad IOP options");
    ExpvImpl &E = e->impl;
    if ((E.iop != 0) == (iop != 0) && (!iop || E.q_iop == q_iop)) return FSP_OK;
    E.iop = iop ? 1 : 0;
    E.q_iop = q_iop;
    if (E.gexec) (void)hipGraphExecDestroy(E.gexec);
    if (E.graph) (void)hipGraphDestroy(E.graph);
    E.gexec = nullptr;
    E.graph = nullptr;
    return FSP_OK;
}

extern "C" fsp_status fsp_expv_set_limits(fsp_expv e, int max_reject, int64_t max_nstep)
{
    using namespace fsp;
    if (!e || max_reject < 1 || max_nstep < 1) return fail(FSP_ERR_INVALID, "bad limits");
    ExpvImpl &E = e->impl;
    E.max_reject = max_reject;
    E.max_nstep = max_nstep;
    if (E.gexec) (void)hipGraphExecDestroy(E.gexec);
    if (E.graph) (void)hipGraphDestroy(E.graph);
    E.gexec = nullptr;
    E.graph = nullptr;
    return FSP_OK;
}

namespace {
__global__ void __launch_bounds__(fsp::kBlock) k_expm_dense(int N, const double *Ag, double *Fg)
{
    FSP_DYNAMIC_LDS(smem);
    const int NN = N * N;
    double *A = smem, *X = A + NN, *T = X + NN, *Nm = T + NN, *Dm = Nm + NN;
    __shared__ double red[fsp::kBlock / fsp::kWave];
    for (int e = threadIdx.x; e < NN; e += blockDim.x) A[e] = Ag[e];
    __syncthreads();
    fsp::expm_block(N, A, X, T, Nm, Dm, red);
    for (int e = threadIdx.x; e < NN; e += blockDim.x) Fg[e] = Nm[e];
}
}  // namespace

extern "C" fsp_status fsp_expm_dense(int N, const double *d_A, double *d_F, void *stream)
{
    using namespace fsp;
    if (N < 1 || N > kMaxKrylov + 2) return fail(FSP_ERR_UNSUPPORTED, "N must be in 1..52");
    const size_t lds = sizeof(double) * 5 * (size_t)N * N;
    k_expm_dense<<<1, kBlock, lds, (hipStream_t)stream>>>(N, d_A, d_F);
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}

