// fsp_batch_lockstep.hip -- a batch of uneven FSP solves advanced in lockstep.
//
// Every instance runs its own KryExpvFSP::step state machine
// (src/KryExpvFSP.cu:38-182): own step sizes, reject loop, happy breakdown
// and final time.  What is shared is the launch sequence: one batch step is
// the same graph of kernels for all instances (step prologue, m + 1 Arnoldi
// iterations, Pade control, solution update), and each kernel covers every
// instance still running.  Row kernels walk a list of 1024-row tiles of all
// running instances (compacted at each step), so one launch fills the
// device however small the instances are; per-instance reductions are kept
// per tile (dot k of tile t at partials[2 (k T + t)]) and summed by one
// workgroup per (instance, dot).  The operator is each instance's generator
// in the HYB layout (FP64 diagonal + one ELL slot per unique net change),
// filled on the device from its model once per batch: a streamed row costs
// a few loads instead of the Kronecker factor products.  With the FP32
// basis (mixed) the slot values are FP32 and the basis is projected out with
// its Gram matrix (one pass over the basis), as in the single-solve engine.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fsp_assemble.h"
#include "fsp_batch_impl.h"
#include "fsp_krylov_dev.h"

namespace fsp {

struct LsInst {
    const DevModel *dm;
    const double *tab;      // FP64 factor tables
    const double *coefs;
    double *hdiag;          // HYB: diagonal (n)
    void *hvals;            // HYB: slot values [U][n], FP64 or FP32 (mixed)
    int *hoff;              // HYB: slot column offsets [U] (col = row + off)
    double *hsink;          // HYB: outflow to the sink (n), or null
    double *p;              // solution (nv)
    const double *p0;       // initial vector (nv)
    double *V64;            // FP64 basis V_0..V_m, or V_0 alone with an FP32 basis
    float *V32;             // FP32 basis V_1..V_m (mixed)
    double *W0, *W1;        // Arnoldi work vectors
    double t_final;
    int n, nv, with_sink, tab_size, tile0, ntiles, U, pad;
};

struct LsArgs {
    const LsInst *inst;
    KState *st;
    const int *tile_inst;   // instance of tile t
    const int *tile_r0;     // first row (within its instance) of tile t
    int *atiles;            // tiles of the running instances
    int *n_atiles;          // their count; [1]: running instances
    double *partials;       // 2 max(2m + 4, RPT) T doubles
    const int *owner;       // per instance: 1 when the persistent kernel solves it
    int T, count, m, M2;
    int j, istart, mode, nk; // Arnoldi iteration, first vector orthogonalized against,
                             // mode (0: dots, 1: avnorm), dots this iteration
};

namespace {

// Slots per tile: dots 0..nk-1, the sink at nk, with an FP32 basis the
// Gram column at nk+1..2nk.
constexpr int kLsSlots = 2 * kMaxKrylov + 4;
constexpr int kLsRedLen = 2 * kLsSlots * (kBlock / kWave);
constexpr int kLsChunk = 4;  // dots per pass over a tile's rows
// Tiles of RPT kBlock rows (RPT = 2, 4 or 8, the autotuner's tile height):
// each of the kBlock threads owns rows r0 + tid + u kBlock (u < RPT), so a
// tile's set-up and its dot reductions are shared by RPT rows per thread.
#define FSP_LS_RPT_SWITCH(rpt, CALL)                    \
    switch (rpt) {                                      \
    case 2: { constexpr int RPT_ = 2; CALL; } break;    \
    case 8: { constexpr int RPT_ = 8; CALL; } break;    \
    default: { constexpr int RPT_ = 4; CALL; } break;   \
    }

__device__ inline double *ls_in(const LsInst &I, int j) { return j == 0 ? I.p : ((j & 1) ? I.W1 : I.W0); }
__device__ inline double *ls_out(const LsInst &I, int j) { return ((j + 1) & 1) ? I.W1 : I.W0; }

template <typename B>
__device__ inline B *ls_vb(const LsInst &I)
{
    return std::is_same<B, double>::value ? (B *)(void *)I.V64 : (B *)(void *)I.V32;
}

// Wave butterflies of slot k into red (kLsRedLen); ls_store then writes,
// after one barrier, the tile's partials of slots 0..last: slot k of tile t
// at 2 (k T + t), and leaves a barrier for the caller's next use of red.
__device__ inline void ls_put_wave(double *red, int k, dd v)
{
    constexpr int NW = kBlock / kWave;
    const dd t = wave_sum_dd(v);
    if (threadIdx.x % kWave == 0) {
        const int w = threadIdx.x / kWave;
        red[2 * (k * NW + w)] = t.s;
        red[2 * (k * NW + w) + 1] = t.c;
    }
}

// FP64 sum (no compensation) into the same slot layout: used by the FP32
// basis, whose dots need no bit-exact match with the oracle.
__device__ inline void ls_put_wave_f(double *red, int k, double v)
{
    constexpr int NW = kBlock / kWave;
    const double t = wave_sum(v);
    if (threadIdx.x % kWave == 0) {
        const int w = threadIdx.x / kWave;
        red[2 * (k * NW + w)] = t;
        red[2 * (k * NW + w) + 1] = 0.0;
    }
}

__device__ inline void ls_store(const LsArgs &a, int t, int last, double *red)
{
    constexpr int NW = kBlock / kWave;
    __syncthreads();
    if ((int)threadIdx.x <= last) {
        const int k = threadIdx.x;
        dd s = dd_zero();
#pragma unroll
        for (int q = 0; q < NW; ++q) s = dd_merge(s, dd{red[2 * (k * NW + q)], red[2 * (k * NW + q) + 1]});
        a.partials[2 * ((size_t)k * a.T + t)] = s.s;
        a.partials[2 * ((size_t)k * a.T + t) + 1] = s.c;
    }
    __syncthreads();
}

// ---------------------------------------------------------- step prologue
// Tiles of the instances still running (t_now < t_final, no failure), in
// instance order; n_atiles[1] = their number of instances.  One workgroup.
__global__ void __launch_bounds__(1024) k_ls_compact(LsArgs a)
{
    __shared__ int flag[1024], ntl[1024], off[1024];
    __shared__ int base, inst_run;
    if (threadIdx.x == 0) {
        base = 0;
        inst_run = 0;
    }
    __syncthreads();
    for (int c0 = 0; c0 < a.count; c0 += 1024) {
        const int i = c0 + threadIdx.x;
        int f = 0, nt = 0;
        if (i < a.count) {
            const KState *st = a.st + i;
            f = st->t_now < st->t_final && st->status == 0 && !a.owner[i];
            nt = a.inst[i].ntiles;
        }
        flag[threadIdx.x] = f;
        ntl[threadIdx.x] = nt;
        __syncthreads();
        if (threadIdx.x == 0) {
            int s = base, r = 0;
            for (int q = 0; q < 1024 && c0 + q < a.count; ++q) {
                off[q] = s;
                if (flag[q]) {
                    s += ntl[q];
                    ++r;
                }
            }
            base = s;
            inst_run += r;
        }
        __syncthreads();
        for (int q = 0; q < 1024 && c0 + q < a.count; ++q) {
            if (!flag[q]) continue;
            const int tile0 = a.inst[c0 + q].tile0;
            for (int u = threadIdx.x; u < ntl[q]; u += blockDim.x) a.atiles[off[q] + u] = tile0 + u;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.n_atiles[0] = base;
        a.n_atiles[1] = inst_run;
    }
}

// |p|^2 per running tile (slot 0).
template <int RPT>
__global__ void __launch_bounds__(kBlock) k_ls_norm_partials(LsArgs a)
{
    __shared__ double red[2 * kBlock / kWave];
    const int na = a.n_atiles[0];
    for (int q = blockIdx.x; q < na; q += gridDim.x) {
        const int t = a.atiles[q];
        const LsInst &I = a.inst[a.tile_inst[t]];
        dd s = dd_zero();
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
            const int r = a.tile_r0[t] + (int)threadIdx.x + u * kBlock;
            if (r < I.nv) s = dd_add_prod(s, I.p[r], I.p[r]);
        }
        s = block_sum_dd(s, red);
        if (threadIdx.x == 0) {
            a.partials[2 * (size_t)t] = s.s;
            a.partials[2 * (size_t)t + 1] = s.c;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_ls_step_begin(LsArgs a, long long max_nstep)
{
    __shared__ double red[2 * kBlock / kWave];
    const int i = blockIdx.x;
    if (a.owner[i]) return;
    const LsInst &I = a.inst[i];
    kry_step_begin(a.st + i, a.partials + 2 * (size_t)I.tile0, I.ntiles, a.m, a.M2, max_nstep, red);
}

// --------------------------------------------------------------- Arnoldi
// Per running tile: V_j = w_j sigma_j, w_{j+1} = A V_j (Kronecker sum of the
// instance's model), partial dots V_k . w_{j+1} (k = istart..j) and the sink
// inflow; mode 1: |A V_m|^2 partials.
template <typename B>
__device__ inline HybF<B> ls_hyb(const LsInst &I, bool with_sink)
{
    HybF<B> A;
    A.diag = I.hdiag;
    A.slot_vals = (const B *)I.hvals;
    A.slot_off = I.hoff;
    A.sink_out = (with_sink && I.with_sink) ? I.hsink : nullptr;
    A.n = (uint32_t)I.n;
    A.U = I.U;
    return A;
}

template <typename B>
__device__ inline double ls_row(const LsInst &I, uint32_t r, const KryX<B> &X, dd &sink,
                                bool with_sink = true)
{
    return hyb_row<B>(ls_hyb<B>(I, with_sink), r, X, sink);
}

// Row sums of the thread's RPT rows of a tile (r0 + u kBlock) for a
// compile-time slot count, RG = 2 rows at a time (the loads of a group
// issued before their sums; two rows keep the kernel at 5 waves per SIMD
// in both precisions: four FP32 rows, with each slot value held for its
// zero test until its gather returns, took 108 VGPRs and the 512-instance
// mixed batch from 28.3 to 30.0 s); rows past the instance read row 0 and
// are discarded.  No sink terms (the caller adds them in row order).
template <int NU, typename B, int RPT>
__device__ inline void ls_rows(const LsInst &I, int r0, const KryX<B> &X, double (&y)[RPT])
{
    constexpr int RG0 = 2;
    constexpr int RG = RG0 < RPT ? RG0 : RPT;  // a group never runs past y[RPT]
    static_assert(RPT % RG == 0, "rows per thread must be a multiple of the group");
    const HybF<B> A = ls_hyb<B>(I, false);
#pragma unroll 1
    for (int u0 = 0; u0 < RPT; u0 += RG) {
#pragma unroll
        for (int u = u0; u < u0 + RG; ++u) {
            const uint32_t r = (uint32_t)(r0 + u * kBlock);
            dd unused = dd_zero();
            y[u] = hyb_row_n<NU>(A, r < (uint32_t)I.n ? r : 0u, X, unused);
        }
    }
}

// At least 5 waves per SIMD (at most 96 VGPRs): the compiler's own choice,
// 98 VGPRs at 4 rows per thread, held 4 waves and took the 512-instance
// batch 28.9 s (FP64) / 29.0 s (mixed) against 28.6 / 27.5 s.
template <typename B, int RPT>
__global__ void __launch_bounds__(kBlock, 5) k_ls_arn(LsArgs a)
{
    __shared__ double red[kLsRedLen];
    const int na = a.n_atiles[0];
    for (int q = blockIdx.x; q < na; q += gridDim.x) {
        const int t = a.atiles[q];
        const int ii = a.tile_inst[t];
        const KState *st = a.st + ii;
        if (!st->active || st->broke || (a.mode == 1 && st->k1 == 0)) continue;
        const LsInst &I = a.inst[ii];
        const int r0 = a.tile_r0[t] + (int)threadIdx.x;
        const double *in = ls_in(I, a.j);
        double *out = ls_out(I, a.j);
        const KryX<B> X{in, st->sigma[a.j], a.j};
        B *Vb = ls_vb<B>(I);
        const int nk = a.nk;
        dd sink = dd_zero(), nrm = dd_zero();
        double y[RPT];
        switch (I.U) {  // the instance's slot count, uniform over the tile
        case 4: ls_rows<4, B, RPT>(I, r0, X, y); break;
        case 5: ls_rows<5, B, RPT>(I, r0, X, y); break;
        case 6: ls_rows<6, B, RPT>(I, r0, X, y); break;
        default:
#pragma unroll
            for (int u = 0; u < RPT; ++u) {
                const uint32_t r = (uint32_t)(r0 + u * kBlock);
                dd unused = dd_zero();
                y[u] = r < (uint32_t)I.n ? ls_row<B>(I, r, X, unused, false) : 0.0;
            }
        }
        // stores and sink terms in row order
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
            const uint32_t r = (uint32_t)(r0 + u * kBlock);
            if (r < (uint32_t)I.n) {
                if (I.with_sink) sink = dd_add_prod(sink, I.hsink[r], X(r));
                store_basis<B>(I.V64, Vb, (size_t)I.nv, a.j, r, X(r));
                if (a.mode == 0)
                    out[r] = y[u];
                else
                    nrm = dd_add_prod(nrm, y[u], y[u]);
            } else {
                y[u] = 0.0;
            }
        }
        constexpr bool GRAM = gram_basis<B>();
        if (a.mode == 0 && !GRAM) {
            for (int c0 = 0; c0 < nk; c0 += kLsChunk) {
                dd acc[kLsChunk];
#pragma unroll
                for (int v = 0; v < kLsChunk; ++v) acc[v] = dd_zero();
#pragma unroll
                for (int u = 0; u < RPT; ++u) {
                    const uint32_t r = (uint32_t)(r0 + u * kBlock);
                    if (r >= (uint32_t)I.n) continue;
                    double vk[kLsChunk];
#pragma unroll
                    for (int v = 0; v < kLsChunk; ++v)
                        vk[v] = basis<B>(I.V64, Vb, (size_t)I.nv, min(a.istart + c0 + v, a.j), r);
#pragma unroll
                    for (int v = 0; v < kLsChunk; ++v)
                        if (c0 + v < nk) acc[v] = dd_add_prod(acc[v], y[u], vk[v]);
                }
#pragma unroll
                for (int v = 0; v < kLsChunk; ++v)
                    if (c0 + v < nk) ls_put_wave(red, c0 + v, acc[v]);
            }
        } else if (a.mode == 0) {
            // FP32 basis: the dots and the Gram column V_k . V_j from the same
            // loads, in plain FP64 (a fixed summation tree, so still
            // deterministic)
            constexpr int CH = kLsChunk;
            for (int c0 = 0; c0 < nk; c0 += CH) {
                double acc[CH], accg[CH];
#pragma unroll
                for (int v = 0; v < CH; ++v) acc[v] = accg[v] = 0.0;
#pragma unroll
                for (int u = 0; u < RPT; ++u) {
                    const uint32_t r = (uint32_t)(r0 + u * kBlock);
                    if (r >= (uint32_t)I.n) continue;
                    const double xj = basis<B>(I.V64, Vb, (size_t)I.nv, a.j, r);
                    double vk[CH];
#pragma unroll
                    for (int v = 0; v < CH; ++v)
                        vk[v] = basis<B>(I.V64, Vb, (size_t)I.nv, min(a.istart + c0 + v, a.j), r);
#pragma unroll
                    for (int v = 0; v < CH; ++v) {
                        acc[v] = fma(y[u], vk[v], acc[v]);
                        accg[v] = fma(xj, vk[v], accg[v]);
                    }
                }
#pragma unroll
                for (int v = 0; v < CH; ++v) {
                    if (c0 + v < nk) {
                        ls_put_wave_f(red, c0 + v, acc[v]);
                        ls_put_wave_f(red, nk + 1 + c0 + v, accg[v]);
                    }
                }
            }
        } else {
            ls_put_wave(red, 0, nrm);
        }
        ls_put_wave(red, nk, sink);
        ls_store(a, t, (GRAM && a.mode == 0) ? 2 * nk : nk, red);
    }
}

// Block (i, k) of count x nk: dot k of instance i from its tiles' partials
// in tile order, with the sink element (src/KryExpvFSP.cu:76-78); mode 1:
// avnorm and H(m+1, m) = 1 (:113-116).
template <typename B>
__global__ void __launch_bounds__(kBlock) k_ls_dots(LsArgs a)
{
    __shared__ double red[2 * kBlock / kWave];
    const int i = blockIdx.x / a.nk, k = blockIdx.x % a.nk;
    if (a.owner[i]) return;
    KState *st = a.st + i;
    if (!st->active || st->broke || (a.mode == 1 && st->k1 == 0)) return;
    const LsInst &I = a.inst[i];
    const double *in = ls_in(I, a.j);
    double *out = ls_out(I, a.j);
    B *Vb = ls_vb<B>(I);
    const size_t nv = (size_t)I.nv;
    double ys = 0.0, xs = 0.0;
    if (I.with_sink) {
        ys = dd_value(sum_dd_partials(a.partials + 2 * ((size_t)a.nk * a.T + I.tile0), I.ntiles, red));
        xs = kry_scale<B>(in[I.n], st->sigma[a.j], a.j);
    }
    dd sd = sum_dd_partials(a.partials + 2 * ((size_t)k * a.T + I.tile0), I.ntiles, red);
    dd sg = dd_zero();
    if (gram_basis<B>() && a.mode == 0)
        sg = sum_dd_partials(a.partials + 2 * ((size_t)(a.nk + 1 + k) * a.T + I.tile0), I.ntiles, red);
    if (threadIdx.x == 0) {
        if (a.mode == 0) {
            const int kk = a.istart + k;
            if (I.with_sink) {
                const double vk = (kk == a.j) ? xs : basis<B>(I.V64, Vb, nv, kk, (uint32_t)I.n);
                sd = dd_add_prod(sd, ys, vk);
                if (gram_basis<B>()) sg = dd_add_prod(sg, xs, vk);
            }
            st->H[kk + a.j * a.M2] = dd_value(sd);
            if (gram_basis<B>()) {
                const double g = dd_value(sg);
                st->G[kk + a.j * kGramLd] = g;
                st->G[a.j + kk * kGramLd] = g;
            }
            if (k == 0) {
                st->n_matvec += 1;
                if (I.with_sink) {
                    out[I.n] = ys;
                    store_basis<B>(I.V64, Vb, nv, a.j, (uint32_t)I.n, xs);
                }
            }
        } else {
            st->avnorm = sqrt(dd_value(dd_add_prod(sd, ys, ys)));
            st->H[(a.j + 1) + a.j * a.M2] = 1.0;  // src/KryExpvFSP.cu:115
            st->n_matvec += 1;
            if (I.with_sink) store_basis<B>(I.V64, Vb, nv, a.j, (uint32_t)I.n, xs);
        }
    }
}

// FP32 basis: per instance, H(., j) <- 2h - G h (gram_coefs).
__global__ void __launch_bounds__(kBlock) k_ls_gram_coef(LsArgs a)
{
    if (a.owner[blockIdx.x]) return;
    KState *st = a.st + blockIdx.x;
    if (!st->active || st->broke) return;
    FSP_DYNAMIC_LDS(lds);
    gram_coefs(st, a.istart, a.j, a.M2, lds);
}

// w_{j+1} -= V[istart..j] H(., j) (all nv entries, the sink element
// included), one row per thread: RPT workgroups per running tile
// (part u of tile t in workgroup slot RPT t + u), |w|^2 partial of each
// part at 2 (RPT t + u).
template <typename B, int RPT>
__global__ void __launch_bounds__(kBlock) k_ls_update(LsArgs a)
{
    __shared__ double red[2 * kBlock / kWave];
    const int nq = a.n_atiles[0] * RPT;
    for (int q = blockIdx.x; q < nq; q += gridDim.x) {
        const int t = a.atiles[q / RPT];
        const int sub = q % RPT;
        const int ii = a.tile_inst[t];
        const KState *st = a.st + ii;
        if (!st->active || st->broke) continue;
        const LsInst &I = a.inst[ii];
        const uint32_t r = (uint32_t)(a.tile_r0[t] + sub * kBlock + (int)threadIdx.x);
        const double *h = st->H + (size_t)a.j * a.M2;
        double *out = ls_out(I, a.j);
        const B *Vb = ls_vb<B>(I);
        dd nrm = dd_zero();
        if (r < (uint32_t)I.nv) {
            // k ascending as one sum; 8 basis loads issued together
            double acc = 0.0;
            int k = a.istart;
            for (; k + kLsChunk * 2 <= a.j + 1; k += kLsChunk * 2) {
                double v[kLsChunk * 2];
#pragma unroll
                for (int q = 0; q < kLsChunk * 2; ++q) v[q] = basis<B>(I.V64, Vb, (size_t)I.nv, k + q, r);
#pragma unroll
                for (int q = 0; q < kLsChunk * 2; ++q) acc += v[q] * h[k + q];
            }
            for (; k <= a.j; ++k) acc += basis<B>(I.V64, Vb, (size_t)I.nv, k, r) * h[k];
            const double w = -1.0 * acc + out[r];
            out[r] = w;
            nrm = dd_add_prod(nrm, w, w);
        }
        nrm = block_sum_dd(nrm, red);
        if (threadIdx.x == 0) {
            const size_t slot = (size_t)t * RPT + sub;
            a.partials[2 * slot] = nrm.s;
            a.partials[2 * slot + 1] = nrm.c;
        }
    }
}

// Per instance: s = |w_{j+1}| from the update's quarter-tile partials,
// happy breakdown (src/KryExpvFSP.cu:90-98) or H(j+1, j) = s and
// sigma_{j+1} = 1 / s.
template <int RPT>
__global__ void __launch_bounds__(kBlock) k_ls_norm(LsArgs a)
{
    __shared__ double red[2 * kBlock / kWave];
    const int i = blockIdx.x;
    if (a.owner[i]) return;
    KState *st = a.st + i;
    if (!st->active || st->broke) return;
    const LsInst &I = a.inst[i];
    const double s = sqrt(dd_value(
        sum_dd_partials(a.partials + 2 * (size_t)I.tile0 * RPT, I.ntiles * RPT, red)));
    if (threadIdx.x == 0) {
        if (s < st->btol) {
            st->k1 = 0;
            st->mb = a.j + 1;
            st->tau = st->t_final - st->t_now;
            st->broke = 1;
            st->n_breakdown += 1;
        } else {
            st->H[(a.j + 1) + a.j * a.M2] = s;
            st->sigma[a.j + 1] = 1.0 / s;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_ls_control(LsArgs a, int max_reject)
{
    if (a.owner[blockIdx.x]) return;
    FSP_DYNAMIC_LDS(smem);
    kry_control(a.st + blockIdx.x, a.m, max_reject, smem);
}

// p = beta V[:, 0:mx] F0 per running tile (src/KryExpvFSP.cu:160-166).
template <typename B, int RPT>
__global__ void __launch_bounds__(kBlock) k_ls_solution(LsArgs a)
{
    const int na = a.n_atiles[0];
    for (int q = blockIdx.x; q < na; q += gridDim.x) {
        const int t = a.atiles[q];
        const int ii = a.tile_inst[t];
        const KState *st = a.st + ii;
        if (!st->active) continue;
        const LsInst &I = a.inst[ii];
        const int r0 = a.tile_r0[t] + (int)threadIdx.x;
        const B *Vb = ls_vb<B>(I);
        const int mx = st->mx;
        const double beta = st->beta;
        const uint32_t nv = (uint32_t)I.nv;
        double acc[RPT];
#pragma unroll
        for (int u = 0; u < RPT; ++u) acc[u] = 0.0;
        for (int k = 0; k < mx; ++k) {
            const double fk = st->F0[k];
#pragma unroll
            for (int u = 0; u < RPT; ++u) {
                const uint32_t r = (uint32_t)(r0 + u * kBlock);
                if (r < nv) acc[u] += basis<B>(I.V64, Vb, (size_t)nv, k, r) * fk;
            }
        }
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
            const uint32_t r = (uint32_t)(r0 + u * kBlock);
            if (r < nv) I.p[r] = beta * acc[u];
        }
    }
}

// ------------------------------------------------------ set-up and export
// HYB generator of every instance from its model (the single-solve HYB
// fill's row values: fsp_assemble.h), one row per thread over all tiles.
template <int D, typename V, int RPT>
__global__ void __launch_bounds__(kBlock) k_ls_fill_hyb(LsArgs a, int mixed)
{
    for (int t = blockIdx.x; t < a.T; t += gridDim.x) {
        const LsInst &I = a.inst[a.tile_inst[t]];
        const DevModel *dm = I.dm;
        V *vals = (V *)I.hvals;
        for (int u = 0; u < RPT; ++u) {
            const uint32_t i = (uint32_t)(a.tile_r0[t] + (int)threadIdx.x + u * kBlock);
            if (i >= (uint32_t)I.n) continue;
            int xs[D];
            decode<D>(i, dm, xs);
            double dg, out;
            int has;
            diag_and_out<D>(dm, I.tab, nullptr, I.coefs, mixed, i, xs, dg, out, has);
            I.hdiag[i] = dg;
            if (I.with_sink) I.hsink[i] = out;
            for (int s = 0; s < I.U; ++s) {
                const size_t o = (size_t)s * I.n + i;
                const bool in = source_in_box<D>(dm->rx[dm->slot_rxn[dm->slot_ptr[s]]], xs, dm);
                vals[o] = (V)(in ? slot_value<D>(dm, I.tab, nullptr, I.coefs, s, i, xs, 1) : 0.0);
            }
        }
    }
}

// Per instance: p = p0 and a fresh step state (t_now = 0, first step size
// from the Expokit formula).
__global__ void __launch_bounds__(kBlock) k_ls_reset(LsArgs a)
{
    const int i = blockIdx.x;
    const LsInst &I = a.inst[i];
    for (int r = threadIdx.x; r < I.nv; r += blockDim.x) I.p[r] = I.p0[r];
    if (threadIdx.x == 0) {
        KState *st = a.st + i;
        st->t_now = 0.0;
        st->t_new = 0.0;
        st->t_final = I.t_final;
        st->i_step = 0;
        st->k1 = 2;
        st->status = 0;
        st->active = 0;
        st->broke = 0;
        st->n_reject = st->n_breakdown = st->n_matvec = 0;
    }
}

__global__ void __launch_bounds__(kBlock) k_ls_export(LsArgs a, BatchInst *out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.count; i += gridDim.x * blockDim.x) {
        const KState *st = a.st + i;
        out[i].steps = st->i_step;
        out[i].rejects = st->n_reject;
        out[i].status = st->status;
    }
}

__global__ void __launch_bounds__(kBlock) k_ls_running(LsArgs a, int *out)
{
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int c = 0;
    for (int i = threadIdx.x; i < a.count; i += blockDim.x) {
        const KState *st = a.st + i;
        c += st->t_now < st->t_final && st->status == 0 && !a.owner[i];
    }
    atomicAdd(&cnt, c);
    __syncthreads();
    if (threadIdx.x == 0) *out = cnt;
}

// ------------------------------------------------------ persistent kernel
// One workgroup solves one instance at a time, from t = 0 to its final
// time, with the lockstep path's per-instance data (HYB generator, basis,
// work vectors, KState) and its arithmetic (row sums, double-double dots
// and norms, Gram-corrected coefficients, kry_control), so an FP64 instance
// takes the same steps on either path.  Workgroups take instances from a
// work queue in decreasing order of estimated cost.  No workgroup ever
// waits for another (the queue's atomic increment is the only
// cross-workgroup operation), every loop is bounded, and a workgroup that
// has run for `budget` wall-clock ticks takes no further step: the instance
// it holds and every one it takes afterwards end with FSP_ERR_TIMEOUT, so a
// launch always returns.
struct PqArgs {
    const int *list;        // instances of the kernel, decreasing estimated cost
    int count;
    unsigned int *queue;    // next list position; 0 at launch
    int max_reject, iop, q_iop;
    long long max_nstep;
    long long budget;       // wall-clock ticks
};

constexpr int kPqChunk = 8;  // dots per pass over the instance's rows

// Slot k of red (kLsRedLen layout), merged over the waves in wave order.
__device__ inline dd pq_slot(const double *red, int k)
{
    constexpr int NW = kBlock / kWave;
    dd s = dd_zero();
#pragma unroll
    for (int q = 0; q < NW; ++q) s = dd_merge(s, dd{red[2 * (k * NW + q)], red[2 * (k * NW + q) + 1]});
    return s;
}

// Step prologue of one instance (k_ls_norm_partials + kry_step_begin):
// returns whether the instance takes a step (uniform over the workgroup).
__device__ inline bool pq_step_begin(const LsInst &I, KState *st, int m, int M2, const PqArgs &q,
                                     long long t0, double *red, int *flag)
{
    dd s2 = dd_zero();
    for (int r = threadIdx.x; r < I.nv; r += kBlock) s2 = dd_add_prod(s2, I.p[r], I.p[r]);
    const double s = dd_value(block_sum_dd(s2, red));
    if (threadIdx.x == 0) {
        int a = st->t_now < st->t_final && st->status == 0;
        if (a && (long long)(wall_clock64() - t0) > q.budget) {
            st->status = FSP_ERR_TIMEOUT;
            a = 0;
        }
        if (a && st->i_step >= q.max_nstep) {
            st->status = FSP_ERR_NOT_CONVERGED;
            a = 0;
        }
        st->active = a;
        if (a) {
            st->beta = sqrt(s);
            if (st->i_step == 0) {
                st->xm = 1.0 / (double)m;
                st->t_new = (1.0 / st->anorm) *
                            pow((st->fact * st->tol) / (4.0 * st->beta * st->anorm), st->xm);
                st->btol = st->anorm * st->tol;
            }
            st->mb = m;
            st->tau = fmin(st->t_final - st->t_now, st->t_new);
            st->broke = 0;
            st->sigma[0] = 1.0 / st->beta;
        }
        *flag = a;
    }
    __syncthreads();
    const bool a = *flag != 0;
    if (a)
        for (int k = threadIdx.x; k < M2 * M2; k += kBlock) st->H[k] = 0.0;
    __syncthreads();
    return a;
}

// Arnoldi iteration j of one instance (k_ls_arn + k_ls_dots): V_j, w_{j+1}
// and column j of H (mode 0), or avnorm (mode 1).
template <typename B>
__device__ inline void pq_arnoldi(const LsInst &I, KState *st, int j, int istart, int mode, int M2,
                                  double *red)
{
    const int nk = mode == 0 ? j - istart + 1 : 1;
    const double *in = ls_in(I, j);
    double *out = ls_out(I, j);
    const double sig = st->sigma[j];
    const KryX<B> X{in, sig, j};
    B *Vb = ls_vb<B>(I);
    const size_t nv = (size_t)I.nv;
    const HybF<B> A = ls_hyb<B>(I, false);
    dd sink = dd_zero(), nrm = dd_zero();
    for (int r = threadIdx.x; r < I.n; r += kBlock) {
        dd unused = dd_zero();
        const double y = hyb_row<B>(A, (uint32_t)r, X, unused);
        const double xr = X((uint32_t)r);
        if (I.with_sink) sink = dd_add_prod(sink, I.hsink[r], xr);
        store_basis<B>(I.V64, Vb, nv, j, (uint32_t)r, xr);
        if (mode == 0)
            out[r] = y;
        else
            nrm = dd_add_prod(nrm, y, y);
    }
    constexpr bool GRAM = gram_basis<B>();
    if (mode == 0) {
        // each thread re-reads the rows it wrote above
        for (int c0 = 0; c0 < nk; c0 += kPqChunk) {
            if (!GRAM) {
                dd acc[kPqChunk];
#pragma unroll
                for (int v = 0; v < kPqChunk; ++v) acc[v] = dd_zero();
                for (int r = threadIdx.x; r < I.n; r += kBlock) {
                    const double y = out[r];
                    double vk[kPqChunk];
#pragma unroll
                    for (int v = 0; v < kPqChunk; ++v)
                        vk[v] = basis<B>(I.V64, Vb, nv, min(istart + c0 + v, j), (uint32_t)r);
#pragma unroll
                    for (int v = 0; v < kPqChunk; ++v)
                        if (c0 + v < nk) acc[v] = dd_add_prod(acc[v], y, vk[v]);
                }
#pragma unroll
                for (int v = 0; v < kPqChunk; ++v)
                    if (c0 + v < nk) ls_put_wave(red, c0 + v, acc[v]);
            } else {
                double acc[kPqChunk], accg[kPqChunk];
#pragma unroll
                for (int v = 0; v < kPqChunk; ++v) acc[v] = accg[v] = 0.0;
                for (int r = threadIdx.x; r < I.n; r += kBlock) {
                    const double y = out[r];
                    const double xj = basis<B>(I.V64, Vb, nv, j, (uint32_t)r);
                    double vk[kPqChunk];
#pragma unroll
                    for (int v = 0; v < kPqChunk; ++v)
                        vk[v] = basis<B>(I.V64, Vb, nv, min(istart + c0 + v, j), (uint32_t)r);
#pragma unroll
                    for (int v = 0; v < kPqChunk; ++v) {
                        acc[v] = fma(y, vk[v], acc[v]);
                        accg[v] = fma(xj, vk[v], accg[v]);
                    }
                }
#pragma unroll
                for (int v = 0; v < kPqChunk; ++v) {
                    if (c0 + v < nk) {
                        ls_put_wave_f(red, c0 + v, acc[v]);
                        ls_put_wave_f(red, nk + 1 + c0 + v, accg[v]);
                    }
                }
            }
        }
    } else {
        ls_put_wave(red, 0, nrm);
    }
    ls_put_wave(red, nk, sink);
    __syncthreads();
    const int k = threadIdx.x;
    if (k < nk) {
        double ys = 0.0, xs = 0.0;
        if (I.with_sink) {
            ys = dd_value(pq_slot(red, nk));
            xs = kry_scale<B>(in[I.n], sig, j);
        }
        dd sd = pq_slot(red, k);
        if (mode == 0) {
            dd sg = GRAM ? pq_slot(red, nk + 1 + k) : dd_zero();
            const int kk = istart + k;
            if (I.with_sink) {
                const double vk = (kk == j) ? xs : basis<B>(I.V64, Vb, nv, kk, (uint32_t)I.n);
                sd = dd_add_prod(sd, ys, vk);
                if (GRAM) sg = dd_add_prod(sg, xs, vk);
            }
            st->H[kk + j * M2] = dd_value(sd);
            if (GRAM) {
                const double g = dd_value(sg);
                st->G[kk + j * kGramLd] = g;
                st->G[j + kk * kGramLd] = g;
            }
        } else {
            st->avnorm = sqrt(dd_value(dd_add_prod(sd, ys, ys)));
            st->H[(j + 1) + j * M2] = 1.0;  // src/KryExpvFSP.cu:115
        }
        if (k == 0) {
            st->n_matvec += 1;
            if (I.with_sink) {
                if (mode == 0) out[I.n] = ys;
                store_basis<B>(I.V64, Vb, nv, j, (uint32_t)I.n, xs);
            }
        }
    }
    __syncthreads();
}

// w_{j+1} -= V[istart..j] H(., j) and |w_{j+1}| (k_ls_update + k_ls_norm);
// returns whether the Krylov space broke down (uniform).
template <typename B>
__device__ inline bool pq_update(const LsInst &I, KState *st, int j, int istart, int M2, double *red)
{
    const double *h = st->H + (size_t)j * M2;
    double *out = ls_out(I, j);
    const B *Vb = ls_vb<B>(I);
    dd nrm = dd_zero();
    for (int r = threadIdx.x; r < I.nv; r += kBlock) {
        // k ascending as one sum; 8 basis loads issued together
        double acc = 0.0;
        int k = istart;
        for (; k + kLsChunk * 2 <= j + 1; k += kLsChunk * 2) {
            double v[kLsChunk * 2];
#pragma unroll
            for (int q = 0; q < kLsChunk * 2; ++q) v[q] = basis<B>(I.V64, Vb, (size_t)I.nv, k + q, (uint32_t)r);
#pragma unroll
            for (int q = 0; q < kLsChunk * 2; ++q) acc += v[q] * h[k + q];
        }
        for (; k <= j; ++k) acc += basis<B>(I.V64, Vb, (size_t)I.nv, k, (uint32_t)r) * h[k];
        const double w = -1.0 * acc + out[r];
        out[r] = w;
        nrm = dd_add_prod(nrm, w, w);
    }
    const double s = sqrt(dd_value(block_sum_dd(nrm, red)));
    const bool broke = s < st->btol;
    if (threadIdx.x == 0) {
        if (broke) {
            st->k1 = 0;
            st->mb = j + 1;
            st->tau = st->t_final - st->t_now;
            st->broke = 1;
            st->n_breakdown += 1;
        } else {
            st->H[(j + 1) + j * M2] = s;
            st->sigma[j + 1] = 1.0 / s;
        }
    }
    __syncthreads();
    return broke;
}

// p = beta V[:, 0:mx] F0 (k_ls_solution).
template <typename B>
__device__ inline void pq_solution(const LsInst &I, const KState *st)
{
    const B *Vb = ls_vb<B>(I);
    const int mx = st->mx;
    const double beta = st->beta;
    for (int r = threadIdx.x; r < I.nv; r += kBlock) {
        double acc = 0.0;
        for (int k = 0; k < mx; ++k) acc += basis<B>(I.V64, Vb, (size_t)I.nv, k, (uint32_t)r) * st->F0[k];
        I.p[r] = beta * acc;
    }
}

// The whole solve of instance ii.  smem: kLsRedLen doubles of reduction
// slots, then the control / Gram workspace.
template <typename B>
__device__ __forceinline__ void pq_solve(const LsArgs &a, const PqArgs &q, int ii, double *smem, long long t0)
{
    __shared__ int flag;
    double *red = smem;
    double *ctl = smem + kLsRedLen;
    KState *st = a.st + ii;
    const LsInst &I = a.inst[ii];
    const int m = a.m, M2 = a.M2;
    while (pq_step_begin(I, st, m, M2, q, t0, red, &flag)) {
        int istart = 0;
        for (int j = 0; j <= m; ++j) {
            const bool last = (j == m);
            if (last && st->k1 == 0) break;
            if (!last && q.iop) istart = (j >= q.q_iop - 1) ? j - q.q_iop + 1 : 0;
            pq_arnoldi<B>(I, st, j, last ? 0 : istart, last ? 1 : 0, M2, red);
            if (last) break;
            if (gram_basis<B>()) {
                gram_coefs(st, istart, j, M2, ctl);
                __syncthreads();
            }
            if (pq_update<B>(I, st, j, istart, M2, red)) break;
        }
        kry_control(st, m, q.max_reject, ctl);
        __syncthreads();
        pq_solution<B>(I, st);
        __syncthreads();
    }
}

template <typename B>
__global__ void __launch_bounds__(kBlock) k_pq(LsArgs a, PqArgs q)
{
    FSP_DYNAMIC_LDS(smem);
    __shared__ int idx;
    __shared__ long long t0;
    if (threadIdx.x == 0) t0 = (long long)wall_clock64();
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) idx = (int)atomicAdd(q.queue, 1u);
        __syncthreads();
        const int qi = idx;
        if (qi >= q.count) return;
        pq_solve<B>(a, q, q.list[qi], smem, t0);
    }
}

__global__ void k_pq_reset(unsigned int *queue) { *queue = 0u; }

}  // namespace

// Device state of the lockstep path (owned by BatchImpl::ls).
struct LsHost {
    LsInst *d_inst = nullptr;
    KState *d_st = nullptr;
    int *d_tile_inst = nullptr, *d_tile_r0 = nullptr, *d_atiles = nullptr, *d_cnt = nullptr;
    int *h_cnt = nullptr;
    double *d_partials = nullptr, *d_basis = nullptr, *d_work = nullptr;
    float *d_basis32 = nullptr;
    double *d_hdiag = nullptr, *d_hsink = nullptr;
    void *d_hvals = nullptr;
    int *d_hoff = nullptr;
    int T = 0, grid = 0, graph_grid = 0;
    int rpt = 0;                    // tile rows / kBlock of the tile tables
    std::vector<LsInst> h_inst;     // host copy of d_inst
    hipStream_t s = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // persistent kernel: instance ownership, its work list and queue, its
    // stream and graphs (the kernel alone for the hybrid solve; reset,
    // kernel and export for the all-persistent launch)
    int *d_owner = nullptr;
    std::vector<int> h_owner, h_list;
    int *d_pq_list = nullptr;
    unsigned int *d_pq_queue = nullptr;
    int pq_count = 0;
    hipStream_t pq_s = nullptr;
    hipEvent_t ev_reset = nullptr;
    hipGraph_t pq_graph = nullptr;
    hipGraphExec_t pq_gexec = nullptr;
    int pq_graph_grid = 0;
    hipGraph_t full_graph = nullptr;
    hipGraphExec_t full_gexec = nullptr;
    int full_graph_grid = 0;
    long long ticks_per_s = 0;
};

static LsArgs ls_args(const BatchImpl &B, const LsHost &L)
{
    LsArgs a;
    a.inst = L.d_inst;
    a.st = L.d_st;
    a.tile_inst = L.d_tile_inst;
    a.tile_r0 = L.d_tile_r0;
    a.atiles = L.d_atiles;
    a.n_atiles = L.d_cnt;
    a.partials = L.d_partials;
    a.owner = L.d_owner;
    a.T = L.T;
    a.count = B.count;
    a.m = B.m;
    a.M2 = B.m + 2;
    a.j = 0;
    a.istart = 0;
    a.mode = 0;
    a.nk = 1;
    return a;
}

void batch_lockstep_free(BatchImpl &B)
{
    LsHost *L = B.ls;
    if (!L) return;
    if (L->gexec) (void)hipGraphExecDestroy(L->gexec);
    if (L->graph) (void)hipGraphDestroy(L->graph);
    if (L->pq_gexec) (void)hipGraphExecDestroy(L->pq_gexec);
    if (L->pq_graph) (void)hipGraphDestroy(L->pq_graph);
    if (L->full_gexec) (void)hipGraphExecDestroy(L->full_gexec);
    if (L->full_graph) (void)hipGraphDestroy(L->full_graph);
    if (L->ev_reset) (void)hipEventDestroy(L->ev_reset);
    if (L->s) (void)hipStreamDestroy(L->s);
    if (L->pq_s) (void)hipStreamDestroy(L->pq_s);
    void *ptrs[] = {L->d_inst, L->d_st, L->d_tile_inst, L->d_tile_r0, L->d_atiles, L->d_cnt,
                    L->d_partials, L->d_basis, L->d_work, L->d_basis32, L->d_hdiag,
                    L->d_hsink, L->d_hvals, L->d_hoff, L->d_owner, L->d_pq_list, L->d_pq_queue};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (L->h_cnt) (void)hipHostFree(L->h_cnt);
    delete L;
    B.ls = nullptr;
}

// Tile tables for tiles of B.ls_rpt * kBlock rows over every instance's nv
// entries: each instance's first tile and tile count (uploaded with the
// instances), the running-tile list and the partials sized for them.  The
// captured batch step holds the tile count, so it is recaptured.
static fsp_status ls_tiles(BatchImpl &B, LsHost &L)
{
    const int rpt = (B.ls_rpt == 2 || B.ls_rpt == 8) ? B.ls_rpt : 4;
    const int rows = rpt * kBlock;
    std::vector<int> tinst, tr0;
    for (int i = 0; i < B.count; ++i) {
        LsInst &I = L.h_inst[i];
        I.tile0 = (int)tinst.size();
        I.ntiles = (I.nv + rows - 1) / rows;
        for (int u = 0; u < I.ntiles; ++u) {
            tinst.push_back(i);
            tr0.push_back(u * rows);
        }
    }
    void *old[] = {L.d_tile_inst, L.d_tile_r0, L.d_atiles, L.d_partials};
    for (void *p : old)
        if (p) (void)hipFree(p);
    L.d_tile_inst = L.d_tile_r0 = L.d_atiles = nullptr;
    L.d_partials = nullptr;
    L.T = (int)tinst.size();
    FSP_HIP(copy_sync(L.d_inst, L.h_inst.data(), sizeof(LsInst) * (size_t)B.count, hipMemcpyHostToDevice,
                      nullptr));
    FSP_HIP(hipMalloc(&L.d_tile_inst, sizeof(int) * (size_t)L.T));
    FSP_HIP(hipMalloc(&L.d_tile_r0, sizeof(int) * (size_t)L.T));
    FSP_HIP(hipMalloc(&L.d_atiles, sizeof(int) * (size_t)L.T));
    FSP_HIP(copy_sync(L.d_tile_inst, tinst.data(), sizeof(int) * (size_t)L.T, hipMemcpyHostToDevice, nullptr));
    FSP_HIP(copy_sync(L.d_tile_r0, tr0.data(), sizeof(int) * (size_t)L.T, hipMemcpyHostToDevice, nullptr));
    // dot partials: 2m + 4 slots of T (dots, sink, Gram column); the
    // update's part-tile norms: rpt T
    const size_t pslots = (size_t)std::max(2 * B.m + 4, rpt) * (size_t)L.T;
    FSP_HIP(hipMalloc(&L.d_partials, sizeof(double) * 2 * pslots));
    if (L.gexec) (void)hipGraphExecDestroy(L.gexec);
    if (L.graph) (void)hipGraphDestroy(L.graph);
    L.gexec = nullptr;
    L.graph = nullptr;
    L.graph_grid = 0;
    L.rpt = rpt;
    return FSP_OK;
}

static fsp_status ls_build(BatchImpl &B)
{
    if (B.ls) return FSP_OK;
    B.ls = new LsHost();
    LsHost &L = *B.ls;
    const int count = B.count, m = B.m;
    L.h_inst.assign((size_t)count, LsInst{});
    std::vector<LsInst> &hi = L.h_inst;
    size_t basis_total = 0, basis32_total = 0, work_total = 0, diag_total = 0, slot_total = 0;
    std::vector<int> offs;  // slot column offsets of every instance, -delta_u
    for (int i = 0; i < count; ++i) {
        const BatchInst &bi = B.h_inst[i];
        LsInst &I = hi[i];
        std::memset(&I, 0, sizeof(I));
        I.n = bi.n;
        I.nv = bi.nv;
        I.with_sink = bi.with_sink;
        I.tab_size = bi.tab_size;
        I.t_final = bi.t_final;
        basis_total += (size_t)bi.nv * (B.mixed ? 1 : (size_t)(m + 1));
        basis32_total += B.mixed ? (size_t)bi.nv * (size_t)m : 0;
        work_total += 2 * (size_t)bi.nv;
        I.U = B.hms[i].dm.U;
        diag_total += (size_t)bi.n;
        slot_total += (size_t)I.U * (size_t)bi.n;
        for (int u = 0; u < I.U; ++u) {
            const int64_t d = -B.hms[i].dm.slot_delta[u];
            offs.push_back((int)std::max<int64_t>(-(int64_t)bi.n, std::min<int64_t>((int64_t)bi.n, d)));
        }
    }
    FSP_HIP(hipMalloc(&L.d_basis, sizeof(double) * basis_total));
    if (B.mixed) FSP_HIP(hipMalloc(&L.d_basis32, sizeof(float) * basis32_total));
    FSP_HIP(hipMalloc(&L.d_work, sizeof(double) * work_total));
    const size_t vbytes = B.mixed ? sizeof(float) : sizeof(double);
    FSP_HIP(hipMalloc(&L.d_hdiag, sizeof(double) * diag_total));
    FSP_HIP(hipMalloc(&L.d_hsink, sizeof(double) * diag_total));
    FSP_HIP(hipMalloc(&L.d_hvals, vbytes * std::max<size_t>(slot_total, 1)));
    FSP_HIP(hipMalloc(&L.d_hoff, sizeof(int) * std::max<size_t>(offs.size(), 1)));
    if (!offs.empty())
        FSP_HIP(copy_sync(L.d_hoff, offs.data(), sizeof(int) * offs.size(), hipMemcpyHostToDevice, nullptr));
    size_t ob = 0, ob32 = 0, ow = 0, od = 0, os = 0, oo = 0;
    for (int i = 0; i < count; ++i) {
        const BatchInst &bi = B.h_inst[i];
        LsInst &I = hi[i];
        I.dm = bi.dm;
        I.tab = bi.tab;
        I.coefs = bi.coefs;
        I.hdiag = L.d_hdiag + od;
        I.hsink = L.d_hsink + od;
        I.hvals = (char *)L.d_hvals + vbytes * os;
        I.hoff = L.d_hoff + oo;
        I.p = bi.p;
        I.p0 = bi.p0;
        I.V64 = L.d_basis + ob;
        I.V32 = B.mixed ? L.d_basis32 + ob32 : nullptr;
        I.W0 = L.d_work + ow;
        I.W1 = I.W0 + bi.nv;
        ob += (size_t)bi.nv * (B.mixed ? 1 : (size_t)(m + 1));
        ob32 += B.mixed ? (size_t)bi.nv * (size_t)m : 0;
        ow += 2 * (size_t)bi.nv;
        od += (size_t)bi.n;
        os += (size_t)I.U * (size_t)bi.n;
        oo += (size_t)I.U;
    }
    FSP_HIP(hipMalloc(&L.d_inst, sizeof(LsInst) * (size_t)count));
    FSP_TRY(ls_tiles(B, L));
    FSP_HIP(hipMalloc(&L.d_cnt, sizeof(int) * 4));
    FSP_HIP(fill_sync(L.d_cnt, 0, sizeof(int) * 4, nullptr));
    FSP_HIP(hipMalloc(&L.d_owner, sizeof(int) * (size_t)count));
    FSP_HIP(fill_sync(L.d_owner, 0, sizeof(int) * (size_t)count, nullptr));
    L.h_owner.assign((size_t)count, 0);
    FSP_HIP(hipMalloc(&L.d_pq_list, sizeof(int) * (size_t)count));
    FSP_HIP(hipMalloc(&L.d_pq_queue, sizeof(unsigned int)));
    {
        int dev = 0, khz = 0;
        FSP_HIP(hipGetDevice(&dev));
        FSP_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        L.ticks_per_s = (long long)std::max(khz, 1) * 1000;
    }
    FSP_HIP(hipHostMalloc((void **)&L.h_cnt, sizeof(int) * 4, hipHostMallocDefault));
    // step states: constants once (tolerance, decimal powers), the rest per solve
    std::vector<double> p10(kP10Hi - kP10Lo + 1);
    FSP_HIP(copy_sync(p10.data(), B.d_p10, sizeof(double) * p10.size(), hipMemcpyDeviceToHost, nullptr));
    std::vector<KState> hs((size_t)count);
    const double fact = std::pow((m + 1) / std::exp(1.0), m + 1) * std::sqrt(2 * (3.1416) * (m + 1));
    for (int i = 0; i < count; ++i) {
        KState &k = hs[i];
        std::memset(&k, 0, sizeof(k));
        k.fact = fact;
        k.tol = B.tol;
        k.anorm = B.anorm;
        k.t_final = B.h_inst[i].t_final;
        k.k1 = 2;
        std::memcpy(k.p10, p10.data(), sizeof(double) * p10.size());
    }
    FSP_HIP(hipMalloc(&L.d_st, sizeof(KState) * (size_t)count));
    FSP_HIP(copy_sync(L.d_st, hs.data(), sizeof(KState) * (size_t)count, hipMemcpyHostToDevice, nullptr));
    FSP_HIP(hipStreamCreateWithFlags(&L.s, hipStreamNonBlocking));
    FSP_HIP(hipStreamCreateWithFlags(&L.pq_s, hipStreamNonBlocking));
    FSP_HIP(hipEventCreateWithFlags(&L.ev_reset, hipEventDisableTiming));
    L.grid = 0;
    // HYB generators
    const LsArgs a = ls_args(B, L);
    const int fgrid = std::max(1, std::min(L.T, 65535));
    FSP_LS_RPT_SWITCH(L.rpt, {
        if (B.dmax <= 4) {
            if (B.mixed)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ls_fill_hyb<4, float, RPT_>), dim3(fgrid), dim3(kBlock), 0, L.s, a, 1);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ls_fill_hyb<4, double, RPT_>), dim3(fgrid), dim3(kBlock), 0, L.s, a, 0);
        } else {
            if (B.mixed)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ls_fill_hyb<8, float, RPT_>), dim3(fgrid), dim3(kBlock), 0, L.s, a, 1);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ls_fill_hyb<8, double, RPT_>), dim3(fgrid), dim3(kBlock), 0, L.s, a, 0);
        }
    });
    FSP_HIP(hipGetLastError());
    FSP_HIP(hipStreamSynchronize(L.s));
    return FSP_OK;
}

template <typename Bt, int RPT>
static fsp_status ls_enqueue_step_t(BatchImpl &B, LsHost &L, int grid, hipStream_t s)
{
    const int m = B.m, M2 = m + 2;
    LsArgs a = ls_args(B, L);
    hipLaunchKernelGGL(k_ls_compact, dim3(1), dim3(1024), 0, s, a);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ls_norm_partials<RPT>), dim3(grid), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_ls_step_begin, dim3(B.count), dim3(kBlock), 0, s, a, (long long)B.max_nstep);
    FSP_HIP(hipGetLastError());
    int istart = 0;
    for (int j = 0; j <= m; ++j) {
        const bool last = (j == m);
        if (!last && B.iop) istart = (j >= B.q_iop - 1) ? j - B.q_iop + 1 : 0;
        a.j = j;
        a.istart = last ? 0 : istart;
        a.mode = last ? 1 : 0;
        a.nk = last ? 1 : j - istart + 1;
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ls_arn<Bt, RPT>), dim3(grid), dim3(kBlock), 0, s, a);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ls_dots<Bt>), dim3(B.count * a.nk), dim3(kBlock), 0, s, a);
        FSP_HIP(hipGetLastError());
        if (last) break;
        if (B.mixed)
            hipLaunchKernelGGL(k_ls_gram_coef, dim3(B.count), dim3(kBlock), kGramLdsBytes, s, a);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ls_update<Bt, RPT>), dim3(grid * RPT), dim3(kBlock), 0, s, a);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ls_norm<RPT>), dim3(B.count), dim3(kBlock), 0, s, a);
        FSP_HIP(hipGetLastError());
    }
    const size_t lds = sizeof(double) * 5 * (size_t)M2 * M2;
    hipLaunchKernelGGL(k_ls_control, dim3(B.count), dim3(kBlock), lds, s, a, B.max_reject);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ls_solution<Bt, RPT>), dim3(grid), dim3(kBlock), 0, s, a);
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}

static fsp_status ls_enqueue_step(BatchImpl &B, LsHost &L, int grid, hipStream_t s)
{
    fsp_status st = FSP_OK;
    FSP_LS_RPT_SWITCH(L.rpt, (st = B.mixed ? ls_enqueue_step_t<float, RPT_>(B, L, grid, s)
                                           : ls_enqueue_step_t<double, RPT_>(B, L, grid, s)));
    return st;
}

// Workgroups of the row kernels: one per tile (blocks_per_cu = 0; those
// past the running tiles return at once), or blocks_per_cu per CU walking
// the running tiles grid-stride.
static int ls_grid(const BatchImpl &B, const LsHost &L, int blocks_per_cu)
{
    if (blocks_per_cu <= 0) return std::max(1, L.T);
    return std::max(1, std::min(L.T, B.cus * blocks_per_cu));
}

// One batch step as a HIP graph (the kernels read all per-step values from
// device state, so one capture serves every step).
static fsp_status ls_graph(BatchImpl &B, LsHost &L, int grid)
{
    if (L.gexec && L.graph_grid == grid) return FSP_OK;
    if (L.gexec) (void)hipGraphExecDestroy(L.gexec);
    if (L.graph) (void)hipGraphDestroy(L.graph);
    L.gexec = nullptr;
    L.graph = nullptr;
    hipStream_t cap = nullptr;
    FSP_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    FSP_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    fsp_status st = ls_enqueue_step(B, L, grid, cap);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(cap, &g);
    (void)hipStreamDestroy(cap);
    FSP_TRY(st);
    if (ce != hipSuccess) return fail(FSP_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
    L.graph = g;
    FSP_HIP(hipGraphInstantiate(&L.gexec, g, nullptr, nullptr, 0));
    L.graph_grid = grid;
    return FSP_OK;
}

// FSP_BATCH_GRAPH=0 launches each batch step's kernels directly instead of
// replaying the captured graph (for tools that do not follow graphs).
static bool ls_use_graph()
{
    const char *g = std::getenv("FSP_BATCH_GRAPH");
    return !(g && std::atoi(g) == 0);
}

fsp_status batch_lockstep_prepare(BatchImpl &B, int blocks_per_cu)
{
    FSP_TRY(ls_build(B));
    if (B.ls->rpt != ((B.ls_rpt == 2 || B.ls_rpt == 8) ? B.ls_rpt : 4)) FSP_TRY(ls_tiles(B, *B.ls));
    if (!ls_use_graph()) return FSP_OK;
    return ls_graph(B, *B.ls, ls_grid(B, *B.ls, blocks_per_cu));
}

// Instances of at most `split` states (split < 0: every instance) are
// owned by the persistent kernel, listed in decreasing estimated cost; the
// lockstep kernels skip them.
static fsp_status pq_assign(BatchImpl &B, LsHost &L, int64_t split)
{
    std::vector<int> own((size_t)B.count, 0), list;
    for (int i : B.order) {
        if (split < 0 || (int64_t)B.h_inst[i].nv <= split) {
            own[i] = 1;
            list.push_back(i);
        }
    }
    if (own != L.h_owner) {
        FSP_HIP(copy_sync(L.d_owner, own.data(), sizeof(int) * own.size(), hipMemcpyHostToDevice, nullptr));
        L.h_owner = own;
    }
    if (list != L.h_list) {
        if (!list.empty())
            FSP_HIP(copy_sync(L.d_pq_list, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice,
                              nullptr));
        L.h_list = list;
        L.pq_graph_grid = L.full_graph_grid = 0;  // graphs hold the count
    }
    L.pq_count = (int)list.size();
    return FSP_OK;
}

static size_t pq_lds_bytes(const BatchImpl &B)
{
    const size_t M2 = (size_t)B.m + 2;
    const size_t ctl = std::max(5 * M2 * M2, kGramLdsBytes / sizeof(double));
    return sizeof(double) * (kLsRedLen + ctl);
}

static int pq_grid(const BatchImpl &B, const LsHost &L, int wg_per_cu)
{
    return std::max(1, std::min(L.pq_count, B.cus * std::max(1, wg_per_cu)));
}

template <typename Bt>
static fsp_status pq_enqueue_t(BatchImpl &B, LsHost &L, int grid, hipStream_t s)
{
    const LsArgs a = ls_args(B, L);
    PqArgs q;
    q.list = L.d_pq_list;
    q.count = L.pq_count;
    q.queue = L.d_pq_queue;
    q.max_reject = B.max_reject;
    q.iop = B.iop;
    q.q_iop = B.q_iop;
    q.max_nstep = B.max_nstep;
    q.budget = (long long)(B.pq_budget_s * (double)L.ticks_per_s);
    hipLaunchKernelGGL(k_pq_reset, dim3(1), dim3(1), 0, s, L.d_pq_queue);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pq<Bt>), dim3(grid), dim3(kBlock), pq_lds_bytes(B), s, a, q);
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}

static fsp_status pq_enqueue(BatchImpl &B, LsHost &L, int grid, hipStream_t s)
{
    return B.mixed ? pq_enqueue_t<float>(B, L, grid, s) : pq_enqueue_t<double>(B, L, grid, s);
}

// Captures `body` on a private stream into (*g, *ge).
template <class F>
static fsp_status capture(hipGraph_t *g, hipGraphExec_t *ge, F body)
{
    if (*ge) (void)hipGraphExecDestroy(*ge);
    if (*g) (void)hipGraphDestroy(*g);
    *ge = nullptr;
    *g = nullptr;
    hipStream_t cap = nullptr;
    FSP_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    FSP_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    fsp_status st = body(cap);
    hipGraph_t gg = nullptr;
    hipError_t ce = hipStreamEndCapture(cap, &gg);
    (void)hipStreamDestroy(cap);
    FSP_TRY(st);
    if (ce != hipSuccess) return fail(FSP_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
    *g = gg;
    FSP_HIP(hipGraphInstantiate(ge, gg, nullptr, nullptr, 0));
    return FSP_OK;
}

static fsp_status instance_failures(BatchImpl &B)
{
    std::vector<BatchInst> h((size_t)B.count);
    FSP_HIP(copy_sync(h.data(), B.d_inst, sizeof(BatchInst) * (size_t)B.count, hipMemcpyDeviceToHost, nullptr));
    for (int i = 0; i < B.count; ++i) {
        if (h[i].status != 0) {
            char buf[200];
            const char *why = h[i].status == FSP_ERR_TIMEOUT ? "persistent-kernel wall-clock limit reached"
                              : h[i].steps >= B.max_nstep   ? "step count reached max_nstep"
                                                            : "Krylov reject loop exhausted max_reject";
            std::snprintf(buf, sizeof buf, "batch instance %d (%d states, t = %f): %s", i, h[i].nv,
                          h[i].t_final, why);
            return fail((fsp_status)h[i].status, buf);
        }
    }
    return FSP_OK;
}

// Allocations, tile tables and captured graphs of a hybrid solve with this
// split (the lockstep step graph when any instance is lockstep, the
// persistent kernel's graph when any is persistent).
static fsp_status hybrid_prepare(BatchImpl &B, int blocks_per_cu, int wg_per_cu, int64_t split)
{
    FSP_TRY(ls_build(B));
    LsHost &L = *B.ls;
    FSP_TRY(pq_assign(B, L, split));
    if (B.count - L.pq_count > 0) FSP_TRY(batch_lockstep_prepare(B, blocks_per_cu));
    const int pgrid = pq_grid(B, L, wg_per_cu);
    if (L.pq_count > 0 && ls_use_graph() && (!L.pq_gexec || L.pq_graph_grid != pgrid)) {
        FSP_TRY(capture(&L.pq_graph, &L.pq_gexec, [&](hipStream_t c) { return pq_enqueue(B, L, pgrid, c); }));
        L.pq_graph_grid = pgrid;
    }
    return FSP_OK;
}

fsp_status batch_solve_hybrid(BatchImpl &B, int blocks_per_cu, int wg_per_cu, int64_t split)
{
    FSP_TRY(hybrid_prepare(B, blocks_per_cu, wg_per_cu, split));
    LsHost &L = *B.ls;
    const int n_pq = L.pq_count, n_ls = B.count - L.pq_count;
    const int pgrid = pq_grid(B, L, wg_per_cu);
    LsArgs a = ls_args(B, L);
    hipLaunchKernelGGL(k_ls_reset, dim3(B.count), dim3(kBlock), 0, L.s, a);
    FSP_HIP(hipGetLastError());
    if (n_pq > 0) {
        // the persistent kernel on its own stream, beside the lockstep steps
        FSP_HIP(hipEventRecord(L.ev_reset, L.s));
        FSP_HIP(hipStreamWaitEvent(L.pq_s, L.ev_reset, 0));
        if (ls_use_graph())
            FSP_HIP(hipGraphLaunch(L.pq_gexec, L.pq_s));
        else
            FSP_TRY(pq_enqueue(B, L, pgrid, L.pq_s));
    }
    // the lockstep steps; whatever happens to them, the persistent kernel is
    // waited for before this call returns
    auto lockstep_loop = [&]() -> fsp_status {
        // steps are launched in groups ahead of the host check; a step with
        // no running instance does nothing
        const int ahead = 4;
        for (long long guard = 0;; guard += ahead) {
            for (int k = 0; k < ahead; ++k) {
                if (L.gexec && ls_use_graph())
                    FSP_HIP(hipGraphLaunch(L.gexec, L.s));
                else
                    FSP_TRY(ls_enqueue_step(B, L, ls_grid(B, L, blocks_per_cu), L.s));
            }
            hipLaunchKernelGGL(k_ls_running, dim3(1), dim3(kBlock), 0, L.s, a, L.d_cnt + 2);
            FSP_HIP(hipGetLastError());
            FSP_HIP(hipMemcpyAsync(L.h_cnt, L.d_cnt, sizeof(int) * 4, hipMemcpyDeviceToHost, L.s));
            FSP_HIP(hipStreamSynchronize(L.s));
            if (L.h_cnt[2] == 0) return FSP_OK;
            if (guard > B.max_nstep + 8) return fail(FSP_ERR_NOT_CONVERGED, "batch steps exceeded max_nstep");
        }
    };
    const fsp_status ls_st = n_ls > 0 ? lockstep_loop() : FSP_OK;
    if (n_pq > 0) FSP_HIP(hipStreamSynchronize(L.pq_s));
    FSP_TRY(ls_st);
    hipLaunchKernelGGL(k_ls_export, dim3((B.count + kBlock - 1) / kBlock), dim3(kBlock), 0, L.s, a, B.d_inst);
    FSP_HIP(hipGetLastError());
    FSP_HIP(hipStreamSynchronize(L.s));
    return instance_failures(B);
}

fsp_status batch_solve_lockstep(BatchImpl &B, int blocks_per_cu)
{
    return batch_solve_hybrid(B, blocks_per_cu, 1, 0);
}

// The all-persistent launch as one graph: instance reset, queue reset, the
// kernel, the stats export.
static fsp_status full_body(BatchImpl &B, LsHost &L, int grid, hipStream_t c)
{
    const LsArgs a = ls_args(B, L);
    hipLaunchKernelGGL(k_ls_reset, dim3(B.count), dim3(kBlock), 0, c, a);
    FSP_TRY(pq_enqueue(B, L, grid, c));
    hipLaunchKernelGGL(k_ls_export, dim3((B.count + kBlock - 1) / kBlock), dim3(kBlock), 0, c, a, B.d_inst);
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}

static fsp_status persistent_prepare(BatchImpl &B, int wg_per_cu)
{
    FSP_TRY(ls_build(B));
    LsHost &L = *B.ls;
    FSP_TRY(pq_assign(B, L, -1));
    const int grid = pq_grid(B, L, wg_per_cu);
    if (ls_use_graph() && (!L.full_gexec || L.full_graph_grid != grid)) {
        FSP_TRY(capture(&L.full_graph, &L.full_gexec, [&](hipStream_t c) { return full_body(B, L, grid, c); }));
        L.full_graph_grid = grid;
    }
    return FSP_OK;
}

fsp_status batch_launch_persistent(BatchImpl &B, int wg_per_cu, hipStream_t stream)
{
    FSP_TRY(persistent_prepare(B, wg_per_cu));
    LsHost &L = *B.ls;
    if (!ls_use_graph()) return full_body(B, L, pq_grid(B, L, wg_per_cu), stream);
    FSP_HIP(hipGraphLaunch(L.full_gexec, stream));
    return FSP_OK;
}

fsp_status batch_prepare(BatchImpl &B, int path)
{
    if (path == FSP_BATCH_LOCKSTEP) return hybrid_prepare(B, B.ls_bpc, B.pq_wg, 0);
    if (path == FSP_BATCH_HYBRID) return hybrid_prepare(B, B.ls_bpc, B.pq_wg, B.pq_split);
    if (path == FSP_BATCH_PERSISTENT) return persistent_prepare(B, B.pq_wg);
    return FSP_OK;
}

}  // namespace fsp
