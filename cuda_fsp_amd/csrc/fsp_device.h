// fsp_device.h -- device helpers shared by the FSP kernels (gfx950, wave64).
//
// All HIP sources are compiled with -ffp-contract=off: every multiply and add
// is a separately rounded IEEE operation unless fma() is written, so
// assembled generator values and CSR row sums match the CPU oracle bit for
// bit.
#pragma once

#include "fsp_internal.h"

// The dynamic LDS of a kernel as doubles at a 16-byte boundary (the static
// LDS before it need not end on one), for 128-bit LDS accesses.
#ifndef FSP_DYNAMIC_LDS
#define FSP_DYNAMIC_LDS(var) extern __shared__ __attribute__((aligned(16))) double var[]
#endif

namespace fsp {

constexpr int kWave = 64;

// ---------------------------------------------------------------- reductions
// Deterministic: fixed butterfly inside the wave, fixed order across waves.
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Orders the LDS accesses of one wavefront's lanes: a wave executes its LDS
// instructions in order, so after this fence (outstanding LDS operations
// complete, no memory access moves across it) every lane sees what the other
// lanes of its wave stored before it.  Called by every lane of the wave.
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Block sum for blockDim.x == kBlock; result valid in every thread.
// `red` is LDS scratch of at least kBlock / kWave doubles.
__device__ inline double block_sum(double v, double *red)
{
    v = wave_sum(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kBlock / kWave; ++k) s += red[k];
    return s;
}

// --------------------------------------------- order-independent reductions
// Cross-row sums (dot products, norms, the sink row) are accumulated as
// unevaluated sums s + c with error-free transformations (Knuth's TwoSum,
// TwoProd through fma), as in Ogita, Rump & Oishi, "Accurate sum and dot
// product", SIAM J. Sci. Comput. 26 (2005), Dot2.  The result fl(s + c) is
// as accurate as a twice-working-precision sum, so it does not depend on the
// reduction tree except when the exact sum lies within ~eps^2 of a rounding
// boundary: the device and the CPU oracle then agree bit for bit.  Needs
// -ffp-contract=off (a contracted TwoSum is not error-free).
struct dd {
    double s, c;
};

__host__ __device__ inline dd dd_zero() { return dd{0.0, 0.0}; }

__host__ __device__ inline void two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

__device__ inline dd dd_add(dd x, double v)
{
    double s, e;
    two_sum(x.s, v, s, e);
    return dd{s, x.c + e};
}

__device__ inline dd dd_add_prod(dd x, double a, double b)
{
    const double p = a * b;
    const double pe = fma(a, b, -p);
    double s, e;
    two_sum(x.s, p, s, e);
    return dd{s, x.c + (e + pe)};
}

__device__ inline dd dd_merge(dd x, dd y)
{
    double s, e;
    two_sum(x.s, y.s, s, e);
    return dd{s, (x.c + y.c) + e};
}

__device__ inline double dd_value(dd x) { return x.s + x.c; }

__device__ inline dd wave_sum_dd(dd v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        dd o;
        o.s = __shfl_xor(v.s, off, kWave);
        o.c = __shfl_xor(v.c, off, kWave);
        v = dd_merge(v, o);
    }
    return v;
}

// Block dd sum for blockDim.x == kBlock; valid in every thread.  `red`:
// LDS scratch of at least 2 * kBlock / kWave doubles.
__device__ inline dd block_sum_dd(dd v, double *red)
{
    v = wave_sum_dd(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) {
        red[2 * w] = v.s;
        red[2 * w + 1] = v.c;
    }
    __syncthreads();
    dd s = dd_zero();
#pragma unroll
    for (int k = 0; k < kBlock / kWave; ++k) s = dd_merge(s, dd{red[2 * k], red[2 * k + 1]});
    return s;
}

// Block sums of CH dd values per thread at once (blockDim.x == kBlock),
// transposed through LDS: the values go to buf (2 CH kBlock doubles) as
// [value][thread]; thread x then merges value q = x / S (S = kBlock / CH
// threads per value) over the threads x % S + S u, u = 0, 1, ... in that
// order, and a butterfly over its S lanes finishes the sum.  Returns true in
// the one thread per value (x % S == 0) that holds it, with q and the sum.
// A fixed order, and CH + log2 S dd merges per thread where a 64-lane
// butterfly per value takes 6 CH.  Starts with a barrier, so earlier reads
// of buf (or other uses of its bytes) are done; the caller reuses buf only
// after another barrier.
template <int CH>
__device__ inline bool block_sums_dd_t(const dd (&v)[CH], double *buf, int &q, dd &sum)
{
    constexpr int S = kBlock / CH;
    static_assert(kBlock % CH == 0 && S <= kWave && (S & (S - 1)) == 0,
                  "a value's threads must be a power-of-two part of one wave");
    __syncthreads();
#pragma unroll
    for (int k = 0; k < CH; ++k) {
        buf[2 * (k * kBlock + threadIdx.x)] = v[k].s;
        buf[2 * (k * kBlock + threadIdx.x) + 1] = v[k].c;
    }
    __syncthreads();
    q = threadIdx.x / S;
    const int sg = threadIdx.x % S;
    dd t = dd_zero();
#pragma unroll
    for (int u = 0; u < kBlock / S; ++u) {
        const int x = q * kBlock + sg + S * u;
        t = dd_merge(t, dd{buf[2 * x], buf[2 * x + 1]});
    }
#pragma unroll
    for (int off = S / 2; off > 0; off >>= 1) {
        dd o;
        o.s = __shfl_xor(t.s, off, kWave);
        o.c = __shfl_xor(t.c, off, kWave);
        t = dd_merge(t, o);
    }
    sum = t;
    return sg == 0;
}

// Wave sums of CH dd values per lane at once, transposed through the wave's
// own LDS (wbuf: 2 CH kWave doubles, used by this wave only): lane l merges
// value q = l / S (S = kWave / CH lanes per value) over the lanes l % S + S u,
// u = 0, 1, ... in that order, and a butterfly over its S lanes finishes the
// wave's sum.  Returns true in the one lane per value (l % S == 0) that holds
// it, with q and the sum.  A fixed order, CH + log2 S dd merges per lane, and
// only wave-level synchronisation (no block barrier).
template <int CH>
__device__ inline bool wave_sums_dd_t(const dd (&v)[CH], double *wbuf, int &q, dd &sum)
{
    constexpr int S = kWave / CH;
    static_assert(kWave % CH == 0 && (S & (S - 1)) == 0, "a value's lanes must be a power of two");
    const int l = threadIdx.x % kWave;
    wave_sync();  // the wave's earlier reads of wbuf are done
#pragma unroll
    for (int k = 0; k < CH; ++k) {
        wbuf[2 * (k * kWave + l)] = v[k].s;
        wbuf[2 * (k * kWave + l) + 1] = v[k].c;
    }
    wave_sync();
    q = l / S;
    const int sg = l % S;
    dd t = dd_zero();
#pragma unroll
    for (int u = 0; u < kWave / S; ++u) {
        const int x = q * kWave + sg + S * u;
        t = dd_merge(t, dd{wbuf[2 * x], wbuf[2 * x + 1]});
    }
#pragma unroll
    for (int off = S / 2; off > 0; off >>= 1) {
        dd o;
        o.s = __shfl_xor(t.s, off, kWave);
        o.c = __shfl_xor(t.c, off, kWave);
        t = dd_merge(t, o);
    }
    sum = t;
    return sg == 0;
}

// Sum of `count` dd partials stored as (s, c) pairs, fixed order.
__device__ inline dd sum_dd_partials(const double *__restrict__ p, int count, double *red)
{
    // each thread merges its partials b = tid, tid + blockDim, ... in that
    // order; four are loaded ahead of their merges
    constexpr int kAhead = 4;
    dd s = dd_zero();
    int b = threadIdx.x;
    const int step = blockDim.x;
    for (; b + (kAhead - 1) * step < count; b += kAhead * step) {
        double v[2 * kAhead];
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            v[2 * u] = p[2 * (b + u * step)];
            v[2 * u + 1] = p[2 * (b + u * step) + 1];
        }
#pragma unroll
        for (int u = 0; u < kAhead; ++u) s = dd_merge(s, dd{v[2 * u], v[2 * u + 1]});
    }
    for (; b < count; b += step) s = dd_merge(s, dd{p[2 * b], p[2 * b + 1]});
    return block_sum_dd(s, red);
}

// ------------------------------------------------------------ state decode
// Mixed-radix digits of row i (species 0 fastest) for a model with
// dm->d <= D species; digits past d are 0.
template <int D>
__device__ inline void decode(uint32_t i, const DevModel *__restrict__ dm, int (&x)[D])
{
    const int d = dm->d;
    uint32_t q = i;
#pragma unroll
    for (int s = 0; s < D; ++s) {
        if (s < d - 1) {
            const uint32_t q2 = fast_div(q, dm->fd[s]);
            x[s] = (int)(q - q2 * (uint32_t)dm->radix[s]);
            q = q2;
        } else if (s == d - 1) {
            x[s] = (int)q;
        } else {
            x[s] = 0;
        }
    }
}

template <int D>
__device__ inline int pick(const int (&x)[D], int s)
{
    int v = x[0];
#pragma unroll
    for (int t = 1; t < D; ++t) v = (t == s) ? x[t] : v;
    return v;
}

// a_r at x (shift = 0) or at the inflow source x - nu_r (shift = 1) from the
// compact factor tables `tab` (LDS or global).  Product in species order.
template <int D, typename TT>
__device__ inline double prop_tab(const RxnDev &rx, const int (&x)[D], const TT *tab, int shift)
{
    double p = 1.0;
    for (int k = 0; k < rx.nsp; ++k) {
        const int fo = rx.foff[k];
        if (fo >= 0) {
            const int c = pick(x, rx.sp[k]) - (shift ? rx.nu[k] : 0);
            p *= (double)tab[fo + c];
        }
    }
    return p;
}

// Inflow source x - nu_r inside the box?
template <int D>
__device__ inline bool source_in_box(const RxnDev &rx, const int (&x)[D],
                                     const DevModel *__restrict__ dm)
{
    bool ok = true;
    for (int k = 0; k < rx.nsp; ++k) {
        const int c = pick(x, rx.sp[k]) - rx.nu[k];
        ok = ok && (c >= 0) && (c < dm->radix[rx.sp[k]]);
    }
    return ok;
}

// Outflow target x + nu_r leaves the box?
template <int D>
__device__ inline bool target_leaves(const RxnDev &rx, const int (&x)[D],
                                     const DevModel *__restrict__ dm)
{
    bool out = false;
    for (int k = 0; k < rx.nsp; ++k) {
        const int c = pick(x, rx.sp[k]) + rx.nu[k];
        out = out || (c < 0) || (c >= dm->radix[rx.sp[k]]);
    }
    return out && !rx.zero_nu;
}

// Cooperative copy of the compact factor table into LDS.
template <typename TT>
__device__ inline void stage_table(TT *lds, const TT *__restrict__ g, int count)
{
    for (int k = threadIdx.x; k < count; k += blockDim.x) lds[k] = g[k];
    __syncthreads();
}

// XCD-aware tile walk.  MI355X dispatches workgroup b to XCD b % 8, and each
// XCD has its own L2: every XCD walks one contiguous range of row tiles (its
// workgroups interleaved within it), so the x gathers of neighbouring rows
// hit the same L2 instead of eight.  Falls back to the plain grid-stride
// order for grids smaller than the XCD count.
constexpr uint32_t kXcds = 8;

struct TileWalk {
    uint32_t t, end, step;
};

__device__ inline TileWalk tile_walk(uint32_t tiles)
{
    const uint32_t G = gridDim.x;
    if (G < kXcds || tiles <= G) return TileWalk{blockIdx.x, tiles, G};
    const uint32_t x = blockIdx.x % kXcds, l = blockIdx.x / kXcds;
    const uint32_t cnt = G / kXcds + (x < G % kXcds ? 1u : 0u);
    const uint32_t start = (uint32_t)(((uint64_t)x * tiles) / kXcds);
    const uint32_t end = (uint32_t)(((uint64_t)(x + 1) * tiles) / kXcds);
    return TileWalk{start + l, end, cnt};
}

// Workgroups are dispatched to the XCDs round-robin (block b on XCD
// b % kXcds).  Renumbers them so that each XCD's blocks cover one contiguous
// range of [0, G), so neighbouring tiles (and the vector entries they
// gather) share that XCD's L2.  A bijection on [0, G).
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t G)
{
    const uint32_t q = G / kXcds, r = G % kXcds, x = b % kXcds, s = b / kXcds;
    return x < r ? x * (q + 1) + s : r * (q + 1) + (x - r) * q + s;
}

inline int grid_for(int64_t tiles, int cap)
{
    int64_t g = tiles < cap ? tiles : cap;
    return (int)(g > 0 ? g : 1);
}

}  // namespace fsp
