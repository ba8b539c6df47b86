// fsp_device.h -- device helpers shared by the FSP kernels (gfx950, wave64).
//
// All HIP sources are compiled with -ffp-contract=off: every multiply and add
// is a separately rounded IEEE operation unless fma() is written, so
// assembled generator values and CSR row sums match the CPU oracle bit for
// bit.
#pragma once

#include "fsp_internal.h"

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

// ------------------------------------------------------------ state decode
template <int D>
__device__ inline void decode(uint32_t i, const DevModel *__restrict__ dm, int (&x)[D])
{
    uint32_t q = i;
#pragma unroll
    for (int s = 0; s < D; ++s) {
        if (s == D - 1) {
            x[s] = (int)q;
        } else {
            uint32_t q2 = fast_div(q, dm->fd[s]);
            x[s] = (int)(q - q2 * (uint32_t)dm->radix[s]);
            q = q2;
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

inline int grid_for(int64_t tiles, int cap)
{
    int64_t g = tiles < cap ? tiles : cap;
    return (int)(g > 0 ? g : 1);
}

}  // namespace fsp
