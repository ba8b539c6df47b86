// fsp_expm.h -- one-workgroup dense Pade exponential and the Expokit
// step-size rounding, shared by the single-solve and batched Krylov kernels.
#pragma once

#include "fsp_device.h"

namespace fsp {

constexpr int kP10Lo = -300, kP10Hi = 300;

// C = A * B for column-major N x N matrices in LDS.  Each thread computes a
// 4 x 4 block of C (4 + 4 LDS reads per 16 products); every element is still
// sum_l A(i, l) B(l, j) accumulated in increasing l, the oracle's order.
__device__ inline void matmul_lds(int N, const double *A, const double *Bm, double *C)
{
    const int nb = (N + 3) / 4;
    for (int b = threadIdx.x; b < nb * nb; b += blockDim.x) {
        const int i0 = (b % nb) * 4, j0 = (b / nb) * 4;
        double acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[a][c] = 0.0;
        for (int l = 0; l < N; ++l) {
            double av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = (i0 + a < N) ? A[(i0 + a) + l * N] : 0.0;
#pragma unroll
            for (int c = 0; c < 4; ++c) bv[c] = (j0 + c < N) ? Bm[l + (j0 + c) * N] : 0.0;
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[a][c] += av[a] * bv[c];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (i0 + a < N && j0 + c < N) C[(i0 + a) + (j0 + c) * N] = acc[a][c];
    }
}

// --------------------------------------------------------------- dense expm
// Golub & Van Loan, Matrix Computations, Alg. 11.3.1 (diagonal Pade q = 6,
// scaling by 2^-j with j = max(0, 1 + floor(log2 |A|_inf)), Gaussian
// elimination with partial pivoting, j squarings), column-major N x N
// (N <= 64) in LDS, one workgroup (one wave is enough and fastest).  Same
// operation order as the oracle's orc_expm, so the two agree bit for bit on
// equal input.  Result in Nm.
__device__ void expm_block(int N, double *A, double *X, double *T, double *Nm, double *Dm,
                           double *red)
{
    (void)red;
    const int NN = N * N;
    const int tid = threadIdx.x;
    __shared__ int js_sh, piv_sh;
    // infinity norm: row sums in column order, max taken in row order
    if (tid < kWave) {
        double rs = 0.0;
        if (tid < N)
            for (int j = 0; j < N; ++j) rs += fabs(A[tid + j * N]);
        // max over lanes (the maximum is exact, its order irrelevant)
        double mx = rs;
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, kWave));
        if (tid == 0) {
            int js = 0;
            if (mx > 0.0) {
                int e;
                frexp(mx, &e);
                js = e > 0 ? e : 0;
            }
            js_sh = js;
        }
    }
    __syncthreads();
    const int js = js_sh;
    const double scale = ldexp(1.0, -js);
    for (int k = tid; k < NN; k += blockDim.x) {
        A[k] = A[k] * scale;
        const double id = ((k % N) == (k / N)) ? 1.0 : 0.0;
        X[k] = id;
        Nm[k] = id;
        Dm[k] = id;
    }
    __syncthreads();
    const int q = 6;
    double c = 1.0;
    for (int kk = 1; kk <= q; ++kk) {
        c = c * (double)(q - kk + 1) / (double)((2 * q - kk + 1) * kk);
        matmul_lds(N, A, X, T);
        __syncthreads();
        const double sgn = (kk % 2) ? -1.0 : 1.0;
        for (int e = tid; e < NN; e += blockDim.x) {
            X[e] = T[e];
            Nm[e] += c * T[e];
            Dm[e] += sgn * c * T[e];
        }
        __syncthreads();
    }
    // solve Dm F = Nm by Gaussian elimination with partial pivoting
    for (int k = 0; k < N; ++k) {
        if (tid < kWave) {
            // first row of maximal |Dm(i, k)|, i >= k (the oracle's choice)
            double v = (tid >= k && tid < N) ? fabs(Dm[tid + k * N]) : -1.0;
            int idx = tid;
#pragma unroll
            for (int off = kWave / 2; off > 0; off >>= 1) {
                const double ov = __shfl_xor(v, off, kWave);
                const int oi = __shfl_xor(idx, off, kWave);
                if (ov > v || (ov == v && oi < idx)) {
                    v = ov;
                    idx = oi;
                }
            }
            if (tid == 0) piv_sh = idx;
        }
        __syncthreads();
        const int p = piv_sh;
        if (p != k) {
            for (int j = tid; j < N; j += blockDim.x) {
                double t = Dm[k + j * N];
                Dm[k + j * N] = Dm[p + j * N];
                Dm[p + j * N] = t;
                t = Nm[k + j * N];
                Nm[k + j * N] = Nm[p + j * N];
                Nm[p + j * N] = t;
            }
            __syncthreads();
        }
        // one thread per row below the pivot: multiplier, row update against
        // the (unchanged, broadcast) pivot row, multiplier stored in place
        const double pv = Dm[k + k * N];
        for (int i = k + 1 + tid; i < N; i += blockDim.x) {
            const double l = Dm[i + k * N] / pv;
            for (int j = k + 1; j < N; ++j) Dm[i + j * N] -= l * Dm[k + j * N];
            for (int j = 0; j < N; ++j) Nm[i + j * N] -= l * Nm[k + j * N];
            Dm[i + k * N] = l;
        }
        __syncthreads();
    }
    // back substitution, one right-hand side per thread, on a row-major copy
    // of Nm in X (consecutive lanes read consecutive LDS words)
    for (int e = tid; e < NN; e += blockDim.x) X[(e / N) + (e % N) * N] = Nm[e];
    __syncthreads();
    for (int j = tid; j < N; j += blockDim.x) {
        for (int i = N - 1; i >= 0; --i) {
            double acc = X[j + i * N];
            for (int k = i + 1; k < N; ++k) acc -= Dm[i + k * N] * X[j + k * N];
            X[j + i * N] = acc / Dm[i + i * N];
        }
    }
    __syncthreads();
    for (int e = tid; e < NN; e += blockDim.x) Nm[e] = X[(e / N) + (e % N) * N];
    __syncthreads();
    for (int sq = 0; sq < js; ++sq) {
        matmul_lds(N, Nm, Nm, T);
        __syncthreads();
        for (int e = tid; e < NN; e += blockDim.x) Nm[e] = T[e];
        __syncthreads();
    }
}

// src/KryExpvFSP.cu:148-150 and :172-174: s = 10^(floor(log10 t) - 1),
// t = ceil(t / s) * s.  The decade is found by comparison with correctly
// rounded decimal powers (built on the host with strtod) instead of
// log10/pow, so the device and the host oracle round identically.
__device__ inline double round_up_2sig(const double *p10, double t)
{
    if (!(t > 0.0) || isinf(t)) return t;
    int lo = kP10Lo, hi = kP10Hi, e = kP10Lo;
    while (lo <= hi) {
        const int mid = (lo + hi) / 2;
        if (p10[mid - kP10Lo] <= t) {
            e = mid;
            lo = mid + 1;
        } else {
            hi = mid - 1;
        }
    }
    if (e - 1 < kP10Lo) return t;
    const double s = p10[e - 1 - kP10Lo];
    return ceil(t / s) * s;
}

}  // namespace fsp
