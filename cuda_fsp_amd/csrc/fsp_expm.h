// fsp_expm.h -- one-workgroup dense Pade exponential and the Expokit
// step-size rounding, shared by the single-solve and batched Krylov kernels.
#pragma once

#include "fsp_device.h"

namespace fsp {

constexpr int kP10Lo = -300, kP10Hi = 300;
constexpr int kElimChunk = 4;  // register chunk of the elimination's row updates
constexpr int kSubChunk = 8;   // loads issued together in back substitution

// C = A * B for column-major N x N matrices in LDS.  Each thread computes a
// BS x BS block of C from BS + BS LDS reads per BS^2 products; every element
// is still sum_l A(i, l) B(l, j) accumulated in increasing l, the oracle's
// order, so the block size changes speed only.
template <int BS>
__device__ inline void matmul_lds_bs(int N, const double *A, const double *Bm, double *C)
{
    const int nb = (N + BS - 1) / BS;
    for (int b = threadIdx.x; b < nb * nb; b += blockDim.x) {
        const int i0 = (b % nb) * BS, j0 = (b / nb) * BS;
        double acc[BS][BS];
#pragma unroll
        for (int a = 0; a < BS; ++a)
#pragma unroll
            for (int c = 0; c < BS; ++c) acc[a][c] = 0.0;
        for (int l = 0; l < N; ++l) {
            double av[BS], bv[BS];
#pragma unroll
            for (int a = 0; a < BS; ++a) av[a] = (i0 + a < N) ? A[(i0 + a) + l * N] : 0.0;
#pragma unroll
            for (int c = 0; c < BS; ++c) bv[c] = (j0 + c < N) ? Bm[l + (j0 + c) * N] : 0.0;
#pragma unroll
            for (int a = 0; a < BS; ++a)
#pragma unroll
                for (int c = 0; c < BS; ++c) acc[a][c] += av[a] * bv[c];
        }
#pragma unroll
        for (int a = 0; a < BS; ++a)
#pragma unroll
            for (int c = 0; c < BS; ++c)
                if (i0 + a < N && j0 + c < N) C[(i0 + a) + (j0 + c) * N] = acc[a][c];
    }
}

// 2 x 2 blocks when the workgroup has a thread for most of them (more
// waves in flight), 4 x 4 otherwise (one wave).
__device__ inline void matmul_lds(int N, const double *A, const double *Bm, double *C)
{
    const int nb2 = (N + 1) / 2;
    if ((int)blockDim.x * 2 >= nb2 * nb2)
        matmul_lds_bs<2>(N, A, Bm, C);
    else
        matmul_lds_bs<4>(N, A, Bm, C);
}

// --------------------------------------------------------------- dense expm
// Golub & Van Loan, Matrix Computations, Alg. 11.3.1 (diagonal Pade q = 6,
// scaling by 2^-j with j = max(0, 1 + floor(log2 |A|_inf)), Gaussian
// elimination with partial pivoting, j squarings), column-major N x N
// (N <= 64) in LDS, one workgroup of any size.  Same
// operation order as the oracle's orc_expm, so the two agree bit for bit on
// equal input.  Result in Nm.
__device__ inline void expm_block(int N, double *A, double *X, double *T, double *Nm, double *Dm,
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
        if (tid == 0) {
            // first row of maximal |Dm(i, k)|, i >= k (the oracle's choice);
            // one thread: N - k independent LDS loads, no shuffle round trips
            int p = k;
            double best = fabs(Dm[k + k * N]);
            for (int i = k + 1; i < N; ++i) {
                const double v = fabs(Dm[i + k * N]);
                if (v > best) {
                    best = v;
                    p = i;
                }
            }
            piv_sh = p;
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
        // Row updates below the pivot by (row, column chunk) pairs: each pair
        // forms its row's multiplier (the same division in every pair, so the
        // same value), loads its chunk of the unchanged pivot row and of its
        // own row, then stores the updated chunk.  Consecutive pairs own
        // consecutive rows (conflict-free LDS).  The multipliers are not kept:
        // only the upper triangle is read afterwards.
        const double pv = Dm[k + k * N];
        const int rows = N - k - 1;
        const int chunks = (N + kElimChunk - 1) / kElimChunk;
        for (int e = tid; e < rows * chunks; e += blockDim.x) {
            const int i = k + 1 + e % rows;
            const int j0 = (e / rows) * kElimChunk;
            const double l = Dm[i + k * N] / pv;
            double pd[kElimChunk], pn[kElimChunk], rd[kElimChunk], rn[kElimChunk];
#pragma unroll
            for (int q = 0; q < kElimChunk; ++q) {
                const int j = min(j0 + q, N - 1);
                pd[q] = Dm[k + j * N];
                pn[q] = Nm[k + j * N];
                rd[q] = Dm[i + j * N];
                rn[q] = Nm[i + j * N];
            }
#pragma unroll
            for (int q = 0; q < kElimChunk; ++q) {
                const int j = j0 + q;
                if (j < N) {
                    if (j > k) Dm[i + j * N] = rd[q] - l * pd[q];
                    Nm[i + j * N] = rn[q] - l * pn[q];
                }
            }
        }
        __syncthreads();
    }
    // back substitution, one right-hand side per thread, on a row-major copy
    // of Nm in X (consecutive lanes read consecutive LDS words)
    for (int e = tid; e < NN; e += blockDim.x) X[(e / N) + (e % N) * N] = Nm[e];
    __syncthreads();
    for (int j = tid; j < N; j += blockDim.x) {
        for (int i = N - 1; i >= 0; --i) {
            // k ascending as one sum; kSubChunk loads issued together
            double acc = X[j + i * N];
            int k = i + 1;
            for (; k + kSubChunk <= N; k += kSubChunk) {
                double d[kSubChunk], x[kSubChunk];
#pragma unroll
                for (int u = 0; u < kSubChunk; ++u) {
                    d[u] = Dm[i + (k + u) * N];
                    x[u] = X[j + (k + u) * N];
                }
#pragma unroll
                for (int u = 0; u < kSubChunk; ++u) acc -= d[u] * x[u];
            }
            for (; k < N; ++k) acc -= Dm[i + k * N] * X[j + k * N];
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
