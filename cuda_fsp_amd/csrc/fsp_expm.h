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
// order, so the block size changes speed only.  A has lower bandwidth lba
// (A(i, l) = 0 for i - l > lba) and B lower bandwidth lbb: the terms outside
// l in [i - lba, j + lbb] are exact zeros and are skipped (adding +-0 to the
// running sum leaves its value unchanged); pass N for dense operands.
template <int BS>
__device__ inline void matmul_lds_bs(int N, const double *A, const double *Bm, double *C,
                                     int lba, int lbb)
{
    const int nb = (N + BS - 1) / BS;
    for (int b = threadIdx.x; b < nb * nb; b += blockDim.x) {
        const int i0 = (b % nb) * BS, j0 = (b / nb) * BS;
        double acc[BS][BS];
#pragma unroll
        for (int a = 0; a < BS; ++a)
#pragma unroll
            for (int c = 0; c < BS; ++c) acc[a][c] = 0.0;
        const int l0 = max(0, i0 - lba);
        const int l1 = min(N - 1, j0 + BS - 1 + lbb);
        for (int l = l0; l <= l1; ++l) {
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
__device__ inline void matmul_lds(int N, const double *A, const double *Bm, double *C,
                                  int lba, int lbb)
{
    const int nb2 = (N + 1) / 2;
    if ((int)blockDim.x * 2 >= nb2 * nb2)
        matmul_lds_bs<2>(N, A, Bm, C, lba, lbb);
    else
        matmul_lds_bs<4>(N, A, Bm, C, lba, lbb);
}

constexpr int kExpmMaxN = 64;

// Back substitution U X = B for the 32 x 32 case, one right-hand side per
// thread with its column in registers (fully unrolled): x_i = (b_i -
// sum_{k > i} U(i, k) x_k) / U(i, i), k ascending, as the oracle.  U(i, k)
// is D(perm[i], k) and B(i, j) is Nm(perm[i], j); the solution overwrites
// column j of Nm in natural row order (each thread reads and writes only
// its own column).
__device__ inline void back_subst_32(const double *Dm, double *Nm, const int *perm)
{
    constexpr int N = 32;
    const int j = threadIdx.x;
    if (j >= N) return;
    double x[N];
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = Nm[perm[i] + j * N];
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        const int ri = perm[i];
        double acc = x[i];
#pragma unroll
        for (int k = i + 1; k < N; ++k) acc -= Dm[ri + k * N] * x[k];
        x[i] = acc / Dm[ri + i * N];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) Nm[i + j * N] = x[i];
}

// The 32 x 32 exponential of a scaled matrix with lower bandwidth <= 1 (the
// Krylov Hessenberg) by the first wavefront, every matrix held by columns in
// registers: lane j and lane 32 + j hold column j of the Pade powers, lane j
// accumulates column j of D and lane 32 + j column j of N, so that [D | N]
// is one column per lane for the elimination.  Values another lane needs
// (the pivot column, the triangular factor, F for a squaring) go through LDS
// with wavefront-scope fences, no workgroup barrier.  Every value follows the
// operation order of the oracle's orc_expm (products with l ascending from
// 0.0, the elimination row by row, back substitution with k ascending, the
// squarings), skipping only terms that are exact zeros: A(i, l) = 0 for
// l < i - 1, and the rows more than 6 below the pivot (D has lower
// bandwidth 6).  A: the scaled matrix (column-major, LDS); At, Ur, Fr: LDS
// scratch of N x N; the result goes to Nm (column-major).
__device__ __forceinline__ void expm_wave32(const double *A, double *At, double *Ur, double *Fr,
                                   double *Nm, int js)
{
    constexpr int N = 32, q = 6, BW = 6;
    __shared__ double colbuf[2][BW + 1];
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) At[(e % N) * N + e / N] = A[e];
    __syncthreads();
    if (threadIdx.x < kWave) {
        const int lane = threadIdx.x;
        const int j = lane & (N - 1);
        const bool nside = lane >= N;
        double x[N], acc[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            x[i] = (i == j) ? 1.0 : 0.0;
            acc[i] = x[i];
        }
        // Pade polynomials: X = A^kk, N += c X, D += (-1)^kk c X.  X <- A X
        // in place: row i of the product reads x[i-1..], so x[i-1] takes its
        // new value once row i is formed.
        double c = 1.0;
        for (int kk = 1; kk <= q; ++kk) {
            // keeps the loads of A inside the loop (hoisted, their 559 values
            // would not fit in registers)
            wave_sync();
            c = c * (double)(q - kk + 1) / (double)((2 * q - kk + 1) * kk);
            const double sgn = (kk % 2) ? -1.0 : 1.0;
            const double cf = nside ? c : sgn * c;
            double prev = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                double sum = 0.0;
#pragma unroll
                for (int l = (i > 0 ? i - 1 : 0); l < N; ++l) sum += At[i * N + l] * x[l];
                if (i > 0) x[i - 1] = prev;
                prev = sum;
                acc[i] += cf * sum;
                // one row's loads in flight at a time (bounded registers)
                __builtin_amdgcn_sched_barrier(0);
            }
            x[N - 1] = prev;
        }
        // D F = N by Gaussian elimination with partial pivoting
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const int kend = (k + BW < N) ? k + BW : N - 1;
            double *cb = colbuf[k & 1];
            if (lane == k) {
#pragma unroll
                for (int i = k; i <= kend; ++i) cb[i - k] = acc[i];
            }
            wave_sync();
            // first row of maximal |D(i, k)|, k <= i <= kend (zeros below)
            double best = fabs(cb[0]);
            int p = k;
#pragma unroll
            for (int i = k + 1; i <= kend; ++i) {
                const double v = fabs(cb[i - k]);
                if (v > best) {
                    best = v;
                    p = i;
                }
            }
#pragma unroll
            for (int i = k + 1; i <= kend; ++i) {
                if (p == i) {
                    const double tmp = acc[k];
                    acc[k] = acc[i];
                    acc[i] = tmp;
                }
            }
            const double pv = cb[p - k];
#pragma unroll
            for (int i = k + 1; i <= kend; ++i) {
                const double dik = (i == p) ? cb[0] : cb[i - k];  // D(i, k) after the exchange
                const double l = dik / pv;
                if (nside || j > k) acc[i] = acc[i] - l * acc[k];
            }
        }
        // back substitution U F = N with U the upper triangle of D (LDS,
        // row-major); every lane runs it, the N lanes keep F
        if (!nside) {
#pragma unroll
            for (int i = 0; i < N; ++i)
                if (i <= j) Ur[i * N + j] = acc[i];
        }
        wave_sync();
#pragma unroll
        for (int i = N - 1; i >= 0; --i) {
            double a = acc[i];
#pragma unroll
            for (int cc = i + 1; cc < N; ++cc) a -= Ur[i * N + cc] * acc[cc];
            acc[i] = a / Ur[i * N + i];
            __builtin_amdgcn_sched_barrier(0);
        }
        // squarings F <- F F (x holds the product)
        for (int sq = 0; sq < js; ++sq) {
            if (nside) {
#pragma unroll
                for (int i = 0; i < N; ++i) Fr[i * N + j] = acc[i];
            }
            wave_sync();
#pragma unroll
            for (int i = 0; i < N; ++i) {
                double sum = 0.0;
#pragma unroll
                for (int l = 0; l < N; ++l) sum += Fr[i * N + l] * acc[l];
                x[i] = sum;
                __builtin_amdgcn_sched_barrier(0);
            }
            wave_sync();
#pragma unroll
            for (int i = 0; i < N; ++i) acc[i] = x[i];
        }
        if (nside) {
#pragma unroll
            for (int i = 0; i < N; ++i) Nm[i + j * N] = acc[i];
        }
    }
    __syncthreads();
}

// --------------------------------------------------------------- dense expm
// Golub & Van Loan, Matrix Computations, Alg. 11.3.1 (diagonal Pade q = 6,
// scaling by 2^-j with j = max(0, 1 + floor(log2 |A|_inf)), Gaussian
// elimination with partial pivoting, j squarings), column-major N x N
// (N <= 64) in LDS, one workgroup of any size.  Same operation order as the
// oracle's orc_expm, so the two agree bit for bit on equal input.  Result in
// Nm.
//
// A 32 x 32 Hessenberg input (the Krylov case) goes to expm_wave32.  Work
// that only adds exact zeros is skipped.  With b the lower bandwidth
// of A (1 for the Krylov Hessenberg), A^k has lower bandwidth k b and the
// Pade denominator D lower bandwidth 6 b; elimination with partial pivoting
// keeps that bandwidth in the multipliers (the rows below it hold zeros in
// the pivot column, so their multiplier is 0 and their update subtracts
// zeros), so each step searches and updates only the 6 b rows under the
// pivot.  Row exchanges are a logical permutation (perm), so a step needs
// one barrier.
__device__ __forceinline__ void expm_block(int N, double *A, double *X, double *T, double *Nm, double *Dm,
                                  double *red)
{
    (void)red;
    const int NN = N * N;
    const int tid = threadIdx.x;
    __shared__ int js_sh, lb_sh;
    __shared__ int perm[2][kExpmMaxN];
    // infinity norm (row sums in column order, max in any order: exact)
    // and the lower bandwidth
    if (tid < kWave) {
        double rs = 0.0;
        int lb = 0;
        if (tid < N)
            for (int j = 0; j < N; ++j) {
                const double v = A[tid + j * N];
                rs += fabs(v);
                if (v != 0.0 && tid - j > lb) lb = tid - j;
            }
        double mx = rs;
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
            mx = fmax(mx, __shfl_xor(mx, off, kWave));
            lb = max(lb, __shfl_xor(lb, off, kWave));
        }
        if (tid == 0) {
            int js = 0;
            if (mx > 0.0) {
                int e;
                frexp(mx, &e);
                js = e > 0 ? e : 0;
            }
            js_sh = js;
            lb_sh = lb;
        }
    }
    __syncthreads();
    const int js = js_sh;
    const int lb = lb_sh;
    const double scale = ldexp(1.0, -js);
    for (int k = tid; k < NN; k += blockDim.x) {
        A[k] = A[k] * scale;
        const double id = ((k % N) == (k / N)) ? 1.0 : 0.0;
        X[k] = id;
        Nm[k] = id;
        Dm[k] = id;
    }
    for (int i = tid; i < N; i += blockDim.x) perm[0][i] = i;
    __syncthreads();
    if (N == 32 && lb <= 1 && blockDim.x >= kWave) {
        expm_wave32(A, T, X, Dm, Nm, js);
        return;
    }
    const int q = 6;
    double c = 1.0;
    for (int kk = 1; kk <= q; ++kk) {
        c = c * (double)(q - kk + 1) / (double)((2 * q - kk + 1) * kk);
        // X = A^(kk-1) has lower bandwidth (kk-1) lb
        matmul_lds(N, A, X, T, lb, min(N, (kk - 1) * lb));
        __syncthreads();
        const double sgn = (kk % 2) ? -1.0 : 1.0;
        for (int e = tid; e < NN; e += blockDim.x) {
            X[e] = T[e];
            Nm[e] += c * T[e];
            Dm[e] += sgn * c * T[e];
        }
        __syncthreads();
    }
    // solve Dm F = Nm by Gaussian elimination with partial pivoting; row
    // exchanges go to the permutation (logical row i is physical row
    // perm[cur][i])
    const int bw = min(N - 1, q * lb);
    int cur = 0;
    for (int k = 0; k < N; ++k) {
        const int *pc = perm[cur];
        int *pn = perm[cur ^ 1];
        const int iend = min(N - 1, k + bw);
        // first logical row of maximal |D(row, k)|, k <= row <= iend (the
        // oracle's choice: the rows below hold zeros), found by every thread
        int p = k;
        double best = fabs(Dm[pc[k] + k * N]);
        for (int i = k + 1; i <= iend; ++i) {
            const double v = fabs(Dm[pc[i] + k * N]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        for (int i = tid; i < N; i += blockDim.x) pn[i] = (i == k) ? pc[p] : (i == p ? pc[k] : pc[i]);
        const int rk = pc[p];
        const double pv = Dm[rk + k * N];
        // updates of the logical rows k+1..iend by (row, column chunk)
        // pairs: each pair forms its row's multiplier (the same division in
        // every pair), loads its chunk of the pivot row and of its own row,
        // then stores the updated chunk.  The multipliers are not kept: only
        // the upper triangle is read afterwards.
        const int rows = iend - k;
        const int chunks = (N + kElimChunk - 1) / kElimChunk;
        for (int e = tid; e < rows * chunks; e += blockDim.x) {
            const int i = k + 1 + e % rows;
            const int ri = (i == p) ? pc[k] : pc[i];
            const int j0 = (e / rows) * kElimChunk;
            const double l = Dm[ri + k * N] / pv;
            double pd[kElimChunk], pn2[kElimChunk], rd[kElimChunk], rn[kElimChunk];
#pragma unroll
            for (int u = 0; u < kElimChunk; ++u) {
                const int j = min(j0 + u, N - 1);
                pd[u] = Dm[rk + j * N];
                pn2[u] = Nm[rk + j * N];
                rd[u] = Dm[ri + j * N];
                rn[u] = Nm[ri + j * N];
            }
#pragma unroll
            for (int u = 0; u < kElimChunk; ++u) {
                const int j = j0 + u;
                if (j < N) {
                    if (j > k) Dm[ri + j * N] = rd[u] - l * pd[u];
                    Nm[ri + j * N] = rn[u] - l * pn2[u];
                }
            }
        }
        __syncthreads();
        cur ^= 1;
    }
    const int *pf = perm[cur];
    if (N == 32) {
        back_subst_32(Dm, Nm, pf);
    } else {
        // one right-hand side per thread on a row-major copy of Nm in X in
        // logical row order (consecutive lanes read consecutive LDS words)
        for (int e = tid; e < NN; e += blockDim.x) {
            const int i = e % N, j = e / N;
            X[j + i * N] = Nm[pf[i] + j * N];
        }
        __syncthreads();
        for (int j = tid; j < N; j += blockDim.x) {
            for (int i = N - 1; i >= 0; --i) {
                const int ri = pf[i];
                double acc = X[j + i * N];
                int k = i + 1;
                for (; k + kSubChunk <= N; k += kSubChunk) {
                    double d[kSubChunk], x[kSubChunk];
#pragma unroll
                    for (int u = 0; u < kSubChunk; ++u) {
                        d[u] = Dm[ri + (k + u) * N];
                        x[u] = X[j + (k + u) * N];
                    }
#pragma unroll
                    for (int u = 0; u < kSubChunk; ++u) acc -= d[u] * x[u];
                }
                for (; k < N; ++k) acc -= Dm[ri + k * N] * X[j + k * N];
                X[j + i * N] = acc / Dm[ri + i * N];
            }
        }
        __syncthreads();
        for (int e = tid; e < NN; e += blockDim.x) Nm[e] = X[(e / N) + (e % N) * N];
    }
    __syncthreads();
    for (int sq = 0; sq < js; ++sq) {
        matmul_lds(N, Nm, Nm, T, N, N);
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
