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

// A pair of doubles at a 16-byte boundary: one 128-bit LDS access.
struct alignas(16) dpair {
    double x, y;
};

// Lane `lane`'s value of v (lane uniform across the wave): two 32-bit
// readlanes, no memory.
__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// The 2 x 2 block (i0..i0+1, j0..j0+1) of the 32 x 32 product P Q with P
// column-major (P(i, l) at P[i + 32 l]) and Q by rows of stride qs (Q(l, j)
// at Q[qs l + j]); every element is sum_l P(i, l) Q(l, j) over all 32 l in
// increasing order from 0.0, the oracle's matmul_cm.  t = {(i0, j0),
// (i0, j0+1), (i0+1, j0), (i0+1, j0+1)}.
__device__ __forceinline__ void prod32_2x2(const double *P, const double *Q, int qs, int i0, int j0,
                                           double (&t)[4])
{
    double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
#pragma unroll 8
    for (int l = 0; l < 32; ++l) {
        const dpair a = *reinterpret_cast<const dpair *>(P + i0 + 32 * l);
        const dpair b = *reinterpret_cast<const dpair *>(Q + qs * l + j0);
        t0 += a.x * b.x;
        t1 += a.x * b.y;
        t2 += a.y * b.x;
        t3 += a.y * b.y;
    }
    t[0] = t0;
    t[1] = t1;
    t[2] = t2;
    t[3] = t3;
}

// Stores a 2 x 2 block by rows of stride rs / column-major (stride 32).
__device__ __forceinline__ void store_rows_2x2(double *Q, int rs, int i0, int j0, const double (&t)[4])
{
    *reinterpret_cast<dpair *>(Q + rs * i0 + j0) = dpair{t[0], t[1]};
    *reinterpret_cast<dpair *>(Q + rs * (i0 + 1) + j0) = dpair{t[2], t[3]};
}
__device__ __forceinline__ void store_cols_2x2(double *P, int i0, int j0, const double (&t)[4])
{
    *reinterpret_cast<dpair *>(P + i0 + 32 * j0) = dpair{t[0], t[2]};
    *reinterpret_cast<dpair *>(P + i0 + 32 * (j0 + 1)) = dpair{t[1], t[3]};
}

// Step k of the banded elimination of expm32_hess on one wavefront: lane c
// holds rows k..k+6 of column c of [D | N] (32 x 64, by rows of 64 in XT).
// The pivot column (lane k's window) comes by readlane; the first row of
// maximal |D(i, k)| is exchanged with row k in every column; lanes 0..5
// divide the six multipliers (one division each, broadcast by readlane);
// rows k+1..k+6 are updated in every column.  Columns c <= k of D are updated
// too: they hold the multipliers in the oracle and are never read again.
// Row k, now final, goes to XT and row k + 7 enters the window.  kTail:
// rows past 31 are skipped (k > 24; k is a constant under unrolling).
template <bool kTail>
__device__ __forceinline__ void elim_step32(int k, int lane, double (&w)[7], double *XT)
{
    constexpr int N = 32, BW = 6, W = 2 * N;
    double v[BW + 1];
#pragma unroll
    for (int r = 0; r <= BW; ++r) v[r] = (!kTail || k + r < N) ? readlane_f64(w[r], k) : 0.0;
    int p = 0;
    double best = fabs(v[0]);
#pragma unroll
    for (int r = 1; r <= BW; ++r) {
        if ((!kTail || k + r < N) && fabs(v[r]) > best) {
            best = fabs(v[r]);
            p = r;
        }
    }
    double wp = w[0], num = 0.0;
#pragma unroll
    for (int r = 1; r <= BW; ++r) {
        if (p == r) {
            wp = w[r];
            w[r] = w[0];
        }
        // D(k + r, k) after the exchange
        if (lane == r - 1) num = (p == r) ? v[0] : v[r];
    }
    w[0] = wp;
    const double lq = num / readlane_f64(wp, k);
#pragma unroll
    for (int r = 1; r <= BW; ++r) {
        if (!kTail || k + r < N) {
            const double l = readlane_f64(lq, r - 1);
            w[r] = w[r] - l * w[0];
        }
    }
    XT[W * k + lane] = w[0];
#pragma unroll
    for (int r = 0; r < BW; ++r) w[r] = w[r + 1];
    w[BW] = (!kTail || k + BW + 1 < N) ? XT[W * (k + BW + 1) + lane] : 0.0;
}

// The 32 x 32 exponential of a scaled matrix with lower bandwidth <= 1 (the
// Krylov Hessenberg) by a workgroup of exactly 256 threads.  Every value
// follows the operation order of the oracle's orc_expm: products sum all 32
// terms in increasing l from 0.0, the elimination runs row by row, back
// substitution with k ascending, then the squarings.
//
//  * Pade polynomials: thread t owns the 2 x 2 block (2 (t % 16), 2 (t / 16))
//    of every matrix; its two rows of A stay in registers for the six
//    products, the powers X = A^k go through LDS by rows, and the blocks of N
//    and D accumulate in registers.
//  * D F = N: one wavefront, lane c holding column c of the 32 x 64 matrix
//    [D | N] (row-major in LDS, so a row is one conflict-free access).  Only
//    the 7 rows k..k+6 change in step k (D has lower bandwidth 6; the rows
//    below hold zeros in the pivot column, their multipliers are 0 and their
//    updates subtract zeros), so the lane keeps them in a register window
//    that slides down one row per step.  The pivot column comes from lane k
//    by readlane; lanes 0..5 divide the 6 multipliers, one division each.
//  * U F = N: lanes 32..63, one column each, U read by rows from LDS.
//  * Squarings: the 2 x 2 blocks again, F by columns and by rows in LDS.
//
// A: the scaled matrix (column-major); XT: 2 N^2 doubles (X then T in the
// caller's layout) for the powers and then [D | N]; Nm: the result
// (column-major).  All three 16-byte aligned.
__device__ __forceinline__ void expm32_hess(const double *A, double *XT, double *Nm, int js)
{
    constexpr int N = 32, q = 6, BW = 6, W = 2 * N;
    const int tid = threadIdx.x;
    const int i0 = 2 * (tid & 15), j0 = 2 * (tid >> 4);
    double ar[2][N];  // rows i0, i0 + 1 of A
#pragma unroll
    for (int l = 0; l < N; ++l) {
        const dpair a = *reinterpret_cast<const dpair *>(A + i0 + N * l);
        ar[0][l] = a.x;
        ar[1][l] = a.y;
    }
    double nb[4], db[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const double id = (i0 + (u >> 1) == j0 + (u & 1)) ? 1.0 : 0.0;
        nb[u] = id;
        db[u] = id;
    }
    // X = I by rows
    for (int e = tid; e < N * N; e += blockDim.x) XT[e] = ((e >> 5) == (e & 31)) ? 1.0 : 0.0;
    __syncthreads();
    double c = 1.0;
    for (int kk = 1; kk <= q; ++kk) {
        c = c * (double)(q - kk + 1) / (double)((2 * q - kk + 1) * kk);
        const double sgn = (kk % 2) ? -1.0 : 1.0;
        double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
#pragma unroll
        for (int l = 0; l < N; ++l) {
            const dpair b = *reinterpret_cast<const dpair *>(XT + N * l + j0);
            t0 += ar[0][l] * b.x;
            t1 += ar[0][l] * b.y;
            t2 += ar[1][l] * b.x;
            t3 += ar[1][l] * b.y;
        }
        const double t[4] = {t0, t1, t2, t3};
        __syncthreads();
        store_rows_2x2(XT, N, i0, j0, t);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            nb[u] += c * t[u];
            db[u] += sgn * c * t[u];
        }
        __syncthreads();
    }
    // [D | N] by rows of W = 64 over XT
    store_rows_2x2(XT, W, i0, j0, db);
    store_rows_2x2(XT + N, W, i0, j0, nb);
    __syncthreads();
    if (tid < kWave) {
        const int lane = tid;
        double w[BW + 1];  // rows k..k+6 of column `lane`
#pragma unroll
        for (int r = 0; r <= BW; ++r) w[r] = XT[W * r + lane];
        // steps with all six rows below the pivot and row k + 7 in the matrix
        for (int k = 0; k < N - BW - 1; ++k) elim_step32<false>(k, lane, w, XT);
#pragma unroll
        for (int k = N - BW - 1; k < N; ++k) elim_step32<true>(k, lane, w, XT);
        wave_sync();
        if (lane >= N) {
            double x[N];
#pragma unroll
            for (int i = 0; i < N; ++i) x[i] = XT[W * i + lane];
#pragma unroll
            for (int i = N - 1; i >= 0; --i) {
                double a = x[i];
#pragma unroll
                for (int cc = i + 1; cc < N; ++cc) a -= XT[W * i + cc] * x[cc];
                x[i] = a / XT[W * i + i];
            }
#pragma unroll
            for (int i = 0; i < N; ++i) XT[W * i + lane] = x[i];
        }
    }
    __syncthreads();
    // F (the right half of XT by rows) by columns into Nm
    double f[4];
    f[0] = XT[W * i0 + N + j0];
    f[1] = XT[W * i0 + N + j0 + 1];
    f[2] = XT[W * (i0 + 1) + N + j0];
    f[3] = XT[W * (i0 + 1) + N + j0 + 1];
    store_cols_2x2(Nm, i0, j0, f);
    __syncthreads();
    for (int sq = 0; sq < js; ++sq) {
        prod32_2x2(Nm, XT + N, W, i0, j0, f);
        __syncthreads();
        store_cols_2x2(Nm, i0, j0, f);
        store_rows_2x2(XT + N, W, i0, j0, f);
        __syncthreads();
    }
}

// --------------------------------------------------------------- dense expm
// Golub & Van Loan, Matrix Computations, Alg. 11.3.1 (diagonal Pade q = 6,
// scaling by 2^-j with j = max(0, 1 + floor(log2 |A|_inf)), Gaussian
// elimination with partial pivoting, j squarings), column-major N x N
// (N <= 64) in LDS, one workgroup of any size.  Same operation order as the
// oracle's orc_expm, so the two agree bit for bit on equal input.  Result in
// Nm.
//
// A 32 x 32 Hessenberg input (the Krylov case) goes to expm32_hess.  Work
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
    if (N == 32 && lb <= 1 && blockDim.x == 4 * kWave && T == X + NN &&
        ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(X) |
          reinterpret_cast<uintptr_t>(Nm)) & 15) == 0) {
        expm32_hess(A, X, Nm, js);
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
