// fsp_batch_kernel.h -- device half of the batched persistent solver: the
// per-instance records, the workgroup-level Krylov expv (solve_one) and the
// work-queue kernel.  The host half (set-up, graph capture, autotuner, C ABI)
// is fsp_batch.hip.
#pragma once

#include <type_traits>

#include "fsp_expm.h"
#include "fsp_formats.h"

namespace fsp {

struct BatchInst {
    const DevModel *dm;
    const double *tab;
    const double *coefs;
    double *p;
    const double *p0;
    double t_final;
    int n, nv, with_sink, tab_size;
    long long steps, rejects;
    int status, pad;
};

struct BatchArgs {
    BatchInst *inst;
    const int *order;
    int count;
    unsigned int *queue;
    double *ws;          // FP64 workspace, ws_stride doubles per workgroup
    float *ws32;         // FP32 basis (mixed), ws32_stride floats per workgroup
    size_t ws_stride, ws32_stride;
    int nv_max, m, iop, q_iop, max_reject, tab_lds;
    long long max_nstep;
    double tol, anorm, fact;
    const double *p10;
};

namespace {

template <int BS>
__device__ inline dd bsum_dd(dd v, double *red)
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
    for (int k = 0; k < BS / kWave; ++k) s = dd_merge(s, dd{red[2 * k], red[2 * k + 1]});
    return s;
}

template <typename B>
__device__ inline double bget(const double *V64, const B *Vb, size_t nv, int k, int i)
{
    if (std::is_same<B, double>::value) return V64[(size_t)k * nv + i];
    return k == 0 ? V64[i] : (double)Vb[(size_t)(k - 1) * nv + i];
}

template <typename B>
__device__ inline void bput(double *V64, B *Vb, size_t nv, int k, int i, double v)
{
    if (std::is_same<B, double>::value)
        V64[(size_t)k * nv + i] = v;
    else if (k == 0)
        V64[i] = v;
    else
        Vb[(size_t)(k - 1) * nv + i] = (B)v;
}

// Krylov vector j = in * sig as stored: rounded to B for j >= 1, FP64 for
// V_0 (the scaled solution vector) with any B.
template <typename B>
__device__ inline double bscale(double v, double sig, int j)
{
    return j > 0 ? (double)(B)(v * sig) : v * sig;
}

template <typename B>
struct BX {
    const double *__restrict__ in;
    double sig;
    int j;
    __device__ double operator()(uint32_t c) const { return bscale<B>(in[c], sig, j); }
};

template <int BS, typename B, int NKB, int DMAX>
__device__ void solve_one(const BatchArgs &a, BatchInst &I, double *sm)
{
    const int m = a.m, M2 = m + 2, NN = M2 * M2;
    double *H = sm;
    double *EA = H + NN, *EX = EA + NN, *ET = EX + NN, *EN = ET + NN, *ED = EN + NN;
    double *red = ED + NN;
    double *F0 = red + 2 * (BS / kWave);
    double *tabl = F0 + M2;
    const int tid = threadIdx.x;
    const int n = I.n, nv = I.nv;
    // workspace of this workgroup
    double *V64 = a.ws + (size_t)blockIdx.x * a.ws_stride;
    B *Vb = std::is_same<B, double>::value ? (B *)(void *)V64
                                           : (B *)(void *)(a.ws32 + (size_t)blockIdx.x * a.ws32_stride);
    const size_t vcount = std::is_same<B, double>::value ? (size_t)(m + 1) : 1;
    double *W[2] = {V64 + vcount * (size_t)a.nv_max, V64 + (vcount + 1) * (size_t)a.nv_max};
    // factor tables
    const double *tab = I.tab;
    if (I.tab && I.tab_size <= a.tab_lds) {
        for (int k = tid; k < I.tab_size; k += BS) tabl[k] = I.tab[k];
        tab = tabl;
    }
    KronF<double> K;
    K.dm = I.dm;
    K.tab_g = I.tab;
    K.prop_table = nullptr;
    K.coefs = I.coefs;
    K.n = (uint32_t)n;
    K.R = I.dm->R;
    K.tab_size = I.tab_size;
    K.with_sink = I.with_sink;
    K.single = -1;
    K.single_coef = 0.0;
    double *p = I.p;
    for (int i = tid; i < nv; i += BS) p[i] = I.p0[i];
    __syncthreads();

    const double tol = a.tol, anorm = a.anorm;
    const double t_final = I.t_final;
    double t_now = 0.0, t_new = 0.0, xm = 1.0 / m, btol = anorm * tol, err = 0.0;
    long long i_step = 0, nrej = 0;
    int k1 = 2, status = 0;
    while (t_now < t_final) {
        if (i_step >= a.max_nstep) {
            status = FSP_ERR_NOT_CONVERGED;
            break;
        }
        dd s2 = dd_zero();
        for (int i = tid; i < nv; i += BS) s2 = dd_add_prod(s2, p[i], p[i]);
        const double beta = sqrt(dd_value(bsum_dd<BS>(s2, red)));
        if (!(beta > 0.0) || !(beta < INFINITY)) {
            // v = 0 stays 0; a non-finite vector is an error
            if (beta != 0.0) status = FSP_ERR_NOT_CONVERGED;
            break;
        }
        if (i_step == 0) {
            xm = 1.0 / (double)m;
            t_new = (1.0 / anorm) * pow((a.fact * tol) / (4.0 * beta * anorm), xm);
            btol = anorm * tol;
        }
        int mb = m;
        double tau = fmin(t_final - t_now, t_new);
        double avnorm = 0.0;
        for (int e = tid; e < NN; e += BS) H[e] = 0.0;
        __syncthreads();
        double sig = 1.0 / beta;
        int istart = 0;
        for (int j = 0; j <= m; ++j) {
            const bool last = (j == m);
            if (last && k1 == 0) break;
            if (!last && a.iop) istart = (j >= a.q_iop - 1) ? j - a.q_iop + 1 : 0;
            const double *in = (j == 0) ? p : W[j & 1];
            double *out = W[(j + 1) & 1];
            const BX<B> X{in, sig, j};
            const int nk = last ? 1 : j - istart + 1;
            dd acc[NKB];
#pragma unroll
            for (int kk = 0; kk < NKB; ++kk) acc[kk] = dd_zero();
            dd sink = dd_zero();
            for (int i = tid; i < n; i += BS) {
                const double xi = X(i);
                bput<B>(V64, Vb, nv, j, i, xi);
                const double y = kron_row<DMAX, double>(K, (uint32_t)i, X, tab, sink);
                if (!last) {
                    out[i] = y;
#pragma unroll
                    for (int kk = 0; kk < NKB; ++kk) {
                        if (kk < nk) {
                            const int k = istart + kk;
                            acc[kk] = dd_add_prod(acc[kk], y, k == j ? xi : bget<B>(V64, Vb, nv, k, i));
                        }
                    }
                } else {
                    acc[0] = dd_add_prod(acc[0], y, y);
                }
            }
            double ys = 0.0, xs = 0.0;
            if (I.with_sink) {
                ys = dd_value(bsum_dd<BS>(sink, red));
                xs = bscale<B>(in[n], sig, j);
            }
#pragma unroll
            for (int kk = 0; kk < NKB; ++kk) {
                if (kk < nk) {
                    dd t = bsum_dd<BS>(acc[kk], red);
                    if (!last) {
                        const int k = istart + kk;
                        if (I.with_sink)
                            t = dd_add_prod(t, ys, k == j ? xs : bget<B>(V64, Vb, nv, k, n));
                        if (tid == 0) H[k + j * M2] = dd_value(t);
                    } else {
                        if (I.with_sink) t = dd_add_prod(t, ys, ys);
                        avnorm = sqrt(dd_value(t));
                    }
                }
            }
            __syncthreads();
            if (tid == 0 && I.with_sink) {
                bput<B>(V64, Vb, nv, j, n, xs);
                if (!last) out[n] = ys;
            }
            if (last) {
                if (tid == 0) H[(m + 1) + m * M2] = 1.0;  // src/KryExpvFSP.cu:115
                break;
            }
            __syncthreads();
            // w_{j+1} -= V[istart..j] H(istart..j, j), |w|
            dd nr = dd_zero();
            for (int i = tid; i < nv; i += BS) {
                double accv = 0.0;
                for (int k = istart; k <= j; ++k) accv += bget<B>(V64, Vb, nv, k, i) * H[k + j * M2];
                const double w = -1.0 * accv + out[i];
                out[i] = w;
                nr = dd_add_prod(nr, w, w);
            }
            if (!std::is_same<B, double>::value) {
                // FP32 basis: a second Gram-Schmidt pass.  With one pass the
                // loss of orthogonality compounds from the FP32 rounding level
                // by the factor |w| / h(j+1, j) per step and can reach 1
                // within m steps; the second pass keeps it at that level.
                // Coefficients h2 in EA (free until the error estimate).
                dd acc2[NKB];
#pragma unroll
                for (int kk = 0; kk < NKB; ++kk) acc2[kk] = dd_zero();
                for (int i = tid; i < nv; i += BS) {
                    const double w = out[i];
#pragma unroll
                    for (int kk = 0; kk < NKB; ++kk)
                        if (kk < nk) acc2[kk] = dd_add_prod(acc2[kk], w, bget<B>(V64, Vb, nv, istart + kk, i));
                }
#pragma unroll
                for (int kk = 0; kk < NKB; ++kk) {
                    if (kk < nk) {
                        const double h2 = dd_value(bsum_dd<BS>(acc2[kk], red));
                        if (tid == 0) {
                            EA[kk] = h2;
                            H[(istart + kk) + j * M2] += h2;
                        }
                    }
                }
                __syncthreads();
                nr = dd_zero();
                for (int i = tid; i < nv; i += BS) {
                    double accv = 0.0;
                    for (int k = istart; k <= j; ++k) accv += bget<B>(V64, Vb, nv, k, i) * EA[k - istart];
                    const double w = -1.0 * accv + out[i];
                    out[i] = w;
                    nr = dd_add_prod(nr, w, w);
                }
            }
            const double s = sqrt(dd_value(bsum_dd<BS>(nr, red)));
            if (s < btol) {  // happy breakdown
                k1 = 0;
                mb = j + 1;
                tau = t_final - t_now;
                break;
            }
            if (tid == 0) H[(j + 1) + j * M2] = s;
            sig = 1.0 / s;
            __syncthreads();
        }
        __syncthreads();
        // error estimate and reject loop (src/KryExpvFSP.cu:122-158)
        int ireject = 0;
        while (ireject < a.max_reject) {
            const int mx = mb + k1;
            for (int e = tid; e < NN; e += BS) EA[e] = tau * H[e];
            __syncthreads();
            expm_block(M2, EA, EX, ET, EN, ED, red);
            if (k1 == 0) {
                err = btol;
                break;
            }
            const double phi1 = fabs(beta * EN[mx - 2]);
            const double phi2 = fabs(beta * EN[mx - 1] * avnorm);
            if (phi1 > phi2 * 10.0) {
                err = phi2;
                xm = 1.0 / (double)mx;
            } else if (phi1 > phi2) {
                err = (phi1 * phi2) / (phi1 - phi2);
                xm = 1.0 / (double)mx;
            } else {
                err = phi1;
                xm = 1.0 / (double)(mx - 1);
            }
            if (err <= 1.2 * tau * tol) break;
            tau = 0.9 * tau * pow(tau * tol / err, xm);
            tau = round_up_2sig(a.p10, tau);
            ++ireject;
            __syncthreads();
        }
        nrej += ireject;
        if (ireject >= a.max_reject) status = FSP_ERR_NOT_CONVERGED;
        const int mx = mb + max(0, k1 - 1);
        for (int k = tid; k < M2; k += BS) F0[k] = (k < mx) ? EN[k] : 0.0;
        __syncthreads();
        for (int i = tid; i < nv; i += BS) {
            double accv = 0.0;
            for (int k = 0; k < mx; ++k) accv += bget<B>(V64, Vb, nv, k, i) * F0[k];
            p[i] = beta * accv;
        }
        if (!(tau > 0.0) || status != 0) {  // no progress possible: stop with an error
            status = FSP_ERR_NOT_CONVERGED;
            break;
        }
        t_now = t_now + tau;
        t_new = round_up_2sig(a.p10, 0.9 * tau * pow(tau * tol / err, xm));
        ++i_step;
        __syncthreads();
    }
    if (tid == 0) {
        I.steps = i_step;
        I.rejects = nrej;
        I.status = status;
    }
}

template <int BS, typename B, int NKB, int DMAX>
__global__ void __launch_bounds__(BS) k_batch(BatchArgs a)
{
    FSP_DYNAMIC_LDS(sm);
    __shared__ int q_sh;
    for (;;) {
        if (threadIdx.x == 0) q_sh = (int)atomicAdd(a.queue, 1u);
        __syncthreads();
        const int q = q_sh;
        __syncthreads();
        if (q >= a.count) return;  // every wave of the workgroup sees the same q
        solve_one<BS, B, NKB, DMAX>(a, a.inst[a.order[q]], sm);
    }
}

__global__ void k_batch_reset(unsigned int *queue) { *queue = 0u; }

}  // namespace

}  // namespace fsp
