// fsp_krylov_dev.h -- device state and per-step bodies of the Krylov expv
// (Expokit scheme, src/KryExpvFSP.cu:38-182), shared by the single-solve
// engine (fsp_krylov.hip) and the lockstep batch (fsp_batch_lockstep.hip).
#pragma once

#include <type_traits>

#include "fsp_expm.h"
#include "fsp_formats.h"

namespace fsp {

// One solve's step state: scalars of KryExpvFSP::step, the (m+2)^2
// Hessenberg matrix (column-major), the scaling of each Arnoldi vector
// (sigma_j = 1 / h(j, j-1)), the first column of exp(tau H) (F0), the Gram
// matrix of an FP32 basis and the decimal powers used by the step-size
// rounding.
struct KState {
    double beta, btol, t_now, t_new, t_final, tau, xm, err_loc, avnorm, fact;
    double tol, anorm;
    double tau_last;
    double sigma[kMaxKrylov + 2];
    double F0[kMaxKrylov + 2];
    double H[(kMaxKrylov + 2) * (kMaxKrylov + 2)];
    // FP32 basis: Gram matrix V_k . V_l of the stored basis, leading
    // dimension kMaxKrylov + 1
    double G[(kMaxKrylov + 1) * (kMaxKrylov + 1)];
    // FP32 basis: the raw dots V^T w of the current column, read by the
    // update kernel while its workgroup 0 replaces H(., j) by 2h - G h
    double hs[kMaxKrylov + 2];
    double p10[kP10Hi - kP10Lo + 1];   // correctly rounded 1e-300 .. 1e300
    double *tau_log;                   // accepted step sizes, by step index
    long long tau_cap;
    int mb, k1, broke, active, mx, status, pad0, pad1;
    long long i_step, n_reject, n_breakdown, n_matvec;
};

constexpr int kGramLd = kMaxKrylov + 1;  // leading dimension of KState::G

// An FP32 basis (B = float) is projected out with its Gram matrix.
template <typename B>
constexpr bool gram_basis() { return !std::is_same<B, double>::value; }

// Gram-corrected Gram-Schmidt coefficients for column j over the window
// istart..j, one workgroup: c_k = 2 h_k - sum_l G(k, l) h_l (l ascending),
// with h = H(., j) = V^T w on entry and c written back as H(., j).  This
// folds the second classical Gram-Schmidt pass, whose coefficients are
// V^T (w - V h) = h - G h, into the first, so the update sweeps the basis
// once.  lds: (kMaxKrylov + 1)^2 + (kMaxKrylov + 2) doubles.
__device__ inline void gram_coefs(KState *st, int istart, int j, int M2, double *lds)
{
    const int w = j - istart + 1;
    double *gw = lds, *hw = lds + (kMaxKrylov + 1) * (kMaxKrylov + 1);
    double *h = st->H + (size_t)j * M2;
    for (int e = threadIdx.x; e < w * w; e += blockDim.x)
        gw[e] = st->G[(istart + e % w) + (istart + e / w) * kGramLd];
    for (int e = threadIdx.x; e < w; e += blockDim.x) hw[e] = h[istart + e];
    __syncthreads();
    for (int k = threadIdx.x; k < w; k += blockDim.x) {
        double s = 0.0;
        for (int l = 0; l < w; ++l) s += gw[k + l * w] * hw[l];
        h[istart + k] = 2.0 * hw[k] - s;
    }
}

constexpr size_t kGramLdsBytes = sizeof(double) * ((kMaxKrylov + 1) * (kMaxKrylov + 1) + kMaxKrylov + 2);

// The same coefficients from the raw dots st->hs into LDS (cw[0..w)), for
// every workgroup of a kernel; nothing is written to global memory.
// lds: kGramLdsBytes.
__device__ inline void gram_coefs_lds(const KState *st, int istart, int j, double *lds, double *cw)
{
    const int w = j - istart + 1;
    double *gw = lds, *hw = lds + (kMaxKrylov + 1) * (kMaxKrylov + 1);
    for (int e = threadIdx.x; e < w * w; e += blockDim.x)
        gw[e] = st->G[(istart + e % w) + (istart + e / w) * kGramLd];
    for (int e = threadIdx.x; e < w; e += blockDim.x) hw[e] = st->hs[istart + e];
    __syncthreads();
    for (int k = threadIdx.x; k < w; k += blockDim.x) {
        double s = 0.0;
        for (int l = 0; l < w; ++l) s += gw[k + l * w] * hw[l];
        cw[k] = 2.0 * hw[k] - s;
    }
    __syncthreads();
}

// Krylov vector j = in * sigma_j as the basis stores it: rounded to B for
// j >= 1; V_0 (the scaled solution vector) is kept in FP64 with any B, so
// no rounding noise enters the solution's own direction.
template <typename B>
__device__ inline double kry_scale(double v, double sig, int j)
{
    return j > 0 ? (double)(B)(v * sig) : v * sig;
}

template <typename B>
struct KryX {
    const double *__restrict__ in;
    double sig;
    int j;
    __device__ double operator()(uint32_t c) const { return kry_scale<B>(in[c], sig, j); }
};

// Basis vector k, entry i: FP64 V_0..V_m (B = double), or V_0 in FP64 and
// V_1..V_m in B.  Both candidate loads are issued (each address is valid
// for any k) and one is selected, so a loop over k has no branch between
// its loads.
template <typename B>
__device__ inline double basis(const double *__restrict__ V64, const B *__restrict__ Vb,
                               size_t nv, int k, uint32_t i)
{
    if (std::is_same<B, double>::value) return V64[(size_t)k * nv + i];
    const double v0 = V64[i];
    const double vb = (double)Vb[(size_t)(k > 0 ? k - 1 : 0) * nv + i];
    return k == 0 ? v0 : vb;
}

template <typename B>
__device__ inline void store_basis(double *__restrict__ V64, B *__restrict__ Vb, size_t nv,
                                   int k, uint32_t i, double v)
{
    if (std::is_same<B, double>::value) {
        V64[(size_t)k * nv + i] = v;
    } else if (k == 0) {
        V64[i] = v;
    } else {
        Vb[(size_t)(k - 1) * nv + i] = (B)v;
    }
}

// Step prologue (src/KryExpvFSP.cu:41-65), one workgroup of kBlock threads:
// beta = |v| from G dd partials of |v|^2, first step size, tau =
// min(t_final - t_now, t_new).  Marks the step inactive when t_final is
// reached or the solve has failed, so steps launched ahead of a host check
// do nothing; a solve that would exceed max_nstep fails here.
__device__ inline void kry_step_begin(KState *st, const double *partials, int G, int m, int M2,
                                      long long max_nstep, double *red)
{
    __shared__ int act;
    if (threadIdx.x == 0) {
        int a = st->t_now < st->t_final && st->status == 0;
        if (a && st->i_step >= max_nstep) {
            st->status = FSP_ERR_NOT_CONVERGED;
            a = 0;
        }
        act = a;
    }
    __syncthreads();
    if (!act) {
        if (threadIdx.x == 0) st->active = 0;
        return;
    }
    const double s = dd_value(sum_dd_partials(partials, G, red));
    for (int k = threadIdx.x; k < M2 * M2; k += kBlock) st->H[k] = 0.0;
    if (threadIdx.x == 0) {
        st->active = 1;
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
}

// Error estimate, reject loop and step update (src/KryExpvFSP.cu:122-174),
// one workgroup; smem: 5 (m+2)^2 doubles of LDS.
__device__ __forceinline__ void kry_control(KState *st, int m, int max_reject, double *smem)
{
    if (!st->active) return;
    const int M2 = m + 2;
    const int NN = M2 * M2;
    double *A = smem, *X = A + NN, *T = X + NN, *Nm = T + NN, *Dm = Nm + NN;
    __shared__ double red[kBlock / kWave];
    const int mb = st->mb, k1 = st->k1;
    const double beta = st->beta, avnorm = st->avnorm, btol = st->btol, tol = st->tol;
    double tau = st->tau, xm = st->xm, err = 0.0;
    int ireject = 0;
    while (ireject < max_reject) {
        const int mx = mb + k1;
        for (int e = threadIdx.x; e < NN; e += blockDim.x) A[e] = tau * st->H[e];
        __syncthreads();
        expm_block(M2, A, X, T, Nm, Dm, red);
        if (k1 == 0) {
            err = btol;
            break;
        }
        const double phi1 = fabs(beta * Nm[(mx - 2) + 0 * M2]);
        const double phi2 = fabs(beta * Nm[(mx - 1) + 0 * M2] * avnorm);
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
        tau = round_up_2sig(st->p10, tau);
        ++ireject;
        __syncthreads();
    }
    const int mx = mb + max(0, k1 - 1);
    for (int k = threadIdx.x; k < M2; k += blockDim.x) st->F0[k] = (k < mx) ? Nm[k] : 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
        st->mx = mx;
        st->n_reject += ireject;
        if (ireject >= max_reject) st->status = FSP_ERR_NOT_CONVERGED;
        st->err_loc = err;
        st->xm = xm;
        st->tau_last = tau;
        if (st->tau_log && st->i_step < st->tau_cap) st->tau_log[st->i_step] = tau;
        st->t_now = st->t_now + tau;
        double tn = 0.9 * tau * pow(tau * tol / err, xm);
        st->t_new = round_up_2sig(st->p10, tn);
        st->i_step += 1;
    }
}

}  // namespace fsp
