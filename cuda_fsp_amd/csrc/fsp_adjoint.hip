// fsp_adjoint.hip -- histogram log-likelihood of FSP solutions and its
// gradient with respect to the log rate constants by the adjoint method.
//
// The intended semantics of the reference's cme::Likelihood::evaluate
// (src/Likelihood.cpp:8-33: sum of log max(floor, p) over observations at
// output times) on a joint marginal histogram:
//     LL(theta) = sum_k sum_b N_k[b] log max(P_k[b], floor),
//     P_k = marginal over the observed species of p(t_k),  p(t) = exp(t A) p0,
//     A = sum_r c_r A_r,  c_r = exp(theta_r).
// Gradient (no reference counterpart):
//     dLL/dtheta_r = c_r * integral_0^{t_K} lambda(s)^T A_r p(s) ds,
//     lambda(s) = sum_{k : t_k > s} exp((t_k - s) A^T) g_k,
//     g_k[i] = N_k[b(i)] / P_k[b(i)] where P_k > floor (else 0).
// lambda is integrated backwards with the transposed generator (the fused
// Krylov expv on A^T, assembled in HYB: row j holds the outflow rates of
// state j at its targets' columns), p forwards; the integral uses
// Gauss-Legendre panels graded geometrically towards both ends of every
// observation interval (the integrand has transients of width ~1/|A| there).
// Checkpointing: the forward pass keeps p at the observation times and,
// when all of them fit in half the free device memory, at every quadrature
// node (dense output from inside the accepted Krylov steps, for lambda as
// well); otherwise the backward pass recomputes p at one interval's nodes
// from the checkpoint at the interval's start, with steps clipped at them.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "fsp_assemble.h"
#include "fsp_formats.h"

namespace fsp {
namespace {

constexpr int kGridMax = 4096;

inline int vgrid(int64_t n)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, kGridMax));
}

// marginal bin of state i over the species with keep[s] != 0 (first fastest)
template <int D>
__device__ inline int64_t bin_of(uint32_t i, const DevModel *dm, const int *keep)
{
    int xs[D];
    decode<D>(i, dm, xs);
    int64_t b = 0, mul = 1;
#pragma unroll
    for (int s = 0; s < D; ++s) {
        if (s < dm->d && keep[s]) {
            b += (int64_t)xs[s] * mul;
            mul *= dm->radix[s];
        }
    }
    return b;
}

template <int D>
__global__ void k_adjoint_seed(const DevModel *dm, const int *keep, const double *P,
                               const double *counts, double floor_val, double *g, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int64_t b = bin_of<D>(i, dm, keep);
        const double pb = P[b];
        g[i] = pb > floor_val ? counts[b] / pb : 0.0;
    }
}

template <int D>
void launch_seed(const DevModel *dm, const int *keep, const double *P, const double *counts,
                 double floor_val, double *g, uint32_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_adjoint_seed<D>, dim3(vgrid(n)), dim3(kBlock), 0, s, dm, keep, P, counts,
                       floor_val, g, n);
}

// y = a y + x, with per-workgroup dd partials of |y|^2
__global__ void k_axpby_sumsq(double a, const double *__restrict__ x, double *__restrict__ y,
                              int64_t n, double *__restrict__ partials)
{
    __shared__ double red[2 * kBlock / kWave];
    dd s = dd_zero();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double v = a * y[i] + x[i];
        y[i] = v;
        s = dd_add_prod(s, v, v);
    }
    s = block_sum_dd(s, red);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = s.s;
        partials[2 * blockIdx.x + 1] = s.c;
    }
}

__global__ void k_reduce_one(const double *partials, int G, double *out)
{
    __shared__ double red[2 * kBlock / kWave];
    const dd s = sum_dd_partials(partials, G, red);
    if (threadIdx.x == 0) out[0] = dd_value(s);
}

__global__ void k_scale(double a, double *y, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] *= a;
}

constexpr int kGradChunk = 16;  // reactions per workgroup of k_grad_terms

// Per-workgroup dd partials of lambda^T A_r p for the reactions
// r = 16 blockIdx.y + k, all of them in one pass over p and lambda, in the
// outflow (column) form of A_r p:
//     lambda^T A_r p = sum_j a_r(x_j) p_j (lambda_{j + delta_r} - lambda_j),
// where lambda_{j + delta_r} is 0 when the target of state j leaves the box
// (the outflow stays on the diagonal).  Partials: [r][block][2].
template <int D>
__global__ void k_grad_terms(const DevModel *__restrict__ dm, const double *__restrict__ tab_g,
                             int use_lds, const double *__restrict__ table,
                             const double *__restrict__ p, const double *__restrict__ lam,
                             uint32_t n, double *__restrict__ partials)
{
    extern __shared__ double smem[];
    __shared__ double red[2 * kBlock / kWave];
    const double *tab = tab_g;
    if (tab_g && use_lds) {
        stage_table(smem, tab_g, dm->compact_size);
        tab = smem;
    }
    const int r0 = (int)blockIdx.y * kGradChunk;
    const int R = dm->R;
    dd acc[kGradChunk];
#pragma unroll
    for (int k = 0; k < kGradChunk; ++k) acc[k] = dd_zero();
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        int xs[D];
        decode<D>(i, dm, xs);
        const double pi = p[i], li = lam[i];
#pragma unroll
        for (int k = 0; k < kGradChunk; ++k) {
            const int r = r0 + k;
            if (r < R && !dm->rx[r].zero_nu) {
                const RxnDev &rx = dm->rx[r];
                const double lt = target_leaves<D>(rx, xs, dm) ? 0.0 : lam[(int64_t)i + rx.delta];
                const double a = prop_at<D>(dm, tab, table, r, i, xs, 0);
                acc[k] = dd_add_prod(acc[k], a * pi, lt - li);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kGradChunk; ++k) {
        const int r = r0 + k;
        if (r >= R) break;
        const dd sk = block_sum_dd(acc[k], red);
        if (threadIdx.x == 0) {
            partials[2 * ((size_t)r * gridDim.x + blockIdx.x)] = sk.s;
            partials[2 * ((size_t)r * gridDim.x + blockIdx.x) + 1] = sk.c;
        }
    }
}

// grad[r] += w * (sum of reaction r's partials), one workgroup per reaction
__global__ void k_grad_accum(const double *__restrict__ partials, int G, double w,
                             double *__restrict__ grad)
{
    __shared__ double red[2 * kBlock / kWave];
    const int r = blockIdx.x;
    const dd s = sum_dd_partials(partials + 2 * (size_t)r * G, G, red);
    if (threadIdx.x == 0) grad[r] += w * dd_value(s);
}

template <int D>
void launch_grad_terms(const MatImpl &M, const double *p, const double *lam, int G,
                       double *partials, hipStream_t s)
{
    const DevModel &dm = M.hm.dm;
    const size_t tb = sizeof(double) * M.hm.compact.size();
    const int use_lds = (M.d_compact && !M.hm.compact.empty() && tb <= kLdsTableMax) ? 1 : 0;
    const dim3 grid((unsigned)G, (unsigned)((dm.R + kGradChunk - 1) / kGradChunk));
    hipLaunchKernelGGL(k_grad_terms<D>, grid, dim3(kBlock), use_lds ? tb : 0, s, M.d_model,
                       M.d_compact, use_lds, M.hm.prop_table_dev, p, lam, (uint32_t)M.n,
                       partials);
}

// Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_q).
void gauss_legendre(int q, std::vector<double> &x, std::vector<double> &w)
{
    x.assign(q, 0.0);
    w.assign(q, 0.0);
    for (int i = 0; i < q; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (q + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= q; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = q * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        x[i] = -z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

// nodes (ascending) and weights of the graded composite rule on [a, b]:
// panel breakpoints a + (b-a) {0, 2^-L, ..., 1/2, ..., 1 - 2^-L, 1}
void graded_rule(double a, double b, int q, int L, std::vector<double> &s, std::vector<double> &w)
{
    std::vector<double> gx, gw;
    gauss_legendre(q, gx, gw);
    std::vector<double> br;
    br.push_back(0.0);
    for (int j = L; j >= 1; --j) br.push_back(std::ldexp(1.0, -j));
    for (int j = 2; j <= L; ++j) br.push_back(1.0 - std::ldexp(1.0, -j));
    br.push_back(1.0);
    s.clear();
    w.clear();
    const double len = b - a;
    for (size_t p = 0; p + 1 < br.size(); ++p) {
        const double lo = a + len * br[p], hi = a + len * br[p + 1];
        for (int i = 0; i < q; ++i) {
            s.push_back(0.5 * (lo + hi) + 0.5 * (hi - lo) * gx[i]);
            w.push_back(0.5 * (hi - lo) * gw[i]);
        }
    }
}

struct Work {
    std::vector<double *> bufs;
    double *alloc(int64_t n)
    {
        double *p = nullptr;
        if (hipMalloc(&p, sizeof(double) * (size_t)n) != hipSuccess) {
            (void)hipGetLastError();  // clear the error, so later launch checks do not see it
            return nullptr;
        }
        bufs.push_back(p);
        return p;
    }
    ~Work()
    {
        for (double *p : bufs) (void)hipFree(p);
    }
};

struct LikRun {
    fsp_mat A = nullptr;
    fsp_mat AT = nullptr;  // transposed generator of the backward pass
    MatImpl *M = nullptr;
    fsp_expv E = nullptr;
    fsp_expv Eb = nullptr;
    ~LikRun()
    {
        if (Eb) fsp_expv_destroy(Eb);
        if (E) fsp_expv_destroy(E);
        if (AT) fsp_mat_destroy(AT);
        if (A) fsp_mat_destroy(A);
    }
};

// node_budget: device bytes the stored node vectors may take (< 0: half the
// free device memory at the time of the call).
fsp_status loglik_impl(const fsp_model *model, const double *log_coefs, const double *d_p0,
                       const fsp_lik_data *data, const fsp_grad_options *opt, double *h_ll,
                       double *h_grad, hipStream_t s, double node_budget = -1.0)
{
    if (!model || !log_coefs || !d_p0 || !data || !h_ll) return fail(FSP_ERR_INVALID, "NULL argument");
    if (data->n_times < 1 || !data->times || !data->keep_mask || !data->counts)
        return fail(FSP_ERR_INVALID, "bad likelihood data");
    for (int k = 0; k < data->n_times; ++k)
        if (!(data->times[k] > (k ? data->times[k - 1] : 0.0)))
            return fail(FSP_ERR_INVALID, "observation times must be positive and increasing");
    if (!model->factors) return fail(FSP_ERR_UNSUPPORTED, "the likelihood gradient needs factor tables");
    const int R = model->n_reactions, d = model->n_species;
    fsp_grad_options o;
    o.format = FSP_CSR;
    o.m = 30;
    o.tol = 1e-10;
    o.quad_nodes = 8;
    o.grading_levels = 10;
    if (opt) {
        o.format = opt->format;
        if (opt->m > 0) o.m = opt->m;
        if (opt->tol > 0) o.tol = opt->tol;
        if (opt->quad_nodes > 0) o.quad_nodes = std::min(opt->quad_nodes, 32);
        if (opt->grading_levels > 0) o.grading_levels = std::min(opt->grading_levels, 40);
    }
    std::vector<double> c((size_t)R);
    for (int r = 0; r < R; ++r) c[r] = std::exp(log_coefs[r]);
    LikRun run;
    FSP_TRY(fsp_mat_create(model, o.format, FSP_F64, 0, c.data(), s, &run.A));
    FSP_TRY(fsp_mat_internal(run.A, &run.M));
    MatImpl &M = *run.M;
    const int64_t n = M.n;
    // bins of the observed marginal
    int64_t nb = 1;
    std::vector<int> keep((size_t)d);
    for (int sp = 0; sp < d; ++sp) {
        keep[sp] = data->keep_mask[sp] ? 1 : 0;
        if (keep[sp]) nb *= (int64_t)model->bounds[sp] + 1;
    }
    const int K = data->n_times;
    Work W;
    double *p = W.alloc(n), *tmp = W.alloc(n), *lam = W.alloc(n), *g = W.alloc(n);
    double *P = W.alloc(nb), *cnt = W.alloc(nb * K);
    std::vector<double *> ckpt((size_t)K + 1, nullptr);
    if (!p || !tmp || !lam || !g || !P || !cnt)
        return fail(FSP_ERR_NOMEM, "likelihood workspace");
    int *d_keep = nullptr;
    FSP_HIP(hipMalloc(&d_keep, sizeof(int) * (size_t)d));
    W.bufs.push_back((double *)d_keep);
    FSP_HIP(copy_sync(d_keep, keep.data(), sizeof(int) * (size_t)d, hipMemcpyHostToDevice, s));
    FSP_HIP(copy_sync(cnt, data->counts, sizeof(double) * (size_t)(nb * K), hipMemcpyHostToDevice, s));

    // quadrature nodes of every observation interval (t_{k-1}, t_k]
    const int nodes_per = 2 * o.grading_levels * o.quad_nodes;
    std::vector<std::vector<double>> sn((size_t)K), wn((size_t)K);
    for (int k = 0; k < K; ++k)
        graded_rule(k ? data->times[k - 1] : 0.0, data->times[k], o.quad_nodes, o.grading_levels,
                    sn[k], wn[k]);
    // With the gradient, p at every node is kept from the forward pass when
    // all K * nodes_per vectors fit in the node budget (default: half the
    // free device memory) and their allocation succeeds; else the backward
    // pass recomputes one interval's nodes from the checkpoint at its start
    // (FSP_ADJOINT_RECOMPUTE=1 forces that).  The two paths evaluate the
    // node values differently (dense output inside the accepted steps, or
    // steps clipped at the nodes), so their gradients agree to the Krylov
    // tolerance, not bit for bit.
    double *node_all = nullptr;
    if (h_grad) {
        if (node_budget < 0.0) {
            size_t free_b = 0, total_b = 0;
            FSP_HIP(hipMemGetInfo(&free_b, &total_b));
            node_budget = 0.5 * (double)free_b;
        }
        const double need = 8.0 * (double)K * nodes_per * (double)n;
        const char *env = std::getenv("FSP_ADJOINT_RECOMPUTE");
        const bool force = env && std::atoi(env) != 0;
        if (!force && need <= node_budget) node_all = W.alloc((int64_t)K * nodes_per * n);
    }
    // forward pass: p at the observation times (checkpoints), LL
    FSP_HIP(hipMemcpyAsync(p, d_p0, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
    ckpt[0] = W.alloc(n);
    if (!ckpt[0]) return fail(FSP_ERR_NOMEM, "likelihood checkpoints");
    FSP_HIP(hipMemcpyAsync(ckpt[0], p, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
    FSP_TRY(fsp_expv_create(run.A, nullptr, nullptr, n, data->times[0], o.m, o.tol, 0, 2, 1.0, s, &run.E));
    FSP_TRY(fsp_expv_set_limits(run.E, 1000, 1000000));
    double ll = 0.0;
    std::vector<double *> outs((size_t)nodes_per, nullptr);
    for (int k = 0; k < K; ++k) {
        FSP_TRY(fsp_expv_set_final_time(run.E, data->times[k]));
        if (node_all) {
            // p at the interval's nodes from inside the accepted steps
            for (int q = 0; q < nodes_per; ++q) outs[q] = node_all + ((int64_t)k * nodes_per + q) * n;
            FSP_TRY(fsp_expv_solve_dense(run.E, p, nodes_per, sn[k].data(), outs.data()));
        } else {
            FSP_TRY(fsp_expv_solve(run.E, p));
        }
        if (h_grad) {
            ckpt[k + 1] = W.alloc(n);
            if (!ckpt[k + 1]) return fail(FSP_ERR_NOMEM, "likelihood checkpoints");
            FSP_HIP(hipMemcpyAsync(ckpt[k + 1], p, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
        }
        FSP_TRY(fsp_marginal(p, d, model->bounds, keep.data(), P, s));
        double llk = 0.0;
        FSP_TRY(fsp_loglik_weighted(P, cnt + (int64_t)k * nb, nb, data->floor_val, &llk, s));
        ll += llk;
    }
    *h_ll = ll;
    if (!h_grad) return FSP_OK;

    // backward pass over the intervals (t_{k-1}, t_k], k = K..1: lambda by
    // the fused Krylov engine on A^T stored in HYB; the sensitivity
    // integrals accumulate on the device
    FSP_TRY(fsp_mat_create_transposed(model, c.data(), s, &run.AT));
    FSP_TRY(fsp_expv_create(run.AT, nullptr, nullptr, n, 1.0, o.m, o.tol, 0, 2, 1.0, s, &run.Eb));
    FSP_TRY(fsp_expv_set_limits(run.Eb, 1000, 1000000));
    FSP_HIP(hipMemsetAsync(lam, 0, sizeof(double) * (size_t)n, s));
    const int Gg = vgrid(n);
    const int Rc = (R + kGradChunk - 1) / kGradChunk * kGradChunk;
    double *gpart = W.alloc(2 * (int64_t)Rc * Gg), *d_grad = W.alloc(R + 1);
    if (!gpart || !d_grad) return fail(FSP_ERR_NOMEM, "likelihood gradient workspace");
    FSP_HIP(hipMemsetAsync(d_grad, 0, sizeof(double) * (size_t)(R + 1), s));
    // The Krylov error test is absolute in |v| (src/KryExpvFSP.cu:122-146),
    // and |lambda| reaches N/P ~ 1e6 or more, so lambda is integrated as
    // sig * lam with |lam|_2 = 1 (tol relative to |lambda|).
    double sig = 1.0;
    std::vector<double *> node_p((size_t)nodes_per, nullptr), node_lam((size_t)nodes_per, nullptr);
    std::vector<double> taus((size_t)nodes_per);
    for (int q = 0; q < nodes_per; ++q) {
        // stored nodes: lambda at every node of one interval; else p there
        double *v = W.alloc(n);
        if (!v) return fail(FSP_ERR_NOMEM, "likelihood node vectors");
        (node_all ? node_lam : node_p)[q] = v;
    }
    const DevModel *dm = M.d_model;
    for (int k = K - 1; k >= 0; --k) {
        // lambda(t_k) += g_k, renormalised: lambda = sig * lam with |lam| = 1
        FSP_TRY(fsp_marginal(ckpt[k + 1], d, model->bounds, keep.data(), P, s));
        FSP_SWITCH_D(d, launch_seed<D_>(dm, d_keep, P, cnt + (int64_t)k * nb, data->floor_val, g,
                                        (uint32_t)n, s));
        hipLaunchKernelGGL(k_axpby_sumsq, dim3(Gg), dim3(kBlock), 0, s, sig, g, lam, n, gpart);
        hipLaunchKernelGGL(k_reduce_one, dim3(1), dim3(kBlock), 0, s, gpart, Gg, d_grad + R);
        FSP_HIP(hipGetLastError());
        double ss = 0.0;
        FSP_HIP(hipMemcpyAsync(&ss, d_grad + R, sizeof(double), hipMemcpyDeviceToHost, s));
        FSP_HIP(hipStreamSynchronize(s));
        sig = std::sqrt(ss);
        const double a = k ? data->times[k - 1] : 0.0, b = data->times[k];
        if (!(sig > 0.0)) {
            // lambda vanishes on this interval: no contribution, and it stays 0
            sig = 1.0;
            continue;
        }
        if (!std::isfinite(sig)) return fail(FSP_ERR_NOT_CONVERGED, "adjoint vector is not finite");
        hipLaunchKernelGGL(k_scale, dim3(Gg), dim3(kBlock), 0, s, 1.0 / sig, lam, n);
        const std::vector<double> &snk = sn[k], &wnk = wn[k];
        if (node_all) {
            for (int q = 0; q < nodes_per; ++q) node_p[q] = node_all + ((int64_t)k * nodes_per + q) * n;
        } else {
            // p at the nodes from the checkpoint at a
            FSP_HIP(hipMemcpyAsync(tmp, ckpt[k], sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
            FSP_TRY(fsp_expv_restart(run.E, snk[0] - a));
            for (int q = 0; q < nodes_per; ++q) {
                FSP_TRY(fsp_expv_set_final_time(run.E, snk[q] - a));
                FSP_TRY(fsp_expv_solve(run.E, tmp));
                FSP_HIP(hipMemcpyAsync(node_p[q], tmp, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
            }
        }
        if (node_all) {
            // lambda backwards over the whole interval (tau = b - s), with
            // its values at the nodes from inside the accepted steps
            for (int q = 0; q < nodes_per; ++q) {
                taus[q] = b - snk[nodes_per - 1 - q];
                outs[q] = node_lam[nodes_per - 1 - q];
            }
            FSP_TRY(fsp_expv_restart(run.Eb, b - a));
            FSP_TRY(fsp_expv_solve_dense(run.Eb, lam, nodes_per, taus.data(), outs.data()));
            for (int q = nodes_per - 1; q >= 0; --q) {
                FSP_SWITCH_D(d, launch_grad_terms<D_>(M, node_p[q], node_lam[q], Gg, gpart, s));
                hipLaunchKernelGGL(k_grad_accum, dim3(R), dim3(kBlock), 0, s, gpart, Gg,
                                   wnk[q] * sig, d_grad);
                FSP_HIP(hipGetLastError());
            }
            continue;
        }
        // lambda backwards through the nodes: tau = b - s
        FSP_TRY(fsp_expv_restart(run.Eb, b - snk[nodes_per - 1]));
        for (int q = nodes_per - 1; q >= 0; --q) {
            FSP_TRY(fsp_expv_set_final_time(run.Eb, b - snk[q]));
            FSP_TRY(fsp_expv_solve(run.Eb, lam));
            FSP_SWITCH_D(d, launch_grad_terms<D_>(M, node_p[q], lam, Gg, gpart, s));
            hipLaunchKernelGGL(k_grad_accum, dim3(R), dim3(kBlock), 0, s, gpart, Gg, wnk[q] * sig,
                               d_grad);
            FSP_HIP(hipGetLastError());
        }
        // lambda down to the interval start
        FSP_TRY(fsp_expv_set_final_time(run.Eb, b - a));
        FSP_TRY(fsp_expv_solve(run.Eb, lam));
    }
    std::vector<double> grad((size_t)R, 0.0);
    FSP_HIP(hipMemcpyAsync(grad.data(), d_grad, sizeof(double) * (size_t)R, hipMemcpyDeviceToHost, s));
    FSP_HIP(hipStreamSynchronize(s));
    for (int r = 0; r < R; ++r) h_grad[r] = c[r] * grad[r];
    return FSP_OK;
}

// Batched gradients: points claimed in order from a shared counter by
// `workers` host threads, each with its own stream (a process has 4 hardware
// queues), so one point's host synchronisations and small kernels overlap
// another's row kernels.  Each point runs exactly the single-point code.
struct GradBatch {
    const fsp_model *model;
    int n_points, R;
    const double *log_coefs, *d_p0;
    const fsp_lik_data *data;
    const fsp_grad_options *opt;
    double *h_ll, *h_grad;
    double node_budget = -1.0;  // device bytes per point for stored node vectors
    std::atomic<int> next{0};
    std::atomic<int> failed{0};
};

void grad_worker(GradBatch *B, fsp_status *st, std::string *err)
{
    *st = FSP_OK;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        *st = fail(FSP_ERR_HIP, "batched gradient: stream creation");
        *err = fsp_last_error();
        B->failed = 1;
        return;
    }
    for (;;) {
        if (B->failed.load()) break;
        const int i = B->next.fetch_add(1);
        if (i >= B->n_points) break;
        const fsp_status r = loglik_impl(B->model, B->log_coefs + (size_t)i * B->R, B->d_p0, B->data,
                                         B->opt, B->h_ll + i, B->h_grad + (size_t)i * B->R, s,
                                         B->node_budget);
        if (r != FSP_OK) {
            *st = r;
            *err = "point " + std::to_string(i) + ": " + fsp_last_error();
            B->failed = 1;
            break;
        }
    }
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
}

fsp_status loglik_grad_batch(const fsp_model *model, int n_points, const double *log_coefs,
                             const double *d_p0, const fsp_lik_data *data,
                             const fsp_grad_options *opt, int workers, double *h_ll,
                             double *h_grad)
{
    if (!model || !log_coefs || !d_p0 || !data || !h_ll || !h_grad)
        return fail(FSP_ERR_INVALID, "NULL argument");
    if (n_points < 0) return fail(FSP_ERR_INVALID, "n_points < 0");
    if (n_points == 0) return FSP_OK;
    if (workers <= 0) workers = 4;
    workers = std::min(workers, n_points);
    GradBatch B;
    B.model = model;
    B.n_points = n_points;
    B.R = model->n_reactions;
    B.log_coefs = log_coefs;
    B.d_p0 = d_p0;
    B.data = data;
    B.opt = opt;
    B.h_ll = h_ll;
    B.h_grad = h_grad;
    // Free device memory is measured once, before any worker allocates:
    // each worker's stored node vectors may take free / (2 workers), so the
    // choice between stored and recomputed nodes (and with it the bits of
    // every point's result) does not depend on how the workers interleave.
    {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess)
            return fail(FSP_ERR_HIP, "batched gradient: hipMemGetInfo");
        B.node_budget = (double)free_b / (2.0 * workers);
    }
    std::vector<fsp_status> st((size_t)workers, FSP_OK);
    std::vector<std::string> err((size_t)workers);
    std::vector<std::thread> th;
    for (int w = 1; w < workers; ++w) th.emplace_back(grad_worker, &B, &st[w], &err[w]);
    grad_worker(&B, &st[0], &err[0]);
    for (std::thread &t : th) t.join();
    for (int w = 0; w < workers; ++w)
        if (st[w] != FSP_OK) return fail(st[w], "batched gradient, " + err[w]);
    return FSP_OK;
}

}  // namespace
}  // namespace fsp

extern "C" fsp_status fsp_loglik(const fsp_model *model, const double *log_coefs,
                                 const double *d_p0, const fsp_lik_data *data,
                                 const fsp_grad_options *opt, double *h_ll, void *stream)
{
    return fsp::loglik_impl(model, log_coefs, d_p0, data, opt, h_ll, nullptr, (hipStream_t)stream);
}

extern "C" fsp_status fsp_loglik_grad(const fsp_model *model, const double *log_coefs,
                                      const double *d_p0, const fsp_lik_data *data,
                                      const fsp_grad_options *opt, double *h_ll, double *h_grad,
                                      void *stream)
{
    if (!h_grad) return fsp::fail(FSP_ERR_INVALID, "h_grad is NULL");
    return fsp::loglik_impl(model, log_coefs, d_p0, data, opt, h_ll, h_grad, (hipStream_t)stream);
}

extern "C" fsp_status fsp_loglik_grad_batch(const fsp_model *model, int n_points,
                                            const double *log_coefs, const double *d_p0,
                                            const fsp_lik_data *data,
                                            const fsp_grad_options *opt, int workers,
                                            double *h_ll, double *h_grad)
{
    return fsp::loglik_grad_batch(model, n_points, log_coefs, d_p0, data, opt, workers, h_ll,
                                  h_grad);
}
