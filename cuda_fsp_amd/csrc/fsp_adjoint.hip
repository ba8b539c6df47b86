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
// lambda is integrated backwards with the transposed generator (Krylov expv
// on A^T, matrix free), p forwards; the integral uses Gauss-Legendre panels
// graded geometrically towards both ends of every observation interval (the
// integrand has transients of width ~1/|A| there).  Checkpointing: the
// forward pass keeps p only at the observation times; the backward pass
// recomputes p at one interval's nodes from its checkpoint.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

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

__global__ void k_axpy_inplace(double a, const double *x, double *y, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}

// partial dd sums of x . y per block
__global__ void k_dot_partials(const double *x, const double *y, int64_t n, double *partials)
{
    __shared__ double red[2 * kBlock / kWave];
    dd s = dd_zero();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        s = dd_add_prod(s, x[i], y[i]);
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

struct TransposeCtx {
    MatImpl *M;
    hipStream_t s;
    fsp_status st;
};

void transpose_cb(void *ctx, const double *x, double *y)
{
    auto *c = (TransposeCtx *)ctx;
    const fsp_status st = mat_apply_transpose(*c->M, x, y, 1, c->s);
    if (st != FSP_OK) c->st = st;
}

struct Work {
    std::vector<double *> bufs;
    double *alloc(int64_t n)
    {
        double *p = nullptr;
        if (hipMalloc(&p, sizeof(double) * (size_t)n) != hipSuccess) return nullptr;
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
    MatImpl *M = nullptr;
    fsp_expv E = nullptr;
    fsp_expv Eb = nullptr;
    ~LikRun()
    {
        if (Eb) fsp_expv_destroy(Eb);
        if (E) fsp_expv_destroy(E);
        if (A) fsp_mat_destroy(A);
    }
};

fsp_status dot_dev(const double *x, const double *y, int64_t n, double *partials, double *dout,
                   double *host, hipStream_t s)
{
    hipLaunchKernelGGL(k_dot_partials, dim3(vgrid(n)), dim3(kBlock), 0, s, x, y, n, partials);
    hipLaunchKernelGGL(k_reduce_one, dim3(1), dim3(kBlock), 0, s, partials, vgrid(n), dout);
    FSP_HIP(hipGetLastError());
    FSP_HIP(hipMemcpyAsync(host, dout, sizeof(double), hipMemcpyDeviceToHost, s));
    FSP_HIP(hipStreamSynchronize(s));
    return FSP_OK;
}

fsp_status loglik_impl(const fsp_model *model, const double *log_coefs, const double *d_p0,
                       const fsp_lik_data *data, const fsp_grad_options *opt, double *h_ll,
                       double *h_grad, hipStream_t s)
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
    double *partials = W.alloc(2 * kGridMax + 2), *dout = W.alloc(1);
    std::vector<double *> ckpt((size_t)K + 1, nullptr);
    if (!p || !tmp || !lam || !g || !P || !cnt || !partials || !dout)
        return fail(FSP_ERR_NOMEM, "likelihood workspace");
    int *d_keep = nullptr;
    FSP_HIP(hipMalloc(&d_keep, sizeof(int) * (size_t)d));
    W.bufs.push_back((double *)d_keep);
    FSP_HIP(hipMemcpy(d_keep, keep.data(), sizeof(int) * (size_t)d, hipMemcpyHostToDevice));
    FSP_HIP(hipMemcpy(cnt, data->counts, sizeof(double) * (size_t)(nb * K), hipMemcpyHostToDevice));
    double *h_pin = nullptr;
    FSP_HIP(hipHostMalloc((void **)&h_pin, sizeof(double), hipHostMallocDefault));
    struct PinFree {
        double *p;
        ~PinFree() { (void)hipHostFree(p); }
    } pin_guard{h_pin};

    // forward pass: p at the observation times (checkpoints), LL
    FSP_HIP(hipMemcpyAsync(p, d_p0, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
    ckpt[0] = W.alloc(n);
    if (!ckpt[0]) return fail(FSP_ERR_NOMEM, "likelihood checkpoints");
    FSP_HIP(hipMemcpyAsync(ckpt[0], p, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
    FSP_TRY(fsp_expv_create(run.A, nullptr, nullptr, n, data->times[0], o.m, o.tol, 0, 2, 1.0, s, &run.E));
    FSP_TRY(fsp_expv_set_limits(run.E, 1000, 1000000));
    double ll = 0.0;
    for (int k = 0; k < K; ++k) {
        FSP_TRY(fsp_expv_set_final_time(run.E, data->times[k]));
        FSP_TRY(fsp_expv_solve(run.E, p));
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

    // backward pass over the intervals (t_{k-1}, t_k], k = K..1
    std::vector<double> grad((size_t)R, 0.0);
    TransposeCtx tctx{run.M, s, FSP_OK};
    FSP_TRY(fsp_expv_create(nullptr, &transpose_cb, &tctx, n, 1.0, o.m, o.tol, 0, 2, 1.0, s, &run.Eb));
    FSP_TRY(fsp_expv_set_limits(run.Eb, 1000, 1000000));
    FSP_HIP(hipMemsetAsync(lam, 0, sizeof(double) * (size_t)n, s));
    const int nodes_per = 2 * o.grading_levels * o.quad_nodes;
    std::vector<double *> node_p((size_t)nodes_per, nullptr);
    for (int q = 0; q < nodes_per; ++q) {
        node_p[q] = W.alloc(n);
        if (!node_p[q]) return fail(FSP_ERR_NOMEM, "likelihood node vectors");
    }
    const DevModel *dm = M.d_model;
    for (int k = K - 1; k >= 0; --k) {
        // lambda(t_k) += g_k
        FSP_TRY(fsp_marginal(ckpt[k + 1], d, model->bounds, keep.data(), P, s));
        FSP_SWITCH_D(d, launch_seed<D_>(dm, d_keep, P, cnt + (int64_t)k * nb, data->floor_val, g,
                                        (uint32_t)n, s));
        FSP_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_axpy_inplace, dim3(vgrid(n)), dim3(kBlock), 0, s, 1.0, g, lam, n);
        const double a = k ? data->times[k - 1] : 0.0, b = data->times[k];
        std::vector<double> sn, wn;
        graded_rule(a, b, o.quad_nodes, o.grading_levels, sn, wn);
        // p at the nodes from the checkpoint at a
        FSP_HIP(hipMemcpyAsync(tmp, ckpt[k], sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
        FSP_TRY(fsp_expv_restart(run.E, sn[0] - a));
        for (int q = 0; q < nodes_per; ++q) {
            FSP_TRY(fsp_expv_set_final_time(run.E, sn[q] - a));
            FSP_TRY(fsp_expv_solve(run.E, tmp));
            FSP_HIP(hipMemcpyAsync(node_p[q], tmp, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
        }
        // lambda backwards through the nodes: tau = b - s
        FSP_TRY(fsp_expv_restart(run.Eb, b - sn[nodes_per - 1]));
        for (int q = nodes_per - 1; q >= 0; --q) {
            FSP_TRY(fsp_expv_set_final_time(run.Eb, b - sn[q]));
            FSP_TRY(fsp_expv_solve(run.Eb, lam));
            if (tctx.st != FSP_OK) return tctx.st;
            for (int r = 0; r < R; ++r) {
                FSP_TRY(mat_apply_term(M, r, 1.0, node_p[q], tmp, 0, s));
                double dv = 0.0;
                FSP_TRY(dot_dev(lam, tmp, n, partials, dout, h_pin, s));
                dv = *h_pin;
                grad[r] += wn[q] * dv;
            }
        }
        // lambda down to the interval start
        FSP_TRY(fsp_expv_set_final_time(run.Eb, b - a));
        FSP_TRY(fsp_expv_solve(run.Eb, lam));
        if (tctx.st != FSP_OK) return tctx.st;
    }
    for (int r = 0; r < R; ++r) h_grad[r] = c[r] * grad[r];
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
