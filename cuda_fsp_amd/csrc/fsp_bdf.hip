// fsp_bdf.hip -- time-varying right-hand side and a native BDF integrator.
//
// fsp_rhs_apply is the body of the reference's CVODE callbacks
// (tests/test_sundials.cu:67-91): the coefficients c_r(t) come from a host
// time function (TcoefFun, src/FSPMat.h:16) and udot = A(t) u overwrites the
// output (the reference zero-fills then accumulates).
//
// fsp_bdf_* integrate dp/dt = A(t) p, the role CVODE's BDF method with an
// unpreconditioned Krylov linear solver plays in tests/test_sundials.cu:131-186
// (CVODE itself is third-party and absent, so the method is this file's own):
//   * variable order 1..5, variable step, fully variable-coefficient BDF: the
//     corrector uses the derivative at t_new of the Lagrange interpolant
//     through t_new and the last k accepted points;
//   * the implicit stage (alpha_0 I - A(t_new)) y = b is linear in y and is
//     solved by restarted GMRES on the error-weighted system D M D^-1 (D = 1 /
//     (rtol |y_n| + atol)), right-preconditioned by its diagonal
//     alpha_0 - A_ii (Jacobi), started from the predictor;
//   * error test on the predictor/corrector difference scaled by the exact
//     variable-step ratio r / (1 + r), r = 1 / (alpha_0 (t_new - t_{n-k}));
//   * order selection from divided differences of the accepted history.
// Host control; every vector operation is a device kernel; cross-row sums use
// double-double accumulation, so results do not depend on the launch shape.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fsp_device.h"
#include "fsp_formats.h"

namespace fsp {
namespace {

constexpr int kMaxOrder = 5;
constexpr int kHist = kMaxOrder + 2;   // accepted points kept (order 5 needs 7)
constexpr int kMaxLc = 8;
constexpr int kRedGrid = 1024;

struct LinComb {
    const double *v[kMaxLc];
    double c[kMaxLc];
    int k;
};

__global__ void k_lincomb(LinComb L, double *__restrict__ out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int j = 0; j < L.k; ++j) s += L.c[j] * L.v[j][i];
        out[i] = s;
    }
}

__global__ void k_ewt(const double *__restrict__ y, double *__restrict__ ewt, int64_t n,
                      double rtol, double atol)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        ewt[i] = 1.0 / (rtol * fabs(y[i]) + atol);
}

// partial dd sums of (x_i w_i)^2 (w == nullptr: x_i^2)
__global__ void k_wsq_partials(const double *__restrict__ x, const double *__restrict__ w,
                               int64_t n, double *partials)
{
    __shared__ double red[2 * kBlock / kWave];
    dd s = dd_zero();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double v = w ? x[i] * w[i] : x[i];
        s = dd_add_prod(s, v, v);
    }
    s = block_sum_dd(s, red);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = s.s;
        partials[2 * blockIdx.x + 1] = s.c;
    }
}

constexpr int kDotChunk = 8;  // dots per pass over the rows

// sum_k c_k V_k[i], k ascending, 8 loads issued together
__device__ inline double combo(const double *__restrict__ V, int64_t ldv, int nk,
                               const double *__restrict__ c, int64_t i)
{
    double acc = 0.0;
    int k = 0;
    for (; k + kDotChunk <= nk; k += kDotChunk) {
        double v[kDotChunk];
#pragma unroll
        for (int q = 0; q < kDotChunk; ++q) v[q] = V[(int64_t)(k + q) * ldv + i];
#pragma unroll
        for (int q = 0; q < kDotChunk; ++q) acc += v[q] * c[k + q];
    }
    for (; k < nk; ++k) acc += V[(int64_t)k * ldv + i] * c[k];
    return acc;
}

// GMRES state of the current restart cycle in device memory.  The host
// enqueues iterations without waiting for them; every kernel of an
// iteration returns at once when `done` is set, so iterations launched past
// the end of a cycle do nothing (the operator product itself still runs on
// stale input and its result is not used).
struct GmState {
    int done;          // the cycle ended: converged, broke down or full
    int conv;          // converged (or broke down: the exact solution is in the space)
    int j;             // iterations taken in this cycle
    int m;             // subspace dimension
    double tol, sqn;   // residual test |g_{j+1}| / sqrt(n) <= tol
    double inv_wn;     // 1 / |w| of the newest basis vector (1 for V_0)
    double *H;         // (m + 1) x m upper triangular after the rotations
    double *G;         // (m + 1) x (m + 1) Gram matrix V^T V
    double *cs, *sn;   // rotations
    double *g;         // rotated right-hand side, m + 1
    double *c;         // Gram-Schmidt coefficients of the current column
    double *z;         // back-substitution result, m
};

__global__ void k_gm_init(GmState *gm, double beta, double tol, double sqn)
{
    if (threadIdx.x == 0) {
        gm->tol = tol;
        gm->sqn = sqn;
        gm->done = 0;
        gm->conv = 0;
        gm->j = 0;
        gm->inv_wn = 1.0;
    }
    for (int i = threadIdx.x; i <= gm->m; i += blockDim.x) gm->g[i] = i == 0 ? beta : 0.0;
}

// V_j <- V_j / |w| (the scaling of the vector the previous iteration
// orthogonalised; V_0 is already normalised) and the preconditioned,
// unscaled input of the operator: t = V_j / ((alpha0 - A_ii) ewt).
__global__ void k_gm_unscale(double *__restrict__ vj, const GmState *gm, const double *__restrict__ ewt,
                             const double *__restrict__ diag, double alpha0, double *__restrict__ t,
                             int64_t n)
{
    if (gm->done) return;
    const double s = gm->inv_wn;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double v = vj[i] * s;
        vj[i] = v;
        t[i] = v / ((alpha0 - diag[i]) * ewt[i]);
    }
}

// w = alpha0 x / q - ewt * u (D (alpha0 I - A) D^-1 Q^-1 applied to x,
// u = A t from k_gm_unscale)
__global__ void k_gm_post(const double *__restrict__ x, const double *__restrict__ ewt,
                          const double *__restrict__ diag, const double *__restrict__ u, double alpha0,
                          double *__restrict__ out, int64_t n, const GmState *gm)
{
    if (gm->done) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = alpha0 * (x[i] / (alpha0 - diag[i])) - ewt[i] * u[i];
}

// k_gm_post fused with the operator product for an FP64 HYB generator
// without a sink: u_i = (A t)_i by the HYB row and, in the same pass,
// w_i = alpha0 x_i / (alpha0 - A_ii) - ewt_i u_i with k_gm_post's
// arithmetic (so w has the same bits), u never stored.
template <int NU>
__global__ void __launch_bounds__(kBlock) k_gm_apply_post(HybF<double> A, const double *__restrict__ t,
                                                          const double *__restrict__ x,
                                                          const double *__restrict__ ewt,
                                                          const double *__restrict__ diag, double alpha0,
                                                          double *__restrict__ out, uint32_t tiles,
                                                          const GmState *gm)
{
    if (gm->done) return;
    for (TileWalk w = tile_walk(tiles); w.t < w.end; w.t += w.step) {
        const uint32_t i = w.t * kBlock + threadIdx.x;
        if (i >= A.n) continue;
        dd unused = dd_zero();
        const double u = hyb_row_n<NU>(A, i, RawX{t}, unused);
        out[i] = alpha0 * (x[i] / (alpha0 - diag[i])) - ewt[i] * u;
    }
}

// Launches k_gm_apply_post when the generator allows it (HYB, FP64 values,
// no sink, at most 12 slots); false: the caller takes the two-kernel path.
static bool launch_gm_apply_post(const MatImpl &M, const double *t, const double *x, const double *ewt,
                                 const double *diag, double alpha0, double *out, int64_t n,
                                 const GmState *gm, hipStream_t s)
{
    if (M.format != FSP_HYB || M.prec != FSP_F64 || M.with_sink || M.transposed || n != M.n) return false;
    const HybF<double> A = hyb_view<double>(M);
    const uint32_t tiles = (uint32_t)((M.n + kBlock - 1) / kBlock);
    const int grid = (int)tiles;  // one tile per workgroup, as the plain product
#define FSP_GM_POST_CASE(K)                                                                   \
    case K:                                                                                   \
        hipLaunchKernelGGL(k_gm_apply_post<K>, dim3(grid), dim3(kBlock), 0, s, A, t, x, ewt, diag, \
                           alpha0, out, tiles, gm);                                           \
        return true;
    switch (A.U) {
        FSP_GM_POST_CASE(1) FSP_GM_POST_CASE(2) FSP_GM_POST_CASE(3) FSP_GM_POST_CASE(4)
        FSP_GM_POST_CASE(5) FSP_GM_POST_CASE(6) FSP_GM_POST_CASE(7) FSP_GM_POST_CASE(8)
        FSP_GM_POST_CASE(9) FSP_GM_POST_CASE(10) FSP_GM_POST_CASE(11) FSP_GM_POST_CASE(12)
    default:
        return false;
    }
#undef FSP_GM_POST_CASE
}

// Partial sums over the rows of block b of h_k = V_k . w (slot k, double-
// double) and of the Gram column V_k . V_j (slot nk + k: each thread's sum in
// plain FP64 in a fixed order, as its entries are of rounding size off the
// diagonal and only correct the coefficients at that size), k < nk = j + 1,
// from the same basis loads: 8 dots per pass over the rows, each pass's
// block sums by the transposed LDS reduction (block_sums_dd_t).
__global__ void __launch_bounds__(kBlock) k_gm_dots(const double *__restrict__ V, int64_t ldv, int nk,
                                                    const double *__restrict__ w, int64_t n, double *partials,
                                                    const GmState *gm)
{
    if (gm->done) return;
    __shared__ double buf[2 * kDotChunk * kBlock];
    const double *vj = V + (int64_t)(nk - 1) * ldv;
    for (int c0 = 0; c0 < nk; c0 += kDotChunk) {
        dd s[kDotChunk];
        double q2[kDotChunk];
#pragma unroll
        for (int q = 0; q < kDotChunk; ++q) {
            s[q] = dd_zero();
            q2[q] = 0.0;
        }
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x) {
            const double wi = w[i], xj = vj[i];
            double v[kDotChunk];
#pragma unroll
            for (int q = 0; q < kDotChunk; ++q) v[q] = V[(int64_t)min(c0 + q, nk - 1) * ldv + i];
#pragma unroll
            for (int q = 0; q < kDotChunk; ++q) {
                s[q] = dd_add_prod(s[q], v[q], wi);
                q2[q] = fma(v[q], xj, q2[q]);
            }
        }
        int q;
        dd t;
        if (block_sums_dd_t<kDotChunk>(s, buf, q, t) && c0 + q < nk) {
            const int64_t o = (int64_t)(c0 + q) * gridDim.x + blockIdx.x;
            partials[2 * o] = t.s;
            partials[2 * o + 1] = t.c;
        }
        dd g[kDotChunk];
#pragma unroll
        for (int k = 0; k < kDotChunk; ++k) g[k] = dd{q2[k], 0.0};
        if (block_sums_dd_t<kDotChunk>(g, buf, q, t) && c0 + q < nk) {
            const int64_t og = (int64_t)(nk + c0 + q) * gridDim.x + blockIdx.x;
            partials[2 * og] = t.s;
            partials[2 * og + 1] = t.c;
        }
    }
}

// w -= V c with the Gram-corrected coefficients c = 2 h - G h (one sweep
// doing the work of classical Gram-Schmidt's two passes: the second pass's
// coefficients V^T (w - V h) are h - G h), |w|^2 partials.  Every workgroup
// forms c in LDS from the reduced dots (hv: h then the new Gram column);
// workgroup 0 also stores the Gram column and c for the rotation kernel.
__global__ void k_gm_update(const double *__restrict__ V, int64_t ldv, int nk, const double *__restrict__ hv,
                            double *__restrict__ w, int64_t n, double *partials, GmState *gm)
{
    if (gm->done) return;
    __shared__ double red[2 * kBlock / kWave];
    FSP_DYNAMIC_LDS(cl);  // nk doubles
    const int j = nk - 1, ld = gm->m + 1;
    for (int k = threadIdx.x; k < nk; k += blockDim.x) {
        double acc = 0.0;
        for (int l = 0; l < nk; ++l) {
            const double gkl = l == j ? hv[nk + k] : (k == j ? hv[nk + l] : gm->G[k + (int64_t)l * ld]);
            acc += gkl * hv[l];
        }
        cl[k] = 2.0 * hv[k] - acc;
    }
    __syncthreads();
    if (blockIdx.x == 0) {
        for (int k = threadIdx.x; k < nk; k += blockDim.x) {
            gm->G[k + (int64_t)j * ld] = hv[nk + k];
            gm->G[j + (int64_t)k * ld] = hv[nk + k];
            gm->c[k] = cl[k];
        }
    }
    dd s = dd_zero();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double acc = combo(V, ldv, nk, cl, i);
        const double v = -1.0 * acc + w[i];
        w[i] = v;
        s = dd_add_prod(s, v, v);
    }
    s = block_sum_dd(s, red);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = s.s;
        partials[2 * blockIdx.x + 1] = s.c;
    }
}

// out[k] = value of the dd sum of partials[k][0..G), one workgroup per k
__global__ void k_reduce_partials(const double *__restrict__ partials, int G, double *out)
{
    __shared__ double red[2 * kBlock / kWave];
    const dd s = sum_dd_partials(partials + 2 * (int64_t)blockIdx.x * G, G, red);
    if (threadIdx.x == 0) out[blockIdx.x] = dd_value(s);
}

// Iteration j's Hessenberg column (c_0..c_j, |w|) through the previous
// rotations and a new one, the rotated right-hand side, the residual test
// and the scale of the next basis vector; one workgroup.  |w|^2 is summed
// here from the update kernel's G partials.
__global__ void k_gm_givens(GmState *gm, int j, const double *partials, int G)
{
    if (gm->done) return;
    __shared__ double red[2 * kBlock / kWave];
    const double wn = sqrt(dd_value(sum_dd_partials(partials, G, red)));
    if (threadIdx.x != 0) return;
    const int m = gm->m;
    double *h = gm->H + (int64_t)j * (m + 1);
    for (int i = 0; i <= j; ++i) h[i] = gm->c[i];
    const double hj1 = wn;  // subdiagonal entry: no earlier rotation reaches it
    for (int i = 0; i < j; ++i) {
        const double a = gm->cs[i] * h[i] + gm->sn[i] * h[i + 1];
        h[i + 1] = -gm->sn[i] * h[i] + gm->cs[i] * h[i + 1];
        h[i] = a;
    }
    const double rr = hypot(h[j], hj1);
    const double cj = rr > 0.0 ? h[j] / rr : 1.0;
    const double sj = rr > 0.0 ? hj1 / rr : 0.0;
    gm->cs[j] = cj;
    gm->sn[j] = sj;
    h[j] = rr;
    const double gj = gm->g[j];
    gm->g[j + 1] = -sj * gj;
    gm->g[j] = cj * gj;
    gm->j = j + 1;
    gm->inv_wn = wn > 0.0 ? 1.0 / wn : 0.0;
    const bool conv = fabs(gm->g[j + 1]) / gm->sqn <= gm->tol || wn == 0.0;
    gm->conv = conv ? 1 : 0;
    gm->done = (conv || j + 1 == m) ? 1 : 0;
}

// z = H(0:j, 0:j)^-1 g(0:j) for the cycle's j iterations (back
// substitution, one thread).
__global__ void k_gm_backsub(GmState *gm)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int j = gm->j, ld = gm->m + 1;
    for (int i = j - 1; i >= 0; --i) {
        double a = gm->g[i];
        for (int c = i + 1; c < j; ++c) a -= gm->H[i + (int64_t)c * ld] * gm->z[c];
        gm->z[i] = a / gm->H[i + (int64_t)i * ld];
    }
}

__global__ void k_scale_copy(const double *__restrict__ x, double s, double *__restrict__ y,
                             int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = x[i] * s;
}

// Right preconditioner: Jacobi, q_i = alpha0 - A_ii (> 0, the diagonal of
// the scaled stage operator D (alpha0 I - A) D^-1); k_gm_unscale and
// k_gm_post apply it around the generator product.

// r = ewt * (b - alpha0 y + u), u = A y   (scaled residual)
__global__ void k_resid(const double *__restrict__ b, const double *__restrict__ y,
                        const double *__restrict__ u, const double *__restrict__ ewt,
                        double alpha0, double *__restrict__ r, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        r[i] = ewt[i] * (b[i] - alpha0 * y[i] + u[i]);
}

// y += Q^-1 D^-1 (sum_k g_k V_k)
__global__ void k_sol_update(const double *__restrict__ V, int64_t ldv, int nk,
                             const double *__restrict__ g, const double *__restrict__ ewt,
                             const double *__restrict__ diag, double alpha0,
                             double *__restrict__ y, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double acc = combo(V, ldv, nk, g, i);
        y[i] += acc / ((alpha0 - diag[i]) * ewt[i]);
    }
}

inline int vec_grid(int64_t n)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 4096));
}

// derivative at x[0] of the Lagrange basis through x[0..m]
void lagrange_deriv_at_first(const double *x, int m, double *w)
{
    // l_0'(x0) = sum_{j>0} 1/(x0 - xj); l_j'(x0) = prod_{i != 0, j}(x0 - xi) / prod_{i != j}(xj - xi)
    double s = 0.0;
    for (int j = 1; j <= m; ++j) s += 1.0 / (x[0] - x[j]);
    w[0] = s;
    for (int j = 1; j <= m; ++j) {
        double num = 1.0, den = 1.0;
        for (int i = 0; i <= m; ++i) {
            if (i == j) continue;
            den *= (x[j] - x[i]);
            if (i != 0) num *= (x[0] - x[i]);
        }
        w[j] = num / den;
    }
}

// value at t of the Lagrange interpolant through x[0..m]: weights w[j]
void lagrange_value(const double *x, int m, double t, double *w)
{
    for (int j = 0; j <= m; ++j) {
        double v = 1.0;
        for (int i = 0; i <= m; ++i)
            if (i != j) v *= (t - x[i]) / (x[j] - x[i]);
        w[j] = v;
    }
}

}  // namespace

struct BdfImpl {
    MatImpl *A = nullptr;
    fsp_tcoef_fn coef = nullptr;
    void *ctx = nullptr;
    hipStream_t s = nullptr;
    int64_t n = 0;
    fsp_bdf_options opt;
    // history: Y[j] at T[j], j = 0 newest
    double *Y[kHist] = {};
    double T[kHist] = {};
    int nh = 0;
    bool started = false;
    double t = 0.0, h = 0.0;
    int k = 1, steps_at_k = 0, fails = 0;
    double *ewt = nullptr, *b = nullptr, *yp = nullptr, *ynew = nullptr, *u = nullptr, *tmp = nullptr;
    double *diag = nullptr;         // generator diagonal at the stage time
    double *V = nullptr;            // (kdim + 1) x n scaled Krylov basis
    double *partials = nullptr;     // 2 x 2 (kdim + 1) x kRedGrid (dd)
    double *dred = nullptr;         // 2 (kdim + 1) reduced values
    double *h_red = nullptr;        // pinned host copy
    // GMRES on the device: its state, arrays, and the flags the host reads
    GmState *d_gm = nullptr;
    double *gm_arr = nullptr;       // H, G, cs, sn, g, c, z
    double *gm_z = nullptr;         // z within gm_arr
    int *h_flags = nullptr;         // pinned: (done, conv, j, m) per iteration
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<double> coefs;
    fsp_bdf_stats st;
};

namespace {

fsp_status set_time(BdfImpl &B, double t)
{
    B.coef(t, B.coefs.data(), B.ctx);
    return mat_revalue(*B.A, B.coefs.data(), B.s);
}

fsp_status lincomb(BdfImpl &B, int k, const double *const *v, const double *c, double *out)
{
    LinComb L;
    L.k = k;
    for (int j = 0; j < k; ++j) {
        L.v[j] = v[j];
        L.c[j] = c[j];
    }
    hipLaunchKernelGGL(k_lincomb, dim3(vec_grid(B.n)), dim3(kBlock), 0, B.s, L, out, B.n);
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}

// reduced values of `count` partial rows -> host
fsp_status pull_reduced(BdfImpl &B, int count, double *host)
{
    hipLaunchKernelGGL(k_reduce_partials, dim3(count), dim3(kBlock), 0, B.s, B.partials,
                       kRedGrid, B.dred);
    FSP_HIP(hipGetLastError());
    FSP_HIP(hipMemcpyAsync(B.h_red, B.dred, sizeof(double) * (size_t)count,
                           hipMemcpyDeviceToHost, B.s));
    FSP_HIP(hipStreamSynchronize(B.s));
    std::memcpy(host, B.h_red, sizeof(double) * (size_t)count);
    return FSP_OK;
}

// weighted RMS norm sqrt(mean((x ewt)^2))
fsp_status wrms(BdfImpl &B, const double *x, const double *w, double *out)
{
    hipLaunchKernelGGL(k_wsq_partials, dim3(kRedGrid), dim3(kBlock), 0, B.s, x, w, B.n,
                       B.partials);
    FSP_HIP(hipGetLastError());
    double s2 = 0.0;
    FSP_TRY(pull_reduced(B, 1, &s2));
    *out = std::sqrt(s2 / (double)B.n);
    return FSP_OK;
}

// Solves (alpha0 I - A) y = b, y starting at the predictor (already in y),
// by restarted GMRES on the ewt-scaled system; converged when the weighted
// RMS of the scaled residual is <= tol.  A restart cycle runs on the device:
// per iteration the operator product, one sweep for the dots and the new
// Gram column, one sweep for the Gram-corrected update, and the rotations in
// one workgroup.  The host enqueues iteration it + 1 before it waits for
// iteration it's flags, so the device never idles on the host.
fsp_status gmres(BdfImpl &B, double alpha0, double tol, double *y, bool *converged)
{
    const int m = B.opt.krylov_dim;
    const int64_t n = B.n;
    const int G = vec_grid(n);
    const double sqn = std::sqrt((double)n);
    *converged = false;
    for (int cycle = 0; cycle <= B.opt.max_restarts; ++cycle) {
        // r0 = D (b - alpha0 y + A y)
        FSP_TRY(mat_apply(*B.A, y, B.u, 0, B.s));
        B.st.n_rhs++;
        double *r = B.V;
        hipLaunchKernelGGL(k_resid, dim3(G), dim3(kBlock), 0, B.s, B.b, y, B.u, B.ewt, alpha0, r, n);
        FSP_HIP(hipGetLastError());
        double beta;
        FSP_TRY(wrms(B, r, nullptr, &beta));
        beta *= sqn;  // Euclidean norm
        if (beta / sqn <= tol) {
            *converged = true;
            return FSP_OK;
        }
        hipLaunchKernelGGL(k_scale_copy, dim3(G), dim3(kBlock), 0, B.s, r, 1.0 / beta, r, n);
        hipLaunchKernelGGL(k_gm_init, dim3(1), dim3(kBlock), 0, B.s, B.d_gm, beta, tol, sqn);
        FSP_HIP(hipGetLastError());
        int last = -1;
        for (int j = 0; j < m; ++j) {
            double *vj = B.V + (int64_t)j * n, *w = B.V + (int64_t)(j + 1) * n;
            const int nk = j + 1;
            hipLaunchKernelGGL(k_gm_unscale, dim3(G), dim3(kBlock), 0, B.s, vj, B.d_gm, B.ewt, B.diag,
                               alpha0, B.tmp, n);
            if (!launch_gm_apply_post(*B.A, B.tmp, vj, B.ewt, B.diag, alpha0, w, n, B.d_gm, B.s)) {
                FSP_TRY(mat_apply(*B.A, B.tmp, B.u, 0, B.s));
                hipLaunchKernelGGL(k_gm_post, dim3(G), dim3(kBlock), 0, B.s, vj, B.ewt, B.diag, B.u, alpha0,
                                   w, n, B.d_gm);
            }
            hipLaunchKernelGGL(k_gm_dots, dim3(kRedGrid), dim3(kBlock), 0, B.s, B.V, n, nk, w, n,
                               B.partials, B.d_gm);
            hipLaunchKernelGGL(k_reduce_partials, dim3(2 * nk), dim3(kBlock), 0, B.s, B.partials, kRedGrid,
                               B.dred);
            hipLaunchKernelGGL(k_gm_update, dim3(kRedGrid), dim3(kBlock), sizeof(double) * (size_t)nk, B.s,
                               B.V, n, nk, B.dred, w, n, B.partials, B.d_gm);
            hipLaunchKernelGGL(k_gm_givens, dim3(1), dim3(kBlock), 0, B.s, B.d_gm, j, B.partials, kRedGrid);
            FSP_HIP(hipGetLastError());
            FSP_HIP(hipMemcpyAsync(B.h_flags + 4 * j, B.d_gm, sizeof(int) * 4, hipMemcpyDeviceToHost, B.s));
            FSP_HIP(hipEventRecord(B.ev[j & 1], B.s));
            last = j;
            if (j >= 1) {
                FSP_HIP(hipEventSynchronize(B.ev[(j - 1) & 1]));
                if (B.h_flags[4 * (j - 1)]) break;  // iteration j was launched past the end
            }
        }
        FSP_HIP(hipStreamSynchronize(B.s));
        const int *fl = B.h_flags + 4 * last;
        const int nit = fl[2];
        const bool conv = fl[1] != 0;
        B.st.n_lin_iters += nit;
        B.st.n_rhs += nit;
        // y += Q^-1 D^-1 V z, z = H^-1 g over the cycle's iterations
        hipLaunchKernelGGL(k_gm_backsub, dim3(1), dim3(kWave), 0, B.s, B.d_gm);
        hipLaunchKernelGGL(k_sol_update, dim3(G), dim3(kBlock), 0, B.s, B.V, n, nit,
                           (const double *)B.gm_z, B.ewt, B.diag, alpha0, y, n);
        FSP_HIP(hipGetLastError());
        if (conv) {
            *converged = true;
            return FSP_OK;
        }
    }
    B.st.n_lin_fails++;
    return FSP_OK;
}

fsp_status bdf_alloc(BdfImpl &B)
{
    const size_t n = (size_t)B.n, kd = (size_t)B.opt.krylov_dim + 1, m = (size_t)B.opt.krylov_dim;
    for (int j = 0; j < kHist; ++j) FSP_HIP(hipMalloc(&B.Y[j], sizeof(double) * n));
    double **vs[] = {&B.ewt, &B.b, &B.yp, &B.ynew, &B.u, &B.tmp, &B.diag};
    for (double **p : vs) FSP_HIP(hipMalloc(p, sizeof(double) * n));
    FSP_HIP(hipMalloc(&B.V, sizeof(double) * n * kd));
    FSP_HIP(hipMalloc(&B.partials, sizeof(double) * 4 * kd * kRedGrid));
    FSP_HIP(hipMalloc(&B.dred, sizeof(double) * 2 * kd));
    FSP_HIP(hipHostMalloc((void **)&B.h_red, sizeof(double) * (2 * kd + 1), hipHostMallocDefault));
    // GMRES state: H ((m+1) m), G ((m+1)^2), cs, sn (m), g, c (m+1), z (m)
    const size_t na = kd * m + kd * kd + 2 * m + 2 * kd + m;
    FSP_HIP(hipMalloc(&B.gm_arr, sizeof(double) * na));
    FSP_HIP(fill_sync(B.gm_arr, 0, sizeof(double) * na, B.s));
    FSP_HIP(hipMalloc(&B.d_gm, sizeof(GmState)));
    GmState hs;
    std::memset(&hs, 0, sizeof(hs));
    hs.m = (int)m;
    hs.inv_wn = 1.0;
    double *q = B.gm_arr;
    hs.H = q;
    q += kd * m;
    hs.G = q;
    q += kd * kd;
    hs.cs = q;
    q += m;
    hs.sn = q;
    q += m;
    hs.g = q;
    q += kd;
    hs.c = q;
    q += kd;
    hs.z = q;
    B.gm_z = q;
    FSP_HIP(copy_sync(B.d_gm, &hs, sizeof(GmState), hipMemcpyHostToDevice, B.s));
    FSP_HIP(hipHostMalloc((void **)&B.h_flags, sizeof(int) * 4 * m, hipHostMallocDefault));
    for (hipEvent_t &e : B.ev) FSP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return FSP_OK;
}

void bdf_free(BdfImpl &B)
{
    for (int j = 0; j < kHist; ++j)
        if (B.Y[j]) (void)hipFree(B.Y[j]);
    double *vs[] = {B.ewt, B.b, B.yp, B.ynew, B.u, B.tmp, B.diag, B.V, B.partials, B.dred, B.gm_arr};
    for (double *p : vs)
        if (p) (void)hipFree(p);
    if (B.d_gm) (void)hipFree(B.d_gm);
    if (B.h_red) (void)hipHostFree(B.h_red);
    if (B.h_flags) (void)hipHostFree(B.h_flags);
    for (hipEvent_t e : B.ev)
        if (e) (void)hipEventDestroy(e);
}

// One attempted step of order k and size h from the newest history point.
// On return *accepted tells whether the error test passed; err is the
// weighted error estimate (<= 1 passes).
fsp_status try_step(BdfImpl &B, double h, int k, bool *accepted, double *err)
{
    const double t_new = B.T[0] + h;
    double x[kHist + 1], w[kHist + 1];
    const double *v[kMaxLc];
    // corrector: alpha_0 y + sum_{j=1..k} alpha_j Y[j-1] = A y
    x[0] = t_new;
    for (int j = 1; j <= k; ++j) x[j] = B.T[j - 1];
    lagrange_deriv_at_first(x, k, w);
    const double alpha0 = w[0];
    double c[kMaxLc];
    for (int j = 1; j <= k; ++j) {
        v[j - 1] = B.Y[j - 1];
        c[j - 1] = -w[j];
    }
    FSP_TRY(lincomb(B, k, v, c, B.b));
    // predictor through the newest k+1 points (k points + derivative at startup)
    double ratio;  // predictor error / corrector error factor inverse, r in r/(1+r)
    if (B.nh >= k + 1) {
        for (int j = 0; j <= k; ++j) x[j] = B.T[j];
        lagrange_value(x, k, t_new, w);
        for (int j = 0; j <= k; ++j) {
            v[j] = B.Y[j];
            c[j] = w[j];
        }
        FSP_TRY(lincomb(B, k + 1, v, c, B.yp));
        ratio = 1.0 / (alpha0 * (t_new - B.T[k]));
    } else {
        // first step (k = 1): explicit Euler predictor y0 + h A(t0) y0
        FSP_TRY(set_time(B, B.T[0]));
        FSP_TRY(mat_apply(*B.A, B.Y[0], B.u, 0, B.s));
        B.st.n_rhs++;
        v[0] = B.Y[0];
        v[1] = B.u;
        c[0] = 1.0;
        c[1] = h;
        FSP_TRY(lincomb(B, 2, v, c, B.yp));
        ratio = 1.0;  // Taylor predictor error h^2 y''/2, BDF1 error h^2 y''/2
    }
    // solve the stage from the predictor
    FSP_HIP(hipMemcpyAsync(B.ynew, B.yp, sizeof(double) * (size_t)B.n, hipMemcpyDeviceToDevice, B.s));
    FSP_TRY(set_time(B, t_new));
    FSP_TRY(mat_diagonal(*B.A, B.diag, B.s));
    bool conv = false;
    FSP_TRY(gmres(B, alpha0, B.opt.lin_tol_factor * std::fabs(alpha0), B.ynew, &conv));
    if (!conv) {
        *accepted = false;
        *err = 1e10;
        return FSP_OK;
    }
    // error estimate
    v[0] = B.ynew;
    v[1] = B.yp;
    c[0] = 1.0;
    c[1] = -1.0;
    FSP_TRY(lincomb(B, 2, v, c, B.tmp));
    double dn;
    FSP_TRY(wrms(B, B.tmp, B.ewt, &dn));
    *err = dn * ratio / (1.0 + ratio);
    *accepted = *err <= 1.0;
    return FSP_OK;
}

// weighted norm of the divided difference of order q over (t_new, T[0..q-1])
// times q!: an estimate of |y^(q)|
fsp_status deriv_norm(BdfImpl &B, const double *ynew, double t_new, int q, double *out)
{
    double x[kHist + 1], c[kMaxLc];
    const double *v[kMaxLc];
    x[0] = t_new;
    v[0] = ynew;
    for (int j = 1; j <= q; ++j) {
        x[j] = B.T[j - 1];
        v[j] = B.Y[j - 1];
    }
    double fact = 1.0;
    for (int j = 2; j <= q; ++j) fact *= j;
    for (int j = 0; j <= q; ++j) {
        double den = 1.0;
        for (int i = 0; i <= q; ++i)
            if (i != j) den *= (x[j] - x[i]);
        c[j] = fact / den;
    }
    FSP_TRY(lincomb(B, q + 1, v, c, B.tmp));
    return wrms(B, B.tmp, B.ewt, out);
}

// BDF-q local error for a constant step h: the residual error is
// -h^(q+1) y^(q+1) / (q+1) and the stage divides it by a_0 = sum_{j<=q} 1/j,
// so |LTE| = h^(q+1) |y^(q+1)| / ((q+1) a_0).
double lte_const(int q)
{
    double a0 = 0.0;
    for (int j = 1; j <= q; ++j) a0 += 1.0 / j;
    return 1.0 / ((q + 1) * a0);
}

fsp_status bdf_advance(BdfImpl &B, double *d_v, double t_out)
{
    if (!B.started) {
        FSP_HIP(hipMemcpyAsync(B.Y[0], d_v, sizeof(double) * (size_t)B.n, hipMemcpyDeviceToDevice, B.s));
        B.T[0] = B.t;
        B.nh = 1;
        B.k = 1;
        B.steps_at_k = 0;
        B.started = true;
        // initial step: 0.5 / |A(t0) y0|_w, bounded by the interval
        FSP_HIP(hipStreamSynchronize(B.s));
        hipLaunchKernelGGL(k_ewt, dim3(vec_grid(B.n)), dim3(kBlock), 0, B.s, B.Y[0], B.ewt, B.n,
                           B.opt.rtol, B.opt.atol);
        FSP_TRY(set_time(B, B.t));
        FSP_TRY(mat_apply(*B.A, B.Y[0], B.u, 0, B.s));
        B.st.n_rhs++;
        double fn;
        FSP_TRY(wrms(B, B.u, B.ewt, &fn));
        double span = std::fabs(t_out - B.t);
        B.h = B.opt.h0 > 0.0 ? B.opt.h0 : (fn > 0.0 ? std::min(0.5 / fn, 0.1 * span) : 0.1 * span);
        if (!(B.h > 0.0)) B.h = 1e-6;
    }
    while (B.T[0] < t_out) {
        if (B.st.n_steps >= B.opt.max_steps)
            return fail(FSP_ERR_NOT_CONVERGED, "BDF: max_steps reached");
        double h = B.h;
        const double rem = t_out - B.T[0];
        const bool clipped = h >= rem || rem - h < 0.05 * h;
        if (clipped) h = rem;
        hipLaunchKernelGGL(k_ewt, dim3(vec_grid(B.n)), dim3(kBlock), 0, B.s, B.Y[0], B.ewt, B.n,
                           B.opt.rtol, B.opt.atol);
        // order k needs k+1 accepted points for its predictor (the first
        // step, with one point, uses an explicit Euler predictor)
        const int k = std::max(1, std::min(B.k, B.nh - 1));
        bool acc = false;
        double err = 0.0;
        FSP_TRY(try_step(B, h, k, &acc, &err));
        if (!acc) {
            B.st.n_rejects++;
            B.fails++;
            const double r = err >= 1e10 ? 0.25 : std::max(0.2, 0.9 * std::pow(err, -1.0 / (k + 1)));
            B.h = h * std::min(r, 0.9);
            if (B.fails >= 3) B.k = 1;
            if (B.h < 1e-14 * std::max(1.0, std::fabs(B.T[0])))
                return fail(FSP_ERR_NOT_CONVERGED, "BDF: step size underflow");
            continue;
        }
        B.fails = 0;
        const double t_new = B.T[0] + h;
        // order selection: errors at k-1, k, k+1 from divided differences
        double best_r = std::pow(0.9 / std::max(err, 1e-10), 1.0 / (k + 1));
        int best_q = k;
        B.steps_at_k++;
        if (B.steps_at_k > k) {
            // order k-1 from |y^(k)| over (t_new, T[0..k-1])
            if (k > 1) {
                double dq;
                FSP_TRY(deriv_norm(B, B.ynew, t_new, k, &dq));
                const double e = lte_const(k - 1) * std::pow(h, k) * dq;
                const double r = std::pow(0.9 / std::max(e, 1e-10), 1.0 / k);
                if (r > best_r) {
                    best_r = r;
                    best_q = k - 1;
                }
            }
            // order k+1 from |y^(k+2)| over (t_new, T[0..k+1])
            if (k < B.opt.max_order && B.nh >= k + 2) {
                double dq;
                FSP_TRY(deriv_norm(B, B.ynew, t_new, k + 2, &dq));
                const double e = lte_const(k + 1) * std::pow(h, k + 2) * dq;
                const double r = std::pow(0.9 / std::max(e, 1e-10), 1.0 / (k + 2));
                if (r > 1.1 * best_r) {
                    best_r = r;
                    best_q = k + 1;
                }
            }
        }
        // accept: shift history
        double *old = B.Y[kHist - 1];
        for (int j = kHist - 1; j > 0; --j) {
            B.Y[j] = B.Y[j - 1];
            B.T[j] = B.T[j - 1];
        }
        B.Y[0] = B.ynew;
        B.ynew = old;
        B.T[0] = t_new;
        B.nh = std::min(B.nh + 1, kHist);
        B.st.n_steps++;
        if (best_q != B.k) {
            B.k = best_q;
            B.steps_at_k = 0;
        }
        double r = std::min(2.0, best_r);
        if (r < 1.2 && r > 0.9) r = 1.0;
        const double h_nominal = B.h;
        B.h = h * std::max(0.2, r);
        // a step shortened to land on t_out does not shrink the next one
        if (clipped) B.h = std::max(B.h, std::min(h_nominal, 2.0 * B.h));
    }
    B.t = B.T[0];
    B.st.order = B.k;
    B.st.h_last = B.h;
    B.st.t = B.t;
    FSP_HIP(hipMemcpyAsync(d_v, B.Y[0], sizeof(double) * (size_t)B.n, hipMemcpyDeviceToDevice, B.s));
    FSP_HIP(hipStreamSynchronize(B.s));
    return FSP_OK;
}

}  // namespace
}  // namespace fsp

// ===================================================================== C ABI
using fsp::BdfImpl;
using fsp::MatImpl;

struct fsp_bdf_s {
    BdfImpl impl;
};

extern "C" fsp_status fsp_rhs_apply(fsp_mat A, double t, fsp_tcoef_fn coef, void *ctx,
                                    const double *d_u, double *d_udot, void *stream)
{
    using namespace fsp;
    if (!A || !coef || !d_u || !d_udot) return fail(FSP_ERR_INVALID, "NULL argument");
    MatImpl *M = nullptr;
    FSP_TRY(fsp_mat_internal(A, &M));
    std::vector<double> c((size_t)M->hm.dm.R);
    coef(t, c.data(), ctx);
    FSP_TRY(mat_revalue(*M, c.data(), (hipStream_t)stream));
    return mat_apply(*M, d_u, d_udot, 0, (hipStream_t)stream);
}

extern "C" fsp_status fsp_bdf_create(fsp_mat A, fsp_tcoef_fn coef, void *ctx, double t0,
                                     const fsp_bdf_options *opt, void *stream, fsp_bdf *out)
{
    using namespace fsp;
    if (!out) return fail(FSP_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!A || !coef) return fail(FSP_ERR_INVALID, "NULL argument");
    auto *h = new fsp_bdf_s();
    BdfImpl &B = h->impl;
    fsp_status st = fsp_mat_internal(A, &B.A);
    if (st != FSP_OK) {
        delete h;
        return st;
    }
    B.coef = coef;
    B.ctx = ctx;
    B.s = (hipStream_t)stream;
    B.n = B.A->n_rows;
    B.t = t0;
    fsp_bdf_options d;
    d.rtol = 1e-6;
    d.atol = 1e-12;
    d.max_order = 5;
    d.krylov_dim = 30;
    d.max_restarts = 5;
    d.h0 = 0.0;
    d.max_steps = 500000;
    d.lin_tol_factor = 0.05;
    if (opt) {
        if (opt->rtol > 0) d.rtol = opt->rtol;
        if (opt->atol > 0) d.atol = opt->atol;
        if (opt->max_order > 0) d.max_order = std::min(opt->max_order, kMaxOrder);
        if (opt->krylov_dim > 0) d.krylov_dim = std::min(opt->krylov_dim, 200);
        if (opt->max_restarts >= 0) d.max_restarts = opt->max_restarts;
        if (opt->h0 > 0) d.h0 = opt->h0;
        if (opt->max_steps > 0) d.max_steps = opt->max_steps;
        if (opt->lin_tol_factor > 0) d.lin_tol_factor = opt->lin_tol_factor;
    }
    B.opt = d;
    B.coefs.assign((size_t)B.A->hm.dm.R, 0.0);
    std::memset(&B.st, 0, sizeof(B.st));
    st = bdf_alloc(B);
    if (st != FSP_OK) {
        bdf_free(B);
        delete h;
        return st;
    }
    *out = h;
    return FSP_OK;
}

extern "C" fsp_status fsp_bdf_destroy(fsp_bdf b)
{
    if (!b) return FSP_OK;
    fsp::bdf_free(b->impl);
    delete b;
    return FSP_OK;
}

extern "C" fsp_status fsp_bdf_reset(fsp_bdf b, double t0)
{
    using namespace fsp;
    if (!b) return fail(FSP_ERR_INVALID, "NULL handle");
    b->impl.t = t0;
    b->impl.started = false;
    b->impl.nh = 0;
    return FSP_OK;
}

extern "C" fsp_status fsp_bdf_advance(fsp_bdf b, double *d_v, double t_out)
{
    using namespace fsp;
    if (!b || !d_v) return fail(FSP_ERR_INVALID, "NULL argument");
    BdfImpl &B = b->impl;
    if (t_out < B.t) return fail(FSP_ERR_INVALID, "t_out is before the current time");
    if (B.started && t_out == B.T[0]) {
        FSP_HIP(hipMemcpyAsync(d_v, B.Y[0], sizeof(double) * (size_t)B.n, hipMemcpyDeviceToDevice, B.s));
        FSP_HIP(hipStreamSynchronize(B.s));
        return FSP_OK;
    }
    return bdf_advance(B, d_v, t_out);
}

extern "C" fsp_status fsp_bdf_get_stats(fsp_bdf b, fsp_bdf_stats *st)
{
    using namespace fsp;
    if (!b || !st) return fail(FSP_ERR_INVALID, "NULL argument");
    *st = b->impl.st;
    return FSP_OK;
}
