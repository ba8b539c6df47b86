/*
 * fsp_oracle.c -- CPU restatement of the reference FSP solve path.
 * TEST INFRASTRUCTURE ONLY (see fsp_oracle.h).  Plain C11 + OpenMP.
 */
#include "fsp_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

/* Order-independent sums: unevaluated s + c with error-free transformations
 * (TwoSum; TwoProd through fma), Ogita, Rump & Oishi, SIAM J. Sci. Comput.
 * 26 (2005), "Dot2".  fl(s + c) is as accurate as a twice-working-precision
 * sum, so it agrees with the GPU's tree reduction bit for bit except when the
 * exact sum lies within ~eps^2 of a rounding boundary. */
typedef struct { double s, c; } ddv;

static inline void two_sum(double a, double b, double *s, double *e)
{
    *s = a + b;
    double bb = *s - a;
    *e = (a - (*s - bb)) + (b - bb);
}

static inline ddv dd_add_prod(ddv x, double a, double b)
{
    double p = a * b, pe = fma(a, b, -p), s, e;
    two_sum(x.s, p, &s, &e);
    ddv r = {s, x.c + (e + pe)};
    return r;
}

static inline ddv dd_merge(ddv x, ddv y)
{
    double s, e;
    two_sum(x.s, y.s, &s, &e);
    ddv r = {s, (x.c + y.c) + e};
    return r;
}

static inline double dd_value(ddv x) { return x.s + x.c; }

/* sum_i a[i]*b[i] (b may equal a), threads merged in thread order */
static double dot_dd(int64_t n, const double *a, const double *b)
{
    ddv tot = {0.0, 0.0};
    int nt = g_threads;
    ddv part[256];
    if (nt > 256) nt = 256;
    for (int t = 0; t < nt; ++t) { part[t].s = 0.0; part[t].c = 0.0; }
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
#else
        int t = 0;
#endif
        ddv acc = {0.0, 0.0};
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) acc = dd_add_prod(acc, a[i], b[i]);
        part[t] = acc;
    }
    for (int t = 0; t < nt; ++t) tot = dd_merge(tot, part[t]);
    return dd_value(tot);
}

/* out[k] = V_{k0+k} . w for k = 0..nk-1 (V column-major with leading
 * dimension n): cublasDgemv with CUBLAS_OP_T (src/KryExpvFSP.cu:76-77) as
 * one pass over the rows.  Every dot is an order-independent dd sum, as in
 * dot_dd; each thread keeps 8 dots' accumulators at a time over its rows
 * (independent chains), then the threads' partials merge in thread order. */
#define ORC_DOT_GROUP 8
#define ORC_MAX_DOTS 64
static void dots_dd(int64_t n, const double *V, int k0, int nk, const double *w, double *out)
{
    int nt = g_threads;
    if (nt > 256) nt = 256;
    static ddv part[256][ORC_MAX_DOTS];
    for (int t = 0; t < nt; ++t)
        for (int k = 0; k < nk; ++k) { part[t][k].s = 0.0; part[t][k].c = 0.0; }
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
#else
        int t = 0;
#endif
        int64_t chunk = (n + nt - 1) / nt;
        int64_t i0 = (int64_t)t * chunk, i1 = i0 + chunk < n ? i0 + chunk : n;
        for (int g = 0; g < nk; g += ORC_DOT_GROUP) {
            int gn = nk - g < ORC_DOT_GROUP ? nk - g : ORC_DOT_GROUP;
            ddv acc[ORC_DOT_GROUP];
            const double *col[ORC_DOT_GROUP];
            for (int q = 0; q < ORC_DOT_GROUP; ++q) {
                acc[q].s = 0.0; acc[q].c = 0.0;
                col[q] = V + (int64_t)(k0 + g + (q < gn ? q : 0)) * n;
            }
            if (gn == ORC_DOT_GROUP) {
                for (int64_t i = i0; i < i1; ++i) {
                    const double wi = w[i];
                    for (int q = 0; q < ORC_DOT_GROUP; ++q) acc[q] = dd_add_prod(acc[q], col[q][i], wi);
                }
            } else {
                for (int64_t i = i0; i < i1; ++i) {
                    const double wi = w[i];
                    for (int q = 0; q < gn; ++q) acc[q] = dd_add_prod(acc[q], col[q][i], wi);
                }
            }
            for (int q = 0; q < gn; ++q) part[t][g + q] = acc[q];
        }
    }
    for (int k = 0; k < nk; ++k) {
        ddv tot = {0.0, 0.0};
        for (int t = 0; t < nt; ++t) tot = dd_merge(tot, part[t][k]);
        out[k] = dd_value(tot);
    }
}

static int g_verbose = 0;
/* verbose != 0: the expv drivers print a line per output time to stderr */
void orc_set_verbose(int verbose) { g_verbose = verbose; }

void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int orc_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------ */
/* State space: src/cme_util.cu                                        */
/* ------------------------------------------------------------------ */

int64_t orc_num_states(int d, const int *bounds)
{
    /* src/cme_util.cu:55-62 */
    int64_t n = 1;
    for (int s = 0; s < d; ++s) n *= (int64_t)(bounds[s] + 1);
    return n;
}

void orc_indx2state(int64_t indx, int *state, int d, const int *bounds)
{
    /* src/cme_util.cu:5-11: species 0 is the fastest digit */
    for (int s = 0; s < d; ++s) {
        state[s] = (int)(indx % (bounds[s] + 1));
        indx /= (bounds[s] + 1);
    }
}

int64_t orc_state2indx(const int *state, int d, const int *bounds)
{
    /* src/cme_util.cu:13-22 */
    int64_t indx = 0, nprod = 1;
    for (int s = 0; s < d; ++s) {
        indx += (int64_t)state[s] * nprod;
        nprod *= (bounds[s] + 1);
    }
    return indx;
}

void orc_enumerate(int d, const int *bounds, int *states)
{
    /* src/cme_util.cu:24-41 (AoS int[n][d]) */
    int64_t n = orc_num_states(d, bounds);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t i = 0; i < n; ++i) orc_indx2state(i, &states[i * d], d, bounds);
}

void orc_net_change(const orc_model *m, int r, int *nu)
{
    /* src/cme_util.cu:43-53: rstate[col] += direction * val over the
     * reaction's stoichiometry row. */
    for (int s = 0; s < m->n_species; ++s) nu[s] = 0;
    for (int k = m->stoich_rowptrs[r]; k < m->stoich_rowptrs[r + 1]; ++k)
        nu[m->stoich_cols[k]] += m->stoich_vals[k];
}

static int factor_offset(const orc_model *m, int s)
{
    int off = 0;
    for (int t = 0; t < s; ++t) off += m->bounds[t] + 1;
    return off;
}

static int factor_stride(const orc_model *m)
{
    return factor_offset(m, m->n_species);
}

double orc_prop(const orc_model *m, const int *x, int64_t idx, int r)
{
    if (m->prop_table) {
        int64_t n = orc_num_states(m->n_species, m->bounds);
        return m->prop_table[(int64_t)r * n + idx];
    }
    /* product over species in species order, starting from 1.0, as the
     * reference's Kronecker kernel forms beta (src/fspmat_kron_kernels.cu:
     * 71-113 with the tables of :170-190). */
    const double *f = m->factors + (int64_t)r * factor_stride(m);
    double p = 1.0;
    int off = 0;
    for (int s = 0; s < m->n_species; ++s) {
        p *= f[off + x[s]];
        off += m->bounds[s] + 1;
    }
    return p;
}

static int in_box(const int *x, int d, const int *bounds)
{
    for (int s = 0; s < d; ++s)
        if (x[s] < 0 || x[s] > bounds[s]) return 0;
    return 1;
}

typedef struct {
    int d, R;
    int64_t n;
    int *nu;        /* R x d */
    int *nu_zero;   /* R     */
} model_geom;

static void geom_init(model_geom *g, const orc_model *m)
{
    g->d = m->n_species;
    g->R = m->n_reactions;
    g->n = orc_num_states(g->d, m->bounds);
    g->nu = (int *)malloc(sizeof(int) * (size_t)(g->R * g->d + 1));
    g->nu_zero = (int *)malloc(sizeof(int) * (size_t)(g->R + 1));
    for (int r = 0; r < g->R; ++r) {
        orc_net_change(m, r, &g->nu[r * g->d]);
        int z = 1;
        for (int s = 0; s < g->d; ++s) z = z && (g->nu[r * g->d + s] == 0);
        g->nu_zero[r] = z;
    }
}

static void geom_free(model_geom *g)
{
    free(g->nu);
    free(g->nu_zero);
}

/* ------------------------------------------------------------------ */
/* Per-reaction CSR, reference layout: src/fspmat_csr_kernels.cu       */
/* ------------------------------------------------------------------ */

int64_t orc_term_csr(const orc_model *m, int r, int *row_ptrs, int *cols,
                     double *vals)
{
    model_geom g;
    geom_init(&g, m);
    const int d = g.d;
    const int64_t n = g.n;
    int64_t *off_idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int64_t nnz = 0;
    /* count: fspmat_csr_get_nnz_per_row, :7-51 */
#pragma omp parallel num_threads(g_threads)
    {
        int x[64];
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            orc_indx2state(i, x, d, m->bounds);
            for (int s = 0; s < d; ++s) x[s] -= g.nu[r * d + s];
            off_idx[i] = in_box(x, d, m->bounds) ? orc_state2indx(x, d, m->bounds) : -1;
        }
    }
    /* inclusive scan over the per-row counts (:206-213) */
    if (row_ptrs) row_ptrs[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        nnz += 1 + (off_idx[i] >= 0);
        if (row_ptrs) row_ptrs[i + 1] = (int)nnz;
    }
    if (cols || vals) {
        /* fill: fspmat_csr_fill_data, :53-94.  Columns ascending: the
         * diagonal goes first when the off-diagonal column is larger. */
#pragma omp parallel num_threads(g_threads)
        {
            int x[64], xs[64];
#pragma omp for schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                int64_t rp = row_ptrs[i];
                int64_t j = off_idx[i];
                orc_indx2state(i, x, d, m->bounds);
                double a_here = orc_prop(m, x, i, r);
                if (j >= 0) {
                    int64_t idg, iof;
                    if (j > i) { idg = rp; iof = rp + 1; }
                    else       { idg = rp + 1; iof = rp; }
                    orc_indx2state(j, xs, d, m->bounds);
                    if (vals) {
                        vals[idg] = -1.0 * a_here;
                        vals[iof] = orc_prop(m, xs, j, r);
                    }
                    if (cols) { cols[idg] = (int)i; cols[iof] = (int)j; }
                } else {
                    if (vals) vals[rp] = -1.0 * a_here;
                    if (cols) cols[rp] = (int)i;
                }
            }
        }
    }
    free(off_idx);
    geom_free(&g);
    return nnz;
}

void orc_csr_apply(int64_t n_rows, const int *row_ptrs, const int *cols,
                   const double *vals, double alpha, const double *x,
                   double *y)
{
    /* y = alpha*A*x + 1*y, as cusparseDcsrmv with beta = 1
     * (src/fspmat_csr_kernels.cu:126-132) */
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t i = 0; i < n_rows; ++i) {
        double acc = 0.0;
        for (int64_t k = row_ptrs[i]; k < row_ptrs[i + 1]; ++k)
            acc += vals[k] * x[cols[k]];
        y[i] += alpha * acc;
    }
}

/* ------------------------------------------------------------------ */
/* Merged CSR and ELL (this project's formats)                         */
/* ------------------------------------------------------------------ */

int orc_unique_shifts(const orc_model *m, int *slot_of_reaction)
{
    model_geom g;
    geom_init(&g, m);
    int U = 0;
    int *first = (int *)malloc(sizeof(int) * (size_t)(g.R + 1));
    for (int r = 0; r < g.R; ++r) {
        int slot = -1;
        if (!g.nu_zero[r]) {
            for (int u = 0; u < U && slot < 0; ++u) {
                int same = 1;
                for (int s = 0; s < g.d; ++s)
                    same = same && (g.nu[first[u] * g.d + s] == g.nu[r * g.d + s]);
                if (same) slot = u;
            }
            if (slot < 0) { first[U] = r; slot = U++; }
        }
        if (slot_of_reaction) slot_of_reaction[r] = slot;
    }
    free(first);
    geom_free(&g);
    return U;
}

/* Per-row entries of the merged generator, columns ascending. */
typedef struct { int64_t col; double val; } centry;

static int cmp_centry(const void *a, const void *b)
{
    int64_t ca = ((const centry *)a)->col, cb = ((const centry *)b)->col;
    return (ca > cb) - (ca < cb);
}

/* Build row i of the merged generator into e[]; returns count.  The value of
 * each merged off-diagonal is the sum over the reactions of its slot in
 * reaction order; the diagonal is -(sum_r c_r a_r(x_i)) in reaction order. */
static int merged_row(const orc_model *m, const model_geom *g,
                      const int *slot, int U, const double *coefs,
                      int64_t i, centry *e, int *x, int *xs)
{
    const int d = g->d;
    orc_indx2state(i, x, d, m->bounds);
    double diag = 0.0;
    int cnt = 0;
    for (int u = 0; u < U; ++u) { e[u + 1].col = -1; e[u + 1].val = 0.0; }
    for (int r = 0; r < g->R; ++r) {
        if (g->nu_zero[r]) continue;
        diag += coefs[r] * orc_prop(m, x, i, r);
        for (int s = 0; s < d; ++s) xs[s] = x[s] - g->nu[r * d + s];
        if (in_box(xs, d, m->bounds)) {
            int64_t j = orc_state2indx(xs, d, m->bounds);
            int u = slot[r];
            double v = coefs[r] * orc_prop(m, xs, j, r);
            if (e[u + 1].col < 0) { e[u + 1].col = j; e[u + 1].val = v; }
            else e[u + 1].val += v;
        }
    }
    e[0].col = i;
    e[0].val = -diag;
    cnt = 1;
    for (int u = 0; u < U; ++u)
        if (e[u + 1].col >= 0) e[cnt++] = e[u + 1];
    qsort(e, (size_t)cnt, sizeof(centry), cmp_centry);
    return cnt;
}

static double sink_out_of(const orc_model *m, const model_geom *g,
                          const double *coefs, int64_t i, int *x, int *xs,
                          int *has)
{
    const int d = g->d;
    orc_indx2state(i, x, d, m->bounds);
    double out = 0.0;
    *has = 0;
    for (int r = 0; r < g->R; ++r) {
        if (g->nu_zero[r]) continue;
        for (int s = 0; s < d; ++s) xs[s] = x[s] + g->nu[r * d + s];
        if (!in_box(xs, d, m->bounds)) {
            out += coefs[r] * orc_prop(m, x, i, r);
            *has = 1;
        }
    }
    return out;
}

int64_t orc_merged_csr(const orc_model *m, const double *coefs, int with_sink,
                       int *row_ptrs, int *cols, double *vals)
{
    model_geom g;
    geom_init(&g, m);
    int *slot = (int *)malloc(sizeof(int) * (size_t)(g.R + 1));
    int U = orc_unique_shifts(m, slot);
    const int64_t n = g.n;
    int *cnt = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    centry *scratch_all = NULL;
    (void)scratch_all;
#pragma omp parallel num_threads(g_threads)
    {
        int x[64], xs[64];
        centry *e = (centry *)malloc(sizeof(centry) * (size_t)(U + 2));
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i)
            cnt[i] = merged_row(m, &g, slot, U, coefs, i, e, x, xs);
        free(e);
    }
    int64_t nnz = 0;
    if (row_ptrs) row_ptrs[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        nnz += cnt[i];
        if (row_ptrs) row_ptrs[i + 1] = (int)nnz;
    }
    int64_t nnz_rows = nnz;
    if (with_sink) {
        int x[64], xs[64];
        for (int64_t i = 0; i < n; ++i) {
            int has;
            double o = sink_out_of(m, &g, coefs, i, x, xs, &has);
            if (has) {
                if (cols) cols[nnz] = (int)i;
                if (vals) vals[nnz] = o;
                ++nnz;
            }
        }
        if (row_ptrs) row_ptrs[n + 1] = (int)nnz;
    }
    if ((cols || vals) && row_ptrs) {
#pragma omp parallel num_threads(g_threads)
        {
            int x[64], xs[64];
            centry *e = (centry *)malloc(sizeof(centry) * (size_t)(U + 2));
#pragma omp for schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                int c = merged_row(m, &g, slot, U, coefs, i, e, x, xs);
                int64_t rp = row_ptrs[i];
                for (int k = 0; k < c; ++k) {
                    if (cols) cols[rp + k] = (int)e[k].col;
                    if (vals) vals[rp + k] = e[k].val;
                }
            }
            free(e);
        }
    }
    (void)nnz_rows;
    free(cnt);
    free(slot);
    geom_free(&g);
    return nnz;
}

void orc_ell(const orc_model *m, const double *coefs, double *diag,
             double *slot_vals, int *slot_cols, double *sink_out)
{
    model_geom g;
    geom_init(&g, m);
    int *slot = (int *)malloc(sizeof(int) * (size_t)(g.R + 1));
    int U = orc_unique_shifts(m, slot);
    const int64_t n = g.n;
    const int d = g.d;
#pragma omp parallel num_threads(g_threads)
    {
        int x[64], xs[64];
        int *seen = (int *)malloc(sizeof(int) * (size_t)(U + 1));
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            orc_indx2state(i, x, d, m->bounds);
            double dg = 0.0;
            for (int u = 0; u < U; ++u) {
                seen[u] = 0;
                slot_vals[(int64_t)u * n + i] = 0.0;
                slot_cols[(int64_t)u * n + i] = (int)i;
            }
            for (int r = 0; r < g.R; ++r) {
                if (g.nu_zero[r]) continue;
                dg += coefs[r] * orc_prop(m, x, i, r);
                for (int s = 0; s < d; ++s) xs[s] = x[s] - g.nu[r * d + s];
                if (in_box(xs, d, m->bounds)) {
                    int64_t j = orc_state2indx(xs, d, m->bounds);
                    int u = slot[r];
                    double v = coefs[r] * orc_prop(m, xs, j, r);
                    if (!seen[u]) { slot_vals[(int64_t)u * n + i] = v; seen[u] = 1; }
                    else slot_vals[(int64_t)u * n + i] += v;
                    slot_cols[(int64_t)u * n + i] = (int)j;
                }
            }
            diag[i] = -dg;
            if (sink_out) {
                int has;
                sink_out[i] = sink_out_of(m, &g, coefs, i, x, xs, &has);
            }
        }
        free(seen);
    }
    free(slot);
    geom_free(&g);
}

void orc_ell_apply(int64_t n, int U, const double *diag,
                   const double *slot_vals, const int *slot_cols,
                   const double *sink_out, const double *x, double *y)
{
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double acc = diag[i] * x[i];
        /* a zero slot (an out-of-box neighbour, or a zero rate) adds +0, as
         * the device's HYB row does, whatever x holds at its column */
        for (int u = 0; u < U; ++u) {
            const double v = slot_vals[(int64_t)u * n + i];
            acc += v != 0.0 ? v * x[slot_cols[(int64_t)u * n + i]] : 0.0;
        }
        y[i] += acc;
    }
    if (sink_out) y[n] += dot_dd(n, sink_out, x);
}

/* ------------------------------------------------------------------ */
/* Reference HYB layout: src/fspmat_hyb_kernels.cu                     */
/* ------------------------------------------------------------------ */

void orc_hyb_ref(const orc_model *m, double *diag, double *off, int *col)
{
    model_geom g;
    geom_init(&g, m);
    const int64_t n = g.n;
    const int d = g.d;
    /* fspmat_hyb_fill_data, :7-64 */
#pragma omp parallel num_threads(g_threads)
    {
        int x[64], xs[64];
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            orc_indx2state(i, x, d, m->bounds);
            for (int r = 0; r < g.R; ++r) {
                diag[(int64_t)r * n + i] = -1.0 * orc_prop(m, x, i, r);
                for (int s = 0; s < d; ++s) xs[s] = x[s] - g.nu[r * d + s];
                if (in_box(xs, d, m->bounds)) {
                    int64_t j = orc_state2indx(xs, d, m->bounds);
                    col[(int64_t)r * n + i] = (int)j;
                    off[(int64_t)r * n + i] = orc_prop(m, xs, j, r);
                } else {
                    col[(int64_t)r * n + i] = 0;
                    off[(int64_t)r * n + i] = 0.0;
                }
            }
        }
    }
    geom_free(&g);
}

void orc_hyb_ref_apply(int64_t n, int R, const double *diag,
                       const double *off, const int *col,
                       const double *coefs, const double *x, double *y)
{
    /* fspmat_hyb_mv, :66-80 */
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double yv = 0.0;
        for (int r = 0; r < R; ++r) {
            yv += coefs[r] * diag[(int64_t)r * n + i] * x[i];
            yv += coefs[r] * off[(int64_t)r * n + i] * x[col[(int64_t)r * n + i]];
        }
        y[i] += yv;
    }
}

/* ------------------------------------------------------------------ */
/* Matrix-free Kronecker-sum apply: src/fspmat_kron_kernels.cu:51-129  */
/* ------------------------------------------------------------------ */

void orc_kron_apply(const orc_model *m, const double *coefs, int with_sink,
                    const double *x, double *y)
{
    model_geom g;
    geom_init(&g, m);
    const int64_t n = g.n;
    const int d = g.d;
    const int R = g.R;
    const int stride = factor_stride(m);
    double *outv = with_sink ? (double *)malloc(sizeof(double) * (size_t)n) : NULL;
#pragma omp parallel num_threads(g_threads)
    {
        int xi[64];
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            orc_indx2state(i, xi, d, m->bounds);
            double yv = 0.0, out = 0.0;
            const double xh = x[i];
            for (int r = 0; r < R; ++r) {
                const double *f = m->factors + (int64_t)r * stride;
                double alpha = 1.0, beta = 1.0;
                int64_t ix = i, nl = 1;
                int off = 0, leaves = 0;
                for (int s = 0; s < d; ++s) {
                    int nh = m->bounds[s] + 1;
                    int sh = -g.nu[r * d + s];
                    int ih = xi[s];
                    if (ih + sh >= 0 && ih + sh < nh) {
                        alpha *= f[off + ih + sh];
                        ix += nl * sh;
                    } else {
                        alpha = 0.0;
                    }
                    beta *= f[off + ih];
                    if (ih - sh < 0 || ih - sh >= nh) leaves = 1;
                    nl *= nh;
                    off += nh;
                }
                if (g.nu_zero[r]) leaves = 0;
                yv += coefs[r] * (x[ix] * alpha - xh * beta);
                if (leaves) out += coefs[r] * beta;
            }
            y[i] += yv;
            if (outv) outv[i] = out;
        }
    }
    if (with_sink) {
        y[n] += dot_dd(n, outv, x);
        free(outv);
    }
    geom_free(&g);
}

/* The same operator grouped by net change: this project's KRON slot form
 * (the reference sums per reaction, orc_kron_apply above; the two agree to
 * rounding).  Slot u is a unique net change (first appearance order) with
 * its member reactions in reaction order; when every member's propensity
 * varies with one species s_u at most (factor tables not identically 1),
 *   h_u(k) = sum_{r in slot u} c_r f_{r,s_u}(k)      (from 0.0, member order)
 *   y_i += acc - dsum x_i,  acc = sum_u [x_i - nu_u in box] h_u(x_{s_u} -
 *   nu_{u,s_u}) x_{i - delta_u},  dsum = sum_u h_u(x_{s_u})   (slot order),
 * and the sink collects sum_u [x_i + nu_u leaves the box] h_u(x_{s_u}).
 * orc_kron_slot_form returns the number of table entries, or -1 when the
 * form does not apply (a slot whose members vary with different species,
 * more than 24 slots, or tables over 96 KiB), as on the device. */
typedef struct {
    int U;
    int *slot;       /* R: slot of each reaction, -1 for zero net change */
    int *sp;         /* U: table species, -1 constant */
    int *toff;       /* U + 1 */
    int *first;      /* U: first member (its net change is the slot's) */
} kron_slots;

static int factor_trivial(const orc_model *m, int r, int s)
{
    int stride = factor_stride(m), off = 0;
    for (int t = 0; t < s; ++t) off += m->bounds[t] + 1;
    const double *f = m->factors + (int64_t)r * stride + off;
    for (int k = 0; k <= m->bounds[s]; ++k)
        if (f[k] != 1.0) return 0;
    return 1;
}

static int slots_init(kron_slots *ks, const orc_model *m, const model_geom *g)
{
    ks->slot = (int *)malloc(sizeof(int) * (size_t)(g->R + 1));
    ks->U = orc_unique_shifts(m, ks->slot);
    ks->sp = (int *)malloc(sizeof(int) * (size_t)(ks->U + 1));
    ks->toff = (int *)malloc(sizeof(int) * (size_t)(ks->U + 2));
    ks->first = (int *)malloc(sizeof(int) * (size_t)(ks->U + 1));
    if (!m->factors || ks->U > 24) return -1;
    int off = 0;
    for (int u = 0; u < ks->U; ++u) {
        ks->sp[u] = -1;
        ks->first[u] = -1;
        for (int r = 0; r < g->R; ++r) {
            if (ks->slot[r] != u) continue;
            if (ks->first[u] < 0) ks->first[u] = r;
            for (int s = 0; s < g->d; ++s) {
                if (factor_trivial(m, r, s)) continue;
                if (ks->sp[u] >= 0 && ks->sp[u] != s) return -1;
                ks->sp[u] = s;
            }
        }
        ks->toff[u] = off;
        off += ks->sp[u] >= 0 ? m->bounds[ks->sp[u]] + 1 : 1;
    }
    ks->toff[ks->U] = off;
    if ((int64_t)off * 8 > 96 * 1024) return -1;
    return off;
}

static void slots_free(kron_slots *ks)
{
    free(ks->slot); free(ks->sp); free(ks->toff); free(ks->first);
}

int orc_kron_slot_form(const orc_model *m)
{
    model_geom g;
    geom_init(&g, m);
    kron_slots ks;
    int rc = slots_init(&ks, m, &g);
    slots_free(&ks);
    geom_free(&g);
    return rc;
}

int orc_kron_slot_apply(const orc_model *m, const double *coefs, int with_sink,
                        const double *x, double *y)
{
    model_geom g;
    geom_init(&g, m);
    kron_slots ks;
    const int size = slots_init(&ks, m, &g);
    if (size < 0) {
        slots_free(&ks);
        geom_free(&g);
        return -1;
    }
    const int d = g.d, U = ks.U, stride = factor_stride(m);
    const int64_t n = g.n;
    /* the tables, member by member in reaction order */
    double *tab = (double *)calloc((size_t)size + 1, sizeof(double));
    for (int u = 0; u < U; ++u) {
        const int s = ks.sp[u];
        const int len = s >= 0 ? m->bounds[s] + 1 : 1;
        int off = 0;
        for (int t = 0; t < s; ++t) off += m->bounds[t] + 1;
        for (int r = 0; r < g.R; ++r) {
            if (ks.slot[r] != u) continue;
            const int triv = s < 0 || factor_trivial(m, r, s);
            const double *f = m->factors + (int64_t)r * stride + off;
            for (int k = 0; k < len; ++k) tab[ks.toff[u] + k] += coefs[r] * (triv ? 1.0 : f[k]);
        }
    }
    /* slot shifts */
    int64_t *delta = (int64_t *)malloc(sizeof(int64_t) * (size_t)(U + 1));
    for (int u = 0; u < U; ++u) {
        const int *nu = &g.nu[ks.first[u] * d];
        int64_t dl = 0, mul = 1;
        for (int s = 0; s < d; ++s) { dl += (int64_t)nu[s] * mul; mul *= m->bounds[s] + 1; }
        delta[u] = dl;
    }
    double *outv = with_sink ? (double *)malloc(sizeof(double) * (size_t)n) : NULL;
#pragma omp parallel num_threads(g_threads)
    {
        int xi[64];
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            orc_indx2state(i, xi, d, m->bounds);
            double acc = 0.0, dsum = 0.0, out = 0.0;
            for (int u = 0; u < U; ++u) {
                const int *nu = &g.nu[ks.first[u] * d];
                int in = 1, leaves = 0;
                for (int s = 0; s < d; ++s) {
                    if (nu[s] == 0) continue;
                    const int src = xi[s] - nu[s], tgt = xi[s] + nu[s];
                    if (src < 0 || src > m->bounds[s]) in = 0;
                    if (tgt < 0 || tgt > m->bounds[s]) leaves = 1;
                }
                const int s = ks.sp[u];
                const double hx = tab[ks.toff[u] + (s >= 0 ? xi[s] : 0)];
                if (in) {
                    const double hs = tab[ks.toff[u] + (s >= 0 ? xi[s] - nu[s] : 0)];
                    acc += hs * x[i - delta[u]];
                }
                dsum += hx;
                if (leaves) out += hx;
            }
            y[i] += acc - dsum * x[i];
            if (outv) outv[i] = out;
        }
    }
    if (with_sink) {
        y[n] += dot_dd(n, outv, x);
        free(outv);
    }
    free(delta);
    free(tab);
    slots_free(&ks);
    geom_free(&g);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Dense expm: Golub & Van Loan Algorithm 11.3.1 (Pade q = 6)           */
/* ------------------------------------------------------------------ */

static void matmul_cm(int N, const double *A, const double *B, double *C)
{
    /* C = A*B, column-major; each C(i,j) sums k in increasing order */
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            double acc = 0.0;
            for (int k = 0; k < N; ++k) acc += A[i + k * N] * B[k + j * N];
            C[i + j * N] = acc;
        }
}

/* Solve D X = B (column-major N x N, N right-hand sides) by Gaussian
 * elimination with partial pivoting; D and B are overwritten. */
static int gesv_cm(int N, double *D, double *B)
{
    for (int k = 0; k < N; ++k) {
        int p = k;
        double best = fabs(D[k + k * N]);
        for (int i = k + 1; i < N; ++i)
            if (fabs(D[i + k * N]) > best) { best = fabs(D[i + k * N]); p = i; }
        if (best == 0.0) return -1;
        if (p != k) {
            for (int j = 0; j < N; ++j) {
                double t = D[k + j * N]; D[k + j * N] = D[p + j * N]; D[p + j * N] = t;
                t = B[k + j * N]; B[k + j * N] = B[p + j * N]; B[p + j * N] = t;
            }
        }
        double piv = D[k + k * N];
        for (int i = k + 1; i < N; ++i) {
            double l = D[i + k * N] / piv;
            D[i + k * N] = l;
            for (int j = k + 1; j < N; ++j) D[i + j * N] -= l * D[k + j * N];
            for (int j = 0; j < N; ++j) B[i + j * N] -= l * B[k + j * N];
        }
    }
    for (int j = 0; j < N; ++j)
        for (int i = N - 1; i >= 0; --i) {
            double acc = B[i + j * N];
            for (int k = i + 1; k < N; ++k) acc -= D[i + k * N] * B[k + j * N];
            B[i + j * N] = acc / D[i + i * N];
        }
    return 0;
}

int orc_expm(int N, const double *A, double *F)
{
    const int q = 6;
    size_t sz = sizeof(double) * (size_t)N * (size_t)N;
    double *As = (double *)malloc(sz), *X = (double *)malloc(sz),
           *T = (double *)malloc(sz), *Nm = (double *)malloc(sz),
           *D = (double *)malloc(sz);
    /* infinity norm (max row sum) */
    double nrm = 0.0;
    for (int i = 0; i < N; ++i) {
        double rs = 0.0;
        for (int j = 0; j < N; ++j) rs += fabs(A[i + j * N]);
        if (rs > nrm) nrm = rs;
    }
    /* j = max(0, 1 + floor(log2 ||A||_inf)) */
    int js = 0;
    if (nrm > 0.0) {
        int e;
        frexp(nrm, &e);           /* nrm = f * 2^e, f in [0.5, 1) */
        js = e;                   /* = 1 + floor(log2 nrm) */
        if (js < 0) js = 0;
    }
    double scale = ldexp(1.0, -js);
    for (int k = 0; k < N * N; ++k) As[k] = A[k] * scale;
    for (int k = 0; k < N * N; ++k) { X[k] = 0.0; Nm[k] = 0.0; D[k] = 0.0; }
    for (int i = 0; i < N; ++i) { X[i + i * N] = 1.0; Nm[i + i * N] = 1.0; D[i + i * N] = 1.0; }
    double c = 1.0;
    for (int k = 1; k <= q; ++k) {
        c = c * (double)(q - k + 1) / (double)((2 * q - k + 1) * k);
        matmul_cm(N, As, X, T);
        memcpy(X, T, sz);
        double sgn = (k % 2) ? -1.0 : 1.0;
        for (int t = 0; t < N * N; ++t) {
            Nm[t] += c * X[t];
            D[t] += sgn * c * X[t];
        }
    }
    int rc = gesv_cm(N, D, Nm);
    for (int k = 0; k < js; ++k) {
        matmul_cm(N, Nm, Nm, T);
        memcpy(Nm, T, sz);
    }
    memcpy(F, Nm, sz);
    free(As); free(X); free(T); free(Nm); free(D);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Krylov expv: src/KryExpvFSP.cu:38-188                              */
/* ------------------------------------------------------------------ */

struct orc_expv {
    int64_t n;
    orc_matvec_fn mv;
    void *ctx;
    double t_final, t_now, t_new;
    double tol, anorm, btol, xm, err_loc, beta, avnorm;
    int m, iop, q_iop, mb, k1, max_reject;
    int64_t i_step, n_matvec, n_reject, n_breakdown;
    double tau_last;
    double *wsp;     /* V: n*(m+2) */
    double *H, *F;   /* (m+2)^2 column-major */
    const double *replay;
    int replay_count, replay_pos;
    double *tau_log;  /* accepted step sizes, optional */
    int64_t tau_cap;
};

/* cublasDdot / cublasDnrm2 (src/KryExpvFSP.cu:41,76,86,117) as
 * order-independent sums */


static double nrm2_n(int64_t n, const double *a) { return sqrt(dot_dd(n, a, a)); }

orc_expv *orc_expv_create(int64_t n, double t_final, orc_matvec_fn mv,
                          void *ctx, int m, double tol, int iop, int q_iop,
                          double anorm)
{
    /* constructor: src/KryExpvFSP.cu:6-36; knobs src/KryExpvFSP.h:64-69 */
    if (m < 1 || m + 1 > ORC_MAX_DOTS) return NULL;
    orc_expv *e = (orc_expv *)calloc(1, sizeof(orc_expv));
    e->n = n; e->mv = mv; e->ctx = ctx;
    e->t_final = t_final; e->t_now = 0.0; e->t_new = 0.0;
    e->m = m; e->tol = tol; e->iop = iop; e->q_iop = q_iop; e->anorm = anorm;
    e->mb = m; e->k1 = 2; e->max_reject = 1000;
    e->wsp = (double *)malloc(sizeof(double) * (size_t)n * (size_t)(m + 2));
    e->H = (double *)malloc(sizeof(double) * (size_t)(m + 2) * (size_t)(m + 2));
    e->F = (double *)malloc(sizeof(double) * (size_t)(m + 2) * (size_t)(m + 2));
    return e;
}

void orc_expv_destroy(orc_expv *e)
{
    if (!e) return;
    free(e->wsp); free(e->H); free(e->F); free(e);
}

void orc_expv_reset_time(orc_expv *e) { e->t_now = 0.0; }
void orc_expv_set_final_time(orc_expv *e, double t_final) { e->t_final = t_final; }
int orc_expv_final_time_reached(const orc_expv *e) { return e->t_now >= e->t_final; }
void orc_expv_set_max_reject(orc_expv *e, int mr) { e->max_reject = mr; }
void orc_expv_set_replay(orc_expv *e, const double *taus, int count)
{
    e->replay = taus; e->replay_count = count; e->replay_pos = 0;
}

void orc_expv_set_tau_log(orc_expv *e, double *taus, int64_t cap)
{
    e->tau_log = taus; e->tau_cap = cap;
}

void orc_expv_get_stats(const orc_expv *e, orc_expv_stats *s)
{
    s->t_now = e->t_now; s->t_new = e->t_new; s->tau_last = e->tau_last;
    s->err_last = e->err_loc; s->beta_last = e->beta;
    s->i_step = e->i_step; s->n_matvec = e->n_matvec;
    s->n_reject = e->n_reject; s->n_breakdown = e->n_breakdown;
}

/* Decimal powers 1e-300 .. 1e300 as correctly rounded literals. */
#define P10_LO (-300)
#define P10_HI 300
static double g_p10[P10_HI - P10_LO + 1];
static int g_p10_init = 0;

static void p10_init(void)
{
    if (g_p10_init) return;
    /* strtod of "1e<k>" is correctly rounded (glibc) */
    char buf[16];
    for (int k = P10_LO; k <= P10_HI; ++k) {
        snprintf(buf, sizeof buf, "1e%d", k);
        g_p10[k - P10_LO] = strtod(buf, NULL);
    }
    g_p10_init = 1;
}

double orc_round_up_2sig(double t)
{
    /* src/KryExpvFSP.cu:148-150 and :172-174: s = 10^(floor(log10 t) - 1),
     * t = ceil(t / s) * s.  The decade exponent is found by comparison with
     * correctly rounded decimal powers rather than by log10/pow, so the CPU
     * and GPU (whose libm may differ by an ulp) round identically. */
    p10_init();
    if (!(t > 0.0) || !isfinite(t)) return t;
    int e = P10_LO;
    int lo = P10_LO, hi = P10_HI;
    while (lo <= hi) { /* largest e with 10^e <= t */
        int mid = (lo + hi) / 2;
        if (g_p10[mid - P10_LO] <= t) { e = mid; lo = mid + 1; }
        else hi = mid - 1;
    }
    if (e - 1 < P10_LO) return t;
    double s = g_p10[e - 1 - P10_LO];
    return ceil(t / s) * s;
}

static double round_up_2sig(double t) { return orc_round_up_2sig(t); }

int orc_expv_step(orc_expv *e, double *v)
{
    const int64_t n = e->n;
    const int m = e->m;
    const int M2 = m + 2;
    double *V = e->wsp;
    double *H = e->H;
    int rc = 0;

    e->beta = nrm2_n(n, v);                                   /* :41 */
    if (e->i_step == 0) {                                     /* :44-51 */
        e->xm = 1.0 / (double)m;
        double fact = pow((m + 1) / exp(1.0), m + 1) * sqrt(2 * (3.1416) * (m + 1));
        e->t_new = (1.0 / e->anorm) * pow((fact * e->tol) / (4.0 * e->beta * e->anorm), e->xm);
        e->btol = e->anorm * e->tol;
    }
    e->mb = m;                                                /* :52 */
    double tau = fmin(e->t_final - e->t_now, e->t_new);       /* :54 */
    if (e->replay && e->replay_pos < e->replay_count) tau = e->replay[e->replay_pos++];

    memset(V, 0, sizeof(double) * (size_t)n * (size_t)M2);   /* :56 */
    memset(H, 0, sizeof(double) * (size_t)M2 * (size_t)M2);
    double betainv = 1.0 / e->beta;                           /* :58-65 */
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t i = 0; i < n; ++i) V[i] = v[i] * betainv;

    int istart = 0;
    for (int j = 0; j < m; ++j) {                             /* :70-109 */
        double *Vj1 = V + (int64_t)(j + 1) * n;
        e->mv(e->ctx, V + (int64_t)j * n, Vj1);
        e->n_matvec++;
        if (e->iop) istart = (j >= e->q_iop - 1) ? j - e->q_iop + 1 : 0;
        {                                                     /* gemv T, :76-77 */
            double hk[ORC_MAX_DOTS];
            dots_dd(n, V, istart, j - istart + 1, Vj1, hk);
            for (int k = istart; k <= j; ++k) H[k + j * M2] += hk[k - istart];
        }
        /* gemv N, :79-81: Vj1 -= V[istart..j] * H(istart..j, j) */
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            double acc = 0.0;
            for (int k = istart; k <= j; ++k) acc += V[(int64_t)k * n + i] * H[k + j * M2];
            Vj1[i] = -1.0 * acc + Vj1[i];
        }
        double s = nrm2_n(n, Vj1);                            /* :86 */
        if (s < e->btol) {                                    /* :90-98 */
            e->k1 = 0;
            e->mb = j + 1;
            tau = e->t_final - e->t_now;
            e->n_breakdown++;
            break;
        }
        H[(j + 1) + j * M2] = s;                              /* :100 */
        double sinv = 1.0 / s;                                /* :102-105 */
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t i = 0; i < n; ++i) Vj1[i] *= sinv;
    }

    if (e->k1 != 0) {                                         /* :114-120 */
        H[(m + 1) + m * M2] = 1.0;
        e->mv(e->ctx, V + (int64_t)e->mb * n, V + (int64_t)(e->mb + 1) * n);
        e->n_matvec++;
        e->avnorm = nrm2_n(n, V + (int64_t)(e->mb + 1) * n);
    }

    double *tH = (double *)malloc(sizeof(double) * (size_t)M2 * (size_t)M2);
    int ireject = 0;
    int mx = 0;
    int replayed = (e->replay && e->replay_pos > 0 && e->replay_pos <= e->replay_count);
    while (ireject < e->max_reject) {                         /* :122-158 */
        mx = e->mb + e->k1;
        for (int k = 0; k < M2 * M2; ++k) tH[k] = tau * H[k];
        if (orc_expm(M2, tH, e->F) != 0) rc = 2;
        if (e->k1 == 0) { e->err_loc = e->btol; break; }
        double phi1 = fabs(e->beta * e->F[(mx - 2) + 0 * M2]);
        double phi2 = fabs(e->beta * e->F[(mx - 1) + 0 * M2] * e->avnorm);
        if (phi1 > phi2 * 10.0) { e->err_loc = phi2; e->xm = 1.0 / (double)mx; }
        else if (phi1 > phi2) { e->err_loc = (phi1 * phi2) / (phi1 - phi2); e->xm = 1.0 / (double)mx; }
        else { e->err_loc = phi1; e->xm = 1.0 / (double)(mx - 1); }
        if (replayed || e->err_loc <= 1.2 * tau * e->tol) break;
        tau = 0.9 * tau * pow(tau * e->tol / e->err_loc, e->xm);
        tau = round_up_2sig(tau);
        ireject++;
        e->n_reject++;
    }
    if (ireject >= e->max_reject) rc = 1;
    free(tH);

    mx = e->mb + (e->k1 - 1 > 0 ? e->k1 - 1 : 0);             /* :160-166 */
    const double beta = e->beta;
    const double *F0 = e->F;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int k = 0; k < mx; ++k) acc += V[(int64_t)k * n + i] * F0[k];
        v[i] = beta * acc;
    }

    e->t_now = e->t_now + tau;                                /* :171-174 */
    e->t_new = 0.9 * tau * pow(tau * e->tol / e->err_loc, e->xm);
    e->t_new = round_up_2sig(e->t_new);
    e->tau_last = tau;
    if (e->tau_log && e->i_step < e->tau_cap) e->tau_log[e->i_step] = tau;
    e->i_step++;
    return rc;
}

int orc_expv_solve(orc_expv *e, double *v)
{
    /* src/KryExpvFSP.cu:184-188 */
    int rc = 0;
    while (!orc_expv_final_time_reached(e)) {
        int r = orc_expv_step(e, v);
        if (r) rc = r;
        if (e->i_step > 100000) return 3;
    }
    return rc;
}

typedef struct {
    int64_t n;
    const int *rp, *cols;
    const double *vals;
} csr_ctx;

static void csr_mv(void *ctx, const double *x, double *y)
{
    csr_ctx *c = (csr_ctx *)ctx;
    orc_csr_apply(c->n, c->rp, c->cols, c->vals, 1.0, x, y);
}

int orc_expv_csr(int64_t n, const int *row_ptrs, const int *cols,
                 const double *vals, double *v, double t_final, int m,
                 double tol, int iop, int q_iop, double anorm,
                 orc_expv_stats *stats)
{
    csr_ctx c = {n, row_ptrs, cols, vals};
    orc_expv *e = orc_expv_create(n, t_final, csr_mv, &c, m, tol, iop, q_iop, anorm);
    int rc = orc_expv_solve(e, v);
    if (stats) orc_expv_get_stats(e, stats);
    orc_expv_destroy(e);
    return rc;
}

int orc_expv_csr_outputs(int64_t n, const int *row_ptrs, const int *cols,
                         const double *vals, double *v, const double *t_out,
                         int n_out, double *out, int m, double tol, int iop,
                         int q_iop, double anorm, const double *replay,
                         int n_replay, orc_expv_stats *stats)
{
    csr_ctx c = {n, row_ptrs, cols, vals};
    orc_expv *e = orc_expv_create(n, t_out[0], csr_mv, &c, m, tol, iop, q_iop, anorm);
    if (replay && n_replay > 0) orc_expv_set_replay(e, replay, n_replay);
    int rc = 0;
    for (int k = 0; k < n_out; ++k) {
        orc_expv_set_final_time(e, t_out[k]);
        int r = orc_expv_solve(e, v);
        if (r) rc = r;
        memcpy(out + (int64_t)k * n, v, sizeof(double) * (size_t)n);
    }
    if (stats) orc_expv_get_stats(e, stats);
    orc_expv_destroy(e);
    return rc;
}

void orc_op_apply(const orc_op *op, const double *x, double *y)
{
    if (op->kind == 0) {
        const int64_t rows = op->n_rows - (op->dd_last ? 1 : 0);
        orc_csr_apply(rows, op->rp, op->cols, op->vals, 1.0, x, y);
        if (op->dd_last) { /* the sink row, summed order-independently */
            ddv s = {0.0, 0.0};
            for (int64_t k = op->rp[rows]; k < op->rp[rows + 1]; ++k)
                s = dd_add_prod(s, op->vals[k], x[op->cols[k]]);
            y[rows] += dd_value(s);
        }
    } else if (op->kind == 1) {
        orc_ell_apply(op->n_rows - (op->sink_out ? 1 : 0), op->U, op->diag, op->slot_vals,
                      op->slot_cols, op->sink_out, x, y);
    } else if (op->kind == 3 ||
               orc_kron_slot_apply(op->model, op->coefs, op->with_sink, x, y) != 0) {
        /* kind 2, the device's KRON form: by slot where it applies, else per
         * reaction; kind 3: per reaction (the reference's formula) */
        orc_kron_apply(op->model, op->coefs, op->with_sink, x, y);
    }
}

static void op_mv(void *ctx, const double *x, double *y) { orc_op_apply((const orc_op *)ctx, x, y); }

int orc_expv_op_outputs(const orc_op *op, double *v, const double *t_out, int n_out,
                        double *out, int m, double tol, int iop, int q_iop, double anorm,
                        const double *replay, int n_replay, orc_expv_stats *stats,
                        double *tau_log, int64_t tau_cap)
{
    const int64_t n = op->n_rows;
    orc_expv *e = orc_expv_create(n, t_out[0], op_mv, (void *)op, m, tol, iop, q_iop, anorm);
    if (replay && n_replay > 0) orc_expv_set_replay(e, replay, n_replay);
    if (tau_log && tau_cap > 0) orc_expv_set_tau_log(e, tau_log, tau_cap);
    int rc = 0;
    for (int k = 0; k < n_out; ++k) {
        orc_expv_set_final_time(e, t_out[k]);
        int r = orc_expv_solve(e, v);
        if (r) rc = r;
        memcpy(out + (int64_t)k * n, v, sizeof(double) * (size_t)n);
        if (g_verbose) {
            fprintf(stderr, "[oracle] t = %g reached after %lld Krylov steps\n", t_out[k],
                    (long long)e->i_step);
            fflush(stderr);
        }
    }
    if (stats) orc_expv_get_stats(e, stats);
    orc_expv_destroy(e);
    return rc;
}

double orc_loglik_indices(const double *p, const int64_t *idx, int64_t count,
                          double floor_val)
{
    /* src/Likelihood.cpp:14-20 */
    double ll = 0.0;
    for (int64_t k = 0; k < count; ++k) {
        double v = p[idx[k]];
        ll += log(v > floor_val ? v : floor_val);
    }
    return ll;
}
