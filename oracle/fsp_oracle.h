/*
 * fsp_oracle.h -- CPU restatement of the reference FSP solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never links,
 * imports or calls it.
 *
 * Every function cites the reference file:line it restates (paths are
 * relative to the voduchuy/CUDA_FSP source tree).  Pins (how each function
 * is checked against something other than itself) are listed in
 * tests/test_oracle.py; functions without an independent pin say so here.
 *
 * Conventions
 *   - states are mixed-radix, species 0 fastest (src/cme_util.cu:5-22)
 *   - the generator is A = sum_r c_r A_r; "apply" always ACCUMULATES
 *     (y += A x), as every reference format does (src/FSPMat.cu:102-105)
 *   - propensities come either from separable factor tables
 *     (a_r(x) = prod_s f_{r,s}(x_s), the layout of
 *     src/fspmat_kron_kernels.cu:170-190) or from a full per-state table
 *     prop_table[r*n + i] (what a PropFun evaluates, src/cme_util.h:25)
 *   - floating point: compiled with -ffp-contract=off so every value is a
 *     sequence of correctly rounded IEEE operations.
 */
#ifndef FSP_ORACLE_H
#define FSP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n_species;
    int n_reactions;
    const int *bounds;          /* [d] upper bounds N_s; box is 0..N_s      */
    const int *stoich_vals;     /* reactions x species CSR (host), as the   */
    const int *stoich_cols;     /* reference's CSRMatInt stoich             */
    const int *stoich_rowptrs;  /* (src/cme_util.h:35-40)                   */
    const double *factors;      /* R * sum_s(N_s+1) or NULL                 */
    const double *prop_table;   /* R * n or NULL                            */
} orc_model;

void orc_set_threads(int n);
int orc_get_threads(void);

/* src/cme_util.cu:55-62 (computed in 64 bits; the reference is int32) */
int64_t orc_num_states(int d, const int *bounds);
/* src/cme_util.cu:5-11 */
void orc_indx2state(int64_t indx, int *state, int d, const int *bounds);
/* src/cme_util.cu:13-22 */
int64_t orc_state2indx(const int *state, int d, const int *bounds);
/* src/cme_util.cu:24-41 : AoS int[n][d] */
void orc_enumerate(int d, const int *bounds, int *states);
/* src/cme_util.cu:43-53 with direction +1 applied to a zero state */
void orc_net_change(const orc_model *m, int r, int *nu);
/* propensity a_r(x) without its coefficient */
double orc_prop(const orc_model *m, const int *x, int64_t idx, int r);

/* Per-reaction CSR in the reference layout: src/fspmat_csr_kernels.cu:7-94
 * and the scan at :206-213.  Returns nnz.  row_ptrs may be NULL (count
 * only); cols/vals may be NULL (structure only). */
int64_t orc_term_csr(const orc_model *m, int r, int *row_ptrs, int *cols,
                     double *vals);
/* y += alpha * A x, sequential row sums (cusparseDcsrmv with beta = 1,
 * src/fspmat_csr_kernels.cu:123-136). */
void orc_csr_apply(int64_t n_rows, const int *row_ptrs, const int *cols,
                   const double *vals, double alpha, const double *x,
                   double *y);

/* Merged CSR of A = sum_r c_r A_r (this project's CSR format; no reference
 * counterpart, derived from the per-reaction structure above).  Entries with
 * equal column are merged; columns ascending; structural zeros kept;
 * reactions with zero net change are omitted (their rows sum to exactly 0).
 * with_sink adds row n holding the outflow sum_{r: x+nu_r outside} c_r a_r(x)
 * for every state with an outward reaction.  Returns nnz.  Any output may be
 * NULL; call once with NULLs to size the arrays. */
int64_t orc_merged_csr(const orc_model *m, const double *coefs, int with_sink,
                       int *row_ptrs, int *cols, double *vals);

/* Unique net-change vectors in order of first appearance (reactions with
 * nu = 0 skipped).  Returns U; slot_of_reaction[r] = slot or -1. */
int orc_unique_shifts(const orc_model *m, int *slot_of_reaction);

/* This project's HYB format: merged diagonal + one ELL slot per unique nu
 * (SoA [U][n]); a missing neighbour stores col = row, val = 0.  sink_out may
 * be NULL.  Mathematically equal to the merged CSR. */
void orc_ell(const orc_model *m, const double *coefs, double *diag,
             double *slot_vals, int *slot_cols, double *sink_out);
void orc_ell_apply(int64_t n, int U, const double *diag,
                   const double *slot_vals, const int *slot_cols,
                   const double *sink_out, const double *x, double *y);

/* Reference HYB layout: src/fspmat_hyb_kernels.cu:7-64 (per-reaction slots,
 * missing neighbour = col 0 / val 0) and its apply :66-80. */
void orc_hyb_ref(const orc_model *m, double *diag, double *off, int *col);
void orc_hyb_ref_apply(int64_t n, int R, const double *diag,
                       const double *off, const int *col,
                       const double *coefs, const double *x, double *y);

/* Matrix-free Kronecker-sum apply: src/fspmat_kron_kernels.cu:51-129
 * (formula only; the reference's tail-block and shared-memory race defects
 * are not reproduced).  Requires factor tables.  If with_sink, x and y have
 * length n+1 and y[n] += sum_i out_i x_i. */
void orc_kron_apply(const orc_model *m, const double *coefs, int with_sink,
                    const double *x, double *y);
/* The same operator grouped by net change (the device's KRON slot form, see
 * fsp_oracle.c): returns -1 and leaves y when the form does not apply.
 * orc_kron_slot_form: table entries, or -1. */
int orc_kron_slot_form(const orc_model *m);
int orc_kron_slot_apply(const orc_model *m, const double *coefs, int with_sink,
                        const double *x, double *y);

/* Dense matrix exponential, Golub & Van Loan, Matrix Computations (3rd ed.)
 * Algorithm 11.3.1: diagonal Pade q = 6 with power-of-two scaling by the
 * infinity norm, linear solve by Gaussian elimination with partial
 * pivoting, then repeated squaring.  Column-major N x N.  Stands in for
 * arma::expmat at src/KryExpvFSP.cu:125 (Armadillo is absent; its own
 * output is parity unpinned -- this routine is pinned against
 * scipy.linalg.expm fixtures). */
int orc_expm(int N, const double *A, double *F);

/* Expokit-style Krylov expv: a state machine that restates
 * KryExpvFSP::step (src/KryExpvFSP.cu:38-182), solve (:184-188) and
 * reset_time (src/KryExpvFSP.h:82-84). */
typedef void (*orc_matvec_fn)(void *ctx, const double *x, double *y);
typedef struct orc_expv orc_expv;

typedef struct {
    double t_now, t_new, tau_last, err_last, beta_last;
    int64_t i_step, n_matvec, n_reject, n_breakdown;
} orc_expv_stats;

orc_expv *orc_expv_create(int64_t n, double t_final, orc_matvec_fn mv,
                          void *ctx, int m, double tol, int iop, int q_iop,
                          double anorm);
void orc_expv_destroy(orc_expv *e);
int orc_expv_step(orc_expv *e, double *v);
int orc_expv_solve(orc_expv *e, double *v);
void orc_expv_reset_time(orc_expv *e);
void orc_expv_set_final_time(orc_expv *e, double t_final);
int orc_expv_final_time_reached(const orc_expv *e);
void orc_expv_get_stats(const orc_expv *e, orc_expv_stats *s);
void orc_expv_set_max_reject(orc_expv *e, int max_reject);
/* Replay mode: force the next step sizes (tau) instead of the adaptive
 * choice; used to separate arithmetic error from control-flow divergence. */
void orc_expv_set_replay(orc_expv *e, const double *taus, int count);
/* Records the accepted step sizes of the following steps into taus[i_step]
 * (the first cap of them). */
void orc_expv_set_tau_log(orc_expv *e, double *taus, int64_t cap);
/* verbose != 0: orc_expv_op_outputs prints a line per output time to stderr. */
void orc_set_verbose(int verbose);

/* Convenience: run expv to t_final with a CSR operator (y += A x). */
int orc_expv_csr(int64_t n, const int *row_ptrs, const int *cols,
                 const double *vals, double *v, double t_final, int m,
                 double tol, int iop, int q_iop, double anorm,
                 orc_expv_stats *stats);
/* Same, stopping at each of n_out output times (ascending, absolute) and
 * copying v into out[k*n ...]; one Krylov object continues across times. */
int orc_expv_csr_outputs(int64_t n, const int *row_ptrs, const int *cols,
                         const double *vals, double *v, const double *t_out,
                         int n_out, double *out, int m, double tol, int iop,
                         int q_iop, double anorm, const double *replay,
                         int n_replay, orc_expv_stats *stats);
/* Operators for the expv restatement (y += A x).  kind 0: CSR rows summed
 * in column order (dd_last: the last row is the sink row, summed
 * order-independently); 1: this project's ELL/HYB layout; 2: Kronecker in
 * the device's form (by slot where that applies); 3: Kronecker per reaction.
 * Sums across rows (the sink element) are order-independent in every kind. */
typedef struct {
    int kind;
    int64_t n_rows;
    int dd_last;
    const int *rp, *cols;
    const double *vals;
    int U;
    const double *diag, *slot_vals, *sink_out;
    const int *slot_cols;
    const orc_model *model;
    const double *coefs;
    int with_sink;
} orc_op;

void orc_op_apply(const orc_op *op, const double *x, double *y);
int orc_expv_op_outputs(const orc_op *op, double *v, const double *t_out, int n_out,
                        double *out, int m, double tol, int iop, int q_iop, double anorm,
                        const double *replay, int n_replay, orc_expv_stats *stats,
                        double *tau_log, int64_t tau_cap);

/* 2-significant-digit upward rounding of a step size (see fsp_oracle.c). */
double orc_round_up_2sig(double t);

/* src/Likelihood.cpp:8-33 : sum over observations of log(max(floor, p)). */
double orc_loglik_indices(const double *p, const int64_t *idx, int64_t count,
                          double floor_val);

#ifdef __cplusplus
}
#endif
#endif
