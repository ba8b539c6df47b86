/*
 * fsp_capi.h -- C ABI of the MI355X FSP engine (HIP kernels for gfx950).
 *
 * Plain pointers and sizes only.  "d_" pointers are device (HBM) pointers,
 * "h_" pointers host pointers.  `stream` is a hipStream_t passed as void*
 * (NULL = the null stream).  Every call returns an fsp_status; on failure
 * fsp_last_error() returns a message (thread-local).  No call exits the
 * process: the reference's exit(EXIT_FAILURE) on CUDA errors
 * (src/cme_util.h:9-22) becomes a status.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the voduchuy/CUDA_FSP source tree).
 */
#ifndef FSP_CAPI_H
#define FSP_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum fsp_status {
    FSP_OK = 0,
    FSP_ERR_INVALID = 1,       /* bad argument                                 */
    FSP_ERR_UNSUPPORTED = 2,   /* format / precision / size not supported      */
    FSP_ERR_OVERFLOW = 3,      /* state or nonzero count beyond int32          */
    FSP_ERR_HIP = 4,           /* HIP runtime error                            */
    FSP_ERR_NOMEM = 5,         /* device allocation failed                     */
    FSP_ERR_NOT_CONVERGED = 6, /* reject / step cap exhausted                  */
    FSP_ERR_NO_DEVICE = 7,     /* no GPU visible                               */
    FSP_ERR_TIMEOUT = 8        /* a batch's persistent kernel hit its time limit */
} fsp_status;

/* src/FSPMat.h:17-19 (enum MatrixFormat {CUDA_CSR, HYB, KRONECKER}) */
typedef enum fsp_format { FSP_CSR = 0, FSP_HYB = 1, FSP_KRON = 2 } fsp_format;

typedef enum fsp_precision {
    FSP_F64 = 0,   /* values, vectors and sums in FP64                           */
    FSP_MIXED = 1  /* FP32 values and Krylov basis, FP64 row sums / reductions  */
} fsp_precision;

/* Model description: the arguments of the FSPMat constructors
 * (src/FSPMat.h:43-52, src/FSPMat.cu:10-76).
 *   bounds[d]            : fsp_dim, box 0..bounds[s]
 *   stoich_*             : CSRMatInt stoich, reactions x species (host)
 *   factors              : host, R * sum_s(bounds[s]+1) separable propensity
 *                          tables in the layout of
 *                          fspmat_sdkron_fill_host_data
 *                          (src/fspmat_kron_kernels.cu:170-190), or NULL
 *   prop_table_dev       : device, R * n values a_r(x_i) (what a PropFun
 *                          evaluates, src/cme_util.h:25), or NULL
 * Exactly one of factors / prop_table_dev must be given.  The KRON format
 * needs factors. */
typedef struct fsp_model {
    int n_species;
    int n_reactions;
    const int *bounds;
    const int *stoich_vals;
    const int *stoich_cols;
    const int *stoich_rowptrs;
    const double *factors;
    const double *prop_table_dev;
} fsp_model;

const char *fsp_last_error(void);
int fsp_version(void);

/* ---- device memory helpers (so C/C++ callers need no HIP headers) ---- */
fsp_status fsp_device_count(int *count);
fsp_status fsp_malloc(void **d_ptr, size_t bytes);
fsp_status fsp_free(void *d_ptr);
fsp_status fsp_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes);
fsp_status fsp_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes);
fsp_status fsp_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, void *stream);
fsp_status fsp_memset_zero(void *d_ptr, size_t bytes, void *stream);
fsp_status fsp_stream_sync(void *stream);
/* Sum / L1 norm of a device FP64 vector (deterministic two-stage reduction).
 * Replaces thrust::reduce in the reference tests (tests/test_fsp_mat.cu:154). */
fsp_status fsp_vec_sum(const double *d_x, int64_t n, double *h_sum, void *stream);
fsp_status fsp_vec_axpy(double a, const double *d_x, double *d_y, int64_t n, void *stream);
fsp_status fsp_vec_fill(double *d_x, double value, int64_t n, void *stream);

/* ---- state space: src/cme_util.cu ---- */
/* rect_fsp_num_states (:55-62); FSP_ERR_OVERFLOW past 2^31-1 states. */
fsp_status fsp_num_states(int n_species, const int *bounds, int64_t *n_states);
/* fsp_get_states kernel (:24-41): d_states is AoS int[n][d]. */
fsp_status fsp_states_enumerate(int n_species, const int *bounds, int *d_states,
                                void *stream);
/* state2indx (:13-22) over a batch of AoS states; states outside the box map
 * to -1. */
fsp_status fsp_states_to_index(int n_species, const int *bounds, const int *d_states,
                               int64_t count, int *d_index, void *stream);

/* ---- generator: src/FSPMat.cu, src/fspmat_{csr,hyb,kron}_kernels.cu ---- */
typedef struct fsp_mat_s *fsp_mat;

typedef struct fsp_mat_info {
    int64_t n_states;   /* box states n                                      */
    int64_t n_rows;     /* n, or n+1 with the sink row                        */
    int64_t nnz;        /* stored entries (CSR incl. sink row; HYB n*(U+1))   */
    int n_slots;        /* HYB: unique net-change slots U                     */
    int format;
    int precision;
    int with_sink;
    int max_row;        /* longest CSR row (sink row excluded)               */
    double bytes_per_apply; /* algorithmic HBM bytes of one apply            */
} fsp_mat_info;

/* FSPMat constructors (src/FSPMat.cu:10-76) with generate_fsp_mats_cuda_csr
 * (src/fspmat_csr_kernels.cu:138-248), generate_fsp_mats_hyb
 * (src/fspmat_hyb_kernels.cu:108-184) and generate_fsp_mats_sdkron
 * (src/fspmat_kron_kernels.cu:131-168).  All assembly runs on the GPU.
 * `coefs` (host, R) are the coefficients the CSR/HYB values are built with;
 * apply() re-values them when it is given different coefficients.
 * with_sink appends state n that collects the outflow leaving the box. */
fsp_status fsp_mat_create(const fsp_model *model, fsp_format format,
                          fsp_precision precision, int with_sink,
                          const double *coefs, void *stream, fsp_mat *out);
/* FSPMat::~FSPMat / destroy (src/FSPMat.cu:79-95) */
fsp_status fsp_mat_destroy(fsp_mat A);
fsp_status fsp_mat_get_info(fsp_mat A, fsp_mat_info *info);
/* Rebuild CSR/HYB values for new coefficients (no-op for KRON). */
fsp_status fsp_mat_set_coefs(fsp_mat A, const double *h_coefs, void *stream);
/* FSPMat::action (src/FSPMat.cu:102-105): y (+)= A(coefs) x.
 * h_coefs NULL = the coefficients last set.  accumulate != 0 gives the
 * reference's y += A x; 0 overwrites y.  Vectors have n_rows entries.
 * FP64 vectors for both precisions (a mixed matrix stores FP32 values and
 * sums rows in FP64). */
fsp_status fsp_mat_apply(fsp_mat A, const double *h_coefs, const double *d_x,
                         double *d_y, int accumulate, void *stream);
/* y (+)= A^T x (adjoint operator; no reference counterpart). */
fsp_status fsp_mat_apply_transpose(fsp_mat A, const double *h_coefs,
                                   const double *d_x, double *d_y, int accumulate,
                                   void *stream);
/* Single-term apply y (+)= c * A_r x (per-reaction generator; the building
 * block of rate gradients; no reference counterpart). */
fsp_status fsp_mat_apply_term(fsp_mat A, int reaction, double c, const double *d_x,
                              double *d_y, int accumulate, void *stream);
/* Copy out the merged CSR (FSP_CSR only): arrays sized from info. */
fsp_status fsp_mat_export_csr(fsp_mat A, int *d_row_ptrs, int *d_cols, double *d_vals,
                              void *stream);
/* Copy out the HYB slots (FSP_HYB only): diag[n], slot_vals[U*n],
 * slot_cols[U*n], sink_out[n] (may be NULL).  The library stores no
 * columns (slot u of row i reads x at i - delta_u); slot_cols is rebuilt:
 * i - delta_u where that source lies in the box, else i. */
fsp_status fsp_mat_export_hyb(fsp_mat A, double *d_diag, double *d_slot_vals,
                              int *d_slot_cols, double *d_sink_out, void *stream);

/* Per-reaction CSR in the reference layout (fspmat_csr_get_nnz_per_row,
 * the row-pointer scan and fspmat_csr_fill_data,
 * src/fspmat_csr_kernels.cu:7-94,206-213), assembled on the GPU.
 * Call with d_cols == NULL to get *h_nnz (row pointers are written if
 * d_row_ptrs != NULL), then again with arrays of h_nnz entries. */
fsp_status fsp_term_csr(const fsp_model *model, int reaction, int *d_row_ptrs,
                        int *d_cols, double *d_vals, int64_t *h_nnz, void *stream);

/* ---- Krylov expv: src/KryExpvFSP.{h,cu} ---- */
typedef struct fsp_expv_s *fsp_expv;

typedef struct fsp_expv_stats {
    double t_now, t_new, tau_last, err_last, beta_last;
    int64_t i_step, n_matvec, n_reject, n_breakdown;
} fsp_expv_stats;

/* Host callback operator: y += A x on device vectors (the reference's MVFun,
 * src/KryExpvFSP.h:14).  Called from the host thread driving the solve. */
typedef void (*fsp_matvec_cb)(void *ctx, const double *d_x, double *d_y);

/* KryExpvFSP(t_final, matvec, v, m, tol, iop, q_iop, anorm)
 * (src/KryExpvFSP.cu:6-36).  Operator: an fsp_mat (A != NULL; fused device
 * path, hipGraph-captured steps) or a callback (A == NULL).  n = vector
 * length.  m <= 50.  For a FSP_MIXED matrix the basis is FP32. */
fsp_status fsp_expv_create(fsp_mat A, fsp_matvec_cb cb, void *cb_ctx, int64_t n,
                           double t_final, int m, double tol, int iop, int q_iop,
                           double anorm, void *stream, fsp_expv *out);
fsp_status fsp_expv_destroy(fsp_expv e);
/* KryExpvFSP::step (src/KryExpvFSP.cu:38-182): one adaptive step, v in place */
fsp_status fsp_expv_step(fsp_expv e, double *d_v);
/* KryExpvFSP::solve (src/KryExpvFSP.cu:184-188) */
fsp_status fsp_expv_solve(fsp_expv e, double *d_v);
/* KryExpvFSP::reset_time (src/KryExpvFSP.h:82-84) */
fsp_status fsp_expv_reset_time(fsp_expv e);
/* New final time (continues from t_now; clears the breakdown latch).
 * No reference counterpart (the reference's FSPSolver calls a missing
 * update_final_time, src/FSPSolver.cpp:12). */
fsp_status fsp_expv_set_final_time(fsp_expv e, double t_final);
/* Fresh solve with the same buffers: t_now = 0, step counter and statistics
 * cleared (the first step size is recomputed from |v|).  No reference
 * counterpart (a new KryExpvFSP object there). */
fsp_status fsp_expv_restart(fsp_expv e, double t_final);
/* KryExpvFSP::final_time_reached (src/KryExpvFSP.h:100-102) */
fsp_status fsp_expv_final_time_reached(fsp_expv e, int *reached);
fsp_status fsp_expv_get_stats(fsp_expv e, fsp_expv_stats *stats);
/* Accepted step sizes so far (the first min(count, cap) are copied).  No
 * reference counterpart: lets a CPU run replay the device's step sequence. */
fsp_status fsp_expv_get_tau_history(fsp_expv e, double *h_taus, int64_t cap,
                                    int64_t *count);
/* Public knobs max_reject / max_nstep (src/KryExpvFSP.h:68-69). */
fsp_status fsp_expv_set_limits(fsp_expv e, int max_reject, int64_t max_nstep);
/* Public knobs IOP / q_iop (src/KryExpvFSP.h:64-65), read by the next step. */
fsp_status fsp_expv_set_iop(fsp_expv e, int iop, int q_iop);

/* Per-kernel timing of the fused Krylov step (no reference counterpart; the
 * reference has no tracing).  While enabled, steps run without the captured
 * HIP graph (one step per host check) and every kernel launch is bracketed
 * by HIP events on the handle's stream; each launch of an active step adds
 * its time and its algorithmic bytes (every stored matrix value and vector
 * entry the operation needs once) to its class.  Enabling (or disabling)
 * clears the totals.  Matrix operators only. */
#define FSP_PROF_PROLOGUE 0  /* |v| partials and the step prologue           */
#define FSP_PROF_ARNOLDI 1   /* row kernel: SpMV, V_j store, dots vs V_k      */
#define FSP_PROF_DOTS 2      /* per-dot finalizers                            */
#define FSP_PROF_UPDATE 3    /* w -= V h and |w|^2 partials                   */
#define FSP_PROF_CONTROL 4   /* Pade exponential, error estimate, step size   */
#define FSP_PROF_SOLUTION 5  /* v = beta V F0                                 */
#define FSP_PROF_CLASSES 6
typedef struct fsp_expv_profile {
    int64_t launches[FSP_PROF_CLASSES];
    double ms[FSP_PROF_CLASSES];
    double bytes[FSP_PROF_CLASSES];
} fsp_expv_profile;
fsp_status fsp_expv_set_profiling(fsp_expv e, int enable);
/* Krylov basis precision for a FSP_MIXED matrix: FSP_MIXED (the default: an
 * FP32 basis, Gram-corrected) or FSP_F64 (FP32 generator values with an FP64
 * basis: the same Arnoldi process as FSP_F64 on the FP32-rounded generator).
 * Call before the first step (it restarts the handle).  No reference
 * counterpart. */
fsp_status fsp_expv_set_basis_precision(fsp_expv e, int precision);
fsp_status fsp_expv_get_profile(fsp_expv e, fsp_expv_profile *profile);

/* ---- batched solves (no reference counterpart: the reference solves one
 * system at a time, src/KryExpvFSP.cu:184-188) ---- */
typedef struct fsp_batch_s *fsp_batch;

typedef struct fsp_batch_options {
    int m;            /* Krylov dimension (default 30, <= 50)                  */
    double tol;       /* Krylov tolerance (default 1e-8)                       */
    double anorm;     /* norm estimate (default 1.0, as the reference)         */
    int precision;    /* FSP_F64, or FSP_MIXED: FP32 Krylov basis              */
    int iop, q_iop;   /* incomplete orthogonalisation                          */
    int lockstep_bpc; /* lockstep row-kernel workgroups per CU (0: one per
                         1024-row tile)                                        */
    int wg_per_cu;    /* persistent-kernel workgroups per CU (0: 2, <= 8)      */
    int max_reject;   /* 0: 1000                                               */
    int64_t max_nstep;/* 0: 100000                                             */
    int path;         /* FSP_BATCH_AUTO (0) or one of the paths below          */
    int64_t split_states; /* HYBRID: instances of at most this many entries go
                         to the persistent kernel (0: 16384; < 0: all)         */
    double time_limit_s;  /* wall-clock seconds a persistent workgroup may run
                         before its instances end with FSP_ERR_TIMEOUT (0: 600) */
    int lockstep_rows;    /* lockstep tile rows per thread: 2, 4 or 8 (tiles of
                         512, 1024 or 2048 rows; 0: 4)                         */
} fsp_batch_options;

/* Batch execution paths.  LOCKSTEP: one graph-captured batch step advances
 * every running instance by one Krylov step (row kernels over the 1024-row
 * tiles of all running instances; per-instance step control).  PERSISTENT:
 * one persistent work-queue kernel in which each workgroup solves whole
 * instances, largest estimated cost first; the launch (reset, kernel, stats
 * export) is one captured HIP graph.  HYBRID: the persistent kernel takes
 * the instances of at most split_states entries and the lockstep steps the
 * others, at the same time on two streams.  STREAMS: each instance through
 * the single-solve engine, up to four at a time on their own streams.
 * AUTO: the autotuner's choice, else LOCKSTEP. */
#define FSP_BATCH_AUTO 0
#define FSP_BATCH_LOCKSTEP 1
#define FSP_BATCH_STREAMS 2
#define FSP_BATCH_PERSISTENT 3
#define FSP_BATCH_HYBRID 4

/* count independent solves v_i <- exp(t_final[i] A_i) v_i (Krylov expv,
 * KryExpvFSP semantics, each instance with its own step sizes).  models[i]
 * needs factor tables; coefs holds the R_i coefficients of every instance
 * in order; with_sink may be NULL.  Initial vectors default to
 * delta(state 0). */
fsp_status fsp_batch_create(const fsp_model *models, const double *coefs,
                            const double *t_final, const int *with_sink, int count,
                            const fsp_batch_options *options, void *stream,
                            fsp_batch *out);
fsp_status fsp_batch_destroy(fsp_batch b);
fsp_status fsp_batch_set_initial(fsp_batch b, int i, const double *h_p0);
/* One solve of every instance from its initial vector.  PERSISTENT: the
 * captured graph launched on `stream` (asynchronous; the getters wait for
 * it, and an instance's status reports its failure).  LOCKSTEP, HYBRID and
 * STREAMS wait for `stream`, run on the batch's own streams and return when
 * every instance is done; an instance that fails (reject loop, step limit,
 * time limit) makes the call return its status and a message naming it,
 * the others' results stay valid. */
fsp_status fsp_batch_solve(fsp_batch b, void *stream);
/* Does the set-up of the path a solve would take (device buffers, HYB
 * generators, tile tables, captured graphs) so that the next solve does
 * none; optional (a solve prepares what is missing).  STREAMS: nothing. */
fsp_status fsp_batch_prepare(fsp_batch b);
fsp_status fsp_batch_get_solution(fsp_batch b, int i, double *h_p, int64_t cap);
fsp_status fsp_batch_get_stats(fsp_batch b, int i, int64_t *steps, int64_t *rejects,
                               int *status);
/* Configuration of the path in use: PERSISTENT (workgroup size, workgroups
 * per CU); LOCKSTEP (-1, row-kernel workgroups per CU, 0 = one per 1024-row
 * tile); HYBRID (-2, row-kernel workgroups per CU); STREAMS (0, concurrent
 * streams). */
fsp_status fsp_batch_get_config(fsp_batch b, int *block, int *wg_per_cu);
/* The path a solve takes and every tuned setting (any pointer may be NULL). */
fsp_status fsp_batch_get_tuning(fsp_batch b, int *path, int *lockstep_bpc, int *lockstep_rows,
                                int *persistent_wg, int64_t *split_states);
/* Times candidate configurations on this batch, one whole solve each after
 * an untimed one (LOCKSTEP: row-kernel workgroups one per tile, 4 or 16 per
 * CU, times tiles of 512, 1024 or 2048 rows; PERSISTENT: 1, 2 or 4
 * workgroups per CU; HYBRID: the lockstep configurations, then splits at
 * 4,096, 16,384 and 65,536 entries; AUTO: all of these and the path;
 * STREAMS: 1, 2 or 4 streams), keeps the fastest and records it in the JSON
 * file cache_path (NULL: no cache); a cached entry for the same device /
 * shape is reused unless force. */
fsp_status fsp_batch_autotune(fsp_batch b, const char *cache_path, int force, double *best_ms);

/* ---- time-varying right-hand side and BDF integration ---- */
/* Host time function: writes the n_reactions coefficients c_r(t) (the
 * reference's TcoefFun, src/FSPMat.h:16, with a context pointer). */
typedef void (*fsp_tcoef_fn)(double t, double *h_coefs, void *ctx);

/* udot = A(t) u with c_r = coef(t): the body of the reference's CVODE RHS
 * and Jacobian-times-vector callbacks (tests/test_sundials.cu:67-91, which
 * zero-fill udot and call FSPMat::action).  Overwrites d_udot. */
fsp_status fsp_rhs_apply(fsp_mat A, double t, fsp_tcoef_fn coef, void *ctx,
                         const double *d_u, double *d_udot, void *stream);

typedef struct fsp_bdf_s *fsp_bdf;
typedef struct fsp_bdf_options {
    double rtol, atol;     /* error weights 1/(rtol |y| + atol) (1e-6, 1e-12)   */
    int max_order;         /* 1..5 (5)                                          */
    int krylov_dim;        /* GMRES subspace (30)                               */
    int max_restarts;      /* GMRES restarts per stage (5; negative: default)   */
    double h0;             /* first step (0: from |A(t0) y0|)                   */
    int64_t max_steps;     /* (500000)                                          */
    double lin_tol_factor; /* stage residual tolerance / alpha_0 (0.05)         */
} fsp_bdf_options;
typedef struct fsp_bdf_stats {
    int64_t n_steps, n_rejects, n_rhs, n_lin_iters, n_lin_fails;
    int order;
    double h_last, t;
} fsp_bdf_stats;
/* dp/dt = A(t) p integrated by a variable-order (1..5), variable-step BDF
 * with GMRES stages: the role of CVODE's BDF + SPBCGS with the RHS above in
 * tests/test_sundials.cu:131-186 (CVODE itself is not used).  The matrix A
 * is revalued at each stage time; the caller keeps A alive. */
fsp_status fsp_bdf_create(fsp_mat A, fsp_tcoef_fn coef, void *ctx, double t0,
                          const fsp_bdf_options *options, void *stream, fsp_bdf *out);
fsp_status fsp_bdf_destroy(fsp_bdf b);
/* Integrates to t_out (CVode with CV_NORMAL, landing on t_out).  The first
 * call after create / reset reads the initial vector from d_v; every call
 * writes p(t_out) to d_v and continues from its own history. */
fsp_status fsp_bdf_advance(fsp_bdf b, double *d_v, double t_out);
/* Restart at t0: the next advance reads a new initial vector. */
fsp_status fsp_bdf_reset(fsp_bdf b, double t0);
fsp_status fsp_bdf_get_stats(fsp_bdf b, fsp_bdf_stats *stats);

/* ---- likelihood: src/Likelihood.cpp:8-33 ---- */
/* sum_k log(max(floor, p[idx[k]])) over device index lists; idx < 0 (a state
 * outside the box) counts as probability 0.  Deterministic reduction. */
fsp_status fsp_loglik_indices(const double *d_p, const int *d_idx, int64_t count,
                              double floor_val, double *h_ll, void *stream);
/* fsp_loglik_indices without any device allocation.  No reference
 * counterpart: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * partial sums live in a fixed 16 KiB device array of the library (one per
 * GPU, held until the process ends), so any count is accepted (at most
 * 1024 partial sums); the result equals fsp_loglik_indices' bit for bit.
 * Calls that overlap on one GPU (other threads or streams) share that array
 * and race: serialise them. */
fsp_status fsp_loglik_indices_noalloc(const double *d_p, const int *d_idx, int64_t count,
                                      double floor_val, double *h_ll, void *stream);
/* Marginal of p over the species with keep_mask[s] != 0: d_out has
 * prod_{kept} (bounds[s]+1) entries, kept species in order, first fastest.
 * No reference counterpart (the joint-histogram likelihood needs it). */
fsp_status fsp_marginal(const double *d_p, int n_species, const int *bounds,
                        const int *keep_mask, double *d_out, void *stream);
/* sum_k w[k] log(max(floor, q[k])) (weighted histogram log-likelihood). */
fsp_status fsp_loglik_weighted(const double *d_q, const double *d_w, int64_t count,
                               double floor_val, double *h_ll, void *stream);
/* Histogram likelihood of FSP solutions: the intended semantics of
 * cme::Likelihood::evaluate (src/Likelihood.cpp:8-33) on joint marginal
 * histograms, LL = sum_k sum_b counts[k][b] log max(P_k[b], floor_val),
 * P_k the marginal over the species with keep_mask[s] != 0 (first kept
 * species fastest) of p(times[k]) = exp(times[k] A) p0, no sink state. */
typedef struct fsp_lik_data {
    int n_times;
    const double *times;        /* increasing, > 0                             */
    const int *keep_mask;       /* [n_species]                                 */
    const double *counts;       /* host [n_times][n_bins]                      */
    double floor_val;           /* e.g. 1e-12 (the reference uses 1e-44)       */
} fsp_lik_data;
typedef struct fsp_grad_options {
    fsp_format format;          /* forward operator (FSP_CSR)                  */
    int m;                      /* Krylov dimension (30)                       */
    double tol;                 /* Krylov tolerance (1e-10)                    */
    int quad_nodes;             /* Gauss-Legendre nodes per panel (8)          */
    int grading_levels;         /* panels per observation interval / 2 (10)    */
} fsp_grad_options;
/* LL at coefficients exp(log_coefs); the model needs factor tables.  d_p0:
 * initial distribution (n states).  options may be NULL. */
fsp_status fsp_loglik(const fsp_model *model, const double *log_coefs, const double *d_p0,
                      const fsp_lik_data *data, const fsp_grad_options *options,
                      double *h_ll, void *stream);
/* LL and dLL/dlog_coefs by the adjoint method: lambda integrated backwards
 * with the transposed generator (Krylov expv on A^T), the sensitivity
 * integrals by graded Gauss-Legendre panels, p recomputed per observation
 * interval from checkpoints.  No reference counterpart. */
fsp_status fsp_loglik_grad(const fsp_model *model, const double *log_coefs,
                           const double *d_p0, const fsp_lik_data *data,
                           const fsp_grad_options *options, double *h_ll, double *h_grad,
                           void *stream);
/* LL and gradient at n_points parameter points: log_coefs and h_grad are
 * n_points x n_reactions (row-major), h_ll has n_points entries.  Points
 * run on `workers` host threads with a stream each (<= 0: 4, the hardware
 * queues of a process).  The gradient keeps p at every quadrature node in
 * device memory when that fits a budget (fsp_loglik_grad: half the free
 * memory; here free / (2 workers), measured once before the workers
 * start) and recomputes the nodes from checkpoints otherwise.  When a
 * point and fsp_loglik_grad take the same path the results are bit for bit
 * equal; across the two paths they agree to the Krylov tolerance.
 * The error names the first failed point.  The reference has no gradient;
 * its batched counterpart would be a loop over cme::Likelihood::evaluate
 * (src/Likelihood.cpp:8-33). */
fsp_status fsp_loglik_grad_batch(const fsp_model *model, int n_points, const double *log_coefs,
                                 const double *d_p0, const fsp_lik_data *data,
                                 const fsp_grad_options *options, int workers, double *h_ll,
                                 double *h_grad);
/* Dense matrix exponential on the device (one workgroup, FP64, Pade q=6
 * with scaling and squaring; stands in for arma::expmat at
 * src/KryExpvFSP.cu:125).  Column-major N x N, N <= 52. */
fsp_status fsp_expm_dense(int N, const double *d_A, double *d_F, void *stream);

#ifdef __cplusplus
}
#endif
#endif
