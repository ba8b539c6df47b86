// fsp_batch_impl.h -- host state of a batch of independent FSP solves,
// shared by the batch set-up and autotuner (fsp_batch.hip), the lockstep and
// persistent-kernel paths (fsp_batch_lockstep.hip) and the stream scheduler
// (fsp_batch_streams.hip).
#pragma once

#include <vector>

#include "fsp_expm.h"
#include "fsp_formats.h"

namespace fsp {

// One instance of a batch as the set-up sees it; the device copy
// (BatchImpl::d_inst) receives each instance's step and reject counts and
// status at the end of a solve.
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

// Owned copy of one instance's fsp_model arrays (the caller's arrays need
// not outlive fsp_batch_create).
struct ModelCopy {
    int d = 0, R = 0;
    std::vector<int> bounds, sv, sc, sr;
    std::vector<double> factors;
    fsp_model view() const
    {
        fsp_model m;
        m.n_species = d;
        m.n_reactions = R;
        m.bounds = bounds.data();
        m.stoich_vals = sv.data();
        m.stoich_cols = sc.data();
        m.stoich_rowptrs = sr.data();
        m.factors = factors.data();
        m.prop_table_dev = nullptr;
        return m;
    }
};

struct LsHost;  // device state of the lockstep path (fsp_batch_lockstep.hip)

struct BatchImpl {
    int count = 0;
    int path_req = FSP_BATCH_AUTO;     // requested path
    int path_tuned = FSP_BATCH_AUTO;   // autotuner's choice for AUTO
    int ls_bpc = 0;                    // lockstep row-kernel workgroups per CU (0: one per tile)
    int ls_rpt = 4;                    // lockstep tile rows / 256 (2, 4 or 8)
    int pq_wg = 2;                     // persistent-kernel workgroups per CU
    int64_t pq_split = 16384;          // HYBRID: instances of at most this many states
                                       // go to the persistent kernel
    double pq_budget_s = 600.0;        // wall-clock limit of a persistent workgroup
    LsHost *ls = nullptr;
    hipStream_t solve_stream = nullptr;  // stream of the last solve (persistent path)
    int m = 30, iop = 0, q_iop = 2, max_reject = 1000;
    long long max_nstep = 100000;
    double tol = 1e-8, anorm = 1.0;
    bool mixed = false;
    int cus = 256, dmax = 4;
    std::vector<HostModel> hms;
    std::vector<BatchInst> h_inst;
    std::vector<int64_t> vec_off;
    std::vector<double> cost;          // estimated work per instance
    std::vector<int> order;            // instances by decreasing cost
    DevModel *d_models = nullptr;
    double *d_tabs = nullptr, *d_coefs = nullptr, *d_vecs = nullptr, *d_p10 = nullptr;
    BatchInst *d_inst = nullptr;
    int nv_max = 0;
    // stream scheduler
    std::vector<ModelCopy> mc;
    std::vector<double> h_coefs;
    std::vector<int64_t> coef_off;
    int workers = 4;                   // concurrent streams (hardware queues)
    int built_workers = 0;
    std::vector<hipStream_t> wstreams;
    std::vector<std::vector<int>> parts;
    std::vector<fsp_mat> smat;
    std::vector<fsp_expv> sexpv;
};

// Stream scheduler (fsp_batch_streams.hip): every instance solved by the
// single-instance Krylov engine, `workers` instances at a time on their own
// streams.
fsp_status batch_solve_streams(BatchImpl &B);
// Lockstep path (fsp_batch_lockstep.hip): all running instances advanced
// one Krylov step per replay of one captured batch step.
fsp_status batch_solve_lockstep(BatchImpl &B, int blocks_per_cu);
// Allocations and the captured batch step for that grid, ahead of a timed solve.
fsp_status batch_lockstep_prepare(BatchImpl &B, int blocks_per_cu);
// Persistent path (fsp_batch_lockstep.hip): instances of at most `split`
// states are solved whole, one per workgroup, by one persistent work-queue
// kernel (a captured graph on its own stream); the others, if any, run on
// the lockstep path at the same time.  split < 0: every instance.
fsp_status batch_solve_hybrid(BatchImpl &B, int blocks_per_cu, int wg_per_cu, int64_t split);
// The persistent kernel alone as one captured graph on `stream`
// (asynchronous: instance reset, queue reset, the kernel, the stats export).
fsp_status batch_launch_persistent(BatchImpl &B, int wg_per_cu, hipStream_t stream);
// Allocations, tile tables and captured graphs of a path (LOCKSTEP, HYBRID,
// PERSISTENT), so that a following solve does no set-up; STREAMS: nothing.
fsp_status batch_prepare(BatchImpl &B, int path);
void batch_lockstep_free(BatchImpl &B);
fsp_status batch_autotune_streams(BatchImpl &B, double *best_ms, int *best_workers);
void batch_streams_free(BatchImpl &B);

}  // namespace fsp
