// fsp_batch_impl.h -- host state of a batch of independent FSP solves,
// shared by the persistent-kernel path (fsp_batch.hip), the lockstep path
// (fsp_batch_lockstep.hip) and the stream scheduler (fsp_batch_streams.hip).
#pragma once

#include <vector>

#include "fsp_batch_kernel.h"

namespace fsp {

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
    LsHost *ls = nullptr;
    hipStream_t solve_stream = nullptr;  // stream of the last solve (persistent path)
    int m = 30, iop = 0, q_iop = 2, max_reject = 1000;
    long long max_nstep = 100000;
    double tol = 1e-8, anorm = 1.0;
    bool mixed = false;
    int block = 512, wg_per_cu = 2, cus = 256, grid = 1, dmax = 4;
    std::vector<HostModel> hms;
    std::vector<BatchInst> h_inst;
    std::vector<int64_t> vec_off;
    std::vector<double> cost;          // estimated work per instance (LPT order)
    std::vector<int> order;            // instances by decreasing cost
    DevModel *d_models = nullptr;
    double *d_tabs = nullptr, *d_coefs = nullptr, *d_vecs = nullptr, *d_p10 = nullptr;
    BatchInst *d_inst = nullptr;
    int *d_order = nullptr;
    unsigned int *d_queue = nullptr;
    double *ws = nullptr;
    float *ws32 = nullptr;
    size_t ws_stride = 0, ws32_stride = 0, ws_slots = 0;
    int nv_max = 0, tab_lds = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_block = 0, graph_wgcu = 0;
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
void batch_lockstep_free(BatchImpl &B);
fsp_status batch_autotune_streams(BatchImpl &B, double *best_ms, int *best_workers);
void batch_streams_free(BatchImpl &B);

}  // namespace fsp
