// fsp_batch_streams.hip -- the batch through the single-instance Krylov
// engine, several instances at a time.
//
// Instances are partitioned over `workers` host threads, largest estimated
// cost first, each to the least loaded worker.  Every thread drives its own
// stream (MI355X processes get 4 hardware queues, so up to 4 streams run
// concurrently).  Each instance keeps its own matrix-free Kronecker operator
// and fused expv handle, whose steps are captured in a HIP graph, across
// solves; a solve copies p0 into p and runs the solve, with no allocation.
// The step sequence of every instance is that of the single-instance engine
// (the reference's KryExpvFSP::step, src/KryExpvFSP.cu:38-182).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "fsp_batch_impl.h"

namespace fsp {

void batch_streams_free(BatchImpl &B)
{
    for (fsp_expv e : B.sexpv)
        if (e) (void)fsp_expv_destroy(e);
    for (fsp_mat a : B.smat)
        if (a) (void)fsp_mat_destroy(a);
    for (hipStream_t s : B.wstreams)
        if (s) (void)hipStreamDestroy(s);
    B.sexpv.clear();
    B.smat.clear();
    B.wstreams.clear();
    B.parts.clear();
    B.built_workers = 0;
}

static fsp_status build_streams(BatchImpl &B, int workers)
{
    if (B.built_workers == workers) return FSP_OK;
    batch_streams_free(B);
    B.wstreams.assign((size_t)workers, nullptr);
    for (hipStream_t &s : B.wstreams) FSP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    B.parts.assign((size_t)workers, {});
    std::vector<double> load((size_t)workers, 0.0);
    for (int i : B.order) {
        const int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        B.parts[w].push_back(i);
        load[w] += B.cost[i];
    }
    B.smat.assign((size_t)B.count, nullptr);
    B.sexpv.assign((size_t)B.count, nullptr);
    for (int w = 0; w < workers; ++w) {
        for (int i : B.parts[w]) {
            const fsp_model m = B.mc[i].view();
            const BatchInst &I = B.h_inst[i];
            FSP_TRY(fsp_mat_create(&m, FSP_KRON, B.mixed ? FSP_MIXED : FSP_F64, I.with_sink,
                                   B.h_coefs.data() + B.coef_off[i], B.wstreams[w], &B.smat[i]));
            FSP_TRY(fsp_expv_create(B.smat[i], nullptr, nullptr, I.nv, I.t_final, B.m, B.tol, B.iop,
                                    B.q_iop, B.anorm, B.wstreams[w], &B.sexpv[i]));
            FSP_TRY(fsp_expv_set_limits(B.sexpv[i], B.max_reject, B.max_nstep));
        }
    }
    B.built_workers = workers;
    return FSP_OK;
}

static void run_worker(BatchImpl *B, int w, fsp_status *st, std::string *err)
{
    *st = FSP_OK;
    const hipStream_t s = B->wstreams[w];
    for (int i : B->parts[w]) {
        BatchInst &I = B->h_inst[i];
        fsp_status r = FSP_OK;
        if (hipMemcpyAsync(I.p, I.p0, sizeof(double) * (size_t)I.nv, hipMemcpyDeviceToDevice, s) !=
            hipSuccess)
            r = fail(FSP_ERR_HIP, "batch: copy of the initial vector");
        if (r == FSP_OK) r = fsp_expv_restart(B->sexpv[i], I.t_final);
        if (r == FSP_OK) r = fsp_expv_solve(B->sexpv[i], I.p);
        fsp_expv_stats es;
        std::memset(&es, 0, sizeof(es));
        if (r == FSP_OK) r = fsp_expv_get_stats(B->sexpv[i], &es);
        I.steps = es.i_step;
        I.rejects = es.n_reject;
        I.status = (int)r;
        if (r != FSP_OK && *st == FSP_OK) {
            *st = r;
            *err = "instance " + std::to_string(i) + " (" + std::to_string(I.nv) + " states, t = " +
                   std::to_string(I.t_final) + "): " + fsp_last_error();
        }
    }
}

fsp_status batch_solve_streams(BatchImpl &B)
{
    const int workers = std::max(1, std::min(B.workers, B.count));
    FSP_TRY(build_streams(B, workers));
    std::vector<fsp_status> st((size_t)workers, FSP_OK);
    std::vector<std::string> err((size_t)workers);
    std::vector<std::thread> th;
    for (int w = 1; w < workers; ++w) th.emplace_back(run_worker, &B, w, &st[w], &err[w]);
    run_worker(&B, 0, &st[0], &err[0]);
    for (std::thread &t : th) t.join();
    for (hipStream_t s : B.wstreams) FSP_HIP(hipStreamSynchronize(s));
    // per-instance statistics into the device records read by fsp_batch_get_stats
    FSP_HIP(copy_sync(B.d_inst, B.h_inst.data(), sizeof(BatchInst) * (size_t)B.count,
                      hipMemcpyHostToDevice, nullptr));
    for (int w = 0; w < workers; ++w)
        if (st[w] != FSP_OK) return fail(st[w], "batch " + err[w]);
    return FSP_OK;
}

fsp_status batch_autotune_streams(BatchImpl &B, double *best_ms, int *best_workers)
{
    const int cands[3] = {1, 2, 4};
    double best = 1e300;
    int bw = B.workers;
    for (int w : cands) {
        if (w > B.count) continue;
        B.workers = w;
        FSP_TRY(build_streams(B, w));  // handles for this partition (not timed)
        const auto t0 = std::chrono::steady_clock::now();
        FSP_TRY(batch_solve_streams(B));
        const double ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (ms < best) {
            best = ms;
            bw = w;
        }
    }
    B.workers = bw;
    *best_ms = best;
    *best_workers = bw;
    return FSP_OK;
}

}  // namespace fsp
