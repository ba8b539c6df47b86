// fsp_batch.hip -- many uneven FSP solves: set-up, path dispatch and the
// autotuner.
//
// Paths (all on the device; fsp_batch_lockstep.hip holds their kernels):
//   LOCKSTEP    one graph-captured batch step advances every running
//               instance by one Krylov step of its own;
//   PERSISTENT  one persistent work-queue kernel: each workgroup solves
//               whole instances (largest estimated cost first), the launch
//               (instance reset, queue reset, kernel, stats export) captured
//               once in a HIP graph and replayed per solve;
//   HYBRID      the persistent kernel takes the instances of at most
//               `split` states, the lockstep steps take the rest, both at
//               once on their own streams;
//   STREAMS     each instance through the single-solve engine, up to four
//               at a time (fsp_batch_streams.hip).
// The autotuner (fsp_batch_autotune) times the candidates on the device and
// caches the choice in a JSON file.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "fsp_batch_impl.h"

namespace fsp {


// Path that fsp_batch_solve takes: the requested one, else the autotuner's
// choice, else LOCKSTEP (on the 512-instance uneven batch of 2,048 to
// 262,144 states the lockstep path took 28.8 s against 28.9 s for the
// hybrid split at 4,096 states and 32.1 s at 16,384: one workgroup streams
// a mid-size instance's basis far slower than the whole device does).
static int batch_path(const BatchImpl &B)
{
    int p = B.path_req;
    if (p == FSP_BATCH_AUTO) p = B.path_tuned;
    if (p == FSP_BATCH_AUTO) p = FSP_BATCH_LOCKSTEP;
    return p;
}

static void batch_free(BatchImpl &B)
{
    batch_streams_free(B);
    batch_lockstep_free(B);
    void *ptrs[] = {B.d_models, B.d_tabs, B.d_coefs, B.d_vecs, B.d_p10, B.d_inst};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
}

// Upper estimate of the outflow rate max_x sum_r c_r a_r(x) from the tables.
static double outflow_bound(const HostModel &hm, const double *coefs)
{
    double tot = 0.0;
    for (int r = 0; r < hm.dm.R; ++r) {
        const RxnDev &rx = hm.dm.rx[r];
        if (rx.zero_nu) continue;
        double a = std::fabs(coefs[r]);
        for (int k = 0; k < rx.nsp; ++k) {
            if (rx.foff[k] < 0) continue;
            double mx = 0.0;
            for (int i = 0; i < hm.dm.radix[rx.sp[k]]; ++i)
                mx = std::max(mx, std::fabs(hm.compact[(size_t)rx.foff[k] + i]));
            a *= mx;
        }
        tot += a;
    }
    return tot;
}

}  // namespace fsp

// ===================================================================== C ABI
using fsp::BatchImpl;

struct fsp_batch_s {
    BatchImpl impl;
};

extern "C" fsp_status fsp_batch_create(const fsp_model *models, const double *coefs,
                                       const double *t_final, const int *with_sink, int count,
                                       const fsp_batch_options *opt, void *stream,
                                       fsp_batch *out)
{
    using namespace fsp;
    if (!out) return fail(FSP_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!models || !coefs || !t_final || count < 1) return fail(FSP_ERR_INVALID, "bad batch arguments");
    auto *h = new fsp_batch_s();
    BatchImpl &B = h->impl;
    auto bail = [&](fsp_status st) {
        batch_free(B);
        delete h;
        return st;
    };
    // a HIP error after the handle exists frees it and its device buffers
#define FSP_BATCH_HIP(call)                                                          \
    do {                                                                             \
        hipError_t e__ = (call);                                                     \
        if (e__ != hipSuccess)                                                       \
            return bail(fail(FSP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e__))); \
    } while (0)
    B.count = count;
    if (opt) {
        B.m = opt->m > 0 ? opt->m : 30;
        B.tol = opt->tol > 0 ? opt->tol : 1e-8;
        B.anorm = opt->anorm > 0 ? opt->anorm : 1.0;
        B.mixed = opt->precision == FSP_MIXED;
        if (opt->wg_per_cu > 0) B.pq_wg = std::min(opt->wg_per_cu, 8);
        if (opt->lockstep_bpc > 0) B.ls_bpc = opt->lockstep_bpc;
        if (opt->lockstep_rows > 0) {
            if (opt->lockstep_rows != 2 && opt->lockstep_rows != 4 && opt->lockstep_rows != 8)
                return bail(fail(FSP_ERR_INVALID, "lockstep_rows must be 2, 4 or 8"));
            B.ls_rpt = opt->lockstep_rows;
        }
        if (opt->split_states != 0) B.pq_split = opt->split_states;
        if (opt->time_limit_s > 0) B.pq_budget_s = opt->time_limit_s;
        if (opt->max_reject > 0) B.max_reject = opt->max_reject;
        if (opt->max_nstep > 0) B.max_nstep = opt->max_nstep;
        B.iop = opt->iop ? 1 : 0;
        if (opt->q_iop > 0) B.q_iop = opt->q_iop;
        if (opt->path < FSP_BATCH_AUTO || opt->path > FSP_BATCH_HYBRID)
            return bail(fail(FSP_ERR_INVALID, "unknown batch path"));
        B.path_req = opt->path;
    }
    if (B.m < 1 || B.m > kMaxKrylov) return bail(fail(FSP_ERR_UNSUPPORTED, "m must be in 1..50"));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return bail(fail(FSP_ERR_NO_DEVICE, "no HIP device visible"));
    int dev = 0;
    FSP_BATCH_HIP(hipGetDevice(&dev));
    FSP_BATCH_HIP(hipDeviceGetAttribute(&B.cus, hipDeviceAttributeMultiprocessorCount, dev));
    // models
    B.hms.resize((size_t)count);
    std::vector<double> tabs, cvec;
    std::vector<int64_t> tab_off((size_t)count), coef_off((size_t)count);
    int64_t vec_total = 0;
    B.vec_off.resize((size_t)count);
    B.h_inst.resize((size_t)count);
    std::vector<double> cost((size_t)count);
    B.dmax = 1;
    for (int i = 0; i < count; ++i) {
        fsp_status st = build_host_model(&models[i], B.hms[i]);
        if (st != FSP_OK) return bail(st);
        const HostModel &hm = B.hms[i];
        if (!hm.has_factors) return bail(fail(FSP_ERR_UNSUPPORTED, "batch needs factor tables"));
        B.dmax = std::max(B.dmax, hm.dm.d);
        tab_off[i] = (int64_t)tabs.size();
        tabs.insert(tabs.end(), hm.compact.begin(), hm.compact.end());
        coef_off[i] = (int64_t)cvec.size();
        int64_t cprev = 0;
        for (int j = 0; j < i; ++j) cprev += B.hms[j].dm.R;
        cvec.insert(cvec.end(), coefs + cprev, coefs + cprev + hm.dm.R);
        // owned copies for the stream scheduler
        {
            const fsp_model &src = models[i];
            ModelCopy mcp;
            mcp.d = src.n_species;
            mcp.R = src.n_reactions;
            mcp.bounds.assign(src.bounds, src.bounds + mcp.d);
            mcp.sr.assign(src.stoich_rowptrs, src.stoich_rowptrs + mcp.R + 1);
            const int snz = src.stoich_rowptrs[mcp.R];
            mcp.sv.assign(src.stoich_vals, src.stoich_vals + snz);
            mcp.sc.assign(src.stoich_cols, src.stoich_cols + snz);
            size_t fw = 0;
            for (int sp = 0; sp < mcp.d; ++sp) fw += (size_t)src.bounds[sp] + 1;
            mcp.factors.assign(src.factors, src.factors + fw * (size_t)mcp.R);
            B.mc.push_back(std::move(mcp));
            B.coef_off.push_back((int64_t)B.h_coefs.size());
            B.h_coefs.insert(B.h_coefs.end(), coefs + cprev, coefs + cprev + hm.dm.R);
        }
        const int sink = with_sink ? (with_sink[i] ? 1 : 0) : 0;
        const int nv = (int)hm.dm.n + sink;
        B.vec_off[i] = vec_total;
        vec_total += 2 * (int64_t)nv;  // p then p0
        B.nv_max = std::max(B.nv_max, nv);
        BatchInst &I = B.h_inst[i];
        std::memset(&I, 0, sizeof(I));
        I.t_final = t_final[i];
        I.n = (int)hm.dm.n;
        I.nv = nv;
        I.with_sink = sink;
        I.tab_size = hm.dm.compact_size;
        cost[i] = (double)nv * (1.0 + t_final[i] * outflow_bound(hm, coefs + cprev) / B.m);
    }
    FSP_BATCH_HIP(hipMalloc(&B.d_models, sizeof(DevModel) * (size_t)count));
    for (int i = 0; i < count; ++i)
        FSP_BATCH_HIP(copy_sync(B.d_models + i, &B.hms[i].dm, sizeof(DevModel), hipMemcpyHostToDevice, nullptr));
    FSP_BATCH_HIP(hipMalloc(&B.d_tabs, sizeof(double) * std::max<size_t>(tabs.size(), 1)));
    if (!tabs.empty())
        FSP_BATCH_HIP(copy_sync(B.d_tabs, tabs.data(), sizeof(double) * tabs.size(), hipMemcpyHostToDevice, nullptr));
    FSP_BATCH_HIP(hipMalloc(&B.d_coefs, sizeof(double) * cvec.size()));
    FSP_BATCH_HIP(copy_sync(B.d_coefs, cvec.data(), sizeof(double) * cvec.size(), hipMemcpyHostToDevice, nullptr));
    FSP_BATCH_HIP(hipMalloc(&B.d_vecs, sizeof(double) * (size_t)vec_total));
    FSP_BATCH_HIP(fill_sync(B.d_vecs, 0, sizeof(double) * (size_t)vec_total, nullptr));
    // default initial vectors: delta at state 0
    for (int i = 0; i < count; ++i) {
        const double one = 1.0;
        FSP_BATCH_HIP(copy_sync(B.d_vecs + B.vec_off[i] + B.h_inst[i].nv, &one, sizeof(double),
                          hipMemcpyHostToDevice, nullptr));
    }
    for (int i = 0; i < count; ++i) {
        BatchInst &I = B.h_inst[i];
        I.dm = B.d_models + i;
        I.tab = I.tab_size ? B.d_tabs + tab_off[i] : nullptr;
        I.coefs = B.d_coefs + coef_off[i];
        I.p = B.d_vecs + B.vec_off[i];
        I.p0 = I.p + I.nv;
    }
    FSP_BATCH_HIP(hipMalloc(&B.d_inst, sizeof(BatchInst) * (size_t)count));
    FSP_BATCH_HIP(copy_sync(B.d_inst, B.h_inst.data(), sizeof(BatchInst) * (size_t)count,
                      hipMemcpyHostToDevice, nullptr));
    std::vector<int> order((size_t)count);
    for (int i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    B.order = order;
    B.cost = cost;
    std::vector<double> p10(kP10Hi - kP10Lo + 1);
    for (int k = kP10Lo; k <= kP10Hi; ++k) {
        char buf[16];
        std::snprintf(buf, sizeof buf, "1e%d", k);
        p10[k - kP10Lo] = std::strtod(buf, nullptr);
    }
    FSP_BATCH_HIP(hipMalloc(&B.d_p10, sizeof(double) * p10.size()));
    FSP_BATCH_HIP(copy_sync(B.d_p10, p10.data(), sizeof(double) * p10.size(), hipMemcpyHostToDevice, nullptr));
    (void)stream;
    *out = h;
    return FSP_OK;
#undef FSP_BATCH_HIP
}

extern "C" fsp_status fsp_batch_destroy(fsp_batch b)
{
    if (!b) return FSP_OK;
    fsp::batch_free(b->impl);
    delete b;
    return FSP_OK;
}

extern "C" fsp_status fsp_batch_set_initial(fsp_batch b, int i, const double *h_p0)
{
    using namespace fsp;
    if (!b || !h_p0 || i < 0 || i >= b->impl.count) return fail(FSP_ERR_INVALID, "bad argument");
    BatchImpl &B = b->impl;
    FSP_HIP(copy_sync(B.d_vecs + B.vec_off[i] + B.h_inst[i].nv, h_p0,
                      sizeof(double) * (size_t)B.h_inst[i].nv, hipMemcpyHostToDevice, nullptr));
    return FSP_OK;
}

extern "C" fsp_status fsp_batch_solve(fsp_batch b, void *stream)
{
    using namespace fsp;
    if (!b) return fail(FSP_ERR_INVALID, "NULL batch");
    BatchImpl &B = b->impl;
    const int path = batch_path(B);
    B.solve_stream = (hipStream_t)stream;
    if (path == FSP_BATCH_PERSISTENT) return batch_launch_persistent(B, B.pq_wg, (hipStream_t)stream);
    // lockstep, hybrid and stream scheduler: synchronous, on their own streams
    FSP_HIP(hipStreamSynchronize((hipStream_t)stream));
    if (path == FSP_BATCH_LOCKSTEP) return batch_solve_lockstep(B, B.ls_bpc);
    if (path == FSP_BATCH_HYBRID) return batch_solve_hybrid(B, B.ls_bpc, B.pq_wg, B.pq_split);
    return batch_solve_streams(B);
}

extern "C" fsp_status fsp_batch_prepare(fsp_batch b)
{
    using namespace fsp;
    if (!b) return fail(FSP_ERR_INVALID, "NULL batch");
    return batch_prepare(b->impl, batch_path(b->impl));
}

extern "C" fsp_status fsp_batch_get_solution(fsp_batch b, int i, double *h_p, int64_t cap)
{
    using namespace fsp;
    if (!b || !h_p || i < 0 || i >= b->impl.count) return fail(FSP_ERR_INVALID, "bad argument");
    BatchImpl &B = b->impl;
    const int64_t nv = B.h_inst[i].nv;
    if (cap < nv) return fail(FSP_ERR_INVALID, "buffer too small");
    FSP_HIP(hipStreamSynchronize(B.solve_stream));
    FSP_HIP(copy_sync(h_p, B.d_vecs + B.vec_off[i], sizeof(double) * (size_t)nv,
                      hipMemcpyDeviceToHost, nullptr));
    return FSP_OK;
}

extern "C" fsp_status fsp_batch_get_stats(fsp_batch b, int i, int64_t *steps, int64_t *rejects,
                                          int *status)
{
    using namespace fsp;
    if (!b || i < 0 || i >= b->impl.count) return fail(FSP_ERR_INVALID, "bad argument");
    BatchImpl &B = b->impl;
    fsp::BatchInst I;
    FSP_HIP(hipStreamSynchronize(B.solve_stream));
    FSP_HIP(copy_sync(&I, B.d_inst + i, sizeof(I), hipMemcpyDeviceToHost, nullptr));
    if (steps) *steps = I.steps;
    if (rejects) *rejects = I.rejects;
    if (status) *status = I.status;
    return FSP_OK;
}

extern "C" fsp_status fsp_batch_get_config(fsp_batch b, int *block, int *wg_per_cu)
{
    using namespace fsp;
    if (!b) return fail(FSP_ERR_INVALID, "NULL batch");
    const BatchImpl &B = b->impl;
    const int path = batch_path(B);
    // persistent: (workgroup size, workgroups per CU); lockstep: (-1,
    // row-kernel workgroups per CU); hybrid: (-2, row-kernel workgroups per
    // CU); streams: (0, concurrent streams)
    int bs = 0, wg = B.workers;
    if (path == FSP_BATCH_PERSISTENT) {
        bs = kBlock;
        wg = B.pq_wg;
    } else if (path == FSP_BATCH_LOCKSTEP) {
        bs = -1;
        wg = B.ls_bpc;
    } else if (path == FSP_BATCH_HYBRID) {
        bs = -2;
        wg = B.ls_bpc;
    }
    if (block) *block = bs;
    if (wg_per_cu) *wg_per_cu = wg;
    return FSP_OK;
}

extern "C" fsp_status fsp_batch_get_tuning(fsp_batch b, int *path, int *lockstep_bpc, int *lockstep_rows,
                                           int *persistent_wg, int64_t *split_states)
{
    using namespace fsp;
    if (!b) return fail(FSP_ERR_INVALID, "NULL batch");
    const BatchImpl &B = b->impl;
    if (path) *path = batch_path(B);
    if (lockstep_bpc) *lockstep_bpc = B.ls_bpc;
    if (lockstep_rows) *lockstep_rows = B.ls_rpt;
    if (persistent_wg) *persistent_wg = B.pq_wg;
    if (split_states) *split_states = B.pq_split;
    return FSP_OK;
}

// Autotuner: times candidate configurations on the batch itself (one whole
// solve each, after one untimed solve of the path) and keeps the fastest:
//   LOCKSTEP    row-kernel workgroups per CU (one per tile, 4 or 16) x
//               tile rows per thread (2, 4 or 8: tiles of 512-2,048 rows);
//   PERSISTENT  persistent workgroups per CU: 1, 2 or 4;
//   HYBRID      the lockstep configurations as above (all instances on the
//               lockstep path), then the split at 4,096 / 16,384 / 65,536
//               states;
//   AUTO        every candidate above, and the path with them;
//   STREAMS     1, 2 or 4 concurrent streams.
// The choice is cached in a JSON file keyed by device, instance count,
// longest vector, species, m, precision and path family.
namespace {

struct Tuned {
    int path = FSP_BATCH_LOCKSTEP, bpc = 0, rpt = 4, wg = 2, workers = 4;
    long long split = 16384;
    double ms = 1e300;
};

fsp_status time_solve(fsp::BatchImpl &B, int path, int bpc, int rpt, int wg, long long split, double *ms)
{
    using namespace fsp;
    B.ls_rpt = rpt;
    auto run = [&]() -> fsp_status {
        if (path == FSP_BATCH_LOCKSTEP) return batch_solve_lockstep(B, bpc);
        if (path == FSP_BATCH_HYBRID) return batch_solve_hybrid(B, bpc, wg, split);
        return batch_solve_hybrid(B, bpc, wg, -1);  // every instance in the persistent kernel
    };
    FSP_TRY(run());  // allocations, graph capture, warm caches
    const auto t0 = std::chrono::steady_clock::now();
    FSP_TRY(run());
    *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return FSP_OK;
}

fsp_status tune_family(fsp::BatchImpl &B, int fam, Tuned &best)
{
    using namespace fsp;
    auto consider = [&](int path, int bpc, int rpt, int wg, long long split) -> fsp_status {
        double ms = 0;
        FSP_TRY(time_solve(B, path, bpc, rpt, wg, split, &ms));
        if (ms < best.ms) {
            best.path = path;
            best.bpc = bpc;
            best.rpt = rpt;
            best.wg = wg;
            best.split = split;
            best.ms = ms;
        }
        return FSP_OK;
    };
    const int bpcs[3] = {0, 4, 16};
    const int rpts[3] = {2, 4, 8};
    const int wgs[3] = {1, 2, 4};
    const long long splits[3] = {4096, 16384, 65536};
    const int rpt0 = B.ls_rpt;
    if (fam == FSP_BATCH_LOCKSTEP || fam == FSP_BATCH_HYBRID || fam == FSP_BATCH_AUTO)
        for (int rpt : rpts)
            for (int bpc : bpcs) FSP_TRY(consider(FSP_BATCH_LOCKSTEP, bpc, rpt, B.pq_wg, 0));
    if (fam == FSP_BATCH_PERSISTENT || fam == FSP_BATCH_AUTO)
        for (int wg : wgs) FSP_TRY(consider(FSP_BATCH_PERSISTENT, B.ls_bpc, rpt0, wg, -1));
    if (fam == FSP_BATCH_HYBRID || fam == FSP_BATCH_AUTO) {
        const Tuned ls = best;
        const int bpc = ls.path == FSP_BATCH_LOCKSTEP ? ls.bpc : B.ls_bpc;
        const int rpt = ls.path == FSP_BATCH_LOCKSTEP ? ls.rpt : rpt0;
        const int wg = ls.path == FSP_BATCH_PERSISTENT ? ls.wg : B.pq_wg;
        for (long long sp : splits) FSP_TRY(consider(FSP_BATCH_HYBRID, bpc, rpt, wg, sp));
    }
    return FSP_OK;
}

void apply_tuned(fsp::BatchImpl &B, const Tuned &t)
{
    B.ls_bpc = t.bpc;
    B.ls_rpt = t.rpt;
    B.pq_wg = t.wg;
    B.pq_split = t.split;
    B.workers = t.workers;
    if (B.path_req == FSP_BATCH_AUTO) B.path_tuned = t.path;
}

}  // namespace

extern "C" fsp_status fsp_batch_autotune(fsp_batch b, const char *cache_path, int force,
                                         double *best_ms)
{
    using namespace fsp;
    if (!b) return fail(FSP_ERR_INVALID, "NULL batch");
    BatchImpl &B = b->impl;
    const int fam = B.path_req;
    hipDeviceProp_t prop;
    int dev = 0;
    FSP_HIP(hipGetDevice(&dev));
    FSP_HIP(hipGetDeviceProperties(&prop, dev));
    std::ostringstream key;
    key << prop.gcnArchName << "|cu" << B.cus << "|n" << B.count << "|nv" << B.nv_max << "|d"
        << B.dmax << "|m" << B.m << "|" << (B.mixed ? "mixed" : "f64") << "|path" << fam;
    std::string cache;
    if (cache_path) {
        std::ifstream f(cache_path);
        std::stringstream ss;
        ss << f.rdbuf();
        cache = ss.str();
        const std::string k = "\"" + key.str() + "\": {";
        const size_t pos = cache.find(k);
        if (pos != std::string::npos && !force) {
            Tuned t;
            if (std::sscanf(cache.c_str() + pos + k.size(),
                            "\"path\": %d, \"bpc\": %d, \"rpt\": %d, \"wg\": %d, \"split\": %lld, "
                            "\"workers\": %d, \"ms\": %lf",
                            &t.path, &t.bpc, &t.rpt, &t.wg, &t.split, &t.workers, &t.ms) == 7) {
                apply_tuned(B, t);
                if (best_ms) *best_ms = t.ms;
                return FSP_OK;
            }
        }
    }
    Tuned best;
    best.bpc = B.ls_bpc;
    best.rpt = B.ls_rpt;
    best.wg = B.pq_wg;
    best.split = B.pq_split;
    best.workers = B.workers;
    if (fam == FSP_BATCH_STREAMS) {
        FSP_TRY(batch_autotune_streams(B, &best.ms, &best.workers));
        best.path = FSP_BATCH_STREAMS;
    } else {
        FSP_TRY(tune_family(B, fam, best));
    }
    apply_tuned(B, best);
    if (best_ms) *best_ms = best.ms;
    if (cache_path) {
        std::ostringstream ent;
        ent << "\"" << key.str() << "\": {\"path\": " << best.path << ", \"bpc\": " << best.bpc
            << ", \"rpt\": " << best.rpt << ", \"wg\": " << best.wg << ", \"split\": " << best.split << ", \"workers\": "
            << best.workers << ", \"ms\": " << best.ms << "}";
        std::string body;
        const std::string k = "\"" + key.str() + "\": {";
        const size_t pos = cache.find(k);
        if (pos != std::string::npos) {
            const size_t end = cache.find('}', pos);
            cache.replace(pos, end - pos + 1, ent.str());
            body = cache;
        } else if (cache.find('{') != std::string::npos && cache.rfind('}') != std::string::npos) {
            const size_t close = cache.rfind('}');
            const bool empty = cache.find('"') == std::string::npos;
            body = cache.substr(0, close) + (empty ? "" : ",\n ") + ent.str() + "\n}\n";
        } else {
            body = "{\n " + ent.str() + "\n}\n";
        }
        std::ofstream f(cache_path);
        f << body;
    }
    return FSP_OK;
}
