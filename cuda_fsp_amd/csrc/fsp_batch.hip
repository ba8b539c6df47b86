// fsp_batch.hip -- many uneven FSP solves: set-up, path dispatch, autotuner
// and the persistent-kernel path.
//
// Persistent path: each workgroup pulls instances from a work queue
// (largest estimated cost first) and runs the whole Krylov expv of that
// instance inside the workgroup: matrix-free Kronecker-sum SpMV from
// LDS-staged factor tables, fused dot products, classical Gram-Schmidt, the
// Pade expm of the Hessenberg matrix and the step control of
// KryExpvFSP::step (src/KryExpvFSP.cu:38-182), with workgroup barriers only.
// The launch (queue reset + persistent kernel) is captured once in a HIP
// graph and replayed per solve.  The lockstep path (fsp_batch_lockstep.hip)
// and the stream scheduler (fsp_batch_streams.hip) are the GPU-device paths.
// The autotuner (fsp_batch_autotune) times the candidates of the path in use
// on the device and caches the choice in a JSON file.
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


static size_t batch_lds_bytes(const BatchImpl &B, int bs)
{
    const int M2 = B.m + 2;
    return sizeof(double) * ((size_t)6 * M2 * M2 + 2 * (bs / kWave) + M2 + (size_t)B.tab_lds);
}

template <int BS, typename Bt, int NKB, int DMAX>
static fsp_status launch_batch_t(BatchImpl &B, const BatchArgs &a, int grid, hipStream_t s)
{
    const size_t lds = batch_lds_bytes(B, BS);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_batch<BS, Bt, NKB, DMAX>), dim3(grid), dim3(BS), lds, s, a);
    FSP_HIP(hipGetLastError());
    return FSP_OK;
}

template <int BS>
static fsp_status launch_batch_bs(BatchImpl &B, const BatchArgs &a, int grid, hipStream_t s)
{
    const bool big = B.m + 1 > 32;
    if (B.dmax <= 4) {
        if (B.mixed)
            return big ? launch_batch_t<BS, float, kMaxKrylov + 1, 4>(B, a, grid, s)
                       : launch_batch_t<BS, float, 32, 4>(B, a, grid, s);
        return big ? launch_batch_t<BS, double, kMaxKrylov + 1, 4>(B, a, grid, s)
                   : launch_batch_t<BS, double, 32, 4>(B, a, grid, s);
    }
    if (B.mixed)
        return big ? launch_batch_t<BS, float, kMaxKrylov + 1, 8>(B, a, grid, s)
                   : launch_batch_t<BS, float, 32, 8>(B, a, grid, s);
    return big ? launch_batch_t<BS, double, kMaxKrylov + 1, 8>(B, a, grid, s)
               : launch_batch_t<BS, double, 32, 8>(B, a, grid, s);
}

static BatchArgs batch_args(const BatchImpl &B)
{
    BatchArgs a;
    a.inst = B.d_inst;
    a.order = B.d_order;
    a.count = B.count;
    a.queue = B.d_queue;
    a.ws = B.ws;
    a.ws32 = B.ws32;
    a.ws_stride = B.ws_stride;
    a.ws32_stride = B.ws32_stride;
    a.nv_max = B.nv_max;
    a.m = B.m;
    a.iop = B.iop;
    a.q_iop = B.q_iop;
    a.max_reject = B.max_reject;
    a.tab_lds = B.tab_lds;
    a.max_nstep = B.max_nstep;
    a.tol = B.tol;
    a.anorm = B.anorm;
    a.fact = std::pow((B.m + 1) / std::exp(1.0), B.m + 1) * std::sqrt(2 * (3.1416) * (B.m + 1));
    a.p10 = B.d_p10;
    return a;
}

static fsp_status enqueue_batch(BatchImpl &B, int block, int wg_per_cu, hipStream_t s)
{
    const int grid = std::max(1, std::min<int>(B.count, B.cus * wg_per_cu));
    if ((size_t)grid > B.ws_slots) return fail(FSP_ERR_INVALID, "batch grid exceeds workspace");
    const BatchArgs a = batch_args(B);
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, s, B.d_queue);
    switch (block) {
    case 256: return launch_batch_bs<256>(B, a, grid, s);
    case 512: return launch_batch_bs<512>(B, a, grid, s);
    case 1024: return launch_batch_bs<1024>(B, a, grid, s);
    default: return fail(FSP_ERR_INVALID, "batch block must be 256, 512 or 1024");
    }
}

// The persistent kernel matches the oracle bit for bit in the CPU (SIMT)
// build of this file, but one run on a gfx950 device stalled without an
// identified cause; until that cause is found it is never launched on a GPU
// device, and batches there go through the stream scheduler instead.
static fsp_status persistent_allowed(bool *ok)
{
    int dev = 0;
    FSP_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    FSP_HIP(hipGetDeviceProperties(&prop, dev));
    *ok = std::strncmp(prop.gcnArchName, "gfx", 3) != 0;
    return FSP_OK;
}

static fsp_status batch_graph(BatchImpl &B, hipStream_t s)
{
    bool ok = false;
    FSP_TRY(persistent_allowed(&ok));
    if (!ok) return fail(FSP_ERR_UNSUPPORTED, "the persistent batch kernel is not launched on GPU devices");
    if (B.gexec && B.graph_block == B.block && B.graph_wgcu == B.wg_per_cu) {
        FSP_HIP(hipGraphLaunch(B.gexec, s));
        return FSP_OK;
    }
    if (B.gexec) (void)hipGraphExecDestroy(B.gexec);
    if (B.graph) (void)hipGraphDestroy(B.graph);
    B.gexec = nullptr;
    B.graph = nullptr;
    if (!B.cap) FSP_HIP(hipStreamCreateWithFlags(&B.cap, hipStreamNonBlocking));
    FSP_HIP(hipStreamSynchronize(s));
    FSP_HIP(hipStreamBeginCapture(B.cap, hipStreamCaptureModeThreadLocal));
    fsp_status st = enqueue_batch(B, B.block, B.wg_per_cu, B.cap);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(B.cap, &g);
    FSP_TRY(st);
    if (ce != hipSuccess) return fail(FSP_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
    B.graph = g;
    FSP_HIP(hipGraphInstantiate(&B.gexec, g, nullptr, nullptr, 0));
    B.graph_block = B.block;
    B.graph_wgcu = B.wg_per_cu;
    FSP_HIP(hipGraphLaunch(B.gexec, s));
    return FSP_OK;
}

// Path that fsp_batch_solve takes: the requested one, else the autotuner's
// choice, else LOCKSTEP on GPU devices and PERSISTENT elsewhere.
static fsp_status batch_path(const BatchImpl &B, int *path)
{
    bool persistent = false;
    FSP_TRY(persistent_allowed(&persistent));
    int p = B.path_req;
    if (p == FSP_BATCH_AUTO) p = B.path_tuned;
    if (p == FSP_BATCH_AUTO) p = persistent ? FSP_BATCH_PERSISTENT : FSP_BATCH_LOCKSTEP;
    if (p == FSP_BATCH_PERSISTENT && !persistent)
        return fail(FSP_ERR_UNSUPPORTED, "the persistent batch kernel is not launched on GPU devices");
    *path = p;
    return FSP_OK;
}

static void batch_free(BatchImpl &B)
{
    batch_streams_free(B);
    batch_lockstep_free(B);
    if (B.gexec) (void)hipGraphExecDestroy(B.gexec);
    if (B.graph) (void)hipGraphDestroy(B.graph);
    if (B.cap) (void)hipStreamDestroy(B.cap);
    void *ptrs[] = {B.d_models, B.d_tabs, B.d_coefs, B.d_vecs, B.d_p10, B.d_inst,
                    B.d_order, B.d_queue, B.ws, B.ws32};
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
        if (opt->block) B.block = opt->block;
        if (opt->wg_per_cu) B.wg_per_cu = opt->wg_per_cu;
        if (opt->max_reject > 0) B.max_reject = opt->max_reject;
        if (opt->max_nstep > 0) B.max_nstep = opt->max_nstep;
        B.iop = opt->iop ? 1 : 0;
        if (opt->q_iop > 0) B.q_iop = opt->q_iop;
        if (opt->path < FSP_BATCH_AUTO || opt->path > FSP_BATCH_PERSISTENT)
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
    int max_tab = 0;
    for (int i = 0; i < count; ++i) {
        fsp_status st = build_host_model(&models[i], B.hms[i]);
        if (st != FSP_OK) return bail(st);
        const HostModel &hm = B.hms[i];
        if (!hm.has_factors) return bail(fail(FSP_ERR_UNSUPPORTED, "batch needs factor tables"));
        B.dmax = std::max(B.dmax, hm.dm.d);
        tab_off[i] = (int64_t)tabs.size();
        tabs.insert(tabs.end(), hm.compact.begin(), hm.compact.end());
        max_tab = std::max(max_tab, hm.dm.compact_size);
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
    // LDS factor tables when they fit next to the Pade workspace
    const size_t base = batch_lds_bytes(B, 1024);
    B.tab_lds = (base + sizeof(double) * max_tab <= 150 * 1024) ? max_tab : 0;
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
    FSP_BATCH_HIP(hipMalloc(&B.d_order, sizeof(int) * (size_t)count));
    FSP_BATCH_HIP(copy_sync(B.d_order, order.data(), sizeof(int) * (size_t)count, hipMemcpyHostToDevice, nullptr));
    FSP_BATCH_HIP(hipMalloc(&B.d_queue, sizeof(unsigned int)));
    std::vector<double> p10(kP10Hi - kP10Lo + 1);
    for (int k = kP10Lo; k <= kP10Hi; ++k) {
        char buf[16];
        std::snprintf(buf, sizeof buf, "1e%d", k);
        p10[k - kP10Lo] = std::strtod(buf, nullptr);
    }
    FSP_BATCH_HIP(hipMalloc(&B.d_p10, sizeof(double) * p10.size()));
    FSP_BATCH_HIP(copy_sync(B.d_p10, p10.data(), sizeof(double) * p10.size(), hipMemcpyHostToDevice, nullptr));
    // workspace for the largest grid any configuration may use (2 per CU);
    // none where the persistent kernel is not launched
    bool persistent = false;
    {
        const fsp_status pst = persistent_allowed(&persistent);
        if (pst != FSP_OK) return bail(pst);
    }
    B.ws_slots = persistent ? (size_t)std::min(count, B.cus * 2) : 0;
    const size_t nvm = (size_t)B.nv_max;
    B.ws_stride = (B.mixed ? 3 : (size_t)(B.m + 3)) * nvm;
    B.ws32_stride = B.mixed ? (size_t)B.m * nvm : 0;
    hipError_t e = hipSuccess;
    if (B.ws_slots) e = hipMalloc(&B.ws, sizeof(double) * B.ws_stride * B.ws_slots);
    if (e != hipSuccess) return bail(fail(FSP_ERR_NOMEM, "batch workspace"));
    if (B.mixed && B.ws_slots) {
        e = hipMalloc(&B.ws32, sizeof(float) * B.ws32_stride * B.ws_slots);
        if (e != hipSuccess) return bail(fail(FSP_ERR_NOMEM, "batch workspace"));
    }
    if (B.wg_per_cu > 2) B.wg_per_cu = 2;
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
    int path = 0;
    FSP_TRY(batch_path(b->impl, &path));
    b->impl.solve_stream = (hipStream_t)stream;
    if (path == FSP_BATCH_PERSISTENT) return batch_graph(b->impl, (hipStream_t)stream);
    // lockstep and stream scheduler: synchronous, on their own streams
    FSP_HIP(hipStreamSynchronize((hipStream_t)stream));
    if (path == FSP_BATCH_LOCKSTEP) return batch_solve_lockstep(b->impl, b->impl.ls_bpc);
    return batch_solve_streams(b->impl);
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
    int path = 0;
    FSP_TRY(batch_path(b->impl, &path));
    const BatchImpl &B = b->impl;
    // persistent: (workgroup size, workgroups per CU); lockstep: (-1,
    // row-kernel workgroups per CU); streams: (0, concurrent streams)
    const int bs = path == FSP_BATCH_PERSISTENT ? B.block : (path == FSP_BATCH_LOCKSTEP ? -1 : 0);
    const int wg = path == FSP_BATCH_PERSISTENT ? B.wg_per_cu
                                                : (path == FSP_BATCH_LOCKSTEP ? B.ls_bpc : B.workers);
    if (block) *block = bs;
    if (wg_per_cu) *wg_per_cu = wg;
    return FSP_OK;
}

// Autotuner: times each candidate configuration on the batch itself and
// keeps the fastest: (workgroup size, workgroups per CU) for the persistent
// kernel, row-kernel workgroups per CU for the lockstep path, the number of
// concurrent streams for the stream scheduler.  The choice is cached in a
// JSON file keyed by device, instance count, longest vector, species, m,
// precision and path family.
static void apply_tuned(BatchImpl &B, int path, int blk, int wg, int wk)
{
    if (path == FSP_BATCH_PERSISTENT) {
        B.block = blk;
        B.wg_per_cu = wg;
    } else if (path == FSP_BATCH_LOCKSTEP) {
        B.ls_bpc = wg;
    } else {
        B.workers = wk;
    }
    if (B.path_req == FSP_BATCH_AUTO) B.path_tuned = path;
}

extern "C" fsp_status fsp_batch_autotune(fsp_batch b, const char *cache_path, int force,
                                         double *best_ms)
{
    using namespace fsp;
    if (!b) return fail(FSP_ERR_INVALID, "NULL batch");
    BatchImpl &B = b->impl;
    bool persistent = false;
    FSP_TRY(persistent_allowed(&persistent));
    // family of candidates: the requested path, or with AUTO the persistent
    // kernel where it runs and the lockstep path on GPU devices
    int fam = B.path_req;
    if (fam == FSP_BATCH_AUTO) fam = persistent ? FSP_BATCH_PERSISTENT : FSP_BATCH_LOCKSTEP;
    if (fam == FSP_BATCH_PERSISTENT && !persistent)
        return fail(FSP_ERR_UNSUPPORTED, "the persistent batch kernel is not launched on GPU devices");
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
            int path = 0, blk = 0, wg = 0, wk = 0;
            double ms = 0;
            if (std::sscanf(cache.c_str() + pos + k.size(),
                            "\"path\": %d, \"block\": %d, \"wg_per_cu\": %d, \"workers\": %d, \"ms\": %lf",
                            &path, &blk, &wg, &wk, &ms) == 5) {
                apply_tuned(B, path, blk, wg, wk);
                if (best_ms) *best_ms = ms;
                return FSP_OK;
            }
        }
    }
    double best = 1e300;
    int bb = B.block, bw = B.wg_per_cu, bk = B.workers;
    if (fam == FSP_BATCH_PERSISTENT) {
        const int blocks[3] = {256, 512, 1024};
        hipEvent_t e0, e1;
        FSP_HIP(hipEventCreate(&e0));
        FSP_HIP(hipEventCreate(&e1));
        for (int bs : blocks) {
            for (int wg = 1; wg <= 2; ++wg) {
                if (batch_lds_bytes(B, bs) * wg > 160 * 1024) continue;
                FSP_TRY(enqueue_batch(B, bs, wg, nullptr));  // warm-up
                FSP_HIP(hipEventRecord(e0, nullptr));
                FSP_TRY(enqueue_batch(B, bs, wg, nullptr));
                FSP_HIP(hipEventRecord(e1, nullptr));
                FSP_HIP(hipEventSynchronize(e1));
                float ms = 0;
                FSP_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    bb = bs;
                    bw = wg;
                }
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    } else if (fam == FSP_BATCH_LOCKSTEP) {
        const int bpcs[3] = {0, 4, 16};  // 0: one workgroup per tile
        for (int bpc : bpcs) {
            FSP_TRY(batch_lockstep_prepare(B, bpc));
            const auto t0 = std::chrono::steady_clock::now();
            FSP_TRY(batch_solve_lockstep(B, bpc));
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (ms < best) {
                best = ms;
                bw = bpc;
            }
        }
        bb = -1;
    } else {
        FSP_TRY(batch_autotune_streams(B, &best, &bk));
        bb = 0;
    }
    apply_tuned(B, fam, bb, bw, bk);
    if (best_ms) *best_ms = best;
    if (cache_path) {
        std::ostringstream ent;
        ent << "\"" << key.str() << "\": {\"path\": " << fam << ", \"block\": " << bb
            << ", \"wg_per_cu\": " << bw << ", \"workers\": " << bk << ", \"ms\": " << best << "}";
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
