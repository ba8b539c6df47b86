"""FSP engine benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A "step" is one complete FSP solve of the 3 x 512 x 512 gene-expression
network (786,432 states): Krylov expv (m = 30, tol = 1e-8) from
p0 = delta(G=0, n=0, m=0) to t = 60 with 30 equispaced output times, FP64
HYB generator (diagonal + one ELL slot per unique net change).  `value` =
solves/s with all inputs resident in HBM.

Also reported on the same line:
  roofline      HYB FP64 SpMV on the 16,777,216-state five-species network
                (the HBM-scale shape): algorithmic bytes / launch time from
                HIP events on the launch stream, against 8 TB/s
  spmv          the same for merged CSR / KRON / mixed precision
  csr_solve     the headline solve with the merged-CSR generator
  mixed         the same solve with the FP32-valued / FP32-basis form:
                solves/s and worst L1 error vs FP64 over the output times
  cpu_baseline  the CPU oracle (restatement of the reference algorithm) on a
                bounded sample of the same solve, extrapolated to solves/s
  chain_mixed   the 16.7M-state chain solved to t = 20 in FP64 and in mixed
                precision: times, L1 of p and of the 1-D marginals
  bdf, adjoint  the time-varying BDF solve; one log-likelihood gradient, and
                gradients at several parameter points one at a time and
                batched on 4 streams
  batch         an uneven batch: sequential plain FP64 solves against the
                lockstep batch path in FP64 and mixed precision
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from cuda_fsp_amd import capi, models as M  # noqa: E402

METRIC = "FSP solves/s + SpMV HBM GB/s; mixed/batched vs plain FP64 at 1e-6 L1"
HBM_PEAK_GBS = 8000.0


class HipEvents:
    """HIP events through the runtime the engine already loaded."""

    def __init__(self):
        capi.lib()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipEventCreate.argtypes = [C.c_void_p]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream=None):
        assert self.h.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = C.c_float(0)
        assert self.h.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return float(ms.value)


def solve_outputs(E, v, p0_dev, n, t_out, out_buf):
    capi.check(capi.lib().fsp_memcpy_d2d(v.data_ptr(), p0_dev.data_ptr(), n * 8, None))
    E.restart(t_out[0])
    for k, t in enumerate(t_out):
        E.set_final_time(t)
        E.solve(v)
        capi.check(capi.lib().fsp_memcpy_d2d(out_buf.offset(k * n), v.data_ptr(), n * 8, None))
    capi.check(capi.lib().fsp_stream_sync(None))


def bench_solve(net, precision, steps, warmup, t_out, fmt="hyb"):
    A = capi.Mat(net, fmt, precision)
    n = A.n_rows
    p0 = M.delta_p0(net)
    p0_dev = capi.DeviceBuffer.from_host(p0)
    v = capi.DeviceBuffer.zeros(n)
    out = capi.DeviceBuffer.zeros(n * len(t_out))
    E = capi.Expv(A, n, t_out[0], m=30, tol=1e-8)
    for _ in range(warmup):
        solve_outputs(E, v, p0_dev, n, t_out, out)
    t0 = time.perf_counter()
    for _ in range(steps):
        solve_outputs(E, v, p0_dev, n, t_out, out)
    dt = time.perf_counter() - t0
    st = E.stats()
    res = out.to_host().reshape(len(t_out), n)
    return dt, st, res


def bench_spmv(net, fmt, precision, reps, ev):
    A = capi.Mat(net, fmt, precision)
    n = A.n_rows
    x = capi.DeviceBuffer.from_host(np.random.default_rng(0).random(n))
    y = capi.DeviceBuffer.zeros(n)
    for _ in range(3):
        A.apply(x, y, accumulate=False)
    a, b = ev.event(), ev.event()
    ev.record(a)
    for _ in range(reps):
        A.apply(x, y, accumulate=False)
    ev.record(b)
    ms = ev.elapsed_ms(a, b) / reps
    moved = A.info.bytes_per_apply  # bytes this implementation must move
    bytes_ = algorithmic_bytes(fmt, precision, n, int(A.info.nnz), moved)
    out = dict(format=fmt, precision=precision, n=n, nnz=int(A.info.nnz),
               us_per_apply=round(ms * 1e3, 2), algorithmic_bytes=int(bytes_),
               achieved_GBs=round(bytes_ / (ms * 1e-3) / 1e9, 1),
               bytes_moved=int(moved), moved_GBs=round(moved / (ms * 1e-3) / 1e9, 1))
    A.close()
    return out


def algorithmic_bytes(fmt, precision, n, nnz, moved):
    """Algorithmic bytes of one product by the standard per-format count.
    HYB: v n (diagonal) + (v + 4) U n (slot values and int32 columns) + 16 n
    (x read once, y written), v = 8 (FP64) or 4 (mixed), U n = nnz - n.
    This library stores no HYB columns (they follow from the slot's net
    change), so it moves 4 U n bytes less (`moved`).  Merged CSR and
    Kronecker: the library's count, which is the standard one (mixed CSR
    adds its separate FP64 diagonal)."""
    if fmt == "hyb":
        v = 8 if precision == "f64" else 4
        return v * n + (v + 4) * (nnz - n) + 16 * n
    return moved


def cpu_baseline(net, t_out, budget_s=15.0, threads=None):
    """Oracle (CPU restatement of KryExpvFSP) on the same solve, bounded:
    the first Krylov steps until ~budget_s, extrapolated to a full solve by
    the step count of the full solve."""
    from oracle import oracle as O
    threads = threads or min(os.cpu_count() or 1, 16)
    O.set_threads(threads)
    rp, c, v = O.merged_csr(net)
    p = M.delta_p0(net)
    lib = O.lib()
    st = O.ExpvStats()

    class _Ctx(C.Structure):
        _fields_ = [("n", C.c_int64), ("rp", C.c_void_p), ("cols", C.c_void_p),
                    ("vals", C.c_void_p)]
    # step the oracle directly through its C API
    L = lib
    L.orc_expv_create.restype = C.c_void_p
    L.orc_expv_create.argtypes = [C.c_int64, C.c_double, C.c_void_p, C.c_void_p, C.c_int,
                                  C.c_double, C.c_int, C.c_int, C.c_double]
    L.orc_expv_step.argtypes = [C.c_void_p, C.c_void_p]
    L.orc_expv_destroy.argtypes = [C.c_void_p]
    L.orc_expv_get_stats.argtypes = [C.c_void_p, C.c_void_p]
    L.orc_expv_final_time_reached.argtypes = [C.c_void_p]
    MV = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p)
    n = len(rp) - 1
    L.orc_csr_apply.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                C.c_void_p, C.c_void_p]
    mv = MV(lambda ctx, x, y: L.orc_csr_apply(n, rp.ctypes.data, c.ctypes.data, v.ctypes.data,
                                              1.0, x, y))
    e = L.orc_expv_create(n, float(t_out[-1]), C.cast(mv, C.c_void_p), None, 30, 1e-8, 0, 2, 1.0)
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < budget_s and not L.orc_expv_final_time_reached(e):
        L.orc_expv_step(e, p.ctypes.data)
        steps += 1
    dt = time.perf_counter() - t0
    L.orc_expv_get_stats(e, C.byref(st))
    L.orc_expv_destroy(e)
    return dt, steps, threads


def bench_chain_mixed():
    """Mixed precision against FP64 at scale: the 16,777,216-state
    five-species chain (rates log-uniform in [1e-3, 10], seed 0) from
    delta(0) to t = 20, Krylov m = 30, tol = 1e-8, HYB generator in FP64 and
    in mixed precision; L1 distance of p(20) and of each 1-D marginal."""
    net = M.five_species_chain()
    n = net.n_states
    p0 = M.delta_p0(net)
    out, info = {}, {}
    for prec in ("f64", "mixed"):
        A = capi.Mat(net, "hyb", prec)
        E = capi.Expv(A, n, 20.0, m=30, tol=1e-8)
        v = capi.DeviceBuffer.from_host(p0)
        capi.check(capi.lib().fsp_stream_sync(None))
        t0 = time.perf_counter()
        E.solve(v)
        capi.check(capi.lib().fsp_stream_sync(None))
        dt = time.perf_counter() - t0
        st = E.stats()
        out[prec] = v.to_host()
        info[prec] = {"solve_s": round(dt, 3), "krylov_steps": int(st.i_step),
                      "mass": float(out[prec].sum())}
        print(f"[bench] chain t=20 {prec}: {dt:.2f} s, {st.i_step} Krylov steps", flush=True)
        v.free()
        E.close()
        A.close()
    dims = [int(b) + 1 for b in net.bounds]
    cube = {k: v.reshape(dims[::-1]) for k, v in out.items()}
    marg = []
    for sp in range(len(dims)):
        ax = tuple(a for a in range(len(dims)) if a != len(dims) - 1 - sp)
        marg.append(float(np.abs(cube["mixed"].sum(axis=ax) - cube["f64"].sum(axis=ax)).sum()))
    return {"workload": f"five_species_chain {'x'.join(str(d) for d in dims)} ({n:,} states), "
                        "delta(0) to t=20, m=30 tol=1e-8, HYB generator",
            "f64": info["f64"], "mixed": info["mixed"],
            "speedup_mixed_vs_f64": round(info["f64"]["solve_s"] / info["mixed"]["solve_s"], 3),
            "l1_p_mixed_vs_f64": float(np.abs(out["mixed"] - out["f64"]).sum()),
            "l1_marginals_mixed_vs_f64": marg}


def bench_bdf(net):
    """The headline network with a time-varying promoter rate, G2->G1 =
    0.5 (1 + sin(2 pi t / 30)), integrated by the native BDF (rtol 1e-6,
    atol 1e-12) through the same 30 output times to t = 60."""
    from cuda_fsp_amd import capi as _c
    base = net.coefs.copy()

    def coefs(t):
        c = base.copy()
        c[3] = 0.5 * (1.0 + np.sin(2.0 * np.pi * t / 30.0))
        return c

    A = _c.Mat(net, "hyb")
    p = _c.DeviceBuffer.from_host(M.delta_p0(net))
    B = _c.Bdf(A, coefs, rtol=1e-6, atol=1e-12)
    t0 = time.perf_counter()
    for t in [2.0 * k for k in range(1, 31)]:
        B.advance(p, t)
    dt = time.perf_counter() - t0
    st = B.stats()
    mass = float(p.to_host().sum())
    B.close()
    A.close()
    p.free()
    return {"workload": "gene_expression 3x512x512, k21(t) = 0.5(1+sin(2 pi t/30)), BDF "
                        "rtol 1e-6 atol 1e-12, 30 outputs to t=60, GMRES(30) Jacobi, HYB "
                        "generator refilled at each new time",
            "solves_per_s": round(1.0 / dt, 4), "s_per_solve": round(dt, 3),
            "steps": int(st.n_steps), "rejects": int(st.n_rejects),
            "gmres_iterations": int(st.n_lin_iters), "final_order": int(st.order),
            "mass_in_box_t60": mass}


def bench_adjoint(net, points):
    """One log-likelihood gradient by the adjoint on the headline network:
    joint (n, m) histograms of 2,000 cells at t = 5, 10, 20, 30, 45, 60
    (multinomial from the FSP marginal, seed 2), 9 log rates, HYB generator.
    Then `points` parameter points (log rates within +-0.5 of nominal,
    seed 3) one at a time and batched on 4 streams."""
    from cuda_fsp_amd import likelihood as _L
    times = [5.0, 10.0, 20.0, 30.0, 45.0, 60.0]
    keep = [0, 1, 1]
    counts = _L.sample_histograms(net, times, keep, 2000, seed=2, tol=1e-10)
    prob = _L.LikelihoodProblem(net, times, keep, counts, tol=1e-10, grading_levels=6, fmt="hyb")
    th = np.log(net.coefs)
    t0 = time.perf_counter()
    ll = prob.value(th)
    t1 = time.perf_counter()
    print(f"[bench] log-likelihood {t1 - t0:.2f} s", flush=True)
    ll2, g = prob.value_and_grad(th)
    t2 = time.perf_counter()
    print(f"[bench] gradient {t2 - t1:.2f} s", flush=True)
    out = {"workload": "gene_expression 3x512x512, 6 times x 2,000 cells, 9 log rates, "
                       "Krylov tol 1e-10, 2x6 graded Gauss-Legendre panels of 8 nodes per "
                       "interval, HYB generator",
           "loglik": ll, "loglik_s": round(t1 - t0, 3), "gradient_s": round(t2 - t1, 3),
           "gradients_per_s": round(1.0 / (t2 - t1), 4), "gradient": [round(x, 6) for x in g]}
    if points > 0:
        rng = np.random.default_rng(3)
        thetas = th[None, :] + rng.uniform(-0.5, 0.5, (points, len(th)))
        t3 = time.perf_counter()
        ll_s, g_s = prob.value_and_grad_batch(thetas, workers=1)
        t4 = time.perf_counter()
        print(f"[bench] {points} gradients one at a time {t4 - t3:.2f} s", flush=True)
        ll_b, g_b = prob.value_and_grad_batch(thetas, workers=4)
        t5 = time.perf_counter()
        print(f"[bench] {points} gradients on 4 streams {t5 - t4:.2f} s", flush=True)
        out["batch"] = {"points": points, "sequential_s": round(t4 - t3, 3),
                        "batched_s": round(t5 - t4, 3), "workers": 4,
                        "speedup": round((t4 - t3) / (t5 - t4), 3),
                        "gradients_per_s": round(points / (t5 - t4), 4),
                        "identical_to_sequential": bool(np.array_equal(g_s, g_b) and
                                                        np.array_equal(ll_s, ll_b))}
    return out


def grad_only(points):
    """Only the adjoint section, with `points` batched parameter points."""
    net = M.gene_expression()
    print(json.dumps(bench_adjoint(net, points)), flush=True)


def bench_batch(count):
    """Uneven batch (toggle switches and gene-expression networks drawn
    50/50, rates within a decade of nominal, n log-uniform in [2,048,
    262,144], t_final uniform in [1, 100], seed 1): every instance solved
    sequentially by the plain FP64 engine (merged CSR, then HYB, one solve
    at a time), then the whole batch by the lockstep path in FP64 and in
    mixed precision.  Solve times exclude assembly; speedups are against
    the merged-CSR and against the faster sequential leg; worst-instance L1
    distances are against the sequential merged-CSR solutions."""
    nets, tf = M.batch_instances(count, seed=1)
    out = {"workload": f"{count} uneven instances (toggle / gene expression 50/50, n in "
                       f"[2048, 262144], t_final in [1, 100], seed 1), m=30 tol=1e-8",
           "count": count, "states_total": int(sum(n.n_states for n in nets))}
    times = {}
    for fmt in ("csr", "hyb"):
        seq_f = []
        t_seq = 0.0
        steps_seq = 0
        for i, net in enumerate(nets):
            A = capi.Mat(net, fmt, "f64")
            v = capi.DeviceBuffer.from_host(M.delta_p0(net))
            E = capi.Expv(A, A.n_rows, float(tf[i]), m=30, tol=1e-8)
            capi.check(capi.lib().fsp_stream_sync(None))
            t0 = time.perf_counter()
            E.solve(v)
            capi.check(capi.lib().fsp_stream_sync(None))
            t_seq += time.perf_counter() - t0
            steps_seq += int(E.stats().i_step)
            seq_f.append(v.to_host())
            E.close()
            A.close()
            v.free()
        if fmt == "csr":
            seq = seq_f
        times[fmt] = t_seq
        key = "sequential_f64" if fmt == "csr" else "sequential_f64_hyb"
        out[key] = {"solve_s": round(t_seq, 3), "solves_per_s": round(count / t_seq, 3),
                    "krylov_steps": steps_seq,
                    "engine": f"single-solve FP64 {'merged CSR' if fmt == 'csr' else 'HYB'}, "
                              "one instance at a time"}
        print(f"[bench] batch {count}: sequential FP64 {fmt} {t_seq:.2f} s", flush=True)
    t_best = min(times.values())
    for prec in ("f64", "mixed"):
        B = capi.Batch(nets, tf, precision=prec, path="lockstep")
        B.solve()  # allocations, HYB fill and graph capture
        t0 = time.perf_counter()
        B.solve()
        dt = time.perf_counter() - t0
        st = [B.stats(i) for i in range(count)]
        worst = max(float(np.abs(B.solution(i) - seq[i]).sum()) for i in range(count))
        out[f"lockstep_{prec}"] = {"solve_s": round(dt, 3), "solves_per_s": round(count / dt, 3),
                                   "speedup_vs_sequential_f64": round(times["csr"] / dt, 3),
                                   "speedup_vs_fastest_sequential_f64": round(t_best / dt, 3),
                                   "krylov_steps": int(sum(x[0] for x in st)),
                                   "failed": sum(1 for x in st if x[2]),
                                   "worst_l1_vs_sequential_f64": worst}
        B.close()
    return out


def pmc_traffic(kernel="k_apply_hyb<10, double>"):
    """HBM bytes per launch of the SpMV kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (counter passes cannot run inside the
    timed bench); None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "chain_16m_spmv_pmc_traffic.md")
    kernel = kernel.replace(" ", "")
    try:
        for line in open(path):
            if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def kernels_only(reps):
    """Short run for hardware-counter passes: calibration kernels of known
    traffic (fsp_vec_fill writes 8n bytes, fsp_vec_sum reads 8n bytes) on
    n = 16,777,216, then the SpMV kernels on the five-species network."""
    big = M.five_species_chain()
    n = big.n_states
    x = capi.DeviceBuffer.zeros(n)
    for _ in range(reps):
        capi.check(capi.lib().fsp_vec_fill(x.data_ptr(), 1.0, n, None))
    for _ in range(reps):
        capi.vec_sum(x, n)
    for fmt, prec in (("csr", "f64"), ("hyb", "f64"), ("csr", "mixed"), ("hyb", "mixed")):
        A = capi.Mat(big, fmt, prec)
        y = capi.DeviceBuffer.zeros(n)
        for _ in range(reps):
            A.apply(x, y, accumulate=False)
        capi.check(capi.lib().fsp_stream_sync(None))
        print(f"[kernels] {fmt}/{prec} bytes_per_apply={int(A.info.bytes_per_apply)}", flush=True)
        A.close()
    print(f"[kernels] calibration bytes per launch: {8 * n}", flush=True)


def batch_only(count, precision):
    """One lockstep solve of the uneven batch (for kernel-trace runs):
    prints the solve time, the Krylov step counts and the state-steps."""
    nets, tf = M.batch_instances(count, seed=1)
    B = capi.Batch(nets, tf, precision=precision, path="lockstep")
    t0 = time.perf_counter()
    B.solve()
    dt = time.perf_counter() - t0
    steps = [B.stats(i)[0] for i in range(count)]
    ns = [n.n_states for n in nets]
    print(json.dumps({"batch": count, "precision": precision, "solve_s": round(dt, 3),
                      "krylov_steps": int(sum(steps)), "batch_steps": int(max(steps)),
                      "state_steps": int(sum(s * n for s, n in zip(steps, ns))),
                      "states": int(sum(ns))}), flush=True)
    B.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the BDF, adjoint-gradient and batch sections")
    ap.add_argument("--batch-count", type=int, default=64,
                    help="instances in the uneven batch section")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the calibration and SpMV kernels (for counter passes)")
    ap.add_argument("--batch-only", type=int, default=0,
                    help="only one lockstep solve of an uneven batch of this many instances")
    ap.add_argument("--batch-precision", default="f64", choices=["f64", "mixed"])
    ap.add_argument("--grad-points", type=int, default=4,
                    help="parameter points in the batched-gradient section")
    ap.add_argument("--grad-only", action="store_true",
                    help="only the adjoint-gradient section")
    ap.add_argument("--chain-only", action="store_true",
                    help="only the 16.7M-state mixed-vs-FP64 solve")
    ap.add_argument("--bdf-only", action="store_true",
                    help="only the time-varying BDF solve")
    args = ap.parse_args()
    if args.kernels_only:
        kernels_only(5)
        return
    if args.batch_only:
        batch_only(args.batch_only, args.batch_precision)
        return
    if args.grad_only:
        grad_only(args.grad_points)
        return
    if args.chain_only:
        print(json.dumps(bench_chain_mixed()), flush=True)
        return
    if args.bdf_only:
        print(json.dumps(bench_bdf(M.gene_expression())), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this engine has no multi-GPU path; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if capi.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    t_out = [2.0 * k for k in range(1, 31)]
    net = M.gene_expression()
    print(f"[bench] FP64 solve: {net.name} n={net.n_states}", flush=True)
    dt, st, ref = bench_solve(net, "f64", args.steps, args.warmup, t_out)
    solves_s = args.steps / dt
    print(f"[bench] FP64 {solves_s:.3f} solves/s, {st.i_step} Krylov steps/solve", flush=True)

    print("[bench] FP64 solve, merged CSR generator", flush=True)
    dtc, stc, refc = bench_solve(net, "f64", args.steps, args.warmup, t_out, fmt="csr")
    l1_csr = float(np.abs(refc - ref).sum(axis=1).max())

    print("[bench] mixed solve", flush=True)
    dtm, stm, mixed = bench_solve(net, "mixed", args.steps, args.warmup, t_out)
    l1_mixed = float(np.abs(mixed - ref).sum(axis=1).max())
    print(f"[bench] mixed {args.steps / dtm:.3f} solves/s, worst L1 {l1_mixed:.2e}", flush=True)

    ev = HipEvents()
    big = M.five_species_chain()
    spmv = []
    for fmt, prec in (("hyb", "f64"), ("csr", "f64"), ("kron", "f64"), ("csr", "mixed"),
                      ("hyb", "mixed"), ("kron", "mixed")):
        print(f"[bench] SpMV {fmt}/{prec} on n={big.n_states}", flush=True)
        spmv.append(bench_spmv(big, fmt, prec, args.spmv_reps, ev))
    head = spmv[0]
    traffic = pmc_traffic()
    roofline = {"bound": "hbm", "achieved": head["achieved_GBs"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(head["achieved_GBs"] / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "traffic_source": "profiles/chain_16m_spmv_pmc_traffic.md (FETCH_SIZE, WRITE_SIZE "
                                  "scaled by a read and a write kernel of known size)",
                "kernel": "k_apply_hyb<10, double> (HYB SpMV, FP64, 10 slots)",
                "workload": f"five_species_chain n={big.n_states} slots={head['nnz']}",
                "algorithmic_bytes": head["algorithmic_bytes"],
                "algorithmic_bytes_formula": "8n + 12 U n + 16n (HYB with int32 columns)",
                "bytes_moved": head["bytes_moved"],
                "bytes_moved_formula": "8n + 8 U n + 16n (columns derived, not stored)",
                "moved_GBs": head["moved_GBs"],
                "hbm_measured_GBs": (round(traffic / (head["us_per_apply"] * 1e-6) / 1e9, 1)
                                     if traffic else None),
                "hbm_measured_frac": (round(traffic / (head["us_per_apply"] * 1e-6) / 1e9
                                            / HBM_PEAK_GBS, 4) if traffic else None),
                "us_per_launch": head["us_per_apply"]}

    cpu = None
    if not args.no_cpu:
        print("[bench] CPU oracle sample", flush=True)
        cdt, csteps, threads = cpu_baseline(net, t_out)
        frac = csteps / max(st.i_step, 1)
        cpu = {"value": round(frac / cdt, 6), "unit": "solves/s", "cores": threads,
               "kind": "port",
               "sample": f"first {csteps} of {st.i_step} Krylov steps of one solve "
                         f"({cdt:.1f} s, {threads} OpenMP threads), extrapolated",
               "cpu": platform.processor() or platform.machine()}

    extras = {}
    if not args.no_extras:
        print("[bench] 16.7M-state chain, mixed vs FP64", flush=True)
        extras["chain_mixed"] = bench_chain_mixed()
        print("[bench] BDF time-varying solve", flush=True)
        extras["bdf"] = bench_bdf(net)
        print("[bench] adjoint gradient", flush=True)
        extras["adjoint"] = bench_adjoint(net, args.grad_points)
        print(f"[bench] uneven batch of {args.batch_count}", flush=True)
        extras["batch"] = bench_batch(args.batch_count)

    line = {
        "metric": METRIC, "value": round(solves_s, 4), "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "gene_expression 3x512x512 (786,432 states): Krylov expv "
                               "m=30 tol=1e-8 to t=60 at 30 output times, FP64 HYB generator",
                   "krylov_steps_per_solve": int(st.i_step),
                   "matvecs_per_solve": int(st.n_matvec)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "spmv": spmv,
        "csr_solve": {"solves_per_s": round(args.steps / dtc, 4),
                      "krylov_steps_per_solve": int(stc.i_step), "worst_l1_vs_hyb": l1_csr},
        "mixed": {"solves_per_s": round(args.steps / dtm, 4),
                  "speedup_vs_f64": round(dt / dtm, 3), "worst_l1_vs_f64": l1_mixed,
                  "krylov_steps_per_solve": int(stm.i_step)},
        **extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
