"""FSP engine benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A "step" is one complete FSP solve of the 3 x 512 x 512 gene-expression
network (786,432 states): Krylov expv (m = 30, tol = 1e-8) from
p0 = delta(G=0, n=0, m=0) to t = 60 with 30 equispaced output times, FP64
HYB generator (diagonal + one ELL slot per unique net change).  `value` =
solves/s with all inputs resident in HBM.

Also reported on the same line:
  roofline      the dominant kernel of that solve (the Arnoldi row kernel:
                SpMV, basis store and Gram-Schmidt dots), timed per launch
                with HIP events on its stream during one more solve;
                algorithmic bytes per launch / launch time, against 8 TB/s;
                traffic from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
                passes
  kernels       the same per-kernel figures for every kernel class of the
                solve
  spmv          every format and precision on the 16,777,216-state
                five-species network (the HBM-scale shape), HIP-event timed
  csr_solve     the headline solve with the merged-CSR generator
  mixed         the same solve with the FP32-valued / FP32-basis form:
                solves/s and worst L1 error vs FP64 over the output times
  cpu_baseline  the CPU oracle (restatement of the reference algorithm): the
                whole solve with the same 30 output stops on the host's
                cores, and a 1-core sample
  chain_mixed   the 16.7M-state chain solved to t = 20 in FP64 and in mixed
                precision: times, L1 of p and of the 1-D marginals
  bdf, adjoint  the time-varying BDF solve; one log-likelihood gradient, and
                gradients at several parameter points batched on 4 streams
  hog1p         the reference's Hog1p problem (5 species, 38.7M states)
  batch         the 512-instance uneven batch: lockstep batch path in FP64
                and mixed precision against sequential plain FP64 solves
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
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


def bench_solve(net, precision, steps, warmup, t_out, fmt="hyb", basis=None):
    A = capi.Mat(net, fmt, precision)
    n = A.n_rows
    p0 = M.delta_p0(net)
    p0_dev = capi.DeviceBuffer.from_host(p0)
    v = capi.DeviceBuffer.zeros(n)
    out = capi.DeviceBuffer.zeros(n * len(t_out))
    E = capi.Expv(A, n, t_out[0], m=30, tol=1e-8, basis=basis)
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
    # algorithmic bytes: what this format must move once per product
    # (HYB: v n diagonal + v U n slot values + 16 n for x and y; the columns
    # are implicit; merged CSR: 4 (n+1) + (v+4) nnz + 16 n; Kronecker: 16 n)
    moved = A.info.bytes_per_apply
    out = dict(format=fmt, precision=precision, n=n, nnz=int(A.info.nnz),
               us_per_apply=round(ms * 1e3, 2), algorithmic_bytes=int(moved),
               achieved_GBs=round(moved / (ms * 1e-3) / 1e9, 1),
               frac_of_peak=round(moved / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
    A.close()
    return out


def bench_expm(ev, reps=200):
    """The device Pade exponential (fsp_expm_dense: the step-control kernels'
    expm, one workgroup) on 32 x 32 upper Hessenberg matrices (the Krylov
    shape, m = 30) at several infinity norms (more squarings for larger
    norms): microseconds per call, HIP events around `reps` calls."""
    rng = np.random.default_rng(5)
    out = []
    for norm in (0.5, 10.0, 200.0):
        H = np.triu(rng.standard_normal((32, 32)), -1)
        H *= norm / np.abs(H).sum(axis=1).max()
        d_A = capi.DeviceBuffer.from_host(np.asfortranarray(H).ravel(order="F"))
        d_F = capi.DeviceBuffer.zeros(32 * 32)
        for _ in range(3):
            capi.expm_dense(d_A, d_F, 32)
        a, b = ev.event(), ev.event()
        ev.record(a)
        for _ in range(reps):
            capi.expm_dense(d_A, d_F, 32)
        ev.record(b)
        out.append({"inf_norm": norm, "squarings": max(0, math.frexp(norm)[1]),
                    "us_per_call": round(ev.elapsed_ms(a, b) / reps * 1e3, 2)})
    return out


def profile_solve(net, t_out, fmt="hyb", precision="f64"):
    """One more solve of the headline workload with per-kernel HIP-event
    timing (fsp_expv_set_profiling): per kernel class, launches, average
    launch time and algorithmic bytes per launch."""
    A = capi.Mat(net, fmt, precision)
    n = A.n_rows
    p0_dev = capi.DeviceBuffer.from_host(M.delta_p0(net))
    v = capi.DeviceBuffer.zeros(n)
    out = capi.DeviceBuffer.zeros(n * len(t_out))
    E = capi.Expv(A, n, t_out[0], m=30, tol=1e-8)
    solve_outputs(E, v, p0_dev, n, t_out, out)   # warm (graphs off below)
    E.set_profiling(True)
    t0 = time.perf_counter()
    solve_outputs(E, v, p0_dev, n, t_out, out)
    dt = time.perf_counter() - t0
    prof = E.profile()
    E.set_profiling(False)
    info = A.info
    res = {}
    for cls, (cnt, ms, b) in prof.items():
        if cnt == 0:
            continue
        us = ms / cnt * 1e3
        res[cls] = {"launches": cnt, "avg_us": round(us, 3), "total_ms": round(ms, 3),
                    "bytes_per_launch": round(b / cnt), "achieved_GBs":
                    round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 and b > 0 else None}
    E.close()
    A.close()
    return res, dt, int(info.n_slots)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(net, t_out, threads=None, one_core_outputs=1):
    """The oracle (CPU restatement of KryExpvFSP, src/KryExpvFSP.cu:38-182,
    with the HYB generator's row arithmetic) on the same solve with the same
    30 output stops, on `threads` host threads; then the first
    `one_core_outputs` output intervals on one thread, and the same
    intervals again on `threads` threads, for the per-core figure."""
    from oracle import oracle as O
    threads = threads or min(int(os.environ.get("OMP_NUM_THREADS", "0")) or
                             (os.cpu_count() or 1), 16)
    opr = O.Operator(net, "hyb")
    p0 = M.delta_p0(net)
    O.set_threads(threads)
    t0 = time.perf_counter()
    rc, out, st = O.expv_outputs(opr, p0.copy(), t_out)
    full_s = time.perf_counter() - t0
    sample = t_out[:one_core_outputs]
    O.set_threads(1)
    t0 = time.perf_counter()
    rc1, _, st1 = O.expv_outputs(opr, p0.copy(), sample)
    one_s = time.perf_counter() - t0
    O.set_threads(threads)
    t0 = time.perf_counter()
    O.expv_outputs(opr, p0.copy(), sample)
    many_s = time.perf_counter() - t0
    return {"full_s": full_s, "steps": int(st.i_step), "rc": rc, "threads": threads,
            "solution": out, "sample_t": sample[-1], "sample_steps": int(st1.i_step),
            "one_core_s": one_s, "sample_threads_s": many_s}


def bench_chain_mixed():
    """Mixed precision against FP64 at scale: the 16,777,216-state
    five-species chain (rates log-uniform in [1e-3, 10], seed 0) from
    delta(0) to t = 20, Krylov m = 30, tol = 1e-8, HYB generator in FP64 and
    in mixed precision; L1 distance of p(20) and of each 1-D marginal."""
    net = M.five_species_chain()
    n = net.n_states
    p0 = M.delta_p0(net)
    out, info = {}, {}
    # key, generator precision, Krylov basis: FP64; FP32 values and basis;
    # FP32 values with an FP64 basis
    for key, prec, basis in (("f64", "f64", None), ("mixed", "mixed", None),
                             ("mixed_values", "mixed", "f64")):
        A = capi.Mat(net, "hyb", prec)
        E = capi.Expv(A, n, 20.0, m=30, tol=1e-8, basis=basis)
        v = capi.DeviceBuffer.from_host(p0)
        capi.check(capi.lib().fsp_stream_sync(None))
        t0 = time.perf_counter()
        E.solve(v)
        capi.check(capi.lib().fsp_stream_sync(None))
        dt = time.perf_counter() - t0
        st = E.stats()
        out[key] = v.to_host()
        info[key] = {"solve_s": round(dt, 3), "krylov_steps": int(st.i_step),
                     "mass": float(out[key].sum())}
        print(f"[bench] chain t=20 {key}: {dt:.2f} s, {st.i_step} Krylov steps", flush=True)
        v.free()
        E.close()
        A.close()
    dims = [int(b) + 1 for b in net.bounds]
    cube = {k: v.reshape(dims[::-1]) for k, v in out.items()}

    def marginals(key):
        res = []
        for sp in range(len(dims)):
            ax = tuple(a for a in range(len(dims)) if a != len(dims) - 1 - sp)
            res.append(float(np.abs(cube[key].sum(axis=ax) - cube["f64"].sum(axis=ax)).sum()))
        return res

    return {"workload": f"five_species_chain {'x'.join(str(d) for d in dims)} ({n:,} states), "
                        "delta(0) to t=20, m=30 tol=1e-8, HYB generator",
            "f64": info["f64"], "mixed": info["mixed"],
            "speedup_mixed_vs_f64": round(info["f64"]["solve_s"] / info["mixed"]["solve_s"], 3),
            "l1_p_mixed_vs_f64": float(np.abs(out["mixed"] - out["f64"]).sum()),
            "l1_marginals_mixed_vs_f64": marginals("mixed"),
            "mixed_values_f64_basis": {
                **info["mixed_values"],
                "speedup_vs_f64": round(info["f64"]["solve_s"] / info["mixed_values"]["solve_s"], 3),
                "l1_p_vs_f64": float(np.abs(out["mixed_values"] - out["f64"]).sum()),
                "l1_marginals_vs_f64": marginals("mixed_values")}}


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
    seed 3) batched on 4 streams."""
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
        t4 = time.perf_counter()
        ll_b, g_b = prob.value_and_grad_batch(thetas, workers=4)
        t5 = time.perf_counter()
        print(f"[bench] {points} gradients on 4 streams {t5 - t4:.2f} s", flush=True)
        out["batch"] = {"points": points, "batched_s": round(t5 - t4, 3), "workers": 4,
                        "gradients_per_s": round(points / (t5 - t4), 4),
                        "all_finite": bool(np.isfinite(g_b).all() and np.isfinite(ll_b).all())}
    return out


def bench_hog1p(rtol=1e-2, atol=1e-8):
    """The reference's Hog1p problem (problems/hog1p_5species.cu:232-348):
    box {3, 50, 50, 60, 60} (38,713,284 states), HYB generator with the sink
    state, p0 = e_0, the signal-repressed coefficient of reaction 2, BDF to
    t = 300 s at the reference's tolerances (rtol 1e-2, atol 1e-8, :246)."""
    net = M.hog1p()
    n = net.n_states
    t0 = time.perf_counter()
    A = capi.Mat(net, "hyb", with_sink=True, coefs=M.hog1p_coefs(0.0))
    p = capi.DeviceBuffer.from_host(M.delta_p0(net, with_sink=True))
    capi.check(capi.lib().fsp_stream_sync(None))
    t1 = time.perf_counter()
    B = capi.Bdf(A, M.hog1p_coefs, rtol=rtol, atol=atol)
    B.advance(p, 300.0)
    capi.check(capi.lib().fsp_stream_sync(None))
    t2 = time.perf_counter()
    st = B.stats()
    got = p.to_host()
    B.close()
    A.close()
    p.free()
    return {"workload": f"hog1p {'x'.join(str(int(b) + 1) for b in net.bounds)} ({n:,} states), "
                        f"BDF rtol {rtol:g} atol {atol:g} to t=300, HYB generator with sink, "
                        "coefficient of reaction 2 refilled at each new time",
            "assembly_s": round(t1 - t0, 3), "solve_s": round(t2 - t1, 3),
            "steps": int(st.n_steps), "rejects": int(st.n_rejects),
            "gmres_iterations": int(st.n_lin_iters), "final_order": int(st.order),
            "mass_box": float(got[:n].sum()), "sink": float(got[n]),
            "mass_total_error": float(got.sum() - 1.0)}


def grad_only(points):
    """Only the adjoint section, with `points` batched parameter points."""
    net = M.gene_expression()
    print(json.dumps(bench_adjoint(net, points)), flush=True)


def bench_batch(count, seq_every=8, cpu_budget_s=15.0):
    """Uneven batch (toggle switches and gene-expression networks drawn
    50/50, rates within a decade of nominal, n log-uniform in [2,048,
    262,144], t_final uniform in [1, 100], seed 1), solved by the lockstep
    path in mixed precision and in FP64: set-up (buffers, HYB generators,
    tile tables, step graph) untimed, then one timed solve.  The sequential
    plain FP64 leg (the single-solve engine, HYB generator, one instance at
    a time, set-up untimed) runs on every `seq_every`-th instance (a fixed
    subset) and is scaled to the whole batch by the subset's share of the
    state-steps of the lockstep FP64 solve; the worst-instance L1 distances
    are over that subset.  The CPU leg runs the oracle (HYB row arithmetic)
    on the subset's instances in order until `cpu_budget_s` has passed and
    scales the same way."""
    nets, tf = M.batch_instances(count, seed=1)
    out = {"workload": f"{count} uneven instances (toggle / gene expression 50/50, n in "
                       f"[2048, 262144], t_final in [1, 100], seed 1), m=30 tol=1e-8",
           "count": count, "states_total": int(sum(n.n_states for n in nets))}
    lock = {}
    for prec in ("mixed", "f64"):
        B = capi.Batch(nets, tf, precision=prec, path="lockstep")
        capi.check(capi.lib().fsp_stream_sync(None))
        t0 = time.perf_counter()
        B.prepare()
        t_setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        B.solve()
        dt = time.perf_counter() - t0
        st = [B.stats(i) for i in range(count)]
        lock[prec] = {"sol": {i: B.solution(i) for i in range(0, count, seq_every)},
                      "steps": [int(x[0]) for x in st]}
        out[f"lockstep_{prec}"] = {"solve_s": round(dt, 3), "setup_s": round(t_setup, 3),
                                   "solves_per_s": round(count / dt, 3),
                                   "tuning": B.tuning(),
                                   "krylov_steps": int(sum(x[0] for x in st)),
                                   "failed": sum(1 for x in st if x[2])}
        print(f"[bench] batch {count}: lockstep {prec} {dt:.2f} s (set-up {t_setup:.2f} s)", flush=True)
        B.close()
    subset = list(range(0, count, seq_every))
    t_seq, worst = 0.0, {"mixed": 0.0, "f64": 0.0}
    steps_seq = 0
    for i in subset:
        net = nets[i]
        A = capi.Mat(net, "hyb", "f64")
        v = capi.DeviceBuffer.from_host(M.delta_p0(net))
        E = capi.Expv(A, A.n_rows, float(tf[i]), m=30, tol=1e-8)
        capi.check(capi.lib().fsp_stream_sync(None))
        t0 = time.perf_counter()
        E.solve(v)
        capi.check(capi.lib().fsp_stream_sync(None))
        t_seq += time.perf_counter() - t0
        steps_seq += int(E.stats().i_step)
        ref = v.to_host()
        for prec in worst:
            worst[prec] = max(worst[prec], float(np.abs(lock[prec]["sol"][i] - ref).sum()))
        E.close()
        A.close()
        v.free()
    # the subset's share of the batch's work (state-steps of the FP64 solve)
    work = [lock["f64"]["steps"][i] * nets[i].n_states for i in range(count)]
    share = sum(work[i] for i in subset) / sum(work)
    t_all = t_seq / share
    out["sequential_f64"] = {"engine": "single-solve FP64 HYB, one instance at a time",
                             "subset": f"instances 0, {seq_every}, {2 * seq_every}, ... "
                                       f"({len(subset)} of {count})",
                             "subset_solve_s": round(t_seq, 3), "subset_work_share":
                             round(share, 4), "krylov_steps_subset": steps_seq,
                             "estimated_solve_s": round(t_all, 3),
                             "solves_per_s": round(count / t_all, 3)}
    for prec in ("mixed", "f64"):
        o = out[f"lockstep_{prec}"]
        o["speedup_vs_sequential_f64"] = round(t_all / o["solve_s"], 3)
        o["worst_l1_vs_sequential_f64_subset"] = worst[prec]
    print(f"[bench] batch {count}: sequential FP64 subset {t_seq:.2f} s "
          f"(share {share:.3f})", flush=True)
    # CPU oracle on the subset's instances, in order, within the budget
    from oracle import oracle as O
    threads = min(int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1), 16)
    O.set_threads(threads)
    t_cpu, done = 0.0, []
    for i in subset:
        if t_cpu >= cpu_budget_s:
            break
        opr = O.Operator(nets[i], "hyb")
        t0 = time.perf_counter()
        rc, res, _ = O.expv_outputs(opr, M.delta_p0(nets[i]), [float(tf[i])])
        t_cpu += time.perf_counter() - t0
        done.append(i)
    cpu_share = sum(work[i] for i in done) / sum(work)
    out["cpu_oracle"] = {"kind": "port", "cores": threads,
                         "sample": f"instances {done[0]}, {done[1] if len(done) > 1 else ''} ... of the "
                                   f"subset ({len(done)} instances, {cpu_share:.4f} of the batch's "
                                   f"state-steps) in {t_cpu:.1f} s",
                         "estimated_solve_s": round(t_cpu / cpu_share, 1),
                         "solves_per_s": round(count * cpu_share / t_cpu, 4)}
    print(f"[bench] batch {count}: CPU oracle {len(done)} instances {t_cpu:.1f} s "
          f"(share {cpu_share:.4f})", flush=True)
    return out


def pmc_traffic(kernel, fname):
    """HBM bytes per launch of `kernel` from a committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary in profiles/ (counter passes cannot run
    inside the timed bench); None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", fname)
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
    for fmt, prec in (("csr", "f64"), ("hyb", "f64"), ("csr", "mixed"), ("hyb", "mixed"),
                      ("kron", "f64")):
        A = capi.Mat(big, fmt, prec)
        y = capi.DeviceBuffer.zeros(n)
        for _ in range(reps):
            A.apply(x, y, accumulate=False)
        capi.check(capi.lib().fsp_stream_sync(None))
        print(f"[kernels] {fmt}/{prec} bytes_per_apply={int(A.info.bytes_per_apply)}", flush=True)
        A.close()
    print(f"[kernels] calibration bytes per launch: {8 * n}", flush=True)


def solve_only():
    """Short run for hardware-counter passes on the headline solve: the
    calibration kernels (fsp_vec_fill writes 8n bytes, fsp_vec_sum reads 8n
    bytes, n = 786,432), then the FP64 HYB solve of the gene-expression
    network through its first 5 output times (t = 2, ..., 10; every Krylov
    step launches the same 31 Arnoldi row kernels, so per-launch averages
    are those of the whole solve)."""
    net = M.gene_expression()
    n = net.n_states
    x = capi.DeviceBuffer.zeros(n)
    for _ in range(5):
        capi.check(capi.lib().fsp_vec_fill(x.data_ptr(), 1.0, n, None))
    for _ in range(5):
        capi.vec_sum(x, n)
    print(f"[solve] calibration bytes per launch: {8 * n}", flush=True)
    t_out = [2.0 * k for k in range(1, 6)]
    dt, st, _ = bench_solve(net, "f64", 1, 0, t_out)
    print(f"[solve] {dt:.2f} s, {st.i_step} Krylov steps", flush=True)


def batch_only(count, precision, path="lockstep", n_max=262144):
    """One solve of the uneven batch by one path (for kernel-trace runs):
    prints the solve time, the Krylov step counts and the state-steps."""
    nets, tf = M.batch_instances(count, seed=1, n_max=n_max)
    B = capi.Batch(nets, tf, precision=precision, path=path, time_limit_s=120.0)
    t0 = time.perf_counter()
    B.solve()
    capi.check(capi.lib().fsp_stream_sync(None))
    dt = time.perf_counter() - t0
    steps = [B.stats(i)[0] for i in range(count)]
    ns = [n.n_states for n in nets]
    print(json.dumps({"batch": count, "precision": precision, "path": path, "n_max": n_max,
                      "solve_s": round(dt, 3),
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
    ap.add_argument("--side-solves", type=int, default=3,
                    help="timed solves of the merged-CSR and mixed-precision legs")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the chain, BDF, adjoint, Hog1p and batch sections")
    ap.add_argument("--batch-count", type=int, default=512,
                    help="instances in the uneven batch section")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the calibration and SpMV kernels (for counter passes)")
    ap.add_argument("--solve-only", action="store_true",
                    help="only the calibration kernels and one headline solve (counter passes)")
    ap.add_argument("--batch-only", type=int, default=0,
                    help="only one lockstep solve of an uneven batch of this many instances")
    ap.add_argument("--batch-precision", default="f64", choices=["f64", "mixed"])
    ap.add_argument("--batch-path", default="lockstep", choices=["lockstep", "persistent", "hybrid"])
    ap.add_argument("--batch-nmax", type=int, default=262144,
                    help="largest instance of the --batch-only batch (states)")
    ap.add_argument("--grad-points", type=int, default=4,
                    help="parameter points in the batched-gradient section")
    ap.add_argument("--grad-only", action="store_true",
                    help="only the adjoint-gradient section")
    ap.add_argument("--chain-only", action="store_true",
                    help="only the 16.7M-state mixed-vs-FP64 solve")
    ap.add_argument("--bdf-only", action="store_true",
                    help="only the time-varying BDF solve")
    ap.add_argument("--hog1p-only", action="store_true",
                    help="only the 38.7M-state Hog1p BDF solve")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this engine has no multi-GPU path (every GPU box "
              "here has one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.kernels_only:
        kernels_only(5)
        return
    if args.solve_only:
        solve_only()
        return
    if args.batch_only:
        batch_only(args.batch_only, args.batch_precision, args.batch_path, args.batch_nmax)
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
    if args.hog1p_only:
        print(json.dumps(bench_hog1p()), flush=True)
        return
    if capi.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    t_out = [2.0 * k for k in range(1, 31)]
    net = M.gene_expression()
    print(f"[bench] FP64 solve: {net.name} n={net.n_states}", flush=True)
    dt, st, ref = bench_solve(net, "f64", args.steps, args.warmup, t_out)
    solves_s = args.steps / dt
    print(f"[bench] FP64 {solves_s:.3f} solves/s, {st.i_step} Krylov steps/solve", flush=True)

    print("[bench] per-kernel timing of one more solve", flush=True)
    kern, dt_prof, n_slots = profile_solve(net, t_out)
    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    kname = f"k_arn_hyb<{n_slots}, double, double>" if dom == "arnoldi" else dom
    ka = kern[dom]
    traffic = pmc_traffic(kname, "gene_expression_786k_arnoldi_pmc_traffic.md")
    achieved = ka["achieved_GBs"]
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kname + " (Arnoldi row kernel of the headline solve: HYB SpMV, "
                                  "V_j store, dots V_k . w)",
                "workload": f"{net.name} n={net.n_states}, one solve (m=30, 30 outputs)",
                "launches": ka["launches"], "avg_us_per_launch": ka["avg_us"],
                "algorithmic_bytes_per_launch": ka["bytes_per_launch"],
                "algorithmic_bytes_formula":
                    "8n (diagonal) + 8Un (slot values) + 8n (x) + 8n (w_{j+1} store) + 8n "
                    "(V_j store) + 8n (j - istart + 1) (basis reads of the dots); the "
                    "avnorm launch stores no w_{j+1} and reads no basis; averaged over the "
                    "solve's launches",
                "share_of_solve_kernel_time": round(
                    ka["total_ms"] / sum(v["total_ms"] for v in kern.values()), 4),
                "traffic_source": "profiles/gene_expression_786k_arnoldi_pmc_traffic.md "
                                  "(FETCH_SIZE, WRITE_SIZE, scaled by kernels of known size)",
                "traffic_over_algorithmic": (round(traffic / ka["bytes_per_launch"], 4)
                                             if traffic else None)}
    print(f"[bench] {kname}: {ka['avg_us']:.2f} us/launch, {achieved} GB/s "
          f"({roofline['frac']:.3f} of peak)", flush=True)

    print("[bench] FP64 solve, merged CSR generator", flush=True)
    k_side = max(1, min(args.steps, args.side_solves))
    dtc, stc, refc = bench_solve(net, "f64", k_side, 1, t_out, fmt="csr")
    l1_csr = float(np.abs(refc - ref).sum(axis=1).max())

    print("[bench] mixed solve", flush=True)
    dtm, stm, mixed = bench_solve(net, "mixed", k_side, 1, t_out)
    l1_mixed = float(np.abs(mixed - ref).sum(axis=1).max())
    print(f"[bench] mixed {k_side / dtm:.3f} solves/s, worst L1 {l1_mixed:.2e}", flush=True)
    print("[bench] FP32 values, FP64 basis", flush=True)
    dtv, stv, mv = bench_solve(net, "mixed", k_side, 1, t_out, basis="f64")
    l1_mv = float(np.abs(mv - ref).sum(axis=1).max())
    print(f"[bench] FP32 values / FP64 basis {k_side / dtv:.3f} solves/s, worst L1 {l1_mv:.2e}", flush=True)

    ev = HipEvents()
    big = M.five_species_chain()
    spmv = []
    for fmt, prec in (("hyb", "f64"), ("csr", "f64"), ("kron", "f64"), ("csr", "mixed"),
                      ("hyb", "mixed"), ("kron", "mixed")):
        print(f"[bench] SpMV {fmt}/{prec} on n={big.n_states}", flush=True)
        spmv.append(bench_spmv(big, fmt, prec, args.spmv_reps, ev))

    print("[bench] device Pade exponential", flush=True)
    expm = bench_expm(ev)
    print(f"[bench] expm 32x32: {expm}", flush=True)

    cpu = None
    if not args.no_cpu:
        print("[bench] CPU oracle: the whole solve", flush=True)
        cb = cpu_baseline(net, t_out)
        l1_cpu = float(np.abs(cb["solution"] - ref).sum(axis=1).max())
        cpu = {"value": round(1.0 / cb["full_s"], 6), "unit": "solves/s",
               "cores": cb["threads"], "kind": "port",
               "sample": f"the whole solve (t = 0 to 60, the same 30 output stops, "
                         f"{cb['steps']} Krylov steps) in {cb['full_s']:.1f} s on "
                         f"{cb['threads']} OpenMP threads; HYB row arithmetic",
               "cpu_model": cpu_model(), "hardware_threads": os.cpu_count(),
               "krylov_steps": cb["steps"], "gpu_krylov_steps": int(st.i_step),
               "worst_l1_vs_gpu": l1_cpu,
               "one_core": {"sample": f"the first output interval (t = 0 to "
                                      f"{cb['sample_t']:g}, {cb['sample_steps']} Krylov steps)",
                            "s_one_thread": round(cb["one_core_s"], 3),
                            f"s_{cb['threads']}_threads": round(cb["sample_threads_s"], 3),
                            "thread_speedup": round(cb["one_core_s"] / cb["sample_threads_s"], 3)},
               "gpu_speedup": round(solves_s * cb["full_s"], 2)}
        print(f"[bench] CPU oracle {cb['full_s']:.1f} s per solve on {cb['threads']} threads",
              flush=True)

    extras = {}
    if not args.no_extras:
        print("[bench] 16.7M-state chain, mixed vs FP64", flush=True)
        extras["chain_mixed"] = bench_chain_mixed()
        print("[bench] BDF time-varying solve", flush=True)
        extras["bdf"] = bench_bdf(net)
        print("[bench] adjoint gradient", flush=True)
        extras["adjoint"] = bench_adjoint(net, args.grad_points)
        print("[bench] Hog1p 38.7M-state BDF solve", flush=True)
        extras["hog1p"] = bench_hog1p()
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
        "kernels": kern,
        "expm_dense": expm,
        "krylov_rejects_per_solve": int(st.n_reject),
        "profiled_solve_s": round(dt_prof, 3),
        "spmv": spmv,
        "csr_solve": {"solves_per_s": round(k_side / dtc, 4),
                      "krylov_steps_per_solve": int(stc.i_step), "worst_l1_vs_hyb": l1_csr},
        "mixed": {"solves_per_s": round(k_side / dtm, 4),
                  "speedup_vs_f64": round((k_side / dtm) / solves_s, 3),
                  "worst_l1_vs_f64": l1_mixed,
                  "krylov_steps_per_solve": int(stm.i_step)},
        "mixed_values_f64_basis": {"solves_per_s": round(k_side / dtv, 4),
                                   "speedup_vs_f64": round((k_side / dtv) / solves_s, 3),
                                   "worst_l1_vs_f64": l1_mv,
                                   "krylov_steps_per_solve": int(stv.i_step)},
        **extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
