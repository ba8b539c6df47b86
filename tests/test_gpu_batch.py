"""Batched solves on the device: the lockstep path (one graph-captured
batch step advances every running instance by one Krylov step of its own),
the persistent work-queue kernel (each workgroup solves whole instances; the
launch is one captured graph), the hybrid of the two and the
stream scheduler (each instance through the single-solve engine, up to four
at a time).  Checks: every instance against the oracle's Krylov
restatement with the same operator (HYB generators on the lockstep,
persistent and hybrid paths, Kronecker for the stream scheduler; same step
counts, <= 1e-10 L1), the mixed form against FP64 (including rates that
need the second Gram-Schmidt pass), repeated solves, the autotuner and its
cache, per-instance initial vectors.  The persistent kernel runs with a
30 s wall-clock limit per workgroup, so a launch always returns."""
import json
import os

import numpy as np
import pytest

from cuda_fsp_amd import capi, models as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _small_batch(count=24, seed=5, sink=False):
    nets, tf = M.batch_instances(count, seed=seed, n_min=300, n_max=6000, t_range=(0.5, 6.0))
    return nets, tf, [sink and (i % 3 == 0) for i in range(count)]


def _oracle(net, t, sink, fmt="kron"):
    opr = O.Operator(net, fmt, with_sink=sink)
    p0 = M.delta_p0(net, with_sink=sink)
    rc, out, st = O.expv_outputs(opr, p0, [t])
    assert rc == 0
    return out[0], st


LIMIT_S = 30.0


@pytest.mark.parametrize("sink,path,rows", [(False, "lockstep", 4), (True, "lockstep", 4),
                                            (True, "lockstep", 2), (False, "lockstep", 8),
                                            (True, "streams", 0)])
def test_batch_f64_matches_oracle(sink, path, rows):
    """rows: lockstep tiles of rows x 256 rows."""
    O.set_threads(8)
    nets, tf, ws = _small_batch(sink=sink)
    print(f"[case] batch f64 sink={sink} path={path} rows={rows} n={[n.n_states for n in nets][:6]}...",
          flush=True)
    B = capi.Batch(nets, tf, with_sink=ws, precision="f64", path=path, lockstep_rows=rows)
    if path == "lockstep":
        assert B.config()[0] == -1
    B.solve()
    print("[batch] solved", flush=True)
    worst = 0.0
    for i, net in enumerate(nets):
        # lockstep: HYB generators; stream scheduler: Kronecker operator
        want, st = _oracle(net, tf[i], ws[i], "hyb" if path == "lockstep" else "kron")
        got = B.solution(i)
        steps, rej, status = B.stats(i)
        assert status == 0
        assert steps == st.i_step, (i, steps, st.i_step)
        worst = max(worst, float(np.abs(got - want).sum()))
    print(f"[batch] worst L1 vs oracle {worst:.3e}", flush=True)
    assert worst <= 1e-10


@pytest.mark.parametrize("path,sink", [("persistent", True), ("hybrid", True), ("persistent", False)])
def test_persistent_and_hybrid_match_oracle(path, sink):
    """The persistent kernel (every instance) and the hybrid split (the
    instances of at most 2,000 states persistent, the rest lockstep, on two
    streams at once): same step counts as the oracle and <= 1e-10 L1."""
    O.set_threads(8)
    nets, tf, ws = _small_batch(count=20, seed=7, sink=sink)
    print(f"[case] batch {path} sink={sink} n={[n.n_states for n in nets]}", flush=True)
    B = capi.Batch(nets, tf, with_sink=ws, precision="f64", path=path, split_states=2000,
                   time_limit_s=LIMIT_S)
    B.prepare()  # set-up (buffers, generators, graphs) ahead of the solve
    B.solve()
    print("[batch] solved", flush=True)
    worst = 0.0
    for i, net in enumerate(nets):
        want, st = _oracle(net, tf[i], ws[i], "hyb")
        steps, rej, status = B.stats(i)
        assert status == 0, (i, status)
        assert (steps, rej) == (st.i_step, st.n_reject), (i, steps, st.i_step)
        worst = max(worst, float(np.abs(B.solution(i) - want).sum()))
    print(f"[batch] worst L1 vs oracle {worst:.3e}", flush=True)
    assert worst <= 1e-10
    # a second launch of the captured graph restarts from the initial vectors
    first = [B.solution(i) for i in range(len(nets))]
    B.solve()
    for i in range(len(nets)):
        assert np.array_equal(B.solution(i), first[i])


TOGGLE_CGS_RATES = (7.838583757547972, 0.5659944527754909, 13.088930136958103, 6.072545924918103)


@pytest.mark.parametrize("path,rows", [("lockstep", 4), ("lockstep", 2), ("lockstep", 8),
                                       ("persistent", 0)])
def test_batch_mixed_vs_f64_and_replay(path, rows):
    nets, tf, ws = _small_batch(count=16, seed=9)
    # rates under which one Gram-Schmidt pass over an FP32 basis degenerates
    nets = nets + [M.toggle_switch((60, 60), rates=TOGGLE_CGS_RATES)]
    tf = np.append(tf, 20.0)
    print(f"[case] batch mixed {path} rows={rows} n={[n.n_states for n in nets][:6]}...", flush=True)
    B64 = capi.Batch(nets, tf, precision="f64", path=path, time_limit_s=LIMIT_S, lockstep_rows=rows)
    B32 = capi.Batch(nets, tf, precision="mixed", path=path, time_limit_s=LIMIT_S, lockstep_rows=rows)
    B64.solve()
    B32.solve()
    ref = [B64.solution(i) for i in range(len(nets))]
    worst = max(float(np.abs(B32.solution(i) - ref[i]).sum()) for i in range(len(nets)))
    print(f"[batch] mixed worst L1 vs f64 {worst:.3e}", flush=True)
    assert worst <= 1e-6
    # a second graph replay recomputes from the initial vectors
    B64.solve()
    for i in range(len(nets)):
        assert np.array_equal(B64.solution(i), ref[i])


def test_batch_autotune_cache(tmp_path):
    nets, tf, ws = _small_batch(count=12, seed=3)
    print(f"[case] autotune n={[n.n_states for n in nets]}", flush=True)
    B = capi.Batch(nets, tf, precision="mixed", time_limit_s=LIMIT_S)
    path = str(tmp_path / "tune.json")
    ms = B.autotune(path)
    assert ms > 0
    data = json.load(open(path))
    assert len(data) == 1
    entry = list(data.values())[0]
    name, bpc, rpt, wg, split = B.tuning()
    assert capi.BATCH_PATHS[name] == entry["path"] and entry["path"] in (1, 3, 4)
    assert (bpc, rpt, wg, split) == (entry["bpc"], entry["rpt"], entry["wg"], entry["split"])
    assert rpt in (2, 4, 8)
    print(f"[batch] tuned {B.tuning()} {ms:.1f} ms", flush=True)
    B2 = capi.Batch(nets, tf, precision="mixed", time_limit_s=LIMIT_S)
    assert B2.autotune(path) == pytest.approx(entry["ms"])
    assert B2.tuning() == B.tuning()
    B2.solve()
    ref = capi.Batch(nets, tf, precision="mixed", path=name, lockstep_bpc=bpc, wg_per_cu=wg,
                     split_states=split, time_limit_s=LIMIT_S, lockstep_rows=rpt)
    ref.solve()
    for i in range(len(nets)):
        assert np.array_equal(B2.solution(i), ref.solution(i))


def test_batch_initial_vectors():
    net = M.birth_death((40, 30), k=3.0, g=0.8)
    p0 = np.zeros(net.n_states)
    p0[5 + 41 * 7] = 1.0
    B = capi.Batch([net, net], [1.5, 1.5], time_limit_s=LIMIT_S)
    B.set_initial(1, p0)
    B.solve()
    opr = O.Operator(net, "hyb")
    rc, out, st = O.expv_outputs(opr, p0.copy(), [1.5])
    assert np.abs(B.solution(1) - out[0]).sum() <= 1e-10
    assert np.abs(B.solution(0) - B.solution(1)).sum() > 1e-3
