"""Batched solves on the device.  On GPU devices fsp_batch_solve takes the
lockstep path by default: one graph-captured batch step advances every
running instance by one Krylov step of its own (HYB generator filled from
its model, own step sizes).  The stream scheduler (each instance through the
single-solve engine, up to four at a time) is checked as well.  (The
persistent one-workgroup-per-instance kernel is not launched on GPU devices;
its source, and the lockstep source, are checked bit for bit against the
oracle on the CPU in test_batch_simt.py.)  Checks: every instance against
the oracle's Krylov restatement with the same operator (HYB for lockstep,
Kronecker for the stream scheduler; same step counts, <= 1e-10 L1), the mixed form against FP64 (including rates that need
the second Gram-Schmidt pass), repeated solves, the autotuner and its cache,
per-instance initial vectors."""
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


@pytest.mark.parametrize("sink,path", [(False, "auto"), (True, "auto"), (True, "streams")])
def test_batch_f64_matches_oracle(sink, path):
    O.set_threads(8)
    nets, tf, ws = _small_batch(sink=sink)
    print(f"[case] batch f64 sink={sink} path={path} n={[n.n_states for n in nets][:6]}...", flush=True)
    B = capi.Batch(nets, tf, with_sink=ws, precision="f64", path=path)
    if path == "auto":
        assert B.config()[0] == -1  # lockstep
    B.solve()
    print("[batch] solved", flush=True)
    worst = 0.0
    for i, net in enumerate(nets):
        # lockstep: HYB generators; stream scheduler: Kronecker operator
        want, st = _oracle(net, tf[i], ws[i], "hyb" if path == "auto" else "kron")
        got = B.solution(i)
        steps, rej, status = B.stats(i)
        assert status == 0
        assert steps == st.i_step, (i, steps, st.i_step)
        worst = max(worst, float(np.abs(got - want).sum()))
    print(f"[batch] worst L1 vs oracle {worst:.3e}", flush=True)
    assert worst <= 1e-10


TOGGLE_CGS_RATES = (7.838583757547972, 0.5659944527754909, 13.088930136958103, 6.072545924918103)


def test_batch_mixed_vs_f64_and_replay():
    nets, tf, ws = _small_batch(count=16, seed=9)
    # rates under which one Gram-Schmidt pass over an FP32 basis degenerates
    nets = nets + [M.toggle_switch((60, 60), rates=TOGGLE_CGS_RATES)]
    tf = np.append(tf, 20.0)
    print(f"[case] batch mixed n={[n.n_states for n in nets][:6]}...", flush=True)
    B64 = capi.Batch(nets, tf, precision="f64")
    B32 = capi.Batch(nets, tf, precision="mixed")
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
    B = capi.Batch(nets, tf, precision="mixed")
    path = str(tmp_path / "tune.json")
    ms = B.autotune(path)
    assert ms > 0
    cfg = B.config()
    data = json.load(open(path))
    assert len(data) == 1
    entry = list(data.values())[0]
    # lockstep: (-1, row-kernel workgroups per CU)
    assert cfg[0] == -1 and cfg[1] in (0, 4, 16)
    assert entry["path"] == 1 and entry["wg_per_cu"] == cfg[1]
    B2 = capi.Batch(nets, tf, precision="mixed")
    assert B2.autotune(path) == pytest.approx(entry["ms"])
    assert B2.config() == cfg
    B2.solve()
    ref = capi.Batch(nets, tf, precision="mixed")
    ref.solve()
    for i in range(len(nets)):
        assert np.array_equal(B2.solution(i), ref.solution(i))


def test_batch_initial_vectors():
    net = M.birth_death((40, 30), k=3.0, g=0.8)
    p0 = np.zeros(net.n_states)
    p0[5 + 41 * 7] = 1.0
    B = capi.Batch([net, net], [1.5, 1.5])
    B.set_initial(1, p0)
    B.solve()
    opr = O.Operator(net, "hyb")
    rc, out, st = O.expv_outputs(opr, p0.copy(), [1.5])
    assert np.abs(B.solution(1) - out[0]).sum() <= 1e-10
    assert np.abs(B.solution(0) - B.solution(1)).sum() > 1e-3
