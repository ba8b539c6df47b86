"""Batched solves on the device.  On GPU devices fsp_batch_solve runs the
stream scheduler: every instance through the single-instance fused Krylov
engine (Kronecker operator, graph-captured steps), up to four instances at a
time on their own streams.  (The persistent one-workgroup-per-instance kernel
is not launched on GPU devices; its source is checked bit for bit against the
oracle on the CPU in test_batch_simt.py.)  Checks: every instance against the
oracle's Krylov restatement with the Kronecker operator (same step counts,
<= 1e-10 L1), the mixed form against FP64, repeated solves, the autotuner and
its cache, per-instance initial vectors."""
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


def _oracle(net, t, sink):
    opr = O.Operator(net, "kron", with_sink=sink)
    p0 = M.delta_p0(net, with_sink=sink)
    rc, out, st = O.expv_outputs(opr, p0, [t])
    assert rc == 0
    return out[0], st


@pytest.mark.parametrize("sink", [False, True])
def test_batch_f64_matches_oracle(sink):
    O.set_threads(8)
    nets, tf, ws = _small_batch(sink=sink)
    print(f"[case] batch f64 sink={sink} n={[n.n_states for n in nets][:6]}...", flush=True)
    B = capi.Batch(nets, tf, with_sink=ws, precision="f64")
    B.solve()
    print("[batch] solved", flush=True)
    worst = 0.0
    for i, net in enumerate(nets):
        want, st = _oracle(net, tf[i], ws[i])
        got = B.solution(i)
        steps, rej, status = B.stats(i)
        assert status == 0
        assert steps == st.i_step, (i, steps, st.i_step)
        worst = max(worst, float(np.abs(got - want).sum()))
    print(f"[batch] worst L1 vs oracle {worst:.3e}", flush=True)
    assert worst <= 1e-10


def test_batch_mixed_vs_f64_and_replay():
    nets, tf, ws = _small_batch(count=16, seed=9)
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
    if cfg[0] == 0:  # stream scheduler: (0, concurrent streams)
        assert cfg[1] in (1, 2, 4) and entry["workers"] == cfg[1]
    else:
        assert cfg[0] in (256, 512, 1024) and cfg[1] in (1, 2)
        assert (entry["block"], entry["wg_per_cu"]) == cfg
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
    opr = O.Operator(net, "kron")
    rc, out, st = O.expv_outputs(opr, p0.copy(), [1.5])
    assert np.abs(B.solution(1) - out[0]).sum() <= 1e-10
    assert np.abs(B.solution(0) - B.solution(1)).sum() > 1e-3
