"""The batched solver (cuda_fsp_amd/csrc/fsp_batch.hip and
fsp_batch_lockstep.hip, compiled unchanged for the CPU against the SIMT
stand-in in tests/simt): the persistent work-queue kernel, the hybrid split
and the lockstep path against the oracle's Krylov restatement with the HYB
operator: bit-identical solutions and step/reject counts, in both lane
orders, with and without a sink state; the mixed form within 1e-6 of FP64;
graph replay; the persistent kernel's time limit; the autotuner cache.  Each
workgroup runs as fibers on one host thread, so every instance here is
small."""
import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from cuda_fsp_amd import capi, models as M
from oracle import oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "simt"))
import simt_build  # noqa: E402


@pytest.fixture(scope="module")
def simlib():
    L = C.CDLL(simt_build.build())
    for name, (res, args) in capi._SIGS.items():
        if hasattr(L, name):
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
    return L


@pytest.fixture
def sim(simlib, monkeypatch):
    monkeypatch.setattr(capi, "_lib", simlib)
    return simlib


def _instances(k, seed=5):
    nets, tf = M.batch_instances(24, seed=seed, n_min=300, n_max=6000, t_range=(0.5, 6.0))
    idx = sorted(range(len(nets)), key=lambda i: nets[i].n_states)[:k]
    return [nets[i] for i in idx], [float(tf[i]) for i in idx]


def _oracle(net, t, sink=False, p0=None):
    # every path stores each generator in the HYB layout
    opr = O.Operator(net, "hyb", with_sink=sink)
    v = M.delta_p0(net, with_sink=sink) if p0 is None else p0.copy()
    rc, out, st = O.expv_outputs(opr, v, [t])
    assert rc == 0
    return out[0], st


@pytest.mark.parametrize("path,wg", [("persistent", 1), ("persistent", 2), ("hybrid", 2)])
def test_simt_batch_f64_bit_exact(sim, path, wg):
    nets, tf = _instances(3)
    sinks = [False, True, False]
    # hybrid: the two smaller instances in the persistent kernel, the largest lockstep
    split = sorted(n.n_states for n in nets)[1] + 1
    B = capi.Batch(nets, tf, with_sink=sinks, precision="f64", path=path, wg_per_cu=wg,
                   split_states=split)
    B.prepare()
    B.solve()
    for i, net in enumerate(nets):
        want, st = _oracle(net, tf[i], sinks[i])
        steps, rej, status = B.stats(i)
        assert status == 0
        assert (steps, rej) == (st.i_step, st.n_reject)
        assert np.array_equal(B.solution(i), want)


def test_simt_batch_mixed_and_replay(sim):
    nets, tf = _instances(2, seed=9)
    B64 = capi.Batch(nets, tf, precision="f64", path="persistent")
    B32 = capi.Batch(nets, tf, precision="mixed", path="persistent")
    B64.solve()
    B32.solve()
    ref = [B64.solution(i) for i in range(len(nets))]
    for i in range(len(nets)):
        assert np.abs(B32.solution(i) - ref[i]).sum() <= 1e-6
    B64.solve()  # a second graph replay restarts from the initial vectors
    for i in range(len(nets)):
        assert np.array_equal(B64.solution(i), ref[i])


def test_simt_batch_initial_vector_and_autotune(sim, tmp_path):
    net = M.birth_death((20, 15), k=3.0, g=0.8)
    p0 = np.zeros(net.n_states)
    p0[5 + 21 * 7] = 1.0
    B = capi.Batch([net, net], [1.5, 1.5], path="persistent")
    B.set_initial(1, p0)
    path = str(tmp_path / "tune.json")
    assert B.autotune(path) > 0
    cfg = B.config()
    entry = list(json.load(open(path)).values())[0]
    assert cfg == (256, entry["wg"]) and entry["path"] == 3 and entry["wg"] in (1, 2, 4)
    B.solve()
    want, _ = _oracle(net, 1.5, p0=p0)
    assert np.array_equal(B.solution(1), want)
    assert np.abs(B.solution(0) - B.solution(1)).sum() > 1e-3


def test_simt_persistent_time_limit(sim):
    """A workgroup past its wall-clock limit takes no further step: its
    instances end with FSP_ERR_TIMEOUT and the launch returns."""
    nets, tf = _instances(2)
    B = capi.Batch(nets, [50.0, 50.0], path="persistent", time_limit_s=1e-9)
    B.solve()
    for i in range(2):
        steps, rej, status = B.stats(i)
        assert status == 8 and steps <= 1
    H = capi.Batch(nets, [50.0, 50.0], path="hybrid", split_states=-1, time_limit_s=1e-9)
    with pytest.raises(capi.FspError, match="FSP_ERR_TIMEOUT"):
        H.solve()


def test_simt_lane_order_independent(simlib):
    """Reversed lane order between barriers gives the same bits (no result
    depends on which lane of a phase runs first), for the persistent and the
    lockstep path."""
    code = (
        "import sys, numpy as np; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "import ctypes as C, simt_build\n"
        "from cuda_fsp_amd import capi, models as M\n"
        "L = C.CDLL(simt_build.build())\n"
        "for n, (r, a) in capi._SIGS.items():\n"
        "    if hasattr(L, n): f = getattr(L, n); f.restype = r; f.argtypes = a\n"
        "capi._lib = L\n"
        "net = M.toggle_switch((14, 14), rates=M.TOGGLE_RATES)\n"
        "B = capi.Batch([net], [2.0], with_sink=[True], path='persistent'); B.solve()\n"
        "L2 = capi.Batch([net, net], [2.0, 1.0], with_sink=[True, False], path='lockstep'); L2.solve()\n"
        "np.save(sys.argv[1], np.concatenate([B.solution(0), L2.solution(0), L2.solution(1)]))\n"
    ) % (os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
         os.path.join(os.path.dirname(os.path.abspath(__file__)), "simt"))
    outs = []
    for rev in ("0", "1"):
        path = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"simt_rev{rev}_{os.getpid()}.npy")
        env = dict(os.environ, SIMT_REVERSE=rev)
        subprocess.run([sys.executable, "-c", code, path], check=True, env=env, timeout=600)
        outs.append(np.load(path))
        os.remove(path)
    assert np.array_equal(outs[0], outs[1])


# Rates under which one classical Gram-Schmidt pass over an FP32 basis loses
# orthogonality completely within 30 Arnoldi steps from a point mass (the loss
# grows about 2.5x per step from the FP32 rounding level), so the Hessenberg
# exponential overflows: the mixed form must reorthogonalize.
TOGGLE_CGS_RATES = (7.838583757547972, 0.5659944527754909, 13.088930136958103, 6.072545924918103)


def test_simt_batch_mixed_reorthogonalizes(sim):
    net = M.toggle_switch((20, 20), rates=TOGGLE_CGS_RATES)
    B64 = capi.Batch([net], [4.0], precision="f64", path="persistent")
    B32 = capi.Batch([net], [4.0], precision="mixed", path="persistent")
    B64.solve()
    B32.solve()
    assert B64.stats(0)[2] == 0 and B32.stats(0)[2] == 0
    assert np.abs(B32.solution(0) - B64.solution(0)).sum() <= 1e-6


# ---------------------------------------------------------------- lockstep
# The lockstep path (fsp_batch_lockstep.hip) in the same CPU build: every
# instance's own step sequence, bit-identical to the oracle with the HYB
# operator, in both lane orders; mixed within 1e-6 of FP64 (including the rates that need the
# second Gram-Schmidt pass); incomplete orthogonalization; the autotuner.

def _oracle_kw(net, t, sink=False, **kw):
    # the lockstep path stores each generator in the HYB layout
    opr = O.Operator(net, "hyb", with_sink=sink)
    v = M.delta_p0(net, with_sink=sink)
    rc, out, st = O.expv_outputs(opr, v, [t], **kw)
    assert rc == 0
    return out[0], st


@pytest.mark.parametrize("iop,rows", [(False, 4), (True, 4), (False, 2), (False, 8)])
def test_simt_lockstep_f64_bit_exact(sim, iop, rows):
    nets, tf = _instances(3)
    sinks = [False, True, False]
    B = capi.Batch(nets, tf, with_sink=sinks, precision="f64", path="lockstep", iop=iop, q_iop=2,
                   lockstep_rows=rows)
    assert B.config()[0] == -1
    B.solve()
    for i, net in enumerate(nets):
        want, st = _oracle_kw(net, tf[i], sinks[i], iop=iop, q_iop=2)
        steps, rej, status = B.stats(i)
        assert status == 0
        assert (steps, rej) == (st.i_step, st.n_reject), (i, steps, rej, st.i_step, st.n_reject)
        assert np.array_equal(B.solution(i), want), i
    # a second solve restarts from the initial vectors
    first = [B.solution(i) for i in range(len(nets))]
    B.solve()
    for i in range(len(nets)):
        assert np.array_equal(B.solution(i), first[i])


@pytest.mark.parametrize("rows", [2, 4, 8])
def test_simt_lockstep_mixed(sim, rows):
    nets, tf = _instances(2, seed=9)
    nets = nets + [M.toggle_switch((20, 20), rates=TOGGLE_CGS_RATES)]
    tf = tf + [4.0]
    B64 = capi.Batch(nets, tf, precision="f64", path="lockstep", lockstep_rows=rows)
    B32 = capi.Batch(nets, tf, precision="mixed", path="lockstep", lockstep_rows=rows)
    B64.solve()
    B32.solve()
    for i in range(len(nets)):
        assert B32.stats(i)[2] == 0
        assert np.abs(B32.solution(i) - B64.solution(i)).sum() <= 1e-6


def test_simt_lockstep_initial_vector_and_autotune(sim, tmp_path):
    net = M.birth_death((30, 20), k=3.0, g=0.8)
    p0 = np.zeros(net.n_states)
    p0[5 + 31 * 7] = 1.0
    B = capi.Batch([net, net], [1.5, 1.5], path="lockstep")
    B.set_initial(1, p0)
    path = str(tmp_path / "tune.json")
    ms = B.autotune(path)
    assert ms > 0
    cfg = B.config()
    entry = list(json.load(open(path)).values())[0]
    assert cfg[0] == -1 and cfg[1] in (0, 4, 16) and entry["path"] == 1 and entry["bpc"] == cfg[1]
    assert entry["rpt"] in (2, 4, 8) and B.tuning()[2] == entry["rpt"]
    B.solve()
    opr = O.Operator(net, "hyb")
    rc, out, st = O.expv_outputs(opr, p0.copy(), [1.5])
    assert np.array_equal(B.solution(1), out[0])


def _expm_cases():
    """32 x 32 upper Hessenberg matrices (the Krylov shape) over infinity
    norms that give 0 to 11 squarings, the Expokit-augmented Krylov matrix,
    and dense and smaller inputs that take the general path."""
    rng = np.random.default_rng(21)
    out = []
    for norm in (0.0, 1e-3, 0.4, 0.99, 3.0, 40.0, 900.0, 1500.0):
        H = np.triu(rng.standard_normal((32, 32)), -1)
        if norm > 0:
            H *= norm / np.abs(H).sum(axis=1).max()
        else:
            H[:] = 0.0
        out.append(H)
    # tau * [H_m, e_1 column block]: H(m+1, m) = h, H(m, m+1) = 1, H(m+1, m+2) = 1
    m = 30
    H = np.zeros((32, 32))
    H[:m + 1, :m] = np.triu(rng.standard_normal((m + 1, m)), -1)
    H[0, m] = 1.0
    H[m + 1, m] = 1.0
    out.append(5.0 * H)
    out.append(rng.standard_normal((32, 32)) * 0.3)          # dense: general path
    out.append(np.triu(rng.standard_normal((12, 12)), -1))    # N = 12
    out.append(np.triu(rng.standard_normal((32, 32)), -2))    # lower bandwidth 2
    return out


@pytest.mark.parametrize("block", [256, 128, 64])
def test_simt_expm_bit_exact(simlib, block):
    """The device Pade exponential against the oracle's orc_expm, bit for
    bit: at 256 lanes the 32 x 32 Hessenberg inputs take the blocked path,
    at 128 and 64 lanes (and for the other inputs) the general one."""
    f = simlib.simt_expm_dense
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int]
    for A in _expm_cases():
        A = np.asfortranarray(A)
        F = np.zeros_like(A, order="F")
        assert f(A.shape[0], A.ctypes.data, F.ctypes.data, block) == 0
        want = O.expm(A)
        assert np.array_equal(F.view(np.uint64), want.view(np.uint64)), (
            A.shape, np.abs(A).sum(axis=1).max(), np.abs(F - want).max())
