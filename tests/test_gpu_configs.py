"""GPU parity on the benchmark workloads themselves.

* The 3 x 512 x 512 gene-expression network (786,432 states) at full size:
  the FP64 Krylov solve through its 30 output times (t = 2, 4, ..., 60)
  against the oracle's restatement of KryExpvFSP::step
  (src/KryExpvFSP.cu:38-182) on the same generator arithmetic.  The oracle
  runs free (its own step control) and records its accepted step sizes: the
  two step sequences must be identical, which makes the free run and a run
  replaying the device's step sizes one and the same computation, and the
  solutions must agree to 1e-10 L1 at every output time.
* The five-species chain (G in 0..3, four counts): FP64 against the oracle
  on a reduced box, mixed precision against the oracle there (1e-6 L1 on p
  and on each 1-D marginal), and mixed against FP64 on the full
  16,777,216-state box.
* The time-varying variant of the gene-expression network (G2 -> G1 rate
  0.5 (1 + sin(2 pi t / 30)), BDF at rtol 1e-6, atol 1e-12) against a
  DOP853 solution of the dense generator (golden fixture).
"""
import os
import time

import numpy as np
import pytest

from cuda_fsp_amd import capi, models as M
from oracle import oracle as O
from tests._util import dev, host, l1

pytestmark = pytest.mark.gpu

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "fixtures.npz"))


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.name}", flush=True)


def _gpu_outputs(net, fmt, prec, p0, t_out, sink=False):
    A = capi.Mat(net, fmt, prec, with_sink=sink)
    v = dev(p0)
    E = capi.Expv(A, A.n_rows, t_out[0], m=30, tol=1e-8)
    outs = []
    for t in t_out:
        E.set_final_time(t)
        E.solve(v)
        outs.append(host(v).copy())
    st, taus = E.stats(), E.tau_history()
    E.close()
    A.close()
    return np.array(outs), st, taus


def _marginals(p, bounds):
    dims = [int(b) + 1 for b in bounds]
    cube = p[:int(np.prod(dims))].reshape(dims[::-1])
    d = len(dims)
    return [cube.sum(axis=tuple(a for a in range(d) if a != d - 1 - s)) for s in range(d)]


def _free_run_parity(net, fmt, t_out):
    """Device FP64 solve vs the oracle's free run: identical accepted step
    sizes and <= 1e-10 L1 at every output time."""
    O.set_threads(min(os.cpu_count() or 1, 16))
    p0 = M.delta_p0(net)
    t0 = time.perf_counter()
    got, st, taus = _gpu_outputs(net, fmt, "f64", p0, t_out)
    t1 = time.perf_counter()
    opr = O.Operator(net, fmt)
    O.set_verbose(True)
    try:
        rc, want, st_o, taus_o = O.expv_outputs(opr, p0.copy(), t_out, taus=True)
    finally:
        O.set_verbose(False)
    t2 = time.perf_counter()
    print(f"[parity] {net.name} n={net.n_states} {fmt}: device {t1 - t0:.1f} s, oracle "
          f"{t2 - t1:.1f} s, steps {st.i_step} / {st_o.i_step}", flush=True)
    assert rc == 0
    assert st.i_step == st_o.i_step
    assert len(taus) == len(taus_o) == st.i_step
    assert np.array_equal(taus, taus_o), int(np.argmax(taus != taus_o))
    worst = 0.0
    for k in range(len(t_out)):
        e = l1(got[k], want[k])
        worst = max(worst, e)
        assert e <= 1e-10, (k, t_out[k], e)
    print(f"[parity] worst L1 over {len(t_out)} outputs {worst:.3e}", flush=True)
    return got


@pytest.mark.timeout(900)
def test_gene_expression_full_box_matches_oracle():
    net = M.gene_expression()
    assert net.n_states == 786_432
    t_out = [2.0 * k for k in range(1, 31)]
    got = _free_run_parity(net, "hyb", t_out)
    # no sink state: 1 - sum(p) is the FSP truncation error of the box,
    # 5.7e-8 at t = 60 (nascent counts reaching 512), growing with t
    mass = [1.0 - g.sum() for g in got]
    print(f"[parity] 1 - sum(p) at t = 2, 20, 40, 60: {mass[0]:.2e} {mass[9]:.2e} "
          f"{mass[19]:.2e} {mass[29]:.2e}", flush=True)
    # the Krylov steps keep the mass to their tolerance; the deficiency is
    # the truncation error, which is below 1e-7 up to t = 60
    assert all(a >= -1e-10 for a in mass)
    assert mass[-1] <= 1e-7


CHAIN_SMALL = (3, 7, 7, 15, 15)


@pytest.mark.parametrize("fmt", ["hyb", "csr"])
def test_chain_reduced_box_f64_matches_oracle(fmt):
    net = M.five_species_chain(CHAIN_SMALL)
    assert net.n_states == 65_536
    _free_run_parity(net, fmt, [5.0, 20.0])


@pytest.mark.parametrize("fmt", ["hyb", "csr", "kron"])
def test_chain_reduced_box_mixed_matches_oracle(fmt):
    net = M.five_species_chain(CHAIN_SMALL)
    O.set_threads(min(os.cpu_count() or 1, 16))
    p0 = M.delta_p0(net)
    t_out = [5.0, 20.0]
    got, st, _ = _gpu_outputs(net, fmt, "mixed", p0, t_out)
    rc, want, st_o = O.expv_outputs(O.Operator(net, fmt), p0.copy(), t_out)
    assert rc == 0
    for k in range(len(t_out)):
        e = l1(got[k], want[k])
        em = [l1(a, b) for a, b in zip(_marginals(got[k], net.bounds),
                                       _marginals(want[k], net.bounds))]
        print(f"[mixed] chain {fmt} t={t_out[k]}: L1 {e:.3e}, marginals "
              f"{max(em):.3e}, steps {st.i_step} vs {st_o.i_step}", flush=True)
        assert e <= 1e-6
        assert max(em) <= 1e-6


@pytest.mark.timeout(600)
def test_chain_full_box_mixed_vs_f64():
    net = M.five_species_chain()
    assert net.n_states == 16_777_216
    p0 = M.delta_p0(net)
    ref, st64, _ = _gpu_outputs(net, "hyb", "f64", p0, [20.0])
    got, st32, _ = _gpu_outputs(net, "hyb", "mixed", p0, [20.0])
    e = l1(got[0], ref[0])
    em = [l1(a, b) for a, b in zip(_marginals(got[0], net.bounds), _marginals(ref[0], net.bounds))]
    print(f"[mixed] chain 16.7M t=20: L1 {e:.3e}, marginals {['%.2e' % x for x in em]}, steps "
          f"{st32.i_step} vs {st64.i_step}", flush=True)
    assert np.isfinite(got[0]).all()
    assert e <= 1e-6
    assert max(em) <= 1e-6


def _gene_sin_coefs(base):
    def c(t):
        out = np.array(base, float)
        out[3] = 0.5 * (1.0 + np.sin(2.0 * np.pi * t / 30.0))
        return out
    return c


@pytest.mark.parametrize("fmt", ["hyb", "csr"])
@pytest.mark.parametrize("rtol,atol,bound", [(1e-6, 1e-12, 1e-4), (1e-10, 1e-16, 1e-8)])
def test_bdf_gene_sin_matches_dop853(fmt, rtol, atol, bound):
    net = M.gene_expression((2, 12, 12))
    A = capi.Mat(net, fmt, with_sink=True)
    p = dev(M.delta_p0(net, with_sink=True))
    B = capi.Bdf(A, _gene_sin_coefs(net.coefs), rtol=rtol, atol=atol)
    for k, t in enumerate(FIX["bdf_gene_sin_t"]):
        B.advance(p, t)
        got = host(p)
        err = l1(got, FIX["bdf_gene_sin_p"][k])
        print(f"[bdf] gene sin {fmt} rtol={rtol} t={t}: L1 {err:.3e}, mass "
              f"{got.sum():.15f}, steps {B.stats().n_steps}", flush=True)
        assert err <= bound
        assert abs(got.sum() - 1.0) <= 10 * rtol
