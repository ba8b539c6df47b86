"""Time-varying right-hand side (the body of the reference's CVODE callbacks,
tests/test_sundials.cu:67-91) and the native BDF integrator.

Pins: the RHS equals FSPMat::action at the frozen coefficients bit for bit;
the BDF solution of the reference's time-varying toggle coefficients
(tests/test_sundials.cu:55-64) on a tiny box against scipy's DOP853 at rtol
1e-12 (golden fixture); constant-rate solutions against scipy's dense expm
(golden fixture) and against the oracle's Krylov restatement on a larger box;
mass conservation with the sink state.  CVODE's own step sequence is not
reproducible here (parity unpinned, as the reference's test checks only
mass)."""
import os

import numpy as np
import pytest

from cuda_fsp_amd import capi, models as M
from oracle import oracle as O
from tests._util import dev, host, l1, zeros

pytestmark = pytest.mark.gpu

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "fixtures.npz"))


def _ref_test_coefs(base):
    """tests/test_sundials.cu:55-64 time function on a coefficient vector."""
    def c(t):
        out = np.array(base, float)
        out[1] = base[1] * (1.0 + np.cos(t / 1000.0))
        out[2] = base[2] * np.exp(-t / 3600.0)
        out[5] = base[5] * np.exp(-t / 3600.0)
        return out
    return c


def test_rhs_apply_equals_action_at_frozen_coefficients():
    net = M.toggle_ref_test((30, 30))
    cf = _ref_test_coefs(net.coefs)
    A = capi.Mat(net, "csr")
    rng = np.random.default_rng(4)
    x = dev(rng.random(A.n_rows))
    y1, y2 = zeros(A.n_rows), zeros(A.n_rows)
    t = 1234.5
    print(f"[case] rhs_apply toggle 31x31 t={t}", flush=True)
    capi.rhs_apply(A, t, cf, x, y1)
    A.apply(x, y2, coefs=cf(t), accumulate=False)
    assert np.array_equal(host(y1), host(y2))


@pytest.mark.parametrize("rtol,atol,bound", [(1e-6, 1e-12, 1e-7), (1e-10, 1e-16, 1e-10)])
def test_bdf_time_varying_toggle_matches_dop853(rtol, atol, bound):
    net = M.toggle_ref_test((6, 6))
    cf = _ref_test_coefs(net.coefs)
    A = capi.Mat(net, "csr", with_sink=False)
    p = dev(M.delta_p0(net))
    print(f"[case] bdf time-varying toggle 7x7 rtol={rtol}", flush=True)
    B = capi.Bdf(A, cf, rtol=rtol, atol=atol)
    for k, t in enumerate(FIX["bdf_toggle_t"]):
        B.advance(p, t)
        err = l1(host(p), FIX["bdf_toggle_p"][k])
        print(f"[bdf] t={t} L1={err:.3e} steps={B.stats().n_steps}", flush=True)
        assert err <= bound


@pytest.mark.parametrize("fmt", ["csr", "hyb", "kron"])
def test_bdf_constant_rates_match_dense_expm(fmt):
    net = M.gene_expression((2, 6, 5))
    A = capi.Mat(net, fmt, with_sink=True)
    p = dev(M.delta_p0(net, with_sink=True))
    print(f"[case] bdf gene 3x7x6 {fmt}", flush=True)
    rtol = 1e-9
    B = capi.Bdf(A, lambda t: net.coefs, rtol=rtol, atol=1e-15)
    for k, t in enumerate(FIX["cme_gene_t"]):
        B.advance(p, t)
        got = host(p)
        assert l1(got, FIX["cme_gene_p"][k]) <= 2e-8
        # 1^T A = 0, so mass changes only by the stage solves' residuals,
        # which GMRES leaves at a fraction of rtol per step
        assert abs(got.sum() - 1.0) <= 10 * rtol


def test_bdf_gene_matches_oracle_krylov_and_uses_high_order():
    net = M.gene_expression((2, 40, 40))
    A = capi.Mat(net, "csr", with_sink=True)
    p = dev(M.delta_p0(net, with_sink=True))
    print("[case] bdf gene 3x41x41 vs oracle expv", flush=True)
    B = capi.Bdf(A, lambda t: net.coefs, rtol=1e-8, atol=1e-14)
    B.advance(p, 10.0)
    st = B.stats()
    opr = O.Operator(net, "csr", with_sink=True)
    rc, out, _ = O.expv_outputs(opr, M.delta_p0(net, with_sink=True), [10.0], tol=1e-12)
    assert rc == 0
    got = host(p)
    print(f"[bdf] L1 vs expv {l1(got, out[0]):.3e} steps {st.n_steps} order {st.order} "
          f"lin iters {st.n_lin_iters}", flush=True)
    assert l1(got, out[0]) <= 2e-6
    assert abs(got.sum() - 1.0) <= 1e-10
    assert st.n_lin_fails == 0
    assert st.order >= 3


def test_bdf_continues_across_outputs_and_resets():
    net = M.birth_death((30, 20), k=3.0, g=0.8)
    A = capi.Mat(net, "hyb", with_sink=True)
    p0 = M.delta_p0(net, with_sink=True)
    p = dev(p0)
    B = capi.Bdf(A, lambda t: net.coefs, rtol=1e-8, atol=1e-14)
    B.advance(p, 1.0)
    B.advance(p, 2.0)
    two_legs = host(p)
    B.reset(0.0)
    q = dev(p0)
    B.advance(q, 2.0)
    assert l1(two_legs, host(q)) <= 1e-7
    opr = O.Operator(net, "hyb", with_sink=True)
    rc, out, _ = O.expv_outputs(opr, p0.copy(), [2.0], tol=1e-12)
    assert l1(host(q), out[0]) <= 1e-6
