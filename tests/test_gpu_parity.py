"""GPU parity: HIP path (through the C ABI) against the CPU oracle.

Integer work (states, CSR structure) and the assembled generator values are
compared bit for bit; SpMV rows are bit-exact for the CSR/HYB/KRON formats
(same operation order as the oracle); Krylov solutions to 1e-10 L1.
"""
import numpy as np
import pytest

from cuda_fsp_amd import capi, models as M
from oracle import oracle as O
from tests._util import dev, host, zeros, l1

pytestmark = pytest.mark.gpu

SMALL_NETS = [
    lambda: M.toggle_ref_test((31, 31)),
    lambda: M.toggle_switch((127, 127)),
    lambda: M.gene_expression((2, 40, 35)),
    lambda: M.five_species_chain((3, 5, 6, 4, 7)),
    lambda: M.birth_death((9, 0, 4)),
]


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.name}", flush=True)


def test_states_enumerate_reference_box():
    # tests/test_fsp_get_states.cu:21-54: d = 3, bounds (1023, 1023, 3)
    bounds = (1023, 1023, 3)
    n = capi.num_states(bounds)
    assert n == 4_194_304
    d_states = zeros(n * 3, np.int32)
    capi.states_enumerate(bounds, d_states)
    got = host(d_states).reshape(n, 3)
    x2, x1, x0 = np.meshgrid(np.arange(4), np.arange(1024), np.arange(1024), indexing="ij")
    want = np.stack([x0.ravel(), x1.ravel(), x2.ravel()], axis=1).astype(np.int32)
    assert np.array_equal(got, want)
    assert np.array_equal(got, O.enumerate_states(bounds))
    # state -> index round trip, plus out-of-box states
    idx = zeros(n, np.int32)
    capi.states_to_index(bounds, d_states, n, idx)
    assert np.array_equal(host(idx), np.arange(n, dtype=np.int32))
    bad = dev(np.array([[-1, 0, 0], [1024, 0, 0], [0, 0, 4], [5, 6, 2]], np.int32))
    out = zeros(4, np.int32)
    capi.states_to_index(bounds, bad, 4, out)
    assert host(out).tolist() == [-1, -1, -1, 5 + 1024 * 6 + 1024 * 1024 * 2]


@pytest.mark.parametrize("mk", SMALL_NETS)
def test_term_csr_bit_exact(mk):
    net = mk()
    n = net.n_states
    for r in range(net.n_reactions):
        rp_o, c_o, v_o = O.term_csr(net, r)
        d_rp = zeros(n + 1, np.int32)
        nnz = capi.term_csr_nnz(net, r, d_rp)
        assert nnz == len(c_o)
        d_c, d_v = zeros(nnz, np.int32), zeros(nnz)
        capi.term_csr(net, r, d_rp, d_c, d_v)
        assert np.array_equal(host(d_rp), rp_o)
        assert np.array_equal(host(d_c), c_o)
        assert np.array_equal(host(d_v).view(np.uint64), v_o.view(np.uint64))


@pytest.mark.parametrize("sink", [False, True])
@pytest.mark.parametrize("mk", SMALL_NETS)
def test_merged_csr_bit_exact(mk, sink):
    net = mk()
    A = capi.Mat(net, "csr", "f64", with_sink=sink)
    rp_o, c_o, v_o = O.merged_csr(net, with_sink=sink)
    assert A.info.nnz == len(c_o)
    rows = A.n_rows
    d_rp, d_c, d_v = zeros(rows + 1, np.int32), zeros(len(c_o), np.int32), zeros(len(c_o))
    A.export_csr(d_rp, d_c, d_v)
    assert np.array_equal(host(d_rp), rp_o)
    assert np.array_equal(host(d_c), c_o)
    assert np.array_equal(host(d_v).view(np.uint64), v_o.view(np.uint64))


@pytest.mark.parametrize("mk", SMALL_NETS)
def test_hyb_bit_exact(mk):
    net = mk()
    A = capi.Mat(net, "hyb", "f64", with_sink=True)
    U, dg, sv, sc, sk = O.ell(net, with_sink=True)
    n = net.n_states
    d_dg, d_sv, d_sc, d_sk = zeros(n), zeros(max(U, 1) * n), zeros(max(U, 1) * n, np.int32), zeros(n)
    A.export_hyb(d_dg, d_sv, d_sc, d_sk)
    assert A.info.n_slots == U
    assert np.array_equal(host(d_dg).view(np.uint64), dg.view(np.uint64))
    assert np.array_equal(host(d_sv)[:U * n].view(np.uint64), sv.ravel().view(np.uint64))
    assert np.array_equal(host(d_sc)[:U * n], sc.ravel())
    assert np.array_equal(host(d_sk).view(np.uint64), sk.view(np.uint64))


def _apply(A, x, accumulate=False, y0=None):
    y = dev(y0) if y0 is not None else zeros(A.n_rows)
    A.apply(dev(x), y, accumulate=accumulate)
    return host(y)


@pytest.mark.parametrize("mk", SMALL_NETS)
def test_spmv_matches_oracle(mk):
    net = mk()
    n = net.n_states
    rng = np.random.default_rng(7)
    x = rng.random(n + 1)
    y0 = rng.random(n + 1)
    # CSR: same per-row order as the oracle -> bit-exact rows, sink to 1e-15
    rp, c, v = O.merged_csr(net, with_sink=True)
    want = y0.copy()
    O.csr_apply(rp, c, v, x, want)
    for fmt in ("csr", "hyb", "kron"):
        A = capi.Mat(net, fmt, "f64", with_sink=True)
        got = _apply(A, x, accumulate=True, y0=y0)
        if fmt == "csr":
            assert np.array_equal(got[:n], want[:n]), fmt
        else:
            assert np.abs(got[:n] - want[:n]).max() <= 1e-13 * np.abs(want[:n]).max(), fmt
        assert abs(got[n] - want[n]) <= 1e-14 * max(1.0, abs(want[n])), fmt
    # HYB and KRON bit-exact against their own oracle restatements
    U, dg, sv, sc, sk = O.ell(net, with_sink=False)
    w = np.zeros(n)
    O.ell_apply(U, dg, sv, sc, None, x[:n], w)
    got = _apply(capi.Mat(net, "hyb", "f64"), x[:n])
    assert np.array_equal(got, w)
    w = np.zeros(n)
    O.kron_apply_device_form(net, x[:n], w)
    got = _apply(capi.Mat(net, "kron", "f64"), x[:n])
    assert np.array_equal(got, w)
    # with the sink state: box rows bit-exact, the sink to 1e-14
    w = y0.copy()
    O.kron_apply_device_form(net, x, w, with_sink=True)
    got = _apply(capi.Mat(net, "kron", "f64", with_sink=True), x, accumulate=True, y0=y0)
    assert np.array_equal(got[:n], w[:n])
    assert abs(got[n] - w[n]) <= 1e-14 * max(1.0, abs(w[n]))


def _connected_rows(rp, c, n, k):
    rows = np.repeat(np.arange(n), np.diff(rp))
    out = np.zeros(n, bool)
    out[rows[c == k]] = True
    return out


@pytest.mark.parametrize("fmt", ["csr", "hyb", "kron"])
def test_nonfinite_entry_reaches_only_connected_rows(fmt):
    """A NaN in x at state k reaches row i of A x only where the stored
    generator has an entry at (i, k) (the merged CSR structure, explicit
    zeros included): the HYB layout's implicit columns of out-of-box
    neighbours, which name other states, carry nothing."""
    net = M.gene_expression((2, 20, 25))
    n = net.n_states
    rp, c, _ = O.merged_csr(net)
    rng = np.random.default_rng(11)
    A = capi.Mat(net, fmt, "f64")
    for k in [0, 1, 2, 3, 62, n - 1] + rng.choice(n, 6, replace=False).tolist():
        x = rng.random(n)
        x[k] = np.nan
        bad = ~np.isfinite(_apply(A, x))
        assert bad[k], (fmt, k)
        assert not (bad & ~_connected_rows(rp, c, n, k)).any(), (fmt, k, np.flatnonzero(bad))


def test_reference_fsp_mat_test():
    # tests/test_fsp_mat.cu: toggle, bounds (31, 31); A.0 = 0 for every
    # format, CSR == HYB (L2 <= 1e-14), KRON vs HYB (L2 <= 1e-12)
    net = M.toggle_ref_test((31, 31))
    n = net.n_states
    mats = {f: capi.Mat(net, f, "f64") for f in ("csr", "hyb", "kron")}
    for f, A in mats.items():
        y = _apply(A, np.zeros(n), accumulate=True, y0=np.zeros(n))
        assert abs(y.sum()) <= 1e-14, f
    ones = np.ones(n)
    ys = {f: _apply(A, ones, accumulate=True, y0=np.zeros(n)) for f, A in mats.items()}
    assert np.linalg.norm(ys["csr"] - ys["hyb"]) <= 1e-14
    assert np.linalg.norm(ys["kron"] - ys["hyb"]) <= 1e-12


@pytest.mark.parametrize("fmt", ["csr", "hyb", "kron"])
def test_sink_generator_conserves_mass(fmt):
    # with the sink row every column sums to 0: 1^T (A x) = 0 for any x
    net = M.gene_expression((2, 30, 30))
    A = capi.Mat(net, fmt, "f64", with_sink=True)
    x = np.random.default_rng(3).random(A.n_rows)
    y = _apply(A, x)
    assert abs(y.sum()) <= 1e-12 * np.abs(y).sum()


def test_kronecker_shift_closed_form():
    # tests/test_kronecker_kernel.cu:247-337: x == 1 and reaction "x0 -> 0"
    # with factor f(i) = i: the off-diagonal part gives y[i] = i0 + 1 for
    # i0 < N0 and 0 at i0 = N0.  Here: full generator term, so y = (A x)
    # with x == 1 equals inflow - outflow = (i0+1) [i0 < N0] - i0.
    net = M.birth_death((1024, 1024))
    net = net.with_coefs([0.0, 1.0, 0.0, 0.0])
    A = capi.Mat(net, "kron", "f64")
    y = _apply(A, np.ones(net.n_states)).reshape(1025, 1025)
    i0 = np.arange(1025, dtype=np.float64)
    want = np.where(i0 < 1024, i0 + 1, 0.0) - i0
    assert np.array_equal(y, np.broadcast_to(want, (1025, 1025)))


def test_transpose_and_term_apply():
    net = M.gene_expression((2, 20, 25))
    n = net.n_states
    rp, c, v = O.merged_csr(net)
    Ad = np.zeros((n, n))
    for i in range(n):
        Ad[i, c[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    x = np.random.default_rng(5).random(n)
    for fmt in ("csr", "hyb", "kron"):
        A = capi.Mat(net, fmt, "f64")
        y = zeros(n)
        A.apply_transpose(dev(x), y, accumulate=False)
        assert np.abs(host(y) - Ad.T @ x).max() <= 1e-12 * np.abs(Ad.T @ x).max()
    A = capi.Mat(net, "kron", "f64")
    tot = np.zeros(n)
    for r in range(net.n_reactions):
        y = zeros(n)
        A.apply_term(r, net.coefs[r], dev(x), y, accumulate=False)
        tot += host(y)
    assert np.abs(tot - Ad @ x).max() <= 1e-12 * np.abs(Ad @ x).max()


def test_expm_dense_matches_oracle():
    rng = np.random.default_rng(11)
    for N in (5, 17, 32, 52):
        A = np.triu(rng.standard_normal((N, N)), -1) * rng.uniform(0.1, 30)
        Af = np.asfortranarray(A)
        d_A, d_F = dev(Af.ravel(order="F")), zeros(N * N)
        capi.expm_dense(d_A, d_F, N)
        F = host(d_F).reshape(N, N, order="F")
        Fo = O.expm(A)
        assert np.abs(F - Fo).max() <= 1e-15 * max(1.0, np.abs(Fo).max()), N


def _gpu_expv(net, fmt, p0, t_out, prec="f64", sink=True, with_taus=False, **kw):
    A = capi.Mat(net, fmt, prec, with_sink=sink)
    v = dev(p0)
    E = capi.Expv(A, A.n_rows, t_out[0], **kw)
    outs = []
    for t in t_out:
        E.set_final_time(t)
        E.solve(v)
        outs.append(host(v).copy())
    if with_taus:
        return np.array(outs), E.stats(), E.tau_history()
    return np.array(outs), E.stats()


def _expv_parity(net, fmt, p0, t_out, sink, **kw):
    """Parity of the device expv with the oracle's restatement of
    KryExpvFSP::step: (1) the oracle replaying the device's accepted step
    sizes agrees to 1e-10 L1 at every output time (arithmetic parity);
    (2) free-running, the step sequences agree and so do the solutions
    (1e-10 L1).  Returns the device outputs and stats."""
    O.set_threads(8)
    opr = O.Operator(net, fmt, with_sink=sink)   # same row arithmetic as the format
    got, st, taus = _gpu_expv(net, fmt, p0, t_out, sink=sink, with_taus=True, **kw)
    rc, rep, st_r = O.expv_outputs(opr, p0.copy(), t_out, replay=taus, **kw)
    assert rc == 0 and st_r.i_step == st.i_step
    for k in range(len(t_out)):
        print(f"[parity] {fmt} t={t_out[k]} replay L1 {l1(got[k], rep[k]):.3e} "
              f"bit-identical {np.array_equal(got[k], rep[k])}", flush=True)
        assert l1(got[k], rep[k]) <= 1e-10, (fmt, k, l1(got[k], rep[k]))
    rc, want, st_o = O.expv_outputs(opr, p0.copy(), t_out, **kw)
    assert rc == 0
    assert st.i_step == st_o.i_step, (fmt, st.i_step, st_o.i_step)
    for k in range(len(t_out)):
        assert l1(got[k], want[k]) <= 1e-10, (fmt, k, l1(got[k], want[k]))
    return got, st


@pytest.mark.parametrize("fmt", ["csr", "hyb", "kron"])
def test_expv_toggle_matches_oracle(fmt):
    # toggle switch 128 x 128 + sink, p0 = delta(0,0), m = 30, tol 1e-8,
    # outputs at t = 1, 2, 5, 10
    net = M.toggle_switch()
    p0 = M.delta_p0(net, with_sink=True)
    got, _ = _expv_parity(net, fmt, p0, [1.0, 2.0, 5.0, 10.0], sink=True)
    assert abs(got[-1].sum() - 1.0) <= 1e-12


def test_expv_gene_expression_matches_oracle():
    net = M.gene_expression((2, 63, 63))
    p0 = M.delta_p0(net, with_sink=True)
    _expv_parity(net, "csr", p0, [2.0, 10.0, 30.0], sink=True)


def test_expv_iop_reference_kry_test():
    # tests/test_kry_expv.cu:80-110: toggle on (256, 256), HYB, p0 = e_0,
    # t = 8 h, m = 30, tol = 1e-8, IOP with q = 2; mass kept to 1e-10
    net = M.toggle_ref_test((256, 256))
    p0 = M.delta_p0(net)
    got, st = _expv_parity(net, "hyb", p0, [8 * 3600.0], sink=False, iop=True, q_iop=2)
    assert abs(1.0 - got[0].sum()) <= 1e-10


def test_expv_callback_operator_matches_fused():
    net = M.gene_expression((2, 30, 30))
    A = capi.Mat(net, "csr", "f64")
    p0 = M.delta_p0(net)
    v1 = dev(p0)
    E1 = capi.Expv(A, A.n_rows, 5.0)
    E1.solve(v1)
    v2 = dev(p0)
    E2 = capi.Expv(None, A.n_rows, 5.0, matvec=lambda x, y: A.apply(x, y, accumulate=True))
    E2.solve(v2)
    assert l1(host(v1), host(v2)) <= 1e-12


def test_happy_breakdown():
    # all coefficients zero: A = 0, the Krylov space is 1-dimensional
    net = M.toggle_switch((15, 15)).with_coefs([0.0, 0.0, 0.0, 0.0])
    p0 = M.delta_p0(net, state=(3, 4))
    got, st = _gpu_expv(net, "csr", p0, [5.0], sink=False)
    assert st.n_breakdown >= 1
    assert np.array_equal(got[0], p0)


def test_failed_solve_reports_stats_and_restarts():
    # max_reject = 1: the first step whose first step size is rejected
    # fails the solve.  The handle still reports its statistics, and a
    # restart with the default limits gives the clean solve's result.
    net = M.toggle_switch((31, 31))
    A = capi.Mat(net, "csr", "f64")
    p0 = M.delta_p0(net)
    clean = dev(p0)
    E0 = capi.Expv(A, A.n_rows, 5.0)
    E0.solve(clean)
    want, st_clean = host(clean), E0.stats()
    assert st_clean.n_reject > 0
    v = dev(p0)
    E = capi.Expv(A, A.n_rows, 5.0)
    E.set_limits(max_reject=1, max_nstep=10000)
    with pytest.raises(capi.FspError):
        E.solve(v)
    st = E.stats()
    assert st.n_reject >= 1 and st.t_now < 5.0
    assert len(E.tau_history()) == st.i_step
    E.set_limits(max_reject=1000, max_nstep=10000)
    v.copy_from(p0)
    E.restart(5.0)
    E.solve(v)
    assert E.stats().i_step == st_clean.i_step
    assert np.array_equal(host(v), want)


@pytest.mark.parametrize("fmt", ["csr", "hyb", "kron"])
def test_mixed_precision_l1(fmt):
    net = M.gene_expression((2, 60, 60))
    p0 = M.delta_p0(net, with_sink=True)
    t_out = [5.0, 20.0]
    ref, _ = _gpu_expv(net, fmt, p0, t_out, prec="f64")
    got, _ = _gpu_expv(net, fmt, p0, t_out, prec="mixed")
    for k in range(len(t_out)):
        err = l1(got[k], ref[k])
        print(f"[mixed] {fmt} t={t_out[k]} L1 vs FP64 = {err:.3e} "
              f"mass err = {abs(got[k].sum() - 1.0):.3e}", flush=True)
        assert err <= 1e-6, (fmt, k, err)
        assert abs(got[k].sum() - 1.0) <= 1e-6


@pytest.mark.parametrize("fmt", ["csr", "hyb", "kron"])
def test_mixed_values_with_f64_basis(fmt):
    """FP32 generator values with an FP64 Krylov basis: the FP64 Arnoldi
    process on the FP32-rounded generator (itself a generator: its diagonal
    is the negated sum of the rounded rates), so the step count stays at
    FP64's and the solution within 1e-6 L1 of FP64."""
    net = M.gene_expression((2, 60, 60))
    p0 = M.delta_p0(net, with_sink=True)
    t_out = [5.0, 20.0]
    ref, st64 = _gpu_expv(net, fmt, p0, t_out, prec="f64")
    got, stv = _gpu_expv(net, fmt, p0, t_out, prec="mixed", basis="f64")
    for k in range(len(t_out)):
        err = l1(got[k], ref[k])
        print(f"[mixed values] {fmt} t={t_out[k]} L1 vs FP64 = {err:.3e} steps {stv.i_step} vs {st64.i_step}",
              flush=True)
        assert err <= 1e-6, (fmt, k, err)
        assert abs(got[k].sum() - 1.0) <= 1e-6
    assert abs(stv.i_step - st64.i_step) <= max(2, st64.i_step // 20)


# Rates under which one classical Gram-Schmidt pass over an FP32 basis loses
# orthogonality completely within 30 Arnoldi steps from a point mass (the
# Hessenberg exponential then overflows and every step size is rejected).
TOGGLE_CGS_RATES = (7.838583757547972, 0.5659944527754909, 13.088930136958103, 6.072545924918103)


@pytest.mark.parametrize("side,t_final", [(20, 4.0), (405, 42.50610831860647)])
@pytest.mark.parametrize("fmt", ["csr", "hyb", "kron"])
def test_mixed_precision_reorthogonalizes(fmt, side, t_final):
    """The mixed form converges where one Gram-Schmidt pass would not, and
    stays within 1e-6 L1 of FP64 over ~800 steps on the 405 x 405 box."""
    net = M.toggle_switch((side, side), rates=TOGGLE_CGS_RATES)
    p0 = M.delta_p0(net, with_sink=False)
    print(f"[case] mixed reorth {fmt} {side}x{side} t={t_final}", flush=True)
    ref, st64 = _gpu_expv(net, fmt, p0, [t_final], prec="f64", sink=False)
    got, st32 = _gpu_expv(net, fmt, p0, [t_final], prec="mixed", sink=False)
    err = l1(got[0], ref[0])
    print(f"[mixed] reorth {fmt} {side}x{side} L1 vs FP64 = {err:.3e} steps "
          f"{st32.i_step} vs {st64.i_step}", flush=True)
    assert np.isfinite(got[0]).all()
    assert err <= 1e-6
