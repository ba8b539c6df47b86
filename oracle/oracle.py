"""ctypes wrapper of the CPU oracle (oracle/libfsp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg, never by the product package.  See
fsp_oracle.h for the reference file:line each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfsp_oracle.so")


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE, "-B" if force else "all"])
    return LIB_PATH


class _Model(C.Structure):
    _fields_ = [("n_species", C.c_int), ("n_reactions", C.c_int),
                ("bounds", C.c_void_p), ("stoich_vals", C.c_void_p),
                ("stoich_cols", C.c_void_p), ("stoich_rowptrs", C.c_void_p),
                ("factors", C.c_void_p), ("prop_table", C.c_void_p)]


class ExpvStats(C.Structure):
    _fields_ = [("t_now", C.c_double), ("t_new", C.c_double), ("tau_last", C.c_double),
                ("err_last", C.c_double), ("beta_last", C.c_double),
                ("i_step", C.c_int64), ("n_matvec", C.c_int64),
                ("n_reject", C.c_int64), ("n_breakdown", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i64, dbl, ci = C.c_void_p, C.c_int64, C.c_double, C.c_int
        L.orc_num_states.restype = i64
        L.orc_num_states.argtypes = [ci, vp]
        L.orc_enumerate.argtypes = [ci, vp, vp]
        L.orc_state2indx.restype = i64
        L.orc_state2indx.argtypes = [vp, ci, vp]
        L.orc_indx2state.argtypes = [i64, vp, ci, vp]
        L.orc_term_csr.restype = i64
        L.orc_term_csr.argtypes = [vp, ci, vp, vp, vp]
        L.orc_csr_apply.argtypes = [i64, vp, vp, vp, dbl, vp, vp]
        L.orc_merged_csr.restype = i64
        L.orc_merged_csr.argtypes = [vp, vp, ci, vp, vp, vp]
        L.orc_unique_shifts.restype = ci
        L.orc_unique_shifts.argtypes = [vp, vp]
        L.orc_ell.argtypes = [vp, vp, vp, vp, vp, vp]
        L.orc_ell_apply.argtypes = [i64, ci, vp, vp, vp, vp, vp, vp]
        L.orc_hyb_ref.argtypes = [vp, vp, vp, vp]
        L.orc_hyb_ref_apply.argtypes = [i64, ci, vp, vp, vp, vp, vp, vp]
        L.orc_kron_apply.argtypes = [vp, vp, ci, vp, vp]
        L.orc_kron_slot_form.restype = ci
        L.orc_kron_slot_form.argtypes = [vp]
        L.orc_kron_slot_apply.restype = ci
        L.orc_kron_slot_apply.argtypes = [vp, vp, ci, vp, vp]
        L.orc_expm.restype = ci
        L.orc_expm.argtypes = [ci, vp, vp]
        L.orc_expv_csr.restype = ci
        L.orc_expv_csr.argtypes = [i64, vp, vp, vp, vp, dbl, ci, dbl, ci, ci, dbl, vp]
        L.orc_expv_csr_outputs.restype = ci
        L.orc_expv_csr_outputs.argtypes = [i64, vp, vp, vp, vp, vp, ci, vp, ci, dbl, ci,
                                           ci, dbl, vp, ci, vp]
        L.orc_round_up_2sig.restype = dbl
        L.orc_round_up_2sig.argtypes = [dbl]
        L.orc_loglik_indices.restype = dbl
        L.orc_loglik_indices.argtypes = [vp, vp, i64, dbl]
        L.orc_set_threads.argtypes = [ci]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def set_verbose(on: bool) -> None:
    """Progress lines (one per output time) from the expv drivers."""
    lib().orc_set_verbose(1 if on else 0)


def _p(a):
    return None if a is None else a.ctypes.data


class OracleModel:
    """Keeps the numpy buffers alive for the C struct."""

    def __init__(self, net, prop_table=None):
        self.net = net
        self.bounds = np.ascontiguousarray(net.bounds, np.int32)
        self.sv, self.sc, self.sr = net.stoich_csr()
        self.factors = None if prop_table is not None else np.ascontiguousarray(net.factors, np.float64)
        self.prop_table = None if prop_table is None else np.ascontiguousarray(prop_table, np.float64)
        self.s = _Model(net.n_species, net.n_reactions, _p(self.bounds), _p(self.sv),
                        _p(self.sc), _p(self.sr), _p(self.factors), _p(self.prop_table))

    @property
    def ptr(self):
        return C.addressof(self.s)


def num_states(bounds) -> int:
    b = np.ascontiguousarray(bounds, np.int32)
    return int(lib().orc_num_states(len(b), _p(b)))


def enumerate_states(bounds) -> np.ndarray:
    b = np.ascontiguousarray(bounds, np.int32)
    n = num_states(b)
    out = np.empty((n, len(b)), np.int32)
    lib().orc_enumerate(len(b), _p(b), _p(out))
    return out


def term_csr(net, r, prop_table=None):
    m = OracleModel(net, prop_table)
    n = net.n_states
    rp = np.empty(n + 1, np.int32)
    nnz = lib().orc_term_csr(m.ptr, r, _p(rp), None, None)
    cols = np.empty(nnz, np.int32)
    vals = np.empty(nnz, np.float64)
    lib().orc_term_csr(m.ptr, r, _p(rp), _p(cols), _p(vals))
    return rp, cols, vals


def csr_apply(rp, cols, vals, x, y, alpha=1.0):
    lib().orc_csr_apply(len(rp) - 1, _p(rp), _p(cols), _p(vals), float(alpha), _p(x), _p(y))


def merged_csr(net, coefs=None, with_sink=False, prop_table=None):
    m = OracleModel(net, prop_table)
    c = np.ascontiguousarray(net.coefs if coefs is None else coefs, np.float64)
    n = net.n_states
    rows = n + (1 if with_sink else 0)
    rp = np.empty(rows + 1, np.int32)
    nnz = lib().orc_merged_csr(m.ptr, _p(c), int(with_sink), _p(rp), None, None)
    cols = np.empty(nnz, np.int32)
    vals = np.empty(nnz, np.float64)
    lib().orc_merged_csr(m.ptr, _p(c), int(with_sink), _p(rp), _p(cols), _p(vals))
    return rp, cols, vals


def unique_shifts(net):
    m = OracleModel(net)
    slot = np.empty(net.n_reactions, np.int32)
    U = lib().orc_unique_shifts(m.ptr, _p(slot))
    return U, slot


def ell(net, coefs=None, with_sink=False, prop_table=None):
    m = OracleModel(net, prop_table)
    c = np.ascontiguousarray(net.coefs if coefs is None else coefs, np.float64)
    U, _ = unique_shifts(net)
    n = net.n_states
    diag = np.empty(n, np.float64)
    sv = np.empty((max(U, 1), n), np.float64)
    sc = np.empty((max(U, 1), n), np.int32)
    sink = np.empty(n, np.float64) if with_sink else None
    lib().orc_ell(m.ptr, _p(c), _p(diag), _p(sv), _p(sc), _p(sink))
    return U, diag, sv[:U], sc[:U], sink


def ell_apply(U, diag, sv, sc, sink, x, y):
    sv = np.ascontiguousarray(sv)
    sc = np.ascontiguousarray(sc)
    lib().orc_ell_apply(len(diag), U, _p(diag), _p(sv), _p(sc), _p(sink), _p(x), _p(y))


def hyb_ref(net, prop_table=None):
    m = OracleModel(net, prop_table)
    n, R = net.n_states, net.n_reactions
    diag = np.empty((R, n), np.float64)
    off = np.empty((R, n), np.float64)
    col = np.empty((R, n), np.int32)
    lib().orc_hyb_ref(m.ptr, _p(diag), _p(off), _p(col))
    return diag, off, col


def hyb_ref_apply(diag, off, col, coefs, x, y):
    R, n = diag.shape
    c = np.ascontiguousarray(coefs, np.float64)
    lib().orc_hyb_ref_apply(n, R, _p(diag), _p(off), _p(col), _p(c), _p(x), _p(y))


def kron_apply(net, x, y, coefs=None, with_sink=False):
    m = OracleModel(net)
    c = np.ascontiguousarray(net.coefs if coefs is None else coefs, np.float64)
    lib().orc_kron_apply(m.ptr, _p(c), int(with_sink), _p(x), _p(y))


def kron_slot_apply(net, x, y, coefs=None, with_sink=False) -> bool:
    """y += A x in the slot form of the Kronecker operator (reactions grouped
    by net change with coefficient-weighted factor tables, see fsp_oracle.c);
    False (y untouched) where that form does not apply."""
    m = OracleModel(net)
    c = np.ascontiguousarray(net.coefs if coefs is None else coefs, np.float64)
    return lib().orc_kron_slot_apply(m.ptr, _p(c), int(with_sink), _p(x), _p(y)) == 0


def kron_apply_device_form(net, x, y, coefs=None, with_sink=False):
    """y += A x with the row arithmetic of the device's KRON format: the slot
    form where it applies, else the reference's per-reaction form."""
    if not kron_slot_apply(net, x, y, coefs, with_sink):
        kron_apply(net, x, y, coefs, with_sink)


def expm(A: np.ndarray) -> np.ndarray:
    A = np.asfortranarray(A, np.float64)
    F = np.empty_like(A, order="F")
    rc = lib().orc_expm(A.shape[0], A.ctypes.data, F.ctypes.data)
    if rc != 0:
        raise RuntimeError("orc_expm: singular Pade denominator")
    return F


def expv_csr(rp, cols, vals, v, t_final, m=30, tol=1e-8, iop=False, q_iop=2, anorm=1.0):
    """Integrates v in place to t_final; returns (rc, stats)."""
    st = ExpvStats()
    rc = lib().orc_expv_csr(len(rp) - 1, _p(rp), _p(cols), _p(vals), _p(v), float(t_final),
                            int(m), float(tol), int(iop), int(q_iop), float(anorm),
                            C.addressof(st))
    return rc, st


def expv_csr_outputs(rp, cols, vals, v, t_out, m=30, tol=1e-8, iop=False, q_iop=2, anorm=1.0,
                     replay=None):
    """Integrates v through the output times; returns (rc, outputs, stats).
    replay: step sizes to force (accepted as given), e.g. a GPU run's."""
    t = np.ascontiguousarray(t_out, np.float64)
    n = len(rp) - 1
    out = np.empty((len(t), n), np.float64)
    st = ExpvStats()
    rep = None if replay is None else np.ascontiguousarray(replay, np.float64)
    rc = lib().orc_expv_csr_outputs(n, _p(rp), _p(cols), _p(vals), _p(v), _p(t), len(t),
                                    _p(out), int(m), float(tol), int(iop), int(q_iop),
                                    float(anorm), _p(rep), 0 if rep is None else len(rep),
                                    C.addressof(st))
    return rc, out, st


def round_up_2sig(t: float) -> float:
    return float(lib().orc_round_up_2sig(float(t)))


def loglik_indices(p, idx, floor_val=1e-44):
    idx = np.ascontiguousarray(idx, np.int64)
    return float(lib().orc_loglik_indices(_p(p), _p(idx), len(idx), float(floor_val)))


class _Op(C.Structure):
    _fields_ = [("kind", C.c_int), ("n_rows", C.c_int64), ("dd_last", C.c_int),
                ("rp", C.c_void_p), ("cols", C.c_void_p), ("vals", C.c_void_p),
                ("U", C.c_int), ("diag", C.c_void_p), ("slot_vals", C.c_void_p),
                ("sink_out", C.c_void_p), ("slot_cols", C.c_void_p),
                ("model", C.c_void_p), ("coefs", C.c_void_p), ("with_sink", C.c_int)]


class Operator:
    """Oracle operator y += A x with the same row arithmetic as one of the
    GPU formats: 'csr' (merged CSR, column order), 'hyb' (diag + ELL slots),
    'kron' (Kronecker, by slot where that form applies), 'kron_reaction'
    (Kronecker per reaction, the reference's formula, as the persistent
    batch kernel computes it).  Keeps its arrays alive."""

    def __init__(self, net, fmt="csr", with_sink=False, coefs=None):
        c = np.ascontiguousarray(net.coefs if coefs is None else coefs, np.float64)
        self.keep = [c]
        n = net.n_states
        rows = n + (1 if with_sink else 0)
        op = _Op()
        op.n_rows = rows
        if fmt == "csr":
            rp, cols, vals = merged_csr(net, coefs=c, with_sink=with_sink)
            self.keep += [rp, cols, vals]
            op.kind, op.dd_last = 0, int(with_sink)
            op.rp, op.cols, op.vals = _p(rp), _p(cols), _p(vals)
        elif fmt == "hyb":
            U, dg, sv, sc, sk = ell(net, coefs=c, with_sink=with_sink)
            sv, sc = np.ascontiguousarray(sv), np.ascontiguousarray(sc)
            self.keep += [dg, sv, sc, sk]
            op.kind, op.U = 1, U
            op.diag, op.slot_vals, op.slot_cols, op.sink_out = _p(dg), _p(sv), _p(sc), _p(sk)
        elif fmt in ("kron", "kron_reaction"):
            # "kron": the device's KRON row arithmetic (the slot form where it
            # applies); "kron_reaction": the reference's per-reaction formula
            m = OracleModel(net)
            self.keep.append(m)
            op.kind, op.with_sink = (2 if fmt == "kron" else 3), int(with_sink)
            op.model, op.coefs = m.ptr, _p(c)
        else:
            raise ValueError(fmt)
        self.op = op
        self.n_rows = rows

    def apply(self, x, y):
        L = lib()
        L.orc_op_apply.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_op_apply(C.addressof(self.op), _p(x), _p(y))


def expv_outputs(opr: Operator, v, t_out, m=30, tol=1e-8, iop=False, q_iop=2, anorm=1.0,
                 replay=None, taus=False, tau_cap=1 << 16):
    """KryExpvFSP restatement through output times with an Operator; returns
    (rc, outputs, stats), plus the accepted step sizes with taus=True.
    replay: step sizes to force (accepted as given)."""
    L = lib()
    L.orc_expv_op_outputs.restype = C.c_int
    L.orc_expv_op_outputs.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                      C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                                      C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int64]
    t = np.ascontiguousarray(t_out, np.float64)
    out = np.empty((len(t), opr.n_rows), np.float64)
    st = ExpvStats()
    rep = None if replay is None else np.ascontiguousarray(replay, np.float64)
    log = np.zeros(tau_cap if taus else 0, np.float64)
    rc = L.orc_expv_op_outputs(C.addressof(opr.op), _p(v), _p(t), len(t), _p(out), int(m),
                               float(tol), int(iop), int(q_iop), float(anorm), _p(rep),
                               0 if rep is None else len(rep), C.addressof(st),
                               _p(log) if taus else None, len(log))
    if taus:
        return rc, out, st, log[:min(int(st.i_step), len(log))].copy()
    return rc, out, st
