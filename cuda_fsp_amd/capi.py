"""ctypes binding of the HIP library's C ABI (include/fsp_capi.h).

This is the Python-side caller of the FSP engine: every compute call goes to
``cuda_fsp_amd/lib/libfsp_hip.so``.  There is no CPU fallback: if the
library is missing, ``lib()`` raises.  Device memory is plain device pointers
(ints); torch CUDA tensors are accepted wherever a pointer is expected.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfsp_hip.so")

FSP_CSR, FSP_HYB, FSP_KRON = 0, 1, 2
FSP_F64, FSP_MIXED = 0, 1
FORMATS = {"csr": FSP_CSR, "hyb": FSP_HYB, "kron": FSP_KRON}
PRECISIONS = {"f64": FSP_F64, "mixed": FSP_MIXED}

STATUS = {0: "FSP_OK", 1: "FSP_ERR_INVALID", 2: "FSP_ERR_UNSUPPORTED", 3: "FSP_ERR_OVERFLOW",
          4: "FSP_ERR_HIP", 5: "FSP_ERR_NOMEM", 6: "FSP_ERR_NOT_CONVERGED",
          7: "FSP_ERR_NO_DEVICE", 8: "FSP_ERR_TIMEOUT"}


class FspError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class FspModel(C.Structure):
    _fields_ = [("n_species", C.c_int), ("n_reactions", C.c_int),
                ("bounds", C.c_void_p), ("stoich_vals", C.c_void_p),
                ("stoich_cols", C.c_void_p), ("stoich_rowptrs", C.c_void_p),
                ("factors", C.c_void_p), ("prop_table_dev", C.c_void_p)]


class FspMatInfo(C.Structure):
    _fields_ = [("n_states", C.c_int64), ("n_rows", C.c_int64), ("nnz", C.c_int64),
                ("n_slots", C.c_int), ("format", C.c_int), ("precision", C.c_int),
                ("with_sink", C.c_int), ("max_row", C.c_int),
                ("bytes_per_apply", C.c_double)]


class FspExpvStats(C.Structure):
    _fields_ = [("t_now", C.c_double), ("t_new", C.c_double), ("tau_last", C.c_double),
                ("err_last", C.c_double), ("beta_last", C.c_double),
                ("i_step", C.c_int64), ("n_matvec", C.c_int64),
                ("n_reject", C.c_int64), ("n_breakdown", C.c_int64)]


PROF_CLASSES = ("prologue", "arnoldi", "dots", "update", "control", "solution")


class FspExpvProfile(C.Structure):
    _fields_ = [("launches", C.c_int64 * 6), ("ms", C.c_double * 6), ("bytes", C.c_double * 6)]


class FspBatchOptions(C.Structure):
    _fields_ = [("m", C.c_int), ("tol", C.c_double), ("anorm", C.c_double),
                ("precision", C.c_int), ("iop", C.c_int), ("q_iop", C.c_int),
                ("lockstep_bpc", C.c_int), ("wg_per_cu", C.c_int), ("max_reject", C.c_int),
                ("max_nstep", C.c_int64), ("path", C.c_int), ("split_states", C.c_int64),
                ("time_limit_s", C.c_double), ("lockstep_rows", C.c_int)]


BATCH_PATHS = {"auto": 0, "lockstep": 1, "streams": 2, "persistent": 3, "hybrid": 4}
BATCH_PATH_NAMES = {v: k for k, v in BATCH_PATHS.items()}


MATVEC_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p)
TCOEF_CB = C.CFUNCTYPE(None, C.c_double, C.POINTER(C.c_double), C.c_void_p)


class FspBdfOptions(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("max_order", C.c_int),
                ("krylov_dim", C.c_int), ("max_restarts", C.c_int), ("h0", C.c_double),
                ("max_steps", C.c_int64), ("lin_tol_factor", C.c_double)]


class FspLikData(C.Structure):
    _fields_ = [("n_times", C.c_int), ("times", C.c_void_p), ("keep_mask", C.c_void_p),
                ("counts", C.c_void_p), ("floor_val", C.c_double)]


class FspGradOptions(C.Structure):
    _fields_ = [("format", C.c_int), ("m", C.c_int), ("tol", C.c_double),
                ("quad_nodes", C.c_int), ("grading_levels", C.c_int)]


class FspBdfStats(C.Structure):
    _fields_ = [("n_steps", C.c_int64), ("n_rejects", C.c_int64), ("n_rhs", C.c_int64),
                ("n_lin_iters", C.c_int64), ("n_lin_fails", C.c_int64), ("order", C.c_int),
                ("h_last", C.c_double), ("t", C.c_double)]

# (name, restype, argtypes); every entry point returns fsp_status unless noted
_vp, _i64, _d, _i = C.c_void_p, C.c_int64, C.c_double, C.c_int
_SIGS = {
    "fsp_last_error": (C.c_char_p, []),
    "fsp_version": (C.c_int, []),
    "fsp_device_count": (C.c_int, [_vp]),
    "fsp_malloc": (C.c_int, [_vp, C.c_size_t]),
    "fsp_free": (C.c_int, [_vp]),
    "fsp_memcpy_h2d": (C.c_int, [_vp, _vp, C.c_size_t]),
    "fsp_memcpy_d2h": (C.c_int, [_vp, _vp, C.c_size_t]),
    "fsp_memcpy_d2d": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    "fsp_memset_zero": (C.c_int, [_vp, C.c_size_t, _vp]),
    "fsp_stream_sync": (C.c_int, [_vp]),
    "fsp_vec_sum": (C.c_int, [_vp, _i64, _vp, _vp]),
    "fsp_vec_axpy": (C.c_int, [_d, _vp, _vp, _i64, _vp]),
    "fsp_vec_fill": (C.c_int, [_vp, _d, _i64, _vp]),
    "fsp_num_states": (C.c_int, [_i, _vp, _vp]),
    "fsp_states_enumerate": (C.c_int, [_i, _vp, _vp, _vp]),
    "fsp_states_to_index": (C.c_int, [_i, _vp, _vp, _i64, _vp, _vp]),
    "fsp_mat_create": (C.c_int, [_vp, _i, _i, _i, _vp, _vp, _vp]),
    "fsp_mat_destroy": (C.c_int, [_vp]),
    "fsp_mat_get_info": (C.c_int, [_vp, _vp]),
    "fsp_mat_set_coefs": (C.c_int, [_vp, _vp, _vp]),
    "fsp_mat_apply": (C.c_int, [_vp, _vp, _vp, _vp, _i, _vp]),
    "fsp_mat_apply_transpose": (C.c_int, [_vp, _vp, _vp, _vp, _i, _vp]),
    "fsp_mat_apply_term": (C.c_int, [_vp, _i, _d, _vp, _vp, _i, _vp]),
    "fsp_mat_export_csr": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "fsp_mat_export_hyb": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "fsp_term_csr": (C.c_int, [_vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "fsp_expv_create": (C.c_int, [_vp, MATVEC_CB, _vp, _i64, _d, _i, _d, _i, _i, _d, _vp, _vp]),
    "fsp_expv_destroy": (C.c_int, [_vp]),
    "fsp_expv_step": (C.c_int, [_vp, _vp]),
    "fsp_expv_solve": (C.c_int, [_vp, _vp]),
    "fsp_expv_reset_time": (C.c_int, [_vp]),
    "fsp_expv_set_final_time": (C.c_int, [_vp, _d]),
    "fsp_expv_restart": (C.c_int, [_vp, _d]),
    "fsp_expv_final_time_reached": (C.c_int, [_vp, _vp]),
    "fsp_expv_get_stats": (C.c_int, [_vp, _vp]),
    "fsp_expv_set_limits": (C.c_int, [_vp, _i, _i64]),
    "fsp_expv_get_tau_history": (C.c_int, [_vp, _vp, _i64, _vp]),
    "fsp_expv_set_iop": (C.c_int, [_vp, _i, _i]),
    "fsp_expv_set_profiling": (C.c_int, [_vp, _i]),
    "fsp_expv_set_basis_precision": (C.c_int, [_vp, _i]),
    "fsp_expv_get_profile": (C.c_int, [_vp, _vp]),
    "fsp_batch_create": (C.c_int, [_vp, _vp, _vp, _vp, _i, _vp, _vp, _vp]),
    "fsp_batch_destroy": (C.c_int, [_vp]),
    "fsp_batch_set_initial": (C.c_int, [_vp, _i, _vp]),
    "fsp_batch_solve": (C.c_int, [_vp, _vp]),
    "fsp_batch_get_solution": (C.c_int, [_vp, _i, _vp, _i64]),
    "fsp_batch_get_stats": (C.c_int, [_vp, _i, _vp, _vp, _vp]),
    "fsp_batch_get_config": (C.c_int, [_vp, _vp, _vp]),
    "fsp_batch_get_tuning": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "fsp_batch_prepare": (C.c_int, [_vp]),
    "fsp_batch_autotune": (C.c_int, [_vp, C.c_char_p, _i, _vp]),
    "fsp_loglik_indices": (C.c_int, [_vp, _vp, _i64, _d, _vp, _vp]),
    "fsp_loglik_indices_noalloc": (C.c_int, [_vp, _vp, _i64, _d, _vp, _vp]),
    "fsp_marginal": (C.c_int, [_vp, _i, _vp, _vp, _vp, _vp]),
    "fsp_loglik_weighted": (C.c_int, [_vp, _vp, _i64, _d, _vp, _vp]),
    "fsp_expm_dense": (C.c_int, [_i, _vp, _vp, _vp]),
    "fsp_rhs_apply": (C.c_int, [_vp, _d, _vp, _vp, _vp, _vp, _vp]),
    "fsp_loglik": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fsp_loglik_grad": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fsp_loglik_grad_batch": (C.c_int, [_vp, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp]),
    "fsp_bdf_create": (C.c_int, [_vp, _vp, _vp, _d, _vp, _vp, _vp]),
    "fsp_bdf_destroy": (C.c_int, [_vp]),
    "fsp_bdf_advance": (C.c_int, [_vp, _vp, _d]),
    "fsp_bdf_reset": (C.c_int, [_vp, _d]),
    "fsp_bdf_get_stats": (C.c_int, [_vp, _vp]),
}

_lib = None


def lib():
    """Loads libfsp_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FspError(-1, f"{LIB_PATH} is missing: run cuda_fsp_amd.build first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    return sorted(_SIGS)


def check(status):
    if status != 0:
        msg = lib().fsp_last_error().decode(errors="replace")
        raise FspError(status, msg)


def ptr(x):
    """Device pointer of a torch CUDA tensor, or an int pointer."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    raise TypeError(f"cannot take a device pointer of {type(x)}")


class DeviceBuffer:
    """Device allocation owned from Python through the C ABI (fsp_malloc).

    The engine brings its own HIP runtime (/opt/rocm); it must not share a
    process with another initialised HIP runtime (e.g. torch's bundled one),
    so Python callers hold device memory through this class."""

    def __init__(self, count, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().fsp_malloc(C.byref(p), max(self.nbytes, 1)))
        self._p = p.value

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        b = cls(a.size, a.dtype)
        if a.nbytes:
            check(lib().fsp_memcpy_h2d(b._p, a.ctypes.data, a.nbytes))
        return b

    @classmethod
    def zeros(cls, count, dtype=np.float64):
        b = cls(count, dtype)
        check(lib().fsp_memset_zero(b._p, max(b.nbytes, 1), None))
        check(lib().fsp_stream_sync(None))
        return b

    def data_ptr(self) -> int:
        return self._p

    def offset(self, k) -> int:
        return self._p + int(k) * self.dtype.itemsize

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size <= self.count
        check(lib().fsp_memcpy_h2d(self._p, a.ctypes.data, a.nbytes))

    def to_host(self) -> np.ndarray:
        out = np.empty(self.count, self.dtype)
        check(lib().fsp_stream_sync(None))
        if self.nbytes:
            check(lib().fsp_memcpy_d2h(out.ctypes.data, self._p, self.nbytes))
        return out

    def free(self):
        if getattr(self, "_p", None):
            lib().fsp_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_count() -> int:
    c = C.c_int(0)
    check(lib().fsp_device_count(C.byref(c)))
    return c.value


class Model:
    """Host-side buffers of an fsp_model built from a ReactionNetwork."""

    def __init__(self, net, prop_table_dev=None):
        self.net = net
        self.bounds = np.ascontiguousarray(net.bounds, np.int32)
        self.sv, self.sc, self.sr = net.stoich_csr()
        self.factors = None if prop_table_dev is not None else np.ascontiguousarray(net.factors, np.float64)
        self._table = prop_table_dev
        self.s = FspModel(net.n_species, net.n_reactions, self.bounds.ctypes.data,
                          self.sv.ctypes.data, self.sc.ctypes.data, self.sr.ctypes.data,
                          None if self.factors is None else self.factors.ctypes.data,
                          ptr(prop_table_dev))

    @property
    def ref(self):
        return C.byref(self.s)


class Mat:
    """fsp_mat handle: the generator A(t) = sum_r c_r A_r on the device."""

    def __init__(self, net, fmt="csr", precision="f64", with_sink=False, coefs=None,
                 stream=None, prop_table_dev=None):
        self.model = Model(net, prop_table_dev)
        self.net = net
        c = np.ascontiguousarray(net.coefs if coefs is None else coefs, np.float64)
        h = C.c_void_p()
        check(lib().fsp_mat_create(self.model.ref, FORMATS[fmt], PRECISIONS[precision],
                                   int(with_sink), c.ctypes.data, stream, C.byref(h)))
        self.h = h
        self.fmt, self.precision, self.with_sink = fmt, precision, with_sink
        self.info = self.get_info()

    def get_info(self) -> FspMatInfo:
        info = FspMatInfo()
        check(lib().fsp_mat_get_info(self.h, C.byref(info)))
        return info

    @property
    def n_rows(self) -> int:
        return int(self.info.n_rows)

    def set_coefs(self, coefs, stream=None):
        c = np.ascontiguousarray(coefs, np.float64)
        check(lib().fsp_mat_set_coefs(self.h, c.ctypes.data, stream))

    def apply(self, x, y, accumulate=True, coefs=None, stream=None):
        c = None if coefs is None else np.ascontiguousarray(coefs, np.float64)
        check(lib().fsp_mat_apply(self.h, None if c is None else c.ctypes.data, ptr(x), ptr(y),
                                  int(accumulate), stream))

    def apply_transpose(self, x, y, accumulate=True, coefs=None, stream=None):
        c = None if coefs is None else np.ascontiguousarray(coefs, np.float64)
        check(lib().fsp_mat_apply_transpose(self.h, None if c is None else c.ctypes.data,
                                            ptr(x), ptr(y), int(accumulate), stream))

    def apply_term(self, r, c, x, y, accumulate=True, stream=None):
        check(lib().fsp_mat_apply_term(self.h, int(r), float(c), ptr(x), ptr(y),
                                       int(accumulate), stream))

    def export_csr(self, rp, cols, vals, stream=None):
        check(lib().fsp_mat_export_csr(self.h, ptr(rp), ptr(cols), ptr(vals), stream))

    def export_hyb(self, diag, slot_vals, slot_cols, sink_out=None, stream=None):
        check(lib().fsp_mat_export_hyb(self.h, ptr(diag), ptr(slot_vals), ptr(slot_cols),
                                       ptr(sink_out), stream))

    def close(self):
        if getattr(self, "h", None):
            check(lib().fsp_mat_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def term_csr_nnz(net, r, rp=None, stream=None) -> int:
    m = Model(net)
    nnz = C.c_int64(0)
    check(lib().fsp_term_csr(m.ref, int(r), ptr(rp), None, None, C.byref(nnz), stream))
    return nnz.value


def term_csr(net, r, rp, cols, vals, stream=None) -> int:
    m = Model(net)
    nnz = C.c_int64(0)
    check(lib().fsp_term_csr(m.ref, int(r), ptr(rp), ptr(cols), ptr(vals), C.byref(nnz), stream))
    return nnz.value


def states_enumerate(bounds, out, stream=None):
    b = np.ascontiguousarray(bounds, np.int32)
    check(lib().fsp_states_enumerate(len(b), b.ctypes.data, ptr(out), stream))


def states_to_index(bounds, states, count, out, stream=None):
    b = np.ascontiguousarray(bounds, np.int32)
    check(lib().fsp_states_to_index(len(b), b.ctypes.data, ptr(states), int(count), ptr(out),
                                    stream))


def num_states(bounds) -> int:
    b = np.ascontiguousarray(bounds, np.int32)
    n = C.c_int64(0)
    check(lib().fsp_num_states(len(b), b.ctypes.data, C.byref(n)))
    return n.value


def vec_sum(x, n, stream=None) -> float:
    s = C.c_double(0.0)
    check(lib().fsp_vec_sum(ptr(x), int(n), C.byref(s), stream))
    return s.value


def loglik_indices(p, idx, count, floor_val=1e-44, stream=None) -> float:
    s = C.c_double(0.0)
    check(lib().fsp_loglik_indices(ptr(p), ptr(idx), int(count), float(floor_val), C.byref(s),
                                   stream))
    return s.value


def loglik_indices_noalloc(p, idx, count, floor_val=1e-44, stream=None) -> float:
    s = C.c_double(0.0)
    check(lib().fsp_loglik_indices_noalloc(ptr(p), ptr(idx), int(count), float(floor_val),
                                           C.byref(s), stream))
    return s.value


def marginal(p, bounds, keep, out, stream=None):
    b = np.ascontiguousarray(bounds, np.int32)
    k = np.ascontiguousarray([1 if v else 0 for v in keep], np.int32)
    check(lib().fsp_marginal(ptr(p), len(b), b.ctypes.data, k.ctypes.data, ptr(out), stream))


def loglik_weighted(q, w, count, floor_val=1e-12, stream=None) -> float:
    s = C.c_double(0.0)
    check(lib().fsp_loglik_weighted(ptr(q), ptr(w), int(count), float(floor_val), C.byref(s),
                                    stream))
    return s.value


def expm_dense(A, F, N, stream=None):
    check(lib().fsp_expm_dense(int(N), ptr(A), ptr(F), stream))


def loglik(net, log_coefs, times, keep_mask, counts, floor_val=1e-12, p0=None, fmt="csr",
           m=30, tol=1e-10, quad_nodes=8, grading_levels=10, grad=False, stream=None):
    """Histogram log-likelihood of the FSP solution (fsp_loglik); with
    grad=True also dLL/dlog_coefs by the adjoint method (fsp_loglik_grad).
    counts: [n_times, n_bins] over the kept species (first fastest)."""
    model = Model(net)
    th = np.ascontiguousarray(log_coefs, np.float64)
    tt = np.ascontiguousarray(times, np.float64)
    km = np.ascontiguousarray([1 if k else 0 for k in keep_mask], np.int32)
    cn = np.ascontiguousarray(counts, np.float64)
    data = FspLikData(len(tt), tt.ctypes.data, km.ctypes.data, cn.ctypes.data, float(floor_val))
    opt = FspGradOptions(FORMATS[fmt], int(m), float(tol), int(quad_nodes), int(grading_levels))
    own = p0 is None
    pbuf = DeviceBuffer.from_host(_delta0(net.n_states)) if own else p0
    ll = C.c_double(0.0)
    try:
        if grad:
            g = np.zeros(net.n_reactions, np.float64)
            check(lib().fsp_loglik_grad(model.ref, th.ctypes.data, ptr(pbuf), C.byref(data),
                                        C.byref(opt), C.byref(ll), g.ctypes.data, stream))
            return ll.value, g
        check(lib().fsp_loglik(model.ref, th.ctypes.data, ptr(pbuf), C.byref(data), C.byref(opt),
                               C.byref(ll), stream))
        return ll.value
    finally:
        if own:
            pbuf.free()


def loglik_grad_batch(net, log_coefs, times, keep_mask, counts, floor_val=1e-12, p0=None,
                      fmt="csr", m=30, tol=1e-10, quad_nodes=8, grading_levels=10, workers=0):
    """LL and dLL/dlog_coefs at every row of log_coefs [points, R]
    (fsp_loglik_grad_batch): returns (ll [points], grad [points, R])."""
    model = Model(net)
    th = np.ascontiguousarray(log_coefs, np.float64)
    if th.ndim != 2 or th.shape[1] != net.n_reactions:
        raise ValueError("log_coefs must be [points, n_reactions]")
    tt = np.ascontiguousarray(times, np.float64)
    km = np.ascontiguousarray([1 if k else 0 for k in keep_mask], np.int32)
    cn = np.ascontiguousarray(counts, np.float64)
    data = FspLikData(len(tt), tt.ctypes.data, km.ctypes.data, cn.ctypes.data, float(floor_val))
    opt = FspGradOptions(FORMATS[fmt], int(m), float(tol), int(quad_nodes), int(grading_levels))
    own = p0 is None
    pbuf = DeviceBuffer.from_host(_delta0(net.n_states)) if own else p0
    ll = np.zeros(th.shape[0], np.float64)
    g = np.zeros(th.shape, np.float64)
    try:
        check(lib().fsp_loglik_grad_batch(model.ref, th.shape[0], th.ctypes.data, ptr(pbuf),
                                          C.byref(data), C.byref(opt), int(workers),
                                          ll.ctypes.data, g.ctypes.data))
        return ll, g
    finally:
        if own:
            pbuf.free()


def _delta0(n):
    p = np.zeros(n, np.float64)
    p[0] = 1.0
    return p


def _tcoef_callback(coef_fn, R):
    """ctypes trampoline for a Python time function t -> R coefficients."""
    def _cb(t, out, ctx):
        c = np.asarray(coef_fn(float(t)), np.float64)
        for r in range(R):
            out[r] = c[r]
    return TCOEF_CB(_cb)


def rhs_apply(A: "Mat", t, coef_fn, x, y, stream=None):
    """udot = A(t) u with coefficients coef_fn(t) (the CVODE RHS body)."""
    cb = _tcoef_callback(coef_fn, A.net.n_reactions)
    check(lib().fsp_rhs_apply(A.h, float(t), C.cast(cb, C.c_void_p), None, ptr(x), ptr(y), stream))


class Bdf:
    """fsp_bdf handle: dp/dt = A(t) p by variable-order BDF with GMRES stages."""

    def __init__(self, A: "Mat", coef_fn, t0=0.0, rtol=1e-6, atol=1e-12, max_order=5,
                 krylov_dim=30, max_restarts=5, h0=0.0, max_steps=500000, lin_tol_factor=0.05,
                 stream=None):
        self.A = A
        self._cb = _tcoef_callback(coef_fn, A.net.n_reactions)
        opt = FspBdfOptions(rtol, atol, max_order, krylov_dim, max_restarts, h0, max_steps,
                            lin_tol_factor)
        h = C.c_void_p()
        check(lib().fsp_bdf_create(A.h, C.cast(self._cb, C.c_void_p), None, float(t0),
                                   C.byref(opt), stream, C.byref(h)))
        self.h = h

    def advance(self, v, t_out):
        check(lib().fsp_bdf_advance(self.h, ptr(v), float(t_out)))

    def reset(self, t0):
        check(lib().fsp_bdf_reset(self.h, float(t0)))

    def stats(self) -> FspBdfStats:
        st = FspBdfStats()
        check(lib().fsp_bdf_get_stats(self.h, C.byref(st)))
        return st

    def close(self):
        if getattr(self, "h", None):
            check(lib().fsp_bdf_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Expv:
    """fsp_expv handle: KryExpvFSP on the device (src/KryExpvFSP.h:75-102)."""

    def __init__(self, A: Mat | None, n, t_final, m=30, tol=1e-8, iop=False, q_iop=2,
                 anorm=1.0, matvec=None, stream=None, basis=None):
        self._cb_keep = None
        if matvec is not None:
            def _cb(ctx, x, y):
                matvec(int(x), int(y))
            self._cb_keep = MATVEC_CB(_cb)
            cb = self._cb_keep
        else:
            cb = MATVEC_CB()
        h = C.c_void_p()
        check(lib().fsp_expv_create(None if A is None else A.h, cb, None, int(n), float(t_final),
                                    int(m), float(tol), int(iop), int(q_iop), float(anorm),
                                    stream, C.byref(h)))
        self.h = h
        self.A = A
        if basis is not None:
            # "f64": FP64 Krylov basis over a mixed (FP32-valued) generator
            check(lib().fsp_expv_set_basis_precision(self.h, PRECISIONS[basis]))

    def step(self, v):
        check(lib().fsp_expv_step(self.h, ptr(v)))

    def solve(self, v):
        check(lib().fsp_expv_solve(self.h, ptr(v)))

    def reset_time(self):
        check(lib().fsp_expv_reset_time(self.h))

    def set_final_time(self, t):
        check(lib().fsp_expv_set_final_time(self.h, float(t)))

    def restart(self, t_final):
        check(lib().fsp_expv_restart(self.h, float(t_final)))

    def final_time_reached(self) -> bool:
        r = C.c_int(0)
        check(lib().fsp_expv_final_time_reached(self.h, C.byref(r)))
        return bool(r.value)

    def stats(self) -> FspExpvStats:
        s = FspExpvStats()
        check(lib().fsp_expv_get_stats(self.h, C.byref(s)))
        return s

    def tau_history(self) -> np.ndarray:
        cnt = C.c_int64(0)
        check(lib().fsp_expv_get_tau_history(self.h, None, 0, C.byref(cnt)))
        out = np.empty(cnt.value, np.float64)
        if cnt.value:
            check(lib().fsp_expv_get_tau_history(self.h, out.ctypes.data, cnt.value,
                                                 C.byref(cnt)))
        return out

    def set_profiling(self, enable=True):
        check(lib().fsp_expv_set_profiling(self.h, 1 if enable else 0))

    def profile(self) -> dict:
        """{class: (launches, total ms, algorithmic bytes)} of the profiled steps."""
        p = FspExpvProfile()
        check(lib().fsp_expv_get_profile(self.h, C.byref(p)))
        return {name: (int(p.launches[k]), float(p.ms[k]), float(p.bytes[k]))
                for k, name in enumerate(PROF_CLASSES)}

    def set_limits(self, max_reject=1000, max_nstep=10000):
        check(lib().fsp_expv_set_limits(self.h, int(max_reject), int(max_nstep)))

    def close(self):
        if getattr(self, "h", None):
            check(lib().fsp_expv_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """fsp_batch handle: many independent solves v_i <- exp(t_i A_i) v_i.
    path: "auto" (the autotuner's choice, else "lockstep"), "lockstep" (one
    graph-captured batch step advances every running instance),
    "persistent" (one persistent work-queue kernel, each workgroup solving
    whole instances), "hybrid" (the persistent kernel for instances of at
    most split_states entries, lockstep for the rest, concurrently) or
    "streams" (single-solve engine, up to four instances at a time)."""

    def __init__(self, nets, t_final, with_sink=None, m=30, tol=1e-8, anorm=1.0,
                 precision="f64", iop=False, q_iop=2, lockstep_bpc=0, wg_per_cu=0,
                 max_reject=0, max_nstep=0, path="auto", split_states=0, time_limit_s=0.0,
                 lockstep_rows=0):
        self.nets = list(nets)
        count = len(self.nets)
        arr = (FspModel * count)()
        self._models = [Model(n) for n in self.nets]
        for i, mdl in enumerate(self._models):
            arr[i] = mdl.s
        self._arr = arr
        coefs = np.ascontiguousarray(np.concatenate([n.coefs for n in self.nets]), np.float64)
        tf = np.ascontiguousarray(t_final, np.float64)
        ws = None if with_sink is None else np.ascontiguousarray([1 if s else 0 for s in with_sink], np.int32)
        self._keep = [coefs, tf, ws]
        opt = FspBatchOptions(m, tol, anorm, PRECISIONS[precision], int(iop), int(q_iop),
                              int(lockstep_bpc), int(wg_per_cu), int(max_reject), int(max_nstep),
                              BATCH_PATHS[path], int(split_states), float(time_limit_s),
                              int(lockstep_rows))
        h = C.c_void_p()
        check(lib().fsp_batch_create(C.cast(arr, C.c_void_p), coefs.ctypes.data, tf.ctypes.data,
                                     None if ws is None else ws.ctypes.data, count, C.byref(opt),
                                     None, C.byref(h)))
        self.h = h
        self.count = count
        self.nv = [n.n_states + (1 if (with_sink is not None and with_sink[i]) else 0)
                   for i, n in enumerate(self.nets)]

    def set_initial(self, i, p0):
        p = np.ascontiguousarray(p0, np.float64)
        check(lib().fsp_batch_set_initial(self.h, int(i), p.ctypes.data))

    def solve(self, stream=None):
        check(lib().fsp_batch_solve(self.h, stream))

    def prepare(self):
        """Set-up of the solve path (buffers, generators, graphs) ahead of a
        timed solve."""
        check(lib().fsp_batch_prepare(self.h))

    def solution(self, i) -> np.ndarray:
        out = np.empty(self.nv[i], np.float64)
        check(lib().fsp_batch_get_solution(self.h, int(i), out.ctypes.data, out.size))
        return out

    def stats(self, i):
        st, rj, ss = C.c_int64(0), C.c_int64(0), C.c_int(0)
        check(lib().fsp_batch_get_stats(self.h, int(i), C.byref(st), C.byref(rj), C.byref(ss)))
        return st.value, rj.value, ss.value

    def config(self):
        b, w = C.c_int(0), C.c_int(0)
        check(lib().fsp_batch_get_config(self.h, C.byref(b), C.byref(w)))
        return b.value, w.value

    def tuning(self):
        """(path name, lockstep workgroups per CU, lockstep tile rows per
        thread, persistent workgroups per CU, split in states) of the next
        solve."""
        p, b, r, w, sp = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0), C.c_int64(0)
        check(lib().fsp_batch_get_tuning(self.h, C.byref(p), C.byref(b), C.byref(r), C.byref(w),
                                         C.byref(sp)))
        return BATCH_PATH_NAMES[p.value], b.value, r.value, w.value, sp.value

    def autotune(self, cache_path=None, force=False) -> float:
        ms = C.c_double(0.0)
        check(lib().fsp_batch_autotune(self.h, None if cache_path is None else cache_path.encode(),
                                       int(force), C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            check(lib().fsp_batch_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
