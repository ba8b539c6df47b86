"""C ABI checks that need no GPU: the library loads, exports every entry
point include/fsp_capi.h declares, the Python binding covers them, and the
argument / overflow error paths return statuses instead of exiting."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from cuda_fsp_amd import capi, models as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fsp_capi.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fsp_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    lib = capi.lib()
    names = _declared()
    assert len(names) > 30
    for name in names:
        assert hasattr(lib, name), name


def test_binding_covers_header():
    assert set(_declared()) <= set(capi.declared_symbols())


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(capi, "_lib", None)
    monkeypatch.setattr(capi, "LIB_PATH", "/nonexistent/libfsp_hip.so")
    with pytest.raises(capi.FspError):
        capi.lib()


def test_num_states_and_overflow():
    assert capi.num_states((3, 31, 31, 63, 63)) == 16_777_216
    with pytest.raises(capi.FspError) as e:
        capi.num_states((65535, 65535))
    assert e.value.status == 3  # FSP_ERR_OVERFLOW
    with pytest.raises(capi.FspError) as e:
        capi.num_states((4, -1))
    assert e.value.status == 1


def _mat_status(net, fmt=capi.FSP_CSR, prec=capi.FSP_F64, factors=True):
    m = capi.Model(net)
    if not factors:
        m.s.factors = None
    c = np.ascontiguousarray(net.coefs)
    h = C.c_void_p()
    return capi.lib().fsp_mat_create(C.byref(m.s), fmt, prec, 0, c.ctypes.data, None, C.byref(h))


def test_model_validation_errors():
    net = M.toggle_switch((7, 7))
    assert _mat_status(net, fmt=7) == 2                 # unsupported format
    assert _mat_status(net, prec=9) == 2                # unsupported precision
    assert _mat_status(net, factors=False) == 1          # no propensity source
    big = M.ReactionNetwork("x", np.full(9, 1, np.int32), np.eye(9, dtype=np.int32),
                            np.ones((9, 18)), np.ones(9))
    assert _mat_status(big) == 2                         # > 8 species
    huge = M.toggle_switch((65535, 65535))
    assert _mat_status(huge) == 3                        # > 2^31-1 states


@pytest.mark.skipif(capi.device_count() > 0, reason="a GPU is visible")
def test_no_device_is_a_status():
    assert _mat_status(M.toggle_switch((7, 7))) == 7      # FSP_ERR_NO_DEVICE


def test_expv_argument_errors():
    h = C.c_void_p()
    L = capi.lib()
    cb = capi.MATVEC_CB()
    assert L.fsp_expv_create(None, cb, None, 10, 1.0, 30, 1e-8, 0, 2, 1.0, None, C.byref(h)) == 1
    good = capi.MATVEC_CB(lambda c, x, y: None)
    assert L.fsp_expv_create(None, good, None, 10, 1.0, 51, 1e-8, 0, 2, 1.0, None, C.byref(h)) == 2
    assert L.fsp_expv_create(None, good, None, 10, 1.0, 30, -1.0, 0, 2, 1.0, None, C.byref(h)) == 1
    assert L.fsp_expv_create(None, good, None, 0, 1.0, 30, 1e-8, 0, 2, 1.0, None, C.byref(h)) == 1


def test_batched_gradient_argument_errors():
    L = capi.lib()
    net = M.toggle_switch((7, 7))
    model = capi.Model(net)
    th = np.zeros((2, net.n_reactions))
    ll = np.zeros(2)
    g = np.zeros((2, net.n_reactions))
    # NULL data / negative point count are refused before any device work
    assert L.fsp_loglik_grad_batch(model.ref, 2, th.ctypes.data, None, None, None, 0,
                                   ll.ctypes.data, g.ctypes.data) == 1
    fake = C.c_void_p(16)
    assert L.fsp_loglik_grad_batch(model.ref, -1, th.ctypes.data, fake, fake, None, 0,
                                   ll.ctypes.data, g.ctypes.data) == 1
    assert L.fsp_loglik_grad_batch(model.ref, 0, th.ctypes.data, fake, fake, None, 0,
                                   ll.ctypes.data, g.ctypes.data) == 0


def test_fastdiv_exact():
    # q / d == (q * mul) >> shift for every q < 2^31 (Granlund-Montgomery);
    # checked for every radix the models use and edge quotients
    for d in list(range(1, 300)) + [511, 512, 513, 1024, 1025, 4096, 65535, 65536, 2**20 + 7]:
        l = 0
        while (1 << l) < d:
            l += 1
        shift = 31 + l
        mul = ((1 << shift) + d - 1) // d
        assert mul < 2**32
        edge = [2**31 - 1, 2**31 - 2, d * 12345, d * 12345 - 1, (2**31 - 1) // d * d,
                (2**31 - 1) // d * d - 1]
        qs = np.concatenate([np.arange(0, 2000), np.random.default_rng(d).integers(0, 2**31, 2000),
                             np.array([q for q in edge if 0 <= q < 2**31])]).astype(object)
        for q in qs:
            assert (int(q) * mul) >> shift == int(q) // d
