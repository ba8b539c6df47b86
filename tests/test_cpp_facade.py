"""The C++ drop-in facade (cuFSP::, cme::): test programs that mirror the
reference's own tests, compiled against libcufsp.so and run on the GPU,
including the reference's device PropFun route (a __device__ propensity
fetched with hipMemcpyFromSymbol, built with hipcc -fgpu-rdc)."""
import os
import subprocess

import numpy as np
import pytest
import scipy.stats as ss

from cuda_fsp_amd import build as B, models as M

HERE = os.path.dirname(os.path.abspath(__file__))
PROGRAMS = ["test_fsp_get_states", "test_fsp_mat", "test_kry_expv", "test_cme", "test_fsp_expand"]


def _exe(name, tmp_path_factory):
    out = str(tmp_path_factory.mktemp("cpp") / name)
    B.compile_cpp_program(os.path.join(HERE, "cpp", name + ".cpp"), out)
    return out


@pytest.mark.parametrize("name", PROGRAMS)
def test_cpp_programs_compile(name, tmp_path_factory):
    assert os.path.exists(_exe(name, tmp_path_factory))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["test_fsp_get_states", "test_fsp_mat", "test_kry_expv",
                                  "test_fsp_expand"])
def test_cpp_reference_tests(name, tmp_path_factory):
    exe = _exe(name, tmp_path_factory)
    print(f"[case] {name}", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0


def _device_propfun_exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hip") / "test_fsp_mat_device")
    B.compile_hip_program([os.path.join(HERE, "cpp", "test_fsp_mat_device.hip"),
                           os.path.join(HERE, "cpp", "toggle_device.hip")], out)
    return out


def test_device_propfun_program_compiles(tmp_path_factory):
    # the reference's device PropFun route, the propensity in a second
    # translation unit: hipcc -fgpu-rdc against the facade
    assert os.path.exists(_device_propfun_exe(tmp_path_factory))


@pytest.mark.gpu
def test_device_propfun_route(tmp_path_factory):
    exe = _device_propfun_exe(tmp_path_factory)
    print("[case] test_fsp_mat_device", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "device PropFun test done" in r.stdout


@pytest.mark.gpu
def test_cpp_cme_layer(tmp_path_factory):
    from oracle import oracle as O
    exe = _exe("test_cme", tmp_path_factory)
    out = str(tmp_path_factory.mktemp("cme") / "cme.bin")
    print("[case] test_cme", flush=True)
    r = subprocess.run([exe, out], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    raw = open(out, "rb").read()
    nt, n, nc = np.frombuffer(raw[:12], np.int32)
    off = 12
    P = np.frombuffer(raw[off:off + 8 * nt * n], np.float64).reshape(nt, n)
    off += 8 * nt * n
    idx = np.frombuffer(raw[off:off + 4 * nt * nc], np.int32).reshape(nt, nc)
    off += 4 * nt * nc
    ll_f, ll_t = np.frombuffer(raw[off:off + 16], np.float64)
    coefs = np.frombuffer(raw[off + 16:off + 16 + 72], np.float64)  # exactly what t_fun gave
    # FSPSolver solution vs the oracle (m = 30, IOP q = 2, tol 1e-8)
    net = M.gene_expression((2, 40, 40), rates=coefs)
    rp, c, v = O.merged_csr(net)
    p0 = M.delta_p0(net)
    rc, want, st = O.expv_csr_outputs(rp, c, v, p0, [1.0, 5.0, 20.0], iop=True, q_iop=2)
    for k in range(nt):
        assert np.abs(P[k] - want[k]).sum() <= 1e-10
        assert P[k].sum() > 1 - 1e-6
    # Likelihood::evaluate = sum log max(1e-44, p[idx]) (src/Likelihood.cpp:14-20)
    ll = sum(np.log(np.maximum(1e-44, np.where(idx[k] >= 0, P[k][np.maximum(idx[k], 0)], 0.0))).sum()
             for k in range(nt))
    assert abs(ll - ll_f) <= 1e-9 * abs(ll)
    # SSA samples follow the FSP distribution (chi-square on the n marginal)
    for k in range(nt):
        pn = P[k].reshape(41, 41, 3).sum(axis=(0, 2))
        nn = (idx[k] // 3) % 41
        obs = np.bincount(nn, minlength=41)
        exp = pn * nc
        keep = exp >= 5
        o = np.append(obs[keep], obs[~keep].sum())
        e = np.append(exp[keep], exp[~keep].sum())
        e *= o.sum() / e.sum()
        stat = ((o - e) ** 2 / np.maximum(e, 1e-300)).sum()
        assert ss.chi2.sf(stat, max(len(o) - 1, 1)) > 1e-4
