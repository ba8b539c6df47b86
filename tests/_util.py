"""Test helpers: device buffers through the engine's own C ABI (no torch on
the GPU path: one HIP runtime per process)."""
import numpy as np

from cuda_fsp_amd.capi import DeviceBuffer


def dev(a, dtype=None):
    """numpy -> device buffer (same dtype)."""
    return DeviceBuffer.from_host(a, dtype)


def host(x):
    return x.to_host()


def zeros(n, dtype=np.float64):
    return DeviceBuffer.zeros(int(n), dtype)


def l1(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).sum())
