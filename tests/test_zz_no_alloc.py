"""fsp_loglik_indices_noalloc (the allocation-free entry point) against
fsp_loglik_indices through the C ABI: bit-identical log-likelihoods on two
calls in a row, the second with a smaller input (the shared scratch keeps the
first call's partials, which the second must not read)."""
import numpy as np
import pytest

from cuda_fsp_amd import capi
from tests._util import dev

pytestmark = pytest.mark.gpu


def _case(rng, n, count, floor_val):
    p = rng.random(n) * 1e-3
    p[rng.integers(0, n, n // 10)] = 0.0        # below the floor
    idx = rng.integers(-1, n, count).astype(np.int32)   # -1: outside the box
    return dev(p), dev(idx), count, floor_val


def test_noalloc_matches_allocating_entry_point():
    rng = np.random.default_rng(12)
    for n, count, floor_val in [(1_000_000, 3_000_000, 1e-12), (5_000, 700, 1e-44)]:
        d_p, d_idx, cnt, fl = _case(rng, n, count, floor_val)
        print(f"[case] noalloc loglik n={n} count={cnt}", flush=True)
        want = capi.loglik_indices(d_p, d_idx, cnt, fl)
        got = capi.loglik_indices_noalloc(d_p, d_idx, cnt, fl)
        assert np.isfinite(want)
        assert got == want, (n, count, got, want)
    assert capi.loglik_indices_noalloc(d_p, d_idx, 0) == 0.0
