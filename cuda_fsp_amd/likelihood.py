"""Histogram likelihood problems over FSP solutions, and the bridge into
torch.autograd.

The reference's cme::Likelihood (src/Likelihood.cpp:8-33) scores observed
cells against the FSP distribution at the output times.  Here the data are
joint histograms of the observed species at each time, the score is
LL = sum counts * log max(P, floor), and the gradient with respect to the
log rate constants comes from the library's adjoint (fsp_loglik_grad).

torch is used with CPU tensors only: the library runs the device work on its
own HIP runtime, so torch never initialises a device here.
"""
from __future__ import annotations

import numpy as np

from . import capi


def marginal(p: np.ndarray, bounds, keep_mask) -> np.ndarray:
    """Marginal of a state vector over the species with keep_mask[s] (first
    kept species fastest), on the host."""
    shape = tuple(int(b) + 1 for b in bounds)
    arr = np.asarray(p[: int(np.prod(shape))]).reshape(shape[::-1])
    drop = tuple(len(shape) - 1 - s for s, k in enumerate(keep_mask) if not k)
    return arr.sum(axis=drop).ravel() if drop else arr.ravel()


def sample_histograms(net, times, keep_mask, cells, seed, log_coefs=None, m=30, tol=1e-12):
    """Joint histograms of `cells` cells per time, drawn by multinomial
    sampling from the FSP marginal at the given parameters (device solve)."""
    coefs = np.exp(log_coefs) if log_coefs is not None else net.coefs
    A = capi.Mat(net.with_coefs(coefs), "csr")
    v = capi.DeviceBuffer.from_host(_delta0(net.n_states))
    E = capi.Expv(A, A.n_rows, float(times[0]), m=m, tol=tol)
    rng = np.random.default_rng(seed)
    out = []
    for t in times:
        E.set_final_time(float(t))
        E.solve(v)
        P = np.clip(marginal(v.to_host(), net.bounds, keep_mask), 0.0, None)
        out.append(rng.multinomial(cells, P / P.sum()).astype(np.float64))
    E.close()
    A.close()
    v.free()
    return np.stack(out)


def _delta0(n):
    p = np.zeros(n, np.float64)
    p[0] = 1.0
    return p


class LikelihoodProblem:
    """One network, observation times, observed species and histograms."""

    def __init__(self, net, times, keep_mask, counts, floor_val=1e-12, fmt="csr", m=30,
                 tol=1e-10, quad_nodes=8, grading_levels=10):
        self.net = net
        self.times = np.asarray(times, np.float64)
        self.keep_mask = list(keep_mask)
        self.counts = np.asarray(counts, np.float64)
        self.kw = dict(floor_val=floor_val, fmt=fmt, m=m, tol=tol, quad_nodes=quad_nodes,
                       grading_levels=grading_levels)

    def value(self, log_coefs) -> float:
        return capi.loglik(self.net, log_coefs, self.times, self.keep_mask, self.counts, **self.kw)

    def value_and_grad(self, log_coefs):
        return capi.loglik(self.net, log_coefs, self.times, self.keep_mask, self.counts,
                           grad=True, **self.kw)

    def value_and_grad_batch(self, log_coefs, workers=0):
        """LL and gradient at every row of log_coefs [points, R], several
        points at a time on their own streams."""
        return capi.loglik_grad_batch(self.net, log_coefs, self.times, self.keep_mask,
                                      self.counts, workers=workers, **self.kw)


def torch_loglik(problem: LikelihoodProblem, theta):
    """LL(theta) as a differentiable torch scalar (theta: CPU tensor of log
    rate constants); backward uses the adjoint gradient."""
    import torch

    class _FspLogLik(torch.autograd.Function):
        @staticmethod
        def forward(ctx, th):
            ll, g = problem.value_and_grad(th.detach().to("cpu", torch.float64).numpy())
            ctx.save_for_backward(torch.from_numpy(g))
            return th.new_tensor(ll)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out * g.to(grad_out.dtype)

    return _FspLogLik.apply(theta)
