"""Histogram log-likelihood (intended semantics of src/Likelihood.cpp:8-33 on
joint marginal histograms) and its adjoint gradient.

Pins: LL against the oracle's Krylov restatement plus a host marginal; the
adjoint gradient against exact forward sensitivities from the dense
block-triangular exponential exp(t [[A, A_r], [0, A]]) (Van Loan 1978) on a
tiny box, and against central finite differences (relative step 1e-5 on the
log rates, bound 1e-5 relative) on a larger one; the torch.autograd bridge
against the C ABI."""
import os
import sys

import numpy as np
import pytest
import scipy.linalg as sl

from cuda_fsp_amd import capi, models as M
from cuda_fsp_amd import likelihood as L
from oracle import oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
from make_golden import dense_generator  # noqa: E402

pytestmark = pytest.mark.gpu

KEEP = [0, 1, 1]  # (n, m) of the gene-expression network


def _data(net, times, cells, seed):
    return L.sample_histograms(net, times, KEEP, cells, seed)


def _host_ll(P_list, counts, floor_val):
    return sum(float((c * np.log(np.maximum(P, floor_val))).sum()) for P, c in zip(P_list, counts))


def test_loglik_matches_oracle():
    net = M.gene_expression((2, 12, 12))
    times = [1.0, 2.5, 6.0]
    counts = _data(net, times, 300, seed=2)
    print(f"[case] loglik gene 3x13x13 times={times}", flush=True)
    ll = capi.loglik(net, np.log(net.coefs), times, KEEP, counts, tol=1e-12)
    opr = O.Operator(net, "csr")
    rc, out, _ = O.expv_outputs(opr, M.delta_p0(net), times, tol=1e-12)
    assert rc == 0
    ref = _host_ll([L.marginal(o, net.bounds, KEEP) for o in out], counts, 1e-12)
    assert abs(ll - ref) <= 1e-9 * abs(ref)


def _exact_gradient(net, times, counts, floor_val=1e-12):
    """dLL/dlog c_r from exp(t [[A, A_r], [0, A]]) [0; p0] = [dp/dc_r; p]."""
    A = dense_generator(net, sink=False)
    n = A.shape[0]
    p0 = np.zeros(n)
    p0[0] = 1.0
    P = [sl.expm(t * A) @ p0 for t in times]
    grad = np.zeros(net.n_reactions)
    for r in range(net.n_reactions):
        Ar = dense_generator(net.with_coefs(np.eye(net.n_reactions)[r]), sink=False)
        Big = np.block([[A, Ar], [np.zeros_like(A), A]])
        for k, t in enumerate(times):
            z = sl.expm(t * Big) @ np.concatenate([np.zeros(n), p0])
            dp = L.marginal(z[:n], net.bounds, KEEP)
            Pm = L.marginal(P[k], net.bounds, KEEP)
            w = np.where(Pm > floor_val, counts[k] / np.where(Pm > floor_val, Pm, 1.0), 0.0)
            grad[r] += net.coefs[r] * float(w @ dp)
    ll = _host_ll([L.marginal(p, net.bounds, KEEP) for p in P], counts, floor_val)
    return ll, grad


def test_adjoint_gradient_matches_exact_sensitivities():
    net = M.gene_expression((2, 6, 5))
    times = [0.5, 1.5, 3.0]
    counts = _data(net, times, 200, seed=3)
    print("[case] adjoint gradient gene 3x7x6 vs dense forward sensitivities", flush=True)
    ll, g = capi.loglik(net, np.log(net.coefs), times, KEEP, counts, tol=1e-13, grad=True)
    ll_ref, g_ref = _exact_gradient(net, times, counts)
    print(f"[adj] LL {ll:.12g} vs {ll_ref:.12g}; grad {g} vs {g_ref}", flush=True)
    assert abs(ll - ll_ref) <= 1e-10 * abs(ll_ref)
    assert np.max(np.abs(g - g_ref)) <= 1e-7 * np.max(np.abs(g_ref))


def test_adjoint_gradient_matches_central_differences():
    net = M.gene_expression((2, 20, 20))
    times = [2.0, 5.0, 10.0]
    theta = np.log(net.coefs)
    counts = _data(net, times, 500, seed=4)
    prob = L.LikelihoodProblem(net, times, KEEP, counts, tol=1e-14)
    print("[case] adjoint gradient gene 3x21x21 vs central differences", flush=True)
    ll, g = prob.value_and_grad(theta)
    fd = np.zeros_like(theta)
    for r in range(len(theta)):
        h = 1e-5 * max(1.0, abs(theta[r]))
        tp, tm = theta.copy(), theta.copy()
        tp[r] += h
        tm[r] -= h
        fd[r] = (prob.value(tp) - prob.value(tm)) / (2 * h)
    err = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    print(f"[adj] relative error vs central FD {err:.3e}", flush=True)
    assert err <= 1e-5


def test_batched_gradients_equal_single_points():
    net = M.gene_expression((2, 10, 9))
    times = [1.0, 2.5]
    counts = _data(net, times, 150, seed=6)
    prob = L.LikelihoodProblem(net, times, KEEP, counts, tol=1e-12, grading_levels=4)
    rng = np.random.default_rng(7)
    thetas = np.log(net.coefs)[None, :] + rng.uniform(-0.5, 0.5, (5, net.n_reactions))
    print("[case] batched gradients, 5 points on 3 workers, gene 3x11x10", flush=True)
    ll, g = prob.value_and_grad_batch(thetas, workers=3)
    for i in range(len(thetas)):
        lli, gi = prob.value_and_grad(thetas[i])
        assert ll[i] == lli
        assert np.array_equal(g[i], gi)


def test_gradient_with_more_than_16_reactions():
    # every reaction twice at half its rate: the same generator, R = 18
    # (two reaction chunks in the sensitivity kernel); each half-rate copy's
    # dLL/dlog c is half of the original reaction's
    net = M.gene_expression((2, 9, 8))
    times = [1.0, 3.0]
    counts = _data(net, times, 150, seed=8)
    dup = M.ReactionNetwork("dup", net.bounds, np.concatenate([net.nu, net.nu]),
                            np.concatenate([net.factors, net.factors]),
                            np.concatenate([0.5 * net.coefs, 0.5 * net.coefs]))
    print("[case] gradient with 18 reactions (duplicated gene 3x10x9)", flush=True)
    ll, g = capi.loglik(net, np.log(net.coefs), times, KEEP, counts, tol=1e-12, grad=True)
    ll2, g2 = capi.loglik(dup, np.log(dup.coefs), times, KEEP, counts, tol=1e-12, grad=True)
    R = net.n_reactions
    assert abs(ll2 - ll) <= 1e-10 * abs(ll)
    assert np.max(np.abs(g2[:R] - 0.5 * g)) <= 1e-8 * np.max(np.abs(g))
    assert np.max(np.abs(g2[R:] - 0.5 * g)) <= 1e-8 * np.max(np.abs(g))


def test_gradient_recomputed_nodes_agree(monkeypatch):
    # node vectors kept from the forward pass (the default when they fit) and
    # recomputed per interval from checkpoints give the same gradient
    net = M.gene_expression((2, 10, 10))
    times = [1.0, 2.5, 4.0]
    counts = _data(net, times, 200, seed=9)
    print("[case] gradient, stored vs recomputed quadrature nodes, gene 3x11x11", flush=True)
    ll, g = capi.loglik(net, np.log(net.coefs), times, KEEP, counts, tol=1e-12, grad=True)
    monkeypatch.setenv("FSP_ADJOINT_RECOMPUTE", "1")
    ll2, g2 = capi.loglik(net, np.log(net.coefs), times, KEEP, counts, tol=1e-12, grad=True)
    assert abs(ll2 - ll) <= 1e-10 * abs(ll)
    assert np.max(np.abs(g2 - g)) <= 1e-8 * np.max(np.abs(g))


def test_torch_autograd_bridge():
    import torch

    net = M.gene_expression((2, 8, 8))
    times = [1.0, 3.0]
    counts = _data(net, times, 100, seed=5)
    prob = L.LikelihoodProblem(net, times, KEEP, counts, tol=1e-12)
    theta = torch.tensor(np.log(net.coefs), dtype=torch.float64, requires_grad=True)
    ll = L.torch_loglik(prob, theta)
    (2.0 * ll).backward()
    ll_ref, g_ref = prob.value_and_grad(np.log(net.coefs))
    assert float(ll.detach()) == pytest.approx(ll_ref, rel=1e-14)
    assert np.allclose(theta.grad.numpy(), 2.0 * g_ref, rtol=1e-14, atol=0)
