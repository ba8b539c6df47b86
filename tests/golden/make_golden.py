"""Writes the golden fixtures in tests/golden/ (run in this container; scipy
1.15 is the ground truth).  Nothing here reads the reference.

fixtures.npz:
  expm_*      random upper-Hessenberg matrices (the shape Krylov steps
              exponentiate, N = 32 = m + 2) at several norms, and
              scipy.linalg.expm of each
  cme_*       tiny FSP boxes: the dense generator built from the CME
              definition (column = source state, outflow leaving the box to
              a sink state n), p(t) = expm(t A) p0 at several times
  bdf_*       a time-varying generator A(t) = sum_r c_r(t) A_r on a tiny box,
              integrated by scipy solve_ivp (DOP853, rtol 1e-12): the
              reference's time-varying toggle coefficients, and the
              gene-expression network with a sin-modulated G2 -> G1 rate
"""
import os
import sys

import numpy as np
import scipy.linalg as sl
from scipy.integrate import solve_ivp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from cuda_fsp_amd import models as M  # noqa: E402


def dense_generator(net, coefs=None, sink=True):
    """CME generator from its definition: for each state x and reaction r,
    rate c_r a_r(x) leaves x; it enters x + nu_r when that is in the box and
    the sink state n otherwise.  Independent of the row-wise (inflow)
    construction of the oracle and the GPU kernels."""
    c = net.coefs if coefs is None else np.asarray(coefs)
    b = net.bounds.astype(int)
    n = net.n_states
    N = n + 1 if sink else n
    A = np.zeros((N, N))
    strides = np.cumprod(np.concatenate([[1], b[:-1] + 1]))
    for j in range(n):
        x = (j // strides) % (b + 1)
        for r in range(net.n_reactions):
            a = c[r]
            for s in range(net.n_species):
                a *= net.factor(r, s)[x[s]]
            if net.nu[r].any() == 0:
                continue
            A[j, j] -= a
            y = x + net.nu[r]
            if np.all(y >= 0) and np.all(y <= b):
                A[int(np.dot(y, strides)), j] += a
            elif sink:
                A[n, j] += a
    return A


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    # dense expm fixtures
    for k, scale in enumerate([0.05, 0.7, 3.0, 25.0, 140.0]):
        H = np.triu(rng.standard_normal((32, 32)), -1) * scale
        out[f"expm_A{k}"] = H
        out[f"expm_F{k}"] = sl.expm(H)
    # tiny CME boxes
    cases = {
        "toggle": (M.toggle_switch((9, 8)), [0.3, 1.0, 2.5]),
        "gene": (M.gene_expression((2, 6, 5)), [0.5, 4.0, 20.0]),
        "bd": (M.birth_death((12, 7), k=2.0, g=0.7), [0.25, 1.5, 6.0]),
        "chain": (M.five_species_chain((2, 3, 2, 3, 2)), [0.5, 2.0]),
    }
    for name, (net, times) in cases.items():
        A = dense_generator(net)
        p0 = np.zeros(net.n_states + 1)
        p0[0] = 1.0
        P = np.stack([sl.expm(t * A) @ p0 for t in times])
        out[f"cme_{name}_A"] = A
        out[f"cme_{name}_t"] = np.asarray(times, float)
        out[f"cme_{name}_p"] = P
    # time-varying toggle (coefficients as the reference's CVODE test,
    # tests/test_sundials.cu:55-64, on a tiny box and short times)
    net = M.toggle_ref_test((6, 6))
    base = net.coefs.copy()

    def coefs_t(t):
        c = base.copy()
        c[1] = base[1] * (1.0 + np.cos(t / 1000.0))
        c[2] = base[2] * np.exp(-t / 3600.0)
        c[5] = base[5] * np.exp(-t / 3600.0)
        return c

    terms = [dense_generator(net.with_coefs(np.eye(6)[r]), sink=False) for r in range(6)]
    p0 = np.zeros(net.n_states)
    p0[0] = 1.0
    t_eval = np.array([600.0, 1800.0, 3600.0])
    sol = solve_ivp(lambda t, p: sum(ci * Ti for ci, Ti in zip(coefs_t(t), terms)) @ p,
                    (0.0, t_eval[-1]), p0, method="DOP853", rtol=1e-12, atol=1e-14,
                    t_eval=t_eval)
    out["bdf_toggle_t"] = t_eval
    out["bdf_toggle_p"] = sol.y.T
    # the gene-expression network with the time-varying promoter rate
    # G2 -> G1 = 0.5 (1 + sin(2 pi t / 30)), on a small box with the sink
    # state, at several output times up to t = 60
    net = M.gene_expression((2, 12, 12))
    gterms = [dense_generator(net.with_coefs(np.eye(9)[r]), sink=True) for r in range(9)]

    def gene_coefs(t):
        c = net.coefs.copy()
        c[3] = 0.5 * (1.0 + np.sin(2.0 * np.pi * t / 30.0))
        return c

    p0 = np.zeros(net.n_states + 1)
    p0[0] = 1.0
    t_eval = np.array([2.0, 10.0, 24.0, 40.0, 60.0])
    sol = solve_ivp(lambda t, p: sum(ci * Ti for ci, Ti in zip(gene_coefs(t), gterms)) @ p,
                    (0.0, t_eval[-1]), p0, method="DOP853", rtol=1e-12, atol=1e-15,
                    t_eval=t_eval)
    assert sol.success
    out["bdf_gene_sin_t"] = t_eval
    out["bdf_gene_sin_p"] = sol.y.T
    np.savez_compressed(os.path.join(HERE, "fixtures.npz"), **out)
    print("wrote", os.path.join(HERE, "fixtures.npz"), sorted(out))


if __name__ == "__main__":
    main()
