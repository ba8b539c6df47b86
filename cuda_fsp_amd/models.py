"""Reaction-network descriptions for the FSP solve path.

A network is the data the reference's ``FSPMat`` constructors take
(reference ``src/FSPMat.h:43-52``): the FSP box bounds, the stoichiometry as a
reactions x species CSR (``CSRMatInt``, ``src/cme_util.h:35-40``), the
per-reaction time coefficients that ``t_func`` writes (``TcoefFun``,
``src/FSPMat.h:16``), and the propensity.  Propensities are given in the
separable form the reference's Kronecker path uses (``PropFactorFun``,
``src/fspmat_kron_kernels.h:13``): ``a_r(x) = prod_s f_{r,s}(x_s)``, stored as
the factor tables ``factors[r, off_s + x_s]`` laid out as
``fspmat_sdkron_fill_host_data`` lays them out (``src/fspmat_kron_kernels.cu:
170-190``).

The builders below define the workloads used by tests and the benchmark:

* ``toggle_ref_test`` -- the 6-reaction toggle switch of the reference tests
  (``tests/test_fsp_mat.cu:19-97``), any box.
* ``toggle_switch`` -- the 4-reaction toggle switch on a 128 x 128 box:
  0->x at 50/(1+y^2), x->0 at x, 0->y at 16/(1+x^2), y->0 at y.
* ``gene_expression`` -- promoter G in {0,1,2}, nascent n, mature m
  (9 terms, 3 x 512 x 512 = 786,432 states by default).
* ``five_species_chain`` -- G in 0..3 and four counts bounded 31, 31, 63, 63
  (16,777,216 states, 14 reactions, rates log-uniform in [1e-3, 10]).
* ``hog1p`` -- the reference's 5-species Hog1p problem on the box
  {3, 50, 50, 60, 60} (38,713,284 states, 9 reactions, one time-varying
  coefficient, ``hog1p_coefs``).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable, Sequence

import numpy as np


@dataclasses.dataclass
class ReactionNetwork:
    name: str
    bounds: np.ndarray          # int32 [d]
    nu: np.ndarray              # int32 [R, d] net change per reaction
    factors: np.ndarray         # float64 [R, sum(N_s + 1)]
    coefs: np.ndarray           # float64 [R] rate constants (t_func output)
    param_names: tuple = ()

    @property
    def n_species(self) -> int:
        return int(self.bounds.shape[0])

    @property
    def n_reactions(self) -> int:
        return int(self.nu.shape[0])

    @property
    def n_states(self) -> int:
        return int(np.prod(self.bounds.astype(np.int64) + 1))

    @property
    def factor_offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.bounds + 1)[:-1]]).astype(np.int64)

    def factor(self, r: int, s: int) -> np.ndarray:
        o = int(self.factor_offsets[s])
        return self.factors[r, o:o + int(self.bounds[s]) + 1]

    def stoich_csr(self):
        """Reactions x species CSR of the net changes, species ascending."""
        vals, cols, rp = [], [], [0]
        for r in range(self.n_reactions):
            for s in range(self.n_species):
                if self.nu[r, s] != 0:
                    vals.append(int(self.nu[r, s]))
                    cols.append(s)
            rp.append(len(vals))
        return (np.asarray(vals, np.int32), np.asarray(cols, np.int32),
                np.asarray(rp, np.int32))

    def with_bounds(self, bounds: Sequence[int]) -> "ReactionNetwork":
        """Same network on another box (factor tables re-tabulated by the
        builder that made it is not possible here, so this only shrinks)."""
        bounds = np.asarray(bounds, np.int32)
        assert np.all(bounds <= self.bounds)
        offs = self.factor_offsets
        cols = np.concatenate([np.arange(offs[s], offs[s] + bounds[s] + 1)
                               for s in range(self.n_species)])
        return ReactionNetwork(self.name, bounds, self.nu.copy(),
                               np.ascontiguousarray(self.factors[:, cols]),
                               self.coefs.copy(), self.param_names)

    def with_coefs(self, coefs) -> "ReactionNetwork":
        return dataclasses.replace(self, coefs=np.asarray(coefs, np.float64).copy())


def _tabulate(bounds, R, fn: Callable[[int, int, np.ndarray], np.ndarray]):
    """factors[r, off_s + i] = fn(r, s, i) over the box, species-major."""
    bounds = np.asarray(bounds, np.int32)
    stride = int(np.sum(bounds + 1))
    F = np.ones((R, stride), np.float64)
    off = 0
    for s, N in enumerate(bounds):
        i = np.arange(N + 1)
        for r in range(R):
            F[r, off:off + N + 1] = fn(r, s, i)
        off += N + 1
    return F


# Parameters of the reference test toggle (tests/test_fsp_mat.cu:19-21).
TOGGLE_REF = dict(ayx=2.6e-3, axy=6.1e-3, nyx=3.0, nxy=2.1, kx0=2.2e-3, kx=1.7e-2,
                  dx=3.8e-4, ky0=6.8e-5, ky=1.6e-2, dy=3.8e-4)


def toggle_ref_test(bounds=(31, 31)) -> ReactionNetwork:
    """The reference tests' 6-reaction toggle (tests/test_fsp_mat.cu:23-114):
    stoichiometry {x+, x+, x-, y+, y+, y-}; factors as toggle_propensity_factor
    (:52-84); coefficients as t_func (:88-97)."""
    p = TOGGLE_REF
    nu = np.array([[1, 0], [1, 0], [-1, 0], [0, 1], [0, 1], [0, -1]], np.int32)

    def fn(r, s, i):
        x = i.astype(np.float64)
        one = np.ones_like(x)
        if r == 1 and s == 1:
            return 1.0 / (1.0 + p["ayx"] * np.power(1.0 * x, p["nyx"]))
        if r == 2 and s == 0:
            return 1.0 * x
        if r == 4 and s == 0:
            return 1.0 / (1.0 + p["axy"] * np.power(1.0 * x, p["nxy"]))
        if r == 5 and s == 1:
            return 1.0 * x
        return one

    coefs = np.array([p["kx0"], p["kx"], p["dx"], p["ky0"], p["ky"], p["dy"]])
    return ReactionNetwork("toggle_ref_test", np.asarray(bounds, np.int32), nu,
                           _tabulate(bounds, 6, fn), coefs,
                           ("kx0", "kx", "dx", "ky0", "ky", "dy"))


def toggle_switch(bounds=(127, 127), rates=(50.0, 1.0, 16.0, 1.0)) -> ReactionNetwork:
    """Two-species toggle switch: 0->x at k0/(1+y^2), x->0 at k1 x,
    0->y at k2/(1+x^2), y->0 at k3 y.  Index x + (Nx+1) y."""
    nu = np.array([[1, 0], [-1, 0], [0, 1], [0, -1]], np.int32)

    def fn(r, s, i):
        x = i.astype(np.float64)
        one = np.ones_like(x)
        if r == 0 and s == 1:
            return 1.0 / (1.0 + x * x)
        if r == 1 and s == 0:
            return x
        if r == 2 and s == 0:
            return 1.0 / (1.0 + x * x)
        if r == 3 and s == 1:
            return x
        return one

    return ReactionNetwork("toggle_switch", np.asarray(bounds, np.int32), nu,
                           _tabulate(bounds, 4, fn), np.asarray(rates, np.float64),
                           ("k_x", "g_x", "k_y", "g_y"))


GENE_RATES = (0.05, 0.1, 0.02, 0.5, 5.0, 50.0, 2.0, 0.01, 0.1)
GENE_NAMES = ("k01", "k10", "k12", "k21", "r1", "r2", "k_nm", "g_n", "g_m")


def gene_expression(bounds=(2, 511, 511), rates=GENE_RATES) -> ReactionNetwork:
    """Promoter G in {0,1,2}, nascent mRNA n, mature mRNA m.
    Terms: G0->G1, G1->G0, G1->G2, G2->G1, 0->n at r_1 [G=1],
    0->n at r_2 [G=2], n->m at k n, n->0 at g_n n, m->0 at g_m m.  Production
    gated by G is split into one term per active promoter state so that each
    rate multiplies one fixed generator term."""
    nu = np.array([[1, 0, 0], [-1, 0, 0], [1, 0, 0], [-1, 0, 0],
                   [0, 1, 0], [0, 1, 0], [0, -1, 1], [0, -1, 0], [0, 0, -1]], np.int32)
    gate = {0: 0, 1: 1, 2: 1, 3: 2, 4: 1, 5: 2}

    def fn(r, s, i):
        x = i.astype(np.float64)
        if s == 0 and r in gate:
            return (i == gate[r]).astype(np.float64)
        if s == 1 and r in (6, 7):
            return x
        if s == 2 and r == 8:
            return x
        return np.ones_like(x)

    return ReactionNetwork("gene_expression", np.asarray(bounds, np.int32), nu,
                           _tabulate(bounds, 9, fn), np.asarray(rates, np.float64),
                           GENE_NAMES)


def five_species_chain(bounds=(3, 31, 31, 63, 63), seed=0) -> ReactionNetwork:
    """G in 0..3 and counts x1..x4.  14 mass-action reactions: six promoter
    switches G_k <-> G_{k+1}; production of x1 at rate k G (gated by the
    promoter level); conversions x1->x2->x3->x4; first-order decay of each
    count.  Rates log-uniform in [1e-3, 10] from ``seed``."""
    d = 5
    R = 14
    nu = np.zeros((R, d), np.int32)
    kinds = []
    r = 0
    for k in range(3):
        nu[r, 0] = 1; kinds.append(("on", k)); r += 1
        nu[r, 0] = -1; kinds.append(("off", k + 1)); r += 1
    nu[r, 1] = 1; kinds.append(("prod", 0)); r += 1
    for s in (1, 2, 3):
        nu[r, s] = -1; nu[r, s + 1] = 1; kinds.append(("conv", s)); r += 1
    for s in (1, 2, 3, 4):
        nu[r, s] = -1; kinds.append(("decay", s)); r += 1
    assert r == R

    def fn(rr, s, i):
        kind, a = kinds[rr]
        x = i.astype(np.float64)
        if kind in ("on", "off"):
            return (i == a).astype(np.float64) if s == 0 else np.ones_like(x)
        if kind == "prod":
            return x if s == 0 else np.ones_like(x)
        if kind in ("conv", "decay"):
            return x if s == a else np.ones_like(x)
        raise AssertionError

    rng = np.random.default_rng(seed)
    rates = np.exp(rng.uniform(math.log(1e-3), math.log(10.0), R))
    return ReactionNetwork("five_species_chain", np.asarray(bounds, np.int32), nu,
                           _tabulate(bounds, R, fn), rates,
                           tuple(f"{k}{a}" for k, a in kinds))


# Parameters of the reference's Hog1p problem (problems/hog1p_5species.cu:22-30).
HOG1P = dict(k12=1.29, k21=1.0e0, k23=0.0067, k32=0.027, k34=0.133, k43=0.0381,
             kr2=0.0116, kr3=0.987, kr4=0.0538, trans=0.01, gamma=0.0049,
             r1=6.9e-5, r2=7.1e-3, eta=3.1, Ahog=9.3e09, Mhog=6.4e-4)


def hog1p(bounds=(3, 50, 50, 60, 60)) -> ReactionNetwork:
    """The reference's 5-species Hog1p problem (problems/hog1p_5species.cu):
    gene state G in 0..3, two transcripts made in the nucleus (species 1, 2)
    and moved to the cytoplasm (species 3, 4), 9 reactions (stoichiometry at
    :239-241): G+1, G-1, G-1 (signal-repressed), x1+1, x2+1, x1->x3, x2->x4,
    x3-1, x4-1.  The propensities are those of the device function at
    :33-99, written as separable factor tables: the G-dependent rate
    constants in the species-0 factor, trans x_s and gamma x_s in the count
    factor.  (The reference's host factor function at :101-184 falls through
    its switch cases and gives reactions 5-7 factors of other species; the
    tables here follow the device function the reference's HYB run uses.)
    Every time coefficient is 1 except reaction 2's, hog1p_coefs(t)
    (:187-197).  Box {3, 50, 50, 60, 60}: 38,713,284 states."""
    p = HOG1P
    nu = np.zeros((9, 5), np.int32)
    nu[0, 0], nu[1, 0], nu[2, 0] = 1, -1, -1
    nu[3, 1], nu[4, 2] = 1, 1
    nu[5, 1], nu[5, 3] = -1, 1
    nu[6, 2], nu[6, 4] = -1, 1
    nu[7, 3], nu[8, 4] = -1, -1
    gtab = {0: (p["k12"], p["k23"], p["k34"], 0.0),
            1: (0.0, 0.0, p["k32"], p["k43"]),
            2: (0.0, p["k21"], 0.0, 0.0),
            3: (0.0, p["kr2"], p["kr3"], p["kr4"]),
            4: (0.0, p["kr2"], p["kr3"], p["kr4"])}
    count = {5: (1, p["trans"]), 6: (2, p["trans"]), 7: (3, p["gamma"]), 8: (4, p["gamma"])}

    def fn(r, s, i):
        x = i.astype(np.float64)
        if s == 0 and r in gtab:
            g = np.asarray(gtab[r], np.float64)
            return g[np.minimum(i, 3)] * (i <= 3)
        if r in count and s == count[r][0]:
            return count[r][1] * x
        return np.ones_like(x)

    return ReactionNetwork("hog1p", np.asarray(bounds, np.int32), nu,
                           _tabulate(bounds, 9, fn), np.ones(9, np.float64),
                           ("G+", "G-", "G-sig", "x1+", "x2+", "x1->x3", "x2->x4", "x3-", "x4-"))


def hog1p_coefs(t: float) -> np.ndarray:
    """Time coefficients of the Hog1p problem (problems/hog1p_5species.cu:
    187-197): 1 for every reaction except reaction 2, which is
    max(0, 3200 - 7710 hog1p(t)) with hog1p(t) = A (h / (1 + h / M))^eta,
    h = (1 - exp(-r1 t)) exp(-r2 t)."""
    p = HOG1P
    out = np.ones(9, np.float64)
    h1 = (1.0 - math.exp(-p["r1"] * t)) * math.exp(-p["r2"] * t)
    sig = math.pow(h1 / (1.0 + h1 / p["Mhog"]), p["eta"]) * p["Ahog"]
    out[2] = max(0.0, 3200.0 - 7710.0 * sig)
    return out


def birth_death(bounds=(63,), k=5.0, g=1.0) -> ReactionNetwork:
    """Independent birth-death processes per species (0->x_s at k, x_s->0 at
    g x_s): from x = 0 the exact CME solution is a product of Poissons with
    mean (k/g)(1 - exp(-g t))."""
    bounds = np.asarray(bounds, np.int32)
    d = len(bounds)
    R = 2 * d
    nu = np.zeros((R, d), np.int32)
    for s in range(d):
        nu[2 * s, s] = 1
        nu[2 * s + 1, s] = -1

    def fn(r, s, i):
        x = i.astype(np.float64)
        if r == 2 * s + 1:
            return x
        return np.ones_like(x)

    coefs = np.array([k, g] * d, np.float64)
    return ReactionNetwork("birth_death", bounds, nu, _tabulate(bounds, R, fn), coefs)


def delta_p0(net: ReactionNetwork, state=None, with_sink=False) -> np.ndarray:
    n = net.n_states
    p = np.zeros(n + (1 if with_sink else 0), np.float64)
    idx = 0
    if state is not None:
        mul = 1
        for s, x in enumerate(state):
            idx += int(x) * mul
            mul *= int(net.bounds[s]) + 1
    p[idx] = 1.0
    return p


TOGGLE_RATES = (50.0, 1.0, 16.0, 1.0)


def batch_instances(count=512, seed=1, n_min=2048, n_max=262144, t_range=(1.0, 100.0)):
    """Uneven batch: toggle switches and gene-expression networks drawn 50/50,
    every rate log-uniform within one decade of its nominal value, box sized
    so the state count is log-uniform in [n_min, n_max], final time uniform
    in t_range.  Returns (networks, t_final)."""
    rng = np.random.default_rng(seed)
    nets, tf = [], []
    for i in range(count):
        target = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        if i % 2 == 0:
            side = max(2, int(round(math.sqrt(target))))
            rates = np.asarray(TOGGLE_RATES) * 10.0 ** rng.uniform(-1, 1, 4)
            net = toggle_switch((side - 1, side - 1), rates=rates)
        else:
            side = max(2, int(round(math.sqrt(target / 3.0))))
            rates = np.asarray(GENE_RATES) * 10.0 ** rng.uniform(-1, 1, 9)
            net = gene_expression((2, side - 1, side - 1), rates=rates)
        nets.append(net)
        tf.append(rng.uniform(*t_range))
    return nets, np.asarray(tf)
