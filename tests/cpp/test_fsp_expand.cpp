// Adaptive FSP (cme::FSPSolver with fsp_tol): two independent birth-death
// species started at 0 on a box too small for the first one.  The exact CME
// solution is a product of Poissons with means (k_s / g_s)(1 - exp(-g_s t)),
// so at every output time
//   * the truncation error 1 - sum(P) stays within fsp_tol * t / t_max,
//   * P differs from the exact solution on the box by at most that error
//     (plus the Krylov tolerance),
//   * the box grew along species 0 only (species 1's edge holds ~1e-8).
// reset_time keeps the grown box (no further expansion, and repeated passes
// agree bit for bit); a Likelihood on the same model remaps its data when
// the box grows.
#include <cassert>
#include <cmath>
#include <cstdio>
#include <vector>

#include "FSPSolver.h"

static const double kBirth[2] = {20.0, 1.0}, kDeath[2] = {1.0, 1.0};

// reactions: 0 -> x0, x0 -> 0, 0 -> x1, x1 -> 0
static double factor(int x, int s, int r)
{
    if (r == 2 * s + 1) return (double)x;
    return 1.0;
}

static std::vector<double> t_fun(double, const std::vector<double> &p) { return p; }

static double poisson(int k, double lam)
{
    return std::exp(k * std::log(lam) - lam - std::lgamma(k + 1.0));
}

static double l1_vs_exact(const std::vector<double> &P, const std::vector<int> &b, double t)
{
    double lam[2];
    for (int s = 0; s < 2; ++s) lam[s] = kBirth[s] / kDeath[s] * (1.0 - std::exp(-kDeath[s] * t));
    double d = 0.0;
    for (int y = 0; y <= b[1]; ++y)
        for (int x = 0; x <= b[0]; ++x)
            d += std::abs(P[(size_t)x + (size_t)(b[0] + 1) * y] - poisson(x, lam[0]) * poisson(y, lam[1]));
    return d;
}

int main()
{
    cme::Model model;
    model.FSPSize = {10, 10};
    model.states_init = {{0, 0}};
    model.probs_init = {1.0};
    model.stoich_mat = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};
    model.t_fun = &t_fun;
    model.parameters = {kBirth[0], kDeath[0], kBirth[1], kDeath[1]};
    model.propensity_factor = &factor;
    const std::vector<double> times = {1.0, 2.0, 5.0};
    const double tol = 1e-6;

    std::vector<double> P;
    cme::FSPSolver fsp(P, model, times, tol, 1e-8, 1e-10, cuFSP::HYB);
    std::vector<std::vector<double>> outs;
    for (double t : times) {
        fsp.next_time();
        const std::vector<int> &b = fsp.bounds();
        double mass = 0.0;
        for (double v : P) mass += v;
        const double err = fsp.truncation_error();
        const double d = l1_vs_exact(P, b, t);
        std::printf("t=%g box (%d, %d) expansions %d: 1 - sum P = %.3e (budget %.3e), "
                    "L1 vs exact %.3e\n",
                    t, b[0], b[1], fsp.n_expansions(), err, tol * t / times.back(), d);
        assert((int)P.size() == (b[0] + 1) * (b[1] + 1));
        assert(err >= -1e-14 && err <= tol * t / times.back());
        assert(std::abs((1.0 - mass) - err) <= 1e-12);
        assert(d <= err + 1e-7);
        outs.push_back(P);
    }
    assert(fsp.n_expansions() >= 1);
    assert(fsp.bounds()[0] > 10);
    assert(fsp.bounds()[1] == 10);
    // the grown box stays: a second pass on it needs no expansion, and a
    // third repeats the second bit for bit
    const std::vector<int> grown = fsp.bounds();
    const int expansions = fsp.n_expansions();
    fsp.reset_time();
    assert(fsp.bounds() == grown);
    for (size_t k = 0; k < times.size(); ++k) fsp.next_time();
    assert(fsp.bounds() == grown);
    assert(fsp.n_expansions() == expansions);
    const std::vector<double> second = P;
    fsp.reset_time();
    for (size_t k = 0; k < times.size(); ++k) fsp.next_time();
    assert(P == second);
    const double d2 = l1_vs_exact(P, grown, times.back());
    std::printf("second pass on the grown box: L1 vs exact %.3e\n", d2);
    assert(d2 <= tol + 1e-7);

    // Likelihood: the data (seeded SSA) are remapped when the box grows
    cme::SingleCellData data = cme::generate_data(model, times, {0, 0}, 300, 7);
    cme::Likelihood L(data, model);
    const double ll1 = L.evaluate(model.parameters);
    const double ll2 = L.evaluate(model.parameters);
    std::printf("log-likelihood %.9f then %.9f\n", ll1, ll2);
    assert(std::isfinite(ll1) && std::isfinite(ll2));
    // the first evaluation grows the box part way through, the second runs
    // on the grown box from the start: they differ by the truncation error
    // only (indices left on the old box would change the sum entirely)
    assert(std::abs(ll1 - ll2) <= 1e-4 * std::abs(ll2));
    std::printf("fsp expand test done\n");
    return 0;
}
