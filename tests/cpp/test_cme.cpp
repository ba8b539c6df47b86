// cme layer on the GPU engine: FSPSolver stepping through output times,
// Likelihood::evaluate on seeded SSA data (generate_data).  Writes its
// results to argv[1] for the Python side to check against the oracle:
//   int32 n_times, n_states, n_cells; double p[n_times][n_states];
//   int32 idx[n_times][n_cells]; double ll_factor, ll_table
#include <cassert>
#include <cmath>
#include <cstdio>
#include <vector>

#include "FSPSolver.h"

// rates chosen so that the (2, 40, 40) box holds essentially all the mass
static const double kRates[9] = {0.5, 0.5, 0.5, 0.5, 2.0, 4.0, 1.0, 0.1, 0.5};

// promoter G in {0,1,2}, nascent n, mature m (9 terms)
static double factor(int x, int s, int r)
{
    static const int gate[6] = {0, 1, 1, 2, 1, 2};
    if (s == 0 && r < 6) return x == gate[r] ? 1.0 : 0.0;
    if (s == 1 && (r == 6 || r == 7)) return (double)x;
    if (s == 2 && r == 8) return (double)x;
    return 1.0;
}

static std::vector<double> prop_all(const std::vector<int> &x)
{
    std::vector<double> a(9, 1.0);
    for (int r = 0; r < 9; ++r)
        for (int s = 0; s < 3; ++s) a[r] *= factor(x[s], s, r);
    return a;
}

static std::vector<double> t_fun(double, const std::vector<double> &p)
{
    std::vector<double> c(9);
    for (int r = 0; r < 9; ++r) c[r] = std::exp(p[r]);
    return c;
}

int main(int argc, char **argv)
{
    cme::Model model;
    model.FSPSize = {2, 40, 40};
    model.states_init = {{0, 0, 0}};
    model.probs_init = {1.0};
    model.stoich_mat = {{1, 0, 0}, {-1, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0},
                        {0, 1, 0}, {0, -1, 1}, {0, -1, 0}, {0, 0, -1}};
    model.t_fun = &t_fun;
    for (double k : kRates) model.parameters.push_back(std::log(k));
    model.propensity_factor = &factor;
    std::vector<double> times = {1.0, 5.0, 20.0};

    std::vector<double> P;
    cme::FSPSolver fsp(P, model, times, 1e-6);
    const int n = fsp.n_states();
    std::vector<std::vector<double>> outs;
    for (size_t k = 0; k < times.size(); ++k) {
        fsp.next_time();
        outs.push_back(P);
    }
    // a second pass after reset_time reproduces the first
    fsp.reset_time();
    fsp.next_time();
    double d = 0.0;
    for (int i = 0; i < n; ++i) d += std::abs(P[i] - outs[0][i]);
    assert(d == 0.0);

    cme::SingleCellData data = cme::generate_data(model, times, {0, 0, 0}, 400, 12345);
    cme::Likelihood L1(data, model);
    const double ll_factor = L1.evaluate(model.parameters);
    cme::Model m2 = model;
    m2.propensity_factor = nullptr;
    m2.propensity = &prop_all;
    cme::Likelihood L2(data, m2);
    const double ll_table = L2.evaluate(model.parameters);
    std::printf("ll factor %.12f table %.12f\n", ll_factor, ll_table);
    assert(std::abs(ll_factor - ll_table) <= 1e-9 * std::abs(ll_factor));

    FILE *f = std::fopen(argc > 1 ? argv[1] : "cme_out.bin", "wb");
    int hdr[3] = {(int)times.size(), n, 400};
    std::fwrite(hdr, sizeof(int), 3, f);
    for (auto &o : outs) std::fwrite(o.data(), sizeof(double), (size_t)n, f);
    for (size_t it = 0; it < times.size(); ++it)
        for (auto &x : data.snapshots[it]) {
            int i = x[0] + 3 * x[1] + 3 * 41 * x[2];
            if (x[1] > 40 || x[2] > 40) i = -1;
            std::fwrite(&i, sizeof(int), 1, f);
        }
    std::fwrite(&ll_factor, sizeof(double), 1, f);
    std::fwrite(&ll_table, sizeof(double), 1, f);
    const std::vector<double> c = model.t_fun(0.0, model.parameters);
    std::fwrite(c.data(), sizeof(double), c.size(), f);
    std::fclose(f);
    std::printf("cme test done\n");
    return 0;
}
