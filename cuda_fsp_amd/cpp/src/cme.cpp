// cme.cpp -- Model / FSPSolver / Likelihood / generate_data on the engine.
#include <algorithm>
#include <cmath>
#include <random>

#include "FSPSolver.h"

namespace cme {

using cuFSP::throw_on;

static int64_t box_states(const std::vector<int> &b)
{
    int64_t n = 0;
    throw_on(fsp_num_states((int)b.size(), b.data(), &n), "cme");
    return n;
}

static int64_t sub2ind(const std::vector<int> &b, const std::vector<int> &x)
{
    int64_t idx = 0, mul = 1;
    for (size_t s = 0; s < b.size(); ++s) {
        if (s >= x.size() || x[s] < 0 || x[s] > b[s]) return -1;
        idx += (int64_t)x[s] * mul;
        mul *= b[s] + 1;
    }
    return idx;
}

FSPSolver::FSPSolver(std::vector<double> &p, const Model &model, const std::vector<double> &times,
                     double fsp_tol, double, double kry_tol, cuFSP::MatrixFormat format)
    : model_(model), P_(p), times_(times), fsp_tol_(fsp_tol), kry_tol_(kry_tol), format_(format)
{
    if (times_.empty()) throw std::runtime_error("FSPSolver: no output times");
    for (size_t k = 0; k < times_.size(); ++k)
        if (!(times_[k] > (k ? times_[k - 1] : 0.0)))
            throw std::runtime_error("FSPSolver: output times must be positive and increasing");
    load_initial();
}

FSPSolver::FSPSolver(std::vector<double> &p, const Model &model, double t_final, double fsp_tol,
                     double, double kry_tol, cuFSP::MatrixFormat format)
    : model_(model), P_(p), times_({t_final}), fsp_tol_(fsp_tol), kry_tol_(kry_tol),
      format_(format)
{
    if (!(t_final > 0.0)) throw std::runtime_error("FSPSolver: t_final must be > 0");
    load_initial();
}

// Generator and Krylov object on the box `bounds` (replacing any earlier).
void FSPSolver::build(const std::vector<int> &bounds)
{
    const Model &m = model_;
    const int d = (int)m.FSPSize.size();
    const int R = (int)m.stoich_mat.size();
    if (d < 1 || R < 1 || !m.t_fun) throw std::runtime_error("FSPSolver: incomplete model");
    if ((int)bounds.size() != d) throw std::runtime_error("FSPSolver: bounds / species mismatch");
    release();
    bounds_ = bounds;
    n_ = (int)box_states(bounds_);
    // stoichiometry as a reactions x species CSR
    std::vector<int> sv, sc, sr(1, 0);
    for (int r = 0; r < R; ++r) {
        for (int sp = 0; sp < d; ++sp)
            if (sp < (int)m.stoich_mat[r].size() && m.stoich_mat[r][sp] != 0) {
                sv.push_back(m.stoich_mat[r][sp]);
                sc.push_back(sp);
            }
        sr.push_back((int)sv.size());
    }
    if (sv.empty()) {  // all reactions with zero net change
        sv.push_back(0);
        sc.push_back(0);
    }
    std::vector<double> factors;
    if (m.propensity_factor) {
        for (int r = 0; r < R; ++r)
            for (int sp = 0; sp < d; ++sp)
                for (int i = 0; i <= bounds_[sp]; ++i) factors.push_back(m.propensity_factor(i, sp, r));
    } else {
        if (!m.propensity) throw std::runtime_error("FSPSolver: model has no propensity");
        std::vector<double> table((size_t)R * n_);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n_; ++i) {
            std::vector<int> x((size_t)d);
            int q = i;
            for (int sp = 0; sp < d; ++sp) {
                x[sp] = q % (bounds_[sp] + 1);
                q /= (bounds_[sp] + 1);
            }
            std::vector<double> a = m.propensity(x);
            for (int r = 0; r < R; ++r) table[(size_t)r * n_ + i] = a.at((size_t)r);
        }
        void *p = nullptr;
        throw_on(fsp_malloc(&p, table.size() * sizeof(double)), "FSPSolver");
        table_dev_ = (double *)p;
        throw_on(fsp_memcpy_h2d(table_dev_, table.data(), table.size() * sizeof(double)), "FSPSolver");
    }
    fsp_model fm;
    fm.n_species = d;
    fm.n_reactions = R;
    fm.bounds = bounds_.data();
    fm.stoich_vals = sv.data();
    fm.stoich_cols = sc.data();
    fm.stoich_rowptrs = sr.data();
    fm.factors = factors.empty() ? nullptr : factors.data();
    fm.prop_table_dev = table_dev_;
    std::vector<double> c = m.t_fun(0.0, m.parameters);
    if ((int)c.size() != R) throw std::runtime_error("FSPSolver: t_fun returned the wrong count");
    fsp_format f = format_ == cuFSP::CUDA_CSR ? FSP_CSR : (format_ == cuFSP::HYB ? FSP_HYB : FSP_KRON);
    throw_on(fsp_mat_create(&fm, f, FSP_F64, 0, c.data(), nullptr, &A_), "FSPSolver");
    // src/FSPSolver.h:71: m = 30, IOP on
    throw_on(fsp_expv_create(A_, nullptr, nullptr, n_, times_[0], 30, kry_tol_, 1, 2, 1.0, nullptr,
                             &expv_),
             "FSPSolver");
}

void FSPSolver::release()
{
    if (expv_) fsp_expv_destroy(expv_);
    if (A_) fsp_mat_destroy(A_);
    if (table_dev_) fsp_free(table_dev_);
    expv_ = nullptr;
    A_ = nullptr;
    table_dev_ = nullptr;
}

FSPSolver::~FSPSolver() { release(); }

// Initial distribution on the model's box (src/FSPSolver.h:74-89), t = 0.
void FSPSolver::load_initial()
{
    const Model &m = model_;
    if (!expv_ || bounds_ != m.FSPSize) {
        n_expand_ = 0;
        build(m.FSPSize);
    }
    if ((int64_t)P_.size() != n_) {
        P_.assign((size_t)n_, 0.0);
        for (size_t k = 0; k < m.states_init.size(); ++k) {
            const int64_t i = sub2ind(bounds_, m.states_init[k]);
            if (i < 0) throw std::runtime_error("FSPSolver: initial state outside the FSP box");
            P_[(size_t)i] = m.probs_init.at(k);
        }
    }
    p_init_ = P_;
    mass0_ = 0.0;
    for (double v : P_) mass0_ += v;
    p_dev_.from_host(P_);
    t_now_ = 0.0;
    t_origin_ = 0.0;
    err_ = 0.0;
    throw_on(fsp_expv_restart(expv_, times_[0]), "FSPSolver");
}

// p_old on old_bounds -> p_new on bounds_ (state by state; the new box
// contains the old one).
void FSPSolver::embed(const std::vector<double> &p_old, const std::vector<int> &old_bounds,
                      std::vector<double> &p_new) const
{
    const int d = (int)bounds_.size();
    p_new.assign((size_t)n_, 0.0);
    std::vector<int64_t> stride((size_t)d);
    int64_t mul = 1;
    for (int sp = 0; sp < d; ++sp) {
        stride[sp] = mul;
        mul *= bounds_[sp] + 1;
    }
    for (size_t i = 0; i < p_old.size(); ++i) {
        if (p_old[i] == 0.0) continue;
        int64_t q = (int64_t)i, j = 0;
        for (int sp = 0; sp < d; ++sp) {
            j += (q % (old_bounds[sp] + 1)) * stride[sp];
            q /= old_bounds[sp] + 1;
        }
        p_new[(size_t)j] = p_old[i];
    }
}

// New bounds from where the failed solution sits: the species whose upper
// edge holds at least a tenth of the largest edge mass grow by half (at
// least by one).
void FSPSolver::grow_box(const std::vector<double> &p_fail)
{
    const int d = (int)bounds_.size();
    std::vector<double> edge((size_t)d, 0.0);
    for (size_t i = 0; i < p_fail.size(); ++i) {
        int64_t q = (int64_t)i;
        for (int sp = 0; sp < d; ++sp) {
            if (q % (bounds_[sp] + 1) == bounds_[sp]) edge[sp] += std::abs(p_fail[i]);
            q /= bounds_[sp] + 1;
        }
    }
    const double top = *std::max_element(edge.begin(), edge.end());
    std::vector<int> nb = bounds_;
    for (int sp = 0; sp < d; ++sp) {
        if (top > 0.0 && edge[sp] < 0.1 * top) continue;
        const int64_t grown = std::max<int64_t>((int64_t)bounds_[sp] + 1,
                                                (int64_t)std::ceil(1.5 * (bounds_[sp] + 1)) - 1);
        nb[sp] = (int)std::min<int64_t>(grown, 2147483646LL);
    }
    int64_t states = 0;
    if (fsp_num_states(d, nb.data(), &states) != FSP_OK)
        throw std::runtime_error("FSPSolver: the FSP box cannot grow past 2^31 - 1 states");
    // the solution at t_now_ on the old box, carried to the new one
    std::vector<double> p_old = p_start_;
    const std::vector<int> old_bounds = bounds_;
    build(nb);
    embed(p_old, old_bounds, p_start_);
    n_expand_ += 1;
}

// Integrates from t_now_ to t with the truncation check; on failure the
// interval is solved again from t_now_ on a larger box.
void FSPSolver::advance_to(double t)
{
    const bool adaptive = fsp_tol_ > 0.0 && fsp_tol_ < 1.0;
    const double t_max = *std::max_element(times_.begin(), times_.end());
    if (adaptive) p_start_ = p_dev_.to_host();
    for (int attempt = 0;; ++attempt) {
        throw_on(fsp_expv_set_final_time(expv_, t - t_origin_), "FSPSolver");
        throw_on(fsp_expv_solve(expv_, p_dev_.data()), "FSPSolver");
        double mass = 0.0;
        throw_on(fsp_vec_sum(p_dev_.data(), n_, &mass, nullptr), "FSPSolver");
        err_ = mass0_ - mass;
        if (!adaptive || err_ <= fsp_tol_ * t / t_max) break;
        if (attempt >= 40) throw std::runtime_error("FSPSolver: box expansion did not meet fsp_tol");
        grow_box(p_dev_.to_host());
        p_dev_.from_host(p_start_);
        // the new Krylov object's clock starts at t_now_
        t_origin_ = t_now_;
        throw_on(fsp_expv_restart(expv_, t - t_origin_), "FSPSolver");
    }
    t_now_ = t;
}

void FSPSolver::next_time()
{
    // src/FSPSolver.cpp:9-15
    if (current_ >= times_.size()) throw std::runtime_error("FSPSolver::next_time: past the last time");
    advance_to(times_[current_]);
    current_ += 1;
    P_ = p_dev_.to_host();
}

void FSPSolver::solve()
{
    // src/FSPSolver.cpp:17-22
    advance_to(*std::max_element(times_.begin(), times_.end()));
    current_ = times_.size() - 1;
    P_ = p_dev_.to_host();
}

void FSPSolver::reset_time()
{
    // src/FSPSolver.h:103-109 (and the reset of Likelihood.cpp:23-30): the
    // box stays as grown; the initial distribution is put into it
    current_ = 0;
    std::vector<double> c = model_.t_fun(0.0, model_.parameters);
    throw_on(fsp_mat_set_coefs(A_, c.data(), nullptr), "FSPSolver::reset_time");
    std::vector<double> p0;
    embed(p_init_, model_.FSPSize, p0);  // the initial vector, in the current box
    P_ = p0;
    mass0_ = 0.0;
    for (double v : p0) mass0_ += v;
    p_dev_.from_host(p0);
    t_now_ = 0.0;
    t_origin_ = 0.0;
    err_ = 0.0;
    throw_on(fsp_expv_restart(expv_, times_[0]), "FSPSolver::reset_time");
}

// ------------------------------------------------------------ Likelihood
Likelihood::Likelihood(SingleCellData &data, Model &model, double floor_val)
    : data_(data), model_(model), floor_(floor_val), fsp_(full_dist_, model, data.times, 1.0e-6)
{
    map_data();
}

// data2fsp (src/Likelihood.h:38-41): each cell's state -> its index in the
// current FSP box (-1 outside), on the device.
void Likelihood::map_data()
{
    for (int *p : idx_dev_) fsp_free(p);
    idx_dev_.clear();
    idx_count_.clear();
    mapped_bounds_ = fsp_.bounds();
    for (size_t it = 0; it < data_.times.size(); ++it) {
        const auto &snap = data_.snapshots.at(it);
        std::vector<int> idx(snap.size());
        for (size_t c = 0; c < snap.size(); ++c) idx[c] = (int)sub2ind(mapped_bounds_, snap[c]);
        void *p = nullptr;
        throw_on(fsp_malloc(&p, std::max<size_t>(idx.size(), 1) * sizeof(int)), "Likelihood");
        if (!idx.empty()) throw_on(fsp_memcpy_h2d(p, idx.data(), idx.size() * sizeof(int)), "Likelihood");
        idx_dev_.push_back((int *)p);
        idx_count_.push_back((int64_t)idx.size());
    }
}

Likelihood::~Likelihood()
{
    for (int *p : idx_dev_) fsp_free(p);
}

double Likelihood::evaluate(const std::vector<double> &parameters)
{
    // src/Likelihood.cpp:8-33
    model_.parameters = parameters;
    fsp_.reset_time();
    double ll = 0.0;
    for (size_t it = 0; it < data_.times.size(); ++it) {
        fsp_.next_time();
        if (fsp_.bounds() != mapped_bounds_) map_data();
        double s = 0.0;
        throw_on(fsp_loglik_indices(fsp_.device_solution().data(), idx_dev_[it], idx_count_[it],
                                    floor_, &s, nullptr),
                 "Likelihood::evaluate");
        ll += s;
    }
    return ll;
}

SingleCellData generate_data(Model &model, std::vector<double> times, std::vector<int> x0,
                             size_t num_samples, uint64_t seed)
{
    // src/Likelihood.cpp:35-84
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const std::vector<double> c = model.t_fun(0.0, model.parameters);
    const size_t R = model.stoich_mat.size();
    SingleCellData data;
    data.times = times;
    data.snapshots.resize(times.size());
    for (size_t it = 0; it < times.size(); ++it) {
        data.snapshots[it].resize(num_samples);
        for (size_t k = 0; k < num_samples; ++k) {
            std::vector<int> x = x0;
            double t = 0.0;
            while (t < times[it]) {
                double r1 = U(rng), r2 = U(rng);
                while (r1 <= 0.0) r1 = U(rng);
                std::vector<double> alpha(R);
                std::vector<double> a;
                if (model.propensity) {
                    a = model.propensity(x);
                } else {
                    a.assign(R, 1.0);
                    for (size_t r = 0; r < R; ++r)
                        for (size_t s = 0; s < x.size(); ++s) a[r] *= model.propensity_factor(x[s], (int)s, (int)r);
                }
                double a0 = 0.0;
                for (size_t r = 0; r < R; ++r) {
                    alpha[r] = c[r] * a[r];
                    a0 += alpha[r];
                }
                const double tau = -std::log(r1) / a0;
                size_t reaction = 0;
                double a1 = alpha[0];
                while (a1 < r2 * a0 && reaction < R - 1) {
                    reaction += 1;
                    a1 += alpha[reaction];
                }
                if (t + tau <= times[it]) {
                    for (size_t s = 0; s < x.size(); ++s) x[s] += model.stoich_mat[reaction][s];
                    t += tau;
                } else {
                    t = times[it];
                }
            }
            data.snapshots[it][k] = x;
        }
    }
    return data;
}

}  // namespace cme
