// FSPSolver.h -- the reference's cme layer (src/Model.h, src/FSPSolver.h,
// src/Likelihood.h) with std containers in place of Armadillo.  The
// reference's versions never compiled (they call an FSPMat(model)
// constructor and expv methods that do not exist); this header gives them
// the semantics their code spells out, on the GPU engine.
#pragma once

#include <cstdint>
#include <vector>

#include "KryExpvFSP.h"

namespace cme {

// src/Model.h:14-26.  States are one std::vector<int> each (the reference
// stored them column-wise in a Mat<int>); stoich_mat[r] is the net change
// of reaction r (a column of the reference's species x reactions matrix).
struct Model {
    // all R propensities at x, without coefficients (Mat<double>(*)(Mat<int>))
    typedef std::vector<double> (*PropFun)(const std::vector<int> &x);
    // R coefficients at time t (Row<double>(*)(double, const Row<double>&))
    typedef std::vector<double> (*InputFun)(double t, const std::vector<double> &parameters);

    std::vector<std::vector<int>> states_init;
    std::vector<double> probs_init;
    std::vector<int> FSPSize;
    std::vector<std::vector<int>> stoich_mat;
    PropFun propensity = nullptr;
    InputFun t_fun = nullptr;
    std::vector<double> parameters;
    // optional separable form a_r(x) = prod_s f(x_s, s, r): used instead of
    // `propensity` when set (no host tabulation over the whole box)
    cuFSP::PropFactorFun propensity_factor = nullptr;
};

// src/FSPSolver.h:16-111, src/FSPSolver.cpp:9-22.
//
// fsp_tol (accepted but never read by the reference, src/FSPSolver.h:36-44)
// drives the adaptive FSP: with 0 < fsp_tol < 1 each output time t_k must
// satisfy the truncation bound 1 - sum(p(t_k)) <= fsp_tol * t_k / t_max
// (the FSP error of a generator whose diagonal carries the outflow leaving
// the box, Munsky & Khammash 2006).  When it does not, the interval is
// solved again from t_{k-1} on a larger box: every species with probability
// on its upper edge (at least a tenth of the largest edge mass) grows by
// half, the probabilities are carried over state by state, and the
// generator is rebuilt.  P then has the new box's size.  fsp_tol <= 0 or
// >= 1 keeps the box fixed.
class FSPSolver {
public:
    FSPSolver(std::vector<double> &p, const Model &model, const std::vector<double> &times,
              double fsp_tol, double mg_tol = 1.0e-8, double kry_tol = 1.0e-8,
              cuFSP::MatrixFormat format = cuFSP::CUDA_CSR);
    FSPSolver(std::vector<double> &p, const Model &model, double t_final, double fsp_tol,
              double mg_tol = 1.0e-8, double kry_tol = 1.0e-8,
              cuFSP::MatrixFormat format = cuFSP::CUDA_CSR);
    ~FSPSolver();
    FSPSolver(const FSPSolver &) = delete;
    FSPSolver &operator=(const FSPSolver &) = delete;

    void next_time();   // integrate to the next output time, P <- p(t_k)
    void solve();       // integrate to max(times), P <- p
    void reset_time();  // back to t = 0 with the initial distribution and
                        // the coefficients t_fun(0, model.parameters)

    const cuFSP::dvec &device_solution() const { return p_dev_; }
    size_t current_time_index() const { return current_; }
    int n_states() const { return n_; }
    // the current FSP box (the model's FSPSize until the box grows)
    const std::vector<int> &bounds() const { return bounds_; }
    // 1 - sum(p) at the last output time: the L1 distance to the CME solution
    double truncation_error() const { return err_; }
    int n_expansions() const { return n_expand_; }

private:
    void build(const std::vector<int> &bounds);
    void release();
    void advance_to(double t);       // one output interval, with expansion
    void grow_box(const std::vector<double> &p_fail);
    void embed(const std::vector<double> &p_old, const std::vector<int> &old_bounds,
               std::vector<double> &p_new) const;
    void load_initial();
    const Model &model_;
    std::vector<double> &P_;
    std::vector<double> times_;
    double fsp_tol_;
    double kry_tol_;
    cuFSP::MatrixFormat format_;
    size_t current_ = 0;
    int n_ = 0;
    std::vector<int> bounds_;
    double t_now_ = 0.0;             // time of the solution in p_dev_
    double t_origin_ = 0.0;          // time the Krylov object's clock starts at
    double mass0_ = 1.0;             // sum of the initial distribution
    double err_ = 0.0;
    int n_expand_ = 0;
    std::vector<double> p_start_;    // host copy of p at t_now_ (adaptive FSP)
    std::vector<double> p_init_;     // initial distribution on the model's box
    cuFSP::dvec p_dev_;
    double *table_dev_ = nullptr;
    fsp_mat A_ = nullptr;
    fsp_expv expv_ = nullptr;
};

// src/Likelihood.h:15-18
struct SingleCellData {
    std::vector<double> times;
    std::vector<std::vector<std::vector<int>>> snapshots;  // [time][cell] -> state
};

// src/Likelihood.h:20-54, src/Likelihood.cpp:8-33
class Likelihood {
public:
    Likelihood(SingleCellData &data, Model &model, double floor_val = 1.0e-44);
    ~Likelihood();
    double evaluate(const std::vector<double> &parameters);

private:
    void map_data();                 // data -> state indices of the current box
    SingleCellData &data_;
    Model &model_;
    double floor_;
    std::vector<double> full_dist_;
    FSPSolver fsp_;
    std::vector<int> mapped_bounds_;
    std::vector<int *> idx_dev_;
    std::vector<int64_t> idx_count_;
};

// src/Likelihood.cpp:35-84: Gillespie direct-method samples at each time
// (independent runs from x0, propensities frozen at t_fun(0, parameters)),
// seeded (the reference seeds from the clock).
SingleCellData generate_data(Model &model, std::vector<double> times, std::vector<int> x0,
                             size_t num_samples, uint64_t seed = 0);

}  // namespace cme
