// FSPMat.h -- drop-in for the reference's src/FSPMat.h (namespace cuFSP).
//
// The generator A(t) = sum_r c_r(t) A_r of the CME on a rectangular FSP box,
// held on the GPU by the engine (fsp_mat).  Same constructors, action()
// semantics (y += A(t) x, t_func called on every action) and getters as
// src/FSPMat.h:38-55; in addition every format accepts both propensity
// forms, a precision and an optional sink state.
#pragma once

#include <functional>

#include "cme_util.h"

namespace cuFSP {

// src/FSPMat.h:16
typedef std::function<void(double, double *)> TcoefFun;
// src/FSPMat.h:17-19
enum MatrixFormat { CUDA_CSR, HYB, KRONECKER };
// FP64 throughout, or FP32 values/basis with FP64 sums (no reference counterpart)
enum Precision { FP64, MIXED };

class FSPMat {
public:
    // src/FSPMat.cu:10-45 (PropFun: evaluated on the host over the box once,
    // the table lives on the GPU)
    explicit FSPMat(int n_reactions, int n_species, int *fsp_dim, CSRMatInt stoich,
                    TcoefFun t_func, PropFun prop_func, MatrixFormat format = CUDA_CSR,
                    Precision precision = FP64, bool with_sink = false);
    // src/FSPMat.cu:49-76 (separable propensity factors, tabulated on the host)
    explicit FSPMat(int n_reactions, int n_species, int *fsp_dim, CSRMatInt stoich,
                    TcoefFun t_func, PropFactorFun pffunc, MatrixFormat format,
                    Precision precision = FP64, bool with_sink = false);
    ~FSPMat();
    FSPMat(const FSPMat &) = delete;
    FSPMat &operator=(const FSPMat &) = delete;

    // src/FSPMat.cu:102-105: y += A(t) x on device vectors of get_n_rows()
    void action(double t, double *x, double *y);

    int get_n_rows();       // box states (+1 with the sink)
    int get_n_species();
    int get_n_reactions();

    // engine handle (for the fused Krylov path) and the coefficients at t
    fsp_mat handle() const { return h_; }
    void coefficients(double t, std::vector<double> &c);
    void set_coefficients(const std::vector<double> &c);

private:
    void init(int n_reactions, int n_species, int *fsp_dim, CSRMatInt stoich,
              const double *factors, const double *table_dev, MatrixFormat format,
              Precision precision, bool with_sink);
    fsp_mat h_ = nullptr;
    int nst_ = 0, ns_ = 0, nr_ = 0, rows_ = 0;
    TcoefFun tcoeffunc_;
    std::vector<double> tcoef_;
    double *table_dev_ = nullptr;
};

}  // namespace cuFSP
