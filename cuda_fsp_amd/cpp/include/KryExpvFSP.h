// KryExpvFSP.h -- drop-in for the reference's src/KryExpvFSP.h (namespace
// cuFSP): v <- exp(t_final A) v by the adaptive Krylov scheme of Sidje's
// Expokit (Arnoldi, or incomplete orthogonalisation), the whole step on the
// GPU.  Public knobs IOP, q_iop, max_nstep, max_reject as in
// src/KryExpvFSP.h:64-69 (read at every step).
#pragma once

#include <functional>

#include "FSPMat.h"

namespace cuFSP {

// src/KryExpvFSP.h:14
using MVFun = std::function<void(double *x, double *y)>;

class KryExpvFSP {
public:
    // src/KryExpvFSP.h:75-76: matvec(x, y) must do y += A x on the device
    KryExpvFSP(double t_final, MVFun &matvec, thrust_dvec &v, int m, double tol = 1.0e-8,
               bool iop = false, int q_iop = 2, double anorm = 1.0);
    // Fused path: the operator is A at the coefficients t_func(t_coef); the
    // Arnoldi loop runs as fused device kernels captured in a HIP graph.
    KryExpvFSP(double t_final, FSPMat &A, double t_coef, thrust_dvec &v, int m,
               double tol = 1.0e-8, bool iop = false, int q_iop = 2, double anorm = 1.0);
    ~KryExpvFSP();
    KryExpvFSP(const KryExpvFSP &) = delete;
    KryExpvFSP &operator=(const KryExpvFSP &) = delete;

    void reset_time();            // src/KryExpvFSP.h:82-84
    void solve();                 // src/KryExpvFSP.cu:184-188
    void step();                  // src/KryExpvFSP.cu:38-182
    bool final_time_reached();    // src/KryExpvFSP.h:100-102

    // no reference counterpart
    void set_final_time(double t_final);
    fsp_expv_stats stats();

    bool IOP = false;
    int q_iop = 2;
    int m = 30;
    int max_nstep = 10000;
    int max_reject = 1000;

private:
    void sync_knobs();
    static void trampoline(void *ctx, const double *x, double *y);
    fsp_expv h_ = nullptr;
    MVFun matvec_;
    thrust_dvec &sol_vec_;
    int last_iop_ = -1, last_q_ = -1, last_reject_ = -1, last_nstep_ = -1;
};

}  // namespace cuFSP
