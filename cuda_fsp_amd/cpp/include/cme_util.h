// cme_util.h -- drop-in for the reference's src/cme_util.h (namespace cuFSP):
// shared types, state-space index maps and a device vector.
//
// Differences a caller sees:
//  * PropFun is a HOST function (double(int*, int)); the reference fetched a
//    __device__ pointer with cudaMemcpyFromSymbol.  Propensities written
//    `__device__ __host__` in the reference tests work unchanged.
//  * thrust_dvec is cuFSP::dvec, a device vector over the engine's C ABI.
//  * Errors throw std::runtime_error instead of exit(EXIT_FAILURE).
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "fsp_capi.h"

namespace cuFSP {

// Species per model the engine supports.
constexpr int kMaxSpeciesFacade = 8;

// src/cme_util.h:25 (device pointer there; host function here).  Called
// from several host threads at once while the box is tabulated, so it must
// be a pure function of its arguments, as the reference's device propensity
// (evaluated by every GPU thread) is.
typedef double (*PropFun)(int *x, int reaction);
// src/fspmat_kron_kernels.h:13
typedef double (*PropFactorFun)(int state, int species, int reaction);

// src/cme_util.h:28-46
struct CSRMat {
    double *vals = nullptr;
    int *col_idxs = nullptr;
    int *row_ptrs = nullptr;
    int n_rows = 0, n_cols = 0, nnz = 0;
};

struct CSRMatInt {
    int *vals = nullptr;
    int *col_idxs = nullptr;
    int *row_ptrs = nullptr;
    int n_rows = 0, n_cols = 0, nnz = 0;
};

struct SDKronMat {
    int d = 0;
    double *vals = nullptr;
    int *offsets = nullptr;
};

// Throws std::runtime_error(where + ": " + fsp_last_error()) on failure.
void throw_on(fsp_status st, const char *where);

// src/cme_util.cu:5-11, 13-22, 43-53, 55-62 (host versions)
void indx2state(int indx, int *state, int dim, int *fsp_bounds);
int state2indx(int *state, int dim, int *fsp_bounds);
void reachable_state(int *state, int *rstate, int reaction, int direction, int n_species,
                     int *stoich_val, int *stoich_colidxs, int *stoich_rowptrs);
int rect_fsp_num_states(int n_species, int *fsp_bounds);

// src/cme_util.cu:24-41: fills the device AoS array d_states[n_states][dim]
// on the GPU (a host call here; the reference launched the kernel itself).
// n_bounds is a HOST array.
void fsp_get_states(int *d_states, int dim, int n_states, int *n_bounds);

// Device vector of doubles (stands in for thrust::device_vector<double>).
class dvec {
public:
    explicit dvec(size_t n = 0, double value = 0.0);
    dvec(const dvec &o);
    dvec(dvec &&o) noexcept;
    dvec &operator=(const dvec &o);
    dvec &operator=(dvec &&o) noexcept;
    ~dvec();

    size_t size() const { return n_; }
    double *data() { return d_; }
    const double *data() const { return d_; }

    void resize(size_t n, double value = 0.0);
    void fill(double value);
    void set(size_t i, double value);
    double get(size_t i) const;
    std::vector<double> to_host() const;
    void from_host(const std::vector<double> &h);
    double sum() const;

    class ref {
    public:
        ref(dvec &v, size_t i) : v_(v), i_(i) {}
        operator double() const { return v_.get(i_); }
        ref &operator=(double x) {
            v_.set(i_, x);
            return *this;
        }

    private:
        dvec &v_;
        size_t i_;
    };
    ref operator[](size_t i) { return ref(*this, i); }
    double operator[](size_t i) const { return get(i); }

private:
    double *d_ = nullptr;
    size_t n_ = 0;
};

typedef dvec thrust_dvec;

// thrust::raw_pointer_cast(&v[0]) in the reference tests
inline double *raw_pointer(dvec &v) { return v.data(); }

}  // namespace cuFSP
