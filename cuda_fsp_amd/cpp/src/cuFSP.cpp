// cuFSP.cpp -- C++ facade (namespace cuFSP) over the engine's C ABI.
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "FSPMat.h"
#include "KryExpvFSP.h"

// Defined (weak) in every hipcc translation unit that includes FSPMat.h:
// evaluates a __device__ PropFun on the GPU.  Null when the program has none.
extern "C" int cufsp_tabulate_device_propensity(cuFSP::PropFun f, int n_species, int n_reactions,
                                                const int *bounds, int n_states,
                                                double *table_dev) __attribute__((weak));

namespace cuFSP {

void throw_on(fsp_status st, const char *where)
{
    if (st != FSP_OK) throw std::runtime_error(std::string(where) + ": " + fsp_last_error());
}

// ------------------------------------------------------------ index maps
void indx2state(int indx, int *state, int dim, int *fsp_bounds)
{
    for (int s = 0; s < dim; ++s) {
        state[s] = indx % (fsp_bounds[s] + 1);
        indx /= (fsp_bounds[s] + 1);
    }
}

int state2indx(int *state, int dim, int *fsp_bounds)
{
    int indx = 0, nprod = 1;
    for (int s = 0; s < dim; ++s) {
        indx += state[s] * nprod;
        nprod *= (fsp_bounds[s] + 1);
    }
    return indx;
}

void reachable_state(int *state, int *rstate, int reaction, int direction, int n_species,
                     int *stoich_val, int *stoich_colidxs, int *stoich_rowptrs)
{
    for (int k = 0; k < n_species; ++k) rstate[k] = state[k];
    for (int i = stoich_rowptrs[reaction]; i < stoich_rowptrs[reaction + 1]; ++i)
        rstate[stoich_colidxs[i]] += direction * stoich_val[i];
}

int rect_fsp_num_states(int n_species, int *fsp_bounds)
{
    int64_t n = 0;
    throw_on(fsp_num_states(n_species, fsp_bounds, &n), "rect_fsp_num_states");
    return (int)n;
}

void fsp_get_states(int *d_states, int dim, int n_states, int *n_bounds)
{
    if (rect_fsp_num_states(dim, n_bounds) != n_states)
        throw std::runtime_error("fsp_get_states: n_states does not match the bounds");
    throw_on(fsp_states_enumerate(dim, n_bounds, d_states, nullptr), "fsp_get_states");
}

// ------------------------------------------------------------------ dvec
dvec::dvec(size_t n, double value) { resize(n, value); }

dvec::dvec(const dvec &o)
{
    resize(o.n_);
    if (n_) throw_on(fsp_memcpy_d2d(d_, o.d_, n_ * sizeof(double), nullptr), "dvec copy");
}

dvec::dvec(dvec &&o) noexcept : d_(o.d_), n_(o.n_)
{
    o.d_ = nullptr;
    o.n_ = 0;
}

dvec &dvec::operator=(const dvec &o)
{
    if (this != &o) {
        resize(o.n_);
        if (n_) throw_on(fsp_memcpy_d2d(d_, o.d_, n_ * sizeof(double), nullptr), "dvec copy");
    }
    return *this;
}

dvec &dvec::operator=(dvec &&o) noexcept
{
    if (this != &o) {
        if (d_) fsp_free(d_);
        d_ = o.d_;
        n_ = o.n_;
        o.d_ = nullptr;
        o.n_ = 0;
    }
    return *this;
}

dvec::~dvec()
{
    if (d_) fsp_free(d_);
}

void dvec::resize(size_t n, double value)
{
    if (d_) fsp_free(d_);
    d_ = nullptr;
    n_ = n;
    if (n) {
        void *p = nullptr;
        throw_on(fsp_malloc(&p, n * sizeof(double)), "dvec");
        d_ = (double *)p;
        fill(value);
    }
}

void dvec::fill(double value)
{
    throw_on(fsp_vec_fill(d_, value, (int64_t)n_, nullptr), "dvec::fill");
    throw_on(fsp_stream_sync(nullptr), "dvec::fill");
}

void dvec::set(size_t i, double value)
{
    if (i >= n_) throw std::out_of_range("dvec::set");
    throw_on(fsp_stream_sync(nullptr), "dvec::set");
    throw_on(fsp_memcpy_h2d(d_ + i, &value, sizeof(double)), "dvec::set");
}

double dvec::get(size_t i) const
{
    if (i >= n_) throw std::out_of_range("dvec::get");
    double v = 0.0;
    throw_on(fsp_stream_sync(nullptr), "dvec::get");
    throw_on(fsp_memcpy_d2h(&v, d_ + i, sizeof(double)), "dvec::get");
    return v;
}

std::vector<double> dvec::to_host() const
{
    std::vector<double> h(n_);
    throw_on(fsp_stream_sync(nullptr), "dvec::to_host");
    if (n_) throw_on(fsp_memcpy_d2h(h.data(), d_, n_ * sizeof(double)), "dvec::to_host");
    return h;
}

void dvec::from_host(const std::vector<double> &h)
{
    if (h.size() != n_) resize(h.size());
    if (n_) throw_on(fsp_memcpy_h2d(d_, h.data(), n_ * sizeof(double)), "dvec::from_host");
}

double dvec::sum() const
{
    double s = 0.0;
    throw_on(fsp_vec_sum(d_, (int64_t)n_, &s, nullptr), "dvec::sum");
    return s;
}

// ---------------------------------------------------------------- FSPMat
static std::vector<int> stoich_rowptrs(const CSRMatInt &st, int n_reactions)
{
    if (!st.row_ptrs || !st.col_idxs || !st.vals)
        throw std::runtime_error("FSPMat: stoichiometry arrays are required");
    return std::vector<int>(st.row_ptrs, st.row_ptrs + n_reactions + 1);
}

void FSPMat::init(int n_reactions, int n_species, int *fsp_dim, CSRMatInt stoich,
                  const double *factors, const double *table_dev, MatrixFormat format,
                  Precision precision, bool with_sink)
{
    std::vector<int> rp = stoich_rowptrs(stoich, n_reactions);
    fsp_model m;
    m.n_species = n_species;
    m.n_reactions = n_reactions;
    m.bounds = fsp_dim;
    m.stoich_vals = stoich.vals;
    m.stoich_cols = stoich.col_idxs;
    m.stoich_rowptrs = rp.data();
    m.factors = factors;
    m.prop_table_dev = table_dev;
    tcoef_.assign((size_t)n_reactions, 0.0);
    tcoeffunc_(0.0, tcoef_.data());
    fsp_format f = format == CUDA_CSR ? FSP_CSR : (format == HYB ? FSP_HYB : FSP_KRON);
    throw_on(fsp_mat_create(&m, f, precision == MIXED ? FSP_MIXED : FSP_F64, with_sink ? 1 : 0,
                            tcoef_.data(), nullptr, &h_),
             "FSPMat::FSPMat");
    fsp_mat_info info;
    throw_on(fsp_mat_get_info(h_, &info), "FSPMat::FSPMat");
    nst_ = (int)info.n_states;
    rows_ = (int)info.n_rows;
}

// a_r(x) over the box into the device table [R][n].
static void tabulate_propensity(PropFun prop_func, int n_species, int n_reactions, int *fsp_dim,
                                int n, double *table_dev)
{
    // A pointer into no loaded host object is the address of a __device__
    // function (hipMemcpyFromSymbol, as tests/test_fsp_mat.cu:123-124 of the
    // reference): evaluated on the GPU, where the reference calls it inside
    // its assembly kernel (src/fspmat_csr_kernels.cu:79).
    Dl_info where;
    if (dladdr(reinterpret_cast<void *>(prop_func), &where) == 0) {
        if (!cufsp_tabulate_device_propensity)
            throw std::runtime_error(
                "FSPMat: prop_func is no host function and the program holds no device "
                "propensity tabulator (compile the caller with hipcc, including FSPMat.h)");
        const int rc = cufsp_tabulate_device_propensity(prop_func, n_species, n_reactions,
                                                        fsp_dim, n, table_dev);
        if (rc != 0)
            throw std::runtime_error("FSPMat: device propensity tabulation failed (code " +
                                     std::to_string(rc) + ")");
        return;
    }
    // A host function: a chunk of states at a time, so the host holds
    // R * chunk values, not R * n.  prop_func is called from several OpenMP
    // threads at once: like the reference's device propensity (called by
    // every GPU thread), it must be a pure function of its arguments.
    constexpr int kChunk = 1 << 20;
    std::vector<double> buf((size_t)n_reactions * (size_t)std::min(n, kChunk));
    for (int c0 = 0; c0 < n; c0 += kChunk) {
        const int cn = std::min(kChunk, n - c0);
#pragma omp parallel for schedule(static)
        for (int k = 0; k < cn; ++k) {
            int x[kMaxSpeciesFacade];
            indx2state(c0 + k, x, n_species, fsp_dim);
            for (int r = 0; r < n_reactions; ++r) {
                int xc[kMaxSpeciesFacade];
                std::memcpy(xc, x, sizeof(int) * (size_t)n_species);
                buf[(size_t)r * cn + k] = prop_func(xc, r);
            }
        }
        for (int r = 0; r < n_reactions; ++r)
            throw_on(fsp_memcpy_h2d(table_dev + (size_t)r * n + c0, buf.data() + (size_t)r * cn,
                                    (size_t)cn * sizeof(double)),
                     "FSPMat::FSPMat");
    }
}

FSPMat::FSPMat(int n_reactions, int n_species, int *fsp_dim, CSRMatInt stoich, TcoefFun t_func,
               PropFun prop_func, MatrixFormat format, Precision precision, bool with_sink)
    : ns_(n_species), nr_(n_reactions), tcoeffunc_(t_func)
{
    if (!prop_func) throw std::runtime_error("FSPMat: propensity function is NULL");
    if (n_species < 1 || n_species > kMaxSpeciesFacade)
        throw std::runtime_error("FSPMat: n_species must be in 1..8");
    if (n_reactions < 1) throw std::runtime_error("FSPMat: n_reactions must be >= 1");
    const int n = rect_fsp_num_states(n_species, fsp_dim);
    void *p = nullptr;
    throw_on(fsp_malloc(&p, (size_t)n_reactions * (size_t)n * sizeof(double)), "FSPMat::FSPMat");
    table_dev_ = (double *)p;
    try {
        tabulate_propensity(prop_func, n_species, n_reactions, fsp_dim, n, table_dev_);
        init(n_reactions, n_species, fsp_dim, stoich, nullptr, table_dev_, format, precision,
             with_sink);
    } catch (...) {
        fsp_free(table_dev_);
        table_dev_ = nullptr;
        throw;
    }
}

FSPMat::FSPMat(int n_reactions, int n_species, int *fsp_dim, CSRMatInt stoich, TcoefFun t_func,
               PropFactorFun pffunc, MatrixFormat format, Precision precision, bool with_sink)
    : ns_(n_species), nr_(n_reactions), tcoeffunc_(t_func)
{
    if (!pffunc) throw std::runtime_error("FSPMat: propensity factor function is NULL");
    // tables vals[r][s][0..N_s] as fspmat_sdkron_fill_host_data
    // (src/fspmat_kron_kernels.cu:170-190)
    size_t stride = 0;
    for (int s = 0; s < n_species; ++s) stride += (size_t)fsp_dim[s] + 1;
    std::vector<double> f((size_t)n_reactions * stride);
    size_t k = 0;
    for (int r = 0; r < n_reactions; ++r)
        for (int s = 0; s < n_species; ++s)
            for (int i = 0; i <= fsp_dim[s]; ++i) f[k++] = pffunc(i, s, r);
    init(n_reactions, n_species, fsp_dim, stoich, f.data(), nullptr, format, precision,
         with_sink);
}

FSPMat::~FSPMat()
{
    if (h_) fsp_mat_destroy(h_);
    if (table_dev_) fsp_free(table_dev_);
}

void FSPMat::coefficients(double t, std::vector<double> &c)
{
    c.assign((size_t)nr_, 0.0);
    tcoeffunc_(t, c.data());
}

void FSPMat::set_coefficients(const std::vector<double> &c)
{
    throw_on(fsp_mat_set_coefs(h_, c.data(), nullptr), "FSPMat::set_coefficients");
}

void FSPMat::action(double t, double *x, double *y)
{
    tcoeffunc_(t, tcoef_.data());
    throw_on(fsp_mat_apply(h_, tcoef_.data(), x, y, 1, nullptr), "FSPMat::action");
}

int FSPMat::get_n_rows() { return rows_; }
int FSPMat::get_n_species() { return ns_; }
int FSPMat::get_n_reactions() { return nr_; }

// ------------------------------------------------------------ KryExpvFSP
void KryExpvFSP::trampoline(void *ctx, const double *x, double *y)
{
    (*(MVFun *)ctx)(const_cast<double *>(x), y);
}

KryExpvFSP::KryExpvFSP(double t_final, MVFun &matvec, thrust_dvec &v, int m_, double tol,
                       bool iop, int q, double anorm)
    : IOP(iop), q_iop(q), m(m_), matvec_(matvec), sol_vec_(v)
{
    throw_on(fsp_expv_create(nullptr, &KryExpvFSP::trampoline, &matvec_, (int64_t)v.size(), t_final,
                             m, tol, iop ? 1 : 0, q, anorm, nullptr, &h_),
             "KryExpvFSP::KryExpvFSP");
    last_iop_ = IOP;
    last_q_ = q_iop;
}

KryExpvFSP::KryExpvFSP(double t_final, FSPMat &A, double t_coef, thrust_dvec &v, int m_,
                       double tol, bool iop, int q, double anorm)
    : IOP(iop), q_iop(q), m(m_), sol_vec_(v)
{
    std::vector<double> c;
    A.coefficients(t_coef, c);
    A.set_coefficients(c);
    throw_on(fsp_expv_create(A.handle(), nullptr, nullptr, (int64_t)v.size(), t_final, m, tol,
                             iop ? 1 : 0, q, anorm, nullptr, &h_),
             "KryExpvFSP::KryExpvFSP");
    last_iop_ = IOP;
    last_q_ = q_iop;
}

KryExpvFSP::~KryExpvFSP()
{
    if (h_) fsp_expv_destroy(h_);
}

void KryExpvFSP::sync_knobs()
{
    if ((int)IOP != last_iop_ || q_iop != last_q_) {
        throw_on(fsp_expv_set_iop(h_, IOP ? 1 : 0, q_iop), "KryExpvFSP");
        last_iop_ = IOP;
        last_q_ = q_iop;
    }
    if (max_reject != last_reject_ || max_nstep != last_nstep_) {
        throw_on(fsp_expv_set_limits(h_, max_reject, max_nstep), "KryExpvFSP");
        last_reject_ = max_reject;
        last_nstep_ = max_nstep;
    }
}

void KryExpvFSP::reset_time() { throw_on(fsp_expv_reset_time(h_), "KryExpvFSP::reset_time"); }

void KryExpvFSP::set_final_time(double t_final)
{
    throw_on(fsp_expv_set_final_time(h_, t_final), "KryExpvFSP::set_final_time");
}

void KryExpvFSP::solve()
{
    sync_knobs();
    throw_on(fsp_expv_solve(h_, sol_vec_.data()), "KryExpvFSP::solve");
}

void KryExpvFSP::step()
{
    sync_knobs();
    throw_on(fsp_expv_step(h_, sol_vec_.data()), "KryExpvFSP::step");
}

bool KryExpvFSP::final_time_reached()
{
    int r = 0;
    throw_on(fsp_expv_final_time_reached(h_, &r), "KryExpvFSP::final_time_reached");
    return r != 0;
}

fsp_expv_stats KryExpvFSP::stats()
{
    fsp_expv_stats s;
    throw_on(fsp_expv_get_stats(h_, &s), "KryExpvFSP::stats");
    return s;
}

}  // namespace cuFSP
