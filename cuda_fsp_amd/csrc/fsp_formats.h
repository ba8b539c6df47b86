// fsp_formats.h -- generator storage formats and their row kernels.
//
// Every row kernel works on tiles of consecutive rows (kBlock, or a larger
// compile-time height for the CSR apply), one thread per row, grid-stride
// over tiles.  An "x accessor" X supplies input values:
// x(c) returns the FP64 value of input entry c (a plain load, or a scaled
// and rounded Krylov vector).  Each tile function returns the row sum for
// row tile*kBlock + threadIdx.x and adds the thread's share of the sink row
// (outflow from the box) to `sink`.
#pragma once

#include <type_traits>

#include "fsp_device.h"

namespace fsp {

// --------------------------------------------------------------------- CSR
// Merged CSR.  F64: the diagonal is an ordinary row entry (columns
// ascending), diag == nullptr.  MIXED: rows hold only off-diagonals in FP32
// and the FP64 diagonal is separate, so every column of the stored
// generator sums to exactly zero.  Sink row: entries sorted by column;
// sink_off[i] = first sink entry with column >= i (i = 0..n).
template <typename V>
struct CsrF {
    const int *__restrict__ rp;
    const int *__restrict__ cols;
    const V *__restrict__ vals;
    const double *__restrict__ diag;
    const int *__restrict__ sink_off;
    const int *__restrict__ sink_cols;
    const double *__restrict__ sink_vals;
    uint32_t n;
    int cap;  // LDS product slots per chunk (>= rows per tile)
};

constexpr int kCsrPer = 16;  // default LDS products per thread per chunk (cap <= PER * tile rows)

// Phase 1 streams the tile's nonzeros (coalesced) into LDS as products: all
// of a thread's column indices and values of the chunk are loaded first
// (non-temporal: they are read once), then all its gathers of x, so a chunk
// costs two memory round trips.  Every load is unconditional (slots past
// the chunk re-read its last entry, a valid index) so that the loads of a
// phase are issued back to back instead of one wait per slot; only the LDS
// stores are predicated.  Phase 2 lets each thread add its row's products
// in column order.
template <typename V, class X, int TB = kBlock, int PER = kCsrPer>
__device__ inline double csr_tile(const CsrF<V> &A, uint32_t tile, const X &x,
                                  double *prod, dd &sink)
{
    const uint32_t r0 = tile * TB;
    const uint32_t r1 = min(r0 + (uint32_t)TB, A.n);
    const uint32_t i = r0 + threadIdx.x;
    const int k0 = A.rp[r0], k1 = A.rp[r1];
    const int kb = (i < A.n) ? A.rp[i] : k1;
    const int ke = (i < A.n) ? A.rp[i + 1] : k1;
    double y = 0.0;
    for (int c0 = k0; c0 < k1; c0 += A.cap) {
        const int c1 = min(c0 + A.cap, k1);  // c1 > c0
        __syncthreads();
        int cl[PER];
        V vl[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int k = min(c0 + (int)threadIdx.x + q * TB, c1 - 1);
            cl[q] = __builtin_nontemporal_load(A.cols + k);
            vl[q] = __builtin_nontemporal_load(A.vals + k);
        }
        double xv[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) xv[q] = x(cl[q]);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int k = c0 + (int)threadIdx.x + q * TB;
            if (k < c1) prod[k - c0] = (double)vl[q] * xv[q];
        }
        __syncthreads();
        const int a = max(kb, c0), b = min(ke, c1);
        for (int k = a; k < b; ++k) y += prod[k - c0];
    }
    if (A.diag && i < A.n) y = A.diag[i] * x(i) + y;
    if (A.sink_off) {
        const int s0 = A.sink_off[r0], s1 = A.sink_off[r1];
        for (int k = s0 + (int)threadIdx.x; k < s1; k += TB)
            sink = dd_add_prod(sink, A.sink_vals[k], x(A.sink_cols[k]));
    }
    return y;
}

// --------------------------------------------------------------------- HYB
// Merged FP64 diagonal + U ELL slots (SoA [U][n]), one per unique net
// change.  The columns are implicit: slot u of row i sits at column
// i + off[u] (off[u] = -delta_u, the inflow source; +delta_u for the
// transpose), so only values are stored.  A neighbour outside the box
// stores value 0; its column is then any valid state (the index itself is
// clamped to row i when it falls outside 0..n-1), and 0 * x adds nothing.
template <typename V>
struct HybF {
    const double *__restrict__ diag;
    const V *__restrict__ slot_vals;
    const int *__restrict__ slot_off;     // U column offsets
    const double *__restrict__ sink_out;  // dense outflow rates or nullptr
    uint32_t n;
    int U;
};

__device__ inline uint32_t hyb_col(uint32_t i, int off, uint32_t n)
{
    const int64_t c = (int64_t)i + off;
    return (c >= 0 && c < (int64_t)n) ? (uint32_t)c : i;
}

// A zero slot value contributes +0 whatever x holds at its implicit column:
// the column of an out-of-box neighbour names some other state (or row i
// itself at the ends of the vector), and a non-finite x there must not reach
// this row.
// Row i with a compile-time slot count: all slot values are loaded first
// (non-temporal: streamed once per product), with all gathers of x at the
// implicit columns, then the sum in slot order, so a row costs one memory
// round trip whatever U is.
// The loads of one such row (hyb_load) and its sum (hyb_sum), apart so that
// a kernel can issue a later row's loads before an earlier row's sum.
template <int NU, typename V>
struct HybLoads {
    double d, xi;
    V v[NU];
    double xc[NU];
};

template <int NU, typename V, class X>
__device__ inline void hyb_load(const HybF<V> &A, uint32_t i, const X &x, HybLoads<NU, V> &L)
{
    L.xi = x(i);
    L.d = __builtin_nontemporal_load(A.diag + i);
#pragma unroll
    for (int u = 0; u < NU; ++u) L.v[u] = __builtin_nontemporal_load(A.slot_vals + (size_t)u * A.n + i);
#pragma unroll
    for (int u = 0; u < NU; ++u) L.xc[u] = x(hyb_col(i, A.slot_off[u], A.n));
}

// hyb_sum of loads whose x entries are still to be mapped by xs (e.g. the
// Krylov scaling, applied to raw loads issued before the scale was known).
template <int NU, typename V, class XS>
__device__ inline double hyb_sum_x(const HybLoads<NU, V> &L, const XS &xs)
{
    double y = L.d * xs(L.xi);
#pragma unroll
    for (int u = 0; u < NU; ++u) y += L.v[u] != (V)0 ? (double)L.v[u] * xs(L.xc[u]) : 0.0;
    return y;
}

struct XIdentity {
    __device__ double operator()(double v) const { return v; }
};

template <int NU, typename V>
__device__ inline double hyb_sum(const HybLoads<NU, V> &L)
{
    return hyb_sum_x<NU>(L, XIdentity{});
}

// Raw entries of a vector (for hyb_load before a scale is known).
struct RawX {
    const double *__restrict__ in;
    __device__ double operator()(uint32_t c) const { return in[c]; }
};

template <int NU, typename V, class X>
__device__ inline double hyb_row_n(const HybF<V> &A, uint32_t i, const X &x, dd &sink)
{
    HybLoads<NU, V> L;
    hyb_load<NU>(A, i, x, L);
    const double y = hyb_sum<NU>(L);
    if (A.sink_out) sink = dd_add_prod(sink, A.sink_out[i], L.xi);
    return y;
}

template <typename V, class X>
__device__ inline double hyb_row(const HybF<V> &A, uint32_t i, const X &x, dd &sink)
{
    if (i >= A.n) return 0.0;
    switch (A.U) {
    case 1: return hyb_row_n<1>(A, i, x, sink);
    case 2: return hyb_row_n<2>(A, i, x, sink);
    case 3: return hyb_row_n<3>(A, i, x, sink);
    case 4: return hyb_row_n<4>(A, i, x, sink);
    case 5: return hyb_row_n<5>(A, i, x, sink);
    case 6: return hyb_row_n<6>(A, i, x, sink);
    case 7: return hyb_row_n<7>(A, i, x, sink);
    case 8: return hyb_row_n<8>(A, i, x, sink);
    case 9: return hyb_row_n<9>(A, i, x, sink);
    case 10: return hyb_row_n<10>(A, i, x, sink);
    case 11: return hyb_row_n<11>(A, i, x, sink);
    case 12: return hyb_row_n<12>(A, i, x, sink);
    default: break;
    }
    const double xi = x(i);
    double y = A.diag[i] * xi;
    for (int u = 0; u < A.U; ++u) {
        const size_t o = (size_t)u * A.n + i;
        const V v = A.slot_vals[o];
        y += v != (V)0 ? (double)v * x(hyb_col(i, A.slot_off[u], A.n)) : 0.0;
    }
    if (A.sink_out) sink = dd_add_prod(sink, A.sink_out[i], xi);
    return y;
}

// -------------------------------------------------------------------- KRON
// Matrix-free Kronecker-sum apply (the reference's fspmat_sdkron_mv formula,
// src/fspmat_kron_kernels.cu:69-128): per reaction, the inflow from
// i - delta_r times the source propensity minus x_i times the local
// propensity.  Factor tables are staged in LDS once per workgroup (`tab`);
// with use_table the propensities come from a dense R x n table instead.
template <typename V>
struct KronF {
    const DevModel *__restrict__ dm;
    const V *__restrict__ tab_g;          // compact factor table (global)
    const double *__restrict__ prop_table; // R x n (use_table)
    const double *__restrict__ coefs;      // device, R
    uint32_t n;
    int R;
    int tab_size;
    int with_sink;
    int single;   // >= 0: only this reaction
    double single_coef;
};

template <int D, typename V, class X>
__device__ inline double kron_row(const KronF<V> &A, uint32_t i, const X &x, const V *tab,
                                  dd &sink)
{
    if (i >= A.n) return 0.0;
    int xs[D];
    decode<D>(i, A.dm, xs);
    const double xh = x(i);
    double y = 0.0, out = 0.0;
    const int rbeg = A.single >= 0 ? A.single : 0;
    const int rend = A.single >= 0 ? A.single + 1 : A.R;
    for (int r = rbeg; r < rend; ++r) {
        const RxnDev &rx = A.dm->rx[r];
        if (rx.zero_nu) continue;
        const double c = A.single >= 0 ? A.single_coef : A.coefs[r];
        double alpha = 0.0, xs_src = 0.0, beta;
        const bool in = source_in_box<D>(rx, xs, A.dm);
        if (A.prop_table) {
            beta = A.prop_table[(size_t)r * A.n + i];
            if (in) {
                const uint32_t src = (uint32_t)((int64_t)i - rx.delta);
                alpha = A.prop_table[(size_t)r * A.n + src];
                xs_src = x(src);
            }
        } else {
            beta = prop_tab<D>(rx, xs, tab, 0);
            if (in) {
                alpha = prop_tab<D>(rx, xs, tab, 1);
                xs_src = x((uint32_t)((int64_t)i - rx.delta));
            }
        }
        y += c * (xs_src * alpha - xh * beta);
        if (A.with_sink && target_leaves<D>(rx, xs, A.dm)) out += c * beta;
    }
    if (A.with_sink) sink = dd_add_prod(sink, out, xh);
    return y;
}

// Per-reaction geometry of the Kronecker apply, passed by value to the
// kernels (so it sits in the kernel-argument segment and every field is a
// scalar load shared by the wave): the inflow source offset, each species'
// net change and factor-table offset (-1: the factor is 1 everywhere), and
// the box radices.  Models with more than kKronPlanR reactions use
// kron_row.
constexpr int kKronPlanR = 24;
struct KronRx {
    int delta;                          // inflow source of row i: i - delta
    int live;                           // 0: net change zero, contributes 0
    int nu[kMaxSpecies];
    int foff[kMaxSpecies];
};
struct KronPlan {
    int R;
    int radix[kMaxSpecies];
    KronRx rx[kKronPlanR];              // one contiguous record per reaction
};

// kron_row with the species loop unrolled (no per-species selects on the
// digits) and the reaction geometry read from a KronPlan: per reaction,
// the source range checks of the species it changes, the factor products
// at x and at x - nu in species order, one gather, then
// y += c (x_src alpha - x_i beta) in reaction order, as the oracle's
// orc_kron_apply does, so rows agree bit for bit.  The gather is issued
// for every reaction (row i itself when the source is outside the box,
// weighted by 0).
template <int D, typename V, class X>
__device__ inline double kron_row_plan(const KronF<V> &A, const KronPlan &P, uint32_t i,
                                       const X &x, const V *tab, dd &sink)
{
    if (i >= A.n) return 0.0;
    int xs[D];
    decode<D>(i, A.dm, xs);
    const double xh = x(i);
    double y = 0.0, out = 0.0;
    for (int r = 0; r < P.R; ++r) {
        const KronRx q = P.rx[r];  // the whole record in scalar registers
        if (!q.live) continue;
        const double c = A.coefs[r];
        // source range checks and target leaves of the changed species
        bool in = true, leaves = false;
#pragma unroll
        for (int s = 0; s < D; ++s) {
            if (q.nu[s] != 0) {
                const int src = xs[s] - q.nu[s], tgt = xs[s] + q.nu[s];
                in = in && src >= 0 && src < P.radix[s];
                leaves = leaves || tgt < 0 || tgt >= P.radix[s];
            }
        }
        // every factor load issued before the products
        double fb[D], fa[D];
#pragma unroll
        for (int s = 0; s < D; ++s) {
            fb[s] = 1.0;
            fa[s] = 1.0;
            if (q.foff[s] >= 0) {
                fb[s] = (double)tab[q.foff[s] + xs[s]];
                if (q.nu[s] != 0)
                    fa[s] = (double)tab[q.foff[s] + min(max(xs[s] - q.nu[s], 0), P.radix[s] - 1)];
            }
        }
        const uint32_t si = in ? (uint32_t)((int64_t)i - q.delta) : i;
        const double xsrc = x(si);
        // products in species order (factors equal to 1 are skipped: exact)
        double alpha = 1.0, beta = 1.0;
#pragma unroll
        for (int s = 0; s < D; ++s) {
            if (q.foff[s] >= 0) {
                beta *= fb[s];
                alpha *= (q.nu[s] != 0) ? fa[s] : fb[s];
            }
        }
        y += c * (xsrc * (in ? alpha : 0.0) - xh * beta);
        if (A.with_sink && leaves) out += c * beta;
    }
    if (A.with_sink) sink = dd_add_prod(sink, out, xh);
    return y;
}

// ------------------------------------------------------ KRON, slot form
// The reference's Kronecker apply sums over reactions (per row R factor
// products).  Grouped by net change (the HYB slots), the same operator is
//   y_i = sum_u [x_i - nu_u in box] h_u(x_{s_u} - nu_{u,s_u}) x_{i - delta_u}
//         - (sum_u h_u(x_{s_u})) x_i,
// with h_u(k) = sum_{r in slot u} c_r f_{r,s_u}(k) when every member's
// propensity varies with one species s_u at most (its other factors are 1):
// per row, U slot lookups in place of R products.  The tables depend on the
// coefficients and are rebuilt with them; a model with a slot whose members
// vary with different species, or with tables too large for LDS, keeps
// kron_row_plan.  The sink collects sum_u [x + nu_u leaves] h_u(x_{s_u}).
constexpr int kKronSlots = 24;
struct KronSlot {
    int delta;                 // inflow source of row i: i - delta
    int sp;                    // species the table runs over, -1: constant
    int nu_sp;                 // net change of that species
    int toff;                  // table offset
    int nu[kMaxSpecies];       // net change per species
};
struct KronSlotPlan {
    int U;
    int tab_size;              // table entries
    int radix[kMaxSpecies];
    KronSlot u[kKronSlots];
};

template <int D, typename V, class X>
__device__ inline double kron_row_slot(const KronF<V> &A, const KronSlotPlan &P, uint32_t i,
                                       const X &x, const V *tab, dd &sink)
{
    if (i >= A.n) return 0.0;
    int xs[D];
    decode<D>(i, A.dm, xs);
    const double xh = x(i);
    double acc = 0.0, dsum = 0.0, out = 0.0;
    for (int u = 0; u < P.U; ++u) {
        const KronSlot q = P.u[u];  // the whole record in scalar registers
        bool in = true, leaves = false;
        int k = 0;
#pragma unroll
        for (int s = 0; s < D; ++s) {
            if (q.nu[s] != 0) {
                const int src = xs[s] - q.nu[s], tgt = xs[s] + q.nu[s];
                in = in && src >= 0 && src < P.radix[s];
                leaves = leaves || tgt < 0 || tgt >= P.radix[s];
            }
            if (s == q.sp) k = xs[s];
        }
        const int ks = q.sp >= 0 ? min(max(k - q.nu_sp, 0), P.radix[q.sp < 0 ? 0 : q.sp] - 1) : 0;
        const double hx = (double)tab[q.toff + k];
        const double hs = (double)tab[q.toff + ks];
        const double xsrc = x(in ? (uint32_t)((int64_t)i - q.delta) : i);
        acc += (in ? hs : 0.0) * xsrc;
        dsum += hx;
        if (A.with_sink && leaves) out += hx;
    }
    if (A.with_sink) sink = dd_add_prod(sink, out, xh);
    return acc - dsum * xh;
}

// Slot plan of a model (false: some slot's members vary with different
// species, more than kKronSlots slots, or no factor tables).
inline bool kron_slot_plan(const HostModel &hm, KronSlotPlan &P)
{
    const DevModel &dm = hm.dm;
    if (!hm.has_factors || dm.U > kKronSlots) return false;
    P = KronSlotPlan{};
    P.U = dm.U;
    for (int s = 0; s < kMaxSpecies; ++s) P.radix[s] = s < dm.d ? dm.radix[s] : 1;
    int off = 0;
    for (int u = 0; u < dm.U; ++u) {
        KronSlot &q = P.u[u];
        q.delta = (int)dm.slot_delta[u];
        q.sp = -1;
        for (int m = dm.slot_ptr[u]; m < dm.slot_ptr[u + 1]; ++m) {
            const RxnDev &rx = dm.rx[dm.slot_rxn[m]];
            for (int k = 0; k < rx.nsp; ++k) {
                q.nu[rx.sp[k]] = rx.nu[k];
                if (rx.foff[k] < 0) continue;
                if (q.sp >= 0 && q.sp != rx.sp[k]) return false;
                q.sp = rx.sp[k];
            }
        }
        q.nu_sp = q.sp >= 0 ? q.nu[q.sp] : 0;
        q.toff = off;
        off += q.sp >= 0 ? dm.radix[q.sp] : 1;
    }
    P.tab_size = off;
    return true;
}

// h_u(k) = sum over the slot's members in reaction order of c_r f_{r,s_u}(k)
// (f = 1 for a member without a factor in s_u), from 0.0.
inline void kron_slot_tables(const HostModel &hm, const KronSlotPlan &P, const double *coefs,
                             std::vector<double> &tab)
{
    const DevModel &dm = hm.dm;
    tab.assign((size_t)P.tab_size, 0.0);
    for (int u = 0; u < P.U; ++u) {
        const KronSlot &q = P.u[u];
        const int len = q.sp >= 0 ? dm.radix[q.sp] : 1;
        for (int m = dm.slot_ptr[u]; m < dm.slot_ptr[u + 1]; ++m) {
            const int r = dm.slot_rxn[m];
            const RxnDev &rx = dm.rx[r];
            int fo = -1;
            for (int k = 0; k < rx.nsp; ++k)
                if (rx.sp[k] == q.sp) fo = rx.foff[k];
            for (int k = 0; k < len; ++k) {
                const double f = fo >= 0 ? hm.compact[(size_t)fo + k] : 1.0;
                tab[(size_t)q.toff + k] += coefs[r] * f;
            }
        }
    }
}

// Transpose (adjoint) row: (A^T x)_j = sum_r c_r a_r(x_j) (x_{j+delta_r}
// [target in box] - x_j).  No sink column (the sink is absorbing).
template <int D, typename V, class X>
__device__ inline double kron_row_t(const KronF<V> &A, uint32_t j, const X &x, const V *tab)
{
    if (j >= A.n) return 0.0;
    int xs[D];
    decode<D>(j, A.dm, xs);
    const double xj = x(j);
    double y = 0.0;
    const int rbeg = A.single >= 0 ? A.single : 0;
    const int rend = A.single >= 0 ? A.single + 1 : A.R;
    for (int r = rbeg; r < rend; ++r) {
        const RxnDev &rx = A.dm->rx[r];
        if (rx.zero_nu) continue;
        const double c = A.single >= 0 ? A.single_coef : A.coefs[r];
        const double a = A.prop_table ? A.prop_table[(size_t)r * A.n + j]
                                      : prop_tab<D>(rx, xs, tab, 0);
        const bool leaves = target_leaves<D>(rx, xs, A.dm);
        const double xt = leaves ? 0.0 : x((uint32_t)((int64_t)j + rx.delta));
        y += c * a * (xt - xj);
    }
    return y;
}

// ------------------------------------------------------------ host views
#define FSP_SWITCH_D(d, CALL)                                                  \
    switch (d) {                                                               \
    case 1: { constexpr int D_ = 1; CALL; } break;                             \
    case 2: { constexpr int D_ = 2; CALL; } break;                             \
    case 3: { constexpr int D_ = 3; CALL; } break;                             \
    case 4: { constexpr int D_ = 4; CALL; } break;                             \
    case 5: { constexpr int D_ = 5; CALL; } break;                             \
    case 6: { constexpr int D_ = 6; CALL; } break;                             \
    case 7: { constexpr int D_ = 7; CALL; } break;                             \
    case 8: { constexpr int D_ = 8; CALL; } break;                             \
    default: return ::fsp::fail(FSP_ERR_UNSUPPORTED, "n_species > 8");         \
    }

static constexpr size_t kLdsTableMax = 96 * 1024;

// tb: rows per tile (threads per workgroup) of the kernel using the view;
// per: its csr_tile PER
template <typename V>
inline CsrF<V> csr_view(const MatImpl &M, int tb = kBlock, int per = kCsrPer)
{
    CsrF<V> A;
    A.rp = M.rp;
    A.cols = M.cols;
    A.vals = (const V *)M.vals;
    A.diag = M.prec == FSP_MIXED ? (const double *)M.diag : nullptr;
    A.sink_off = M.with_sink ? M.sink_block_ptr : nullptr;
    A.sink_cols = M.sink_cols;
    A.sink_vals = (const double *)M.sink_vals;
    A.n = (uint32_t)M.n;
    int cap = tb * (M.max_row > 1 ? M.max_row : 1);
    // at most per products per thread per chunk
    A.cap = cap > per * tb ? per * tb : cap;
    return A;
}

template <typename V>
inline HybF<V> hyb_view(const MatImpl &M)
{
    HybF<V> A;
    A.diag = (const double *)M.diag;
    A.slot_vals = (const V *)M.slot_vals;
    A.slot_off = M.slot_off;
    A.sink_out = M.with_sink ? (const double *)M.sink_out : nullptr;
    A.n = (uint32_t)M.n;
    A.U = M.hm.dm.U;
    return A;
}

// Plan of a model with at most kKronPlanR reactions (false otherwise).
inline bool kron_plan(const HostModel &hm, KronPlan &P)
{
    const DevModel &dm = hm.dm;
    if (dm.R > kKronPlanR || !hm.has_factors) return false;
    P = KronPlan{};
    P.R = dm.R;
    for (int s = 0; s < kMaxSpecies; ++s) P.radix[s] = s < dm.d ? dm.radix[s] : 1;
    for (int r = 0; r < dm.R; ++r) {
        const RxnDev &rx = dm.rx[r];
        KronRx &q = P.rx[r];
        q.delta = (int)rx.delta;
        q.live = rx.zero_nu ? 0 : 1;
        for (int s = 0; s < kMaxSpecies; ++s) q.foff[s] = -1;
        for (int k = 0; k < rx.nsp; ++k) {
            q.nu[rx.sp[k]] = rx.nu[k];
            q.foff[rx.sp[k]] = rx.foff[k];
        }
    }
    return true;
}

template <typename V>
inline KronF<V> kron_view(const MatImpl &M, int single, double single_coef)
{
    KronF<V> A;
    A.dm = M.d_model;
    A.tab_g = std::is_same<V, float>::value ? (const V *)(const void *)M.d_compact_f
                                            : (const V *)(const void *)M.d_compact;
    A.prop_table = M.hm.prop_table_dev;
    A.coefs = M.d_coefs;
    A.n = (uint32_t)M.n;
    A.R = M.hm.dm.R;
    A.tab_size = M.hm.dm.compact_size;
    A.with_sink = M.with_sink;
    A.single = single;
    A.single_coef = single_coef;
    return A;
}

}  // namespace fsp
