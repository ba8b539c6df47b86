// fsp_assemble.h -- per-row generator values from a device model, shared by
// the matrix assembly kernels (fsp_mat.hip) and the lockstep batch's HYB
// fill (fsp_batch_lockstep.hip).
#pragma once

#include "fsp_device.h"

namespace fsp {

// Propensity a_r at state x (shift = 0) or at the inflow source x - nu_r
// (shift = 1): from the dense R x n table when given, else the product of
// the compact factor tables.
template <int D>
__device__ inline double prop_at(const DevModel *__restrict__ dm, const double *tab,
                                 const double *__restrict__ table, int r, uint32_t i,
                                 const int (&xs)[D], int shift)
{
    if (table) {
        const uint32_t idx = shift ? (uint32_t)((int64_t)i - dm->rx[r].delta) : i;
        return table[(size_t)r * dm->n + idx];
    }
    return prop_tab<D>(dm->rx[r], xs, tab, shift);
}

// Row values shared by CSR and HYB fills.  F64 (mixed = 0): slot value =
// sum_{r in u} c_r a_r(x - nu) (first term assigned, reaction order),
// diag = -(sum_r c_r a_r(x)), out = sum_{r leaves} c_r a_r(x).  Mixed: the
// per-slot outflow rates w_u(x) = sum_{r in u} c_r a_r(x) are rounded to
// FP32 once, and diag / out are exact FP64 sums of those rounded rates, so
// each column of the stored generator sums to exactly zero.
template <int D>
__device__ inline double slot_value(const DevModel *__restrict__ dm, const double *tab,
                                    const double *__restrict__ table,
                                    const double *__restrict__ coefs, int u, uint32_t i,
                                    const int (&xs)[D], int shift)
{
    double v = 0.0;
    for (int k = dm->slot_ptr[u]; k < dm->slot_ptr[u + 1]; ++k) {
        const int r = dm->slot_rxn[k];
        const double t = coefs[r] * prop_at<D>(dm, tab, table, r, i, xs, shift);
        v = (k == dm->slot_ptr[u]) ? t : v + t;
    }
    return v;
}

template <int D>
__device__ inline void diag_and_out(const DevModel *__restrict__ dm, const double *tab,
                                    const double *__restrict__ table,
                                    const double *__restrict__ coefs, int mixed, uint32_t i,
                                    const int (&xs)[D], double &diag, double &out, int &has)
{
    double acc = 0.0, o = 0.0;
    has = 0;
    if (!mixed) {
        for (int r = 0; r < dm->R; ++r) {
            if (dm->rx[r].zero_nu) continue;
            const double t = coefs[r] * prop_at<D>(dm, tab, table, r, i, xs, 0);
            acc += t;
            if (target_leaves<D>(dm->rx[r], xs, dm)) {
                o += t;
                has = 1;
            }
        }
    } else {
        for (int u = 0; u < dm->U; ++u) {
            const double w = (double)(float)slot_value<D>(dm, tab, table, coefs, u, i, xs, 0);
            acc += w;
            if (target_leaves<D>(dm->rx[dm->slot_rxn[dm->slot_ptr[u]]], xs, dm)) {
                o += w;
                has = 1;
            }
        }
    }
    diag = -acc;
    out = o;
}

}  // namespace fsp
