// fsp_model.hip -- host-side validation and geometry of an fsp_model
// (state-space box, mixed-radix strides, compact factor tables, unique
// net-change slots).  Pure host code, shared by the matrix and the batch
// paths; it replaces the host half of the reference's model set-up
// (src/cme_util.cu:56-62 box size, src/fspmat_kron_kernels.cu:170-192
// factor tables and per-reaction offsets, src/fspmat_csr_kernels.cu:152-160
// stoichiometry upload).
#include <algorithm>
#include <cstring>
#include <utility>
#include <vector>

#include "fsp_internal.h"

namespace fsp {

fsp_status build_host_model(const fsp_model *m, HostModel &hm)
{
    if (!m) return fail(FSP_ERR_INVALID, "model is NULL");
    const int d = m->n_species, R = m->n_reactions;
    if (d < 1 || d > kMaxSpecies)
        return fail(FSP_ERR_UNSUPPORTED, "n_species must be in 1.." + std::to_string(kMaxSpecies));
    if (R < 1 || R > kMaxReactions)
        return fail(FSP_ERR_UNSUPPORTED, "n_reactions must be in 1.." + std::to_string(kMaxReactions));
    if (!m->bounds || !m->stoich_vals || !m->stoich_cols || !m->stoich_rowptrs)
        return fail(FSP_ERR_INVALID, "bounds and stoichiometry are required");
    if ((m->factors == nullptr) == (m->prop_table_dev == nullptr))
        return fail(FSP_ERR_INVALID, "give exactly one of factors / prop_table_dev");
    DevModel &dm = hm.dm;
    std::memset(&dm, 0, sizeof(dm));
    dm.d = d;
    dm.R = R;
    int64_t n = 1;
    for (int s = 0; s < d; ++s) {
        if (m->bounds[s] < 0) return fail(FSP_ERR_INVALID, "negative bound");
        dm.radix[s] = m->bounds[s] + 1;
        dm.stride[s] = n;
        dm.fd[s] = make_fastdiv((uint32_t)dm.radix[s]);
        n *= dm.radix[s];
        if (n > 2147483647LL)
            return fail(FSP_ERR_OVERFLOW, "more than 2^31-1 states in the FSP box");
    }
    dm.n = n;
    dm.use_table = m->prop_table_dev != nullptr;
    hm.prop_table_dev = m->prop_table_dev;
    hm.has_factors = m->factors != nullptr;
    hm.nu_dense.assign((size_t)R * d, 0);
    for (int r = 0; r < R; ++r) {
        if (m->stoich_rowptrs[r + 1] < m->stoich_rowptrs[r])
            return fail(FSP_ERR_INVALID, "stoich row pointers decrease");
        for (int k = m->stoich_rowptrs[r]; k < m->stoich_rowptrs[r + 1]; ++k) {
            const int c = m->stoich_cols[k];
            if (c < 0 || c >= d) return fail(FSP_ERR_INVALID, "stoich column out of range");
            hm.nu_dense[(size_t)r * d + c] += m->stoich_vals[k];
        }
    }
    int foff_stride = 0;
    for (int s = 0; s < d; ++s) foff_stride += dm.radix[s];
    hm.compact.clear();
    for (int r = 0; r < R; ++r) {
        RxnDev &rx = dm.rx[r];
        rx.delta = 0;
        rx.zero_nu = 1;
        rx.nsp = 0;
        int off = 0;
        for (int s = 0; s < d; ++s) {
            const int nu = hm.nu_dense[(size_t)r * d + s];
            rx.delta += (int64_t)nu * dm.stride[s];
            if (nu != 0) rx.zero_nu = 0;
            bool trivial = true;
            const double *f = nullptr;
            if (m->factors) {
                f = m->factors + (size_t)r * foff_stride + off;
                for (int i = 0; i < dm.radix[s]; ++i) trivial = trivial && (f[i] == 1.0);
            }
            if (nu != 0 || !trivial) {
                const int k = rx.nsp++;
                rx.sp[k] = s;
                rx.nu[k] = nu;
                rx.foff[k] = -1;
                if (!trivial) {
                    rx.foff[k] = (int)hm.compact.size();
                    hm.compact.insert(hm.compact.end(), f, f + dm.radix[s]);
                }
            }
            off += dm.radix[s];
        }
    }
    dm.compact_size = (int)hm.compact.size();
    // unique net-change slots in order of first appearance
    int U = 0;
    std::vector<int> first;
    for (int r = 0; r < R; ++r) {
        dm.rx[r].slot = -1;
        if (dm.rx[r].zero_nu) continue;
        int slot = -1;
        for (int u = 0; u < U && slot < 0; ++u) {
            bool same = true;
            for (int s = 0; s < d; ++s)
                same = same && hm.nu_dense[(size_t)first[u] * d + s] == hm.nu_dense[(size_t)r * d + s];
            if (same) slot = u;
        }
        if (slot < 0) {
            first.push_back(r);
            slot = U++;
        }
        dm.rx[r].slot = slot;
    }
    dm.U = U;
    int pos = 0;
    for (int u = 0; u < U; ++u) {
        dm.slot_ptr[u] = pos;
        dm.slot_delta[u] = dm.rx[first[u]].delta;
        for (int r = 0; r < R; ++r)
            if (dm.rx[r].slot == u) dm.slot_rxn[pos++] = r;
    }
    dm.slot_ptr[U] = pos;
    // column order of a merged row: column i - delta ascending = delta descending
    std::vector<std::pair<int64_t, int>> ent;
    ent.push_back({0, -1});
    for (int u = 0; u < U; ++u) ent.push_back({dm.slot_delta[u], u});
    std::sort(ent.begin(), ent.end(), [](auto &a, auto &b) { return a.first > b.first; });
    for (int k = 0; k <= U; ++k) dm.csr_order[k] = ent[k].second;
    return FSP_OK;
}

}  // namespace fsp
