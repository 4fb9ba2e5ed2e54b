// plan_shots.cpp -- the shots of a batch: slot order (longest record first), source
// injection groups, receiver lists and the per-tile lists the step kernels read.
#include "plan.h"

namespace psvplan {

template <typename T>
psvfwi_status Plan<T>::set_shots(int n, const psvfwi_shot_desc* sd)
{
    if (n < 1 || !sd) return fail(PSVFWI_E_ARG, "set_shots: need at least one shot");
    std::vector<ShotHost> shots((size_t)n);
    for (int i = 0; i < n; i++) {
        const psvfwi_shot_desc& d = sd[i];
        if (d.nt < 1 || d.nsrc < 0 || d.nrec < 0) return fail(PSVFWI_E_ARG, "set_shots: bad nt/nsrc/nrec");
        ShotHost& s = shots[(size_t)i];
        s.nt = d.nt;
        for (int k = 0; k < d.nsrc; k++) {
            if (!in_grid(d.src_z[k], d.src_x[k])) return fail(PSVFWI_E_ARG, "set_shots: source outside the grid");
            if (d.src_comp[k] != 1 && d.src_comp[k] != 2 && d.src_comp[k] != 4)
                return fail(PSVFWI_E_ARG, "set_shots: source component must be 1, 2 or 4");
            s.sx.push_back(d.src_x[k]);
            s.sz.push_back(d.src_z[k]);
            s.sc.push_back(d.src_comp[k]);
        }
        if (d.nsrc) s.ssig.assign(d.src_signal, d.src_signal + (size_t)d.nsrc * d.nt);
        for (int k = 0; k < d.nrec; k++) {
            if (!in_grid(d.rec_z[k], d.rec_x[k])) return fail(PSVFWI_E_ARG, "set_shots: receiver outside the grid");
            if (d.rec_comp[k] != 1 && d.rec_comp[k] != 2)
                return fail(PSVFWI_E_ARG, "set_shots: receiver component must be 1 or 2");
            s.rx.push_back(d.rec_x[k]);
            s.rz.push_back(d.rec_z[k]);
            s.rc.push_back(d.rec_comp[k]);
        }
    }
    shots_.swap(shots);
    nslots_ = n;
    all_slots();
    shot_of_slot_.resize((size_t)n);
    std::iota(shot_of_slot_.begin(), shot_of_slot_.end(), 0);
    std::stable_sort(shot_of_slot_.begin(), shot_of_slot_.end(),
                     [&](int a, int b) { return shots_[(size_t)a].nt > shots_[(size_t)b].nt; });
    slot_of_shot_.resize((size_t)n);
    for (int s = 0; s < n; s++) slot_of_shot_[(size_t)shot_of_slot_[(size_t)s]] = s;
    ntmax_ = shots_[(size_t)shot_of_slot_[0]].nt;
    nt_slot_.resize((size_t)n);
    for (int s = 0; s < n; s++) nt_slot_[(size_t)s] = shots_[(size_t)shot_of_slot_[(size_t)s]].nt;
    PSV_TRY(upload_tables());
    drop_results();
    have_obs_ = false;
    adj_ready_.assign((size_t)n, 0);
    return PSVFWI_OK;
}

// injection groups: entries sorted by (z, x, comp), original order kept
// inside a group, so co-located sources are added in index order
template <typename T>
psvfwi_status Plan<T>::build_injection(Tables& t, bool adjoint_sources, bool with_signals)
{
    std::vector<int> off(1, 0), grp, row;
    std::vector<size_t> sig_off;
    std::vector<T> sig;
    for (int s = 0; s < nslots_; s++) {
        const ShotHost& sh = shots_[(size_t)shot_of_slot_[(size_t)s]];
        const std::vector<int>& X = adjoint_sources ? sh.rx : sh.sx;
        const std::vector<int>& Z = adjoint_sources ? sh.rz : sh.sz;
        const std::vector<int>& Cc = adjoint_sources ? sh.rc : sh.sc;
        std::vector<int> idx(X.size());
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
            if (Z[(size_t)a] != Z[(size_t)b]) return Z[(size_t)a] < Z[(size_t)b];
            if (X[(size_t)a] != X[(size_t)b]) return X[(size_t)a] < X[(size_t)b];
            return Cc[(size_t)a] < Cc[(size_t)b];
        });
        const int base_row = (int)row.size();
        size_t k = 0;
        while (k < idx.size()) {
            size_t e = k;
            while (e < idx.size() && Z[(size_t)idx[e]] == Z[(size_t)idx[k]] &&
                   X[(size_t)idx[e]] == X[(size_t)idx[k]] && Cc[(size_t)idx[e]] == Cc[(size_t)idx[k]])
                e++;
            grp.push_back(Z[(size_t)idx[k]]);
            grp.push_back(X[(size_t)idx[k]]);
            grp.push_back(Cc[(size_t)idx[k]]);
            grp.push_back(base_row + (int)k);
            grp.push_back((int)(e - k));
            k = e;
        }
        for (size_t q = 0; q < idx.size(); q++) row.push_back(idx[q]);
        off.push_back((int)(grp.size() / 5));
        if (with_signals) {
            sig_off.push_back(sig.size());
            for (double v : sh.ssig) sig.push_back((T)v);
        }
    }
    if (with_signals) fwd_sig_h_ = sig;
    // inj_row holds signal rows relative to the shot; the kernel adds sig_off[slot]
    t.grp_off = off;
    t.grp_zx.clear();
    if (with_signals) fwd_grp_comp_.clear();
    for (size_t q = 0; q < grp.size(); q += 5) {
        t.grp_zx.push_back(grp[q]);
        t.grp_zx.push_back(grp[q + 1]);
        if (with_signals) fwd_grp_comp_.push_back(grp[q + 2]);
    }
    t.built_tz = -1;
    PSV_HIP(t.inj_grp.alloc(std::max<size_t>(grp.size(), 1)));
    PSV_HIP(t.inj_row.alloc(std::max<size_t>(row.size(), 1)));
    if (!grp.empty()) PSV_HIP(hipMemcpy(t.inj_grp.p, grp.data(), grp.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!row.empty()) PSV_HIP(hipMemcpy(t.inj_row.p, row.data(), row.size() * sizeof(int), hipMemcpyHostToDevice));
    if (with_signals) {
        PSV_HIP(t.sig.alloc(std::max<size_t>(sig.size(), 1)));
        if (!sig.empty()) PSV_HIP(hipMemcpy(t.sig.p, sig.data(), sig.size() * sizeof(T), hipMemcpyHostToDevice));
        PSV_HIP(t.sig_off.alloc(sig_off.size()));
        PSV_HIP(hipMemcpy(t.sig_off.p, sig_off.data(), sig_off.size() * sizeof(size_t), hipMemcpyHostToDevice));
    }
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::build_receivers(Tables& t, bool none)
{
    std::vector<int> off(1, 0), zxc;
    t.rec_zx.clear();
    for (int s = 0; s < nslots_; s++) {
        const ShotHost& sh = shots_[(size_t)shot_of_slot_[(size_t)s]];
        if (!none)
            for (size_t r = 0; r < sh.rx.size(); r++) {
                zxc.push_back(sh.rz[r]);
                zxc.push_back(sh.rx[r]);
                zxc.push_back(sh.rc[r]);
                t.rec_zx.push_back(sh.rz[r]);
                t.rec_zx.push_back(sh.rx[r]);
            }
        off.push_back((int)(zxc.size() / 3));
    }
    t.rec_off = off;
    t.built_tz = -1;
    PSV_HIP(t.rec_first.alloc(off.size()));
    PSV_HIP(t.rec_zxc.alloc(std::max<size_t>(zxc.size(), 1)));
    PSV_HIP(hipMemcpy(t.rec_first.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!zxc.empty()) PSV_HIP(hipMemcpy(t.rec_zxc.p, zxc.data(), zxc.size() * sizeof(int), hipMemcpyHostToDevice));
    return PSVFWI_OK;
}

// CSR lists of the injection groups and receivers of every (slot, tile)
// for the current tile height, so a workgroup touches only its own
template <typename T>
psvfwi_status Plan<T>::ensure_tile_lists(Tables& t)
{
    const int tzh = tz_abs();
    if (t.built_tz == tzh) return PSVFWI_OK;
    const int tiles_x = (g_.nxt + psv::kTX - 1) / psv::kTX;
    const int tiles_z = (g_.nzt + tzh - 1) / tzh;
    const size_t ntile = (size_t)tiles_x * tiles_z;
    auto build = [&](const std::vector<int>& soff, const std::vector<int>& zx, DBuf<int>& doff,
                     DBuf<int>& didx) -> psvfwi_status {
        std::vector<int> cnt((size_t)nslots_ * ntile + 1, 0);
        std::vector<std::pair<size_t, int>> items;
        for (int s = 0; s < nslots_; s++)
            for (int k = soff[(size_t)s]; k < soff[(size_t)s + 1]; k++) {
                const int z = zx[2 * (size_t)k], x = zx[2 * (size_t)k + 1];
                const size_t tix = (size_t)s * ntile + (size_t)(z / tzh) * tiles_x + (size_t)(x / psv::kTX);
                items.emplace_back(tix, k);
            }
        std::stable_sort(items.begin(), items.end(),
                         [](const std::pair<size_t, int>& a, const std::pair<size_t, int>& b) { return a.first < b.first; });
        for (const auto& it : items) cnt[it.first + 1]++;
        for (size_t i = 1; i < cnt.size(); i++) cnt[i] += cnt[i - 1];
        std::vector<int> idx;
        for (const auto& it : items) idx.push_back(it.second);
        PSV_HIP(doff.alloc(cnt.size()));
        PSV_HIP(didx.alloc(std::max<size_t>(idx.size(), 1)));
        PSV_HIP(hipMemcpy(doff.p, cnt.data(), cnt.size() * sizeof(int), hipMemcpyHostToDevice));
        if (!idx.empty()) PSV_HIP(hipMemcpy(didx.p, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
        return PSVFWI_OK;
    };
    PSV_TRY(build(t.grp_off, t.grp_zx, t.tinj_off, t.tinj_idx));
    PSV_TRY(build(t.rec_off, t.rec_zx, t.trec_off, t.trec_idx));
    t.built_tz = tzh;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::ensure_all_tile_lists()
{
    PSV_TRY(ensure_tile_lists(fwd_tab_));
    PSV_TRY(ensure_tile_lists(adj_tab_));
    PSV_TRY(ensure_tile_lists(fwd_tab_noreceivers_));
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::upload_tables()
{
    PSV_TRY(build_injection(fwd_tab_, false, true));
    PSV_TRY(build_injection(adj_tab_, true, false));
    PSV_TRY(build_injection(fwd_tab_noreceivers_, false, false));
    PSV_TRY(build_receivers(fwd_tab_, false));
    PSV_TRY(build_receivers(adj_tab_, true));
    PSV_TRY(build_receivers(fwd_tab_noreceivers_, true));
    {
        // (group, slot) pairs of the stress (explosive) sources
        std::vector<int> ex;
        for (int s = 0; s < nslots_; s++)
            for (int k = fwd_tab_.grp_off[(size_t)s]; k < fwd_tab_.grp_off[(size_t)s + 1]; k++)
                if (fwd_grp_comp_[(size_t)k] == 4) {
                    ex.push_back(k);
                    ex.push_back(s);
                }
        nexpl_ = (int)(ex.size() / 2);
        PSV_HIP(expl_grp_.alloc(std::max<size_t>(ex.size(), 2)));
        if (!ex.empty()) PSV_HIP(hipMemcpy(expl_grp_.p, ex.data(), ex.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    PSV_HIP(shot_nt_.alloc((size_t)nslots_));
    PSV_HIP(hipMemcpy(shot_nt_.p, nt_slot_.data(), (size_t)nslots_ * sizeof(int), hipMemcpyHostToDevice));
    // receiver traces and adjoint sources share one layout: slot s at rec_out_off[s]
    rec_out_off_h_.assign((size_t)nslots_, 0);
    std::vector<size_t> len((size_t)nslots_);
    size_t tot = 0;
    for (int s = 0; s < nslots_; s++) {
        const ShotHost& sh = shots_[(size_t)shot_of_slot_[(size_t)s]];
        rec_out_off_h_[(size_t)s] = tot;
        len[(size_t)s] = sh.rx.size() * (size_t)sh.nt;
        tot += len[(size_t)s];
    }
    PSV_HIP(rec_out_off_.alloc((size_t)nslots_));
    PSV_HIP(rec_len_.alloc((size_t)nslots_));
    PSV_HIP(hipMemcpy(rec_out_off_.p, rec_out_off_h_.data(), (size_t)nslots_ * sizeof(size_t), hipMemcpyHostToDevice));
    PSV_HIP(hipMemcpy(rec_len_.p, len.data(), (size_t)nslots_ * sizeof(size_t), hipMemcpyHostToDevice));
    PSV_HIP(rec_.alloc(std::max<size_t>(tot, 1)));
    PSV_HIP(adj_sig_.alloc(std::max<size_t>(tot, 1)));
    PSV_HIP(psv::launch_fill<T>(adj_sig_.p, adj_sig_.n, (T)0, stream_));
    PSV_HIP(hipStreamSynchronize(stream_));
    return PSVFWI_OK;
}

#define PSV_INSTANTIATE(T) \
    template psvfwi_status Plan<T>::set_shots(int, const psvfwi_shot_desc*); \
    template psvfwi_status Plan<T>::build_injection(Tables&, bool, bool); \
    template psvfwi_status Plan<T>::build_receivers(Tables&, bool); \
    template psvfwi_status Plan<T>::ensure_tile_lists(Tables&); \
    template psvfwi_status Plan<T>::ensure_all_tile_lists(); \
    template psvfwi_status Plan<T>::upload_tables();
PSV_FOR_EACH_PRECISION(PSV_INSTANTIATE)

}  // namespace psvplan
