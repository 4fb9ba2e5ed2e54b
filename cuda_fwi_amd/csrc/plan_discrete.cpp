// plan_discrete.cpp -- the discrete (exact) gradient: checkpointed segments replayed
// with every state kept, then the transposed step run back through them.
#include "plan.h"

namespace psvplan {

// Discrete gradient: per segment (last first), restore its checkpoint,
// replay it keeping the state before every step and after the last, then
// run the transposed steps back through it.  With segments of one step the
// forward kept every state, and the transposed steps read those directly.
template <typename T>
psvfwi_status Plan<T>::adjoint_discrete()
{
    if (!fwd_grad_done_ || !fwd_disc_)
        return fail(PSVFWI_E_STATE, "adjoint: run forward with PSVFWI_GRADIENT in the discrete gradient mode first");
    for (int i = 0; i < nslots_; i++)
        if (!adj_ready_[(size_t)i] && !shots_[(size_t)i].rx.empty())
            return fail(PSVFWI_E_STATE, "adjoint: no adjoint sources (call misfit or set_adjoint_sources)");
    if (nexpl_ > 0) return fail(PSVFWI_E_ARG, "adjoint: the discrete gradient takes vx / vz point sources only");
    if (g_.npml > 0 && (h_pml_[(size_t)g_.npml] != 0.0 || h_pml_[(size_t)3 * tp_ + g_.npml] != 0.0))
        return fail(PSVFWI_E_ARG, "adjoint: the discrete gradient needs a CPML profile with a = a_half = 0 at "
                                  "index npml, as psvfwi_cpml_profiles makes");
    const int kk = ck_used_k_;
    PSV_TRY(alloc_state(adj_));
    PSV_TRY(zero_state(adj_));
    const size_t nf = (size_t)nslots_ * fst_;
    for (int i = 0; i < 8; i++) {
        PSV_HIP(dw_[i].alloc(nf));
        PSV_HIP(psv::launch_fill<T>(dw_[i].p, nf, (T)0, stream_));
    }
    for (int i = 0; i < 5; i++) {
        PSV_HIP(dg_[i].alloc(nf));
        PSV_HIP(psv::launch_fill<T>(dg_[i].p, nf, (T)0, stream_));
    }
    // segments of one step: the forward kept every state, no replay
    const bool every = kk == 1;
    if (!every) {
        PSV_TRY(alloc_state(replay_));
        if (fused()) {
            PSV_TRY(alloc_state(replay2_));
            PSV_TRY(zero_state(replay2_));
        }
        PSV_HIP(seg_.reserve((size_t)(std::min(kk, ntmax_) + 1) * state_elems()));
    }
    // receiver groups of each slot, for the adjoint sources
    const std::vector<int>& go = adj_tab_.grp_off;
    int maxg = 0;
    for (int sl = 0; sl < nslots_; sl++) maxg = std::max(maxg, go[(size_t)sl + 1] - go[(size_t)sl]);
    PSV_HIP(dgrp_off_.alloc(go.size()));
    PSV_HIP(hipMemcpy(dgrp_off_.p, go.data(), go.size() * sizeof(int), hipMemcpyHostToDevice));
    // the receiver groups of each (slot, 64 x 4 workgroup of the reverse-step kernels)
    {
        const int bx = (g_.nxt + psv::kDiscreteBX - 1) / psv::kDiscreteBX;
        const int bz = (g_.nzt + psv::kDiscreteBZ - 1) / psv::kDiscreteBZ;
        dnblk_ = bx * bz;
        std::vector<int> cnt((size_t)nslots_ * dnblk_ + 1, 0);
        std::vector<std::pair<size_t, int>> items;
        for (int sl = 0; sl < nslots_; sl++)
            for (int g = go[(size_t)sl]; g < go[(size_t)sl + 1]; g++) {
                const int z = adj_tab_.grp_zx[2 * (size_t)g], x = adj_tab_.grp_zx[2 * (size_t)g + 1];
                const size_t tix = (size_t)sl * dnblk_ + (size_t)(z / psv::kDiscreteBZ) * bx +
                                   (size_t)(x / psv::kDiscreteBX);
                items.emplace_back(tix, g);
            }
        std::stable_sort(items.begin(), items.end(),
                         [](const std::pair<size_t, int>& a, const std::pair<size_t, int>& b) { return a.first < b.first; });
        for (const auto& q : items) cnt[q.first + 1]++;
        for (size_t i = 1; i < cnt.size(); i++) cnt[i] += cnt[i - 1];
        std::vector<int> idx;
        for (const auto& q : items) idx.push_back(q.second);
        PSV_HIP(dblk_off_.alloc(cnt.size()));
        PSV_HIP(dblk_idx_.alloc(std::max<size_t>(idx.size(), 1)));
        PSV_HIP(hipMemcpy(dblk_off_.p, cnt.data(), cnt.size() * sizeof(int), hipMemcpyHostToDevice));
        if (!idx.empty())
            PSV_HIP(hipMemcpy(dblk_idx_.p, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
    }

    PSV_HIP(hipEventRecord(ev0_, stream_));
    launches_ = 0;
    AllSlotsAfter restore{this};
    for (int w0 = 0; w0 < nslots_; w0 += group_size()) {
        set_group(w0);
        const int ntg = group_nt();
        const int nseg = (ntg + kk - 1) / kk;
        for (int c = nseg - 1; c >= 0; c--) {
            const int t0 = c * kk, t1 = std::min(t0 + kk, ntg);
            // the states before steps t0 .. t1 - 1 and after t1 - 1
            const T* base = every ? ckpt_.p + (size_t)t0 * state_elems() : seg_.p;
            if (!every) {
                PSV_TRY(restore_checkpoint(c, active(t0)));
                PSV_TRY(state_copy<true>(seg_.p, active(t0), replay_));
                for (int it = t0; it < t1; it++) {
                    psv::StepParams<T> p = base_params(replay_, fwd_tab_noreceivers_, true);
                    p.it = it;
                    p.nslots = active(it);
                    p.flags = 0;
                    p.We = nullptr;
                    PSV_TRY(launch_plain(p, psv::kForward, replay_, replay2_));
                    PSV_TRY(state_copy<true>(seg_.p + (size_t)(it - t0 + 1) * state_elems(), active(it), replay_));
                    launches_++;
                }
            }
            for (int it = t1 - 1; it >= t0; it--) {
                psv::DiscreteParams<T> d = dparams(it, entry(base, it - t0), entry(base, it - t0 + 1));
                d.max_groups = maxg;
                // receivers: the stress pass of step it + 1 added the sources
                // of the shots active there; the rest start at step it
                d.inj_lo = active(it + 1);
                d.inj_hi = d.nslots;
                d.gather_prev = it + 1 < ntg;
                d.inj_next = it > 0 ? it - 1 : -1;
                d.nblk = dnblk_;
                d.blk_off = plus(dblk_off_.p, (size_t)w0_ * dnblk_);
                d.blk_idx = dblk_idx_.p;
                PSV_HIP(psv::launch_discrete_step<T>(d, g_.fdorder, stream_));
                launches_++;
            }
        }
    }
    all_slots();
    PSV_HIP(hipEventRecord(ev1_, stream_));
    PSV_HIP(hipStreamSynchronize(stream_));
    float ms = 0;
    PSV_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    last_ms_ = ms;
    adj_grad_done_ = false;
    disc_done_ = true;
    return PSVFWI_OK;
}

template <typename T>
typename Plan<T>::View Plan<T>::entry(const T* states, int e) const
{
    const T* base = states + (size_t)e * state_elems();
    View v{};
    size_t off = 0;
    const size_t w0 = (size_t)w0_;
    for (int f = 0; f < 5; f++) {
        v.f[f] = base + off + org_ + w0 * fst_;
        off += (size_t)nslots_ * fst_;
    }
    if (g_.npml > 0) {
        const size_t st[4] = {xst_, xst_, zst_, zst_};
        for (int m = 0; m < 4; m++) {
            v.ms[m] = base + off + w0 * st[m];
            off += (size_t)nslots_ * st[m];
        }
        for (int m = 0; m < 4; m++) {
            v.mv[m] = base + off + w0 * st[m];
            off += (size_t)nslots_ * st[m];
        }
    }
    return v;
}

template <typename T>
psv::DiscreteParams<T> Plan<T>::dparams(int it, const View& pre, const View& post)
{
    psv::DiscreteParams<T> d{};
    d.nzt = g_.nzt;
    d.nxt = g_.nxt;
    d.pitch = pitch_;
    d.ppad = ppad_;
    d.npml = g_.npml;
    d.tp = tp_;
    d.fsurf = g_.fsurf;
    d.pnx = pnx_;
    d.pnz = pnz_;
    d.nslots = active(it);
    d.it = it;
    d.dt = (T)g_.dt;
    d.dxi = (T)1.0 / (T)g_.dx;
    d.dzi = (T)1.0 / (T)g_.dz;
    d.hc1 = (T)g_.hc[1];
    d.hc2 = (T)g_.hc[2];
    d.lam = at(lam_);
    d.mu = at(mu_);
    d.mu_zx = at(mu_zx_);
    d.rho_zp = at(rho_zp_);
    d.rho_xp = at(rho_xp_);
    d.pml = pml_.p;
    d.fst = fst_;
    d.xst = xst_;
    d.zst = zst_;
    d.pre_vx = pre.f[kF_VX];
    d.pre_vz = pre.f[kF_VZ];
    d.pre_mvxx = pre.ms[0];
    d.pre_mvzx = pre.ms[1];
    d.pre_mvxz = pre.ms[2];
    d.pre_mvzz = pre.ms[3];
    d.pre_msxxx = pre.mv[0];
    d.pre_mszxx = pre.mv[1];
    d.pre_mszzz = pre.mv[2];
    d.pre_mszxz = pre.mv[3];
    d.post_sxx = post.f[kF_SXX];
    d.post_szx = post.f[kF_SZX];
    d.post_szz = post.f[kF_SZZ];
    d.vbx = atw(adj_.f[kF_VX], fst_);
    d.vbz = atw(adj_.f[kF_VZ], fst_);
    d.sbxx = atw(adj_.f[kF_SXX], fst_);
    d.sbzx = atw(adj_.f[kF_SZX], fst_);
    d.sbzz = atw(adj_.f[kF_SZZ], fst_);
    d.bvxx = offw(adj_.ms[0], xst_);
    d.bvzx = offw(adj_.ms[1], xst_);
    d.bvxz = offw(adj_.ms[2], zst_);
    d.bvzz = offw(adj_.ms[3], zst_);
    d.bsxxx = offw(adj_.mv[0], xst_);
    d.bszxx = offw(adj_.mv[1], xst_);
    d.bszzz = offw(adj_.mv[2], zst_);
    d.bszxz = offw(adj_.mv[3], zst_);
    d.w0 = atw(dw_[0], fst_);
    d.w1 = atw(dw_[1], fst_);
    d.w2 = atw(dw_[2], fst_);
    d.w3 = atw(dw_[3], fst_);
    d.ws0 = atw(dw_[4], fst_);
    d.ws1 = atw(dw_[5], fst_);
    d.ws2 = atw(dw_[6], fst_);
    d.ws3 = atw(dw_[7], fst_);
    d.g_lam = atw(dg_[0], fst_);
    d.g_mu = atw(dg_[1], fst_);
    d.g_mu_zx = atw(dg_[2], fst_);
    d.g_rho_zp = atw(dg_[3], fst_);
    d.g_rho_xp = atw(dg_[4], fst_);
    d.grp = adj_tab_.inj_grp.p;
    d.grp_off = plus(dgrp_off_.p, w0_);
    d.grp_row = adj_tab_.inj_row.p;
    d.sig = adj_sig_.p;
    d.sig_off = plus(rec_out_off_.p, w0_);
    d.shot_nt = plus(shot_nt_.p, w0_);
    return d;
}

template <typename T>
psvfwi_status Plan<T>::get_coefficient_gradient(int shot, double* gl, double* gm, double* gmzx, double* grzp, double* grxp) const
{
    if (!disc_done_)
        return fail(PSVFWI_E_STATE, "get_coefficient_gradient: run adjoint with PSVFWI_GRADIENT in the "
                                    "discrete gradient mode first");
    if (shot < -1 || shot >= nslots_ || !gl || !gm || !gmzx || !grzp || !grxp)
        return fail(PSVFWI_E_ARG, "get_coefficient_gradient: bad shot or null array");
    const int nzt = g_.nzt, nxt = g_.nxt;
    double* out[5] = {gl, gm, gmzx, grzp, grxp};
    std::vector<double> h(fst_);
    for (int a = 0; a < 5; a++) {
        const bool avg = a >= 2;
        const int rows = avg ? nzt - 1 : nzt, cols = avg ? nxt - 1 : nxt;
        std::fill(out[a], out[a] + (size_t)rows * cols, 0.0);
        // FP64 sum over the shots in shot order
        const int s0 = shot < 0 ? 0 : shot, s1 = shot < 0 ? nslots_ : shot + 1;
        for (int i = s0; i < s1; i++) {
            const int sl = slot_of_shot_[(size_t)i];
            PSV_TRY(download(dg_[a].p + (size_t)sl * fst_, fst_, h.data()));
            for (int iz = 0; iz < rows; iz++)
                for (int ix = 0; ix < cols; ix++)
                    out[a][(size_t)iz * cols + ix] += h[(size_t)org_ + (size_t)iz * pitch_ + ix];
        }
    }
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::get_model_gradient(int shot, double* gl, double* gm, double* gr) const
{
    if (!gl || !gm || !gr) return fail(PSVFWI_E_ARG, "get_model_gradient: null array");
    const size_t n = (size_t)g_.nzt * g_.nxt, na = (size_t)(g_.nzt - 1) * (g_.nxt - 1);
    std::vector<double> c0(n), c1(n), c2(na), c3(na), c4(na);
    PSV_TRY(get_coefficient_gradient(shot, c0.data(), c1.data(), c2.data(), c3.data(), c4.data()));
    return psvfwi_chain_gradient(g_.nzt, g_.nxt, h_mu_.data(), h_mu_zx_.data(), h_rho_zp_.data(),
                                 h_rho_xp_.data(), c0.data(), c1.data(), c2.data(), c3.data(), c4.data(),
                                 gl, gm, gr);
}

#define PSV_INSTANTIATE(T) \
    template psvfwi_status Plan<T>::adjoint_discrete(); \
    template Plan<T>::View Plan<T>::entry(const T*, int) const; \
    template psv::DiscreteParams<T> Plan<T>::dparams(int, const View&, const View&); \
    template psvfwi_status Plan<T>::get_coefficient_gradient(int, double*, double*, double*, double*, double*) const; \
    template psvfwi_status Plan<T>::get_model_gradient(int, double*, double*, double*) const;
PSV_FOR_EACH_PRECISION(PSV_INSTANTIATE)

}  // namespace psvplan
