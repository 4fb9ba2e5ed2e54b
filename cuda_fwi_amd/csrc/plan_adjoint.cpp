// plan_adjoint.cpp -- the adjoint time loop (adjoint_kernel_PSV,
// CUDA_week2/adjoint_kernel_PSV_GPU.cu:688-1102) over the decimated store, over
// checkpoints with segment replay, or with random-boundary reconstruction.
#include "plan.h"

namespace psvplan {

template <typename T>
psvfwi_status Plan<T>::set_adjoint_sources(int shot, const double* sig)
{
    if (shot < 0 || shot >= nslots_ || !sig) return fail(PSVFWI_E_ARG, "set_adjoint_sources: bad shot");
    const int s = slot_of_shot_[(size_t)shot];
    const size_t len = shots_[(size_t)shot].rx.size() * (size_t)shots_[(size_t)shot].nt;
    std::vector<T> h(len);
    for (size_t i = 0; i < len; i++) h[i] = (T)sig[i];
    if (len)
        PSV_HIP(hipMemcpyAsync(adj_sig_.p + rec_out_off_h_[(size_t)s], h.data(), len * sizeof(T),
                               hipMemcpyHostToDevice, stream_));
    PSV_HIP(hipStreamSynchronize(stream_));
    adj_ready_[(size_t)shot] = 1;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::set_store(int shot, const double* st)
{
    if (shot < 0 || shot >= nslots_ || !st) return fail(PSVFWI_E_ARG, "set_store: bad shot");
    if (!win_on_) return fail(PSVFWI_E_STATE, "set_store: no window");
    if (store_mode_ != PSVFWI_STORE_FULL) return fail(PSVFWI_E_STATE, "set_store: needs PSVFWI_STORE_FULL");
    const int ntf = 1 + (ntmax_ - 1) / w_.dt;
    if (store_.n != (size_t)nslots_ * ntf * 5 * nfz_ * nfx_) {
        PSV_TRY(alloc_store(ntf));
        PSV_HIP(psv::launch_fill<T>(store_.p, store_.n, (T)0, stream_));
    }
    const int s = slot_of_shot_[(size_t)shot];
    const int nft = 1 + (shots_[(size_t)shot].nt - 1) / w_.dt;
    const size_t plane = (size_t)nfz_ * nfx_;
    std::vector<T> h(sst_, (T)0);
    for (int f = 0; f < 5; f++)
        for (int t = 0; t < nft; t++)
            for (size_t k = 0; k < plane; k++)
                h[((size_t)t * 5 + f) * plane + k] = (T)st[((size_t)f * nft + t) * plane + k];
    PSV_HIP(hipMemcpyAsync(store_.p + (size_t)s * sst_, h.data(), sst_ * sizeof(T), hipMemcpyHostToDevice, stream_));
    PSV_HIP(hipStreamSynchronize(stream_));
    store_loaded_ = true;
    fwd_grad_done_ = true;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::adjoint(int flags)
{
    PSV_TRY(ready());
    const bool energy = flags & PSVFWI_ENERGY;
    const bool grad = (flags & PSVFWI_GRADIENT) != 0;
    if (grad && grad_mode_ == PSVFWI_GRADIENT_DISCRETE) return adjoint_discrete();
    if (grad && store_mode_ == PSVFWI_STORE_RANDOM_BOUNDARY) return adjoint_rb(flags);
    if (grad && !win_on_) return fail(PSVFWI_E_STATE, "adjoint: PSVFWI_GRADIENT needs a window");
    if (grad && fwd_disc_)
        return fail(PSVFWI_E_STATE, "adjoint: the forward ran for the discrete gradient; run it again");
    if (grad && !fwd_grad_done_)
        return fail(PSVFWI_E_STATE, "adjoint: run forward with PSVFWI_GRADIENT (or set_store) first");
    for (int i = 0; i < nslots_; i++)
        if (!adj_ready_[(size_t)i] && !shots_[(size_t)i].rx.empty())
            return fail(PSVFWI_E_STATE, "adjoint: no adjoint sources (call misfit or set_adjoint_sources)");
    const bool ckpt = grad && store_mode_ == PSVFWI_STORE_CHECKPOINT;
    if (grad && !ckpt && !store_loaded_) return fail(PSVFWI_E_STATE, "adjoint: forward store missing");

    PSV_TRY(alloc_state(main_));
    PSV_TRY(zero_state(main_));
    if (fused()) {
        PSV_TRY(alloc_state(main2_));
        PSV_TRY(zero_state(main2_));
    }
    PSV_HIP(We_.alloc((size_t)nslots_ * fst_));
    PSV_HIP(psv::launch_fill<T>(We_.p, We_.n, (T)0, stream_));
    PSV_HIP(ux_.alloc((size_t)nslots_ * fst_));
    PSV_HIP(uz_.alloc((size_t)nslots_ * fst_));
    PSV_HIP(psv::launch_fill<T>(ux_.p, ux_.n, (T)0, stream_));
    PSV_HIP(psv::launch_fill<T>(uz_.p, uz_.n, (T)0, stream_));
    if (grad) {
        gst_ = (size_t)3 * nfz_ * nfx_;
        PSV_HIP(grad_.alloc((size_t)nslots_ * gst_));
        PSV_HIP(psv::launch_fill<T>(grad_.p, grad_.n, (T)0, stream_));
    }
    const int aflags = (energy ? psv::kEnergy : 0) | (grad ? psv::kStore : 0);
    PSV_TRY(ensure_all_tile_lists());

    PSV_TRY(begin_samples());
    PSV_HIP(hipEventRecord(ev0_, stream_));
    launches_ = 0;
    std::vector<hipGraphExec_t> graphs;
    AllSlotsAfter restore{this};
    if (ckpt) {
        PSV_TRY(alloc_state(replay_));
        if (fused()) {
            PSV_TRY(alloc_state(replay2_));
            PSV_TRY(zero_state(replay2_));
        }
        PSV_TRY(alloc_store((ck_used_k_ + w_.dt - 1) / w_.dt + 1));
    }
    for (int w0 = 0; w0 < nslots_; w0 += group_size()) {
        set_group(w0);
        const int ntg = group_nt();
        if (!ckpt) {
            for (int it = ntg - 1; it >= 0; it--) {
                if (graph_steps_ > 0) {
                    const int na = active(it);
                    int end = it;  // steps it, it-1, ..., end+1 have the same active shots
                    while (end >= 0 && active(end) == na) end--;
                    if (it - end >= graph_steps_) {
                        psv::StepParams<T> p = base_params(main_, adj_tab_, false);
                        p.nslots = na;
                        p.flags = aflags;
                        p.tf0 = 0;
                        hipGraphExec_t ex = nullptr;
                        PSV_TRY(capture_steps(p, psv::kAdjoint, -1, main_, main2_, adj_tab_, false, &ex));
                        graphs.push_back(ex);
                        for (; it - end >= graph_steps_; it -= graph_steps_) {
                            PSV_HIP(psv::launch_fill<int>(it_dev_.p, 1, it, stream_));
                            PSV_HIP(hipGraphLaunch(ex, stream_));
                            launches_ += graph_steps_;
                        }
                        if (it < 0) break;
                    }
                }
                PSV_TRY(adjoint_step(it, aflags, 0));
            }
            continue;
        }
        const int kk = ck_used_k_;
        const int nseg = (ntg + kk - 1) / kk;
        for (int c = nseg - 1; c >= 0; c--) {
            const int t0 = c * kk, t1 = std::min(t0 + kk, ntg);
            const int tf0 = (t0 + w_.dt - 1) / w_.dt;
            // replay the segment forward from its checkpoint into the segment store
            PSV_TRY(restore_checkpoint(c, active(t0)));
            for (int it = t0; it < t1; it++) {
                psv::StepParams<T> p = base_params(replay_, fwd_tab_noreceivers_, true);
                p.it = it;
                p.nslots = active(it);
                p.flags = psv::kStore;
                p.tf0 = tf0;
                p.We = nullptr;
                PSV_TRY(launch_timed(p, psv::kForward, 2, replay_, replay2_));
            }
            for (int it = t1 - 1; it >= t0; it--) PSV_TRY(adjoint_step(it, aflags, tf0));
        }
    }
    all_slots();
    PSV_HIP(hipEventRecord(ev1_, stream_));
    PSV_HIP(hipStreamSynchronize(stream_));
    for (hipGraphExec_t ex : graphs) PSV_HIP(hipGraphExecDestroy(ex));
    float ms = 0;
    PSV_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    last_ms_ = ms;
    PSV_TRY(collect_samples());
    ran_ = true;
    adj_grad_done_ = grad;
    parity_ = false;  // every shot ends at it = 0
    return PSVFWI_OK;
}

// The conditions under which the step can be run backward: no CPML (its
// damping is not reversible; the random boundary layer of the model takes
// its place), the split kernels (they update in place), and no order-2
// free surface (the reference's override sets sxx from the velocities
// alone, so the previous sxx cannot be recovered).
template <typename T>
psvfwi_status Plan<T>::rb_check() const
{
    if (g_.npml != 0)
        return fail(PSVFWI_E_ARG, "random-boundary reconstruction needs npml = 0 (a random boundary "
                                  "layer in the model instead of CPML)");
    if (fused()) return fail(PSVFWI_E_ARG, "random-boundary reconstruction needs the split step (tile > 0)");
    if (g_.fsurf && g_.fdorder == 2)
        return fail(PSVFWI_E_ARG, "random-boundary reconstruction: the order-2 free surface is not reversible");
    return PSVFWI_OK;
}

// Adjoint with the forward state rebuilt backward from the forward's
// final state: per step it = ntmax-1 .. 0, the reverse step u(it+1) ->
// u(it), the store plane of u(it) when it is a correlation step, then the
// adjoint step.
template <typename T>
psvfwi_status Plan<T>::adjoint_rb(int flags)
{
    if (!fwd_grad_done_ || !fwd_rb_)
        return fail(PSVFWI_E_STATE, "adjoint: run forward with PSVFWI_GRADIENT in the random-boundary store "
                                    "mode first");
    if (!win_on_) return fail(PSVFWI_E_STATE, "adjoint: PSVFWI_GRADIENT needs a window");
    PSV_TRY(rb_check());
    for (int i = 0; i < nslots_; i++)
        if (!adj_ready_[(size_t)i] && !shots_[(size_t)i].rx.empty())
            return fail(PSVFWI_E_STATE, "adjoint: no adjoint sources (call misfit or set_adjoint_sources)");
    const bool energy = flags & PSVFWI_ENERGY;
    // the reconstruction starts from the forward's final state
    PSV_TRY(alloc_state(replay_));
    {
        psv::CopyList l{};
        for (int f = 0; f < 5; f++) {
            l.src[l.n] = main_.f[f].p;
            l.dst[l.n] = replay_.f[f].p;
            l.bytes[l.n] = (size_t)nslots_ * fst_ * sizeof(T);
            l.n++;
        }
        PSV_HIP(psv::launch_copies(l, stream_));
    }
    // negated source signals for the reverse step
    PSV_HIP(neg_sig_.alloc(std::max<size_t>(fwd_sig_h_.size(), 1)));
    if (!fwd_sig_h_.empty()) {
        std::vector<T> neg(fwd_sig_h_.size());
        for (size_t i = 0; i < neg.size(); i++) neg[i] = -fwd_sig_h_[i];
        PSV_HIP(hipMemcpy(neg_sig_.p, neg.data(), neg.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    PSV_TRY(zero_state(main_));
    PSV_HIP(We_.alloc((size_t)nslots_ * fst_));
    PSV_HIP(psv::launch_fill<T>(We_.p, We_.n, (T)0, stream_));
    PSV_HIP(ux_.alloc((size_t)nslots_ * fst_));
    PSV_HIP(uz_.alloc((size_t)nslots_ * fst_));
    PSV_HIP(psv::launch_fill<T>(ux_.p, ux_.n, (T)0, stream_));
    PSV_HIP(psv::launch_fill<T>(uz_.p, uz_.n, (T)0, stream_));
    gst_ = (size_t)3 * nfz_ * nfx_;
    PSV_HIP(grad_.alloc((size_t)nslots_ * gst_));
    PSV_HIP(psv::launch_fill<T>(grad_.p, grad_.n, (T)0, stream_));
    PSV_TRY(alloc_store(1));
    PSV_TRY(ensure_all_tile_lists());
    const int aflags = (energy ? psv::kEnergy : 0) | psv::kStore;

    PSV_TRY(begin_samples());
    PSV_HIP(hipEventRecord(ev0_, stream_));
    launches_ = 0;
    AllSlotsAfter restore{this};
    for (int w0 = 0; w0 < nslots_; w0 += group_size()) {
        set_group(w0);
        for (int it = group_nt() - 1; it >= 0; it--) {
            const int na = active(it);
            psv::StepParams<T> r = base_params(replay_, fwd_tab_noreceivers_, true);
            r.it = it;
            r.nslots = na;
            r.flags = 0;
            r.We = nullptr;
            r.dt = -r.dt;
            r.sig = neg_sig_.p;
            PSV_HIP(psv::launch_stress_sources<T>(r, stream_));
            PSV_HIP(psv::launch_half<T>(r, g_.fdorder, psv::kForward, tz_, 1, stream_));
            PSV_HIP(psv::launch_half<T>(r, g_.fdorder, psv::kForward, tz_, 0, stream_));
            if (it % w_.dt == 0) {
                psv::StepParams<T> q = base_params(replay_, fwd_tab_noreceivers_, true);
                q.nslots = na;
                PSV_HIP(psv::launch_store_plane<T>(q, stream_));
            }
            PSV_TRY(adjoint_step(it, aflags, it / w_.dt));
        }
    }
    all_slots();
    PSV_HIP(hipEventRecord(ev1_, stream_));
    PSV_HIP(hipStreamSynchronize(stream_));
    float ms = 0;
    PSV_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    last_ms_ = ms;
    PSV_TRY(collect_samples());
    ran_ = true;
    adj_grad_done_ = true;
    fwd_grad_done_ = false;  // the forward's final state has been consumed
    parity_ = false;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::adjoint_step(int it, int flags, int tf0)
{
    psv::StepParams<T> p = base_params(main_, adj_tab_, false);
    p.it = it;
    p.nslots = active(it);
    p.flags = flags;
    p.tf0 = tf0;
    PSV_TRY(launch_timed(p, psv::kAdjoint, 1, main_, main2_));
    return PSVFWI_OK;
}

#define PSV_INSTANTIATE(T) \
    template psvfwi_status Plan<T>::set_adjoint_sources(int, const double*); \
    template psvfwi_status Plan<T>::set_store(int, const double*); \
    template psvfwi_status Plan<T>::adjoint(int); \
    template psvfwi_status Plan<T>::rb_check() const; \
    template psvfwi_status Plan<T>::adjoint_rb(int); \
    template psvfwi_status Plan<T>::adjoint_step(int, int, int);
PSV_FOR_EACH_PRECISION(PSV_INSTANTIATE)

}  // namespace psvplan
