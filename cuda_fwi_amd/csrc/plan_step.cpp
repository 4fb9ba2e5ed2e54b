// plan_step.cpp -- one time step: kernel parameters of the current launch group, the
// split / fused / z-marching step launch, sampled kernel timing, captured step
// graphs and the autotuner.
#include "plan.h"

namespace psvplan {

// Times `steps` forward launches of the current batch for every tile
// height and keeps the fastest.  Runs on the plan's own state buffers
// (every forward/adjoint starts from a zeroed state, so nothing leaks).
template <typename T>
psvfwi_status Plan<T>::autotune(int steps, int* best)
{
    PSV_TRY(ready());
    if (steps < 1) return fail(PSVFWI_E_ARG, "autotune: steps must be >= 1");
    PSV_TRY(alloc_state(main_));
    PSV_TRY(alloc_state(main2_));
    PSV_HIP(We_.alloc((size_t)nslots_ * fst_));
    // with a window, the candidates are timed on forward steps and on
    // adjoint steps with the gradient correlation (one store plane), the
    // two kinds of step of a gradient; without one on forward steps
    const bool adj = win_on_;
    if (adj) {
        PSV_HIP(ux_.alloc((size_t)nslots_ * fst_));
        PSV_HIP(uz_.alloc((size_t)nslots_ * fst_));
        gst_ = (size_t)3 * nfz_ * nfx_;
        PSV_HIP(grad_.alloc((size_t)nslots_ * gst_));
        PSV_TRY(alloc_store(1));
    }
    auto time_steps = [&](float* out) -> psvfwi_status {
        PSV_TRY(ensure_all_tile_lists());
        AllSlotsAfter restore{this};
        float ms = 0;
        for (int rep = 0; rep < 2; rep++) {  // the first repetition warms up
            PSV_TRY(zero_state(main_));
            PSV_TRY(zero_state(main2_));
            PSV_HIP(hipEventRecord(ev0_, stream_));
            for (int w0 = 0; w0 < nslots_; w0 += group_size()) {
                set_group(w0);
                const int nstep = std::min(steps, group_nt());
                for (int it = 0; it < nstep; it++) {
                    psv::StepParams<T> p = base_params(main_, fwd_tab_noreceivers_, true);
                    p.it = it;
                    p.nslots = active(it);
                    p.flags = 0;
                    PSV_TRY(launch_plain(p, psv::kForward, main_, main2_));
                }
                for (int it = 0; adj && it < nstep; it++) {
                    psv::StepParams<T> p = base_params(main_, adj_tab_, false);
                    p.it = it;
                    p.nslots = active(it);
                    p.flags = psv::kStore;
                    p.tf0 = it / w_.dt;
                    PSV_TRY(launch_plain(p, psv::kAdjoint, main_, main2_));
                }
            }
            all_slots();
            PSV_HIP(hipEventRecord(ev1_, stream_));
            PSV_HIP(hipStreamSynchronize(stream_));
            PSV_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
        }
        *out = ms;
        return PSVFWI_OK;
    };
    const int cand[kNumCand] = {4, 8, 16, 32, -8, -16, -32, -64};
    float best_ms = 1e30f;
    int best_tz = tz_;
    for (int c = 0; c < kNumCand; c++) {
        tune_ms_[c] = -1;
        if (!tile_ok(cand[c])) continue;
        // the reverse steps of the random-boundary reconstruction need the split kernels
        if (cand[c] < 0 && store_mode_ == PSVFWI_STORE_RANDOM_BOUNDARY) continue;
        tz_ = cand[c];
        float ms = 0;
        PSV_TRY(time_steps(&ms));
        tune_ms_[c] = ms;
        if (ms < best_ms) {
            best_ms = ms;
            best_tz = tz_;
        }
    }
    tz_ = best_tz;
    // the shot-to-workgroup order of a batch: shot-major, or bands of 1 or
    // 4 tile rows with the shots of a band together
    if (nslots_ > 1) {
        int best_band = band_;
        const int bands[3] = {0, 1, 4};
        for (int b : bands) {
            band_ = b;
            float ms = 0;
            PSV_TRY(time_steps(&ms));
            if (b == 0 || ms < best_ms) {
                best_ms = ms;
                best_band = b;
            }
        }
        band_ = best_band;
        // shots per launch group (split step only): the whole batch, or
        // groups small enough for their wavefields to stay in the
        // Infinity Cache between steps
        if (!fused()) {
            int best_lg = 0;
            const int sizes[6] = {16, 12, 8, 6, 4, 2};
            lg_ = 0;
            PSV_TRY(time_steps(&best_ms));
            for (int k : sizes) {
                if (k >= nslots_) continue;
                lg_ = k;
                float ms = 0;
                PSV_TRY(time_steps(&ms));
                if (ms < best_ms) {
                    best_ms = ms;
                    best_lg = k;
                }
            }
            lg_ = best_lg;
        }
    }
    if (best) *best = best_tz;
    // the candidates ran in the plan's own state, store and gradient buffers
    drop_results();
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::timing(double* ms, long* launches) const
{
    if (ms) *ms = last_ms_;
    if (launches) *launches = launches_;
    return PSVFWI_OK;
}

// Captures graph_steps_ consecutive steps starting at the step held in
// it_dev_ (step k of the chunk is *it_dev_ + dir * k) into an executable
// graph.  An even count leaves the ping-pong buffers of the fused step in
// their original roles.  The chunk is captured on the plan's stream and
// holds only kernel launches.
template <typename T>
psvfwi_status Plan<T>::capture_steps(psv::StepParams<T> p0, int mode, int dir, State<T>& st, State<T>& alt, const Tables& tab, bool fwd_sig, hipGraphExec_t* out)
{
    hipGraph_t g = nullptr;
    PSV_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    psvfwi_status rc = PSVFWI_OK;
    for (int k = 0; k < graph_steps_ && rc == PSVFWI_OK; k++) {
        psv::StepParams<T> p = base_params(st, tab, fwd_sig);
        p.nslots = p0.nslots;
        p.flags = p0.flags;
        p.tf0 = p0.tf0;
        p.it = dir * k;
        p.it_dev = it_dev_.p;
        rc = launch_plain(p, mode, st, alt);
    }
    const hipError_t e = hipStreamEndCapture(stream_, &g);
    if (rc != PSVFWI_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    PSV_HIP(e);
    const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    PSV_HIP(ei);
    return PSVFWI_OK;
}

// one step on st: the split kernels update st in place; the fused kernel
// reads st, writes alt, and the two swap roles (st is current after it)
template <typename T>
psvfwi_status Plan<T>::launch_plain(psv::StepParams<T> p, int mode, State<T>& st, State<T>& alt, hipEvent_t e_mid)
{
    if (!fused()) {
        PSV_HIP(psv::launch_step<T>(p, g_.fdorder, mode, tz_, stream_, e_mid));
        return PSVFWI_OK;
    }
    p.vx_w = at(alt.f[kF_VX]);
    p.vz_w = at(alt.f[kF_VZ]);
    p.sxx_w = at(alt.f[kF_SXX]);
    p.szx_w = at(alt.f[kF_SZX]);
    p.szz_w = at(alt.f[kF_SZZ]);
    p.mvxx_w = alt.ms[0].p;
    p.mvzx_w = alt.ms[1].p;
    p.mvxz_w = alt.ms[2].p;
    p.mvzz_w = alt.ms[3].p;
    if (tz_abs() >= 32)
        PSV_HIP(psv::launch_march<T>(p, g_.fdorder, mode, tz_abs(), stream_, use_side_ ? &side_ : nullptr));
    else PSV_HIP(psv::launch_fused<T>(p, g_.fdorder, mode, tz_abs(), stream_));
    if (e_mid) PSV_HIP(hipEventRecord(e_mid, stream_));
    for (int f = 0; f < 5; f++) swap_buf(st.f[f], alt.f[f]);
    for (int m = 0; m < 4; m++) swap_buf(st.ms[m], alt.ms[m]);
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::launch_timed(const psv::StepParams<T>& p, int mode, int kind, State<T>& st, State<T>& alt)
{
    const bool sample = sample_ev_[0][0] && (launches_ % kSampleEvery == 0) && nsamp_ < kMaxSamples;
    if (sample) {
        PSV_HIP(hipEventRecord(sample_ev_[nsamp_][0], stream_));
        PSV_TRY(launch_plain(p, mode, st, alt, sample_ev_[nsamp_][1]));
        PSV_HIP(hipEventRecord(sample_ev_[nsamp_][2], stream_));
        samp_kind_[nsamp_] = kind;
        nsamp_++;
    } else {
        PSV_TRY(launch_plain(p, mode, st, alt));
    }
    launches_++;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::begin_samples()
{
    if (!sample_ev_[0][0])
        for (int i = 0; i < kMaxSamples; i++)
            for (int j = 0; j < 3; j++) PSV_HIP(hipEventCreate(&sample_ev_[i][j]));
    nsamp_ = 0;
    return PSVFWI_OK;
}

// after the stream has drained: mean microseconds of each kernel per kind
template <typename T>
psvfwi_status Plan<T>::collect_samples()
{
    for (int k = 0; k < 3; k++) {
        double a = 0, b = 0;
        int n = 0;
        for (int i = 0; i < nsamp_; i++) {
            if (samp_kind_[i] != k) continue;
            float m1 = 0, m2 = 0;
            PSV_HIP(hipEventElapsedTime(&m1, sample_ev_[i][0], sample_ev_[i][1]));
            PSV_HIP(hipEventElapsedTime(&m2, sample_ev_[i][1], sample_ev_[i][2]));
            a += m1;
            b += m2;
            n++;
        }
        kus_[k][0] = n ? a / n * 1e3 : 0;
        kus_[k][1] = n ? b / n * 1e3 : 0;
        kn_[k] = n;
    }
    return PSVFWI_OK;
}

template <typename T>
psv::StepParams<T> Plan<T>::base_params(State<T>& st, const Tables& tab, bool fwd_sig)
{
    psv::StepParams<T> p{};
    p.nzt = g_.nzt;
    p.nxt = g_.nxt;
    p.pitch = pitch_;
    p.ppad = ppad_;
    p.npml = g_.npml;
    p.tp = tp_;
    p.fsurf = g_.fsurf;
    p.pnx = pnx_;
    p.pnz = pnz_;
    p.tiles_x = (g_.nxt + psv::kTX - 1) / psv::kTX;
    p.tiles_z = (g_.nzt + tz_abs() - 1) / tz_abs();
    p.band = band_;
    p.org = org_;
    p.it_dev = nullptr;
    p.dt = (T)g_.dt;
    p.dxi = (T)1.0 / (T)g_.dx;
    p.dzi = (T)1.0 / (T)g_.dz;
    p.hc1 = (T)g_.hc[1];
    p.hc2 = (T)g_.hc[2];
    p.lam = at(lam_);
    p.mu = at(mu_);
    p.mu_zx = at(mu_zx_);
    p.rho_zp = at(rho_zp_);
    p.rho_xp = at(rho_xp_);
    p.pml = pml_.p;
    p.vx = atw(st.f[kF_VX], fst_);
    p.vz = atw(st.f[kF_VZ], fst_);
    p.sxx = atw(st.f[kF_SXX], fst_);
    p.szx = atw(st.f[kF_SZX], fst_);
    p.szz = atw(st.f[kF_SZZ], fst_);
    p.vx_w = p.vx;
    p.vz_w = p.vz;
    p.sxx_w = p.sxx;
    p.szx_w = p.szx;
    p.szz_w = p.szz;
    p.fst = fst_;
    p.We = atw(We_, fst_);
    p.ux = atw(ux_, fst_);
    p.uz = atw(uz_, fst_);
    p.mvxx = offw(st.ms[0], xst_);
    p.mvzx = offw(st.ms[1], xst_);
    p.mvxz = offw(st.ms[2], zst_);
    p.mvzz = offw(st.ms[3], zst_);
    p.mvxx_w = p.mvxx;
    p.mvzx_w = p.mvzx;
    p.mvxz_w = p.mvxz;
    p.mvzz_w = p.mvzz;
    p.msxxx = offw(st.mv[0], xst_);
    p.mszxx = offw(st.mv[1], xst_);
    p.mszzz = offw(st.mv[2], zst_);
    p.mszxz = offw(st.mv[3], zst_);
    p.xst = xst_;
    p.zst = zst_;
    p.store = offw(store_, sst_);
    p.sst = sst_;
    if (win_on_) {
        p.fwi_dt = w_.dt;
        p.fwi_dx = w_.dx;
        p.fwi_dz = w_.dz;
        p.fwi_x1 = w_.x1;
        p.fwi_x2 = w_.x2;
        p.fwi_z1 = w_.z1;
        p.fwi_z2 = w_.z2;
        p.nfz = nfz_;
        p.nfx = nfx_;
    } else {
        p.fwi_dt = p.fwi_dx = p.fwi_dz = 1;
    }
    p.grad = offw(grad_, gst_);
    p.gst = gst_;
    p.shot_nt = plus(shot_nt_.p, w0_);
    p.inj_grp = tab.inj_grp.p;
    p.inj_row = tab.inj_row.p;
    const size_t ntile = (size_t)p.tiles_x * p.tiles_z;
    p.tinj_off = plus(tab.tinj_off.p, (size_t)w0_ * ntile);
    p.tinj_idx = tab.tinj_idx.p;
    p.trec_off = plus(tab.trec_off.p, (size_t)w0_ * ntile);
    p.trec_idx = tab.trec_idx.p;
    if (fwd_sig) {
        p.sig = fwd_tab_.sig.p;
        p.sig_off = plus(fwd_tab_.sig_off.p, w0_);
    } else {
        p.sig = adj_sig_.p;
        p.sig_off = plus(rec_out_off_.p, w0_);
    }
    p.rec_first = plus(tab.rec_first.p, w0_);
    p.rec_zxc = tab.rec_zxc.p;
    p.rec = rec_.p;
    p.rec_out_off = plus(rec_out_off_.p, w0_);
    p.nexpl = fwd_sig ? nexpl_ : 0;
    p.expl_grp = expl_grp_.p;  // (group, slot) pairs of the whole batch
    p.slot0 = w0_;
    return p;
}

template <typename T>
psvfwi_status Plan<T>::kernel_timing(int kind, double* us_stress, double* us_velocity, int* samples) const
{
    if (kind < 0 || kind > 2) return fail(PSVFWI_E_ARG, "kernel_timing: kind must be 0, 1 or 2");
    if (us_stress) *us_stress = kus_[kind][0];
    if (us_velocity) *us_velocity = kus_[kind][1];
    if (samples) *samples = kn_[kind];
    return PSVFWI_OK;
}

#define PSV_INSTANTIATE(T) \
    template psvfwi_status Plan<T>::autotune(int, int*); \
    template psvfwi_status Plan<T>::timing(double*, long*) const; \
    template psvfwi_status Plan<T>::capture_steps(psv::StepParams<T>, int, int, State<T>&, State<T>&, const Tables&, bool, hipGraphExec_t*); \
    template psvfwi_status Plan<T>::launch_plain(psv::StepParams<T>, int, State<T>&, State<T>&, hipEvent_t); \
    template psvfwi_status Plan<T>::launch_timed(const psv::StepParams<T>&, int, int, State<T>&, State<T>&); \
    template psvfwi_status Plan<T>::begin_samples(); \
    template psvfwi_status Plan<T>::collect_samples(); \
    template psv::StepParams<T> Plan<T>::base_params(State<T>&, const Tables&, bool); \
    template psvfwi_status Plan<T>::kernel_timing(int, double*, double*, int*) const;
PSV_FOR_EACH_PRECISION(PSV_INSTANTIATE)

}  // namespace psvplan
