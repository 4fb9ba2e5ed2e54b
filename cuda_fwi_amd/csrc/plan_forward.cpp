// plan_forward.cpp -- the forward time loop (forward_kernel_PSV,
// CUDA_week2/forward_kernel_PSV_GPU.cu:681-1084) and the FP64 misfit.
#include "plan.h"

namespace psvplan {

template <typename T>
psvfwi_status Plan<T>::forward(int flags)
{
    PSV_TRY(ready());
    const bool energy = flags & PSVFWI_ENERGY;
    const bool grad = (flags & PSVFWI_GRADIENT) != 0;
    const bool discrete = grad && grad_mode_ == PSVFWI_GRADIENT_DISCRETE;
    if (grad && !discrete && !win_on_) return fail(PSVFWI_E_STATE, "forward: PSVFWI_GRADIENT needs a window");
    PSV_TRY(alloc_state(main_));
    PSV_TRY(zero_state(main_));
    if (fused()) {
        PSV_TRY(alloc_state(main2_));
        PSV_TRY(zero_state(main2_));
    }
    PSV_HIP(We_.alloc((size_t)nslots_ * fst_));
    PSV_HIP(psv::launch_fill<T>(We_.p, We_.n, (T)0, stream_));
    PSV_HIP(psv::launch_fill<T>(rec_.p, rec_.n, (T)0, stream_));
    const bool full = grad && !discrete && store_mode_ == PSVFWI_STORE_FULL;
    const bool ckpt = grad && (discrete || store_mode_ == PSVFWI_STORE_CHECKPOINT);
    const bool rb = grad && !discrete && store_mode_ == PSVFWI_STORE_RANDOM_BOUNDARY;
    if (rb) PSV_TRY(rb_check());
    const int kint = discrete ? disc_k_ : ck_k_;
    // discrete gradient with segments of one step: every state is kept (entry
    // t = the state before step t, entry nt = the final state) and the adjoint
    // reads them without a replay
    const bool every = discrete && kint == 1;
    // the store is not cleared: the adjoint reads exactly the entries the
    // forward writes, and get_store zeroes the ones the window loops never
    // write
    if (full) PSV_TRY(alloc_store(1 + (ntmax_ - 1) / w_.dt));
    if (ckpt) PSV_TRY(alloc_checkpoints(every ? 0 : kint));
    const bool snap = snap_interval_ > 0;
    if (snap) PSV_HIP(snaps_.alloc((size_t)((ntmax_ - 1) / snap_interval_ + 1) * nslots_ * fst_));
    snaps_taken_ = snap;

    PSV_TRY(ensure_all_tile_lists());
    PSV_TRY(begin_samples());
    PSV_HIP(hipEventRecord(ev0_, stream_));
    launches_ = 0;
    const int fflags = (energy ? psv::kEnergy : 0) | (full ? psv::kStore : 0);
    std::vector<hipGraphExec_t> graphs;
    AllSlotsAfter restore{this};
    for (int w0 = 0; w0 < nslots_; w0 += group_size()) {
        set_group(w0);
        const int ntg = group_nt();
        for (int it = 0; it < ntg; it++) {
            const int na = active(it);
            if (graph_steps_ > 0 && !snap && !ckpt) {
                // a run of steps with the same active shots: replay a captured
                // chunk of graph_steps_ steps, the step index read from it_dev_
                int end = it;
                while (end < ntg && active(end) == na) end++;
                if (end - it >= graph_steps_) {
                    psv::StepParams<T> p = base_params(main_, fwd_tab_, true);
                    p.nslots = na;
                    p.flags = fflags;
                    p.tf0 = 0;
                    hipGraphExec_t ex = nullptr;
                    PSV_TRY(capture_steps(p, psv::kForward, +1, main_, main2_, fwd_tab_, true, &ex));
                    graphs.push_back(ex);
                    for (; end - it >= graph_steps_; it += graph_steps_) {
                        PSV_HIP(psv::launch_fill<int>(it_dev_.p, 1, it, stream_));
                        PSV_HIP(hipGraphLaunch(ex, stream_));
                        launches_ += graph_steps_;
                    }
                    if (it >= ntg) break;
                }
            }
            if (ckpt && (every ? it == 0 : it % kint == 0)) PSV_TRY(save_checkpoint(it / kint, na));
            psv::StepParams<T> p = base_params(main_, fwd_tab_, true);
            p.it = it;
            p.nslots = active(it);
            p.flags = fflags;
            p.tf0 = 0;
            PSV_TRY(launch_timed(p, psv::kForward, 0, main_, main2_));
            // after the step, for the shots it advanced: a shot's final state is
            // entry nt_s, which the next step's shots (a prefix) may not include
            if (every) PSV_TRY(save_checkpoint(it + 1, na));
            if (snap && it % snap_interval_ == 0) {
                // forward_kernel_PSV_GPU.cu:1067-1080: after the step, every snap_interval
                const size_t k = (size_t)(it / snap_interval_);
                PSV_HIP(hipMemcpyAsync(snaps_.p + (k * nslots_ + w0_) * fst_, offw(main_.f[snap_field_], fst_),
                                       (size_t)na * fst_ * sizeof(T), hipMemcpyDeviceToDevice, stream_));
            }
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
    fwd_done_ = true;
    fwd_grad_done_ = grad;
    fwd_disc_ = discrete;
    fwd_rb_ = rb;
    ck_used_k_ = kint;
    store_loaded_ = full;
    // with the fused kernel a shot that ended early holds its final state in
    // the buffer set it last wrote: main_ when ntmax - nt_s is even
    parity_ = fused();
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::set_observed(const double* obs)
{
    if (nslots_ < 1) return fail(PSVFWI_E_STATE, "set_observed: set_shots first");
    if (!obs) return fail(PSVFWI_E_ARG, "set_observed: null argument");
    // observed traces in slot layout, FP64 on the device
    std::vector<double> h(rec_.n, 0.0);
    size_t o = 0;
    for (int i = 0; i < nslots_; i++) {
        const int s = slot_of_shot_[(size_t)i];
        const size_t len = shots_[(size_t)i].rx.size() * (size_t)shots_[(size_t)i].nt;
        std::memcpy(h.data() + rec_out_off_h_[(size_t)s], obs + o, len * sizeof(double));
        o += len;
    }
    PSV_HIP(obs_.alloc(h.size()));
    PSV_HIP(hipMemcpyAsync(obs_.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
    PSV_HIP(hipStreamSynchronize(stream_));
    have_obs_ = true;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::misfit(const double* obs, double* out)
{
    if (!fwd_done_) return fail(PSVFWI_E_STATE, "misfit: run forward first");
    if (!out) return fail(PSVFWI_E_ARG, "misfit: null argument");
    if (obs) PSV_TRY(set_observed(obs));
    if (!have_obs_) return fail(PSVFWI_E_STATE, "misfit: no observed traces");
    const int nchunk = 64;
    PSV_HIP(partial_.alloc((size_t)nslots_ * nchunk));
    PSV_HIP(misfit_.alloc((size_t)nslots_));
    PSV_HIP(flag_.alloc(1));
    PSV_HIP(psv::launch_fill<int>(flag_.p, 1, 0, stream_));
    PSV_HIP(psv::launch_misfit<T>(rec_.p, obs_.p, adj_sig_.p, rec_out_off_.p, rec_len_.p, nslots_,
                                  nchunk, 0.5 * g_.dt, partial_.p, misfit_.p, flag_.p, stream_));
    std::vector<double> m((size_t)nslots_);
    int flag = 0;
    PSV_HIP(hipMemcpyAsync(m.data(), misfit_.p, m.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    PSV_HIP(hipMemcpyAsync(&flag, flag_.p, sizeof(int), hipMemcpyDeviceToHost, stream_));
    PSV_HIP(hipStreamSynchronize(stream_));
    for (int i = 0; i < nslots_; i++) out[i] = m[(size_t)slot_of_shot_[(size_t)i]];
    std::fill(adj_ready_.begin(), adj_ready_.end(), 1);
    if (flag) return fail(PSVFWI_E_NONFINITE, "misfit: |residual| > 1e10 (the reference aborts here)");
    for (int i = 0; i < nslots_; i++)
        if (!std::isfinite(out[i])) return fail(PSVFWI_E_NONFINITE, "misfit: non-finite misfit");
    return PSVFWI_OK;
}

// misfit() on the resident observed traces, with the partial sums, per-shot misfits and
// range flag in the library's static scratch instead of plan buffers: no device
// allocation (the kernels are misfit()'s, so the results are the same bit for bit)
template <typename T>
psvfwi_status Plan<T>::misfit_noalloc(double* out)
{
    if (!fwd_done_) return fail(PSVFWI_E_STATE, "misfit_noalloc: run forward first");
    if (!out) return fail(PSVFWI_E_ARG, "misfit_noalloc: null argument");
    if (!have_obs_) return fail(PSVFWI_E_STATE, "misfit_noalloc: no observed traces (psvfwi_set_observed)");
    const int nchunk = 64;
    if (nslots_ > PSVFWI_NOALLOC_MAX_SHOTS)
        return fail(PSVFWI_E_ARG, "misfit_noalloc: more than PSVFWI_NOALLOC_MAX_SHOTS shots in the batch");
    const size_t need = ((size_t)nslots_ * (nchunk + 1)) * sizeof(double) + sizeof(int);
    if (need > psv::kNoAllocBytes) return fail(PSVFWI_E_ARG, "misfit_noalloc: batch exceeds the scratch");
    void* base = nullptr;
    PSV_HIP(psv::noalloc_scratch(&base));
    double* partial = static_cast<double*>(base);
    double* mis = partial + (size_t)nslots_ * nchunk;
    int* flag = reinterpret_cast<int*>(mis + nslots_);
    // the flag starts at zero on every call (a fill kernel, not a memset)
    PSV_HIP(psv::launch_fill<int>(flag, 1, 0, stream_));
    PSV_HIP(psv::launch_misfit<T>(rec_.p, obs_.p, adj_sig_.p, rec_out_off_.p, rec_len_.p, nslots_,
                                  nchunk, 0.5 * g_.dt, partial, mis, flag, stream_));
    std::vector<double> m((size_t)nslots_);
    int hflag = 0;
    PSV_HIP(hipMemcpyAsync(m.data(), mis, m.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    PSV_HIP(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, stream_));
    PSV_HIP(hipStreamSynchronize(stream_));
    for (int i = 0; i < nslots_; i++) out[i] = m[(size_t)slot_of_shot_[(size_t)i]];
    std::fill(adj_ready_.begin(), adj_ready_.end(), 1);
    if (hflag) return fail(PSVFWI_E_NONFINITE, "misfit_noalloc: |residual| > 1e10 (the reference aborts here)");
    for (int i = 0; i < nslots_; i++)
        if (!std::isfinite(out[i])) return fail(PSVFWI_E_NONFINITE, "misfit_noalloc: non-finite misfit");
    return PSVFWI_OK;
}

#define PSV_INSTANTIATE(T) \
    template psvfwi_status Plan<T>::forward(int); \
    template psvfwi_status Plan<T>::set_observed(const double*); \
    template psvfwi_status Plan<T>::misfit(const double*, double*); \
    template psvfwi_status Plan<T>::misfit_noalloc(double*);
PSV_FOR_EACH_PRECISION(PSV_INSTANTIATE)

}  // namespace psvplan
