// plan_setup.cpp -- construction of a plan, its configuration (model, CPML, window, store and
// gradient modes, tiles, launch groups) and its state and store buffers.
#include <cstdlib>

#include "plan.h"

namespace psvplan {

template <typename T>
Plan<T>::Plan(const psvfwi_grid_desc& g) : g_(g)
{
    fpad_ = g.fdorder / 2;
    ppad_ = g.npml + fpad_;
    tp_ = 2 * (g.npml + 1);
    // guard-padded rows: kGuardLeft zero columns before cell 0, room for the
    // last tile and kGuardRight columns after it; kGuardRows zero rows above
    // and below the rows of the tallest tiles (psv_step.h)
    const int tiles_x = (g.nxt + psv::kTX - 1) / psv::kTX;
    pitch_ = (int)round_up((size_t)(psv::kGuardLeft + tiles_x * psv::kTX + psv::kGuardRight), 32);
    rows_ = psv::kGuardRows + (int)round_up((size_t)g.nzt, psv::kMaxTz) + psv::kGuardRows;
    fst_ = (size_t)rows_ * pitch_;
    org_ = psv::kGuardRows * pitch_ + psv::kGuardLeft;
    xst_ = round_up((size_t)g.nzt * tp_, 64);
    zst_ = round_up((size_t)tp_ * pitch_, 64);
    pnx_ = g.nxt - 2 * ppad_ + fpad_ - 1;
    // bottom strip offset; equals the reference's nzt - 2 ppad - 1 at fpad == 1
    pnz_ = g.fsurf ? g.nzt - 2 * ppad_ + fpad_ - 2 : g.nzt - 2 * ppad_ + fpad_ - 1;
    tz_ = kDefaultTz;
}

template <typename T>
Plan<T>::~Plan()
{
    for (int i = 0; i < kMaxSamples; i++)
        for (int j = 0; j < 3; j++)
            if (sample_ev_[i][j]) (void)hipEventDestroy(sample_ev_[i][j]);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (side_.fork) (void)hipEventDestroy(side_.fork);
    if (side_.join) (void)hipEventDestroy(side_.join);
    if (side_.stream) (void)hipStreamDestroy(side_.stream);
    if (stream_) (void)hipStreamDestroy(stream_);
}

template <typename T>
psvfwi_status Plan<T>::init()
{
    PSV_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    PSV_HIP(hipStreamCreateWithFlags(&side_.stream, hipStreamNonBlocking));
    PSV_HIP(hipEventCreateWithFlags(&side_.fork, hipEventDisableTiming));
    PSV_HIP(hipEventCreateWithFlags(&side_.join, hipEventDisableTiming));
    {
        const char* e = std::getenv("PSVFWI_MARCH_STREAMS");
        use_side_ = !(e && std::string(e) == "1");
    }
    PSV_HIP(hipEventCreate(&ev0_));
    PSV_HIP(hipEventCreate(&ev1_));
    PSV_HIP(pml_.alloc((size_t)6 * tp_));
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::set_pml(const double* prof)
{
    if (!prof) return fail(PSVFWI_E_ARG, "set_pml: null profiles");
    std::vector<T> h((size_t)6 * tp_);
    for (size_t i = 0; i < h.size(); i++) h[i] = (T)prof[i];
    h_pml_.assign(prof, prof + h.size());
    PSV_HIP(hipMemcpyAsync(pml_.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, stream_));
    PSV_HIP(hipStreamSynchronize(stream_));
    have_pml_ = true;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::set_model(const double* lam, const double* mu, const double* mu_zx, const double* rho_zp, const double* rho_xp)
{
    if (!lam || !mu || !mu_zx || !rho_zp || !rho_xp) return fail(PSVFWI_E_ARG, "set_model: null array");
    const int nzt = g_.nzt, nxt = g_.nxt;
    DBuf<T>* dst[5] = {&lam_, &mu_, &mu_zx_, &rho_zp_, &rho_xp_};
    const double* src[5] = {lam, mu, mu_zx, rho_zp, rho_xp};
    std::vector<T> h(fst_);
    for (int a = 0; a < 5; a++) {
        const bool avg = a >= 2;  // [nzt-1][nxt-1]
        const int rows = avg ? nzt - 1 : nzt, cols = avg ? nxt - 1 : nxt;
        std::fill(h.begin(), h.end(), (T)0);
        for (int iz = 0; iz < rows; iz++)
            for (int ix = 0; ix < cols; ix++)
                h[(size_t)org_ + (size_t)iz * pitch_ + ix] = (T)src[a][(size_t)iz * cols + ix];
        PSV_HIP(dst[a]->alloc(h.size()));
        PSV_HIP(hipMemcpyAsync(dst[a]->p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, stream_));
        PSV_HIP(hipStreamSynchronize(stream_));
    }
    // host copies for the chain rule of the discrete gradient
    h_mu_.assign(mu, mu + (size_t)nzt * nxt);
    const size_t na = (size_t)(nzt - 1) * (nxt - 1);
    h_mu_zx_.assign(mu_zx, mu_zx + na);
    h_rho_zp_.assign(rho_zp, rho_zp + na);
    h_rho_xp_.assign(rho_xp, rho_xp + na);
    have_model_ = true;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::set_window(const psvfwi_window_desc* w)
{
    if (!w) {
        win_on_ = false;
        drop_results();
        return PSVFWI_OK;
    }
    if (w->dt < 1 || w->dx < 1 || w->dz < 1) return fail(PSVFWI_E_ARG, "set_window: strides must be >= 1");
    if (w->x1 < 0 || w->z1 < 0 || w->x2 > g_.nxt || w->z2 > g_.nzt || w->x2 < w->x1 || w->z2 < w->z1)
        return fail(PSVFWI_E_ARG, "set_window: window outside the padded grid");
    w_ = *w;
    nfz_ = 1 + (w->z2 - w->z1) / w->dz;
    nfx_ = 1 + (w->x2 - w->x1) / w->dx;
    win_on_ = true;
    drop_results();
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::set_store_mode(int mode, int interval)
{
    if (mode != PSVFWI_STORE_FULL && mode != PSVFWI_STORE_CHECKPOINT && mode != PSVFWI_STORE_RANDOM_BOUNDARY)
        return fail(PSVFWI_E_ARG, "set_store_mode: unknown mode");
    if (mode == PSVFWI_STORE_CHECKPOINT && interval < 1)
        return fail(PSVFWI_E_ARG, "set_store_mode: checkpoint interval must be >= 1");
    store_mode_ = mode;
    ck_k_ = interval;
    drop_results();
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::set_gradient_mode(int mode, int interval)
{
    if (mode != PSVFWI_GRADIENT_REFERENCE && mode != PSVFWI_GRADIENT_DISCRETE)
        return fail(PSVFWI_E_ARG, "set_gradient_mode: unknown mode");
    if (mode == PSVFWI_GRADIENT_DISCRETE && interval < 1)
        return fail(PSVFWI_E_ARG, "set_gradient_mode: segment length must be >= 1");
    grad_mode_ = mode;
    if (mode == PSVFWI_GRADIENT_DISCRETE) disc_k_ = interval;
    drop_results();
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::set_snapshots(int interval, int field)
{
    if (interval < 0 || field < 0 || field > 4) return fail(PSVFWI_E_ARG, "set_snapshots: bad interval or field");
    snap_interval_ = interval;
    snap_field_ = field;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::set_tile(int tz)
{
    if (tz == 0) tz = kDefaultTz;
    if (!tile_ok(tz))
        return fail(PSVFWI_E_ARG, "set_tile: tile height must be 4, 8, 16 or (FP32) 32 (split step), -8 or -16 (fused step), "
                                  "-32 or -64 (z-marching step)");
    tz_ = tz;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::set_graph_steps(int n)
{
    if (n < 0 || (n % 2) != 0) return fail(PSVFWI_E_ARG, "set_graph_steps: need 0 or an even number of steps");
    if (n > 0) PSV_HIP(it_dev_.alloc(1));
    graph_steps_ = n;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::set_shot_order(int band)
{
    if (band < 0) return fail(PSVFWI_E_ARG, "set_shot_order: band must be >= 0");
    band_ = band;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::set_launch_group(int shots)
{
    if (shots < 0) return fail(PSVFWI_E_ARG, "set_launch_group: shots must be >= 0");
    lg_ = shots;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::launch_group(int* shots) const
{
    if (!shots) return fail(PSVFWI_E_ARG, "launch_group: null argument");
    *shots = lg_;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::shot_order(int* band) const
{
    if (!band) return fail(PSVFWI_E_ARG, "shot_order: null argument");
    *band = band_;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::ready() const
{
    if (!have_model_) return fail(PSVFWI_E_STATE, "set_model has not been called");
    if (g_.npml > 0 && !have_pml_) return fail(PSVFWI_E_STATE, "set_pml has not been called");
    if (nslots_ < 1) return fail(PSVFWI_E_STATE, "set_shots has not been called");
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::alloc_state(State<T>& st)
{
    const size_t nf = (size_t)nslots_ * fst_;
    for (int f = 0; f < 5; f++) PSV_HIP(st.f[f].alloc(nf));
    const size_t nx = g_.npml > 0 ? (size_t)nslots_ * xst_ : 1;
    const size_t nz = g_.npml > 0 ? (size_t)nslots_ * zst_ : 1;
    for (State<T>* q : {&st}) {
        PSV_HIP(q->ms[0].alloc(nx));
        PSV_HIP(q->ms[1].alloc(nx));
        PSV_HIP(q->ms[2].alloc(nz));
        PSV_HIP(q->ms[3].alloc(nz));
        PSV_HIP(q->mv[0].alloc(nx));
        PSV_HIP(q->mv[1].alloc(nx));
        PSV_HIP(q->mv[2].alloc(nz));
        PSV_HIP(q->mv[3].alloc(nz));
    }
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::zero_state(State<T>& st)
{
    for (int f = 0; f < 5; f++) PSV_HIP(psv::launch_fill<T>(st.f[f].p, st.f[f].n, (T)0, stream_));
    for (int m = 0; m < 4; m++) {
        PSV_HIP(psv::launch_fill<T>(st.ms[m].p, st.ms[m].n, (T)0, stream_));
        PSV_HIP(psv::launch_fill<T>(st.mv[m].p, st.mv[m].n, (T)0, stream_));
    }
    return PSVFWI_OK;
}

// Forgets every result held: a change of window, shots, store or gradient mode, or the
// autotuner's runs in the plan's buffers, leave them unreadable (the getters refuse).
template <typename T>
void Plan<T>::drop_results()
{
    ran_ = false;
    fwd_done_ = fwd_grad_done_ = store_loaded_ = adj_grad_done_ = false;
    fwd_disc_ = disc_done_ = fwd_rb_ = false;
}

template <typename T>
psvfwi_status Plan<T>::alloc_store(int ntf)
{
    sst_ = (size_t)ntf * 5 * nfz_ * nfx_;
    PSV_HIP(store_.reserve((size_t)nslots_ * sst_));
    return PSVFWI_OK;
}

// checkpoint c holds, for every slot, the 5 wavefields (current set), the
// 4 stress-side CPML arrays (current set) and the 4 velocity-side arrays
template <typename T>
size_t Plan<T>::ck_bytes_per_slot() const
{
    const size_t m = g_.npml > 0 ? 4 * (xst_ + zst_) : 0;
    return (5 * fst_ + m) * sizeof(T);
}

template <typename T>
psvfwi_status Plan<T>::alloc_checkpoints(int k)
{
    ncp_ = k > 0 ? (ntmax_ + k - 1) / k : ntmax_ + 1;
    const hipError_t e = ckpt_.alloc((size_t)ncp_ * state_elems());
    if (e == hipErrorOutOfMemory && k == 0) {
        (void)hipGetLastError();
        return fail(PSVFWI_E_NOMEM, "forward: keeping every state (discrete gradient, segment 1) needs " +
                                        std::to_string((double)ncp_ * state_elems() * sizeof(T) / 1e9) +
                                        " GB of device memory; use a longer segment");
    }
    PSV_HIP(e);
    return PSVFWI_OK;
}

#define PSV_INSTANTIATE(T) \
    template Plan<T>::Plan(const psvfwi_grid_desc&); \
    template Plan<T>::~Plan(); \
    template psvfwi_status Plan<T>::init(); \
    template psvfwi_status Plan<T>::set_pml(const double*); \
    template psvfwi_status Plan<T>::set_model(const double*, const double*, const double*, const double*, const double*); \
    template psvfwi_status Plan<T>::set_window(const psvfwi_window_desc*); \
    template psvfwi_status Plan<T>::set_store_mode(int, int); \
    template psvfwi_status Plan<T>::set_gradient_mode(int, int); \
    template psvfwi_status Plan<T>::set_snapshots(int, int); \
    template psvfwi_status Plan<T>::set_tile(int); \
    template psvfwi_status Plan<T>::set_graph_steps(int); \
    template psvfwi_status Plan<T>::set_shot_order(int); \
    template psvfwi_status Plan<T>::set_launch_group(int); \
    template psvfwi_status Plan<T>::launch_group(int*) const; \
    template psvfwi_status Plan<T>::shot_order(int*) const; \
    template psvfwi_status Plan<T>::ready() const; \
    template psvfwi_status Plan<T>::alloc_state(State<T>&); \
    template psvfwi_status Plan<T>::zero_state(State<T>&); \
    template void Plan<T>::drop_results(); \
    template psvfwi_status Plan<T>::alloc_store(int); \
    template size_t Plan<T>::ck_bytes_per_slot() const; \
    template psvfwi_status Plan<T>::alloc_checkpoints(int);
PSV_FOR_EACH_PRECISION(PSV_INSTANTIATE)

}  // namespace psvplan
