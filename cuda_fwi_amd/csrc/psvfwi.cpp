// psvfwi.cpp -- plan, device memory and time loops behind the C ABI (psvfwi.h).
//
// A plan owns one HIP stream and every device buffer of a batch of shots on
// one model.  Shots are kept in "slots" sorted by record length (longest
// first), so the shots still running at step it are always a prefix of the
// slots and one launch of the step kernel covers exactly them.
//
// Forward (forward_kernel_PSV, CUDA_week2/forward_kernel_PSV_GPU.cu:681-1084):
//   it = 0 .. ntmax-1, state ping-pong between two buffer sets.
// Adjoint (adjoint_kernel_PSV, CUDA_week2/adjoint_kernel_PSV_GPU.cu:688-1102):
//   it = ntmax-1 .. 0 on a zeroed state, residuals injected at the receivers.
// Gradient storage: either the reference's decimated store of every fwi_dt-th
// forward state (PSVFWI_STORE_FULL), or checkpoints of the full state
// (wavefields + CPML memory) every k steps; the adjoint then replays each
// segment forward into a small segment store just before it needs it
// (PSVFWI_STORE_CHECKPOINT).  Both give the same gradient bit for bit.
// Random-boundary reconstruction (PSVFWI_STORE_RANDOM_BOUNDARY, an
// extension): without CPML the scheme runs backward exactly up to rounding,
// so the adjoint rebuilds the forward state step by step from the forward's
// final state (reverse step: sources off, velocity kernel with -dt, stress
// kernel with -dt) and keeps one store plane; nothing is stored in the forward.
// Discrete gradient (PSVFWI_GRADIENT_DISCRETE, an extension): checkpoints of
// the full state every k steps; the adjoint replays each segment keeping
// every state of it and runs the transposed step (psv_discrete.hip) back
// through the segment, accumulating the exact misfit gradient with respect to
// the five coefficient arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "psv_aux.h"
#include "psv_discrete.h"
#include "psv_step.h"
#include "psvfwi.h"

namespace {

thread_local std::string g_err;

// A ROCTX range around a C ABI call (forward, misfit, adjoint, autotune), seen by
// rocprofv3 --marker-trace beside the kernels it launched
struct Range {
    explicit Range(const char* what) { roctxRangePush(what); }
    ~Range() { roctxRangePop(); }
    Range(const Range&) = delete;
    Range& operator=(const Range&) = delete;
};

psvfwi_status fail(psvfwi_status st, const std::string& msg)
{
    g_err = msg;
    return st;
}

#define PSV_HIP(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(e_ == hipErrorOutOfMemory ? PSVFWI_E_NOMEM : PSVFWI_E_HIP,       \
                        std::string(#expr) + ": " + hipGetErrorString(e_));              \
    } while (0)

#define PSV_TRY(expr)                     \
    do {                                  \
        psvfwi_status s_ = (expr);        \
        if (s_ != PSVFWI_OK) return s_;   \
    } while (0)

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;    // elements in use
    size_t cap = 0;  // elements allocated
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = cap = 0;
    }
    // exactly count elements
    hipError_t alloc(size_t count)
    {
        if (p && count == n && count == cap) return hipSuccess;
        release();
        if (count == 0) return hipSuccess;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
        if (e == hipSuccess) n = cap = count;
        else p = nullptr;
        return e;
    }
    // at least count elements; keeps a larger allocation (large buffers whose
    // size changes between runs are not freed and mapped again every time)
    hipError_t reserve(size_t count)
    {
        if (p && count <= cap) {
            n = count;
            return hipSuccess;
        }
        return alloc(count);
    }
};

struct ShotHost {
    int nt = 0;
    std::vector<int> sx, sz, sc;
    std::vector<double> ssig;  // [nsrc][nt]
    std::vector<int> rx, rz, rc;
};

size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

}  // namespace

struct psvfwi_plan {
    virtual ~psvfwi_plan() = default;
    virtual psvfwi_status set_pml(const double*) = 0;
    virtual psvfwi_status set_model(const double*, const double*, const double*, const double*,
                                    const double*) = 0;
    virtual psvfwi_status set_window(const psvfwi_window_desc*) = 0;
    virtual psvfwi_status set_store_mode(int, int) = 0;
    virtual psvfwi_status set_shots(int, const psvfwi_shot_desc*) = 0;
    virtual psvfwi_status set_tile(int) = 0;
    virtual psvfwi_status set_graph_steps(int) = 0;
    virtual psvfwi_status set_shot_order(int) = 0;
    virtual psvfwi_status shot_order(int*) const = 0;
    virtual psvfwi_status set_launch_group(int) = 0;
    virtual psvfwi_status launch_group(int*) const = 0;
    virtual psvfwi_status set_snapshots(int, int) = 0;
    virtual psvfwi_status get_snapshot(int, int, double*) const = 0;
    virtual psvfwi_status snapshot_count(int, int*) const = 0;
    virtual psvfwi_status autotune(int, int*) = 0;
    virtual psvfwi_status forward(int) = 0;
    virtual psvfwi_status set_observed(const double*) = 0;
    virtual psvfwi_status misfit(const double*, double*) = 0;
    virtual psvfwi_status misfit_noalloc(double*) = 0;
    virtual psvfwi_status set_adjoint_sources(int, const double*) = 0;
    virtual psvfwi_status set_store(int, const double*) = 0;
    virtual psvfwi_status adjoint(int) = 0;
    virtual psvfwi_status get_records(int, double*) const = 0;
    virtual psvfwi_status get_field(int, int, double*) const = 0;
    virtual psvfwi_status get_store(int, double*) const = 0;
    virtual psvfwi_status get_gradient(int, double*, double*, double*) const = 0;
    virtual psvfwi_status set_gradient_mode(int, int) = 0;
    virtual psvfwi_status get_coefficient_gradient(int, double*, double*, double*, double*, double*) const = 0;
    virtual psvfwi_status get_model_gradient(int, double*, double*, double*) const = 0;
    virtual psvfwi_status timing(double*, long*) const = 0;
    virtual psvfwi_status kernel_timing(int, double*, double*, int*) const = 0;
};

namespace {

enum { kF_VX, kF_VZ, kF_SXX, kF_SZX, kF_SZZ };
// stress-side CPML memory: 0 mem_vx_x, 1 mem_vz_x (x-type), 2 mem_vx_z, 3 mem_vz_z (z-type)
// velocity-side: 0 mem_sxx_x, 1 mem_szx_x (x-type), 2 mem_szz_z, 3 mem_szx_z (z-type)

template <typename T>
struct State {
    DBuf<T> f[5];   // vx, vz, sxx, szx, szz
    DBuf<T> ms[4];  // stress-side CPML memory
    DBuf<T> mv[4];  // velocity-side CPML memory
};

template <typename T>
class Plan final : public psvfwi_plan {
public:
    static constexpr int kDefaultTz = 8;

    explicit Plan(const psvfwi_grid_desc& g) : g_(g)
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

    ~Plan() override
    {
        for (int i = 0; i < kMaxSamples; i++)
            for (int j = 0; j < 3; j++)
                if (sample_ev_[i][j]) (void)hipEventDestroy(sample_ev_[i][j]);
        if (ev0_) (void)hipEventDestroy(ev0_);
        if (ev1_) (void)hipEventDestroy(ev1_);
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    psvfwi_status init()
    {
        PSV_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        PSV_HIP(hipEventCreate(&ev0_));
        PSV_HIP(hipEventCreate(&ev1_));
        PSV_HIP(pml_.alloc((size_t)6 * tp_));
        return PSVFWI_OK;
    }

    // ------------------------------------------------------------------ setup
    psvfwi_status set_pml(const double* prof) override
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

    psvfwi_status set_model(const double* lam, const double* mu, const double* mu_zx,
                            const double* rho_zp, const double* rho_xp) override
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

    psvfwi_status set_window(const psvfwi_window_desc* w) override
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

    psvfwi_status set_store_mode(int mode, int interval) override
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

    psvfwi_status set_gradient_mode(int mode, int interval) override
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

    // tz > 0: split step kernels with 64 x tz tiles; tz = -8, -16: the fused step
    // kernel with 64 x |tz| tiles; tz = -32, -64: the z-marching fused step on
    // 64-column segments of |tz| rows
    static bool tile_ok(int tz)
    {
        if (tz < 0) return tz == -8 || tz == -16 || tz == -32 || tz == -64;
        return tz == 4 || tz == 8 || tz == 16 || (tz == 32 && sizeof(T) == 4);
    }
    int tz_abs() const { return tz_ < 0 ? -tz_ : tz_; }
    bool fused() const { return tz_ < 0; }

    psvfwi_status set_snapshots(int interval, int field) override
    {
        if (interval < 0 || field < 0 || field > 4) return fail(PSVFWI_E_ARG, "set_snapshots: bad interval or field");
        snap_interval_ = interval;
        snap_field_ = field;
        return PSVFWI_OK;
    }

    psvfwi_status get_snapshot(int shot, int k, double* out) const override
    {
        if (shot < 0 || shot >= nslots_ || !out) return fail(PSVFWI_E_ARG, "get_snapshot: bad shot");
        const int s = slot_of_shot_[(size_t)shot];
        const int n = nsnap_slot(s);
        if (k < 0 || k >= n) return fail(PSVFWI_E_ARG, "get_snapshot: no such snapshot");
        std::vector<double> h((size_t)g_.nzt * pitch_);
        PSV_TRY(download(snaps_.p + ((size_t)k * nslots_ + s) * fst_ + org_, h.size(), h.data()));
        for (int iz = 0; iz < g_.nzt; iz++)
            std::memcpy(out + (size_t)iz * g_.nxt, h.data() + (size_t)iz * pitch_, (size_t)g_.nxt * sizeof(double));
        return PSVFWI_OK;
    }

    psvfwi_status snapshot_count(int shot, int* n) const override
    {
        if (shot < 0 || shot >= nslots_ || !n) return fail(PSVFWI_E_ARG, "snapshot_count: bad shot");
        *n = snaps_taken_ ? nsnap_slot(slot_of_shot_[(size_t)shot]) : 0;
        return PSVFWI_OK;
    }

    psvfwi_status set_tile(int tz) override
    {
        if (tz == 0) tz = kDefaultTz;
        if (!tile_ok(tz))
            return fail(PSVFWI_E_ARG, "set_tile: tile height must be 4, 8, 16 or (FP32) 32 (split step), -8 or -16 (fused step)");
        tz_ = tz;
        return PSVFWI_OK;
    }

    psvfwi_status set_graph_steps(int n) override
    {
        if (n < 0 || (n % 2) != 0) return fail(PSVFWI_E_ARG, "set_graph_steps: need 0 or an even number of steps");
        if (n > 0) PSV_HIP(it_dev_.alloc(1));
        graph_steps_ = n;
        return PSVFWI_OK;
    }

    // Times `steps` forward launches of the current batch for every tile
    // height and keeps the fastest.  Runs on the plan's own state buffers
    // (every forward/adjoint starts from a zeroed state, so nothing leaks).
    psvfwi_status autotune(int steps, int* best) override
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

    psvfwi_status set_shot_order(int band) override
    {
        if (band < 0) return fail(PSVFWI_E_ARG, "set_shot_order: band must be >= 0");
        band_ = band;
        return PSVFWI_OK;
    }

    psvfwi_status set_launch_group(int shots) override
    {
        if (shots < 0) return fail(PSVFWI_E_ARG, "set_launch_group: shots must be >= 0");
        lg_ = shots;
        return PSVFWI_OK;
    }

    psvfwi_status launch_group(int* shots) const override
    {
        if (!shots) return fail(PSVFWI_E_ARG, "launch_group: null argument");
        *shots = lg_;
        return PSVFWI_OK;
    }

    psvfwi_status shot_order(int* band) const override
    {
        if (!band) return fail(PSVFWI_E_ARG, "shot_order: null argument");
        *band = band_;
        return PSVFWI_OK;
    }

    psvfwi_status set_shots(int n, const psvfwi_shot_desc* sd) override
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

    // ------------------------------------------------------------------ runs
    psvfwi_status forward(int flags) override
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
        // the store is not cleared: the adjoint reads exactly the entries the
        // forward writes, and get_store zeroes the ones the window loops never
        // write (see window_written)
        if (full) PSV_TRY(alloc_store(1 + (ntmax_ - 1) / w_.dt));
        if (ckpt) PSV_TRY(alloc_checkpoints(kint));
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
                if (ckpt && it % kint == 0) PSV_TRY(save_checkpoint(it / kint, na));
                psv::StepParams<T> p = base_params(main_, fwd_tab_, true);
                p.it = it;
                p.nslots = active(it);
                p.flags = fflags;
                p.tf0 = 0;
                PSV_TRY(launch_timed(p, psv::kForward, 0, main_, main2_));
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

    psvfwi_status set_observed(const double* obs) override
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

    psvfwi_status misfit(const double* obs, double* out) override
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
    psvfwi_status misfit_noalloc(double* out) override
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

    psvfwi_status set_adjoint_sources(int shot, const double* sig) override
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

    psvfwi_status set_store(int shot, const double* st) override
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

    psvfwi_status adjoint(int flags) override
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
    psvfwi_status rb_check() const
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
    psvfwi_status adjoint_rb(int flags)
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

    // Discrete gradient: per segment (last first), restore its checkpoint,
    // replay it keeping the state before every step and after the last, then
    // run the transposed steps back through it.
    psvfwi_status adjoint_discrete()
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
        PSV_TRY(alloc_state(replay_));
        if (fused()) {
            PSV_TRY(alloc_state(replay2_));
            PSV_TRY(zero_state(replay2_));
        }
        PSV_HIP(seg_.reserve((size_t)(std::min(kk, ntmax_) + 1) * state_elems()));
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
                for (int it = t1 - 1; it >= t0; it--) {
                    psv::DiscreteParams<T> d = dparams(it, entry(it - t0), entry(it - t0 + 1));
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

    // device pointers into one saved state (state_copy layout)
    struct View {
        const T* f[5];
        const T* ms[4];
        const T* mv[4];
    };

    View entry(int e) const
    {
        const T* base = seg_.p + (size_t)e * state_elems();
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

    psv::DiscreteParams<T> dparams(int it, const View& pre, const View& post)
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

    // ------------------------------------------------------------------ getters
    psvfwi_status get_coefficient_gradient(int shot, double* gl, double* gm, double* gmzx, double* grzp,
                                           double* grxp) const override
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

    psvfwi_status get_model_gradient(int shot, double* gl, double* gm, double* gr) const override
    {
        if (!gl || !gm || !gr) return fail(PSVFWI_E_ARG, "get_model_gradient: null array");
        const size_t n = (size_t)g_.nzt * g_.nxt, na = (size_t)(g_.nzt - 1) * (g_.nxt - 1);
        std::vector<double> c0(n), c1(n), c2(na), c3(na), c4(na);
        PSV_TRY(get_coefficient_gradient(shot, c0.data(), c1.data(), c2.data(), c3.data(), c4.data()));
        return psvfwi_chain_gradient(g_.nzt, g_.nxt, h_mu_.data(), h_mu_zx_.data(), h_rho_zp_.data(),
                                     h_rho_xp_.data(), c0.data(), c1.data(), c2.data(), c3.data(), c4.data(),
                                     gl, gm, gr);
    }

    psvfwi_status get_records(int shot, double* out) const override
    {
        if (shot < 0 || shot >= nslots_ || !out) return fail(PSVFWI_E_ARG, "get_records: bad shot");
        if (!fwd_done_) return fail(PSVFWI_E_STATE, "get_records: run forward first");
        const int s = slot_of_shot_[(size_t)shot];
        const size_t len = shots_[(size_t)shot].rx.size() * (size_t)shots_[(size_t)shot].nt;
        return download(rec_.p + rec_out_off_h_[(size_t)s], len, out);
    }

    psvfwi_status get_field(int shot, int field, double* out) const override
    {
        if (shot < 0 || shot >= nslots_ || !out) return fail(PSVFWI_E_ARG, "get_field: bad shot");
        if (!ran_) return fail(PSVFWI_E_STATE, "get_field: nothing has run");
        const int s = slot_of_shot_[(size_t)shot];
        const int nzt = g_.nzt, nxt = g_.nxt;
        const State<T>& cur = (parity_ && ((ntmax_ - nt_slot_[(size_t)s]) & 1)) ? main2_ : main_;
        if (field >= 0 && field <= 7) {
            const T* base = nullptr;
            if (field <= 4) base = cur.f[field].p;
            else if (field == 5) base = We_.p;
            else if (field == 6) base = ux_.p;
            else base = uz_.p;
            if (!base) return fail(PSVFWI_E_STATE, "get_field: field not allocated");
            std::vector<double> h((size_t)nzt * pitch_);
            PSV_TRY(download(base + (size_t)s * fst_ + org_, h.size(), h.data()));
            for (int iz = 0; iz < nzt; iz++)
                std::memcpy(out + (size_t)iz * nxt, h.data() + (size_t)iz * pitch_, (size_t)nxt * sizeof(double));
            return PSVFWI_OK;
        }
        if (field >= 8 && field <= 15) {
            if (g_.npml <= 0) return fail(PSVFWI_E_ARG, "get_field: no CPML");
            const int k = field - 8;
            // 8 mem_vx_x, 9 mem_vz_x (stress side, x), 10 mem_sxx_x, 11 mem_szx_x (velocity side, x)
            // 12 mem_vx_z, 13 mem_vz_z (stress side, z), 14 mem_szz_z, 15 mem_szx_z (velocity side, z)
            const T* base = nullptr;
            bool xtype = k < 4;
            switch (k) {
            case 0: base = cur.ms[0].p; break;
            case 1: base = cur.ms[1].p; break;
            case 2: base = main_.mv[0].p; break;
            case 3: base = main_.mv[1].p; break;
            case 4: base = cur.ms[2].p; break;
            case 5: base = cur.ms[3].p; break;
            case 6: base = main_.mv[2].p; break;
            default: base = main_.mv[3].p; break;
            }
            if (xtype) return download(base + (size_t)s * xst_, (size_t)nzt * tp_, out);
            std::vector<double> h((size_t)tp_ * pitch_);
            PSV_TRY(download(base + (size_t)s * zst_, h.size(), h.data()));
            for (int r = 0; r < tp_; r++)
                std::memcpy(out + (size_t)r * nxt, h.data() + (size_t)r * pitch_, (size_t)nxt * sizeof(double));
            return PSVFWI_OK;
        }
        return fail(PSVFWI_E_ARG, "get_field: unknown field");
    }

    psvfwi_status get_store(int shot, double* out) const override
    {
        if (shot < 0 || shot >= nslots_ || !out) return fail(PSVFWI_E_ARG, "get_store: bad shot");
        if (!store_loaded_ || store_mode_ != PSVFWI_STORE_FULL)
            return fail(PSVFWI_E_STATE, "get_store: no full forward store held");
        const int s = slot_of_shot_[(size_t)shot];
        const int nft = 1 + (shots_[(size_t)shot].nt - 1) / w_.dt;
        const size_t plane = (size_t)nfz_ * nfx_;
        if (sst_ < (size_t)nft * 5 * plane || store_.n < (size_t)nslots_ * sst_)
            return fail(PSVFWI_E_STATE, "get_store: the store does not match the current window and shots");
        std::vector<double> h(sst_);
        PSV_TRY(download(store_.p + (size_t)s * sst_, sst_, h.data()));
        // the window loops run over [z1, z2) x [x1, x2): when an extent is a
        // multiple of its stride the last row / column of the store is never
        // written (CUDA_week2/forward_kernel_PSV_GPU.cu:803-821) and reads 0
        const int nzw = (w_.z2 - w_.z1 + w_.dz - 1) / w_.dz, nxw = (w_.x2 - w_.x1 + w_.dx - 1) / w_.dx;
        for (int f = 0; f < 5; f++)
            for (int t = 0; t < nft; t++) {
                double* o = out + ((size_t)f * nft + t) * plane;
                const double* src = h.data() + ((size_t)t * 5 + f) * plane;
                for (int zf = 0; zf < nfz_; zf++)
                    for (int xf = 0; xf < nfx_; xf++) {
                        const size_t k = (size_t)zf * nfx_ + xf;
                        o[k] = (zf < nzw && xf < nxw) ? src[k] : 0.0;
                    }
            }
        return PSVFWI_OK;
    }

    psvfwi_status get_gradient(int shot, double* gl, double* gm, double* gr) const override
    {
        if (!adj_grad_done_) return fail(PSVFWI_E_STATE, "get_gradient: run adjoint with PSVFWI_GRADIENT first");
        if (shot < -1 || shot >= nslots_ || !gl || !gm || !gr) return fail(PSVFWI_E_ARG, "get_gradient: bad shot");
        const size_t plane = (size_t)nfz_ * nfx_;
        if (gst_ < 3 * plane || grad_.n < (size_t)nslots_ * gst_)
            return fail(PSVFWI_E_STATE, "get_gradient: the gradient does not match the current window and shots");
        std::vector<double> h(3 * plane);
        if (shot >= 0) {
            const int s = slot_of_shot_[(size_t)shot];
            PSV_TRY(download(grad_.p + (size_t)s * gst_, 3 * plane, h.data()));
        } else {
            DBuf<double> sum;
            DBuf<int> order;
            PSV_HIP(sum.alloc(3 * plane));
            PSV_HIP(order.alloc((size_t)nslots_));
            PSV_HIP(hipMemcpy(order.p, slot_of_shot_.data(), (size_t)nslots_ * sizeof(int), hipMemcpyHostToDevice));
            PSV_HIP(psv::launch_grad_sum<T>(grad_.p, gst_, order.p, nslots_, 3 * plane, sum.p, stream_));
            PSV_HIP(hipMemcpyAsync(h.data(), sum.p, 3 * plane * sizeof(double), hipMemcpyDeviceToHost, stream_));
            PSV_HIP(hipStreamSynchronize(stream_));
        }
        std::memcpy(gl, h.data(), plane * sizeof(double));
        std::memcpy(gm, h.data() + plane, plane * sizeof(double));
        std::memcpy(gr, h.data() + 2 * plane, plane * sizeof(double));
        return PSVFWI_OK;
    }

    psvfwi_status timing(double* ms, long* launches) const override
    {
        if (ms) *ms = last_ms_;
        if (launches) *launches = launches_;
        return PSVFWI_OK;
    }

private:
    struct Tables {
        DBuf<int> inj_grp, inj_row, rec_zxc, rec_first;
        DBuf<T> sig;
        DBuf<size_t> sig_off;
        // host copies: per-slot ranges of groups / receivers and their cells
        std::vector<int> grp_off, grp_zx, rec_off, rec_zx;
        // per-tile lists for the tile height they were built for
        DBuf<int> tinj_off, tinj_idx, trec_off, trec_idx;
        int built_tz = -1;
    };

    // Captures graph_steps_ consecutive steps starting at the step held in
    // it_dev_ (step k of the chunk is *it_dev_ + dir * k) into an executable
    // graph.  An even count leaves the ping-pong buffers of the fused step in
    // their original roles.  The chunk is captured on the plan's stream and
    // holds only kernel launches.
    psvfwi_status capture_steps(psv::StepParams<T> p0, int mode, int dir, State<T>& st, State<T>& alt,
                                const Tables& tab, bool fwd_sig, hipGraphExec_t* out)
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

    bool in_grid(int z, int x) const { return z >= 0 && z < g_.nzt && x >= 0 && x < g_.nxt; }

    // shots of the current launch group [w0_, w1_) still running at step it
    // (slots are sorted by decreasing nt, so they are a prefix of the group)
    int active(int it) const
    {
        int n = 0;
        while (w0_ + n < w1_ && nt_slot_[(size_t)(w0_ + n)] > it) n++;
        return n;
    }

    // Launch groups: the time loops of a batch run group by group, lg_
    // consecutive slots at a time (0: the whole batch), so that one group's
    // wavefields stay in the 256 MB Infinity Cache from one step to the next.
    // The fused step swaps whole buffer sets after each step, so it runs the
    // batch as one group.
    int group_size() const { return (lg_ > 0 && !fused()) ? std::min(lg_, nslots_) : nslots_; }

    void set_group(int w0)
    {
        w0_ = w0;
        w1_ = std::min(nslots_, w0 + group_size());
    }

    void all_slots()
    {
        w0_ = 0;
        w1_ = nslots_;
    }

    // restores the whole batch as the current group when a run ends, also on
    // an early error return
    struct AllSlotsAfter {
        Plan* q;
        ~AllSlotsAfter() { q->all_slots(); }
    };

    // length of the current group's time loop
    int group_nt() const { return nt_slot_[(size_t)w0_]; }

    template <typename U>
    static U* plus(U* q, size_t o)
    {
        return q ? q + o : nullptr;
    }

    // device pointer to the current group's first slot of a per-slot array
    T* atw(DBuf<T>& b, size_t stride) const { return b.p ? b.p + org_ + (size_t)w0_ * stride : nullptr; }
    T* offw(DBuf<T>& b, size_t stride) const { return b.p ? b.p + (size_t)w0_ * stride : nullptr; }

    psvfwi_status ready() const
    {
        if (!have_model_) return fail(PSVFWI_E_STATE, "set_model has not been called");
        if (g_.npml > 0 && !have_pml_) return fail(PSVFWI_E_STATE, "set_pml has not been called");
        if (nslots_ < 1) return fail(PSVFWI_E_STATE, "set_shots has not been called");
        return PSVFWI_OK;
    }

    psvfwi_status download(const T* src, size_t n, double* out) const
    {
        if (n == 0) return PSVFWI_OK;
        std::vector<T> h(n);
        PSV_HIP(hipMemcpyAsync(h.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, stream_));
        PSV_HIP(hipStreamSynchronize(stream_));
        for (size_t i = 0; i < n; i++) out[i] = (double)h[i];
        return PSVFWI_OK;
    }

    // injection groups: entries sorted by (z, x, comp), original order kept
    // inside a group, so co-located sources are added in index order
    psvfwi_status build_injection(Tables& t, bool adjoint_sources, bool with_signals)
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

    psvfwi_status build_receivers(Tables& t, bool none)
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
    psvfwi_status ensure_tile_lists(Tables& t)
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

    psvfwi_status ensure_all_tile_lists()
    {
        PSV_TRY(ensure_tile_lists(fwd_tab_));
        PSV_TRY(ensure_tile_lists(adj_tab_));
        PSV_TRY(ensure_tile_lists(fwd_tab_noreceivers_));
        return PSVFWI_OK;
    }

    psvfwi_status upload_tables()
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

    psvfwi_status alloc_state(State<T>& st)
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

    psvfwi_status zero_state(State<T>& st)
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
    void drop_results()
    {
        ran_ = false;
        fwd_done_ = fwd_grad_done_ = store_loaded_ = adj_grad_done_ = false;
        fwd_disc_ = disc_done_ = fwd_rb_ = false;
    }

    psvfwi_status alloc_store(int ntf)
    {
        sst_ = (size_t)ntf * 5 * nfz_ * nfx_;
        PSV_HIP(store_.reserve((size_t)nslots_ * sst_));
        return PSVFWI_OK;
    }

    // checkpoint c holds, for every slot, the 5 wavefields (current set), the
    // 4 stress-side CPML arrays (current set) and the 4 velocity-side arrays
    size_t ck_bytes_per_slot() const
    {
        const size_t m = g_.npml > 0 ? 4 * (xst_ + zst_) : 0;
        return (5 * fst_ + m) * sizeof(T);
    }

    psvfwi_status alloc_checkpoints(int k)
    {
        ncp_ = (ntmax_ + k - 1) / k;
        PSV_HIP(ckpt_.alloc((size_t)ncp_ * state_elems()));
        return PSVFWI_OK;
    }

    // elements of one saved state of every slot (checkpoint or segment entry)
    size_t state_elems() const { return (size_t)nslots_ * (ck_bytes_per_slot() / sizeof(T)); }

    template <bool SAVE>
    psvfwi_status ck_copy(int c, int na, State<T>& st)
    {
        return state_copy<SAVE>(ckpt_.p + (size_t)c * state_elems(), na, st);
    }

    // copy/restore the first na slots of the current launch group of a saved
    // state at base: each array is [nslots][stride], so a run of slots is one
    // contiguous range
    template <bool SAVE>
    psvfwi_status state_copy(T* base, int na, State<T>& st)
    {
        // one launch for all 13 arrays (sizes and offsets are multiples of 32 elements)
        psv::CopyList l{};
        size_t off = 0;
        auto one = [&](T* arr0, size_t stride) {
            T* ck = base + off + (size_t)w0_ * stride;
            T* arr = arr0 + (size_t)w0_ * stride;
            l.src[l.n] = SAVE ? (const void*)arr : (const void*)ck;
            l.dst[l.n] = SAVE ? (void*)ck : (void*)arr;
            l.bytes[l.n] = (size_t)na * stride * sizeof(T);
            l.n++;
            off += (size_t)nslots_ * stride;
        };
        for (int f = 0; f < 5; f++) one(st.f[f].p, fst_);
        if (g_.npml > 0) {
            one(st.ms[0].p, xst_);
            one(st.ms[1].p, xst_);
            one(st.ms[2].p, zst_);
            one(st.ms[3].p, zst_);
            one(st.mv[0].p, xst_);
            one(st.mv[1].p, xst_);
            one(st.mv[2].p, zst_);
            one(st.mv[3].p, zst_);
        }
        PSV_HIP(psv::launch_copies(l, stream_));
        return PSVFWI_OK;
    }

    psvfwi_status save_checkpoint(int c, int na) { return ck_copy<true>(c, na, main_); }

    psvfwi_status restore_checkpoint(int c, int na) { return ck_copy<false>(c, na, replay_); }

    psvfwi_status adjoint_step(int it, int flags, int tf0)
    {
        psv::StepParams<T> p = base_params(main_, adj_tab_, false);
        p.it = it;
        p.nslots = active(it);
        p.flags = flags;
        p.tf0 = tf0;
        PSV_TRY(launch_timed(p, psv::kAdjoint, 1, main_, main2_));
        return PSVFWI_OK;
    }

    // Launches one step; on every kSampleEvery-th step of a timed loop the
    // two kernels are bracketed by events (kind 0 forward, 1 adjoint,
    // 2 checkpoint replay), read back by collect_samples().
    static constexpr int kSampleEvery = 64;
    static constexpr int kMaxSamples = 512;

    // one step on st: the split kernels update st in place; the fused kernel
    // reads st, writes alt, and the two swap roles (st is current after it)
    psvfwi_status launch_plain(psv::StepParams<T> p, int mode, State<T>& st, State<T>& alt,
                               hipEvent_t e_mid = nullptr)
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
        if (tz_abs() >= 32) PSV_HIP(psv::launch_march<T>(p, g_.fdorder, mode, tz_abs(), stream_));
        else PSV_HIP(psv::launch_fused<T>(p, g_.fdorder, mode, tz_abs(), stream_));
        if (e_mid) PSV_HIP(hipEventRecord(e_mid, stream_));
        for (int f = 0; f < 5; f++) swap_buf(st.f[f], alt.f[f]);
        for (int m = 0; m < 4; m++) swap_buf(st.ms[m], alt.ms[m]);
        return PSVFWI_OK;
    }

    static void swap_buf(DBuf<T>& a, DBuf<T>& b)
    {
        std::swap(a.p, b.p);
        std::swap(a.n, b.n);
    }

    T* at(DBuf<T>& b) const { return b.p ? b.p + org_ : nullptr; }

    psvfwi_status launch_timed(const psv::StepParams<T>& p, int mode, int kind, State<T>& st, State<T>& alt)
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

    psvfwi_status begin_samples()
    {
        if (!sample_ev_[0][0])
            for (int i = 0; i < kMaxSamples; i++)
                for (int j = 0; j < 3; j++) PSV_HIP(hipEventCreate(&sample_ev_[i][j]));
        nsamp_ = 0;
        return PSVFWI_OK;
    }

    // after the stream has drained: mean microseconds of each kernel per kind
    psvfwi_status collect_samples()
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

    psv::StepParams<T> base_params(State<T>& st, const Tables& tab, bool fwd_sig)
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

    psvfwi_grid_desc g_;
    int fpad_ = 1, ppad_ = 0, tp_ = 2, pitch_ = 0, rows_ = 0, org_ = 0, pnx_ = 0, pnz_ = 0, tz_ = 16;
    bool parity_ = false;
    int graph_steps_ = 0;  // > 0: time loops replay captured chunks of this many steps
    int band_ = 0;         // block order of a batch (StepParams::band)
    int lg_ = 0;           // shots per launch group (0: the whole batch)
    int w0_ = 0, w1_ = 0;  // the current launch group: slots [w0_, w1_)
    DBuf<int> it_dev_;
    size_t fst_ = 0, xst_ = 0, zst_ = 0;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;

    DBuf<T> lam_, mu_, mu_zx_, rho_zp_, rho_xp_, pml_;
    bool have_model_ = false, have_pml_ = false;

    bool win_on_ = false;
    psvfwi_window_desc w_{};
    int nfz_ = 0, nfx_ = 0;
    int store_mode_ = PSVFWI_STORE_FULL, ck_k_ = 200, ncp_ = 0;

    std::vector<ShotHost> shots_;
    std::vector<int> shot_of_slot_, slot_of_shot_, nt_slot_;
    int nslots_ = 0, ntmax_ = 0;

    State<T> main_, replay_;
    State<T> main2_, replay2_;  // partners of the fused (ping-pong) step
    DBuf<T> We_, ux_, uz_, store_, grad_, rec_, adj_sig_, ckpt_;
    size_t sst_ = 0, gst_ = 0;
    Tables fwd_tab_, adj_tab_, fwd_tab_noreceivers_;
    DBuf<int> shot_nt_, flag_;
    DBuf<size_t> rec_out_off_, rec_len_;
    std::vector<size_t> rec_out_off_h_;
    DBuf<double> obs_, partial_, misfit_;

    bool ran_ = false;
    std::vector<char> adj_ready_;
    int nexpl_ = 0;
    DBuf<int> expl_grp_;
    int snap_interval_ = 0, snap_field_ = 4;
    bool snaps_taken_ = false;
    DBuf<T> snaps_;
    int nsnap_slot(int s) const { return snap_interval_ > 0 ? (nt_slot_[(size_t)s] - 1) / snap_interval_ + 1 : 0; }
    std::vector<int> fwd_grp_comp_;
    bool fwd_done_ = false, fwd_grad_done_ = false, store_loaded_ = false, adj_grad_done_ = false;
    // discrete gradient: mode, segment length, the forward's checkpoint
    // interval, adjoint state, derivative-term weights, coefficient gradients
    // (lam, mu, mu_zx, rho_zp, rho_xp), segment states, host model copies
    int grad_mode_ = PSVFWI_GRADIENT_REFERENCE, disc_k_ = 64, ck_used_k_ = 200;
    bool fwd_disc_ = false, disc_done_ = false;
    // random-boundary reconstruction: the forward ran for it; the host copy of
    // the source signals and their negation on the device
    bool fwd_rb_ = false;
    std::vector<T> fwd_sig_h_;
    DBuf<T> neg_sig_;
    State<T> adj_;
    DBuf<T> dw_[8], dg_[5], seg_;  // dw_: velocity weights 0-3, stress weights 4-7
    DBuf<int> dgrp_off_, dblk_off_, dblk_idx_;
    int dnblk_ = 0;
    std::vector<double> h_pml_, h_mu_, h_mu_zx_, h_rho_zp_, h_rho_xp_;
    bool have_obs_ = false;
    double last_ms_ = 0;
    long launches_ = 0;
    static constexpr int kNumCand = 8;
    float tune_ms_[kNumCand] = {};
    hipEvent_t sample_ev_[kMaxSamples][3] = {};
    int samp_kind_[kMaxSamples] = {};
    int nsamp_ = 0;
    double kus_[3][2] = {};
    int kn_[3] = {};

public:
    psvfwi_status kernel_timing(int kind, double* us_stress, double* us_velocity, int* samples) const override
    {
        if (kind < 0 || kind > 2) return fail(PSVFWI_E_ARG, "kernel_timing: kind must be 0, 1 or 2");
        if (us_stress) *us_stress = kus_[kind][0];
        if (us_velocity) *us_velocity = kus_[kind][1];
        if (samples) *samples = kn_[kind];
        return PSVFWI_OK;
    }
};

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char* psvfwi_last_error(void) { return g_err.c_str(); }

psvfwi_status psvfwi_plan_create(const psvfwi_grid_desc* g, psvfwi_precision prec, psvfwi_plan** out)
{
    if (!g || !out) return fail(PSVFWI_E_ARG, "plan_create: null argument");
    *out = nullptr;
    if (g->fdorder != 2 && g->fdorder != 4)
        return fail(PSVFWI_E_ORDER, "FD order mismatch: fdorder must be 2 or 4");
    const int fpad = g->fdorder / 2, ppad = g->npml + fpad;
    if (g->npml < 0 || g->nzt < 2 * fpad + 1 || g->nxt < 2 * fpad + 1)
        return fail(PSVFWI_E_ARG, "plan_create: grid smaller than the stencil");
    if (g->npml > 0 && (g->nxt < 2 * ppad || g->nzt < (g->fsurf ? ppad + fpad : 2 * ppad)))
        return fail(PSVFWI_E_ARG, "plan_create: CPML wider than the grid");
    if (!(g->dt > 0) || !(g->dx > 0) || !(g->dz > 0)) return fail(PSVFWI_E_ARG, "plan_create: dt, dx, dz must be > 0");
    if (prec != PSVFWI_FP32 && prec != PSVFWI_FP64 && prec != PSVFWI_FP32X2 && prec != PSVFWI_FP32_FMA)
        return fail(PSVFWI_E_ARG, "plan_create: unknown precision");
    {
        // the kernels address a shot's full-grid arrays with 32-bit byte offsets
        const long long tiles_x = (g->nxt + psv::kTX - 1) / psv::kTX;
        const long long pitch = (psv::kGuardLeft + tiles_x * psv::kTX + psv::kGuardRight + 31) / 32 * 32;
        const long long rows = 2 * psv::kGuardRows + ((long long)g->nzt + psv::kMaxTz - 1) / psv::kMaxTz * psv::kMaxTz;
        const long long bytes = rows * pitch * (prec == PSVFWI_FP32 || prec == PSVFWI_FP32_FMA ? 4 : 8);
        if (bytes >= (1LL << 31)) return fail(PSVFWI_E_ARG, "plan_create: grid too large (a field of one shot must stay below 2 GiB)");
    }
    std::unique_ptr<psvfwi_plan> p;
    if (prec == PSVFWI_FP32) {
        auto q = std::make_unique<Plan<float>>(*g);
        PSV_TRY(q->init());
        p = std::move(q);
    } else if (prec == PSVFWI_FP32_FMA) {
        auto q = std::make_unique<Plan<psv::ffloat>>(*g);
        PSV_TRY(q->init());
        p = std::move(q);
    } else if (prec == PSVFWI_FP32X2) {
        auto q = std::make_unique<Plan<psv::dfloat>>(*g);
        PSV_TRY(q->init());
        p = std::move(q);
    } else {
        auto q = std::make_unique<Plan<double>>(*g);
        PSV_TRY(q->init());
        p = std::move(q);
    }
    *out = p.release();
    return PSVFWI_OK;
}

psvfwi_status psvfwi_plan_destroy(psvfwi_plan* plan)
{
    delete plan;
    return PSVFWI_OK;
}

#define PSV_PLAN(plan) \
    if (!(plan)) return fail(PSVFWI_E_ARG, "null plan")

psvfwi_status psvfwi_set_pml(psvfwi_plan* plan, const double* profiles)
{
    PSV_PLAN(plan);
    return plan->set_pml(profiles);
}

psvfwi_status psvfwi_set_model(psvfwi_plan* plan, const double* lam, const double* mu,
                               const double* mu_zx, const double* rho_zp, const double* rho_xp)
{
    PSV_PLAN(plan);
    return plan->set_model(lam, mu, mu_zx, rho_zp, rho_xp);
}

psvfwi_status psvfwi_set_window(psvfwi_plan* plan, const psvfwi_window_desc* win)
{
    PSV_PLAN(plan);
    return plan->set_window(win);
}

psvfwi_status psvfwi_set_store_mode(psvfwi_plan* plan, int mode, int interval)
{
    PSV_PLAN(plan);
    return plan->set_store_mode(mode, interval);
}

psvfwi_status psvfwi_set_shots(psvfwi_plan* plan, int nshots, const psvfwi_shot_desc* shots)
{
    PSV_PLAN(plan);
    return plan->set_shots(nshots, shots);
}

psvfwi_status psvfwi_set_tile(psvfwi_plan* plan, int tz)
{
    PSV_PLAN(plan);
    return plan->set_tile(tz);
}

psvfwi_status psvfwi_autotune(psvfwi_plan* plan, int steps, int* best_tile)
{
    PSV_PLAN(plan);
    const Range r("psvfwi_autotune");
    return plan->autotune(steps, best_tile);
}

psvfwi_status psvfwi_set_graph_steps(psvfwi_plan* plan, int steps)
{
    if (!plan) return fail(PSVFWI_E_ARG, "null plan");
    return plan->set_graph_steps(steps);
}

psvfwi_status psvfwi_set_shot_order(psvfwi_plan* plan, int band)
{
    PSV_PLAN(plan);
    return plan->set_shot_order(band);
}

psvfwi_status psvfwi_get_shot_order(const psvfwi_plan* plan, int* band)
{
    PSV_PLAN(plan);
    return plan->shot_order(band);
}

psvfwi_status psvfwi_set_launch_group(psvfwi_plan* plan, int shots)
{
    PSV_PLAN(plan);
    return plan->set_launch_group(shots);
}

psvfwi_status psvfwi_get_launch_group(const psvfwi_plan* plan, int* shots)
{
    PSV_PLAN(plan);
    return plan->launch_group(shots);
}

psvfwi_status psvfwi_set_snapshots(psvfwi_plan* plan, int interval, int field)
{
    PSV_PLAN(plan);
    return plan->set_snapshots(interval, field);
}

psvfwi_status psvfwi_snapshot_count(const psvfwi_plan* plan, int shot, int* count)
{
    PSV_PLAN(plan);
    return plan->snapshot_count(shot, count);
}

psvfwi_status psvfwi_get_snapshot(const psvfwi_plan* plan, int shot, int k, double* out)
{
    PSV_PLAN(plan);
    return plan->get_snapshot(shot, k, out);
}

psvfwi_status psvfwi_forward(psvfwi_plan* plan, int flags)
{
    PSV_PLAN(plan);
    const Range r("psvfwi_forward");
    return plan->forward(flags);
}

psvfwi_status psvfwi_set_observed(psvfwi_plan* plan, const double* observed)
{
    PSV_PLAN(plan);
    return plan->set_observed(observed);
}

psvfwi_status psvfwi_misfit(psvfwi_plan* plan, const double* observed, double* misfit_out)
{
    PSV_PLAN(plan);
    const Range r("psvfwi_misfit");
    return plan->misfit(observed, misfit_out);
}

psvfwi_status psvfwi_misfit_noalloc(psvfwi_plan* plan, double* misfit_out)
{
    PSV_PLAN(plan);
    const Range r("psvfwi_misfit_noalloc");
    return plan->misfit_noalloc(misfit_out);
}

psvfwi_status psvfwi_set_adjoint_sources(psvfwi_plan* plan, int shot, const double* signal)
{
    PSV_PLAN(plan);
    return plan->set_adjoint_sources(shot, signal);
}

psvfwi_status psvfwi_set_store(psvfwi_plan* plan, int shot, const double* store)
{
    PSV_PLAN(plan);
    return plan->set_store(shot, store);
}

psvfwi_status psvfwi_adjoint(psvfwi_plan* plan, int flags)
{
    PSV_PLAN(plan);
    const Range r((flags & PSVFWI_GRADIENT) ? "psvfwi_adjoint_gradient" : "psvfwi_adjoint");
    return plan->adjoint(flags);
}

psvfwi_status psvfwi_get_records(const psvfwi_plan* plan, int shot, double* out)
{
    PSV_PLAN(plan);
    return plan->get_records(shot, out);
}

psvfwi_status psvfwi_get_field(const psvfwi_plan* plan, int shot, int field, double* out)
{
    PSV_PLAN(plan);
    return plan->get_field(shot, field, out);
}

psvfwi_status psvfwi_get_store(const psvfwi_plan* plan, int shot, double* out)
{
    PSV_PLAN(plan);
    return plan->get_store(shot, out);
}

psvfwi_status psvfwi_get_gradient(const psvfwi_plan* plan, int shot, double* grad_lam,
                                  double* grad_mu, double* grad_rho)
{
    PSV_PLAN(plan);
    return plan->get_gradient(shot, grad_lam, grad_mu, grad_rho);
}

psvfwi_status psvfwi_set_gradient_mode(psvfwi_plan* plan, int mode, int segment)
{
    PSV_PLAN(plan);
    return plan->set_gradient_mode(mode, segment);
}

psvfwi_status psvfwi_get_coefficient_gradient(const psvfwi_plan* plan, int shot, double* g_lam, double* g_mu,
                                              double* g_mu_zx, double* g_rho_zp, double* g_rho_xp)
{
    PSV_PLAN(plan);
    return plan->get_coefficient_gradient(shot, g_lam, g_mu, g_mu_zx, g_rho_zp, g_rho_xp);
}

psvfwi_status psvfwi_get_model_gradient(const psvfwi_plan* plan, int shot, double* g_lam, double* g_mu,
                                        double* g_rho)
{
    PSV_PLAN(plan);
    return plan->get_model_gradient(shot, g_lam, g_mu, g_rho);
}

psvfwi_status psvfwi_chain_gradient(int nzt, int nxt, const double* mu, const double* mu_zx,
                                    const double* rho_zp, const double* rho_xp, const double* g_lam,
                                    const double* g_mu, const double* g_mu_zx, const double* g_rho_zp,
                                    const double* g_rho_xp, double* out_lam, double* out_mu, double* out_rho)
{
    if (nzt < 2 || nxt < 2) return fail(PSVFWI_E_ARG, "chain_gradient: grid smaller than 2 x 2");
    if (!mu || !mu_zx || !rho_zp || !rho_xp || !g_lam || !g_mu || !g_mu_zx || !g_rho_zp || !g_rho_xp ||
        !out_lam || !out_mu || !out_rho)
        return fail(PSVFWI_E_ARG, "chain_gradient: null array");
    // mu_zx = 4 / sum(1 / mu) over the 4 corners, rho_zp = 2 / (rho(iz) + rho(iz + 1)),
    // rho_xp = 2 / (rho(ix) + rho(ix + 1)) (average_parameters, CUDA_week2/fd_medium.cu:128-159)
    const int ax = nxt - 1;
    for (int iz = 0; iz < nzt; iz++) {
        for (int ix = 0; ix < nxt; ix++) {
            const size_t q = (size_t)iz * nxt + ix;
            double gm = g_mu[q], gr = 0.0;
            // averaged cells with this cell as corner: (iz, ix), (iz, ix-1), (iz-1, ix), (iz-1, ix-1)
            for (int dz = 0; dz <= 1; dz++) {
                for (int dx = 0; dx <= 1; dx++) {
                    const int oz = iz - dz, ox = ix - dx;
                    if (oz < 0 || ox < 0 || oz >= nzt - 1 || ox >= ax) continue;
                    const size_t o = (size_t)oz * ax + ox;
                    const double m = mu_zx[o];
                    if (m != 0.0) gm += g_mu_zx[o] * (m * m / (4.0 * mu[q] * mu[q]));
                    if (dx == 0) gr += g_rho_zp[o] * (-0.5 * rho_zp[o] * rho_zp[o]);
                    if (dz == 0) gr += g_rho_xp[o] * (-0.5 * rho_xp[o] * rho_xp[o]);
                }
            }
            out_lam[q] = g_lam[q];
            out_mu[q] = gm;
            out_rho[q] = gr;
        }
    }
    return PSVFWI_OK;
}

psvfwi_status psvfwi_last_timing(const psvfwi_plan* plan, double* ms, long* launches)
{
    PSV_PLAN(plan);
    return plan->timing(ms, launches);
}

psvfwi_status psvfwi_kernel_timing(const psvfwi_plan* plan, int kind, double* us_stress,
                                   double* us_velocity, int* samples)
{
    PSV_PLAN(plan);
    return plan->kernel_timing(kind, us_stress, us_velocity, samples);
}

psvfwi_status psvfwi_parabola_step(const double* L2, const double* step, double* opt)
{
    if (!L2 || !step || !opt) return fail(PSVFWI_E_ARG, "parabola_step: null argument");
    // Newton divided differences of the interpolating parabola, FP64
    const double d01 = (L2[1] - L2[0]) / (step[1] - step[0]);
    const double d02 = (L2[2] - L2[0]) / (step[2] - step[0]);
    const double x2 = (d02 - d01) / (step[2] - step[1]);
    const double x1 = d01 - x2 * (step[0] + step[1]);
    *opt = -x1 / (2.0 * x2);
    if (!std::isfinite(*opt)) return fail(PSVFWI_E_NONFINITE, "parabola_step: degenerate parabola");
    return PSVFWI_OK;
}

}  // extern "C"
