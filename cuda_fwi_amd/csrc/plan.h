// plan.h -- the plan behind the C ABI (psvfwi.h): device memory and time loops
// of a batch of shots on one model.
//
// A plan owns one HIP stream and every device buffer of a batch.  Shots are
// kept in "slots" sorted by record length (longest first), so the shots still
// running at step it are always a prefix of the slots and one launch of the
// step kernel covers exactly them.
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
//
// The member functions of Plan<T> are defined by topic in plan_setup.cpp,
// plan_shots.cpp, plan_step.cpp, plan_forward.cpp, plan_adjoint.cpp,
// plan_discrete.cpp and plan_results.cpp; each file instantiates its own for
// the four arithmetic types (PSV_FOR_EACH_PRECISION).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "psv_aux.h"
#include "psv_discrete.h"
#include "psv_step.h"
#include "psvfwi.h"

// FP32, FP64, compensated FP32 (PSVFWI_FP32X2), FP32 with FMA contraction (PSVFWI_FP32_FMA)
#define PSV_FOR_EACH_PRECISION(X) X(float) X(double) X(psv::dfloat) X(psv::ffloat)

// The abstract plan the C ABI's opaque handle points to; Plan<T> implements it
// for one arithmetic type.
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

namespace psvplan {

// records msg as the calling thread's last error (psvfwi_last_error) and returns st
psvfwi_status fail(psvfwi_status st, const std::string& msg);

#define PSV_HIP(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::psvplan::fail(e_ == hipErrorOutOfMemory ? PSVFWI_E_NOMEM : PSVFWI_E_HIP, \
                                   std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

#define PSV_TRY(expr)                     \
    do {                                  \
        psvfwi_status s_ = (expr);        \
        if (s_ != PSVFWI_OK) return s_;   \
    } while (0)

// A device array owned by the plan.
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

// one shot as the caller described it (grid indices, components, signals)
struct ShotHost {
    int nt = 0;
    std::vector<int> sx, sz, sc;
    std::vector<double> ssig;  // [nsrc][nt]
    std::vector<int> rx, rz, rc;
};

inline size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

enum { kF_VX, kF_VZ, kF_SXX, kF_SZX, kF_SZZ };
// stress-side CPML memory: 0 mem_vx_x, 1 mem_vz_x (x-type), 2 mem_vx_z, 3 mem_vz_z (z-type)
// velocity-side: 0 mem_sxx_x, 1 mem_szx_x (x-type), 2 mem_szz_z, 3 mem_szx_z (z-type)

// the wavefields and CPML memory of every slot of a batch
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

    // ---------------------------------------------------------- plan_setup.cpp
    explicit Plan(const psvfwi_grid_desc& g);
    ~Plan() override;
    psvfwi_status init();
    psvfwi_status set_pml(const double* prof) override;
    psvfwi_status set_model(const double* lam, const double* mu, const double* mu_zx, const double* rho_zp,
                            const double* rho_xp) override;
    psvfwi_status set_window(const psvfwi_window_desc* w) override;
    psvfwi_status set_store_mode(int mode, int interval) override;
    psvfwi_status set_gradient_mode(int mode, int interval) override;
    psvfwi_status set_snapshots(int interval, int field) override;
    psvfwi_status set_tile(int tz) override;
    psvfwi_status set_graph_steps(int n) override;
    psvfwi_status set_shot_order(int band) override;
    psvfwi_status shot_order(int* band) const override;
    psvfwi_status set_launch_group(int shots) override;
    psvfwi_status launch_group(int* shots) const override;

    // ---------------------------------------------------------- plan_shots.cpp
    psvfwi_status set_shots(int n, const psvfwi_shot_desc* sd) override;

    // ----------------------------------------------------------- plan_step.cpp
    psvfwi_status autotune(int steps, int* best) override;
    psvfwi_status timing(double* ms, long* launches) const override;
    psvfwi_status kernel_timing(int kind, double* us_stress, double* us_velocity, int* samples) const override;

    // -------------------------------------------------------- plan_forward.cpp
    psvfwi_status forward(int flags) override;
    psvfwi_status set_observed(const double* obs) override;
    psvfwi_status misfit(const double* obs, double* out) override;
    psvfwi_status misfit_noalloc(double* out) override;

    // -------------------------------------------------------- plan_adjoint.cpp
    psvfwi_status set_adjoint_sources(int shot, const double* sig) override;
    psvfwi_status set_store(int shot, const double* st) override;
    psvfwi_status adjoint(int flags) override;

    // ------------------------------------------------------- plan_discrete.cpp
    psvfwi_status get_coefficient_gradient(int shot, double* gl, double* gm, double* gmzx, double* grzp,
                                           double* grxp) const override;
    psvfwi_status get_model_gradient(int shot, double* gl, double* gm, double* gr) const override;

    // -------------------------------------------------------- plan_results.cpp
    psvfwi_status get_snapshot(int shot, int k, double* out) const override;
    psvfwi_status snapshot_count(int shot, int* n) const override;
    psvfwi_status get_records(int shot, double* out) const override;
    psvfwi_status get_field(int shot, int field, double* out) const override;
    psvfwi_status get_store(int shot, double* out) const override;
    psvfwi_status get_gradient(int shot, double* gl, double* gm, double* gr) const override;

private:
    // source injection groups, receivers and their per-tile lists for one kind of step
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

    // device pointers into one saved state (state_copy layout)
    struct View {
        const T* f[5];
        const T* ms[4];
        const T* mv[4];
    };

    // ---------------------------------------------------------- plan_setup.cpp
    psvfwi_status ready() const;
    psvfwi_status alloc_state(State<T>& st);
    psvfwi_status zero_state(State<T>& st);
    void drop_results();
    psvfwi_status alloc_store(int ntf);
    size_t ck_bytes_per_slot() const;
    // k = 0: one entry per step plus the final state (every state kept)
    psvfwi_status alloc_checkpoints(int k);

    // ---------------------------------------------------------- plan_shots.cpp
    psvfwi_status build_injection(Tables& t, bool adjoint_sources, bool with_signals);
    psvfwi_status build_receivers(Tables& t, bool none);
    psvfwi_status ensure_tile_lists(Tables& t);
    psvfwi_status ensure_all_tile_lists();
    psvfwi_status upload_tables();

    // ----------------------------------------------------------- plan_step.cpp
    psv::StepParams<T> base_params(State<T>& st, const Tables& tab, bool fwd_sig);
    psvfwi_status launch_plain(psv::StepParams<T> p, int mode, State<T>& st, State<T>& alt,
                               hipEvent_t e_mid = nullptr);
    psvfwi_status launch_timed(const psv::StepParams<T>& p, int mode, int kind, State<T>& st, State<T>& alt);
    psvfwi_status begin_samples();
    psvfwi_status collect_samples();
    psvfwi_status capture_steps(psv::StepParams<T> p0, int mode, int dir, State<T>& st, State<T>& alt,
                                const Tables& tab, bool fwd_sig, hipGraphExec_t* out);

    // -------------------------------------------------------- plan_adjoint.cpp
    psvfwi_status adjoint_step(int it, int flags, int tf0);
    psvfwi_status rb_check() const;
    psvfwi_status adjoint_rb(int flags);

    // ------------------------------------------------------- plan_discrete.cpp
    psvfwi_status adjoint_discrete();
    View entry(const T* base, int e) const;
    psv::DiscreteParams<T> dparams(int it, const View& pre, const View& post);

    // -------------------------------------------------------- plan_results.cpp
    psvfwi_status download(const T* src, size_t n, double* out) const;

    // ------------------------------------------------------------ inline helpers
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
    T* at(DBuf<T>& b) const { return b.p ? b.p + org_ : nullptr; }

    static void swap_buf(DBuf<T>& a, DBuf<T>& b)
    {
        std::swap(a.p, b.p);
        std::swap(a.n, b.n);
    }

    // elements of one saved state of every slot (checkpoint or segment entry)
    size_t state_elems() const { return (size_t)nslots_ * (ck_bytes_per_slot() / sizeof(T)); }

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

    template <bool SAVE>
    psvfwi_status ck_copy(int c, int na, State<T>& st)
    {
        return state_copy<SAVE>(ckpt_.p + (size_t)c * state_elems(), na, st);
    }

    psvfwi_status save_checkpoint(int c, int na) { return ck_copy<true>(c, na, main_); }
    psvfwi_status restore_checkpoint(int c, int na) { return ck_copy<false>(c, na, replay_); }

    int nsnap_slot(int s) const { return snap_interval_ > 0 ? (nt_slot_[(size_t)s] - 1) / snap_interval_ + 1 : 0; }

    // ------------------------------------------------------------------- state
    // grid layout (guard-padded rows, psv_step.h)
    psvfwi_grid_desc g_;
    int fpad_ = 1, ppad_ = 0, tp_ = 2, pitch_ = 0, rows_ = 0, org_ = 0, pnx_ = 0, pnz_ = 0, tz_ = 16;
    size_t fst_ = 0, xst_ = 0, zst_ = 0;
    bool parity_ = false;
    int graph_steps_ = 0;  // > 0: time loops replay captured chunks of this many steps
    int band_ = 0;         // block order of a batch (StepParams::band)
    int lg_ = 0;           // shots per launch group (0: the whole batch)
    int w0_ = 0, w1_ = 0;  // the current launch group: slots [w0_, w1_)
    DBuf<int> it_dev_;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    // the z-marching step's edge segments run on a second stream (PSVFWI_MARCH_STREAMS=1: on
    // the plan's stream after the interior ones)
    psv::MarchSide side_{};
    bool use_side_ = true;

    // model and CPML profiles
    DBuf<T> lam_, mu_, mu_zx_, rho_zp_, rho_xp_, pml_;
    bool have_model_ = false, have_pml_ = false;

    // gradient window and store mode
    bool win_on_ = false;
    psvfwi_window_desc w_{};
    int nfz_ = 0, nfx_ = 0;
    int store_mode_ = PSVFWI_STORE_FULL, ck_k_ = 200, ncp_ = 0;

    // shots and slots
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

    // timing: the last run, and kernel samples every kSampleEvery-th launch
    // (kind 0 forward, 1 adjoint, 2 checkpoint replay)
    static constexpr int kSampleEvery = 64;
    static constexpr int kMaxSamples = 512;
    double last_ms_ = 0;
    long launches_ = 0;
    static constexpr int kNumCand = 8;
    float tune_ms_[kNumCand] = {};
    hipEvent_t sample_ev_[kMaxSamples][3] = {};
    int samp_kind_[kMaxSamples] = {};
    int nsamp_ = 0;
    double kus_[3][2] = {};
    int kn_[3] = {};
};

}  // namespace psvplan
