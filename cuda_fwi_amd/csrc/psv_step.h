// psv_step.h -- launch parameters of the fused P-SV time-step kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "psv_dfloat.h"

namespace psv {

constexpr int kTX = 64;        // tile width: one wave64 row
constexpr int kThreads = 256;  // four waves per workgroup
constexpr int kNumXcd = 8;     // MI355X: 8 XCDs, 32 CUs each

enum { kForward = 0, kAdjoint = 1 };
enum { kStore = 1, kEnergy = 2 };

// All full-grid arrays share one row pitch (elements) and are addressed as
// base + slot * stride + iz * pitch + ix.  Averaged materials use the same
// pitch (their last row and column are unused).  A step is two launches
// (stress, then velocity) that update the state in place.
template <typename T>
struct StepParams {
    int nzt, nxt, pitch;
    int ppad, npml, tp, fsurf, pnx, pnz;
    int tiles_x, tiles_z, nslots;
    int band;            // block order: 0 shot-major (slot, tile); k > 0 bands of k tile rows,
                         // within a band the shots one after another (band, slot, tile)
    int it;              // time index (plus *it_dev when it_dev != nullptr)
    const int* it_dev;   // device-side time counter for graph replays
    int flags;           // kStore | kEnergy
    T dt, dxi, dzi, hc1, hc2;

    // model, shared by every shot of the batch
    const T *lam, *mu, *mu_zx, *rho_zp, *rho_xp;
    const T* pml;  // a, b, K, a_half, b_half, K_half, each tp long

    // wavefields; per-shot stride fst.  The split kernels update them in
    // place; the fused kernel reads these and writes the *_w set (ping-pong).
    T *vx, *vz, *sxx, *szx, *szz;
    T *vx_w, *vz_w, *sxx_w, *szx_w, *szz_w;
    size_t fst;
    // Full-grid arrays (fields, model, We, ux, uz) are guard-padded: every
    // pointer above points at cell (0, 0), which sits org elements into its
    // slot's allocation, with at least 4 guard rows above and below the
    // grid's tiles and 32 guard columns on the left and 4 on the right, all
    // zero.  Halo loads up to 4 cells outside the grid stay inside the slot.
    int org;
    T* We;         // energy weight, stride fst
    T *ux, *uz;    // adjoint displacement, stride fst

    // CPML memory: x-type [nzt][tp] (stride xst), z-type [tp][pitch] (stride zst)
    T *mvxx, *mvzx, *mvxz, *mvzz;      // stress side: mem_vx_x, mem_vz_x, mem_vx_z, mem_vz_z
    T *mvxx_w, *mvzx_w, *mvxz_w, *mvzz_w;  // fused kernel: stress-side memory written here
    T *msxxx, *mszxx, *mszzz, *mszxz;  // velocity side: mem_sxx_x, mem_szx_x, mem_szz_z, mem_szx_z
    size_t xst, zst;

    // forward store [ntf][5][nfz][nfx] per shot (stride sst), first tf held = tf0
    T* store;
    size_t sst;
    int tf0;
    int fwi_dt, fwi_dx, fwi_dz, fwi_x1, fwi_x2, fwi_z1, fwi_z2, nfz, nfx;
    T* grad;       // [3][nfz][nfx] per shot (stride gst): lambda, mu, rho
    size_t gst;

    // per-slot tables
    const int* shot_nt;      // [nslots]
    const int* inj_grp;      // 5 ints per group: z, x, comp, first entry, count
    const int* inj_row;      // signal row (within the shot) of each entry
    const T* sig;            // signals, shot s at sig_off[s], row-major [row][nt_s]
    const size_t* sig_off;
    const int* rec_zxc;      // 3 ints per receiver: z, x, comp
    const int* rec_first;    // [nslots] index of the slot's first receiver in rec_zxc
    T* rec;                  // traces, shot s at rec_out_off[s], [nrec_s][nt_s]
    const size_t* rec_out_off;
    // per-tile lists (CSR over tiles, slot-major: entry slot * ntile + tile)
    const int* tinj_off;     // [nslots * ntile + 1] into tinj_idx
    const int* tinj_idx;     // injection group ids
    const int* trec_off;     // [nslots * ntile + 1] into trec_idx
    const int* trec_idx;     // receiver ids (global index into rec_zxc)
    // explosive (stress) source groups: pairs (group id, slot of the whole
    // batch); slot0 is the batch slot of this launch's slot 0
    int nexpl;
    const int* expl_grp;
    int slot0;
};

// Launches one step (stress kernel, velocity kernel, and the stress-source
// kernel when the batch has explosive sources).  e_mid, when not null, is
// recorded between the stress and the velocity kernel.
template <typename T>
hipError_t launch_step(const StepParams<T>& p, int order, int mode, int tz, hipStream_t s,
                       hipEvent_t e_mid = nullptr);

// One kernel of a step: which 0 the stress kernel, 1 the velocity kernel
// (with its point sources and receivers).  The reverse step of the
// random-boundary reconstruction runs them in the opposite order with a
// negated dt and negated source signals.
template <typename T>
hipError_t launch_half(const StepParams<T>& p, int order, int mode, int tz, int which, hipStream_t s);

// The FWI window cells of the active shots' current state into store plane
// 0 (v / dt and the stresses, as the forward stress kernel stores them).
template <typename T>
hipError_t launch_store_plane(const StepParams<T>& p, hipStream_t s);

// The explosive (sxx + szz) sources of the active shots, added after a step.
template <typename T>
hipError_t launch_stress_sources(const StepParams<T>& p, hipStream_t s);

// Guard layout of the full-grid arrays (see StepParams::org).
constexpr int kGuardRows = 4;
constexpr int kGuardLeft = 32;
constexpr int kGuardRight = 4;
constexpr int kMaxTz = 32;

// One step as a single launch (fused stress + velocity, halo ring of the
// stresses recomputed), reading the state through vx.. and writing it
// through vx_w..; the explosive-source launch follows when needed.  tz is 8
// or 16.  Returns hipErrorInvalidValue for other tile heights.
template <typename T>
hipError_t launch_fused(const StepParams<T>& p, int order, int mode, int tz, hipStream_t s);

// One step as a single launch in which each workgroup marches down a segment
// of 64 columns x `rows` rows (32 or 64), keeping the rows it needs in LDS
// rings, followed by the point-source / receiver launch (and the explosive
// one when needed); reads vx.. and writes vx_w.. like launch_fused.  Returns
// hipErrorInvalidValue for other segment heights.
// The edge segments of a z-marching step run on a second stream, concurrently with
// the interior ones: `fork` is recorded on the step's stream and waited for on
// `stream`, `join` the other way round, before the point-source launch.
struct MarchSide {
    hipStream_t stream;
    hipEvent_t fork, join;
};

template <typename T>
hipError_t launch_march(const StepParams<T>& p, int order, int mode, int rows, hipStream_t s,
                        const MarchSide* side = nullptr);

}  // namespace psv
