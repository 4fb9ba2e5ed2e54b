// psv_discrete.h -- launch parameters of the discrete-adjoint step kernels.
//
// The discrete adjoint is the exact gradient of the misfit
// sum 0.5 dt (d - d_obs)^2 of this scheme with respect to the coefficient
// arrays lam, mu, mu_zx, rho_zp, rho_xp (an extension: the reference's
// gradient is the correlation of adjoint_kernel_PSV_GPU.cu:830-863).  One
// reverse time step is: receiver adjoint sources, the transposed velocity
// half-step (weights, then gather into the stress adjoints), the transposed
// stress mirror, the transposed stress half-step (weights, then gather into
// the velocity adjoints).  The forward quantities each step needs are
// recomputed from the forward states before and after the step.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "psv_dfloat.h"

namespace psv {

// workgroup shape of the reverse-step kernels: 64 x 4 cells
constexpr int kDiscreteBX = 64, kDiscreteBZ = 4;

template <typename T>
struct DiscreteParams {
    int nzt, nxt, pitch;
    int ppad, npml, tp, fsurf, pnx, pnz, nslots, it;
    T dt, dxi, dzi, hc1, hc2;
    // model and CPML profile (full-grid arrays guard-padded as in StepParams)
    const T *lam, *mu, *mu_zx, *rho_zp, *rho_xp;
    const T* pml;
    size_t fst, xst, zst;  // per-slot strides: full-grid arrays, x-type and z-type CPML arrays
    // forward state before the step: velocities and both CPML sets
    const T *pre_vx, *pre_vz;
    const T *pre_mvxx, *pre_mvzx, *pre_mvxz, *pre_mvzz;
    const T *pre_msxxx, *pre_mszxx, *pre_mszzz, *pre_mszxz;
    // forward state after the step: its stresses
    const T *post_sxx, *post_szx, *post_szz;
    // adjoint state: fields and the adjoints of the CPML memories
    T *vbx, *vbz, *sbxx, *sbzx, *sbzz;
    T *bvxx, *bvzx, *bvxz, *bvzz;
    T *bsxxx, *bszxx, *bszzz, *bszxz;
    // weights of the derivative terms (full-grid layout, zero outside the
    // interior): w* of the velocity half-step, ws* of the stress half-step
    // (separate arrays: the stress pass reads the velocity weights of
    // neighbouring cells while it writes its own cell's stress weights)
    T *w0, *w1, *w2, *w3;
    T *ws0, *ws1, *ws2, *ws3;
    // coefficient gradients (full-grid layout)
    T *g_lam, *g_mu, *g_mu_zx, *g_rho_zp, *g_rho_xp;
    // receiver adjoint sources: groups of co-located receivers per slot
    const int* grp;      // 5 ints per group: z, x, comp, first entry, count
    const int* grp_off;  // [nslots + 1]
    const int* grp_row;  // receiver row (within the shot) of each entry
    const T* sig;        // residuals, slot s at sig_off[s], [nrec_s][nt_s]
    const size_t* sig_off;
    const int* shot_nt;
    int max_groups;
    // receiver adjoint sources added by the receiver kernel at p.it: slots
    // [inj_lo, inj_hi) (the others get theirs from the previous launch's
    // stress pass)
    int inj_lo, inj_hi;
    // velocity pass: first add the velocity adjoints gathered from the stress
    // weights of the previous reverse step (0 for a batch's first step)
    int gather_prev;
    // stress pass: also add the receiver adjoint sources of time index
    // inj_next (the next reverse step; -1: none), through the receiver
    // groups of each 64 x 4 workgroup (CSR over slot * nblk + block)
    int inj_next, nblk;
    const int* blk_off;
    const int* blk_idx;
};

// One reverse time step of the active slots (p.nslots) at time index p.it:
// the receiver kernel (slots [inj_lo, inj_hi)), the velocity pass (gather of
// the previous step's stress weights, then the velocity weights) and the
// stress pass (gather, transposed mirror, stress weights, and the next step's
// receiver sources).  The velocity adjoints gathered from the last step's
// stress weights are never added: nothing reads them.
template <typename T>
hipError_t launch_discrete_step(const DiscreteParams<T>& p, int order, hipStream_t s);

}  // namespace psv
