// psv_discrete.hip -- discrete-adjoint step kernels (gfx950).
//
// The transpose of one forward step (psv_step.hip, the reference's
// forward_kernel_PSV_GPU.cu:830-1063) applied in reverse order:
//   receivers -> velocity half-step -> stress mirror -> stress half-step.
// Each half-step has a weight pass, one thread per interior cell, that
// recomputes the forward derivative terms of the cell, adds the cell's
// gradient contributions and forms the adjoint weights of the four
// derivative terms (through the transposed CPML recursions), and a gather
// pass that applies the transposed difference stencils to those weights.
// Splitting at the weights keeps every adjoint update a gather (no
// scatter, no atomics) and the order of every sum fixed.  A weight pass
// needs the adjoints of its own cell only, so each gather runs in the
// launch of the next weight pass:
//   velocity pass: velocity adjoints += gather of the previous reverse
//     step's stress weights (after that step's receiver sources), then the
//     velocity weights;
//   stress pass: stress adjoints += gather of the velocity weights, the
//     transposed mirror, the stress weights, and the next reverse step's
//     receiver sources.
// A reverse step is these two launches, plus a receiver launch for the
// shots whose first reverse step it is.
//
// The per-cell expressions are fixed: the CPU restatement evaluates the
// same ones in the same order.
#include <hip/hip_runtime.h>

#include "psv_discrete.h"

namespace psv {

namespace {

constexpr int kDX = kDiscreteBX, kDZ = kDiscreteBZ;

template <typename T>
struct Cell {
    const T* b;
    int pitch;
    __device__ T operator()(int iz, int ix) const { return b[(long)iz * pitch + ix]; }
};

// transposed "psi = b psi + a t; t' = t / K + psi" for the weight w of t'
template <typename T>
__device__ __forceinline__ T cpml_t(T w, T* psib, T a, T b, T K)
{
    const T tot = *psib + w;
    const T wn = w / K + a * tot;
    if (a != (T)0) *psib = b * tot;  // a == 0: psi stays zero, its adjoint reaches nothing
    return wn;
}

// the receiver adjoint sources of one group of co-located receivers, time index it
template <typename T>
__device__ __forceinline__ void add_receiver_group(const DiscreteParams<T>& p, int slot, int gid, int it)
{
    const int* e = p.grp + 5 * (size_t)gid;
    const int comp = e[2];
    if (comp != 1 && comp != 2) return;
    const int nt_s = p.shot_nt[slot];
    T* v = (comp == 1 ? p.vbx : p.vbz) + (size_t)slot * p.fst + (long)e[0] * p.pitch + e[1];
    const T* sig = p.sig + p.sig_off[slot] + it;
    T acc = *v;
    for (int q = 0; q < e[4]; q++) acc = acc + p.dt * sig[(size_t)p.grp_row[e[3] + q] * nt_s];
    *v = acc;
}

template <typename T>
__global__ void d_inject_kernel(const DiscreteParams<T> p)
{
    const int slot = p.inj_lo + (int)blockIdx.y;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int g0 = p.grp_off[slot], g1 = p.grp_off[slot + 1];
    if (g0 + g >= g1) return;
    add_receiver_group(p, slot, g0 + g, p.it);
}

// increments of the velocity adjoints at offset k: the transposed velocity
// differences of the stress half-step's weights (interior cells: the
// velocities elsewhere never change)
template <typename T, int ORDER>
__device__ __forceinline__ void stress_gather_inc(const DiscreteParams<T>& p, size_t k, T& ix_, T& iz_)
{
    const Cell<T> W0{p.ws0 + k, p.pitch}, W1{p.ws1 + k, p.pitch}, W2{p.ws2 + k, p.pitch},
        W3{p.ws3 + k, p.pitch};
    const T dxi = p.dxi, dzi = p.dzi, hc1 = p.hc1, hc2 = p.hc2;
    if (ORDER == 2) {
        const T xx = dxi * hc1 * (W0(0, 0) - W0(0, 1));
        const T xz = dzi * hc1 * (W2(-1, 0) - W2(0, 0));
        ix_ = xx + xz;
        const T zx = dxi * hc1 * (W1(0, -1) - W1(0, 0));
        const T zz = dzi * hc1 * (W3(0, 0) - W3(1, 0));
        iz_ = zx + zz;
    } else {
        const T xx = dxi * (hc1 * (W0(0, 0) - W0(0, 1)) + hc2 * (W0(0, -1) - W0(0, 2)));
        const T xz = dzi * (hc1 * (W2(-1, 0) - W2(0, 0)) + hc2 * (W2(-2, 0) - W2(1, 0)));
        ix_ = xx + xz;
        const T zx = dxi * (hc1 * (W1(0, -1) - W1(0, 0)) + hc2 * (W1(0, -2) - W1(0, 1)));
        const T zz = dzi * (hc1 * (W3(0, 0) - W3(1, 0)) + hc2 * (W3(-1, 0) - W3(2, 0)));
        iz_ = zx + zz;
    }
}

// The velocity side of a reverse step: the velocity adjoints gathered from
// the previous reverse step's stress weights (when p.gather_prev), then the
// transposed velocity half-step: weights and the rho_xp / rho_zp gradients
template <typename T, int ORDER>
__global__ __launch_bounds__(kDX* kDZ) void d_vel_pass_kernel(const DiscreteParams<T> p)
{
    constexpr int H = ORDER / 2;
    const int ix = blockIdx.x * kDX + threadIdx.x, iz = blockIdx.y * kDZ + threadIdx.y;
    const int slot = blockIdx.z;
    const int nzt = p.nzt, nxt = p.nxt, ppad = p.ppad;
    if (iz < H || iz >= nzt - H || ix < H || ix >= nxt - H) return;
    const size_t so = (size_t)slot * p.fst;
    const long c = (long)iz * p.pitch + ix;
    const Cell<T> sxx{p.post_sxx + so + c, p.pitch}, szx{p.post_szx + so + c, p.pitch},
        szz{p.post_szz + so + c, p.pitch};
    const T dxi = p.dxi, dzi = p.dzi, hc1 = p.hc1, hc2 = p.hc2, dt = p.dt;
    T sxx_x, szx_x, szx_z, szz_z;
    if (ORDER == 2) {
        sxx_x = dxi * hc1 * (sxx(0, 1) - sxx(0, 0));
        szx_x = dxi * hc1 * (szx(0, 0) - szx(0, -1));
        szx_z = dzi * hc1 * (szx(0, 0) - szx(-1, 0));
        szz_z = dzi * hc1 * (szz(1, 0) - szz(0, 0));
    } else {
        sxx_x = dxi * (hc1 * (sxx(0, 1) - sxx(0, 0)) + hc2 * (sxx(0, 2) - sxx(0, -1)));
        szx_x = dxi * (hc1 * (szx(0, 0) - szx(0, -1)) + hc2 * (szx(0, 1) - szx(0, -2)));
        szx_z = dzi * (hc1 * (szx(0, 0) - szx(-1, 0)) + hc2 * (szx(1, 0) - szx(-2, 0)));
        szz_z = dzi * (hc1 * (szz(1, 0) - szz(0, 0)) + hc2 * (szz(2, 0) - szz(-1, 0)));
    }
    const int tp = p.tp;
    const T* pa = p.pml;
    const T* pb = p.pml + tp;
    const T* pK = p.pml + 2 * tp;
    const T* pah = p.pml + 3 * tp;
    const T* pbh = p.pml + 4 * tp;
    const T* pKh = p.pml + 5 * tp;
    const bool fsurf = p.fsurf != 0;
    const size_t xo = (size_t)slot * p.xst, zo = (size_t)slot * p.zst;
    if (p.npml > 0) {
        // forward derivative terms after the velocity-side CPML (psi before the step)
        if (ix < ppad) {
            const int px = ix - H;
            const size_t q = xo + (size_t)iz * tp + px;
            const T m1 = pb[px] * p.pre_msxxx[q] + pa[px] * sxx_x;
            const T m2 = pbh[px] * p.pre_mszxx[q] + pah[px] * szx_x;
            sxx_x = sxx_x / pK[px] + m1;
            szx_x = szx_x / pKh[px] + m2;
        }
        if (ix >= nxt - ppad - 1) {
            const int px = ix - p.pnx;
            const size_t q = xo + (size_t)iz * tp + px;
            const T m1 = pb[px] * p.pre_msxxx[q] + pa[px] * sxx_x;
            const T m2 = pbh[px] * p.pre_mszxx[q] + pah[px] * szx_x;
            sxx_x = sxx_x / pK[px] + m1;
            szx_x = szx_x / pKh[px] + m2;
        }
        if (iz < ppad && !fsurf) {
            const int pz = iz - H;
            const size_t q = zo + (size_t)pz * p.pitch + ix;
            const T m1 = pb[pz] * p.pre_mszzz[q] + pa[pz] * szz_z;
            const T m2 = pbh[pz] * p.pre_mszxz[q] + pah[pz] * szx_z;
            szz_z = szz_z / pK[pz] + m1;
            szx_z = szx_z / pKh[pz] + m2;
        }
        if (iz >= nzt - ppad - 1) {
            const int pz = iz - p.pnz;
            const size_t q = zo + (size_t)pz * p.pitch + ix;
            const T m1 = pb[pz] * p.pre_mszzz[q] + pa[pz] * szz_z;
            const T m2 = pbh[pz] * p.pre_mszxz[q] + pah[pz] * szx_z;
            szz_z = szz_z / pK[pz] + m1;
            szx_z = szx_z / pKh[pz] + m2;
        }
    }
    const size_t k = so + c;
    T bx = p.vbx[k], bz = p.vbz[k];
    if (p.gather_prev) {
        T gx, gz;
        stress_gather_inc<T, ORDER>(p, k, gx, gz);
        bx = bx + gx;
        bz = bz + gz;
        p.vbx[k] = bx;
        p.vbz[k] = bz;
    }
    p.g_rho_xp[k] = p.g_rho_xp[k] + bx * (dt * (sxx_x + szx_z));
    p.g_rho_zp[k] = p.g_rho_zp[k] + bz * (dt * (szx_x + szz_z));
    const T ax = dt * p.rho_xp[c] * bx;
    const T az = dt * p.rho_zp[c] * bz;
    T w_sxx_x = ax, w_szx_z = ax, w_szx_x = az, w_szz_z = az;
    if (p.npml > 0) {
        // transposed recursions, in the reverse of the forward order
        if (iz >= nzt - ppad - 1) {
            const int pz = iz - p.pnz;
            const size_t q = zo + (size_t)pz * p.pitch + ix;
            w_szz_z = cpml_t(w_szz_z, &p.bszzz[q], pa[pz], pb[pz], pK[pz]);
            w_szx_z = cpml_t(w_szx_z, &p.bszxz[q], pah[pz], pbh[pz], pKh[pz]);
        }
        if (iz < ppad && !fsurf) {
            const int pz = iz - H;
            const size_t q = zo + (size_t)pz * p.pitch + ix;
            w_szz_z = cpml_t(w_szz_z, &p.bszzz[q], pa[pz], pb[pz], pK[pz]);
            w_szx_z = cpml_t(w_szx_z, &p.bszxz[q], pah[pz], pbh[pz], pKh[pz]);
        }
        if (ix >= nxt - ppad - 1) {
            const int px = ix - p.pnx;
            const size_t q = xo + (size_t)iz * tp + px;
            w_sxx_x = cpml_t(w_sxx_x, &p.bsxxx[q], pa[px], pb[px], pK[px]);
            w_szx_x = cpml_t(w_szx_x, &p.bszxx[q], pah[px], pbh[px], pKh[px]);
        }
        if (ix < ppad) {
            const int px = ix - H;
            const size_t q = xo + (size_t)iz * tp + px;
            w_sxx_x = cpml_t(w_sxx_x, &p.bsxxx[q], pa[px], pb[px], pK[px]);
            w_szx_x = cpml_t(w_szx_x, &p.bszxx[q], pah[px], pbh[px], pKh[px]);
        }
    }
    p.w0[k] = w_sxx_x;
    p.w1[k] = w_szx_x;
    p.w2[k] = w_szx_z;
    p.w3[k] = w_szz_z;
}

// increments of the stress adjoints at offset k: the transposed stress
// differences of the velocity half-step's weights
template <typename T, int ORDER>
__device__ __forceinline__ void vel_gather_inc(const DiscreteParams<T>& p, size_t k, T& ixx, T& izx, T& izz)
{
    const Cell<T> W0{p.w0 + k, p.pitch}, W1{p.w1 + k, p.pitch}, W2{p.w2 + k, p.pitch},
        W3{p.w3 + k, p.pitch};
    const T dxi = p.dxi, dzi = p.dzi, hc1 = p.hc1, hc2 = p.hc2;
    if (ORDER == 2) {
        ixx = dxi * hc1 * (W0(0, -1) - W0(0, 0));
        const T tx = dxi * hc1 * (W1(0, 0) - W1(0, 1));
        const T tz = dzi * hc1 * (W2(0, 0) - W2(1, 0));
        izx = tx + tz;
        izz = dzi * hc1 * (W3(-1, 0) - W3(0, 0));
    } else {
        ixx = dxi * (hc1 * (W0(0, -1) - W0(0, 0)) + hc2 * (W0(0, -2) - W0(0, 1)));
        const T tx = dxi * (hc1 * (W1(0, 0) - W1(0, 1)) + hc2 * (W1(0, -1) - W1(0, 2)));
        const T tz = dzi * (hc1 * (W2(0, 0) - W2(1, 0)) + hc2 * (W2(-1, 0) - W2(2, 0)));
        izx = tx + tz;
        izz = dzi * (hc1 * (W3(-1, 0) - W3(0, 0)) + hc2 * (W3(-2, 0) - W3(1, 0)));
    }
}

// The stress side of a reverse step in one pass over the interior cells:
//   1. stress adjoints += the transposed stress differences of the velocity
//      half-step's weights;
//   2. with a free surface, the transposed mirror s[fpad - k] = -s[fpad + k]
//      of szx and szz: the rows that the ghost rows mirror receive minus the
//      ghost rows' adjoints, which are recomputed here as 0 + their gathered
//      increments (the ghost rows' adjoints are zero before every gather and
//      are never stored);
//   3. the transposed stress half-step: weights and the lam / mu / mu_zx
//      gradients;
// and the receiver adjoint sources of the next reverse step.
// Every value is the one the three steps give as separate passes over the
// whole grid, bit for bit.
template <typename T, int ORDER>
__global__ __launch_bounds__(kDX* kDZ) void d_stress_pass_kernel(const DiscreteParams<T> p)
{
    constexpr int H = ORDER / 2;
    const int ix = blockIdx.x * kDX + threadIdx.x, iz = blockIdx.y * kDZ + threadIdx.y;
    const int slot = blockIdx.z;
    if (p.inj_next >= 0) {
        // the next reverse step's receiver sources in this workgroup's cells
        // (velocity adjoints: nothing else in this pass touches them)
        const int tix = slot * p.nblk + (int)(blockIdx.y * gridDim.x + blockIdx.x);
        const int q1 = p.blk_off[tix + 1];
        for (int q = p.blk_off[tix] + (int)(threadIdx.y * kDX + threadIdx.x); q < q1; q += kDX * kDZ)
            add_receiver_group(p, slot, p.blk_idx[q], p.inj_next);
    }
    const int nzt = p.nzt, nxt = p.nxt, ppad = p.ppad;
    if (iz < H || iz >= nzt - H || ix < H || ix >= nxt - H) return;
    const size_t so = (size_t)slot * p.fst;
    const long c = (long)iz * p.pitch + ix;
    const Cell<T> vx{p.pre_vx + so + c, p.pitch}, vz{p.pre_vz + so + c, p.pitch};
    const T dxi = p.dxi, dzi = p.dzi, hc1 = p.hc1, hc2 = p.hc2, dt = p.dt;
    T vx_x, vz_x, vx_z, vz_z;
    if (ORDER == 2) {
        vx_x = dxi * hc1 * (vx(0, 0) - vx(0, -1));
        vz_x = dxi * hc1 * (vz(0, 1) - vz(0, 0));
        vx_z = dzi * hc1 * (vx(1, 0) - vx(0, 0));
        vz_z = dzi * hc1 * (vz(0, 0) - vz(-1, 0));
    } else {
        vx_x = dxi * (hc1 * (vx(0, 0) - vx(0, -1)) + hc2 * (vx(0, 1) - vx(0, -2)));
        vz_x = dxi * (hc1 * (vz(0, 1) - vz(0, 0)) + hc2 * (vz(0, 2) - vz(0, -1)));
        vx_z = dzi * (hc1 * (vx(1, 0) - vx(0, 0)) + hc2 * (vx(2, 0) - vx(-1, 0)));
        vz_z = dzi * (hc1 * (vz(0, 0) - vz(-1, 0)) + hc2 * (vz(1, 0) - vz(-2, 0)));
    }
    const int tp = p.tp;
    const T* pa = p.pml;
    const T* pb = p.pml + tp;
    const T* pK = p.pml + 2 * tp;
    const T* pah = p.pml + 3 * tp;
    const T* pbh = p.pml + 4 * tp;
    const T* pKh = p.pml + 5 * tp;
    const bool fsurf = p.fsurf != 0;
    const size_t xo = (size_t)slot * p.xst, zo = (size_t)slot * p.zst;
    if (p.npml > 0) {
        // forward derivative terms after the stress-side CPML (psi before the step)
        if (ix <= ppad) {
            const int px = ix - H;
            const size_t q = xo + (size_t)iz * tp + px;
            const T m1 = pb[px] * p.pre_mvxx[q] + pa[px] * vx_x;
            const T m2 = pbh[px] * p.pre_mvzx[q] + pah[px] * vz_x;
            vx_x = vx_x / pK[px] + m1;
            vz_x = vz_x / pKh[px] + m2;
        }
        if (ix >= nxt - ppad - 1) {
            const int px = ix - p.pnx;
            const size_t q = xo + (size_t)iz * tp + px;
            const T m1 = pb[px] * p.pre_mvxx[q] + pa[px] * vx_x;
            const T m2 = pbh[px] * p.pre_mvzx[q] + pah[px] * vz_x;
            vx_x = vx_x / pK[px] + m1;
            vz_x = vz_x / pKh[px] + m2;
        }
        if (iz <= ppad && !fsurf) {
            const int pz = iz - H;
            const size_t q = zo + (size_t)pz * p.pitch + ix;
            const T m1 = pb[pz] * p.pre_mvzz[q] + pa[pz] * vz_z;
            const T m2 = pbh[pz] * p.pre_mvxz[q] + pah[pz] * vx_z;
            vz_z = vz_z / pK[pz] + m1;
            vx_z = vx_z / pKh[pz] + m2;
        }
        if (iz >= nzt - ppad - 1) {
            const int pz = iz - p.pnz;
            const size_t q = zo + (size_t)pz * p.pitch + ix;
            const T m1 = pb[pz] * p.pre_mvzz[q] + pa[pz] * vz_z;
            const T m2 = pbh[pz] * p.pre_mvxz[q] + pah[pz] * vx_z;
            vz_z = vz_z / pK[pz] + m1;
            vx_z = vx_z / pKh[pz] + m2;
        }
    }
    const size_t k = so + c;
    const T l = p.lam[c], m = p.mu[c];
    T bxx, bzx, bzz;
    {
        T ixx, izx, izz;
        vel_gather_inc<T, ORDER>(p, k, ixx, izx, izz);
        bxx = p.sbxx[k] + ixx;
        bzx = p.sbzx[k] + izx;
        bzz = p.sbzz[k] + izz;
    }
    if (fsurf && iz <= 2 * H) {
#pragma unroll
        for (int kk = 1; kk <= H; kk++) {
            // szz of ghost row H - kk goes to row H + kk; szx to row H + kk
            // (order 2) or H + kk - 1 (order 4)
            const bool to_zz = iz == H + kk, to_zx = iz == (ORDER == 2 ? H + kk : H + kk - 1);
            if (to_zz || to_zx) {
                T gxx, gzx, gzz;
                vel_gather_inc<T, ORDER>(p, so + (long)(H - kk) * p.pitch + ix, gxx, gzx, gzz);
                if (to_zz) bzz = bzz - ((T)0 + gzz);
                if (to_zx) bzx = bzx - ((T)0 + gzx);
            }
        }
    }
    T wxx, wzz, wzx;
    if (fsurf && iz == H) {
        // free surface, order 2: sxx = 4 dt vx_x (l m + m^2) / (l + 2 m), szz = szx = 0;
        // order 4: sxx += the same, szz = 0, szx updated as elsewhere
        const T l2m = l + (T)2.0 * m;
        const T den = l2m * l2m;
        wxx = bxx * ((T)4.0 * dt * (l * m + m * m) / l2m);
        wzz = (T)0;
        p.g_lam[k] = p.g_lam[k] + bxx * ((T)4.0 * dt * vx_x * (m * m) / den);
        p.g_mu[k] = p.g_mu[k] + bxx * ((T)4.0 * dt * vx_x * (l * l + (T)2.0 * l * m + (T)2.0 * m * m) / den);
        if (ORDER == 2) {
            wzx = (T)0;
            p.sbxx[k] = (T)0;
            p.sbzx[k] = (T)0;
        } else {
            wzx = dt * p.mu_zx[c] * bzx;
            p.g_mu_zx[k] = p.g_mu_zx[k] + bzx * (dt * (vz_x + vx_z));
            p.sbxx[k] = bxx;
            p.sbzx[k] = bzx;
        }
        p.sbzz[k] = (T)0;
    } else {
        p.sbxx[k] = bxx;
        p.sbzx[k] = bzx;
        p.sbzz[k] = bzz;
        wxx = dt * (l * (bxx + bzz) + (T)2.0 * m * bxx);
        wzz = dt * (l * (bxx + bzz) + (T)2.0 * m * bzz);
        wzx = dt * p.mu_zx[c] * bzx;
        p.g_lam[k] = p.g_lam[k] + (bxx + bzz) * (dt * (vx_x + vz_z));
        p.g_mu[k] = p.g_mu[k] + dt * ((T)2.0 * (vx_x * bxx + vz_z * bzz));
        p.g_mu_zx[k] = p.g_mu_zx[k] + bzx * (dt * (vz_x + vx_z));
    }
    T w_vx_x = wxx, w_vz_x = wzx, w_vx_z = wzx, w_vz_z = wzz;
    if (p.npml > 0) {
        if (iz >= nzt - ppad - 1) {
            const int pz = iz - p.pnz;
            const size_t q = zo + (size_t)pz * p.pitch + ix;
            w_vz_z = cpml_t(w_vz_z, &p.bvzz[q], pa[pz], pb[pz], pK[pz]);
            w_vx_z = cpml_t(w_vx_z, &p.bvxz[q], pah[pz], pbh[pz], pKh[pz]);
        }
        if (iz <= ppad && !fsurf) {
            const int pz = iz - H;
            const size_t q = zo + (size_t)pz * p.pitch + ix;
            w_vz_z = cpml_t(w_vz_z, &p.bvzz[q], pa[pz], pb[pz], pK[pz]);
            w_vx_z = cpml_t(w_vx_z, &p.bvxz[q], pah[pz], pbh[pz], pKh[pz]);
        }
        if (ix >= nxt - ppad - 1) {
            const int px = ix - p.pnx;
            const size_t q = xo + (size_t)iz * tp + px;
            w_vx_x = cpml_t(w_vx_x, &p.bvxx[q], pa[px], pb[px], pK[px]);
            w_vz_x = cpml_t(w_vz_x, &p.bvzx[q], pah[px], pbh[px], pKh[px]);
        }
        if (ix <= ppad) {
            const int px = ix - H;
            const size_t q = xo + (size_t)iz * tp + px;
            w_vx_x = cpml_t(w_vx_x, &p.bvxx[q], pa[px], pb[px], pK[px]);
            w_vz_x = cpml_t(w_vz_x, &p.bvzx[q], pah[px], pbh[px], pKh[px]);
        }
    }
    p.ws0[k] = w_vx_x;
    p.ws1[k] = w_vz_x;
    p.ws2[k] = w_vx_z;
    p.ws3[k] = w_vz_z;
}

template <typename T, int ORDER>
hipError_t discrete_step(const DiscreteParams<T>& p, hipStream_t s)
{
    const dim3 blk(kDX, kDZ);
    const dim3 grd((p.nxt + kDX - 1) / kDX, (p.nzt + kDZ - 1) / kDZ, p.nslots);
    if (p.max_groups > 0 && p.inj_hi > p.inj_lo)
        hipLaunchKernelGGL((d_inject_kernel<T>), dim3((p.max_groups + 63) / 64, p.inj_hi - p.inj_lo),
                           dim3(64), 0, s, p);
    hipLaunchKernelGGL((d_vel_pass_kernel<T, ORDER>), grd, blk, 0, s, p);
    hipLaunchKernelGGL((d_stress_pass_kernel<T, ORDER>), grd, blk, 0, s, p);
    return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t launch_discrete_step(const DiscreteParams<T>& p, int order, hipStream_t s)
{
    if (p.nslots <= 0) return hipSuccess;
    if (order == 2) return discrete_step<T, 2>(p, s);
    if (order == 4) return discrete_step<T, 4>(p, s);
    return hipErrorInvalidValue;
}

template hipError_t launch_discrete_step<float>(const DiscreteParams<float>&, int, hipStream_t);
template hipError_t launch_discrete_step<double>(const DiscreteParams<double>&, int, hipStream_t);
template hipError_t launch_discrete_step<dfloat>(const DiscreteParams<dfloat>&, int, hipStream_t);
template hipError_t launch_discrete_step<ffloat>(const DiscreteParams<ffloat>&, int, hipStream_t);

}  // namespace psv
