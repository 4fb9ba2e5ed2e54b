// psv_step.hip -- the fused time-step kernel of the 2D elastic P-SV
// velocity-stress scheme for gfx950 (CDNA4, wave64, 160 KB LDS per CU).
//
// One launch advances every active shot of a batch by one time step:
//   (adjoint only) adjoint displacement ux += dt vx, uz += dt vz and the
//                  lambda/mu/rho gradient correlation on the old state
//   (forward only) decimated store of the old state (vx/dt, vz/dt, sxx, szx, szz)
//   stress half-step with CPML on the strips and the free-surface override
//   stress mirroring above the free surface
//   velocity half-step with CPML and the energy weight
//   point-source injection, then receiver recording
// It restates forward_kernel_PSV / adjoint_kernel_PSV
// (CUDA_week2/forward_kernel_PSV_GPU.cu:803-1063,
//  CUDA_week2/adjoint_kernel_PSV_GPU.cu:818-1087) with the same floating-point
// operation order; the library is compiled with -ffp-contract=off, so the
// FP64 instantiation reproduces the FP64 CPU path bit for bit.
//
// Work decomposition: a workgroup of 256 threads (4 waves) owns a 64 x TZ
// tile.  It stages vx, vz on the tile plus a 2*fpad halo in LDS, computes the
// new stresses on the tile plus an fpad halo (the halo is recomputed, never
// exchanged), then the new velocities on the tile, so each time step is one
// pass over HBM.  Wavefields and the stress-side CPML memory variables are
// double-buffered (read buffer 0, write buffer 1) because neighbouring tiles
// read the old values of the cells a tile rewrites; the velocity-side memory
// variables are touched by their owner only and are updated in place.
#include <hip/hip_runtime.h>

#include "psv_step.h"

namespace psv {

// Flat tile index -> logical tile, so that the blocks the dispatcher sends to
// one XCD (block ids congruent mod 8) form one contiguous band of tiles and
// their halos meet in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int total)
{
    const int q = total / kNumXcd, r = total % kNumXcd;
    const int x = bid % kNumXcd, k = bid / kNumXcd;
    return x * q + (x < r ? x : r) + k;
}

template <typename T, int ORDER, int MODE, int TZ>
__global__ __launch_bounds__(kThreads) void psv_step_kernel(const StepParams<T> p)
{
    constexpr int H = ORDER / 2;          // fpad
    constexpr int SH = H;                 // stress halo
    constexpr int VH = 2 * H;             // velocity halo
    constexpr int SW = kTX + 2 * SH, SHT = TZ + 2 * SH;
    constexpr int VW = kTX + 2 * VH, VHT = TZ + 2 * VH;

    __shared__ T lvx[VHT][VW];
    __shared__ T lvz[VHT][VW];
    __shared__ T lsxx[SHT][SW];
    __shared__ T lszx[SHT][SW];
    __shared__ T lszz[SHT][SW];

    const int ntile = p.tiles_x * p.tiles_z;
    const int bid = xcd_remap(blockIdx.x, ntile * p.nslots);
    const int slot = bid / ntile;
    const int tile = bid - slot * ntile;
    const int tzi = tile / p.tiles_x;
    const int z0 = tzi * TZ;
    const int x0 = (tile - tzi * p.tiles_x) * kTX;
    const int it = p.it_dev ? (*p.it_dev + p.it) : p.it;
    const int nt_s = p.shot_nt[slot];
    if (it >= nt_s) return;  // this shot has finished (uniform over the block)

    const int nzt = p.nzt, nxt = p.nxt, pitch = p.pitch;
    const int fpad = H, ppad = p.ppad;
    const int tid = threadIdx.x;
    const int tx = tid & (kTX - 1), ty = tid / kTX;
    const size_t fo = (size_t)slot * p.fst;
    const T dt = p.dt, dxi = p.dxi, dzi = p.dzi, hc1 = p.hc1, hc2 = p.hc2;

    // ---- stage old velocities (tile + 2*fpad halo), zero outside the grid
    for (int i = tid; i < VW * VHT; i += kThreads) {
        const int r = i / VW, c = i - r * VW;
        const int gz = z0 - VH + r, gx = x0 - VH + c;
        T a = 0, b = 0;
        if (gz >= 0 && gz < nzt && gx >= 0 && gx < nxt) {
            const size_t g = fo + (size_t)gz * pitch + gx;
            a = p.vx0[g];
            b = p.vz0[g];
        }
        lvx[r][c] = a;
        lvz[r][c] = b;
    }
    __syncthreads();

    // ---- per-cell work on the old state: forward store / adjoint correlation
    const bool fwi_step = (p.flags & kStore) && (it % p.fwi_dt == 0);
    if (MODE == kAdjoint || fwi_step) {
        const int tf = it / p.fwi_dt - p.tf0;
        for (int rr = ty; rr < TZ; rr += kThreads / kTX) {
            const int iz = z0 + rr, ix = x0 + tx;
            if (iz >= nzt || ix >= nxt) continue;
            const size_t g = (size_t)iz * pitch + ix;
            const T vxo = lvx[rr + VH][tx + VH];
            const T vzo = lvz[rr + VH][tx + VH];
            T uxn = 0, uzn = 0;
            if (MODE == kAdjoint) {
                if (iz >= fpad && iz < nzt - fpad && ix >= fpad && ix < nxt - fpad) {
                    uzn = p.uz[fo + g] + dt * vzo;
                    uxn = p.ux[fo + g] + dt * vxo;
                    p.uz[fo + g] = uzn;
                    p.ux[fo + g] = uxn;
                } else {
                    uzn = p.uz[fo + g];
                    uxn = p.ux[fo + g];
                }
            }
            if (!fwi_step) continue;
            const int dzw = iz - p.fwi_z1, dxw = ix - p.fwi_x1;
            if (dzw < 0 || dxw < 0 || iz >= p.fwi_z2 || ix >= p.fwi_x2) continue;
            if (dzw % p.fwi_dz || dxw % p.fwi_dx) continue;
            const size_t plane = (size_t)p.nfz * p.nfx;
            const size_t w = (size_t)(dzw / p.fwi_dz) * p.nfx + dxw / p.fwi_dx;
            T* st = p.store + (size_t)slot * p.sst + (size_t)tf * 5 * plane + w;
            const T sxxo = p.sxx0[fo + g], szxo = p.szx0[fo + g], szzo = p.szz0[fo + g];
            if (MODE == kForward) {
                st[0 * plane] = vxo / dt;
                st[1 * plane] = vzo / dt;
                st[2 * plane] = sxxo;
                st[3 * plane] = szxo;
                st[4 * plane] = szzo;
            } else {
                const T fvx = st[0 * plane], fvz = st[1 * plane];
                const T fsxx = st[2 * plane], fszx = st[3 * plane], fszz = st[4 * plane];
                const T lam = p.lam[g], mu = p.mu[g];
                const T lm = lam + mu;
                const T s1 = (fsxx + fszz) * (sxxo + szzo) * (T)0.25 / (lm * lm);
                const T s2 = (fsxx - fszz) * (sxxo - szzo) / (mu * mu);
                const T s3 = (fszx * szxo) / (mu * mu);
                const T s4 = uxn * fvx + uzn * fvz;
                const T wdt = (T)p.fwi_dt * dt;
                T* gr = p.grad + (size_t)slot * p.gst + w;
                gr[0 * plane] = gr[0 * plane] + wdt * s1;
                gr[1 * plane] = gr[1 * plane] + wdt * (s3 + s1 + s2);
                gr[2 * plane] = gr[2 * plane] - wdt * s4;
            }
        }
    }

    // ---- stress half-step on tile + fpad halo (CUDA_week2/forward_kernel_PSV_GPU.cu:830-931)
    const int tp = p.tp;
    const int npml = p.npml;
    const bool fsurf = p.fsurf != 0;
    // does any cell of the stress region touch a CPML strip?
    const bool pml_blk = npml > 0 &&
        ((x0 - SH <= ppad) || (x0 + kTX + SH - 1 >= nxt - ppad - 1) ||
         (!fsurf && z0 - SH <= ppad) || (z0 + TZ + SH - 1 >= nzt - ppad - 1));
    const T* pa = p.pml;
    const T* pb = p.pml + tp;
    const T* pK = p.pml + 2 * tp;
    const T* pah = p.pml + 3 * tp;
    const T* pbh = p.pml + 4 * tp;
    const T* pKh = p.pml + 5 * tp;

    for (int i = tid; i < SW * SHT; i += kThreads) {
        const int r = i / SW, c = i - r * SW;
        const int iz = z0 - SH + r, ix = x0 - SH + c;
        T sxxn = 0, szxn = 0, szzn = 0;
        if (iz >= 0 && iz < nzt && ix >= 0 && ix < nxt) {
            const size_t g = (size_t)iz * pitch + ix;
            sxxn = p.sxx0[fo + g];
            szxn = p.szx0[fo + g];
            szzn = p.szz0[fo + g];
            if (iz >= fpad && iz < nzt - fpad && ix >= fpad && ix < nxt - fpad) {
                const int vr = r + (VH - SH), vc = c + (VH - SH);
                T vx_x, vz_x, vx_z, vz_z;
                if (ORDER == 2) {
                    vx_x = dxi * hc1 * (lvx[vr][vc] - lvx[vr][vc - 1]);
                    vz_x = dxi * hc1 * (lvz[vr][vc + 1] - lvz[vr][vc]);
                    vx_z = dzi * hc1 * (lvx[vr + 1][vc] - lvx[vr][vc]);
                    vz_z = dzi * hc1 * (lvz[vr][vc] - lvz[vr - 1][vc]);
                } else {
                    vx_x = dxi * (hc1 * (lvx[vr][vc] - lvx[vr][vc - 1]) +
                                  hc2 * (lvx[vr][vc + 1] - lvx[vr][vc - 2]));
                    vz_x = dxi * (hc1 * (lvz[vr][vc + 1] - lvz[vr][vc]) +
                                  hc2 * (lvz[vr][vc + 2] - lvz[vr][vc - 1]));
                    vx_z = dzi * (hc1 * (lvx[vr + 1][vc] - lvx[vr][vc]) +
                                  hc2 * (lvx[vr + 2][vc] - lvx[vr - 1][vc]));
                    vz_z = dzi * (hc1 * (lvz[vr][vc] - lvz[vr - 1][vc]) +
                                  hc2 * (lvz[vr + 1][vc] - lvz[vr - 2][vc]));
                }
                if (pml_blk) {
                    // the tile owner writes the new memory variables
                    const bool own = (iz >= z0 && iz < z0 + TZ && ix >= x0 && ix < x0 + kTX);
                    const size_t xo = (size_t)slot * p.xst, zo = (size_t)slot * p.zst;
                    if (ix <= ppad) {                                  // left strip
                        const int px = ix - fpad;
                        const size_t q = xo + (size_t)iz * tp + px;
                        const T m1 = pb[px] * p.mvxx0[q] + pa[px] * vx_x;
                        const T m2 = pbh[px] * p.mvzx0[q] + pah[px] * vz_x;
                        if (own) { p.mvxx1[q] = m1; p.mvzx1[q] = m2; }
                        vx_x = vx_x / pK[px] + m1;
                        vz_x = vz_x / pKh[px] + m2;
                    }
                    if (ix >= nxt - ppad - 1) {                        // right strip
                        const int px = ix - p.pnx;
                        const size_t q = xo + (size_t)iz * tp + px;
                        const T m1 = pb[px] * p.mvxx0[q] + pa[px] * vx_x;
                        const T m2 = pbh[px] * p.mvzx0[q] + pah[px] * vz_x;
                        if (own) { p.mvxx1[q] = m1; p.mvzx1[q] = m2; }
                        vx_x = vx_x / pK[px] + m1;
                        vz_x = vz_x / pKh[px] + m2;
                    }
                    if (iz <= ppad && !fsurf) {                        // top strip
                        const int pz = iz - fpad;
                        const size_t q = zo + (size_t)pz * pitch + ix;
                        const T m1 = pb[pz] * p.mvzz0[q] + pa[pz] * vz_z;
                        const T m2 = pbh[pz] * p.mvxz0[q] + pah[pz] * vx_z;
                        if (own) { p.mvzz1[q] = m1; p.mvxz1[q] = m2; }
                        vz_z = vz_z / pK[pz] + m1;
                        vx_z = vx_z / pKh[pz] + m2;
                    }
                    if (iz >= nzt - ppad - 1) {                        // bottom strip
                        const int pz = iz - p.pnz;
                        const size_t q = zo + (size_t)pz * pitch + ix;
                        const T m1 = pb[pz] * p.mvzz0[q] + pa[pz] * vz_z;
                        const T m2 = pbh[pz] * p.mvxz0[q] + pah[pz] * vx_z;
                        if (own) { p.mvzz1[q] = m1; p.mvxz1[q] = m2; }
                        vz_z = vz_z / pK[pz] + m1;
                        vx_z = vx_z / pKh[pz] + m2;
                    }
                }
                const T lam = p.lam[g], mu = p.mu[g];
                szxn = szxn + dt * p.mu_zx[g] * (vz_x + vx_z);
                sxxn = sxxn + dt * (lam * (vx_x + vz_z) + ((T)2.0 * mu * vx_x));
                szzn = szzn + dt * (lam * (vx_x + vz_z) + ((T)2.0 * mu * vz_z));
                if (fsurf && iz == fpad) {
                    szzn = (T)0.0;
                    szxn = (T)0.0;
                    sxxn = (T)4.0 * dt * vx_x * (lam * mu + mu * mu) / (lam + (T)2.0 * mu);
                }
            }
        }
        lsxx[r][c] = sxxn;
        lszx[r][c] = szxn;
        lszz[r][c] = szzn;
    }
    __syncthreads();

    // ---- stress mirroring above the free surface (forward_kernel_PSV_GPU.cu:935-942)
    if (fsurf && z0 - SH < fpad) {
        for (int i = tid; i < SW * SHT; i += kThreads) {
            const int r = i / SW, c = i - r * SW;
            const int iz = z0 - SH + r, ix = x0 - SH + c;
            if (iz >= 0 && iz < fpad && ix >= fpad && ix < nxt - fpad) {
                const int rm = r + 2 * (fpad - iz);
                lszx[r][c] = -lszx[rm][c];
                lszz[r][c] = -lszz[rm][c];
            }
        }
        __syncthreads();
    }

    // ---- velocity half-step on the tile (forward_kernel_PSV_GPU.cu:947-1023)
    const bool pml_vel = npml > 0 &&
        ((x0 < ppad) || (x0 + kTX - 1 >= nxt - ppad - 1) ||
         (!fsurf && z0 < ppad) || (z0 + TZ - 1 >= nzt - ppad - 1));
    for (int rr = ty; rr < TZ; rr += kThreads / kTX) {
        const int iz = z0 + rr, ix = x0 + tx;
        if (!(iz >= fpad && iz < nzt - fpad && ix >= fpad && ix < nxt - fpad)) continue;
        const int sr = rr + SH, sc = tx + SH;
        T sxx_x, szx_x, szx_z, szz_z;
        if (ORDER == 2) {
            sxx_x = dxi * hc1 * (lsxx[sr][sc + 1] - lsxx[sr][sc]);
            szx_x = dxi * hc1 * (lszx[sr][sc] - lszx[sr][sc - 1]);
            szx_z = dzi * hc1 * (lszx[sr][sc] - lszx[sr - 1][sc]);
            szz_z = dzi * hc1 * (lszz[sr + 1][sc] - lszz[sr][sc]);
        } else {
            sxx_x = dxi * (hc1 * (lsxx[sr][sc + 1] - lsxx[sr][sc]) +
                           hc2 * (lsxx[sr][sc + 2] - lsxx[sr][sc - 1]));
            szx_x = dxi * (hc1 * (lszx[sr][sc] - lszx[sr][sc - 1]) +
                           hc2 * (lszx[sr][sc + 1] - lszx[sr][sc - 2]));
            szx_z = dzi * (hc1 * (lszx[sr][sc] - lszx[sr - 1][sc]) +
                           hc2 * (lszx[sr + 1][sc] - lszx[sr - 2][sc]));
            szz_z = dzi * (hc1 * (lszz[sr + 1][sc] - lszz[sr][sc]) +
                           hc2 * (lszz[sr + 2][sc] - lszz[sr - 1][sc]));
        }
        const size_t g = (size_t)iz * pitch + ix;
        if (pml_vel) {
            const size_t xo = (size_t)slot * p.xst, zo = (size_t)slot * p.zst;
            if (ix < ppad) {                                           // left strip (strict)
                const int px = ix - fpad;
                const size_t q = xo + (size_t)iz * tp + px;
                const T m1 = pb[px] * p.msxxx[q] + pa[px] * sxx_x;
                const T m2 = pbh[px] * p.mszxx[q] + pah[px] * szx_x;
                p.msxxx[q] = m1;
                p.mszxx[q] = m2;
                sxx_x = sxx_x / pK[px] + m1;
                szx_x = szx_x / pKh[px] + m2;
            }
            if (ix >= nxt - ppad - 1) {                                // right strip
                const int px = ix - p.pnx;
                const size_t q = xo + (size_t)iz * tp + px;
                const T m1 = pb[px] * p.msxxx[q] + pa[px] * sxx_x;
                const T m2 = pbh[px] * p.mszxx[q] + pah[px] * szx_x;
                p.msxxx[q] = m1;
                p.mszxx[q] = m2;
                sxx_x = sxx_x / pK[px] + m1;
                szx_x = szx_x / pKh[px] + m2;
            }
            if (iz < ppad && !fsurf) {                                 // top strip (strict)
                const int pz = iz - fpad;
                const size_t q = zo + (size_t)pz * pitch + ix;
                const T m1 = pb[pz] * p.mszzz[q] + pa[pz] * szz_z;
                const T m2 = pbh[pz] * p.mszxz[q] + pah[pz] * szx_z;
                p.mszzz[q] = m1;
                p.mszxz[q] = m2;
                szz_z = szz_z / pK[pz] + m1;
                szx_z = szx_z / pKh[pz] + m2;
            }
            if (iz >= nzt - ppad - 1) {                                // bottom strip
                const int pz = iz - p.pnz;
                const size_t q = zo + (size_t)pz * pitch + ix;
                const T m1 = pb[pz] * p.mszzz[q] + pa[pz] * szz_z;
                const T m2 = pbh[pz] * p.mszxz[q] + pah[pz] * szx_z;
                p.mszzz[q] = m1;
                p.mszxz[q] = m2;
                szz_z = szz_z / pK[pz] + m1;
                szx_z = szx_z / pKh[pz] + m2;
            }
        }
        const int vr = rr + VH, vc = tx + VH;
        const T vxn = lvx[vr][vc] + dt * p.rho_xp[g] * (sxx_x + szx_z);
        const T vzn = lvz[vr][vc] + dt * p.rho_zp[g] * (szx_x + szz_z);
        if (p.flags & kEnergy) p.We[fo + g] = p.We[fo + g] + (vxn * vxn + vzn * vzn);
        lvx[vr][vc] = vxn;
        lvz[vr][vc] = vzn;
    }

    // ---- point sources after the velocity update (forward_kernel_PSV_GPU.cu:1035-1047)
    const int g0 = p.inj_off[slot], g1 = p.inj_off[slot + 1];
    const int r0 = p.rec_off[slot], r1 = p.rec_off[slot + 1];
    if (g1 > g0 || r1 > r0) __syncthreads();
    if (g1 > g0) {
        for (int gi = g0 + tid; gi < g1; gi += kThreads) {
            const int* e = p.inj_grp + 5 * (size_t)gi;
            const int rz = e[0] - z0 + VH, rx = e[1] - x0 + VH;
            if (e[0] < z0 || e[0] >= z0 + TZ || e[1] < x0 || e[1] >= x0 + kTX) continue;
            const T* sig = p.sig + p.sig_off[slot] + it;
            const int comp = e[2], first = e[3], cnt = e[4];
            if (comp == 4) {
                T a = lsxx[rz - VH + SH][rx - VH + SH], b = lszz[rz - VH + SH][rx - VH + SH];
                for (int k = 0; k < cnt; k++) {
                    const T s = sig[(size_t)p.inj_row[first + k] * nt_s];
                    a = a + s;
                    b = b + s;
                }
                lsxx[rz - VH + SH][rx - VH + SH] = a;
                lszz[rz - VH + SH][rx - VH + SH] = b;
            } else {
                T v = comp == 1 ? lvx[rz][rx] : lvz[rz][rx];
                for (int k = 0; k < cnt; k++) v = v + sig[(size_t)p.inj_row[first + k] * nt_s];
                if (comp == 1) lvx[rz][rx] = v;
                else lvz[rz][rx] = v;
            }
        }
        __syncthreads();
    }
    // ---- receivers (forward_kernel_PSV_GPU.cu:1052-1063)
    for (int ri = r0 + tid; ri < r1; ri += kThreads) {
        const int* e = p.rec_zxc + 3 * (size_t)ri;
        if (e[0] < z0 || e[0] >= z0 + TZ || e[1] < x0 || e[1] >= x0 + kTX) continue;
        const int rz = e[0] - z0 + VH, rx = e[1] - x0 + VH;
        const T v = e[2] == 1 ? lvx[rz][rx] : lvz[rz][rx];
        p.rec[p.rec_out_off[slot] + (size_t)(ri - r0) * nt_s + it] = v;
    }
    // The tile write below reads only cells this thread wrote itself or
    // cells finished before an earlier barrier.

    // ---- write the tile of the new state
    for (int rr = ty; rr < TZ; rr += kThreads / kTX) {
        const int iz = z0 + rr, ix = x0 + tx;
        if (iz >= nzt || ix >= nxt) continue;
        const size_t g = fo + (size_t)iz * pitch + ix;
        p.vx1[g] = lvx[rr + VH][tx + VH];
        p.vz1[g] = lvz[rr + VH][tx + VH];
        p.sxx1[g] = lsxx[rr + SH][tx + SH];
        p.szx1[g] = lszx[rr + SH][tx + SH];
        p.szz1[g] = lszz[rr + SH][tx + SH];
    }
}

template <typename T, int ORDER, int MODE, int TZ>
static hipError_t launch_one(const StepParams<T>& p, hipStream_t s)
{
    const int blocks = p.tiles_x * p.tiles_z * p.nslots;
    hipLaunchKernelGGL((psv_step_kernel<T, ORDER, MODE, TZ>), dim3(blocks), dim3(kThreads), 0, s, p);
    return hipGetLastError();
}

#define PSV_CASE(O, M, Z) \
    if (order == O && mode == M && tz == Z) return launch_one<T, O, M, Z>(p, s);

template <typename T>
hipError_t launch_step(const StepParams<T>& p, int order, int mode, int tz, hipStream_t s);

template <>
hipError_t launch_step<float>(const StepParams<float>& p, int order, int mode, int tz, hipStream_t s)
{
    using T = float;
    PSV_CASE(2, kForward, 16) PSV_CASE(2, kAdjoint, 16)
    PSV_CASE(4, kForward, 16) PSV_CASE(4, kAdjoint, 16)
    PSV_CASE(2, kForward, 8) PSV_CASE(2, kAdjoint, 8)
    PSV_CASE(4, kForward, 8) PSV_CASE(4, kAdjoint, 8)
    PSV_CASE(2, kForward, 32) PSV_CASE(2, kAdjoint, 32)
    PSV_CASE(4, kForward, 32) PSV_CASE(4, kAdjoint, 32)
    return hipErrorInvalidValue;
}

// FP64 tiles stop at TZ = 16 so that one workgroup's LDS stays below 64 KB.
template <>
hipError_t launch_step<double>(const StepParams<double>& p, int order, int mode, int tz, hipStream_t s)
{
    using T = double;
    PSV_CASE(2, kForward, 16) PSV_CASE(2, kAdjoint, 16)
    PSV_CASE(4, kForward, 16) PSV_CASE(4, kAdjoint, 16)
    PSV_CASE(2, kForward, 8) PSV_CASE(2, kAdjoint, 8)
    PSV_CASE(4, kForward, 8) PSV_CASE(4, kAdjoint, 8)
    return hipErrorInvalidValue;
}
#undef PSV_CASE


}  // namespace psv
