// psv_fused.hip -- one time step of the 2D elastic P-SV scheme as a single
// launch, for gfx950 (CDNA4, wave64).
//
// A workgroup (256 threads, 4 waves) owns a 64 x TZ tile of one shot.  It
//   1. loads vx, vz on the tile plus a 2*fpad ring, and sxx, szx, szz and
//      lambda, mu, mu_zx on the tile plus an fpad ring (one coalesced load
//      per row; the 64 tile columns and the ring columns are separate passes,
//      so no wave runs a mostly idle pass);
//   2. does the work on the old state for its own cells (forward: decimated
//      store; adjoint: ux += dt vx, uz += dt vz and the gradient correlation),
//      then computes the new stresses on the tile and its ring into LDS (the
//      ring is recomputed, never exchanged), with CPML on the strips and the
//      free-surface override, and writes its own cells;
//   3. mirrors the new stresses above the free surface in LDS;
//   4. computes the new velocities and the energy weight on the tile, adds
//      point sources, records receivers and writes its own cells.
// Neighbouring tiles read the old values of the cells a tile rewrites, so the
// five wavefields and the stress-side CPML memory are read from one buffer
// set (vx.., mvxx..) and written to the other (vx_w.., mvxx_w..); every
// in-grid cell of a tile is written, updated or not, so the two sets agree
// outside the updated cells.  The velocity-side CPML memory, We, ux and uz
// are touched by their owner only and are updated in place.
//
// The arithmetic is that of psv_step.hip (forward_kernel_PSV /
// adjoint_kernel_PSV, CUDA_week2/forward_kernel_PSV_GPU.cu:803-1063,
// CUDA_week2/adjoint_kernel_PSV_GPU.cu:818-1087) in the same operation
// order, so the results equal the split kernels' bit for bit.
//
// Addressing: the full-grid arrays are guard-padded (StepParams::org), so
// every load of the tile and its rings stays inside the slot's allocation and
// needs no bounds test; values loaded from outside the grid only ever feed
// cells that are not updated.  Offsets are 32-bit byte offsets from uniform
// per-slot bases.
#include <hip/hip_runtime.h>

#include "psv_step.h"

namespace psv {

namespace {

template <typename T>
__device__ __forceinline__ T ldg(const T* base, unsigned boff)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}

template <typename T>
__device__ __forceinline__ void stg(T* base, unsigned boff, T v)
{
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff) = v;
}

// Flat block index -> logical tile; the blocks an XCD receives (ids congruent
// mod 8) form one contiguous band of tiles.
__device__ __forceinline__ int xcd_band(int bid, int total)
{
    const int q = total / kNumXcd, r = total % kNumXcd;
    const int x = bid % kNumXcd, k = bid / kNumXcd;
    return x * q + (x < r ? x : r) + k;
}

}  // namespace

template <typename T, int ORDER, int MODE, int TZ>
__global__ __launch_bounds__(kThreads) void psv_fused_kernel(const StepParams<T> p)
{
    constexpr int H = ORDER / 2;            // fpad
    constexpr int SH = H, VH = 2 * H;       // stress ring, velocity ring
    constexpr int VW = kTX + 2 * VH, VR = TZ + 2 * VH;
    constexpr int SW = kTX + 2 * SH, SR = TZ + 2 * SH;
    constexpr int NVR = (VR + 3) / 4;       // velocity-region rows per thread (tile columns)
    constexpr int NSR = (SR + 3) / 4;       // stress-region rows per thread (tile columns)
    constexpr int NR = TZ / 4;              // own rows per thread
    constexpr int NVH = 2 * VH * VR;        // ring-column cells of the velocity region
    constexpr int NSH = 2 * SH * SR;        // ring-column cells of the stress region
    static_assert(NVH <= kThreads && NSH <= kThreads, "ring columns in one pass");
    __shared__ T lvx[VR][VW];
    __shared__ T lvz[VR][VW];
    __shared__ T lsxx[SR][SW];
    __shared__ T lszx[SR][SW];
    __shared__ T lszz[SR][SW];
    __shared__ int lmeta[4];

    const int ntile = p.tiles_x * p.tiles_z;
    const int bid = xcd_band(blockIdx.x, ntile * p.nslots);
    int slot, tile;
    if (p.band > 0) {
        const int per_band = p.band * p.tiles_x * p.nslots;
        const int b = bid / per_band;
        const int rem = bid - b * per_band;
        const int tb = min(p.band, p.tiles_z - b * p.band) * p.tiles_x;
        slot = rem / tb;
        tile = b * p.band * p.tiles_x + (rem - slot * tb);
    } else {
        slot = bid / ntile;
        tile = bid - slot * ntile;
    }
    const int tzi = tile / p.tiles_x;
    const int z0 = tzi * TZ, x0 = (tile - tzi * p.tiles_x) * kTX;
    const int it = p.it_dev ? (*p.it_dev + p.it) : p.it;
    const int nt_s = p.shot_nt[slot];
    if (it >= nt_s) return;  // this shot has finished (uniform over the block)

    const int nzt = p.nzt, nxt = p.nxt, pitch = p.pitch, org = p.org;
    const int fpad = H, ppad = p.ppad;
    const int tid = threadIdx.x;
    const int tx = tid & (kTX - 1), ty = tid / kTX;
    const bool fsurf = p.fsurf != 0;
    const bool energy = (p.flags & kEnergy) != 0;
    const bool fwi_step = (p.flags & kStore) && (it % p.fwi_dt == 0);
    const bool unit_stride = p.fwi_dz == 1 && p.fwi_dx == 1;
    const int tf_cur = it / p.fwi_dt - p.tf0;
    const bool pre = MODE == kAdjoint || fwi_step;

    // uniform bases at the start of the slot's allocation: byte offsets from
    // them are non-negative and below 2^32
    const ptrdiff_t so = (ptrdiff_t)slot * (ptrdiff_t)p.fst - org;
    const T* vxA = p.vx + so;
    const T* vzA = p.vz + so;
    const T* sxxA = p.sxx + so;
    const T* szxA = p.szx + so;
    const T* szzA = p.szz + so;
    T* vxB = p.vx_w + so;
    T* vzB = p.vz_w + so;
    T* sxxB = p.sxx_w + so;
    T* szxB = p.szx_w + so;
    T* szzB = p.szz_w + so;
    T* WeS = p.We + so;
    T* uxS = p.ux + so;
    T* uzS = p.uz + so;
    const T* lamM = p.lam - org;
    const T* muM = p.mu - org;
    const T* mzxM = p.mu_zx - org;
    const T* rxpM = p.rho_xp - org;
    const T* rzpM = p.rho_zp - org;
    auto boff = [&](int gz, int gx) -> unsigned { return (unsigned)((org + gz * pitch + gx) * (int)sizeof(T)); };

    // ---- 1. loads, all issued before any is used
    T hvx[NVR], hvz[NVR];
#pragma unroll
    for (int k = 0; k < NVR; k++) {
        const int r = ty + 4 * k;
        hvx[k] = hvz[k] = 0;
        if (r < VR) {
            const unsigned o = boff(z0 - VH + r, x0 + tx);
            hvx[k] = ldg(vxA, o);
            hvz[k] = ldg(vzA, o);
        }
    }
    T rvh_x = 0, rvh_z = 0;
    int vhr = 0, vhc = 0;
    if (tid < NVH) {
        vhr = tid / (2 * VH);
        const int c = tid - vhr * (2 * VH);
        vhc = c < VH ? c : kTX + c;
        const unsigned o = boff(z0 - VH + vhr, x0 - VH + vhc);
        rvh_x = ldg(vxA, o);
        rvh_z = ldg(vzA, o);
    }
    // stress region: old stresses and lambda, mu, mu_zx (plus ux, uz on own rows)
    T rs[NSR][6];
    T rux[NSR], ruz[NSR];
#pragma unroll
    for (int k = 0; k < NSR; k++) {
        const int r = ty + 4 * k;
#pragma unroll
        for (int q = 0; q < 6; q++) rs[k][q] = 0;
        rux[k] = ruz[k] = 0;
        if (r < SR) {
            const unsigned o = boff(z0 - SH + r, x0 + tx);
            rs[k][0] = ldg(sxxA, o);
            rs[k][1] = ldg(szxA, o);
            rs[k][2] = ldg(szzA, o);
            rs[k][3] = ldg(lamM, o);
            rs[k][4] = ldg(muM, o);
            rs[k][5] = ldg(mzxM, o);
            if (MODE == kAdjoint && r >= SH && r < SH + TZ) {
                rux[k] = ldg((const T*)uxS, o);
                ruz[k] = ldg((const T*)uzS, o);
            }
        }
    }
    T rsh[6] = {0, 0, 0, 0, 0, 0};
    int shr = 0, shc = 0;
    if (tid < NSH) {
        shr = tid / (2 * SH);
        const int c = tid - shr * (2 * SH);
        shc = c < SH ? c : kTX + c;
        const unsigned o = boff(z0 - SH + shr, x0 - SH + shc);
        rsh[0] = ldg(sxxA, o);
        rsh[1] = ldg(szxA, o);
        rsh[2] = ldg(szzA, o);
        rsh[3] = ldg(lamM, o);
        rsh[4] = ldg(muM, o);
        rsh[5] = ldg(mzxM, o);
    }
    // own cells of the velocity update
    T rrx[NR], rrz[NR], rwe[NR];
#pragma unroll
    for (int m = 0; m < NR; m++) {
        const unsigned o = boff(z0 + ty + 4 * m, x0 + tx);
        rrx[m] = ldg(rxpM, o);
        rrz[m] = ldg(rzpM, o);
        rwe[m] = energy ? ldg((const T*)WeS, o) : (T)0;
    }
    int meta = 0;
    {
        const size_t tix = (size_t)slot * ntile + tile;
        if (tid == 0) meta = p.tinj_off[tix];
        else if (tid == 1) meta = p.tinj_off[tix + 1];
        else if (tid == 2) meta = p.trec_off[tix];
        else if (tid == 3) meta = p.trec_off[tix + 1];
    }

#pragma unroll
    for (int k = 0; k < NVR; k++) {
        const int r = ty + 4 * k;
        if (r < VR) {
            lvx[r][VH + tx] = hvx[k];
            lvz[r][VH + tx] = hvz[k];
        }
    }
    if (tid < NVH) {
        lvx[vhr][vhc] = rvh_x;
        lvz[vhr][vhc] = rvh_z;
    }
    if (tid < 4) lmeta[tid] = meta;
    __syncthreads();

    const T dt = p.dt, dxi = p.dxi, dzi = p.dzi, hc1 = p.hc1, hc2 = p.hc2;
    const int tp = p.tp, npml = p.npml;
    const T* pa = p.pml;
    const T* pb = p.pml + tp;
    const T* pK = p.pml + 2 * tp;
    const T* pah = p.pml + 3 * tp;
    const T* pbh = p.pml + 4 * tp;
    const T* pKh = p.pml + 5 * tp;
    const size_t xo = (size_t)slot * p.xst, zo = (size_t)slot * p.zst;

    // ---- 2. stresses on the tile and its ring
    const bool spml = npml > 0 &&
        ((x0 - SH <= ppad) || (x0 + kTX + SH - 1 >= nxt - ppad - 1) || (!fsurf && z0 - SH <= ppad) ||
         (z0 + TZ + SH - 1 >= nzt - ppad - 1));
    // r, c: stress-region cell; s: old sxx, szx, szz, lambda, mu, mu_zx;
    // own: a cell of this tile (writes back); uxo/uzo: its adjoint displacement
    auto stress_cell = [&](int r, int c, const T* s, bool own, T uxo, T uzo) {
        const int gz = z0 - SH + r, gx = x0 - SH + c;
        const int vr = r + H, vc = c + H;  // the same cell in the velocity region
        const bool ingrid = gz >= 0 && gz < nzt && gx >= 0 && gx < nxt;
        T sxxn = s[0], szxn = s[1], szzn = s[2];
        if (own && pre && ingrid) {
            T uxn = 0, uzn = 0;
            if (MODE == kAdjoint) {
                uzn = uzo;
                uxn = uxo;
                if (gz >= fpad && gz < nzt - fpad && gx >= fpad && gx < nxt - fpad) {
                    uzn = uzn + dt * lvz[vr][vc];
                    uxn = uxn + dt * lvx[vr][vc];
                    const unsigned o = boff(gz, gx);
                    stg(uzS, o, uzn);
                    stg(uxS, o, uxn);
                }
            }
            const int dzw = gz - p.fwi_z1, dxw = gx - p.fwi_x1;
            // unit strides (the common case) need no integer division
            const bool on = unit_stride ? true : (dzw % p.fwi_dz == 0 && dxw % p.fwi_dx == 0);
            if (fwi_step && dzw >= 0 && dxw >= 0 && gz < p.fwi_z2 && gx < p.fwi_x2 && on) {
                const int tf = tf_cur;
                const size_t plane = (size_t)p.nfz * p.nfx;
                const int zf = unit_stride ? dzw : dzw / p.fwi_dz, xf = unit_stride ? dxw : dxw / p.fwi_dx;
                const size_t w = (size_t)zf * p.nfx + xf;
                T* st = p.store + (size_t)slot * p.sst + (size_t)tf * 5 * plane + w;
                if (MODE == kForward) {
                    st[0 * plane] = lvx[vr][vc] / dt;
                    st[1 * plane] = lvz[vr][vc] / dt;
                    st[2 * plane] = s[0];
                    st[3 * plane] = s[1];
                    st[4 * plane] = s[2];
                } else {
                    const T fvx = st[0 * plane], fvz = st[1 * plane];
                    const T fsxx = st[2 * plane], fszx = st[3 * plane], fszz = st[4 * plane];
                    const T lam = s[3], mu = s[4];
                    const T lm = lam + mu;
                    const T s1 = (fsxx + fszz) * (s[0] + s[2]) * (T)0.25 / (lm * lm);
                    const T s2 = (fsxx - fszz) * (s[0] - s[2]) / (mu * mu);
                    const T s3 = (fszx * s[1]) / (mu * mu);
                    const T s4 = uxn * fvx + uzn * fvz;
                    const T wdt = (T)p.fwi_dt * dt;
                    T* gr = p.grad + (size_t)slot * p.gst + w;
                    gr[0 * plane] = gr[0 * plane] + wdt * s1;
                    gr[1 * plane] = gr[1 * plane] + wdt * (s3 + s1 + s2);
                    gr[2 * plane] = gr[2 * plane] - wdt * s4;
                }
            }
        }
        if (gz >= fpad && gz < nzt - fpad && gx >= fpad && gx < nxt - fpad) {
            T vx_x, vz_x, vx_z, vz_z;
            if (ORDER == 2) {
                vx_x = dxi * hc1 * (lvx[vr][vc] - lvx[vr][vc - 1]);
                vz_x = dxi * hc1 * (lvz[vr][vc + 1] - lvz[vr][vc]);
                vx_z = dzi * hc1 * (lvx[vr + 1][vc] - lvx[vr][vc]);
                vz_z = dzi * hc1 * (lvz[vr][vc] - lvz[vr - 1][vc]);
            } else {
                vx_x = dxi * (hc1 * (lvx[vr][vc] - lvx[vr][vc - 1]) + hc2 * (lvx[vr][vc + 1] - lvx[vr][vc - 2]));
                vz_x = dxi * (hc1 * (lvz[vr][vc + 1] - lvz[vr][vc]) + hc2 * (lvz[vr][vc + 2] - lvz[vr][vc - 1]));
                vx_z = dzi * (hc1 * (lvx[vr + 1][vc] - lvx[vr][vc]) + hc2 * (lvx[vr + 2][vc] - lvx[vr - 1][vc]));
                vz_z = dzi * (hc1 * (lvz[vr][vc] - lvz[vr - 1][vc]) + hc2 * (lvz[vr + 1][vc] - lvz[vr - 2][vc]));
            }
            if (spml) {
                if (gx <= ppad) {                                  // left strip
                    const int px = gx - fpad;
                    const size_t q = xo + (size_t)gz * tp + px;
                    const T m1 = pb[px] * p.mvxx[q] + pa[px] * vx_x;
                    const T m2 = pbh[px] * p.mvzx[q] + pah[px] * vz_x;
                    if (own) {
                        p.mvxx_w[q] = m1;
                        p.mvzx_w[q] = m2;
                    }
                    vx_x = vx_x / pK[px] + m1;
                    vz_x = vz_x / pKh[px] + m2;
                }
                if (gx >= nxt - ppad - 1) {                        // right strip
                    const int px = gx - p.pnx;
                    const size_t q = xo + (size_t)gz * tp + px;
                    const T m1 = pb[px] * p.mvxx[q] + pa[px] * vx_x;
                    const T m2 = pbh[px] * p.mvzx[q] + pah[px] * vz_x;
                    if (own) {
                        p.mvxx_w[q] = m1;
                        p.mvzx_w[q] = m2;
                    }
                    vx_x = vx_x / pK[px] + m1;
                    vz_x = vz_x / pKh[px] + m2;
                }
                if (gz <= ppad && !fsurf) {                        // top strip
                    const int pz = gz - fpad;
                    const size_t q = zo + (size_t)pz * pitch + gx;
                    const T m1 = pb[pz] * p.mvzz[q] + pa[pz] * vz_z;
                    const T m2 = pbh[pz] * p.mvxz[q] + pah[pz] * vx_z;
                    if (own) {
                        p.mvzz_w[q] = m1;
                        p.mvxz_w[q] = m2;
                    }
                    vz_z = vz_z / pK[pz] + m1;
                    vx_z = vx_z / pKh[pz] + m2;
                }
                if (gz >= nzt - ppad - 1) {                        // bottom strip
                    const int pz = gz - p.pnz;
                    const size_t q = zo + (size_t)pz * pitch + gx;
                    const T m1 = pb[pz] * p.mvzz[q] + pa[pz] * vz_z;
                    const T m2 = pbh[pz] * p.mvxz[q] + pah[pz] * vx_z;
                    if (own) {
                        p.mvzz_w[q] = m1;
                        p.mvxz_w[q] = m2;
                    }
                    vz_z = vz_z / pK[pz] + m1;
                    vx_z = vx_z / pKh[pz] + m2;
                }
            }
            const T lam = s[3], mu = s[4];
            szxn = s[1] + dt * s[5] * (vz_x + vx_z);
            sxxn = s[0] + dt * (lam * (vx_x + vz_z) + ((T)2.0 * mu * vx_x));
            szzn = s[2] + dt * (lam * (vx_x + vz_z) + ((T)2.0 * mu * vz_z));
            if (fsurf && gz == fpad) {
                szzn = (T)0.0;
                if (ORDER == 2) {
                    szxn = (T)0.0;
                    sxxn = (T)4.0 * dt * vx_x * (lam * mu + mu * mu) / (lam + (T)2.0 * mu);
                } else {
                    sxxn = s[0] + (T)4.0 * dt * vx_x * (lam * mu + mu * mu) / (lam + (T)2.0 * mu);
                }
            }
        }
        lsxx[r][c] = sxxn;
        lszx[r][c] = szxn;
        lszz[r][c] = szzn;
        if (own && ingrid) {
            const unsigned o = boff(gz, gx);
            stg(sxxB, o, sxxn);
            stg(szxB, o, szxn);
            stg(szzB, o, szzn);
        }
    };
#pragma unroll
    for (int k = 0; k < NSR; k++) {
        const int r = ty + 4 * k;
        if (r < SR) stress_cell(r, SH + tx, rs[k], r >= SH && r < SH + TZ, rux[k], ruz[k]);
    }
    if (tid < NSH) stress_cell(shr, shc, rsh, false, (T)0, (T)0);
    __syncthreads();

    // ---- 3. ghost rows above the free surface: mirror of the new stresses
    // (CUDA_week2/forward_kernel_PSV_GPU.cu:935-942)
    if (fsurf && z0 - SH < fpad) {
        auto mirror = [&](int r, int c) {
            const int gz = z0 - SH + r, gx = x0 - SH + c;
            if (gz >= 0 && gz < fpad && gx >= fpad && gx < nxt - fpad) {
                const int mr = r + 2 * (fpad - gz);
                lszx[r][c] = -lszx[ORDER == 2 ? mr : mr - 1][c];
                lszz[r][c] = -lszz[mr][c];
            }
        };
#pragma unroll
        for (int k = 0; k < NSR; k++) {
            const int r = ty + 4 * k;
            if (r < SR) mirror(r, SH + tx);
        }
        if (tid < NSH) mirror(shr, shc);
        __syncthreads();
        // own ghost cells carry the mirrored values in memory
#pragma unroll
        for (int k = 0; k < NSR; k++) {
            const int r = ty + 4 * k;
            const int gz = z0 - SH + r, gx = x0 + tx;
            if (r >= SH && r < SH + TZ && gz < fpad && gx >= fpad && gx < nxt - fpad) {
                const unsigned o = boff(gz, gx);
                stg(szxB, o, lszx[r][SH + tx]);
                stg(szzB, o, lszz[r][SH + tx]);
            }
        }
    }

    // ---- 4. velocities on the tile (CUDA_week2/forward_kernel_PSV_GPU.cu:947-1023)
    const bool vpml = npml > 0 &&
        ((x0 < ppad) || (x0 + kTX - 1 >= nxt - ppad - 1) || (!fsurf && z0 < ppad) ||
         (z0 + TZ - 1 >= nzt - ppad - 1));
    T rvx[NR], rvz[NR];
#pragma unroll
    for (int m = 0; m < NR; m++) {
        const int j = ty + 4 * m;
        const int iz = z0 + j, ix = x0 + tx;
        const int sr = SH + j, sc = SH + tx;
        T vxn = lvx[VH + j][VH + tx], vzn = lvz[VH + j][VH + tx];
        if (iz >= fpad && iz < nzt - fpad && ix >= fpad && ix < nxt - fpad) {
            T sxx_x, szx_x, szx_z, szz_z;
            if (ORDER == 2) {
                sxx_x = dxi * hc1 * (lsxx[sr][sc + 1] - lsxx[sr][sc]);
                szx_x = dxi * hc1 * (lszx[sr][sc] - lszx[sr][sc - 1]);
                szx_z = dzi * hc1 * (lszx[sr][sc] - lszx[sr - 1][sc]);
                szz_z = dzi * hc1 * (lszz[sr + 1][sc] - lszz[sr][sc]);
            } else {
                sxx_x = dxi * (hc1 * (lsxx[sr][sc + 1] - lsxx[sr][sc]) + hc2 * (lsxx[sr][sc + 2] - lsxx[sr][sc - 1]));
                szx_x = dxi * (hc1 * (lszx[sr][sc] - lszx[sr][sc - 1]) + hc2 * (lszx[sr][sc + 1] - lszx[sr][sc - 2]));
                szx_z = dzi * (hc1 * (lszx[sr][sc] - lszx[sr - 1][sc]) + hc2 * (lszx[sr + 1][sc] - lszx[sr - 2][sc]));
                szz_z = dzi * (hc1 * (lszz[sr + 1][sc] - lszz[sr][sc]) + hc2 * (lszz[sr + 2][sc] - lszz[sr - 1][sc]));
            }
            if (vpml) {
                if (ix < ppad) {                                   // left strip (strict)
                    const int px = ix - fpad;
                    const size_t q = xo + (size_t)iz * tp + px;
                    const T m1 = pb[px] * p.msxxx[q] + pa[px] * sxx_x;
                    const T m2 = pbh[px] * p.mszxx[q] + pah[px] * szx_x;
                    p.msxxx[q] = m1;
                    p.mszxx[q] = m2;
                    sxx_x = sxx_x / pK[px] + m1;
                    szx_x = szx_x / pKh[px] + m2;
                }
                if (ix >= nxt - ppad - 1) {                        // right strip
                    const int px = ix - p.pnx;
                    const size_t q = xo + (size_t)iz * tp + px;
                    const T m1 = pb[px] * p.msxxx[q] + pa[px] * sxx_x;
                    const T m2 = pbh[px] * p.mszxx[q] + pah[px] * szx_x;
                    p.msxxx[q] = m1;
                    p.mszxx[q] = m2;
                    sxx_x = sxx_x / pK[px] + m1;
                    szx_x = szx_x / pKh[px] + m2;
                }
                if (iz < ppad && !fsurf) {                         // top strip (strict)
                    const int pz = iz - fpad;
                    const size_t q = zo + (size_t)pz * pitch + ix;
                    const T m1 = pb[pz] * p.mszzz[q] + pa[pz] * szz_z;
                    const T m2 = pbh[pz] * p.mszxz[q] + pah[pz] * szx_z;
                    p.mszzz[q] = m1;
                    p.mszxz[q] = m2;
                    szz_z = szz_z / pK[pz] + m1;
                    szx_z = szx_z / pKh[pz] + m2;
                }
                if (iz >= nzt - ppad - 1) {                        // bottom strip
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
            vxn = vxn + dt * rrx[m] * (sxx_x + szx_z);
            vzn = vzn + dt * rrz[m] * (szx_x + szz_z);
            if (energy) stg(WeS, boff(iz, ix), rwe[m] + (vxn * vxn + vzn * vzn));
        }
        rvx[m] = vxn;
        rvz[m] = vzn;
    }

    // ---- point sources, then receivers (CUDA_week2/forward_kernel_PSV_GPU.cu:1035-1063),
    // through the velocity tile in LDS for tiles that have any
    const int g0 = lmeta[0], g1 = lmeta[1];
    const int r0 = lmeta[2], r1 = lmeta[3];
    if (g1 > g0 || r1 > r0) {
#pragma unroll
        for (int m = 0; m < NR; m++) {
            lvx[VH + ty + 4 * m][VH + tx] = rvx[m];
            lvz[VH + ty + 4 * m][VH + tx] = rvz[m];
        }
        __syncthreads();
        for (int gi = g0 + tid; gi < g1; gi += kThreads) {
            const int* e = p.inj_grp + 5 * (size_t)p.tinj_idx[gi];
            const int comp = e[2];
            if (comp != 1 && comp != 2) continue;  // stress sources: psv_stress_source_kernel
            const int rz = VH + e[0] - z0, rx = VH + e[1] - x0;
            const T* sig = p.sig + p.sig_off[slot] + it;
            const int first = e[3], cnt = e[4];
            T v = comp == 1 ? lvx[rz][rx] : lvz[rz][rx];
            for (int q = 0; q < cnt; q++) v = v + sig[(size_t)p.inj_row[first + q] * nt_s];
            if (comp == 1) lvx[rz][rx] = v;
            else lvz[rz][rx] = v;
        }
        __syncthreads();
        for (int ri = r0 + tid; ri < r1; ri += kThreads) {
            const int rid = p.trec_idx[ri];
            const int* e = p.rec_zxc + 3 * (size_t)rid;
            const int rz = VH + e[0] - z0, rx = VH + e[1] - x0;
            const T v = e[2] == 1 ? lvx[rz][rx] : lvz[rz][rx];
            p.rec[p.rec_out_off[slot] + (size_t)(rid - p.rec_first[slot]) * nt_s + it] = v;
        }
#pragma unroll
        for (int m = 0; m < NR; m++) {
            rvx[m] = lvx[VH + ty + 4 * m][VH + tx];
            rvz[m] = lvz[VH + ty + 4 * m][VH + tx];
        }
    }
#pragma unroll
    for (int m = 0; m < NR; m++) {
        const int iz = z0 + ty + 4 * m, ix = x0 + tx;
        if (iz < nzt && ix < nxt) {
            const unsigned o = boff(iz, ix);
            stg(vxB, o, rvx[m]);
            stg(vzB, o, rvz[m]);
        }
    }
}

template <typename T, int ORDER, int MODE, int TZ>
static hipError_t launch_fused_one(const StepParams<T>& p, hipStream_t s)
{
    const int blocks = p.tiles_x * p.tiles_z * p.nslots;
    hipLaunchKernelGGL((psv_fused_kernel<T, ORDER, MODE, TZ>), dim3(blocks), dim3(kThreads), 0, s, p);
    if (p.nexpl > 0) {
        // explosive sources act on the stresses just written
        StepParams<T> q = p;
        q.sxx = p.sxx_w;
        q.szz = p.szz_w;
        const hipError_t e = launch_stress_sources<T>(q, s);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

#define PSV_FCASE(O, M, Z) \
    if (order == O && mode == M && tz == Z) return launch_fused_one<T, O, M, Z>(p, s);

template <typename T>
hipError_t launch_fused(const StepParams<T>& p, int order, int mode, int tz, hipStream_t s)
{
    PSV_FCASE(2, kForward, 8) PSV_FCASE(2, kAdjoint, 8)
    PSV_FCASE(4, kForward, 8) PSV_FCASE(4, kAdjoint, 8)
    PSV_FCASE(2, kForward, 16) PSV_FCASE(2, kAdjoint, 16)
    PSV_FCASE(4, kForward, 16) PSV_FCASE(4, kAdjoint, 16)
    return hipErrorInvalidValue;
}
#undef PSV_FCASE

template hipError_t launch_fused<float>(const StepParams<float>&, int, int, int, hipStream_t);
template hipError_t launch_fused<double>(const StepParams<double>&, int, int, int, hipStream_t);
template hipError_t launch_fused<dfloat>(const StepParams<dfloat>&, int, int, int, hipStream_t);
template hipError_t launch_fused<ffloat>(const StepParams<ffloat>&, int, int, int, hipStream_t);

}  // namespace psv
