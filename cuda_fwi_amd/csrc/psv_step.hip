// psv_step.hip -- the time step of the 2D elastic P-SV velocity-stress scheme
// for gfx950 (CDNA4, wave64).
//
// One step of every active shot of a batch is two launches, each a streaming
// pass over the grid with one thread per cell (and TZ/4 cells per thread):
//   stress kernel   (adjoint) ux += dt vx, uz += dt vz and the lambda/mu/rho
//                   gradient correlation, or (forward) the decimated store,
//                   all on the old state; then the stress update with CPML on
//                   the strips and the free-surface override
//   velocity kernel stress mirroring above the free surface (done while the
//                   stress tile is staged), velocity update with CPML and the
//                   energy weight, point-source injection, receiver recording
// plus, only when a batch has explosive (stress) sources, a tiny launch that
// adds them after the velocity update.  Each kernel writes only the fields it
// updates and reads its neighbours' values of the other fields, so every
// array is updated in place.  The arithmetic restates forward_kernel_PSV /
// adjoint_kernel_PSV (CUDA_week2/forward_kernel_PSV_GPU.cu:803-1063,
// CUDA_week2/adjoint_kernel_PSV_GPU.cu:818-1087) in the same floating-point
// operation order; the library is built with -ffp-contract=off, so the FP64
// instantiation reproduces the FP64 CPU path bit for bit.
//
// A workgroup is 256 threads (4 waves) on a 64 x TZ tile.  The neighbour
// values a tile needs (the other field group within 2 cells) are staged in
// LDS in the shape of the stencils: full rows of the 64 tile columns above
// and below the tile, and the 4 halo columns only on the tile's own rows (one
// pass of the first wave).  Everything else is one coalesced load per cell.
// The arrays are guard-padded (StepParams::org), so loads are not predicated;
// offsets are 32-bit byte offsets from uniform per-slot bases.
#include <hip/hip_runtime.h>

#include "psv_step.h"

namespace psv {

constexpr int kHalo = 2;  // staged halo: covers the order-4 stencils (order 2 uses 1)
constexpr int kRowsPerWave = kThreads / kTX;

// Flat block index -> logical tile, so that the blocks the dispatcher sends
// to one XCD (block ids congruent mod 8) form one contiguous band of tiles
// whose halos meet in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int total)
{
    const int q = total / kNumXcd, r = total % kNumXcd;
    const int x = bid % kNumXcd, k = bid / kNumXcd;
    return x * q + (x < r ? x : r) + k;
}

struct TileId {
    int slot, tile, z0, x0;
};

template <int TZ, typename T>
__device__ __forceinline__ TileId tile_of(const StepParams<T>& p)
{
    const int ntile = p.tiles_x * p.tiles_z;
    const int bid = xcd_remap(blockIdx.x, ntile * p.nslots);
    TileId t;
    if (p.band > 0) {
        // bands of p.band tile rows; the shots of a band run one after
        // another, so the band's model tiles can be L2 hits after the first
        const int per_band = p.band * p.tiles_x * p.nslots;
        const int b = bid / per_band;
        const int rem = bid - b * per_band;
        const int tb = min(p.band, p.tiles_z - b * p.band) * p.tiles_x;
        t.slot = rem / tb;
        t.tile = b * p.band * p.tiles_x + (rem - t.slot * tb);
    } else {
        t.slot = bid / ntile;
        t.tile = bid - t.slot * ntile;
    }
    const int tzi = t.tile / p.tiles_x;
    t.z0 = tzi * TZ;
    t.x0 = (t.tile - tzi * p.tiles_x) * kTX;
    return t;
}

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

// ---------------------------------------------------------------- stress
template <typename T, int ORDER, int MODE, int TZ>
__global__ __launch_bounds__(kThreads) void psv_stress_kernel(const StepParams<T> p)
{
    constexpr int H = ORDER / 2;                 // fpad
    constexpr int VW = kTX + 2 * kHalo, VHT = TZ + 2 * kHalo;
    constexpr int NR = TZ / kRowsPerWave;
    constexpr int NH = (VHT + kRowsPerWave - 1) / kRowsPerWave;
    constexpr int NHC = 2 * kHalo * TZ;          // halo-column cells (own rows only)
    static_assert(NHC <= kThreads, "halo columns in one pass");
    // vx and vz: the x derivatives need columns -2..+2 on the own rows, the
    // z derivatives rows -2..+2 on the own columns; the corners are not used
    __shared__ T lvx[VHT][VW];
    __shared__ T lvz[VHT][VW];

    const TileId t = tile_of<TZ>(p);
    const int slot = t.slot, z0 = t.z0, x0 = t.x0;
    const int it = p.it_dev ? (*p.it_dev + p.it) : p.it;
    const int nt_s = p.shot_nt[slot];
    if (it >= nt_s) return;  // this shot has finished (uniform over the block)
    const int nzt = p.nzt, nxt = p.nxt, pitch = p.pitch, org = p.org;
    const int fpad = H, ppad = p.ppad;
    const int tid = threadIdx.x;
    const int tx = tid & (kTX - 1), ty = tid / kTX;
    const bool fwi_step = (p.flags & kStore) && (it % p.fwi_dt == 0);
    const bool unit_stride = p.fwi_dz == 1 && p.fwi_dx == 1;
    const int tf_cur = it / p.fwi_dt - p.tf0;

    // uniform bases at the start of the slot's allocation (guard-padded
    // layout): every load below stays inside it, so none is predicated
    const ptrdiff_t so = (ptrdiff_t)slot * (ptrdiff_t)p.fst - org;
    const T* vxA = p.vx + so;
    const T* vzA = p.vz + so;
    T* sxxS = p.sxx + so;
    T* szxS = p.szx + so;
    T* szzS = p.szz + so;
    T* uxS = p.ux + so;
    T* uzS = p.uz + so;
    const T* lamM = p.lam - org;
    const T* muM = p.mu - org;
    const T* mzxM = p.mu_zx - org;
    auto boff = [&](int gz, int gx) -> unsigned { return (unsigned)((org + gz * pitch + gx) * (int)sizeof(T)); };

    // ---- loads: own cells (stresses, model, adjoint displacement) and the
    // velocity stencil region, all issued before any is used
    T rsxx[NR], rszx[NR], rszz[NR], rlam[NR], rmu[NR], rmzx[NR], rux[NR], ruz[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const unsigned o = boff(z0 + ty + k * kRowsPerWave, x0 + tx);
        rsxx[k] = ldg((const T*)sxxS, o);
        rszx[k] = ldg((const T*)szxS, o);
        rszz[k] = ldg((const T*)szzS, o);
        rlam[k] = ldg(lamM, o);
        rmu[k] = ldg(muM, o);
        rmzx[k] = ldg(mzxM, o);
        rux[k] = ruz[k] = 0;
        if (MODE == kAdjoint) {
            rux[k] = ldg((const T*)uxS, o);
            ruz[k] = ldg((const T*)uzS, o);
        }
    }
    T hvx[NH], hvz[NH];
#pragma unroll
    for (int k = 0; k < NH; k++) {
        const int r = ty + k * kRowsPerWave;
        hvx[k] = hvz[k] = 0;
        if (r < VHT) {
            const unsigned o = boff(z0 - kHalo + r, x0 + tx);
            hvx[k] = ldg(vxA, o);
            hvz[k] = ldg(vzA, o);
        }
    }
    T cvx = 0, cvz = 0;
    int cr = 0, cc = 0;
    if (tid < NHC) {
        cr = kHalo + tid / (2 * kHalo);
        const int c = tid % (2 * kHalo);
        cc = c < kHalo ? c : kTX + c;
        const unsigned o = boff(z0 - kHalo + cr, x0 - kHalo + cc);
        cvx = ldg(vxA, o);
        cvz = ldg(vzA, o);
    }
#pragma unroll
    for (int k = 0; k < NH; k++) {
        const int r = ty + k * kRowsPerWave;
        if (r < VHT) {
            lvx[r][kHalo + tx] = hvx[k];
            lvz[r][kHalo + tx] = hvz[k];
        }
    }
    if (tid < NHC) {
        lvx[cr][cc] = cvx;
        lvz[cr][cc] = cvz;
    }
    __syncthreads();

    const T dt = p.dt, dxi = p.dxi, dzi = p.dzi, hc1 = p.hc1, hc2 = p.hc2;
    const int tp = p.tp, npml = p.npml;
    const bool fsurf = p.fsurf != 0;
    const bool pml_blk = npml > 0 &&
        ((x0 <= ppad) || (x0 + kTX - 1 >= nxt - ppad - 1) || (!fsurf && z0 <= ppad) ||
         (z0 + TZ - 1 >= nzt - ppad - 1));
    const T* pa = p.pml;
    const T* pb = p.pml + tp;
    const T* pK = p.pml + 2 * tp;
    const T* pah = p.pml + 3 * tp;
    const T* pbh = p.pml + 4 * tp;
    const T* pKh = p.pml + 5 * tp;

#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int rr = ty + k * kRowsPerWave;
        const int iz = z0 + rr, ix = x0 + tx;
        if (iz >= nzt || ix >= nxt) continue;
        const unsigned o = boff(iz, ix);
        const int vr = rr + kHalo, vc = tx + kHalo;

        // -- work on the old state: forward store / adjoint correlation
        if (MODE == kAdjoint || fwi_step) {
            T uxn = 0, uzn = 0;
            if (MODE == kAdjoint) {
                uzn = ruz[k];
                uxn = rux[k];
                if (iz >= fpad && iz < nzt - fpad && ix >= fpad && ix < nxt - fpad) {
                    uzn = uzn + dt * lvz[vr][vc];
                    uxn = uxn + dt * lvx[vr][vc];
                    stg(uzS, o, uzn);
                    stg(uxS, o, uxn);
                }
            }
            const int dzw = iz - p.fwi_z1, dxw = ix - p.fwi_x1;
            // unit strides (the common case) need no integer division
            const bool on = unit_stride ? true : (dzw % p.fwi_dz == 0 && dxw % p.fwi_dx == 0);
            if (fwi_step && dzw >= 0 && dxw >= 0 && iz < p.fwi_z2 && ix < p.fwi_x2 && on) {
                const int tf = tf_cur;
                const size_t plane = (size_t)p.nfz * p.nfx;
                const int zf = unit_stride ? dzw : dzw / p.fwi_dz, xf = unit_stride ? dxw : dxw / p.fwi_dx;
                const size_t w = (size_t)zf * p.nfx + xf;
                T* st = p.store + (size_t)slot * p.sst + (size_t)tf * 5 * plane + w;
                if (MODE == kForward) {
                    st[0 * plane] = lvx[vr][vc] / dt;
                    st[1 * plane] = lvz[vr][vc] / dt;
                    st[2 * plane] = rsxx[k];
                    st[3 * plane] = rszx[k];
                    st[4 * plane] = rszz[k];
                } else {
                    const T fvx = st[0 * plane], fvz = st[1 * plane];
                    const T fsxx = st[2 * plane], fszx = st[3 * plane], fszz = st[4 * plane];
                    const T lam = rlam[k], mu = rmu[k];
                    const T lm = lam + mu;
                    const T s1 = (fsxx + fszz) * (rsxx[k] + rszz[k]) * (T)0.25 / (lm * lm);
                    const T s2 = (fsxx - fszz) * (rsxx[k] - rszz[k]) / (mu * mu);
                    const T s3 = (fszx * rszx[k]) / (mu * mu);
                    const T s4 = uxn * fvx + uzn * fvz;
                    const T wdt = (T)p.fwi_dt * dt;
                    T* gr = p.grad + (size_t)slot * p.gst + w;
                    gr[0 * plane] = gr[0 * plane] + wdt * s1;
                    gr[1 * plane] = gr[1 * plane] + wdt * (s3 + s1 + s2);
                    gr[2 * plane] = gr[2 * plane] - wdt * s4;
                }
            }
        }

        // -- stress half-step (CUDA_week2/forward_kernel_PSV_GPU.cu:830-931)
        if (!(iz >= fpad && iz < nzt - fpad && ix >= fpad && ix < nxt - fpad)) continue;
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
        if (pml_blk) {
            const size_t xo = (size_t)slot * p.xst, zo = (size_t)slot * p.zst;
            if (ix <= ppad) {                                  // left strip
                const int px = ix - fpad;
                const size_t q = xo + (size_t)iz * tp + px;
                const T m1 = pb[px] * p.mvxx[q] + pa[px] * vx_x;
                const T m2 = pbh[px] * p.mvzx[q] + pah[px] * vz_x;
                p.mvxx[q] = m1;
                p.mvzx[q] = m2;
                vx_x = vx_x / pK[px] + m1;
                vz_x = vz_x / pKh[px] + m2;
            }
            if (ix >= nxt - ppad - 1) {                        // right strip
                const int px = ix - p.pnx;
                const size_t q = xo + (size_t)iz * tp + px;
                const T m1 = pb[px] * p.mvxx[q] + pa[px] * vx_x;
                const T m2 = pbh[px] * p.mvzx[q] + pah[px] * vz_x;
                p.mvxx[q] = m1;
                p.mvzx[q] = m2;
                vx_x = vx_x / pK[px] + m1;
                vz_x = vz_x / pKh[px] + m2;
            }
            if (iz <= ppad && !fsurf) {                        // top strip
                const int pz = iz - fpad;
                const size_t q = zo + (size_t)pz * pitch + ix;
                const T m1 = pb[pz] * p.mvzz[q] + pa[pz] * vz_z;
                const T m2 = pbh[pz] * p.mvxz[q] + pah[pz] * vx_z;
                p.mvzz[q] = m1;
                p.mvxz[q] = m2;
                vz_z = vz_z / pK[pz] + m1;
                vx_z = vx_z / pKh[pz] + m2;
            }
            if (iz >= nzt - ppad - 1) {                        // bottom strip
                const int pz = iz - p.pnz;
                const size_t q = zo + (size_t)pz * pitch + ix;
                const T m1 = pb[pz] * p.mvzz[q] + pa[pz] * vz_z;
                const T m2 = pbh[pz] * p.mvxz[q] + pah[pz] * vx_z;
                p.mvzz[q] = m1;
                p.mvxz[q] = m2;
                vz_z = vz_z / pK[pz] + m1;
                vx_z = vx_z / pKh[pz] + m2;
            }
        }
        const T lam = rlam[k], mu = rmu[k];
        T szxn = rszx[k] + dt * rmzx[k] * (vz_x + vx_z);
        T sxxn = rsxx[k] + dt * (lam * (vx_x + vz_z) + ((T)2.0 * mu * vx_x));
        T szzn = rszz[k] + dt * (lam * (vx_x + vz_z) + ((T)2.0 * mu * vz_z));
        if (fsurf && iz == fpad) {
            szzn = (T)0.0;
            if (ORDER == 2) {
                szxn = (T)0.0;
                sxxn = (T)4.0 * dt * vx_x * (lam * mu + mu * mu) / (lam + (T)2.0 * mu);
            } else {
                // order 4 (extension): sxx integrates the surface relation and
                // szx keeps its update; the order-2 form is unstable here
                sxxn = rsxx[k] + (T)4.0 * dt * vx_x * (lam * mu + mu * mu) / (lam + (T)2.0 * mu);
            }
        }
        stg(sxxS, o, sxxn);
        stg(szxS, o, szxn);
        stg(szzS, o, szzn);
    }
}

// ---------------------------------------------------------------- velocity
template <typename T, int ORDER, int TZ>
__global__ __launch_bounds__(kThreads) void psv_velocity_kernel(const StepParams<T> p)
{
    constexpr int H = ORDER / 2;
    constexpr int SW = kTX + 2 * kHalo, SHT = TZ + 2 * kHalo;
    constexpr int NR = TZ / kRowsPerWave;
    constexpr int NH = (SHT + kRowsPerWave - 1) / kRowsPerWave;
    constexpr int NHC = 2 * kHalo * TZ;          // halo-column cells (own rows only)
    static_assert(NHC <= kThreads, "halo columns in one pass");
    // the stencil each stress needs: sxx columns -1..+2 on the own rows,
    // szz rows -1..+2 on the own columns, szx both (no corners)
    __shared__ T lsxx[TZ][SW];
    __shared__ T lszx[SHT][SW];
    __shared__ T lszz[SHT][kTX];
    __shared__ T lvx[TZ][kTX];
    __shared__ T lvz[TZ][kTX];
    __shared__ int lmeta[4];

    const TileId t = tile_of<TZ>(p);
    const int slot = t.slot, z0 = t.z0, x0 = t.x0;
    const int it = p.it_dev ? (*p.it_dev + p.it) : p.it;
    const int nt_s = p.shot_nt[slot];
    if (it >= nt_s) return;  // uniform over the block
    const int nzt = p.nzt, nxt = p.nxt, pitch = p.pitch, org = p.org;
    const int fpad = H, ppad = p.ppad;
    const int tid = threadIdx.x;
    const int tx = tid & (kTX - 1), ty = tid / kTX;
    const bool fsurf = p.fsurf != 0;
    const bool energy = (p.flags & kEnergy) != 0;
    // rows above the free surface are staged only by the top tiles
    const bool top = fsurf && z0 - kHalo < fpad;

    const ptrdiff_t so = (ptrdiff_t)slot * (ptrdiff_t)p.fst - org;
    const T* sxxA = p.sxx + so;
    T* szxS = p.szx + so;
    T* szzS = p.szz + so;
    T* vxS = p.vx + so;
    T* vzS = p.vz + so;
    T* WeS = p.We + so;
    const T* rxpM = p.rho_xp - org;
    const T* rzpM = p.rho_zp - org;
    auto boff = [&](int gz, int gx) -> unsigned { return (unsigned)((org + gz * pitch + gx) * (int)sizeof(T)); };

    // ---- loads: own cells and the stress stencil region
    T rvx[NR], rvz[NR], rrxp[NR], rrzp[NR], rwe[NR], hsxx[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const unsigned o = boff(z0 + ty + k * kRowsPerWave, x0 + tx);
        rvx[k] = ldg((const T*)vxS, o);
        rvz[k] = ldg((const T*)vzS, o);
        rrxp[k] = ldg(rxpM, o);
        rrzp[k] = ldg(rzpM, o);
        rwe[k] = energy ? ldg((const T*)WeS, o) : (T)0;
        hsxx[k] = ldg(sxxA, o);
    }
    int meta = 0;
    {
        const size_t tix = (size_t)slot * p.tiles_x * p.tiles_z + t.tile;
        if (tid == 0) meta = p.tinj_off[tix];
        else if (tid == 1) meta = p.tinj_off[tix + 1];
        else if (tid == 2) meta = p.trec_off[tix];
        else if (tid == 3) meta = p.trec_off[tix + 1];
    }
    // szx, szz on the own columns, rows -2..TZ+1; above the free surface the
    // rows hold the mirror of the new stresses below it
    // (forward_kernel_PSV_GPU.cu:935-942), so they are read there
    T hszx[NH], hszz[NH];
#pragma unroll
    for (int k = 0; k < NH; k++) {
        const int r = ty + k * kRowsPerWave;
        const int gz = z0 - kHalo + r, gx = x0 + tx;
        hszx[k] = hszz[k] = 0;
        if (r < SHT) {
            if (top && gz >= 0 && gz < fpad && gx >= fpad && gx < nxt - fpad) {
                // order 4: szx (half a cell below szz) mirrors about row fpad - 1/2
                const unsigned om = boff(2 * fpad - gz, gx);
                const unsigned omx = ORDER == 2 ? om : boff(2 * fpad - gz - 1, gx);
                hszx[k] = -ldg((const T*)szxS, omx);
                hszz[k] = -ldg((const T*)szzS, om);
            } else {
                const unsigned o = boff(gz, gx);
                hszx[k] = ldg((const T*)szxS, o);
                hszz[k] = ldg((const T*)szzS, o);
            }
        }
    }
    // sxx and szx on the halo columns of the own rows
    T csxx = 0, cszx = 0;
    int cj = 0, cc = 0;
    if (tid < NHC) {
        cj = tid / (2 * kHalo);
        const int c = tid % (2 * kHalo);
        cc = c < kHalo ? c : kTX + c;
        const unsigned o = boff(z0 + cj, x0 - kHalo + cc);
        csxx = ldg(sxxA, o);
        cszx = ldg((const T*)szxS, o);
    }
#pragma unroll
    for (int k = 0; k < NR; k++) lsxx[ty + k * kRowsPerWave][kHalo + tx] = hsxx[k];
#pragma unroll
    for (int k = 0; k < NH; k++) {
        const int r = ty + k * kRowsPerWave;
        if (r < SHT) {
            lszx[r][kHalo + tx] = hszx[k];
            lszz[r][tx] = hszz[k];
        }
    }
    if (tid < NHC) {
        lsxx[cj][cc] = csxx;
        lszx[kHalo + cj][cc] = cszx;
    }
    if (tid < 4) lmeta[tid] = meta;
    __syncthreads();

    const T dt = p.dt, dxi = p.dxi, dzi = p.dzi, hc1 = p.hc1, hc2 = p.hc2;
    const int tp = p.tp, npml = p.npml;
    const bool pml_blk = npml > 0 &&
        ((x0 < ppad) || (x0 + kTX - 1 >= nxt - ppad - 1) || (!fsurf && z0 < ppad) ||
         (z0 + TZ - 1 >= nzt - ppad - 1));
    const T* pa = p.pml;
    const T* pb = p.pml + tp;
    const T* pK = p.pml + 2 * tp;
    const T* pah = p.pml + 3 * tp;
    const T* pbh = p.pml + 4 * tp;
    const T* pKh = p.pml + 5 * tp;

    // ---- velocity half-step (CUDA_week2/forward_kernel_PSV_GPU.cu:947-1023)
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int rr = ty + k * kRowsPerWave;
        const int iz = z0 + rr, ix = x0 + tx;
        if (!(iz >= fpad && iz < nzt - fpad && ix >= fpad && ix < nxt - fpad)) continue;
        const int sr = rr + kHalo, sc = tx + kHalo;
        T sxx_x, szx_x, szx_z, szz_z;
        if (ORDER == 2) {
            sxx_x = dxi * hc1 * (lsxx[rr][sc + 1] - lsxx[rr][sc]);
            szx_x = dxi * hc1 * (lszx[sr][sc] - lszx[sr][sc - 1]);
            szx_z = dzi * hc1 * (lszx[sr][sc] - lszx[sr - 1][sc]);
            szz_z = dzi * hc1 * (lszz[sr + 1][tx] - lszz[sr][tx]);
        } else {
            sxx_x = dxi * (hc1 * (lsxx[rr][sc + 1] - lsxx[rr][sc]) + hc2 * (lsxx[rr][sc + 2] - lsxx[rr][sc - 1]));
            szx_x = dxi * (hc1 * (lszx[sr][sc] - lszx[sr][sc - 1]) + hc2 * (lszx[sr][sc + 1] - lszx[sr][sc - 2]));
            szx_z = dzi * (hc1 * (lszx[sr][sc] - lszx[sr - 1][sc]) + hc2 * (lszx[sr + 1][sc] - lszx[sr - 2][sc]));
            szz_z = dzi * (hc1 * (lszz[sr + 1][tx] - lszz[sr][tx]) + hc2 * (lszz[sr + 2][tx] - lszz[sr - 1][tx]));
        }
        if (pml_blk) {
            const size_t xo = (size_t)slot * p.xst, zo = (size_t)slot * p.zst;
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
        const T vxn = rvx[k] + dt * rrxp[k] * (sxx_x + szx_z);
        const T vzn = rvz[k] + dt * rrzp[k] * (szx_x + szz_z);
        if (energy) stg(WeS, boff(iz, ix), rwe[k] + (vxn * vxn + vzn * vzn));
        rvx[k] = vxn;
        rvz[k] = vzn;
    }

    // ---- point sources, then receivers (forward_kernel_PSV_GPU.cu:1035-1063);
    // a tile with sources or receivers goes through LDS, the others store
    // their registers directly
    const int g0 = lmeta[0], g1 = lmeta[1];
    const int r0 = lmeta[2], r1 = lmeta[3];
    if (g1 > g0 || r1 > r0) {
#pragma unroll
        for (int k = 0; k < NR; k++) {
            lvx[ty + k * kRowsPerWave][tx] = rvx[k];
            lvz[ty + k * kRowsPerWave][tx] = rvz[k];
        }
        __syncthreads();
        for (int gi = g0 + tid; gi < g1; gi += kThreads) {
            const int* e = p.inj_grp + 5 * (size_t)p.tinj_idx[gi];
            const int comp = e[2];
            if (comp != 1 && comp != 2) continue;  // stress sources: psv_stress_source_kernel
            const int rz = e[0] - z0, rx = e[1] - x0;
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
            const int rz = e[0] - z0, rx = e[1] - x0;
            const T v = e[2] == 1 ? lvx[rz][rx] : lvz[rz][rx];
            p.rec[p.rec_out_off[slot] + (size_t)(rid - p.rec_first[slot]) * nt_s + it] = v;
        }
#pragma unroll
        for (int k = 0; k < NR; k++) {
            rvx[k] = lvx[ty + k * kRowsPerWave][tx];
            rvz[k] = lvz[ty + k * kRowsPerWave][tx];
        }
    }

    // ---- write the new velocities, and the mirrored stresses of the ghost rows
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int rr = ty + k * kRowsPerWave;
        const int iz = z0 + rr, ix = x0 + tx;
        if (iz >= nzt || ix >= nxt) continue;
        const unsigned o = boff(iz, ix);
        stg(vxS, o, rvx[k]);
        stg(vzS, o, rvz[k]);
        if (top && iz < fpad && ix >= fpad && ix < nxt - fpad) {
            stg(szxS, o, lszx[rr + kHalo][tx + kHalo]);
            stg(szzS, o, lszz[rr + kHalo][tx]);
        }
    }
}

// Explosive (stress) sources after the velocity update: one thread per
// source cell of the active shots (forward_kernel_PSV_GPU.cu:1035-1047 adds
// sources at this point of the step; the stress type is an extension).
template <typename T>
__global__ void psv_stress_source_kernel(const StepParams<T> p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.nexpl) return;
    const int gid = p.expl_grp[2 * i], slot = p.expl_grp[2 * i + 1] - p.slot0;
    if (slot < 0 || slot >= p.nslots) return;
    const int it = p.it_dev ? (*p.it_dev + p.it) : p.it;
    const int nt_s = p.shot_nt[slot];
    if (it >= nt_s) return;
    const int* e = p.inj_grp + 5 * (size_t)gid;
    const size_t g = (size_t)slot * p.fst + (size_t)e[0] * p.pitch + e[1];
    const T* sig = p.sig + p.sig_off[slot] + it;
    T a = p.sxx[g], b = p.szz[g];
    for (int q = 0; q < e[4]; q++) {
        const T s = sig[(size_t)p.inj_row[e[3] + q] * nt_s];
        a = a + s;
        b = b + s;
    }
    p.sxx[g] = a;
    p.szz[g] = b;
}

template <typename T>
hipError_t launch_stress_sources(const StepParams<T>& p, hipStream_t s)
{
    if (p.nexpl <= 0) return hipSuccess;
    hipLaunchKernelGGL((psv_stress_source_kernel<T>), dim3((p.nexpl + 63) / 64), dim3(64), 0, s, p);
    return hipGetLastError();
}

template hipError_t launch_stress_sources<float>(const StepParams<float>&, hipStream_t);
template hipError_t launch_stress_sources<double>(const StepParams<double>&, hipStream_t);
template hipError_t launch_stress_sources<dfloat>(const StepParams<dfloat>&, hipStream_t);
template hipError_t launch_stress_sources<ffloat>(const StepParams<ffloat>&, hipStream_t);

template <typename T, int ORDER, int MODE, int TZ>
static hipError_t launch_one(const StepParams<T>& p, hipStream_t s, hipEvent_t e_mid)
{
    const int blocks = p.tiles_x * p.tiles_z * p.nslots;
    hipLaunchKernelGGL((psv_stress_kernel<T, ORDER, MODE, TZ>), dim3(blocks), dim3(kThreads), 0, s, p);
    if (e_mid) {
        const hipError_t e = hipEventRecord(e_mid, s);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((psv_velocity_kernel<T, ORDER, TZ>), dim3(blocks), dim3(kThreads), 0, s, p);
    if (p.nexpl > 0)
        hipLaunchKernelGGL((psv_stress_source_kernel<T>), dim3((p.nexpl + 63) / 64), dim3(64), 0, s, p);
    return hipGetLastError();
}

template <typename T, int ORDER, int MODE, int TZ>
static hipError_t launch_half_one(const StepParams<T>& p, int which, hipStream_t s)
{
    const int blocks = p.tiles_x * p.tiles_z * p.nslots;
    if (which == 0)
        hipLaunchKernelGGL((psv_stress_kernel<T, ORDER, MODE, TZ>), dim3(blocks), dim3(kThreads), 0, s, p);
    else
        hipLaunchKernelGGL((psv_velocity_kernel<T, ORDER, TZ>), dim3(blocks), dim3(kThreads), 0, s, p);
    return hipGetLastError();
}

// The window cells of the active shots' current state into store plane 0,
// the values the forward stress kernel stores: v / dt and the stresses.
template <typename T>
__global__ void store_plane_kernel(const StepParams<T> p, int nzw, int nxw)
{
    const int slot = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nzw * nxw) return;
    const int zf = i / nxw, xf = i - zf * nxw;
    const int iz = p.fwi_z1 + zf * p.fwi_dz, ix = p.fwi_x1 + xf * p.fwi_dx;
    const size_t c = (size_t)slot * p.fst + (size_t)iz * p.pitch + ix;
    const size_t plane = (size_t)p.nfz * p.nfx;
    T* st = p.store + (size_t)slot * p.sst + (size_t)zf * p.nfx + xf;
    st[0 * plane] = p.vx[c] / p.dt;
    st[1 * plane] = p.vz[c] / p.dt;
    st[2 * plane] = p.sxx[c];
    st[3 * plane] = p.szx[c];
    st[4 * plane] = p.szz[c];
}

template <typename T>
hipError_t launch_store_plane(const StepParams<T>& p, hipStream_t s)
{
    const int nzw = (p.fwi_z2 - p.fwi_z1 + p.fwi_dz - 1) / p.fwi_dz;
    const int nxw = (p.fwi_x2 - p.fwi_x1 + p.fwi_dx - 1) / p.fwi_dx;
    if (nzw <= 0 || nxw <= 0 || p.nslots <= 0) return hipSuccess;
    hipLaunchKernelGGL((store_plane_kernel<T>), dim3((nzw * nxw + 255) / 256, p.nslots), dim3(256), 0, s, p,
                       nzw, nxw);
    return hipGetLastError();
}

template hipError_t launch_store_plane<float>(const StepParams<float>&, hipStream_t);
template hipError_t launch_store_plane<double>(const StepParams<double>&, hipStream_t);
template hipError_t launch_store_plane<dfloat>(const StepParams<dfloat>&, hipStream_t);
template hipError_t launch_store_plane<ffloat>(const StepParams<ffloat>&, hipStream_t);

#define PSV_HALF(O, M, Z) \
    if (order == O && mode == M && tz == Z) return launch_half_one<T, O, M, Z>(p, which, s);

template <typename T>
hipError_t launch_half(const StepParams<T>& p, int order, int mode, int tz, int which, hipStream_t s)
{
    PSV_HALF(2, kForward, 4) PSV_HALF(2, kAdjoint, 4)
    PSV_HALF(4, kForward, 4) PSV_HALF(4, kAdjoint, 4)
    PSV_HALF(2, kForward, 8) PSV_HALF(2, kAdjoint, 8)
    PSV_HALF(4, kForward, 8) PSV_HALF(4, kAdjoint, 8)
    PSV_HALF(2, kForward, 16) PSV_HALF(2, kAdjoint, 16)
    PSV_HALF(4, kForward, 16) PSV_HALF(4, kAdjoint, 16)
    PSV_HALF(2, kForward, 32) PSV_HALF(2, kAdjoint, 32)
    PSV_HALF(4, kForward, 32) PSV_HALF(4, kAdjoint, 32)
    return hipErrorInvalidValue;
}
#undef PSV_HALF

template hipError_t launch_half<float>(const StepParams<float>&, int, int, int, int, hipStream_t);
template hipError_t launch_half<double>(const StepParams<double>&, int, int, int, int, hipStream_t);
template hipError_t launch_half<dfloat>(const StepParams<dfloat>&, int, int, int, int, hipStream_t);
template hipError_t launch_half<ffloat>(const StepParams<ffloat>&, int, int, int, int, hipStream_t);

#define PSV_CASE(O, M, Z) \
    if (order == O && mode == M && tz == Z) return launch_one<T, O, M, Z>(p, s, e_mid);

template <typename T>
hipError_t launch_step(const StepParams<T>& p, int order, int mode, int tz, hipStream_t s,
                       hipEvent_t e_mid)
{
    PSV_CASE(2, kForward, 4) PSV_CASE(2, kAdjoint, 4)
    PSV_CASE(4, kForward, 4) PSV_CASE(4, kAdjoint, 4)
    PSV_CASE(2, kForward, 8) PSV_CASE(2, kAdjoint, 8)
    PSV_CASE(4, kForward, 8) PSV_CASE(4, kAdjoint, 8)
    PSV_CASE(2, kForward, 16) PSV_CASE(2, kAdjoint, 16)
    PSV_CASE(4, kForward, 16) PSV_CASE(4, kAdjoint, 16)
    PSV_CASE(2, kForward, 32) PSV_CASE(2, kAdjoint, 32)
    PSV_CASE(4, kForward, 32) PSV_CASE(4, kAdjoint, 32)
    return hipErrorInvalidValue;
}
#undef PSV_CASE

template hipError_t launch_step<float>(const StepParams<float>&, int, int, int, hipStream_t, hipEvent_t);
template hipError_t launch_step<double>(const StepParams<double>&, int, int, int, hipStream_t, hipEvent_t);
template hipError_t launch_step<dfloat>(const StepParams<dfloat>&, int, int, int, hipStream_t, hipEvent_t);
template hipError_t launch_step<ffloat>(const StepParams<ffloat>&, int, int, int, hipStream_t, hipEvent_t);

}  // namespace psv
