// psv_march.hip -- one time step of the 2D elastic P-SV scheme as a single
// launch in which each workgroup marches down a column segment, for gfx950
// (CDNA4, wave64).
//
// A workgroup (256 threads, 4 waves) owns 64 columns x S rows of one shot.
// It walks the segment from the top in iterations of 4 rows, one row per wave:
//   iteration i: load old vx, vz of 4 new rows into an LDS ring of rows;
//                compute the new stresses of 4 rows (the stress front); then
//                compute the new velocities of the 4 rows two rows behind the
//                front, whose stress stencil (rows -2 .. +2) is now complete.
// Every old value is read from HBM once and every new value written once; the
// stresses of a row are computed once, in LDS, and used by the velocity rows
// below and above it.  What the segment needs from outside its 64 columns
// (the stresses of 2 columns on each side, for the velocity x derivatives) is
// computed once per segment in a prologue and kept in LDS; the two stress
// rows above and below the segment are recomputed (not written).  So an HBM
// pass moves the algorithm's 60 bytes per cell (5 fields read and written, 5
// model arrays read) plus the segment edges, where the split step moves about
// 80 (each half reads the other half's fields).
//
// Neighbouring segments read the old values of cells a segment rewrites, so
// the five wavefields and the stress-side CPML memory are read from one
// buffer set (vx.., mvxx..) and written to the other (vx_w.., mvxx_w..), as
// the fused tile kernel does; every in-grid cell of a segment is written,
// updated or not.  Point sources and receivers act after the launch
// (psv_point_kernel), on the written set.
//
// The arithmetic is that of psv_step.hip (forward_kernel_PSV /
// adjoint_kernel_PSV, CUDA_week2/forward_kernel_PSV_GPU.cu:803-1063,
// CUDA_week2/adjoint_kernel_PSV_GPU.cu:818-1087) in the same operation order:
// with contraction off the results equal the split kernels' bit for bit.
#include <hip/hip_runtime.h>

#include "psv_step.h"

namespace psv {

namespace {

template <typename T>
__device__ __forceinline__ T ldm(const T* base, unsigned boff)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}

template <typename T>
__device__ __forceinline__ void stm(T* base, unsigned boff, T v)
{
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff) = v;
}

__device__ __forceinline__ int xcd_spread(int bid, int total)
{
    const int q = total / kNumXcd, r = total % kNumXcd;
    const int x = bid % kNumXcd, k = bid / kNumXcd;
    return x * q + (x < r ? x : r) + k;
}

// CPML profile pointers and the strip geometry of a step
template <typename T>
struct Pml {
    const T *a, *b, *K, *ah, *bh, *Kh;
};

}  // namespace

constexpr int kMarchR = 4;     // rows per iteration, one per wave
constexpr int kMarchVB = 16;   // rows of the velocity ring (power of two)
constexpr int kMarchSB = 8;    // rows of the stress ring (power of two)
constexpr int kMarchW = kTX + 4;  // ring columns x0-2 .. x0+65

template <typename T, int S>
struct MarchLds {
    T lvx[kMarchVB][kMarchW];
    T lvz[kMarchVB][kMarchW];
    T lsxx[kMarchSB][kMarchW];
    T lszx[kMarchSB][kMarchW];
    T lszz[kMarchSB][kTX];
    T lring[S][6];  // sxx at x0-1, x0+64, x0+65; szx at x0-2, x0-1, x0+64
};

// A segment is interior when every stress cell it computes (rows zs-2 .. ze+1, columns
// x0-2 .. x0+65) is an updated cell outside every CPML strip and below the free-surface
// row: then none of the boundary code runs (EDGE = false below).
template <int ORDER>
__device__ __forceinline__ bool march_interior(int zs, int ze, int x0, int nzt, int nxt, int npml, int ppad,
                                               bool fsurf)
{
    constexpr int H = ORDER / 2;
    const int zlo = fsurf ? H + 1 : (npml > 0 ? ppad + 1 : H);
    const int zhi = npml > 0 ? nzt - ppad - 1 : nzt - H;
    const int xlo = npml > 0 ? ppad + 1 : H;
    const int xhi = npml > 0 ? nxt - ppad - 1 : nxt - H;
    return zs - 2 >= zlo && ze + 2 <= zhi && x0 - 2 >= xlo && x0 + kTX + 2 <= xhi;
}

template <typename T, int ORDER, int MODE, int S, bool EDGE>
__device__ __forceinline__ void march_segment(const StepParams<T>& p, MarchLds<T, S>& L, int slot, int zs,
                                              int x0, int it)
{
    constexpr int H = ORDER / 2;       // fpad
    constexpr int R = kMarchR;
    constexpr int VB = kMarchVB;
    constexpr int SB = kMarchSB;
    constexpr int NRING = 6 * S;       // prologue cells: 3 sxx and 3 szx columns per row
    static_assert(S % R == 0, "segment rows");
    auto& lvx = L.lvx;
    auto& lvz = L.lvz;
    auto& lsxx = L.lsxx;
    auto& lszx = L.lszx;
    auto& lszz = L.lszz;
    auto& lring = L.lring;

    const int nzt = p.nzt, nxt = p.nxt, pitch = p.pitch, org = p.org;
    const int ze = min(zs + S, nzt);
    const int fpad = H, ppad = p.ppad;
    const int tid = threadIdx.x;
    const int tx = tid & (kTX - 1), wv = tid / kTX;
    const bool fsurf = EDGE && p.fsurf != 0;
    const bool energy = (p.flags & kEnergy) != 0;
    const bool fwi_step = (p.flags & kStore) && (it % p.fwi_dt == 0);
    const bool unit_stride = p.fwi_dz == 1 && p.fwi_dx == 1;
    const int tf_cur = it / p.fwi_dt - p.tf0;
    const int npml = EDGE ? p.npml : 0, tp = p.tp;
    const T dt = p.dt, dxi = p.dxi, dzi = p.dzi, hc1 = p.hc1, hc2 = p.hc2;
    const Pml<T> pm{p.pml, p.pml + tp, p.pml + 2 * tp, p.pml + 3 * tp, p.pml + 4 * tp, p.pml + 5 * tp};
    const size_t xo = (size_t)slot * p.xst, zo = (size_t)slot * p.zst;
    // strips this segment's columns touch (stress side: left strip ix <= ppad)
    const bool xpml = npml > 0 && (x0 - 3 <= ppad || x0 + kTX + 2 >= nxt - ppad - 1);

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
    auto interior = [&](int gz, int gx) {
        return !EDGE || (gz >= fpad && gz < nzt - fpad && gx >= fpad && gx < nxt - fpad);
    };

    // ---- the CPML of the stress half-step: x strips (left ix <= ppad, then right) and z
    // strips (top iz <= ppad without a free surface, then bottom), each applied in turn as
    // the split kernel does; memory read from the current set, written to the other by the
    // cell's owner only.  d / dh: the derivative on full / half profile points.
    auto pml_x = [&](int gz, int gx, T& d, T& dh, bool sx, bool sh, bool own) {
        auto one = [&](int px) {
            const size_t q = xo + (size_t)gz * tp + px;
            if (sx) {
                const T m1 = pm.b[px] * p.mvxx[q] + pm.a[px] * d;
                if (own) p.mvxx_w[q] = m1;
                d = d / pm.K[px] + m1;
            }
            if (sh) {
                const T m2 = pm.bh[px] * p.mvzx[q] + pm.ah[px] * dh;
                if (own) p.mvzx_w[q] = m2;
                dh = dh / pm.Kh[px] + m2;
            }
        };
        if (gx <= ppad) one(gx - fpad);
        if (gx >= nxt - ppad - 1) one(gx - p.pnx);
    };
    auto pml_z = [&](int gz, int gx, T& d, T& dh, bool sz, bool sh, bool own) {
        auto one = [&](int pz) {
            const size_t q = zo + (size_t)pz * pitch + gx;
            if (sz) {
                const T m1 = pm.b[pz] * p.mvzz[q] + pm.a[pz] * d;
                if (own) p.mvzz_w[q] = m1;
                d = d / pm.K[pz] + m1;
            }
            if (sh) {
                const T m2 = pm.bh[pz] * p.mvxz[q] + pm.ah[pz] * dh;
                if (own) p.mvxz_w[q] = m2;
                dh = dh / pm.Kh[pz] + m2;
            }
        };
        if (gz <= ppad && !fsurf) one(gz - fpad);
        if (gz >= nzt - ppad - 1) one(gz - p.pnz);
    };

    // ---- prologue: the stresses of the 2 + 2 columns beside the segment (rows zs .. ze-1),
    // from the old velocities in memory (L2 hits: the neighbour strips load them too)
    for (int c = tid; c < NRING; c += kThreads) {
        const int r = c / 6, k = c - 6 * r;
        const int gz = zs + r;
        if (gz >= ze) continue;
        const bool is_sxx = k < 3;
        const int gx = is_sxx ? (k == 0 ? x0 - 1 : x0 + 63 + k) : (k == 3 ? x0 - 2 : (k == 4 ? x0 - 1 : x0 + 64));
        const unsigned o = boff(gz, gx);
        T v = is_sxx ? ldm(sxxA, o) : ldm(szxA, o);
        if (interior(gz, gx)) {
            if (is_sxx) {
                T vx_x, vz_z;
                if (ORDER == 2) {
                    vx_x = dxi * hc1 * (ldm(vxA, o) - ldm(vxA, boff(gz, gx - 1)));
                    vz_z = dzi * hc1 * (ldm(vzA, o) - ldm(vzA, boff(gz - 1, gx)));
                } else {
                    vx_x = dxi * (hc1 * (ldm(vxA, o) - ldm(vxA, boff(gz, gx - 1))) +
                                  hc2 * (ldm(vxA, boff(gz, gx + 1)) - ldm(vxA, boff(gz, gx - 2))));
                    vz_z = dzi * (hc1 * (ldm(vzA, o) - ldm(vzA, boff(gz - 1, gx))) +
                                  hc2 * (ldm(vzA, boff(gz + 1, gx)) - ldm(vzA, boff(gz - 2, gx))));
                }
                if (npml > 0) {
                    T unused = 0;
                    pml_x(gz, gx, vx_x, unused, true, false, false);
                    pml_z(gz, gx, vz_z, unused, true, false, false);
                }
                const T lam = ldm(lamM, o), mu = ldm(muM, o);
                const T old = v;
                v = old + dt * (lam * (vx_x + vz_z) + ((T)2.0 * mu * vx_x));
                if (fsurf && gz == fpad) {
                    if (ORDER == 2) v = (T)4.0 * dt * vx_x * (lam * mu + mu * mu) / (lam + (T)2.0 * mu);
                    else v = old + (T)4.0 * dt * vx_x * (lam * mu + mu * mu) / (lam + (T)2.0 * mu);
                }
            } else {
                T vz_x, vx_z;
                if (ORDER == 2) {
                    vz_x = dxi * hc1 * (ldm(vzA, boff(gz, gx + 1)) - ldm(vzA, o));
                    vx_z = dzi * hc1 * (ldm(vxA, boff(gz + 1, gx)) - ldm(vxA, o));
                } else {
                    vz_x = dxi * (hc1 * (ldm(vzA, boff(gz, gx + 1)) - ldm(vzA, o)) +
                                  hc2 * (ldm(vzA, boff(gz, gx + 2)) - ldm(vzA, boff(gz, gx - 1))));
                    vx_z = dzi * (hc1 * (ldm(vxA, boff(gz + 1, gx)) - ldm(vxA, o)) +
                                  hc2 * (ldm(vxA, boff(gz + 2, gx)) - ldm(vxA, boff(gz - 1, gx))));
                }
                if (npml > 0) {
                    T unused = 0;
                    pml_x(gz, gx, unused, vz_x, false, true, false);
                    pml_z(gz, gx, unused, vx_z, false, true, false);
                }
                v = v + dt * ldm(mzxM, o) * (vz_x + vx_z);
                if (fsurf && gz == fpad && ORDER == 2) v = (T)0.0;
            }
        }
        lring[r][k] = v;
    }

    // ---- operands of one iteration, loaded one iteration ahead
    // the stress front of iteration i is rows zr .. zr+3, zr = zs - 2 + 4 i; the velocity
    // rows are two rows behind; the velocity ring receives rows zr+2 .. zr+5
    const int niter = (ze - zs + 3 + 4) / R;  // until the velocity rows pass ze - 1
    struct Ops {
        T vx, vz, vxh, vzh;                    // ring row: own column and one halo column
        T sxx, szx, szz, lam, mu, mzx;         // stress row (own column)
        T ux, uz;                              // adjoint displacement (stress row)
        T rxp, rzp, we;                        // velocity row
    };
    const int hc = tx < 2 ? x0 - 2 + tx : x0 + 62 + tx;  // halo column loaded by lanes 0..3
    // Every load of an iteration is issued on every path, so that the number of memory
    // operations between a register set's loads and its use is the same on all paths and
    // the wait before the use leaves the next set's loads in flight; nothing is computed
    // from a loaded value here (a select would wait for it at once).  A row outside the rows
    // an iteration uses is read at the nearest row it may read (velocity ring rows up to
    // ze + 3, stress rows up to ze + 1: the bottom segment's later rows would lie beyond the
    // guard rows; the segment's own rows for the rest), and its value is never used: stress
    // rows below ze + 1 and their velocity ring rows feed no velocity row, and the other
    // operands are used for the segment's own rows only.  Lanes 4..63 read their own column
    // in place of a halo column.
    auto load_ops = [&](int i, Ops& q) {
        const int zr = zs - 2 + R * i;
        {
            const int gz = min(zr + 2 + wv, ze + 3);
            const unsigned o = boff(gz, x0 + tx);
            const unsigned oh = tx < 4 ? boff(gz, hc) : o;
            q.vx = ldm(vxA, o);
            q.vz = ldm(vzA, o);
            q.vxh = ldm(vxA, oh);
            q.vzh = ldm(vzA, oh);
        }
        {
            const int gz = min(zr + wv, ze + 1);
            const unsigned o = boff(gz, x0 + tx);
            q.sxx = ldm(sxxA, o);
            q.szx = ldm(szxA, o);
            q.szz = ldm(szzA, o);
            q.lam = ldm(lamM, o);
            q.mu = ldm(muM, o);
            q.mzx = ldm(mzxM, o);
            if (MODE == kAdjoint) {
                const unsigned oo = boff(max(zs, min(zr + wv, ze - 1)), x0 + tx);
                q.ux = ldm((const T*)uxS, oo);
                q.uz = ldm((const T*)uzS, oo);
            }
        }
        {
            const unsigned o = boff(max(zs, min(zr - 2 + wv, ze - 1)), x0 + tx);
            q.rxp = ldm(rxpM, o);
            q.rzp = ldm(rzpM, o);
            if (energy) q.we = ldm((const T*)WeS, o);
        }
    };

    // velocity ring rows zs-4 .. zs-1 (before the first iteration)
    {
        const int gz = zs - 4 + wv;
        const int rs = gz & (VB - 1);
        const unsigned o = boff(gz, x0 + tx);
        lvx[rs][2 + tx] = ldm(vxA, o);
        lvz[rs][2 + tx] = ldm(vzA, o);
        if (tx < 4) {
            const int lc = tx < 2 ? tx : 64 + tx;
            const unsigned oh = boff(gz, hc);
            lvx[rs][lc] = ldm(vxA, oh);
            lvz[rs][lc] = ldm(vzA, oh);
        }
    }
    // operands are loaded two iterations ahead (two register sets in flight), so a
    // workgroup waits on memory once per two iterations at most
    Ops qa{}, qb{};
    load_ops(0, qa);
    load_ops(1, qb);  // unconditional (its rows are clamped): see the loop below
    auto iteration = [&](int i, const Ops& cu) {
        const int zr = zs - 2 + R * i;
        {
            const int rs = (zr + 2 + wv) & (VB - 1);
            lvx[rs][2 + tx] = cu.vx;
            lvz[rs][2 + tx] = cu.vz;
            if (tx < 4) {
                const int lc = tx < 2 ? tx : 64 + tx;
                lvx[rs][lc] = cu.vxh;
                lvz[rs][lc] = cu.vzh;
            }
        }
        __syncthreads();

        // ---- stress row gz = zr + wv
        {
            const int gz = zr + wv, gx = x0 + tx;
            const int vr = gz & (VB - 1), vc = 2 + tx;
            const bool own = gz >= zs && gz < ze;
            const bool ingrid = !EDGE || (gz >= 0 && gz < nzt && gx < nxt);
            // ghost row above the free surface: its szz and szx come from the mirror below
            const bool mcol = gx >= fpad && gx < nxt - fpad;
            const bool ghost = fsurf && gz >= 0 && gz < fpad && mcol;
            T sxxn = cu.sxx, szxn = cu.szx, szzn = cu.szz;
            // work on the old state of the segment's own cells
            if (own && ingrid && (MODE == kAdjoint || fwi_step)) {
                T uxn = 0, uzn = 0;
                const unsigned o = boff(gz, gx);
                if (MODE == kAdjoint) {
                    uzn = cu.uz;
                    uxn = cu.ux;
                    if (interior(gz, gx)) {
                        uzn = uzn + dt * lvz[vr][vc];
                        uxn = uxn + dt * lvx[vr][vc];
                        stm(uzS, o, uzn);
                        stm(uxS, o, uxn);
                    }
                }
                const int dzw = gz - p.fwi_z1, dxw = gx - p.fwi_x1;
                const bool on = unit_stride ? true : (dzw % p.fwi_dz == 0 && dxw % p.fwi_dx == 0);
                if (fwi_step && dzw >= 0 && dxw >= 0 && gz < p.fwi_z2 && gx < p.fwi_x2 && on) {
                    const size_t plane = (size_t)p.nfz * p.nfx;
                    const int zf = unit_stride ? dzw : dzw / p.fwi_dz, xf = unit_stride ? dxw : dxw / p.fwi_dx;
                    const size_t w = (size_t)zf * p.nfx + xf;
                    T* st = p.store + (size_t)slot * p.sst + (size_t)tf_cur * 5 * plane + w;
                    if (MODE == kForward) {
                        st[0 * plane] = lvx[vr][vc] / dt;
                        st[1 * plane] = lvz[vr][vc] / dt;
                        st[2 * plane] = cu.sxx;
                        st[3 * plane] = cu.szx;
                        st[4 * plane] = cu.szz;
                    } else {
                        const T fvx = st[0 * plane], fvz = st[1 * plane];
                        const T fsxx = st[2 * plane], fszx = st[3 * plane], fszz = st[4 * plane];
                        const T lam = cu.lam, mu = cu.mu;
                        const T lm = lam + mu;
                        const T s1 = (fsxx + fszz) * (cu.sxx + cu.szz) * (T)0.25 / (lm * lm);
                        const T s2 = (fsxx - fszz) * (cu.sxx - cu.szz) / (mu * mu);
                        const T s3 = (fszx * cu.szx) / (mu * mu);
                        const T s4 = uxn * fvx + uzn * fvz;
                        const T wdt = (T)p.fwi_dt * dt;
                        T* gr = p.grad + (size_t)slot * p.gst + w;
                        gr[0 * plane] = gr[0 * plane] + wdt * s1;
                        gr[1 * plane] = gr[1 * plane] + wdt * (s3 + s1 + s2);
                        gr[2 * plane] = gr[2 * plane] - wdt * s4;
                    }
                }
            }
            if (interior(gz, gx)) {
                T vx_x, vz_x, vx_z, vz_z;
                if (ORDER == 2) {
                    vx_x = dxi * hc1 * (lvx[vr][vc] - lvx[vr][vc - 1]);
                    vz_x = dxi * hc1 * (lvz[vr][vc + 1] - lvz[vr][vc]);
                    vx_z = dzi * hc1 * (lvx[(gz + 1) & (VB - 1)][vc] - lvx[vr][vc]);
                    vz_z = dzi * hc1 * (lvz[vr][vc] - lvz[(gz - 1) & (VB - 1)][vc]);
                } else {
                    vx_x = dxi * (hc1 * (lvx[vr][vc] - lvx[vr][vc - 1]) + hc2 * (lvx[vr][vc + 1] - lvx[vr][vc - 2]));
                    vz_x = dxi * (hc1 * (lvz[vr][vc + 1] - lvz[vr][vc]) + hc2 * (lvz[vr][vc + 2] - lvz[vr][vc - 1]));
                    vx_z = dzi * (hc1 * (lvx[(gz + 1) & (VB - 1)][vc] - lvx[vr][vc]) +
                                  hc2 * (lvx[(gz + 2) & (VB - 1)][vc] - lvx[(gz - 1) & (VB - 1)][vc]));
                    vz_z = dzi * (hc1 * (lvz[vr][vc] - lvz[(gz - 1) & (VB - 1)][vc]) +
                                  hc2 * (lvz[(gz + 1) & (VB - 1)][vc] - lvz[(gz - 2) & (VB - 1)][vc]));
                }
                if (npml > 0) {
                    if (xpml) pml_x(gz, gx, vx_x, vz_x, true, true, own);
                    pml_z(gz, gx, vz_z, vx_z, true, true, own);
                }
                const T lam = cu.lam, mu = cu.mu;
                szxn = cu.szx + dt * cu.mzx * (vz_x + vx_z);
                sxxn = cu.sxx + dt * (lam * (vx_x + vz_z) + ((T)2.0 * mu * vx_x));
                szzn = cu.szz + dt * (lam * (vx_x + vz_z) + ((T)2.0 * mu * vz_z));
                if (fsurf && gz == fpad) {
                    szzn = (T)0.0;
                    if (ORDER == 2) {
                        szxn = (T)0.0;
                        sxxn = (T)4.0 * dt * vx_x * (lam * mu + mu * mu) / (lam + (T)2.0 * mu);
                    } else {
                        sxxn = cu.sxx + (T)4.0 * dt * vx_x * (lam * mu + mu * mu) / (lam + (T)2.0 * mu);
                    }
                }
            }
            const int sr = gz & (SB - 1);
            lsxx[sr][2 + tx] = sxxn;
            lszx[sr][2 + tx] = szxn;
            lszz[sr][tx] = szzn;
            if (own && tx < 4) {
                // the stresses beside the segment, from the prologue
                const int r = gz - zs;
                if (tx == 0) lszx[sr][0] = lring[r][3];
                else if (tx == 1) {
                    lsxx[sr][1] = lring[r][0];
                    lszx[sr][1] = lring[r][4];
                } else if (tx == 2) {
                    lsxx[sr][66] = lring[r][1];
                    lszx[sr][66] = lring[r][5];
                } else {
                    lsxx[sr][67] = lring[r][2];
                }
            }
            // ghost rows above the free surface hold the mirror of the new stresses
            // (CUDA_week2/forward_kernel_PSV_GPU.cu:935-942): szz about row fpad, szx
            // about row fpad at order 2 and about fpad - 1/2 at order 4
            T gszz = 0, gszx = 0;
            int ggz = -1, ggx = -1;
            if (EDGE && fsurf && zs == 0 && mcol) {
                if (gz > fpad && gz <= 2 * fpad) {
                    ggz = 2 * fpad - gz;
                    lszz[ggz & (SB - 1)][tx] = -szzn;
                }
                const int gx_row = ORDER == 2 ? 2 * fpad - gz : 2 * fpad - gz - 1;
                if ((ORDER == 2 && gz > fpad && gz <= 2 * fpad) || (ORDER == 4 && gz >= fpad && gz < 2 * fpad)) {
                    ggx = gx_row;
                    lszx[ggx & (SB - 1)][2 + tx] = -szxn;
                }
                gszz = -szzn;
                gszx = -szxn;
            }
            if (own && ingrid) {
                const unsigned o = boff(gz, gx);
                stm(sxxB, o, sxxn);
                if (!ghost) {
                    stm(szxB, o, szxn);
                    stm(szzB, o, szzn);
                }
            }
            // ghost rows belong to the top segment; their cells carry the mirrored values
            if (ggz >= 0 && ggz < ze) stm(szzB, boff(ggz, gx), gszz);
            if (ggx >= 0 && ggx < ze) stm(szxB, boff(ggx, gx), gszx);
        }
        __syncthreads();

        // ---- velocity row gz = zr - 2 + wv
        {
            const int gz = zr - 2 + wv, gx = x0 + tx;
            if (gz >= zs && gz < ze && (!EDGE || gx < nxt)) {
                const int vr = gz & (VB - 1), vc = 2 + tx;
                const int sr = gz & (SB - 1), sc = 2 + tx;
                T vxn = lvx[vr][vc], vzn = lvz[vr][vc];
                if (interior(gz, gx)) {
                    T sxx_x, szx_x, szx_z, szz_z;
                    const int zm1 = (gz - 1) & (SB - 1), zp1 = (gz + 1) & (SB - 1);
                    if (ORDER == 2) {
                        sxx_x = dxi * hc1 * (lsxx[sr][sc + 1] - lsxx[sr][sc]);
                        szx_x = dxi * hc1 * (lszx[sr][sc] - lszx[sr][sc - 1]);
                        szx_z = dzi * hc1 * (lszx[sr][sc] - lszx[zm1][sc]);
                        szz_z = dzi * hc1 * (lszz[zp1][tx] - lszz[sr][tx]);
                    } else {
                        const int zm2 = (gz - 2) & (SB - 1), zp2 = (gz + 2) & (SB - 1);
                        sxx_x = dxi * (hc1 * (lsxx[sr][sc + 1] - lsxx[sr][sc]) + hc2 * (lsxx[sr][sc + 2] - lsxx[sr][sc - 1]));
                        szx_x = dxi * (hc1 * (lszx[sr][sc] - lszx[sr][sc - 1]) + hc2 * (lszx[sr][sc + 1] - lszx[sr][sc - 2]));
                        szx_z = dzi * (hc1 * (lszx[sr][sc] - lszx[zm1][sc]) + hc2 * (lszx[zp1][sc] - lszx[zm2][sc]));
                        szz_z = dzi * (hc1 * (lszz[zp1][tx] - lszz[sr][tx]) + hc2 * (lszz[zp2][tx] - lszz[zm1][tx]));
                    }
                    if (npml > 0) {
                        // velocity-side CPML memory is touched by its owner only: in place
                        auto vx_strip = [&](int px) {
                            const size_t q = xo + (size_t)gz * tp + px;
                            const T m1 = pm.b[px] * p.msxxx[q] + pm.a[px] * sxx_x;
                            const T m2 = pm.bh[px] * p.mszxx[q] + pm.ah[px] * szx_x;
                            p.msxxx[q] = m1;
                            p.mszxx[q] = m2;
                            sxx_x = sxx_x / pm.K[px] + m1;
                            szx_x = szx_x / pm.Kh[px] + m2;
                        };
                        auto vz_strip = [&](int pz) {
                            const size_t q = zo + (size_t)pz * pitch + gx;
                            const T m1 = pm.b[pz] * p.mszzz[q] + pm.a[pz] * szz_z;
                            const T m2 = pm.bh[pz] * p.mszxz[q] + pm.ah[pz] * szx_z;
                            p.mszzz[q] = m1;
                            p.mszxz[q] = m2;
                            szz_z = szz_z / pm.K[pz] + m1;
                            szx_z = szx_z / pm.Kh[pz] + m2;
                        };
                        if (xpml) {
                            if (gx < ppad) vx_strip(gx - fpad);              // left strip (strict)
                            if (gx >= nxt - ppad - 1) vx_strip(gx - p.pnx);  // right strip
                        }
                        if (gz < ppad && !fsurf) vz_strip(gz - fpad);        // top strip (strict)
                        if (gz >= nzt - ppad - 1) vz_strip(gz - p.pnz);      // bottom strip
                    }
                    vxn = vxn + dt * cu.rxp * (sxx_x + szx_z);
                    vzn = vzn + dt * cu.rzp * (szx_x + szz_z);
                    if (energy) stm(WeS, boff(gz, gx), cu.we + (vxn * vxn + vzn * vzn));
                }
                const unsigned o = boff(gz, gx);
                stm(vxB, o, vxn);
                stm(vzB, o, vzn);
            }
        }
    };
    // Each register set is consumed, then refilled for the iteration two ahead: no set is
    // copied while its loads are in flight (a copy of registers still being loaded makes the
    // compiler wait for every outstanding load at the copy), so the loads for i + 2 stay in
    // flight through iteration i + 1.  The loop runs while both refills happen, so every
    // path through it issues the same loads (a path that skips one would make the compiler
    // wait for all loads in flight at the loop head); the last one to three iterations
    // follow it.
    int i = 0;
    for (; i + 3 < niter; i += 2) {
        iteration(i, qa);
        load_ops(i + 2, qa);
        iteration(i + 1, qb);
        load_ops(i + 3, qb);
    }
    if (i < niter) {
        iteration(i, qa);
        if (i + 2 < niter) load_ops(i + 2, qa);
    }
    if (i + 1 < niter) iteration(i + 1, qb);
    if (i + 2 < niter) iteration(i + 2, qa);
}

// PATH 0: every segment of the launch, interior ones on the lean path; 1: interior segments
// only; 2: the other segments only.
template <typename T, int ORDER, int MODE, int S, int PATH>
__global__ __launch_bounds__(kThreads) void psv_march_kernel(const StepParams<T> p)
{
    __shared__ MarchLds<T, S> lds;
    const int ntile = p.tiles_x * p.tiles_z;
    const int bid = xcd_spread(blockIdx.x, ntile * p.nslots);
    const int slot = bid / ntile, tile = bid - slot * ntile;
    const int tzi = tile / p.tiles_x;
    const int zs = tzi * S, x0 = (tile - tzi * p.tiles_x) * kTX;
    const int it = p.it_dev ? (*p.it_dev + p.it) : p.it;
    if (it >= p.shot_nt[slot]) return;  // this shot has finished (uniform over the block)
    const int ze = min(zs + S, p.nzt);
    const bool inner = march_interior<ORDER>(zs, ze, x0, p.nzt, p.nxt, p.npml, p.ppad, p.fsurf != 0);
    if (inner) {
        if (PATH != 2) march_segment<T, ORDER, MODE, S, false>(p, lds, slot, zs, x0, it);
    } else {
        if (PATH != 1) march_segment<T, ORDER, MODE, S, true>(p, lds, slot, zs, x0, it);
    }
}

// Point sources, then receivers, of one step on the written buffer set (the
// order of CUDA_week2/forward_kernel_PSV_GPU.cu:1035-1063), after the march
// launch: one workgroup per active slot.  Explosive (stress) sources stay with
// psv_stress_source_kernel.
template <typename T>
__global__ __launch_bounds__(256) void psv_point_kernel(const StepParams<T> p)
{
    const int slot = blockIdx.x;
    const int it = p.it_dev ? (*p.it_dev + p.it) : p.it;
    const int nt_s = p.shot_nt[slot];
    if (it >= nt_s) return;
    const int ntile = p.tiles_x * p.tiles_z;
    const int g0 = p.tinj_off[(size_t)slot * ntile], g1 = p.tinj_off[(size_t)(slot + 1) * ntile];
    const int r0 = p.trec_off[(size_t)slot * ntile], r1 = p.trec_off[(size_t)(slot + 1) * ntile];
    const size_t base = (size_t)slot * p.fst;
    for (int gi = g0 + (int)threadIdx.x; gi < g1; gi += blockDim.x) {
        const int* e = p.inj_grp + 5 * (size_t)p.tinj_idx[gi];
        const int comp = e[2];
        if (comp != 1 && comp != 2) continue;
        const size_t c = base + (size_t)e[0] * p.pitch + e[1];
        T* f = comp == 1 ? p.vx_w : p.vz_w;
        const T* sig = p.sig + p.sig_off[slot] + it;
        T v = f[c];
        for (int q = 0; q < e[4]; q++) v = v + sig[(size_t)p.inj_row[e[3] + q] * nt_s];
        f[c] = v;
    }
    __syncthreads();
    for (int ri = r0 + (int)threadIdx.x; ri < r1; ri += blockDim.x) {
        const int rid = p.trec_idx[ri];
        const int* e = p.rec_zxc + 3 * (size_t)rid;
        const size_t c = base + (size_t)e[0] * p.pitch + e[1];
        p.rec[p.rec_out_off[slot] + (size_t)(rid - p.rec_first[slot]) * nt_s + it] = e[2] == 1 ? p.vx_w[c] : p.vz_w[c];
    }
}

template <typename T, int ORDER, int MODE, int S>
static hipError_t launch_march_one(const StepParams<T>& p, hipStream_t s, const MarchSide* side)
{
    const int blocks = p.tiles_x * p.tiles_z * p.nslots;
    // interior segments on the lean instance (no boundary code: a third of the registers,
    // twice the waves per SIMD) and the segments along the grid's edges on the full one;
    // both read only the old buffer set and write disjoint cells of the new one, so with a
    // second stream they run side by side, and the edge segments' long serial walk overlaps
    // the interior launch
    if (side) {
        hipError_t e = hipEventRecord(side->fork, s);
        if (e == hipSuccess) e = hipStreamWaitEvent(side->stream, side->fork, 0);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((psv_march_kernel<T, ORDER, MODE, S, 2>), dim3(blocks), dim3(kThreads), 0,
                           side->stream, p);
        hipLaunchKernelGGL((psv_march_kernel<T, ORDER, MODE, S, 1>), dim3(blocks), dim3(kThreads), 0, s, p);
        e = hipEventRecord(side->join, side->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, side->join, 0);
        if (e != hipSuccess) return e;
    } else {
        hipLaunchKernelGGL((psv_march_kernel<T, ORDER, MODE, S, 1>), dim3(blocks), dim3(kThreads), 0, s, p);
        hipLaunchKernelGGL((psv_march_kernel<T, ORDER, MODE, S, 2>), dim3(blocks), dim3(kThreads), 0, s, p);
    }
    hipLaunchKernelGGL((psv_point_kernel<T>), dim3(p.nslots), dim3(256), 0, s, p);
    if (p.nexpl > 0) {
        StepParams<T> q = p;
        q.sxx = p.sxx_w;
        q.szz = p.szz_w;
        const hipError_t e = launch_stress_sources<T>(q, s);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

#define PSV_MCASE(O, M, Z) \
    if (order == O && mode == M && rows == Z) return launch_march_one<T, O, M, Z>(p, s, side);

template <typename T>
hipError_t launch_march(const StepParams<T>& p, int order, int mode, int rows, hipStream_t s, const MarchSide* side)
{
    PSV_MCASE(2, kForward, 32) PSV_MCASE(2, kAdjoint, 32)
    PSV_MCASE(4, kForward, 32) PSV_MCASE(4, kAdjoint, 32)
    PSV_MCASE(2, kForward, 64) PSV_MCASE(2, kAdjoint, 64)
    PSV_MCASE(4, kForward, 64) PSV_MCASE(4, kAdjoint, 64)
    return hipErrorInvalidValue;
}
#undef PSV_MCASE

template hipError_t launch_march<float>(const StepParams<float>&, int, int, int, hipStream_t, const MarchSide*);
template hipError_t launch_march<double>(const StepParams<double>&, int, int, int, hipStream_t, const MarchSide*);
template hipError_t launch_march<dfloat>(const StepParams<dfloat>&, int, int, int, hipStream_t, const MarchSide*);
template hipError_t launch_march<ffloat>(const StepParams<ffloat>&, int, int, int, hipStream_t, const MarchSide*);

}  // namespace psv
