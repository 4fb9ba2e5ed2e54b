// fd_setup.cpp -- host-side setup of the P-SV hot path (FP64).
#include "fd_setup.hpp"

#include <cmath>

namespace psvhost {

namespace {

// Holberg (1987) staggered-grid first-derivative coefficients.  Rows: maximum
// relative group-velocity error 0 (Taylor), 0.1%, 0.5%, 1%, 3%; columns: FD
// order 2..12; entry 0 is points per wavelength.  The numbers define the
// scheme, so they are the table of CUDA_week2/holbergcoeff.cu:10-52.
const double kHolberg[5][6][7] = {
    {{23.0, 1.0, 0, 0, 0, 0, 0},
     {8.0, 9.0 / 8.0, -1.0 / 24.0, 0, 0, 0, 0},
     {6.0, 75.0 / 64.0, -25.0 / 384.0, 3.0 / 640.0, 0, 0, 0},
     {5.0, 1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0, 0, 0},
     {5.0, 19845.0 / 16384.0, -735.0 / 8192.0, 567.0 / 40960.0, -405.0 / 229376.0, 35.0 / 294912.0, 0},
     {4.0, 160083.0 / 131072.0, -12705.0 / 131072.0, 22869.0 / 1310720.0, -5445.0 / 1835008.0,
      847.0 / 2359296.0, -63.0 / 2883584.0}},
    {{49.7, 1.0010, 0, 0, 0, 0, 0},
     {8.32, 1.1382, -0.046414, 0, 0, 0, 0},
     {4.77, 1.1965, -0.078804, 0.0081781, 0, 0, 0},
     {3.69, 1.2257, -0.099537, 0.018063, -0.0026274, 0, 0},
     {3.19, 1.2415, -0.11231, 0.026191, -0.0064682, 0.001191, 0},
     {2.91, 1.2508, -0.12034, 0.032131, -0.010142, 0.0029857, -0.00066667}},
    {{22.2, 1.0050, 0, 0, 0, 0, 0},
     {5.65, 1.1534, -0.052806, 0, 0, 0, 0},
     {3.74, 1.2111, -0.088313, 0.011768, 0, 0, 0},
     {3.11, 1.2367, -0.10815, 0.023113, -0.0046905, 0, 0},
     {2.80, 1.2496, -0.11921, 0.031130, -0.0093272, 0.0025161, 0},
     {2.62, 1.2568, -0.12573, 0.036423, -0.013132, 0.0047484, -0.0015979}},
    {{15.8, 1.0100, 0, 0, 0, 0, 0},
     {4.80, 1.1640, -0.057991, 0, 0, 0, 0},
     {3.39, 1.2192, -0.094070, 0.014608, 0, 0, 0},
     {2.90, 1.2422, -0.11269, 0.026140, -0.0064054, 0, 0},
     {2.65, 1.2534, -0.12257, 0.033755, -0.011081, 0.0036784, 0},
     {2.51, 1.2596, -0.12825, 0.038550, -0.014763, 0.0058619, -0.0024538}},
    {{9.16, 1.0300, 0, 0, 0, 0, 0},
     {3.47, 1.1876, -0.072518, 0, 0, 0, 0},
     {2.91, 1.2341, -0.10569, 0.022589, 0, 0, 0},
     {2.61, 1.2516, -0.12085, 0.032236, -0.011459, 0, 0},
     {2.45, 1.2596, -0.12829, 0.038533, -0.014681, 0.0072580, 0},
     {2.36, 1.2640, -0.13239, 0.042217, -0.017803, 0.0081959, -0.0051848}},
};

// damping d, scaling K and frequency shift alpha at one CPML abscissa
struct Profile {
    double d = 0.0, K = 1.0, alpha = 0.0;
};

Profile profile_at(double abscissa, double thickness, double d0, const CpmlParams& p, double alpha_max)
{
    Profile r;
    if (abscissa >= 0) {
        const double xn = abscissa / thickness;
        const double xp = std::pow(xn, p.npower);
        r.d = d0 * xp;
        r.K = 1.0 + (p.k_max - 1.0) * xp;
        r.alpha = alpha_max * (1.0 - xn);
    }
    if (r.alpha < 0.0) r.alpha = 0.0;
    return r;
}

}  // namespace

bool holberg_coefficients(int fdorder, int err, double hc[7])
{
    if (fdorder < 2 || fdorder > 12 || fdorder % 2 || err < 0 || err > 4) return false;
    for (int i = 0; i < 7; i++) hc[i] = kHolberg[err][fdorder / 2 - 1][i];
    return true;
}

std::vector<double> cpml_profiles(const CpmlParams& p, double dt)
{
    const int tp = 2 * (p.npml + 1);
    std::vector<double> out((size_t)6 * tp);
    double* a = out.data();
    double* b = a + tp;
    double* K = b + tp;
    double* ah = K + tp;
    double* bh = ah + tp;
    double* Kh = bh + tp;
    // Collino & Tsogka / Komatitsch & Martin CPML: d0 from the target
    // reflection coefficient, polynomial grading of d and K, alpha linear.
    const double alpha_max = 2.0 * kPi * (p.fpml / 2.0);
    const double thickness = p.npml * p.dh;
    const double d0 = -(p.npower + 1) * p.damp_v * std::log(p.rcoef) / (2.0 * thickness);
    for (int i = 0; i < tp; i++) {
        const double x = p.dh * (double)i;
        // left side measures from the outer edge inwards, right side from npml+1 outwards
        const bool left = i <= p.npml;
        const double full = left ? thickness - x : x - p.dh * (p.npml + 1);
        const double half = left ? thickness - (x + p.dh / 2.0) : (x + p.dh / 2.0) - p.dh * (p.npml + 1);
        const Profile f = profile_at(full, thickness, d0, p, alpha_max);
        const Profile h = profile_at(half, thickness, d0, p, alpha_max);
        K[i] = f.K;
        Kh[i] = h.K;
        b[i] = std::exp(-(f.d / f.K + f.alpha) * dt);
        bh[i] = std::exp(-(h.d / h.K + h.alpha) * dt);
        a[i] = std::fabs(f.d) > 1.0e-6 ? f.d * (b[i] - 1.0) / (f.K * (f.d + f.K * f.alpha)) : 0.0;
        ah[i] = std::fabs(h.d) > 1.0e-6 ? h.d * (bh[i] - 1.0) / (h.K * (h.d + h.K * h.alpha)) : 0.0;
    }
    return out;
}

void velocity_to_lame(double vp, double vs, double rho, double& lam, double& mu)
{
    mu = std::pow(vs, 2) * rho;
    lam = std::pow(vp, 2) * rho - 2.0 * mu;
}

void average_parameters(int nzt, int nxt, const double* mu, const double* rho, double* mu_zx,
                        double* rho_zp, double* rho_xp)
{
    const int nza = nzt - 1, nxa = nxt - 1;
    for (int iz = 0; iz < nza; iz++) {
        const double* m0 = mu + (size_t)iz * nxt;
        const double* m1 = m0 + nxt;
        const double* r0 = rho + (size_t)iz * nxt;
        const double* r1 = r0 + nxt;
        for (int ix = 0; ix < nxa; ix++) {
            const size_t o = (size_t)iz * nxa + ix;
            const bool zero = m0[ix] == 0.0 || m0[ix + 1] == 0.0 || m1[ix] == 0.0 || m1[ix + 1] == 0.0;
            mu_zx[o] = zero ? 0.0
                            : 4.0 / ((1.0 / m0[ix]) + (1.0 / m0[ix + 1]) + (1.0 / m1[ix]) + (1.0 / m1[ix + 1]));
            rho_zp[o] = 1.0 / (0.5 * (r0[ix] + r1[ix]));
            rho_xp[o] = 1.0 / (0.5 * (r0[ix] + r0[ix + 1]));
            // air guard: the reference clears rho_zp in both tests (fd_medium.cu:148-154)
            if ((r0[ix] < 1e-4 && r1[ix] < 1e-4) || (r0[ix] < 1e-4 && r0[ix + 1] < 1e-4)) rho_zp[o] = 0.0;
        }
    }
}

void create_wavelet(double* signal, int nt, double dt, double amp, double fc, double ts, int shape)
{
    if (shape != 1) return;
    const double fci = 1.0 / fc;
    for (int it = 0; it < nt; it++) {
        const double t = it * dt;
        const double tau = kPi * (t - 1.5 * fci - ts) / (1.5 * fci);
        signal[it] = amp * (1.0 - 2.0 * tau * tau) * std::exp(-2.0 * tau * tau);
    }
}

PaddedGrid total_grid(int nx, int nz, int fdorder, int npml, bool fsurf)
{
    PaddedGrid g;
    g.fpad = fdorder / 2;
    g.ppad = npml + g.fpad;
    g.nxt = nx + 2 * g.ppad;
    g.nzt = fsurf ? nz + g.ppad + g.fpad : nz + 2 * g.ppad;
    return g;
}

Stability check_stability(int n, const double* lam, const double* mu, const double* rho, double dh,
                          double fc, int fdorder, const double* hc)
{
    Stability s{0.0, 1e30, 0.0, 0.0};
    for (int i = 0; i < n; i++) {
        if (!(rho[i] > 0)) continue;
        const double vp = std::sqrt((lam[i] + 2.0 * mu[i]) / rho[i]);
        if (vp > s.vp_max) s.vp_max = vp;
        if (mu[i] > 1e-9) {
            const double vs = std::sqrt(mu[i] / rho[i]);
            if (vs < s.vs_min) s.vs_min = vs;
        }
    }
    double sum = 0.0;
    for (int k = 1; k <= fdorder / 2; k++) sum += std::fabs(hc[k]);
    s.dt_stab = dh / (std::sqrt(2.0) * sum * s.vp_max);
    s.dh_stab = s.vs_min / (12.0 * 2.0 * fc);
    return s;
}

}  // namespace psvhost
