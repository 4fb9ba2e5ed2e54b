// capi_setup.cpp -- C ABI of the host-side setup functions (include/psvfwi.h).
#include <algorithm>
#include <cstring>
#include <string>

#include "fd_setup.hpp"
#include "psvfwi.h"

extern "C" {

psvfwi_status psvfwi_holberg(int fdorder, int max_relative_error, double* hc)
{
    if (!hc || !psvhost::holberg_coefficients(fdorder, max_relative_error, hc)) return PSVFWI_E_ARG;
    return PSVFWI_OK;
}

psvfwi_status psvfwi_cpml_profiles(int npml, double dh, double fpml, double npower, double k_max,
                                   double rcoef, double damp_v, double dt, double* profiles)
{
    if (!profiles || npml < 1 || !(dh > 0) || !(dt > 0) || !(rcoef > 0)) return PSVFWI_E_ARG;
    psvhost::CpmlParams p;
    p.npml = npml;
    p.dh = dh;
    p.fpml = fpml;
    p.npower = npower;
    p.k_max = k_max;
    p.rcoef = rcoef;
    p.damp_v = damp_v;
    const std::vector<double> v = psvhost::cpml_profiles(p, dt);
    std::memcpy(profiles, v.data(), v.size() * sizeof(double));
    return PSVFWI_OK;
}

psvfwi_status psvfwi_average_parameters(int nzt, int nxt, const double* mu, const double* rho,
                                        double* mu_zx, double* rho_zp, double* rho_xp)
{
    if (nzt < 2 || nxt < 2 || !mu || !rho || !mu_zx || !rho_zp || !rho_xp) return PSVFWI_E_ARG;
    psvhost::average_parameters(nzt, nxt, mu, rho, mu_zx, rho_zp, rho_xp);
    return PSVFWI_OK;
}

psvfwi_status psvfwi_ricker(int nt, double dt, double amp, double fc, double ts, double* signal)
{
    if (nt < 1 || !signal || !(fc > 0)) return PSVFWI_E_ARG;
    psvhost::create_wavelet(signal, nt, dt, amp, fc, ts, 1);
    return PSVFWI_OK;
}

// A random boundary layer (for the random-boundary reconstruction; not in
// the reference): splitmix64 of (seed, iz, ix) gives u in [0, 1).
static double hash_uniform(unsigned long long seed, int iz, int ix)
{
    unsigned long long z = seed ^ (0x9E3779B97F4A7C15ULL * (unsigned long long)(iz + 1)) ^
                           (0xC2B2AE3D27D4EB4FULL * (unsigned long long)(ix + 1));
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

psvfwi_status psvfwi_random_boundary(int nzt, int nxt, int width, int fsurf, unsigned long long seed,
                                     double max_increase, double* lam, double* mu, double* rho)
{
    if (nzt < 1 || nxt < 1 || width < 1 || !(max_increase >= 0) || !lam || !mu || !rho) return PSVFWI_E_ARG;
    (void)rho;  // density is kept
    for (int iz = 0; iz < nzt; iz++) {
        for (int ix = 0; ix < nxt; ix++) {
            int d = 0;  // depth into the layer, 1 .. width
            d = std::max(d, width - ix);
            d = std::max(d, width - (nxt - 1 - ix));
            d = std::max(d, width - (nzt - 1 - iz));
            if (!fsurf) d = std::max(d, width - iz);
            if (d <= 0) continue;
            const double f = 1.0 + max_increase * ((double)d / width) * hash_uniform(seed, iz, ix);
            const size_t q = (size_t)iz * nxt + ix;
            lam[q] *= f * f;
            mu[q] *= f * f;
        }
    }
    return PSVFWI_OK;
}

}  // extern "C"
