// capi_setup.cpp -- C ABI of the host-side setup functions (include/psvfwi.h).
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

}  // extern "C"
