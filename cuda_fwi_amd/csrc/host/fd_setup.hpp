// fd_setup.hpp -- host-side setup of the P-SV hot path: Holberg
// coefficients, CPML profiles, Lame parameters, staggered averages, Ricker
// wavelet, padded-grid bookkeeping.  FP64, as the reference (real_sim = double).
#pragma once

#include <vector>

namespace psvhost {

// PI as the reference defines it (CUDA_week2/globvar.cuh:18); the wavelet and
// the CPML alpha profile are defined with this value.
constexpr double kPi = 3.14159265;

// hc[0..6] for FD order 2..12 and error class 0..4 (CUDA_week2/holbergcoeff.cu:6-65).
// Returns false for an order or class outside the table.
bool holberg_coefficients(int fdorder, int max_relative_error, double hc[7]);

struct CpmlParams {
    int npml = 10;
    double dh = 0.1, fpml = 20.0, npower = 2.0, k_max = 1.0, rcoef = 0.001, damp_v = 0.0;
};

// a, b, K, a_half, b_half, K_half, each 2(npml+1) long, concatenated
// (FDCpml::cpml_pro, CUDA_week2/fd_cpml.cu:59-192).
std::vector<double> cpml_profiles(const CpmlParams& p, double dt);

// mu = vs^2 rho, lam = vp^2 rho - 2 mu (CUDA_week2/fd_medium.cu:54-59)
void velocity_to_lame(double vp, double vs, double rho, double& lam, double& mu);

// mu_zx (harmonic, 4 corners), rho_zp, rho_xp (inverse arithmetic) on
// [nzt-1][nxt-1] from mu, rho on [nzt][nxt] (CUDA_week2/fd_medium.cu:128-159).
void average_parameters(int nzt, int nxt, const double* mu, const double* rho, double* mu_zx,
                        double* rho_zp, double* rho_xp);

// Ricker wavelet (shape 1), CUDA_week2/fd_seismogram.cu:166-186.
void create_wavelet(double* signal, int nt, double dt, double amp, double fc, double ts, int shape);

// Padded sizes (FDGrid::total_grid, CUDA_week2/fd_grid.cu:52-71).
struct PaddedGrid {
    int fpad, ppad, nxt, nzt;
};
PaddedGrid total_grid(int nx, int nz, int fdorder, int npml, bool fsurf);

// Largest stable time step and smallest grid spacing for a model
// (CUDA_week2/checkfd_ssg_elastic.cu:6-113, without the printout):
// dt_stab = dh / (sqrt(2) sum|hc[1..]| vp_max), dh_stab = vs_min / (12 * 2 fc)
struct Stability {
    double vp_max, vs_min, dt_stab, dh_stab;
};
Stability check_stability(int n, const double* lam, const double* mu, const double* rho, double dh,
                          double fc, int fdorder, const double* hc);

}  // namespace psvhost
