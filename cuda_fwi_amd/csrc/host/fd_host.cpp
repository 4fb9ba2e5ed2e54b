// fd_host.cpp -- configuration and host arrays of the P-SV FWI drop-in:
// FDGrid, FDMedium, FDCpml, FDWave, FDFWI, FDSeismogram and the free
// functions the reference's host code calls (create_wavelet, holbergcoeff,
// checkfd_ssg_elastic).  INI keys, defaults and index conventions are the
// reference's; each method cites the reference function it stands in for.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>

#include "fd_host.hpp"
#include "fd_setup.hpp"

namespace {

INIReader open_ini(const char* path)
{
    INIReader r(path);
    if (r.ParseError() < 0) throw std::runtime_error(std::string("cannot load .ini file '") + path + "'");
    return r;
}

}  // namespace

// ---------------------------------------------------------------- FDGrid

// CUDA_week2/fd_grid.cu:15-50
void FDGrid::parse_configuration(const char* path)
{
    const INIReader r = open_ini(path);
    const double t = r.GetReal("grid", "time", 0);
    nx = (int)r.GetInteger("grid", "nx", 201);
    nz = (int)r.GetInteger("grid", "nz", 101);
    dx = r.GetReal("grid", "dx", 1.25);
    dz = r.GetReal("grid", "dz", 1.25);
    dt = r.GetReal("grid", "dt", 0.00025);
    nt = (int)r.GetInteger("grid", "nt", (long)std::ceil(t / dt));
    snap_interval = (int)r.GetInteger("grid", "snap_interval", nt - 1);
    fsurf = r.GetInteger("grid", "free_surface", 0) != 0;
    std::cout << "Grid: nt = " << nt << ", nx = " << nx << ", nz = " << nz << ", dt = " << dt
              << ", dx = " << dx << ", dz = " << dz << std::endl;
}

// CUDA_week2/fd_grid.cu:52-71
void FDGrid::total_grid(int fdorder, int npml)
{
    const psvhost::PaddedGrid g = psvhost::total_grid(nx, nz, fdorder, npml, fsurf);
    fpad = g.fpad;
    ppad = g.ppad;
    nxt = g.nxt;
    nzt = g.nzt;
}

// ---------------------------------------------------------------- FDMedium

// CUDA_week2/fd_medium.cu:29-52
void FDMedium::parse_configuration(const char* path)
{
    const INIReader r = open_ini(path);
    scalar_rho = r.GetReal("medium", "scalar_rho", 0.0);
    scalar_vp = r.GetReal("medium", "scalar_vp", 0.0);
    scalar_vs = r.GetReal("medium", "scalar_vs", 0.0);
    velocity_to_lame();
    std::cout << "Medium: rho = " << scalar_rho << ", vp = " << scalar_vp << ", vs = " << scalar_vs
              << ", lambda = " << scalar_lam << ", mu = " << scalar_mu << std::endl;
}

// CUDA_week2/fd_medium.cu:54-59
void FDMedium::velocity_to_lame()
{
    psvhost::velocity_to_lame(scalar_vp, scalar_vs, scalar_rho, scalar_lam, scalar_mu);
}

// CUDA_week2/fd_medium.cu:62-75: homogeneous fill
void FDMedium::stagger_over_grid(int dimz, int dimx)
{
    for (int iz = 0; iz < dimz; iz++)
        for (int ix = 0; ix < dimx; ix++) {
            rho[iz][ix] = scalar_rho;
            lam[iz][ix] = scalar_lam;
            mu[iz][ix] = scalar_mu;
        }
}

// CUDA_week2/fd_medium.cu:78-126: homogeneous fill with lambda raised by 30 %
// inside the disc (iz - dimz/2)^2 + (ix - dimx/2)^2 < dimz^2 / 49 (integer
// arithmetic).  The reference then writes rho (not lambda) to
// ./io/snap_data/lamda_snap.csv; the caller does that (write_lamda_snap).
void FDMedium::stagger_over_grid_modify(int dimz, int dimx)
{
    stagger_over_grid(dimz, dimx);
    for (int iz = 0; iz < dimz; iz++)
        for (int ix = 0; ix < dimx; ix++)
            if ((iz - dimz / 2) * (iz - dimz / 2) + (ix - dimx / 2) * (ix - dimx / 2) < dimz * dimz / 49)
                lam[iz][ix] = 1.3 * scalar_lam;
}

// CUDA_week2/fd_medium.cu:128-159 (dimz = nzt - 1, dimx = nxt - 1; rho, mu
// are [dimz + 1][dimx + 1])
void FDMedium::average_parameters(int dimz, int dimx)
{
    psvhost::average_parameters(dimz + 1, dimx + 1, mu[0], rho[0], mu_zx[0], rho_zp[0], rho_xp[0]);
}

// CUDA_week2/fd_medium.cu:161-240.  mat_new = mat + s grad over the window,
// s = step_length * max(mat) / grad_max, clipped to the material's bounds
// (rho [1.5, 3000], lambda [0, 48e9], mu [0, 27e9]).  As in the reference:
// grad_max is the signed value of the entry of largest magnitude, and the
// window loop runs over full-grid indices iz in [z1, dimz), ix in [x1, dimx)
// that index the [nfz][nfx] gradient directly.  override copies the new
// values into mat.
void FDMedium::fwi_grad_update(real_sim** mat, real_sim** mat_new, real_sim** grad_mat, real_sim step_length,
                               int dimz, int dimx, int z1, int x1, int fwi_dz, int fwi_dx, bool override,
                               int matType)
{
    double hi = 0, lo = 0;
    if (matType == 1) { hi = 3000.0; lo = 1.5; }
    else if (matType == 2) { hi = 48.0e+9; lo = 0.0; }
    else if (matType == 3) { hi = 27.0e+9; lo = 0.0; }
    double mat_max = 0.0, grad_max = 0.0;
    for (int iz = z1; iz < dimz; iz += fwi_dz)
        for (int ix = x1; ix < dimx; ix += fwi_dx) {
            if (mat[iz][ix] > mat_max) mat_max = mat[iz][ix];
            if (std::fabs(grad_mat[iz][ix]) > grad_max) grad_max = grad_mat[iz][ix];
        }
    const double s = step_length * (mat_max / grad_max);
    for (int iz = z1; iz < dimz; iz += fwi_dz)
        for (int ix = x1; ix < dimx; ix += fwi_dx) {
            double v = mat[iz][ix] + s * grad_mat[iz][ix];
            if (v < lo) v = lo;
            if (v > hi) v = hi;
            if (v < 0.0) v = mat[iz][ix];
            mat_new[iz][ix] = v;
        }
    if (override)
        for (int iz = z1; iz < dimz; iz += fwi_dz)
            for (int ix = x1; ix < dimx; ix += fwi_dx) mat[iz][ix] = mat_new[iz][ix];
}

// CUDA_week2/fd_medium.cu:242-251
void FDMedium::allocate_medium(int dimz, int dimx)
{
    allocate_array_2d(rho, dimz, dimx);
    allocate_array_2d(lam, dimz, dimx);
    allocate_array_2d(mu, dimz, dimx);
    allocate_array_2d(rho_old, dimz, dimx);
    allocate_array_2d(lam_old, dimz, dimx);
    allocate_array_2d(mu_old, dimz, dimx);
}

// CUDA_week2/fd_medium.cu:253-277: plain means over [dimz][dimx]
void FDMedium::medium_average(real_sim& C_lam, real_sim& C_mu, real_sim& C_rho, int dimz, int dimx)
{
    double l = 0, m = 0, r = 0;
    for (int iz = 0; iz < dimz; iz++)
        for (int ix = 0; ix < dimx; ix++) {
            l += lam[iz][ix];
            m += mu[iz][ix];
            r += rho[iz][ix];
        }
    C_lam = l / (dimz * dimx);
    C_mu = m / (dimz * dimx);
    C_rho = r / (dimz * dimx);
}

void FDMedium::allocate_medium_av(int dimz, int dimx)
{
    allocate_array_2d(rho_zp, dimz, dimx);
    allocate_array_2d(rho_xp, dimz, dimx);
    allocate_array_2d(mu_zx, dimz, dimx);
}

void FDMedium::deallocate_medium(int dimz)
{
    deallocate_array_2d(rho, dimz);
    deallocate_array_2d(lam, dimz);
    deallocate_array_2d(mu, dimz);
    deallocate_array_2d(rho_old, dimz);
    deallocate_array_2d(lam_old, dimz);
    deallocate_array_2d(mu_old, dimz);
}

void FDMedium::deallocate_medium_av(int dimz)
{
    deallocate_array_2d(rho_zp, dimz);
    deallocate_array_2d(rho_xp, dimz);
    deallocate_array_2d(mu_zx, dimz);
}

// ---------------------------------------------------------------- FDCpml

// CUDA_week2/fd_cpml.cu:28-56
void FDCpml::parse_configuration(const char* path)
{
    const INIReader r = open_ini(path);
    npml = (int)r.GetInteger("PML_boundary", "npml", 10);
    dh = r.GetReal("PML_boundary", "dh", 0.1);
    fpml = r.GetReal("PML_boundary", "fpml", 20.0);
    npower = r.GetReal("PML_boundary", "npower", 2.0);
    k_max_PML = r.GetReal("PML_boundary", "k_max_PML", 1.0);
    rcoef = r.GetReal("PML_boundary", "rcoef", 0.001);
    damp_v_PML = r.GetReal("PML_boundary", "damp_v_PML", 0.0);
    tpml = 2 * npml + 2;
}

// CUDA_week2/fd_cpml.cu:59-192
void FDCpml::cpml_pro(real_sim dt)
{
    psvhost::CpmlParams p;
    p.npml = npml;
    p.dh = dh;
    p.fpml = fpml;
    p.npower = npower;
    p.k_max = k_max_PML;
    p.rcoef = rcoef;
    p.damp_v = damp_v_PML;
    const std::vector<double> prof = psvhost::cpml_profiles(p, dt);
    const int tp = 2 * (npml + 1);
    real_sim* dst[6] = {a, b, K, a_half, b_half, K_half};
    for (int k = 0; k < 6; k++)
        for (int i = 0; i < tp; i++) dst[k][i] = prof[(size_t)k * tp + i];
}

// CUDA_week2/fd_cpml.cu:194-205
void FDCpml::allocate_pml()
{
    const int tp = 2 * (npml + 1);
    allocate_array_1d(a, tp);
    allocate_array_1d(b, tp);
    allocate_array_1d(K, tp);
    allocate_array_1d(a_half, tp);
    allocate_array_1d(b_half, tp);
    allocate_array_1d(K_half, tp);
}

// CUDA_week2/fd_cpml.cu:207-225: x-type [dimz][tp], z-type [tp][dimx]
void FDCpml::allocate_memory_arrays(int dimz, int dimx)
{
    const int tp = 2 * (npml + 1);
    allocate_array_2d(mem_vx_x, dimz, tp);
    allocate_array_2d(mem_vz_x, dimz, tp);
    allocate_array_2d(mem_sxx_x, dimz, tp);
    allocate_array_2d(mem_szx_x, dimz, tp);
    allocate_array_2d(mem_vx_z, tp, dimx);
    allocate_array_2d(mem_vz_z, tp, dimx);
    allocate_array_2d(mem_szz_z, tp, dimx);
    allocate_array_2d(mem_szx_z, tp, dimx);
}

void FDCpml::reset_memory_arrays(int dimz, int dimx)
{
    const int tp = 2 * (npml + 1);
    reset_array_2d(mem_vx_x, dimz, tp);
    reset_array_2d(mem_vz_x, dimz, tp);
    reset_array_2d(mem_sxx_x, dimz, tp);
    reset_array_2d(mem_szx_x, dimz, tp);
    reset_array_2d(mem_vx_z, tp, dimx);
    reset_array_2d(mem_vz_z, tp, dimx);
    reset_array_2d(mem_szz_z, tp, dimx);
    reset_array_2d(mem_szx_z, tp, dimx);
}

void FDCpml::deallocate_pml()
{
    deallocate_array_1d(a);
    deallocate_array_1d(b);
    deallocate_array_1d(K);
    deallocate_array_1d(a_half);
    deallocate_array_1d(b_half);
    deallocate_array_1d(K_half);
}

void FDCpml::deallocate_memory_arrays(int dimz)
{
    deallocate_array_2d(mem_vx_x, dimz);
    deallocate_array_2d(mem_vz_x, dimz);
    deallocate_array_2d(mem_sxx_x, dimz);
    deallocate_array_2d(mem_szx_x, dimz);
    deallocate_array_2d(mem_vx_z);
    deallocate_array_2d(mem_vz_z);
    deallocate_array_2d(mem_szz_z);
    deallocate_array_2d(mem_szx_z);
}

// ---------------------------------------------------------------- FDWave

// CUDA_week2/fd_wave.cu:14-46
void FDWave::allocate_wave(int dimz, int dimx)
{
    allocate_array_2d(vx, dimz, dimx);
    allocate_array_2d(vz, dimz, dimx);
    allocate_array_2d(sxx, dimz, dimx);
    allocate_array_2d(szx, dimz, dimx);
    allocate_array_2d(szz, dimz, dimx);
    allocate_array_2d(We, dimz, dimx);
}

void FDWave::reset_kernel(int dimz, int dimx)
{
    reset_array_2d(vx, dimz, dimx);
    reset_array_2d(vz, dimz, dimx);
    reset_array_2d(sxx, dimz, dimx);
    reset_array_2d(szx, dimz, dimx);
    reset_array_2d(szz, dimz, dimx);
    reset_array_2d(We, dimz, dimx);
}

void FDWave::deallocate_wave(int dimz)
{
    deallocate_array_2d(vx, dimz);
    deallocate_array_2d(vz, dimz);
    deallocate_array_2d(sxx, dimz);
    deallocate_array_2d(szx, dimz);
    deallocate_array_2d(szz, dimz);
    deallocate_array_2d(We, dimz);
}

// ---------------------------------------------------------------- FDFWI

// CUDA_week2/fd_fwi.cu:16-58.  The vertical decimation is read from the key
// grid_interval_y, as the reference reads it (a grid_interval_z key, as in
// the reference's test_configuration.ini, is ignored there and here).  Zero
// right/bottom bounds mean nxt - 1 / nzt - 1 before padding, as in the
// reference; the padded window is then clamped to the grid, which the
// reference does not do (its default window reaches past the grid).
void FDFWI::parse_configuration(const char* path, const FDGrid& grid)
{
    const INIReader r = open_ini(path);
    const char* s = "full_waveform_inversion";
    dt = (int)r.GetInteger(s, "time_interval", 1);
    dx = (int)r.GetInteger(s, "grid_interval_x", 1);
    dz = (int)r.GetInteger(s, "grid_interval_y", 1);
    x1 = (int)r.GetInteger(s, "left_grid", 0);
    x2 = (int)r.GetInteger(s, "right_grid", 0);
    z1 = (int)r.GetInteger(s, "top_grid", 0);
    z2 = (int)r.GetInteger(s, "bottom_grid", 0);
    if (z2 == 0) z2 = grid.nzt - 1;
    if (x2 == 0) x2 = grid.nxt - 1;
    x1 += grid.ppad;
    x2 += grid.ppad;
    const int zo = grid.fsurf ? grid.fpad : grid.ppad;
    z1 += zo;
    z2 += zo;
    if (dt < 1 || dx < 1 || dz < 1) throw std::runtime_error("full_waveform_inversion: intervals must be >= 1");
    // the window loops run over [x1, x2) x [z1, z2): x2 = nxt covers the grid
    if (x2 > grid.nxt) x2 = grid.nxt;
    if (z2 > grid.nzt) z2 = grid.nzt;
    if (x1 > x2 || z1 > z2) throw std::runtime_error("full_waveform_inversion: empty window");
    std::cout << "FWI: dt = " << dt << ", dx = " << dx << ", dz = " << dz << ", x1 = " << x1 << ", x2 = " << x2
              << ", z1 = " << z1 << ", z2 = " << z2 << std::endl;
}

// CUDA_week2/fd_fwi.cu:63-79
void FDFWI::allocate_fwi(int dimt, int dimz, int dimx)
{
    allocate_array_3d(vx, dimt, dimz, dimx);
    allocate_array_3d(vz, dimt, dimz, dimx);
    allocate_array_3d(sxx, dimt, dimz, dimx);
    allocate_array_3d(szx, dimt, dimz, dimx);
    allocate_array_3d(szz, dimt, dimz, dimx);
    allocate_array_2d(grad_lam, dimz, dimx);
    allocate_array_2d(grad_mu, dimz, dimx);
    allocate_array_2d(grad_rho, dimz, dimx);
}

void FDFWI::reset_fwi_kernel(int dimt, int dimz, int dimx, bool grad_reset)
{
    reset_array_3d(vx, dimt, dimz, dimx);
    reset_array_3d(vz, dimt, dimz, dimx);
    reset_array_3d(sxx, dimt, dimz, dimx);
    reset_array_3d(szx, dimt, dimz, dimx);
    reset_array_3d(szz, dimt, dimz, dimx);
    if (grad_reset) {
        reset_array_2d(grad_rho, dimz, dimx);
        reset_array_2d(grad_lam, dimz, dimx);
        reset_array_2d(grad_mu, dimz, dimx);
    }
}

void FDFWI::deallocate_fwi(int dimt, int dimz)
{
    deallocate_array_3d(vx, dimt, dimz);
    deallocate_array_3d(vz, dimt, dimz);
    deallocate_array_3d(sxx, dimt, dimz);
    deallocate_array_3d(szx, dimt, dimz);
    deallocate_array_3d(szz, dimt, dimz);
    deallocate_array_2d(grad_lam, dimz);
    deallocate_array_2d(grad_mu, dimz);
    deallocate_array_2d(grad_rho, dimz);
}

// ---------------------------------------------------------------- FDSeismogram

// CUDA_week2/fd_seismogram.cu:18-151.  sr = 's' parses [source] and builds
// the wavelets, 'r' parses [receiver] and allocates the observed traces when
// measurement_records is set.  A shot_to_fire list shorter than nseis is
// padded with shot 0 (the reference reads past the end of the list).
void FDSeismogram::parse_configuration(const char* path, const char sr, const FDGrid& grid)
{
    const INIReader r = open_ini(path);
    const std::string sec = sr == 's' ? "source" : "receiver";
    nseis = (unsigned int)r.GetInteger(sec, "nseis", 1);
    read_from_data_files = (int)r.GetInteger(sec, "read_from_data_files", 0);
    seis_file = r.Get(sec, "seis_file", "inbuilt");
    signal_file = r.Get(sec, "signal_file", "inbuilt");
    if (read_from_data_files == 1)
        throw std::runtime_error("[" + sec + "] read_from_data_files = 1 is not supported (as in the reference)");
    parse_string_to_vector(r.Get(sec, "x_seis", "{0};"), &x_seis);
    parse_string_to_vector(r.Get(sec, "z_seis", "{0};"), &z_seis);
    parse_string_to_vector(r.Get(sec, "velocity_or_stress", "{0};"), &v_or_s);
    parse_string_to_vector(r.Get(sec, "component_type", "{0};"), &seis_comp);
    parse_string_to_vector(r.Get(sec, "shot_to_fire", "{0};"), &shot_to_fire);
    if (x_seis.size() < nseis || z_seis.size() < nseis || seis_comp.size() < nseis)
        throw std::runtime_error("[" + sec + "] x_seis / z_seis / component_type hold fewer than nseis entries");
    while (shot_to_fire.size() < nseis) shot_to_fire.push_back(0);
    while (v_or_s.size() < nseis) v_or_s.push_back(0);
    allocate_array_2d(signal, (int)nseis, grid.nt);
    if (sr == 's') {
        parse_string_to_vector(r.Get(sec, "verticle_angle_deg", "{0.0};"), &angle_v);
        parse_string_to_vector(r.Get(sec, "shape", "{0};"), &shape);
        parse_string_to_vector(r.Get(sec, "amplitude", "{0.0};"), &amp);
        parse_string_to_vector(r.Get(sec, "peak_frequency", "{0.0};"), &fc);
        parse_string_to_vector(r.Get(sec, "time_shift", "{0.0};"), &ts);
        if (shape.size() < nseis || amp.size() < nseis || fc.size() < nseis || ts.size() < nseis)
            throw std::runtime_error("[source] shape / amplitude / peak_frequency / time_shift hold fewer than nseis entries");
        for (unsigned int i = 0; i < nseis; i++)
            create_wavelet(signal[i], grid.nt, grid.dt, amp[i], fc[i], ts[i], shape[i]);
    } else {
        measurement_records = (int)r.GetInteger(sec, "measurement_records", 0);
        if (measurement_records) allocate_array_2d(signal_meas, (int)nseis, grid.nt);
    }
    std::cout << (sr == 's' ? "Sources: " : "Receivers: ") << nseis << std::endl;
}

// CUDA_week2/fd_seismogram.cu:153-164
void FDSeismogram::grid_adjustment(const FDGrid& grid)
{
    for (unsigned int i = 0; i < nseis; i++) {
        x_seis[i] += grid.ppad;
        z_seis[i] += grid.fsurf ? grid.fpad : grid.ppad;
    }
}

// ---------------------------------------------------------------- free functions

// CUDA_week2/fd_seismogram.cu:166-186 (shape 1: Ricker; other shapes leave
// the signal untouched)
void create_wavelet(real_sim* signal, int nt, real_sim dt, real_sim amp, real_sim fc, real_sim ts, int shape)
{
    psvhost::create_wavelet(signal, nt, dt, amp, fc, ts, shape);
}

// CUDA_week2/holbergcoeff.cu:6-65
void holbergcoeff(const int FDORDER, const int MAXRELERROR, real_sim* hc)
{
    double h[7];
    if (!psvhost::holberg_coefficients(FDORDER, MAXRELERROR, h))
        throw std::runtime_error("holbergcoeff: no coefficients for this order / error class");
    for (int i = 0; i < 7; i++) hc[i] = h[i];
}

// CUDA_week2/checkfd_ssg_elastic.cu:6-113: extreme phase velocities over the
// cells [1, NY] x [1, NX] of the padded arrays, then the dispersion criterion
// dh_stab = v_min / (12 * 2 TS) and the stability limit
// dt_stab = DH / (sqrt(2) sum|hc[1..6]| v_max).  Reports only; an absorbing
// frame wider than the domain is an error.
void checkfd_ssg_elastic(int NX, int NY, real_sim DH, real_sim DT, real_sim TS, int FW, real_sim** prho,
                         real_sim** ppi, real_sim** pu, real_sim* hc)
{
    double cmax_p = 0, cmin_p = 1e9, cmax_s = 0, cmin_s = 1e9, c = 0;
    bool zero_cs = false;
    for (int i = 1; i <= NX; i++)
        for (int j = 1; j <= NY; j++) {
            if (std::fabs(pu[j][i]) < 1.0e-9) zero_cs = true;
            else c = std::sqrt(pu[j][i] / prho[j][i]);
            if (cmax_s < c) cmax_s = c;
            if (cmin_s > c) cmin_s = c;
        }
    for (int i = 1; i <= NX; i++)
        for (int j = 1; j <= NY; j++) {
            c = std::sqrt((ppi[j][i] + 2.0 * pu[j][i]) / prho[j][i]);
            if (cmax_p < c) cmax_p = c;
            if (cmin_p > c) cmin_p = c;
        }
    const double cmax = cmax_s > cmax_p ? cmax_s : cmax_p;
    const double cmin = cmin_s < cmin_p ? cmin_s : cmin_p;
    const double dhstab = cmin / (12.0 * 2.0 * TS);
    double gamma = 0;
    for (int k = 1; k <= 6; k++) gamma += std::fabs(hc[k]);
    const double dtstab = DH / (std::sqrt(2.0) * gamma * cmax);
    if (zero_cs) std::cout << "WARNING: zero shear waves neglected" << std::endl;
    if (DH > dhstab) std::cout << "WARNING: grid dispersion will influence wave propagation, choose smaller grid spacing (DH)." << std::endl;
    std::cout << "DHSTAB = " << dhstab << ", DH = " << DH << std::endl;
    if (DT > dtstab) std::cout << "The simulation will get unstable, choose smaller DT. DTSTAB = " << dtstab << " DT = " << DT << std::endl;
    else std::cout << "The simulation will be stable. DTSTAB = " << dtstab << " DT = " << DT << std::endl;
    if (FW > NX - FW || FW > NY - FW)
        throw std::runtime_error("checkfd_ssg_elastic: absorbing frame wider than the computational domain");
    if (FW < 10) std::cout << "WARNING: be aware of artificial reflections from grid boundaries" << std::endl;
}

// ---------------------------------------------------------------- device options

void MI355XOptions::parse(const char* path)
{
    const INIReader r = open_ini(path);
    const char* s = "mi355x";
    precision = r.Get(s, "precision", precision);
    gradient_store = r.Get(s, "gradient_store", gradient_store);
    checkpoint_interval = (int)r.GetInteger(s, "checkpoint_interval", checkpoint_interval);
    tile = (int)r.GetInteger(s, "tile", tile);
    observed = r.Get(s, "observed", observed);
    model_file = r.Get(s, "model_file", model_file);
    true_model_file = r.Get(s, "true_model_file", true_model_file);
    gradient_scaling = r.Get(s, "gradient_scaling", gradient_scaling);
    gradient = r.Get(s, "gradient", gradient);
    gradient_segment = (int)r.GetInteger(s, "gradient_segment", gradient_segment);
    illumination = r.Get(s, "illumination", illumination);
    direction_smooth = r.GetReal(s, "direction_smooth", direction_smooth);
    direction_depth_power = r.GetReal(s, "direction_depth_power", direction_depth_power);
    direction_mute = (int)r.GetInteger(s, "direction_mute", direction_mute);
    direction_taper = (int)r.GetInteger(s, "direction_taper", direction_taper);
    output_dir = r.Get(s, "output_dir", output_dir);
    write_csv = r.GetBoolean(s, "write_csv", write_csv);
    if (precision != "fp32" && precision != "fp64" && precision != "fp32x2" && precision != "fp32fma")
        throw std::runtime_error("[mi355x] precision: fp32, fp64, fp32x2 or fp32fma");
    if (gradient_store != "full" && gradient_store != "checkpoint")
        throw std::runtime_error("[mi355x] gradient_store: full or checkpoint");
    if (gradient_scaling != "reference" && gradient_scaling != "clean")
        throw std::runtime_error("[mi355x] gradient_scaling: reference or clean");
    if (!observed.empty() && observed != "modified" && observed != "file" && observed != "model")
        throw std::runtime_error("[mi355x] observed: empty, modified, file or model");
    if (observed == "model" && true_model_file.empty())
        throw std::runtime_error("[mi355x] observed = model needs true_model_file");
    if (checkpoint_interval < 1) throw std::runtime_error("[mi355x] checkpoint_interval must be >= 1");
    if (gradient != "reference" && gradient != "discrete")
        throw std::runtime_error("[mi355x] gradient: reference or discrete");
    if (gradient_segment < 1) throw std::runtime_error("[mi355x] gradient_segment must be >= 1");
    if (illumination != "source" && illumination != "source_receiver")
        throw std::runtime_error("[mi355x] illumination: source or source_receiver");
    if (!(direction_smooth >= 0.0) || !(direction_depth_power >= 0.0) || direction_mute < 0 || direction_taper < 0)
        throw std::runtime_error("[mi355x] direction_smooth, direction_depth_power, direction_mute and "
                                 "direction_taper must be >= 0");
}
