// fd_kernel_dropin.cpp -- the reference's propagation entry points
// (CUDA_week2/fd_kernel.cuh:12-173) on top of the psvfwi C ABI.
//
// Each call uploads the host model and CPML profiles to a plan cached per
// grid and precision, runs the whole time loop on the GPU and writes every
// output back to the caller's host arrays.  Behaviour kept from the
// reference:
//   * only vz sources and vz receivers act (component 2; the reference's
//     injection and recording switches have no other case,
//     CUDA_week2/forward_kernel_PSV_GPU.cu:1035-1060)
//   * a source fires, and a receiver records, in shot ishot when its
//     shot_to_fire entry equals ishot; receiver rows of other shots are left
//     untouched
//   * the wavefields, We and the FWI store start from zero (the reference's
//     kernels reset them, CUDA_week2/forward_kernel_PSV_GPU.cu:760-792)
//   * the adjoint writes this shot's gradient into grad_lam / grad_mu / grad_rho,
//     overwriting what they held (the reference zeroes them on entry,
//     CUDA_week2/adjoint_kernel_PSV_GPU.cu:799-806)
//   * forward snapshots (when !fwinv): szz every snap_interval steps into
//     <dir>/vz_snap<k>.csv, k = 0, 1, ... (the reference's file name and
//     ", "-separated rows, CUDA_week2/forward_kernel_PSV_GPU.cu:1064-1080)
// Difference: the CPML memory variables also start from zero on every call
// (the reference keeps whatever the caller's arrays hold; its FWI loop
// resets them before each call, CUDA_week2/fd_wave_model.cu:790-793).
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fd_kernel.hpp"
#include "psvfwi.h"

namespace {

int g_precision = -1;  // -1: from PSVFWI_PRECISION (default fp32)
std::string g_snapdir = "./io/snap_data";

psvfwi_precision precision()
{
    if (g_precision >= 0) return g_precision ? PSVFWI_FP64 : PSVFWI_FP32;
    const char* e = std::getenv("PSVFWI_PRECISION");
    return (e && std::strcmp(e, "fp64") == 0) ? PSVFWI_FP64 : PSVFWI_FP32;
}

void check(psvfwi_status s, const char* what)
{
    if (s != PSVFWI_OK) throw std::runtime_error(std::string(what) + ": " + psvfwi_last_error());
}

struct PlanCache {
    psvfwi_plan* plan = nullptr;
    psvfwi_grid_desc grid{};
    psvfwi_precision prec = PSVFWI_FP32;
    ~PlanCache()
    {
        if (plan) psvfwi_plan_destroy(plan);
    }
};

bool same_grid(const psvfwi_grid_desc& a, const psvfwi_grid_desc& b)
{
    return std::memcmp(&a, &b, sizeof(a)) == 0;
}

psvfwi_plan* plan_for(const psvfwi_grid_desc& g)
{
    thread_local PlanCache c;
    const psvfwi_precision p = precision();
    if (c.plan && same_grid(c.grid, g) && c.prec == p) return c.plan;
    if (c.plan) {
        psvfwi_plan_destroy(c.plan);
        c.plan = nullptr;
    }
    check(psvfwi_plan_create(&g, p, &c.plan), "plan_create");
    c.grid = g;
    c.prec = p;
    return c.plan;
}

struct Common {
    int nzt, nxt, npml;
    psvfwi_plan* plan;
};

// grid, CPML profiles and model of one call
Common setup(int nt, int nzt, int nxt, int ppad, real_sim dt, real_sim dx, real_sim dz, bool fsurf,
             const real_sim* hc, int fdorder, real_sim** lam, real_sim** mu, real_sim** mu_zx,
             real_sim** rho_zp, real_sim** rho_xp, int npml, const real_sim* a, const real_sim* b,
             const real_sim* K, const real_sim* a_half, const real_sim* b_half, const real_sim* K_half)
{
    (void)nt;
    (void)ppad;
    psvfwi_grid_desc g;
    std::memset(&g, 0, sizeof(g));
    g.nzt = nzt;
    g.nxt = nxt;
    g.fdorder = fdorder;
    g.npml = npml;
    g.fsurf = fsurf ? 1 : 0;
    g.dt = dt;
    g.dx = dx;
    g.dz = dz;
    for (int i = 0; i < 7; i++) g.hc[i] = hc[i];
    psvfwi_plan* plan = plan_for(g);
    const int tp = 2 * (npml + 1);
    std::vector<double> prof((size_t)6 * tp);
    const real_sim* src[6] = {a, b, K, a_half, b_half, K_half};
    const double dflt[6] = {0.0, 1.0, 1.0, 0.0, 1.0, 1.0};
    for (int k = 0; k < 6; k++)
        for (int i = 0; i < tp; i++) prof[(size_t)k * tp + i] = src[k] ? src[k][i] : dflt[k];
    check(psvfwi_set_pml(plan, prof.data()), "set_pml");
    check(psvfwi_set_model(plan, lam[0], mu[0], mu_zx[0], rho_zp[0], rho_xp[0]), "set_model");
    return Common{nzt, nxt, npml, plan};
}

void set_window(psvfwi_plan* plan, bool fwinv, int nt, int fwi_dt, int fwi_dx, int fwi_dz, int x1, int x2, int z1,
                int z2)
{
    psvfwi_window_desc w{fwi_dt, fwi_dx, fwi_dz, x1, x2, z1, z2};
    if (!fwinv) w = psvfwi_window_desc{nt, 1, 1, 0, 0, 0, 0};  // nothing stored
    check(psvfwi_set_window(plan, &w), "set_window");
    check(psvfwi_set_store_mode(plan, PSVFWI_STORE_FULL, 1), "set_store_mode");
}

int shot_of(const ivec& to_fire, size_t i)
{
    return i < to_fire.size() ? to_fire[i] : 0;
}

// indices of the entries that act in shot ishot (vz component only)
std::vector<int> select(int n, const ivec& comp, const ivec& to_fire, int ishot)
{
    std::vector<int> out;
    for (int i = 0; i < n; i++)
        if ((size_t)i < comp.size() && comp[(size_t)i] == 2 && shot_of(to_fire, (size_t)i) == ishot) out.push_back(i);
    return out;
}

void gather_xz(const std::vector<int>& idx, const ivec& x, const ivec& z, std::vector<int>& gx, std::vector<int>& gz,
               std::vector<int>& gc)
{
    for (int i : idx) {
        gx.push_back(x[(size_t)i]);
        gz.push_back(z[(size_t)i]);
        gc.push_back(2);
    }
}

void get_fields(const Common& c, real_sim** vx, real_sim** vz, real_sim** sxx, real_sim** szx, real_sim** szz,
                real_sim** We, real_sim** mem_vx_x, real_sim** mem_vx_z, real_sim** mem_vz_x, real_sim** mem_vz_z,
                real_sim** mem_sxx_x, real_sim** mem_szx_x, real_sim** mem_szz_z, real_sim** mem_szx_z)
{
    real_sim** f[6] = {vx, vz, sxx, szx, szz, We};
    for (int k = 0; k < 6; k++)
        if (f[k]) check(psvfwi_get_field(c.plan, 0, k, f[k][0]), "get_field");
    if (c.npml <= 0) return;
    // field ids 8..15: mem_vx_x, mem_vz_x, mem_sxx_x, mem_szx_x, mem_vx_z, mem_vz_z, mem_szz_z, mem_szx_z
    real_sim** m[8] = {mem_vx_x, mem_vz_x, mem_sxx_x, mem_szx_x, mem_vx_z, mem_vz_z, mem_szz_z, mem_szx_z};
    for (int k = 0; k < 8; k++)
        if (m[k]) check(psvfwi_get_field(c.plan, 0, 8 + k, m[k][0]), "get_field");
}

void write_snapshots(const Common& c)
{
    int n = 0;
    check(psvfwi_snapshot_count(c.plan, 0, &n), "snapshot_count");
    std::vector<double> buf((size_t)c.nzt * c.nxt);
    for (int k = 0; k < n; k++) {
        check(psvfwi_get_snapshot(c.plan, 0, k, buf.data()), "get_snapshot");
        std::ofstream out(g_snapdir + "/vz_snap" + std::to_string(k) + ".csv");
        if (!out) throw std::runtime_error("cannot write snapshot file in " + g_snapdir);
        for (int j = 0; j < c.nzt; j++) {
            for (int i = 0; i < c.nxt; i++) out << buf[(size_t)j * c.nxt + i] << ", ";
            out << std::endl;
        }
    }
}

}  // namespace

void fd_kernel_set_precision(int fp64) { g_precision = fp64 ? 1 : 0; }

void fd_kernel_set_snapshot_dir(const char* dir) { g_snapdir = dir ? dir : ""; }

// CUDA_week2/fd_kernel.cuh:12-40, CUDA_week2/forward_kernel_PSV_GPU.cu:681-1084
void forward_kernel_PSV(int ishot, int nt, int nzt, int nxt, int fpad, int ppad, real_sim dt, real_sim dx,
                        real_sim dz, int snap_interval, bool fsurf, real_sim* hc, int fdorder, real_sim** vx,
                        real_sim** vz, real_sim** sxx, real_sim** szx, real_sim** szz, real_sim** We,
                        real_sim** lam, real_sim** mu, real_sim** mu_zx, real_sim** rho_zp, real_sim** rho_xp,
                        int npml, real_sim* a, real_sim* b, real_sim* K, real_sim* a_half, real_sim* b_half,
                        real_sim* K_half, real_sim** mem_vx_x, real_sim** mem_vx_z, real_sim** mem_vz_x,
                        real_sim** mem_vz_z, real_sim** mem_sxx_x, real_sim** mem_szx_x, real_sim** mem_szz_z,
                        real_sim** mem_szx_z, int nsrc, ivec src_x, ivec src_z, ivec src_comp,
                        real_sim** src_signal, ivec source_to_fire_shot, int nrec, ivec rec_x, ivec rec_z,
                        ivec rec_comp, real_sim** rec_signal, ivec receiver_to_record_shot, bool fwinv, int fwi_dt,
                        int fwi_dx, int fwi_dz, int fwi_x1, int fwi_x2, int fwi_z1, int fwi_z2,
                        real_sim*** fwi_vx, real_sim*** fwi_vz, real_sim*** fwi_sxx, real_sim*** fwi_szx,
                        real_sim*** fwi_szz)
{
    (void)fpad;
    const Common c = setup(nt, nzt, nxt, ppad, dt, dx, dz, fsurf, hc, fdorder, lam, mu, mu_zx, rho_zp, rho_xp, npml,
                           a, b, K, a_half, b_half, K_half);
    set_window(c.plan, fwinv, nt, fwi_dt, fwi_dx, fwi_dz, fwi_x1, fwi_x2, fwi_z1, fwi_z2);

    const std::vector<int> si = select(nsrc, src_comp, source_to_fire_shot, ishot);
    const std::vector<int> ri = select(nrec, rec_comp, receiver_to_record_shot, ishot);
    std::vector<int> sx, sz, sc, rx, rz, rc;
    gather_xz(si, src_x, src_z, sx, sz, sc);
    gather_xz(ri, rec_x, rec_z, rx, rz, rc);
    std::vector<double> sig(si.size() * (size_t)nt);
    for (size_t k = 0; k < si.size(); k++) std::memcpy(&sig[k * nt], src_signal[si[k]], sizeof(double) * nt);
    psvfwi_shot_desc shot{nt,        (int)si.size(), sx.data(), sz.data(), sc.data(), sig.data(),
                          (int)ri.size(), rx.data(), rz.data(), rc.data()};
    check(psvfwi_set_shots(c.plan, 1, &shot), "set_shots");
    const bool snaps = !fwinv && snap_interval > 0 && !g_snapdir.empty();
    check(psvfwi_set_snapshots(c.plan, snaps ? snap_interval : 0, 4), "set_snapshots");
    check(psvfwi_forward(c.plan, PSVFWI_ENERGY | (fwinv ? PSVFWI_GRADIENT : 0)), "forward");

    if (!ri.empty()) {
        std::vector<double> rec(ri.size() * (size_t)nt);
        check(psvfwi_get_records(c.plan, 0, rec.data()), "get_records");
        for (size_t k = 0; k < ri.size(); k++) std::memcpy(rec_signal[ri[k]], &rec[k * nt], sizeof(double) * nt);
    }
    get_fields(c, vx, vz, sxx, szx, szz, We, mem_vx_x, mem_vx_z, mem_vz_x, mem_vz_z, mem_sxx_x, mem_szx_x,
               mem_szz_z, mem_szx_z);
    if (fwinv && fwi_vx) {
        const size_t nft = 1 + (size_t)(nt - 1) / fwi_dt;
        const size_t nfz = 1 + (size_t)(fwi_z2 - fwi_z1) / fwi_dz, nfx = 1 + (size_t)(fwi_x2 - fwi_x1) / fwi_dx;
        const size_t vol = nft * nfz * nfx;
        std::vector<double> st(5 * vol);
        check(psvfwi_get_store(c.plan, 0, st.data()), "get_store");
        real_sim*** dst[5] = {fwi_vx, fwi_vz, fwi_sxx, fwi_szx, fwi_szz};
        for (int k = 0; k < 5; k++) std::memcpy(dst[k][0][0], &st[k * vol], sizeof(double) * vol);
    }
    if (snaps) write_snapshots(c);
}

// CUDA_week2/fd_kernel.cuh:43-70, CUDA_week2/adjoint_kernel_PSV_GPU.cu:702-1102
void adjoint_kernel_PSV(int ishot, int nt, int nzt, int nxt, int fpad, int ppad, real_sim dt, real_sim dx,
                        real_sim dz, int snap_interval, bool fsurf, real_sim* hc, int fdorder, real_sim** vx,
                        real_sim** vz, real_sim** sxx, real_sim** szx, real_sim** szz, real_sim** We,
                        real_sim** lam, real_sim** mu, real_sim** mu_zx, real_sim** rho_zp, real_sim** rho_xp,
                        int npml, real_sim* a, real_sim* b, real_sim* K, real_sim* a_half, real_sim* b_half,
                        real_sim* K_half, real_sim** mem_vx_x, real_sim** mem_vx_z, real_sim** mem_vz_x,
                        real_sim** mem_vz_z, real_sim** mem_sxx_x, real_sim** mem_szx_x, real_sim** mem_szz_z,
                        real_sim** mem_szx_z, int nsrc, ivec src_x, ivec src_z, ivec src_comp,
                        real_sim** src_signal, ivec source_to_fire_shot, bool fwinv, int fwi_dt, int fwi_dx,
                        int fwi_dz, int fwi_x1, int fwi_x2, int fwi_z1, int fwi_z2, real_sim*** fwi_vx,
                        real_sim*** fwi_vz, real_sim*** fwi_sxx, real_sim*** fwi_szx, real_sim*** fwi_szz,
                        real_sim** grad_lam, real_sim** grad_mu, real_sim** grad_rho)
{
    (void)fpad;
    (void)snap_interval;
    const Common c = setup(nt, nzt, nxt, ppad, dt, dx, dz, fsurf, hc, fdorder, lam, mu, mu_zx, rho_zp, rho_xp, npml,
                           a, b, K, a_half, b_half, K_half);
    set_window(c.plan, fwinv, nt, fwi_dt, fwi_dx, fwi_dz, fwi_x1, fwi_x2, fwi_z1, fwi_z2);

    // the adjoint sources are the residual traces injected at the receivers
    const std::vector<int> ai = select(nsrc, src_comp, source_to_fire_shot, ishot);
    std::vector<int> ax, az, ac;
    gather_xz(ai, src_x, src_z, ax, az, ac);
    psvfwi_shot_desc shot{nt, 0, nullptr, nullptr, nullptr, nullptr, (int)ai.size(), ax.data(), az.data(), ac.data()};
    check(psvfwi_set_shots(c.plan, 1, &shot), "set_shots");
    check(psvfwi_set_snapshots(c.plan, 0, 4), "set_snapshots");
    if (!ai.empty()) {
        std::vector<double> sig(ai.size() * (size_t)nt);
        for (size_t k = 0; k < ai.size(); k++) std::memcpy(&sig[k * nt], src_signal[ai[k]], sizeof(double) * nt);
        check(psvfwi_set_adjoint_sources(c.plan, 0, sig.data()), "set_adjoint_sources");
    }
    const size_t nfz = 1 + (size_t)(fwi_z2 - fwi_z1) / fwi_dz, nfx = 1 + (size_t)(fwi_x2 - fwi_x1) / fwi_dx;
    if (fwinv) {
        const size_t nft = 1 + (size_t)(nt - 1) / fwi_dt;
        const size_t vol = nft * nfz * nfx;
        std::vector<double> st(5 * vol);
        real_sim*** srcs[5] = {fwi_vx, fwi_vz, fwi_sxx, fwi_szx, fwi_szz};
        for (int k = 0; k < 5; k++) std::memcpy(&st[k * vol], srcs[k][0][0], sizeof(double) * vol);
        check(psvfwi_set_store(c.plan, 0, st.data()), "set_store");
    }
    check(psvfwi_adjoint(c.plan, PSVFWI_ENERGY | (fwinv ? PSVFWI_GRADIENT : 0)), "adjoint");

    get_fields(c, vx, vz, sxx, szx, szz, We, mem_vx_x, mem_vx_z, mem_vz_x, mem_vz_z, mem_sxx_x, mem_szx_x,
               mem_szz_z, mem_szx_z);
    if (fwinv) {
        const size_t n = nfz * nfx;
        std::vector<double> gl(n), gm(n), gr(n);
        check(psvfwi_get_gradient(c.plan, 0, gl.data(), gm.data(), gr.data()), "get_gradient");
        // one shot's gradient overwrites the arrays, as the reference sets them to zero on
        // entry (CUDA_week2/adjoint_kernel_PSV_GPU.cu:799-806)
        for (size_t i = 0; i < n; i++) {
            grad_lam[0][i] = gl[i];
            grad_mu[0][i] = gm[i];
            grad_rho[0][i] = gr[i];
        }
    }
}

// CUDA_week2/fd_kernel.cuh:72, CUDA_week2/fwi_step_length_estimation.cu:226-271
real_sim calc_opt_step(real_sim* L2_test, real_sim* step_length)
{
    double opt = 0;
    check(psvfwi_parabola_step(L2_test, step_length, &opt), "calc_opt_step");
    return opt;
}

// CUDA_week2/fd_kernel.cuh:75-120: same computation; the reference's device
// buffers are not touched (the plan owns the device memory), and the energy
// weight goes to a scratch host array as the reference keeps it on the device
void forward_kernel_PSV_GPU(int ishot, int nt, int nzt, int nxt, int fpad, int ppad, real_sim dt, real_sim dx,
                            real_sim dz, int snap_interval, bool fsurf, real_sim* hc, int fdorder, real_sim** vx,
                            real_sim** vz, real_sim** sxx, real_sim** szx, real_sim** szz, real_sim** lam,
                            real_sim** mu, real_sim** mu_zx, real_sim** rho_zp, real_sim** rho_xp, int npml,
                            real_sim* a, real_sim* b, real_sim* K, real_sim* a_half, real_sim* b_half,
                            real_sim* K_half, real_sim** mem_vx_x, real_sim** mem_vx_z, real_sim** mem_vz_x,
                            real_sim** mem_vz_z, real_sim** mem_sxx_x, real_sim** mem_szx_x, real_sim** mem_szz_z,
                            real_sim** mem_szx_z, int nsrc, ivec src_x, ivec src_z, ivec src_comp,
                            real_sim** src_signal, ivec source_to_fire_shot, int nrec, ivec rec_x, ivec rec_z,
                            ivec rec_comp, real_sim** rec_signal, ivec receiver_to_record_shot, bool fwinv,
                            int fwi_dt, int fwi_dx, int fwi_dz, int fwi_x1, int fwi_x2, int fwi_z1, int fwi_z2,
                            real_sim*** fwi_vx, real_sim*** fwi_vz, real_sim*** fwi_sxx, real_sim*** fwi_szx,
                            real_sim*** fwi_szz, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*,
                            real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*,
                            real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*,
                            real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*)
{
    forward_kernel_PSV(ishot, nt, nzt, nxt, fpad, ppad, dt, dx, dz, snap_interval, fsurf, hc, fdorder, vx, vz, sxx,
                       szx, szz, nullptr, lam, mu, mu_zx, rho_zp, rho_xp, npml, a, b, K, a_half, b_half, K_half,
                       mem_vx_x, mem_vx_z, mem_vz_x, mem_vz_z, mem_sxx_x, mem_szx_x, mem_szz_z, mem_szx_z, nsrc,
                       src_x, src_z, src_comp, src_signal, source_to_fire_shot, nrec, rec_x, rec_z, rec_comp,
                       rec_signal, receiver_to_record_shot, fwinv, fwi_dt, fwi_dx, fwi_dz, fwi_x1, fwi_x2, fwi_z1,
                       fwi_z2, fwi_vx, fwi_vz, fwi_sxx, fwi_szx, fwi_szz);
}

// CUDA_week2/fd_kernel.cuh:123-173
void adjoint_kernel_PSV_GPU(int ishot, int nt, int nzt, int nxt, int fpad, int ppad, real_sim dt, real_sim dx,
                            real_sim dz, int snap_interval, bool fsurf, real_sim* hc, int fdorder, real_sim** vx,
                            real_sim** vz, real_sim** sxx, real_sim** szx, real_sim** szz, real_sim** lam,
                            real_sim** mu, real_sim** mu_zx, real_sim** rho_zp, real_sim** rho_xp, int npml,
                            real_sim* a, real_sim* b, real_sim* K, real_sim* a_half, real_sim* b_half,
                            real_sim* K_half, real_sim** mem_vx_x, real_sim** mem_vx_z, real_sim** mem_vz_x,
                            real_sim** mem_vz_z, real_sim** mem_sxx_x, real_sim** mem_szx_x, real_sim** mem_szz_z,
                            real_sim** mem_szx_z, int nsrc, ivec src_x, ivec src_z, ivec src_comp,
                            real_sim** src_signal, ivec source_to_fire_shot, bool fwinv, int fwi_dt, int fwi_dx,
                            int fwi_dz, int fwi_x1, int fwi_x2, int fwi_z1, int fwi_z2, real_sim*** fwi_vx,
                            real_sim*** fwi_vz, real_sim*** fwi_sxx, real_sim*** fwi_szx, real_sim*** fwi_szz,
                            real_sim** grad_lam, real_sim** grad_mu, real_sim** grad_rho, real_sim*, real_sim*,
                            real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*,
                            real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*,
                            real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*,
                            real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*, real_sim*)
{
    adjoint_kernel_PSV(ishot, nt, nzt, nxt, fpad, ppad, dt, dx, dz, snap_interval, fsurf, hc, fdorder, vx, vz, sxx,
                       szx, szz, nullptr, lam, mu, mu_zx, rho_zp, rho_xp, npml, a, b, K, a_half, b_half, K_half,
                       mem_vx_x, mem_vx_z, mem_vz_x, mem_vz_z, mem_sxx_x, mem_szx_x, mem_szz_z, mem_szx_z, nsrc,
                       src_x, src_z, src_comp, src_signal, source_to_fire_shot, fwinv, fwi_dt, fwi_dx, fwi_dz,
                       fwi_x1, fwi_x2, fwi_z1, fwi_z2, fwi_vx, fwi_vz, fwi_sxx, fwi_szx, fwi_szz, grad_lam, grad_mu,
                       grad_rho);
}
