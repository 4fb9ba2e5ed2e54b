// fd_wave_model.cpp -- FDWaveModel, the reference's driver
// (CUDA_week2/fd_wave_model.cu:33-1052, fwi_step_length_estimation.cu:11-224),
// on the MI355X path: all shots of a model run as one batch through a
// psvfwi plan (forward, FP64 misfit on the GPU, adjoint, per-shot
// gradients), so an FWI iteration is one batched forward, one batched
// adjoint and one batched forward per step-length trial.
//
// Sequence per FWI iteration (as the reference's CPU branch,
// CUDA_week2/fd_wave_model.cu:753-966): average the model, the medium means,
// forward + residual (L2 = sum 0.5 dt r^2), adjoint, energy scaling of the
// gradients, step-length search with trial forwards and the parabola fit,
// model update (rho with half the step), CSV dumps of rho / lambda / mu.
// With several shots the reference updates the model after every shot and
// compares the first trial's misfit of the last shot against the current
// shot's misfit; here one update per iteration uses the summed misfit and
// the sum of the shots' scaled gradients (identical for one shot).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "fd_host.hpp"
#include "fd_kernel.hpp"
#include "psvfwi.h"

namespace {

void check(psvfwi_status s, const char* what)
{
    if (s != PSVFWI_OK) throw std::runtime_error(std::string(what) + ": " + psvfwi_last_error());
}

void write_csv(const std::string& path, real_sim** a, int rows, int cols)
{
    std::ofstream out(path);
    if (!out) throw std::runtime_error("cannot write " + path);
    for (int j = 0; j < rows; j++) {
        for (int i = 0; i < cols; i++) out << a[j][i] << ", ";
        out << std::endl;
    }
}

// rows of ", "-separated numbers (the format this library and the reference write)
// vp, vs, rho [nz][nx] (raw FP64, row-major) into lam, mu, rho [nzt][nxt]: the model
// cells at their padded indices (x + ppad, z + fpad with a free surface, z + ppad without),
// the padding filled with the nearest model cell; mu = vs^2 rho, lam = vp^2 rho - 2 mu
// (FDMedium::velocity_to_lame, CUDA_week2/fd_medium.cu:54-59, per cell).
void load_model(const FDWaveModel& m, const std::string& path, real_sim** lam, real_sim** mu, real_sim** rho)
{
    const int nz = m.grid.nz, nx = m.grid.nx, nzt = m.grid.nzt, nxt = m.grid.nxt;
    const size_t n = (size_t)nz * nx;
    std::vector<double> v(3 * n);
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open model file " + path);
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
    if ((size_t)f.gcount() != v.size() * sizeof(double))
        throw std::runtime_error("model file " + path + ": expected 3 x nz x nx FP64 values (vp, vs, rho)");
    const int zoff = m.grid.fsurf ? m.grid.fpad : m.grid.ppad, xoff = m.grid.ppad;
    for (int iz = 0; iz < nzt; iz++)
        for (int ix = 0; ix < nxt; ix++) {
            const int z = std::min(std::max(iz - zoff, 0), nz - 1), x = std::min(std::max(ix - xoff, 0), nx - 1);
            const size_t k = (size_t)z * nx + x;
            const double vp = v[k], vs = v[n + k], r = v[2 * n + k];
            if (!(vp > 0 && vs >= 0 && r > 0)) throw std::runtime_error("model file " + path + ": vp, rho must be > 0");
            rho[iz][ix] = r;
            mu[iz][ix] = vs * vs * r;
            lam[iz][ix] = vp * vp * r - 2.0 * mu[iz][ix];
        }
}

void read_csv(const std::string& path, real_sim** a, int rows, int cols)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot read " + path);
    std::string line;
    for (int j = 0; j < rows; j++) {
        if (!std::getline(in, line)) throw std::runtime_error(path + ": fewer rows than receivers");
        std::stringstream ss(line);
        std::string tok;
        int i = 0;
        while (i < cols && std::getline(ss, tok, ',')) {
            if (tok.find_first_not_of(" \t\r") == std::string::npos) continue;
            a[j][i++] = std::strtod(tok.c_str(), nullptr);
        }
        if (i < cols) throw std::runtime_error(path + ": row shorter than nt");
    }
}

// Energy scaling of one shot's gradients, CUDA_week2/fd_wave_model.cu:1004-1052:
// W = sqrt(Ws Wr) + 0.005 max(sqrt(Ws Wr)); grad /= W C^2.
// reference: loops over full-grid indices [z1, nfz) x [x1, nfx) that index the
// [nfz][nfx] gradients and the [nzt][nxt] energies alike (as the reference);
// clean: gradient cell (zf, xf) uses the energies at (z1 + zf dz, x1 + xf dx).
void scale_energy(double* gl, double* gm, double* gr, const double* Ws, const double* Wr, int nzt, int nxt,
                  double C_rho, double C_lam, double C_mu, int nfz, int nfx, int z1, int x1, int fdz, int fdx,
                  bool clean)
{
    const double eps = 0.005;
    std::vector<double> W((size_t)nfz * nfx, 0.0);
    std::vector<char> on((size_t)nfz * nfx, 0);
    double wmax = 0.0;
    auto visit = [&](auto f) {
        if (!clean) {
            for (int iz = z1; iz < nfz; iz += fdz)
                for (int ix = x1; ix < nfx; ix += fdx) f((size_t)iz * nfx + ix, (size_t)iz * nxt + ix);
        } else {
            for (int zf = 0; zf < nfz; zf++)
                for (int xf = 0; xf < nfx; xf++) {
                    const int iz = z1 + zf * fdz, ix = x1 + xf * fdx;
                    if (iz < nzt && ix < nxt) f((size_t)zf * nfx + xf, (size_t)iz * nxt + ix);
                }
        }
    };
    visit([&](size_t g, size_t p) {
        W[g] = std::sqrt(Ws[p] * Wr[p]);
        on[g] = 1;
        if (W[g] > wmax) wmax = W[g];
    });
    visit([&](size_t g, size_t) { W[g] += eps * wmax; });
    visit([&](size_t g, size_t) {
        gl[g] = gl[g] / (W[g] * C_lam * C_lam);
        gm[g] = gm[g] / (W[g] * C_mu * C_mu);
        gr[g] = gr[g] / (W[g] * C_rho * C_rho);
    });
}

// Window-correct model update: cell (z1 + zf dz, x1 + xf dx) moves by
// s grad[zf][xf], s = step max(mat) / max|grad|, clipped as the reference.
void clean_update(real_sim** mat, real_sim** mat_new, real_sim** grad, double step, int nfz, int nfx, int nzt,
                  int nxt, int z1, int x1, int dz, int dx, bool override, int mat_type)
{
    double hi = 0, lo = 0;
    if (mat_type == 1) { hi = 3000.0; lo = 1.5; }
    else if (mat_type == 2) { hi = 48.0e+9; lo = 0.0; }
    else { hi = 27.0e+9; lo = 0.0; }
    double mat_max = 0, grad_max = 0;
    for (int zf = 0; zf < nfz; zf++)
        for (int xf = 0; xf < nfx; xf++) {
            const int iz = z1 + zf * dz, ix = x1 + xf * dx;
            if (iz >= nzt || ix >= nxt) continue;
            mat_max = std::max(mat_max, mat[iz][ix]);
            grad_max = std::max(grad_max, std::fabs(grad[zf][xf]));
        }
    if (grad_max == 0.0) return;
    const double s = step * mat_max / grad_max;
    for (int zf = 0; zf < nfz; zf++)
        for (int xf = 0; xf < nfx; xf++) {
            const int iz = z1 + zf * dz, ix = x1 + xf * dx;
            if (iz >= nzt || ix >= nxt) continue;
            double v = mat[iz][ix] + s * grad[zf][xf];
            if (v < lo) v = lo;
            if (v > hi) v = hi;
            mat_new[iz][ix] = v;
            if (override) mat[iz][ix] = v;
        }
}

}  // namespace

struct FDWaveModel::Impl {
    psvfwi_plan* plan = nullptr;
    bool tuned = false;                               // the autotuner has run on this plan
    std::vector<std::vector<int>> src_idx, rec_idx;  // per shot: source / receiver entries
    ~Impl()
    {
        if (plan) psvfwi_plan_destroy(plan);
    }
};

// CUDA_week2/fd_wave_model.cu:33-75
FDWaveModel::FDWaveModel(const char* path) : impl_(new Impl)
{
    parse_configuration(path);
    grid.parse_configuration(path);
    medium.parse_configuration(path);
    pml.parse_configuration(path);
    dev.parse(path);
    grid.total_grid((int)fdorder, pml.npml);
    source.parse_configuration(path, 's', grid);
    receiver.parse_configuration(path, 'r', grid);
    source.grid_adjustment(grid);
    receiver.grid_adjustment(grid);
    if (fwinv) fwi.parse_configuration(path, grid);
    if (pml.npml > 0) {
        pml.allocate_pml();
        pml.cpml_pro(grid.dt);
    }
    holbergcoeff((int)fdorder, max_rel_err, hc);
    fd_kernel_set_precision(dev.precision == "fp64");
}

// CUDA_week2/fd_wave_model.cu:79-102
FDWaveModel::~FDWaveModel()
{
    if (pml.npml > 0) {
        pml.deallocate_memory_arrays(grid.nzt);
        pml.deallocate_pml();
    }
    wave.deallocate_wave(grid.nzt);
    medium.deallocate_medium_av(grid.nzt - 1);
    medium.deallocate_medium(grid.nzt);
    fwi.deallocate_fwi(0, 0);
    deallocate_array_2d(source.signal);
    deallocate_array_2d(receiver.signal);
    deallocate_array_2d(receiver.signal_meas);
    delete impl_;
}

// CUDA_week2/fd_wave_model.cu:106-134
void FDWaveModel::parse_configuration(const char* path)
{
    INIReader r(path);
    if (r.ParseError() < 0) throw std::runtime_error(std::string("cannot load .ini file '") + path + "'");
    fdorder = (unsigned int)r.GetInteger("FDM", "fdorder", 2);
    max_rel_err = (int)r.GetInteger("FDM", "max_relative_error", 1);
    fwinv = r.GetInteger("FDM", "full_waveform_inversion", 0) != 0;
    nshots = (unsigned int)r.GetInteger("FDM", "nshots", 1);
    if (fwinv) {
        maxIter = (int)r.GetInteger("FDM", "max_iteration", 500);
        maxError = r.GetReal("FDM", "max_error", 1.0e-3);
    }
}

namespace {

std::vector<int> acting(const FDSeismogram& s, unsigned int ishot)
{
    std::vector<int> out;
    for (unsigned int i = 0; i < s.nseis; i++)
        if (s.seis_comp[i] == 2 && s.shot_to_fire[i] == (int)ishot) out.push_back((int)i);
    return out;
}

}  // namespace

// One plan holding every shot of the model: sources and receivers of shot i
// are the entries whose shot_to_fire is i (vz component, as the reference).
static void build_plan(FDWaveModel& m, FDWaveModel::Impl* impl, bool with_window)
{
    psvfwi_grid_desc g;
    std::memset(&g, 0, sizeof(g));
    g.nzt = m.grid.nzt;
    g.nxt = m.grid.nxt;
    g.fdorder = (int)m.fdorder;
    g.npml = m.pml.npml;
    g.fsurf = m.grid.fsurf;
    g.dt = m.grid.dt;
    g.dx = m.grid.dx;
    g.dz = m.grid.dz;
    for (int i = 0; i < 7; i++) g.hc[i] = m.hc[i];
    if (impl->plan) psvfwi_plan_destroy(impl->plan);
    impl->plan = nullptr;
    impl->tuned = false;
    const psvfwi_precision prec = m.dev.precision == "fp64"     ? PSVFWI_FP64
                                  : m.dev.precision == "fp32x2" ? PSVFWI_FP32X2
                                  : m.dev.precision == "fp32fma" ? PSVFWI_FP32_FMA
                                                                 : PSVFWI_FP32;
    check(psvfwi_plan_create(&g, prec, &impl->plan), "plan_create");
    psvfwi_plan* p = impl->plan;
    if (m.pml.npml > 0) {
        const int tp = 2 * (m.pml.npml + 1);
        std::vector<double> prof((size_t)6 * tp);
        const real_sim* src[6] = {m.pml.a, m.pml.b, m.pml.K, m.pml.a_half, m.pml.b_half, m.pml.K_half};
        for (int k = 0; k < 6; k++) std::memcpy(&prof[(size_t)k * tp], src[k], sizeof(double) * tp);
        check(psvfwi_set_pml(p, prof.data()), "set_pml");
    }
    if (with_window) {
        psvfwi_window_desc w{m.fwi.dt, m.fwi.dx, m.fwi.dz, m.fwi.x1, m.fwi.x2, m.fwi.z1, m.fwi.z2};
        check(psvfwi_set_window(p, &w), "set_window");
        if (m.dev.gradient_store == "checkpoint")
            check(psvfwi_set_store_mode(p, PSVFWI_STORE_CHECKPOINT, m.dev.checkpoint_interval), "set_store_mode");
        if (m.dev.gradient == "discrete")
            check(psvfwi_set_gradient_mode(p, PSVFWI_GRADIENT_DISCRETE, m.dev.gradient_segment), "set_gradient_mode");
    }
    const int nt = m.grid.nt;
    impl->src_idx.clear();
    impl->rec_idx.clear();
    std::vector<std::vector<int>> buf(6 * (size_t)m.nshots);
    std::vector<std::vector<double>> sig(m.nshots);
    std::vector<psvfwi_shot_desc> shots(m.nshots);
    for (unsigned int s = 0; s < m.nshots; s++) {
        const std::vector<int> si = acting(m.source, s), ri = acting(m.receiver, s);
        impl->src_idx.push_back(si);
        impl->rec_idx.push_back(ri);
        std::vector<int>* b = &buf[6 * (size_t)s];
        for (int i : si) {
            b[0].push_back(m.source.x_seis[i]);
            b[1].push_back(m.source.z_seis[i]);
            b[2].push_back(2);
            sig[s].insert(sig[s].end(), m.source.signal[i], m.source.signal[i] + nt);
        }
        for (int i : ri) {
            b[3].push_back(m.receiver.x_seis[i]);
            b[4].push_back(m.receiver.z_seis[i]);
            b[5].push_back(2);
        }
        shots[s] = psvfwi_shot_desc{nt,         (int)si.size(), b[0].data(), b[1].data(), b[2].data(), sig[s].data(),
                                    (int)ri.size(), b[3].data(), b[4].data(), b[5].data()};
    }
    check(psvfwi_set_shots(p, (int)m.nshots, shots.data()), "set_shots");
    if (m.dev.tile != 0) check(psvfwi_set_tile(p, m.dev.tile), "set_tile");
}

static void upload_model(FDWaveModel& m, psvfwi_plan* p)
{
    m.medium.average_parameters(m.grid.nzt - 1, m.grid.nxt - 1);
    check(psvfwi_set_model(p, m.medium.lam[0], m.medium.mu[0], m.medium.mu_zx[0], m.medium.rho_zp[0],
                           m.medium.rho_xp[0]),
          "set_model");
}

// [mi355x] tile = 0: once per plan, after the first model upload, the
// autotuner picks the step kernel, tile height, shot order and launch group on
// the plan's own shots (results do not depend on the choice)
static void tune_once(FDWaveModel& m, FDWaveModel::Impl* impl)
{
    if (m.dev.tile != 0 || impl->tuned) return;
    check(psvfwi_autotune(impl->plan, std::min(100, (int)m.grid.nt), nullptr), "autotune");
    impl->tuned = true;
}

static void download_records(FDWaveModel& m, FDWaveModel::Impl* impl)
{
    const int nt = m.grid.nt;
    for (unsigned int s = 0; s < m.nshots; s++) {
        const std::vector<int>& ri = impl->rec_idx[s];
        if (ri.empty()) continue;
        std::vector<double> rec(ri.size() * (size_t)nt);
        check(psvfwi_get_records(impl->plan, (int)s, rec.data()), "get_records");
        for (size_t k = 0; k < ri.size(); k++)
            std::memcpy(m.receiver.signal[ri[k]], &rec[k * nt], sizeof(double) * nt);
    }
}

static void download_fields(FDWaveModel& m, psvfwi_plan* p, int shot)
{
    real_sim** f[6] = {m.wave.vx, m.wave.vz, m.wave.sxx, m.wave.szx, m.wave.szz, m.wave.We};
    for (int k = 0; k < 6; k++) check(psvfwi_get_field(p, shot, k, f[k][0]), "get_field");
    if (m.pml.npml <= 0) return;
    real_sim** mm[8] = {m.pml.mem_vx_x, m.pml.mem_vz_x, m.pml.mem_sxx_x, m.pml.mem_szx_x,
                        m.pml.mem_vx_z, m.pml.mem_vz_z, m.pml.mem_szz_z, m.pml.mem_szx_z};
    for (int k = 0; k < 8; k++) check(psvfwi_get_field(p, shot, 8 + k, mm[k][0]), "get_field");
}

static std::string io_dir(const FDWaveModel& m, const char* sub)
{
    const std::string d = m.dev.output_dir + "/" + sub;
    std::filesystem::create_directories(d);
    return d;
}

// CUDA_week2/fd_wave_model.cu:138-231: the modified (lambda disc) model,
// stability report, all shots forward, receiver traces to
// <output_dir>/receiver/rec_signal.csv, szz snapshots of the last shot to
// <output_dir>/snap_data/vz_snap<k>.csv
void FDWaveModel::forward_simulate()
{
    const int nzt = grid.nzt, nxt = grid.nxt;
    medium.allocate_medium(nzt, nxt);
    medium.velocity_to_lame();
    if (dev.model_file.empty()) medium.stagger_over_grid_modify(nzt, nxt);
    else load_model(*this, dev.model_file, medium.lam, medium.mu, medium.rho);
    if (dev.write_csv) write_csv(io_dir(*this, "snap_data") + "/lamda_snap.csv", medium.rho, nzt, nxt);
    checkfd_ssg_elastic(grid.nx, grid.nz, grid.dx, grid.dt, source.fc.empty() ? 0.0 : source.fc[0], pml.npml,
                        medium.rho, medium.lam, medium.mu, hc);
    medium.allocate_medium_av(nzt - 1, nxt - 1);
    wave.allocate_wave(nzt, nxt);
    if (pml.npml > 0) pml.allocate_memory_arrays(nzt, nxt);

    build_plan(*this, impl_, false);
    psvfwi_plan* p = impl_->plan;
    upload_model(*this, p);
    tune_once(*this, impl_);
    const bool snaps = dev.write_csv && grid.snap_interval > 0;
    check(psvfwi_set_snapshots(p, snaps ? grid.snap_interval : 0, 4), "set_snapshots");
    check(psvfwi_forward(p, PSVFWI_ENERGY), "forward");
    download_records(*this, impl_);
    download_fields(*this, p, (int)nshots - 1);
    double ms = 0;
    long launches = 0;
    check(psvfwi_last_timing(p, &ms, &launches), "last_timing");
    std::cout << "Forward: " << nshots << " shot(s), " << grid.nt << " steps, " << ms << " ms on the GPU" << std::endl;
    if (!dev.write_csv) return;
    if (snaps) {
        int n = 0;
        check(psvfwi_snapshot_count(p, (int)nshots - 1, &n), "snapshot_count");
        real_sim** buf = nullptr;
        allocate_array_2d(buf, nzt, nxt);
        const std::string d = io_dir(*this, "snap_data");
        for (int k = 0; k < n; k++) {
            check(psvfwi_get_snapshot(p, (int)nshots - 1, k, buf[0]), "get_snapshot");
            write_csv(d + "/vz_snap" + std::to_string(k) + ".csv", buf, nzt, nxt);
        }
        deallocate_array_2d(buf);
    }
    write_csv(io_dir(*this, "receiver") + "/rec_signal.csv", receiver.signal, (int)receiver.nseis, grid.nt);
}

// CUDA_week2/fd_wave_model.cu:972-1002 (host version; the FWI loop uses the
// GPU misfit).  A residual above 1e10 is an error (the reference exits).
real_sim FDWaveModel::calculate_l2_adjoint_sources(unsigned int nr, int nt, real_sim dt, real_sim** signal,
                                                   real_sim** signal_meas)
{
    double m = 0.0;
    for (unsigned int ir = 0; ir < nr; ir++)
        for (int it = 0; it < nt; it++) {
            const double d = signal[ir][it] - signal_meas[ir][it];
            m += 0.5 * dt * d * d;
            signal[ir][it] = d;
            if (std::fabs(d) > 1.0e+10) throw std::runtime_error("residual exceptionally high");
        }
    return m;
}

// CUDA_week2/fd_wave_model.cu:1004-1052, on fwi.grad_* with the reference's indexing
void FDWaveModel::scale_gradients_with_energy_weights(real_sim** Ws, real_sim** Wr, real_sim C_rho, real_sim C_lam,
                                                      real_sim C_mu, int nz, int nx, int z1, int x1, int fdz, int fdx)
{
    scale_energy(fwi.grad_lam[0], fwi.grad_mu[0], fwi.grad_rho[0], Ws[0], Wr[0], grid.nzt, grid.nxt, C_rho, C_lam,
                 C_mu, nz, nx, z1, x1, fdz, fdx, false);
}

// summed misfit of a batched forward of the current (trial) model
static double batch_misfit(FDWaveModel& m, FDWaveModel::Impl* impl, int flags)
{
    check(psvfwi_forward(impl->plan, flags), "forward");
    std::vector<double> per(m.nshots);
    check(psvfwi_misfit(impl->plan, nullptr, per.data()), "misfit");
    double s = 0;
    for (double v : per) s += v;
    return s;
}

// CUDA_week2/fwi_step_length_estimation.cu:11-224: bracketing search over
// trial models (step halved until the misfit drops, then grown by 1.5 until
// it rises again, at most 6 tries each) and the parabola through
// (0, L2_0), (s1, L2_1), (s2, L2_2).  Each trial is one batched forward.
real_sim FDWaveModel::fwi_step_length_estimation(real_sim est, real_sim* L2_test)
{
    const int nfz = 1 + (fwi.z2 - fwi.z1) / fwi.dz, nfx = 1 + (fwi.x2 - fwi.x1) / fwi.dx;
    const bool clean = dev.gradient_scaling == "clean" || dev.gradient == "discrete";
    real_sim step_length[4] = {0, 0, 0, 0};
    const double rho_factor = 0.5, grow = 2.0;
    const unsigned stepmax = 6;
    unsigned step1 = 0, step2 = 0, itest = 1, countstep = 0;
    int checkstep = 0;
    while (step2 != 1 || step1 != 1) {
        checkstep++;
        auto upd = [&](real_sim** old, real_sim** cur, real_sim** g, double s, int type) {
            if (clean)
                clean_update(old, cur, g, s, nfz, nfx, grid.nzt, grid.nxt, fwi.z1, fwi.x1, fwi.dz, fwi.dx, false, type);
            else
                medium.fwi_grad_update(old, cur, g, s, nfz, nfx, fwi.z1, fwi.x1, fwi.dz, fwi.dx, false, type);
        };
        upd(medium.lam_old, medium.lam, fwi.grad_lam, est, 2);
        upd(medium.mu_old, medium.mu, fwi.grad_mu, est, 3);
        upd(medium.rho_old, medium.rho, fwi.grad_rho, rho_factor * est, 1);
        upload_model(*this, impl_->plan);
        tune_once(*this, impl_);
        L2_test[itest] = batch_misfit(*this, impl_, 0);
        step_length[itest] = est;
        step_length[0] = 0.0;
        if (dev.write_csv) {
            download_records(*this, impl_);
            const std::string d = io_dir(*this, "receiver"), tag = std::to_string(iter) + "_" + std::to_string(checkstep);
            write_csv(d + "/rec_signal_" + tag + ".csv", receiver.signal, (int)receiver.nseis, grid.nt);
            write_csv(d + "/src_signal_" + tag + ".csv", source.signal, (int)source.nseis, grid.nt);
        }
        std::cout << "Step length: " << est << ", L2 = " << L2_test[itest] << std::endl;
        if (step1 == 1 && L2_test[1] < L2_test[2]) {
            step_length[2] = est;
            step2 = 1;
        } else if (step1 == 1 && L2_test[1] >= L2_test[2]) {
            step_length[2] = est;
            est = est + est / grow;
            countstep++;
        } else if (step1 == 0 && L2_test[0] > L2_test[1]) {
            step_length[1] = est;
            step1 = 1;
            itest = 2;
            countstep = 0;
            est = est + est / grow;
        } else if (step1 == 0 && L2_test[0] <= L2_test[1]) {
            est = est / grow;
            countstep++;
        }
        if (step1 == 0 && countstep > stepmax) throw std::runtime_error("Step length estimation failed");
        if (step1 == 1 && countstep > stepmax) {
            std::cout << "Could not find a third step length that brackets the minimum" << std::endl;
            step2 = 1;
        }
    }
    return calc_opt_step(L2_test, step_length);
}

// Gaussian smoothing of one [nz][nx] array in place: a separable correlation with the
// normalised weights exp(-k^2 / (2 sigma^2)), |k| <= r = int(4 sigma + 0.5), rows first, then
// columns, indices beyond an edge clamped to it (the definition of scipy.ndimage's
// gaussian_filter with mode "nearest" and truncate 4).
static void gaussian_smooth(double* a, int nz, int nx, double sigma)
{
    const int r = (int)(4.0 * sigma + 0.5);
    std::vector<double> w((size_t)(2 * r + 1));
    double sum = 0.0;
    for (int k = -r; k <= r; k++) sum += (w[(size_t)(k + r)] = std::exp(-0.5 * k * k / (sigma * sigma)));
    for (double& v : w) v /= sum;
    std::vector<double> t((size_t)nz * nx);
    for (int z = 0; z < nz; z++)
        for (int x = 0; x < nx; x++) {
            double acc = 0.0;
            for (int k = -r; k <= r; k++) acc += w[(size_t)(k + r)] * a[(size_t)std::min(std::max(z + k, 0), nz - 1) * nx + x];
            t[(size_t)z * nx + x] = acc;
        }
    for (int z = 0; z < nz; z++)
        for (int x = 0; x < nx; x++) {
            double acc = 0.0;
            for (int k = -r; k <= r; k++) acc += w[(size_t)(k + r)] * t[(size_t)z * nx + std::min(std::max(x + k, 0), nx - 1)];
            a[(size_t)z * nx + x] = acc;
        }
}

void shape_direction(double* d[3], int nfz, int nfx, const MI355XOptions& o)
{
    for (int k = 0; k < 3; k++) {
        double* a = d[k];
        if (o.direction_smooth > 0.0) gaussian_smooth(a, nfz, nfx, o.direction_smooth);
        for (int z = 0; z < nfz; z++) {
            double g = 1.0;
            if (o.direction_depth_power > 0.0) g = std::pow((z + 1.0) / nfz, o.direction_depth_power);
            if (z < o.direction_mute) g = 0.0;
            else if (z < o.direction_mute + o.direction_taper)
                g *= 0.5 * (1.0 - std::cos(M_PI * (z - o.direction_mute + 0.5) / o.direction_taper));
            if (g != 1.0)
                for (int x = 0; x < nfx; x++) a[(size_t)z * nfx + x] *= g;
        }
    }
}

// Discrete gradient (an extension): the descent direction on the FWI window
// is -dJ/dm of the batch, divided by the illumination (the source energy
// sum_shots Ws, or sqrt(sum Ws * sum Wr) with the adjoint wavefield's energy
// Wr) plus 0.005 of its maximum over the window (a positive diagonal
// preconditioner, in place of the reference's energy scaling), then shaped by
// the [mi355x] direction options.
static void discrete_direction(const FDWaveModel& m, psvfwi_plan* p, const std::vector<double>& Ws,
                               const std::vector<double>* Wr, int nfz, int nfx)
{
    const int nzt = m.grid.nzt, nxt = m.grid.nxt;
    const size_t plane = (size_t)nzt * nxt;
    std::vector<double> g[3] = {std::vector<double>(plane), std::vector<double>(plane), std::vector<double>(plane)};
    check(psvfwi_get_model_gradient(p, -1, g[0].data(), g[1].data(), g[2].data()), "get_model_gradient");
    std::vector<double> ill(plane, 0.0);
    for (unsigned int s = 0; s < m.nshots; s++)
        for (size_t i = 0; i < plane; i++) ill[i] += Ws[s * plane + i];
    if (Wr)
        for (size_t i = 0; i < plane; i++) ill[i] = std::sqrt(ill[i] * (*Wr)[i]);
    double imax = 0.0;
    for (int zf = 0; zf < nfz; zf++)
        for (int xf = 0; xf < nfx; xf++) {
            const int iz = m.fwi.z1 + zf * m.fwi.dz, ix = m.fwi.x1 + xf * m.fwi.dx;
            if (iz < nzt && ix < nxt) imax = std::max(imax, ill[(size_t)iz * nxt + ix]);
        }
    real_sim** out[3] = {m.fwi.grad_lam, m.fwi.grad_mu, m.fwi.grad_rho};
    for (int zf = 0; zf < nfz; zf++)
        for (int xf = 0; xf < nfx; xf++) {
            const int iz = m.fwi.z1 + zf * m.fwi.dz, ix = m.fwi.x1 + xf * m.fwi.dx;
            for (int k = 0; k < 3; k++) out[k][zf][xf] = 0.0;
            if (iz >= nzt || ix >= nxt) continue;
            const size_t q = (size_t)iz * nxt + ix;
            const double w = ill[q] + 0.005 * imax;
            if (w <= 0.0) continue;
            for (int k = 0; k < 3; k++) out[k][zf][xf] = -g[k][q] / w;
        }
    double* d[3] = {out[0][0], out[1][0], out[2][0]};
    shape_direction(d, nfz, nfx, m.dev);
}

// CUDA_week2/fd_wave_model.cu:237-968 (the CPU branch's sequence), batched
void FDWaveModel::fwi_simulate()
{
    const int nzt = grid.nzt, nxt = grid.nxt, nt = grid.nt;
    const int nfz = 1 + (fwi.z2 - fwi.z1) / fwi.dz, nfx = 1 + (fwi.x2 - fwi.x1) / fwi.dx;
    const bool discrete = dev.gradient == "discrete";
    const bool clean = dev.gradient_scaling == "clean" || discrete;
    medium.allocate_medium(nzt, nxt);
    medium.velocity_to_lame();
    medium.allocate_medium_av(nzt - 1, nxt - 1);
    wave.allocate_wave(nzt, nxt);
    if (pml.npml > 0) pml.allocate_memory_arrays(nzt, nxt);
    // gradients only: the forward store stays on the device
    allocate_array_2d(fwi.grad_lam, nfz, nfx);
    allocate_array_2d(fwi.grad_mu, nfz, nfx);
    allocate_array_2d(fwi.grad_rho, nfz, nfx);

    build_plan(*this, impl_, true);
    psvfwi_plan* p = impl_->plan;

    // observed traces
    if (!receiver.signal_meas) allocate_array_2d(receiver.signal_meas, (int)receiver.nseis, nt);
    if (dev.observed == "modified") {
        medium.stagger_over_grid_modify(nzt, nxt);
        upload_model(*this, p);
        tune_once(*this, impl_);
        check(psvfwi_forward(p, 0), "forward");
        download_records(*this, impl_);
        for (unsigned int i = 0; i < receiver.nseis; i++)
            std::memcpy(receiver.signal_meas[i], receiver.signal[i], sizeof(double) * nt);
    } else if (dev.observed == "model") {
        load_model(*this, dev.true_model_file, medium.lam, medium.mu, medium.rho);
        upload_model(*this, p);
        tune_once(*this, impl_);
        check(psvfwi_forward(p, 0), "forward");
        download_records(*this, impl_);
        for (unsigned int i = 0; i < receiver.nseis; i++)
            std::memcpy(receiver.signal_meas[i], receiver.signal[i], sizeof(double) * nt);
    } else if (dev.observed == "file") {
        read_csv(receiver.signal_file, receiver.signal_meas, (int)receiver.nseis, nt);
    } else if (!receiver.measurement_records) {
        throw std::runtime_error("FWI needs observed traces: [receiver] measurement_records = 1 or [mi355x] observed");
    }
    {
        std::vector<double> obs;
        for (unsigned int s = 0; s < nshots; s++)
            for (int i : impl_->rec_idx[s]) obs.insert(obs.end(), receiver.signal_meas[i], receiver.signal_meas[i] + nt);
        check(psvfwi_set_observed(p, obs.data()), "set_observed");
    }
    if (dev.model_file.empty()) medium.stagger_over_grid(nzt, nxt);
    else load_model(*this, dev.model_file, medium.lam, medium.mu, medium.rho);
    for (int iz = 0; iz < nzt; iz++)
        for (int ix = 0; ix < nxt; ix++) {
            medium.rho_old[iz][ix] = medium.rho[iz][ix];
            medium.lam_old[iz][ix] = medium.lam[iz][ix];
            medium.mu_old[iz][ix] = medium.mu[iz][ix];
        }

    const size_t plane = (size_t)nzt * nxt, gsz = (size_t)nfz * nfx;
    std::vector<double> Ws(plane * nshots), Wr(plane), WrSum(plane), gl(gsz), gm(gsz), gr(gsz);
    real_sim step = 0.01;
    iter = 0;
    L2_norm = 1000.0;
    misfit_history.clear();
    while (L2_norm >= maxError) {
        if (iter >= maxIter) {
            std::cout << "Error: convergence not reached within maximum iteration" << std::endl;
            break;
        }
        std::cout << "Full Waveform Inversion iteration " << iter++ << std::endl;
        real_sim C_lam = 0, C_mu = 0, C_rho = 0;
        upload_model(*this, p);
        tune_once(*this, impl_);
        medium.medium_average(C_lam, C_mu, C_rho, nzt, nxt);

        real_sim L2_test[4] = {0, 0, 0, 0};
        L2_norm = batch_misfit(*this, impl_, PSVFWI_ENERGY | PSVFWI_GRADIENT);
        L2_test[0] = L2_test[3] = L2_norm;
        misfit_history.push_back(L2_norm);
        std::cout << "L2 norm = " << L2_norm << ", max error = " << maxError << std::endl;
        for (unsigned int s = 0; s < nshots; s++) check(psvfwi_get_field(p, (int)s, 5, &Ws[s * plane]), "get_field");
        const bool sr = discrete && dev.illumination == "source_receiver";
        if (sr) {
            // the discrete adjoint keeps no energy: the adjoint wavefield's, from a run
            // without the gradient
            check(psvfwi_adjoint(p, PSVFWI_ENERGY), "adjoint");
            std::fill(WrSum.begin(), WrSum.end(), 0.0);
            for (unsigned int s = 0; s < nshots; s++) {
                check(psvfwi_get_field(p, (int)s, 5, Wr.data()), "get_field");
                for (size_t i = 0; i < plane; i++) WrSum[i] += Wr[i];
            }
        }
        check(psvfwi_adjoint(p, PSVFWI_ENERGY | PSVFWI_GRADIENT), "adjoint");

        if (discrete) {
            discrete_direction(*this, p, Ws, sr ? &WrSum : nullptr, nfz, nfx);
        } else {
        reset_array_2d(fwi.grad_lam, nfz, nfx);
        reset_array_2d(fwi.grad_mu, nfz, nfx);
        reset_array_2d(fwi.grad_rho, nfz, nfx);
        for (unsigned int s = 0; s < nshots; s++) {
            check(psvfwi_get_field(p, (int)s, 5, Wr.data()), "get_field");
            check(psvfwi_get_gradient(p, (int)s, gl.data(), gm.data(), gr.data()), "get_gradient");
            scale_energy(gl.data(), gm.data(), gr.data(), &Ws[s * plane], Wr.data(), nzt, nxt, C_rho, C_lam, C_mu,
                         nfz, nfx, fwi.z1, fwi.x1, fwi.dz, fwi.dx, clean);
            for (size_t i = 0; i < gsz; i++) {
                fwi.grad_lam[0][i] += gl[i];
                fwi.grad_mu[0][i] += gm[i];
                fwi.grad_rho[0][i] += gr[i];
            }
        }
        std::memcpy(wave.We[0], Wr.data(), sizeof(double) * plane);
        }

        step = fwi_step_length_estimation(step, L2_test);
        std::cout << "Optimum step length: " << step << std::endl;
        if (clean) {
            clean_update(medium.rho_old, medium.rho, fwi.grad_rho, 0.5 * step, nfz, nfx, nzt, nxt, fwi.z1, fwi.x1,
                         fwi.dz, fwi.dx, true, 1);
            clean_update(medium.mu_old, medium.mu, fwi.grad_mu, step, nfz, nfx, nzt, nxt, fwi.z1, fwi.x1, fwi.dz,
                         fwi.dx, true, 3);
            clean_update(medium.lam_old, medium.lam, fwi.grad_lam, step, nfz, nfx, nzt, nxt, fwi.z1, fwi.x1, fwi.dz,
                         fwi.dx, true, 2);
        } else {
            medium.fwi_grad_update(medium.rho_old, medium.rho, fwi.grad_rho, 0.5 * step, nfz, nfx, fwi.z1, fwi.x1,
                                   fwi.dz, fwi.dx, true, 1);
            medium.fwi_grad_update(medium.mu_old, medium.mu, fwi.grad_mu, step, nfz, nfx, fwi.z1, fwi.x1, fwi.dz,
                                   fwi.dx, true, 3);
            medium.fwi_grad_update(medium.lam_old, medium.lam, fwi.grad_lam, step, nfz, nfx, fwi.z1, fwi.x1, fwi.dz,
                                   fwi.dx, true, 2);
        }
        if (dev.write_csv) {
            const std::string d = io_dir(*this, "density_snap"), k = std::to_string(iter);
            write_csv(d + "/density_snap" + k + ".csv", medium.rho, nzt, nxt);
            write_csv(d + "/lam_snap" + k + ".csv", medium.lam, nzt, nxt);
            write_csv(d + "/mu_snap" + k + ".csv", medium.mu, nzt, nxt);
        }
    }
}
