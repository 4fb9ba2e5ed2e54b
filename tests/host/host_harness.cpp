// host_harness.cpp -- test driver of the C++ host layer (fd_host.hpp,
// fd_kernel.hpp), built by the tests with g++ against libpsvfwi.so.
//   host_harness ini <file.ini> <out.txt> <section> <name> [...]
//       ParseError, then per key: Get | GetInteger(-7) | GetReal(-7.5) | GetBoolean(default -1)
//   host_harness config <file.ini> <out.txt>
//       the configuration FDWaveModel parses, as "key value..." lines
//   host_harness dropin <dir> [gpu]
//       forward_kernel_PSV + adjoint_kernel_PSV (with "gpu": the _GPU entry points,
//       null device pointers) on the problem in <dir>/in.txt and <dir>/*.bin (raw
//       float64), outputs to <dir>/out_*.bin; the gradient arrays hold 7.0 on entry;
//       with snap_interval > 0 a second forward (fwinv = false) writes
//       <dir>/snap/vz_snap<k>.csv
//   host_harness opt <L0> <L1> <L2> <s0> <s1> <s2> <out.txt>
//       calc_opt_step of the three trial misfits and steps
//   host_harness shape <in.bin> <nfz> <nfx> <smooth> <power> <mute> <taper> <out.bin>
//       shape_direction on three [nfz][nfx] float64 arrays
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "fd_host.hpp"
#include "fd_kernel.hpp"

namespace {

std::vector<double> load(const std::string& path, size_t n)
{
    std::vector<double> v(n);
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(double), n, f) != n) throw std::runtime_error("cannot read " + path);
    std::fclose(f);
    return v;
}

void save(const std::string& path, const double* p, size_t n)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(p, sizeof(double), n, f) != n) throw std::runtime_error("cannot write " + path);
    std::fclose(f);
}

// row-pointer table over a flat vector
std::vector<double*> rows(std::vector<double>& v, int d1, int d2)
{
    std::vector<double*> r((size_t)d1);
    for (int i = 0; i < d1; i++) r[(size_t)i] = v.data() + (size_t)i * d2;
    return r;
}

template <class T>
void put(std::ofstream& o, const char* k, const std::vector<T>& v)
{
    o << k;
    for (const T& x : v) o << " " << x;
    o << "\n";
}

int run_ini(int argc, char** argv)
{
    INIReader r(argv[2]);
    std::ofstream o(argv[3]);
    o << r.ParseError() << "\n";
    for (int i = 4; i + 1 < argc; i += 2) {
        const bool b = r.GetBoolean(argv[i], argv[i + 1], true), c = r.GetBoolean(argv[i], argv[i + 1], false);
        std::string g = r.Get(argv[i], argv[i + 1], "<default>");
        for (char& ch : g)
            if (ch == '\n') ch = '|';
        o.precision(17);
        o << g << "\t" << r.GetInteger(argv[i], argv[i + 1], -7) << "\t" << r.GetReal(argv[i], argv[i + 1], -7.5)
          << "\t" << (b == c ? (b ? 1 : 0) : -1) << "\n";
    }
    return 0;
}

int run_config(char** argv)
{
    FDWaveModel m(argv[2]);
    std::ofstream o(argv[3]);
    o.precision(17);
    o << "grid " << m.grid.nt << " " << m.grid.nx << " " << m.grid.nz << " " << m.grid.dx << " " << m.grid.dz << " "
      << m.grid.dt << " " << m.grid.snap_interval << " " << m.grid.fsurf << " " << m.grid.fpad << " " << m.grid.ppad
      << " " << m.grid.nxt << " " << m.grid.nzt << "\n";
    o << "medium " << m.medium.scalar_rho << " " << m.medium.scalar_vp << " " << m.medium.scalar_vs << " "
      << m.medium.scalar_lam << " " << m.medium.scalar_mu << "\n";
    o << "pml " << m.pml.npml << " " << m.pml.dh << " " << m.pml.fpml << " " << m.pml.npower << " " << m.pml.k_max_PML
      << " " << m.pml.rcoef << " " << m.pml.damp_v_PML << "\n";
    o << "fdm " << m.fdorder << " " << m.max_rel_err << " " << m.fwinv << " " << m.nshots << " " << m.maxIter << " "
      << m.maxError << "\n";
    put(o, "hc", std::vector<double>(m.hc, m.hc + 7));
    put(o, "src_x", m.source.x_seis);
    put(o, "src_z", m.source.z_seis);
    put(o, "src_comp", m.source.seis_comp);
    put(o, "src_shot", m.source.shot_to_fire);
    put(o, "src_fc", m.source.fc);
    put(o, "rec_x", m.receiver.x_seis);
    put(o, "rec_z", m.receiver.z_seis);
    put(o, "rec_comp", m.receiver.seis_comp);
    put(o, "rec_shot", m.receiver.shot_to_fire);
    if (m.fwinv)
        o << "fwi " << m.fwi.dt << " " << m.fwi.dx << " " << m.fwi.dz << " " << m.fwi.x1 << " " << m.fwi.x2 << " "
          << m.fwi.z1 << " " << m.fwi.z2 << "\n";
    o << "dev " << m.dev.precision << " " << m.dev.gradient_store << " " << m.dev.checkpoint_interval << " "
      << m.dev.tile << " " << (m.dev.observed.empty() ? "-" : m.dev.observed) << " " << m.dev.gradient_scaling
      << "\n";
    o << "direction " << m.dev.illumination << " " << m.dev.direction_smooth << " " << m.dev.direction_depth_power
      << " " << m.dev.direction_mute << " " << m.dev.direction_taper << "\n";
    if (m.source.nseis > 0) put(o, "src_signal0", std::vector<double>(m.source.signal[0], m.source.signal[0] + m.grid.nt));
    if (m.pml.npml > 0) {
        const int tp = 2 * (m.pml.npml + 1);
        std::vector<double> prof;
        for (real_sim* a : {m.pml.a, m.pml.b, m.pml.K, m.pml.a_half, m.pml.b_half, m.pml.K_half})
            prof.insert(prof.end(), a, a + tp);
        put(o, "profiles", prof);
    }
    return 0;
}

int run_opt(char** argv)
{
    real_sim L2[3], st[3];
    for (int i = 0; i < 3; i++) {
        L2[i] = std::stod(argv[2 + i]);
        st[i] = std::stod(argv[5 + i]);
    }
    std::ofstream o(argv[8]);
    o.precision(17);
    o << calc_opt_step(L2, st) << "\n";
    return 0;
}

int run_shape(char** argv)
{
    const int nfz = std::stoi(argv[3]), nfx = std::stoi(argv[4]);
    MI355XOptions o;
    o.direction_smooth = std::stod(argv[5]);
    o.direction_depth_power = std::stod(argv[6]);
    o.direction_mute = std::stoi(argv[7]);
    o.direction_taper = std::stoi(argv[8]);
    const size_t n = (size_t)nfz * nfx;
    std::vector<double> a(3 * n);
    std::ifstream in(argv[2], std::ios::binary);
    in.read(reinterpret_cast<char*>(a.data()), (std::streamsize)(a.size() * sizeof(double)));
    if (!in) throw std::runtime_error("shape: short input");
    double* d[3] = {a.data(), a.data() + n, a.data() + 2 * n};
    shape_direction(d, nfz, nfx, o);
    std::ofstream out(argv[9], std::ios::binary);
    out.write(reinterpret_cast<const char*>(a.data()), (std::streamsize)(a.size() * sizeof(double)));
    return 0;
}

int run_dropin(const std::string& dir, bool gpu)
{
    std::ifstream in(dir + "/in.txt");
    int ishot, nt, nzt, nxt, fpad, ppad, snap, fsurf, fdorder, npml, nsrc, nrec, fwinv, fdt, fdx, fdz, x1, x2, z1, z2, fp64;
    double dt, dx, dz, hc[7];
    in >> ishot >> nt >> nzt >> nxt >> fpad >> ppad >> dt >> dx >> dz >> snap >> fsurf >> fdorder >> npml >> nsrc >>
        nrec >> fwinv >> fdt >> fdx >> fdz >> x1 >> x2 >> z1 >> z2 >> fp64;
    for (double& h : hc) in >> h;
    ivec sx(nsrc), sz(nsrc), sc(nsrc), ss(nsrc), rx(nrec), rz(nrec), rc(nrec), rs(nrec);
    for (ivec* v : {&sx, &sz, &sc, &ss})
        for (int& x : *v) in >> x;
    for (ivec* v : {&rx, &rz, &rc, &rs})
        for (int& x : *v) in >> x;
    if (!in) throw std::runtime_error("bad in.txt");
    fd_kernel_set_precision(fp64);
    const size_t n = (size_t)nzt * nxt, na = (size_t)(nzt - 1) * (nxt - 1);
    const int tp = 2 * (npml + 1);
    std::vector<double> lam = load(dir + "/lam.bin", n), mu = load(dir + "/mu.bin", n);
    std::vector<double> muzx = load(dir + "/mu_zx.bin", na), rzp = load(dir + "/rho_zp.bin", na),
                        rxp = load(dir + "/rho_xp.bin", na);
    std::vector<double> prof = load(dir + "/pml.bin", (size_t)6 * tp);
    std::vector<double> ssig = load(dir + "/src_signal.bin", (size_t)nsrc * nt);
    std::vector<double> asig = load(dir + "/adj_signal.bin", (size_t)nrec * nt);
    std::vector<double> rsig((size_t)nrec * nt, -1.0);  // rows of other shots must stay untouched
    const int nft = 1 + (nt - 1) / fdt, nfz = 1 + (z2 - z1) / fdz, nfx = 1 + (x2 - x1) / fdx;
    const size_t vol = (size_t)nft * nfz * nfx;
    std::vector<double> fld(6 * n), mem(4 * (size_t)nzt * tp + 4 * (size_t)tp * nxt), store(5 * vol),
        grad((size_t)nfz * nfx, 7.0);  // overwritten by the adjoint
    std::vector<std::vector<double*>> F;
    for (int k = 0; k < 6; k++) {
        std::vector<double> tmp;
        F.push_back(std::vector<double*>(nzt));
        for (int i = 0; i < nzt; i++) F.back()[i] = fld.data() + k * n + (size_t)i * nxt;
    }
    std::vector<std::vector<double*>> M;
    for (int k = 0; k < 8; k++) {
        const bool xt = k < 4;
        const int d1 = xt ? nzt : tp, d2 = xt ? tp : nxt;
        double* base = mem.data() + (xt ? (size_t)k * nzt * tp : 4 * (size_t)nzt * tp + (size_t)(k - 4) * tp * nxt);
        M.push_back(std::vector<double*>(d1));
        for (int i = 0; i < d1; i++) M.back()[i] = base + (size_t)i * d2;
    }
    // fields 0..5: vx vz sxx szx szz We; memory order: mem_vx_x mem_vz_x mem_sxx_x mem_szx_x | mem_vx_z mem_vz_z mem_szz_z mem_szx_z
    std::vector<double**> S3[5];
    std::vector<std::vector<double*>> S2;
    double*** st[5];
    for (int k = 0; k < 5; k++) {
        S2.push_back(std::vector<double*>((size_t)nft * nfz));
        for (size_t r = 0; r < (size_t)nft * nfz; r++) S2.back()[r] = store.data() + k * vol + r * nfx;
        S3[k].resize(nft);
        for (int t = 0; t < nft; t++) S3[k][t] = S2.back().data() + (size_t)t * nfz;
        st[k] = S3[k].data();
    }
    auto L = rows(lam, nzt, nxt), U = rows(mu, nzt, nxt), MZ = rows(muzx, nzt - 1, nxt - 1),
         RZ = rows(rzp, nzt - 1, nxt - 1), RX = rows(rxp, nzt - 1, nxt - 1), SS = rows(ssig, nsrc, nt),
         AS = rows(asig, nrec, nt), RS = rows(rsig, nrec, nt), GL = rows(grad, nfz, nfx);
    std::vector<double> g2(grad), g3(grad);
    auto GM = rows(g2, nfz, nfx), GR = rows(g3, nfz, nfx);
    double* pr[6];
    for (int k = 0; k < 6; k++) pr[k] = prof.data() + (size_t)k * tp;

    fd_kernel_set_snapshot_dir("");
    if (gpu) {
        // the _GPU variants take the device buffers the reference's caller cudaMallocs;
        // here the plan owns device memory, so they are passed as null
        double* d = nullptr;
        forward_kernel_PSV_GPU(ishot, nt, nzt, nxt, fpad, ppad, dt, dx, dz, snap, fsurf, hc, fdorder, F[0].data(),
                               F[1].data(), F[2].data(), F[3].data(), F[4].data(), L.data(), U.data(), MZ.data(),
                               RZ.data(), RX.data(), npml, pr[0], pr[1], pr[2], pr[3], pr[4], pr[5], M[0].data(),
                               M[4].data(), M[1].data(), M[5].data(), M[2].data(), M[3].data(), M[6].data(),
                               M[7].data(), nsrc, sx, sz, sc, SS.data(), ss, nrec, rx, rz, rc, RS.data(), rs,
                               fwinv != 0, fdt, fdx, fdz, x1, x2, z1, z2, st[0], st[1], st[2], st[3], st[4], d, d,
                               d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d);
    } else {
        forward_kernel_PSV(ishot, nt, nzt, nxt, fpad, ppad, dt, dx, dz, snap, fsurf, hc, fdorder, F[0].data(),
                           F[1].data(), F[2].data(), F[3].data(), F[4].data(), F[5].data(), L.data(), U.data(),
                           MZ.data(), RZ.data(), RX.data(), npml, pr[0], pr[1], pr[2], pr[3], pr[4], pr[5],
                           M[0].data(), M[4].data(), M[1].data(), M[5].data(), M[2].data(), M[3].data(), M[6].data(),
                           M[7].data(), nsrc, sx, sz, sc, SS.data(), ss, nrec, rx, rz, rc, RS.data(), rs,
                           fwinv != 0, fdt, fdx, fdz, x1, x2, z1, z2, st[0], st[1], st[2], st[3], st[4]);
    }
    save(dir + "/out_rec.bin", rsig.data(), rsig.size());
    save(dir + "/out_fwd_fields.bin", fld.data(), fld.size());
    save(dir + "/out_fwd_mem.bin", mem.data(), mem.size());
    save(dir + "/out_store.bin", store.data(), store.size());
    if (gpu) {
        double* d = nullptr;
        adjoint_kernel_PSV_GPU(ishot, nt, nzt, nxt, fpad, ppad, dt, dx, dz, snap, fsurf, hc, fdorder, F[0].data(),
                               F[1].data(), F[2].data(), F[3].data(), F[4].data(), L.data(), U.data(), MZ.data(),
                               RZ.data(), RX.data(), npml, pr[0], pr[1], pr[2], pr[3], pr[4], pr[5], M[0].data(),
                               M[4].data(), M[1].data(), M[5].data(), M[2].data(), M[3].data(), M[6].data(),
                               M[7].data(), nrec, rx, rz, rc, AS.data(), rs, fwinv != 0, fdt, fdx, fdz, x1, x2, z1,
                               z2, st[0], st[1], st[2], st[3], st[4], GL.data(), GM.data(), GR.data(), d, d, d, d, d,
                               d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d, d);
    } else {
        adjoint_kernel_PSV(ishot, nt, nzt, nxt, fpad, ppad, dt, dx, dz, snap, fsurf, hc, fdorder, F[0].data(),
                           F[1].data(), F[2].data(), F[3].data(), F[4].data(), F[5].data(), L.data(), U.data(),
                           MZ.data(), RZ.data(), RX.data(), npml, pr[0], pr[1], pr[2], pr[3], pr[4], pr[5],
                           M[0].data(), M[4].data(), M[1].data(), M[5].data(), M[2].data(), M[3].data(), M[6].data(),
                           M[7].data(), nrec, rx, rz, rc, AS.data(), rs, fwinv != 0, fdt, fdx, fdz, x1, x2, z1, z2,
                           st[0], st[1], st[2], st[3], st[4], GL.data(), GM.data(), GR.data());
    }
    save(dir + "/out_adj_fields.bin", fld.data(), fld.size());
    std::vector<double> g(grad);
    g.insert(g.end(), g2.begin(), g2.end());
    g.insert(g.end(), g3.begin(), g3.end());
    save(dir + "/out_grad.bin", g.data(), g.size());
    if (snap > 0) {
        std::filesystem::create_directories(dir + "/snap");
        fd_kernel_set_snapshot_dir((dir + "/snap").c_str());
        forward_kernel_PSV(ishot, nt, nzt, nxt, fpad, ppad, dt, dx, dz, snap, fsurf, hc, fdorder, F[0].data(),
                           F[1].data(), F[2].data(), F[3].data(), F[4].data(), F[5].data(), L.data(), U.data(),
                           MZ.data(), RZ.data(), RX.data(), npml, pr[0], pr[1], pr[2], pr[3], pr[4], pr[5],
                           M[0].data(), M[4].data(), M[1].data(), M[5].data(), M[2].data(), M[3].data(), M[6].data(),
                           M[7].data(), nsrc, sx, sz, sc, SS.data(), ss, nrec, rx, rz, rc, RS.data(), rs, false, fdt,
                           fdx, fdz, x1, x2, z1, z2, st[0], st[1], st[2], st[3], st[4]);
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        const std::string mode = argc > 1 ? argv[1] : "";
        if (mode == "ini" && argc >= 4) return run_ini(argc, argv);
        if (mode == "config" && argc >= 4) return run_config(argv);
        if (mode == "dropin" && argc >= 3) return run_dropin(argv[2], argc >= 4 && std::string(argv[3]) == "gpu");
        if (mode == "opt" && argc >= 9) return run_opt(argv);
        if (mode == "shape" && argc >= 10) return run_shape(argv);
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    std::cerr << "usage: host_harness ini|config|dropin|opt|shape ..." << std::endl;
    return 2;
}
