// fd_host.hpp -- C++ host API of the P-SV FWI hot path, drop-in for the
// reference's host classes:
//   INIReader      CUDA_week2/INIReader.cuh:304-355   (own parser, same syntax)
//   FDGrid         CUDA_week2/fd_grid.cuh             (fd_grid.cu:15-71)
//   FDMedium       CUDA_week2/fd_medium.cuh           (fd_medium.cu:29-302)
//   FDCpml         CUDA_week2/fd_cpml.cuh:13-48       (fd_cpml.cu:28-284)
//   FDWave         CUDA_week2/fd_wave.cuh:14-29       (fd_wave.cu:14-46)
//   FDFWI          CUDA_week2/fd_fwi.cuh:14-35        (fd_fwi.cu:16-111)
//   FDSeismogram   CUDA_week2/fd_seismogram.cuh:19-52 (fd_seismogram.cu:18-186)
//   FDWaveModel    CUDA_week2/fd_wave_model.cuh:22-68 (fd_wave_model.cu:33-1052)
// Member and method names, argument meanings and the allocate -> reset ->
// kernel -> deallocate call sequence are the reference's; arrays are
// contiguous row-major with row-pointer tables (arr[0] is the flat buffer,
// CUDA_week2/util.cu:24-74).  Differences, all listed in the design notes:
// the CPML members are per instance (the reference's are static), errors
// throw std::runtime_error instead of exit(), the FWI window is clamped to
// the grid, and a [mi355x] INI section selects the device options.
#pragma once

#include <map>
#include <set>
#include <string>
#include <vector>

using real_sim = double;
using ivec = std::vector<int>;

// ---------------------------------------------------------------- INI
// Same syntax as the vendored inih (CUDA_week2/INIReader.cuh): [section]
// lines, name = value or name : value, ';' / '#' comment lines, ';' inline
// comments only after whitespace, indented continuation lines joined with
// '\n', repeated keys joined with '\n', case-insensitive section and key
// names, prefix parsing of numbers (strtol base 0 / strtod).  No line-length
// limit (inih stops at 200 characters, INIReader.cuh:86-87).
class INIReader {
public:
    explicit INIReader(const std::string& filename);
    int ParseError() const { return error_; }  // 0, -1 (cannot open) or first bad line
    std::string Get(const std::string& section, const std::string& name,
                    const std::string& default_value) const;
    long GetInteger(const std::string& section, const std::string& name, long default_value) const;
    double GetReal(const std::string& section, const std::string& name, double default_value) const;
    bool GetBoolean(const std::string& section, const std::string& name, bool default_value) const;
    bool Has(const std::string& section, const std::string& name) const;
    const std::set<std::string>& Sections() const { return sections_; }

private:
    int error_ = 0;
    std::map<std::string, std::string> values_;
    std::set<std::string> sections_;
};

// "{a, b, c}" -> numbers (CUDA_week2/util.cu:140-163); integers truncate
template <class T>
void parse_string_to_vector(const std::string& text, std::vector<T>* out);

// ---------------------------------------------------------------- arrays
template <class T> void allocate_array_1d(T*& a, int n);
template <class T> void allocate_array_2d(T**& a, int d1, int d2);
template <class T> void allocate_array_3d(T***& a, int d1, int d2, int d3);
template <class T> void reset_array_2d(T** a, int d1, int d2);
template <class T> void reset_array_3d(T*** a, int d1, int d2, int d3);
template <class T> void deallocate_array_1d(T*& a);
template <class T> void deallocate_array_2d(T**& a, int d1 = 0);
template <class T> void deallocate_array_3d(T***& a, int d1 = 0, int d2 = 0);

// ---------------------------------------------------------------- classes
class FDGrid {
public:
    int nt = 0, nx = 0, nz = 0;
    real_sim dx = 0, dz = 0, dt = 0;
    int snap_interval = 0;
    bool fsurf = false;
    int fpad = 0, ppad = 0, nxt = 0, nzt = 0;
    void parse_configuration(const char* path);
    void total_grid(int fdorder, int npml);
};

class FDMedium {
public:
    real_sim scalar_rho = 0, scalar_vp = 0, scalar_vs = 0, scalar_lam = 0, scalar_mu = 0;
    real_sim **rho = nullptr, **lam = nullptr, **mu = nullptr;
    real_sim **rho_old = nullptr, **lam_old = nullptr, **mu_old = nullptr;
    real_sim **mu_zx = nullptr, **rho_zp = nullptr, **rho_xp = nullptr;
    void parse_configuration(const char* path);
    void velocity_to_lame();
    void stagger_over_grid(int dimz, int dimx);
    void stagger_over_grid_modify(int dimz, int dimx);
    void average_parameters(int dimz, int dimx);
    // reference-mode material update (CUDA_week2/fd_medium.cu:161-240)
    void fwi_grad_update(real_sim** mat, real_sim** mat_new, real_sim** grad_mat, real_sim step_length,
                         int dimz, int dimx, int z1, int x1, int fwi_dz, int fwi_dx, bool override,
                         int matType);
    void allocate_medium(int dimz, int dimx);
    void medium_average(real_sim& C_lam, real_sim& C_mu, real_sim& C_rho, int dimz, int dimx);
    void allocate_medium_av(int dimz, int dimx);
    void deallocate_medium(int dimz);
    void deallocate_medium_av(int dimz);
};

class FDCpml {
public:
    int npml = 10, tpml = 22;
    real_sim dh = 0.1, fpml = 20.0, npower = 2.0, k_max_PML = 1.0, rcoef = 0.001, damp_v_PML = 0.0;
    real_sim *a = nullptr, *b = nullptr, *K = nullptr, *a_half = nullptr, *b_half = nullptr, *K_half = nullptr;
    real_sim **mem_vx_x = nullptr, **mem_vx_z = nullptr, **mem_vz_x = nullptr, **mem_vz_z = nullptr;
    real_sim **mem_sxx_x = nullptr, **mem_szx_x = nullptr, **mem_szz_z = nullptr, **mem_szx_z = nullptr;
    void parse_configuration(const char* path);
    void cpml_pro(real_sim dt);
    void allocate_pml();
    void allocate_memory_arrays(int dimz, int dimx);
    void reset_memory_arrays(int dimz, int dimx);
    void deallocate_pml();
    void deallocate_memory_arrays(int dimz);
};

class FDWave {
public:
    real_sim **vx = nullptr, **vz = nullptr, **sxx = nullptr, **szx = nullptr, **szz = nullptr, **We = nullptr;
    void allocate_wave(int dimz, int dimx);
    void reset_kernel(int dimz, int dimx);
    void deallocate_wave(int dimz);
};

class FDFWI {
public:
    int dt = 1, dx = 1, dz = 1, x1 = 0, x2 = 0, z1 = 0, z2 = 0;
    real_sim ***vx = nullptr, ***vz = nullptr, ***sxx = nullptr, ***szx = nullptr, ***szz = nullptr;
    real_sim **grad_lam = nullptr, **grad_mu = nullptr, **grad_rho = nullptr;
    void parse_configuration(const char* path, const FDGrid& grid);
    void allocate_fwi(int dimt, int dimz, int dimx);
    void reset_fwi_kernel(int dimt, int dimz, int dimx, bool grad_reset);
    void deallocate_fwi(int dimt, int dimz);
};

class FDSeismogram {
public:
    unsigned int nseis = 0;
    int read_from_data_files = 0;
    std::string seis_file, signal_file;
    ivec x_seis, z_seis, v_or_s, seis_comp, shot_to_fire, shape;
    std::vector<real_sim> angle_v, amp, fc, ts;
    real_sim **signal = nullptr, **signal_meas = nullptr;
    int measurement_records = 0;
    void parse_configuration(const char* path, const char sr, const FDGrid& grid);
    void grid_adjustment(const FDGrid& grid);
};

void create_wavelet(real_sim* signal, int nt, real_sim dt, real_sim amp, real_sim fc, real_sim ts, int shape);
void holbergcoeff(const int FDORDER, const int MAXRELERROR, real_sim* hc);
void checkfd_ssg_elastic(int NX, int NY, real_sim DH, real_sim DT, real_sim TS, int FW,
                         real_sim** prho, real_sim** ppi, real_sim** pu, real_sim* hc);

// device options of the [mi355x] INI section (all optional)
struct MI355XOptions {
    std::string precision = "fp32";     // fp32 | fp64 | fp32x2 (compensated FP32) | fp32fma
    std::string gradient_store = "full"; // full | checkpoint
    int checkpoint_interval = 200;
    int tile = 0;                        // 0: autotune
    std::string observed = "";           // "" (reference: zeros), "modified" (forward of the
                                         // stagger_over_grid_modify model), "file" (receiver signal_file)
                                         // or "model" (forward of true_model_file)
    std::string model_file = "";         // raw FP64 vp, vs, rho, each [nz][nx] row-major (the unpadded
                                         // model; edges repeated into the padding): the model in place of
                                         // the [medium] scalars (the FWI start model)
    std::string true_model_file = "";    // the same layout: the model observed = model forwards
    std::string gradient_scaling = "reference";  // reference | clean (energy weights, model update)
    std::string gradient = "reference";  // reference (correlation) | discrete (exact misfit gradient;
                                         // implies the clean, window-correct update)
    int gradient_segment = 64;           // discrete: steps between checkpoints (1: every state kept)
    // discrete: shaping of the descent direction on the FWI window grid (an extension; the
    // reference's loop has none of it): illumination "source" (sum_shots Ws) or
    // "source_receiver" (sqrt(sum Ws * sum Wr), one more adjoint run per iteration); Gaussian
    // smoothing of sigma window cells; depth gain ((row + 1) / rows)^power; the first `mute`
    // rows zeroed and the `taper` rows below them ramped up by a cosine
    std::string illumination = "source";
    double direction_smooth = 0.0;
    double direction_depth_power = 0.0;
    int direction_mute = 0;
    int direction_taper = 0;
    std::string output_dir = "./io";
    bool write_csv = true;
    void parse(const char* path);
};

// Applies the direction shaping of `o` (smoothing, depth gain, mute, taper, in that order) to
// the three [nfz][nfx] direction arrays in place.
void shape_direction(double* d[3], int nfz, int nfx, const MI355XOptions& o);

class FDWaveModel {
public:
    FDGrid grid;
    FDMedium medium;
    FDCpml pml;
    FDWave wave;
    FDSeismogram source, receiver;
    FDFWI fwi;
    MI355XOptions dev;
    bool fwinv = false;
    real_sim misfit = 0;
    real_sim hc[7] = {0, 0, 0, 0, 0, 0, 0};
    int max_rel_err = 1;
    real_sim L2_norm = 0;
    int iter = 0;
    unsigned int fdorder = 2, nshots = 1;
    int maxIter = 500;
    real_sim maxError = 1e-3;
    std::vector<real_sim> misfit_history;  // per FWI iteration

    explicit FDWaveModel(const char* path);
    ~FDWaveModel();
    void parse_configuration(const char* path);
    void forward_simulate();
    void fwi_simulate();
    real_sim calculate_l2_adjoint_sources(unsigned int nr, int nt, real_sim dt, real_sim** signal,
                                          real_sim** signal_meas);
    real_sim fwi_step_length_estimation(real_sim est_step_length, real_sim* L2_test);
    void scale_gradients_with_energy_weights(real_sim** Ws, real_sim** Wr, real_sim C_rho, real_sim C_lam,
                                             real_sim C_mu, int nz, int nx, int z1, int x1, int fdz, int fdx);

    struct Impl;  // device plan and per-shot source / receiver lists

private:
    Impl* impl_ = nullptr;
};

// ---------------------------------------------------------------- template bodies
#include <algorithm>
#include <cctype>
#include <cstdlib>

template <class T>
void parse_string_to_vector(const std::string& text, std::vector<T>* out)
{
    // spaces removed, text up to the first '}', the leading character (the
    // '{') dropped, comma-separated tokens read as reals (atof semantics)
    std::string s;
    for (char c : text)
        if (!std::isspace(static_cast<unsigned char>(c))) s += c;
    const size_t close = s.find('}');
    if (close != std::string::npos) s.erase(close);
    if (!s.empty()) s.erase(0, 1);
    size_t from = 0;
    while (true) {
        const size_t comma = s.find(',', from);
        const std::string tok = s.substr(from, comma == std::string::npos ? std::string::npos : comma - from);
        out->push_back(static_cast<T>(std::atof(tok.c_str())));
        if (comma == std::string::npos) break;
        from = comma + 1;
    }
}

template <class T>
void allocate_array_1d(T*& a, int n)
{
    a = new T[(size_t)(n > 0 ? n : 0)]();
}

template <class T>
void allocate_array_2d(T**& a, int d1, int d2)
{
    a = new T*[(size_t)(d1 > 0 ? d1 : 1)];
    a[0] = new T[(size_t)d1 * (size_t)d2]();
    for (int i = 1; i < d1; i++) a[i] = a[0] + (size_t)i * d2;
}

template <class T>
void allocate_array_3d(T***& a, int d1, int d2, int d3)
{
    a = new T**[(size_t)(d1 > 0 ? d1 : 1)];
    T** rows = new T*[(size_t)d1 * d2 + 1];
    T* data = new T[(size_t)d1 * d2 * d3]();
    for (size_t r = 0; r < (size_t)d1 * d2; r++) rows[r] = data + r * d3;
    if ((size_t)d1 * d2 == 0) rows[0] = data;
    for (int i = 0; i < d1; i++) a[i] = rows + (size_t)i * d2;
    if (d1 == 0) a[0] = rows;
}

template <class T>
void reset_array_2d(T** a, int d1, int d2)
{
    std::fill(a[0], a[0] + (size_t)d1 * d2, T(0));
}

template <class T>
void reset_array_3d(T*** a, int d1, int d2, int d3)
{
    std::fill(a[0][0], a[0][0] + (size_t)d1 * d2 * d3, T(0));
}

template <class T>
void deallocate_array_1d(T*& a)
{
    delete[] a;
    a = nullptr;
}

template <class T>
void deallocate_array_2d(T**& a, int)
{
    if (!a) return;
    delete[] a[0];
    delete[] a;
    a = nullptr;
}

template <class T>
void deallocate_array_3d(T***& a, int, int)
{
    if (!a) return;
    delete[] a[0][0];
    delete[] a[0];
    delete[] a;
    a = nullptr;
}
