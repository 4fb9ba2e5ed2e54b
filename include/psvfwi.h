/*
 * psvfwi.h -- C ABI of the MI355X (gfx950) 2D elastic P-SV FWI hot path.
 *
 * This is the drop-in boundary for the reference's propagation entry points
 * (CUDA_week2/fd_kernel.cuh:12-173).  It covers:
 *   forward_kernel_PSV  (fd_kernel.cuh:12-40)   -> psvfwi_forward + getters
 *   adjoint_kernel_PSV  (fd_kernel.cuh:43-70)   -> psvfwi_adjoint + getters
 *   forward_kernel_PSV_GPU / adjoint_kernel_PSV_GPU (fd_kernel.cuh:75-173):
 *                        the same two calls; device memory is owned by the plan
 *   calc_opt_step       (fd_kernel.cuh:72)      -> psvfwi_parabola_step
 *   calculate_l2_adjoint_sources (CUDA_week2/fd_wave_model.cu:972-1002)
 *                                               -> psvfwi_misfit (FP64 on the GPU)
 *
 * Plain C types only: flat row-major arrays and sizes.  Host arrays are FP64
 * (real_sim = double, CUDA_week2/globvar.cuh:16) and are converted to the
 * plan's arithmetic type on upload.  Every call returns a status; nothing
 * exits the process.  One plan per host thread; a plan owns its HIP stream.
 *
 * Index conventions follow the reference: a padded grid of nzt x nxt cells,
 * fpad = fdorder/2, ppad = npml + fpad, interior [fpad, nzt-fpad) x
 * [fpad, nxt-fpad) (CUDA_week2/fd_grid.cu:52-64); source and receiver
 * coordinates are padded-grid indices (after grid_adjustment,
 * CUDA_week2/fd_seismogram.cu:153-164).
 */
#ifndef PSVFWI_H
#define PSVFWI_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum psvfwi_status {
    PSVFWI_OK = 0,
    PSVFWI_E_ARG = -1,        /* invalid argument or inconsistent sizes */
    PSVFWI_E_HIP = -2,        /* HIP runtime error (message in psvfwi_last_error) */
    PSVFWI_E_ORDER = -3,      /* fdorder other than 2 or 4 */
    PSVFWI_E_NONFINITE = -4,  /* |residual| > 1e10 or a non-finite misfit */
    PSVFWI_E_STATE = -5,      /* call out of sequence (e.g. adjoint before forward) */
    PSVFWI_E_NOMEM = -6       /* device allocation failed */
} psvfwi_status;

/* Arithmetic of a plan's kernels (the reference computes in FP64,
 * CUDA_week2/globvar.cuh:16).  PSVFWI_FP32 and PSVFWI_FP64 round every operation on its
 * own, as the CPU path does (FP64 equals it bit for bit).  PSVFWI_FP32_FMA is the throughput
 * precision: FP32 whose multiply-adds fuse into FMA instructions (fewer instructions and
 * roundings; not bit-identical to PSVFWI_FP32).  PSVFWI_FP32X2 is compensated FP32: every
 * value is an unevaluated sum of two floats and every operation a double-word algorithm
 * built from FP32 operations (about 48 significant bits, 8 bytes per value), for cases whose
 * residual is far below the wavefield (a small anomaly), where plain FP32 rounds the
 * difference of two forward runs away. */
typedef enum psvfwi_precision {
    PSVFWI_FP32 = 0,
    PSVFWI_FP64 = 1,
    PSVFWI_FP32X2 = 2,
    PSVFWI_FP32_FMA = 3
} psvfwi_precision;

/* flags of psvfwi_forward / psvfwi_adjoint */
#define PSVFWI_ENERGY 1   /* accumulate the energy weight We (fd_wave.cuh:14-29) */
#define PSVFWI_GRADIENT 2 /* forward: keep what the adjoint needs; adjoint: correlate */

/* how the adjoint gets the forward wavefield at the correlation steps */
#define PSVFWI_STORE_FULL 0   /* decimated store of every fwi_dt-th step (reference) */
#define PSVFWI_STORE_CHECKPOINT 1 /* full state every k steps, segments recomputed */
/* No store: the adjoint rebuilds the forward state by running the scheme backward from the
 * forward's final state (no counterpart in the reference, which stores).  Needs npml = 0
 * with a random boundary layer in the model (psvfwi_random_boundary) in place of CPML, a
 * positive tile height, and no order-2 free surface.  Device memory: one state and one
 * store plane per shot.  The gradient equals the stored one up to the rounding of the
 * reverse steps. */
#define PSVFWI_STORE_RANDOM_BOUNDARY 2

typedef struct psvfwi_grid_desc {
    int nzt, nxt;       /* padded grid (FDGrid::total_grid) */
    int fdorder;        /* 2 (reference) or 4 */
    int npml;           /* CPML width; 0 disables the strips */
    int fsurf;          /* free surface on the top row iz == fpad */
    double dt, dx, dz;
    double hc[7];       /* Holberg coefficients (holbergcoeff) */
} psvfwi_grid_desc;

/* FWI window and decimation (FDFWI, CUDA_week2/fd_fwi.cu:16-61), padded
 * indices, half-open [x1, x2) x [z1, z2), strides dx, dz, time stride dt. */
typedef struct psvfwi_window_desc {
    int dt, dx, dz, x1, x2, z1, z2;
} psvfwi_window_desc;

/* One shot: its record length, its sources (with signals) and receivers.
 * comp: 1 = vx, 2 = vz (the reference's only type), 4 = explosive
 * (sxx and szz; sources only). */
typedef struct psvfwi_shot_desc {
    int nt;
    int nsrc;
    const int* src_x;
    const int* src_z;
    const int* src_comp;
    const double* src_signal;   /* [nsrc][nt] */
    int nrec;
    const int* rec_x;
    const int* rec_z;
    const int* rec_comp;
} psvfwi_shot_desc;

typedef struct psvfwi_plan psvfwi_plan;

/* message of the last failing call on this thread */
const char* psvfwi_last_error(void);

psvfwi_status psvfwi_plan_create(const psvfwi_grid_desc* grid, psvfwi_precision prec,
                                 psvfwi_plan** out);
psvfwi_status psvfwi_plan_destroy(psvfwi_plan* plan);

/* CPML profiles a, b, K, a_half, b_half, K_half, each 2(npml+1) long
 * (FDCpml::cpml_pro, CUDA_week2/fd_cpml.cu:59-192). */
psvfwi_status psvfwi_set_pml(psvfwi_plan* plan, const double* profiles);

/* lam, mu: [nzt][nxt]; mu_zx, rho_zp, rho_xp: [nzt-1][nxt-1]
 * (FDMedium::average_parameters, CUDA_week2/fd_medium.cu:128-159). */
psvfwi_status psvfwi_set_model(psvfwi_plan* plan, const double* lam, const double* mu,
                               const double* mu_zx, const double* rho_zp, const double* rho_xp);

/* NULL disables gradient work.  The window must lie inside the grid. */
psvfwi_status psvfwi_set_window(psvfwi_plan* plan, const psvfwi_window_desc* win);

/* Store policy for gradients: PSVFWI_STORE_FULL, or PSVFWI_STORE_CHECKPOINT
 * with a checkpoint every `interval` steps. */
psvfwi_status psvfwi_set_store_mode(psvfwi_plan* plan, int mode, int interval);

/* Replaces the batch of shots.  Shot i keeps index i in every getter. */
psvfwi_status psvfwi_set_shots(psvfwi_plan* plan, int nshots, const psvfwi_shot_desc* shots);

/* Step kernel and tile height.  tz > 0: two launches per step (stress, then
 * velocity, in place) on 64 x tz tiles, tz = 4, 8, 16 or 32 (FP64: up to 16);
 * tz = -8 or -16: one fused launch per step on 64 x |tz| tiles (ping-pong
 * state, stress ring recomputed); tz = -32 or -64: one launch per step in
 * which each workgroup marches down a segment of 64 columns x |tz| rows,
 * every stress row computed once (ping-pong state), plus a small launch for
 * the point sources and receivers; 0 = default (8).  Every choice gives the
 * same results bit for bit. */
psvfwi_status psvfwi_set_tile(psvfwi_plan* plan, int tz);

/* Autotuner: times `steps` forward steps of the current batch with every
 * step kernel and tile height of psvfwi_set_tile, keeps the fastest and
 * returns its psvfwi_set_tile value in *best_tile (may be NULL); for a batch it
 * then chooses the shot order (psvfwi_set_shot_order) and, with the split
 * step, the launch group (psvfwi_set_launch_group) the same way.  In the
 * random-boundary store mode only the split kernels (tz > 0) are tried.
 * Needs set_model / set_pml / set_shots first.  (No counterpart in the
 * reference, whose launch shape is fixed at 16 x 16,
 * CUDA_week2/forward_kernel_PSV_GPU.cu:543-545.) */
psvfwi_status psvfwi_autotune(psvfwi_plan* plan, int steps, int* best_tile);

/* Order of the workgroups of a batched launch (no counterpart in the reference, which runs
 * one shot per launch): band 0 = shot-major (every tile of shot 0, then shot 1, ...); band
 * k > 0 = bands of k tile rows, and within a band the shots one after another.  psvfwi_autotune
 * chooses among 0, 1 and 4 for batches of more than one shot.  Results do not depend on it. */
psvfwi_status psvfwi_set_shot_order(psvfwi_plan* plan, int band);
psvfwi_status psvfwi_get_shot_order(const psvfwi_plan* plan, int* band);

/* Launch groups (no counterpart in the reference, which runs one shot per launch): with
 * shots > 0 the time loops of a batch run `shots` slots at a time, all steps of one group
 * before the next (slots are the shots sorted by decreasing nt), so that a group's
 * wavefields stay in the GPU's 256 MB Infinity Cache from one step to the next; 0 (default)
 * steps the whole batch together.  The fused step (tile < 0) always runs the batch as one
 * group.  psvfwi_autotune chooses among 0, 16, 12, 8, 6, 4 and 2 for batches.  Results do not
 * depend on it. */
psvfwi_status psvfwi_set_launch_group(psvfwi_plan* plan, int shots);
psvfwi_status psvfwi_get_launch_group(const psvfwi_plan* plan, int* shots);

/* Graph-captured time loop: with steps > 0 (even), forward and adjoint
 * runs replay HIP graphs of `steps` consecutive step launches (the step index
 * is read from device memory) over every run of steps in which the set of
 * active shots does not change; checkpointed stores and snapshots keep the
 * direct launches.  0 (default) launches every step directly.  Results are
 * identical either way; per-kernel timing samples are taken only from direct
 * launches.  (No counterpart in the reference, which launches two kernels per
 * step and synchronizes after each, CUDA_week2/forward_kernel_PSV_GPU.cu:585-600.) */
psvfwi_status psvfwi_set_graph_steps(psvfwi_plan* plan, int steps);

/* Snapshots of one wavefield (field 0 vx, 1 vz, 2 sxx, 3 szx, 4 szz) kept on
 * the device after every step it with it % interval == 0 of the next forward
 * runs (interval 0 switches them off).  The reference writes szz there
 * (CUDA_week2/forward_kernel_PSV_GPU.cu:1067-1080). */
psvfwi_status psvfwi_set_snapshots(psvfwi_plan* plan, int interval, int field);
psvfwi_status psvfwi_snapshot_count(const psvfwi_plan* plan, int shot, int* count);
psvfwi_status psvfwi_get_snapshot(const psvfwi_plan* plan, int shot, int k, double* out); /* [nzt][nxt] */

/* Runs all shots forward from a zero state: forward_kernel_PSV. */
psvfwi_status psvfwi_forward(psvfwi_plan* plan, int flags);

/* Observed traces kept on the device in FP64: shots concatenated, shot i as
 * [nrec_i][nt_i] (the reference's signal_meas, CUDA_week2/fd_seismogram.cu:127-131). */
psvfwi_status psvfwi_set_observed(psvfwi_plan* plan, const double* observed);

/* Residual (synthetic - observed) becomes the adjoint source of each shot;
 * misfit_out[i] = sum 0.5 dt (d - d_obs)^2 accumulated in FP64 on the GPU.
 * observed: as in psvfwi_set_observed, or NULL to use the traces set there. */
psvfwi_status psvfwi_misfit(psvfwi_plan* plan, const double* observed, double* misfit_out);

/* psvfwi_misfit without device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It reads the observed traces resident from
 * psvfwi_set_observed (an allocating setup call) and computes what psvfwi_misfit(plan, NULL,
 * misfit_out) computes, bit for bit, with its partial sums, per-shot misfits and range flag
 * in a static device array of the library (about 0.5 MB per GPU, held until the process
 * ends) instead of plan buffers.  At most PSVFWI_NOALLOC_MAX_SHOTS shots per batch
 * (PSVFWI_E_ARG beyond).  Every plan on one GPU shares that array, which keeps the last
 * call's contents: calls that overlap on one GPU race. */
#define PSVFWI_NOALLOC_MAX_SHOTS 1024
psvfwi_status psvfwi_misfit_noalloc(psvfwi_plan* plan, double* misfit_out);

/* Explicit adjoint sources for shot i: [nrec_i][nt_i], injected at the shot's
 * receivers (the reference passes the receivers as sources,
 * CUDA_week2/fd_wave_model.cu:848-859). */
psvfwi_status psvfwi_set_adjoint_sources(psvfwi_plan* plan, int shot, const double* signal);

/* Replaces the forward store of one shot ([5][nft][nfz][nfx]: vx/dt, vz/dt,
 * sxx, szx, szz as fwi_vx..fwi_szz), for a caller that stored it itself. */
psvfwi_status psvfwi_set_store(psvfwi_plan* plan, int shot, const double* store);

/* Runs all shots' adjoint from a zero state: adjoint_kernel_PSV. */
psvfwi_status psvfwi_adjoint(psvfwi_plan* plan, int flags);

/* getters (device -> host, converted to FP64) */
psvfwi_status psvfwi_get_records(const psvfwi_plan* plan, int shot, double* out);   /* [nrec][nt] */
/* field: 0 vx, 1 vz, 2 sxx, 3 szx, 4 szz, 5 We, 6 ux, 7 uz -> [nzt][nxt];
 * 8..15 CPML memory mem_vx_x, mem_vz_x, mem_sxx_x, mem_szx_x ([nzt][2(npml+1)]),
 * mem_vx_z, mem_vz_z, mem_szz_z, mem_szx_z ([2(npml+1)][nxt]) */
psvfwi_status psvfwi_get_field(const psvfwi_plan* plan, int shot, int field, double* out);
psvfwi_status psvfwi_get_store(const psvfwi_plan* plan, int shot, double* out); /* [5][nft][nfz][nfx] */
/* shot >= 0: that shot's gradient; shot == -1: sum over the batch (FP64 sum in shot order) */
psvfwi_status psvfwi_get_gradient(const psvfwi_plan* plan, int shot, double* grad_lam,
                                  double* grad_mu, double* grad_rho);   /* [nfz][nfx] each */

/* ---- discrete gradient (no counterpart in the reference) ----
 * The reference's gradient (adjoint_kernel_PSV_GPU.cu:830-863) correlates a
 * time-reversed wavefield with stored forward fields; it is not the
 * derivative of the misfit of the discrete scheme.  In the discrete mode,
 * psvfwi_forward(PSVFWI_GRADIENT) keeps checkpoints of the full state every
 * `segment` steps, and psvfwi_adjoint(PSVFWI_GRADIENT) runs the exact
 * transpose of the scheme (CPML, free surface and mirror included), which
 * gives the gradient of J = sum 0.5 dt (d - d_obs)^2 (the misfit of
 * psvfwi_misfit) with respect to the five coefficient arrays.  The adjoint
 * source of receiver r at step t is dt * residual[r][t] (residuals from
 * psvfwi_misfit, or signals from psvfwi_set_adjoint_sources).  No FWI window
 * is used: the gradients cover the whole grid.  Sources must be vx or vz
 * forces (comp 1 or 2), and the CPML profile must have a = a_half = 0 at index
 * npml (psvfwi_cpml_profiles does).  Device memory: (segment + 1) full states
 * per shot besides the checkpoints.  segment == 1: the forward keeps every
 * state (nt + 1 full states per shot) and the adjoint runs no replay. */
#define PSVFWI_GRADIENT_REFERENCE 0 /* correlation gradient of the reference (default) */
#define PSVFWI_GRADIENT_DISCRETE 1  /* exact gradient of the discrete misfit */
psvfwi_status psvfwi_set_gradient_mode(psvfwi_plan* plan, int mode, int segment);
/* d J / d lam, d J / d mu [nzt][nxt]; d J / d mu_zx, rho_zp, rho_xp [nzt-1][nxt-1];
 * shot == -1: sum over the batch (FP64, shot order) */
psvfwi_status psvfwi_get_coefficient_gradient(const psvfwi_plan* plan, int shot, double* g_lam,
                                              double* g_mu, double* g_mu_zx, double* g_rho_zp,
                                              double* g_rho_xp);
/* d J / d lam, d J / d mu, d J / d rho [nzt][nxt], for averages made by
 * psvfwi_average_parameters from the model given to psvfwi_set_model */
psvfwi_status psvfwi_get_model_gradient(const psvfwi_plan* plan, int shot, double* g_lam,
                                        double* g_mu, double* g_rho);
/* The chain rule of psvfwi_get_model_gradient on the host: out_* [nzt][nxt]
 * from the coefficient gradients and mu, mu_zx, rho_zp, rho_xp. */
psvfwi_status psvfwi_chain_gradient(int nzt, int nxt, const double* mu, const double* mu_zx,
                                    const double* rho_zp, const double* rho_xp, const double* g_lam,
                                    const double* g_mu, const double* g_mu_zx, const double* g_rho_zp,
                                    const double* g_rho_xp, double* out_lam, double* out_mu,
                                    double* out_rho);

/* Timing of the last forward/adjoint call (milliseconds, HIP events) and the
 * number of step-kernel launches it made. */
psvfwi_status psvfwi_last_timing(const psvfwi_plan* plan, double* ms, long* launches);

/* Per-kernel durations sampled (HIP events on the plan's stream, every 64th
 * step) during the last forward / adjoint call: kind 0 forward steps, 1
 * adjoint steps, 2 checkpoint replay steps; mean microseconds of the stress
 * kernel and of the velocity kernel, and the number of samples.  (A
 * measurement hook with no counterpart in the reference, which only prints
 * clock() totals, CUDA_week2/fd_wave_model.cu:593-595.) */
psvfwi_status psvfwi_kernel_timing(const psvfwi_plan* plan, int kind, double* us_stress,
                                   double* us_velocity, int* samples);

/* ---- host-side setup (FP64): the inputs the kernels read ---- */

/* hc[0..6] for FD order 2..12 and error class 0..4: holbergcoeff
 * (CUDA_week2/holbergcoeff.cu:6-65) */
psvfwi_status psvfwi_holberg(int fdorder, int max_relative_error, double* hc);

/* a, b, K, a_half, b_half, K_half into profiles[6 * 2(npml+1)]:
 * FDCpml::cpml_pro (CUDA_week2/fd_cpml.cu:59-192) */
psvfwi_status psvfwi_cpml_profiles(int npml, double dh, double fpml, double npower, double k_max,
                                   double rcoef, double damp_v, double dt, double* profiles);

/* mu_zx, rho_zp, rho_xp [nzt-1][nxt-1] from mu, rho [nzt][nxt]:
 * FDMedium::average_parameters (CUDA_week2/fd_medium.cu:128-159) */
psvfwi_status psvfwi_average_parameters(int nzt, int nxt, const double* mu, const double* rho,
                                        double* mu_zx, double* rho_zp, double* rho_xp);

/* Random boundary layer for PSVFWI_STORE_RANDOM_BOUNDARY (no counterpart in the reference):
 * in place on lam, mu, rho [nzt][nxt], the cells within `width` of the left, right and bottom
 * edges (and the top one when fsurf == 0) get vp and vs scaled by 1 + max_increase * (d /
 * width) * u, d the depth into the layer and u uniform in [0, 1) from a hash of (seed, iz,
 * ix); lam and mu scale by the square, rho is kept.  The waves that reach the layer scatter
 * incoherently instead of being absorbed.  The time step must stay stable for the largest
 * velocity, (1 + max_increase) times the model's. */
psvfwi_status psvfwi_random_boundary(int nzt, int nxt, int width, int fsurf, unsigned long long seed,
                                     double max_increase, double* lam, double* mu, double* rho);

/* Ricker wavelet: create_wavelet shape 1 (CUDA_week2/fd_seismogram.cu:166-186) */
psvfwi_status psvfwi_ricker(int nt, double dt, double amp, double fc, double ts, double* signal);

/* Vertex of the parabola through (step[i], L2[i]), i = 0..2, in FP64:
 * calc_opt_step (CUDA_week2/fwi_step_length_estimation.cu:226-271). */
psvfwi_status psvfwi_parabola_step(const double* L2, const double* step, double* opt);

#ifdef __cplusplus
}
#endif
#endif
