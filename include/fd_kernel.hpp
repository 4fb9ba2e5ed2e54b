// fd_kernel.hpp -- drop-in replacement of CUDA_week2/fd_kernel.cuh.
//
// Same four propagation entry points and calc_opt_step, with the reference's
// parameter lists (FP64 host arrays as row-pointer tables over contiguous
// storage, geometry as ivec by value).  They run on the MI355X through the
// C ABI of psvfwi.h: the host arrays are uploaded, the time loop runs on the
// GPU, and every output the reference writes in place is written back:
//   forward:  vx, vz, sxx, szx, szz, We, the 8 CPML memory arrays,
//             rec_signal rows of this shot's receivers, fwi_* (when fwinv),
//             ./io/snap_data/vz_snap<k>.csv snapshots of szz (when !fwinv)
//   adjoint:  vx, vz, sxx, szx, szz, We, the CPML memory arrays, grad_*
// The *_GPU variants take the reference's device pointers and ignore them:
// the device buffers belong to the library (a plan cached per grid).
// Arithmetic type: FP32 unless fd_kernel_set_precision(1) or the environment
// variable PSVFWI_PRECISION=fp64 selects the FP64 build of the same kernels.
#pragma once

#include <vector>

using real_sim = double;
using ivec = std::vector<int>;

// CUDA_week2/fd_kernel.cuh:12-40
void forward_kernel_PSV(int ishot, int nt, int nzt, int nxt, int fpad, int ppad, real_sim dt,
                        real_sim dx, real_sim dz, int snap_interval, bool fsurf, real_sim* hc,
                        int fdorder, real_sim** vx, real_sim** vz, real_sim** sxx, real_sim** szx,
                        real_sim** szz, real_sim** We, real_sim** lam, real_sim** mu,
                        real_sim** mu_zx, real_sim** rho_zp, real_sim** rho_xp, int npml,
                        real_sim* a, real_sim* b, real_sim* K, real_sim* a_half, real_sim* b_half,
                        real_sim* K_half, real_sim** mem_vx_x, real_sim** mem_vx_z,
                        real_sim** mem_vz_x, real_sim** mem_vz_z, real_sim** mem_sxx_x,
                        real_sim** mem_szx_x, real_sim** mem_szz_z, real_sim** mem_szx_z,
                        int nsrc, ivec src_x, ivec src_z, ivec src_comp, real_sim** src_signal,
                        ivec source_to_fire_shot, int nrec, ivec rec_x, ivec rec_z,
                        ivec rec_comp, real_sim** rec_signal, ivec receiver_to_record_shot,
                        bool fwinv, int fwi_dt, int fwi_dx, int fwi_dz, int fwi_x1, int fwi_x2,
                        int fwi_z1, int fwi_z2, real_sim*** fwi_vx, real_sim*** fwi_vz,
                        real_sim*** fwi_sxx, real_sim*** fwi_szx, real_sim*** fwi_szz);

// CUDA_week2/fd_kernel.cuh:43-70
void adjoint_kernel_PSV(int ishot, int nt, int nzt, int nxt, int fpad, int ppad, real_sim dt,
                        real_sim dx, real_sim dz, int snap_interval, bool fsurf, real_sim* hc,
                        int fdorder, real_sim** vx, real_sim** vz, real_sim** sxx, real_sim** szx,
                        real_sim** szz, real_sim** We, real_sim** lam, real_sim** mu,
                        real_sim** mu_zx, real_sim** rho_zp, real_sim** rho_xp, int npml,
                        real_sim* a, real_sim* b, real_sim* K, real_sim* a_half, real_sim* b_half,
                        real_sim* K_half, real_sim** mem_vx_x, real_sim** mem_vx_z,
                        real_sim** mem_vz_x, real_sim** mem_vz_z, real_sim** mem_sxx_x,
                        real_sim** mem_szx_x, real_sim** mem_szz_z, real_sim** mem_szx_z,
                        int nsrc, ivec src_x, ivec src_z, ivec src_comp, real_sim** src_signal,
                        ivec source_to_fire_shot, bool fwinv, int fwi_dt, int fwi_dx, int fwi_dz,
                        int fwi_x1, int fwi_x2, int fwi_z1, int fwi_z2, real_sim*** fwi_vx,
                        real_sim*** fwi_vz, real_sim*** fwi_sxx, real_sim*** fwi_szx,
                        real_sim*** fwi_szz, real_sim** grad_lam, real_sim** grad_mu,
                        real_sim** grad_rho);

// CUDA_week2/fd_kernel.cuh:72
real_sim calc_opt_step(real_sim* L2_test, real_sim* step_length);

// CUDA_week2/fd_kernel.cuh:75-120 (device pointers accepted and unused)
void forward_kernel_PSV_GPU(int ishot, int nt, int nzt, int nxt, int fpad, int ppad, real_sim dt,
                            real_sim dx, real_sim dz, int snap_interval, bool fsurf, real_sim* hc,
                            int fdorder, real_sim** vx, real_sim** vz, real_sim** sxx,
                            real_sim** szx, real_sim** szz, real_sim** lam, real_sim** mu,
                            real_sim** mu_zx, real_sim** rho_zp, real_sim** rho_xp, int npml,
                            real_sim* a, real_sim* b, real_sim* K, real_sim* a_half,
                            real_sim* b_half, real_sim* K_half, real_sim** mem_vx_x,
                            real_sim** mem_vx_z, real_sim** mem_vz_x, real_sim** mem_vz_z,
                            real_sim** mem_sxx_x, real_sim** mem_szx_x, real_sim** mem_szz_z,
                            real_sim** mem_szx_z, int nsrc, ivec src_x, ivec src_z, ivec src_comp,
                            real_sim** src_signal, ivec source_to_fire_shot, int nrec, ivec rec_x,
                            ivec rec_z, ivec rec_comp, real_sim** rec_signal,
                            ivec receiver_to_record_shot, bool fwinv, int fwi_dt, int fwi_dx,
                            int fwi_dz, int fwi_x1, int fwi_x2, int fwi_z1, int fwi_z2,
                            real_sim*** fwi_vx, real_sim*** fwi_vz, real_sim*** fwi_sxx,
                            real_sim*** fwi_szx, real_sim*** fwi_szz, real_sim* d_a, real_sim* d_b,
                            real_sim* d_K, real_sim* d_a_half, real_sim* d_b_half,
                            real_sim* d_K_half, real_sim* d_vx, real_sim* d_vz, real_sim* d_sxx,
                            real_sim* d_szx, real_sim* d_szz, real_sim* d_We, real_sim* d_fwi_vx,
                            real_sim* d_fwi_vz, real_sim* d_fwi_sxx, real_sim* d_fwi_szx,
                            real_sim* d_fwi_szz, real_sim* d_mem_vx_x, real_sim* d_mem_vx_z,
                            real_sim* d_mem_vz_x, real_sim* d_mem_vz_z, real_sim* d_mem_sxx_x,
                            real_sim* d_mem_szx_x, real_sim* d_mem_szz_z, real_sim* d_mem_szx_z,
                            real_sim* d_lam, real_sim* d_mu, real_sim* d_mu_zx,
                            real_sim* d_rho_zp, real_sim* d_rho_xp);

// CUDA_week2/fd_kernel.cuh:123-173 (device pointers accepted and unused)
void adjoint_kernel_PSV_GPU(int ishot, int nt, int nzt, int nxt, int fpad, int ppad, real_sim dt,
                            real_sim dx, real_sim dz, int snap_interval, bool fsurf, real_sim* hc,
                            int fdorder, real_sim** vx, real_sim** vz, real_sim** sxx,
                            real_sim** szx, real_sim** szz, real_sim** lam, real_sim** mu,
                            real_sim** mu_zx, real_sim** rho_zp, real_sim** rho_xp, int npml,
                            real_sim* a, real_sim* b, real_sim* K, real_sim* a_half,
                            real_sim* b_half, real_sim* K_half, real_sim** mem_vx_x,
                            real_sim** mem_vx_z, real_sim** mem_vz_x, real_sim** mem_vz_z,
                            real_sim** mem_sxx_x, real_sim** mem_szx_x, real_sim** mem_szz_z,
                            real_sim** mem_szx_z, int nsrc, ivec src_x, ivec src_z, ivec src_comp,
                            real_sim** src_signal, ivec source_to_fire_shot, bool fwinv, int fwi_dt,
                            int fwi_dx, int fwi_dz, int fwi_x1, int fwi_x2, int fwi_z1, int fwi_z2,
                            real_sim*** fwi_vx, real_sim*** fwi_vz, real_sim*** fwi_sxx,
                            real_sim*** fwi_szx, real_sim*** fwi_szz, real_sim** grad_lam,
                            real_sim** grad_mu, real_sim** grad_rho, real_sim* d_a, real_sim* d_b,
                            real_sim* d_K, real_sim* d_a_half, real_sim* d_b_half,
                            real_sim* d_K_half, real_sim* d_vx, real_sim* d_vz, real_sim* d_sxx,
                            real_sim* d_szx, real_sim* d_szz, real_sim* d_We, real_sim* d_fwi_vx,
                            real_sim* d_fwi_vz, real_sim* d_fwi_sxx, real_sim* d_fwi_szx,
                            real_sim* d_fwi_szz, real_sim* d_mem_vx_x, real_sim* d_mem_vx_z,
                            real_sim* d_mem_vz_x, real_sim* d_mem_vz_z, real_sim* d_mem_sxx_x,
                            real_sim* d_mem_szx_x, real_sim* d_mem_szz_z, real_sim* d_mem_szx_z,
                            real_sim* d_grad_lam, real_sim* d_grad_mu, real_sim* d_grad_rho,
                            real_sim* d_lam, real_sim* d_mu, real_sim* d_mu_zx,
                            real_sim* d_rho_zp, real_sim* d_rho_xp);

// 0 = FP32 kernels (default), 1 = FP64 build of the same kernels
void fd_kernel_set_precision(int fp64);
// directory snapshots go to (default "./io/snap_data"); empty disables them
void fd_kernel_set_snapshot_dir(const char* dir);
