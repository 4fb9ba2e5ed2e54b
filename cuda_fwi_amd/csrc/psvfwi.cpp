// psvfwi.cpp -- the C ABI (psvfwi.h) over the plans of plan.h: argument checks,
// plan creation per arithmetic type, the per-thread last error, ROCTX ranges
// around the runs, and the host-side helpers (chain rule of the coefficient
// gradient, parabola step).
#include <memory>
#include <string>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "plan.h"

namespace psvplan {

namespace {
thread_local std::string g_err;
}  // namespace

psvfwi_status fail(psvfwi_status st, const std::string& msg)
{
    g_err = msg;
    return st;
}

}  // namespace psvplan

namespace {

using psvplan::fail;
using psvplan::Plan;

// A ROCTX range around a C ABI call (forward, misfit, adjoint, autotune), seen by
// rocprofv3 --marker-trace beside the kernels it launched
struct Range {
    explicit Range(const char* what) { roctxRangePush(what); }
    ~Range() { roctxRangePop(); }
    Range(const Range&) = delete;
    Range& operator=(const Range&) = delete;
};

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char* psvfwi_last_error(void) { return psvplan::g_err.c_str(); }

psvfwi_status psvfwi_plan_create(const psvfwi_grid_desc* g, psvfwi_precision prec, psvfwi_plan** out)
{
    if (!g || !out) return fail(PSVFWI_E_ARG, "plan_create: null argument");
    *out = nullptr;
    if (g->fdorder != 2 && g->fdorder != 4)
        return fail(PSVFWI_E_ORDER, "FD order mismatch: fdorder must be 2 or 4");
    const int fpad = g->fdorder / 2, ppad = g->npml + fpad;
    if (g->npml < 0 || g->nzt < 2 * fpad + 1 || g->nxt < 2 * fpad + 1)
        return fail(PSVFWI_E_ARG, "plan_create: grid smaller than the stencil");
    if (g->npml > 0 && (g->nxt < 2 * ppad || g->nzt < (g->fsurf ? ppad + fpad : 2 * ppad)))
        return fail(PSVFWI_E_ARG, "plan_create: CPML wider than the grid");
    if (!(g->dt > 0) || !(g->dx > 0) || !(g->dz > 0)) return fail(PSVFWI_E_ARG, "plan_create: dt, dx, dz must be > 0");
    if (prec != PSVFWI_FP32 && prec != PSVFWI_FP64 && prec != PSVFWI_FP32X2 && prec != PSVFWI_FP32_FMA)
        return fail(PSVFWI_E_ARG, "plan_create: unknown precision");
    {
        // the kernels address a shot's full-grid arrays with 32-bit byte offsets
        const long long tiles_x = (g->nxt + psv::kTX - 1) / psv::kTX;
        const long long pitch = (psv::kGuardLeft + tiles_x * psv::kTX + psv::kGuardRight + 31) / 32 * 32;
        const long long rows = 2 * psv::kGuardRows + ((long long)g->nzt + psv::kMaxTz - 1) / psv::kMaxTz * psv::kMaxTz;
        const long long bytes = rows * pitch * (prec == PSVFWI_FP32 || prec == PSVFWI_FP32_FMA ? 4 : 8);
        if (bytes >= (1LL << 31)) return fail(PSVFWI_E_ARG, "plan_create: grid too large (a field of one shot must stay below 2 GiB)");
    }
    std::unique_ptr<psvfwi_plan> p;
    if (prec == PSVFWI_FP32) {
        auto q = std::make_unique<Plan<float>>(*g);
        PSV_TRY(q->init());
        p = std::move(q);
    } else if (prec == PSVFWI_FP32_FMA) {
        auto q = std::make_unique<Plan<psv::ffloat>>(*g);
        PSV_TRY(q->init());
        p = std::move(q);
    } else if (prec == PSVFWI_FP32X2) {
        auto q = std::make_unique<Plan<psv::dfloat>>(*g);
        PSV_TRY(q->init());
        p = std::move(q);
    } else {
        auto q = std::make_unique<Plan<double>>(*g);
        PSV_TRY(q->init());
        p = std::move(q);
    }
    *out = p.release();
    return PSVFWI_OK;
}

psvfwi_status psvfwi_plan_destroy(psvfwi_plan* plan)
{
    delete plan;
    return PSVFWI_OK;
}

#define PSV_PLAN(plan) \
    if (!(plan)) return fail(PSVFWI_E_ARG, "null plan")

psvfwi_status psvfwi_set_pml(psvfwi_plan* plan, const double* profiles)
{
    PSV_PLAN(plan);
    return plan->set_pml(profiles);
}

psvfwi_status psvfwi_set_model(psvfwi_plan* plan, const double* lam, const double* mu,
                               const double* mu_zx, const double* rho_zp, const double* rho_xp)
{
    PSV_PLAN(plan);
    return plan->set_model(lam, mu, mu_zx, rho_zp, rho_xp);
}

psvfwi_status psvfwi_set_window(psvfwi_plan* plan, const psvfwi_window_desc* win)
{
    PSV_PLAN(plan);
    return plan->set_window(win);
}

psvfwi_status psvfwi_set_store_mode(psvfwi_plan* plan, int mode, int interval)
{
    PSV_PLAN(plan);
    return plan->set_store_mode(mode, interval);
}

psvfwi_status psvfwi_set_shots(psvfwi_plan* plan, int nshots, const psvfwi_shot_desc* shots)
{
    PSV_PLAN(plan);
    return plan->set_shots(nshots, shots);
}

psvfwi_status psvfwi_set_tile(psvfwi_plan* plan, int tz)
{
    PSV_PLAN(plan);
    return plan->set_tile(tz);
}

psvfwi_status psvfwi_autotune(psvfwi_plan* plan, int steps, int* best_tile)
{
    PSV_PLAN(plan);
    const Range r("psvfwi_autotune");
    return plan->autotune(steps, best_tile);
}

psvfwi_status psvfwi_set_graph_steps(psvfwi_plan* plan, int steps)
{
    if (!plan) return fail(PSVFWI_E_ARG, "null plan");
    return plan->set_graph_steps(steps);
}

psvfwi_status psvfwi_set_shot_order(psvfwi_plan* plan, int band)
{
    PSV_PLAN(plan);
    return plan->set_shot_order(band);
}

psvfwi_status psvfwi_get_shot_order(const psvfwi_plan* plan, int* band)
{
    PSV_PLAN(plan);
    return plan->shot_order(band);
}

psvfwi_status psvfwi_set_launch_group(psvfwi_plan* plan, int shots)
{
    PSV_PLAN(plan);
    return plan->set_launch_group(shots);
}

psvfwi_status psvfwi_get_launch_group(const psvfwi_plan* plan, int* shots)
{
    PSV_PLAN(plan);
    return plan->launch_group(shots);
}

psvfwi_status psvfwi_set_snapshots(psvfwi_plan* plan, int interval, int field)
{
    PSV_PLAN(plan);
    return plan->set_snapshots(interval, field);
}

psvfwi_status psvfwi_snapshot_count(const psvfwi_plan* plan, int shot, int* count)
{
    PSV_PLAN(plan);
    return plan->snapshot_count(shot, count);
}

psvfwi_status psvfwi_get_snapshot(const psvfwi_plan* plan, int shot, int k, double* out)
{
    PSV_PLAN(plan);
    return plan->get_snapshot(shot, k, out);
}

psvfwi_status psvfwi_forward(psvfwi_plan* plan, int flags)
{
    PSV_PLAN(plan);
    const Range r("psvfwi_forward");
    return plan->forward(flags);
}

psvfwi_status psvfwi_set_observed(psvfwi_plan* plan, const double* observed)
{
    PSV_PLAN(plan);
    return plan->set_observed(observed);
}

psvfwi_status psvfwi_misfit(psvfwi_plan* plan, const double* observed, double* misfit_out)
{
    PSV_PLAN(plan);
    const Range r("psvfwi_misfit");
    return plan->misfit(observed, misfit_out);
}

psvfwi_status psvfwi_misfit_noalloc(psvfwi_plan* plan, double* misfit_out)
{
    PSV_PLAN(plan);
    const Range r("psvfwi_misfit_noalloc");
    return plan->misfit_noalloc(misfit_out);
}

psvfwi_status psvfwi_set_adjoint_sources(psvfwi_plan* plan, int shot, const double* signal)
{
    PSV_PLAN(plan);
    return plan->set_adjoint_sources(shot, signal);
}

psvfwi_status psvfwi_set_store(psvfwi_plan* plan, int shot, const double* store)
{
    PSV_PLAN(plan);
    return plan->set_store(shot, store);
}

psvfwi_status psvfwi_adjoint(psvfwi_plan* plan, int flags)
{
    PSV_PLAN(plan);
    const Range r((flags & PSVFWI_GRADIENT) ? "psvfwi_adjoint_gradient" : "psvfwi_adjoint");
    return plan->adjoint(flags);
}

psvfwi_status psvfwi_get_records(const psvfwi_plan* plan, int shot, double* out)
{
    PSV_PLAN(plan);
    return plan->get_records(shot, out);
}

psvfwi_status psvfwi_get_field(const psvfwi_plan* plan, int shot, int field, double* out)
{
    PSV_PLAN(plan);
    return plan->get_field(shot, field, out);
}

psvfwi_status psvfwi_get_store(const psvfwi_plan* plan, int shot, double* out)
{
    PSV_PLAN(plan);
    return plan->get_store(shot, out);
}

psvfwi_status psvfwi_get_gradient(const psvfwi_plan* plan, int shot, double* grad_lam,
                                  double* grad_mu, double* grad_rho)
{
    PSV_PLAN(plan);
    return plan->get_gradient(shot, grad_lam, grad_mu, grad_rho);
}

psvfwi_status psvfwi_set_gradient_mode(psvfwi_plan* plan, int mode, int segment)
{
    PSV_PLAN(plan);
    return plan->set_gradient_mode(mode, segment);
}

psvfwi_status psvfwi_get_coefficient_gradient(const psvfwi_plan* plan, int shot, double* g_lam, double* g_mu,
                                              double* g_mu_zx, double* g_rho_zp, double* g_rho_xp)
{
    PSV_PLAN(plan);
    return plan->get_coefficient_gradient(shot, g_lam, g_mu, g_mu_zx, g_rho_zp, g_rho_xp);
}

psvfwi_status psvfwi_get_model_gradient(const psvfwi_plan* plan, int shot, double* g_lam, double* g_mu,
                                        double* g_rho)
{
    PSV_PLAN(plan);
    return plan->get_model_gradient(shot, g_lam, g_mu, g_rho);
}

psvfwi_status psvfwi_chain_gradient(int nzt, int nxt, const double* mu, const double* mu_zx,
                                    const double* rho_zp, const double* rho_xp, const double* g_lam,
                                    const double* g_mu, const double* g_mu_zx, const double* g_rho_zp,
                                    const double* g_rho_xp, double* out_lam, double* out_mu, double* out_rho)
{
    if (nzt < 2 || nxt < 2) return fail(PSVFWI_E_ARG, "chain_gradient: grid smaller than 2 x 2");
    if (!mu || !mu_zx || !rho_zp || !rho_xp || !g_lam || !g_mu || !g_mu_zx || !g_rho_zp || !g_rho_xp ||
        !out_lam || !out_mu || !out_rho)
        return fail(PSVFWI_E_ARG, "chain_gradient: null array");
    // mu_zx = 4 / sum(1 / mu) over the 4 corners, rho_zp = 2 / (rho(iz) + rho(iz + 1)),
    // rho_xp = 2 / (rho(ix) + rho(ix + 1)) (average_parameters, CUDA_week2/fd_medium.cu:128-159)
    const int ax = nxt - 1;
    for (int iz = 0; iz < nzt; iz++) {
        for (int ix = 0; ix < nxt; ix++) {
            const size_t q = (size_t)iz * nxt + ix;
            double gm = g_mu[q], gr = 0.0;
            // averaged cells with this cell as corner: (iz, ix), (iz, ix-1), (iz-1, ix), (iz-1, ix-1)
            for (int dz = 0; dz <= 1; dz++) {
                for (int dx = 0; dx <= 1; dx++) {
                    const int oz = iz - dz, ox = ix - dx;
                    if (oz < 0 || ox < 0 || oz >= nzt - 1 || ox >= ax) continue;
                    const size_t o = (size_t)oz * ax + ox;
                    const double m = mu_zx[o];
                    if (m != 0.0) gm += g_mu_zx[o] * (m * m / (4.0 * mu[q] * mu[q]));
                    if (dx == 0) gr += g_rho_zp[o] * (-0.5 * rho_zp[o] * rho_zp[o]);
                    if (dz == 0) gr += g_rho_xp[o] * (-0.5 * rho_xp[o] * rho_xp[o]);
                }
            }
            out_lam[q] = g_lam[q];
            out_mu[q] = gm;
            out_rho[q] = gr;
        }
    }
    return PSVFWI_OK;
}

psvfwi_status psvfwi_last_timing(const psvfwi_plan* plan, double* ms, long* launches)
{
    PSV_PLAN(plan);
    return plan->timing(ms, launches);
}

psvfwi_status psvfwi_kernel_timing(const psvfwi_plan* plan, int kind, double* us_stress,
                                   double* us_velocity, int* samples)
{
    PSV_PLAN(plan);
    return plan->kernel_timing(kind, us_stress, us_velocity, samples);
}

psvfwi_status psvfwi_parabola_step(const double* L2, const double* step, double* opt)
{
    if (!L2 || !step || !opt) return fail(PSVFWI_E_ARG, "parabola_step: null argument");
    // Newton divided differences of the interpolating parabola, FP64
    const double d01 = (L2[1] - L2[0]) / (step[1] - step[0]);
    const double d02 = (L2[2] - L2[0]) / (step[2] - step[0]);
    const double x2 = (d02 - d01) / (step[2] - step[1]);
    const double x1 = d01 - x2 * (step[0] + step[1]);
    *opt = -x1 / (2.0 * x2);
    if (!std::isfinite(*opt)) return fail(PSVFWI_E_NONFINITE, "parabola_step: degenerate parabola");
    return PSVFWI_OK;
}

}  // extern "C"
