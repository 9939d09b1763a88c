// solver.hip - the C ABI of the solver (include/thermal_hip.h) over
// th::Solver (solver.h).
#include "solver.h"

#include <mutex>

// Scratch of th_solver_colorize_device_noalloc: the die map's per-block
// partial keys and the final key, in a file-scope device array instead of
// memory allocated at the first call. Every GPU the library runs on holds its
// own copy (136 KB) until the process ends. Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the
// symbol): each device's address comes from hipGetSymbolAddress, taken once
// with that device current.
namespace {
constexpr long kNoAllocBlocks = 16384;  // 64 x 4 columns each: grids of up to 2048 x 2048 columns
constexpr int kNoAllocDevices = 64;
__device__ __align__(256) unsigned long long scratch[kNoAllocBlocks + 1];
std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kNoAllocDevices] = {};

// The device address of `scratch` on `dev` (current), looked up on the first
// call for that device.
int scratch_on(int dev, unsigned long long** out) {
    if (dev < 0 || dev >= kNoAllocDevices) return th::fail(TH_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return th::fail(TH_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned long long*)p;
    }
    *out = g_scratch_ptr[dev];
    return TH_OK;
}
}  // namespace

// ===========================================================================
// C ABI: solver
// ===========================================================================
using namespace th;

struct th_solver {
    SolverBase* impl;
};

extern "C" {

void th_solver_default_opts(th_solver_opts* o) {
    o->precond = TH_PRECOND_MG;
    o->precision = TH_PREC_MIXED;
    o->beta = TH_BETA_FLEXIBLE;
    o->use_graphs = 1;
    o->device = 0;
    o->true_residual = 0;
    o->restriction = TH_RESTRICT_LATERAL;
}

static int check_desc(const th_model_desc* m) {
    if (!m) return fail(TH_ERR_ARG, "null model descriptor");
    if (m->nx < 1 || m->ny < 1 || m->nz < 1 || m->nz > TH_MAXZ)
        return fail(TH_ERR_ARG, "model dimensions out of range (1 <= nz <= 64)");
    if (!m->gE || !m->gN || !m->gU || !m->gB || !m->cap || !m->P)
        return fail(TH_ERR_NOT_FINALIZED, "Model3D used before finalize()");
    if (!(m->dx > 0) || !(m->dy > 0)) return fail(TH_ERR_ARG, "cell size must be positive");
    return TH_OK;
}

int th_solver_create(const th_model_desc* m, const th_solver_opts* o, th_solver** out) {
    clear_error();
    int rc = check_desc(m);
    if (rc) return rc;
    if (!out) return fail(TH_ERR_ARG, "null out");
    th_solver_opts opt;
    if (o) opt = *o;
    else th_solver_default_opts(&opt);
    if (opt.precond != TH_PRECOND_MG && opt.precond != TH_PRECOND_JACOBI) return fail(TH_ERR_ARG, "unknown preconditioner");
    SolverBase* impl = nullptr;
    if (opt.precision == TH_PREC_FP64) impl = new Solver<double>();
    else if (opt.precision == TH_PREC_MIXED) impl = new Solver<float>();
    else return fail(TH_ERR_ARG, "unknown precision mode");
    rc = impl->init(*m, opt);
    if (rc) {
        delete impl;
        return rc;
    }
    // flexible flag into the device state
    if (opt.precision == TH_PREC_FP64) {
        Solver<double>* sv = (Solver<double>*)impl;
        k_set_flexible<<<1, 1, 0, sv->s>>>(sv->st, opt.beta == TH_BETA_FLEXIBLE);
    } else {
        Solver<float>* sv = (Solver<float>*)impl;
        k_set_flexible<<<1, 1, 0, sv->s>>>(sv->st, opt.beta == TH_BETA_FLEXIBLE);
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(impl->stream()) != hipSuccess) {
        delete impl;
        return fail(TH_ERR_HIP, "solver setup failed");
    }
    *out = new th_solver{impl};
    return TH_OK;
}

void th_solver_destroy(th_solver* s) {
    if (!s) return;
    delete s->impl;
    delete s;
}

#define TH_S(s) \
    if (!(s)) return fail(TH_ERR_ARG, "null solver"); \
    clear_error();

int th_solver_set_dt(th_solver* s, double dt) { TH_S(s); return s->impl->set_dt(dt); }
int th_solver_set_x_zero(th_solver* s) { TH_S(s); return s->impl->set_x_zero(); }
int th_solver_set_x(th_solver* s, const double* x) {
    TH_S(s);
    if (!x) return fail(TH_ERR_ARG, "null x");
    return s->impl->set_x(x);
}
int th_solver_set_rhs_steady(th_solver* s) { TH_S(s); return s->impl->set_rhs_steady(); }
int th_solver_set_rhs_transient(th_solver* s) { TH_S(s); return s->impl->set_rhs_transient(); }
int th_solver_solve(th_solver* s, double tol, int maxit, th_stats* st) {
    TH_S(s);
    if (maxit < 0) return fail(TH_ERR_ARG, "maxit must be >= 0");
    return s->impl->solve(tol, maxit, st);
}
int th_solver_get_x(th_solver* s, double* x) {
    TH_S(s);
    if (!x) return fail(TH_ERR_ARG, "null x");
    return s->impl->get_x(x);
}
int th_solver_set_power(th_solver* s, const double* P) {
    TH_S(s);
    if (!P) return fail(TH_ERR_ARG, "null P");
    return s->impl->set_power(P);
}
int th_solver_device_x(th_solver* s, const double** d) {
    TH_S(s);
    if (!d) return fail(TH_ERR_ARG, "null out");
    return s->impl->device_x(d);
}
int th_solver_true_relres(th_solver* s, double* out) {
    TH_S(s);
    if (!out) return fail(TH_ERR_ARG, "null out");
    return s->impl->true_relres(out);
}
int th_solver_heat_out(th_solver* s, double* out) {
    TH_S(s);
    if (!out) return fail(TH_ERR_ARG, "null out");
    return s->impl->energy(nullptr, out, nullptr);
}
int th_solver_nlevels(th_solver* s) { return s ? s->impl->nlevels() : -1; }
int th_solver_column_types(th_solver* s) { return s ? s->impl->column_types() : -1; }
int th_solver_setup_times(th_solver* s, th_setup_times* out) {
    TH_S(s);
    if (!out) return fail(TH_ERR_ARG, "null argument");
    return s->impl->setup_times(out);
}
int th_solver_level_info(th_solver* s, int k, int* nx, int* ny, int* nz, int* ntypes, int* kind) {
    TH_S(s);
    return s->impl->level_info(k, nx, ny, nz, ntypes, kind);
}
int th_solver_colorize(th_solver* s, int l0, int l1, float rise_hi, unsigned char* rgba, float* hot_rise,
                       int* hot_cell) {
    TH_S(s);
    return s->impl->colorize(l0, l1, rise_hi, rgba, false, hot_rise, hot_cell);
}
int th_solver_colorize_device(th_solver* s, int l0, int l1, float rise_hi, unsigned char* d_rgba, float* hot_rise,
                              int* hot_cell) {
    TH_S(s);
    if (!d_rgba) return fail(TH_ERR_ARG, "null device image");
    return s->impl->colorize(l0, l1, rise_hi, d_rgba, true, hot_rise, hot_cell);
}
int th_solver_colorize_device_noalloc(th_solver* s, int l0, int l1, float rise_hi, unsigned char* d_rgba,
                                      float* hot_rise, int* hot_cell) {
    TH_S(s);
    int rc = s->impl->check_colorize_args(l0, l1, rise_hi);
    if (rc) return rc;
    const long nb = s->impl->colorize_blocks();
    if (nb > kNoAllocBlocks)
        return fail(TH_ERR_ARG, "colorize_device_noalloc: grid of " + std::to_string(nb) +
                                    " blocks of 64 x 4 columns exceeds the static scratch (" +
                                    std::to_string(kNoAllocBlocks) + " blocks: at most 2048 x 2048 columns)");
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess) return fail(TH_ERR_HIP, "hipGetDevice failed");
    const int dev = s->impl->device();
    if (dev != prev && hipSetDevice(dev) != hipSuccess) return fail(TH_ERR_HIP, "hipSetDevice failed");
    unsigned long long* sc = nullptr;
    rc = scratch_on(dev, &sc);
    // Every partial slot is written before k_max_key reads it, and the key
    // slot is written by k_max_key: nothing in the scratch has to start at 0.
    if (!rc) rc = s->impl->colorize_into(l0, l1, rise_hi, (unsigned*)d_rgba, sc, sc + kNoAllocBlocks, nullptr,
                                         hot_rise, hot_cell);
    if (dev != prev) (void)hipSetDevice(prev);
    return rc;
}
int th_solver_precondition(th_solver* s, const double* r, double* z) {
    TH_S(s);
    if (!r || !z) return fail(TH_ERR_ARG, "null argument");
    return s->impl->precondition(r, z);
}
int th_solver_apply_operator(th_solver* s, const double* x, double* y) {
    TH_S(s);
    if (!x || !y) return fail(TH_ERR_ARG, "null argument");
    return s->impl->apply_operator(x, y);
}
int th_solver_time_kernel(th_solver* s, int which, int reps, double* avg_ms, double* bytes) {
    TH_S(s);
    if (!avg_ms || !bytes) return fail(TH_ERR_ARG, "null argument");
    return s->impl->time_kernel(which, reps, avg_ms, bytes);
}

int th_steady_solve(const th_model_desc* m, const th_solver_opts* o, double tol, int maxit, double* T,
                    th_solve_result* res) {
    clear_error();
    using clk = std::chrono::high_resolution_clock;
    auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    const auto t0 = clk::now();
    th_solver* s = nullptr;
    int rc = th_solver_create(m, o, &s);
    if (rc) return rc;
    if ((rc = th_solver_set_x_zero(s)) || (rc = th_solver_set_rhs_steady(s))) {
        th_solver_destroy(s);
        return rc;
    }
    (void)hipStreamSynchronize(s->impl->stream());
    const auto t1 = clk::now();
    th_stats st{};
    rc = th_solver_solve(s, tol, maxit, &st);
    (void)hipStreamSynchronize(s->impl->stream());
    const auto t2 = clk::now();
    if (!rc && T) rc = th_solver_get_x(s, T);
    th_solver_destroy(s);
    const auto t3 = clk::now();
    if (rc) return rc;
    if (res) {
        res->iters = st.iters;
        res->relres = st.relres;
        res->converged = st.converged;
        res->setup_ms = ms(t0, t1);
        res->solve_ms = ms(t1, t2);
        res->total_ms = ms(t0, t3);
    }
    return TH_OK;
}

void th_transient_default_opts(th_transient_opts* o) {
    o->dt0 = 1e-5;
    o->growth = 1.1;
    o->dt_max = 0.05;
    o->t_end = 1.0;
    o->tol = 1e-9;
    o->maxit = 50000;
    o->max_steps = 0;
    o->predictor = 0;
}

// Plane-to-channel map of a lift: plane l at zc_fine[l] between the coarse
// planes at zc_coarse[lo] and zc_coarse[lo + 1] (nearest outside them).
static LiftMap lift_map(int nz, int nzc, const double* zc_coarse, const double* zc_fine) {
    LiftMap zm{};
    for (int l = 0; l < nz; ++l) {
        const double z = zc_fine[l];
        int k = 0;
        while (k + 1 < nzc && zc_coarse[k + 1] <= z) ++k;
        if (k + 1 >= nzc || z <= zc_coarse[0]) {
            zm.lo[l] = z <= zc_coarse[0] ? 0 : nzc - 1;
            zm.w[l] = 0.0f;
        } else {
            zm.lo[l] = k;
            zm.w[l] = (float)((z - zc_coarse[k]) / (zc_coarse[k + 1] - zc_coarse[k]));
        }
    }
    return zm;
}

int th_solver_set_guide(th_solver* s, const double* G) {
    TH_S(s);
    if (!G) return fail(TH_ERR_ARG, "null argument");
    return s->impl->set_guide(G);
}
int th_solver_set_guide_lifted(th_solver* s, const float* field, int Hc, int Wc, int C, int nzc,
                               const double* zc_coarse, const double* zc_fine, double scale, int passes) {
    TH_S(s);
    if (!field || !zc_coarse || !zc_fine) return fail(TH_ERR_ARG, "null argument");
    if (Hc < 1 || Wc < 1 || nzc < 1 || nzc > C || passes < 0) return fail(TH_ERR_ARG, "bad coarse field shape");
    return s->impl->set_guide_lifted(field, Hc, Wc, C, lift_map(s->impl->nz(), nzc, zc_coarse, zc_fine), scale,
                                     passes);
}
int th_solver_guide_stats(th_solver* s, int* steps, double* w_mean) {
    TH_S(s);
    return s->impl->guide_stats(steps, w_mean);
}

int th_solver_set_x_lifted(th_solver* s, const float* field, int Hc, int Wc, int C, int nzc, const double* zc_coarse,
                           const double* zc_fine, double scale, int passes, int* used, double* relres_plain,
                           double* relres_lift) {
    TH_S(s);
    if (!field || !zc_coarse || !zc_fine) return fail(TH_ERR_ARG, "null argument");
    if (Hc < 1 || Wc < 1 || nzc < 1 || nzc > C || passes < 0) return fail(TH_ERR_ARG, "bad coarse field shape");
    const int nz = s->impl->nz();
    const LiftMap zm = lift_map(nz, nzc, zc_coarse, zc_fine);
    return s->impl->set_x_lifted(field, Hc, Wc, C, zm, scale, passes, used, relres_plain, relres_lift);
}

int th_solver_set_x_learned(th_solver* s, th_surrogate* sur, const float* d_in, int nzc, const double* zc_coarse,
                            const double* zc_fine, double scale, int passes) {
    TH_S(s);
    if (!sur || !d_in || !zc_coarse || !zc_fine) return fail(TH_ERR_ARG, "null argument");
    if (nzc < 1 || nzc > 16 || passes < 0) return fail(TH_ERR_ARG, "bad coarse field shape");
    return s->impl->set_x_learned(sur, d_in, lift_map(s->impl->nz(), nzc, zc_coarse, zc_fine), scale, passes);
}
int th_solver_lifted_stats(th_solver* s, int* used, double* relres_plain, double* relres_lift) {
    TH_S(s);
    return s->impl->lifted_stats(used, relres_plain, relres_lift);
}

// transient3d.h:48-90, stepped on the device (Solver::run_transient).
int th_run_transient(th_solver* s, const th_transient_opts* o, th_step_record* recs, int rec_cap, int* nsteps,
                     double* T_final, int* all_converged, double* solve_ms) {
    TH_S(s);
    if (!o || !nsteps) return fail(TH_ERR_ARG, "null argument");
    if (!(o->dt0 > 0) || !(o->growth > 0) || !(o->dt_max > 0) || o->maxit < 0 || o->predictor < 0 || o->predictor > 3)
        return fail(TH_ERR_ARG, "transient options out of range");
    int rc = s->impl->run_transient(*o, recs, rec_cap, nsteps, all_converged, solve_ms);
    if (rc) return rc;
    if (T_final) return s->impl->get_x(T_final);
    return TH_OK;
}

}  // extern "C"
