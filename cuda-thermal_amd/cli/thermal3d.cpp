// thermal3d - command-line front end of the MI355X solver, with the modes of
// the reference's CLI (main3d.cu:609-647), over the C ABI.
//
//   thermal3d                          verify, then the 256 x 256 benchmark
//   thermal3d <n>                      steady benchmark on the n x n hotspot stack
//   thermal3d transient [n] [jacobi]   heat-up after power-on; writes transient.csv
//   thermal3d case <dir> [n]           solve a case folder; writes tmap_layer<k>.csv
//   thermal3d profile [n]              one timed multigrid solve
//   thermal3d export3dice <dir> <cell_um> <outdir>   write the case for 3D-ICE
//   thermal3d compare3dice <dir> <cell_um> <icedir>  compare with 3D-ICE's results
//
// Library errors print the message and exit(1), as the reference's
// CUDA_CHECK does (cuda_check.h:8-16). The CPU path of the reference is not
// part of this library, so where the reference's verify compares CPU with
// GPU, this one compares the GPU's multigrid and Jacobi solves, the exact 1D
// chain and the conservation checks (the CPU comparison lives in the test
// suite). export3dice writes the 3D-ICE input files and compare3dice reads
// the Tmap / Tflp files 3D-ICE wrote for them; neither runs 3D-ICE itself.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/thermal_hip.h"

namespace {

const double kTol = 1e-6;      // steady benchmarks stop at ||r|| / ||P|| < kTol (main3d.cu:36)
const int kMaxIt = 50000;
const double kTrueTol = 1e-5;  // the true residual is allowed a 10x margin (main3d.cu:39)
const double kMatchTol = 1e-6; // two FP64 solves to 1e-9 agree to this, relative to Tmax

void die(const char* what) {
    std::fprintf(stderr, "thermal3d: %s failed: %s\n", what, th_last_error());
    std::exit(1);
}
#define CK(call)                          \
    do {                                  \
        if ((call) != TH_OK) die(#call);  \
    } while (0)

struct Model {
    th_model* h = nullptr;
    int nx = 0, ny = 0, nz = 0;
    double dx = 0, dy = 0;
    explicit Model(th_model* m) : h(m) { CK(th_model_dims(h, &nx, &ny, &nz, &dx, &dy)); }
    ~Model() { th_model_destroy(h); }
    Model(const Model&) = delete;
    size_t n() const { return (size_t)nx * ny * nz; }
    size_t nxy() const { return (size_t)nx * ny; }
    double T_amb() const {
        double a, b, c, d, t;
        CK(th_model_settings(h, &a, &b, &c, &d, &t));
        return t;
    }
    double power() const { return th_model_total_power(h); }
};

th_model* default_model(int n) {
    th_model* m = nullptr;
    CK(th_model_create_default(n, n, &m));
    return m;
}

// 25 W per tier, 3 W of the bottom tier in a 1 mm hotspot at (3, 3) mm
// (main3d.cu:60-67).
th_model* hotspot_model(int n) {
    th_model* m = default_model(n);
    CK(th_model_add_uniform_power(m, 0, 22.0));
    CK(th_model_add_hotspot(m, 0, 3e-3, 3e-3, 1e-3, 3.0));
    CK(th_model_add_uniform_power(m, 2, 25.0));
    CK(th_model_finalize(m));
    return m;
}

struct Result {
    std::vector<double> T;
    int iters = 0;
    bool converged = false;
    double setup_ms = 0, solve_ms = 0;
};

Result solve(const Model& m, double tol, int precond, int graphs = 1, int precision = TH_PREC_MIXED) {
    th_model_desc d;
    CK(th_model_describe(m.h, &d));
    th_solver_opts o;
    th_solver_default_opts(&o);
    o.precond = precond;
    o.use_graphs = graphs;
    o.precision = precision;
    Result r;
    r.T.resize(m.n());
    th_solve_result s{};
    CK(th_steady_solve(&d, &o, tol, kMaxIt, r.T.data(), &s));
    r.iters = s.iters;
    r.converged = s.converged;
    r.setup_ms = s.setup_ms;
    r.solve_ms = s.solve_ms;
    return r;
}

struct Transient {
    std::vector<th_step_record> steps;
    std::vector<double> T;
    bool converged = false;
    double ms = 0;
    long iters() const {
        long k = 0;
        for (const th_step_record& r : steps) k += r.iters;
        return k;
    }
};

Transient transient(const Model& m, int precond, double t_end, double growth, double dt_max, int graphs = 1) {
    th_model_desc d;
    CK(th_model_describe(m.h, &d));
    th_solver_opts o;
    th_solver_default_opts(&o);
    o.precond = precond;
    o.use_graphs = graphs;
    th_solver* s = nullptr;
    CK(th_solver_create(&d, &o, &s));
    th_transient_opts to;
    th_transient_default_opts(&to);
    to.t_end = t_end;
    to.growth = growth;
    to.dt_max = dt_max;
    Transient tr;
    tr.steps.resize(100000);
    tr.T.resize(m.n());
    int nsteps = 0, conv = 0;
    CK(th_run_transient(s, &to, tr.steps.data(), (int)tr.steps.size(), &nsteps, tr.T.data(), &conv, &tr.ms));
    th_solver_destroy(s);
    tr.steps.resize(nsteps);
    tr.converged = conv != 0;
    return tr;
}

// ||P - A T|| / ||P|| of the steady operator in FP64 on the host, in flux
// form (face fluxes G (T_c - T_nb) plus the ambient links).
double true_relres(const Model& m, const std::vector<double>& T) {
    std::vector<float> gE(m.n()), gN(m.n()), gU(m.n()), gB(m.nxy());
    std::vector<double> P(m.n());
    CK(th_model_get_f32(m.h, "gE", gE.data()));
    CK(th_model_get_f32(m.h, "gN", gN.data()));
    CK(th_model_get_f32(m.h, "gU", gU.data()));
    CK(th_model_get_f32(m.h, "gB", gB.data()));
    CK(th_model_get_f64(m.h, "P", P.data()));
    double rr = 0, pp = 0;
    for (int l = 0; l < m.nz; ++l)
        for (int j = 0; j < m.ny; ++j)
            for (int i = 0; i < m.nx; ++i) {
                const size_t c = ((size_t)l * m.ny + j) * m.nx + i;
                double f = 0;
                if (i > 0) f += gE[c - 1] * (T[c] - T[c - 1]);
                if (i + 1 < m.nx) f += gE[c] * (T[c] - T[c + 1]);
                if (j > 0) f += gN[c - m.nx] * (T[c] - T[c - m.nx]);
                if (j + 1 < m.ny) f += gN[c] * (T[c] - T[c + m.nx]);
                f += (l > 0) ? gU[c - m.nxy()] * (T[c] - T[c - m.nxy()]) : gB[c] * T[c];
                f += (l + 1 < m.nz) ? gU[c] * (T[c] - T[c + m.nxy()]) : gU[c] * T[c];
                const double r = P[c] - f;
                rr += r * r;
                pp += P[c] * P[c];
            }
    return std::sqrt(rr / pp);
}

double max_of(const std::vector<double>& v) { return *std::max_element(v.begin(), v.end()); }
double max_abs_diff(const std::vector<double>& a, const std::vector<double>& b) {
    double d = 0;
    for (size_t i = 0; i < a.size(); ++i) d = std::max(d, std::fabs(a[i] - b[i]));
    return d;
}

bool check(const char* what, double value, double limit) {
    const bool ok = value < limit;
    std::printf("  %-34s = %.3e  %s\n", what, value, ok ? "PASS" : "FAIL");
    return ok;
}

void print_layers(const Model& m, const std::vector<double>& T) {
    std::vector<int> pl(m.nz);
    CK(th_model_plane_layer(m.h, pl.data()));
    const double amb = m.T_amb();
    std::printf("  %-20s %10s %10s %10s\n", "layer", "thick(um)", "min(C)", "max(C)");
    for (int L = th_model_layer_count(m.h) - 1; L >= 0; --L) {  // top to bottom
        double th, k, rc;
        int planes, active;
        const char* name = nullptr;
        CK(th_model_layer_info(m.h, L, &th, &k, &rc, &planes, &active, &name));
        double lo = 1e30, hi = -1e30;
        for (int l = 0; l < m.nz; ++l) {
            if (pl[l] != L) continue;
            for (size_t c = 0; c < m.nxy(); ++c) {
                lo = std::min(lo, T[(size_t)l * m.nxy() + c]);
                hi = std::max(hi, T[(size_t)l * m.nxy() + c]);
            }
        }
        std::printf("  %-20s %10.0f %10.2f %10.2f\n", name, th * 1e6, lo + amb, hi + amb);
    }
}

double worst_energy_balance(const Transient& t, double P) {
    double eb = 0;
    for (const th_step_record& r : t.steps) eb = std::max(eb, std::fabs(r.q_out + r.q_stored - P) / P);
    return eb;
}

// Default stack with one layer's k replaced (the exact-1D TSV check).
th_model* default_stack_with_k(int n, int layer, double k_new) {
    th_model* d = default_model(n);
    const int nl = th_model_layer_count(d);
    std::vector<double> t(nl), k(nl), rc(nl);
    std::vector<int> pl(nl), act(nl);
    std::vector<const char*> names(nl);
    for (int L = 0; L < nl; ++L) CK(th_model_layer_info(d, L, &t[L], &k[L], &rc[L], &pl[L], &act[L], &names[L]));
    double Lx, Ly, ht, hb, amb;
    CK(th_model_settings(d, &Lx, &Ly, &ht, &hb, &amb));
    k[layer] = k_new;
    th_model* m = nullptr;
    CK(th_model_create(nl, names.data(), t.data(), k.data(), rc.data(), pl.data(), act.data(), n, n, Lx, Ly, ht, hb,
                       &m));
    th_model_destroy(d);
    return m;
}

bool verify() {
    bool ok = true;
    const int n = 64;
    {
        Model m(default_model(n));
        std::printf("== verify 1: uniform power, %d x %d x %d, vs exact 1D solution ==\n", n, n, m.nz);
        CK(th_model_add_uniform_power(m.h, 0, 25.0));
        CK(th_model_add_uniform_power(m.h, 2, 25.0));
        CK(th_model_finalize(m.h));
        Result j = solve(m, kTol, TH_PRECOND_JACOBI), g = solve(m, kTol, TH_PRECOND_MG);
        std::vector<double> Tz(m.nz);
        CK(th_model_analytic_1d(m.h, Tz.data()));
        const double tmax = *std::max_element(Tz.begin(), Tz.end());
        double ej = 0, eg = 0;
        for (int l = 0; l < m.nz; ++l)
            for (size_t c = 0; c < m.nxy(); ++c) {
                ej = std::max(ej, std::fabs(j.T[(size_t)l * m.nxy() + c] - Tz[l]));
                eg = std::max(eg, std::fabs(g.T[(size_t)l * m.nxy() + c] - Tz[l]));
            }
        std::printf("  iterations: Jacobi %d, multigrid %d\n", j.iters, g.iters);
        ok &= j.converged && g.converged;
        ok &= check("Jacobi true residual", true_relres(m, j.T), kTrueTol);
        ok &= check("multigrid true residual", true_relres(m, g.T), kTrueTol);
        ok &= check("Jacobi vs exact, max err / Tmax", ej / tmax, 1e-4);
        ok &= check("multigrid vs exact, max err / Tmax", eg / tmax, 1e-4);
        const double P = m.power();
        ok &= check("energy balance |out - in| / in", std::fabs(th_model_heat_out(m.h, g.T.data()) - P) / P, 1e-4);
    }
    {
        std::printf("\n== verify 2: hotspot in bottom tier, %d x %d ==\n", n, n);
        Model m(hotspot_model(n));
        Result j = solve(m, 1e-9, TH_PRECOND_JACOBI), g = solve(m, 1e-9, TH_PRECOND_MG);
        const double P = m.power();
        std::printf("  iterations to 1e-9: Jacobi %d, multigrid %d\n", j.iters, g.iters);
        ok &= j.converged && g.converged;
        ok &= check("Jacobi true residual", true_relres(m, j.T), kTrueTol);
        ok &= check("multigrid true residual", true_relres(m, g.T), kTrueTol);
        ok &= check("energy balance |out - in| / in", std::fabs(th_model_heat_out(m.h, g.T.data()) - P) / P, 1e-4);
        ok &= check("max |multigrid - Jacobi| / Tmax", max_abs_diff(g.T, j.T) / max_of(j.T), kMatchTol);
    }
    {
        std::printf("\n== verify 3: transient heat-up, hotspot stack, %d x %d ==\n", n, n);
        Model m(hotspot_model(n));
        const double P = m.power();
        Transient tj = transient(m, TH_PRECOND_JACOBI, 0.02, 1.2, 0.05);
        Transient tm = transient(m, TH_PRECOND_MG, 0.02, 1.2, 0.05);
        Transient tn = transient(m, TH_PRECOND_MG, 0.02, 1.2, 0.05, 0);
        std::printf("  first 20 ms: %zu steps, PCG iterations multigrid %ld, Jacobi %ld; graphs %ld, direct %ld\n",
                    tm.steps.size(), tm.iters(), tj.iters(), tm.iters(), tn.iters());
        ok &= tj.converged && tm.converged && tn.converged;
        ok &= check("multigrid vs Jacobi / Tmax at 20 ms", max_abs_diff(tm.T, tj.T) / max_of(tj.T), kMatchTol);
        ok &= check("graphs vs direct launches / Tmax", max_abs_diff(tm.T, tn.T) / max_of(tn.T), kMatchTol);
        ok &= check("energy balance, worst step", worst_energy_balance(tm, P), 1e-6);
        Transient lg = transient(m, TH_PRECOND_MG, 3.0, 1.1, 0.1);
        Result ss = solve(m, 1e-10, TH_PRECOND_MG);
        ok &= lg.converged;
        ok &= check("energy balance, worst step (3 s)", worst_energy_balance(lg, P), 1e-6);
        ok &= check("T at 3 s vs steady state / Tmax", max_abs_diff(lg.T, ss.T) / max_of(ss.T), 1e-5);
    }
    {
        std::printf("\n== verify 4: precision modes, hotspot stack, %d x %d ==\n", n, n);
        Model m(hotspot_model(n));
        Result mx = solve(m, 1e-10, TH_PRECOND_MG, 1, TH_PREC_MIXED);
        Result fp = solve(m, 1e-10, TH_PRECOND_MG, 1, TH_PREC_FP64);
        Result nog = solve(m, 1e-10, TH_PRECOND_MG, 0, TH_PREC_MIXED);
        std::printf("  iterations to 1e-10: mixed %d, FP64 %d, mixed without graphs %d\n", mx.iters, fp.iters,
                    nog.iters);
        ok &= mx.converged && fp.converged && nog.converged;
        ok &= check("max |mixed - FP64| / Tmax", max_abs_diff(mx.T, fp.T) / max_of(fp.T), kMatchTol);
        ok &= check("graphs vs direct launches / Tmax", max_abs_diff(mx.T, nog.T) / max_of(nog.T), 1e-15);
    }
    {
        std::printf("\n== verify 5: floorplan and TSVs, %d x %d ==\n", n, n);
        th_model* exh = nullptr;
        CK(th_model_example_case(n, n, 1, &exh));
        Model ex(exh);
        // power conserved with the blocks shifted off the cell edges
        Model mb(default_model(n));
        double want = 0, Lx, Ly, a, b, c;
        CK(th_model_settings(mb.h, &Lx, &Ly, &a, &b, &c));
        for (int k = 0; k < th_model_block_count(ex.h); ++k) {
            int layer;
            double x0, y0, w, h, W;
            CK(th_model_block_info(ex.h, k, nullptr, &layer, &x0, &y0, &w, &h, &W));
            const double sx = (x0 + w + 0.037e-3 <= Lx) ? 0.037e-3 : 0.0;
            const double sy = (y0 + h + 0.011e-3 <= Ly) ? 0.011e-3 : 0.0;
            CK(th_model_add_power_block(mb.h, layer, x0 + sx, y0 + sy, w, h, W));
            want += W;
        }
        ok &= check("floorplan power conserved, rel. err", std::fabs(mb.power() - want) / want, 1e-12);
        // TSVs over the whole bond layer act as a bond layer with k = k_z(eff)
        const int bond = 1;
        const double f = 0.2, k_bond = 2.0, k_cu = 400.0;
        Model mt(default_model(n));
        CK(th_model_add_uniform_power(mt.h, 0, 25.0));
        CK(th_model_add_uniform_power(mt.h, 2, 25.0));
        CK(th_model_add_tsv_region(mt.h, bond, 0, 0, Lx, Ly, f));
        CK(th_model_finalize(mt.h));
        Model ma(default_stack_with_k(n, bond, f * k_cu + (1 - f) * k_bond));
        CK(th_model_add_uniform_power(ma.h, 0, 25.0));
        CK(th_model_add_uniform_power(ma.h, 2, 25.0));
        std::vector<double> Tz(ma.nz);
        CK(th_model_analytic_1d(ma.h, Tz.data()));
        Result gt = solve(mt, 1e-10, TH_PRECOND_MG);
        double e = 0;
        for (int l = 0; l < mt.nz; ++l)
            for (size_t q = 0; q < mt.nxy(); ++q) e = std::max(e, std::fabs(gt.T[(size_t)l * mt.nxy() + q] - Tz[l]));
        ok &= gt.converged;
        ok &= check("TSV layer vs exact 1D, err / Tmax", e / *std::max_element(Tz.begin(), Tz.end()), 1e-5);
        // the example case: blocks plus TSV fields
        const double P = ex.power();
        Result g = solve(ex, 1e-9, TH_PRECOND_MG), gj = solve(ex, 1e-9, TH_PRECOND_JACOBI);
        std::printf("  example case, iterations to 1e-9: multigrid %d, Jacobi %d\n", g.iters, gj.iters);
        ok &= g.converged && gj.converged;
        ok &= check("true residual", true_relres(ex, g.T), kTrueTol);
        ok &= check("max |multigrid - Jacobi| / Tmax", max_abs_diff(g.T, gj.T) / max_of(gj.T), kMatchTol);
        ok &= check("energy balance |out - in| / in", std::fabs(th_model_heat_out(ex.h, g.T.data()) - P) / P, 1e-6);
    }
    return ok;
}

void benchmark(int n) {
    Model m(hotspot_model(n));
    std::printf("\n== steady state: %d x %d x %d = %zu cells, %.0f W, ambient %.0f C ==\n", n, n, m.nz, m.n(),
                m.power(), m.T_amb());
    {
        Model w(hotspot_model(32));  // first use of the GPU pays context setup
        solve(w, kTol, TH_PRECOND_JACOBI);
    }
    Result gm = solve(m, kTol, TH_PRECOND_MG);
    Result gj = solve(m, kTol, TH_PRECOND_JACOBI);
    Result gn = solve(m, kTol, TH_PRECOND_MG, 0);
    Result g64 = solve(m, kTol, TH_PRECOND_MG, 1, TH_PREC_FP64);
    print_layers(m, gm.T);
    std::printf("\n  %-28s %6s %10s %10s %10s %10s\n", "", "iters", "setup ms", "solve ms", "ms / iter", "speedup");
    const double ref = gj.solve_ms;
    auto row = [&](const char* name, const Result& r) {
        std::printf("  %-28s %6d %10.1f %10.2f %10.3f %9.1fx\n", name, r.iters, r.setup_ms, r.solve_ms,
                    r.solve_ms / std::max(1, r.iters), ref / r.solve_ms);
    };
    std::printf("  (solve time only; speedup is relative to GPU Jacobi)\n");
    row("GPU Jacobi", gj);
    row("GPU multigrid FP64", g64);
    row("GPU multigrid, no graphs", gn);
    row("GPU multigrid (mixed)", gm);
    std::printf("  graphs: multigrid solve %.2f -> %.2f ms (%.2fx)\n", gn.solve_ms, gm.solve_ms,
                gn.solve_ms / gm.solve_ms);
    std::printf("  max |graphs - no graphs| / Tmax = %.3e\n", max_abs_diff(gm.T, gn.T) / max_of(gm.T));
    std::printf("  max |multigrid - Jacobi| / Tmax = %.3e\n", max_abs_diff(gm.T, gj.T) / max_of(gj.T));
    const double P = m.power(), tr = true_relres(m, gm.T);
    std::printf("  GPU multigrid: true residual = %.2e   energy balance |out - in| / in = %.2e\n", tr,
                std::fabs(th_model_heat_out(m.h, gm.T.data()) - P) / P);
    if (!gm.converged || !gj.converged) std::printf("  WARNING: a solve hit the iteration limit\n");
    if (tr > kTrueTol) std::printf("  WARNING: true residual above %.0e - result is not accurate\n", kTrueTol);
}

// Time at which T_max first reaches `target`, linear between steps (-1: never).
double time_to(const std::vector<th_step_record>& steps, double target) {
    double t_prev = 0, T_prev = 0;
    for (const th_step_record& r : steps) {
        if (r.T_max >= target) return t_prev + (r.t - t_prev) * (target - T_prev) / (r.T_max - T_prev);
        t_prev = r.t;
        T_prev = r.T_max;
    }
    return -1;
}

void transient_demo(int n, int precond) {
    Model m(hotspot_model(n));
    const double P = m.power(), amb = m.T_amb();
    std::printf("\n== transient: power switched on at t = 0, %d x %d x %d = %zu cells, %.0f W, %s ==\n", n, n, m.nz,
                m.n(), P, precond == TH_PRECOND_MG ? "multigrid" : "Jacobi");
    Result ss = solve(m, 1e-8, TH_PRECOND_MG);
    const double Tss = max_of(ss.T);
    Transient tr = transient(m, precond, 2.0, 1.1, 0.1);
    std::printf("  %10s %10s %10s %8s %12s %12s\n", "t (ms)", "T_max (C)", "% of rise", "iters", "heat out(W)",
                "stored (W)");
    const double marks[] = {0.01e-3, 0.1e-3, 0.3e-3, 1e-3, 3e-3, 10e-3, 30e-3, 0.1, 0.3, 1.0, 2.0};
    size_t k = 0;
    for (double mark : marks) {
        while (k < tr.steps.size() && tr.steps[k].t < mark * (1 - 1e-9)) ++k;
        if (k >= tr.steps.size()) break;
        const th_step_record& r = tr.steps[k];
        std::printf("  %10.3f %10.2f %10.1f %8d %12.3f %12.3f\n", r.t * 1e3, r.T_max + amb, 100 * r.T_max / Tss,
                    r.iters, r.q_out, r.q_stored);
    }
    std::printf("\n  steady state T_max = %.2f C. Hotspot reaches 50%% / 90%% / 99%% of its rise at\n"
                "  %.2f ms / %.0f ms / %.0f ms\n",
                Tss + amb, 1e3 * time_to(tr.steps, 0.5 * Tss), 1e3 * time_to(tr.steps, 0.9 * Tss),
                1e3 * time_to(tr.steps, 0.99 * Tss));
    std::printf("\n  %zu steps, %ld PCG iterations, %.0f ms (%.3f ms / iteration)\n", tr.steps.size(), tr.iters(),
                tr.ms, tr.ms / std::max(1L, tr.iters()));
    std::printf("  worst energy balance |out + stored - in| / in = %.2e\n", worst_energy_balance(tr, P));
    if (!tr.converged) std::printf("  WARNING: some steps hit the iteration limit\n");
    std::printf("\n  temperatures at t = %.1f s:\n", tr.steps.back().t);
    print_layers(m, tr.T);
    FILE* f = std::fopen("transient.csv", "w");
    if (!f) {
        std::printf("  could not write transient.csv\n");
        return;
    }
    std::fprintf(f, "t_s,dt_s,pcg_iters,T_max_C,heat_out_W,heat_stored_W\n");
    for (const th_step_record& r : tr.steps)
        std::fprintf(f, "%.6e,%.6e,%d,%.6f,%.6f,%.6f\n", r.t, r.dt, r.iters, r.T_max + amb, r.q_out, r.q_stored);
    std::fclose(f);
    std::printf("\n  wrote transient.csv (%zu rows)\n", tr.steps.size());
}

// Hottest temperature through each active layer's planes, one row per y from
// the top, deg C (main3d.cu:304-323).
void write_tmaps(const Model& m, const std::vector<double>& T) {
    std::vector<int> pl(m.nz);
    CK(th_model_plane_layer(m.h, pl.data()));
    const double amb = m.T_amb();
    for (int L = 0; L < th_model_layer_count(m.h); ++L) {
        double th, k, rc;
        int planes, active;
        const char* name = nullptr;
        CK(th_model_layer_info(m.h, L, &th, &k, &rc, &planes, &active, &name));
        if (!active) continue;
        const std::string file = "tmap_layer" + std::to_string(L) + ".csv";
        FILE* f = std::fopen(file.c_str(), "w");
        if (!f) {
            std::printf("  could not write %s\n", file.c_str());
            continue;
        }
        for (int j = m.ny - 1; j >= 0; --j) {
            for (int i = 0; i < m.nx; ++i) {
                double t = -1e30;
                for (int l = 0; l < m.nz; ++l)
                    if (pl[l] == L) t = std::max(t, T[((size_t)l * m.ny + j) * m.nx + i]);
                std::fprintf(f, i ? ",%.3f" : "%.3f", t + amb);
            }
            std::fprintf(f, "\n");
        }
        std::fclose(f);
        std::printf("  wrote %s  (%s, %d x %d, deg C)\n", file.c_str(), name, m.nx, m.ny);
    }
}

int run_case(const std::string& dir, int n) {
    th_model* h = nullptr;
    if (th_model_load_case(dir.c_str(), n, n, 1, &h) != TH_OK) {
        std::printf("Could not load case '%s': %s\n", dir.c_str(), th_last_error());
        return 1;
    }
    Model m(h);
    const double P = m.power(), amb = m.T_amb();
    const int nb = th_model_block_count(m.h), nt = th_model_tsv_count(m.h);
    std::printf("\n== case %s: %d layers, %d power blocks, %d TSV regions, %.1f W, %d x %d x %d cells ==\n", dir.c_str(),
                th_model_layer_count(m.h), nb, nt, P, n, n, m.nz);
    Result g = solve(m, 1e-8, TH_PRECOND_MG);
    std::printf("  solved in %d multigrid PCG iterations, %.1f ms (+ %.0f ms setup); true residual %.1e, energy "
                "balance %.1e\n",
                g.iters, g.solve_ms, g.setup_ms, true_relres(m, g.T),
                std::fabs(th_model_heat_out(m.h, g.T.data()) - P) / P);
    if (!g.converged) std::printf("  WARNING: did not converge\n");
    std::vector<double> T0;  // the same case without TSVs, to show their effect
    if (nt > 0) {
        th_model* h0 = nullptr;
        CK(th_model_load_case(dir.c_str(), n, n, 0, &h0));
        Model m0(h0);
        T0 = solve(m0, 1e-8, TH_PRECOND_MG).T;
    }
    std::printf("\n");
    print_layers(m, g.T);
    std::printf("\n  %-14s %-20s %7s %9s %9s %9s", "block", "layer", "W", "W/cm^2", "max(C)", "avg(C)");
    if (!T0.empty()) std::printf(" %13s", "TSVs change");
    std::printf("\n");
    for (int k = 0; k < nb; ++k) {
        const char* name = nullptr;
        int layer;
        double x0, y0, w, hh, W, tmax, tavg;
        CK(th_model_block_info(m.h, k, &name, &layer, &x0, &y0, &w, &hh, &W));
        CK(th_model_block_temps(m.h, k, g.T.data(), &tmax, &tavg));
        double lt, lk, lrc;
        int lp, la;
        const char* lname = nullptr;
        CK(th_model_layer_info(m.h, layer, &lt, &lk, &lrc, &lp, &la, &lname));
        std::printf("  %-14s %-20s %7.2f %9.1f %9.2f %9.2f", name, lname, W, W / (w * hh * 1e4), tmax + amb, tavg + amb);
        if (!T0.empty()) {
            double t0max, t0avg;
            CK(th_model_block_temps(m.h, k, T0.data(), &t0max, &t0avg));
            std::printf(" %+10.2f C", tmax - t0max);
        }
        std::printf("\n");
    }
    const size_t hot = (size_t)(std::max_element(g.T.begin(), g.T.end()) - g.T.begin());
    const int l = (int)(hot / m.nxy()), j = (int)((hot % m.nxy()) / m.nx), i = (int)(hot % m.nx);
    std::vector<int> pl(m.nz);
    CK(th_model_plane_layer(m.h, pl.data()));
    double lt, lk, lrc;
    int lp, la;
    const char* lname = nullptr;
    CK(th_model_layer_info(m.h, pl[l], &lt, &lk, &lrc, &lp, &la, &lname));
    std::printf("\n  hottest point: %.2f C in %s at x = %.2f mm, y = %.2f mm\n", g.T[hot] + amb, lname,
                (i + 0.5) * m.dx * 1e3, (j + 0.5) * m.dy * 1e3);
    if (!T0.empty()) std::printf("  without TSVs:  %.2f C\n", max_of(T0) + amb);
    std::printf("\n");
    write_tmaps(m, g.T);
    return 0;
}

// export_case_3dice (main3d.cu:405-416)
int export_case_3dice(const std::string& dir, double cell_um, const std::string& outdir) {
    int nx = 0, ny = 0;
    if (th_case_export_3dice(dir.c_str(), cell_um, outdir.c_str(), &nx, &ny) != TH_OK) {
        std::printf("Export failed: %s\n", th_last_error());
        return 1;
    }
    std::printf("Wrote %s/case.stk and floorplans (%g um cells, %d x %d).\n", outdir.c_str(), cell_um, nx, ny);
    std::printf("Run 3D-ICE there:  3D-ICE-Emulator case.stk\n");
    return 0;
}

// compare_case_3dice (main3d.cu:418-488): the case solved on the grid of the
// export, each die's top plane against 3D-ICE's Tmap (K, row y = 0 first)
// and each block's max / average against its Tflp files.
int compare_case_3dice(const std::string& dir, double cell_um, const std::string& icedir) {
    th_model* probe = nullptr;
    if (th_model_load_case(dir.c_str(), 1, 1, 1, &probe) != TH_OK) {
        std::printf("Could not load case '%s': %s\n", dir.c_str(), th_last_error());
        return 1;
    }
    double Lx, Ly, ht, hb, amb;
    CK(th_model_settings(probe, &Lx, &Ly, &ht, &hb, &amb));
    th_model_destroy(probe);
    const int nx = (int)std::lround(Lx * 1e6 / cell_um), ny = (int)std::lround(Ly * 1e6 / cell_um);
    if (nx < 1 || ny < 1) {
        std::printf("Could not load case '%s': cell size %g um gives no cells\n", dir.c_str(), cell_um);
        return 1;
    }
    th_model* h = nullptr;
    CK(th_model_load_case(dir.c_str(), nx, ny, 1, &h));
    Model m(h);
    Result g = solve(m, 1e-10, TH_PRECOND_MG);
    std::printf("\n== compare with 3D-ICE: %s, %d x %d x %d cells, %.1f W ==\n", dir.c_str(), nx, ny, m.nz, m.power());
    std::printf("  ours: %d multigrid iterations, %.1f ms (+ %.0f ms setup), true residual %.1e\n", g.iters,
                g.solve_ms, g.setup_ms, true_relres(m, g.T));
    std::vector<int> pl(m.nz);
    CK(th_model_plane_layer(m.h, pl.data()));
    int failures = 0;
    for (int L = 0; L < th_model_layer_count(m.h); ++L) {
        double lt, lk, lrc;
        int lp, active;
        const char* lname = nullptr;
        CK(th_model_layer_info(m.h, L, &lt, &lk, &lrc, &lp, &active, &lname));
        if (!active) continue;
        int plane = -1;  // the layer's top plane
        for (int l = 0; l < m.nz; ++l)
            if (pl[l] == L) plane = l;
        const std::string path = icedir + "/tmap_L" + std::to_string(L) + ".txt";
        int rows = 0, cols = 0;
        std::vector<double> ice;
        bool ok = th_3dice_read_tmap(path.c_str(), &rows, &cols, nullptr, 0) == TH_OK && rows == ny && cols == nx;
        if (ok) {
            ice.resize((size_t)rows * cols);
            ok = th_3dice_read_tmap(path.c_str(), &rows, &cols, ice.data(), (long)ice.size()) == TH_OK;
        }
        if (!ok) {
            std::printf("  %s: missing, or not %d rows x %d columns\n", path.c_str(), ny, nx);
            ++failures;
            continue;
        }
        double maxd = 0, sumd = 0, ours_max = -1e30, ice_max = -1e30;
        int wi = 0, wj = 0;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double a = g.T[((size_t)plane * ny + j) * nx + i] + amb, b = ice[(size_t)j * nx + i] - 273.15;
                const double d = a - b;
                sumd += d;
                if (std::fabs(d) > maxd) {
                    maxd = std::fabs(d);
                    wi = i;
                    wj = j;
                }
                ours_max = std::max(ours_max, a);
                ice_max = std::max(ice_max, b);
            }
        std::printf("\n  %s (whole map, %d cells):\n", lname, nx * ny);
        std::printf("    hottest cell: ours %.3f C, 3D-ICE %.3f C\n", ours_max, ice_max);
        std::printf("    max |difference| = %.4f C at x = %.2f mm, y = %.2f mm; mean difference %+.4f C\n", maxd,
                    (wi + 0.5) * m.dx * 1e3, (wj + 0.5) * m.dy * 1e3, sumd / (nx * ny));
        // Tflp files: block names and one value each
        auto read_tflp = [](const std::string& p, std::vector<std::string>& names, std::vector<double>& vals) {
            int n = 0;
            if (th_3dice_read_tflp(p.c_str(), nullptr, 0, nullptr, 0, &n) != TH_OK) return false;
            const int len = 256;
            std::vector<char> buf((size_t)n * len);
            vals.resize(n);
            if (th_3dice_read_tflp(p.c_str(), buf.data(), len, vals.data(), n, &n) != TH_OK) return false;
            names.clear();
            for (int k = 0; k < n; ++k) names.emplace_back(buf.data() + (size_t)k * len);
            return true;
        };
        std::vector<std::string> nmax, navg;
        std::vector<double> vmax, vavg;
        if (!read_tflp(icedir + "/tflp_max_L" + std::to_string(L) + ".txt", nmax, vmax) ||
            !read_tflp(icedir + "/tflp_avg_L" + std::to_string(L) + ".txt", navg, vavg)) {
            std::printf("    (no Tflp files)\n");
            continue;
        }
        std::printf("    %-14s %10s %10s %8s %10s %10s %8s\n", "block", "max ours", "max 3DICE", "diff", "avg ours",
                    "avg 3DICE", "diff");
        for (int k = 0; k < th_model_block_count(m.h); ++k) {
            const char* bname = nullptr;
            int layer;
            double x0, y0, w, hh, W;
            CK(th_model_block_info(m.h, k, &bname, &layer, &x0, &y0, &w, &hh, &W));
            if (layer != L) continue;
            std::string id(bname);  // 3D-ICE element names: letters and digits, '_' for the rest
            for (char& c : id)
                if (!std::isalnum((unsigned char)c)) c = '_';
            const size_t a = std::find(nmax.begin(), nmax.end(), id) - nmax.begin();
            const size_t b = std::find(navg.begin(), navg.end(), id) - navg.begin();
            if (a >= vmax.size() || b >= vavg.size()) continue;
            double tmax, tavg;
            CK(th_model_block_temps(m.h, k, g.T.data(), &tmax, &tavg));
            tmax += amb;
            tavg += amb;
            const double imax = vmax[a] - 273.15, iavg = vavg[b] - 273.15;
            std::printf("    %-14s %10.3f %10.3f %+8.4f %10.3f %10.3f %+8.4f\n", bname, tmax, imax, tmax - imax, tavg,
                        iavg, tavg - iavg);
        }
    }
    return failures ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 5 && std::strcmp(argv[1], "export3dice") == 0)
        return export_case_3dice(argv[2], std::atof(argv[3]), argv[4]);
    if (argc >= 5 && std::strcmp(argv[1], "compare3dice") == 0)
        return compare_case_3dice(argv[2], std::atof(argv[3]), argv[4]);
    if (argc >= 3 && std::strcmp(argv[1], "case") == 0) return run_case(argv[2], argc >= 4 ? std::atoi(argv[3]) : 200);
    if (argc >= 2 && std::strcmp(argv[1], "profile") == 0) {
        const int n = argc >= 3 ? std::atoi(argv[2]) : 800;
        Model m(hotspot_model(n));
        Result g = solve(m, kTol, TH_PRECOND_MG);
        std::printf("profile: %d x %d x %d, %d multigrid iterations, setup %.1f ms, solve %.2f ms (%.3f ms / "
                    "iteration)\n",
                    n, n, m.nz, g.iters, g.setup_ms, g.solve_ms, g.solve_ms / std::max(1, g.iters));
        return 0;
    }
    if (argc >= 2 && std::strcmp(argv[1], "transient") == 0) {
        const int n = argc >= 3 ? std::atoi(argv[2]) : 200;
        const bool jacobi = argc >= 4 && std::strcmp(argv[3], "jacobi") == 0;
        transient_demo(n, jacobi ? TH_PRECOND_JACOBI : TH_PRECOND_MG);
        return 0;
    }
    if (argc == 2) {
        benchmark(std::atoi(argv[1]));
        return 0;
    }
    std::printf("%s\n", th_version());
    if (!verify()) {
        std::printf("\nVerification FAILED - fix correctness before benchmarking.\n");
        return 1;
    }
    benchmark(256);
    std::printf("\nNext: thermal3d transient   (heat-up after power-on)\n");
    return 0;
}
