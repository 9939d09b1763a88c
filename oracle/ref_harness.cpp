// ref_harness.cpp - a C ABI over the reference's own CPU path, compiled from
// the reference sources where they lie (see oracle/Makefile, target `ref`).
//
// TEST INFRASTRUCTURE ONLY: used to write the golden fixtures under
// tests/golden/ and, where the built library is present, to cross-check the
// oracle and to time the reference CPU path. No reference source is copied:
// this file only includes the reference headers and links pcg3d_cpu.cpp.
//
// The V-cycle and the level list are private members of CpuPcg
// (pcg3d.h:81-98); `private` is widened to `public` for this translation unit
// only, so a single preconditioner application can be captured as a fixture.
// Standard headers first, so the access widening below only reaches the
// reference's own classes.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>
#define private public
#include "case3d.h"
#include "mg3d.h"
#include "pcg3d.h"
#include "transient3d.h"
#undef private

extern "C" {

// --- models -----------------------------------------------------------------
void* ref_model_default(int nx, int ny) { return new Model3D(default_model(nx, ny)); }
void* ref_model_new(int nlayers, const double* thickness, const double* k, const double* rho_c,
                    const int* ncells, const int* active, int nx, int ny, double Lx, double Ly,
                    double h_top, double h_bot) {
    std::vector<Layer> layers;
    for (int L = 0; L < nlayers; ++L)
        layers.push_back({"L" + std::to_string(L), thickness[L], k[L], rho_c[L], ncells[L], active[L] != 0});
    return new Model3D(build_model(layers, nx, ny, Lx, Ly, h_top, h_bot));
}
// Loads a case folder with the reference loader; returns nullptr on error
// (message copied into err).
void* ref_model_case(const char* dir, int nx, int ny, int with_tsv, char* err, int errlen) {
    CaseSpec cs;
    std::string e;
    if (!load_case(dir, cs, e)) {
        std::strncpy(err, e.c_str(), errlen - 1);
        err[errlen - 1] = 0;
        return nullptr;
    }
    return new Model3D(build_case_model(cs, nx, ny, with_tsv != 0));
}
// The reference's 3D-ICE exporter (case3d.h:224-383) on a case folder: 0 on
// success, else 1 with the exporter's message copied into err.
int ref_export_3dice(const char* dir, double cell_um, const char* outdir, char* err, int errlen) {
    CaseSpec cs;
    std::string e;
    if (!load_case(dir, cs, e) || !export_3dice(cs, cell_um, outdir, e)) {
        std::strncpy(err, e.c_str(), errlen - 1);
        err[errlen - 1] = 0;
        return 1;
    }
    return 0;
}
void* ref_model_example_case(int nx, int ny, int with_tsv) {
    return new Model3D(build_case_model(example_case(), nx, ny, with_tsv != 0));
}
void ref_model_free(void* h) { delete (Model3D*)h; }
void ref_model_dims(void* h, int* nx, int* ny, int* nz, double* dx, double* dy) {
    const Model3D* m = (Model3D*)h;
    *nx = m->nx;
    *ny = m->ny;
    *nz = m->nz;
    *dx = m->dx;
    *dy = m->dy;
}
void ref_add_power_block(void* h, int layer, double x0, double y0, double w, double hh, double W) {
    add_power_block(*(Model3D*)h, layer, x0, y0, w, hh, W);
}
void ref_add_uniform_power(void* h, int layer, double W) { add_uniform_power(*(Model3D*)h, layer, W); }
void ref_add_hotspot(void* h, int layer, double xc, double yc, double size, double W) {
    add_hotspot(*(Model3D*)h, layer, xc, yc, size, W);
}
void ref_add_tsv_region(void* h, int layer, double x0, double y0, double w, double hh, double f) {
    add_tsv_region(*(Model3D*)h, layer, x0, y0, w, hh, f);
}
void ref_finalize(void* h) { finalize(*(Model3D*)h); }
void ref_model_get_f(void* h, int which, float* out) {
    const Model3D* m = (Model3D*)h;
    const std::vector<float>* v[] = {&m->gE, &m->gN, &m->gU, &m->gB, &m->cap};
    std::copy(v[which]->begin(), v[which]->end(), out);
}
void ref_model_get_d(void* h, int which, double* out) {
    const Model3D* m = (Model3D*)h;
    const std::vector<double>& v = which == 5 ? m->P : which == 6 ? m->tsv_frac : m->dz;
    std::copy(v.begin(), v.end(), out);
}
double ref_total_power(void* h) { return total_power(*(Model3D*)h); }
void ref_analytic_1d(void* h, double* out) {
    std::vector<double> T = analytic_1d(*(Model3D*)h);
    std::copy(T.begin(), T.end(), out);
}
double ref_true_relres(void* h, const double* T) {
    const Model3D* m = (Model3D*)h;
    std::vector<Real> t(T, T + m->ncells());
    return true_relres(*m, t);
}
double ref_heat_out(void* h, const double* T) {
    const Model3D* m = (Model3D*)h;
    std::vector<double> t(T, T + m->ncells());
    return heat_out(*m, t);
}

// --- multigrid levels ---------------------------------------------------------
// Writes the arrays of level k (natural layout) and returns the level count.
int ref_levels(void* h, int k, int* nx, int* ny, float* gE, float* gN, float* gU, float* gB, float* cap) {
    std::vector<MgLevelHost> L = mg_build_levels(*(Model3D*)h);
    if (k >= 0 && k < (int)L.size()) {
        const MgLevelHost& c = L[k];
        *nx = c.nx;
        *ny = c.ny;
        if (gE) {
            std::copy(c.gE.begin(), c.gE.end(), gE);
            std::copy(c.gN.begin(), c.gN.end(), gN);
            std::copy(c.gU.begin(), c.gU.end(), gU);
            std::copy(c.gB.begin(), c.gB.end(), gB);
            std::copy(c.cap.begin(), c.cap.end(), cap);
        }
    }
    return (int)L.size();
}

// --- solver -----------------------------------------------------------------
// Steady solve from T = 0 (pcg3d_cpu.cpp:187-205). precond: 0 Jacobi, 1 MG.
int ref_cpu_pcg(void* h, double tol, int maxit, int precond, double* T, int* iters, double* relres,
                double* setup_ms, double* solve_ms) {
    SolveResult r = cpu_pcg(*(Model3D*)h, tol, maxit, precond ? Precond::Multigrid : Precond::Jacobi);
    std::copy(r.T.begin(), r.T.end(), T);
    *iters = r.iters;
    *relres = r.relres;
    *setup_ms = r.setup_ms;
    *solve_ms = r.solve_ms;
    return r.converged ? 1 : 0;
}

// One preconditioner application z = M^-1 r (the V-cycle in MgReal), with the
// operator for time step dt (0 = steady).
void ref_precondition(void* h, double dt, const double* r, double* z) {
    const Model3D* m = (Model3D*)h;
    CpuPcg s(*m, Precond::Multigrid);
    s.set_dt(dt);
    std::copy(r, r + m->ncells(), s.r_.begin());
    s.precondition();
    std::copy(s.z_.begin(), s.z_.end(), z);
}

// Backward Euler (transient3d.h:48-90). rec[6 * step + k] = t, dt, iters,
// T_max, q_out, q_stored. Returns the step count (or -1 if rec_cap is short).
int ref_run_transient(void* h, int precond, double dt0, double growth, double dt_max, double t_end,
                      double tol, int maxit, double* rec, int rec_cap, double* T_final, double* solve_ms) {
    const Model3D* m = (Model3D*)h;
    TransientOptions o;
    o.dt0 = dt0;
    o.growth = growth;
    o.dt_max = dt_max;
    o.t_end = t_end;
    o.tol = tol;
    o.maxit = maxit;
    o.precond = precond ? Precond::Multigrid : Precond::Jacobi;
    TransientResult tr = run_transient<CpuPcg>(*m, o);
    if ((int)tr.steps.size() > rec_cap) return -1;
    for (size_t k = 0; k < tr.steps.size(); ++k) {
        const StepRecord& s = tr.steps[k];
        double* R = rec + 6 * k;
        R[0] = s.t;
        R[1] = s.dt;
        R[2] = s.iters;
        R[3] = s.T_max;
        R[4] = s.q_out;
        R[5] = s.q_stored;
    }
    std::copy(tr.T.begin(), tr.T.end(), T_final);
    *solve_ms = tr.solve_ms;
    return (int)tr.steps.size();
}

}  // extern "C"
