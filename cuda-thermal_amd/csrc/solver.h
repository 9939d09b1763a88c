// solver.h - MG-preconditioned CG for layered stacks on MI355X (gfx950): the
// host side of one solve (levels, buffers, launch sequences, graphs).
//
// Follows the reference's GPU/CPU solver (src/pcg3d_gpu.cu, src/pcg3d_cpu.cpp,
// src/mg3d.h) in what it computes; the device design is this file's own:
//  * every vector and level in one colour-split layout (kernels.h), so PCG
//    and multigrid hand vectors over without re-layout kernels;
//  * one thread per column for stencil work, vertical neighbours in registers;
//  * the V-cycle needs no clearing pass: the first (red) pre-smoothing sweep
//    runs with the black neighbours taken as zero, which is what the
//    reference's memset-then-sweep computes, and prolongation is folded into
//    the first post-smoothing sweep, which adds the parent's correction to
//    each neighbour it reads (the cells it writes are overwritten anyway);
//  * only 1/pivot is stored per cell: the upper Thomas factor is rebuilt in
//    the sweep as (-gU) * (1/pivot), the same product the reference stores;
//  * PCG scalars, tolerance test and iteration count live on the device;
//    every kernel returns at once when the solve is done, so iterations can
//    be queued ahead of the host's convergence check and replayed as graphs.
// Precision is a runtime choice: T = float (FP32 conductances, factors and
// V-cycle inside the FP64 PCG) or T = double (FP64 throughout).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "kernels.h"
#include "levels.h"
#include "mg_kernels.h"
#include "model.h"
#include "pcg_kernels.h"

namespace th {

// ===========================================================================
// Host side
// ===========================================================================

#define HIPTRY(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) return fail(TH_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

static dim3 col_block() { return dim3(kColsX, kColsY); }
static dim3 col_grid(int slots, int rows) {
    return dim3((unsigned)((slots + kColsX - 1) / kColsX), (unsigned)((rows + kColsY - 1) / kColsY));
}
static unsigned vec_blocks(long n) {
    return (unsigned)std::max(1L, std::min((long)kMaxVecBlocks, (n + kVecThreads - 1) / kVecThreads));
}
static unsigned lin_blocks(long n) { return (unsigned)((n + 255) / 256); }
static dim3 stencil_block() { return dim3(kColsX, kStencilRows); }
static dim3 stencil_grid(int slots, int rows) {
    return dim3((unsigned)((slots + kColsX - 1) / kColsX), (unsigned)((rows + kStencilRows - 1) / kStencilRows));
}

template <class T>
struct DevLevel {
    Geo g;
    T *gE = nullptr, *gN = nullptr, *gU = nullptr, *gB = nullptr, *cap = nullptr;
    T *finv = nullptr, *x = nullptr, *b = nullptr;
    // column types (levels.h: column_types) of a level run by the column
    // walk; 0 types: the level's kernels read the per-cell arrays
    int ntypes = 0;
    unsigned short* ty = nullptr;
    T* tab = nullptr;
    unsigned char* tfl = nullptr;
    Ops<T> ops() const { return {gE, gN, gU, gB, cap}; }
    Typed<T> typed() const { return Typed<T>{ty, tab, g.nz}; }
};

struct SolverBase {
    virtual ~SolverBase() {}
    virtual int init(const th_model_desc& m, const th_solver_opts& o) = 0;
    virtual int set_dt(double dt) = 0;
    virtual int set_x_zero() = 0;
    virtual int set_x(const double* x) = 0;
    virtual int set_rhs_steady() = 0;
    virtual int set_rhs_transient() = 0;
    virtual int solve(double tol, int maxit, th_stats* st) = 0;
    virtual int get_x(double* x) = 0;
    virtual int set_power(const double* P) = 0;
    virtual int device_x(const double** p) = 0;
    virtual int true_relres(double* out) = 0;
    virtual int energy(double* heat, double* out, double* tmax) = 0;
    virtual int nlevels() const = 0;
    virtual hipStream_t stream() const = 0;
    virtual int precondition(const double* r, double* z) = 0;
    virtual int apply_operator(const double* x, double* y) = 0;
    virtual int time_kernel(int which, int reps, double* avg_ms, double* bytes) = 0;
    virtual int run_transient(const th_transient_opts& o, th_step_record* recs, int rec_cap, int* nsteps,
                              int* all_converged, double* ms) = 0;
    virtual int set_x_lifted(const float* field, int Hc, int Wc, int C, const LiftMap& zm, double scale, int passes,
                             int* used, double* rp, double* rl) = 0;
    virtual int nz() const = 0;
    virtual int column_types() const = 0;
    virtual int level_info(int k, int* nx, int* ny, int* nz, int* ntypes, int* kind) const = 0;
    virtual int colorize(int l0, int l1, float rise_hi, unsigned char* rgba, bool device_rgba, float* hot_rise,
                         int* hot_cell) = 0;
    virtual int check_colorize_args(int l0, int l1, float rise_hi) const = 0;
    virtual long colorize_blocks() const = 0;
    virtual int colorize_into(int l0, int l1, float rise_hi, unsigned* dst, unsigned long long* part,
                              unsigned long long* key, unsigned char* host_rgba, float* hot_rise, int* hot_cell) = 0;
    virtual int device() const = 0;
    virtual int setup_times(th_setup_times* out) const = 0;
    virtual int set_guide(const double* host) = 0;
    virtual int set_guide_lifted(const float* field, int Hc, int Wc, int C, const LiftMap& zm, double scale,
                                 int passes) = 0;
    virtual int guide_stats(int* steps, double* w_mean) = 0;
    virtual int set_x_learned(th_surrogate* sur, const float* d_in, const LiftMap& zm, double scale, int passes) = 0;
    virtual int lifted_stats(int* used, double* rp, double* rl) = 0;
};

template <class T>
struct Solver final : SolverBase {
    th_solver_opts opt{};
    hipStream_t s = nullptr;
    Geo g0{};
    long n0 = 0;       // cells in the split layout (with padding)
    long nnat = 0;     // cells in the natural layout
    double inv_dt = 0;
    double *P = nullptr, *x = nullptr, *b = nullptr, *r = nullptr, *p = nullptr, *Ap = nullptr;
    double *z = nullptr, *invd = nullptr, *xnat = nullptr;
    T* pm = nullptr;  // multigrid search direction, in the V-cycle's precision (p itself in FP64)
    bool allow_types = true;  // column types on the levels that run the column walk
    double* part = nullptr;
    State* st = nullptr;
    State* h_st = nullptr;  // pinned mirror
    std::vector<DevLevel<T>> L;
    hipGraphExec_t g_start = nullptr, g_iter = nullptr;
    hipGraphExec_t g_start_P = nullptr;  // g_start reading P as the right-hand side
    bool rhs_P = false;                  // set_rhs_steady: b is P, not copied (rhs())
    const double* rhs() const { return rhs_P ? P : b; }
    // An iteration in two parts (split_iter: multigrid with a V-cycle whose
    // head is a 256-thread sweep): apply, update and the convergence test;
    // then V-cycle, dot and p update. mid_iter: the first part of the last
    // queued iteration has run, the second has not.
    hipGraphExec_t g_iter_a = nullptr, g_iter_b = nullptr;
    bool split_iter = false, mid_iter = false;
    std::vector<void*> owned;

    ~Solver() override {
        if (s) (void)hipStreamSynchronize(s);
        if (g_start) (void)hipGraphExecDestroy(g_start);
        if (g_start_P) (void)hipGraphExecDestroy(g_start_P);
        if (g_iter) (void)hipGraphExecDestroy(g_iter);
        if (g_iter_a) (void)hipGraphExecDestroy(g_iter_a);
        if (g_iter_b) (void)hipGraphExecDestroy(g_iter_b);
        for (void* ptr : owned) (void)hipFree(ptr);
        if (h_st) (void)hipHostFree(h_st);
        if (s) (void)hipStreamDestroy(s);
    }

    template <class U>
    int dalloc(U** ptr, size_t count) {
        void* v = nullptr;
        HIPTRY(hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(U)));
        owned.push_back(v);
        HIPTRY(hipMemsetAsync(v, 0, std::max<size_t>(count, 1) * sizeof(U), s));
        *ptr = (U*)v;
        return TH_OK;
    }

    bool mg() const { return opt.precond == TH_PRECOND_MG; }

    // natural-layout host array -> split-layout device array of type D,
    // through a staging buffer kept for later uploads (set_x, set_power)
    void* stage = nullptr;
    size_t stage_bytes = 0;
    template <class S, class D>
    int upload_split(const Geo& g, int planes, const S* host, D* dev) {
        const long n = (long)g.nx * g.ny * planes;
        if ((size_t)n * sizeof(S) > stage_bytes) {
            if (stage) {
                HIPTRY(hipStreamSynchronize(s));
                owned.erase(std::find(owned.begin(), owned.end(), stage));
                HIPTRY(hipFree(stage));
                stage = nullptr;
            }
            HIPTRY(hipMalloc(&stage, n * sizeof(S)));
            owned.push_back(stage);
            stage_bytes = n * sizeof(S);
        }
        hipError_t e = hipMemcpyAsync(stage, host, n * sizeof(S), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            k_to_split<S, D><<<lin_blocks(n), 256, 0, s>>>(g, planes, (const S*)stage, dev);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return fail(TH_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
        return TH_OK;
    }

    // Where the set-up time went (GpuSetupTimes, pcg3d.h:101-109), in ms.
    th_setup_times setup{};

    int init(const th_model_desc& m, const th_solver_opts& o) override {
        using clk = std::chrono::steady_clock;
        const auto t_start = clk::now();
        auto t_mark = t_start;
        auto lap = [&](double& acc) {
            const auto t = clk::now();
            acc += std::chrono::duration<double, std::milli>(t - t_mark).count();
            t_mark = t;
        };
        opt = o;
        HIPTRY(hipSetDevice(o.device));
        HIPTRY(hipStreamCreate(&s));
        g0 = make_geo(m.nx, m.ny, m.nz);
        n0 = g0.cells();
        nnat = (long)m.nx * m.ny * m.nz;
        int rc = 0;
        for (double** v : {&P, &x, &b, &r, &p, &Ap})
            if ((rc = dalloc(v, n0))) return rc;
        if ((rc = dalloc(&xnat, nnat))) return rc;
        if (!mg())
            for (double** v : {&z, &invd})
                if ((rc = dalloc(v, n0))) return rc;
        if (mg()) {
            if constexpr (std::is_same<T, double>::value) pm = p;
            else if ((rc = dalloc(&pm, n0))) return rc;
        }
        if ((rc = dalloc(&st, 1))) return rc;
        HIPTRY(hipHostMalloc((void**)&h_st, sizeof(State), hipHostMallocDefault));
        std::memset(h_st, 0, sizeof(State));
        lap(setup.alloc_ms);

        // Levels: coarse operators aggregated on the host (levels.cpp), each
        // level stored in the split layout in precision T.
        std::vector<HostLevel> H;
        if (mg()) {
            H = build_hierarchy(m.nx, m.ny, m.nz, m.dx, m.dy, m.gE, m.gN, m.gU, m.gB, m.cap);
        } else {  // Jacobi still needs the fine operator
            HostLevel f;
            f.nx = m.nx;
            f.ny = m.ny;
            f.nz = m.nz;
            H.push_back(f);
        }
        lap(setup.levels_ms);
        L.resize(H.size());
        long maxblk = kMaxVecBlocks;
        for (size_t k = 0; k < H.size(); ++k) {
            DevLevel<T>& d = L[k];
            d.g = make_geo(H[k].nx, H[k].ny, H[k].nz);
            const long nc = d.g.cells();
            for (T** v : {&d.gE, &d.gN, &d.gU, &d.cap, &d.finv, &d.x, &d.b})
                if ((rc = dalloc(v, nc))) return rc;
            if ((rc = dalloc(&d.gB, d.g.plane))) return rc;
            lap(setup.alloc_ms);
            const float* src[5];
            if (k == 0) {
                src[0] = m.gE, src[1] = m.gN, src[2] = m.gU, src[3] = m.cap, src[4] = m.gB;
            } else {
                src[0] = H[k].gE.data(), src[1] = H[k].gN.data(), src[2] = H[k].gU.data(),
                src[3] = H[k].cap.data(), src[4] = H[k].gB.data();
            }
            T* dst[5] = {d.gE, d.gN, d.gU, d.cap, d.gB};
            for (int a = 0; a < 5; ++a)
                if ((rc = upload_split<float, T>(d.g, a == 4 ? 1 : d.g.nz, src[a], dst[a]))) return rc;
            lap(k == 0 ? setup.upload_ms : setup.levels_ms);
            // column types where the level runs the column walk
            if (allow_types && mg() && (long)d.g.hx * d.g.ny > kPPCols) {
                const ColumnTypes ct = th::column_types(d.g.nx, d.g.ny, d.g.nz, src[0], src[1], src[2], src[4],
                                                        src[3], kMaxColumnTypes);
                if (ct.ntypes > 0) {
                    std::vector<unsigned short> slot((size_t)d.g.plane, 0);  // padding slots: type 0, never read
                    for (int j = 0; j < d.g.ny; ++j)
                        for (int i = 0; i < d.g.nx; ++i) slot[d.g.pos(i, j)] = ct.type[(size_t)j * d.g.nx + i];
                    const std::vector<T> tab(ct.tab.begin(), ct.tab.end());
                    if ((rc = dalloc(&d.ty, slot.size())) || (rc = dalloc(&d.tab, tab.size())) ||
                        (rc = dalloc(&d.tfl, ct.flags.size())))
                        return rc;
                    HIPTRY(hipMemcpyAsync(d.ty, slot.data(), slot.size() * sizeof(unsigned short),
                                          hipMemcpyHostToDevice, s));
                    HIPTRY(hipMemcpyAsync(d.tab, tab.data(), tab.size() * sizeof(T), hipMemcpyHostToDevice, s));
                    HIPTRY(hipMemcpyAsync(d.tfl, ct.flags.data(), ct.flags.size(), hipMemcpyHostToDevice, s));
                    HIPTRY(hipStreamSynchronize(s));
                    d.ntypes = ct.ntypes;
                }
            }
            lap(setup.types_ms);
            const dim3 cg = col_grid(2 * d.g.hx, d.g.ny);
            maxblk = std::max(maxblk, (long)cg.x * cg.y);
            if (!mg()) break;
        }
        if ((rc = dalloc(&part, maxblk * 4))) return rc;
        {
            const dim3 sg = stencil_grid(2 * g0.hx, g0.ny);
            // launch_apply_kernel's and launch_residual's grid: 64 x 4 blocks with column types
            const dim3 ag = L[0].ntypes ? col_grid(2 * g0.hx, g0.ny) : sg;
            nblk_stencil = sg.x * sg.y;
            nblk_apply = ag.x * ag.y;
            nblk_res = ag.x * ag.y;
            nblk_vec = vec_blocks(n0);
            if ((rc = dalloc(&part_res, (size_t)std::max(nblk_stencil, nblk_res) * 3)) ||
                (rc = dalloc(&part_apply, nblk_apply)) ||
                (rc = dalloc(&part_upd, (size_t)nblk_vec * 2)) || (rc = dalloc(&part_dot, nblk_vec)))
                return rc;
        }
        // The lowest levels run in one workgroup. FP32 with the lateral
        // restriction and nz <= 16: as many levels as fit in LDS (k_tail_lds).
        // Otherwise: levels with at most kTailCols columns (k_tail, L2).
        fast = opt.restriction == TH_RESTRICT_LATERAL;
        tail_k0 = L.size() - 1;
        lds_tail = fast && g0.nz <= 16;
        if (lds_tail) {
            auto bytes_from = [&](size_t k0) {
                std::vector<Geo> gs;
                for (size_t k = k0; k < L.size(); ++k) gs.push_back(L[k].g);
                return tail_lds_bytes<T>(gs.data(), (int)gs.size(), pp_lanes(g0.nz));
            };
            while (tail_k0 > 0 && bytes_from(tail_k0 - 1) <= kTailLdsBytes && L.size() - (tail_k0 - 1) <= (size_t)kMaxTail)
                --tail_k0;
            lds_bytes = bytes_from(tail_k0);
            HIPTRY(hipFuncSetAttribute((const void*)k_tail_lds<T, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, kTailLdsBytes));
            HIPTRY(hipFuncSetAttribute((const void*)k_tail_lds<T, 8>, hipFuncAttributeMaxDynamicSharedMemorySize, kTailLdsBytes));
            HIPTRY(hipFuncSetAttribute((const void*)k_tail_lds<T, 12>, hipFuncAttributeMaxDynamicSharedMemorySize, kTailLdsBytes));
            HIPTRY(hipFuncSetAttribute((const void*)k_tail_lds<T, 16>, hipFuncAttributeMaxDynamicSharedMemorySize, kTailLdsBytes));
        } else {
            while (tail_k0 > 0 && (long)L[tail_k0 - 1].g.nx * L[tail_k0 - 1].g.ny <= kTailCols &&
                   L.size() - (tail_k0 - 1) <= (size_t)kMaxTail)
                --tail_k0;
        }
        lap(setup.alloc_ms);
        if ((rc = upload_split<double, double>(g0, g0.nz, m.P, P))) return rc;
        lap(setup.upload_ms);
        if ((rc = set_dt(0.0))) return rc;
        HIPTRY(hipStreamSynchronize(s));
        lap(setup.factor_ms);
        split_iter = mg() && tail_k0 > 0;
        if (opt.use_graphs && mg()) {
            if ((rc = capture(&g_start, true)) ||
                (rc = capture_fn(&g_start_P, [&]() { return enqueue_start(false, true); })))
                return rc;
            if ((rc = capture(&g_iter, false))) return rc;
            if (split_iter && ((rc = capture_fn(&g_iter_a, [&]() { return enqueue_iter_a(); })) ||
                               (rc = capture_fn(&g_iter_b, [&]() { return enqueue_iter_b(); }))))
                return rc;
        }
        lap(setup.graphs_ms);
        setup.total_ms = std::chrono::duration<double, std::milli>(clk::now() - t_start).count();
        return TH_OK;
    }
    int setup_times(th_setup_times* out) const override {
        *out = setup;
        return TH_OK;
    }

    int capture(hipGraphExec_t* exec, bool start) {
        return capture_fn(exec, [&]() { return start ? enqueue_start() : enqueue_iter(); });
    }
    template <class F>
    int capture_fn(hipGraphExec_t* exec, F&& enqueue) {
        hipGraph_t gr = nullptr;
        HIPTRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = enqueue();
        hipError_t e = hipStreamEndCapture(s, &gr);
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        if (e != hipSuccess) return fail(TH_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(exec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e != hipSuccess) return fail(TH_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
        return TH_OK;
    }

    int nlevels() const override { return mg() ? (int)L.size() : 0; }
    hipStream_t stream() const override { return s; }

    int set_dt(double dt) override {
        inv_dt = dt > 0 ? 1.0 / dt : 0.0;
        k_set_inv_dt<<<1, 1, 0, s>>>(st, inv_dt);
        HIPTRY(hipGetLastError());
        return enqueue_factors();
    }
    // Level factors (MG) or 1/diag (Jacobi) for the state's 1/dt; the kernels
    // return at once unless the state asks for new factors.
    int enqueue_factors() {
        if (mg()) {
            // a level with column types reads its factors from the type table
            // only: its per-cell 1/pivot is never read. The other levels'
            // factors, and the typed levels' types, one launch each per
            // kFactorLevels levels (every level has the fine level's nz).
            FactorLevels<T> fl{};
            TypeLevels<T> tl{};
            auto flush = [&]() -> int {
                if (fl.n) {
                    const unsigned nb = (unsigned)fl.blk0[fl.n];
#define TH_FACTOR(NZ) k_factor_levels<T, NZ><<<nb, col_block(), 0, s>>>(fl, st)
                    if (g0.nz <= 4) TH_FACTOR(4);
                    else if (g0.nz <= 8) TH_FACTOR(8);
                    else if (g0.nz <= 12) TH_FACTOR(12);
                    else if (g0.nz <= 16) TH_FACTOR(16);
                    else if (g0.nz <= 24) TH_FACTOR(24);
                    else if (g0.nz <= 32) TH_FACTOR(32);
                    else if (g0.nz <= 48) TH_FACTOR(48);
                    else TH_FACTOR(64);
#undef TH_FACTOR
                    HIPTRY(hipGetLastError());
                    fl.n = 0;
                }
                if (tl.n) {
                    const unsigned nb = (unsigned)tl.blk0[tl.n];
#define TH_FTYPES(NZ) k_factor_types_levels<T, NZ><<<nb, 64, 0, s>>>(tl, g0.nz, st)
                    if (g0.nz <= 4) TH_FTYPES(4);
                    else if (g0.nz <= 8) TH_FTYPES(8);
                    else if (g0.nz <= 12) TH_FTYPES(12);
                    else if (g0.nz <= 16) TH_FTYPES(16);
                    else if (g0.nz <= 24) TH_FTYPES(24);
                    else if (g0.nz <= 32) TH_FTYPES(32);
                    else if (g0.nz <= 48) TH_FTYPES(48);
                    else TH_FTYPES(64);
#undef TH_FTYPES
                    HIPTRY(hipGetLastError());
                    tl.n = 0;
                }
                return TH_OK;
            };
            int rc;
            for (DevLevel<T>& d : L) {
                if (d.ntypes) {
                    if (tl.n == kFactorLevels && (rc = flush())) return rc;
                    const int k = tl.n++;
                    tl.blk0[k] = k ? tl.blk0[k] : 0;
                    tl.nt[k] = d.ntypes;
                    tl.tab[k] = d.tab;
                    tl.fl[k] = d.tfl;
                    tl.blk0[k + 1] = tl.blk0[k] + (d.ntypes + 63) / 64;
                } else {
                    if (fl.n == kFactorLevels && (rc = flush())) return rc;
                    const dim3 grid = col_grid(2 * d.g.hx, d.g.ny);
                    const int k = fl.n++;
                    fl.blk0[k] = k ? fl.blk0[k] : 0;
                    fl.gx[k] = (int)grid.x;
                    fl.g[k] = d.g;
                    fl.o[k] = d.ops();
                    fl.finv[k] = d.finv;
                    fl.blk0[k + 1] = fl.blk0[k] + (int)(grid.x * grid.y);
                }
            }
            if ((rc = flush())) return rc;
        } else {
            const DevLevel<T>& d = L[0];
            k_invdiag<T><<<col_grid(2 * g0.hx, g0.ny), col_block(), 0, s>>>(g0, d.ops(), invd, st);
            HIPTRY(hipGetLastError());
        }
        return TH_OK;
    }
    bool x_zero = true;      // x is known to be 0 (set_x_zero, nothing since)
    bool xz_device = false;  // the device decided whether x is still 0 (set_x_learned)
    int set_x_zero() override {
        HIPTRY(hipMemsetAsync(x, 0, n0 * sizeof(double), s));
        x_zero = true;
        xz_device = false;
        return TH_OK;
    }
    int set_x(const double* host) override {
        x_zero = false;
        xz_device = false;
        return upload_split<double, double>(g0, g0.nz, host, x);
    }
    // b = P (pcg3d_gpu.cu's copy) without the copy: until b is next written,
    // the kernels that read the right-hand side read P itself (rhs()), the
    // start from a graph of its own. set_power keeps the copy's meaning.
    int set_rhs_steady() override {
        rhs_P = true;
        return TH_OK;
    }
    int set_rhs_transient() override {
        rhs_P = false;
        k_rhs_transient<T><<<vec_blocks(n0), kVecThreads, 0, s>>>(n0, L[0].cap, P, x, b, st);
        HIPTRY(hipGetLastError());
        return TH_OK;
    }
    int set_power(const double* host) override {
        if (rhs_P) {  // b keeps the power it was set from
            HIPTRY(hipMemcpyAsync(b, P, n0 * sizeof(double), hipMemcpyDeviceToDevice, s));
            rhs_P = false;
        }
        HIPTRY(hipStreamSynchronize(s));
        return upload_split<double, double>(g0, g0.nz, host, P);
    }
    int device_x(const double** out) override {
        k_to_natural<<<lin_blocks(nnat), 256, 0, s>>>(g0, x, xnat);
        HIPTRY(hipGetLastError());
        HIPTRY(hipStreamSynchronize(s));
        *out = xnat;
        return TH_OK;
    }
    int get_x(double* host) override {
        k_to_natural<<<lin_blocks(nnat), 256, 0, s>>>(g0, x, xnat);
        HIPTRY(hipGetLastError());
        HIPTRY(hipMemcpyAsync(host, xnat, nnat * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        return TH_OK;
    }

    // ---- launch sequences -------------------------------------------------

    size_t tail_k0 = 0;     // first level handled by the single-workgroup tail
    bool fast = true;       // lateral (fast) restriction
    bool lds_tail = false;  // tail levels held in LDS
    size_t lds_bytes = 0;

    // Partial buffers, one per producing kernel, so a consumer can still read
    // its producer's partials while the next reduction runs.
    double *part_res = nullptr, *part_apply = nullptr, *part_upd = nullptr, *part_dot = nullptr;
    unsigned nblk_stencil = 0, nblk_apply = 0, nblk_res = 0, nblk_vec = 0;

    template <int NV, int EPI, unsigned KMAX = 0>
    int finalize(unsigned nblk, const double* pp = nullptr) {
        k_finalize<NV, EPI, KMAX><<<1, kFinThreads, 0, s>>>(pp ? pp : part, (int)nblk, st);
        HIPTRY(hipGetLastError());
        return TH_OK;
    }

    // Consumer of the previous kernel's partials, carried by the first kernel
    // of the V-cycle (see consume_convergence).
    struct Cons {
        const double* part;
        int nblk;
        int mode;
    };

    // Levels with at most kPPCols columns per colour run the plane-parallel
    // sweep (k_sweep_pp); larger ones the column walk (k_sweep).
    // PRE: the pre-smoothing plain sweep (its inputs are not in the caches).
    template <int MODE, bool PRE = false>
    int launch_sweep(size_t k, int colour, Cons cs = {nullptr, 0, CONS_NONE}) {
        const DevLevel<T>& d = L[k];
        const DevLevel<T>& c = L[std::min(k + 1, L.size() - 1)];
        const long cols = (long)d.g.hx * d.g.ny;
        if (cols <= kPPCols && d.g.nz <= 16) {
            const unsigned nb = (unsigned)((cols * pp_lanes(d.g.nz) + kPPThreads - 1) / kPPThreads);
#define TH_PP(LP)                                                                                                   \
    k_sweep_pp<T, LP, MODE><<<nb, kPPThreads, 0, s>>>(d.g, d.ops(), d.finv, d.b, d.x, colour, c.g, c.x, st, cs.part, \
                                                      cs.nblk, cs.mode)
            if (d.g.nz <= 4) TH_PP(4);
            else if (d.g.nz <= 8) TH_PP(8);
            else TH_PP(16);
#undef TH_PP
            HIPTRY(hipGetLastError());
            return TH_OK;
        }
        const dim3 grid = col_grid(d.g.hx, d.g.ny);
        if (d.ntypes) {
            // FP32 prolongation and pre-smoothing plain sweeps up to 16 planes:
            // the type-independent loads first; FP32 prolongation past that:
            // three coarse loads and an occupancy floor
#define TH_SWEEP(NZ)                                                                                                  \
    do {                                                                                                              \
        if constexpr ((MODE == SW_PROLONG || (MODE == SW_PLAIN && PRE)) && std::is_same<T, float>::value &&           \
                      NZ <= 16)                                                                                       \
            k_sweep_typed_lf<T, NZ, MODE><<<grid, col_block(), 0, s>>>(d.g, d.typed(), d.b, d.x, colour, c.g, c.x,     \
                                                                       st, cs.part, cs.nblk, cs.mode);                \
        else if constexpr (MODE == SW_PROLONG && std::is_same<T, float>::value)                                       \
            k_sweep_typed_prolong3<T, NZ><<<grid, col_block(), 0, s>>>(d.g, d.typed(), d.b, d.x, colour, c.g, c.x, st, \
                                                                       cs.part, cs.nblk, cs.mode);                    \
        else                                                                                                          \
            k_sweep_typed<T, NZ, MODE><<<grid, col_block(), 0, s>>>(d.g, d.typed(), d.b, d.x, colour, c.g, c.x, st,    \
                                                                    cs.part, cs.nblk, cs.mode);                       \
    } while (0)
            if (d.g.nz <= 4) TH_SWEEP(4);
            else if (d.g.nz <= 8) TH_SWEEP(8);
            else if (d.g.nz <= 12) TH_SWEEP(12);
            else if (d.g.nz <= 16) TH_SWEEP(16);
            else if (d.g.nz <= 24) TH_SWEEP(24);
            else if (d.g.nz <= 32) TH_SWEEP(32);
            else if (d.g.nz <= 48) TH_SWEEP(48);
            else TH_SWEEP(64);
#undef TH_SWEEP
            HIPTRY(hipGetLastError());
            return TH_OK;
        }
#define TH_SWEEP(NZ)                                                                                         \
    k_sweep<T, NZ, MODE><<<grid, col_block(), 0, s>>>(d.g, d.ops(), d.finv, d.b, d.x, colour, c.g, c.x, st, \
                                                      cs.part, cs.nblk, cs.mode)
        if (d.g.nz <= 4) TH_SWEEP(4);
        else if (d.g.nz <= 8) TH_SWEEP(8);
        else if (d.g.nz <= 12) TH_SWEEP(12);
        else if (d.g.nz <= 16) TH_SWEEP(16);
        else if (d.g.nz <= 24) TH_SWEEP(24);
        else if (d.g.nz <= 32) TH_SWEEP(32);
        else if (d.g.nz <= 48) TH_SWEEP(48);
        else TH_SWEEP(64);
#undef TH_SWEEP
        HIPTRY(hipGetLastError());
        return TH_OK;
    }

    int launch_tail(Cons cs) {
        Tail<T> t{};
        t.n = (int)(L.size() - tail_k0);
        for (int k = 0; k < t.n; ++k) {
            const DevLevel<T>& d = L[tail_k0 + k];
            t.L[k] = TailLevel<T>{d.g, d.ops(), d.finv, d.x, d.b};
        }
        const int nz = g0.nz;
        if (lds_tail) {
#define TH_TAIL(NZ) k_tail_lds<T, NZ><<<1, kTailLdsThreads, lds_bytes, s>>>(t, st, cs.part, cs.nblk, cs.mode)
            if (nz <= 4) TH_TAIL(4);
            else if (nz <= 8) TH_TAIL(8);
            else if (nz <= 12) TH_TAIL(12);
            else TH_TAIL(16);
#undef TH_TAIL
            HIPTRY(hipGetLastError());
            return TH_OK;
        }
#define TH_TAIL(NZ) \
    k_tail<T, NZ><<<1, tail_threads<T, NZ>(), 0, s>>>(t, st, cs.part, cs.nblk, cs.mode, fast ? 1 : 0)
        if (nz <= 4) TH_TAIL(4);
        else if (nz <= 8) TH_TAIL(8);
        else if (nz <= 12) TH_TAIL(12);
        else if (nz <= 16) TH_TAIL(16);
        else if (nz <= 24) TH_TAIL(24);
        else if (nz <= 32) TH_TAIL(32);
        else if (nz <= 48) TH_TAIL(48);
        else TH_TAIL(64);
#undef TH_TAIL
        HIPTRY(hipGetLastError());
        return TH_OK;
    }

    // Level c (lateral restriction, plane-parallel, above the tail) gets its
    // restricted b and its first sweep from one launch (k_restrict_zero_pp).
    bool fuse_restrict_zero(size_t c) const {
        const Geo& g = L[c].g;
        return fast && c != tail_k0 && (long)g.hx * g.ny <= kPPCols && g.nz <= 16;
    }
    int launch_restrict_zero(size_t k) {
        const DevLevel<T>& f = L[k];
        const DevLevel<T>& c = L[k + 1];
        const long cols2 = 2L * c.g.hx * c.g.ny;
        const unsigned nb = (unsigned)((cols2 * pp_lanes(c.g.nz) + kPPThreads - 1) / kPPThreads);
#define TH_RZ(LP)                                                                                                  \
    do {                                                                                                           \
        if (f.ntypes)                                                                                              \
            k_restrict_zero_pp<T, LP, true><<<nb, kPPThreads, 0, s>>>(f.g, f.ops(), f.typed(), f.x, c.g, c.ops(),  \
                                                                      c.finv, c.b, c.x, st);                       \
        else                                                                                                       \
            k_restrict_zero_pp<T, LP, false><<<nb, kPPThreads, 0, s>>>(f.g, f.ops(), f.typed(), f.x, c.g, c.ops(), \
                                                                       c.finv, c.b, c.x, st);                      \
    } while (0)
        if (c.g.nz <= 4) TH_RZ(4);
        else if (c.g.nz <= 8) TH_RZ(8);
        else TH_RZ(16);
#undef TH_RZ
        HIPTRY(hipGetLastError());
        return TH_OK;
    }

    int launch_restrict(size_t k) {
        const DevLevel<T>& f = L[k];
        const DevLevel<T>& c = L[k + 1];
        dim3 grid = col_grid(c.g.nx, c.g.ny);
        grid.z = c.g.nz;
        if (fast && f.ntypes) {
            grid.z = (c.g.nz + kRestrictPlanes - 1) / kRestrictPlanes;
            k_restrict_fast_typed<T><<<grid, col_block(), 0, s>>>(f.g, f.typed(), f.x, c.g, c.b, st);
        } else if (fast)
            k_restrict_fast<T><<<grid, col_block(), 0, s>>>(f.g, f.ops(), f.x, c.g, c.b, st);
        else
            k_restrict<T><<<grid, col_block(), 0, s>>>(f.g, f.ops(), f.b, f.x, c.g, c.b, st);
        HIPTRY(hipGetLastError());
        return TH_OK;
    }

    // One V-cycle on level k (pcg3d_cpu.cpp:69-113): pre-smooth red, black;
    // restrict; coarse V-cycle; post-smooth black (with the prolongation),
    // red. Level k's x starts from zero. Levels from tail_k0 down run in the
    // single-workgroup tail kernel. `cs`: convergence test carried by the
    // V-cycle's first kernel (level 0 only).
    // zeroed: level k's first sweep was run by the restriction into it.
    int vcycle(size_t k, Cons cs = {nullptr, 0, CONS_NONE}, bool zeroed = false) {
        int rc;
        if (k == tail_k0) return launch_tail(cs);
        if (!zeroed && (rc = launch_sweep<SW_ZERO>(k, 0, cs))) return rc;
        if ((rc = launch_sweep<SW_PLAIN, true>(k, 1))) return rc;
        const bool fz = fuse_restrict_zero(k + 1);
        if ((rc = fz ? launch_restrict_zero(k) : launch_restrict(k))) return rc;
        if ((rc = vcycle(k + 1, Cons{nullptr, 0, CONS_NONE}, fz))) return rc;
        if ((rc = launch_sweep<SW_PROLONG>(k, 1))) return rc;
        return launch_sweep<SW_PLAIN>(k, 0);
    }

    // Ap = (A + C/dt) v with partials v.Ap in part_apply.
    template <class PT>
    int launch_apply_kernel(const PT* v) {
        const dim3 grid = stencil_grid(2 * g0.hx, g0.ny);
        const Ops<T> o = L[0].ops();
        if (L[0].ntypes) {
            // 64 x 4 blocks (nblk_apply partials); FP32 faces with an occupancy floor
#define TH_APPLY(NZ)                                                                                                  \
    do {                                                                                                              \
        const dim3 ag = col_grid(2 * g0.hx, g0.ny);                                                                   \
        if constexpr (std::is_same<T, float>::value)                                                                  \
            k_apply_typed_occ<T, NZ, PT><<<ag, col_block(), 0, s>>>(g0, L[0].typed(), v, Ap, st, part_apply);         \
        else                                                                                                          \
            k_apply_typed<T, NZ, PT><<<ag, col_block(), 0, s>>>(g0, L[0].typed(), v, Ap, st, part_apply);             \
    } while (0)
            if (g0.nz <= 4) TH_APPLY(4);
            else if (g0.nz <= 8) TH_APPLY(8);
            else if (g0.nz <= 12) TH_APPLY(12);
            else if (g0.nz <= 16) TH_APPLY(16);
            else if (g0.nz <= 24) TH_APPLY(24);
            else if (g0.nz <= 32) TH_APPLY(32);
            else if (g0.nz <= 48) TH_APPLY(48);
            else TH_APPLY(64);
#undef TH_APPLY
            HIPTRY(hipGetLastError());
            return TH_OK;
        }
#define TH_APPLY(NZ) k_apply<T, NZ, PT><<<grid, stencil_block(), 0, s>>>(g0, o, v, Ap, st, part_apply)
        if (g0.nz <= 4) TH_APPLY(4);
        else if (g0.nz <= 8) TH_APPLY(8);
        else if (g0.nz <= 12) TH_APPLY(12);
        else if (g0.nz <= 16) TH_APPLY(16);
        else if (g0.nz <= 24) TH_APPLY(24);
        else if (g0.nz <= 32) TH_APPLY(32);
        else if (g0.nz <= 48) TH_APPLY(48);
        else TH_APPLY(64);
#undef TH_APPLY
        HIPTRY(hipGetLastError());
        return TH_OK;
    }
    int launch_apply() {  // Ap = A p with alpha in the state (Jacobi path and tests)
        int rc;
        if ((rc = launch_apply_kernel(p))) return rc;
        return finalize<1, FIN_ALPHA>(nblk_apply, part_apply);
    }

    // r = b - (A + C/dt) x with b0 = (T) r (RES_MG), or the steady ||P - A v||
    // (RES_CHECK, no writes), nblk_res partials into pp: the typed column
    // walk (64 x 4 blocks) when the fine level has column types.
    // RES_RHS: bb is P and the step's b = P + C/dt x is written to bw.
    template <int MODE>
    int launch_residual(const double* bb, const double* xx, double* pp, double* bw = nullptr) {
        const dim3 grid = stencil_grid(2 * g0.hx, g0.ny);
        double* rr = MODE == RES_MG || MODE == RES_RHS ? r : nullptr;
        T* b0 = MODE == RES_MG || MODE == RES_RHS ? L[0].b : nullptr;
        if (L[0].ntypes) {
            const dim3 tg = col_grid(2 * g0.hx, g0.ny);
#define TH_RES(NZ) \
    k_residual_typed<T, NZ, MODE><<<tg, col_block(), 0, s>>>(g0, L[0].typed(), bb, xx, rr, b0, st, pp, bw)
            if (g0.nz <= 4) TH_RES(4);
            else if (g0.nz <= 8) TH_RES(8);
            else if (g0.nz <= 12) TH_RES(12);
            else if (g0.nz <= 16) TH_RES(16);
            else if (g0.nz <= 24) TH_RES(24);
            else if (g0.nz <= 32) TH_RES(32);
            else if (g0.nz <= 48) TH_RES(48);
            else TH_RES(64);
#undef TH_RES
        } else {
            k_residual<T, MODE><<<grid, stencil_block(), 0, s>>>(g0, L[0].ops(), bb, xx, rr, b0, nullptr, nullptr,
                                                                 nullptr, st, pp, bw);
        }
        HIPTRY(hipGetLastError());
        return TH_OK;
    }

    // Start: residual, first preconditioner application, rz, p = z. rhs: a
    // backward-Euler step's b = P + C/dt x formed by the residual kernel
    // itself (multigrid).
    // from_P: the right-hand side is P (set_rhs_steady).
    int enqueue_start(bool rhs = false, bool from_P = false) {
        const dim3 grid = stencil_grid(2 * g0.hx, g0.ny);
        const unsigned vb = vec_blocks(n0);
        const double* bb = from_P ? P : b;
        int rc;
        if (mg()) {
            if ((rc = rhs ? launch_residual<RES_RHS>(P, x, part_res, b) : launch_residual<RES_MG>(bb, x, part_res)))
                return rc;
            if ((rc = vcycle(0, Cons{part_res, (int)nblk_res, CONS_START}))) return rc;
            k_dot<T><<<vb, kVecThreads, 0, s>>>(n0, r, L[0].x, st, part_dot);
            HIPTRY(hipGetLastError());
            k_pupdate<T, 1, T><<<vb, kVecThreads, 0, s>>>(n0, L[0].x, pm, nullptr, st, part_dot, (int)vb);
            HIPTRY(hipGetLastError());
            return TH_OK;
        }
        k_residual<T, RES_JACOBI><<<grid, stencil_block(), 0, s>>>(g0, L[0].ops(), bb, x, r, nullptr, invd, z, p, st,
                                                                   part_res, nullptr);
        HIPTRY(hipGetLastError());
        return finalize<3, FIN_RES_JACOBI>(nblk_stencil, part_res);
    }

    // Multigrid iteration, first part: Ap; x, r update (alpha formed from the
    // apply partials).
    int enqueue_apply_update() {
        const unsigned vb = vec_blocks(n0);
        int rc;
        if ((rc = launch_apply_kernel(pm))) return rc;
        if (opt.beta == TH_BETA_FLEXIBLE)
            k_update<T, UP_MG_FLEX><<<vb, kVecThreads, 0, s>>>(n0, nullptr, Ap, x, r, L[0].b, L[0].x, nullptr, nullptr,
                                                               st, part_upd, part_apply, (int)nblk_apply);
        else
            k_update<T, UP_MG><<<vb, kVecThreads, 0, s>>>(n0, nullptr, Ap, x, r, L[0].b, nullptr, nullptr, nullptr, st,
                                                          part_upd, part_apply, (int)nblk_apply);
        HIPTRY(hipGetLastError());
        return TH_OK;
    }
    // Second part: z = M^-1 r (V-cycle, its head consuming the update's
    // partials: the convergence test, unless k_consume_iter ran it); r.z;
    // p = z + beta p (beta formed from the dot partials) with the lagged
    // x += alpha p.
    int enqueue_precondition(Cons cs) {
        const unsigned vb = vec_blocks(n0);
        int rc;
        if ((rc = vcycle(0, cs))) return rc;
        k_dot<T><<<vb, kVecThreads, 0, s>>>(n0, r, L[0].x, st, part_dot);
        HIPTRY(hipGetLastError());
        k_pupdate<T, 0, T><<<vb, kVecThreads, 0, s>>>(n0, L[0].x, pm, x, st, part_dot, (int)vb);
        HIPTRY(hipGetLastError());
        return TH_OK;
    }
    int enqueue_iter_a() {
        int rc;
        if ((rc = enqueue_apply_update())) return rc;
        k_consume_iter<<<1, 256, 0, s>>>(part_upd, (int)vec_blocks(n0), st);
        HIPTRY(hipGetLastError());
        return TH_OK;
    }
    int enqueue_iter_b() { return enqueue_precondition(Cons{nullptr, 0, CONS_NONE}); }

    // One PCG iteration: Ap; x, r update (alpha formed from the apply
    // partials); convergence test + z = M^-1 r (V-cycle); r.z; p = z + beta p
    // (beta formed from the dot partials). Multigrid: 3 + V-cycle launches.
    int enqueue_iter() {
        const unsigned vb = vec_blocks(n0);
        int rc;
        if (mg()) {
            if ((rc = enqueue_apply_update())) return rc;
            return enqueue_precondition(Cons{part_upd, (int)vb, CONS_ITER});
        }
        if ((rc = launch_apply())) return rc;
        k_update<T, UP_JACOBI><<<vb, kVecThreads, 0, s>>>(n0, p, Ap, x, r, nullptr, nullptr, invd, z, st, part_upd,
                                                          nullptr, 0);
        HIPTRY(hipGetLastError());
        if ((rc = finalize<2, FIN_UPDATE_JACOBI>(vb, part_upd))) return rc;
        k_pupdate<double, 0, double><<<vb, kVecThreads, 0, s>>>(n0, z, p, nullptr, st, nullptr, 0);
        HIPTRY(hipGetLastError());
        return TH_OK;
    }

    int fetch_state() {
        HIPTRY(hipMemcpyAsync(h_st, st, sizeof(State), hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        return TH_OK;
    }

    int run_start() {
        mid_iter = false;
        if (hipGraphExec_t gs = rhs_P ? g_start_P : g_start) {
            HIPTRY(hipGraphLaunch(gs, s));
            return TH_OK;
        }
        return enqueue_start(false, rhs_P);
    }
    int run_iter() {
        if (g_iter) {
            HIPTRY(hipGraphLaunch(g_iter, s));
            return TH_OK;
        }
        return enqueue_iter();
    }
    // The last iteration of a queued batch: its first part only where the
    // iteration splits (the host then looks at the state), else whole.
    int run_iter_last() {
        if (!split_iter) return run_iter();
        mid_iter = true;
        if (g_iter_a) {
            HIPTRY(hipGraphLaunch(g_iter_a, s));
            return TH_OK;
        }
        return enqueue_iter_a();
    }
    // The second part of an iteration run_iter_last left open.
    int finish_iter() {
        if (!mid_iter) return TH_OK;
        mid_iter = false;
        if (g_iter_b) {
            HIPTRY(hipGraphLaunch(g_iter_b, s));
            return TH_OK;
        }
        return enqueue_iter_b();
    }

    // Residual replacement: r := b - A x, the Krylov state (p, rz, beta)
    // kept. The recursive residual drifts from the true one by rounding in
    // x += alpha p and r -= alpha Ap, most of it while x still moves a lot;
    // at 1024^2 x 16 the true relative residual of a solve stalls near 1e-10
    // however far the recursive one goes, and the true-residual check then
    // restarts the solve (a start V-cycle and the Krylov history lost).
    // Replacing r once, when the recursive relative residual first falls
    // below kReplaceAt * tol, leaves a gap far below tol.
    static constexpr double kReplaceAt = 1e4;
    int replace_residual() {
        k_clear_xz<<<1, 1, 0, s>>>(st);  // x is no longer 0
        HIPTRY(hipGetLastError());
        return launch_residual<RES_MG>(rhs(), x, part_res);
    }

    // PCG iterations until the device reports done. Iterations are queued in
    // batches sized from the observed convergence rate, so the host waits
    // once per batch rather than once per iteration; iterations queued past
    // convergence return at once. `first`: size of the first batch.
    // `replace`: one residual replacement on the way (multigrid solves with
    // the true-residual check, tol < 1e-8); the batch before it stops where
    // the replacement is due. After an iteration x is complete (the p update
    // applies the lagged x += alpha p).
    // fresh: the solve's start was queued without the host looking at it
    // (h_st holds k_begin's fields only): the first batch's rate is taken
    // from the start's residual, relres0.
    int iterate(int maxit, int first = 1, bool replace = false, bool fresh = false) {
        int rc;
        double rate = 0.0;  // relres reduction per iteration, once known
        replace = replace && opt.true_residual && mg() && h_st->tol < 1e-8;
        bool ran = false;  // a batch of this call has run (the replacement waits for one)
        while (!h_st->done && h_st->iters < maxit) {
            if ((rc = finish_iter())) return rc;  // x complete again
            if (ran && replace && h_st->relres < kReplaceAt * h_st->tol) {
                if ((rc = replace_residual())) return rc;
                replace = false;
            }
            int k = first;
            first = 1;
            if (rate > 0.0 && rate < 1.0 && h_st->relres > 0.0) {
                const double need = std::log(h_st->tol / h_st->relres) / std::log(rate);
                k = (int)std::max(1.0, std::min(16.0, std::ceil(need) - 1.0));
                if (replace && h_st->relres >= kReplaceAt * h_st->tol) {
                    const double nr = std::log(kReplaceAt * h_st->tol / h_st->relres) / std::log(rate);
                    k = std::min(k, (int)std::max(1.0, std::ceil(nr)));
                }
            }
            k = std::min(k, maxit - h_st->iters);
            const int it0 = h_st->iters;
            double rel0 = h_st->relres;
            for (int q = 0; q < k - 1; ++q)
                if ((rc = run_iter())) return rc;
            if ((rc = run_iter_last())) return rc;
            if ((rc = fetch_state())) return rc;
            if (fresh) rel0 = h_st->relres0;
            fresh = false;
            const int did = h_st->iters - it0;
            if (did > 0 && rel0 > 0.0 && h_st->relres > 0.0) rate = std::pow(h_st->relres / rel0, 1.0 / did);
            if (did == 0) break;  // nothing ran: the device is done
            ran = true;
        }
        mid_iter = false;  // done (k_finish applies the lagged x) or out of iterations
        return TH_OK;
    }

    int solve(double tol, int maxit, th_stats* out) override {
        int rc;
        const int xz = xz_device ? -1 : x_zero ? 1 : 0;  // 1: the start residual is b
        x_zero = false;
        xz_device = false;
        k_begin<<<1, 1, 0, s>>>(st, tol, maxit, xz);
        HIPTRY(hipGetLastError());
        if ((rc = run_start())) return rc;
        // The first iteration is queued behind the start without a look at
        // it (a start that already meets tol leaves it returning at once);
        // the host's mirror gets k_begin's fields meanwhile, and the batches
        // that follow are sized as before from the start's residual and the
        // first iteration's (the batch ends decide where the residual
        // replacement falls: a first batch of two moved it and cost an
        // iteration at 1024^2 x 16).
        h_st->tol = tol;
        h_st->maxit = maxit;
        h_st->iters = 0;
        h_st->done = 0;
        h_st->converged = 0;
        h_st->zero_x = 0;
        if ((rc = converge(maxit, 1, true))) return rc;
        k_finish<T><<<vec_blocks(n0), kVecThreads, 0, s>>>(n0, x, pm, st);
        HIPTRY(hipGetLastError());
        if (out) {
            out->iters = h_st->iters;
            out->relres = h_st->relres;
            out->converged = h_st->converged;
        }
        return TH_OK;
    }

    // Iterations after the start of a solve (h_st fresh), then the optional
    // true-residual check: the recurrence ||r|| drifts from ||b - A x|| by
    // rounding; restart from x with the true residual until that one is
    // below tol too.
    int converge(int maxit, int first, bool fresh = false) {
        int rc;
        if (maxit > 0 && (rc = iterate(maxit, first, true, fresh))) return rc;
        if (maxit <= 0 && fresh && (rc = fetch_state())) return rc;  // the start's own state
        for (int restart = 0; opt.true_residual && mg() && h_st->converged && !h_st->zero_x && restart < 4; ++restart) {
            k_finish<T><<<vec_blocks(n0), kVecThreads, 0, s>>>(n0, x, pm, st);  // x complete
            HIPTRY(hipGetLastError());
            k_reopen<<<1, 1, 0, s>>>(st);
            HIPTRY(hipGetLastError());
            // The true residual of x and the convergence test on it, in one
            // kernel and a one-block finalize: when it is below tol the start
            // graph (a V-cycle's worth of launches) is not replayed. The
            // kernel writes nothing (RES_TRUE): the start graph of a restart
            // computes r and the V-cycle's b itself. (With RES_MG's writes
            // the test took 135 us at 1024^2 x 16, without them 102.)
            if ((rc = launch_residual<RES_TRUE>(rhs(), x, part_res)) || (rc = finalize<3, FIN_RES>(nblk_res, part_res)) ||
                (rc = fetch_state()))
                return rc;
            if (h_st->done) break;
            if ((rc = run_start()) || (rc = fetch_state())) return rc;
            if (h_st->done) break;  // true residual below tol
            if (h_st->iters >= maxit) break;
            if ((rc = iterate(maxit))) return rc;
        }
        return TH_OK;
    }

    // ---- transient (transient3d.h:48-90) -----------------------------------
    // A step is three launch sequences, each replayed as one graph:
    //   head: advance dt on the device, new factors if dt changed, b = P +
    //         C/dt x, start of the solve warm-started from x;
    //   iterations, queued ahead as in a steady solve;
    //   tail: energy terms of the new x and the step's record.
    // The host keeps its own copy of t and dt (the same double arithmetic
    // as the device) to know when to stop, and waits once per step.
    double* xprev = nullptr;  // T_older for the step predictor (allocated on first use)
    double* guide = nullptr;  // guide field of the learned step start (set_guide*)
    // predictor: 0 T_old (the reference's warm start), 1 extrapolated, 2
    // learned (T_old moved towards the guide field by the residual-minimising
    // w >= 0), 3 extrapolated, then learned.
    int enqueue_step_head(double tol, int maxit, int predictor) {
        k_step_begin<<<1, 1, 0, s>>>(st, tol, maxit);
        HIPTRY(hipGetLastError());
        int rc;
        if ((rc = enqueue_factors())) return rc;
        // the plain warm start: b = P + C/dt x inside the start's residual
        // kernel (one pass over x and P less than k_rhs_transient + residual)
        if (predictor == 0 && mg()) return enqueue_start(true);
        if ((rc = set_rhs_transient())) return rc;
        if (predictor == 1 || predictor == 3) {
            k_predict<<<vec_blocks(n0), kVecThreads, 0, s>>>(n0, x, xprev, st);
            HIPTRY(hipGetLastError());
        }
        if (predictor >= 2) {
            const dim3 grid = stencil_grid(2 * g0.hx, g0.ny);
            k_guide_terms<T><<<grid, stencil_block(), 0, s>>>(g0, L[0].ops(), b, x, guide, st, part);
            HIPTRY(hipGetLastError());
            k_guide_weight<<<1, kFinThreads, 0, s>>>(part, (int)(grid.x * grid.y), st);
            HIPTRY(hipGetLastError());
            k_guide_step<<<vec_blocks(n0), kVecThreads, 0, s>>>(n0, x, guide, st);
            HIPTRY(hipGetLastError());
        }
        return enqueue_start();
    }
    int set_guide(const double* host) override {
        int rc;
        if (!guide && (rc = dalloc(&guide, n0))) return rc;
        return upload_split<double, double>(g0, g0.nz, host, guide);
    }
    int guide_stats(int* steps, double* w_mean) override {
        int rc;
        if ((rc = fetch_state())) return rc;
        if (steps) *steps = h_st->guide_steps;
        if (w_mean) *w_mean = h_st->guide_steps ? h_st->guide_w_sum / h_st->guide_steps : 0.0;
        return TH_OK;
    }
    int enqueue_step_tail() {
        const unsigned vb = vec_blocks(n0);
        k_finish_energy<T, T><<<vb, kVecThreads, 0, s>>>(g0, L[0].ops(), x, pm, st, part);
        HIPTRY(hipGetLastError());
        int rc;
        if ((rc = finalize<3, FIN_ENERGY, 4u>(vb))) return rc;
        k_step_end<<<1, 1, 0, s>>>(st);
        HIPTRY(hipGetLastError());
        return TH_OK;
    }
    int run_transient(const th_transient_opts& o, th_step_record* recs, int rec_cap, int* nsteps, int* all_converged,
                      double* ms) override {
        static_assert(sizeof(StepRec) == sizeof(th_step_record), "step record layout");
        using clk = std::chrono::high_resolution_clock;
        int rc;
        hipGraphExec_t gh = nullptr, gt = nullptr;
        rhs_P = false;  // every step's head writes b
        auto cleanup = [&]() {
            if (gh) (void)hipGraphExecDestroy(gh);
            if (gt) (void)hipGraphExecDestroy(gt);
        };
        if ((o.predictor == 1 || o.predictor == 3) && !xprev && (rc = dalloc(&xprev, n0))) return rc;
        if (o.predictor >= 2 && !guide) return fail(TH_ERR_ARG, "learned step start: no guide field set");
        if (opt.use_graphs && mg()) {
            if ((rc = capture_fn(&gh, [&]() { return enqueue_step_head(o.tol, o.maxit, o.predictor); })) ||
                (rc = capture_fn(&gt, [&]() { return enqueue_step_tail(); }))) {
                cleanup();
                return rc;
            }
        }
        auto body = [&]() -> int {
            if ((rc = set_x_zero())) return rc;
            k_step_init<<<1, 1, 0, s>>>(st, o.dt0, o.growth, o.dt_max, o.t_end);
            HIPTRY(hipGetLastError());
            HIPTRY(hipStreamSynchronize(s));
            const auto t0 = clk::now();
            double t = 0.0, dt = o.dt0;
            int steps = 0, conv = 1, prev_iters = 2;
            auto take_record = [&](int k) -> int {  // h_st->last is the record of step k
                if (!recs || k < 0) return TH_OK;
                if (k >= rec_cap) return fail(TH_ERR_ARG, "record buffer too small");
                std::memcpy(&recs[k], &h_st->last, sizeof(StepRec));
                return TH_OK;
            };
            while (t < o.t_end * (1 - 1e-12) && (o.max_steps <= 0 || steps < o.max_steps)) {
                dt = std::min(dt, o.t_end - t);
                if (gh) HIPTRY(hipGraphLaunch(gh, s));
                else if ((rc = enqueue_step_head(o.tol, o.maxit, o.predictor))) return rc;
                const int nq = std::min(prev_iters, o.maxit);
                for (int q = 0; q < nq - 1; ++q)
                    if ((rc = run_iter())) return rc;
                if (nq > 0 && (rc = run_iter_last())) return rc;
                if ((rc = fetch_state())) return rc;
                if ((rc = take_record(steps - 1))) return rc;
                if ((rc = converge(o.maxit, 1))) return rc;
                if (!h_st->converged) conv = 0;
                prev_iters = std::max(1, h_st->iters);
                if (gt) HIPTRY(hipGraphLaunch(gt, s));
                else if ((rc = enqueue_step_tail())) return rc;
                t += dt;
                dt = std::min(dt * o.growth, o.dt_max);
                ++steps;
            }
            if ((rc = fetch_state())) return rc;
            if ((rc = take_record(steps - 1))) return rc;
            if (ms) *ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
            *nsteps = steps;
            if (all_converged) *all_converged = conv;
            return TH_OK;
        };
        rc = body();
        x_zero = false;
        cleanup();
        return rc;
    }

    // Initial guess from a coarse field with the plain warm start as the
    // fallback (th_solver_set_x_lifted). Ap holds the plain x meanwhile.
    int set_x_lifted(const float* field, int Hc, int Wc, int C, const LiftMap& zm, double scale, int passes,
                     int* used, double* rp, double* rl) override {
        int rc;
        double r_plain = 1.0, r_lift = 0.0;  // x = 0: P - A x = P
        if (!x_zero && (rc = relres_of(x, &r_plain))) return rc;
        // The lifted guess is built and judged in Ap (free between solves):
        // x is only overwritten when the guess is kept.
        double* v = Ap;
        if ((rc = lift_into(v, field, Hc, Wc, C, zm, scale, passes))) return rc;
        if ((rc = relres_of(v, &r_lift))) return rc;
        const bool keep = r_lift < r_plain;
        if (keep) {
            HIPTRY(hipMemcpyAsync(x, v, n0 * sizeof(double), hipMemcpyDeviceToDevice, s));
            x_zero = false;
            xz_device = false;
        }
        HIPTRY(hipStreamSynchronize(s));
        if (used) *used = keep;
        if (rp) *rp = r_plain;
        if (rl) *rl = r_lift;
        return TH_OK;
    }
    // The guide field of the learned step start from a coarse field, lifted
    // as set_x_lifted lifts its guess (no judging: each step judges its w).
    int set_guide_lifted(const float* field, int Hc, int Wc, int C, const LiftMap& zm, double scale,
                         int passes) override {
        int rc;
        if (!guide && (rc = dalloc(&guide, n0))) return rc;
        if ((rc = lift_into(guide, field, Hc, Wc, C, zm, scale, passes))) return rc;
        HIPTRY(hipStreamSynchronize(s));
        return TH_OK;
    }
    // The learned steady start in one asynchronous call: the surrogate's
    // forward pass on this solver's stream, the lift and set_x_lifted's
    // fallback judged on the device (k_lift_plain, k_lift_decide,
    // k_copy_if_kept), so nothing waits for the host before the solve; the
    // judgement is read afterwards (lifted_stats).
    int set_x_learned(th_surrogate* sur, const float* d_in, const LiftMap& zm, double scale, int passes) override {
        int rc;
        const float* field = nullptr;
        int Hc = 0, Wc = 0, C = 0;
        if ((rc = surrogate_forward_on(sur, d_in, (void*)s, opt.device, &field, &Hc, &Wc, &C))) return rc;
        const int xz0 = x_zero ? 1 : 0;
        if (!x_zero && ((rc = launch_residual<RES_CHECK>(P, x, part)) || (rc = finalize<3, FIN_CHECK>(nblk_res))))
            return rc;
        k_lift_plain<<<1, 1, 0, s>>>(st, xz0);
        HIPTRY(hipGetLastError());
        double* v = Ap;  // free between solves
        if ((rc = lift_into(v, field, Hc, Wc, C, zm, scale, passes))) return rc;
        if ((rc = launch_residual<RES_CHECK>(P, v, part)) || (rc = finalize<3, FIN_CHECK>(nblk_res))) return rc;
        k_lift_decide<<<1, 1, 0, s>>>(st, xz0);
        HIPTRY(hipGetLastError());
        k_copy_if_kept<<<vec_blocks(n0), kVecThreads, 0, s>>>(n0, v, x, st);
        HIPTRY(hipGetLastError());
        x_zero = false;
        xz_device = true;
        return TH_OK;
    }
    int lifted_stats(int* used, double* rp, double* rl) override {
        int rc;
        if ((rc = fetch_state())) return rc;
        if (used) *used = h_st->lift_used;
        if (rp) *rp = h_st->lift_rp;
        if (rl) *rl = h_st->lift_rl;
        return TH_OK;
    }
    // v = the coarse field lifted to the fine grid (its planes first taken to
    // the solver's planes on the coarse grid, k_lift_z, then bilinear over the
    // die, k_lift), then `passes` red-black column line relaxations of the
    // steady operator: with passes > 0 only the black columns are lifted,
    // since the first (red) relaxation overwrites every red column.
    double* lift_fz = nullptr;
    size_t lift_fz_n = 0;
    int lift_into(double* v, const float* field, int Hc, int Wc, int C, const LiftMap& zm, double scale, int passes) {
        int rc;
        const size_t nfz = (size_t)Hc * Wc * g0.nz;
        if (nfz > lift_fz_n) {
            if ((rc = dalloc(&lift_fz, nfz))) return rc;
            lift_fz_n = nfz;
        }
        k_lift_z<<<(unsigned)((nfz + 255) / 256), 256, 0, s>>>(field, Hc, Wc, C, g0.nz, zm, lift_fz);
        HIPTRY(hipGetLastError());
        const int colour = passes > 0 ? 1 : -1;
        k_lift<<<col_grid(passes > 0 ? g0.hx : 2 * g0.hx, g0.ny), col_block(), 0, s>>>(g0, lift_fz, Hc, Wc, scale, colour,
                                                                                      v);
        HIPTRY(hipGetLastError());
        const Ops<T> o = L[0].ops();
        for (int k = 0; k < passes; ++k)
            for (int colour = 0; colour < 2; ++colour) {
                const dim3 grid = col_grid(g0.hx, g0.ny);
#define TH_RELAX(NZ)                                                                                       \
    do {                                                                                                   \
        if (L[0].ntypes)                                                                                   \
            k_line_relax_typed<T, NZ><<<grid, col_block(), 0, s>>>(g0, L[0].typed(), P, v, colour);       \
        else                                                                                               \
            k_line_relax<T, NZ><<<grid, col_block(), 0, s>>>(g0, o, P, v, colour);                        \
    } while (0)
                if (g0.nz <= 4) TH_RELAX(4);
                else if (g0.nz <= 8) TH_RELAX(8);
                else if (g0.nz <= 12) TH_RELAX(12);
                else if (g0.nz <= 16) TH_RELAX(16);
                else if (g0.nz <= 24) TH_RELAX(24);
                else if (g0.nz <= 32) TH_RELAX(32);
                else if (g0.nz <= 48) TH_RELAX(48);
                else TH_RELAX(64);
#undef TH_RELAX
                HIPTRY(hipGetLastError());
            }
        return TH_OK;
    }
    int nz() const override { return g0.nz; }
    int column_types() const override { return L.empty() ? 0 : L[0].ntypes; }
    // Level k's size, column types and which kernels run it: 0 the column
    // walk, 1 the plane-parallel sweeps, 2 the single-workgroup tail.
    int level_info(int k, int* nx, int* ny, int* nz, int* ntypes, int* kind) const override {
        if (k < 0 || k >= nlevels()) return fail(TH_ERR_ARG, "level out of range");
        const Geo& g = L[k].g;
        if (nx) *nx = g.nx;
        if (ny) *ny = g.ny;
        if (nz) *nz = g.nz;
        if (ntypes) *ntypes = L[k].ntypes;
        if (kind) *kind = (size_t)k >= tail_k0 ? 2 : ((long)g.hx * g.ny <= kPPCols && g.nz <= 16 ? 1 : 0);
        return TH_OK;
    }

    int true_relres(double* out) override { return relres_of(x, out); }

    // Die map of x (viewer.cu:53-87): RGBA image (host or device memory, or
    // none) and the hottest cell of planes l0..l1.
    unsigned* map_rgba = nullptr;
    unsigned long long *map_part = nullptr, *map_key = nullptr;
    int colorize(int l0, int l1, float rise_hi, unsigned char* rgba, bool device_rgba, float* hot_rise,
                 int* hot_cell) override {
        int rc;
        if ((rc = check_colorize(l0, l1, rise_hi))) return rc;
        if (!map_part && ((rc = dalloc(&map_part, colorize_blocks())) || (rc = dalloc(&map_key, 1)))) return rc;
        unsigned* dst = nullptr;
        if (rgba && device_rgba) dst = (unsigned*)rgba;
        else if (rgba) {
            if (!map_rgba && (rc = dalloc(&map_rgba, (size_t)g0.nx * g0.ny))) return rc;
            dst = map_rgba;
        }
        return colorize_into(l0, l1, rise_hi, dst, map_part, map_key, rgba && !device_rgba ? rgba : nullptr, hot_rise,
                             hot_cell);
    }
    int check_colorize(int l0, int l1, float rise_hi) const {
        if (l0 < 0 || l1 < l0 || l1 >= g0.nz) return fail(TH_ERR_ARG, "colorize: planes out of range");
        if (!(rise_hi > 0.f) || !std::isfinite(rise_hi)) return fail(TH_ERR_ARG, "colorize: rise_hi must be > 0");
        return TH_OK;
    }
    dim3 colorize_grid() const { return dim3((unsigned)((g0.nx + 63) / 64), (unsigned)((g0.ny + 3) / 4)); }
    long colorize_blocks() const override {
        const dim3 g = colorize_grid();
        return (long)g.x * g.y;
    }
    // k_colorize into dst (device image or null) with one partial per block
    // in part and the key in key, both device memory of the caller's; the
    // host image (host_rgba, from dst) and the hottest cell to the host.
    int colorize_into(int l0, int l1, float rise_hi, unsigned* dst, unsigned long long* part, unsigned long long* key,
                      unsigned char* host_rgba, float* hot_rise, int* hot_cell) override {
        const dim3 grid = colorize_grid();
        k_colorize<<<grid, 256, 0, s>>>(g0, x, l0, l1, rise_hi, dst, part);
        HIPTRY(hipGetLastError());
        k_max_key<<<1, 256, 0, s>>>(part, (int)colorize_blocks(), key);
        HIPTRY(hipGetLastError());
        unsigned long long kv = 0;
        HIPTRY(hipMemcpyAsync(&kv, key, sizeof(kv), hipMemcpyDeviceToHost, s));
        if (host_rgba) HIPTRY(hipMemcpyAsync(host_rgba, dst, (size_t)g0.nx * g0.ny * 4, hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        const unsigned bits = (unsigned)(kv >> 32);
        float v;
        std::memcpy(&v, &bits, 4);
        if (hot_rise) *hot_rise = v;
        if (hot_cell) *hot_cell = (int)(kv & 0xffffffffu);
        return TH_OK;
    }
    int check_colorize_args(int l0, int l1, float rise_hi) const override { return check_colorize(l0, l1, rise_hi); }
    int device() const override { return opt.device; }
    // ||P - A v|| / ||P|| of the steady operator (flux form)
    int relres_of(const double* v, double* out) {
        int rc;
        if ((rc = launch_residual<RES_CHECK>(P, v, part)) || (rc = finalize<3, FIN_CHECK>(nblk_res)) ||
            (rc = fetch_state()))
            return rc;
        *out = std::sqrt(h_st->chk_rr / h_st->chk_bb);
        return TH_OK;
    }

    int energy(double* heat, double* outq, double* tmax) override {
        const unsigned vb = vec_blocks(n0);
        k_energy<T><<<vb, kVecThreads, 0, s>>>(g0, L[0].ops(), x, part);
        HIPTRY(hipGetLastError());
        int rc;
        if ((rc = finalize<3, FIN_ENERGY, 4u>(vb)) || (rc = fetch_state())) return rc;
        if (heat) *heat = h_st->e_heat;
        if (outq) *outq = h_st->e_out;
        if (tmax) *tmax = h_st->e_tmax;
        return TH_OK;
    }

    // Test entry points: one preconditioner application / one operator apply
    // on host vectors. They use r, p and Ap as staging and leave x, b alone.
    int precondition(const double* host_r, double* host_z) override {
        int rc;
        if ((rc = upload_split<double, double>(g0, g0.nz, host_r, r))) return rc;
        k_begin<<<1, 1, 0, s>>>(st, 0.0, 1, 0);
        HIPTRY(hipGetLastError());
        const unsigned vb = vec_blocks(n0);
        double* zd = Ap;  // z as FP64 in the split layout
        if (mg()) {
            k_cast<T, double><<<vb, kVecThreads, 0, s>>>(n0, r, L[0].b);
            HIPTRY(hipGetLastError());
            if ((rc = vcycle(0))) return rc;
            k_cast<double, T><<<vb, kVecThreads, 0, s>>>(n0, L[0].x, zd);
        } else {
            k_mul<<<vb, kVecThreads, 0, s>>>(n0, r, invd, zd);
        }
        HIPTRY(hipGetLastError());
        return download_natural(zd, host_z);
    }
    int apply_operator(const double* host_x, double* host_y) override {
        int rc;
        if ((rc = upload_split<double, double>(g0, g0.nz, host_x, p))) return rc;
        k_begin<<<1, 1, 0, s>>>(st, 0.0, 1, 0);
        HIPTRY(hipGetLastError());
        if ((rc = launch_apply())) return rc;
        return download_natural(Ap, host_y);
    }
    int time_kernel(int which, int reps, double* avg_ms, double* bytes) override {
        if (reps < 1) return fail(TH_ERR_ARG, "reps must be >= 1");
        if (which != TH_K_APPLY && which != TH_K_COPY && !mg())
            return fail(TH_ERR_ARG, "multigrid kernels need a multigrid solver");
        if (which == TH_K_RESTRICT && L.size() < 2) return fail(TH_ERR_ARG, "no coarse level");
        const double N = (double)g0.nx * g0.ny * g0.nz, sz = (double)sizeof(T);
        auto launch = [&]() -> int {
            switch (which) {
                case TH_K_SWEEP:
                    return launch_sweep<SW_PLAIN>(0, 0);
                case TH_K_APPLY:
                    return mg() ? launch_apply_kernel(pm) : launch_apply_kernel(p);
                case TH_K_RESTRICT:
                    return launch_restrict(0);
                case TH_K_COPY:
                    k_cast<double, double><<<vec_blocks(n0), kVecThreads, 0, s>>>(n0, p, Ap);
                    HIPTRY(hipGetLastError());
                    return TH_OK;
                case TH_K_VCYCLE:
                    return vcycle(0);
                case TH_K_UPDATE: {
                    const unsigned vb = vec_blocks(n0);
                    if (opt.beta == TH_BETA_FLEXIBLE)
                        k_update<T, UP_MG_FLEX><<<vb, kVecThreads, 0, s>>>(n0, nullptr, Ap, x, r, L[0].b, L[0].x,
                                                                           nullptr, nullptr, st, part_upd, part_apply,
                                                                           (int)nblk_apply);
                    else
                        k_update<T, UP_MG><<<vb, kVecThreads, 0, s>>>(n0, nullptr, Ap, x, r, L[0].b, nullptr, nullptr,
                                                                      nullptr, st, part_upd, part_apply,
                                                                      (int)nblk_apply);
                    HIPTRY(hipGetLastError());
                    return TH_OK;
                }
                default:
                    return fail(TH_ERR_ARG, "unknown kernel id");
            }
        };
        switch (which) {
            case TH_K_SWEEP:  // b, x(other), gE, gN, gU, 1/pivot; x. Typed: b, x(other), x, 2-byte types
                *bytes = L[0].ntypes ? 1.5 * sz * N + 1.0 * g0.nx * g0.ny : 4.5 * sz * N;
                break;
            case TH_K_APPLY:  // Ap, p; gE, gN, gU (typed: 2-byte types)
                *bytes = (8.0 + (mg() ? sz : 8.0)) * N + (L[0].ntypes ? 2.0 * g0.nx * g0.ny : 3.0 * sz * N);
                break;
            case TH_K_RESTRICT:  // lateral: x (one colour), gE, gN (typed: 2-byte types of the red
                                 // columns); residual form: b, x, gE, gN, gU; coarse b
                *bytes = fast && L[0].ntypes ? 0.75 * sz * N + 1.0 * g0.nx * g0.ny
                                        : (fast ? 2.5 : 5.0) * sz * N + 0.25 * sz * N;
                break;
            case TH_K_COPY: *bytes = 16.0 * (double)n0; break;
            case TH_K_UPDATE:  // r, Ap, z_old (flexible beta); r, b0
                *bytes = (24.0 + (opt.beta == TH_BETA_FLEXIBLE ? 2.0 : 1.0) * sz) * N;
                break;
            default: *bytes = 0.0;
        }
        k_reopen<<<1, 1, 0, s>>>(st);  // kernels return at once on a finished solve
        HIPTRY(hipGetLastError());
        int rc;
        if (which == TH_K_UPDATE && (rc = launch_apply_kernel(pm))) return rc;  // the partials alpha is formed from
        if ((rc = launch())) return rc;
        hipEvent_t e0, e1;
        HIPTRY(hipEventCreate(&e0));
        HIPTRY(hipEventCreate(&e1));
        hipError_t e = hipEventRecord(e0, s);
        for (int k = 0; k < reps && e == hipSuccess; ++k)
            if ((rc = launch())) break;
        if (e == hipSuccess && !rc) e = hipEventRecord(e1, s);
        if (e == hipSuccess && !rc) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess && !rc) e = hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return rc;
        if (e != hipSuccess) return fail(TH_ERR_HIP, std::string("kernel timing: ") + hipGetErrorString(e));
        *avg_ms = ms / reps;
        return TH_OK;
    }

    int download_natural(const double* dev, double* host) {
        k_to_natural<<<lin_blocks(nnat), 256, 0, s>>>(g0, dev, xnat);
        HIPTRY(hipGetLastError());
        HIPTRY(hipMemcpyAsync(host, xnat, nnat * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        return TH_OK;
    }
};

}  // namespace th

