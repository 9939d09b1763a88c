// oracle.cpp - CPU restatement of the reference's MG-PCG hot path.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or
// calls this file; only tests/, __graft_entry__.smoke() and the cpu_baseline
// leg of bench.py use it, and only as the test oracle / the CPU baseline.
//
// It follows the reference's CPU path step by step, in the same operation
// order, so that a float-for-float comparison with the reference (the golden
// fixtures under tests/golden/) is meaningful:
//   model, conductances, power, TSVs, checks   src/stack3d.h:81-298
//   7-point operator (natural layout)          src/operator3d.h:50-112
//   column factors, line solve, coarsening     src/mg3d.h:45-183
//   CpuPcg (Jacobi / multigrid, FP64 PCG)      src/pcg3d_cpu.cpp:20-220
//   backward Euler driver                      src/transient3d.h:48-90
// Two additions with no reference code: the flexible (Polak-Ribiere) beta,
// and a step-count mode for the transient driver (no dt cap).
//
// Pins (what each part is checked against, in tests/test_oracle.py):
//   model arrays      bitwise vs reference fixtures; closed-form harmonic mean
//   operator          row-sum identity, symmetry, dense assembly
//   factors/line      dense tridiagonal solve per column
//   coarse levels     bitwise vs reference fixtures; uniform material = finalize
//                     on the coarse grid; sum invariants
//   V-cycle           bitwise vs reference fixture; symmetry of M^-1
//   PCG               reference iteration counts and fields; dense solve
//   transient         reference trajectory; one step vs dense solve; energy
// Built with -ffp-contract=off so the float/double rounding matches the
// reference's x86-64 build.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace orc {

constexpr int kMaxZ = 64;  // stack3d.h:34

struct Layer {
    double thickness, k, rho_c;
    int ncells;
    bool active;
};

struct Model {
    int nx = 0, ny = 0, nz = 0;
    double Lx = 0, Ly = 0, dx = 0, dy = 0, h_top = 0, h_bot = 0, T_amb = 25.0;
    double tsv_k = 400.0, tsv_rho_c = 3.45e6;
    std::vector<Layer> layers;
    std::vector<int> plane_layer;
    std::vector<double> dz, tsv_frac, P;
    std::vector<float> gE, gN, gU, gB, cap;
    bool finalized = false;
    size_t nxy() const { return (size_t)nx * ny; }
    size_t ncells() const { return nxy() * nz; }
    size_t idx(int i, int j, int l) const { return ((size_t)l * ny + j) * nx + i; }
};

// stack3d.h:81-102
static bool build_model(Model& m, const std::vector<Layer>& layers, int nx, int ny, double Lx,
                        double Ly, double h_top, double h_bot) {
    m = Model();
    m.layers = layers;
    m.nx = nx;
    m.ny = ny;
    m.Lx = Lx;
    m.Ly = Ly;
    m.dx = Lx / nx;
    m.dy = Ly / ny;
    m.h_top = h_top;
    m.h_bot = h_bot;
    for (size_t L = 0; L < layers.size(); ++L)
        for (int c = 0; c < layers[L].ncells; ++c) {
            m.plane_layer.push_back((int)L);
            m.dz.push_back(layers[L].thickness / layers[L].ncells);
        }
    m.nz = (int)m.dz.size();
    if (m.nz > kMaxZ) return false;
    m.tsv_frac.assign(m.ncells(), 0.0);
    m.P.assign(m.ncells(), 0.0);
    return true;
}

// stack3d.h:105-108
static double eff_kz(double k, double kc, double f) { return f * kc + (1 - f) * k; }
static double eff_kxy(double k, double kc, double f) {
    return k * ((1 + f) * kc + (1 - f) * k) / ((1 - f) * kc + (1 + f) * k);
}

// stack3d.h:112-148: harmonic-mean face conductances, computed in double and
// stored in float.
static void finalize(Model& m) {
    const size_t n = m.ncells(), sxy = m.nxy();
    const double A = m.dx * m.dy;
    std::vector<double> kxy(n), kz(n);
    m.cap.assign(n, 0.0f);
    for (int l = 0; l < m.nz; ++l) {
        const Layer& L = m.layers[m.plane_layer[l]];
        for (size_t c = 0; c < sxy; ++c) {
            const size_t id = (size_t)l * sxy + c;
            const double f = m.tsv_frac[id];
            kxy[id] = eff_kxy(L.k, m.tsv_k, f);
            kz[id] = eff_kz(L.k, m.tsv_k, f);
            m.cap[id] = (float)((f * m.tsv_rho_c + (1 - f) * L.rho_c) * A * m.dz[l]);
        }
    }
    m.gE.assign(n, 0.0f);
    m.gN.assign(n, 0.0f);
    m.gU.assign(n, 0.0f);
    m.gB.assign(sxy, 0.0f);
    for (int l = 0; l < m.nz; ++l)
        for (int j = 0; j < m.ny; ++j)
            for (int i = 0; i < m.nx; ++i) {
                const size_t id = m.idx(i, j, l);
                const double dz = m.dz[l];
                if (i < m.nx - 1)
                    m.gE[id] = (float)(m.dy * dz / (m.dx / (2 * kxy[id]) + m.dx / (2 * kxy[id + 1])));
                if (j < m.ny - 1)
                    m.gN[id] = (float)(m.dx * dz / (m.dy / (2 * kxy[id]) + m.dy / (2 * kxy[id + m.nx])));
                const double R_up = (l < m.nz - 1) ? dz / (2 * kz[id]) + m.dz[l + 1] / (2 * kz[id + sxy])
                                                   : dz / (2 * kz[id]) + 1.0 / m.h_top;
                m.gU[id] = (float)(A / R_up);
                if (l == 0 && m.h_bot > 0) m.gB[id] = (float)(A / (dz / (2 * kz[id]) + 1.0 / m.h_bot));
            }
    m.finalized = true;
}

// stack3d.h:171-191
static double overlap(double a0, double a1, double b0, double b1) {
    return std::max(0.0, std::min(a1, b1) - std::max(a0, b0));
}
template <class Fn>
static void for_each_overlap(const Model& m, double x0, double y0, double w, double h, Fn fn) {
    const int i0 = std::max(0, (int)std::floor(x0 / m.dx));
    const int i1 = std::min(m.nx - 1, (int)std::ceil((x0 + w) / m.dx));
    const int j0 = std::max(0, (int)std::floor(y0 / m.dy));
    const int j1 = std::min(m.ny - 1, (int)std::ceil((y0 + h) / m.dy));
    for (int j = j0; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i) {
            const double a = overlap(i * m.dx, (i + 1) * m.dx, x0, x0 + w) *
                             overlap(j * m.dy, (j + 1) * m.dy, y0, y0 + h);
            if (a > 1e-9 * m.dx * m.dy) fn((size_t)j * m.nx + i, a / (m.dx * m.dy));
        }
}

// stack3d.h:196-206
static void add_power_block(Model& m, int layer, double x0, double y0, double w, double h, double watts) {
    const double density = watts / (w * h);
    const int planes = m.layers[layer].ncells;
    for (int l = 0; l < m.nz; ++l) {
        if (m.plane_layer[l] != layer) continue;
        for_each_overlap(m, x0, y0, w, h, [&](size_t c, double frac) {
            m.P[(size_t)l * m.nxy() + c] += density * frac * m.dx * m.dy / planes;
        });
    }
}

// stack3d.h:210-220
static void add_tsv_region(Model& m, int layer, double x0, double y0, double w, double h, double f) {
    for (int l = 0; l < m.nz; ++l) {
        if (m.plane_layer[l] != layer) continue;
        for_each_overlap(m, x0, y0, w, h, [&](size_t c, double frac) {
            double& t = m.tsv_frac[(size_t)l * m.nxy() + c];
            t = std::min(1.0, t + f * frac);
        });
    }
    m.finalized = false;
}

// stack3d.h:230-244 (cell-centre rule). Returns false if no cell is covered.
static bool add_hotspot(Model& m, int layer, double xc, double yc, double size, double watts) {
    std::vector<size_t> cells;
    for (int l = 0; l < m.nz; ++l) {
        if (m.plane_layer[l] != layer) continue;
        for (int j = 0; j < m.ny; ++j)
            for (int i = 0; i < m.nx; ++i) {
                const double x = (i + 0.5) * m.dx, y = (j + 0.5) * m.dy;
                if (std::fabs(x - xc) <= size / 2 && std::fabs(y - yc) <= size / 2) cells.push_back(m.idx(i, j, l));
            }
    }
    if (cells.empty()) return false;
    for (size_t c : cells) m.P[c] += watts / (double)cells.size();
    return true;
}

// stack3d.h:279-298
static std::vector<double> analytic_1d(const Model& m) {
    std::vector<double> flux(m.nz), T(m.nz), gz(m.nz);
    const double A = m.dx * m.dy;
    for (int l = 0; l < m.nz; ++l) {
        const double k = m.layers[m.plane_layer[l]].k;
        const double R = (l < m.nz - 1) ? m.dz[l] / (2 * k) + m.dz[l + 1] / (2 * m.layers[m.plane_layer[l + 1]].k)
                                        : m.dz[l] / (2 * k) + 1.0 / m.h_top;
        gz[l] = A / R;
    }
    double F = 0;
    for (int l = 0; l < m.nz; ++l) {
        for (size_t c = 0; c < m.nxy(); ++c) F += m.P[(size_t)l * m.nxy() + c];
        flux[l] = F;
    }
    const double nxy = (double)m.nxy();
    T[m.nz - 1] = flux[m.nz - 1] / (gz[m.nz - 1] * nxy);
    for (int l = m.nz - 2; l >= 0; --l) T[l] = T[l + 1] + flux[l] / (gz[l] * nxy);
    return T;
}

// ---- operator, natural layout (operator3d.h:50-112) ----------------------
struct Op {
    int nx, ny, nz;
    const float *gE, *gN, *gU, *gB, *cap;
    size_t plane() const { return (size_t)nx * ny; }
    size_t col(int i, int j) const { return (size_t)j * nx + i; }
    size_t idx(int i, int j, int l) const { return (size_t)l * plane() + col(i, j); }
};

template <class R>
static R lat_diag(const Op& o, int i, int j, int l) {
    const size_t p = (size_t)l * o.plane();
    R d = 0;
    if (i > 0) d += (R)o.gE[p + o.col(i - 1, j)];
    if (i < o.nx - 1) d += (R)o.gE[p + o.col(i, j)];
    if (j > 0) d += (R)o.gN[p + o.col(i, j - 1)];
    if (j < o.ny - 1) d += (R)o.gN[p + o.col(i, j)];
    return d;
}
template <class R>
static R lat_sum(const Op& o, const R* x, int i, int j, int l) {
    const size_t p = (size_t)l * o.plane();
    R s = 0;
    if (i > 0) {
        const size_t w = p + o.col(i - 1, j);
        s += (R)o.gE[w] * x[w];
    }
    if (i < o.nx - 1) s += (R)o.gE[p + o.col(i, j)] * x[p + o.col(i + 1, j)];
    if (j > 0) {
        const size_t sth = p + o.col(i, j - 1);
        s += (R)o.gN[sth] * x[sth];
    }
    if (j < o.ny - 1) s += (R)o.gN[p + o.col(i, j)] * x[p + o.col(i, j + 1)];
    return s;
}
template <class R>
static R diag(const Op& o, int i, int j, int l, R inv_dt) {
    const size_t id = o.idx(i, j, l);
    R d = lat_diag<R>(o, i, j, l);
    d += (R)o.gU[id];
    d += (l > 0) ? (R)o.gU[id - o.plane()] : (R)o.gB[o.col(i, j)];
    return d + (R)o.cap[id] * inv_dt;
}
template <class R>
static R offdiag(const Op& o, const R* x, int i, int j, int l) {
    const size_t id = o.idx(i, j, l), sxy = o.plane();
    R s = lat_sum(o, x, i, j, l);
    if (l > 0) s += (R)o.gU[id - sxy] * x[id - sxy];
    if (l < o.nz - 1) s += (R)o.gU[id] * x[id + sxy];
    return s;
}
template <class R>
static R apply(const Op& o, const R* x, int i, int j, int l, R inv_dt) {
    return diag<R>(o, i, j, l, inv_dt) * x[o.idx(i, j, l)] - offdiag(o, x, i, j, l);
}

// ---- multigrid (mg3d.h:45-183) ------------------------------------------
template <class R>
static void factor_column(const Op& o, R inv_dt, R* fcp, R* finv, int i, int j) {
    const size_t sxy = o.plane(), col = o.col(i, j);
    R cp_prev = 0;
    for (int l = 0; l < o.nz; ++l) {
        const size_t id = (size_t)l * sxy + col;
        const R d = diag<R>(o, i, j, l, inv_dt);
        const R lower = (l > 0) ? -(R)o.gU[id - sxy] : R(0);
        const R upper = (l < o.nz - 1) ? -(R)o.gU[id] : R(0);
        const R inv = R(1) / (d - lower * cp_prev);
        finv[id] = inv;
        cp_prev = upper * inv;
        fcp[id] = cp_prev;
    }
}
template <class R>
static void line_solve(const Op& o, const R* fcp, const R* finv, const R* b, R* x, int i, int j) {
    const size_t sxy = o.plane(), col = o.col(i, j);
    R prev = 0;
    for (int l = 0; l < o.nz; ++l) {
        const size_t id = (size_t)l * sxy + col;
        const R s = b[id] + lat_sum(o, x, i, j, l);
        const R lower = (l > 0) ? -(R)o.gU[id - sxy] : R(0);
        prev = (s - lower * prev) * finv[id];
        x[id] = prev;
    }
    for (int l = o.nz - 2; l >= 0; --l) {
        const size_t id = (size_t)l * sxy + col;
        x[id] -= fcp[id] * x[id + sxy];
    }
}

struct LevelOp {
    int nx, ny, nz;
    std::vector<double> dx, dy;
    std::vector<float> gE, gN, gU, gB, cap;
    Op op() const { return {nx, ny, nz, gE.data(), gN.data(), gU.data(), gB.data(), cap.data()}; }
};

// mg3d.h:100-136: one coarse cell from its (up to) 2x2 children.
static void coarse_cell(const LevelOp& f, const LevelOp& c, int I, int J, int l, LevelOp& out) {
    const size_t fxy = (size_t)f.nx * f.ny, cxy = (size_t)c.nx * c.ny;
    const size_t cid = (size_t)l * cxy + (size_t)J * c.nx + I;
    const int i0 = 2 * I, i1 = (2 * I + 1 < f.nx - 1) ? 2 * I + 1 : f.nx - 1;
    const int j0 = 2 * J, j1 = (2 * J + 1 < f.ny - 1) ? 2 * J + 1 : f.ny - 1;
    double gu = 0, cp = 0, gb = 0;
    for (int j = j0; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i) {
            const size_t fid = (size_t)l * fxy + (size_t)j * f.nx + i;
            gu += f.gU[fid];
            cp += f.cap[fid];
            if (l == 0) gb += f.gB[(size_t)j * f.nx + i];
        }
    out.gU[cid] = (float)gu;
    out.cap[cid] = (float)cp;
    if (l == 0) out.gB[(size_t)J * c.nx + I] = (float)gb;
    float e = 0.0f, n = 0.0f;
    if (I < c.nx - 1) {
        double sum = 0;
        for (int j = j0; j <= j1; ++j) sum += f.gE[(size_t)l * fxy + (size_t)j * f.nx + 2 * I + 1];
        const double df = 0.5 * (f.dx[2 * I + 1] + f.dx[2 * I + 2]);
        const double dc = 0.5 * (c.dx[I] + c.dx[I + 1]);
        e = (float)(sum * df / dc);
    }
    if (J < c.ny - 1) {
        double sum = 0;
        for (int i = i0; i <= i1; ++i) sum += f.gN[(size_t)l * fxy + (size_t)(2 * J + 1) * f.nx + i];
        const double df = 0.5 * (f.dy[2 * J + 1] + f.dy[2 * J + 2]);
        const double dc = 0.5 * (c.dy[J] + c.dy[J + 1]);
        n = (float)(sum * df / dc);
    }
    out.gE[cid] = e;
    out.gN[cid] = n;
}

// mg3d.h:139-144
static std::vector<double> coarse_widths(const std::vector<double>& f) {
    std::vector<double> c;
    const int n = (int)f.size();
    for (int I = 0; I < (n + 1) / 2; ++I) c.push_back(f[2 * I] + (2 * I + 1 < n ? f[2 * I + 1] : 0.0));
    return c;
}

// mg3d.h:147-183: finest first, down to 1x1.
static std::vector<LevelOp> build_levels(int nx, int ny, int nz, double dx, double dy, const float* gE,
                                         const float* gN, const float* gU, const float* gB,
                                         const float* cap) {
    const size_t n = (size_t)nx * ny * nz, nxy = (size_t)nx * ny;
    std::vector<LevelOp> L(1);
    L[0].nx = nx;
    L[0].ny = ny;
    L[0].nz = nz;
    L[0].dx.assign(nx, dx);
    L[0].dy.assign(ny, dy);
    L[0].gE.assign(gE, gE + n);
    L[0].gN.assign(gN, gN + n);
    L[0].gU.assign(gU, gU + n);
    L[0].gB.assign(gB, gB + nxy);
    L[0].cap.assign(cap, cap + n);
    while (L.back().nx > 1 || L.back().ny > 1) {
        const LevelOp& f = L.back();
        LevelOp c;
        c.nx = (f.nx + 1) / 2;
        c.ny = (f.ny + 1) / 2;
        c.nz = f.nz;
        c.dx = coarse_widths(f.dx);
        c.dy = coarse_widths(f.dy);
        const size_t cn = (size_t)c.nx * c.ny * c.nz;
        c.gE.assign(cn, 0.0f);
        c.gN.assign(cn, 0.0f);
        c.gU.assign(cn, 0.0f);
        c.cap.assign(cn, 0.0f);
        c.gB.assign((size_t)c.nx * c.ny, 0.0f);
        for (int l = 0; l < c.nz; ++l)
            for (int J = 0; J < c.ny; ++J)
                for (int I = 0; I < c.nx; ++I) coarse_cell(f, c, I, J, l, c);
        L.push_back(std::move(c));
    }
    return L;
}

// ---- PCG (pcg3d_cpu.cpp:20-220) -----------------------------------------
enum { PC_JACOBI = 0, PC_MG = 1 };

struct SolverBase {
    virtual ~SolverBase() {}
    virtual void set_dt(double dt) = 0;
    virtual void set_x_zero() = 0;
    virtual void set_x(const double* x) = 0;
    virtual void set_rhs_steady() = 0;
    virtual void set_rhs_transient() = 0;
    virtual void set_power(const double* P) = 0;
    virtual bool solve(double tol, int maxit, int* iters, double* relres) = 0;
    virtual void get_x(double* x) const = 0;
    virtual void precond_apply(const double* r, double* z) = 0;
    virtual void op_apply(const double* x, double* y) = 0;
    virtual int nlevels() const = 0;
    virtual const LevelOp& level(int k) const = 0;
    virtual void level_factors(int k, double* fcp, double* finv) const = 0;
    bool fast_restrict = false;  // restriction from the red lateral sums (see vcycle)
};

template <class MgR>
struct Solver : SolverBase {
    struct Level {
        LevelOp h;
        std::vector<MgR> fcp, finv, x, b;
    };
    int nx, ny, nz;
    size_t n;
    int pc;
    bool flexible;
    double inv_dt = 0;
    std::vector<float> gE, gN, gU, gB, cap;
    Op op;
    std::vector<double> P, x, b, r, z, zold, p, Ap, invd;
    std::vector<Level> levels;

    Solver(int nx_, int ny_, int nz_, double dx, double dy, const float* gE_, const float* gN_,
           const float* gU_, const float* gB_, const float* cap_, const double* P_, int pc_, bool flex)
        : nx(nx_), ny(ny_), nz(nz_), pc(pc_), flexible(flex) {
        n = (size_t)nx * ny * nz;
        const size_t nxy = (size_t)nx * ny;
        gE.assign(gE_, gE_ + n);
        gN.assign(gN_, gN_ + n);
        gU.assign(gU_, gU_ + n);
        gB.assign(gB_, gB_ + nxy);
        cap.assign(cap_, cap_ + n);
        op = {nx, ny, nz, gE.data(), gN.data(), gU.data(), gB.data(), cap.data()};
        P.assign(P_, P_ + n);
        for (auto* v : {&x, &b, &r, &z, &zold, &p, &Ap, &invd}) v->assign(n, 0.0);
        if (pc == PC_MG)
            for (LevelOp& h : build_levels(nx, ny, nz, dx, dy, gE_, gN_, gU_, gB_, cap_)) {
                Level L;
                const size_t ln = (size_t)h.nx * h.ny * h.nz;
                L.h = std::move(h);
                for (auto* v : {&L.x, &L.b, &L.fcp, &L.finv}) v->assign(ln, MgR(0));
                levels.push_back(std::move(L));
            }
        set_dt(0.0);
    }

    // pcg3d_cpu.cpp:41-57
    void set_dt(double dt) override {
        inv_dt = dt > 0 ? (1.0 / dt) : 0.0;
        for (int l = 0; l < nz; ++l)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) invd[op.idx(i, j, l)] = 1.0 / diag<double>(op, i, j, l, inv_dt);
        const MgR mg_inv_dt = (MgR)inv_dt;
        for (Level& L : levels) {
            const Op o = L.h.op();
            for (int j = 0; j < L.h.ny; ++j)
                for (int i = 0; i < L.h.nx; ++i) factor_column<MgR>(o, mg_inv_dt, L.fcp.data(), L.finv.data(), i, j);
        }
    }
    void set_x_zero() override { std::fill(x.begin(), x.end(), 0.0); }
    void set_x(const double* v) override { std::copy(v, v + n, x.begin()); }
    void set_rhs_steady() override { b = P; }
    // pcg3d_cpu.cpp:63-65
    void set_rhs_transient() override {
        for (size_t c = 0; c < n; ++c) b[c] = P[c] + (double)cap[c] * inv_dt * x[c];
    }
    void set_power(const double* v) override { std::copy(v, v + n, P.begin()); }
    void get_x(double* out) const override { std::copy(x.begin(), x.end(), out); }

    // pcg3d_cpu.cpp:69-113
    void vcycle(size_t k) {
        Level& L = levels[k];
        std::fill(L.x.begin(), L.x.end(), MgR(0));
        const MgR mg_inv_dt = (MgR)inv_dt;
        const Op o = L.h.op();
        const int lnx = L.h.nx, lny = L.h.ny, lnz = L.h.nz;
        auto sweep = [&](int color) {
            for (int j = 0; j < lny; ++j)
                for (int i = 0; i < lnx; ++i)
                    if (((i + j) & 1) == color) line_solve<MgR>(o, L.fcp.data(), L.finv.data(), L.b.data(), L.x.data(), i, j);
        };
        if (k + 1 == levels.size()) {
            line_solve<MgR>(o, L.fcp.data(), L.finv.data(), L.b.data(), L.x.data(), 0, 0);
            return;
        }
        sweep(0);
        sweep(1);
        Level& C = levels[k + 1];
        const size_t fxy = (size_t)lnx * lny, cxy = (size_t)C.h.nx * C.h.ny;
        if (fast_restrict) {
            // After red (on x = 0) then black, every black column satisfies its
            // equation exactly and each red column's residual is its lateral
            // sum: coarse b = sum over the red children (2I, 2J), (2I+1, 2J+1)
            // of sum_lat G x. Equal to the restriction below in exact
            // arithmetic; no reference counterpart.
            for (int l = 0; l < lnz; ++l)
                for (int J = 0; J < C.h.ny; ++J)
                    for (int I = 0; I < C.h.nx; ++I) {
                        MgR sum = 0;
                        sum += lat_sum<MgR>(o, L.x.data(), 2 * I, 2 * J, l);
                        if (2 * I + 1 < lnx && 2 * J + 1 < lny) sum += lat_sum<MgR>(o, L.x.data(), 2 * I + 1, 2 * J + 1, l);
                        C.b[(size_t)l * cxy + (size_t)J * C.h.nx + I] = sum;
                    }
        } else
        for (int l = 0; l < lnz; ++l)
            for (int J = 0; J < C.h.ny; ++J)
                for (int I = 0; I < C.h.nx; ++I) {
                    MgR sum = 0;
                    for (int j = 2 * J; j < std::min(2 * J + 2, lny); ++j)
                        for (int i = 2 * I; i < std::min(2 * I + 2, lnx); ++i) {
                            const size_t id = (size_t)l * fxy + (size_t)j * lnx + i;
                            sum += L.b[id] - apply<MgR>(o, L.x.data(), i, j, l, mg_inv_dt);
                        }
                    C.b[(size_t)l * cxy + (size_t)J * C.h.nx + I] = sum;
                }
        vcycle(k + 1);
        for (int l = 0; l < lnz; ++l)
            for (int j = 0; j < lny; ++j)
                for (int i = 0; i < lnx; ++i)
                    L.x[(size_t)l * fxy + (size_t)j * lnx + i] += C.x[(size_t)l * cxy + (size_t)(j / 2) * C.h.nx + i / 2];
        sweep(1);
        sweep(0);
    }

    // pcg3d_cpu.cpp:115-124
    void precondition() {
        if (pc == PC_JACOBI) {
            for (size_t c = 0; c < n; ++c) z[c] = r[c] * invd[c];
        } else {
            Level& L0 = levels[0];
            for (size_t c = 0; c < n; ++c) L0.b[c] = (MgR)r[c];
            vcycle(0);
            for (size_t c = 0; c < n; ++c) z[c] = (double)L0.x[c];
        }
    }
    void precond_apply(const double* rin, double* zout) override {
        std::copy(rin, rin + n, r.begin());
        precondition();
        std::copy(z.begin(), z.end(), zout);
    }
    void op_apply(const double* xin, double* yout) override {
        for (int l = 0; l < nz; ++l)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) yout[op.idx(i, j, l)] = apply<double>(op, xin, i, j, l, inv_dt);
    }

    double dot(const std::vector<double>& a, const std::vector<double>& c) const {
        double s = 0;
        for (size_t k = 0; k < n; ++k) s += a[k] * c[k];
        return s;
    }

    // pcg3d_cpu.cpp:126-185; with `flexible`, beta = r.(z - z_old) / rz_old.
    bool solve(double tol, int maxit, int* iters_out, double* relres_out) override {
        for (int l = 0; l < nz; ++l)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const size_t c = op.idx(i, j, l);
                    r[c] = b[c] - apply<double>(op, x.data(), i, j, l, inv_dt);
                }
        const double bb = dot(b, b);
        double rr = dot(r, r);
        int iters = 0;
        double relres = 0;
        bool converged = false;
        if (bb == 0) {
            set_x_zero();
            *iters_out = 0;
            *relres_out = 0;
            return true;
        }
        relres = std::sqrt(rr / bb);
        if (relres < tol) {
            *iters_out = 0;
            *relres_out = relres;
            return true;
        }
        precondition();
        p = z;
        zold = z;
        double rz = dot(r, z);
        for (int it = 1; it <= maxit; ++it) {
            double pAp = 0;
            for (int l = 0; l < nz; ++l)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        const size_t c = op.idx(i, j, l);
                        Ap[c] = apply<double>(op, p.data(), i, j, l, inv_dt);
                        pAp += p[c] * Ap[c];
                    }
            const double alpha = rz / pAp;
            rr = 0;
            for (size_t c = 0; c < n; ++c) {
                x[c] += alpha * p[c];
                r[c] -= alpha * Ap[c];
                rr += r[c] * r[c];
            }
            iters = it;
            relres = std::sqrt(rr / bb);
            if (relres < tol) {
                converged = true;
                break;
            }
            precondition();
            const double rz_new = dot(r, z);
            double beta;
            if (flexible) {
                const double rzo = dot(r, zold);
                beta = (rz_new - rzo) / rz;
                zold = z;
            } else {
                beta = rz_new / rz;
            }
            rz = rz_new;
            for (size_t c = 0; c < n; ++c) p[c] = z[c] + beta * p[c];
        }
        *iters_out = iters;
        *relres_out = relres;
        return converged;
    }

    int nlevels() const override { return (int)levels.size(); }
    const LevelOp& level(int k) const override { return levels[k].h; }
    void level_factors(int k, double* fcp, double* finv) const override {
        const Level& L = levels[k];
        for (size_t c = 0; c < L.fcp.size(); ++c) {
            fcp[c] = (double)L.fcp[c];
            finv[c] = (double)L.finv[c];
        }
    }
};

// pcg3d_cpu.cpp:207-220
static double true_relres(const Op& o, const double* P, const double* T) {
    double rr = 0, bb = 0;
    for (int l = 0; l < o.nz; ++l)
        for (int j = 0; j < o.ny; ++j)
            for (int i = 0; i < o.nx; ++i) {
                const size_t c = o.idx(i, j, l);
                const double r = P[c] - apply<double>(o, T, i, j, l, 0.0);
                rr += r * r;
                bb += P[c] * P[c];
            }
    return std::sqrt(rr / bb);
}

// ||P - A T|| / ||P|| with A T evaluated in flux form (sum of G (T_c - T_nb)
// over faces plus the ambient links): the same operator as `apply`, with a
// rounding floor set by the temperature differences instead of T itself.
// No reference counterpart; pinned against true_relres where both floors are
// far below the residual, and by the exact 1D chain.
static double true_relres_flux(const Op& o, const double* P, const double* T) {
    double rr = 0, bb = 0;
    const size_t sxy = o.plane();
    for (int l = 0; l < o.nz; ++l)
        for (int j = 0; j < o.ny; ++j)
            for (int i = 0; i < o.nx; ++i) {
                const size_t c = o.idx(i, j, l);
                const double x = T[c];
                double f = 0;
                if (i > 0) f += (double)o.gE[c - 1] * (x - T[c - 1]);
                if (i < o.nx - 1) f += (double)o.gE[c] * (x - T[c + 1]);
                if (j > 0) f += (double)o.gN[c - o.nx] * (x - T[c - o.nx]);
                if (j < o.ny - 1) f += (double)o.gN[c] * (x - T[c + o.nx]);
                f += (l > 0) ? (double)o.gU[c - sxy] * (x - T[c - sxy]) : (double)o.gB[o.col(i, j)] * x;
                f += (l < o.nz - 1) ? (double)o.gU[c] * (x - T[c + sxy]) : (double)o.gU[c] * x;
                const double r = P[c] - f;
                rr += r * r;
                bb += P[c] * P[c];
            }
    return std::sqrt(rr / bb);
}

// stack3d.h:254-273
static double heat_out(const Op& o, const double* T) {
    double q = 0;
    const size_t top = (size_t)(o.nz - 1) * o.plane();
    for (size_t c = 0; c < o.plane(); ++c) {
        q += (double)o.gU[top + c] * T[top + c];
        q += (double)o.gB[c] * T[c];
    }
    return q;
}
static double heat_stored(const Op& o, const double* Told, const double* Tnew, double dt) {
    double q = 0;
    const size_t n = o.plane() * o.nz;
    for (size_t c = 0; c < n; ++c) q += (double)o.cap[c] * (Tnew[c] - Told[c]);
    return q / dt;
}

}  // namespace orc

// ===========================================================================
// C ABI for ctypes (oracle/oracle.py).
// ===========================================================================
using namespace orc;

extern "C" {

void* or_model_new(int nlayers, const double* thickness, const double* k, const double* rho_c,
                   const int* ncells, const int* active, int nx, int ny, double Lx, double Ly,
                   double h_top, double h_bot) {
    std::vector<Layer> layers;
    for (int L = 0; L < nlayers; ++L) layers.push_back({thickness[L], k[L], rho_c[L], ncells[L], active[L] != 0});
    Model* m = new Model();
    if (!build_model(*m, layers, nx, ny, Lx, Ly, h_top, h_bot)) {
        delete m;
        return nullptr;
    }
    return m;
}
void or_model_free(void* h) { delete (Model*)h; }
void or_model_dims(void* h, int* nx, int* ny, int* nz, double* dx, double* dy) {
    const Model* m = (Model*)h;
    *nx = m->nx;
    *ny = m->ny;
    *nz = m->nz;
    *dx = m->dx;
    *dy = m->dy;
}
void or_model_add_power_block(void* h, int layer, double x0, double y0, double w, double hh, double W) {
    add_power_block(*(Model*)h, layer, x0, y0, w, hh, W);
}
void or_model_add_uniform_power(void* h, int layer, double W) {
    Model* m = (Model*)h;
    add_power_block(*m, layer, 0, 0, m->Lx, m->Ly, W);
}
int or_model_add_hotspot(void* h, int layer, double xc, double yc, double size, double W) {
    return add_hotspot(*(Model*)h, layer, xc, yc, size, W) ? 0 : -1;
}
void or_model_add_tsv_region(void* h, int layer, double x0, double y0, double w, double hh, double f) {
    add_tsv_region(*(Model*)h, layer, x0, y0, w, hh, f);
}
void or_model_finalize(void* h) { finalize(*(Model*)h); }
// which: 0 gE, 1 gN, 2 gU, 3 gB, 4 cap (float); 5 P, 6 tsv_frac, 7 dz (double)
void or_model_get_f(void* h, int which, float* out) {
    const Model* m = (Model*)h;
    const std::vector<float>* v[] = {&m->gE, &m->gN, &m->gU, &m->gB, &m->cap};
    std::copy(v[which]->begin(), v[which]->end(), out);
}
void or_model_get_d(void* h, int which, double* out) {
    const Model* m = (Model*)h;
    const std::vector<double>& v = which == 5 ? m->P : which == 6 ? m->tsv_frac : m->dz;
    std::copy(v.begin(), v.end(), out);
}
void or_model_set_P(void* h, const double* P) {
    Model* m = (Model*)h;
    std::copy(P, P + m->ncells(), m->P.begin());
}
double or_model_total_power(void* h) {
    double s = 0;
    for (double p : ((Model*)h)->P) s += p;
    return s;
}
void or_model_analytic_1d(void* h, double* out) {
    const std::vector<double> T = analytic_1d(*(Model*)h);
    std::copy(T.begin(), T.end(), out);
}
double or_eff_kz(double k, double kc, double f) { return eff_kz(k, kc, f); }
double or_eff_kxy(double k, double kc, double f) { return eff_kxy(k, kc, f); }

void* or_solver_new(int nx, int ny, int nz, double dx, double dy, const float* gE, const float* gN,
                    const float* gU, const float* gB, const float* cap, const double* P, int precond,
                    int mg_double, int flexible) {
    if (mg_double)
        return (SolverBase*)new Solver<double>(nx, ny, nz, dx, dy, gE, gN, gU, gB, cap, P, precond, flexible != 0);
    return (SolverBase*)new Solver<float>(nx, ny, nz, dx, dy, gE, gN, gU, gB, cap, P, precond, flexible != 0);
}
void or_solver_free(void* s) { delete (SolverBase*)s; }
void or_solver_set_dt(void* s, double dt) { ((SolverBase*)s)->set_dt(dt); }
void or_solver_set_x_zero(void* s) { ((SolverBase*)s)->set_x_zero(); }
void or_solver_set_x(void* s, const double* x) { ((SolverBase*)s)->set_x(x); }
void or_solver_set_rhs_steady(void* s) { ((SolverBase*)s)->set_rhs_steady(); }
void or_solver_set_rhs_transient(void* s) { ((SolverBase*)s)->set_rhs_transient(); }
void or_solver_set_power(void* s, const double* P) { ((SolverBase*)s)->set_power(P); }
int or_solver_solve(void* s, double tol, int maxit, int* iters, double* relres) {
    return ((SolverBase*)s)->solve(tol, maxit, iters, relres) ? 1 : 0;
}
void or_solver_get_x(void* s, double* x) { ((SolverBase*)s)->get_x(x); }
void or_solver_precond(void* s, const double* r, double* z) { ((SolverBase*)s)->precond_apply(r, z); }
void or_solver_apply(void* s, const double* x, double* y) { ((SolverBase*)s)->op_apply(x, y); }
int or_solver_nlevels(void* s) { return ((SolverBase*)s)->nlevels(); }
void or_solver_level_dims(void* s, int k, int* nx, int* ny, int* nz) {
    const LevelOp& L = ((SolverBase*)s)->level(k);
    *nx = L.nx;
    *ny = L.ny;
    *nz = L.nz;
}
// which: 0 gE, 1 gN, 2 gU, 3 gB, 4 cap
void or_solver_level_get(void* s, int k, int which, float* out) {
    const LevelOp& L = ((SolverBase*)s)->level(k);
    const std::vector<float>* v[] = {&L.gE, &L.gN, &L.gU, &L.gB, &L.cap};
    std::copy(v[which]->begin(), v[which]->end(), out);
}
void or_solver_set_fast_restrict(void* s, int on) { ((SolverBase*)s)->fast_restrict = on != 0; }
void or_solver_level_factors(void* s, int k, double* fcp, double* finv) {
    ((SolverBase*)s)->level_factors(k, fcp, finv);
}

double or_true_relres(int nx, int ny, int nz, const float* gE, const float* gN, const float* gU,
                      const float* gB, const float* cap, const double* P, const double* T) {
    const Op o{nx, ny, nz, gE, gN, gU, gB, cap};
    return true_relres(o, P, T);
}
double or_true_relres_flux(int nx, int ny, int nz, const float* gE, const float* gN, const float* gU,
                           const float* gB, const float* cap, const double* P, const double* T) {
    const Op o{nx, ny, nz, gE, gN, gU, gB, cap};
    return true_relres_flux(o, P, T);
}
double or_heat_out(int nx, int ny, int nz, const float* gU, const float* gB, const double* T) {
    const Op o{nx, ny, nz, nullptr, nullptr, gU, gB, nullptr};
    return heat_out(o, T);
}
double or_heat_stored(int nx, int ny, int nz, const float* cap, const double* Told, const double* Tnew, double dt) {
    const Op o{nx, ny, nz, nullptr, nullptr, nullptr, nullptr, cap};
    return heat_stored(o, Told, Tnew, dt);
}

// transient3d.h:48-90. Runs until t_end, or for max_steps steps if max_steps > 0
// (whichever comes first). Writes one record per step (t, dt, iters, T_max,
// q_out, q_stored) into rec[6 * step + k]; returns the step count, or -1 if
// rec_cap is too small. T_final receives the last temperature field.
int or_run_transient(int nx, int ny, int nz, double dx, double dy, const float* gE, const float* gN,
                     const float* gU, const float* gB, const float* cap, const double* P, int precond,
                     int mg_double, int flexible, double dt0, double growth, double dt_max, double t_end,
                     int max_steps, double tol, int maxit, double* rec, int rec_cap, double* T_final,
                     int* all_converged) {
    SolverBase* s = (SolverBase*)or_solver_new(nx, ny, nz, dx, dy, gE, gN, gU, gB, cap, P, precond, mg_double, flexible);
    const Op o{nx, ny, nz, gE, gN, gU, gB, cap};
    const size_t n = (size_t)nx * ny * nz;
    s->set_x_zero();
    std::vector<double> Told(n, 0.0), Tnew(n);
    double t = 0, dt = dt0, last_dt = -1;
    int steps = 0;
    *all_converged = 1;
    while (t < t_end * (1 - 1e-12) && (max_steps <= 0 || steps < max_steps)) {
        dt = std::min(dt, t_end - t);
        if (dt != last_dt) {
            s->set_dt(dt);
            last_dt = dt;
        }
        s->set_rhs_transient();
        int iters = 0;
        double relres = 0;
        const bool conv = s->solve(tol, maxit, &iters, &relres);
        s->get_x(Tnew.data());
        t += dt;
        if (steps >= rec_cap) {
            delete s;
            return -1;
        }
        double* R = rec + 6 * (size_t)steps;
        R[0] = t;
        R[1] = dt;
        R[2] = iters;
        R[3] = *std::max_element(Tnew.begin(), Tnew.end());
        R[4] = heat_out(o, Tnew.data());
        R[5] = heat_stored(o, Told.data(), Tnew.data(), dt);
        if (!conv) *all_converged = 0;
        Told.swap(Tnew);
        dt = std::min(dt * growth, dt_max);
        ++steps;
    }
    std::copy(Told.begin(), Told.end(), T_final);
    delete s;
    return steps;
}

}  // extern "C"
