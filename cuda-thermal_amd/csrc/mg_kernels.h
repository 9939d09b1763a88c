// mg_kernels.h - multigrid V-cycle kernels (src/mg3d.h, pcg3d_cpu.cpp:69-113
// in what they compute; the launch design is this file's own).
//
//  * k_sweep: red-black column line smoother, one thread per column of the
//    active colour; the column's Thomas solve runs in registers (NZ is the
//    plane-count bucket), every load issued before the recurrence needs it.
//  * k_restrict: one thread per (coarse cell, plane), so the coarse levels
//    expose nz times more parallelism than a column walk.
//  * k_tail: every level at or below 32 x 32 columns in a single workgroup,
//    phases separated by barriers - one launch instead of ~5 per level for
//    levels whose kernels would be pure launch latency.
//  * The V-cycle needs no clearing pass: the first red sweep takes its black
//    neighbours as zero (what the reference's memset-then-sweep computes) and
//    the prolongation is folded into the first post-smoothing sweep, which
//    adds the parent's coarse correction to each neighbour it reads.
//  * Only 1/pivot is stored per cell; the upper factor is rebuilt as
//    (-gU) * (1/pivot), the same product the reference stores.
#pragma once
#include "kernels.h"

namespace th {

enum { SW_PLAIN = 0, SW_ZERO = 1, SW_PROLONG = 2 };
constexpr int kTailThreads = 512;
constexpr int kMaxTail = 8;
constexpr int kTailCols = 32 * 32;  // levels with at most this many columns go to the L2 tail
constexpr size_t kTailLdsBytes = 148 * 1024;  // LDS budget of the LDS-resident tail (+ 5 or 10 KB static)
constexpr int kTailLdsThreads = 1024;           // plane-parallel sweeps: 64 / LP columns per wave

// Coarse parents of the 4 lateral neighbours of column (i, j) (SW_PROLONG).
struct Parents {
    int w, e, s, n;
};
__device__ __forceinline__ Parents parents(const Geo& gc, const Col& c, int i, int j) {
    return Parents{gc.pos((c.hw ? i - 1 : i) >> 1, j >> 1), gc.pos((c.he ? i + 1 : i) >> 1, j >> 1),
                   gc.pos(i >> 1, (c.hs ? j - 1 : j) >> 1), gc.pos(i >> 1, (c.hn ? j + 1 : j) >> 1)};
}

// The same four parents in three loads: one of the W/E neighbours and one of
// the S/N neighbours always share the column's own parent p0; px and py are
// the other two, and wx/ex/sy/nn say which neighbour's parent is px or py.
struct Parents3 {
    int p0, px, py;
    bool wx, ex, sy, nn;
};
__device__ __forceinline__ Parents3 parents3(const Geo& gc, const Col& c, int i, int j) {
    Parents3 q;
    q.wx = c.hw && !(i & 1);
    q.ex = c.he && (i & 1);
    q.sy = c.hs && !(j & 1);
    q.nn = c.hn && (j & 1);
    q.p0 = gc.pos(i >> 1, j >> 1);
    q.px = gc.pos(q.wx ? (i - 1) >> 1 : q.ex ? (i + 1) >> 1 : i >> 1, j >> 1);
    q.py = gc.pos(i >> 1, q.sy ? (j - 1) >> 1 : q.nn ? (j + 1) >> 1 : j >> 1);
    return q;
}

// Right-hand side of plane l of column c with its lateral neighbours held
// fixed: b + sum G x_nb (SW_PROLONG: each neighbour plus its coarse parent's
// correction; SW_ZERO: neighbours taken as 0).
template <class T, int MODE>
__device__ __forceinline__ T column_rhs(const Geo& g, const T* __restrict__ gE, const T* __restrict__ gN,
                                        const T* __restrict__ b, const T* __restrict__ x, const Col& c, int l,
                                        const Geo& gc, const T* __restrict__ xc, const Parents& pa) {
    const long base = (long)l * g.plane;
    const long id = base + c.c0;
    T v = b[id];
    if (MODE != SW_ZERO) {
        T xw = x[base + c.pw], xe = x[base + c.pe], xsv = x[base + c.ps], xn = x[base + c.pn];
        if (MODE == SW_PROLONG) {
            const long cb = (long)l * gc.plane;
            xw = xw + xc[cb + pa.w];
            xe = xe + xc[cb + pa.e];
            xsv = xsv + xc[cb + pa.s];
            xn = xn + xc[cb + pa.n];
        }
        const T Gw = gE[base + c.pw], Ge = gE[id], Gs = gN[base + c.ps], Gn = gN[id];
        T ls = 0;
        if (c.hw) ls += Gw * xw;
        if (c.he) ls += Ge * xe;
        if (c.hs) ls += Gs * xsv;
        if (c.hn) ls += Gn * xn;
        v = v + ls;
    }
    return v;
}

// Thomas forward elimination and back substitution of one column in place
// in s (planes 0..nz-1), the upper factor rebuilt as (-gU) * (1/pivot).
template <class T, int NZ>
__device__ __forceinline__ void column_thomas(int nz, T (&s)[NZ], const T (&inv)[NZ], const T (&gu)[NZ]) {
    T prev = 0;
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        if (l < nz) {
            const T lower = (l > 0) ? -gu[l - 1] : T(0);
            prev = (s[l] - lower * prev) * inv[l];
            s[l] = prev;
        }
    }
#pragma unroll
    for (int l = NZ - 2; l >= 0; --l)
        if (l < nz - 1) s[l] = s[l] - ((-gu[l]) * inv[l]) * s[l + 1];
}

// Exact solve of column (i, j) with its lateral neighbours fixed
// (mg3d.h:67-84), in place in x. MODE: see above. xc/gc: the coarse level
// (SW_PROLONG only). Every plane's right-hand side is gathered first (no
// plane depends on another, so all loads are in flight at once), then the
// recurrence runs in registers.
template <class T, int NZ, int MODE>
__device__ __forceinline__ void sweep_column(const Geo& g, const T* __restrict__ gE, const T* __restrict__ gN,
                                             const T* __restrict__ gU, const T* __restrict__ finv,
                                             const T* __restrict__ b, T* __restrict__ x, int i, int j,
                                             const Geo& gc, const T* __restrict__ xc) {
    const int nz = g.nz;
    const Col c = make_col(g, i, j);
    const Parents pa = MODE == SW_PROLONG ? parents(gc, c, i, j) : Parents{0, 0, 0, 0};
    T s[NZ], inv[NZ], gu[NZ];
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        if (l < nz) {
            const long id = (long)l * g.plane + c.c0;
            s[l] = column_rhs<T, MODE>(g, gE, gN, b, x, c, l, gc, xc, pa);
            inv[l] = finv[id];
            gu[l] = gU[id];
        }
    }
    column_thomas<T, NZ>(nz, s, inv, gu);
#pragma unroll
    for (int l = 0; l < NZ; ++l)
        if (l < nz) x[(long)l * g.plane + c.c0] = s[l];
}

// Plane-parallel sweep for levels with few columns, where the column walk
// above is latency-bound (one thread issuing ~11 loads per plane): LP lanes
// per column, one per plane, 256 / LP columns per 256-thread block (a column
// never spans two wavefronts). Each lane gathers its plane's right-hand
// side, 1/pivot and gU into LDS; thread k < 256 / LP then runs column k's
// Thomas recurrence from LDS (the same arithmetic in the same order as
// sweep_column): one wave issues the serial part, not every wave for its
// own few columns. Every lane writes its plane back.
constexpr int kPPThreads = 256;
constexpr long kPPCols = 8192;  // levels with at most this many columns per colour sweep plane-parallel
__host__ __device__ constexpr int pp_lanes(int nz) { return nz <= 4 ? 4 : nz <= 8 ? 8 : 16; }
template <class T, int LP, int MODE>
__device__ __forceinline__ void sweep_pp_body(const Blk& B, const Geo& g, const Ops<T>& o, const T* __restrict__ finv,
                                              const T* __restrict__ b, T* __restrict__ x, int colour, const Geo& gc,
                                              const T* __restrict__ xc, State* st,
                                              const double* __restrict__ cons_part, int cons_nblk, int cons_mode) {
    // Scratch rows by plane, CPB + 1 apart: the Thomas lanes (one per
    // column) read consecutive words instead of words LP apart in one bank.
    constexpr int CPB = kPPThreads / LP, RS = CPB + 1;
    __shared__ __align__(16) T sh_s[LP * RS], sh_i[LP * RS], sh_u[LP * RS];
    __shared__ int sh_ok[kPPThreads / 4];
    if (st->done) return;
    if (cons_mode != CONS_NONE && consume_convergence(cons_part, cons_nblk, cons_mode, st, B.first())) return;
    const int tid = threadIdx.x, l = tid % LP, cb = tid / LP;
    const long q = (long)B.x * CPB + cb;  // column of this colour: t + hx * j
    const int j = (int)(q / g.hx), t = (int)(q - (long)j * g.hx);
    const int i = 2 * t + ((j + colour) & 1);
    const bool colok = j < g.ny && i < g.nx;
    const bool ok = colok && l < g.nz;
    Col c{};
    if (colok) c = make_col(g, i, j);
    if (ok) {
        const Parents pa = MODE == SW_PROLONG ? parents(gc, c, i, j) : Parents{0, 0, 0, 0};
        const long id = (long)l * g.plane + c.c0;
        sh_s[l * RS + cb] = column_rhs<T, MODE>(g, o.gE, o.gN, b, x, c, l, gc, xc, pa);
        sh_i[l * RS + cb] = finv[id];
        sh_u[l * RS + cb] = o.gU[id];
    }
    if (l == 0) sh_ok[cb] = colok;
    __syncthreads();
    if (tid < CPB && sh_ok[tid]) {
        T s[LP], inv[LP], gu[LP];
#pragma unroll
        for (int k = 0; k < LP; ++k) {
            s[k] = sh_s[k * RS + tid];
            inv[k] = sh_i[k * RS + tid];
            gu[k] = sh_u[k * RS + tid];
        }
        column_thomas<T, LP>(g.nz, s, inv, gu);
#pragma unroll
        for (int k = 0; k < LP; ++k) sh_s[k * RS + tid] = s[k];
    }
    __syncthreads();
    if (ok) x[(long)l * g.plane + c.c0] = sh_s[l * RS + cb];
}
template <class T, int LP, int MODE>
__global__ __launch_bounds__(kPPThreads) void k_sweep_pp(Geo g, Ops<T> o, const T* __restrict__ finv,
                                                          const T* __restrict__ b, T* __restrict__ x, int colour,
                                                          Geo gc, const T* __restrict__ xc, State* st,
                                                          const double* __restrict__ cons_part, int cons_nblk,
                                                          int cons_mode) {
    sweep_pp_body<T, LP, MODE>(hw_blk(), g, o, finv, b, x, colour, gc, xc, st, cons_part, cons_nblk, cons_mode);
}

// Coarse b at (I, J, l): sum over the (up to 4) children, row-major, of the
// fine residual b - A x (pcg3d_cpu.cpp:87-100).
template <class T>
__device__ __forceinline__ void restrict_cell(const Geo& f, const Ops<T>& o, const T* __restrict__ fb,
                                              const T* __restrict__ fx, const Geo& cg, T* __restrict__ cb, int I,
                                              int J, int l, T inv_dt) {
    const long base = (long)l * f.plane;
    T r[4];
    bool ok[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = 2 * I + (k & 1), j = 2 * J + (k >> 1);
        ok[k] = i < f.nx && j < f.ny;
        const Col c = make_col(f, ok[k] ? i : 2 * I, ok[k] ? j : 2 * J);
        r[k] = fb[base + c.c0] - apply_cell<T>(f, o.gE, o.gN, o.gU, o.gB, o.cap, fx, c, l, inv_dt);
    }
    T sum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (ok[k]) sum += r[k];
    cb[(long)l * cg.plane + cg.pos(I, J)] = sum;
}

// Coarse b at (I, J, l) from the red lateral sums. After the red sweep on
// x = 0 and the black sweep, every black column satisfies its equation
// exactly and a red column's residual is just its lateral term sum G x
// (the black neighbours it was solved without). So the restricted residual is
// the sum over the red children (2I, 2J) and (2I+1, 2J+1) of sum_lat G x:
// equal to restrict_cell in exact arithmetic, with 11 instead of 21 bytes
// per fine cell (x of one colour, gE, gN; no b, gU, cap).
template <class T>
__device__ __forceinline__ T restrict_sum_fast(const Geo& f, const T* __restrict__ gE, const T* __restrict__ gN,
                                               const T* __restrict__ fx, int I, int J, int l) {
    const long base = (long)l * f.plane;
    const bool two = 2 * I + 1 < f.nx && 2 * J + 1 < f.ny;
    T r[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const Col c = make_col(f, 2 * I + (two ? k : 0), 2 * J + (two ? k : 0));
        const T Gw = gE[base + c.pw], Ge = gE[base + c.c0], Gs = gN[base + c.ps], Gn = gN[base + c.c0];
        const T xw = fx[base + c.pw], xe = fx[base + c.pe], xs = fx[base + c.ps], xn = fx[base + c.pn];
        T s = 0;
        if (c.hw) s += Gw * xw;
        if (c.he) s += Ge * xe;
        if (c.hs) s += Gs * xs;
        if (c.hn) s += Gn * xn;
        r[k] = s;
    }
    T sum = 0;
    sum += r[0];
    if (two) sum += r[1];
    return sum;
}
template <class T>
__device__ __forceinline__ void restrict_cell_fast(const Geo& f, const T* __restrict__ gE, const T* __restrict__ gN,
                                                   const T* __restrict__ fx, const Geo& cg, T* __restrict__ cb, int I,
                                                   int J, int l) {
    cb[(long)l * cg.plane + cg.pos(I, J)] = restrict_sum_fast<T>(f, gE, gN, fx, I, J, l);
}

// Block (B.x, B.y) covers blockDim.x x blockDim.y coarse cells of plane B.z.
template <class T>
__device__ __forceinline__ void restrict_fast_body(const Blk& B, const Geo& f, const Ops<T>& o,
                                                   const T* __restrict__ fx, const Geo& cg, T* __restrict__ cb,
                                                   const State* __restrict__ st) {
    if (st->done) return;
    const int I = B.x * blockDim.x + threadIdx.x;
    const int J = B.y * blockDim.y + threadIdx.y;
    if (I >= cg.nx || J >= cg.ny) return;
    restrict_cell_fast<T>(f, o.gE, o.gN, fx, cg, cb, I, J, B.z);
}
template <class T>
__global__ __launch_bounds__(256) void k_restrict_fast(Geo f, Ops<T> o, const T* __restrict__ fx, Geo cg,
                                                        T* __restrict__ cb, const State* __restrict__ st) {
    restrict_fast_body<T>(hw_blk(), f, o, fx, cg, cb, st);
}

// One colour sweep of a level: block (B.x, B.y) covers blockDim.x columns of the
// colour in each of blockDim.y rows. The first sweep of a V-cycle can also
// carry the convergence test of the kernel before it (cons_mode, see
// consume_convergence): a converged solve ends here, with no separate launch.
template <class T, int NZ, int MODE>
__device__ __forceinline__ void sweep_body(const Blk& B, const Geo& g, const Ops<T>& o, const T* __restrict__ finv,
                                           const T* __restrict__ b, T* __restrict__ x, int colour, const Geo& gc,
                                           const T* __restrict__ xc, State* st, const double* __restrict__ cons_part,
                                           int cons_nblk, int cons_mode) {
    if (st->done) return;
    if (cons_mode != CONS_NONE && consume_convergence(cons_part, cons_nblk, cons_mode, st, B.first())) return;
    const int t = B.x * blockDim.x + threadIdx.x;
    const int j = B.y * blockDim.y + threadIdx.y;
    if (j >= g.ny) return;
    const int i = 2 * t + ((j + colour) & 1);
    if (i >= g.nx) return;
    sweep_column<T, NZ, MODE>(g, o.gE, o.gN, o.gU, finv, b, x, i, j, gc, xc);
}
template <class T, int NZ, int MODE>
__global__ __launch_bounds__(256) void k_sweep(Geo g, Ops<T> o, const T* __restrict__ finv,
                                                const T* __restrict__ b, T* __restrict__ x, int colour, Geo gc,
                                                const T* __restrict__ xc, State* st, const double* __restrict__ cons_part,
                                                int cons_nblk, int cons_mode) {
    sweep_body<T, NZ, MODE>(hw_blk(), g, o, finv, b, x, colour, gc, xc, st, cons_part, cons_nblk, cons_mode);
}

template <class T>
__global__ __launch_bounds__(256) void k_restrict(Geo f, Ops<T> o, const T* __restrict__ fb,
                                                   const T* __restrict__ fx, Geo cg, T* __restrict__ cb,
                                                   const State* __restrict__ st) {
    if (st->done) return;
    const int I = blockIdx.x * kColsX + threadIdx.x;
    const int J = blockIdx.y * kColsY + threadIdx.y;
    const int l = blockIdx.z;
    if (I >= cg.nx || J >= cg.ny) return;
    restrict_cell<T>(f, o, fb, fx, cg, cb, I, J, l, (T)st->inv_dt);
}

// Column factors (mg3d.h:46-61) for every column of a level: 1/pivot. NZ
// is the plane-count bucket: planes past nz re-read the top plane, so every
// plane's loads are unconditional and in flight together before the serial
// recurrence (absent lateral faces point at the cell itself and add
// nothing, as before). With the loads inside the recurrence, each plane
// waited for its own: 30 us per launch on the 128^2 ... 16^2 levels at
// nz = 16 whatever their size. The same arithmetic in the same order.
template <class T, int NZ>
__device__ __forceinline__ void factor_body(const Geo& g, const Ops<T>& o, T* __restrict__ finv,
                                            const State* __restrict__ st, int bx, int by) {
    const int q = bx * kColsX + threadIdx.x;
    const int j = by * kColsY + threadIdx.y;
    int i;
    if (!st->refactor || j >= g.ny || q >= 2 * g.hx || !slot_column(g, q, j, i)) return;
    const Col c = make_col(g, i, j);
    const T inv_dt = (T)st->inv_dt;
    const int nz = g.nz;
    T dg[NZ], gu[NZ], gd[NZ];
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        const int lc = l < nz ? l : nz - 1;
        const long base = (long)lc * g.plane, id = base + c.c0;
        const T ew = o.gE[base + c.pw], ee = o.gE[id], ns = o.gN[base + c.ps], nn = o.gN[id];
        const T u = o.gU[id], dn = lc > 0 ? o.gU[id - g.plane] : o.gB[c.c0], cp = o.cap[id];
        T d = 0;
        if (c.hw) d += ew;
        if (c.he) d += ee;
        if (c.hs) d += ns;
        if (c.hn) d += nn;
        d += u;
        d += dn;
        dg[l] = d + cp * inv_dt;
        gu[l] = u;
        gd[l] = dn;
    }
    T cp_prev = 0;
#pragma unroll
    for (int l = 0; l < NZ; ++l)
        if (l < nz) {
            const T lower = (l > 0) ? -gd[l] : T(0);
            const T upper = (l < nz - 1) ? -gu[l] : T(0);
            const T inv = T(1) / (dg[l] - lower * cp_prev);
            finv[(long)l * g.plane + c.c0] = inv;
            cp_prev = upper * inv;
        }
}

// The factors of up to kFactorLevels levels in one launch (every level of a
// hierarchy has the same plane count): a flat grid whose blocks find their
// level in blk0 and their place in its own 64 x 4-block grid. One launch
// instead of one per level for the small levels, whose launches cost the
// same whatever their size.
constexpr int kFactorLevels = 16;
template <class T>
struct FactorLevels {
    int n;
    int blk0[kFactorLevels + 1];  // first flat block of each level; blk0[n]: all blocks
    int gx[kFactorLevels];        // blocks along a row of each level's grid
    Geo g[kFactorLevels];
    Ops<T> o[kFactorLevels];
    T* finv[kFactorLevels];
};
template <class T, int NZ>
__global__ __launch_bounds__(256) void k_factor_levels(FactorLevels<T> L, const State* __restrict__ st) {
    const int b = blockIdx.x;
    int k = 0;
    while (k + 1 < L.n && b >= L.blk0[k + 1]) ++k;
    const int local = b - L.blk0[k];
    factor_body<T, NZ>(L.g[k], L.o[k], L.finv[k], st, local % L.gx[k], local / L.gx[k]);
}

// ---- typed fine level ----------------------------------------------------------
// The fine level of a layered stack has few distinct columns (see levels.h:
// column_types): each plane slot holds a 2-byte type and the faces come from
// the type's table, so a sweep reads b and x per cell and little else. The
// values are the per-cell arrays' own (checked bit for bit when the types
// are built), and every kernel below does its dense counterpart's arithmetic
// in the same order: results are identical.

// 1/pivot of every type for the current dt (factor_body's arithmetic), every
// plane's row read before the recurrence (NZ: the plane-count bucket).
template <class T, int NZ>
__device__ __forceinline__ void factor_types_body(T* __restrict__ tab, const unsigned char* __restrict__ fl, int nt,
                                                  int nz, int nzp, const State* __restrict__ st, int t) {
    if (!st->refactor || t >= nt) return;
    T* rows = tab + (long)t * nzp * kTypeFields;  // rows[l * kTypeFields + field]
    const T gB = rows[TF_B];
    const unsigned fb = fl[t];
    const bool hw = fb & 1, he = fb & 2, hs = fb & 4, hn = fb & 8;
    const T inv_dt = (T)st->inv_dt;
    T dg[NZ], gu[NZ], gd[NZ];
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        const int lc = l < nz ? l : nz - 1;
        const T* r = rows + lc * kTypeFields;
        const T dn = lc > 0 ? r[TF_U - kTypeFields] : gB;
        T d = 0;
        if (hw) d += r[TF_W];
        if (he) d += r[TF_E];
        if (hs) d += r[TF_S];
        if (hn) d += r[TF_N];
        d += r[TF_U];
        d += dn;
        dg[l] = d + r[TF_C] * inv_dt;
        gu[l] = r[TF_U];
        gd[l] = dn;
    }
    T cp_prev = 0;
#pragma unroll
    for (int l = 0; l < NZ; ++l)
        if (l < nz) {
            const T lower = (l > 0) ? -gd[l] : T(0);
            const T upper = (l < nz - 1) ? -gu[l] : T(0);
            const T inv = T(1) / (dg[l] - lower * cp_prev);
            rows[l * kTypeFields + TF_FI] = inv;
            cp_prev = upper * inv;
        }
}

// The types of up to kFactorLevels levels in one launch: block b takes 64
// types of the level whose range of blocks holds it.
template <class T>
struct TypeLevels {
    int n;
    int blk0[kFactorLevels + 1];
    int nt[kFactorLevels];
    T* tab[kFactorLevels];
    const unsigned char* fl[kFactorLevels];
};
template <class T, int NZ>
__global__ __launch_bounds__(64) void k_factor_types_levels(TypeLevels<T> L, int nz, const State* __restrict__ st) {
    const int b = blockIdx.x;
    int k = 0;
    while (k + 1 < L.n && b >= L.blk0[k + 1]) ++k;
    factor_types_body<T, NZ>(L.tab[k], L.fl[k], L.nt[k], nz, nz, st, (b - L.blk0[k]) * 64 + (int)threadIdx.x);
}

// sweep_column with the faces from the column's type. UNI: every column of
// the wave has type t (a wave-uniform value), so the rows come through the
// scalar cache. Planes past nz (NZ is a size bucket) re-read the top plane,
// so every load is unconditional and in bounds and the loads of all planes
// can be in flight together; those planes' values are never used. C3
// (SW_PROLONG): the coarse parents in three loads per plane (parents3).
template <class T, int NZ, int MODE, bool UNI, bool C3>
__device__ __forceinline__ void sweep_column_typed(const Geo& g, const Typed<T>& y, const T* __restrict__ b,
                                                   T* __restrict__ x, const Col& c, int t, const Geo& gc,
                                                   const T* __restrict__ xc, const Parents& pa,
                                                   const Parents3& p3) {
    const int nz = g.nz;
    T s[NZ], inv[NZ], gu[NZ];
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        const int lc = l < nz ? l : nz - 1;
        const long base = (long)lc * g.plane, id = base + c.c0;
        const TypeRow<T> r = UNI ? y.urow(t, lc) : y.row(t, lc);
        T v = b[id];
        if (MODE != SW_ZERO) {  // column_rhs
            T xw = x[base + c.pw], xe = x[base + c.pe], xsv = x[base + c.ps], xn = x[base + c.pn];
            if (MODE == SW_PROLONG && C3) {
                const long cb = (long)lc * gc.plane;
                const T c0 = xc[cb + p3.p0], cx = xc[cb + p3.px], cy = xc[cb + p3.py];
                xw = xw + (p3.wx ? cx : c0);
                xe = xe + (p3.ex ? cx : c0);
                xsv = xsv + (p3.sy ? cy : c0);
                xn = xn + (p3.nn ? cy : c0);
            } else if (MODE == SW_PROLONG) {
                const long cb = (long)lc * gc.plane;
                xw = xw + xc[cb + pa.w];
                xe = xe + xc[cb + pa.e];
                xsv = xsv + xc[cb + pa.s];
                xn = xn + xc[cb + pa.n];
            }
            // Absent faces add (+0) * x: ls starts at +0 and never becomes -0,
            // so the sum is the one with those terms skipped, without the
            // branches that would hold a load back until its use.
            const T Gw = c.hw ? r.v[TF_W] : T(0), Ge = c.he ? r.v[TF_E] : T(0);
            const T Gs = c.hs ? r.v[TF_S] : T(0), Gn = c.hn ? r.v[TF_N] : T(0);
            T ls = 0;
            ls += Gw * xw;
            ls += Ge * xe;
            ls += Gs * xsv;
            ls += Gn * xn;
            v = v + ls;
        }
        s[l] = v;
        inv[l] = r.v[TF_FI];
        gu[l] = r.v[TF_U];
    }
    column_thomas<T, NZ>(nz, s, inv, gu);
#pragma unroll
    for (int l = 0; l < NZ; ++l)
        if (l < nz) x[(long)l * g.plane + c.c0] = s[l];
}
// Block (B.x, B.y) covers blockDim.x columns of the colour in each of
// blockDim.y rows (sweep_body's mapping; the batched engine runs it with its
// own block shapes: the type tests are per wave, whatever rows it spans).
template <class T, int NZ, int MODE, bool C3, int BX = 0>
__device__ __forceinline__ void sweep_typed_body(const Blk& B, const Geo& g, const Typed<T>& y,
                                                 const T* __restrict__ b, T* __restrict__ x, int colour, const Geo& gc,
                                                 const T* __restrict__ xc, State* st,
                                                 const double* __restrict__ cons_part, int cons_nblk, int cons_mode) {
    if (st->done) return;
    if (cons_mode != CONS_NONE && consume_convergence(cons_part, cons_nblk, cons_mode, st, B.first())) return;
    const int t = B.x0<BX>() + threadIdx.x;
    const int j = B.y * blockDim.y + threadIdx.y;
    if (j >= g.ny) return;
    const int i = 2 * t + ((j + colour) & 1);
    if (i >= g.nx) return;
    const Col c = make_col(g, i, j);
    const Parents pa = MODE == SW_PROLONG && !C3 ? parents(gc, c, i, j) : Parents{0, 0, 0, 0};
    const Parents3 p3 = MODE == SW_PROLONG && C3 ? parents3(gc, c, i, j) : Parents3{0, 0, 0, false, false, false, false};
    const int ty = y.ty[c.c0];
    const int rep = wave_rep_type(ty);
    if (__all(ty == rep))
        sweep_column_typed<T, NZ, MODE, true, C3>(g, y, b, x, c, rep, gc, xc, pa, p3);
    else
        sweep_column_typed<T, NZ, MODE, false, C3>(g, y, b, x, c, ty, gc, xc, pa, p3);
}
template <class T, int NZ, int MODE>
__global__ __launch_bounds__(256) void k_sweep_typed(Geo g, Typed<T> y, const T* __restrict__ b, T* __restrict__ x,
                                                     int colour, Geo gc, const T* __restrict__ xc, State* st,
                                                     const double* __restrict__ cons_part, int cons_nblk,
                                                     int cons_mode) {
    sweep_typed_body<T, NZ, MODE, false, kColsX>(hw_blk(), g, y, b, x, colour, gc, xc, st, cons_part, cons_nblk, cons_mode);
}
// The FP32 prolongation sweep: three coarse loads per plane and at least 4
// waves per SIMD. With every plane's loads unconditional the compiler would
// otherwise hold them all at 160 VGPRs; at 1024^2 x 16 with the caches
// flushed this runs in 52 us against 62 for the per-plane form and 75 at the
// compiler's own register choice. (FP64 runs fastest at its own choice.)
template <class T, int NZ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_sweep_typed_prolong3(
    Geo g, Typed<T> y, const T* __restrict__ b, T* __restrict__ x, int colour, Geo gc, const T* __restrict__ xc,
    State* st, const double* __restrict__ cons_part, int cons_nblk, int cons_mode) {
    sweep_typed_body<T, NZ, SW_PROLONG, true, kColsX>(hw_blk(), g, y, b, x, colour, gc, xc, st, cons_part, cons_nblk, cons_mode);
}

// FP32 plain and prolongation sweeps (NZ <= 16) with every load that does
// not depend on the column's type (b, the neighbours' x, the three coarse
// parents) issued before the wave's type test, so no wave waits for its
// 2-byte type before them; then sweep_column_typed's arithmetic, in its
// order, from the loaded values. At 1024^2 x 16, prolongation: 31.3 -> 29.1
// us warm and 51.4 -> 48.5 us with the caches flushed (against
// k_sweep_typed_prolong3); plain: 36.9 -> 34.4 us flushed but 21.3 -> 22.4
// warm, so only the pre-smoothing plain sweep, which runs on inputs the
// caches do not hold, uses it. (FP64 and the zero sweep lose; past NZ = 16
// the five register columns would spill.)
template <class T, int NZ, bool UNI>
__device__ __forceinline__ void sweep_column_typed_loaded(const Geo& g, const Typed<T>& y, T* __restrict__ x,
                                                          const Col& c, int t, const T (&bv)[NZ], const T (&xw)[NZ],
                                                          const T (&xe)[NZ], const T (&xs)[NZ], const T (&xn)[NZ]) {
    const int nz = g.nz;
    T s[NZ], inv[NZ], gu[NZ];
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        const int lc = l < nz ? l : nz - 1;
        const TypeRow<T> r = UNI ? y.urow(t, lc) : y.row(t, lc);
        const T Gw = c.hw ? r.v[TF_W] : T(0), Ge = c.he ? r.v[TF_E] : T(0);
        const T Gs = c.hs ? r.v[TF_S] : T(0), Gn = c.hn ? r.v[TF_N] : T(0);
        T ls = 0;  // absent faces add (+0) * x, as in sweep_column_typed
        ls += Gw * xw[l];
        ls += Ge * xe[l];
        ls += Gs * xs[l];
        ls += Gn * xn[l];
        s[l] = bv[l] + ls;
        inv[l] = r.v[TF_FI];
        gu[l] = r.v[TF_U];
    }
    column_thomas<T, NZ>(nz, s, inv, gu);
#pragma unroll
    for (int l = 0; l < NZ; ++l)
        if (l < nz) x[(long)l * g.plane + c.c0] = s[l];
}
template <class T, int NZ, int MODE, int BX = 0>
__device__ __forceinline__ void sweep_typed_lf_body(const Blk& B, const Geo& g, const Typed<T>& y,
                                                    const T* __restrict__ b, T* __restrict__ x, int colour,
                                                    const Geo& gc, const T* __restrict__ xc, State* st,
                                                    const double* __restrict__ cons_part, int cons_nblk,
                                                    int cons_mode) {
    static_assert(NZ <= 16, "five register columns of NZ values");
    static_assert(MODE == SW_PLAIN || MODE == SW_PROLONG, "plain or prolongation sweep");
    if (st->done) return;
    if (cons_mode != CONS_NONE && consume_convergence(cons_part, cons_nblk, cons_mode, st, B.first())) return;
    const int tt = B.x0<BX>() + threadIdx.x;
    const int j = B.y * blockDim.y + threadIdx.y;
    if (j >= g.ny) return;
    const int i = 2 * tt + ((j + colour) & 1);
    if (i >= g.nx) return;
    const Col c = make_col(g, i, j);
    const int ty = y.ty[c.c0];
    const int nz = g.nz;
    const Parents3 p3 = MODE == SW_PROLONG ? parents3(gc, c, i, j) : Parents3{0, 0, 0, false, false, false, false};
    T bv[NZ], xw[NZ], xe[NZ], xs[NZ], xn[NZ];
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        const int lc = l < nz ? l : nz - 1;
        const long base = (long)lc * g.plane;
        bv[l] = b[base + c.c0];
        xw[l] = x[base + c.pw];
        xe[l] = x[base + c.pe];
        xs[l] = x[base + c.ps];
        xn[l] = x[base + c.pn];
        if (MODE == SW_PROLONG) {
            const long cb = (long)lc * gc.plane;
            const T c0 = xc[cb + p3.p0], cx = xc[cb + p3.px], cy = xc[cb + p3.py];
            xw[l] = xw[l] + (p3.wx ? cx : c0);
            xe[l] = xe[l] + (p3.ex ? cx : c0);
            xs[l] = xs[l] + (p3.sy ? cy : c0);
            xn[l] = xn[l] + (p3.nn ? cy : c0);
        }
    }
    const int rep = wave_rep_type(ty);
    if (__all(ty == rep))
        sweep_column_typed_loaded<T, NZ, true>(g, y, x, c, rep, bv, xw, xe, xs, xn);
    else
        sweep_column_typed_loaded<T, NZ, false>(g, y, x, c, ty, bv, xw, xe, xs, xn);
}
template <class T, int NZ, int MODE>
__global__ __launch_bounds__(256) void k_sweep_typed_lf(Geo g, Typed<T> y, const T* __restrict__ b,
                                                        T* __restrict__ x, int colour, Geo gc,
                                                        const T* __restrict__ xc, State* st,
                                                        const double* __restrict__ cons_part, int cons_nblk,
                                                        int cons_mode) {
    sweep_typed_lf_body<T, NZ, MODE, kColsX>(hw_blk(), g, y, b, x, colour, gc, xc, st, cons_part, cons_nblk, cons_mode);
}

// restrict_sum_fast with the red children's faces from their types (each
// lane its own row).
template <class T>
__device__ __forceinline__ T restrict_sum_fast_typed(const Geo& f, const Typed<T>& y, const T* __restrict__ fx, int I,
                                                     int J, int l) {
    const long base = (long)l * f.plane;
    const bool two = 2 * I + 1 < f.nx && 2 * J + 1 < f.ny;
    T r[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const Col c = make_col(f, 2 * I + (two ? k : 0), 2 * J + (two ? k : 0));
        const TypeRow<T> w = y.row(y.ty[c.c0], l);
        const T Gw = c.hw ? w.v[TF_W] : T(0), Ge = c.he ? w.v[TF_E] : T(0);
        const T Gs = c.hs ? w.v[TF_S] : T(0), Gn = c.hn ? w.v[TF_N] : T(0);
        const T xw = fx[base + c.pw], xe = fx[base + c.pe], xs = fx[base + c.ps], xn = fx[base + c.pn];
        T s = 0;  // absent faces add (+0) * x, as in sweep_column_typed
        s += Gw * xw;
        s += Ge * xe;
        s += Gs * xs;
        s += Gn * xn;
        r[k] = s;
    }
    T sum = 0;
    sum += r[0];
    if (two) sum += r[1];
    return sum;
}

// The lateral restriction of level f into a plane-parallel level g and g's
// first (red, x = 0) sweep in one launch: lane (column of either colour,
// plane) forms its coarse b (restrict_sum_fast's sum, FT: f's faces from its
// column types) and stores it; the red columns then solve as sweep_pp_body
// does in SW_ZERO, whose right-hand side is b itself. One launch and one
// pass over the coarse b less per level; results identical.
template <class T, int LP, bool FT>
__global__ __launch_bounds__(kPPThreads) void k_restrict_zero_pp(Geo f, Ops<T> fo, Typed<T> fy,
                                                                 const T* __restrict__ fx, Geo g, Ops<T> o,
                                                                 const T* __restrict__ finv, T* __restrict__ b,
                                                                 T* __restrict__ x, const State* __restrict__ st) {
    constexpr int CPB = kPPThreads / LP, RS = CPB + 1;
    __shared__ __align__(16) T sh_s[LP * RS], sh_i[LP * RS], sh_u[LP * RS];
    __shared__ int sh_ok[kPPThreads / 4];
    if (st->done) return;
    const int tid = threadIdx.x, l = tid % LP, cb = tid / LP;
    const long half = (long)g.hx * g.ny;             // columns of one colour
    const long qq = (long)blockIdx.x * CPB + cb;     // red columns, then black
    const int colour = qq >= half ? 1 : 0;
    const long q = qq - colour * half;
    const int j = (int)(q / g.hx), t = (int)(q - (long)j * g.hx);
    const int i = 2 * t + ((j + colour) & 1);
    const bool colok = qq < 2 * half && j < g.ny && i < g.nx;
    const bool ok = colok && l < g.nz;
    if (ok) {
        const T sum = FT ? restrict_sum_fast_typed<T>(f, fy, fx, i, j, l)
                         : restrict_sum_fast<T>(f, fo.gE, fo.gN, fx, i, j, l);
        const long id = (long)l * g.plane + g.pos(i, j);
        b[id] = sum;
        if (colour == 0) {
            sh_s[l * RS + cb] = sum;
            sh_i[l * RS + cb] = finv[id];
            sh_u[l * RS + cb] = o.gU[id];
        }
    }
    if (l == 0) sh_ok[cb] = colok && colour == 0;
    __syncthreads();
    if (tid < CPB && sh_ok[tid]) {
        T sv[LP], inv[LP], gu[LP];
#pragma unroll
        for (int k = 0; k < LP; ++k) {
            sv[k] = sh_s[k * RS + tid];
            inv[k] = sh_i[k * RS + tid];
            gu[k] = sh_u[k * RS + tid];
        }
        column_thomas<T, LP>(g.nz, sv, inv, gu);
#pragma unroll
        for (int k = 0; k < LP; ++k) sh_s[k * RS + tid] = sv[k];
    }
    __syncthreads();
    if (ok && colour == 0) x[(long)l * g.plane + g.pos(i, j)] = sh_s[l * RS + cb];
}

// restrict_cell_fast with the red children's faces from their types, for
// kRestrictPlanes planes per thread (grid z: the plane groups): every x of
// the thread's planes is loaded before the first type row is read. At
// 1024^2 x 16 against one plane per thread and the loads after the type:
// 20.4 -> 13.7 us back to back, 33.9 -> 26.5 us with the caches flushed; the
// same sums, so the same values.
constexpr int kRestrictPlanes = 2;
template <class T, int P, int BX = 0>
__device__ __forceinline__ void restrict_fast_typed_body(const Blk& B, const Geo& f, const Typed<T>& y,
                                                         const T* __restrict__ fx, const Geo& cg, T* __restrict__ cb,
                                                         const State* __restrict__ st) {
    if (st->done) return;
    const int I = B.x0<BX>() + threadIdx.x;
    const int J = B.y * blockDim.y + threadIdx.y;
    if (I >= cg.nx || J >= cg.ny) return;
    const bool two = 2 * I + 1 < f.nx && 2 * J + 1 < f.ny;
    Col c[2];
    int t[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        c[k] = make_col(f, 2 * I + (two ? k : 0), 2 * J + (two ? k : 0));
        t[k] = y.ty[c[k].c0];
    }
    const int l0 = B.z * P;
    T xv[P][2][4];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int l = l0 + p < f.nz ? l0 + p : f.nz - 1;  // past nz: the top plane again, unused
        const long base = (long)l * f.plane;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            xv[p][k][0] = fx[base + c[k].pw];
            xv[p][k][1] = fx[base + c[k].pe];
            xv[p][k][2] = fx[base + c[k].ps];
            xv[p][k][3] = fx[base + c[k].pn];
        }
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int l = l0 + p;
        if (l >= f.nz) break;
        T r[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const TypeRow<T> w = y.row(t[k], wave_rep_type(t[k]), l);
            const T Gw = c[k].hw ? w.v[TF_W] : T(0), Ge = c[k].he ? w.v[TF_E] : T(0);
            const T Gs = c[k].hs ? w.v[TF_S] : T(0), Gn = c[k].hn ? w.v[TF_N] : T(0);
            T s = 0;  // absent faces add (+0) * x, as in sweep_column_typed
            s += Gw * xv[p][k][0];
            s += Ge * xv[p][k][1];
            s += Gs * xv[p][k][2];
            s += Gn * xv[p][k][3];
            r[k] = s;
        }
        T sum = 0;
        sum += r[0];
        if (two) sum += r[1];
        cb[(long)l * cg.plane + cg.pos(I, J)] = sum;
    }
}
template <class T>
__global__ __launch_bounds__(256) void k_restrict_fast_typed(Geo f, Typed<T> y, const T* __restrict__ fx, Geo cg,
                                                             T* __restrict__ cb, const State* __restrict__ st) {
    restrict_fast_typed_body<T, kRestrictPlanes, kColsX>(hw_blk(), f, y, fx, cg, cb, st);
}

// ---- coarse tail: the lowest levels of the V-cycle in one workgroup --------

template <class T>
struct TailLevel {
    Geo g;
    Ops<T> o;
    const T* finv;
    T* x;
    T* b;
};
template <class T>
struct Tail {
    int n;  // levels in the tail; L[n-1] is the single column
    TailLevel<T> L[kMaxTail];
};

template <class T, int NZ, int MODE, int TT>
__device__ __forceinline__ void tail_sweep(const TailLevel<T>& L, int colour, const TailLevel<T>& C) {
    const int cols = L.g.hx * L.g.ny;
    for (int q = threadIdx.x; q < cols; q += TT) {
        const int t = q % L.g.hx, j = q / L.g.hx;
        const int i = 2 * t + ((j + colour) & 1);
        if (i < L.g.nx)
            sweep_column<T, NZ, MODE>(L.g, L.o.gE, L.o.gN, L.o.gU, L.finv, L.b, L.x, i, j, C.g, C.x);
    }
    __syncthreads();
}

// Threads of the tail workgroup: 1024 where the column registers fit in
// 128 VGPRs (FP32, nz <= 32), else 512.
template <class T, int NZ>
constexpr int tail_threads() {
    return (sizeof(T) == 4 && NZ <= 32) ? 1024 : 512;
}

// vcycle(k) for the levels of the tail (pcg3d_cpu.cpp:69-113); its first
// level's b has been restricted by the caller (or, for a grid of at most
// 32 x 32 columns, written by the PCG kernel whose convergence test the tail
// then carries: cons_mode).
// The V-cycle body over levels L[0..n-1] of a tail (global or LDS copies).
template <class T, int NZ, int TT>
__device__ void tail_vcycle(const TailLevel<T>* L, int n, T inv_dt, int fast) {
    for (int k = 0; k + 1 < n; ++k) {
        const TailLevel<T>&F = L[k], &C = L[k + 1];
        tail_sweep<T, NZ, SW_ZERO, TT>(F, 0, C);
        tail_sweep<T, NZ, SW_PLAIN, TT>(F, 1, C);
        const int cells = C.g.nx * C.g.ny * C.g.nz;
        for (int q = threadIdx.x; q < cells; q += TT) {
            const int I = q % C.g.nx, rest = q / C.g.nx;
            if (fast)
                restrict_cell_fast<T>(F.g, F.o.gE, F.o.gN, F.x, C.g, C.b, I, rest % C.g.ny, rest / C.g.ny);
            else
                restrict_cell<T>(F.g, F.o, F.b, F.x, C.g, C.b, I, rest % C.g.ny, rest / C.g.ny, inv_dt);
        }
        __syncthreads();
    }
    tail_sweep<T, NZ, SW_ZERO, TT>(L[n - 1], 0, L[n - 1]);
    for (int k = n - 2; k >= 0; --k) {
        tail_sweep<T, NZ, SW_PROLONG, TT>(L[k], 1, L[k + 1]);
        tail_sweep<T, NZ, SW_PLAIN, TT>(L[k], 0, L[k + 1]);
    }
}

template <class T, int NZ>
__global__ __launch_bounds__((tail_threads<T, NZ>())) void k_tail(Tail<T> t, State* st,
                                                                 const double* __restrict__ cons_part,
                                                                 int cons_nblk, int cons_mode, int fast) {
    constexpr int TT = tail_threads<T, NZ>();
    if (st->done) return;
    if (cons_mode != CONS_NONE && consume_convergence(cons_part, cons_nblk, cons_mode, st)) return;
    tail_vcycle<T, NZ, TT>(t.L, t.n, (T)st->inv_dt, fast);
}

// A level copied into the LDS tail keeps its layout except for the plane
// stride: the smallest stride >= the used slots that is 64 / lp modulo 32
// words. The lp lanes of a column read lp planes of it at once; with a
// stride of 2 hx ny (256 words at 16 x 16) they would all fall in one bank,
// with this one the 64 lanes of a wave (64 / lp columns) fall in distinct
// banks pairwise at most.
__host__ __device__ inline Geo tail_lds_geo(Geo g, int lp) {
    const long slots = 2L * g.hx * g.ny, r = (64 / lp) % 32;
    g.plane = slots + (((r - slots) % 32) + 32) % 32;
    return g;
}

// LDS bytes of an LDS-resident tail over levels g[0..n-1] (lp lanes per
// column): per cell gE, gN, gU, 1/pivot, x, b (the lateral restriction reads
// no gB and no capacity).
template <class T>
inline size_t tail_lds_bytes(const Geo* g, int n, int lp) {
    size_t b = 0;
    for (int k = 0; k < n; ++k) b += 6 * sizeof(T) * (size_t)tail_lds_geo(g[k], lp).cells();
    return b;
}

// The same tail with every level copied into LDS first (lateral
// restriction), every sweep plane-parallel: the lanes of a column gather the
// planes' right-hand sides at once, one thread per column runs the Thomas
// recurrence (sweep_column's arithmetic), every lane writes its planes back.
// Level arrays are addressed from the dynamic LDS base, so every access is
// an LDS access. Only the top level's x goes back to global memory, for the
// prolongation into the level above.
template <class T>
struct LdsLevel {
    Geo g;
    T *gE, *gN, *gU, *fi, *x, *b;
};
// The level whose arrays (gE, gN, gU, 1/pivot, x, b, each g.cells() long)
// start `off` values into the dynamic LDS.
template <class T>
__device__ __forceinline__ LdsLevel<T> lds_level_at(const Geo& g, T* sm, long off) {
    const long cells = g.cells();
    T* p = sm + off;
    return LdsLevel<T>{g, p, p + cells, p + 2 * cells, p + 3 * cells, p + 4 * cells, p + 5 * cells};
}

// Values in the sweeps' scratch: LP rows (one per plane) of a pass's columns
// plus one. FP32 levels with more columns of a colour than one pass of LP
// lanes per column holds run with LP / 2 lanes per column (two planes each),
// so that a 16 x 16 level is one pass: 2048 + LP values.
template <class T>
constexpr int tail_scratch() {
    return sizeof(T) == 4 ? 2 * kTailLdsThreads + 16 : kTailLdsThreads + 64;
}

// One colour sweep of an LDS level with LPV lanes per column: lane l0 of a
// column gathers planes l0, l0 + LPV, ... (LP / LPV of them), so a pass
// covers kTailLdsThreads / LPV columns; thread k of the pass then solves
// column k from the scratch.
template <class T, int LP, int LPV, int MODE>
__device__ __forceinline__ void tail_sweep_pp(const LdsLevel<T>& F, int colour, const LdsLevel<T>& C, T* scr_s,
                                              int* scr_c) {
    constexpr int CPP = kTailLdsThreads / LPV;  // columns per pass
    constexpr int RS = CPP + 1;                  // scratch row stride (rows by plane, see sweep_pp_body)
    constexpr int PPL = LP / LPV;                // planes per lane
    static_assert(LP * RS <= tail_scratch<T>() && CPP <= kTailLdsThreads / 4, "tail scratch too small");
    const int tid = threadIdx.x, l0 = tid % LPV, cp = tid / LPV;
    const Geo& g = F.g;
    const int ncols = g.hx * g.ny;
    for (int q0 = 0; q0 < ncols; q0 += CPP) {
        const int q = q0 + cp;
        const int j = q / g.hx, t = q - j * g.hx;
        const int i = 2 * t + ((j + colour) & 1);
        const bool colok = q < ncols && i < g.nx;
        Col c{};
        if (colok) {
            c = make_col(g, i, j);
            const Parents pa = MODE == SW_PROLONG ? parents(C.g, c, i, j) : Parents{0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < PPL; ++r) {
                const int l = l0 + r * LPV;
                if (l < g.nz) scr_s[l * RS + cp] = column_rhs<T, MODE>(g, F.gE, F.gN, F.b, F.x, c, l, C.g, C.x, pa);
            }
        }
        if (l0 == 0) scr_c[cp] = colok ? c.c0 : -1;
        __syncthreads();
        // Thread k < CPP solves column k of the pass (one wave at LPV = 16).
        const int c0 = tid < CPP ? scr_c[tid] : -1;
        if (c0 >= 0) {
            T* sc = scr_s + tid;  // plane k at sc[k * RS]
            if constexpr (sizeof(T) == 4) {
                T sv[LP], iv[LP], uv[LP];
#pragma unroll
                for (int k = 0; k < LP; ++k) {
                    sv[k] = sc[k * RS];
                    if (k < g.nz) {
                        const long id = (long)k * g.plane + c0;
                        iv[k] = F.fi[id];
                        uv[k] = F.gU[id];
                    }
                }
                column_thomas<T, LP>(g.nz, sv, iv, uv);
#pragma unroll
                for (int k = 0; k < LP; ++k) sc[k * RS] = sv[k];
            } else {
                // Three FP64 columns of 16 planes exceed a 1024-thread
                // block's 128 registers: the recurrence (column_thomas's
                // arithmetic, in its order) runs in place in LDS instead.
                const int nz = g.nz;
                T prev = 0;
                for (int k = 0; k < nz; ++k) {
                    const long id = (long)k * g.plane + c0;
                    const T lower = (k > 0) ? -F.gU[id - g.plane] : T(0);
                    prev = (sc[k * RS] - lower * prev) * F.fi[id];
                    sc[k * RS] = prev;
                }
                for (int k = nz - 2; k >= 0; --k) {
                    const long id = (long)k * g.plane + c0;
                    sc[k * RS] = sc[k * RS] - ((-F.gU[id]) * F.fi[id]) * sc[(k + 1) * RS];
                }
            }
        }
        __syncthreads();
        if (colok) {
#pragma unroll
            for (int r = 0; r < PPL; ++r) {
                const int l = l0 + r * LPV;
                if (l < g.nz) F.x[(long)l * g.plane + c.c0] = scr_s[l * RS + cp];
            }
        }
    }
    __syncthreads();
}
// A sweep in one pass where two planes per lane make it one (FP32, LP >= 8).
// At 16 x 16 x 16 one pass instead of two takes each sweep from ~3.6 to ~2.4 us.
template <class T, int LP, int MODE>
__device__ __forceinline__ void tail_sweep_lds(const LdsLevel<T>& F, int colour, const LdsLevel<T>& C, T* scr_s,
                                               int* scr_c) {
    if constexpr (sizeof(T) == 4 && LP >= 8) {
        if (F.g.hx * F.g.ny > kTailLdsThreads / LP) {
            tail_sweep_pp<T, LP, LP / 2, MODE>(F, colour, C, scr_s, scr_c);
            return;
        }
    }
    tail_sweep_pp<T, LP, LP, MODE>(F, colour, C, scr_s, scr_c);
}

// The levels' LDS descriptors are walked level to level (the next one from
// the current one's offset and one geometry read from the tail argument):
// rebuilding each from level 0 read every level's geometry in a chain of
// dependent scalar loads, ~0.8 us per level visit. Staging works in 32-bit
// indices (the tail's levels are small) with a thread's cells of a level
// loaded before any is stored. 16^2 .. 1^2 x 16: 55 -> 47 us per launch,
// results unchanged bit for bit.
template <class T, int NZ>
__device__ __forceinline__ void tail_lds_body(const Tail<T>& t, State* st, const double* __restrict__ cons_part,
                                              int cons_nblk, int cons_mode, T* sm) {
    constexpr int LP = pp_lanes(NZ);
    __shared__ __align__(16) T scr_s[tail_scratch<T>()];
    __shared__ int scr_c[kTailLdsThreads / 4];  // pass columns' c0 (-1: none)
    const int done = st->done;  // in flight with the first level's loads when no test is carried
    if (cons_mode != CONS_NONE) {
        if (done) return;
        if (consume_convergence(cons_part, cons_nblk, cons_mode, st, true)) return;
    }
    {
        long off = 0;
        for (int k = 0; k < t.n; ++k) {
            const TailLevel<T>& G = t.L[k];
            const Geo dg = tail_lds_geo(G.g, LP);
            const int cells = (int)G.g.cells(), gp = (int)G.g.plane, dp = (int)dg.plane, slots = (int)dg.slots();
            const int dc = (int)dg.cells();
            T* p = sm + off;
            constexpr int SPT = 4;
            for (int base = 0; base < cells; base += SPT * kTailLdsThreads) {
                T vE[SPT], vN[SPT], vU[SPT], vF[SPT], vB[SPT];
                int dd[SPT];
#pragma unroll
                for (int r = 0; r < SPT; ++r) {
                    const int q = base + (int)threadIdx.x + r * kTailLdsThreads;
                    dd[r] = -1;
                    if (q < cells) {
                        const int l = q / gp, sl = q - l * gp;
                        if (sl < slots) dd[r] = l * dp + sl;  // pad slots are never read
                        vE[r] = G.o.gE[q];
                        vN[r] = G.o.gN[q];
                        vU[r] = G.o.gU[q];
                        vF[r] = G.finv[q];
                        vB[r] = k == 0 ? G.b[q] : T(0);
                    }
                }
                if (done) return;
#pragma unroll
                for (int r = 0; r < SPT; ++r) {
                    if (dd[r] < 0) continue;
                    p[dd[r]] = vE[r];
                    p[dc + dd[r]] = vN[r];
                    p[2 * dc + dd[r]] = vU[r];
                    p[3 * dc + dd[r]] = vF[r];
                    if (k == 0) p[5 * dc + dd[r]] = vB[r];
                }
            }
            off += 6L * dc;
        }
    }
    __syncthreads();
    const int n = t.n;
    long offF = 0;
    LdsLevel<T> F = lds_level_at<T>(tail_lds_geo(t.L[0].g, LP), sm, 0);
    for (int k = 0; k + 1 < n; ++k) {
        const long offC = offF + 6 * F.g.cells();
        const LdsLevel<T> C = lds_level_at<T>(tail_lds_geo(t.L[k + 1].g, LP), sm, offC);
        tail_sweep_lds<T, LP, SW_ZERO>(F, 0, C, scr_s, scr_c);
        tail_sweep_lds<T, LP, SW_PLAIN>(F, 1, C, scr_s, scr_c);
        const int cells = C.g.nx * C.g.ny * C.g.nz;
        for (int q = threadIdx.x; q < cells; q += kTailLdsThreads) {
            const int I = q % C.g.nx, rest = q / C.g.nx;
            restrict_cell_fast<T>(F.g, F.gE, F.gN, F.x, C.g, C.b, I, rest % C.g.ny, rest / C.g.ny);
        }
        __syncthreads();
        F = C;
        offF = offC;
    }
    tail_sweep_lds<T, LP, SW_ZERO>(F, 0, F, scr_s, scr_c);
    LdsLevel<T> C = F;
    long offC = offF;
    for (int k = n - 2; k >= 0; --k) {
        const Geo fg = tail_lds_geo(t.L[k].g, LP);
        offF = offC - 6 * fg.cells();
        F = lds_level_at<T>(fg, sm, offF);
        tail_sweep_lds<T, LP, SW_PROLONG>(F, 1, C, scr_s, scr_c);
        tail_sweep_lds<T, LP, SW_PLAIN>(F, 0, C, scr_s, scr_c);
        C = F;
        offC = offF;
    }
    // F is level 0 again. Only slots of real columns come back: a padding
    // slot (odd nx: the half-row's last slot in the rows whose colour has no
    // column there) is never written in LDS, so it holds whatever an earlier
    // workgroup left there, and it must stay 0 in global memory: the PCG's
    // r.z and r.z_old run over every slot, and r = 0 there times an Inf or
    // NaN left in LDS would make them NaN.
    const Geo G0 = t.L[0].g;
    const int cells0 = (int)G0.cells(), gp0 = (int)G0.plane, fp0 = (int)F.g.plane, slots0 = (int)G0.slots();
    for (int q = threadIdx.x; q < cells0; q += kTailLdsThreads) {
        const int l = q / gp0, sl = q - l * gp0;
        const int row = sl / G0.hx, ts = sl - row * G0.hx;
        const bool real = sl < slots0 && 2 * ts + (((row >> 1) + (row & 1)) & 1) < G0.nx;
        t.L[0].x[q] = real ? F.x[l * fp0 + sl] : T(0);
    }
}
template <class T, int NZ>
__global__ __launch_bounds__(kTailLdsThreads) void k_tail_lds(Tail<T> t, State* st, const double* __restrict__ cons_part,
                                                   int cons_nblk, int cons_mode) {
    extern __shared__ __align__(16) unsigned char smem[];
    tail_lds_body<T, NZ>(t, st, cons_part, cons_nblk, cons_mode, reinterpret_cast<T*>(smem));
}

}  // namespace th
