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
constexpr size_t kTailLdsBytes = 150 * 1024;  // LDS budget of the LDS-resident tail
constexpr int kTailLdsThreads = 512;            // 256 VGPRs per lane: no spills for nz <= 16

// Exact solve of column (i, j) with its lateral neighbours fixed
// (mg3d.h:67-84), in place in x. MODE: see above. xc/gc: the coarse level
// (SW_PROLONG only).
template <class T, int NZ, int MODE>
__device__ __forceinline__ void sweep_column(const Geo& g, const T* __restrict__ gE, const T* __restrict__ gN,
                                             const T* __restrict__ gU, const T* __restrict__ finv,
                                             const T* __restrict__ b, T* __restrict__ x, int i, int j,
                                             const Geo& gc, const T* __restrict__ xc) {
    const Col c = make_col(g, i, j);
    int qw = 0, qe = 0, qs = 0, qn = 0;
    if (MODE == SW_PROLONG) {  // coarse parents of the 4 neighbours
        qw = gc.pos((c.hw ? i - 1 : i) >> 1, j >> 1);
        qe = gc.pos((c.he ? i + 1 : i) >> 1, j >> 1);
        qs = gc.pos(i >> 1, (c.hs ? j - 1 : j) >> 1);
        qn = gc.pos(i >> 1, (c.hn ? j + 1 : j) >> 1);
    }
    // Pass 1: every plane's right-hand side with the lateral neighbours held
    // fixed. No plane depends on another, so all loads are in flight at once,
    // which is what hides memory latency on the small coarse levels, where a
    // SIMD holds one or two waves.
    T s[NZ], inv[NZ], gu[NZ];
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        if (l < g.nz) {
            const long base = (long)l * g.plane;
            const long id = base + c.c0;
            T v = b[id];
            if (MODE != SW_ZERO) {
                T xw = x[base + c.pw], xe = x[base + c.pe], xsv = x[base + c.ps], xn = x[base + c.pn];
                if (MODE == SW_PROLONG) {
                    const long cb = (long)l * gc.plane;
                    xw = xw + xc[cb + qw];
                    xe = xe + xc[cb + qe];
                    xsv = xsv + xc[cb + qs];
                    xn = xn + xc[cb + qn];
                }
                const T Gw = gE[base + c.pw], Ge = gE[id], Gs = gN[base + c.ps], Gn = gN[id];
                T ls = 0;
                if (c.hw) ls += Gw * xw;
                if (c.he) ls += Ge * xe;
                if (c.hs) ls += Gs * xsv;
                if (c.hn) ls += Gn * xn;
                v = v + ls;
            }
            s[l] = v;
            inv[l] = finv[id];
            gu[l] = gU[id];
        }
    }
    // Pass 2: Thomas forward elimination and back substitution in registers,
    // upper factor rebuilt as (-gU) * (1/pivot).
    T prev = 0;
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        if (l < g.nz) {
            const T lower = (l > 0) ? -gu[l - 1] : T(0);
            prev = (s[l] - lower * prev) * inv[l];
            s[l] = prev;
        }
    }
#pragma unroll
    for (int l = NZ - 2; l >= 0; --l)
        if (l < g.nz - 1) s[l] = s[l] - ((-gu[l]) * inv[l]) * s[l + 1];
#pragma unroll
    for (int l = 0; l < NZ; ++l)
        if (l < g.nz) x[(long)l * g.plane + c.c0] = s[l];
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
__device__ __forceinline__ void restrict_cell_fast(const Geo& f, const T* __restrict__ gE, const T* __restrict__ gN,
                                                   const T* __restrict__ fx, const Geo& cg, T* __restrict__ cb, int I,
                                                   int J, int l) {
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
    cb[(long)l * cg.plane + cg.pos(I, J)] = sum;
}

// Block (B.x, B.y) covers kColsX x blockDim.y coarse cells of plane B.z.
template <class T>
__device__ __forceinline__ void restrict_fast_body(const Blk& B, const Geo& f, const Ops<T>& o,
                                                   const T* __restrict__ fx, const Geo& cg, T* __restrict__ cb,
                                                   const State* __restrict__ st) {
    if (st->done) return;
    const int I = B.x * kColsX + threadIdx.x;
    const int J = B.y * blockDim.y + threadIdx.y;
    if (I >= cg.nx || J >= cg.ny) return;
    restrict_cell_fast<T>(f, o.gE, o.gN, fx, cg, cb, I, J, B.z);
}
template <class T>
__global__ __launch_bounds__(256) void k_restrict_fast(Geo f, Ops<T> o, const T* __restrict__ fx, Geo cg,
                                                        T* __restrict__ cb, const State* __restrict__ st) {
    restrict_fast_body<T>(hw_blk(), f, o, fx, cg, cb, st);
}

// One colour sweep of a level: block (B.x, B.y) covers kColsX columns of the
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
    const int t = B.x * kColsX + threadIdx.x;
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

// Column factors (mg3d.h:46-61) for every column of a level: 1/pivot.
template <class T>
__global__ __launch_bounds__(256) void k_factor(Geo g, Ops<T> o, T* __restrict__ finv, const State* __restrict__ st) {
    const int q = blockIdx.x * kColsX + threadIdx.x;
    const int j = blockIdx.y * kColsY + threadIdx.y;
    int i;
    if (!st->refactor || j >= g.ny || q >= 2 * g.hx || !slot_column(g, q, j, i)) return;
    const Col c = make_col(g, i, j);
    const T inv_dt = (T)st->inv_dt;
    T cp_prev = 0;
    for (int l = 0; l < g.nz; ++l) {
        const long base = (long)l * g.plane, id = base + c.c0;
        T d = 0;
        if (c.hw) d += o.gE[base + c.pw];
        if (c.he) d += o.gE[id];
        if (c.hs) d += o.gN[base + c.ps];
        if (c.hn) d += o.gN[id];
        d += o.gU[id];
        d += (l > 0) ? o.gU[id - g.plane] : o.gB[c.c0];
        d = d + o.cap[id] * inv_dt;
        const T lower = (l > 0) ? -o.gU[id - g.plane] : T(0);
        const T upper = (l < g.nz - 1) ? -o.gU[id] : T(0);
        const T inv = T(1) / (d - lower * cp_prev);
        finv[id] = inv;
        cp_prev = upper * inv;
    }
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

// LDS bytes of an LDS-resident tail over levels g[0..n-1]: per cell gE, gN,
// gU, 1/pivot, x, b (the fast restriction reads no gB and no capacity).
inline size_t tail_lds_bytes(const Geo* g, int n) {
    size_t b = 0;
    for (int k = 0; k < n; ++k) b += 6 * sizeof(float) * (size_t)g[k].cells();
    return b;
}

// The same tail with every level copied into LDS first (FP32, fast
// restriction): the 20-odd barrier-separated phases then wait on LDS
// (~100-cycle) instead of L2 latency. Only the top level's x goes back to
// global memory, for the prolongation into the level above.
template <int NZ>
__device__ __forceinline__ void tail_lds_body(const Tail<float>& t, State* st, const double* __restrict__ cons_part,
                                              int cons_nblk, int cons_mode, float* sm, TailLevel<float>* S) {
    if (st->done) return;
    if (cons_mode != CONS_NONE && consume_convergence(cons_part, cons_nblk, cons_mode, st, true)) return;
    size_t off = 0;
    for (int k = 0; k < t.n; ++k) {
        const TailLevel<float>& G = t.L[k];
        const long cells = G.g.cells();
        float* gE = sm + off;
        float* gN = gE + cells;
        float* gU = gN + cells;
        float* fi = gU + cells;
        float* x = fi + cells;
        float* b = x + cells;
        off += 6 * (size_t)cells;
        for (long q = threadIdx.x; q < cells; q += kTailLdsThreads) {
            gE[q] = G.o.gE[q];
            gN[q] = G.o.gN[q];
            gU[q] = G.o.gU[q];
            fi[q] = G.finv[q];
            if (k == 0) b[q] = G.b[q];
        }
        if (threadIdx.x == 0) S[k] = TailLevel<float>{G.g, Ops<float>{gE, gN, gU, nullptr, nullptr}, fi, x, b};
    }
    __syncthreads();
    tail_vcycle<float, NZ, kTailLdsThreads>(S, t.n, 0.0f, 1);
    const long cells0 = t.L[0].g.cells();
    for (long q = threadIdx.x; q < cells0; q += kTailLdsThreads) t.L[0].x[q] = S[0].x[q];
}
template <int NZ>
__global__ __launch_bounds__(kTailLdsThreads) void k_tail_lds(Tail<float> t, State* st, const double* __restrict__ cons_part,
                                                   int cons_nblk, int cons_mode) {
    extern __shared__ float sm[];
    __shared__ TailLevel<float> S[kMaxTail];  // level views into LDS (not in scratch)
    tail_lds_body<NZ>(t, st, cons_part, cons_nblk, cons_mode, sm, S);
}

}  // namespace th
