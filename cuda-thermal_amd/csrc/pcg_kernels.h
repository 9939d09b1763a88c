// pcg_kernels.h - fine-grid PCG kernels in FP64 (src/pcg3d_gpu.cu:90-264 and
// pcg3d_cpu.cpp:126-185 in what they compute).
//
// Stencil kernels run one thread per column (both colours of a row slot) and
// walk the planes with the vertical neighbours in registers. Every reduction
// writes one partial per block. On the multigrid path the next kernel adds
// them (every block, in block order) and forms its scalar itself: alpha in
// k_update, the convergence test at the head of the V-cycle, beta in
// k_pupdate. The Jacobi path and the checks use a k_finalize launch.
#pragma once
#include "kernels.h"

namespace th {

// RES_TRUE: RES_MG's residual and partials without its writes (the
// true-residual test at the end of a solve, whose r and b0 nothing reads: a
// restart recomputes them in its start graph).
// RES_RHS: a backward-Euler step's start in one pass: b = P + C/dt x
// (k_rhs_transient's arithmetic, stored to bw) and then RES_MG from it.
enum { RES_MG = 0, RES_JACOBI = 1, RES_CHECK = 2, RES_TRUE = 3, RES_RHS = 4 };
enum { UP_MG = 0, UP_MG_FLEX = 1, UP_JACOBI = 2 };
enum {
    FIN_RES = 0,
    FIN_RES_JACOBI,
    FIN_CHECK,
    FIN_ALPHA,
    FIN_UPDATE,
    FIN_UPDATE_FLEX,
    FIN_UPDATE_JACOBI,
    FIN_DOT_FIRST,
    FIN_DOT,
    FIN_ENERGY
};

// Start (or restart) of a solve: r = b - (A + C/dt) x in flux form, with
// b0 = (T) r for the V-cycle (MG) or z = p = r / diag (Jacobi); partials
// r.r, b.b (and r.z). RES_CHECK: the true residual ||P - A x|| of the steady
// operator, no writes.
template <class T, int MODE>
__device__ __forceinline__ void residual_body(const Blk& B, const Geo& g, const Ops<T>& o, const double* __restrict__ b,
                                              const double* __restrict__ x, double* __restrict__ r,
                                              T* __restrict__ b0, const double* __restrict__ invd,
                                              double* __restrict__ z, double* __restrict__ p,
                                              const State* __restrict__ st, double* __restrict__ part,
                                              double* __restrict__ bw = nullptr) {
    if (MODE != RES_CHECK && st->done) return;
    const int q = B.x * blockDim.x + threadIdx.x;
    const int j = B.y * blockDim.y + threadIdx.y;
    double v[3] = {0.0, 0.0, 0.0};  // rr, bb, rz
    const double inv_dt = MODE == RES_CHECK ? 0.0 : st->inv_dt;
    // x = 0 (multigrid start after set_x_zero): every flux term is +-0, so
    // r = b exactly and x is not read.
    const bool xz = MODE == RES_MG && st->xz;
    int i;
    if (j < g.ny && q < 2 * g.hx && slot_column(g, q, j, i)) {
        const Col c = make_col(g, i, j);
        for (int l = 0; l < g.nz; ++l) {
            const long id = (long)l * g.plane + c.c0;
            const double bi = MODE == RES_RHS ? b[id] + (double)o.cap[id] * inv_dt * x[id] : b[id];
            if (MODE == RES_RHS) bw[id] = bi;
            double ri = bi;
            if (!xz) ri = bi - flux_apply_cell(g, o.gE, o.gN, o.gU, o.gB, o.cap, x, c, l, inv_dt);
            v[0] += ri * ri;
            v[1] += bi * bi;
            if (MODE == RES_MG || MODE == RES_RHS) {
                r[id] = ri;
                b0[id] = (T)ri;
            } else if (MODE == RES_JACOBI) {
                const double zi = ri * invd[id];
                r[id] = ri;
                z[id] = zi;
                p[id] = zi;
                v[2] += ri * zi;
            }
        }
    }
    block_partials<3>(v, part, B.flat());
}
template <class T, int MODE>
__global__ __launch_bounds__(1024) void k_residual(Geo g, Ops<T> o, const double* __restrict__ b,
                                                   const double* __restrict__ x, double* __restrict__ r,
                                                   T* __restrict__ b0, const double* __restrict__ invd,
                                                   double* __restrict__ z, double* __restrict__ p,
                                                   const State* __restrict__ st, double* __restrict__ part,
                                                   double* __restrict__ bw) {
    residual_body<T, MODE>(hw_blk(), g, o, b, x, r, b0, invd, z, p, st, part, bw);
}

// residual_body for a level with column types (k_residual_typed): the same
// flux-form arithmetic per cell, walking the column with the planes below and
// above in registers and the faces from the column's type (UNI: through the
// scalar cache). Absent lateral faces add (+0) * dx: f starts at +0 and never
// becomes -0, so the sums are flux_apply_cell's. XZ: x = 0, r = b (x is not
// read). The per-column sums are k_residual's; the blocks are 64 x 4.
template <class T, int NZ, int MODE, bool TR, bool UNI, bool XZ>
__device__ __forceinline__ void residual_column_typed(const Geo& g, const Typed<T>& y, const double* __restrict__ b,
                                                      const double* __restrict__ x, double* __restrict__ r,
                                                      T* __restrict__ b0, const Col& c, int t, double inv_dt,
                                                      double (&v)[3], double* __restrict__ bw) {
    const int nz = g.nz;
    if (XZ) {
#pragma unroll
        for (int l = 0; l < NZ; ++l) {
            const int lc = l < nz ? l : nz - 1;
            const long id = (long)lc * g.plane + c.c0;
            const double bi = b[id];
            if (plane_live<NZ>(l, nz)) {
                v[0] += bi * bi;
                v[1] += bi * bi;
                if (MODE == RES_MG) {
                    r[id] = bi;
                    b0[id] = (T)bi;
                }
            }
        }
        return;
    }
    double xd = 0.0, xc = x[c.c0];
    double gdn = (double)(UNI ? y.urow(t, 0) : y.row(t, 0)).v[TF_B];
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        const int lc = l < nz ? l : nz - 1;
        const bool up = l < nz - 1;
        const long base = (long)lc * g.plane, id = base + c.c0;
        const TypeRow<T> rw = UNI ? y.urow(t, lc) : y.row(t, lc);
        // RES_RHS: k_rhs_transient's b, with C from the type (the same float)
        const double bi = MODE == RES_RHS ? b[id] + (double)rw.v[TF_C] * inv_dt * xc : b[id];
        const double xu = x[id + (up ? g.plane : 0)];
        const double xw = x[base + c.pw], xe = x[base + c.pe], xs = x[base + c.ps], xn = x[base + c.pn];
        const double Gw = c.hw ? (double)rw.v[TF_W] : 0.0, Ge = c.he ? (double)rw.v[TF_E] : 0.0;
        const double Gs = c.hs ? (double)rw.v[TF_S] : 0.0, Gn = c.hn ? (double)rw.v[TF_N] : 0.0;
        const double Gu = (double)rw.v[TF_U];
        double f = 0.0;
        f += Gw * (xc - xw);
        f += Ge * (xc - xe);
        f += Gs * (xc - xs);
        f += Gn * (xc - xn);
        f += (l > 0) ? gdn * (xc - xd) : gdn * xc;  // below, or bottom ambient
        f += up ? Gu * (xc - xu) : Gu * xc;         // above, or the heat sink
        if (TR) f += (double)rw.v[TF_C] * inv_dt * xc;
        const double ri = bi - f;
        if (plane_live<NZ>(l, nz)) {
            v[0] += ri * ri;
            v[1] += bi * bi;
            if (MODE == RES_RHS) bw[id] = bi;
            if (MODE == RES_MG || MODE == RES_RHS) {
                r[id] = ri;
                b0[id] = (T)ri;
            }
        }
        xd = xc;
        xc = xu;
        gdn = Gu;
    }
}
// ROWS: blocks of 64 x (m kColsY) threads whose partials are those of m
// blocks of 64 x kColsY (row_group_partials).
template <class T, int NZ, int MODE, bool ROWS = false>
__device__ __forceinline__ void residual_typed_body(const Blk& B, const Geo& g, const Typed<T>& y,
                                                    const double* __restrict__ b, const double* __restrict__ x,
                                                    double* __restrict__ r, T* __restrict__ b0,
                                                    const State* __restrict__ st, double* __restrict__ part,
                                                    double* __restrict__ bw) {
    if (MODE != RES_CHECK && st->done) return;
    const int q = B.x * blockDim.x + threadIdx.x;
    const int j = B.y * blockDim.y + threadIdx.y;
    double v[3] = {0.0, 0.0, 0.0};  // rr, bb, (rz)
    const double inv_dt = MODE == RES_CHECK ? 0.0 : st->inv_dt;
    const bool xz = MODE == RES_MG && st->xz;
    int i;
    if (j < g.ny && q < 2 * g.hx && slot_column(g, q, j, i)) {
        const Col c = make_col(g, i, j);
        const int ty = y.ty[c.c0];
        const int rep = wave_rep_type(ty);
        if (xz)
            residual_column_typed<T, NZ, MODE, false, true, true>(g, y, b, x, r, b0, c, 0, inv_dt, v, bw);
        else if (__all(ty == rep)) {
            if (inv_dt != 0.0) residual_column_typed<T, NZ, MODE, true, true, false>(g, y, b, x, r, b0, c, rep, inv_dt, v, bw);
            else residual_column_typed<T, NZ, MODE, false, true, false>(g, y, b, x, r, b0, c, rep, inv_dt, v, bw);
        } else {
            if (inv_dt != 0.0) residual_column_typed<T, NZ, MODE, true, false, false>(g, y, b, x, r, b0, c, ty, inv_dt, v, bw);
            else residual_column_typed<T, NZ, MODE, false, false, false>(g, y, b, x, r, b0, c, ty, inv_dt, v, bw);
        }
    }
    if constexpr (ROWS) row_group_partials<3>(v, part, B, (g.ny + kColsY - 1) / kColsY);
    else block_partials<3>(v, part, B.flat_xy());
}
template <class T, int NZ, int MODE>
__global__ __launch_bounds__(256) void k_residual_typed(Geo g, Typed<T> y, const double* __restrict__ b,
                                                         const double* __restrict__ x, double* __restrict__ r,
                                                         T* __restrict__ b0, const State* __restrict__ st,
                                                         double* __restrict__ part, double* __restrict__ bw) {
    residual_typed_body<T, NZ, MODE>(hw_blk(), g, y, b, x, r, b0, st, part, bw);
}

// Ap = (A + C/dt) p; partial p.Ap.
// One column of Ap and its p.Ap terms. TR: the C/dt term is in the diagonal
// (steady solves never load the capacities).
template <class T, int NZ, class PT, bool TR>
__device__ __forceinline__ double apply_column(const Geo& g, const Ops<T>& o, const PT* __restrict__ p,
                                               double* __restrict__ Ap, const Col& c, double inv_dt) {
    const int nz = g.nz;
    const T* __restrict__ gE = o.gE;
    const T* __restrict__ gN = o.gN;
    const T* __restrict__ gU = o.gU;
    const T* __restrict__ cap = o.cap;
    double v = 0.0;
    double pb = 0.0, pc = (double)p[c.c0], gdn = (double)o.gB[c.c0];
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        if (l < nz) {
            const long base = (long)l * g.plane, id = base + c.c0;
            const double pu = (l < nz - 1) ? (double)p[id + g.plane] : 0.0;
            const double gu = (double)gU[id];
            const double Gw = (double)gE[base + c.pw], Ge = (double)gE[id];
            const double Gs = (double)gN[base + c.ps], Gn = (double)gN[id];
            const double xw = (double)p[base + c.pw], xe = (double)p[base + c.pe], xs = (double)p[base + c.ps],
                         xn = (double)p[base + c.pn];
            double d = 0.0, s = 0.0;
            if (c.hw) d += Gw;
            if (c.he) d += Ge;
            if (c.hs) d += Gs;
            if (c.hn) d += Gn;
            d += gu;
            d += gdn;  // gU of the plane below, or gB on the bottom plane
            if (TR) d = d + (double)cap[id] * inv_dt;
            if (c.hw) s += Gw * xw;
            if (c.he) s += Ge * xe;
            if (c.hs) s += Gs * xs;
            if (c.hn) s += Gn * xn;
            if (l > 0) s += gdn * pb;
            if (l < nz - 1) s += gu * pu;
            const double a = d * pc - s;
            Ap[id] = a;
            v += pc * a;
            pb = pc;
            pc = pu;
            gdn = gu;
        }
    }
    return v;
}
template <class T, int NZ, class PT = double>
__device__ __forceinline__ void apply_body(const Blk& B, const Geo& g, const Ops<T>& o, const PT* __restrict__ p,
                                           double* __restrict__ Ap, const State* __restrict__ st,
                                           double* __restrict__ part) {
    if (st->done) return;
    const int q = B.x * blockDim.x + threadIdx.x;
    const int j = B.y * blockDim.y + threadIdx.y;
    double v[1] = {0.0};
    const double inv_dt = st->inv_dt;
    int i;
    if (j < g.ny && q < 2 * g.hx && slot_column(g, q, j, i)) {
        const Col c = make_col(g, i, j);
        v[0] = inv_dt != 0.0 ? apply_column<T, NZ, PT, true>(g, o, p, Ap, c, inv_dt)
                             : apply_column<T, NZ, PT, false>(g, o, p, Ap, c, inv_dt);
    }
    block_partials<1>(v, part, B.flat());
}
// apply_column with the faces from the column's type (fine level). UNI: the
// wave's columns all have type t, read through the scalar cache. Planes past
// nz (NZ is a size bucket) and the plane above the top one re-read planes
// that exist, so every load is unconditional and the loads of all planes can
// be in flight together; the values they give are never used.
template <class T, int NZ, class PT, bool TR, bool UNI>
__device__ __forceinline__ double apply_column_typed(const Geo& g, const Typed<T>& y, const PT* __restrict__ p,
                                                     double* __restrict__ Ap, const Col& c, int t, double inv_dt) {
    const int nz = g.nz;
    double v = 0.0;
    double pb = 0.0, pc = (double)p[c.c0];
    double gdn = (double)(UNI ? y.urow(t, 0) : y.row(t, 0)).v[TF_B];
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        const int lc = l < nz ? l : nz - 1;
        const bool up = l < nz - 1;
        const long base = (long)lc * g.plane, id = base + c.c0;
        const TypeRow<T> r = UNI ? y.urow(t, lc) : y.row(t, lc);
        const double pu_ld = (double)p[id + (up ? g.plane : 0)];
        const double pu = up ? pu_ld : 0.0;
        const double gu = (double)r.v[TF_U];
        // Absent faces contribute (+0) and (+0) * x: d and s start at +0 and
        // never become -0, so both sums equal the ones with those terms
        // skipped, and no load waits behind a branch.
        const double Gw = c.hw ? (double)r.v[TF_W] : 0.0, Ge = c.he ? (double)r.v[TF_E] : 0.0;
        const double Gs = c.hs ? (double)r.v[TF_S] : 0.0, Gn = c.hn ? (double)r.v[TF_N] : 0.0;
        const double xw = (double)p[base + c.pw], xe = (double)p[base + c.pe], xs = (double)p[base + c.ps],
                     xn = (double)p[base + c.pn];
        double d = 0.0, s = 0.0;
        d += Gw;
        d += Ge;
        d += Gs;
        d += Gn;
        d += gu;
        d += gdn;
        if (TR) d = d + (double)r.v[TF_C] * inv_dt;
        s += Gw * xw;
        s += Ge * xe;
        s += Gs * xs;
        s += Gn * xn;
        if (l > 0) s += gdn * pb;
        if (up) s += gu * pu;
        const double a = d * pc - s;
        if (plane_live<NZ>(l, nz)) {
            Ap[id] = a;
            v += pc * a;
        }
        pb = pc;
        pc = pu;
        gdn = gu;
    }
    return v;
}
template <class T, int NZ, class PT, bool ROWS = false>  // ROWS: as residual_typed_body
__device__ __forceinline__ void apply_typed_body(const Blk& B, const Geo& g, const Typed<T>& y,
                                                 const PT* __restrict__ p, double* __restrict__ Ap,
                                                 const State* __restrict__ st, double* __restrict__ part) {
    if (st->done) return;
    const int q = B.x * blockDim.x + threadIdx.x;
    const int j = B.y * blockDim.y + threadIdx.y;
    double v[1] = {0.0};
    const double inv_dt = st->inv_dt;
    int i;
    if (j < g.ny && q < 2 * g.hx && slot_column(g, q, j, i)) {
        const Col c = make_col(g, i, j);
        const int ty = y.ty[c.c0];
        const int rep = wave_rep_type(ty);
        if (__all(ty == rep))
            v[0] = inv_dt != 0.0 ? apply_column_typed<T, NZ, PT, true, true>(g, y, p, Ap, c, rep, inv_dt)
                                 : apply_column_typed<T, NZ, PT, false, true>(g, y, p, Ap, c, rep, inv_dt);
        else
            v[0] = inv_dt != 0.0 ? apply_column_typed<T, NZ, PT, true, false>(g, y, p, Ap, c, ty, inv_dt)
                                 : apply_column_typed<T, NZ, PT, false, false>(g, y, p, Ap, c, ty, inv_dt);
    }
    if constexpr (ROWS) row_group_partials<1>(v, part, B, (g.ny + kColsY - 1) / kColsY);
    else block_partials<1>(v, part, B.flat_xy());
}
// 64 x 4 blocks. FP64 faces: the compiler's own register choice (106 us at
// 1024^2 x 16 with the caches flushed, against 116 for 64 x 16 blocks).
// FP32 faces (k_apply_typed_occ): at least 6 waves per SIMD, 62 us against
// 72 us at the compiler's own choice.
template <class T, int NZ, class PT>
__global__ __launch_bounds__(256) void k_apply_typed(Geo g, Typed<T> y, const PT* __restrict__ p,
                                                     double* __restrict__ Ap, const State* __restrict__ st,
                                                     double* __restrict__ part) {
    apply_typed_body<T, NZ, PT>(hw_blk(), g, y, p, Ap, st, part);
}
template <class T, int NZ, class PT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_apply_typed_occ(
    Geo g, Typed<T> y, const PT* __restrict__ p, double* __restrict__ Ap, const State* __restrict__ st,
    double* __restrict__ part) {
    apply_typed_body<T, NZ, PT>(hw_blk(), g, y, p, Ap, st, part);
}
template <class T, int NZ, class PT = double>
__global__ __launch_bounds__(1024) void k_apply(Geo g, Ops<T> o, const PT* __restrict__ p,
                                                double* __restrict__ Ap, const State* __restrict__ st,
                                                double* __restrict__ part) {
    apply_body<T, NZ, PT>(hw_blk(), g, o, p, Ap, st, part);
}

// r -= alpha Ap, and b0 = (T) r for the V-cycle (MG) or z = r / diag and
// x += alpha p (Jacobi); partials r.r and r.z_old (flexible beta: z_old is the
// previous V-cycle output, still in x0) or r.z (Jacobi).
// Multigrid path: alpha = rz / p.Ap is formed in every block from the apply
// kernel's partials (apart), so no reduction launch sits between the two;
// x += alpha p is left to the p update, which reads p anyway (xlag marks it
// pending until then; k_finish applies it when the solve stops first).
template <class T, int MODE>
__device__ __forceinline__ void update_body(const Blk& B, long n, const double* __restrict__ p,
                                            const double* __restrict__ Ap, double* __restrict__ x,
                                            double* __restrict__ r, T* __restrict__ b0,
                                            const T* __restrict__ zold, const double* __restrict__ invd,
                                            double* __restrict__ z, State* st, double* __restrict__ part,
                                            const double* __restrict__ apart, int anblk) {
    if (st->done) return;
    double alpha;
    if (apart) {
        double pAp[1];
        block_sum_partials<1>(apart, anblk, pAp);
        alpha = st->rzb[st->iters & 1] / pAp[0];
        if (B.x == 0 && threadIdx.x == 0) {
            st->pAp = pAp[0];
            st->alpha = alpha;
            if (MODE != UP_JACOBI) st->xlag = 1;
        }
    } else {
        alpha = st->alpha;
    }
    double v[2] = {0.0, 0.0};
    const long stride = (long)B.gx * blockDim.x;
    const long k0 = (long)B.x * blockDim.x + threadIdx.x;
    if (MODE == UP_JACOBI) {
        for (long k = k0; k < n; k += stride) {
            x[k] += alpha * p[k];
            const double ri = r[k] - alpha * Ap[k];
            r[k] = ri;
            v[0] += ri * ri;
            const double zi = ri * invd[k];
            z[k] = zi;
            v[1] += ri * zi;
        }
    } else {
        // Two cells per lane (16-byte loads of r and Ap): 98 against 105 us
        // at 1024^2 x 16. A level's planes hold an even number of slots, so n
        // is even; an odd last cell would go to lane 0 of block 0. Ap, r and
        // z_old are read, and r is written, with the nontemporal hint (the
        // V-cycle next reads b, not them; r is next read after it). The loads:
        // 97 -> 86 us back to back, 110 -> 99 us after a cache flush, at
        // 1024^2 x 16; the r store (with the p update's x store) took the
        // whole mixed solve from 9.0 to 8.7 ms.
        typedef typename Vec2<T>::type T2;
        const long n2 = n >> 1;
        for (long k = k0; k < n2; k += stride) {
            const double2 a = load2_nt(Ap, k);
            const double2 ro = load2_nt(r, k);
            double2 ri;
            ri.x = ro.x - alpha * a.x;
            ri.y = ro.y - alpha * a.y;
            store2_nt(r, k, ri);
            v[0] += ri.x * ri.x;
            v[0] += ri.y * ri.y;
            if (MODE == UP_MG_FLEX) {
                const T2 zo = load2_nt(zold, k);
                v[1] += ri.x * (double)zo.x;
                v[1] += ri.y * (double)zo.y;
            }
            T2 bv;
            bv.x = (T)ri.x;
            bv.y = (T)ri.y;
            reinterpret_cast<T2*>(b0)[k] = bv;
        }
        if ((n & 1) && k0 == 0) {
            const long k = n - 1;
            const double ri = r[k] - alpha * Ap[k];
            r[k] = ri;
            v[0] += ri * ri;
            if (MODE == UP_MG_FLEX) v[1] += ri * (double)zold[k];
            b0[k] = (T)ri;
        }
    }
    block_partials<2>(v, part, B.x);
}
template <class T, int MODE>
__global__ __launch_bounds__(kVecThreads) void k_update(long n, const double* __restrict__ p,
                                                         const double* __restrict__ Ap, double* __restrict__ x,
                                                         double* __restrict__ r, T* __restrict__ b0,
                                                         const T* __restrict__ zold, const double* __restrict__ invd,
                                                         double* __restrict__ z, State* st, double* __restrict__ part,
                                                         const double* __restrict__ apart, int anblk) {
    update_body<T, MODE>(hw_blk(), n, p, Ap, x, r, b0, zold, invd, z, st, part, apart, anblk);
}

// Partial r.z with z = the V-cycle output.
template <class T>
__device__ __forceinline__ void dot_body(const Blk& B, long n, const double* __restrict__ r, const T* __restrict__ z,
                                         const State* __restrict__ st, double* __restrict__ part) {
    if (st->done) return;
    double v[1] = {0.0};
    const long stride = (long)B.gx * blockDim.x;
    for (long k = (long)B.x * blockDim.x + threadIdx.x; k < n; k += stride)
        v[0] += load_nt(r, k) * (double)load_nt(z, k);
    block_partials<1>(v, part, B.x);
}
template <class T>
__global__ __launch_bounds__(kVecThreads) void k_dot(long n, const double* __restrict__ r, const T* __restrict__ z,
                                                      const State* __restrict__ st, double* __restrict__ part) {
    dot_body<T>(hw_blk(), n, r, z, st, part);
}

// p = z + beta p (first: p = z). Multigrid path: every block forms r.z from
// the dot kernel's partials (dpart) and beta from it - Fletcher-Reeves
// rz_new / rz_old or flexible (rz_new - r.z_old) / rz_old - with rz kept by
// iteration parity, so blocks never read a value another block rewrites.
// x (multigrid, not FIRST): the update's x += alpha p is applied here, from
// the p it reads anyway. p is stored in PT (FP32 in the mixed solver: x and
// r both move by the rounded p, so r stays the residual of x).
template <class Z, int FIRST, class PT = double>
__device__ __forceinline__ void pupdate_body(const Blk& B, long n, const Z* __restrict__ z, PT* __restrict__ p,
                                             double* __restrict__ x, State* st, const double* __restrict__ dpart,
                                             int dnblk) {
    if (st->done) return;
    double beta = 0.0;
    if (dpart) {
        double rz[1];
        block_sum_partials<1>(dpart, dnblk, rz);
        const int it = st->iters;
        if (!FIRST) {
            const double old = st->rzb[(it + 1) & 1];
            beta = st->flexible ? (rz[0] - st->rzo) / old : rz[0] / old;
        }
        if (B.x == 0 && threadIdx.x == 0) {
            st->rzb[it & 1] = rz[0];
            st->rz = rz[0];
            st->beta = beta;
        }
    } else {
        beta = st->beta;
    }
    const long stride = (long)B.gx * blockDim.x;
    if (!FIRST && x) {
        // Two cells per lane, as in update_body (n even; an odd last cell
        // goes to lane 0 of block 0).
        typedef typename Vec2<PT>::type P2;
        typedef typename Vec2<Z>::type Z2;
        const double alpha = st->alpha;
        const long k0 = (long)B.x * blockDim.x + threadIdx.x, n2 = n >> 1;
        for (long k = k0; k < n2; k += stride) {
            const P2 po = reinterpret_cast<const P2*>(p)[k];
            // z and x with the nontemporal hint (no kernel reads them again
            // before the next p update); p, which the apply reads next, not
            const Z2 zz = load2_nt(z, k);
            double2 xx = load2_nt(x, k);
            xx.x += alpha * (double)po.x;
            xx.y += alpha * (double)po.y;
            store2_nt(x, k, xx);
            P2 pn;
            pn.x = (PT)((double)zz.x + beta * (double)po.x);
            pn.y = (PT)((double)zz.y + beta * (double)po.y);
            reinterpret_cast<P2*>(p)[k] = pn;
        }
        if ((n & 1) && k0 == 0) {
            const long k = n - 1;
            const double po = (double)p[k];
            x[k] += alpha * po;
            p[k] = (PT)((double)z[k] + beta * po);
        }
        if (B.x == 0 && threadIdx.x == 0) st->xlag = 0;
        return;
    }
    for (long k = (long)B.x * blockDim.x + threadIdx.x; k < n; k += stride)
        p[k] = FIRST ? (PT)z[k] : (PT)((double)z[k] + beta * (double)p[k]);
}
template <class Z, int FIRST, class PT = double>
__global__ __launch_bounds__(kVecThreads) void k_pupdate(long n, const Z* __restrict__ z, PT* __restrict__ p,
                                                          double* __restrict__ x, State* st,
                                                          const double* __restrict__ dpart, int dnblk) {
    pupdate_body<Z, FIRST, PT>(hw_blk(), n, z, p, x, st, dpart, dnblk);
}

// Adds the partials of the previous kernel and runs its epilogue.
template <int NV, int EPI, unsigned KMAX = 0>
__global__ __launch_bounds__(kFinThreads) void k_finalize(const double* __restrict__ part, int nblk, State* st) {
    if (EPI != FIN_CHECK && EPI != FIN_ENERGY && st->done) return;
    double tn[NV];
    sum_partials<NV, KMAX>(part, nblk, tn);
    if (threadIdx.x != 0) return;
    double tot[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NV; ++k) tot[k] = tn[k];
    if (EPI == FIN_RES || EPI == FIN_RES_JACOBI) {  // pcg3d_cpu.cpp:141-149
        st->bb = tot[1];
        st->rr = tot[0];
        st->rzo = 0.0;
        if (tot[1] == 0.0) {  // b = 0: x = 0 is the answer
            st->relres = 0.0;
            st->relres0 = 0.0;
            st->done = 1;
            st->converged = 1;
            st->zero_x = 1;
            return;
        }
        st->relres = sqrt(tot[0] / tot[1]);
        st->relres0 = st->relres;
        if (st->relres < st->tol) {
            st->done = 1;
            st->converged = 1;
        }
        if (EPI == FIN_RES_JACOBI) st->rz = tot[2];
    } else if (EPI == FIN_CHECK) {
        st->chk_rr = tot[0];
        st->chk_bb = tot[1];
    } else if (EPI == FIN_ALPHA) {
        st->pAp = tot[0];
        st->alpha = st->rz / tot[0];
    } else if (EPI == FIN_UPDATE || EPI == FIN_UPDATE_FLEX || EPI == FIN_UPDATE_JACOBI) {
        st->rr = tot[0];
        st->iters += 1;
        st->relres = sqrt(tot[0] / st->bb);
        if (st->relres < st->tol) {
            st->done = 1;
            st->converged = 1;
        } else if (st->iters >= st->maxit) {
            st->done = 1;
        }
        if (EPI == FIN_UPDATE_FLEX) st->rzo = tot[1];
        if (EPI == FIN_UPDATE_JACOBI && !st->done) {
            st->beta = tot[1] / st->rz;
            st->rz = tot[1];
        }
    } else if (EPI == FIN_DOT_FIRST) {
        st->rz = tot[0];
    } else if (EPI == FIN_DOT) {  // Fletcher-Reeves, or flexible (r.z - r.z_old) / rz_old
        st->beta = st->flexible ? (tot[0] - st->rzo) / st->rz : tot[0] / st->rz;
        st->rz = tot[0];
    } else if (EPI == FIN_ENERGY) {
        st->e_heat = tot[0];
        st->e_out = tot[1];
        st->e_tmax = tot[2];
    }
}

// The convergence test of an iteration's x, r update on its own, for the
// last iteration of a batch the host queues ahead (its V-cycle, dot and p
// update wait for the host's look at the state, so that a converged solve
// leaves no queue of early-returning launches behind): consume_convergence
// as the head of the V-cycle runs it, with its thread count (256), so the
// sums and the state are the same.
static __global__ __launch_bounds__(256) void k_consume_iter(const double* __restrict__ part, int nblk, State* st) {
    if (st->done) return;
    consume_convergence(part, nblk, CONS_ITER, st, true);
}

// Start of a solve: tolerance, iteration cap, cleared flags.
// xz < 0: keep the device's flag (set by k_lift_decide).
static __global__ void k_begin(State* st, double tol, int maxit, int xz) {
    st->xlag = 0;
    if (xz >= 0) st->xz = xz;
    st->tol = tol;
    st->maxit = maxit;
    st->iters = 0;
    st->done = 0;
    st->converged = 0;
    st->zero_x = 0;
    st->relres = 0.0;
}
// Reopens a finished solve for a restart from the current x (iterations kept).
static __global__ void k_reopen(State* st) {
    st->xlag = 0;
    st->xz = 0;
    st->done = 0;
    st->converged = 0;
}
static __global__ void k_clear_xz(State* st) { st->xz = 0; }
static __global__ void k_set_inv_dt(State* st, double inv_dt) {
    st->inv_dt = inv_dt;
    st->refactor = 1;
}
static __global__ void k_set_flexible(State* st, int f) { st->flexible = f; }

// Transient stepping on the device (transient3d.h:57-84), so that a step is
// a replayable graph whose size of step is not a launch argument.
// k_step_init: t = 0, first dt, stepping rule, heat content of x = 0.
static __global__ void k_step_init(State* st, double dt0, double growth, double dt_max, double t_end) {
    st->t = 0.0;
    st->dt = -1.0;
    st->dt_prev = -1.0;
    st->dt_next = dt0;
    st->growth = growth;
    st->dt_max = dt_max;
    st->t_end = t_end;
    st->heat_prev = 0.0;
    st->steps = 0;
    st->guide_w = 0.0;
    st->guide_w_sum = 0.0;
    st->guide_steps = 0;
}
// dt = min(dt, t_end - t) (transient3d.h:60); a changed dt sets the new
// operator's 1/dt and asks for new factors (set_dt only if dt changed, :61).
// Then tolerance and flags of the step's solve, as k_begin.
static __global__ void k_step_begin(State* st, double tol, int maxit) {
    st->xlag = 0;
    st->xz = 0;
    const double dt = fmin(st->dt_next, st->t_end - st->t);
    st->dt_prev = st->dt;
    st->refactor = dt != st->dt;
    if (st->refactor) {
        st->dt = dt;
        st->inv_dt = 1.0 / dt;
    }
    st->tol = tol;
    st->maxit = maxit;
    st->iters = 0;
    st->done = 0;
    st->converged = 0;
    st->zero_x = 0;
    st->relres = 0.0;
}
// After the step's energy terms (k_finalize FIN_ENERGY): t += dt, the step's
// record with q_stored = (sum C T_new - sum C T_old) / dt, and the next dt
// = min(dt * growth, dt_max) (transient3d.h:73-84).
static __global__ void k_step_end(State* st) {
    const double dt = st->dt;
    st->t += dt;
    st->last.t = st->t;
    st->last.dt = dt;
    st->last.iters = st->iters;
    st->last.T_max = st->e_tmax;
    st->last.q_out = st->e_out;
    st->last.q_stored = (st->e_heat - st->heat_prev) / dt;
    st->heat_prev = st->e_heat;
    st->dt_next = fmin(dt * st->growth, st->dt_max);
    st->steps += 1;
}

// Predicted start of a backward-Euler step, after its right-hand side has
// taken T_old from x: linear extrapolation of the last two steps,
// x = T_old + (dt / dt_prev) (T_old - T_older); xprev keeps T_old. The first
// step (no previous dt) starts from T_old.
static __global__ __launch_bounds__(kVecThreads) void k_predict(long n, double* __restrict__ x, double* __restrict__ xprev,
                                                         const State* __restrict__ st) {
    const double r = st->dt_prev > 0.0 ? st->dt / st->dt_prev : 0.0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const double xo = x[k];
        x[k] = xo + r * (xo - xprev[k]);
        xprev[k] = xo;
    }
}

// Learned start of a backward-Euler step, after its right-hand side b = P +
// C/dt T_old (x = T_old, or the extrapolated start): along the direction
// d = g - x towards the guide field g (the surrogate's predicted field), the
// start x + w d with the smallest residual ||b - M (x + w d)||, M = C/dt + A:
// w = (r0 . M d) / (M d . M d), r0 = b - M x; w <= 0 (no descent along d)
// keeps x. Per cell, in flux form: M d = M g - M x. Partials: r0.Md, Md.Md,
// r0.r0 per block.
template <class T>
__global__ __launch_bounds__(1024) void k_guide_terms(Geo g, Ops<T> o, const double* __restrict__ b,
                                                      const double* __restrict__ x, const double* __restrict__ gd,
                                                      const State* __restrict__ st, double* __restrict__ part) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    double v[3] = {0.0, 0.0, 0.0};
    const double inv_dt = st->inv_dt;
    int i;
    if (j < g.ny && q < 2 * g.hx && slot_column(g, q, j, i)) {
        const Col c = make_col(g, i, j);
        for (int l = 0; l < g.nz; ++l) {
            const long id = (long)l * g.plane + c.c0;
            const double mx = flux_apply_cell(g, o.gE, o.gN, o.gU, o.gB, o.cap, x, c, l, inv_dt);
            const double mg = flux_apply_cell(g, o.gE, o.gN, o.gU, o.gB, o.cap, gd, c, l, inv_dt);
            const double r0 = b[id] - mx, md = mg - mx;
            v[0] += r0 * md;
            v[1] += md * md;
            v[2] += r0 * r0;
        }
    }
    block_partials<3>(v, part, blockIdx.x + gridDim.x * blockIdx.y);
}
// w from k_guide_terms' partials (one block), and the run's tally.
static __global__ __launch_bounds__(kFinThreads) void k_guide_weight(const double* __restrict__ part, int nblk, State* st) {
    double t[3];
    sum_partials<3>(part, nblk, t);
    if (threadIdx.x != 0) return;
    const double w = (t[0] > 0.0 && t[1] > 0.0) ? t[0] / t[1] : 0.0;
    st->guide_w = w;
    if (w > 0.0) {
        st->guide_steps += 1;
        st->guide_w_sum += w;
    }
}
// x += w (g - x).
static __global__ __launch_bounds__(kVecThreads) void k_guide_step(long n, double* __restrict__ x,
                                                            const double* __restrict__ gd,
                                                            const State* __restrict__ st) {
    const double w = st->guide_w;
    if (w == 0.0) return;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) x[k] += w * (gd[k] - x[k]);
}

// x = 0 when b = 0 was detected (the reference's set_x_zero on bb == 0).
template <class PT>
__global__ void k_finish(long n, double* __restrict__ x, const PT* __restrict__ p, const State* __restrict__ st) {
    const long stride = (long)gridDim.x * blockDim.x;
    if (st->zero_x) {
        for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) x[k] = 0.0;
    } else if (st->xlag) {  // the last update's x += alpha p
        const double alpha = st->alpha;
        for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) x[k] += alpha * (double)p[k];
    }
}

// Jacobi: 1 / diag of (A + C/dt) (pcg3d_cpu.cpp:41-48).
template <class T>
__global__ __launch_bounds__(256) void k_invdiag(Geo g, Ops<T> o, double* __restrict__ invd, const State* st) {
    const int q = blockIdx.x * kColsX + threadIdx.x;
    const int j = blockIdx.y * kColsY + threadIdx.y;
    int i;
    if (j >= g.ny || q >= 2 * g.hx || !slot_column(g, q, j, i)) return;
    const Col c = make_col(g, i, j);
    if (!st->refactor) return;
    const double inv_dt = st->inv_dt;
    for (int l = 0; l < g.nz; ++l) {
        const long base = (long)l * g.plane, id = base + c.c0;
        double d = 0;
        if (c.hw) d += (double)o.gE[base + c.pw];
        if (c.he) d += (double)o.gE[id];
        if (c.hs) d += (double)o.gN[base + c.ps];
        if (c.hn) d += (double)o.gN[id];
        d += (double)o.gU[id];
        d += (l > 0) ? (double)o.gU[id - g.plane] : (double)o.gB[c.c0];
        d = d + (double)o.cap[id] * inv_dt;
        invd[id] = 1.0 / d;
    }
}

// b = P + C/dt x (pcg3d_gpu.cu:104-109).
template <class T>
__global__ __launch_bounds__(kVecThreads) void k_rhs_transient(long n, const T* __restrict__ cap, const double* __restrict__ P,
                                                        const double* __restrict__ x, double* __restrict__ b,
                                                        const State* __restrict__ st) {
    const double inv_dt = st->inv_dt;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
        b[k] = P[k] + (double)cap[k] * inv_dt * x[k];
}

// Energy terms of the current x: sum C x (heat content), heat out through the
// top and bottom (stack3d.h:254-263), and max x over cells (not padding).
template <class T>
__global__ __launch_bounds__(kVecThreads) void k_energy(Geo g, Ops<T> o, const double* __restrict__ x,
                                                 double* __restrict__ part) {
    double v[3] = {0.0, 0.0, -1.0e300};
    const long n = g.cells(), top = (long)(g.nz - 1) * g.plane;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const double xv = x[k];
        const long q = k % g.plane;
        int i;
        const bool cell = q < g.slots() && slot_column(g, (int)(q % (2 * g.hx)), (int)(q / (2 * g.hx)), i);
        v[0] += (double)o.cap[k] * xv;
        if (k >= top) v[1] += (double)o.gU[k] * xv;
        if (k < g.plane) v[1] += (double)o.gB[k] * xv;
        if (cell) v[2] = fmax(v[2], xv);
    }
    block_partials<3, 4u>(v, part);
}

// The end of a backward-Euler step in one pass: k_finish (the lagged
// x += alpha p, or x = 0 when b was 0) and k_energy of the completed x, each
// cell's terms added in k_energy's order by the same thread, so the partials
// are k_energy's. Saves the second pass over x and a launch per step.
template <class T, class PT>
__global__ __launch_bounds__(kVecThreads) void k_finish_energy(Geo g, Ops<T> o, double* __restrict__ x,
                                                        const PT* __restrict__ p, const State* __restrict__ st,
                                                        double* __restrict__ part) {
    double v[3] = {0.0, 0.0, -1.0e300};
    const long n = g.cells(), top = (long)(g.nz - 1) * g.plane;
    const long stride = (long)gridDim.x * blockDim.x;
    const int zero = st->zero_x, lag = st->xlag;
    const double alpha = st->alpha;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        double xv;
        if (zero) {
            xv = 0.0;
            x[k] = xv;
        } else if (lag) {
            xv = x[k] + alpha * (double)p[k];
            x[k] = xv;
        } else {
            xv = x[k];
        }
        const long q = k % g.plane;
        int i;
        const bool cell = q < g.slots() && slot_column(g, (int)(q % (2 * g.hx)), (int)(q / (2 * g.hx)), i);
        v[0] += (double)o.cap[k] * xv;
        if (k >= top) v[1] += (double)o.gU[k] * xv;
        if (k < g.plane) v[1] += (double)o.gB[k] * xv;
        if (cell) v[2] = fmax(v[2], xv);
    }
    block_partials<3, 4u>(v, part);
}

// ---- die map: colour image and hottest cell (viewer.cu:53-87) -----------------

// The viewer's palette: five RGB stops, linear in between, s clamped to
// [0, 1] (the same float operations as the reference's colormap, so the bytes
// agree).
__device__ __forceinline__ unsigned palette5(float s) {
    constexpr float kStops[5][3] = {
        {0.f, 0.f, 4.f}, {87.f, 16.f, 110.f}, {188.f, 55.f, 84.f}, {249.f, 142.f, 9.f}, {252.f, 255.f, 164.f}};
    s = s < 0.f ? 0.f : (s > 1.f ? 1.f : s);
    float f = s * 4.f;
    int k = (int)f;
    if (k > 3) k = 3;
    f -= (float)k;
    unsigned rgba = 255u << 24;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float lo = kStops[k][c], hi = kStops[k + 1][c];
        rgba |= (unsigned)(unsigned char)(lo + f * (hi - lo)) << (8 * c);
    }
    return rgba;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned lo = __shfl_xor((unsigned)v, off, 64), hi = __shfl_xor((unsigned)(v >> 32), off, 64);
        const unsigned long long o = ((unsigned long long)hi << 32) | lo;
        v = o > v ? o : v;
    }
    return v;
}

// Column (i, j): the hottest rise v of planes l0..l1 (at least 0), its colour
// on [0, rise_hi] as RGBA bytes (rgba may be null), and the block's largest
// key (bits of v << 32 | j * nx + i): non-negative floats order like their
// bits, so the largest key is the hottest cell, the last one among equals.
// Block: 64 x 4 columns (blockDim 256).
static __global__ __launch_bounds__(256) void k_colorize(Geo g, const double* __restrict__ x, int l0, int l1,
                                                         float rise_hi, unsigned* __restrict__ rgba,
                                                         unsigned long long* __restrict__ part) {
    __shared__ unsigned long long sk[4];
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    unsigned long long key = 0;
    if (i < g.nx && j < g.ny) {
        const int c0 = g.pos(i, j);
        float v = 0.f;
        for (int l = l0; l <= l1; ++l) v = fmaxf(v, (float)x[(long)l * g.plane + c0]);
        const int cell = j * g.nx + i;
        if (rgba) rgba[cell] = palette5(v / rise_hi);
        key = ((unsigned long long)__float_as_uint(v) << 32) | (unsigned)cell;
    }
    key = wave_max_u64(key);
    if ((threadIdx.x & 63) == 0) sk[threadIdx.x >> 6] = key;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = sk[0];
        for (int w = 1; w < 4; ++w) m = sk[w] > m ? sk[w] : m;
        part[blockIdx.y * gridDim.x + blockIdx.x] = m;
    }
}
// Largest of n keys (one block of 256 threads).
static __global__ __launch_bounds__(256) void k_max_key(const unsigned long long* __restrict__ part, int n,
                                                        unsigned long long* __restrict__ out) {
    __shared__ unsigned long long sk[4];
    unsigned long long m = 0;
    for (int k = threadIdx.x; k < n; k += 256) m = part[k] > m ? part[k] : m;
    m = wave_max_u64(m);
    if ((threadIdx.x & 63) == 0) sk[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) m = sk[w] > m ? sk[w] : m;
        out[0] = m > sk[0] ? m : sk[0];
    }
}

// ---- initial guess from a coarse field ----------------------------------------

// Plane-to-channel map of the lift: plane l of the solver's grid takes
// (1 - w[l]) f[lo[l]] + w[l] f[lo[l] + 1] of the coarse field.
struct LiftMap {
    int lo[64];
    float w[64];
};

// The coarse field taken to the solver's planes at every coarse cell, in
// FP64: fz[l][y][x] = (1 - w[l]) f[y][x][lo[l]] + w[l] f[y][x][lo[l] + 1]
// (f[y][x][lo[l]] alone where w[l] = 0).
static __global__ __launch_bounds__(256) void k_lift_z(const float* __restrict__ f, int Hc, int Wc, int C, int nz,
                                                       LiftMap zm, double* __restrict__ fz) {
    const long k = (long)blockIdx.x * 256 + threadIdx.x;
    const long hw = (long)Hc * Wc;
    if (k >= hw * nz) return;
    const int l = (int)(k / hw);
    const long pix = k - l * hw;
    const float* p = f + pix * C + zm.lo[l];
    const double w = zm.w[l];
    fz[k] = w > 0.0 ? (1.0 - w) * p[0] + w * p[1] : (double)p[0];
}

// x(i, j, l) = scale * fz[l] at the centre of cell (i, j), bilinear over the
// Hc x Wc coarse cells (clamped at the edges). One thread per column of
// `colour` (-1: both colours): the bilinear weights are formed once per
// column, and a wave stores each plane as one run of a half-row and reads
// the two coarse rows it needs as short runs of fz. Before a line
// relaxation only the black columns are lifted: the red relaxation solves
// every red column from its black neighbours alone, so lifted red values
// would never be read. 1024^2 x 16 from 256^2: 83 us for both colours with
// 8 gathers per plane and cell from the channel-interleaved field, against
// 23 us for the black columns from fz (plus 7 us for k_lift_z); the same
// values.
static __global__ __launch_bounds__(256) void k_lift(Geo g, const double* __restrict__ fz, int Hc, int Wc,
                                                     double scale, int colour, double* __restrict__ x) {
    const int q0 = blockIdx.x * kColsX + threadIdx.x;
    const int j = blockIdx.y * kColsY + threadIdx.y;
    const int q = colour < 0 ? q0 : colour * g.hx + q0;
    int i;
    if (j >= g.ny || q0 >= (colour < 0 ? 2 * g.hx : g.hx) || !slot_column(g, q, j, i)) return;
    const double xc = fmin(fmax((i + 0.5) * Wc / g.nx - 0.5, 0.0), (double)(Wc - 1));
    const double yc = fmin(fmax((j + 0.5) * Hc / g.ny - 0.5, 0.0), (double)(Hc - 1));
    const int x0 = (int)xc, y0 = (int)yc;
    const int x1 = x0 + 1 < Wc ? x0 + 1 : x0, y1 = y0 + 1 < Hc ? y0 + 1 : y0;
    const double fx = xc - x0, fy = yc - y0;
    const long hw = (long)Hc * Wc;
    const double* r00 = fz + (long)y0 * Wc + x0;
    const double* r01 = fz + (long)y0 * Wc + x1;
    const double* r10 = fz + (long)y1 * Wc + x0;
    const double* r11 = fz + (long)y1 * Wc + x1;
    const long c0 = g.pos(i, j);
    for (int l = 0; l < g.nz; ++l) {
        const long o = l * hw;
        const double v = (1.0 - fy) * ((1.0 - fx) * r00[o] + fx * r01[o]) + fy * ((1.0 - fx) * r10[o] + fx * r11[o]);
        x[(long)l * g.plane + c0] = scale * v;
    }
}

// The learned start judged on the device (the fallback of set_x_lifted
// without a host round trip): the plain start's relative residual (1 for
// x = 0) and the lifted guess's, each from the FIN_CHECK totals of the
// residual kernel just run; the guess is kept if its residual is the
// smaller, and the next solve's start then reads x (xz = 0).
static __global__ void k_lift_plain(State* st, int x_zero) {
    st->lift_rp = x_zero ? 1.0 : sqrt(st->chk_rr / st->chk_bb);
}
static __global__ void k_lift_decide(State* st, int x_zero) {
    const double rl = sqrt(st->chk_rr / st->chk_bb);
    const int keep = rl < st->lift_rp;
    st->lift_rl = rl;
    st->lift_used = keep;
    st->xz = keep ? 0 : x_zero;
}
// x = v where the guess was kept.
static __global__ void k_copy_if_kept(long n, const double* __restrict__ v, double* __restrict__ x,
                                      const State* __restrict__ st) {
    if (!st->lift_used) return;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) x[k] = v[k];
}

// One colour of a column line relaxation of the steady operator in FP64:
// each column of `colour` is solved exactly (Thomas) with its lateral
// neighbours fixed at their current values and right-hand side b.
template <class T, int NZ>
__global__ __launch_bounds__(256) void k_line_relax(Geo g, Ops<T> o, const double* __restrict__ b,
                                                    double* __restrict__ x, int colour) {
    const int t = blockIdx.x * kColsX + threadIdx.x;
    const int j = blockIdx.y * kColsY + threadIdx.y;
    if (j >= g.ny || t >= g.hx) return;
    const int i = 2 * t + ((j + colour) & 1);
    if (i >= g.nx) return;
    const Col c = make_col(g, i, j);
    double cp[NZ], dp[NZ];
    double pcp = 0.0, pdp = 0.0;
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        if (l < g.nz) {
            const long base = (long)l * g.plane, id = base + c.c0;
            const double Gw = (double)o.gE[base + c.pw], Ge = (double)o.gE[id];
            const double Gs = (double)o.gN[base + c.ps], Gn = (double)o.gN[id];
            const double Gu = (double)o.gU[id], Gd = (l > 0) ? (double)o.gU[id - g.plane] : (double)o.gB[c.c0];
            double d = 0.0, rhs = b[id];
            if (c.hw) d += Gw, rhs += Gw * x[base + c.pw];
            if (c.he) d += Ge, rhs += Ge * x[base + c.pe];
            if (c.hs) d += Gs, rhs += Gs * x[base + c.ps];
            if (c.hn) d += Gn, rhs += Gn * x[base + c.pn];
            d += Gu + Gd;
            const double lower = (l > 0) ? -Gd : 0.0, upper = (l < g.nz - 1) ? -Gu : 0.0;
            const double den = d - lower * pcp;
            cp[l] = upper / den;
            dp[l] = (rhs - lower * pdp) / den;
            pcp = cp[l];
            pdp = dp[l];
        }
    }
    double xu = 0.0;
#pragma unroll
    for (int l = NZ - 1; l >= 0; --l) {
        if (l < g.nz) {
            const double v = (l == g.nz - 1) ? dp[l] : dp[l] - cp[l] * xu;
            x[(long)l * g.plane + c.c0] = v;
            xu = v;
        }
    }
}

// k_line_relax on a level with column types: the faces from the column's
// type (UNI: wave-uniform, through the scalar cache), absent lateral faces
// masked to +0 (d and rhs start at +0 or b and never become -0, so the sums
// are k_line_relax's), every plane's loads unconditional; the same Thomas
// arithmetic.
template <class T, int NZ, bool UNI>
__device__ __forceinline__ void line_relax_typed_column(const Geo& g, const Typed<T>& y, const double* __restrict__ b,
                                                        double* __restrict__ x, const Col& c, int t) {
    const int nz = g.nz;
    double cp[NZ], dp[NZ];
    double pcp = 0.0, pdp = 0.0;
    double gdn = (double)(UNI ? y.urow(t, 0) : y.row(t, 0)).v[TF_B];
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        const int lc = l < nz ? l : nz - 1;
        const long base = (long)lc * g.plane, id = base + c.c0;
        const TypeRow<T> rw = UNI ? y.urow(t, lc) : y.row(t, lc);
        const double Gw = c.hw ? (double)rw.v[TF_W] : 0.0, Ge = c.he ? (double)rw.v[TF_E] : 0.0;
        const double Gs = c.hs ? (double)rw.v[TF_S] : 0.0, Gn = c.hn ? (double)rw.v[TF_N] : 0.0;
        const double Gu = (double)rw.v[TF_U], Gd = gdn;
        const double xw = x[base + c.pw], xe = x[base + c.pe], xs = x[base + c.ps], xn = x[base + c.pn];
        double d = 0.0, rhs = b[id];
        d += Gw, rhs += Gw * xw;
        d += Ge, rhs += Ge * xe;
        d += Gs, rhs += Gs * xs;
        d += Gn, rhs += Gn * xn;
        d += Gu + Gd;
        const double lower = (l > 0) ? -Gd : 0.0, upper = (l < nz - 1) ? -Gu : 0.0;
        const double den = d - lower * pcp;
        cp[l] = upper / den;
        dp[l] = (rhs - lower * pdp) / den;
        pcp = cp[l];
        pdp = dp[l];
        gdn = Gu;
    }
    double xu = 0.0;
#pragma unroll
    for (int l = NZ - 1; l >= 0; --l) {
        if (l < nz) {
            const double v = (l == nz - 1) ? dp[l] : dp[l] - cp[l] * xu;
            x[(long)l * g.plane + c.c0] = v;
            xu = v;
        }
    }
}
template <class T, int NZ>
__global__ __launch_bounds__(256) void k_line_relax_typed(Geo g, Typed<T> y, const double* __restrict__ b,
                                                          double* __restrict__ x, int colour) {
    const int t = blockIdx.x * kColsX + threadIdx.x;
    const int j = blockIdx.y * kColsY + threadIdx.y;
    if (j >= g.ny || t >= g.hx) return;
    const int i = 2 * t + ((j + colour) & 1);
    if (i >= g.nx) return;
    const Col c = make_col(g, i, j);
    const int ty = y.ty[c.c0];
    const int rep = wave_rep_type(ty);
    if (__all(ty == rep))
        line_relax_typed_column<T, NZ, true>(g, y, b, x, c, rep);
    else
        line_relax_typed_column<T, NZ, false>(g, y, b, x, c, ty);
}

// Natural <-> split layout. Padding slots of `out` are left untouched
// (they were cleared at allocation).
template <class S, class D>
__global__ void k_to_split(Geo g, int planes, const S* __restrict__ in, D* __restrict__ out) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)g.nx * g.ny * planes;
    if (k >= n) return;
    const int i = (int)(k % g.nx);
    const long rest = k / g.nx;
    const int j = (int)(rest % g.ny), l = (int)(rest / g.ny);
    out[(long)l * g.plane + g.pos(i, j)] = (D)in[k];
}
static __global__ void k_to_natural(Geo g, const double* __restrict__ in, double* __restrict__ out) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)g.nx * g.ny * g.nz;
    if (k >= n) return;
    const int i = (int)(k % g.nx);
    const long rest = k / g.nx;
    const int j = (int)(rest % g.ny), l = (int)(rest / g.ny);
    out[k] = in[(long)l * g.plane + g.pos(i, j)];
}
template <class D, class S>
__global__ void k_cast(long n, const S* __restrict__ in, D* __restrict__ out) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) out[k] = (D)in[k];
}
static __global__ void k_mul(long n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) out[k] = a[k] * b[k];
}

}  // namespace th
