// solver.hip - MG-preconditioned CG for layered stacks on MI355X (gfx950),
// and the C ABI of the solver (include/thermal_hip.h).
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
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.h"
#include "levels.h"
#include "model.h"

namespace th {

// ===========================================================================
// Kernels
// ===========================================================================

enum { SW_PLAIN = 0, SW_ZERO = 1, SW_PROLONG = 2 };

// Red-black column line smoother (mg3d.h:67-84): solves every column of one
// colour exactly with its lateral neighbours held fixed, using the
// precomputed 1/pivot. SW_ZERO: neighbours taken as zero (first sweep of a
// V-cycle on a cleared level). SW_PROLONG: each neighbour reads
// x + x_coarse(parent) (prolongation folded in).
template <class T, int NZ, int MODE>
__global__ __launch_bounds__(256) void k_sweep(Geo g, Ops<T> o, const T* __restrict__ finv,
                                                const T* __restrict__ b, T* __restrict__ x, int colour, Geo gc,
                                                const T* __restrict__ xc, const State* __restrict__ st) {
    if (st->done) return;
    const int t = blockIdx.x * kColsX + threadIdx.x;
    const int j = blockIdx.y * kColsY + threadIdx.y;
    if (j >= g.ny) return;
    const int i = 2 * t + ((j + colour) & 1);
    if (i >= g.nx) return;
    const Col c = make_col(g, i, j);
    int qw = 0, qe = 0, qs = 0, qn = 0;
    if (MODE == SW_PROLONG) {  // parents of the 4 neighbours on the coarse level
        qw = c.hw ? gc.pos((i - 1) >> 1, j >> 1) : 0;
        qe = c.he ? gc.pos((i + 1) >> 1, j >> 1) : 0;
        qs = c.hs ? gc.pos(i >> 1, (j - 1) >> 1) : 0;
        qn = c.hn ? gc.pos(i >> 1, (j + 1) >> 1) : 0;
    }
    T xs[NZ], fcp[NZ];
    T prev = 0, gd = 0;
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        if (l < g.nz) {
            const long base = (long)l * g.plane;
            const long id = base + c.c0;
            T s = b[id];
            if (MODE != SW_ZERO) {
                const long cb = (long)l * gc.plane;
                T ls = 0;
                if (c.hw) {
                    T xv = x[base + c.pw];
                    if (MODE == SW_PROLONG) xv = xv + xc[cb + qw];
                    ls += o.gE[base + c.pw] * xv;
                }
                if (c.he) {
                    T xv = x[base + c.pe];
                    if (MODE == SW_PROLONG) xv = xv + xc[cb + qe];
                    ls += o.gE[id] * xv;
                }
                if (c.hs) {
                    T xv = x[base + c.ps];
                    if (MODE == SW_PROLONG) xv = xv + xc[cb + qs];
                    ls += o.gN[base + c.ps] * xv;
                }
                if (c.hn) {
                    T xv = x[base + c.pn];
                    if (MODE == SW_PROLONG) xv = xv + xc[cb + qn];
                    ls += o.gN[id] * xv;
                }
                s = s + ls;
            }
            const T lower = (l > 0) ? -gd : T(0);
            const T inv = finv[id];
            prev = (s - lower * prev) * inv;
            xs[l] = prev;
            const T gu = o.gU[id];
            fcp[l] = (l < g.nz - 1) ? (-gu) * inv : T(0);
            gd = gu;
        }
    }
#pragma unroll
    for (int l = NZ - 2; l >= 0; --l)
        if (l < g.nz - 1) xs[l] = xs[l] - fcp[l] * xs[l + 1];
#pragma unroll
    for (int l = 0; l < NZ; ++l)
        if (l < g.nz) x[(long)l * g.plane + c.c0] = xs[l];
}

// Column factors (mg3d.h:46-61) for every column of a level: 1/pivot.
template <class T>
__global__ __launch_bounds__(256) void k_factor(Geo g, Ops<T> o, T* __restrict__ finv, const State* __restrict__ st) {
    const int q = blockIdx.x * kColsX + threadIdx.x;  // slot in the row (both colours)
    const int j = blockIdx.y * kColsY + threadIdx.y;
    if (j >= g.ny || q >= 2 * g.hx) return;
    const int colour = q >= g.hx ? 1 : 0, t = q - colour * g.hx;
    const int i = 2 * t + ((j + colour) & 1);
    if (i >= g.nx) return;
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

// Restriction (pcg3d_cpu.cpp:87-100): coarse b = sum over the (up to 4)
// children of the fine residual b - A x, children in row-major order. One
// thread per coarse column; fixed summation order.
template <class T>
__global__ __launch_bounds__(256) void k_restrict(Geo f, Ops<T> o, const T* __restrict__ fb,
                                                   const T* __restrict__ fx, Geo cg, T* __restrict__ cb,
                                                   const State* __restrict__ st) {
    if (st->done) return;
    const int I = blockIdx.x * kColsX + threadIdx.x;
    const int J = blockIdx.y * kColsY + threadIdx.y;
    if (I >= cg.nx || J >= cg.ny) return;
    const T inv_dt = (T)st->inv_dt;
    // children in row-major order: (2I, 2J), (2I+1, 2J), (2I, 2J+1), (2I+1, 2J+1)
    Col ch[4];
    bool ok[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = 2 * I + (k & 1), j = 2 * J + (k >> 1);
        ok[k] = i < f.nx && j < f.ny;
        ch[k] = make_col(f, ok[k] ? i : 2 * I, ok[k] ? j : 2 * J);
    }
    const int cpos = cg.pos(I, J);
    for (int l = 0; l < f.nz; ++l) {
        const long base = (long)l * f.plane;
        T sum = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (ok[k]) sum += fb[base + ch[k].c0] - apply_cell<T>(f, o, fx, ch[k], l, inv_dt);
        cb[(long)l * cg.plane + cpos] = sum;
    }
}

// ---- fine-grid PCG kernels (FP64) ------------------------------------------

enum { RES_MG = 0, RES_JACOBI = 1, RES_CHECK = 2 };

// Start of a solve: r = b - (A + C/dt) x, with b0 = (T) r for the V-cycle
// (MG) or z = p = r / diag (Jacobi); sums r.r and b.b (and r.z).
// RES_CHECK: the true residual ||P - A x|| of the steady operator, no writes.
template <class T, int MODE>
__global__ __launch_bounds__(256) void k_residual(Geo g, Ops<T> o, const double* __restrict__ b,
                                                   const double* __restrict__ x, double* __restrict__ r,
                                                   T* __restrict__ b0, const double* __restrict__ invd,
                                                   double* __restrict__ z, double* __restrict__ p, State* st,
                                                   double* part) {
    if (MODE != RES_CHECK && st->done) return;
    const int q = blockIdx.x * kColsX + threadIdx.x;
    const int j = blockIdx.y * kColsY + threadIdx.y;
    double v[3] = {0.0, 0.0, 0.0};  // rr, bb, rz
    const double inv_dt = MODE == RES_CHECK ? 0.0 : st->inv_dt;
    if (j < g.ny && q < 2 * g.hx) {
        const int colour = q >= g.hx ? 1 : 0, t = q - colour * g.hx;
        const int i = 2 * t + ((j + colour) & 1);
        if (i < g.nx) {
            const Col c = make_col(g, i, j);
            for (int l = 0; l < g.nz; ++l) {
                const long id = (long)l * g.plane + c.c0;
                const double bi = b[id];
                const double ri = bi - apply_cell<double>(g, o, x, c, l, inv_dt);
                v[0] += ri * ri;
                v[1] += bi * bi;
                if (MODE == RES_MG) {
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
    }
    double tot[3];
    if (grid_reduce<3>(v, part, &st->ticket[MODE == RES_CHECK ? TK_CHECK : TK_RES], tot) &&
        threadIdx.x == 0 && threadIdx.y == 0) {
        if (MODE == RES_CHECK) {
            st->chk_rr = tot[0];
            st->chk_bb = tot[1];
            return;
        }
        st->bb = tot[1];
        st->rr = tot[0];
        st->iters = 0;
        st->rzo = 0.0;
        if (tot[1] == 0.0) {  // b = 0: x = 0 is the answer (pcg3d_cpu.cpp:146)
            st->relres = 0.0;
            st->done = 1;
            st->converged = 1;
            st->zero_x = 1;
            return;
        }
        st->relres = sqrt(tot[0] / tot[1]);
        if (st->relres < st->tol) {
            st->done = 1;
            st->converged = 1;
        }
        if (MODE == RES_JACOBI) st->rz = tot[2];
    }
}

// Ap = (A + C/dt) p and p.Ap; epilogue: alpha = rz / pAp.
template <class T, int NZ>
__global__ __launch_bounds__(256) void k_apply(Geo g, Ops<T> o, const double* __restrict__ p,
                                                double* __restrict__ Ap, State* st, double* part) {
    if (st->done) return;
    const int q = blockIdx.x * kColsX + threadIdx.x;
    const int j = blockIdx.y * kColsY + threadIdx.y;
    double v[1] = {0.0};
    const double inv_dt = st->inv_dt;
    if (j < g.ny && q < 2 * g.hx) {
        const int colour = q >= g.hx ? 1 : 0, t = q - colour * g.hx;
        const int i = 2 * t + ((j + colour) & 1);
        if (i < g.nx) {
            const Col c = make_col(g, i, j);
            double pb = 0.0, pc = p[c.c0], gdn = 0.0;
#pragma unroll
            for (int l = 0; l < NZ; ++l) {
                if (l < g.nz) {
                    const long base = (long)l * g.plane, id = base + c.c0;
                    const double pu = (l < g.nz - 1) ? p[id + g.plane] : 0.0;
                    const double gu = (double)o.gU[id];
                    double d = 0.0, s = 0.0;
                    double gw = 0, ge = 0, gs = 0, gn = 0;
                    if (c.hw) gw = (double)o.gE[base + c.pw];
                    if (c.he) ge = (double)o.gE[id];
                    if (c.hs) gs = (double)o.gN[base + c.ps];
                    if (c.hn) gn = (double)o.gN[id];
                    if (c.hw) d += gw;
                    if (c.he) d += ge;
                    if (c.hs) d += gs;
                    if (c.hn) d += gn;
                    d += gu;
                    d += (l > 0) ? gdn : (double)o.gB[c.c0];
                    if (inv_dt != 0.0) d = d + (double)o.cap[id] * inv_dt;
                    if (c.hw) s += gw * p[base + c.pw];
                    if (c.he) s += ge * p[base + c.pe];
                    if (c.hs) s += gs * p[base + c.ps];
                    if (c.hn) s += gn * p[base + c.pn];
                    if (l > 0) s += gdn * pb;
                    if (l < g.nz - 1) s += gu * pu;
                    const double a = d * pc - s;
                    Ap[id] = a;
                    v[0] += pc * a;
                    pb = pc;
                    pc = pu;
                    gdn = gu;
                }
            }
        }
    }
    double tot[1];
    if (grid_reduce<1>(v, part, &st->ticket[TK_APPLY], tot) && threadIdx.x == 0 && threadIdx.y == 0) {
        st->pAp = tot[0];
        st->alpha = st->rz / tot[0];
    }
}

enum { UP_MG = 0, UP_MG_FLEX = 1, UP_JACOBI = 2 };

// x += alpha p, r -= alpha Ap, and b0 = (T) r for the V-cycle (MG) or
// z = r / diag (Jacobi); sums r.r (and r.z_old for the flexible beta, where
// z_old is the previous V-cycle output still held in x0, or r.z for Jacobi).
// Epilogue: convergence test on ||r|| / ||b||, iteration count, Jacobi beta.
template <class T, int MODE>
__global__ __launch_bounds__(256) void k_update(long n, const double* __restrict__ p, const double* __restrict__ Ap,
                                                 double* __restrict__ x, double* __restrict__ r, T* __restrict__ b0,
                                                 const T* __restrict__ zold, const double* __restrict__ invd,
                                                 double* __restrict__ z, State* st, double* part) {
    if (st->done) return;
    const double alpha = st->alpha;
    double v[2] = {0.0, 0.0};
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        x[k] += alpha * p[k];
        const double ri = r[k] - alpha * Ap[k];
        r[k] = ri;
        v[0] += ri * ri;
        if (MODE == UP_JACOBI) {
            const double zi = ri * invd[k];
            z[k] = zi;
            v[1] += ri * zi;
        } else {
            if (MODE == UP_MG_FLEX) v[1] += ri * (double)zold[k];
            b0[k] = (T)ri;
        }
    }
    double tot[2];
    if (grid_reduce<2>(v, part, &st->ticket[TK_UPDATE], tot) && threadIdx.x == 0) {
        st->rr = tot[0];
        st->iters += 1;
        st->relres = sqrt(tot[0] / st->bb);
        if (st->relres < st->tol) {
            st->done = 1;
            st->converged = 1;
        } else if (st->iters >= st->maxit) {
            st->done = 1;
        }
        if (MODE == UP_MG_FLEX) st->rzo = tot[1];
        if (MODE == UP_JACOBI && !st->done) {
            st->beta = tot[1] / st->rz;
            st->rz = tot[1];
        }
    }
}

// r.z with z = the V-cycle output; epilogue: beta and rz (first: rz only).
// Flexible beta: (r.z - r.z_old) / rz_old.
template <class T, int FIRST>
__global__ __launch_bounds__(256) void k_dot(long n, const double* __restrict__ r, const T* __restrict__ z,
                                              State* st, double* part) {
    if (st->done) return;
    double v[1] = {0.0};
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) v[0] += r[k] * (double)z[k];
    double tot[1];
    if (grid_reduce<1>(v, part, &st->ticket[TK_DOT], tot) && threadIdx.x == 0) {
        if (!FIRST) st->beta = st->flexible ? (tot[0] - st->rzo) / st->rz : tot[0] / st->rz;
        st->rz = tot[0];
    }
}

// p = z + beta p (first: p = z).
template <class Z, int FIRST>
__global__ __launch_bounds__(256) void k_pupdate(long n, const Z* __restrict__ z, double* __restrict__ p,
                                                  const State* __restrict__ st) {
    if (st->done) return;
    const double beta = st->beta;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
        p[k] = FIRST ? (double)z[k] : (double)z[k] + beta * p[k];
}

// Element-wise helpers for the test entry points.
template <class D, class S>
__global__ void k_cast(long n, const S* __restrict__ in, D* __restrict__ out) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) out[k] = (D)in[k];
}
__global__ void k_mul(long n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) out[k] = a[k] * b[k];
}

// Start of a solve: tolerance, iteration cap, cleared flags.
__global__ void k_begin(State* st, double tol, int maxit) {
    st->tol = tol;
    st->maxit = maxit;
    st->iters = 0;
    st->done = 0;
    st->converged = 0;
    st->zero_x = 0;
    st->relres = 0.0;
}
__global__ void k_set_inv_dt(State* st, double inv_dt) { st->inv_dt = inv_dt; }

// x = 0 when b = 0 was detected (the reference's set_x_zero on bb == 0).
__global__ void k_finish(long n, double* __restrict__ x, const State* __restrict__ st) {
    if (!st->zero_x) return;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) x[k] = 0.0;
}

// Jacobi: 1 / diag of (A + C/dt) (pcg3d_cpu.cpp:41-48).
template <class T>
__global__ __launch_bounds__(256) void k_invdiag(Geo g, Ops<T> o, double* __restrict__ invd, const State* st) {
    const int q = blockIdx.x * kColsX + threadIdx.x;
    const int j = blockIdx.y * kColsY + threadIdx.y;
    if (j >= g.ny || q >= 2 * g.hx) return;
    const int colour = q >= g.hx ? 1 : 0, t = q - colour * g.hx;
    const int i = 2 * t + ((j + colour) & 1);
    if (i >= g.nx) return;
    const Col c = make_col(g, i, j);
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
__global__ __launch_bounds__(256) void k_rhs_transient(long n, const T* __restrict__ cap, const double* __restrict__ P,
                                                        const double* __restrict__ x, double* __restrict__ b,
                                                        const State* __restrict__ st) {
    const double inv_dt = st->inv_dt;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
        b[k] = P[k] + (double)cap[k] * inv_dt * x[k];
}

// Energy terms of the current x: sum C x (heat content), heat out through
// the top and bottom (stack3d.h:254-263), and max x.
template <class T>
__global__ __launch_bounds__(256) void k_energy(Geo g, Ops<T> o, const double* __restrict__ x, State* st,
                                                 double* part) {
    double v[3] = {0.0, 0.0, -1.0e300};
    const long n = g.cells(), top = (long)(g.nz - 1) * g.plane;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const double xv = x[k];
        // padding slots (odd nx) hold zeros and are not cells: skip them for the max
        const long q = k % g.plane;
        const int row = (int)(q / (2 * g.hx)), slot = (int)(q % (2 * g.hx));
        const int colour = slot >= g.hx ? 1 : 0;
        const bool cell = 2 * (slot - colour * g.hx) + ((row + colour) & 1) < g.nx;
        v[0] += (double)o.cap[k] * xv;
        if (k >= top) v[1] += (double)o.gU[k] * xv;
        if (k < g.plane) v[1] += (double)o.gB[k] * xv;
        if (cell) v[2] = fmax(v[2], xv);
    }
    double tot[3];
    if (grid_reduce<3, 4u>(v, part, &st->ticket[TK_ENERGY], tot) && threadIdx.x == 0) {
        st->e_heat = tot[0];
        st->e_out = tot[1];
        st->e_tmax = tot[2];
    }
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
__global__ void k_to_natural(Geo g, const double* __restrict__ in, double* __restrict__ out) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)g.nx * g.ny * g.nz;
    if (k >= n) return;
    const int i = (int)(k % g.nx);
    const long rest = k / g.nx;
    const int j = (int)(rest % g.ny), l = (int)(rest / g.ny);
    out[k] = in[(long)l * g.plane + g.pos(i, j)];
}

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

template <class T>
struct DevLevel {
    Geo g;
    T *gE = nullptr, *gN = nullptr, *gU = nullptr, *gB = nullptr, *cap = nullptr;
    T *finv = nullptr, *x = nullptr, *b = nullptr;
    Ops<T> ops() const { return {gE, gN, gU, gB, cap}; }
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
    double* part = nullptr;
    State* st = nullptr;
    State* h_st = nullptr;  // pinned mirror
    std::vector<DevLevel<T>> L;
    hipGraphExec_t g_start = nullptr, g_iter = nullptr;
    std::vector<void*> owned;

    ~Solver() override {
        if (s) (void)hipStreamSynchronize(s);
        if (g_start) (void)hipGraphExecDestroy(g_start);
        if (g_iter) (void)hipGraphExecDestroy(g_iter);
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

    // natural-layout host array -> split-layout device array of type D
    template <class S, class D>
    int upload_split(const Geo& g, int planes, const S* host, D* dev) {
        const long n = (long)g.nx * g.ny * planes;
        void* tmp = nullptr;
        HIPTRY(hipMalloc(&tmp, n * sizeof(S)));
        hipError_t e = hipMemcpyAsync(tmp, host, n * sizeof(S), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            k_to_split<S, D><<<lin_blocks(n), 256, 0, s>>>(g, planes, (const S*)tmp, dev);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(tmp);
        if (e != hipSuccess) return fail(TH_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
        return TH_OK;
    }

    int init(const th_model_desc& m, const th_solver_opts& o) override {
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
        if ((rc = dalloc(&st, 1))) return rc;
        HIPTRY(hipHostMalloc((void**)&h_st, sizeof(State), hipHostMallocDefault));
        std::memset(h_st, 0, sizeof(State));

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
        L.resize(H.size());
        long maxblk = kMaxVecBlocks;
        for (size_t k = 0; k < H.size(); ++k) {
            DevLevel<T>& d = L[k];
            d.g = make_geo(H[k].nx, H[k].ny, H[k].nz);
            const long nc = d.g.cells();
            for (T** v : {&d.gE, &d.gN, &d.gU, &d.cap, &d.finv, &d.x, &d.b})
                if ((rc = dalloc(v, nc))) return rc;
            if ((rc = dalloc(&d.gB, d.g.plane))) return rc;
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
            const dim3 cg = col_grid(2 * d.g.hx, d.g.ny);
            maxblk = std::max(maxblk, (long)cg.x * cg.y);
            if (!mg()) break;
        }
        if ((rc = dalloc(&part, maxblk * 4))) return rc;
        if ((rc = upload_split<double, double>(g0, g0.nz, m.P, P))) return rc;
        if ((rc = set_dt(0.0))) return rc;
        HIPTRY(hipStreamSynchronize(s));
        if (opt.use_graphs && mg()) {
            if ((rc = capture(&g_start, true))) return rc;
            if ((rc = capture(&g_iter, false))) return rc;
        }
        return TH_OK;
    }

    int capture(hipGraphExec_t* exec, bool start) {
        hipGraph_t gr = nullptr;
        HIPTRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = start ? enqueue_start() : enqueue_iter();
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
        if (mg()) {
            for (DevLevel<T>& d : L) {
                k_factor<T><<<col_grid(2 * d.g.hx, d.g.ny), col_block(), 0, s>>>(d.g, d.ops(), d.finv, st);
                HIPTRY(hipGetLastError());
            }
        } else {
            const DevLevel<T>& d = L[0];
            k_invdiag<T><<<col_grid(2 * g0.hx, g0.ny), col_block(), 0, s>>>(g0, d.ops(), invd, st);
            HIPTRY(hipGetLastError());
        }
        return TH_OK;
    }
    int set_x_zero() override {
        HIPTRY(hipMemsetAsync(x, 0, n0 * sizeof(double), s));
        return TH_OK;
    }
    int set_x(const double* host) override { return upload_split<double, double>(g0, g0.nz, host, x); }
    int set_rhs_steady() override {
        HIPTRY(hipMemcpyAsync(b, P, n0 * sizeof(double), hipMemcpyDeviceToDevice, s));
        return TH_OK;
    }
    int set_rhs_transient() override {
        k_rhs_transient<T><<<vec_blocks(n0), kVecThreads, 0, s>>>(n0, L[0].cap, P, x, b, st);
        HIPTRY(hipGetLastError());
        return TH_OK;
    }
    int set_power(const double* host) override {
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

    template <int MODE>
    int launch_sweep(size_t k, int colour) {
        const DevLevel<T>& d = L[k];
        const DevLevel<T>& c = L[std::min(k + 1, L.size() - 1)];
        const dim3 grid = col_grid(d.g.hx, d.g.ny);
#define TH_SWEEP(NZ) k_sweep<T, NZ, MODE><<<grid, col_block(), 0, s>>>(d.g, d.ops(), d.finv, d.b, d.x, colour, c.g, c.x, st)
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

    // One V-cycle on level k (pcg3d_cpu.cpp:69-113): pre-smooth red, black;
    // restrict; coarse V-cycle; post-smooth black (with the prolongation),
    // red. Level k's x starts from zero.
    int vcycle(size_t k) {
        int rc;
        if (k + 1 == L.size()) return launch_sweep<SW_ZERO>(k, 0);  // one column: exact
        if ((rc = launch_sweep<SW_ZERO>(k, 0))) return rc;
        if ((rc = launch_sweep<SW_PLAIN>(k, 1))) return rc;
        const DevLevel<T>& f = L[k];
        const DevLevel<T>& c = L[k + 1];
        k_restrict<T><<<col_grid(c.g.nx, c.g.ny), col_block(), 0, s>>>(f.g, f.ops(), f.b, f.x, c.g, c.b, st);
        HIPTRY(hipGetLastError());
        if ((rc = vcycle(k + 1))) return rc;
        if ((rc = launch_sweep<SW_PROLONG>(k, 1))) return rc;
        return launch_sweep<SW_PLAIN>(k, 0);
    }

    int launch_apply() {
        const dim3 grid = col_grid(2 * g0.hx, g0.ny);
        const Ops<T> o = L[0].ops();
#define TH_APPLY(NZ) k_apply<T, NZ><<<grid, col_block(), 0, s>>>(g0, o, p, Ap, st, part)
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

    // Start: residual, first preconditioner application, rz, p = z.
    int enqueue_start() {
        const dim3 grid = col_grid(2 * g0.hx, g0.ny);
        const unsigned vb = vec_blocks(n0);
        int rc;
        if (mg()) {
            k_residual<T, RES_MG><<<grid, col_block(), 0, s>>>(g0, L[0].ops(), b, x, r, L[0].b, nullptr, nullptr,
                                                               nullptr, st, part);
            HIPTRY(hipGetLastError());
            if ((rc = vcycle(0))) return rc;
            k_dot<T, 1><<<vb, kVecThreads, 0, s>>>(n0, r, L[0].x, st, part);
            HIPTRY(hipGetLastError());
            k_pupdate<T, 1><<<vb, kVecThreads, 0, s>>>(n0, L[0].x, p, st);
        } else {
            k_residual<T, RES_JACOBI><<<grid, col_block(), 0, s>>>(g0, L[0].ops(), b, x, r, nullptr, invd, z, p, st,
                                                                   part);
        }
        HIPTRY(hipGetLastError());
        return TH_OK;
    }

    // One PCG iteration: Ap and alpha; x, r update and the convergence test;
    // z = M^-1 r; beta; p = z + beta p.
    int enqueue_iter() {
        const unsigned vb = vec_blocks(n0);
        int rc;
        if ((rc = launch_apply())) return rc;
        if (mg()) {
            if (opt.beta == TH_BETA_FLEXIBLE)
                k_update<T, UP_MG_FLEX><<<vb, kVecThreads, 0, s>>>(n0, p, Ap, x, r, L[0].b, L[0].x, nullptr, nullptr,
                                                                   st, part);
            else
                k_update<T, UP_MG><<<vb, kVecThreads, 0, s>>>(n0, p, Ap, x, r, L[0].b, nullptr, nullptr, nullptr, st,
                                                              part);
            HIPTRY(hipGetLastError());
            if ((rc = vcycle(0))) return rc;
            k_dot<T, 0><<<vb, kVecThreads, 0, s>>>(n0, r, L[0].x, st, part);
            HIPTRY(hipGetLastError());
            k_pupdate<T, 0><<<vb, kVecThreads, 0, s>>>(n0, L[0].x, p, st);
        } else {
            k_update<T, UP_JACOBI><<<vb, kVecThreads, 0, s>>>(n0, p, Ap, x, r, nullptr, nullptr, invd, z, st, part);
            HIPTRY(hipGetLastError());
            k_pupdate<double, 0><<<vb, kVecThreads, 0, s>>>(n0, z, p, st);
        }
        HIPTRY(hipGetLastError());
        return TH_OK;
    }

    int fetch_state() {
        HIPTRY(hipMemcpyAsync(h_st, st, sizeof(State), hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        return TH_OK;
    }

    int solve(double tol, int maxit, th_stats* out) override {
        int rc;
        // Flexible beta is a property of the solve; it rides in the state.
        k_begin<<<1, 1, 0, s>>>(st, tol, maxit);
        HIPTRY(hipGetLastError());
        if (g_start) {
            HIPTRY(hipGraphLaunch(g_start, s));
        } else if ((rc = enqueue_start())) {
            return rc;
        }
        if ((rc = fetch_state())) return rc;
        // The host checks convergence after every iteration, like the
        // reference; kernels of a finished solve return at once.
        for (int it = 0; !h_st->done && it < maxit; ++it) {
            if (g_iter) {
                HIPTRY(hipGraphLaunch(g_iter, s));
            } else if ((rc = enqueue_iter())) {
                return rc;
            }
            if ((rc = fetch_state())) return rc;
        }
        k_finish<<<vec_blocks(n0), kVecThreads, 0, s>>>(n0, x, st);
        HIPTRY(hipGetLastError());
        if (out) {
            out->iters = h_st->iters;
            out->relres = h_st->relres;
            out->converged = h_st->converged;
        }
        return TH_OK;
    }

    int true_relres(double* out) override {
        const dim3 grid = col_grid(2 * g0.hx, g0.ny);
        k_residual<T, RES_CHECK><<<grid, col_block(), 0, s>>>(g0, L[0].ops(), P, x, nullptr, nullptr, nullptr,
                                                              nullptr, nullptr, st, part);
        HIPTRY(hipGetLastError());
        int rc;
        if ((rc = fetch_state())) return rc;
        *out = std::sqrt(h_st->chk_rr / h_st->chk_bb);
        return TH_OK;
    }

    int energy(double* heat, double* outq, double* tmax) override {
        k_energy<T><<<vec_blocks(n0), kVecThreads, 0, s>>>(g0, L[0].ops(), x, st, part);
        HIPTRY(hipGetLastError());
        int rc;
        if ((rc = fetch_state())) return rc;
        if (heat) *heat = h_st->e_heat;
        if (outq) *outq = h_st->e_out;
        if (tmax) *tmax = h_st->e_tmax;
        return TH_OK;
    }

    // Test entry points: one preconditioner application / one operator apply
    // on host vectors. They use r and p as staging and leave x, b untouched.
    int precondition(const double* host_r, double* host_z) override {
        int rc;
        if ((rc = upload_split<double, double>(g0, g0.nz, host_r, r))) return rc;
        k_begin<<<1, 1, 0, s>>>(st, 0.0, 1);
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
        k_begin<<<1, 1, 0, s>>>(st, 0.0, 1);
        HIPTRY(hipGetLastError());
        if ((rc = launch_apply())) return rc;
        return download_natural(Ap, host_y);
    }
    int download_natural(const double* dev, double* host) {
        k_to_natural<<<lin_blocks(nnat), 256, 0, s>>>(g0, dev, xnat);
        HIPTRY(hipGetLastError());
        HIPTRY(hipMemcpyAsync(host, xnat, nnat * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        return TH_OK;
    }
};

// The flexible flag lives in the device state; set once at creation.
__global__ void k_set_flexible(State* st, int f) { st->flexible = f; }

}  // namespace th

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
}

// transient3d.h:48-90, with the energy terms on the device: heat content
// sum C T before and after the step gives the stored heat, heat out through
// the boundaries, and T_max.
int th_run_transient(th_solver* s, const th_transient_opts* o, th_step_record* recs, int rec_cap, int* nsteps,
                     double* T_final, int* all_converged, double* solve_ms) {
    TH_S(s);
    if (!o || !nsteps) return fail(TH_ERR_ARG, "null argument");
    using clk = std::chrono::high_resolution_clock;
    SolverBase* S = s->impl;
    int rc;
    if ((rc = S->set_x_zero())) return rc;
    double heat_old = 0, q, tmax, heat_new;
    if ((rc = S->energy(&heat_old, &q, &tmax))) return rc;
    double t = 0, dt = o->dt0, last_dt = -1, ms = 0;
    int steps = 0;
    int conv_all = 1;
    while (t < o->t_end * (1 - 1e-12) && (o->max_steps <= 0 || steps < o->max_steps)) {
        dt = std::min(dt, o->t_end - t);
        const auto ts = clk::now();
        if (dt != last_dt) {
            if ((rc = S->set_dt(dt))) return rc;
            last_dt = dt;
        }
        if ((rc = S->set_rhs_transient())) return rc;
        th_stats st{};
        if ((rc = S->solve(o->tol, o->maxit, &st))) return rc;
        (void)hipStreamSynchronize(S->stream());
        ms += std::chrono::duration<double, std::milli>(clk::now() - ts).count();
        if ((rc = S->energy(&heat_new, &q, &tmax))) return rc;
        t += dt;
        if (recs) {
            if (steps >= rec_cap) return fail(TH_ERR_ARG, "record buffer too small");
            th_step_record& R = recs[steps];
            R.t = t;
            R.dt = dt;
            R.iters = st.iters;
            R.T_max = tmax;
            R.q_out = q;
            R.q_stored = (heat_new - heat_old) / dt;
        }
        heat_old = heat_new;
        if (!st.converged) conv_all = 0;
        dt = std::min(dt * o->growth, o->dt_max);
        ++steps;
    }
    *nsteps = steps;
    if (all_converged) *all_converged = conv_all;
    if (solve_ms) *solve_ms = ms;
    if (T_final) return S->get_x(T_final);
    return TH_OK;
}

}  // extern "C"
