// kernels.h - device-side building blocks shared by the multigrid and PCG
// kernels of the gfx950 MG-PCG solver.
//
// Device layout (every vector and every multigrid level): colour-split,
// plane-major. Within z-plane l, row j stores its red columns ((i+j) even)
// then its black ones, hx = ceil(nx/2) slots each:
//     cell (i, j, l)  ->  l * plane + (2 j + ((i + j) & 1)) * hx + i / 2,
//     plane = 2 hx ny.
// A colour sweep then reads and writes contiguous runs of one colour, and the
// 4 lateral neighbours of a cell all sit in the other colour's half-rows
// (the colour-split layout of src/operator3d.h:15-23). Padding slots (odd nx)
// hold zeros in every array and are never written by the sweeps.
//
// Reductions are deterministic and fence-free: a wave64 butterfly, the waves
// of a block through LDS, one partial per block in a fixed slot; a separate
// one-block kernel then adds the partials in block order and runs the scalar
// epilogue (alpha, beta, ||r||, convergence). On gfx950 a device-scope fence
// in every block (the "last block" pattern) writes back the XCD's L2 each
// time and cost the fused kernels 3-5x of their bandwidth; the kernel
// boundary gives the same ordering once per launch.
#pragma once
#include <hip/hip_runtime.h>

#include "coltypes.h"

#include <cstdint>

namespace th {

constexpr int kColsX = 64;  // threads per block along a half-row (one wave)
constexpr int kColsY = 4;   // rows per block
constexpr int kVecThreads = 1024;
// Grid-stride vector kernels: at most 4 blocks of 1024 threads per CU.
// 1024 blocks against 512: the residual update 86.8 -> 85.4 us at
// 1024^2 x 16, the whole solve within its noise.
constexpr int kMaxVecBlocks = 1024;
constexpr int kStencilRows = 16;  // rows per block of the FP64 stencil kernels (64 x 16 threads)
constexpr int kFinThreads = 256;

// Split-layout geometry of one level.
struct Geo {
    int nx, ny, nz, hx;
    long plane;  // plane stride: 2 * hx * ny slots, plus kPlanePad on large levels
    __host__ __device__ int pos(int i, int j) const { return (2 * j + ((i + j) & 1)) * hx + (i >> 1); }
    __host__ __device__ long cells() const { return plane * nz; }
    __host__ __device__ long slots() const { return 2L * hx * ny; }  // used slots of a plane
};

// A plane stride that is a power of two (4 MB at 1024x1024 in FP32) puts a
// column's planes on the same HBM channels, and a column kernel reads all of
// them at once: padding the stride by 64 slots takes the fine sweep from 60
// to 51 us at 1024x1024x16. Pad slots are zero in every array: the column
// kernels never write them and the flat vector kernels keep zeros zero.
// Small levels (the LDS tail's) are not padded.
constexpr long kPlanePad = 64, kPlanePadMin = 16384;

inline Geo make_geo(int nx, int ny, int nz) {
    Geo g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.hx = (nx + 1) / 2;
    g.plane = 2L * g.hx * ny;
    if (g.plane >= kPlanePadMin) g.plane += kPlanePad;
    return g;
}

// Device-resident PCG scalars and flags (the role of the reference's sc[]
// slots, pcg3d_gpu.cu:80-131).
// One backward-Euler step as recorded on the device (layout of
// th_step_record: StepRecord, src/transient3d.h:29-35).
struct StepRec {
    double t, dt;
    int iters;
    double T_max, q_out, q_stored;
};

struct State {
    double inv_dt;  // 1/dt of the current operator (0 = steady)
    double tol;
    double bb, rr, rz, rzo, pAp, alpha, beta, relres;
    double relres0;  // relres at the start of the solve (its first residual)
    double rzb[2];                 // r.z by iteration parity (multigrid path)
    double chk_rr, chk_bb;         // true-residual check
    double e_heat, e_out, e_tmax;  // energy: sum C T, heat out, max T
    int maxit, iters, done, converged, zero_x, flexible;
    int xlag;  // multigrid path: x += alpha p of the last update is still to be applied
    int xz;    // x is known to be 0 at the start of this solve: its residual is b
    // Transient stepping on the device (transient3d.h:57-84): the time, the
    // step size of the current operator and of the next step, the stepping
    // rule, the heat content after the last step and that step's record.
    double t, dt, dt_next, dt_prev, growth, dt_max, t_end, heat_prev;
    int refactor;  // 1: the level factors must be rebuilt for inv_dt
    int steps;
    StepRec last;
    // Learned step start: this step's weight w on the guide direction, and
    // over the run the steps that took it (w > 0) and the sum of their w.
    double guide_w, guide_w_sum;
    int guide_steps;
    // Learned steady start judged on the device: both starting relative
    // residuals and whether the lifted guess was kept.
    double lift_rp, lift_rl;
    int lift_used;
};

// Block coordinates of a kernel body: the hardware ones (hw_blk), or, in a
// batched launch over many instances, those of the block within its
// instance's own grid.
struct Blk {
    int x, y, z, gx, gy, gz;
    __device__ int flat() const { return x + gx * (y + gy * z); }
    __device__ int flat_xy() const { return x + gx * y; }  // 2D grids (z = 0)
    // first thread index along x: blocks of BX threads, or blockDim.x (BX = 0)
    template <int BX>
    __device__ int x0() const { return BX ? x * BX : x * (int)blockDim.x; }
    __device__ bool first() const { return x == 0 && y == 0 && z == 0; }
};
__device__ __forceinline__ Blk hw_blk() {
    return Blk{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x, (int)gridDim.y, (int)gridDim.z};
}

// Conductance arrays of one level (split layout).
template <class T>
struct Ops {
    const T *gE, *gN, *gU, *gB, *cap;
};

// Column types of the fine level (levels.h: column_types): the type of each
// plane slot and the types' faces, table[type][plane][field]: one plane's
// faces are one aligned row, read with wide loads.
template <class T>
struct alignas(sizeof(T) * kTypeFields) TypeRow {
    T v[kTypeFields];
};
template <class T>
struct Typed {
    const unsigned short* __restrict__ ty;
    const T* __restrict__ tab;
    int nzp;
    __device__ __forceinline__ TypeRow<T> row(int t, int l) const {
        return reinterpret_cast<const TypeRow<T>*>(tab)[(long)t * nzp + l];
    }
    // The same row for a wave-uniform t, read once per wave through the
    // scalar cache (constant address space: the table does not change while
    // a kernel runs) instead of once per lane through L1.
    __device__ __forceinline__ TypeRow<T> urow(int t, int l) const {
        typedef const __attribute__((address_space(4))) T* CPtr;
        const CPtr c = (CPtr)tab + ((long)t * nzp + l) * kTypeFields;
        TypeRow<T> r;
#pragma unroll
        for (int k = 0; k < kTypeFields; ++k) r.v[k] = c[k];
        return r;
    }
    // The wave's representative type through urow, then the lanes whose
    // column has another type load their own row.
    __device__ __forceinline__ TypeRow<T> row(int t, int rep, int l) const {
        TypeRow<T> r = urow(rep, l);
        if (t != rep) r = row(t, l);
        return r;
    }
};

// A wave-uniform type from the active lanes: the first active lane of the
// wave's upper half (else of the lower half), so that a domain-edge column,
// which has a type of its own and sits in the first or last lane of a row,
// is not the representative of its wave.
__device__ __forceinline__ int wave_rep_type(int t) {
    const unsigned long long m = __ballot(1);
    const unsigned hi = (unsigned)(m >> 32);
    const int lane = hi ? 32 + __builtin_ctz(hi) : __builtin_ctzll(m);
    return __builtin_amdgcn_readlane(t, lane);
}

// Plane-count buckets of the fine-level column kernels: an instantiation for
// NZ in {4, 8, 12, 16, 24, 32, 48, 64} runs the levels with
// nz_floor(NZ) < nz <= NZ (the launch ladders pick the smallest bucket that
// holds nz), so planes l <= nz_floor(NZ) exist in every level it runs and
// need no run-time test.
constexpr int nz_floor(int NZ) {
    return NZ <= 4 ? 0 : NZ <= 8 ? 4 : NZ <= 12 ? 8 : NZ <= 16 ? 12 : NZ <= 24 ? 16 : NZ <= 32 ? 24 : NZ <= 48 ? 32 : 48;
}
template <int NZ>
__device__ __forceinline__ bool plane_live(int l, int nz) {
    return l <= nz_floor(NZ) || l < nz;
}

// Two-element vector of T (16-byte loads of double pairs, 8-byte of float).
template <class T>
struct Vec2;
template <>
struct Vec2<float> {
    typedef float2 type;
};
template <>
struct Vec2<double> {
    typedef double2 type;
};

// Loads and stores with the nontemporal hint (global_load / global_store ...
// nt): their lines are the first the caches give up. For streams that no
// kernel soon after reads again; the caches keep the arrays the next kernels
// reuse, and fewer dirty lines are evicted early. Same values as plain
// loads and stores.
template <class T>
struct NtVec2 {
    typedef T type __attribute__((ext_vector_type(2)));
};
template <class T>
__device__ __forceinline__ typename Vec2<T>::type load2_nt(const T* __restrict__ a, long k) {
    typedef typename NtVec2<T>::type V;
    const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(a) + k);
    typename Vec2<T>::type r;
    r.x = v.x;
    r.y = v.y;
    return r;
}
template <class T>
__device__ __forceinline__ T load_nt(const T* __restrict__ a, long k) {
    return __builtin_nontemporal_load(a + k);
}
template <class T>
__device__ __forceinline__ void store2_nt(T* __restrict__ a, long k, const typename Vec2<T>::type& v) {
    typedef typename NtVec2<T>::type V;
    V w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<V*>(a) + k);
}

// ---- reductions -------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Block totals of NV values into part[block * NV + k] (bit k of KMAX set:
// max instead of sum). Every thread of the block must call it.
template <int NV, unsigned KMAX = 0>
__device__ void block_partials(double (&v)[NV], double* __restrict__ part, int bid) {
    __shared__ double sh[NV][16];
    const int nthr = blockDim.x * blockDim.y;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int lane = tid & 63, w = tid >> 6, nw = (nthr + 63) >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = ((KMAX >> k) & 1) ? wave_max(v[k]) : wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][w] = v[k];
    __syncthreads();
    if (tid < NV) {
        double s = sh[tid][0];
        for (int q = 1; q < nw; ++q) s = ((KMAX >> tid) & 1) ? fmax(s, sh[tid][q]) : s + sh[tid][q];
        part[(size_t)bid * NV + tid] = s;
    }
}
// The partials of blocks of 64 x kColsY threads from a block of 64 x
// (m kColsY) threads (one wave per row) that runs in place of m of them: its
// rows in m groups of kColsY waves, each group summed as block_partials sums
// such a block, into the slot that block has (x + gx * row block), for the
// row blocks below nrb.
template <int NV>
__device__ void row_group_partials(double (&v)[NV], double* __restrict__ part, const Blk& B, int nrb) {
    __shared__ double sh[NV][16];
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][w] = v[k];
    __syncthreads();
    const int m = blockDim.y / kColsY;
    const int q = tid / NV, k = tid - q * NV, rb = B.y * m + q;
    if (q < m && rb < nrb) {
        double s = sh[k][q * kColsY];
        for (int i = 1; i < kColsY; ++i) s += sh[k][q * kColsY + i];
        part[((size_t)B.x + (size_t)B.gx * rb) * NV + k] = s;
    }
}
template <int NV, unsigned KMAX = 0>
__device__ void block_partials(double (&v)[NV], double* __restrict__ part) {
    block_partials<NV, KMAX>(v, part, (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)));
}

// Grid totals from nblk block partials, in block order (one block of
// kFinThreads threads). Result valid in thread 0.
template <int NV, unsigned KMAX = 0>
__device__ void sum_partials(const double* __restrict__ part, int nblk, double (&tot)[NV]) {
    __shared__ double sh[NV][kFinThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = ((KMAX >> k) & 1) ? -1.0e300 : 0.0;
    for (int b = tid; b < nblk; b += kFinThreads)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const double x = part[(size_t)b * NV + k];
            acc[k] = ((KMAX >> k) & 1) ? fmax(acc[k], x) : acc[k] + x;
        }
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = ((KMAX >> k) & 1) ? wave_max(acc[k]) : wave_sum(acc[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][w] = acc[k];
    __syncthreads();
    if (tid == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = sh[k][0];
            for (int q = 1; q < kFinThreads / 64; ++q) s = ((KMAX >> k) & 1) ? fmax(s, sh[k][q]) : s + sh[k][q];
            tot[k] = s;
        }
}

// Grid totals of a previous kernel's nblk partials, in block order, in every
// thread of the calling block (each block of a consumer kernel computes the
// same totals, so no separate reduction launch is needed).
template <int NV>
__device__ void block_sum_partials(const double* __restrict__ part, int nblk, double (&tot)[NV]) {
    __shared__ double sh[NV][16];
    const int nthr = blockDim.x * blockDim.y;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int lane = tid & 63, w = tid >> 6, nw = (nthr + 63) >> 6;
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
#pragma unroll 4
    for (int b = tid; b < nblk; b += nthr)
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] += part[(size_t)b * NV + k];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = wave_sum(acc[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][w] = acc[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = sh[k][0];
        for (int q = 1; q < nw; ++q) s += sh[k][q];
        tot[k] = s;
    }
    __syncthreads();
}

// Convergence test at the head of the V-cycle, from the partials of the
// kernel before it: the start residual (CONS_START: r.r, b.b) or the x, r
// update of an iteration (CONS_ITER: r.r, r.z_old). All blocks reach the same
// decision; block 0 records it. Returns true when the solve is finished
// (pcg3d_cpu.cpp:141-176: b = 0, ||r||/||b|| < tol, or maxit reached).
enum { CONS_NONE = 0, CONS_START = 1, CONS_ITER = 2 };
__device__ inline bool consume_convergence(const double* __restrict__ part, int nblk, int mode, State* st,
                                           bool first_block) {
    const bool lead = first_block && threadIdx.x == 0 && threadIdx.y == 0;
    double tot[3] = {0.0, 0.0, 0.0};
    if (mode == CONS_START) {
        block_sum_partials<3>(part, nblk, tot);
    } else {
        double t2[2];
        block_sum_partials<2>(part, nblk, t2);
        tot[0] = t2[0];
        tot[1] = t2[1];
    }
    const double tol = st->tol;
    if (mode == CONS_START) {
        const double bb = tot[1];
        const double relres = bb == 0.0 ? 0.0 : sqrt(tot[0] / bb);
        const bool fin = bb == 0.0 || relres < tol;
        if (lead) {
            st->bb = bb;
            st->rr = tot[0];
            st->rzo = 0.0;
            st->relres = relres;
            st->relres0 = relres;
            st->zero_x = bb == 0.0;
            st->converged = fin;
            st->done = fin;
        }
        return fin;
    }
    const double relres = sqrt(tot[0] / st->bb);
    const bool conv = relres < tol;
    if (lead) {
        const int it = st->iters + 1;
        st->iters = it;
        st->rr = tot[0];
        st->rzo = tot[1];
        st->relres = relres;
        st->converged = conv;
        st->done = conv || it >= st->maxit;
    }
    return conv;
}
__device__ inline bool consume_convergence(const double* __restrict__ part, int nblk, int mode, State* st) {
    return consume_convergence(part, nblk, mode, st, blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0);
}

// ---- column addressing ------------------------------------------------------

// A column (i, j) and the in-plane slots of its 4 lateral neighbours.
struct Col {
    int c0, pw, pe, ps, pn;
    bool hw, he, hs, hn;
};
__device__ __forceinline__ Col make_col(const Geo& g, int i, int j) {
    Col c;
    const int colour = (i + j) & 1, oc = colour ^ 1, t = i >> 1;
    c.c0 = (2 * j + colour) * g.hx + t;
    c.hw = i > 0;
    c.he = i < g.nx - 1;
    c.hs = j > 0;
    c.hn = j < g.ny - 1;
    c.pw = (2 * j + oc) * g.hx + ((i - 1) >> 1);
    c.pe = (2 * j + oc) * g.hx + ((i + 1) >> 1);
    c.ps = (2 * j - 2 + oc) * g.hx + t;
    c.pn = (2 * j + 2 + oc) * g.hx + t;
    // Absent neighbours point at the cell itself, so loads stay in bounds.
    if (!c.hw) c.pw = c.c0;
    if (!c.he) c.pe = c.c0;
    if (!c.hs) c.ps = c.c0;
    if (!c.hn) c.pn = c.c0;
    return c;
}

// Column of the row slot q (both colours) of row j; false for padding.
__device__ __forceinline__ bool slot_column(const Geo& g, int q, int j, int& i) {
    const int colour = q >= g.hx ? 1 : 0, t = q - colour * g.hx;
    i = 2 * t + ((j + colour) & 1);
    return i < g.nx;
}

// (A + C/dt) x at cell (column c, plane l) in precision R, in the
// reference's operation order (operator3d.h:81-112): diagonal = lateral faces
// (W, E, S, N) + up + (down or bottom ambient) + C/dt; off-diagonal = lateral
// G x (W, E, S, N) + below + above. All loads are issued before the
// arithmetic (absent neighbours read the cell itself and are masked out).
template <class R, class G, class V>
__device__ __forceinline__ R apply_cell(const Geo& g, const G* __restrict__ gE, const G* __restrict__ gN,
                                        const G* __restrict__ gU, const G* __restrict__ gB,
                                        const G* __restrict__ cap, const V* __restrict__ x, const Col& c, int l,
                                        R inv_dt) {
    const long base = (long)l * g.plane;
    const long id = base + c.c0;
    const long idd = (l > 0) ? id - g.plane : id;
    const long idu = (l < g.nz - 1) ? id + g.plane : id;
    const R Gw = (R)gE[base + c.pw], Ge = (R)gE[id], Gs = (R)gN[base + c.ps], Gn = (R)gN[id];
    const R Gu = (R)gU[id], Gd = (l > 0) ? (R)gU[idd] : (R)gB[c.c0];
    const R xw = (R)x[base + c.pw], xe = (R)x[base + c.pe], xs = (R)x[base + c.ps], xn = (R)x[base + c.pn];
    const R xc = (R)x[id], xd = (R)x[idd], xu = (R)x[idu];
    R d = 0, s = 0;
    if (c.hw) d += Gw;
    if (c.he) d += Ge;
    if (c.hs) d += Gs;
    if (c.hn) d += Gn;
    d += Gu;
    d += Gd;
    if (inv_dt != R(0)) d = d + (R)cap[id] * inv_dt;
    if (c.hw) s += Gw * xw;
    if (c.he) s += Ge * xe;
    if (c.hs) s += Gs * xs;
    if (c.hn) s += Gn * xn;
    if (l > 0) s += Gd * xd;
    if (l < g.nz - 1) s += Gu * xu;
    return d * xc - s;
}

// (A + C/dt) x at cell (column c, plane l) in FP64, in flux form:
//   sum over faces G (x_c - x_nb) + ambient links x_c + C/dt x_c.
// Mathematically the same operator as apply_cell, but the terms are face
// fluxes instead of the large d x and G x_nb products, so the rounding floor
// of a residual b - A x drops from ~eps * max|G x| per cell (about 2e-10 of
// ||P|| at 1024^2, where the 250 um copper planes have G ~ 0.1 W/K) to
// ~eps * |G dx|. Used for residuals that decide convergence.
template <class G>
__device__ __forceinline__ double flux_apply_cell(const Geo& g, const G* __restrict__ gE, const G* __restrict__ gN,
                                                  const G* __restrict__ gU, const G* __restrict__ gB,
                                                  const G* __restrict__ cap, const double* __restrict__ x,
                                                  const Col& c, int l, double inv_dt) {
    const long base = (long)l * g.plane;
    const long id = base + c.c0;
    const long idd = (l > 0) ? id - g.plane : id;
    const long idu = (l < g.nz - 1) ? id + g.plane : id;
    const double Gw = (double)gE[base + c.pw], Ge = (double)gE[id], Gs = (double)gN[base + c.ps],
                 Gn = (double)gN[id];
    const double Gu = (double)gU[id], Gd = (l > 0) ? (double)gU[idd] : (double)gB[c.c0];
    const double xc = x[id];
    const double xw = x[base + c.pw], xe = x[base + c.pe], xs = x[base + c.ps], xn = x[base + c.pn];
    const double xd = x[idd], xu = x[idu];
    double f = 0.0;
    if (c.hw) f += Gw * (xc - xw);
    if (c.he) f += Ge * (xc - xe);
    if (c.hs) f += Gs * (xc - xs);
    if (c.hn) f += Gn * (xc - xn);
    f += (l > 0) ? Gd * (xc - xd) : Gd * xc;         // below, or bottom ambient
    f += (l < g.nz - 1) ? Gu * (xc - xu) : Gu * xc;  // above, or the heat sink
    if (inv_dt != 0.0) f += (double)cap[id] * inv_dt * xc;
    return f;
}

}  // namespace th
