// kernels.h - gfx950 kernels of the MG-PCG solver.
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
// One thread per column: it walks the nz planes, so vertical neighbours stay
// in registers and each plane's loads are issued back to back. The column
// count per launch (e.g. 2^20 at 1024^2) fills the chip's 256 CUs many
// times over.
//
// Reductions are deterministic: wave64 butterfly, then the waves of a block
// through LDS, then one partial per block; the last block to finish (an
// atomic ticket) adds the partials in block order and runs the scalar
// epilogue (alpha, beta, ||r||, convergence), so PCG scalars never leave the
// device and no floating-point atomics are used.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace th {

constexpr int kColsX = 64;  // threads per block along a half-row (one wave)
constexpr int kColsY = 4;   // rows per block
constexpr int kVecThreads = 256;
constexpr int kMaxVecBlocks = 2048;

// Split-layout geometry of one level.
struct Geo {
    int nx, ny, nz, hx;
    long plane;  // 2 * hx * ny
    __host__ __device__ int pos(int i, int j) const { return (2 * j + ((i + j) & 1)) * hx + (i >> 1); }
    __host__ __device__ long cells() const { return plane * nz; }
};

inline Geo make_geo(int nx, int ny, int nz) {
    Geo g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.hx = (nx + 1) / 2;
    g.plane = 2L * g.hx * ny;
    return g;
}

// Device-resident PCG scalars and flags (the role of the reference's sc[]
// slots, pcg3d_gpu.cu:80-131), plus the tickets of the last-block reductions.
struct State {
    double inv_dt;       // 1/dt of the current operator (0 = steady)
    double tol;
    double bb, rr, rz, rzo, pAp, alpha, beta, relres;
    double chk_rr, chk_bb;          // true-residual check
    double e_heat, e_out, e_tmax;   // energy: sum C T, heat out, max T
    int maxit, iters, done, converged, zero_x, flexible;
    unsigned ticket[8];
};
enum { TK_RES = 0, TK_APPLY, TK_UPDATE, TK_DOT, TK_CHECK, TK_ENERGY };

// Conductance arrays of one level (split layout).
template <class T>
struct Ops {
    const T *gE, *gN, *gU, *gB, *cap;
};

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

// Reduces NV values over the grid. Every thread of every block must call it.
// Returns true in the one block that finished last; there, thread 0 holds the
// grid totals in tot[]. KMAX marks which entries reduce by max instead of sum.
template <int NV, unsigned KMAX = 0>
__device__ bool grid_reduce(double (&v)[NV], double* part, unsigned* ticket, double (&tot)[NV]) {
    __shared__ double sh[NV][16];
    __shared__ int last;
    const int nthr = blockDim.x * blockDim.y;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int lane = tid & 63, w = tid >> 6, nw = (nthr + 63) >> 6;
    const int nblk = gridDim.x * gridDim.y * gridDim.z;
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = ((KMAX >> k) & 1) ? wave_max(v[k]) : wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][w] = v[k];
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = sh[k][0];
            for (int q = 1; q < nw; ++q) s = ((KMAX >> k) & 1) ? fmax(s, sh[k][q]) : s + sh[k][q];
            part[(size_t)bid * NV + k] = s;
        }
        __threadfence();
        const unsigned t = atomicAdd(ticket, 1u);
        last = (t == (unsigned)nblk - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double acc[NV];
    const volatile double* vp = part;
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = ((KMAX >> k) & 1) ? -1.0e300 : 0.0;
    for (int b = tid; b < nblk; b += nthr)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const double x = vp[(size_t)b * NV + k];
            acc[k] = ((KMAX >> k) & 1) ? fmax(acc[k], x) : acc[k] + x;
        }
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = ((KMAX >> k) & 1) ? wave_max(acc[k]) : wave_sum(acc[k]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][w] = acc[k];
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = sh[k][0];
            for (int q = 1; q < nw; ++q) s = ((KMAX >> k) & 1) ? fmax(s, sh[k][q]) : s + sh[k][q];
            tot[k] = s;
        }
        *ticket = 0u;
    }
    return true;
}

// ---- column addressing ------------------------------------------------------

// A column (i, j) and the in-plane slots of its 4 lateral neighbours.
struct Col {
    int i, j, c0, pw, pe, ps, pn;
    bool hw, he, hs, hn;
};
__device__ __forceinline__ Col make_col(const Geo& g, int i, int j) {
    Col c;
    c.i = i;
    c.j = j;
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
    return c;
}

// (A + C/dt) applied at cell (column c, plane l), in precision R, with the
// reference's operation order (operator3d.h:81-112): diagonal = lateral faces
// (W, E, S, N) + up + (down or bottom ambient) + C/dt; off-diagonal = lateral
// G x (W, E, S, N) + below + above.
template <class R, class G, class V>
__device__ __forceinline__ R apply_cell(const Geo& g, const Ops<G>& o, const V* __restrict__ x, const Col& c,
                                        int l, R inv_dt) {
    const long base = (long)l * g.plane;
    const long id = base + c.c0;
    R d = 0, s = 0;
    if (c.hw) {
        const R gw = (R)o.gE[base + c.pw];
        d += gw;
    }
    if (c.he) d += (R)o.gE[id];
    if (c.hs) d += (R)o.gN[base + c.ps];
    if (c.hn) d += (R)o.gN[id];
    d += (R)o.gU[id];
    d += (l > 0) ? (R)o.gU[id - g.plane] : (R)o.gB[c.c0];
    if (inv_dt != R(0)) d = d + (R)o.cap[id] * inv_dt;
    if (c.hw) s += (R)o.gE[base + c.pw] * (R)x[base + c.pw];
    if (c.he) s += (R)o.gE[id] * (R)x[base + c.pe];
    if (c.hs) s += (R)o.gN[base + c.ps] * (R)x[base + c.ps];
    if (c.hn) s += (R)o.gN[id] * (R)x[base + c.pn];
    if (l > 0) s += (R)o.gU[id - g.plane] * (R)x[id - g.plane];
    if (l < g.nz - 1) s += (R)o.gU[id] * (R)x[id + g.plane];
    return d * (R)x[id] - s;
}

}  // namespace th
