// batch.hip - batched steady MG-PCG: many independent stacks (grid sizes may
// differ; the plane count bucket is shared) solved together on one GPU, and
// its C ABI (include/thermal_hip.h). The reference solves one stack per
// solver; this is the batched engine of the project's benchmark description.
//
// Each stage of a PCG iteration is ONE launch over all instances: the flat
// block index is mapped through a per-stage table (instance, first block,
// grid) to the instance and a block of its own grid, and the block runs the
// same body as the single-instance kernel (solver.h) on that instance's
// buffers. Instances that have converged are masked by their own device
// state: their blocks return at once, so an iteration's launches cost the
// unconverged instances only. The V-cycle descends level by level for all
// instances at once (instances whose coarse tail starts at that level drop
// out), runs every instance's tail as one workgroup per instance with its
// levels in LDS, and climbs back. Start and iteration are captured as graphs.
//
// Setup reuses th::Solver<float> per instance (levels, factors, buffers).
// The autotuner times the V-cycle stages for blocks of 1, 2, 4 or 8
// wavefronts whose wavefront covers 64 columns of one row, 32 of two or 16
// of four (columns per wavefront), and the vector stages for 256- or
// 1024-thread blocks, on the batch itself, and keeps the fastest.
#include "solver.h"

namespace th {
namespace bat {

constexpr int kMaxLev = 16;

struct LevelView {
    Geo g;
    Ops<float> o;
    const float* finv;
    float* x;
    float* b;
    Typed<float> y;  // column types (ntypes > 0): the sweeps read them instead of o and finv
    int ntypes;
};

struct Inst {
    State* st;
    Geo g0;
    long n0;
    const double* P;
    double *x, *r, *Ap;
    float* p;  // search direction (FP32, as the single-instance mixed solver)
    double *part_res, *part_apply, *part_upd, *part_dot;
    int nblk_stencil, nblk_vec;
    int nlev, tail_k0;
    LevelView L[kMaxLev];
};

// An instance taking part in a stage: its first flat block and its grid.
struct Entry {
    int inst, begin, gx, gy, gz;
};

// Instance and block coordinates of this block in a stage's table (the
// table covers the launch exactly, its entries in block order): the last
// entry whose first block is <= this one. Each wave tests 64 entries with one
// load per lane and a ballot (the entries that pass are a prefix), so a
// block start waits for one load per 64 instances instead of a binary
// search's chain of dependent loads. Every lane of the wave must be active.
__device__ __forceinline__ const Inst& locate(const Inst* __restrict__ I, const Entry* __restrict__ tab, int n,
                                              Blk& B) {
    const int b = blockIdx.x;
    const int lane = (int)((threadIdx.y * blockDim.x + threadIdx.x) & 63);
    int lo = 0;
    for (int base = 0; base < n; base += 64) {
        const int k = base + lane;
        const unsigned long long m = __ballot(k < n && tab[k].begin <= b);
        const int c = __popcll(m);
        if (c) lo = base + c - 1;
        if (c < 64) break;
    }
    const Entry e = tab[lo];
    int r = b - e.begin;
    B.gx = e.gx;
    B.gy = e.gy;
    B.gz = e.gz;
    B.x = r % e.gx;
    r /= e.gx;
    B.y = r % e.gy;
    B.z = r / e.gy;
    return I[e.inst];
}

// Typed levels run the prolongation sweeps up to 16 planes with the
// type-independent loads first (sweep_typed_lf_body), as the single solves.
// (The single solves' loads-first pre-smoothing sweep is slower here.)
template <int NZ, int MODE>
__global__ __launch_bounds__(512) void kb_sweep(const Inst* __restrict__ I, const Entry* __restrict__ tab, int ntab,
                                                int k, int colour, int cons) {
    Blk B;
    const Inst& in = locate(I, tab, ntab, B);
    // descriptors copied out of the table: values the stores cannot alias,
    // so the compiler keeps them in scalar registers instead of reloading
    const LevelView L = in.L[k];
    const int kc = k + 1 < in.nlev ? k + 1 : k;
    const Geo gc = in.L[kc].g;
    const float* xc = in.L[kc].x;
    State* st = in.st;
    const double* cp = cons == CONS_START ? in.part_res : in.part_upd;
    const int cn = cons == CONS_START ? in.nblk_stencil : in.nblk_vec;
    const int cm = k == 0 ? cons : (int)CONS_NONE;
    if constexpr (MODE == SW_PROLONG && NZ <= 16) {
        if (L.ntypes) {
            sweep_typed_lf_body<float, NZ, MODE>(B, L.g, L.y, L.b, L.x, colour, gc, xc, st, cp, cn, cm);
            return;
        }
    }
    if (L.ntypes)
        sweep_typed_body<float, NZ, MODE, false>(B, L.g, L.y, L.b, L.x, colour, gc, xc, st, cp, cn, cm);
    else
        sweep_body<float, NZ, MODE>(B, L.g, L.o, L.finv, L.b, L.x, colour, gc, xc, st, cp, cn, cm);
}

__global__ __launch_bounds__(512) void kb_restrict(const Inst* __restrict__ I, const Entry* __restrict__ tab, int ntab,
                                                   int k) {
    Blk B;
    const Inst& in = locate(I, tab, ntab, B);
    const LevelView L = in.L[k];  // copies, as in kb_sweep
    const Geo gc = in.L[k + 1].g;
    float* bc = in.L[k + 1].b;
    const State* st = in.st;
    if (L.ntypes)  // kRestrictPlanes planes per block z (the table's z), as k_restrict_fast_typed
        restrict_fast_typed_body<float, kRestrictPlanes>(B, L.g, L.y, L.x, gc, bc, st);
    else
        restrict_fast_body<float>(B, L.g, L.o, L.x, gc, bc, st);
}

// One workgroup per instance; tails[i] describes instance i's tail levels.
template <int NZ>
__global__ __launch_bounds__(kTailLdsThreads) void kb_tail(const Inst* __restrict__ I,
                                                           const Tail<float>* __restrict__ tails, int cons) {
    extern __shared__ __align__(16) unsigned char smem[];
    const Inst& in = I[blockIdx.x];
    const double* cp = cons == CONS_START ? in.part_res : in.part_upd;
    const int cn = cons == CONS_START ? in.nblk_stencil : in.nblk_vec;
    tail_lds_body<float, NZ>(tails[blockIdx.x], in.st, cp, cn, in.tail_k0 == 0 ? cons : (int)CONS_NONE,
                             reinterpret_cast<float*>(smem));
}

// Level-0 stencils in blocks of 64 x 16. Instances whose level 0 has column
// types run the typed bodies from a table of their own, each block writing
// the partials of the four 64 x 4 blocks the single-instance solver runs in
// its place (the same sums); the others the per-cell bodies.
static_assert(kStencilRows % kColsY == 0 && kStencilRows <= 16, "whole row groups, at most 16 waves per block");
__global__ __launch_bounds__(1024) void kb_residual(const Inst* __restrict__ I, const Entry* __restrict__ tab,
                                                    int ntab) {
    Blk B;
    const Inst& in = locate(I, tab, ntab, B);
    const Geo g = in.g0;  // copies, as in kb_sweep
    const Ops<float> o = in.L[0].o;
    residual_body<float, RES_MG>(B, g, o, in.P, in.x, in.r, in.L[0].b, nullptr, nullptr, nullptr, in.st, in.part_res);
}
template <int NZ>
__global__ __launch_bounds__(1024) void kb_residual_typed(const Inst* __restrict__ I, const Entry* __restrict__ tab,
                                                          int ntab) {
    Blk B;
    const Inst& in = locate(I, tab, ntab, B);
    const Geo g = in.g0;  // copies, as in kb_sweep
    const Typed<float> y = in.L[0].y;
    residual_typed_body<float, NZ, RES_MG, true>(B, g, y, in.P, in.x, in.r, in.L[0].b, in.st, in.part_res, nullptr);
}

template <int NZ>
__global__ __launch_bounds__(1024) void kb_apply(const Inst* __restrict__ I, const Entry* __restrict__ tab, int ntab) {
    Blk B;
    const Inst& in = locate(I, tab, ntab, B);
    const Geo g = in.g0;  // copies, as in kb_sweep
    const Ops<float> o = in.L[0].o;
    apply_body<float, NZ, float>(B, g, o, in.p, in.Ap, in.st, in.part_apply);
}
template <int NZ>
__global__ __launch_bounds__(1024) void kb_apply_typed(const Inst* __restrict__ I, const Entry* __restrict__ tab,
                                                       int ntab) {
    Blk B;
    const Inst& in = locate(I, tab, ntab, B);
    const Geo g = in.g0;  // copies, as in kb_sweep
    const Typed<float> y = in.L[0].y;
    apply_typed_body<float, NZ, float, true>(B, g, y, in.p, in.Ap, in.st, in.part_apply);
}

__global__ __launch_bounds__(1024) void kb_update(const Inst* __restrict__ I, const Entry* __restrict__ tab, int ntab) {
    Blk B;
    const Inst& in = locate(I, tab, ntab, B);
    update_body<float, UP_MG_FLEX>(B, in.n0, nullptr, in.Ap, in.x, in.r, in.L[0].b, in.L[0].x, nullptr, nullptr, in.st,
                                   in.part_upd, in.part_apply, in.nblk_stencil);
}

__global__ __launch_bounds__(1024) void kb_dot(const Inst* __restrict__ I, const Entry* __restrict__ tab, int ntab) {
    Blk B;
    const Inst& in = locate(I, tab, ntab, B);
    dot_body<float>(B, in.n0, in.r, in.L[0].x, in.st, in.part_dot);
}

template <int FIRST>
__global__ __launch_bounds__(1024) void kb_pupdate(const Inst* __restrict__ I, const Entry* __restrict__ tab,
                                                   int ntab) {
    Blk B;
    const Inst& in = locate(I, tab, ntab, B);
    pupdate_body<float, FIRST, float>(B, in.n0, in.L[0].x, in.p, FIRST ? nullptr : in.x, in.st, in.part_dot,
                                      in.nblk_vec);
}

// Zero start for every instance: x = 0 (the start residual reads P as b).
__global__ __launch_bounds__(1024) void kb_reset(const Inst* __restrict__ I, const Entry* __restrict__ tab, int ntab) {
    Blk B;
    const Inst& in = locate(I, tab, ntab, B);
    const long stride = (long)B.gx * blockDim.x;
    for (long k = (long)B.x * blockDim.x + threadIdx.x; k < in.n0; k += stride) {
        in.x[k] = 0.0;
    }
}

// x = 0 where b = 0 was detected, else the last update's pending
// x += alpha p (as k_finish).
__global__ __launch_bounds__(1024) void kb_finish(const Inst* __restrict__ I, const Entry* __restrict__ tab, int ntab) {
    Blk B;
    const Inst& in = locate(I, tab, ntab, B);
    const long stride = (long)B.gx * blockDim.x;
    if (in.st->zero_x) {
        for (long k = (long)B.x * blockDim.x + threadIdx.x; k < in.n0; k += stride) in.x[k] = 0.0;
    } else if (in.st->xlag) {
        const double alpha = in.st->alpha;
        for (long k = (long)B.x * blockDim.x + threadIdx.x; k < in.n0; k += stride) in.x[k] += alpha * (double)in.p[k];
    }
}

// xz: the start residual is b = P itself (x = 0, reset before the start).
__global__ void kb_begin(const Inst* __restrict__ I, int n, double tol, int maxit, int xz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    State* st = I[i].st;
    st->xz = xz;
    st->xlag = 0;
    st->tol = tol;
    st->maxit = maxit;
    st->iters = 0;
    st->done = 0;
    st->converged = 0;
    st->zero_x = 0;
    st->relres = 0.0;
}

__global__ void kb_gather(const Inst* __restrict__ I, int n, State* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = *I[i].st;
}

// ---- host ----------------------------------------------------------------------

struct Table {
    std::vector<Entry> h;
    Entry* d = nullptr;
    int blocks = 0;
};

struct Batch {
    std::vector<Solver<float>*> S;
    hipStream_t s = nullptr;
    Inst* dI = nullptr;
    Tail<float>* dTail = nullptr;
    State* dSt = nullptr;
    State* hSt = nullptr;
    std::vector<void*> own;
    int n = 0, nz = 0, maxk = 0;  // maxk: deepest level with separate kernels + 1
    size_t lds = 0;
    int rows = 4, cols = 64, vthreads = 1024;  // V-cycle blocks: rows x (64 / cols) waves of cols columns
    Table t_stencil, t_typed, t_vec;  // level-0 stencils: per-cell instances, typed instances
    std::vector<Table> t_sweep, t_restrict;
    hipGraphExec_t g_start = nullptr, g_iter = nullptr;

    ~Batch() {
        if (s) (void)hipStreamSynchronize(s);
        drop_graphs();
        free_tables();
        for (void* p : own) (void)hipFree(p);
        if (hSt) (void)hipHostFree(hSt);
        if (s) (void)hipStreamDestroy(s);
        for (Solver<float>* x : S) delete x;
    }
    void drop_graphs() {
        if (g_start) (void)hipGraphExecDestroy(g_start);
        if (g_iter) (void)hipGraphExecDestroy(g_iter);
        g_start = g_iter = nullptr;
    }
    void free_table(Table& t) {
        if (t.d) (void)hipFree(t.d);
        t.d = nullptr;
        t.h.clear();
        t.blocks = 0;
    }
    void free_tables() {
        free_table(t_stencil);
        free_table(t_typed);
        free_table(t_vec);
        for (Table& t : t_sweep) free_table(t);
        for (Table& t : t_restrict) free_table(t);
        t_sweep.clear();
        t_restrict.clear();
    }
    int upload_table(Table& t) {
        if (t.h.empty()) return TH_OK;
        HIPTRY(hipMalloc(&t.d, t.h.size() * sizeof(Entry)));
        HIPTRY(hipMemcpy(t.d, t.h.data(), t.h.size() * sizeof(Entry), hipMemcpyHostToDevice));
        return TH_OK;
    }
    static void add(Table& t, int inst, int gx, int gy, int gz) {
        t.h.push_back(Entry{inst, t.blocks, gx, gy, gz});
        t.blocks += gx * gy * gz;
    }

    int init(int count, const th_model_desc* const* m, const th_solver_opts& o) {
        HIPTRY(hipSetDevice(o.device));
        HIPTRY(hipStreamCreate(&s));
        n = count;
        th_solver_opts io = o;
        io.use_graphs = 0;
        io.true_residual = 0;
        for (int i = 0; i < n; ++i) {
            Solver<float>* x = new Solver<float>();
            S.push_back(x);
            int rc = x->init(*m[i], io);
            if (rc) return rc;
            k_set_flexible<<<1, 1, 0, x->s>>>(x->st, 1);
            HIPTRY(hipGetLastError());
            HIPTRY(hipStreamSynchronize(x->s));
            if (!x->lds_tail) return fail(TH_ERR_ARG, "batched solver: every instance needs the LDS tail (nz <= 16)");
            if ((int)x->L.size() > kMaxLev) return fail(TH_ERR_ARG, "batched solver: too many levels");
            if (i == 0) nz = x->g0.nz;
            if (bucket(x->g0.nz) != bucket(nz)) return fail(TH_ERR_ARG, "batched solver: plane counts must share a bucket");
        }
        // device descriptors; partials for the vector stages sized for any block size
        std::vector<Inst> hI(n);
        std::vector<Tail<float>> hT(n);
        for (int i = 0; i < n; ++i) {
            Solver<float>& x = *S[i];
            Inst& d = hI[i];
            d.st = x.st;
            d.g0 = x.g0;
            d.n0 = x.n0;
            d.P = x.P;
            d.x = x.x;
            d.r = x.r;
            d.p = x.pm;
            d.Ap = x.Ap;
            d.part_res = x.part_res;
            d.part_apply = x.part_apply;
            // partials of the level-0 stencils: blocks of 64 x 4 with types (nblk_apply = nblk_res)
            d.nblk_stencil = (int)(x.L[0].ntypes ? x.nblk_apply : x.nblk_stencil);
            int rc;
            if ((rc = dalloc(&d.part_upd, (size_t)kMaxVecBlocks * 2)) || (rc = dalloc(&d.part_dot, kMaxVecBlocks)))
                return rc;
            d.nblk_vec = 0;  // set with the vector block size
            d.nlev = (int)x.L.size();
            d.tail_k0 = (int)x.tail_k0;
            for (size_t k = 0; k < x.L.size(); ++k)
                d.L[k] = LevelView{x.L[k].g, x.L[k].ops(), x.L[k].finv, x.L[k].x, x.L[k].b, x.L[k].typed(),
                                   x.L[k].ntypes};
            Tail<float>& t = hT[i];
            t.n = d.nlev - d.tail_k0;
            for (int k = 0; k < t.n; ++k) {
                const DevLevel<float>& L = x.L[d.tail_k0 + k];
                // the tail reads per-cell factors, which typed levels do not have
                if (L.ntypes) return fail(TH_ERR_ARG, "batched solver: a typed level in the LDS tail");
                t.L[k] = TailLevel<float>{L.g, L.ops(), L.finv, L.x, L.b};
            }
            lds = std::max(lds, x.lds_bytes);
            maxk = std::max(maxk, d.tail_k0);
        }
        hInst = hI;
        int rc;
        if ((rc = dalloc(&dI, n)) || (rc = dalloc(&dTail, n)) || (rc = dalloc(&dSt, n))) return rc;
        HIPTRY(hipMemcpy(dTail, hT.data(), n * sizeof(Tail<float>), hipMemcpyHostToDevice));
        HIPTRY(hipHostMalloc((void**)&hSt, n * sizeof(State), hipHostMallocDefault));
        HIPTRY(hipFuncSetAttribute((const void*)kb_tail<4>, hipFuncAttributeMaxDynamicSharedMemorySize, kTailLdsBytes));
        HIPTRY(hipFuncSetAttribute((const void*)kb_tail<8>, hipFuncAttributeMaxDynamicSharedMemorySize, kTailLdsBytes));
        HIPTRY(hipFuncSetAttribute((const void*)kb_tail<12>, hipFuncAttributeMaxDynamicSharedMemorySize, kTailLdsBytes));
        HIPTRY(hipFuncSetAttribute((const void*)kb_tail<16>, hipFuncAttributeMaxDynamicSharedMemorySize, kTailLdsBytes));
        return configure(4, 64, 1024);
    }
    std::vector<Inst> hInst;

    static int bucket(int z) { return z <= 4 ? 4 : z <= 8 ? 8 : z <= 12 ? 12 : z <= 16 ? 16 : 64; }

    template <class U>
    int dalloc(U** p, size_t count) {
        void* v = nullptr;
        HIPTRY(hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(U)));
        own.push_back(v);
        HIPTRY(hipMemset(v, 0, std::max<size_t>(count, 1) * sizeof(U)));
        *p = (U*)v;
        return TH_OK;
    }

    // Tables and graphs for V-cycle blocks of cw columns by r * 64 / cw rows
    // (r wavefronts, each cw columns of 64 / cw rows) and `vt`-thread vector
    // blocks.
    int configure(int r, int cw, int vt) {
        HIPTRY(hipStreamSynchronize(s));
        drop_graphs();
        free_tables();
        rows = r;
        cols = cw;
        vthreads = vt;
        const int br = r * kColsX / cw;  // rows of a block
        t_sweep.resize(maxk);
        t_restrict.resize(maxk);
        for (int i = 0; i < n; ++i) {
            Inst& d = hInst[i];
            const int sb = (int)((2L * d.g0.hx + kColsX - 1) / kColsX), sr = (d.g0.ny + kStencilRows - 1) / kStencilRows;
            // the consumers read this many partials: of 64 x 4 blocks with types
            const int pr = d.L[0].ntypes ? (d.g0.ny + kColsY - 1) / kColsY : sr;
            if (sb * pr != d.nblk_stencil)
                return fail(TH_ERR_ARG, "batched solver: level-0 blocks differ from the instance's partials");
            add(d.L[0].ntypes ? t_typed : t_stencil, i, sb, sr, 1);
            const int vb = (int)std::max(1L, std::min((long)kMaxVecBlocks, (d.n0 + vt - 1) / vt));
            add(t_vec, i, vb, 1, 1);
            d.nblk_vec = vb;
            for (int k = 0; k < d.tail_k0; ++k) {
                const Geo& g = d.L[k].g;
                const Geo& c = d.L[k + 1].g;
                add(t_sweep[k], i, (g.hx + cw - 1) / cw, (g.ny + br - 1) / br, 1);
                add(t_restrict[k], i, (c.nx + cw - 1) / cw, (c.ny + br - 1) / br,
                    d.L[k].ntypes ? (c.nz + kRestrictPlanes - 1) / kRestrictPlanes : c.nz);
            }
        }
        HIPTRY(hipMemcpy(dI, hInst.data(), n * sizeof(Inst), hipMemcpyHostToDevice));
        int rc;
        if ((rc = upload_table(t_stencil)) || (rc = upload_table(t_typed)) || (rc = upload_table(t_vec))) return rc;
        for (int k = 0; k < maxk; ++k)
            if ((rc = upload_table(t_sweep[k])) || (rc = upload_table(t_restrict[k]))) return rc;
        if ((rc = capture(&g_start, true)) || (rc = capture(&g_iter, false))) return rc;
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
        if (e != hipSuccess) return fail(TH_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(exec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e != hipSuccess) return fail(TH_ERR_HIP, std::string("batch graph instantiate: ") + hipGetErrorString(e));
        return TH_OK;
    }

    // ---- launch sequences ----------------------------------------------------
    template <int MODE>
    int sweep(int k, int colour, int cons) {
        const Table& t = t_sweep[k];
        if (!t.blocks) return TH_OK;
        const dim3 blk(cols, rows * kColsX / cols);
#define TB_SW(NZ) kb_sweep<NZ, MODE><<<t.blocks, blk, 0, s>>>(dI, t.d, (int)t.h.size(), k, colour, cons)
        switch (bucket(nz)) {
            case 4: TB_SW(4); break;
            case 8: TB_SW(8); break;
            case 12: TB_SW(12); break;
            default: TB_SW(16); break;
        }
#undef TB_SW
        HIPTRY(hipGetLastError());
        return TH_OK;
    }
    int restrict_(int k) {
        const Table& t = t_restrict[k];
        if (!t.blocks) return TH_OK;
        kb_restrict<<<t.blocks, dim3(cols, rows * kColsX / cols), 0, s>>>(dI, t.d, (int)t.h.size(), k);
        HIPTRY(hipGetLastError());
        return TH_OK;
    }
    int tail(int cons) {
#define TB_TL(NZ) kb_tail<NZ><<<n, kTailLdsThreads, lds, s>>>(dI, dTail, cons)
        switch (bucket(nz)) {
            case 4: TB_TL(4); break;
            case 8: TB_TL(8); break;
            case 12: TB_TL(12); break;
            default: TB_TL(16); break;
        }
#undef TB_TL
        HIPTRY(hipGetLastError());
        return TH_OK;
    }
    // V-cycle of every instance: descend all levels above the tails, tails,
    // climb. The level-0 red sweep (or a level-0 tail) carries `cons`.
    int vcycle(int cons) {
        int rc;
        for (int k = 0; k < maxk; ++k)
            if ((rc = sweep<SW_ZERO>(k, 0, cons)) || (rc = sweep<SW_PLAIN>(k, 1, CONS_NONE)) || (rc = restrict_(k)))
                return rc;
        if ((rc = tail(cons))) return rc;
        for (int k = maxk - 1; k >= 0; --k)
            if ((rc = sweep<SW_PROLONG>(k, 1, CONS_NONE)) || (rc = sweep<SW_PLAIN>(k, 0, CONS_NONE))) return rc;
        return TH_OK;
    }
    int vec(void (*kern)(const Inst*, const Entry*, int)) {
        kern<<<t_vec.blocks, vthreads, 0, s>>>(dI, t_vec.d, (int)t_vec.h.size());
        HIPTRY(hipGetLastError());
        return TH_OK;
    }
    int enqueue_start() {
        if (t_stencil.blocks)
            kb_residual<<<t_stencil.blocks, dim3(kColsX, kStencilRows), 0, s>>>(dI, t_stencil.d, (int)t_stencil.h.size());
        if (t_typed.blocks) {
#define TB_RS(NZ) kb_residual_typed<NZ><<<t_typed.blocks, dim3(kColsX, kStencilRows), 0, s>>>(dI, t_typed.d, (int)t_typed.h.size())
            switch (bucket(nz)) {
                case 4: TB_RS(4); break;
                case 8: TB_RS(8); break;
                case 12: TB_RS(12); break;
                default: TB_RS(16); break;
            }
#undef TB_RS
        }
        HIPTRY(hipGetLastError());
        int rc;
        if ((rc = vcycle(CONS_START)) || (rc = vec(kb_dot)) || (rc = vec(kb_pupdate<1>))) return rc;
        return TH_OK;
    }
    int enqueue_iter() {
#define TB_AP(K, TB, ROWS, NZ) K<NZ><<<TB.blocks, dim3(kColsX, ROWS), 0, s>>>(dI, TB.d, (int)TB.h.size())
#define TB_APZ(K, TB, ROWS)                      \
    switch (bucket(nz)) {                        \
        case 4: TB_AP(K, TB, ROWS, 4); break;   \
        case 8: TB_AP(K, TB, ROWS, 8); break;   \
        case 12: TB_AP(K, TB, ROWS, 12); break; \
        default: TB_AP(K, TB, ROWS, 16); break; \
    }
        if (t_stencil.blocks) TB_APZ(kb_apply, t_stencil, kStencilRows);
        if (t_typed.blocks) TB_APZ(kb_apply_typed, t_typed, kStencilRows);
#undef TB_APZ
#undef TB_AP
        HIPTRY(hipGetLastError());
        int rc;
        if ((rc = vec(kb_update)) || (rc = vcycle(CONS_ITER)) || (rc = vec(kb_dot)) || (rc = vec(kb_pupdate<0>)))
            return rc;
        return TH_OK;
    }

    int fetch() {
        kb_gather<<<(n + 63) / 64, 64, 0, s>>>(dI, n, dSt);
        HIPTRY(hipGetLastError());
        HIPTRY(hipMemcpyAsync(hSt, dSt, n * sizeof(State), hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        return TH_OK;
    }

    // Steady solves of every instance from T = 0.
    int solve(double tol, int maxit, int* iters, double* relres, int* conv, double* ms) {
        using clk = std::chrono::high_resolution_clock;
        HIPTRY(hipStreamSynchronize(s));
        const auto t0 = clk::now();
        kb_begin<<<(n + 63) / 64, 64, 0, s>>>(dI, n, tol, maxit, 1);
        HIPTRY(hipGetLastError());
        int rc;
        if ((rc = vec(kb_reset))) return rc;
        HIPTRY(hipGraphLaunch(g_start, s));
        if ((rc = fetch())) return rc;
        double rate = 0.0;
        int it = 0;
        auto all_done = [&]() {
            for (int i = 0; i < n; ++i)
                if (!hSt[i].done) return false;
            return true;
        };
        auto worst = [&]() {  // largest relres / tol among unfinished instances
            double w = 0.0;
            for (int i = 0; i < n; ++i)
                if (!hSt[i].done && hSt[i].relres > 0.0) w = std::max(w, hSt[i].relres);
            return w;
        };
        while (!all_done() && it < maxit) {
            int k = 1;
            const double w0 = worst();
            if (rate > 0.0 && rate < 1.0 && w0 > 0.0) {
                const double need = std::log(tol / w0) / std::log(rate);
                k = (int)std::max(1.0, std::min(16.0, std::ceil(need) - 1.0));
            }
            k = std::min(k, maxit - it);
            for (int q = 0; q < k; ++q) HIPTRY(hipGraphLaunch(g_iter, s));
            it += k;
            if ((rc = fetch())) return rc;
            const double w1 = worst();
            if (w0 > 0.0 && w1 > 0.0) rate = std::pow(w1 / w0, 1.0 / k);
        }
        if ((rc = vec(kb_finish))) return rc;
        HIPTRY(hipStreamSynchronize(s));
        if (ms) *ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        for (int i = 0; i < n; ++i) {
            if (iters) iters[i] = hSt[i].iters;
            if (relres) relres[i] = hSt[i].relres;
            if (conv) conv[i] = hSt[i].converged;
        }
        return TH_OK;
    }

    // Times the V-cycle for every block shape (1, 2, 4 or 8 wavefronts of
    // 64, 32 or 16 columns of a row each) and the vector stages (256- or
    // 1024-thread blocks) on the batch, and keeps the fastest. The solution
    // vectors are left arbitrary (each solve starts from T = 0).
    int autotune(int reps, int* best_rows, int* best_cols, int* best_vt, double* best_ms) {
        hipEvent_t e0, e1;
        HIPTRY(hipEventCreate(&e0));
        HIPTRY(hipEventCreate(&e1));
        auto timed = [&](auto&& fn, double* out) -> int {
            kb_begin<<<(n + 63) / 64, 64, 0, s>>>(dI, n, 0.0, 1 << 30, 0);
            HIPTRY(hipGetLastError());
            int rc = fn();
            if (rc) return rc;
            HIPTRY(hipEventRecord(e0, s));
            for (int q = 0; q < reps && !rc; ++q) rc = fn();
            if (rc) return rc;
            HIPTRY(hipEventRecord(e1, s));
            HIPTRY(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPTRY(hipEventElapsedTime(&ms, e0, e1));
            *out = ms / reps;
            return TH_OK;
        };
        int rc = TH_OK, br = 4, bc = 64, bv = 1024;
        double bt = 1e30, vbest = 1e30;
        for (int cw : {64, 32, 16})
            for (int r : {1, 2, 4, 8}) {
                if (rc) break;
                if ((rc = configure(r, cw, 1024))) break;
                double t = 0;
                if ((rc = timed([&]() { return vcycle(CONS_NONE); }, &t))) break;
                if (t < bt) bt = t, br = r, bc = cw;
            }
        for (int vt : {256, 1024}) {
            if (rc) break;
            if ((rc = configure(br, bc, vt))) break;
            double t = 0;
            if ((rc = timed([&]() {
                     int q;
                     if ((q = vec(kb_dot)) || (q = vec(kb_reset))) return q;
                     return (int)TH_OK;
                 }, &t)))
                break;
            if (t < vbest) vbest = t, bv = vt;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return rc;
        if ((rc = configure(br, bc, bv))) return rc;
        if (best_rows) *best_rows = br;
        if (best_cols) *best_cols = bc;
        if (best_vt) *best_vt = bv;
        if (best_ms) *best_ms = bt;
        return TH_OK;
    }
};

}  // namespace bat
}  // namespace th

using namespace th;
using namespace th::bat;

struct th_batch {
    Batch b;
};

extern "C" {

int th_batch_create(int count, const th_model_desc* const* models, const th_solver_opts* o, th_batch** out) {
    clear_error();
    if (count < 1 || !models || !out) return fail(TH_ERR_ARG, "bad argument");
    th_solver_opts opt;
    if (o) opt = *o;
    else th_solver_default_opts(&opt);
    if (opt.precond != TH_PRECOND_MG || opt.precision != TH_PREC_MIXED || opt.restriction != TH_RESTRICT_LATERAL)
        return fail(TH_ERR_ARG, "batched solver: multigrid, mixed precision and lateral restriction only");
    for (int i = 0; i < count; ++i) {
        const th_model_desc* m = models[i];
        if (!m || m->nx < 1 || m->ny < 1 || m->nz < 1 || m->nz > 16)
            return fail(TH_ERR_ARG, "batched solver: model " + std::to_string(i) + " invalid (1 <= nz <= 16)");
        if (!m->gE || !m->gN || !m->gU || !m->gB || !m->cap || !m->P)
            return fail(TH_ERR_NOT_FINALIZED, "Model3D used before finalize()");
    }
    th_batch* b = new th_batch;
    int rc = b->b.init(count, models, opt);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return TH_OK;
}

void th_batch_destroy(th_batch* b) { delete b; }

int th_batch_count(th_batch* b) { return b ? b->b.n : -1; }

int th_batch_set_power(th_batch* b, int i, const double* P) {
    if (!b || !P || i < 0 || i >= b->b.n) return fail(TH_ERR_ARG, "bad argument");
    clear_error();
    HIPTRY(hipStreamSynchronize(b->b.s));
    return b->b.S[i]->set_power(P);
}

int th_batch_solve(th_batch* b, double tol, int maxit, int* iters, double* relres, int* converged, double* ms) {
    if (!b || maxit < 0) return fail(TH_ERR_ARG, "bad argument");
    clear_error();
    return b->b.solve(tol, maxit, iters, relres, converged, ms);
}

int th_batch_get_x(th_batch* b, int i, double* x) {
    if (!b || !x || i < 0 || i >= b->b.n) return fail(TH_ERR_ARG, "bad argument");
    clear_error();
    HIPTRY(hipStreamSynchronize(b->b.s));
    return b->b.S[i]->get_x(x);
}

int th_batch_autotune(th_batch* b, int reps, int* rows, int* cols, int* vec_threads, double* vcycle_ms) {
    if (!b || reps < 1) return fail(TH_ERR_ARG, "bad argument");
    clear_error();
    return b->b.autotune(reps, rows, cols, vec_threads, vcycle_ms);
}

int th_batch_shape(th_batch* b, int* rows, int* cols, int* vec_threads) {
    if (!b) return fail(TH_ERR_ARG, "null batch");
    if (rows) *rows = b->b.rows;
    if (cols) *cols = b->b.cols;
    if (vec_threads) *vec_threads = b->b.vthreads;
    return TH_OK;
}

}  // extern "C"
