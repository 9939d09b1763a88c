// surrogate.hip - the learned warm start of the solver: a small U-Net that
// maps the power density of the two dies and the TSV copper fraction on a
// 256 x 256 grid to the temperature rise of the 10 planes of the two-tier
// stack. Forward, backward and Adam kernels for gfx950 (BF16 MFMA with FP32
// accumulation), the training loop in C++, and the C ABI
// (include/thermal_hip.h). The reference has no learned part; the network
// is the one the project's benchmark description specifies (4 down/up
// levels, 32-64 channels, ~230k parameters).
//
// Tensors are NHWC with 16, 32 or 64 channels, pixels flattened over
// (image, row, column). A 3x3 convolution is an implicit GEMM over
// K = 9 taps x Cin: each wavefront owns 16 consecutive pixels and all output
// channels, loads the 16 x K input slice of one tap straight from global
// memory (neighbouring taps and waves hit the same cache lines) and issues
// one MFMA per 16 output channels: v_mfma_f32_16x16x32_bf16 when Cin is a
// multiple of 32, v_mfma_f32_16x16x16_bf16 for the 16-channel input layer.
// The data gradient is the same kernel on flipped, transposed weights; the
// weight gradient is an MFMA GEMM over pixels, one wave per (pixel chunk,
// tap), with the bias gradient as an extra tap whose input is 1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "model.h"


namespace th {
namespace sur {

using bf = unsigned short;  // bfloat16 storage
typedef short s4v __attribute__((ext_vector_type(4)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
typedef unsigned int u2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float bf2f(bf h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ bf f2bf(float f) {  // nearest even
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf)(u >> 16);
}
static bf host_f2bf(float f) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf)(u >> 16);
}

// (level, cin, cout, kernel size, relu); the input layer's 16 channels carry
// 3 features, the head's 16 outputs carry 10 planes.
struct LayerSpec {
    int level, cin, cout, ks, relu;
};
constexpr int kLayers = 12;
static const LayerSpec kSpec[kLayers] = {
    {0, 16, 32, 3, 1}, {0, 32, 32, 3, 1}, {1, 32, 32, 3, 1}, {2, 32, 64, 3, 1},
    {2, 64, 64, 3, 1}, {3, 64, 64, 3, 1}, {4, 64, 64, 3, 1}, {3, 64, 64, 3, 1},
    {2, 64, 32, 3, 1}, {1, 32, 32, 3, 1}, {0, 32, 32, 3, 1}, {0, 32, 16, 1, 0}};
enum { E0A, E0B, E1A, E2A, E2B, E3A, B0, D3, D2, D1, D0, HEAD };
constexpr int kIn = 3, kOut = 10, kInStore = 4;

static long layer_params(const LayerSpec& l) { return (long)l.cout * l.ks * l.ks * l.cin + l.cout; }
static long total_params() {
    long n = 0;
    for (const LayerSpec& l : kSpec) n += layer_params(l);
    return n;
}

// ---- kernels -----------------------------------------------------------------

enum { EPI_RELU = 0, EPI_LINEAR = 1, EPI_F32 = 2 };

// Y = epi(conv(X, W) + b); Wf[tap][cout][cin]. Each wavefront computes MT
// tiles of 16 consecutive pixels for every output channel: per tap and
// K-slice it loads the MT input fragments and the COUT / 16 weight
// fragments once and issues MT x COUT / 16 MFMAs, so every weight fragment
// feeds MT of them (with MT = 1 a weight fragment read through L1 fed one
// MFMA, and the kernel was bound by those reads, not by the matrix cores).
// On the small levels a wave takes TW of the COUT / 16 output tiles (waves
// (pixel tile, output group)): with every tile in one wave, its 9 taps'
// weight loads ran one after another (~14 us per 3x3 64-channel layer at
// 16 x 16 pixels whatever the size). Each output's MFMA chain runs over
// (tap, K-slice) in the same order for any MT and TW, so the results do not
// depend on them.
template <int KS, int CIN, int COUT, int EPI, int MT = 4, int TW = COUT / 16>
__global__ __launch_bounds__(256) void k_conv(const bf* __restrict__ X, const bf* __restrict__ Wf,
                                              const float* __restrict__ bias, void* __restrict__ Yv, long npix,
                                              int H, int W) {
    constexpr int KC = (CIN % 32 == 0) ? 32 : 16;  // K of one MFMA
    constexpr int NT = TW, NG = COUT / 16 / TW;    // tiles per wave, output groups
    static_assert(NG * TW * 16 == COUT, "output groups");
    const int lane = threadIdx.x & 63;
    const long wv = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long p0 = (wv / NG) * (16 * MT);
    const int t0 = (int)(wv % NG) * TW;  // first output tile of the wave
    if (p0 >= npix) return;
    const int m = lane & 15, q = lane >> 4;
    const long hw = (long)H * W;
    long img[MT];
    int py[MT], px[MT];
    bool valid[MT];
#pragma unroll
    for (int u = 0; u < MT; ++u) {
        const long pm = p0 + 16 * u + m;
        valid[u] = pm < npix;
        img[u] = valid[u] ? pm / hw : 0;
        const int rem = valid[u] ? (int)(pm - img[u] * hw) : 0;
        py[u] = rem / W;
        px[u] = rem - py[u] * W;
    }
    f4v acc[MT][NT];
#pragma unroll
    for (int u = 0; u < MT; ++u)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[u][t] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int tap = 0; tap < KS * KS; ++tap) {
        const bf* xp[MT];
        bool inb[MT];
#pragma unroll
        for (int u = 0; u < MT; ++u) {
            const int yy = py[u] + tap / KS - KS / 2, xx = px[u] + tap % KS - KS / 2;
            inb[u] = valid[u] && yy >= 0 && yy < H && xx >= 0 && xx < W;
            xp[u] = X + (img[u] * hw + (long)(inb[u] ? yy : 0) * W + (inb[u] ? xx : 0)) * CIN;
        }
        const bf* wp = Wf + ((long)tap * COUT + 16 * t0 + m) * CIN;
#pragma unroll
        for (int c0 = 0; c0 < CIN; c0 += KC) {
            if constexpr (KC == 32) {
                b8v a[MT];
#pragma unroll
                for (int u = 0; u < MT; ++u) {
                    const u4v av = inb[u] ? *(const u4v*)(xp[u] + c0 + 8 * q) : u4v{0u, 0u, 0u, 0u};
                    a[u] = __builtin_bit_cast(b8v, av);
                }
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const b8v bw = __builtin_bit_cast(b8v, *(const u4v*)(wp + (long)t * 16 * CIN + c0 + 8 * q));
#pragma unroll
                    for (int u = 0; u < MT; ++u)
                        acc[u][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw, a[u], acc[u][t], 0, 0, 0);
                }
            } else {
                s4v a[MT];
#pragma unroll
                for (int u = 0; u < MT; ++u) {
                    const u2v av = inb[u] ? *(const u2v*)(xp[u] + c0 + 4 * q) : u2v{0u, 0u};
                    a[u] = __builtin_bit_cast(s4v, av);
                }
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const s4v bw = __builtin_bit_cast(s4v, *(const u2v*)(wp + (long)t * 16 * CIN + c0 + 4 * q));
#pragma unroll
                    for (int u = 0; u < MT; ++u)
                        acc[u][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bw, a[u], acc[u][t], 0, 0, 0);
                }
            }
        }
    }
    // The weights are the MFMA's A operand (rows: output channels) and the
    // input its B operand (columns: pixels), so lane (m, q) holds output
    // channels 4q .. 4q + 3 of tile t for pixel m: one 8-byte (BF16) or
    // 16-byte (FP32) store per tile instead of four 2-byte ones.
#pragma unroll
    for (int u = 0; u < MT; ++u) {
        const long pr = p0 + 16 * u + m;
        if (pr >= npix) continue;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int co = (t0 + t) * 16 + 4 * q;
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                v[i] = acc[u][t][i] + (EPI == EPI_LINEAR ? 0.f : bias[co + i]);
                if (EPI == EPI_RELU) v[i] = fmaxf(v[i], 0.f);
            }
            if (EPI == EPI_F32) {
                *(f4v*)((float*)Yv + pr * COUT + co) = f4v{v[0], v[1], v[2], v[3]};
            } else {
                const u2v w2{(unsigned)f2bf(v[0]) | ((unsigned)f2bf(v[1]) << 16),
                             (unsigned)f2bf(v[2]) | ((unsigned)f2bf(v[3]) << 16)};
                *(u2v*)((bf*)Yv + pr * COUT + co) = w2;
            }
        }
    }
}

// k_conv for 3x3 layers of 32 input and 32 output channels on images whose
// width is a multiple of 64 and height a multiple of 4 (the two top levels):
// the image is cut into tiles of 4 rows x 64 columns, which the blocks walk.
// Each wave holds its 9 taps x 2 output tiles of weight fragments in
// registers for all its block's tiles; per tile the block stages the 6 x 66
// input pixels with their halo in LDS (zeros outside the image, 16-byte
// loads along the pixel-contiguous rows) and every MFMA's input fragment is
// one 16-byte LDS read (a wave's 64 lanes read 1 KB of consecutive bytes:
// no bank conflicts) instead of a load through L1. Wave w computes row w of
// the tile, 4 x 16 pixels x 32 output channels, with the same accumulation
// order per output tile as k_conv (tap by tap, one 32-channel slice), so the
// outputs are k_conv's bit for bit.
template <int EPI>
__global__ __launch_bounds__(256) void k_conv_lds(const bf* __restrict__ X, const bf* __restrict__ Wf,
                                                  const float* __restrict__ bias, void* __restrict__ Yv, int H, int W,
                                                  int B) {
    constexpr int CIN = 32, COUT = 32, TR = 4, TC = 64, SR = TR + 2, SC = TC + 2;
    __shared__ __align__(16) bf sx[SR * SC * CIN];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, m = lane & 15, q = lane >> 4;
    const int tiles_x = W / TC, tiles_y = H / TR, ntiles = B * tiles_x * tiles_y;
    // this lane's weight fragments: tap t, output tile ot, channels 8q..8q+7
    b8v wr[9][2];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int ot = 0; ot < 2; ++ot)
            wr[t][ot] = __builtin_bit_cast(b8v, *(const u4v*)(Wf + ((long)t * COUT + 16 * ot + m) * CIN + 8 * q));
    float bv[2][4];
#pragma unroll
    for (int ot = 0; ot < 2; ++ot)
#pragma unroll
        for (int i = 0; i < 4; ++i) bv[ot][i] = EPI == EPI_LINEAR ? 0.f : bias[ot * 16 + 4 * q + i];
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int bx = tile % tiles_x, by = (tile / tiles_x) % tiles_y, img = tile / (tiles_x * tiles_y);
        const int x0 = bx * TC, y0 = by * TR;
        const long ibase = (long)img * H * W;
        __syncthreads();  // the previous tile's reads of sx are done
        for (int e = tid; e < SR * SC * (CIN / 8); e += 256) {
            const int pix = e >> 2, ch = e & 3;
            const int py = pix / SC, px = pix - py * SC;
            const int gy = y0 + py - 1, gx = x0 + px - 1;
            u4v v{0u, 0u, 0u, 0u};
            if (gy >= 0 && gy < H && gx >= 0 && gx < W) v = *(const u4v*)(X + (ibase + (long)gy * W + gx) * CIN + 8 * ch);
            *(u4v*)(sx + pix * CIN + 8 * ch) = v;
        }
        __syncthreads();
        f4v acc[4][2];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) acc[u][ot] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int dy = t / 3, dx = t % 3;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const b8v a = *(const b8v*)(sx + ((wave + dy) * SC + 16 * u + m + dx) * CIN + 8 * q);
#pragma unroll
                for (int ot = 0; ot < 2; ++ot)
                    acc[u][ot] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[t][ot], a, acc[u][ot], 0, 0, 0);
            }
        }
        // epilogue of k_conv: lane (m, q) holds output channels 4q..4q+3 of tile ot for pixel m
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long pr = ibase + (long)(y0 + wave) * W + x0 + 16 * u + m;
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) {
                const int co = ot * 16 + 4 * q;
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    v[i] = acc[u][ot][i] + bv[ot][i];
                    if (EPI == EPI_RELU) v[i] = fmaxf(v[i], 0.f);
                }
                const u2v w2{(unsigned)f2bf(v[0]) | ((unsigned)f2bf(v[1]) << 16),
                             (unsigned)f2bf(v[2]) | ((unsigned)f2bf(v[3]) << 16)};
                *(u2v*)((bf*)Yv + pr * COUT + co) = w2;
            }
        }
    }
}

// Partial weight gradients of one pixel chunk: part[chunk][tap][cout][cin]
// = sum_p GZ[p][co] X[p + off(tap)][ci] (zero outside the image); tap ==
// KS*KS is the bias, sum_p GZ[p][co], in column ci = 0. A GEMM over pixels
// per tap (M = COUT, N = CIN, K = pixels) whose operands are pixel-major in
// memory, while an MFMA lane takes 8 consecutive K values of one row or
// column: the block stages each K-block of 32 pixels transposed in LDS,
// GZ^T [COUT][32] once and X^T [CIN][32] per tap (the tap's shifted pixels),
// so that each operand fragment is one 16-byte LDS read, and issues
// v_mfma_f32_16x16x32_bf16 on them; wave w < KS*KS accumulates tap w's
// COUT x CIN tile set, the last wave the bias (B = 1 in column 0). Staging
// lanes run along the 32 pixels of a K-block, so a row of the transposed
// tile is written by consecutive lanes. The chunk is a multiple of 32
// pixels.
template <int KS, int CIN, int COUT>
__global__ __launch_bounds__((KS * KS + 1) * 64) void k_wgrad(const bf* __restrict__ X, const bf* __restrict__ GZ,
                                                              float* __restrict__ part, int npix, int H, int W,
                                                              int chunk) {
    constexpr int KK = KS * KS, NWAVE = KK + 1, NT = COUT / 16, NU = CIN / 16, KB = 32;
    constexpr int RS = KB + 8;  // row stride in BF16 values: 16-byte aligned rows
    constexpr int NTH = NWAVE * 64;
    __shared__ __align__(16) bf sg[COUT * RS];
    __shared__ __align__(16) bf sx[KK * CIN * RS];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, m = lane & 15, q = lane >> 4;
    const int pb = blockIdx.x * chunk, pe = pb + chunk < npix ? pb + chunk : npix;
    const int hw = H * W;
    f4v acc[NT][NU];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int u = 0; u < NU; ++u) acc[t][u] = f4v{0.f, 0.f, 0.f, 0.f};
    for (int p0 = pb; p0 < pe; p0 += KB) {
        // GZ^T: element e -> pixel k = e % 32, channel group g = e / 32 (8 channels)
        for (int e = tid; e < KB * (COUT / 8); e += NTH) {
            const int k = e % KB, g = e / KB, p = p0 + k;
            u4v v = p < pe ? *(const u4v*)(GZ + (long)p * COUT + 8 * g) : u4v{0u, 0u, 0u, 0u};
#pragma unroll
            for (int c = 0; c < 8; ++c) sg[(8 * g + c) * RS + k] = (bf)((v[c >> 1] >> (16 * (c & 1))) & 0xFFFFu);
        }
        // X^T of every tap: element e -> tap, pixel k, channel group g
        for (int e = tid; e < KK * KB * (CIN / 8); e += NTH) {
            const int k = e % KB, r = e / KB, g = r % (CIN / 8), tap = r / (CIN / 8), p = p0 + k;
            u4v v{0u, 0u, 0u, 0u};
            if (p < pe) {
                const int img = p / hw, rem = p - img * hw, y = rem / W, x = rem - y * W;
                const int yy = y + tap / KS - KS / 2, xx = x + tap % KS - KS / 2;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                    v = *(const u4v*)(X + ((long)img * hw + (long)yy * W + xx) * CIN + 8 * g);
            }
            bf* d = sx + (tap * CIN + 8 * g) * RS + k;
#pragma unroll
            for (int c = 0; c < 8; ++c) d[c * RS] = (bf)((v[c >> 1] >> (16 * (c & 1))) & 0xFFFFu);
        }
        __syncthreads();
        b8v a[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) a[t] = *(const b8v*)(sg + (16 * t + m) * RS + 8 * q);
        if (wave < KK) {
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                const b8v bx = *(const b8v*)(sx + (wave * CIN + 16 * u + m) * RS + 8 * q);
#pragma unroll
                for (int t = 0; t < NT; ++t) acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], bx, acc[t][u], 0, 0, 0);
            }
        } else {
            const unsigned one = m == 0 ? 0x3F803F80u : 0u;  // BF16 1.0 pairs in column 0
            const b8v ones = __builtin_bit_cast(b8v, u4v{one, one, one, one});
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], ones, acc[t][0], 0, 0, 0);
        }
        __syncthreads();
    }
    // lane (m, q) holds D[co = 16 t + 4 q + i][ci = 16 u + m]
    float* out = part + ((long)blockIdx.x * (KK + 1) + wave) * COUT * CIN;
    const int nu = wave < KK ? NU : 1;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int u = 0; u < NU; ++u)
            if (u < nu)
#pragma unroll
                for (int i = 0; i < 4; ++i) out[(16 * t + 4 * q + i) * CIN + 16 * u + m] = acc[t][u][i];
}

// Sums the chunks' partials into the canonical gradient layout
// gW[cout][tap][cin], gb[cout]. A block takes 64 gradient values; its 4
// waves each sum every 4th chunk (4 loads in flight per thread), then the
// four partial sums are added in a fixed order: the result does not depend
// on the launch.
__global__ __launch_bounds__(256) void k_reduce_grad(const float* __restrict__ part, int nchunks, int KK, int COUT,
                                                     int CIN, float* __restrict__ gW, float* __restrict__ gb) {
    __shared__ float sh[4][64];
    const int n = COUT * KK * CIN + COUT;
    const int e = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + e;
    long src = 0;
    if (k < n) {
        if (k < COUT * KK * CIN) {
            const int co = k / (KK * CIN), tap = (k / CIN) % KK, ci = k % CIN;
            src = ((long)tap * COUT + co) * CIN + ci;
        } else {
            src = ((long)KK * COUT + (k - COUT * KK * CIN)) * CIN;
        }
    }
    const long stride = (long)(KK + 1) * COUT * CIN;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    if (k < n) {
        int c = w;
        for (; c + 12 < nchunks; c += 16) {
            s0 += part[c * stride + src];
            s1 += part[(c + 4) * stride + src];
            s2 += part[(c + 8) * stride + src];
            s3 += part[(c + 12) * stride + src];
        }
        for (; c < nchunks; c += 4) s0 += part[c * stride + src];
    }
    sh[w][e] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (w == 0 && k < n) {
        const float s = ((sh[0][e] + sh[1][e]) + sh[2][e]) + sh[3][e];
        if (k < COUT * KK * CIN) gW[k] = s;
        else gb[k - COUT * KK * CIN] = s;
    }
}

// Elementwise kernels over NHWC tensors, 8 channels (16 bytes) per thread:
// C is a multiple of 8, so a thread's 8 values share one pixel. Each value
// is computed as the one-value-per-thread form would (the same float
// operations and rounding).
typedef unsigned short bf8 __attribute__((ext_vector_type(8)));

__global__ void k_relu_bwd(bf* __restrict__ g, const bf* __restrict__ y, long n) {
    const long k = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 8;
    if (k >= n) return;
    bf8 gv = *(const bf8*)(g + k);
    const bf8 yv = *(const bf8*)(y + k);
#pragma unroll
    for (int e = 0; e < 8; ++e)
        if (!(bf2f(yv[e]) > 0.f)) gv[e] = 0;
    *(bf8*)(g + k) = gv;
}

// Y = average of the 2x2 children of X.
__global__ void k_pool(const bf* __restrict__ X, bf* __restrict__ Y, long nout, int Ho, int Wo, int C) {
    const long k = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 8;
    if (k >= nout) return;
    const int c = (int)(k % C);
    const long pix = k / C;
    const int xo = (int)(pix % Wo);
    const long r = pix / Wo;
    const int yo = (int)(r % Ho);
    const long img = r / Ho;
    const int W = 2 * Wo;
    const long base = ((img * 2 * Ho + 2 * yo) * (long)W + 2 * xo) * C + c;
    const bf8 a = *(const bf8*)(X + base), b = *(const bf8*)(X + base + C);
    const bf8 cc = *(const bf8*)(X + base + (long)W * C), d = *(const bf8*)(X + base + (long)W * C + C);
    bf8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) o[e] = f2bf(0.25f * (bf2f(a[e]) + bf2f(b[e]) + bf2f(cc[e]) + bf2f(d[e])));
    *(bf8*)(Y + k) = o;
}

// Y = up(Xc) + S (nearest, x2).
__global__ void k_upadd(const bf* __restrict__ Xc, const bf* __restrict__ S, bf* __restrict__ Y, long n, int H, int W,
                        int C) {
    const long k = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 8;
    if (k >= n) return;
    const int c = (int)(k % C);
    const long pix = k / C;
    const int x = (int)(pix % W);
    const long r = pix / W;
    const int y = (int)(r % H);
    const long img = r / H;
    const long ck = ((img * (H / 2) + y / 2) * (long)(W / 2) + x / 2) * C + c;
    const bf8 a = *(const bf8*)(Xc + ck), b = *(const bf8*)(S + k);
    bf8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) o[e] = f2bf(bf2f(a[e]) + bf2f(b[e]));
    *(bf8*)(Y + k) = o;
}

// Gc = sum over the 2x2 children of G.
__global__ void k_upbwd(const bf* __restrict__ G, bf* __restrict__ Gc, long nc, int Hc, int Wc, int C) {
    const long k = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 8;
    if (k >= nc) return;
    const int c = (int)(k % C);
    const long pix = k / C;
    const int xo = (int)(pix % Wc);
    const long r = pix / Wc;
    const int yo = (int)(r % Hc);
    const long img = r / Hc;
    const int W = 2 * Wc;
    const long base = ((img * 2 * Hc + 2 * yo) * (long)W + 2 * xo) * C + c;
    const bf8 a = *(const bf8*)(G + base), b = *(const bf8*)(G + base + C);
    const bf8 cc = *(const bf8*)(G + base + (long)W * C), d = *(const bf8*)(G + base + (long)W * C + C);
    bf8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) o[e] = f2bf(bf2f(a[e]) + bf2f(b[e]) + bf2f(cc[e]) + bf2f(d[e]));
    *(bf8*)(Gc + k) = o;
}

// G = 0.25 up(Gp) + Gs: gradient of a skip tensor that also fed a pool.
__global__ void k_poolbwd_add(const bf* __restrict__ Gp, const bf* __restrict__ Gs, bf* __restrict__ G, long n, int H,
                              int W, int C) {
    const long k = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 8;
    if (k >= n) return;
    const int c = (int)(k % C);
    const long pix = k / C;
    const int x = (int)(pix % W);
    const long r = pix / W;
    const int y = (int)(r % H);
    const long img = r / H;
    const long ck = ((img * (H / 2) + y / 2) * (long)(W / 2) + x / 2) * C + c;
    const bf8 a = *(const bf8*)(Gp + ck), b = *(const bf8*)(Gs + k);
    bf8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) o[e] = f2bf(0.25f * bf2f(a[e]) + bf2f(b[e]));
    *(bf8*)(G + k) = o;
}

// Network input of a batch from the stored features: 16 channels, 3 used.
__global__ void k_gather(const bf* __restrict__ ds, const int* __restrict__ idx, bf* __restrict__ X0, long n, long hw) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int c = (int)(k & 15);
    const long pix = k >> 4;
    const long img = pix / hw, p = pix - img * hw;
    X0[k] = c < kIn ? ds[((long)idx[img] * hw + p) * kInStore + c] : (bf)0;
}

// Network input from FP32 features [B][H][W][3]: 16 BF16 channels, 3 used.
__global__ void k_pack_input(const float* __restrict__ in, bf* __restrict__ X0, long n) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int c = (int)(k & 15);
    const long pix = k >> 4;
    X0[k] = c < kIn ? f2bf(in[pix * kIn + c]) : (bf)0;
}

// MSE over the 10 used outputs: gradient 2 (out - tgt) / count in BF16 and
// per-block partial sums of the squared error (a grid-stride loop, so the
// partials number the blocks, not the elements). idx == nullptr: targets
// are batch-contiguous.
constexpr int kMseBlocks = 1024;
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ out, const float* __restrict__ tgt,
                                             const int* __restrict__ idx, long hw, long npix, double inv_cnt,
                                             bf* __restrict__ g, double* __restrict__ part) {
    double e = 0.0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < npix * 16; k += stride) {
        const int c = (int)(k & 15);
        const long pix = k >> 4;
        float gv = 0.f;
        if (c < kOut) {
            const long img = pix / hw, p = pix - img * hw;
            const long t = ((idx ? (long)idx[img] : img) * hw + p) * kOut + c;
            const float d = out[k] - tgt[t];
            e += (double)d * d;
            gv = (float)(2.0 * d * inv_cnt);
        }
        g[k] = f2bf(gv);
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_sum_to(const double* __restrict__ part, int n, double scale,
                                                double* __restrict__ dst) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) s += part[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *dst = sh[0] * scale;
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    w[k] -= lr * (mk * c1) / (sqrtf(vk * c2) + eps);
}

// BF16 copies of one layer's weights: Wf[tap][cout][cin] for the forward
// pass, Wb[KK-1-tap][cin][cout] (flipped, transposed) for the data gradient.
__global__ void k_pack(const float* __restrict__ Wc, int KK, int COUT, int CIN, bf* __restrict__ Wf,
                       bf* __restrict__ Wb) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= COUT * KK * CIN) return;
    const int co = k / (KK * CIN), tap = (k / CIN) % KK, ci = k % CIN;
    const bf v = f2bf(Wc[k]);
    Wf[((long)tap * COUT + co) * CIN + ci] = v;
    Wb[((long)(KK - 1 - tap) * CIN + ci) * COUT + co] = v;
}

// ---- host ------------------------------------------------------------------

#define SURTRY(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail(TH_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

static unsigned grid1(long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {  // Box-Muller
        const double u1 = std::max(uniform(), 1e-300), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct Net {
    int H = 0, W = 0, Bcap = 0;
    int dev = 0;    // the device it was created on
    long cus = 256;  // its compute units
    hipStream_t s = nullptr;
    long nparam = 0;
    long off[kLayers];
    float *prm = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    bf *Wf[kLayers] = {}, *Wb[kLayers] = {};
    long adam_t = 0;
    // activations and their gradients (sized for Bcap)
    bf *x0 = nullptr, *a0, *s0, *p0, *s1, *p1, *a2, *s2, *p2, *s3, *p3, *bb, *u3, *c3, *u2, *c2, *u1, *c1, *u0, *c0;
    float* out = nullptr;
    bf *g_out, *g_c0, *g_u0, *g_c1, *g_u1, *g_c2, *g_u2, *g_c3, *g_u3, *g_bb, *g_p3, *g_s3, *g_p2, *g_s2, *g_a2,
        *g_p1, *g_s1, *g_p0, *g_s0, *g_a0;
    float* wpart = nullptr;
    long wpart_n = 0;
    double* lpart = nullptr;
    double* dloss = nullptr;  // per-step losses (device)
    int dloss_cap = 0;
    float* tgt_b = nullptr;  // batch-contiguous targets (gradient checks)
    float* in_stage = nullptr;  // FP32 features of a batch (device)
    float* in_pin = nullptr;    // and their pinned host staging
    size_t in_pin_n = 0;
    int* d_idx = nullptr;
    // dataset
    bf* ds_in = nullptr;
    float* ds_tgt = nullptr;
    int ds_cap = 0, ds_n = 0;
    std::vector<void*> act_mem, own;

    long hw() const { return (long)H * W; }
    long npix(int level, int B) const { return (long)B * (H >> level) * (W >> level); }

    ~Net() {
        if (s) (void)hipStreamSynchronize(s);
        free_acts();
        for (void* p : own) (void)hipFree(p);
        if (in_pin) (void)hipHostFree(in_pin);
        if (s) (void)hipStreamDestroy(s);
    }
    void free_acts() {
        for (void* p : act_mem) (void)hipFree(p);
        act_mem.clear();
        Bcap = 0;
    }
    template <class U>
    int dalloc(U** p, size_t n, std::vector<void*>& list) {
        void* q = nullptr;
        SURTRY(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(U)));
        list.push_back(q);
        SURTRY(hipMemsetAsync(q, 0, std::max<size_t>(n, 1) * sizeof(U), s));
        *p = (U*)q;
        return TH_OK;
    }

    int init(int h, int w, uint64_t seed) {
        H = h;
        W = w;
        SURTRY(hipGetDevice(&dev));
        int ncu = 0;
        SURTRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        cus = ncu > 0 ? ncu : 256;
        SURTRY(hipStreamCreate(&s));
        nparam = 0;
        for (int k = 0; k < kLayers; ++k) {
            off[k] = nparam;
            nparam += layer_params(kSpec[k]);
        }
        int rc;
        for (float** p : {&prm, &grad, &m, &v})
            if ((rc = dalloc(p, nparam, own))) return rc;
        for (int k = 0; k < kLayers; ++k) {
            const long nw = (long)kSpec[k].cout * kSpec[k].ks * kSpec[k].ks * kSpec[k].cin;
            if ((rc = dalloc(&Wf[k], nw, own)) || (rc = dalloc(&Wb[k], nw, own))) return rc;
        }
        // wgrad partials: at most kMaxChunks chunks of (taps + 1) x cout x cin
        wpart_n = 0;
        for (const LayerSpec& l : kSpec)
            wpart_n = std::max(wpart_n, (long)kMaxChunks * (l.ks * l.ks + 1) * l.cout * l.cin);
        if ((rc = dalloc(&wpart, wpart_n, own))) return rc;
        // He-normal weights (fan-in counts the used input channels), zero bias
        std::vector<float> h0(nparam, 0.f);
        SplitMix rng{seed};
        for (int k = 0; k < kLayers; ++k) {
            const LayerSpec& l = kSpec[k];
            const int cin_used = (k == E0A) ? kIn : l.cin;
            const double sd = (k == HEAD) ? std::sqrt(1.0 / l.cin) : std::sqrt(2.0 / (l.ks * l.ks * cin_used));
            const int KK = l.ks * l.ks;
            for (int co = 0; co < l.cout; ++co)
                for (int t = 0; t < KK; ++t)
                    for (int ci = 0; ci < l.cin; ++ci)
                        h0[off[k] + ((long)co * KK + t) * l.cin + ci] =
                            (ci < cin_used && (k != HEAD || co < kOut)) ? (float)(sd * rng.normal()) : 0.f;
        }
        return set_params(h0.data());
    }
    static constexpr int kMaxChunks = 512;

    int set_params(const float* host) {
        SURTRY(hipMemcpyAsync(prm, host, nparam * sizeof(float), hipMemcpyHostToDevice, s));
        SURTRY(hipMemsetAsync(m, 0, nparam * sizeof(float), s));
        SURTRY(hipMemsetAsync(v, 0, nparam * sizeof(float), s));
        adam_t = 0;
        int rc = pack();
        if (rc) return rc;
        SURTRY(hipStreamSynchronize(s));
        return TH_OK;
    }
    int pack() {
        for (int k = 0; k < kLayers; ++k) {
            const LayerSpec& l = kSpec[k];
            const int KK = l.ks * l.ks;
            k_pack<<<grid1((long)l.cout * KK * l.cin), 256, 0, s>>>(prm + off[k], KK, l.cout, l.cin, Wf[k], Wb[k]);
            SURTRY(hipGetLastError());
        }
        return TH_OK;
    }

    int ensure_batch(int B) {
        if (B <= Bcap) return TH_OK;
        SURTRY(hipStreamSynchronize(s));
        free_acts();
        const long n0 = npix(0, B), n1 = npix(1, B), n2 = npix(2, B), n3 = npix(3, B), n4 = npix(4, B);
        struct A {
            bf** p;
            long n;
        };
        const A acts[] = {{&x0, n0 * 16}, {&a0, n0 * 32}, {&s0, n0 * 32}, {&p0, n1 * 32}, {&s1, n1 * 32},
                          {&p1, n2 * 32}, {&a2, n2 * 64}, {&s2, n2 * 64}, {&p2, n3 * 64}, {&s3, n3 * 64},
                          {&p3, n4 * 64}, {&bb, n4 * 64}, {&u3, n3 * 64}, {&c3, n3 * 64}, {&u2, n2 * 64},
                          {&c2, n2 * 32}, {&u1, n1 * 32}, {&c1, n1 * 32}, {&u0, n0 * 32}, {&c0, n0 * 32},
                          {&g_out, n0 * 16}, {&g_c0, n0 * 32}, {&g_u0, n0 * 32}, {&g_c1, n1 * 32}, {&g_u1, n1 * 32},
                          {&g_c2, n2 * 32}, {&g_u2, n2 * 64}, {&g_c3, n3 * 64}, {&g_u3, n3 * 64}, {&g_bb, n4 * 64},
                          {&g_p3, n4 * 64}, {&g_s3, n3 * 64}, {&g_p2, n3 * 64}, {&g_s2, n2 * 64}, {&g_a2, n2 * 64},
                          {&g_p1, n2 * 32}, {&g_s1, n1 * 32}, {&g_p0, n1 * 32}, {&g_s0, n0 * 32}, {&g_a0, n0 * 32}};
        int rc;
        for (const A& a : acts)
            if ((rc = dalloc(a.p, a.n, act_mem))) return rc;
        if ((rc = dalloc(&out, n0 * 16, act_mem)) || (rc = dalloc(&tgt_b, n0 * kOut, act_mem)) ||
            (rc = dalloc(&lpart, grid1(n0 * 16), act_mem)) || (rc = dalloc(&d_idx, B, act_mem)) ||
            (rc = dalloc(&in_stage, n0 * kIn, act_mem)))
            return rc;
        Bcap = B;
        return TH_OK;
    }

    // ---- launches ---------------------------------------------------------
    // 64 pixels per wave where that still gives every SIMD (1024 on the chip)
    // four waves; 16 otherwise (a batch of one 256 x 256 image at the top
    // level would give each SIMD a single wave).
    static bool conv_wide(long npix) { return npix >= 64L * 4 * 1024; }
    // below this many pixels a wave takes one output tile (k_conv's TW = 1)
    static bool conv_narrow(long npix) { return npix < 65536; }
    // 32 -> 32 channel 3x3 layers on images of 4-row x 64-column tiles: the
    // LDS-staged k_conv_lds
    static bool conv_lds(int h, int w) { return h % 4 == 0 && w % 64 == 0; }
    // its blocks walk the 4 x 64 tiles, so that each block's weight
    // fragments serve several tiles: as many blocks as the CUs hold at once
    // (3 of 4 waves each at its registers). 256^2 x 32, batch 16: 71 us for
    // k_conv, 54 us for one block per tile, 41 us for 768 blocks (3 per CU),
    // 50 us for 1024 (the fourth block per CU waits for a slot).
    unsigned conv_lds_grid(int B, int h, int w) const {
        const long nt = (long)B * (h / 4) * (w / 64);
        return (unsigned)std::min(nt, 3L * cus);
    }
    int conv(int k, const bf* in, void* outp, int B, int epi) {
        const LayerSpec& l = kSpec[k];
        const long n = npix(l.level, B);
        const int h = H >> l.level, w = W >> l.level;
        const bool wide = conv_wide(n), narrow = conv_narrow(n);
        const unsigned g = grid1(n, wide ? 256 : 64);  // 4 waves of 64 or 16 pixels
        const unsigned gn = grid1(n * (l.cout / 16), 64);  // 4 waves of 16 pixels x 16 channels
        const float* bias = prm + off[k] + (long)l.cout * l.ks * l.ks * l.cin;
#define SUR_CONV(KS, CI, CO, EPI)                                                                 \
    do {                                                                                          \
        if (wide) k_conv<KS, CI, CO, EPI, 4><<<g, 256, 0, s>>>(in, Wf[k], bias, outp, n, h, w);    \
        else if (narrow) k_conv<KS, CI, CO, EPI, 1, 1><<<gn, 256, 0, s>>>(in, Wf[k], bias, outp, n, h, w); \
        else k_conv<KS, CI, CO, EPI, 1><<<g, 256, 0, s>>>(in, Wf[k], bias, outp, n, h, w);         \
    } while (0)
        if (epi == EPI_F32) {
            if (l.ks == 1 && l.cin == 32 && l.cout == 16) SUR_CONV(1, 32, 16, EPI_F32);
            else return fail(TH_ERR_ARG, "surrogate: no FP32-output conv for this shape");
        } else if (l.ks != 3) {
            return fail(TH_ERR_ARG, "surrogate: no ReLU conv for this shape");
        } else if (l.cin == 16 && l.cout == 32) SUR_CONV(3, 16, 32, EPI_RELU);
        else if (l.cin == 32 && l.cout == 32 && conv_lds(h, w)) {
            k_conv_lds<EPI_RELU><<<conv_lds_grid(B, h, w), 256, 0, s>>>(in, Wf[k], bias, outp, h, w, B);
        } else if (l.cin == 32 && l.cout == 32) SUR_CONV(3, 32, 32, EPI_RELU);
        else if (l.cin == 32 && l.cout == 64) SUR_CONV(3, 32, 64, EPI_RELU);
        else if (l.cin == 64 && l.cout == 64) SUR_CONV(3, 64, 64, EPI_RELU);
        else if (l.cin == 64 && l.cout == 32) SUR_CONV(3, 64, 32, EPI_RELU);
        else return fail(TH_ERR_ARG, "surrogate: no conv for this shape");
#undef SUR_CONV
        SURTRY(hipGetLastError());
        return TH_OK;
    }
    // dX = conv(GZ, flipped W^T): input channels = the layer's cout
    int dgrad(int k, const bf* gz, bf* gx, int B) {
        const LayerSpec& l = kSpec[k];
        const long n = npix(l.level, B);
        const int h = H >> l.level, w = W >> l.level;
        const bool wide = conv_wide(n), narrow = conv_narrow(n);
        const unsigned g = grid1(n, wide ? 256 : 64);
        const unsigned gn = grid1(n * (l.cin / 16), 64);  // the data gradient's outputs: cin channels
#define SUR_DG(KS, CI, CO)                                                                                       \
    do {                                                                                                         \
        if (wide) k_conv<KS, CI, CO, EPI_LINEAR, 4><<<g, 256, 0, s>>>(gz, Wb[k], nullptr, gx, n, h, w);          \
        else if (narrow) k_conv<KS, CI, CO, EPI_LINEAR, 1, 1><<<gn, 256, 0, s>>>(gz, Wb[k], nullptr, gx, n, h, w); \
        else k_conv<KS, CI, CO, EPI_LINEAR, 1><<<g, 256, 0, s>>>(gz, Wb[k], nullptr, gx, n, h, w);               \
    } while (0)
        if (l.ks == 1 && l.cout == 16 && l.cin == 32) SUR_DG(1, 16, 32);
        else if (l.ks == 3 && l.cout == 32 && l.cin == 32 && conv_lds(h, w))
            k_conv_lds<EPI_LINEAR><<<conv_lds_grid(B, h, w), 256, 0, s>>>(gz, Wb[k], nullptr, gx, h, w, B);
        else if (l.ks == 3 && l.cout == 32 && l.cin == 32) SUR_DG(3, 32, 32);
        else if (l.ks == 3 && l.cout == 64 && l.cin == 32) SUR_DG(3, 64, 32);
        else if (l.ks == 3 && l.cout == 64 && l.cin == 64) SUR_DG(3, 64, 64);
        else if (l.ks == 3 && l.cout == 32 && l.cin == 64) SUR_DG(3, 32, 64);
        else return fail(TH_ERR_ARG, "surrogate: no data-gradient conv for this shape");
#undef SUR_DG
        SURTRY(hipGetLastError());
        return TH_OK;
    }
    int wgrad(int k, const bf* xin, const bf* gz, int B) {
        const LayerSpec& l = kSpec[k];
        const long n = npix(l.level, B);
        const int h = H >> l.level, w = W >> l.level;
        const int KK = l.ks * l.ks;
        // chunks of a multiple of 32 pixels (the kernel's K-blocks), at most
        // kMaxChunks and at least 512 pixels each: small layers get enough
        // blocks, large ones no more partials than the reduction reads fast
        long chunk = (n + kMaxChunks - 1) / kMaxChunks;
        chunk = std::max(512L, (chunk + 31) / 32 * 32);
        const int nch = (int)((n + chunk - 1) / chunk);
        const int nthr = (KK + 1) * 64;
#define SUR_WG(KS, CI, CO) \
    k_wgrad<KS, CI, CO><<<nch, nthr, 0, s>>>(xin, gz, wpart, (int)n, h, w, (int)chunk)
        if (l.ks == 1 && l.cin == 32 && l.cout == 16) SUR_WG(1, 32, 16);
        else if (l.ks == 3 && l.cin == 16 && l.cout == 32) SUR_WG(3, 16, 32);
        else if (l.ks == 3 && l.cin == 32 && l.cout == 32) SUR_WG(3, 32, 32);
        else if (l.ks == 3 && l.cin == 32 && l.cout == 64) SUR_WG(3, 32, 64);
        else if (l.ks == 3 && l.cin == 64 && l.cout == 64) SUR_WG(3, 64, 64);
        else if (l.ks == 3 && l.cin == 64 && l.cout == 32) SUR_WG(3, 64, 32);
        else return fail(TH_ERR_ARG, "surrogate: no weight-gradient kernel for this shape");
#undef SUR_WG
        SURTRY(hipGetLastError());
        float* gW = grad + off[k];
        float* gb = gW + (long)l.cout * KK * l.cin;
        k_reduce_grad<<<grid1((long)l.cout * KK * l.cin + l.cout, 64), 256, 0, s>>>(wpart, nch, KK, l.cout, l.cin, gW, gb);
        SURTRY(hipGetLastError());
        return TH_OK;
    }
    int relu_bwd(bf* g, const bf* y, long n) {
        k_relu_bwd<<<grid1(n / 8), 256, 0, s>>>(g, y, n);
        SURTRY(hipGetLastError());
        return TH_OK;
    }
    int pool(const bf* x, bf* y, int lvl_out, int B, int C) {
        const long n = npix(lvl_out, B) * C;
        k_pool<<<grid1(n / 8), 256, 0, s>>>(x, y, n, H >> lvl_out, W >> lvl_out, C);
        SURTRY(hipGetLastError());
        return TH_OK;
    }
    int upadd(const bf* xc, const bf* sk, bf* y, int lvl, int B, int C) {
        const long n = npix(lvl, B) * C;
        k_upadd<<<grid1(n / 8), 256, 0, s>>>(xc, sk, y, n, H >> lvl, W >> lvl, C);
        SURTRY(hipGetLastError());
        return TH_OK;
    }
    int upbwd(const bf* g, bf* gc, int lvl_c, int B, int C) {
        const long n = npix(lvl_c, B) * C;
        k_upbwd<<<grid1(n / 8), 256, 0, s>>>(g, gc, n, H >> lvl_c, W >> lvl_c, C);
        SURTRY(hipGetLastError());
        return TH_OK;
    }
    int poolbwd_add(const bf* gp, const bf* gs, bf* g, int lvl, int B, int C) {
        const long n = npix(lvl, B) * C;
        k_poolbwd_add<<<grid1(n / 8), 256, 0, s>>>(gp, gs, g, n, H >> lvl, W >> lvl, C);
        SURTRY(hipGetLastError());
        return TH_OK;
    }

    int forward(int B) {
        int rc;
        if ((rc = conv(E0A, x0, a0, B, EPI_RELU)) || (rc = conv(E0B, a0, s0, B, EPI_RELU)) ||
            (rc = pool(s0, p0, 1, B, 32)) || (rc = conv(E1A, p0, s1, B, EPI_RELU)) || (rc = pool(s1, p1, 2, B, 32)) ||
            (rc = conv(E2A, p1, a2, B, EPI_RELU)) || (rc = conv(E2B, a2, s2, B, EPI_RELU)) ||
            (rc = pool(s2, p2, 3, B, 64)) || (rc = conv(E3A, p2, s3, B, EPI_RELU)) || (rc = pool(s3, p3, 4, B, 64)) ||
            (rc = conv(B0, p3, bb, B, EPI_RELU)) || (rc = upadd(bb, s3, u3, 3, B, 64)) ||
            (rc = conv(D3, u3, c3, B, EPI_RELU)) || (rc = upadd(c3, s2, u2, 2, B, 64)) ||
            (rc = conv(D2, u2, c2, B, EPI_RELU)) || (rc = upadd(c2, s1, u1, 1, B, 32)) ||
            (rc = conv(D1, u1, c1, B, EPI_RELU)) || (rc = upadd(c1, s0, u0, 0, B, 32)) ||
            (rc = conv(D0, u0, c0, B, EPI_RELU)) || (rc = conv(HEAD, c0, out, B, EPI_F32)))
            return rc;
        return TH_OK;
    }

    int backward(int B) {
        const long n0 = npix(0, B), n1 = npix(1, B), n2 = npix(2, B), n3 = npix(3, B), n4 = npix(4, B);
        int rc;
        // head (linear)
        if ((rc = wgrad(HEAD, c0, g_out, B)) || (rc = dgrad(HEAD, g_out, g_c0, B))) return rc;
        // d0 .. d3: the input of each is up(prev) + skip
        if ((rc = relu_bwd(g_c0, c0, n0 * 32)) || (rc = wgrad(D0, u0, g_c0, B)) || (rc = dgrad(D0, g_c0, g_u0, B)) ||
            (rc = upbwd(g_u0, g_c1, 1, B, 32)))
            return rc;
        if ((rc = relu_bwd(g_c1, c1, n1 * 32)) || (rc = wgrad(D1, u1, g_c1, B)) || (rc = dgrad(D1, g_c1, g_u1, B)) ||
            (rc = upbwd(g_u1, g_c2, 2, B, 32)))
            return rc;
        if ((rc = relu_bwd(g_c2, c2, n2 * 32)) || (rc = wgrad(D2, u2, g_c2, B)) || (rc = dgrad(D2, g_c2, g_u2, B)) ||
            (rc = upbwd(g_u2, g_c3, 3, B, 64)))
            return rc;
        if ((rc = relu_bwd(g_c3, c3, n3 * 64)) || (rc = wgrad(D3, u3, g_c3, B)) || (rc = dgrad(D3, g_c3, g_u3, B)) ||
            (rc = upbwd(g_u3, g_bb, 4, B, 64)))
            return rc;
        // bottleneck and encoder; a skip tensor's gradient adds its pool's
        if ((rc = relu_bwd(g_bb, bb, n4 * 64)) || (rc = wgrad(B0, p3, g_bb, B)) || (rc = dgrad(B0, g_bb, g_p3, B)) ||
            (rc = poolbwd_add(g_p3, g_u3, g_s3, 3, B, 64)))
            return rc;
        if ((rc = relu_bwd(g_s3, s3, n3 * 64)) || (rc = wgrad(E3A, p2, g_s3, B)) || (rc = dgrad(E3A, g_s3, g_p2, B)) ||
            (rc = poolbwd_add(g_p2, g_u2, g_s2, 2, B, 64)))
            return rc;
        if ((rc = relu_bwd(g_s2, s2, n2 * 64)) || (rc = wgrad(E2B, a2, g_s2, B)) || (rc = dgrad(E2B, g_s2, g_a2, B)))
            return rc;
        if ((rc = relu_bwd(g_a2, a2, n2 * 64)) || (rc = wgrad(E2A, p1, g_a2, B)) || (rc = dgrad(E2A, g_a2, g_p1, B)) ||
            (rc = poolbwd_add(g_p1, g_u1, g_s1, 1, B, 32)))
            return rc;
        if ((rc = relu_bwd(g_s1, s1, n1 * 32)) || (rc = wgrad(E1A, p0, g_s1, B)) || (rc = dgrad(E1A, g_s1, g_p0, B)) ||
            (rc = poolbwd_add(g_p0, g_u0, g_s0, 0, B, 32)))
            return rc;
        if ((rc = relu_bwd(g_s0, s0, n0 * 32)) || (rc = wgrad(E0B, a0, g_s0, B)) || (rc = dgrad(E0B, g_s0, g_a0, B)))
            return rc;
        if ((rc = relu_bwd(g_a0, a0, n0 * 32)) || (rc = wgrad(E0A, x0, g_a0, B))) return rc;
        return TH_OK;
    }

    // loss of the current outputs into *dst (device), gradient into g_out
    int loss(int B, const float* tgt, const int* idx, double* dst) {
        const long n0 = npix(0, B);
        const unsigned g = std::min<unsigned>(grid1(n0 * 16), kMseBlocks);
        k_mse<<<g, 256, 0, s>>>(out, tgt, idx, hw(), n0, 1.0 / ((double)n0 * kOut), g_out, lpart);
        SURTRY(hipGetLastError());
        k_sum_to<<<1, 256, 0, s>>>(lpart, (int)g, 1.0 / ((double)n0 * kOut), dst);
        SURTRY(hipGetLastError());
        return TH_OK;
    }

    int adam(float lr) {
        ++adam_t;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)adam_t)));
        const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)adam_t)));
        k_adam<<<grid1(nparam), 256, 0, s>>>(prm, grad, m, v, nparam, lr, b1, b2, 1e-8f, c1, c2);
        SURTRY(hipGetLastError());
        return pack();
    }

    // host features [B][H][W][3] -> x0 (BF16, 16 channels)
    // host features [B][H][W][3] -> pinned staging -> device, packed to x0
    // on the device; returns with the staging buffer free again.
    int upload_input(const float* in, int B) {
        const size_t n = (size_t)npix(0, B) * kIn;
        if (n > in_pin_n) {
            SURTRY(hipStreamSynchronize(s));
            if (in_pin) (void)hipHostFree(in_pin);
            in_pin = nullptr;
            in_pin_n = 0;
            SURTRY(hipHostMalloc((void**)&in_pin, n * sizeof(float), hipHostMallocDefault));
            in_pin_n = n;
        }
        std::memcpy(in_pin, in, n * sizeof(float));
        SURTRY(hipMemcpyAsync(in_stage, in_pin, n * sizeof(float), hipMemcpyHostToDevice, s));
        k_pack_input<<<grid1(npix(0, B) * 16), 256, 0, s>>>(in_stage, x0, npix(0, B) * 16);
        SURTRY(hipGetLastError());
        SURTRY(hipStreamSynchronize(s));
        return TH_OK;
    }
    int download_output(float* host, int B) {
        const long n0 = npix(0, B);
        std::vector<float> h((size_t)n0 * 16);
        SURTRY(hipMemcpyAsync(h.data(), out, h.size() * sizeof(float), hipMemcpyDeviceToHost, s));
        SURTRY(hipStreamSynchronize(s));
        for (long p = 0; p < n0; ++p)
            for (int c = 0; c < kOut; ++c) host[p * kOut + c] = h[p * 16 + c];
        return TH_OK;
    }
};

}  // namespace sur
}  // namespace th

using namespace th;
using namespace th::sur;

struct th_surrogate {
    Net net;
};

extern "C" {

long th_surrogate_param_count(void) { return total_params(); }

int th_surrogate_create(int H, int W, unsigned long long seed, th_surrogate** out) {
    clear_error();
    if (!out) return fail(TH_ERR_ARG, "null out");
    if (H < 16 || W < 16 || H % 16 || W % 16) return fail(TH_ERR_ARG, "surrogate grid must be a multiple of 16");
    th_surrogate* s = new th_surrogate;
    int rc = s->net.init(H, W, seed);
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return TH_OK;
}

void th_surrogate_destroy(th_surrogate* s) { delete s; }

#define TH_SUR(s) \
    if (!(s)) return fail(TH_ERR_ARG, "null surrogate"); \
    clear_error();

int th_surrogate_get_params(th_surrogate* s, float* out) {
    TH_SUR(s);
    if (!out) return fail(TH_ERR_ARG, "null out");
    SURTRY(hipMemcpyAsync(out, s->net.prm, s->net.nparam * sizeof(float), hipMemcpyDeviceToHost, s->net.s));
    SURTRY(hipStreamSynchronize(s->net.s));
    return TH_OK;
}

int th_surrogate_set_params(th_surrogate* s, const float* in) {
    TH_SUR(s);
    if (!in) return fail(TH_ERR_ARG, "null params");
    return s->net.set_params(in);
}

int th_surrogate_predict(th_surrogate* s, const float* in, int batch, float* out) {
    TH_SUR(s);
    if (!in || !out || batch < 1) return fail(TH_ERR_ARG, "bad argument");
    Net& n = s->net;
    int rc;
    if ((rc = n.ensure_batch(batch)) || (rc = n.upload_input(in, batch)) || (rc = n.forward(batch))) return rc;
    return n.download_output(out, batch);
}

int th_surrogate_predict_device(th_surrogate* s, const float* in, const float** d_out) {
    TH_SUR(s);
    if (!in || !d_out) return fail(TH_ERR_ARG, "bad argument");
    Net& n = s->net;
    int rc;
    if ((rc = n.ensure_batch(1)) || (rc = n.upload_input(in, 1)) || (rc = n.forward(1))) return rc;
    SURTRY(hipStreamSynchronize(n.s));
    *d_out = n.out;
    return TH_OK;
}

int th_surrogate_predict_device_input(th_surrogate* s, const float* d_in, const float** d_out) {
    TH_SUR(s);
    if (!d_in || !d_out) return fail(TH_ERR_ARG, "bad argument");
    Net& n = s->net;
    int rc;
    if ((rc = n.ensure_batch(1))) return rc;
    const long npx = n.npix(0, 1) * 16;
    k_pack_input<<<grid1(npx), 256, 0, n.s>>>(d_in, n.x0, npx);
    SURTRY(hipGetLastError());
    if ((rc = n.forward(1))) return rc;
    SURTRY(hipStreamSynchronize(n.s));
    *d_out = n.out;
    return TH_OK;
}

}  // extern "C"

namespace th {
// The forward pass of th_surrogate_predict_device_input queued on another
// stream: that stream first waits for the surrogate's own queue, the
// network's kernels run on it, and the surrogate's queue then waits for
// them, so the two stay ordered without the host.
int surrogate_forward_on(th_surrogate* sur, const float* d_in, void* stream, int device, const float** out, int* H,
                         int* W, int* C) {
    if (!sur || !d_in || !stream || !out) return fail(TH_ERR_ARG, "bad argument");
    sur::Net& n = sur->net;
    if (n.dev != device) return fail(TH_ERR_ARG, "the surrogate and the solver are on different devices");
    int rc;
    if ((rc = n.ensure_batch(1))) return rc;
    hipStream_t other = (hipStream_t)stream;
    hipEvent_t ev = nullptr;
    SURTRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    hipError_t e = hipEventRecord(ev, n.s);
    if (e == hipSuccess) e = hipStreamWaitEvent(other, ev, 0);
    hipStream_t own = n.s;
    if (e == hipSuccess) {
        n.s = other;
        const long npx = n.npix(0, 1) * 16;
        sur::k_pack_input<<<sur::grid1(npx), 256, 0, n.s>>>(d_in, n.x0, npx);
        e = hipGetLastError();
        if (e == hipSuccess) rc = n.forward(1);
        n.s = own;
    }
    if (e == hipSuccess && !rc) e = hipEventRecord(ev, other);
    if (e == hipSuccess && !rc) e = hipStreamWaitEvent(own, ev, 0);
    (void)hipEventDestroy(ev);  // released once the waits are resolved
    if (rc) return rc;
    if (e != hipSuccess) return fail(TH_ERR_HIP, std::string("surrogate on the solver's stream: ") + hipGetErrorString(e));
    *out = n.out;
    if (H) *H = n.H;
    if (W) *W = n.W;
    if (C) *C = 16;
    return TH_OK;
}
}  // namespace th

extern "C" {

int th_surrogate_grad(th_surrogate* s, const float* in, const float* tgt, int batch, double* loss, float* grad) {
    TH_SUR(s);
    if (!in || !tgt || batch < 1) return fail(TH_ERR_ARG, "bad argument");
    Net& n = s->net;
    int rc;
    if ((rc = n.ensure_batch(batch)) || (rc = n.upload_input(in, batch))) return rc;
    SURTRY(hipMemcpyAsync(n.tgt_b, tgt, (size_t)n.npix(0, batch) * kOut * sizeof(float), hipMemcpyHostToDevice, n.s));
    double* dsum = nullptr;
    SURTRY(hipMalloc(&dsum, sizeof(double)));
    rc = n.forward(batch);
    if (!rc) rc = n.loss(batch, n.tgt_b, nullptr, dsum);
    if (!rc) rc = n.backward(batch);
    double h = 0.0;
    hipError_t e = hipSuccess;
    if (!rc) {
        e = hipMemcpyAsync(&h, dsum, sizeof(double), hipMemcpyDeviceToHost, n.s);
        if (e == hipSuccess && grad) e = hipMemcpyAsync(grad, n.grad, n.nparam * sizeof(float), hipMemcpyDeviceToHost, n.s);
        if (e == hipSuccess) e = hipStreamSynchronize(n.s);
    }
    (void)hipFree(dsum);
    if (rc) return rc;
    if (e != hipSuccess) return fail(TH_ERR_HIP, std::string("surrogate grad: ") + hipGetErrorString(e));
    if (loss) *loss = h;
    return TH_OK;
}

int th_surrogate_dataset_reserve(th_surrogate* s, int capacity) {
    TH_SUR(s);
    Net& n = s->net;
    if (capacity < 1) return fail(TH_ERR_ARG, "capacity must be >= 1");
    if (n.ds_cap) return fail(TH_ERR_ARG, "dataset already reserved");
    int rc;
    if ((rc = n.dalloc(&n.ds_in, (size_t)capacity * n.hw() * kInStore, n.own)) ||
        (rc = n.dalloc(&n.ds_tgt, (size_t)capacity * n.hw() * kOut, n.own)))
        return rc;
    n.ds_cap = capacity;
    n.ds_n = 0;
    return TH_OK;
}

int th_surrogate_dataset_add(th_surrogate* s, const float* in, const float* tgt, int count) {
    TH_SUR(s);
    Net& n = s->net;
    if (!in || !tgt || count < 1) return fail(TH_ERR_ARG, "bad argument");
    if (n.ds_n + count > n.ds_cap) return fail(TH_ERR_ARG, "dataset capacity exceeded");
    const long hw = n.hw();
    std::vector<bf> h((size_t)count * hw * kInStore, 0);
    for (long p = 0; p < (long)count * hw; ++p)
        for (int c = 0; c < kIn; ++c) h[p * kInStore + c] = host_f2bf(in[p * kIn + c]);
    SURTRY(hipMemcpyAsync(n.ds_in + (size_t)n.ds_n * hw * kInStore, h.data(), h.size() * sizeof(bf),
                          hipMemcpyHostToDevice, n.s));
    SURTRY(hipMemcpyAsync(n.ds_tgt + (size_t)n.ds_n * hw * kOut, tgt, (size_t)count * hw * kOut * sizeof(float),
                          hipMemcpyHostToDevice, n.s));
    SURTRY(hipStreamSynchronize(n.s));
    n.ds_n += count;
    return TH_OK;
}

int th_surrogate_dataset_size(th_surrogate* s) { return s ? s->net.ds_n : -1; }

// Adam on random batches of samples [0, n_train): `steps` steps, learning
// rate decaying as a cosine from lr to lr_final; per-step losses to host.
int th_surrogate_train(th_surrogate* s, int steps, int batch, float lr, float lr_final, unsigned long long seed,
                       int n_train, float* losses, double* ms) {
    TH_SUR(s);
    Net& n = s->net;
    if (steps < 1 || batch < 1 || n_train < batch || n_train > n.ds_n)
        return fail(TH_ERR_ARG, "bad training arguments (n_train must be in [batch, dataset size])");
    int rc;
    if ((rc = n.ensure_batch(batch))) return rc;
    if (n.dloss_cap < steps) {
        if (n.dloss) (void)hipFree(n.dloss);
        n.dloss = nullptr;
        SURTRY(hipMalloc(&n.dloss, steps * sizeof(double)));
        n.dloss_cap = steps;
    }
    SplitMix rng{seed};
    std::vector<int> order(n_train);
    std::iota(order.begin(), order.end(), 0);
    size_t pos = order.size();
    std::vector<int> idx(batch);
    int* pinned = nullptr;
    SURTRY(hipHostMalloc((void**)&pinned, (size_t)steps * batch * sizeof(int), hipHostMallocDefault));
    for (int t = 0; t < steps; ++t)
        for (int b = 0; b < batch; ++b) {
            if (pos >= order.size()) {  // new epoch: Fisher-Yates shuffle
                for (size_t i = order.size() - 1; i > 0; --i) std::swap(order[i], order[rng.next() % (i + 1)]);
                pos = 0;
            }
            pinned[(size_t)t * batch + b] = order[pos++];
        }
    int* didx = nullptr;
    hipError_t e = hipMalloc(&didx, (size_t)steps * batch * sizeof(int));
    if (e == hipSuccess)
        e = hipMemcpyAsync(didx, pinned, (size_t)steps * batch * sizeof(int), hipMemcpyHostToDevice, n.s);
    if (e != hipSuccess) {
        (void)hipHostFree(pinned);
        if (didx) (void)hipFree(didx);
        return fail(TH_ERR_HIP, std::string("surrogate train: ") + hipGetErrorString(e));
    }
    const auto t0 = std::chrono::high_resolution_clock::now();
    const long n0 = n.npix(0, batch);
    for (int t = 0; t < steps && !rc; ++t) {
        const int* bi = didx + (size_t)t * batch;
        k_gather<<<grid1(n0 * 16), 256, 0, n.s>>>(n.ds_in, bi, n.x0, n0 * 16, n.hw());
        if ((e = hipGetLastError()) != hipSuccess) break;
        const double frac = steps > 1 ? (double)t / (steps - 1) : 1.0;
        const float lr_t = (float)(lr_final + 0.5 * (lr - lr_final) * (1.0 + std::cos(3.141592653589793 * frac)));
        if ((rc = n.forward(batch)) || (rc = n.loss(batch, n.ds_tgt, bi, n.dloss + t)) || (rc = n.backward(batch)) ||
            (rc = n.adam(lr_t)))
            break;
    }
    std::vector<double> h(steps, 0.0);
    if (!rc && e == hipSuccess) e = hipMemcpyAsync(h.data(), n.dloss, steps * sizeof(double), hipMemcpyDeviceToHost, n.s);
    if (e == hipSuccess) e = hipStreamSynchronize(n.s);
    const auto t1 = std::chrono::high_resolution_clock::now();
    (void)hipFree(didx);
    (void)hipHostFree(pinned);
    if (rc) return rc;
    if (e != hipSuccess) return fail(TH_ERR_HIP, std::string("surrogate train: ") + hipGetErrorString(e));
    if (losses)
        for (int t = 0; t < steps; ++t) losses[t] = (float)h[t];
    if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return TH_OK;
}

// Mean squared error (scaled units) over dataset samples [first, first+count).
int th_surrogate_eval(th_surrogate* s, int first, int count, double* mse) {
    TH_SUR(s);
    Net& n = s->net;
    if (!mse || first < 0 || count < 1 || first + count > n.ds_n) return fail(TH_ERR_ARG, "bad range");
    int rc;
    const int B = std::min(count, std::max(1, n.Bcap > 0 ? n.Bcap : 8));
    if ((rc = n.ensure_batch(B))) return rc;
    double* dsum = nullptr;
    SURTRY(hipMalloc(&dsum, sizeof(double) * ((count + B - 1) / B)));
    std::vector<int> idx(B);
    int* pinned = nullptr;
    hipError_t e = hipHostMalloc((void**)&pinned, (size_t)count * sizeof(int), hipHostMallocDefault);
    double total = 0.0;
    long npx = 0;
    int nb = 0;
    std::vector<int> sizes;
    if (e == hipSuccess) {
        for (int i = 0; i < count; ++i) pinned[i] = first + i;
        int* didx = nullptr;
        e = hipMalloc(&didx, (size_t)count * sizeof(int));
        if (e == hipSuccess) e = hipMemcpyAsync(didx, pinned, (size_t)count * sizeof(int), hipMemcpyHostToDevice, n.s);
        for (int b0 = 0; b0 < count && e == hipSuccess && !rc; b0 += B, ++nb) {
            const int bs = std::min(B, count - b0);
            const long n0 = n.npix(0, bs);
            k_gather<<<grid1(n0 * 16), 256, 0, n.s>>>(n.ds_in, didx + b0, n.x0, n0 * 16, n.hw());
            e = hipGetLastError();
            if (e == hipSuccess && !(rc = n.forward(bs))) rc = n.loss(bs, n.ds_tgt, didx + b0, dsum + nb);
            sizes.push_back(bs);
        }
        std::vector<double> h(nb);
        if (e == hipSuccess && !rc) e = hipMemcpyAsync(h.data(), dsum, nb * sizeof(double), hipMemcpyDeviceToHost, n.s);
        if (e == hipSuccess) e = hipStreamSynchronize(n.s);
        for (int k = 0; k < nb; ++k) {
            total += h[k] * sizes[k];
            npx += sizes[k];
        }
        if (didx) (void)hipFree(didx);
    }
    if (pinned) (void)hipHostFree(pinned);
    (void)hipFree(dsum);
    if (rc) return rc;
    if (e != hipSuccess) return fail(TH_ERR_HIP, std::string("surrogate eval: ") + hipGetErrorString(e));
    *mse = total / npx;
    return TH_OK;
}

// Average time of one forward pass (batch B, inputs as they are) and its
// MFMA work: 2 H W cin cout k^2 per layer and image.
int th_surrogate_time_forward(th_surrogate* s, int batch, int reps, double* avg_ms, double* flops) {
    TH_SUR(s);
    Net& n = s->net;
    if (batch < 1 || reps < 1 || !avg_ms || !flops) return fail(TH_ERR_ARG, "bad argument");
    int rc;
    if ((rc = n.ensure_batch(batch)) || (rc = n.forward(batch))) return rc;
    hipEvent_t e0, e1;
    SURTRY(hipEventCreate(&e0));
    SURTRY(hipEventCreate(&e1));
    hipError_t e = hipEventRecord(e0, n.s);
    for (int r = 0; r < reps && !rc && e == hipSuccess; ++r) rc = n.forward(batch);
    if (!rc && e == hipSuccess) e = hipEventRecord(e1, n.s);
    if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (e != hipSuccess) return fail(TH_ERR_HIP, std::string("surrogate timing: ") + hipGetErrorString(e));
    double f = 0.0;
    for (const LayerSpec& l : kSpec) f += 2.0 * n.npix(l.level, batch) * l.cin * l.cout * l.ks * l.ks;
    *avg_ms = ms / reps;
    *flops = f;
    return TH_OK;
}

}  // extern "C"
