// levels.cpp - see levels.h.
#include "levels.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_map>

namespace th {

static std::vector<double> merge_pairs(const std::vector<double>& w) {
    const int n = (int)w.size();
    std::vector<double> c((n + 1) / 2);
    for (int I = 0; I < (int)c.size(); ++I) c[I] = w[2 * I] + (2 * I + 1 < n ? w[2 * I + 1] : 0.0);
    return c;
}

// Runs fn(k) for k = 0 .. n-1 on up to hardware_concurrency threads (at
// most 16), each taking a contiguous range of k.
template <class Fn>
static void parallel_for(long n, Fn fn) {
    const long hw = std::max(1u, std::thread::hardware_concurrency());
    const long nt = std::max(1L, std::min({n, hw, 16L}));
    if (nt == 1) {
        for (long k = 0; k < n; ++k) fn(k);
        return;
    }
    std::vector<std::thread> pool;
    for (long t = 0; t < nt; ++t)
        pool.emplace_back([&, t]() {
            for (long k = n * t / nt; k < n * (t + 1) / nt; ++k) fn(k);
        });
    for (std::thread& th : pool) th.join();
}

// The fine arrays of one level, natural layout (not owned).
struct LevelView {
    int nx, ny, nz;
    const std::vector<double>*wx, *wy;
    const float *gE, *gN, *gU, *gB, *cap;
};

// Coarse operator of a level (mg3d.h:100-136), plane by plane in parallel;
// every coarse value is one double sum in the children's natural order.
static HostLevel coarsen(const LevelView& f) {
    HostLevel c;
    c.nx = (f.nx + 1) / 2;
    c.ny = (f.ny + 1) / 2;
    c.nz = f.nz;
    c.wx = merge_pairs(*f.wx);
    c.wy = merge_pairs(*f.wy);
    const size_t fp = (size_t)f.nx * f.ny, cp = (size_t)c.nx * c.ny, cn = cp * c.nz;
    c.gE.assign(cn, 0.0f);
    c.gN.assign(cn, 0.0f);
    c.gU.assign(cn, 0.0f);
    c.cap.assign(cn, 0.0f);
    c.gB.assign(cp, 0.0f);
    const std::vector<double>&fwx = *f.wx, &fwy = *f.wy;
    parallel_for(c.nz, [&](long l) {
        const float* fU = f.gU + l * fp;
        const float* fC = f.cap + l * fp;
        const float* fE = f.gE + l * fp;
        const float* fN = f.gN + l * fp;
        for (int J = 0; J < c.ny; ++J) {
            const int j0 = 2 * J, j1 = std::min(2 * J + 1, f.ny - 1);
            for (int I = 0; I < c.nx; ++I) {
                const int i0 = 2 * I, i1 = std::min(2 * I + 1, f.nx - 1);
                const size_t cid = l * cp + (size_t)J * c.nx + I;
                double up = 0, heat = 0, bot = 0;
                for (int j = j0; j <= j1; ++j)
                    for (int i = i0; i <= i1; ++i) {
                        const size_t k = (size_t)j * f.nx + i;
                        up += fU[k];
                        heat += fC[k];
                        if (l == 0) bot += f.gB[k];
                    }
                c.gU[cid] = (float)up;
                c.cap[cid] = (float)heat;
                if (l == 0) c.gB[(size_t)J * c.nx + I] = (float)bot;
                if (I + 1 < c.nx) {  // the fine faces between columns 2I+1 and 2I+2
                    double sum = 0;
                    for (int j = j0; j <= j1; ++j) sum += fE[(size_t)j * f.nx + 2 * I + 1];
                    // times fine over coarse centre distance across the face
                    const double df = 0.5 * (fwx[2 * I + 1] + fwx[2 * I + 2]);
                    const double dc = 0.5 * (c.wx[I] + c.wx[I + 1]);
                    c.gE[cid] = (float)(sum * df / dc);
                }
                if (J + 1 < c.ny) {
                    double sum = 0;
                    for (int i = i0; i <= i1; ++i) sum += fN[(size_t)(2 * J + 1) * f.nx + i];
                    const double df = 0.5 * (fwy[2 * J + 1] + fwy[2 * J + 2]);
                    const double dc = 0.5 * (c.wy[J] + c.wy[J + 1]);
                    c.gN[cid] = (float)(sum * df / dc);
                }
            }
        }
    });
    return c;
}

std::vector<HostLevel> build_hierarchy(int nx, int ny, int nz, double dx, double dy, const float* gE,
                                       const float* gN, const float* gU, const float* gB, const float* cap) {
    // Level 0 keeps only its size and cell widths: its arrays are the caller's.
    std::vector<HostLevel> L(1);
    HostLevel& f = L[0];
    f.nx = nx;
    f.ny = ny;
    f.nz = nz;
    f.wx.assign(nx, dx);
    f.wy.assign(ny, dy);
    LevelView v{nx, ny, nz, &L[0].wx, &L[0].wy, gE, gN, gU, gB, cap};
    while (L.back().nx > 1 || L.back().ny > 1) {
        HostLevel c = coarsen(v);
        L.push_back(std::move(c));
        const HostLevel& b = L.back();
        v = LevelView{b.nx, b.ny, b.nz, &b.wx, &b.wy, b.gE.data(), b.gN.data(), b.gU.data(), b.gB.data(), b.cap.data()};
    }
    return L;
}


namespace {
inline uint64_t mix64(uint64_t h, uint64_t v) {  // splitmix64 finaliser of h ^ v
    uint64_t z = (h ^ v) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
}  // namespace

ColumnTypes column_types(int nx, int ny, int nz, const float* gE, const float* gN, const float* gU,
                         const float* gB, const float* cap, int max_types) {
    ColumnTypes out;
    const long cols = (long)nx * ny, plane = cols;
    auto value = [&](long c, int f, int l) -> float {  // one value (representatives only)
        const int i = (int)(c % nx), j = (int)(c / nx);
        const long id = (long)l * plane + c;
        switch (f) {
            case TF_E: return gE[id];
            case TF_N: return gN[id];
            case TF_U: return gU[id];
            case TF_C: return cap[id];
            case TF_W: return i > 0 ? gE[id - 1] : 0.f;
            case TF_S: return j > 0 ? gN[id - nx] : 0.f;
            default: return 0.f;
        }
    };
    auto flags_of = [&](long c) -> unsigned char {
        const int i = (int)(c % nx), j = (int)(c / nx);
        return (unsigned char)((i > 0) | ((i < nx - 1) << 1) | ((j > 0) << 2) | ((j < ny - 1) << 3));
    };
    // Hashes: rows of columns in parallel, each row's columns plane by plane
    // and field by field (the same sequence of values per column as a
    // sequential pass, so the same hashes).
    std::vector<uint64_t> h1(cols);
    // The six fields of row j's columns on plane l, into v[f][i].
    auto row_fields = [&](long j, int l, std::vector<float> (&v)[TF_S + 1]) {
        const long o = (long)l * plane + j * nx;
        for (int f = TF_E; f <= TF_S; ++f) v[f].resize(nx);
        std::memcpy(v[TF_E].data(), gE + o, nx * sizeof(float));
        std::memcpy(v[TF_N].data(), gN + o, nx * sizeof(float));
        std::memcpy(v[TF_U].data(), gU + o, nx * sizeof(float));
        std::memcpy(v[TF_C].data(), cap + o, nx * sizeof(float));
        v[TF_W][0] = 0.f;
        std::memcpy(v[TF_W].data() + 1, gE + o, (nx - 1) * sizeof(float));
        if (j > 0) std::memcpy(v[TF_S].data(), gN + o - nx, nx * sizeof(float));
        else std::fill(v[TF_S].begin(), v[TF_S].end(), 0.f);
    };
    parallel_for(ny, [&](long j) {
        const long r0 = j * nx;
        std::vector<float> v[TF_S + 1];
        for (long c = r0; c < r0 + nx; ++c) h1[c] = mix64(0x243F6A8885A308D3ull, flags_of(c) | ((uint64_t)fbits(gB[c]) << 8));
        for (int l = 0; l < nz; ++l) {
            row_fields(j, l, v);
            for (int f = TF_E; f <= TF_S; ++f) {
                const uint64_t tag = (uint64_t)(f * 64 + l) << 32;
                for (int i = 0; i < nx; ++i) h1[r0 + i] = mix64(h1[r0 + i], fbits(v[f][i]) | tag);
            }
        }
    });
    std::unordered_map<uint64_t, int> ids;
    std::vector<long> rep;
    out.type.resize(cols);
    for (long c = 0; c < cols; ++c) {
        auto it = ids.find(h1[c]);
        int t;
        if (it == ids.end()) {
            t = (int)rep.size();
            if (t >= max_types) return ColumnTypes{};
            ids.emplace(h1[c], t);
            rep.push_back(c);
        } else {
            t = it->second;  // a hash collision fails the check below
        }
        out.type[c] = (unsigned short)t;
    }
    const int nt = (int)rep.size();
    out.tab.assign((size_t)nt * kTypeFields * nz, 0.f);
    out.flags.resize(nt);
    for (int t = 0; t < nt; ++t) {
        out.flags[t] = flags_of(rep[t]);
        for (int f = TF_E; f <= TF_S; ++f)
            for (int l = 0; l < nz; ++l) out.tab[((size_t)t * nz + l) * kTypeFields + f] = value(rep[t], f, l);
        out.tab[(size_t)t * nz * kTypeFields + TF_B] = gB[rep[t]];
    }
    // every column against its type, bit for bit
    for (long c = 0; c < cols; ++c) {
        const int t = out.type[c];
        if (out.flags[t] != flags_of(c) || fbits(out.tab[(size_t)t * nz * kTypeFields + TF_B]) != fbits(gB[c]))
            return ColumnTypes{};
    }
    std::vector<char> bad(ny, 0);  // per row: a column that differs from its type
    parallel_for(ny, [&](long j) {
        std::vector<float> v[TF_S + 1];
        for (int l = 0; l < nz && !bad[j]; ++l) {
            row_fields(j, l, v);
            for (int i = 0; i < nx && !bad[j]; ++i) {
                const float* tv = &out.tab[((size_t)out.type[j * nx + i] * nz + l) * kTypeFields];
                for (int f = TF_E; f <= TF_S; ++f)
                    if (fbits(tv[f]) != fbits(v[f][i])) {
                        bad[j] = 1;
                        break;
                    }
            }
        }
    });
    for (char b : bad)
        if (b) return ColumnTypes{};
    out.ntypes = nt;
    return out;
}

}  // namespace th
