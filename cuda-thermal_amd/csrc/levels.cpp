// levels.cpp - see levels.h.
#include "levels.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <unordered_map>

namespace th {

static std::vector<double> merge_pairs(const std::vector<double>& w) {
    const int n = (int)w.size();
    std::vector<double> c((n + 1) / 2);
    for (int I = 0; I < (int)c.size(); ++I) c[I] = w[2 * I] + (2 * I + 1 < n ? w[2 * I + 1] : 0.0);
    return c;
}

static HostLevel coarsen(const HostLevel& f) {
    HostLevel c;
    c.nx = (f.nx + 1) / 2;
    c.ny = (f.ny + 1) / 2;
    c.nz = f.nz;
    c.wx = merge_pairs(f.wx);
    c.wy = merge_pairs(f.wy);
    const size_t fp = (size_t)f.nx * f.ny, cp = (size_t)c.nx * c.ny, cn = cp * c.nz;
    c.gE.assign(cn, 0.0f);
    c.gN.assign(cn, 0.0f);
    c.gU.assign(cn, 0.0f);
    c.cap.assign(cn, 0.0f);
    c.gB.assign(cp, 0.0f);
    for (int l = 0; l < c.nz; ++l) {
        const float* fU = f.gU.data() + l * fp;
        const float* fC = f.cap.data() + l * fp;
        const float* fE = f.gE.data() + l * fp;
        const float* fN = f.gN.data() + l * fp;
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
                if (I + 1 < c.nx) {  // fine faces between columns 2I+1 and 2I+2
                    double s = 0;
                    for (int j = j0; j <= j1; ++j) s += fE[(size_t)j * f.nx + 2 * I + 1];
                    // scaled by fine / coarse centre distance across the face
                    const double df = 0.5 * (f.wx[2 * I + 1] + f.wx[2 * I + 2]);
                    const double dc = 0.5 * (c.wx[I] + c.wx[I + 1]);
                    c.gE[cid] = (float)(s * df / dc);
                }
                if (J + 1 < c.ny) {
                    double s = 0;
                    for (int i = i0; i <= i1; ++i) s += fN[(size_t)(2 * J + 1) * f.nx + i];
                    const double df = 0.5 * (f.wy[2 * J + 1] + f.wy[2 * J + 2]);
                    const double dc = 0.5 * (c.wy[J] + c.wy[J + 1]);
                    c.gN[cid] = (float)(s * df / dc);
                }
            }
        }
    }
    return c;
}

std::vector<HostLevel> build_hierarchy(int nx, int ny, int nz, double dx, double dy, const float* gE,
                                       const float* gN, const float* gU, const float* gB, const float* cap) {
    std::vector<HostLevel> L(1);
    const size_t n = (size_t)nx * ny * nz, p = (size_t)nx * ny;
    HostLevel& f = L[0];
    f.nx = nx;
    f.ny = ny;
    f.nz = nz;
    f.wx.assign(nx, dx);
    f.wy.assign(ny, dy);
    f.gE.assign(gE, gE + n);
    f.gN.assign(gN, gN + n);
    f.gU.assign(gU, gU + n);
    f.gB.assign(gB, gB + p);
    f.cap.assign(cap, cap + n);
    while (L.back().nx > 1 || L.back().ny > 1) L.push_back(coarsen(L.back()));
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
    // field f of every column on plane l (west / south faces are 0 where the
    // neighbour is absent), into buf
    std::vector<float> buf(cols);
    auto plane_values = [&](int f, int l) -> const float* {
        const long o = (long)l * plane;
        if (f == TF_W) {
            for (int j = 0; j < ny; ++j) {
                const long r = (long)j * nx;
                buf[r] = 0.f;
                std::memcpy(&buf[r + 1], gE + o + r, (size_t)(nx - 1) * sizeof(float));
            }
            return buf.data();
        }
        if (f == TF_S) {
            std::fill(buf.begin(), buf.begin() + nx, 0.f);
            std::memcpy(&buf[nx], gN + o, (size_t)(cols - nx) * sizeof(float));
            return buf.data();
        }
        return (f == TF_E ? gE : f == TF_N ? gN : f == TF_U ? gU : cap) + o;
    };
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
    std::vector<uint64_t> h1(cols);
    for (long c = 0; c < cols; ++c) h1[c] = mix64(0x243F6A8885A308D3ull, flags_of(c) | ((uint64_t)fbits(gB[c]) << 8));
    for (int l = 0; l < nz; ++l)  // plane by plane: every array is read in order
        for (int f = TF_E; f <= TF_S; ++f) {
            const float* vals = plane_values(f, l);
            const uint64_t tag = (uint64_t)(f * 64 + l) << 32;
            for (long c = 0; c < cols; ++c) h1[c] = mix64(h1[c], fbits(vals[c]) | tag);
        }
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
    for (int l = 0; l < nz; ++l)
        for (int f = TF_E; f <= TF_S; ++f) {
            const float* vals = plane_values(f, l);
            const float* col = &out.tab[(size_t)l * kTypeFields + f];  // type t's value: col[t * nz * kTypeFields]
            for (long c = 0; c < cols; ++c)
                if (fbits(col[(size_t)out.type[c] * kTypeFields * nz]) != fbits(vals[c])) return ColumnTypes{};
        }
    out.ntypes = nt;
    return out;
}

}  // namespace th
