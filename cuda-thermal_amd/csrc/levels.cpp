// levels.cpp - see levels.h.
#include "levels.h"

#include <algorithm>

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

}  // namespace th
