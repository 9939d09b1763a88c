// levels.h - the multigrid hierarchy's coarse operators, built on the host.
//
// Semi-coarsening in x and y only (every level keeps all z-planes), 2 x 2
// columns per coarse column, down to a single column; odd sizes leave a
// last coarse cell over one fine cell. Coarse operators come from
// aggregating the fine faces (the rule of src/mg3d.h:100-183):
//   gU, cap, gB   sum over the children (parallel paths)
//   gE (gN)       sum of the fine faces crossing the coarse face, scaled by
//                 (fine centre distance / coarse centre distance)
// Sums are taken in double in the children's natural order and rounded to
// float, so the coarse arrays match the reference's bit for bit.
#pragma once
#include <vector>

namespace th {

struct HostLevel {
    int nx, ny, nz;
    std::vector<double> wx, wy;  // cell widths along x and y (m)
    std::vector<float> gE, gN, gU, gB, cap;  // natural layout
};

// Level 0 is the given fine operator; the last level is 1 x 1.
std::vector<HostLevel> build_hierarchy(int nx, int ny, int nz, double dx, double dy, const float* gE,
                                       const float* gN, const float* gU, const float* gB, const float* cap);

}  // namespace th
