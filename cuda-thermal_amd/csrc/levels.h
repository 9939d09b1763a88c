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

#include "coltypes.h"

namespace th {

struct HostLevel {
    int nx, ny, nz;
    std::vector<double> wx, wy;  // cell widths along x and y (m)
    std::vector<float> gE, gN, gU, gB, cap;  // natural layout
};

// Level 0 is the given fine operator; the last level is 1 x 1.
std::vector<HostLevel> build_hierarchy(int nx, int ny, int nz, double dx, double dy, const float* gE,
                                       const float* gN, const float* gU, const float* gB, const float* cap);

// Column types of a level. A layered stack repeats a few columns over and
// over: a column is described by its faces (gE, gN, gU, cap on every plane,
// the west neighbour's gE and the south neighbour's gN, which are its own
// west and south faces, and gB) and by which of its 4 lateral neighbours
// exist. Columns with equal descriptions get one type; the kernels then
// read a 2-byte type per column and the type's faces from a small table
// instead of 6 arrays per cell. Types are found by a 64-bit hash per column
// and every column is then compared, value by value, with its type's
// description: any mismatch (a hash collision) or more than max_types types
// returns ntypes = 0, and the caller keeps the per-cell arrays.
struct ColumnTypes {
    int ntypes = 0;
    std::vector<unsigned short> type;  // natural column order j * nx + i
    std::vector<float> tab;            // [type][plane][field]; TF_FI left 0 (device-computed), TF_B on plane 0
    std::vector<unsigned char> flags;  // [type]: bit 0 west, 1 east, 2 south, 3 north neighbour present
};
ColumnTypes column_types(int nx, int ny, int nz, const float* gE, const float* gN, const float* gU,
                         const float* gB, const float* cap, int max_types);

}  // namespace th
