// coltypes.h - the fields of a column type (see levels.h: column_types),
// shared by the host builder and the kernels that read the type table.
#pragma once

namespace th {
// gE, gN, gU, cap on every plane; the west neighbour's gE and the south
// neighbour's gN (the column's own west and south faces); 1/pivot (computed
// on the device for the current dt); gB (plane 0 only).
enum { TF_E = 0, TF_N, TF_U, TF_C, TF_W, TF_S, TF_FI, TF_B, kTypeFields };
constexpr int kMaxColumnTypes = 4096;
}  // namespace th
