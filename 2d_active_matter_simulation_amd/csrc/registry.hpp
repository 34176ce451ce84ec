// Launch table: one row-pass and one column-pass kernel per (precision, size,
// brick shape).  A size's brick shape and radix are fixed here, so every run
// of a given (precision, Nx, Ny) uses the same arithmetic whether it is
// stepped alone or inside an ensemble launch.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace af {

using RowLaunch = hipError_t (*)(const void* descs, int ninst, int tiles, const void* pool, hipStream_t s);
using ColLaunch = hipError_t (*)(const void* descs, int ninst, int tiles, int stage, const void* pool,
                                 hipStream_t s);
using ConvLaunch = hipError_t (*)(const void* src, void* dst, int nx, int ny, int by, int bx, hipStream_t s);

struct RowCfg {
  int nx2, rb, r, bx, threads;
  RowLaunch launch;
};
struct ColCfg {
  int ny, w, r, by, threads;
  ColLaunch launch;
};
struct PrecCfg {
  size_t elem_bytes;      // sizeof(complex)
  ConvLaunch to_bricks;   // reference-layout doubles -> bricks
  ConvLaunch from_bricks; // bricks -> reference-layout doubles
};

// Brick shape for a grid: BY follows the row length (rows per row block),
// BX follows the column length (columns per column block).
int brick_by(int precision, int nx2);
int brick_bx(int precision, int ny);
const RowCfg* find_row_cfg(int precision, int nx2, int bx);
const ColCfg* find_col_cfg(int precision, int ny, int by);
const PrecCfg* find_prec_cfg(int precision);

}  // namespace af
