// Launch table and schedules.  A schedule of a grid size fixes the brick shape
// (BY x BX), and for each pass the radix and the number of fields per LDS
// exchange.  Every run of a given (precision, Nx, Ny, schedule) uses the same
// kernels and the same arithmetic, whether stepped alone or in an ensemble.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace af {

using RowLaunch = hipError_t (*)(const void* descs, int ninst, int tiles, const void* pool, hipStream_t s);
using ColLaunch = hipError_t (*)(const void* descs, int ninst, int tiles, int stage, const void* pool,
                                 hipStream_t s);
using ConvLaunch = hipError_t (*)(const void* src, void* dst, int nx, int ny, int by, int bx, hipStream_t s);
using IfacLaunch = hipError_t (*)(const void* desc, int nx2, int ny, int w, hipStream_t s);

// Schedule option bits (the kernels' PP template argument): 1 double-buffered
// LDS exchanges (both passes), 2 streaming column-pass stores, 4 streaming
// row-pass stores (otherwise device-scope write-through), 8 streaming loads of
// P, Q in the column pass, 16 FP32 integrating factors from the per-run table
// also at radix 16 (radix <= 8 always reads it), 32 paired rows in the row
// pass (two rows as one full-length complex transform), 64 half-brick row
// blocks (a row block covers half of a brick row: by/2 rows), 128 P, Q in
// bricks of 2 by rows (both passes), 256 (with 64) U, V, W in bricks of by/2
// rows (both passes).  A row kernel is instantiated for the bits it reads
// only, as is a column kernel.
constexpr int kOptLdsDouble = 1, kOptColStream = 2, kOptRowStream = 4, kOptPqStream = 8, kOptIfTable = 16,
              kOptPairedRows = 32, kOptHalfRows = 64, kOptTallPq = 128, kOptShortUvw = 256;
constexpr int kRowOptions =
    kOptLdsDouble | kOptRowStream | kOptPairedRows | kOptHalfRows | kOptTallPq | kOptShortUvw;
constexpr int kColOptions = kOptLdsDouble | kOptColStream | kOptPqStream | kOptIfTable | kOptTallPq | kOptShortUvw;

struct RowCfg {
  int nx2, rb, r, bx, threads;
  RowLaunch launch;
  int g, pp;   // pp: option bits of the row kernel
};
struct ColCfg {
  int ny, w, r, by, bx, threads;   // w: columns per block (divides the brick width bx)
  ColLaunch launch;
  int gf, gi, pp;   // pp: option bits of the column kernel
};
struct Schedule {
  int nx2, ny, by, bx, row_r, row_g, col_r, col_gf, col_gi, col_w, pp, is_default;   // pp: option bits
};

// Schedules instantiated for (precision, nx2, ny); index 0 is the default.
int schedule_count(int precision, int nx2, int ny);
const Schedule* schedule_at(int precision, int nx2, int ny, int index);
// The schedule new runs of this size use: a process-wide selection, else the
// tuned table beside the library (schedules_batched.txt first for runs of the
// ensemble engine, then schedules.txt), else the default.
const Schedule* active_schedule(int precision, int nx2, int ny, bool batched = false);
bool select_schedule(int precision, int nx2, int ny, int index);
const RowCfg* find_row_cfg(int precision, const Schedule* s);
const ColCfg* find_col_cfg(int precision, const Schedule* s);
// Mixed launches (one launch per pass over several size classes, the
// ensemble's single-launch mode; multi_pass.hpp).  Arguments: per class, in
// launch order, its descriptor array, twiddle pool, tiles per instance and
// blocks in the launch (entries x tiles).
constexpr int kMaxMulti = 4;
struct MultiArgs {
  const void* descs[kMaxMulti];
  const void* pool[kMaxMulti];
  int tiles[kMaxMulti];
  int nblocks[kMaxMulti];
};
using MultiRowLaunch = hipError_t (*)(const MultiArgs& a, hipStream_t s);
using MultiColLaunch = hipError_t (*)(const MultiArgs& a, int stage, hipStream_t s);
// An instantiated mix: its classes' schedules in launch order (the row
// arguments use the row tiles, the column arguments the column tiles).
struct MultiCfg {
  int precision, nclass;
  Schedule cls[kMaxMulti];
  MultiRowLaunch row;
  MultiColLaunch col;
};
const MultiCfg* multi_cfgs(int* n);
// When set (profiling), the next row/column launch on this thread records
// these events at the kernel's own start and end (hipExtLaunchKernel).
struct LaunchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local LaunchEvents g_launch_events;
ConvLaunch to_bricks_launch(int precision);
IfacLaunch ifac_init_launch();   // FP32 runs only
ConvLaunch from_bricks_launch(int precision);

}  // namespace af
