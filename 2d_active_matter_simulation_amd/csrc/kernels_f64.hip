// FP64 schedule and launch tables (the kernels are instantiated in
// kernels_f64_k*.hip, compiled in parallel).
// Generated by gen_kernel_tables.py -- edit that script, not this file.
// Row kernel:    (Nx/2, RB = brick height, R, BX = brick width, G, PP)
// Column kernel: (Ny, W = columns per block, R, BY, BX = brick height/width, GF, GI, PP)
// Schedule:      (Nx/2, Ny, BY, BX, row R, row G, col R, col GF, col GI, col W, PP, default)
// PP: option bits (1 double-buffered LDS, 2 / 4 streaming column / row stores,
// 8 streaming column-pass P, Q loads, 16 integrating-factor table at radix 16)
#include "kernels_inst.hpp"

namespace af {
extern template hipError_t launch_row<double, 16, 4, 4, 1, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 16, 4, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 32, 4, 8, 1, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 32, 4, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 64, 4, 8, 1, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 64, 4, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 128, 2, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 128, 2, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 128, 4, 4, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 128, 4, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 128, 4, 4, 2, 3, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 128, 4, 8, 1, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 128, 4, 8, 1, 3, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 128, 4, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 128, 4, 8, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 128, 4, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 256, 2, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 256, 2, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 256, 4, 4, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 256, 4, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 256, 4, 4, 2, 3, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 256, 4, 8, 1, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 256, 4, 8, 1, 3, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 256, 4, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 256, 4, 8, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 256, 4, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 512, 2, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 512, 2, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 512, 4, 4, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 512, 4, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 512, 4, 8, 1, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 512, 4, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 512, 4, 8, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 512, 4, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 1024, 2, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 1024, 4, 8, 1, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 1024, 4, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 2048, 2, 8, 1, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<double, 2048, 2, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 32, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 32, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 64, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 64, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 128, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 128, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 256, 1, 4, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 256, 1, 8, 4, 1, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 256, 1, 8, 4, 1, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 256, 1, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 256, 1, 8, 4, 2, 1, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 256, 1, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 256, 2, 4, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 256, 2, 4, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 256, 2, 4, 4, 2, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 256, 2, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 256, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 256, 2, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 256, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 512, 1, 4, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 512, 1, 8, 4, 1, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 512, 1, 8, 4, 1, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 512, 1, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 512, 1, 8, 4, 2, 1, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 512, 1, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 512, 2, 4, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 512, 2, 4, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 512, 2, 4, 4, 2, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 512, 2, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 512, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 512, 2, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 512, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 1024, 1, 4, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 1024, 1, 8, 4, 1, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 1024, 1, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 1024, 1, 8, 4, 2, 1, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 1024, 1, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 1024, 2, 4, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 1024, 2, 4, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 1024, 2, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 1024, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 1024, 2, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 1024, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 2048, 1, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 2048, 2, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 2048, 2, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 4096, 1, 8, 2, 1, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<double, 4096, 1, 8, 4, 1, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
namespace {
using T = double;
const RowCfg kRow[] = {
    {16, 4, 4, 1, 16, launch_row<T, 16, 4, 4, 1, 3, 0>, 3, 0},
    {16, 4, 4, 2, 16, launch_row<T, 16, 4, 4, 2, 3, 0>, 3, 0},
    {32, 4, 8, 1, 16, launch_row<T, 32, 4, 8, 1, 3, 0>, 3, 0},
    {32, 4, 8, 2, 16, launch_row<T, 32, 4, 8, 2, 3, 0>, 3, 0},
    {64, 4, 8, 1, 32, launch_row<T, 64, 4, 8, 1, 3, 0>, 3, 0},
    {64, 4, 8, 2, 32, launch_row<T, 64, 4, 8, 2, 3, 0>, 3, 0},
    {128, 2, 4, 2, 64, launch_row<T, 128, 2, 4, 2, 3, 0>, 3, 0},
    {128, 2, 8, 2, 32, launch_row<T, 128, 2, 8, 2, 3, 0>, 3, 0},
    {128, 4, 4, 2, 128, launch_row<T, 128, 4, 4, 2, 1, 0>, 1, 0},
    {128, 4, 4, 2, 128, launch_row<T, 128, 4, 4, 2, 3, 0>, 3, 0},
    {128, 4, 4, 2, 128, launch_row<T, 128, 4, 4, 2, 3, 1>, 3, 1},
    {128, 4, 8, 1, 64, launch_row<T, 128, 4, 8, 1, 3, 0>, 3, 0},
    {128, 4, 8, 1, 64, launch_row<T, 128, 4, 8, 1, 3, 1>, 3, 1},
    {128, 4, 8, 2, 64, launch_row<T, 128, 4, 8, 2, 1, 0>, 1, 0},
    {128, 4, 8, 2, 64, launch_row<T, 128, 4, 8, 2, 1, 1>, 1, 1},
    {128, 4, 8, 2, 64, launch_row<T, 128, 4, 8, 2, 3, 0>, 3, 0},
    {256, 2, 4, 2, 128, launch_row<T, 256, 2, 4, 2, 3, 0>, 3, 0},
    {256, 2, 8, 2, 64, launch_row<T, 256, 2, 8, 2, 3, 0>, 3, 0},
    {256, 4, 4, 2, 256, launch_row<T, 256, 4, 4, 2, 1, 0>, 1, 0},
    {256, 4, 4, 2, 256, launch_row<T, 256, 4, 4, 2, 3, 0>, 3, 0},
    {256, 4, 4, 2, 256, launch_row<T, 256, 4, 4, 2, 3, 1>, 3, 1},
    {256, 4, 8, 1, 128, launch_row<T, 256, 4, 8, 1, 3, 0>, 3, 0},
    {256, 4, 8, 1, 128, launch_row<T, 256, 4, 8, 1, 3, 1>, 3, 1},
    {256, 4, 8, 2, 128, launch_row<T, 256, 4, 8, 2, 1, 0>, 1, 0},
    {256, 4, 8, 2, 128, launch_row<T, 256, 4, 8, 2, 1, 1>, 1, 1},
    {256, 4, 8, 2, 128, launch_row<T, 256, 4, 8, 2, 3, 0>, 3, 0},
    {512, 2, 4, 2, 256, launch_row<T, 512, 2, 4, 2, 3, 0>, 3, 0},
    {512, 2, 8, 2, 128, launch_row<T, 512, 2, 8, 2, 3, 0>, 3, 0},
    {512, 4, 4, 2, 512, launch_row<T, 512, 4, 4, 2, 1, 0>, 1, 0},
    {512, 4, 4, 2, 512, launch_row<T, 512, 4, 4, 2, 3, 0>, 3, 0},
    {512, 4, 8, 1, 256, launch_row<T, 512, 4, 8, 1, 3, 0>, 3, 0},
    {512, 4, 8, 2, 256, launch_row<T, 512, 4, 8, 2, 1, 0>, 1, 0},
    {512, 4, 8, 2, 256, launch_row<T, 512, 4, 8, 2, 1, 1>, 1, 1},
    {512, 4, 8, 2, 256, launch_row<T, 512, 4, 8, 2, 3, 0>, 3, 0},
    {1024, 2, 4, 2, 512, launch_row<T, 1024, 2, 4, 2, 3, 0>, 3, 0},
    {1024, 4, 8, 1, 512, launch_row<T, 1024, 4, 8, 1, 1, 0>, 1, 0},
    {1024, 4, 8, 2, 512, launch_row<T, 1024, 4, 8, 2, 1, 0>, 1, 0},
    {2048, 2, 8, 1, 512, launch_row<T, 2048, 2, 8, 1, 1, 0>, 1, 0},
    {2048, 2, 8, 2, 512, launch_row<T, 2048, 2, 8, 2, 1, 0>, 1, 0},
};
const ColCfg kCol[] = {
    {32, 2, 8, 2, 2, 8, launch_col<T, 32, 2, 8, 2, 2, 2, 3, 0>, 2, 3, 0},
    {32, 2, 8, 4, 2, 8, launch_col<T, 32, 2, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {64, 2, 8, 2, 2, 16, launch_col<T, 64, 2, 8, 2, 2, 2, 3, 0>, 2, 3, 0},
    {64, 2, 8, 4, 2, 16, launch_col<T, 64, 2, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {128, 2, 8, 2, 2, 32, launch_col<T, 128, 2, 8, 2, 2, 2, 3, 0>, 2, 3, 0},
    {128, 2, 8, 4, 2, 32, launch_col<T, 128, 2, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {256, 1, 4, 4, 2, 64, launch_col<T, 256, 1, 4, 4, 2, 2, 3, 0>, 2, 3, 0},
    {256, 1, 8, 4, 1, 32, launch_col<T, 256, 1, 8, 4, 1, 2, 3, 0>, 2, 3, 0},
    {256, 1, 8, 4, 1, 32, launch_col<T, 256, 1, 8, 4, 1, 2, 3, 1>, 2, 3, 1},
    {256, 1, 8, 4, 2, 32, launch_col<T, 256, 1, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {256, 1, 8, 4, 2, 32, launch_col<T, 256, 1, 8, 4, 2, 1, 1, 1>, 1, 1, 1},
    {256, 1, 8, 4, 2, 32, launch_col<T, 256, 1, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {256, 2, 4, 4, 2, 128, launch_col<T, 256, 2, 4, 4, 2, 1, 1, 0>, 1, 1, 0},
    {256, 2, 4, 4, 2, 128, launch_col<T, 256, 2, 4, 4, 2, 2, 3, 0>, 2, 3, 0},
    {256, 2, 4, 4, 2, 128, launch_col<T, 256, 2, 4, 4, 2, 2, 3, 1>, 2, 3, 1},
    {256, 2, 8, 2, 2, 64, launch_col<T, 256, 2, 8, 2, 2, 1, 1, 0>, 1, 1, 0},
    {256, 2, 8, 2, 2, 64, launch_col<T, 256, 2, 8, 2, 2, 2, 3, 0>, 2, 3, 0},
    {256, 2, 8, 4, 2, 64, launch_col<T, 256, 2, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {256, 2, 8, 4, 2, 64, launch_col<T, 256, 2, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {512, 1, 4, 4, 2, 128, launch_col<T, 512, 1, 4, 4, 2, 2, 3, 0>, 2, 3, 0},
    {512, 1, 8, 4, 1, 64, launch_col<T, 512, 1, 8, 4, 1, 2, 3, 0>, 2, 3, 0},
    {512, 1, 8, 4, 1, 64, launch_col<T, 512, 1, 8, 4, 1, 2, 3, 1>, 2, 3, 1},
    {512, 1, 8, 4, 2, 64, launch_col<T, 512, 1, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {512, 1, 8, 4, 2, 64, launch_col<T, 512, 1, 8, 4, 2, 1, 1, 1>, 1, 1, 1},
    {512, 1, 8, 4, 2, 64, launch_col<T, 512, 1, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {512, 2, 4, 4, 2, 256, launch_col<T, 512, 2, 4, 4, 2, 1, 1, 0>, 1, 1, 0},
    {512, 2, 4, 4, 2, 256, launch_col<T, 512, 2, 4, 4, 2, 2, 3, 0>, 2, 3, 0},
    {512, 2, 4, 4, 2, 256, launch_col<T, 512, 2, 4, 4, 2, 2, 3, 1>, 2, 3, 1},
    {512, 2, 8, 2, 2, 128, launch_col<T, 512, 2, 8, 2, 2, 1, 1, 0>, 1, 1, 0},
    {512, 2, 8, 2, 2, 128, launch_col<T, 512, 2, 8, 2, 2, 2, 3, 0>, 2, 3, 0},
    {512, 2, 8, 4, 2, 128, launch_col<T, 512, 2, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {512, 2, 8, 4, 2, 128, launch_col<T, 512, 2, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {1024, 1, 4, 4, 2, 256, launch_col<T, 1024, 1, 4, 4, 2, 2, 3, 0>, 2, 3, 0},
    {1024, 1, 8, 4, 1, 128, launch_col<T, 1024, 1, 8, 4, 1, 2, 3, 0>, 2, 3, 0},
    {1024, 1, 8, 4, 2, 128, launch_col<T, 1024, 1, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {1024, 1, 8, 4, 2, 128, launch_col<T, 1024, 1, 8, 4, 2, 1, 1, 1>, 1, 1, 1},
    {1024, 1, 8, 4, 2, 128, launch_col<T, 1024, 1, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {1024, 2, 4, 4, 2, 512, launch_col<T, 1024, 2, 4, 4, 2, 1, 1, 0>, 1, 1, 0},
    {1024, 2, 4, 4, 2, 512, launch_col<T, 1024, 2, 4, 4, 2, 2, 3, 0>, 2, 3, 0},
    {1024, 2, 8, 2, 2, 256, launch_col<T, 1024, 2, 8, 2, 2, 1, 1, 0>, 1, 1, 0},
    {1024, 2, 8, 2, 2, 256, launch_col<T, 1024, 2, 8, 2, 2, 2, 3, 0>, 2, 3, 0},
    {1024, 2, 8, 4, 2, 256, launch_col<T, 1024, 2, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {1024, 2, 8, 4, 2, 256, launch_col<T, 1024, 2, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {2048, 1, 8, 4, 2, 256, launch_col<T, 2048, 1, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {2048, 2, 8, 2, 2, 512, launch_col<T, 2048, 2, 8, 2, 2, 1, 1, 0>, 1, 1, 0},
    {2048, 2, 8, 4, 2, 512, launch_col<T, 2048, 2, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {4096, 1, 8, 2, 1, 512, launch_col<T, 4096, 1, 8, 2, 1, 1, 1, 0>, 1, 1, 0},
    {4096, 1, 8, 4, 1, 512, launch_col<T, 4096, 1, 8, 4, 1, 1, 1, 0>, 1, 1, 0},
};
const Schedule kSched[] = {
    {16, 32, 4, 2, 4, 3, 8, 2, 3, 2, 0, 1},
    {16, 64, 4, 2, 4, 3, 8, 2, 3, 2, 0, 1},
    {16, 128, 4, 2, 4, 3, 8, 2, 3, 2, 0, 1},
    {16, 256, 4, 2, 4, 3, 8, 2, 3, 2, 0, 1},
    {16, 512, 4, 2, 4, 3, 8, 2, 3, 2, 0, 1},
    {16, 1024, 4, 2, 4, 3, 8, 2, 3, 2, 0, 1},
    {16, 2048, 4, 2, 4, 3, 8, 1, 1, 2, 0, 1},
    {16, 4096, 4, 1, 4, 3, 8, 1, 1, 1, 0, 1},
    {32, 32, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {32, 64, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {32, 128, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {32, 256, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {32, 512, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {32, 1024, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {32, 2048, 4, 2, 8, 3, 8, 1, 1, 2, 0, 1},
    {32, 4096, 4, 1, 8, 3, 8, 1, 1, 1, 0, 1},
    {64, 32, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {64, 64, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {64, 128, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {64, 256, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {64, 512, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {64, 1024, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {64, 2048, 4, 2, 8, 3, 8, 1, 1, 2, 0, 1},
    {64, 4096, 4, 1, 8, 3, 8, 1, 1, 1, 0, 1},
    {128, 32, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {128, 64, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {128, 128, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {128, 256, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {128, 256, 4, 2, 4, 3, 4, 2, 3, 2, 0, 0},
    {128, 256, 4, 1, 8, 3, 8, 2, 3, 1, 0, 0},
    {128, 256, 2, 2, 8, 3, 8, 2, 3, 2, 0, 0},
    {128, 256, 4, 2, 8, 1, 8, 1, 1, 2, 0, 0},
    {128, 256, 4, 2, 4, 1, 4, 1, 1, 2, 0, 0},
    {128, 256, 2, 2, 4, 3, 8, 1, 1, 2, 0, 0},
    {128, 256, 4, 2, 4, 3, 4, 2, 3, 1, 0, 0},
    {128, 256, 4, 2, 8, 3, 8, 2, 3, 1, 0, 0},
    {128, 256, 4, 2, 8, 1, 8, 1, 1, 1, 0, 0},
    {128, 256, 4, 2, 4, 3, 4, 2, 3, 2, 1, 0},
    {128, 256, 4, 1, 8, 3, 8, 2, 3, 1, 1, 0},
    {128, 256, 4, 2, 8, 1, 8, 1, 1, 1, 1, 0},
    {128, 512, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {128, 1024, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {128, 2048, 4, 2, 8, 3, 8, 1, 1, 2, 0, 1},
    {128, 4096, 4, 1, 8, 3, 8, 1, 1, 1, 0, 1},
    {256, 32, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {256, 64, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {256, 128, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {256, 256, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {256, 512, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {256, 512, 4, 2, 4, 3, 4, 2, 3, 2, 0, 0},
    {256, 512, 4, 1, 8, 3, 8, 2, 3, 1, 0, 0},
    {256, 512, 2, 2, 8, 3, 8, 2, 3, 2, 0, 0},
    {256, 512, 4, 2, 8, 1, 8, 1, 1, 2, 0, 0},
    {256, 512, 4, 2, 4, 1, 4, 1, 1, 2, 0, 0},
    {256, 512, 2, 2, 4, 3, 8, 1, 1, 2, 0, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 3, 1, 0, 0},
    {256, 512, 4, 2, 8, 3, 8, 2, 3, 1, 0, 0},
    {256, 512, 4, 2, 8, 1, 8, 1, 1, 1, 0, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 3, 2, 1, 0},
    {256, 512, 4, 1, 8, 3, 8, 2, 3, 1, 1, 0},
    {256, 512, 4, 2, 8, 1, 8, 1, 1, 1, 1, 0},
    {256, 1024, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {256, 2048, 4, 2, 8, 3, 8, 1, 1, 2, 0, 1},
    {256, 4096, 4, 1, 8, 3, 8, 1, 1, 1, 0, 1},
    {512, 32, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {512, 64, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {512, 128, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {512, 256, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {512, 512, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {512, 1024, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {512, 1024, 4, 2, 4, 3, 4, 2, 3, 2, 0, 0},
    {512, 1024, 4, 1, 8, 3, 8, 2, 3, 1, 0, 0},
    {512, 1024, 2, 2, 8, 3, 8, 2, 3, 2, 0, 0},
    {512, 1024, 4, 2, 8, 1, 8, 1, 1, 2, 0, 0},
    {512, 1024, 4, 2, 4, 1, 4, 1, 1, 2, 0, 0},
    {512, 1024, 2, 2, 4, 3, 8, 1, 1, 2, 0, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 3, 1, 0, 0},
    {512, 1024, 4, 2, 8, 3, 8, 2, 3, 1, 0, 0},
    {512, 1024, 4, 2, 8, 1, 8, 1, 1, 1, 0, 0},
    {512, 1024, 4, 2, 8, 1, 8, 1, 1, 1, 1, 0},
    {512, 2048, 4, 2, 8, 3, 8, 1, 1, 2, 0, 1},
    {512, 4096, 4, 1, 8, 3, 8, 1, 1, 1, 0, 1},
    {1024, 32, 4, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {1024, 64, 4, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {1024, 128, 4, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {1024, 256, 4, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {1024, 512, 4, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {1024, 1024, 4, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {1024, 2048, 4, 2, 8, 1, 8, 1, 1, 2, 0, 1},
    {1024, 2048, 2, 2, 4, 3, 8, 1, 1, 2, 0, 0},
    {1024, 2048, 4, 2, 8, 1, 8, 1, 1, 1, 0, 0},
    {1024, 4096, 4, 1, 8, 1, 8, 1, 1, 1, 0, 1},
    {2048, 32, 2, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {2048, 64, 2, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {2048, 128, 2, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {2048, 256, 2, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {2048, 512, 2, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {2048, 1024, 2, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {2048, 2048, 2, 2, 8, 1, 8, 1, 1, 2, 0, 1},
    {2048, 4096, 2, 1, 8, 1, 8, 1, 1, 1, 0, 1},
};
}  // namespace

const RowCfg* row_cfgs_f64(int* n) { *n = (int)(sizeof(kRow) / sizeof(kRow[0])); return kRow; }
const ColCfg* col_cfgs_f64(int* n) { *n = (int)(sizeof(kCol) / sizeof(kCol[0])); return kCol; }
const Schedule* schedules_f64(int* n) { *n = (int)(sizeof(kSched) / sizeof(kSched[0])); return kSched; }
ConvLaunch to_bricks_f64() { return launch_to_bricks<T>; }
ConvLaunch from_bricks_f64() { return launch_from_bricks<T>; }

}  // namespace af
