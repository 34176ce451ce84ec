// FP32 schedule and launch tables (the kernels are instantiated in
// kernels_f32_k*.hip, compiled in parallel).
// Generated by gen_kernel_tables.py -- edit that script, not this file.
// Row kernel:    (Nx/2, RB = brick height, R, BX = brick width, G, PP)
// Column kernel: (Ny, W = columns per block, R, BY, BX = brick height/width, GF, GI, PP)
// Schedule:      (Nx/2, Ny, BY, BX, row R, row G, col R, col GF, col GI, col W, PP, default)
// PP: option bits (1 double-buffered LDS, 2 / 4 streaming column / row stores,
// 8 streaming column-pass P, Q loads, 16 integrating-factor table at radix 16)
#include "kernels_inst.hpp"

namespace af {
extern template hipError_t launch_row<float, 16, 4, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 32, 4, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 64, 4, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 2, 4, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 2, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 2, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 2, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 2, 8, 4, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 2, 16, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 4, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 4, 4, 2, 3, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 4, 4, 2, 3, 5>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 4, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 4, 8, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 4, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 4, 8, 2, 3, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 4, 8, 4, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 4, 8, 4, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 4, 16, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 4, 16, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 4, 16, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 128, 4, 16, 4, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 2, 4, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 2, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 2, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 2, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 2, 8, 4, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 2, 16, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 4, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 4, 2, 1, 5>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 4, 2, 3, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 4, 2, 3, 4>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 4, 2, 3, 5>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 8, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 8, 2, 3, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 8, 4, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 8, 4, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 8, 4, 3, 4>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 16, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 16, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 16, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 256, 4, 16, 4, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 2, 4, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 2, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 2, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 2, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 2, 8, 4, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 2, 16, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 4, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 4, 2, 1, 5>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 4, 2, 3, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 4, 2, 3, 5>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 8, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 8, 2, 1, 4>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 8, 2, 3, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 8, 4, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 8, 4, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 16, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 16, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 16, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 512, 4, 16, 4, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 2, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 2, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 2, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 2, 16, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 4, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 4, 8, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 4, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 4, 8, 4, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 4, 8, 4, 1, 4>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 4, 8, 4, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 4, 16, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 4, 16, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 4, 16, 2, 1, 4>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 4, 16, 2, 1, 5>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 4, 16, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 1024, 4, 16, 4, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 2048, 2, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 2048, 2, 16, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 2048, 4, 16, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 2048, 4, 16, 4, 1, 0>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_row<float, 2048, 4, 16, 4, 1, 4>(const void*, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 32, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 32, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 64, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 64, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 128, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 128, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 1, 4, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 1, 4, 4, 2, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 1, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 1, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 1, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 1, 16, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 1, 16, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 4, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 4, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 4, 4, 2, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 4, 4, 2, 2, 3, 3>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 4, 4, 2, 2, 3, 9>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 4, 4, 2, 2, 3, 11>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 8, 4, 2, 1, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 8, 4, 2, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 8, 4, 4, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 16, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 16, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 2, 16, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 4, 8, 2, 4, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 4, 8, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 4, 8, 4, 4, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 256, 4, 16, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 4, 4, 2, 1, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 4, 4, 2, 1, 1, 3>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 4, 4, 2, 1, 1, 11>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 4, 4, 2, 2, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 4, 4, 2, 2, 1, 3>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 4, 4, 2, 2, 1, 11>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 4, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 4, 4, 2, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 4, 4, 2, 2, 3, 2>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 4, 4, 2, 2, 3, 3>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 4, 4, 2, 2, 3, 8>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 4, 4, 2, 2, 3, 10>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 4, 4, 2, 2, 3, 11>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 16, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 1, 16, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 2, 4, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 2, 4, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 2, 4, 4, 2, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 2, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 2, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 2, 8, 4, 2, 1, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 2, 8, 4, 2, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 2, 8, 4, 4, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 2, 16, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 2, 16, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 2, 16, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 4, 8, 2, 4, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 4, 8, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 4, 8, 4, 4, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 4, 8, 4, 4, 2, 3, 2>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 4, 8, 4, 4, 2, 3, 8>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 4, 8, 4, 4, 2, 3, 10>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 512, 4, 16, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 1, 4, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 1, 4, 4, 2, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 1, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 1, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 1, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 1, 16, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 1, 16, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 4, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 4, 4, 2, 1, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 4, 4, 2, 1, 1, 3>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 4, 4, 2, 1, 1, 11>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 4, 4, 2, 2, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 4, 4, 2, 2, 1, 3>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 4, 4, 2, 2, 1, 11>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 4, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 4, 4, 2, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 4, 4, 2, 2, 3, 3>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 4, 4, 2, 2, 3, 9>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 4, 4, 2, 2, 3, 11>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 8, 4, 2, 1, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 8, 4, 2, 1, 1, 2>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 8, 4, 2, 1, 1, 8>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 8, 4, 2, 1, 1, 10>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 8, 4, 2, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 8, 4, 4, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 16, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 16, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 2, 16, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 4, 8, 2, 4, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 4, 8, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 4, 8, 4, 4, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 1024, 4, 16, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 1, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 1, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 1, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 1, 16, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 1, 16, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 8, 4, 2, 1, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 8, 4, 4, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 2>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 3>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 8>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 9>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 10>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 11>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 16>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 17>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 18>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 19>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 24>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 25>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 26>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 27>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 2, 16, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 4, 16, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 4, 16, 4, 4, 1, 1, 2>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 4, 16, 4, 4, 1, 1, 8>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 4, 16, 4, 4, 1, 1, 10>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 4, 16, 4, 4, 1, 1, 16>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 4, 16, 4, 4, 1, 1, 18>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 4, 16, 4, 4, 1, 1, 24>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 2048, 4, 16, 4, 4, 1, 1, 26>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 4096, 1, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 4096, 1, 16, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 4096, 1, 16, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 4096, 2, 16, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 4096, 2, 16, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 4096, 2, 16, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 4096, 2, 16, 4, 4, 1, 1, 2>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 4096, 2, 16, 4, 4, 1, 1, 8>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 4096, 2, 16, 4, 4, 1, 1, 10>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 4096, 2, 16, 4, 4, 1, 1, 16>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 4096, 2, 16, 4, 4, 1, 1, 18>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 4096, 2, 16, 4, 4, 1, 1, 24>(const void*, int, int, int, const void*, hipStream_t);
extern template hipError_t launch_col<float, 4096, 2, 16, 4, 4, 1, 1, 26>(const void*, int, int, int, const void*, hipStream_t);
namespace {
using T = float;
const RowCfg kRow[] = {
    {16, 4, 4, 2, 16, launch_row<T, 16, 4, 4, 2, 3, 0>, 3, 0},
    {32, 4, 8, 2, 16, launch_row<T, 32, 4, 8, 2, 3, 0>, 3, 0},
    {64, 4, 8, 2, 32, launch_row<T, 64, 4, 8, 2, 3, 0>, 3, 0},
    {128, 2, 4, 2, 64, launch_row<T, 128, 2, 4, 2, 1, 0>, 1, 0},
    {128, 2, 4, 2, 64, launch_row<T, 128, 2, 4, 2, 3, 0>, 3, 0},
    {128, 2, 8, 2, 32, launch_row<T, 128, 2, 8, 2, 1, 0>, 1, 0},
    {128, 2, 8, 2, 32, launch_row<T, 128, 2, 8, 2, 3, 0>, 3, 0},
    {128, 2, 8, 4, 32, launch_row<T, 128, 2, 8, 4, 3, 0>, 3, 0},
    {128, 2, 16, 2, 16, launch_row<T, 128, 2, 16, 2, 1, 0>, 1, 0},
    {128, 4, 4, 2, 128, launch_row<T, 128, 4, 4, 2, 3, 0>, 3, 0},
    {128, 4, 4, 2, 128, launch_row<T, 128, 4, 4, 2, 3, 1>, 3, 1},
    {128, 4, 4, 2, 128, launch_row<T, 128, 4, 4, 2, 3, 5>, 3, 5},
    {128, 4, 8, 2, 64, launch_row<T, 128, 4, 8, 2, 1, 0>, 1, 0},
    {128, 4, 8, 2, 64, launch_row<T, 128, 4, 8, 2, 1, 1>, 1, 1},
    {128, 4, 8, 2, 64, launch_row<T, 128, 4, 8, 2, 3, 0>, 3, 0},
    {128, 4, 8, 2, 64, launch_row<T, 128, 4, 8, 2, 3, 1>, 3, 1},
    {128, 4, 8, 4, 64, launch_row<T, 128, 4, 8, 4, 1, 0>, 1, 0},
    {128, 4, 8, 4, 64, launch_row<T, 128, 4, 8, 4, 3, 0>, 3, 0},
    {128, 4, 16, 2, 32, launch_row<T, 128, 4, 16, 2, 1, 0>, 1, 0},
    {128, 4, 16, 2, 32, launch_row<T, 128, 4, 16, 2, 1, 1>, 1, 1},
    {128, 4, 16, 2, 32, launch_row<T, 128, 4, 16, 2, 3, 0>, 3, 0},
    {128, 4, 16, 4, 32, launch_row<T, 128, 4, 16, 4, 1, 0>, 1, 0},
    {256, 2, 4, 2, 128, launch_row<T, 256, 2, 4, 2, 1, 0>, 1, 0},
    {256, 2, 4, 2, 128, launch_row<T, 256, 2, 4, 2, 3, 0>, 3, 0},
    {256, 2, 8, 2, 64, launch_row<T, 256, 2, 8, 2, 1, 0>, 1, 0},
    {256, 2, 8, 2, 64, launch_row<T, 256, 2, 8, 2, 3, 0>, 3, 0},
    {256, 2, 8, 4, 64, launch_row<T, 256, 2, 8, 4, 3, 0>, 3, 0},
    {256, 2, 16, 2, 32, launch_row<T, 256, 2, 16, 2, 1, 0>, 1, 0},
    {256, 4, 4, 2, 256, launch_row<T, 256, 4, 4, 2, 1, 1>, 1, 1},
    {256, 4, 4, 2, 256, launch_row<T, 256, 4, 4, 2, 1, 5>, 1, 5},
    {256, 4, 4, 2, 256, launch_row<T, 256, 4, 4, 2, 3, 0>, 3, 0},
    {256, 4, 4, 2, 256, launch_row<T, 256, 4, 4, 2, 3, 1>, 3, 1},
    {256, 4, 4, 2, 256, launch_row<T, 256, 4, 4, 2, 3, 4>, 3, 4},
    {256, 4, 4, 2, 256, launch_row<T, 256, 4, 4, 2, 3, 5>, 3, 5},
    {256, 4, 8, 2, 128, launch_row<T, 256, 4, 8, 2, 1, 0>, 1, 0},
    {256, 4, 8, 2, 128, launch_row<T, 256, 4, 8, 2, 1, 1>, 1, 1},
    {256, 4, 8, 2, 128, launch_row<T, 256, 4, 8, 2, 3, 0>, 3, 0},
    {256, 4, 8, 2, 128, launch_row<T, 256, 4, 8, 2, 3, 1>, 3, 1},
    {256, 4, 8, 4, 128, launch_row<T, 256, 4, 8, 4, 1, 0>, 1, 0},
    {256, 4, 8, 4, 128, launch_row<T, 256, 4, 8, 4, 3, 0>, 3, 0},
    {256, 4, 8, 4, 128, launch_row<T, 256, 4, 8, 4, 3, 4>, 3, 4},
    {256, 4, 16, 2, 64, launch_row<T, 256, 4, 16, 2, 1, 0>, 1, 0},
    {256, 4, 16, 2, 64, launch_row<T, 256, 4, 16, 2, 1, 1>, 1, 1},
    {256, 4, 16, 2, 64, launch_row<T, 256, 4, 16, 2, 3, 0>, 3, 0},
    {256, 4, 16, 4, 64, launch_row<T, 256, 4, 16, 4, 1, 0>, 1, 0},
    {512, 2, 4, 2, 256, launch_row<T, 512, 2, 4, 2, 1, 0>, 1, 0},
    {512, 2, 4, 2, 256, launch_row<T, 512, 2, 4, 2, 3, 0>, 3, 0},
    {512, 2, 8, 2, 128, launch_row<T, 512, 2, 8, 2, 1, 0>, 1, 0},
    {512, 2, 8, 2, 128, launch_row<T, 512, 2, 8, 2, 3, 0>, 3, 0},
    {512, 2, 8, 4, 128, launch_row<T, 512, 2, 8, 4, 3, 0>, 3, 0},
    {512, 2, 16, 2, 64, launch_row<T, 512, 2, 16, 2, 1, 0>, 1, 0},
    {512, 4, 4, 2, 512, launch_row<T, 512, 4, 4, 2, 1, 1>, 1, 1},
    {512, 4, 4, 2, 512, launch_row<T, 512, 4, 4, 2, 1, 5>, 1, 5},
    {512, 4, 4, 2, 512, launch_row<T, 512, 4, 4, 2, 3, 0>, 3, 0},
    {512, 4, 4, 2, 512, launch_row<T, 512, 4, 4, 2, 3, 1>, 3, 1},
    {512, 4, 4, 2, 512, launch_row<T, 512, 4, 4, 2, 3, 5>, 3, 5},
    {512, 4, 8, 2, 256, launch_row<T, 512, 4, 8, 2, 1, 0>, 1, 0},
    {512, 4, 8, 2, 256, launch_row<T, 512, 4, 8, 2, 1, 1>, 1, 1},
    {512, 4, 8, 2, 256, launch_row<T, 512, 4, 8, 2, 1, 4>, 1, 4},
    {512, 4, 8, 2, 256, launch_row<T, 512, 4, 8, 2, 3, 0>, 3, 0},
    {512, 4, 8, 2, 256, launch_row<T, 512, 4, 8, 2, 3, 1>, 3, 1},
    {512, 4, 8, 4, 256, launch_row<T, 512, 4, 8, 4, 1, 0>, 1, 0},
    {512, 4, 8, 4, 256, launch_row<T, 512, 4, 8, 4, 3, 0>, 3, 0},
    {512, 4, 16, 2, 128, launch_row<T, 512, 4, 16, 2, 1, 0>, 1, 0},
    {512, 4, 16, 2, 128, launch_row<T, 512, 4, 16, 2, 1, 1>, 1, 1},
    {512, 4, 16, 2, 128, launch_row<T, 512, 4, 16, 2, 3, 0>, 3, 0},
    {512, 4, 16, 4, 128, launch_row<T, 512, 4, 16, 4, 1, 0>, 1, 0},
    {1024, 2, 4, 2, 512, launch_row<T, 1024, 2, 4, 2, 3, 0>, 3, 0},
    {1024, 2, 8, 2, 256, launch_row<T, 1024, 2, 8, 2, 1, 0>, 1, 0},
    {1024, 2, 8, 2, 256, launch_row<T, 1024, 2, 8, 2, 3, 0>, 3, 0},
    {1024, 2, 16, 2, 128, launch_row<T, 1024, 2, 16, 2, 1, 0>, 1, 0},
    {1024, 4, 8, 2, 512, launch_row<T, 1024, 4, 8, 2, 1, 0>, 1, 0},
    {1024, 4, 8, 2, 512, launch_row<T, 1024, 4, 8, 2, 1, 1>, 1, 1},
    {1024, 4, 8, 2, 512, launch_row<T, 1024, 4, 8, 2, 3, 0>, 3, 0},
    {1024, 4, 8, 4, 512, launch_row<T, 1024, 4, 8, 4, 1, 0>, 1, 0},
    {1024, 4, 8, 4, 512, launch_row<T, 1024, 4, 8, 4, 1, 4>, 1, 4},
    {1024, 4, 8, 4, 512, launch_row<T, 1024, 4, 8, 4, 3, 0>, 3, 0},
    {1024, 4, 16, 2, 256, launch_row<T, 1024, 4, 16, 2, 1, 0>, 1, 0},
    {1024, 4, 16, 2, 256, launch_row<T, 1024, 4, 16, 2, 1, 1>, 1, 1},
    {1024, 4, 16, 2, 256, launch_row<T, 1024, 4, 16, 2, 1, 4>, 1, 4},
    {1024, 4, 16, 2, 256, launch_row<T, 1024, 4, 16, 2, 1, 5>, 1, 5},
    {1024, 4, 16, 2, 256, launch_row<T, 1024, 4, 16, 2, 3, 0>, 3, 0},
    {1024, 4, 16, 4, 256, launch_row<T, 1024, 4, 16, 4, 1, 0>, 1, 0},
    {2048, 2, 8, 2, 512, launch_row<T, 2048, 2, 8, 2, 1, 0>, 1, 0},
    {2048, 2, 16, 2, 256, launch_row<T, 2048, 2, 16, 2, 1, 0>, 1, 0},
    {2048, 4, 16, 2, 512, launch_row<T, 2048, 4, 16, 2, 1, 0>, 1, 0},
    {2048, 4, 16, 4, 512, launch_row<T, 2048, 4, 16, 4, 1, 0>, 1, 0},
    {2048, 4, 16, 4, 512, launch_row<T, 2048, 4, 16, 4, 1, 4>, 1, 4},
};
const ColCfg kCol[] = {
    {32, 2, 8, 2, 2, 8, launch_col<T, 32, 2, 8, 2, 2, 2, 3, 0>, 2, 3, 0},
    {32, 2, 8, 4, 2, 8, launch_col<T, 32, 2, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {64, 2, 8, 2, 2, 16, launch_col<T, 64, 2, 8, 2, 2, 2, 3, 0>, 2, 3, 0},
    {64, 2, 8, 4, 2, 16, launch_col<T, 64, 2, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {128, 2, 8, 2, 2, 32, launch_col<T, 128, 2, 8, 2, 2, 2, 3, 0>, 2, 3, 0},
    {128, 2, 8, 4, 2, 32, launch_col<T, 128, 2, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {256, 1, 4, 4, 2, 64, launch_col<T, 256, 1, 4, 4, 2, 2, 3, 0>, 2, 3, 0},
    {256, 1, 4, 4, 2, 64, launch_col<T, 256, 1, 4, 4, 2, 2, 3, 1>, 2, 3, 1},
    {256, 1, 8, 2, 2, 32, launch_col<T, 256, 1, 8, 2, 2, 1, 1, 0>, 1, 1, 0},
    {256, 1, 8, 4, 2, 32, launch_col<T, 256, 1, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {256, 1, 8, 4, 2, 32, launch_col<T, 256, 1, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {256, 1, 16, 2, 2, 16, launch_col<T, 256, 1, 16, 2, 2, 1, 1, 0>, 1, 1, 0},
    {256, 1, 16, 4, 2, 16, launch_col<T, 256, 1, 16, 4, 2, 1, 1, 0>, 1, 1, 0},
    {256, 2, 4, 2, 2, 128, launch_col<T, 256, 2, 4, 2, 2, 1, 1, 0>, 1, 1, 0},
    {256, 2, 4, 4, 2, 128, launch_col<T, 256, 2, 4, 4, 2, 2, 3, 0>, 2, 3, 0},
    {256, 2, 4, 4, 2, 128, launch_col<T, 256, 2, 4, 4, 2, 2, 3, 1>, 2, 3, 1},
    {256, 2, 4, 4, 2, 128, launch_col<T, 256, 2, 4, 4, 2, 2, 3, 3>, 2, 3, 3},
    {256, 2, 4, 4, 2, 128, launch_col<T, 256, 2, 4, 4, 2, 2, 3, 9>, 2, 3, 9},
    {256, 2, 4, 4, 2, 128, launch_col<T, 256, 2, 4, 4, 2, 2, 3, 11>, 2, 3, 11},
    {256, 2, 8, 2, 2, 64, launch_col<T, 256, 2, 8, 2, 2, 1, 1, 0>, 1, 1, 0},
    {256, 2, 8, 2, 2, 64, launch_col<T, 256, 2, 8, 2, 2, 2, 3, 0>, 2, 3, 0},
    {256, 2, 8, 4, 2, 64, launch_col<T, 256, 2, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {256, 2, 8, 4, 2, 64, launch_col<T, 256, 2, 8, 4, 2, 1, 1, 1>, 1, 1, 1},
    {256, 2, 8, 4, 2, 64, launch_col<T, 256, 2, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {256, 2, 8, 4, 2, 64, launch_col<T, 256, 2, 8, 4, 2, 2, 3, 1>, 2, 3, 1},
    {256, 2, 8, 4, 4, 64, launch_col<T, 256, 2, 8, 4, 4, 2, 3, 0>, 2, 3, 0},
    {256, 2, 16, 2, 2, 32, launch_col<T, 256, 2, 16, 2, 2, 1, 1, 0>, 1, 1, 0},
    {256, 2, 16, 4, 2, 32, launch_col<T, 256, 2, 16, 4, 2, 1, 1, 0>, 1, 1, 0},
    {256, 2, 16, 4, 4, 32, launch_col<T, 256, 2, 16, 4, 4, 1, 1, 0>, 1, 1, 0},
    {256, 4, 8, 2, 4, 128, launch_col<T, 256, 4, 8, 2, 4, 2, 3, 0>, 2, 3, 0},
    {256, 4, 8, 4, 4, 128, launch_col<T, 256, 4, 8, 4, 4, 1, 1, 0>, 1, 1, 0},
    {256, 4, 8, 4, 4, 128, launch_col<T, 256, 4, 8, 4, 4, 2, 3, 0>, 2, 3, 0},
    {256, 4, 16, 4, 4, 64, launch_col<T, 256, 4, 16, 4, 4, 1, 1, 0>, 1, 1, 0},
    {512, 1, 4, 4, 2, 128, launch_col<T, 512, 1, 4, 4, 2, 1, 1, 1>, 1, 1, 1},
    {512, 1, 4, 4, 2, 128, launch_col<T, 512, 1, 4, 4, 2, 1, 1, 3>, 1, 1, 3},
    {512, 1, 4, 4, 2, 128, launch_col<T, 512, 1, 4, 4, 2, 1, 1, 11>, 1, 1, 11},
    {512, 1, 4, 4, 2, 128, launch_col<T, 512, 1, 4, 4, 2, 2, 1, 1>, 2, 1, 1},
    {512, 1, 4, 4, 2, 128, launch_col<T, 512, 1, 4, 4, 2, 2, 1, 3>, 2, 1, 3},
    {512, 1, 4, 4, 2, 128, launch_col<T, 512, 1, 4, 4, 2, 2, 1, 11>, 2, 1, 11},
    {512, 1, 4, 4, 2, 128, launch_col<T, 512, 1, 4, 4, 2, 2, 3, 0>, 2, 3, 0},
    {512, 1, 4, 4, 2, 128, launch_col<T, 512, 1, 4, 4, 2, 2, 3, 1>, 2, 3, 1},
    {512, 1, 4, 4, 2, 128, launch_col<T, 512, 1, 4, 4, 2, 2, 3, 2>, 2, 3, 2},
    {512, 1, 4, 4, 2, 128, launch_col<T, 512, 1, 4, 4, 2, 2, 3, 3>, 2, 3, 3},
    {512, 1, 4, 4, 2, 128, launch_col<T, 512, 1, 4, 4, 2, 2, 3, 8>, 2, 3, 8},
    {512, 1, 4, 4, 2, 128, launch_col<T, 512, 1, 4, 4, 2, 2, 3, 10>, 2, 3, 10},
    {512, 1, 4, 4, 2, 128, launch_col<T, 512, 1, 4, 4, 2, 2, 3, 11>, 2, 3, 11},
    {512, 1, 8, 2, 2, 64, launch_col<T, 512, 1, 8, 2, 2, 1, 1, 0>, 1, 1, 0},
    {512, 1, 8, 4, 2, 64, launch_col<T, 512, 1, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {512, 1, 8, 4, 2, 64, launch_col<T, 512, 1, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {512, 1, 16, 2, 2, 32, launch_col<T, 512, 1, 16, 2, 2, 1, 1, 0>, 1, 1, 0},
    {512, 1, 16, 4, 2, 32, launch_col<T, 512, 1, 16, 4, 2, 1, 1, 0>, 1, 1, 0},
    {512, 2, 4, 2, 2, 256, launch_col<T, 512, 2, 4, 2, 2, 1, 1, 0>, 1, 1, 0},
    {512, 2, 4, 4, 2, 256, launch_col<T, 512, 2, 4, 4, 2, 2, 3, 0>, 2, 3, 0},
    {512, 2, 4, 4, 2, 256, launch_col<T, 512, 2, 4, 4, 2, 2, 3, 1>, 2, 3, 1},
    {512, 2, 8, 2, 2, 128, launch_col<T, 512, 2, 8, 2, 2, 1, 1, 0>, 1, 1, 0},
    {512, 2, 8, 2, 2, 128, launch_col<T, 512, 2, 8, 2, 2, 2, 3, 0>, 2, 3, 0},
    {512, 2, 8, 4, 2, 128, launch_col<T, 512, 2, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {512, 2, 8, 4, 2, 128, launch_col<T, 512, 2, 8, 4, 2, 1, 1, 1>, 1, 1, 1},
    {512, 2, 8, 4, 2, 128, launch_col<T, 512, 2, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {512, 2, 8, 4, 2, 128, launch_col<T, 512, 2, 8, 4, 2, 2, 3, 1>, 2, 3, 1},
    {512, 2, 8, 4, 4, 128, launch_col<T, 512, 2, 8, 4, 4, 2, 3, 0>, 2, 3, 0},
    {512, 2, 16, 2, 2, 64, launch_col<T, 512, 2, 16, 2, 2, 1, 1, 0>, 1, 1, 0},
    {512, 2, 16, 4, 2, 64, launch_col<T, 512, 2, 16, 4, 2, 1, 1, 0>, 1, 1, 0},
    {512, 2, 16, 4, 4, 64, launch_col<T, 512, 2, 16, 4, 4, 1, 1, 0>, 1, 1, 0},
    {512, 4, 8, 2, 4, 256, launch_col<T, 512, 4, 8, 2, 4, 2, 3, 0>, 2, 3, 0},
    {512, 4, 8, 4, 4, 256, launch_col<T, 512, 4, 8, 4, 4, 1, 1, 0>, 1, 1, 0},
    {512, 4, 8, 4, 4, 256, launch_col<T, 512, 4, 8, 4, 4, 2, 3, 0>, 2, 3, 0},
    {512, 4, 8, 4, 4, 256, launch_col<T, 512, 4, 8, 4, 4, 2, 3, 2>, 2, 3, 2},
    {512, 4, 8, 4, 4, 256, launch_col<T, 512, 4, 8, 4, 4, 2, 3, 8>, 2, 3, 8},
    {512, 4, 8, 4, 4, 256, launch_col<T, 512, 4, 8, 4, 4, 2, 3, 10>, 2, 3, 10},
    {512, 4, 16, 4, 4, 128, launch_col<T, 512, 4, 16, 4, 4, 1, 1, 0>, 1, 1, 0},
    {1024, 1, 4, 4, 2, 256, launch_col<T, 1024, 1, 4, 4, 2, 2, 3, 0>, 2, 3, 0},
    {1024, 1, 4, 4, 2, 256, launch_col<T, 1024, 1, 4, 4, 2, 2, 3, 1>, 2, 3, 1},
    {1024, 1, 8, 2, 2, 128, launch_col<T, 1024, 1, 8, 2, 2, 1, 1, 0>, 1, 1, 0},
    {1024, 1, 8, 4, 2, 128, launch_col<T, 1024, 1, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {1024, 1, 8, 4, 2, 128, launch_col<T, 1024, 1, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {1024, 1, 16, 2, 2, 64, launch_col<T, 1024, 1, 16, 2, 2, 1, 1, 0>, 1, 1, 0},
    {1024, 1, 16, 4, 2, 64, launch_col<T, 1024, 1, 16, 4, 2, 1, 1, 0>, 1, 1, 0},
    {1024, 2, 4, 2, 2, 512, launch_col<T, 1024, 2, 4, 2, 2, 1, 1, 0>, 1, 1, 0},
    {1024, 2, 4, 4, 2, 512, launch_col<T, 1024, 2, 4, 4, 2, 1, 1, 1>, 1, 1, 1},
    {1024, 2, 4, 4, 2, 512, launch_col<T, 1024, 2, 4, 4, 2, 1, 1, 3>, 1, 1, 3},
    {1024, 2, 4, 4, 2, 512, launch_col<T, 1024, 2, 4, 4, 2, 1, 1, 11>, 1, 1, 11},
    {1024, 2, 4, 4, 2, 512, launch_col<T, 1024, 2, 4, 4, 2, 2, 1, 1>, 2, 1, 1},
    {1024, 2, 4, 4, 2, 512, launch_col<T, 1024, 2, 4, 4, 2, 2, 1, 3>, 2, 1, 3},
    {1024, 2, 4, 4, 2, 512, launch_col<T, 1024, 2, 4, 4, 2, 2, 1, 11>, 2, 1, 11},
    {1024, 2, 4, 4, 2, 512, launch_col<T, 1024, 2, 4, 4, 2, 2, 3, 0>, 2, 3, 0},
    {1024, 2, 4, 4, 2, 512, launch_col<T, 1024, 2, 4, 4, 2, 2, 3, 1>, 2, 3, 1},
    {1024, 2, 4, 4, 2, 512, launch_col<T, 1024, 2, 4, 4, 2, 2, 3, 3>, 2, 3, 3},
    {1024, 2, 4, 4, 2, 512, launch_col<T, 1024, 2, 4, 4, 2, 2, 3, 9>, 2, 3, 9},
    {1024, 2, 4, 4, 2, 512, launch_col<T, 1024, 2, 4, 4, 2, 2, 3, 11>, 2, 3, 11},
    {1024, 2, 8, 2, 2, 256, launch_col<T, 1024, 2, 8, 2, 2, 1, 1, 0>, 1, 1, 0},
    {1024, 2, 8, 2, 2, 256, launch_col<T, 1024, 2, 8, 2, 2, 2, 3, 0>, 2, 3, 0},
    {1024, 2, 8, 4, 2, 256, launch_col<T, 1024, 2, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {1024, 2, 8, 4, 2, 256, launch_col<T, 1024, 2, 8, 4, 2, 1, 1, 1>, 1, 1, 1},
    {1024, 2, 8, 4, 2, 256, launch_col<T, 1024, 2, 8, 4, 2, 1, 1, 2>, 1, 1, 2},
    {1024, 2, 8, 4, 2, 256, launch_col<T, 1024, 2, 8, 4, 2, 1, 1, 8>, 1, 1, 8},
    {1024, 2, 8, 4, 2, 256, launch_col<T, 1024, 2, 8, 4, 2, 1, 1, 10>, 1, 1, 10},
    {1024, 2, 8, 4, 2, 256, launch_col<T, 1024, 2, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {1024, 2, 8, 4, 2, 256, launch_col<T, 1024, 2, 8, 4, 2, 2, 3, 1>, 2, 3, 1},
    {1024, 2, 8, 4, 4, 256, launch_col<T, 1024, 2, 8, 4, 4, 2, 3, 0>, 2, 3, 0},
    {1024, 2, 16, 2, 2, 128, launch_col<T, 1024, 2, 16, 2, 2, 1, 1, 0>, 1, 1, 0},
    {1024, 2, 16, 4, 2, 128, launch_col<T, 1024, 2, 16, 4, 2, 1, 1, 0>, 1, 1, 0},
    {1024, 2, 16, 4, 4, 128, launch_col<T, 1024, 2, 16, 4, 4, 1, 1, 0>, 1, 1, 0},
    {1024, 4, 8, 2, 4, 512, launch_col<T, 1024, 4, 8, 2, 4, 2, 3, 0>, 2, 3, 0},
    {1024, 4, 8, 4, 4, 512, launch_col<T, 1024, 4, 8, 4, 4, 1, 1, 0>, 1, 1, 0},
    {1024, 4, 8, 4, 4, 512, launch_col<T, 1024, 4, 8, 4, 4, 2, 3, 0>, 2, 3, 0},
    {1024, 4, 16, 4, 4, 256, launch_col<T, 1024, 4, 16, 4, 4, 1, 1, 0>, 1, 1, 0},
    {2048, 1, 8, 2, 2, 256, launch_col<T, 2048, 1, 8, 2, 2, 1, 1, 0>, 1, 1, 0},
    {2048, 1, 8, 4, 2, 256, launch_col<T, 2048, 1, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {2048, 1, 8, 4, 2, 256, launch_col<T, 2048, 1, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {2048, 1, 16, 2, 2, 128, launch_col<T, 2048, 1, 16, 2, 2, 1, 1, 0>, 1, 1, 0},
    {2048, 1, 16, 4, 2, 128, launch_col<T, 2048, 1, 16, 4, 2, 1, 1, 0>, 1, 1, 0},
    {2048, 2, 8, 2, 2, 512, launch_col<T, 2048, 2, 8, 2, 2, 1, 1, 0>, 1, 1, 0},
    {2048, 2, 8, 2, 2, 512, launch_col<T, 2048, 2, 8, 2, 2, 2, 3, 0>, 2, 3, 0},
    {2048, 2, 8, 4, 2, 512, launch_col<T, 2048, 2, 8, 4, 2, 1, 1, 0>, 1, 1, 0},
    {2048, 2, 8, 4, 2, 512, launch_col<T, 2048, 2, 8, 4, 2, 1, 1, 1>, 1, 1, 1},
    {2048, 2, 8, 4, 2, 512, launch_col<T, 2048, 2, 8, 4, 2, 2, 3, 0>, 2, 3, 0},
    {2048, 2, 8, 4, 4, 512, launch_col<T, 2048, 2, 8, 4, 4, 2, 3, 0>, 2, 3, 0},
    {2048, 2, 16, 2, 2, 256, launch_col<T, 2048, 2, 16, 2, 2, 1, 1, 0>, 1, 1, 0},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 0>, 1, 1, 0},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 1>, 1, 1, 1},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 2>, 1, 1, 2},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 3>, 1, 1, 3},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 8>, 1, 1, 8},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 9>, 1, 1, 9},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 10>, 1, 1, 10},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 11>, 1, 1, 11},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 16>, 1, 1, 16},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 17>, 1, 1, 17},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 18>, 1, 1, 18},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 19>, 1, 1, 19},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 24>, 1, 1, 24},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 25>, 1, 1, 25},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 26>, 1, 1, 26},
    {2048, 2, 16, 4, 2, 256, launch_col<T, 2048, 2, 16, 4, 2, 1, 1, 27>, 1, 1, 27},
    {2048, 2, 16, 4, 4, 256, launch_col<T, 2048, 2, 16, 4, 4, 1, 1, 0>, 1, 1, 0},
    {2048, 4, 16, 4, 4, 512, launch_col<T, 2048, 4, 16, 4, 4, 1, 1, 0>, 1, 1, 0},
    {2048, 4, 16, 4, 4, 512, launch_col<T, 2048, 4, 16, 4, 4, 1, 1, 2>, 1, 1, 2},
    {2048, 4, 16, 4, 4, 512, launch_col<T, 2048, 4, 16, 4, 4, 1, 1, 8>, 1, 1, 8},
    {2048, 4, 16, 4, 4, 512, launch_col<T, 2048, 4, 16, 4, 4, 1, 1, 10>, 1, 1, 10},
    {2048, 4, 16, 4, 4, 512, launch_col<T, 2048, 4, 16, 4, 4, 1, 1, 16>, 1, 1, 16},
    {2048, 4, 16, 4, 4, 512, launch_col<T, 2048, 4, 16, 4, 4, 1, 1, 18>, 1, 1, 18},
    {2048, 4, 16, 4, 4, 512, launch_col<T, 2048, 4, 16, 4, 4, 1, 1, 24>, 1, 1, 24},
    {2048, 4, 16, 4, 4, 512, launch_col<T, 2048, 4, 16, 4, 4, 1, 1, 26>, 1, 1, 26},
    {4096, 1, 8, 2, 2, 512, launch_col<T, 4096, 1, 8, 2, 2, 1, 1, 0>, 1, 1, 0},
    {4096, 1, 16, 2, 2, 256, launch_col<T, 4096, 1, 16, 2, 2, 1, 1, 0>, 1, 1, 0},
    {4096, 1, 16, 4, 2, 256, launch_col<T, 4096, 1, 16, 4, 2, 1, 1, 0>, 1, 1, 0},
    {4096, 2, 16, 2, 2, 512, launch_col<T, 4096, 2, 16, 2, 2, 1, 1, 0>, 1, 1, 0},
    {4096, 2, 16, 4, 2, 512, launch_col<T, 4096, 2, 16, 4, 2, 1, 1, 0>, 1, 1, 0},
    {4096, 2, 16, 4, 4, 512, launch_col<T, 4096, 2, 16, 4, 4, 1, 1, 0>, 1, 1, 0},
    {4096, 2, 16, 4, 4, 512, launch_col<T, 4096, 2, 16, 4, 4, 1, 1, 2>, 1, 1, 2},
    {4096, 2, 16, 4, 4, 512, launch_col<T, 4096, 2, 16, 4, 4, 1, 1, 8>, 1, 1, 8},
    {4096, 2, 16, 4, 4, 512, launch_col<T, 4096, 2, 16, 4, 4, 1, 1, 10>, 1, 1, 10},
    {4096, 2, 16, 4, 4, 512, launch_col<T, 4096, 2, 16, 4, 4, 1, 1, 16>, 1, 1, 16},
    {4096, 2, 16, 4, 4, 512, launch_col<T, 4096, 2, 16, 4, 4, 1, 1, 18>, 1, 1, 18},
    {4096, 2, 16, 4, 4, 512, launch_col<T, 4096, 2, 16, 4, 4, 1, 1, 24>, 1, 1, 24},
    {4096, 2, 16, 4, 4, 512, launch_col<T, 4096, 2, 16, 4, 4, 1, 1, 26>, 1, 1, 26},
};
const Schedule kSched[] = {
    {16, 32, 4, 2, 4, 3, 8, 2, 3, 2, 0, 1},
    {16, 64, 4, 2, 4, 3, 8, 2, 3, 2, 0, 1},
    {16, 128, 4, 2, 4, 3, 8, 2, 3, 2, 0, 1},
    {16, 256, 4, 2, 4, 3, 8, 2, 3, 2, 0, 1},
    {16, 512, 4, 2, 4, 3, 8, 2, 3, 2, 0, 1},
    {16, 1024, 4, 2, 4, 3, 8, 2, 3, 2, 0, 1},
    {16, 2048, 4, 2, 4, 3, 8, 2, 3, 2, 0, 1},
    {16, 4096, 4, 2, 4, 3, 16, 1, 1, 2, 0, 1},
    {32, 32, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {32, 64, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {32, 128, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {32, 256, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {32, 512, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {32, 1024, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {32, 2048, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {32, 4096, 4, 2, 8, 3, 16, 1, 1, 2, 0, 1},
    {64, 32, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {64, 64, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {64, 128, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {64, 256, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {64, 512, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {64, 1024, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {64, 2048, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {64, 4096, 4, 2, 8, 3, 16, 1, 1, 2, 0, 1},
    {128, 32, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {128, 64, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {128, 128, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {128, 256, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {128, 256, 4, 2, 4, 3, 4, 2, 3, 2, 0, 0},
    {128, 256, 4, 4, 8, 3, 8, 2, 3, 4, 0, 0},
    {128, 256, 2, 4, 8, 3, 8, 2, 3, 4, 0, 0},
    {128, 256, 4, 2, 8, 1, 8, 1, 1, 2, 0, 0},
    {128, 256, 4, 2, 8, 3, 16, 1, 1, 2, 0, 0},
    {128, 256, 2, 2, 8, 3, 8, 2, 3, 2, 0, 0},
    {128, 256, 4, 2, 4, 3, 8, 2, 3, 2, 0, 0},
    {128, 256, 4, 2, 8, 3, 4, 2, 3, 2, 0, 0},
    {128, 256, 4, 2, 4, 3, 4, 2, 3, 1, 0, 0},
    {128, 256, 4, 2, 8, 1, 8, 1, 1, 1, 0, 0},
    {128, 256, 4, 2, 8, 3, 8, 2, 3, 1, 0, 0},
    {128, 256, 4, 4, 8, 3, 8, 2, 3, 2, 0, 0},
    {128, 256, 2, 2, 8, 1, 16, 1, 1, 1, 0, 0},
    {128, 256, 4, 2, 4, 3, 8, 2, 3, 1, 0, 0},
    {128, 256, 4, 2, 4, 3, 16, 1, 1, 1, 0, 0},
    {128, 256, 4, 2, 8, 3, 16, 1, 1, 1, 0, 0},
    {128, 256, 4, 2, 4, 3, 8, 1, 1, 1, 0, 0},
    {128, 256, 2, 2, 8, 1, 8, 1, 1, 2, 0, 0},
    {128, 256, 2, 2, 8, 1, 8, 1, 1, 1, 0, 0},
    {128, 256, 2, 2, 4, 3, 8, 1, 1, 2, 0, 0},
    {128, 256, 2, 2, 4, 1, 4, 1, 1, 2, 0, 0},
    {128, 256, 4, 4, 8, 1, 16, 1, 1, 4, 0, 0},
    {128, 256, 4, 4, 8, 1, 8, 1, 1, 4, 0, 0},
    {128, 256, 4, 4, 8, 3, 16, 1, 1, 4, 0, 0},
    {128, 256, 4, 2, 16, 1, 8, 1, 1, 2, 0, 0},
    {128, 256, 4, 2, 16, 3, 8, 1, 1, 2, 0, 0},
    {128, 256, 4, 2, 16, 1, 4, 2, 3, 2, 0, 0},
    {128, 256, 4, 2, 16, 1, 16, 1, 1, 2, 0, 0},
    {128, 256, 4, 2, 16, 1, 16, 1, 1, 1, 0, 0},
    {128, 256, 2, 2, 16, 1, 16, 1, 1, 2, 0, 0},
    {128, 256, 4, 4, 16, 1, 16, 1, 1, 2, 0, 0},
    {128, 256, 4, 4, 16, 1, 16, 1, 1, 4, 0, 0},
    {128, 256, 4, 2, 4, 3, 4, 2, 3, 2, 1, 0},
    {128, 256, 4, 2, 4, 3, 4, 2, 3, 1, 1, 0},
    {128, 256, 4, 2, 8, 3, 8, 2, 3, 2, 1, 0},
    {128, 256, 4, 2, 16, 1, 8, 1, 1, 2, 1, 0},
    {128, 256, 4, 2, 8, 1, 8, 1, 1, 2, 1, 0},
    {128, 256, 4, 2, 4, 3, 4, 2, 3, 2, 3, 0},
    {128, 256, 4, 2, 4, 3, 4, 2, 3, 2, 5, 0},
    {128, 256, 4, 2, 4, 3, 4, 2, 3, 2, 7, 0},
    {128, 256, 4, 2, 4, 3, 4, 2, 3, 2, 9, 0},
    {128, 256, 4, 2, 4, 3, 4, 2, 3, 2, 11, 0},
    {128, 256, 4, 2, 4, 3, 4, 2, 3, 2, 13, 0},
    {128, 256, 4, 2, 4, 3, 4, 2, 3, 2, 15, 0},
    {128, 512, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {128, 1024, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {128, 2048, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {128, 4096, 4, 2, 8, 3, 16, 1, 1, 2, 0, 1},
    {256, 32, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {256, 64, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {256, 128, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {256, 256, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {256, 512, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {256, 512, 4, 2, 4, 3, 4, 2, 3, 2, 0, 0},
    {256, 512, 4, 4, 8, 3, 8, 2, 3, 4, 0, 0},
    {256, 512, 2, 4, 8, 3, 8, 2, 3, 4, 0, 0},
    {256, 512, 4, 2, 8, 1, 8, 1, 1, 2, 0, 0},
    {256, 512, 4, 2, 8, 3, 16, 1, 1, 2, 0, 0},
    {256, 512, 2, 2, 8, 3, 8, 2, 3, 2, 0, 0},
    {256, 512, 4, 2, 4, 3, 8, 2, 3, 2, 0, 0},
    {256, 512, 4, 2, 8, 3, 4, 2, 3, 2, 0, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 3, 1, 0, 0},
    {256, 512, 4, 2, 8, 1, 8, 1, 1, 1, 0, 0},
    {256, 512, 4, 2, 8, 3, 8, 2, 3, 1, 0, 0},
    {256, 512, 4, 4, 8, 3, 8, 2, 3, 2, 0, 0},
    {256, 512, 2, 2, 8, 1, 16, 1, 1, 1, 0, 0},
    {256, 512, 4, 2, 4, 3, 8, 2, 3, 1, 0, 0},
    {256, 512, 4, 2, 4, 3, 16, 1, 1, 1, 0, 0},
    {256, 512, 4, 2, 8, 3, 16, 1, 1, 1, 0, 0},
    {256, 512, 4, 2, 4, 3, 8, 1, 1, 1, 0, 0},
    {256, 512, 2, 2, 8, 1, 8, 1, 1, 2, 0, 0},
    {256, 512, 2, 2, 8, 1, 8, 1, 1, 1, 0, 0},
    {256, 512, 2, 2, 4, 3, 8, 1, 1, 2, 0, 0},
    {256, 512, 2, 2, 4, 1, 4, 1, 1, 2, 0, 0},
    {256, 512, 4, 4, 8, 1, 16, 1, 1, 4, 0, 0},
    {256, 512, 4, 4, 8, 1, 8, 1, 1, 4, 0, 0},
    {256, 512, 4, 4, 8, 3, 16, 1, 1, 4, 0, 0},
    {256, 512, 4, 2, 16, 1, 8, 1, 1, 2, 0, 0},
    {256, 512, 4, 2, 16, 3, 8, 1, 1, 2, 0, 0},
    {256, 512, 4, 2, 16, 1, 4, 2, 3, 2, 0, 0},
    {256, 512, 4, 2, 16, 1, 16, 1, 1, 2, 0, 0},
    {256, 512, 4, 2, 16, 1, 16, 1, 1, 1, 0, 0},
    {256, 512, 2, 2, 16, 1, 16, 1, 1, 2, 0, 0},
    {256, 512, 4, 4, 16, 1, 16, 1, 1, 2, 0, 0},
    {256, 512, 4, 4, 16, 1, 16, 1, 1, 4, 0, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 3, 2, 1, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 3, 1, 1, 0},
    {256, 512, 4, 2, 8, 3, 8, 2, 3, 2, 1, 0},
    {256, 512, 4, 2, 16, 1, 8, 1, 1, 2, 1, 0},
    {256, 512, 4, 2, 8, 1, 8, 1, 1, 2, 1, 0},
    {256, 512, 4, 2, 4, 1, 4, 2, 3, 1, 1, 0},
    {256, 512, 4, 2, 4, 1, 4, 2, 3, 1, 3, 0},
    {256, 512, 4, 2, 4, 1, 4, 2, 3, 1, 7, 0},
    {256, 512, 4, 2, 4, 1, 4, 2, 3, 1, 11, 0},
    {256, 512, 4, 2, 4, 1, 4, 2, 3, 1, 15, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 1, 1, 1, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 1, 1, 3, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 1, 1, 7, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 1, 1, 11, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 1, 1, 15, 0},
    {256, 512, 4, 2, 4, 3, 4, 1, 1, 1, 1, 0},
    {256, 512, 4, 2, 4, 3, 4, 1, 1, 1, 3, 0},
    {256, 512, 4, 2, 4, 3, 4, 1, 1, 1, 7, 0},
    {256, 512, 4, 2, 4, 3, 4, 1, 1, 1, 11, 0},
    {256, 512, 4, 2, 4, 3, 4, 1, 1, 1, 15, 0},
    {256, 512, 4, 2, 4, 1, 4, 2, 1, 1, 1, 0},
    {256, 512, 4, 2, 4, 1, 4, 2, 1, 1, 3, 0},
    {256, 512, 4, 2, 4, 1, 4, 2, 1, 1, 7, 0},
    {256, 512, 4, 2, 4, 1, 4, 2, 1, 1, 11, 0},
    {256, 512, 4, 2, 4, 1, 4, 2, 1, 1, 15, 0},
    {256, 512, 4, 2, 4, 1, 4, 1, 1, 1, 1, 0},
    {256, 512, 4, 2, 4, 1, 4, 1, 1, 1, 3, 0},
    {256, 512, 4, 2, 4, 1, 4, 1, 1, 1, 7, 0},
    {256, 512, 4, 2, 4, 1, 4, 1, 1, 1, 11, 0},
    {256, 512, 4, 2, 4, 1, 4, 1, 1, 1, 15, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 3, 1, 2, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 3, 1, 4, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 3, 1, 6, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 3, 1, 8, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 3, 1, 10, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 3, 1, 12, 0},
    {256, 512, 4, 2, 4, 3, 4, 2, 3, 1, 14, 0},
    {256, 512, 4, 4, 8, 3, 8, 2, 3, 4, 2, 0},
    {256, 512, 4, 4, 8, 3, 8, 2, 3, 4, 4, 0},
    {256, 512, 4, 4, 8, 3, 8, 2, 3, 4, 6, 0},
    {256, 512, 4, 4, 8, 3, 8, 2, 3, 4, 8, 0},
    {256, 512, 4, 4, 8, 3, 8, 2, 3, 4, 10, 0},
    {256, 512, 4, 4, 8, 3, 8, 2, 3, 4, 12, 0},
    {256, 512, 4, 4, 8, 3, 8, 2, 3, 4, 14, 0},
    {256, 1024, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {256, 2048, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {256, 4096, 4, 2, 8, 3, 16, 1, 1, 2, 0, 1},
    {512, 32, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {512, 64, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {512, 128, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {512, 256, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {512, 512, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {512, 1024, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {512, 1024, 4, 2, 4, 3, 4, 2, 3, 2, 0, 0},
    {512, 1024, 4, 4, 8, 3, 8, 2, 3, 4, 0, 0},
    {512, 1024, 2, 4, 8, 3, 8, 2, 3, 4, 0, 0},
    {512, 1024, 4, 2, 8, 1, 8, 1, 1, 2, 0, 0},
    {512, 1024, 4, 2, 8, 3, 16, 1, 1, 2, 0, 0},
    {512, 1024, 2, 2, 8, 3, 8, 2, 3, 2, 0, 0},
    {512, 1024, 4, 2, 4, 3, 8, 2, 3, 2, 0, 0},
    {512, 1024, 4, 2, 8, 3, 4, 2, 3, 2, 0, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 3, 1, 0, 0},
    {512, 1024, 4, 2, 8, 1, 8, 1, 1, 1, 0, 0},
    {512, 1024, 4, 2, 8, 3, 8, 2, 3, 1, 0, 0},
    {512, 1024, 4, 4, 8, 3, 8, 2, 3, 2, 0, 0},
    {512, 1024, 2, 2, 8, 1, 16, 1, 1, 1, 0, 0},
    {512, 1024, 4, 2, 4, 3, 8, 2, 3, 1, 0, 0},
    {512, 1024, 4, 2, 4, 3, 16, 1, 1, 1, 0, 0},
    {512, 1024, 4, 2, 8, 3, 16, 1, 1, 1, 0, 0},
    {512, 1024, 4, 2, 4, 3, 8, 1, 1, 1, 0, 0},
    {512, 1024, 2, 2, 8, 1, 8, 1, 1, 2, 0, 0},
    {512, 1024, 2, 2, 8, 1, 8, 1, 1, 1, 0, 0},
    {512, 1024, 2, 2, 4, 3, 8, 1, 1, 2, 0, 0},
    {512, 1024, 2, 2, 4, 1, 4, 1, 1, 2, 0, 0},
    {512, 1024, 4, 4, 8, 1, 16, 1, 1, 4, 0, 0},
    {512, 1024, 4, 4, 8, 1, 8, 1, 1, 4, 0, 0},
    {512, 1024, 4, 4, 8, 3, 16, 1, 1, 4, 0, 0},
    {512, 1024, 4, 2, 16, 1, 8, 1, 1, 2, 0, 0},
    {512, 1024, 4, 2, 16, 3, 8, 1, 1, 2, 0, 0},
    {512, 1024, 4, 2, 16, 1, 4, 2, 3, 2, 0, 0},
    {512, 1024, 4, 2, 16, 1, 16, 1, 1, 2, 0, 0},
    {512, 1024, 4, 2, 16, 1, 16, 1, 1, 1, 0, 0},
    {512, 1024, 2, 2, 16, 1, 16, 1, 1, 2, 0, 0},
    {512, 1024, 4, 4, 16, 1, 16, 1, 1, 2, 0, 0},
    {512, 1024, 4, 4, 16, 1, 16, 1, 1, 4, 0, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 3, 2, 1, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 3, 1, 1, 0},
    {512, 1024, 4, 2, 8, 3, 8, 2, 3, 2, 1, 0},
    {512, 1024, 4, 2, 16, 1, 8, 1, 1, 2, 1, 0},
    {512, 1024, 4, 2, 8, 1, 8, 1, 1, 2, 1, 0},
    {512, 1024, 4, 2, 4, 1, 4, 2, 3, 2, 1, 0},
    {512, 1024, 4, 2, 4, 1, 4, 2, 3, 2, 3, 0},
    {512, 1024, 4, 2, 4, 1, 4, 2, 3, 2, 7, 0},
    {512, 1024, 4, 2, 4, 1, 4, 2, 3, 2, 11, 0},
    {512, 1024, 4, 2, 4, 1, 4, 2, 3, 2, 15, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 1, 2, 1, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 1, 2, 3, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 1, 2, 7, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 1, 2, 11, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 1, 2, 15, 0},
    {512, 1024, 4, 2, 4, 3, 4, 1, 1, 2, 1, 0},
    {512, 1024, 4, 2, 4, 3, 4, 1, 1, 2, 3, 0},
    {512, 1024, 4, 2, 4, 3, 4, 1, 1, 2, 7, 0},
    {512, 1024, 4, 2, 4, 3, 4, 1, 1, 2, 11, 0},
    {512, 1024, 4, 2, 4, 3, 4, 1, 1, 2, 15, 0},
    {512, 1024, 4, 2, 4, 1, 4, 2, 1, 2, 1, 0},
    {512, 1024, 4, 2, 4, 1, 4, 2, 1, 2, 3, 0},
    {512, 1024, 4, 2, 4, 1, 4, 2, 1, 2, 7, 0},
    {512, 1024, 4, 2, 4, 1, 4, 2, 1, 2, 11, 0},
    {512, 1024, 4, 2, 4, 1, 4, 2, 1, 2, 15, 0},
    {512, 1024, 4, 2, 4, 1, 4, 1, 1, 2, 1, 0},
    {512, 1024, 4, 2, 4, 1, 4, 1, 1, 2, 3, 0},
    {512, 1024, 4, 2, 4, 1, 4, 1, 1, 2, 7, 0},
    {512, 1024, 4, 2, 4, 1, 4, 1, 1, 2, 11, 0},
    {512, 1024, 4, 2, 4, 1, 4, 1, 1, 2, 15, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 3, 2, 3, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 3, 2, 5, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 3, 2, 7, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 3, 2, 9, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 3, 2, 11, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 3, 2, 13, 0},
    {512, 1024, 4, 2, 4, 3, 4, 2, 3, 2, 15, 0},
    {512, 1024, 4, 2, 8, 1, 8, 1, 1, 2, 2, 0},
    {512, 1024, 4, 2, 8, 1, 8, 1, 1, 2, 4, 0},
    {512, 1024, 4, 2, 8, 1, 8, 1, 1, 2, 6, 0},
    {512, 1024, 4, 2, 8, 1, 8, 1, 1, 2, 8, 0},
    {512, 1024, 4, 2, 8, 1, 8, 1, 1, 2, 10, 0},
    {512, 1024, 4, 2, 8, 1, 8, 1, 1, 2, 12, 0},
    {512, 1024, 4, 2, 8, 1, 8, 1, 1, 2, 14, 0},
    {512, 2048, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {512, 4096, 4, 2, 8, 3, 16, 1, 1, 2, 0, 1},
    {1024, 32, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {1024, 64, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {1024, 128, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {1024, 256, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {1024, 512, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {1024, 1024, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {1024, 2048, 4, 2, 8, 3, 8, 2, 3, 2, 0, 1},
    {1024, 2048, 4, 2, 8, 1, 8, 1, 1, 2, 0, 0},
    {1024, 2048, 4, 2, 8, 3, 16, 1, 1, 2, 0, 0},
    {1024, 2048, 2, 2, 8, 3, 8, 2, 3, 2, 0, 0},
    {1024, 2048, 4, 2, 8, 1, 8, 1, 1, 1, 0, 0},
    {1024, 2048, 4, 2, 8, 3, 8, 2, 3, 1, 0, 0},
    {1024, 2048, 4, 4, 8, 3, 8, 2, 3, 2, 0, 0},
    {1024, 2048, 2, 2, 8, 1, 16, 1, 1, 1, 0, 0},
    {1024, 2048, 4, 2, 8, 3, 16, 1, 1, 1, 0, 0},
    {1024, 2048, 2, 2, 8, 1, 8, 1, 1, 2, 0, 0},
    {1024, 2048, 2, 2, 8, 1, 8, 1, 1, 1, 0, 0},
    {1024, 2048, 2, 2, 4, 3, 8, 1, 1, 2, 0, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 0, 0},
    {1024, 2048, 4, 4, 8, 3, 16, 1, 1, 4, 0, 0},
    {1024, 2048, 4, 2, 16, 1, 8, 1, 1, 2, 0, 0},
    {1024, 2048, 4, 2, 16, 3, 8, 1, 1, 2, 0, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 0, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 1, 0, 0},
    {1024, 2048, 2, 2, 16, 1, 16, 1, 1, 2, 0, 0},
    {1024, 2048, 4, 4, 16, 1, 16, 1, 1, 2, 0, 0},
    {1024, 2048, 4, 4, 16, 1, 16, 1, 1, 4, 0, 0},
    {1024, 2048, 4, 2, 16, 1, 8, 1, 1, 2, 1, 0},
    {1024, 2048, 4, 2, 8, 1, 8, 1, 1, 2, 1, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 2, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 4, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 6, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 8, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 10, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 12, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 14, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 16, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 18, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 20, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 22, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 24, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 26, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 28, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 30, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 3, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 5, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 7, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 9, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 11, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 13, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 15, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 17, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 19, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 21, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 23, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 25, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 27, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 29, 0},
    {1024, 2048, 4, 2, 16, 1, 16, 1, 1, 2, 31, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 2, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 4, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 6, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 8, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 10, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 12, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 14, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 16, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 18, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 20, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 22, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 24, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 26, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 28, 0},
    {1024, 2048, 4, 4, 8, 1, 16, 1, 1, 4, 30, 0},
    {1024, 4096, 4, 2, 8, 3, 16, 1, 1, 2, 0, 1},
    {2048, 32, 2, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {2048, 64, 2, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {2048, 128, 2, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {2048, 256, 2, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {2048, 512, 2, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {2048, 1024, 2, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {2048, 2048, 2, 2, 8, 1, 8, 2, 3, 2, 0, 1},
    {2048, 4096, 2, 2, 8, 1, 16, 1, 1, 2, 0, 1},
    {2048, 4096, 2, 2, 8, 1, 16, 1, 1, 1, 0, 0},
    {2048, 4096, 2, 2, 8, 1, 8, 1, 1, 1, 0, 0},
    {2048, 4096, 4, 2, 16, 1, 16, 1, 1, 2, 0, 0},
    {2048, 4096, 4, 2, 16, 1, 16, 1, 1, 1, 0, 0},
    {2048, 4096, 2, 2, 16, 1, 16, 1, 1, 2, 0, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 0, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 2, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 4, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 6, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 8, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 10, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 12, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 14, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 16, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 18, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 20, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 22, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 24, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 26, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 28, 0},
    {2048, 4096, 4, 4, 16, 1, 16, 1, 1, 2, 30, 0},
};
}  // namespace

const RowCfg* row_cfgs_f32(int* n) { *n = (int)(sizeof(kRow) / sizeof(kRow[0])); return kRow; }
const ColCfg* col_cfgs_f32(int* n) { *n = (int)(sizeof(kCol) / sizeof(kCol[0])); return kCol; }
const Schedule* schedules_f32(int* n) { *n = (int)(sizeof(kSched) / sizeof(kSched[0])); return kSched; }
ConvLaunch to_bricks_f32() { return launch_to_bricks<T>; }
ConvLaunch from_bricks_f32() { return launch_from_bricks<T>; }
IfacLaunch ifac_init_f32() { return launch_ifac_init; }

}  // namespace af
