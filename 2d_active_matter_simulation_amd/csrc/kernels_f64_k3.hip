// FP64 stage-kernel instantiations, part 3 of 4.
// Generated by gen_kernel_tables.py -- edit that script, not this file.
#include "kernels_inst.hpp"

namespace af {
template hipError_t launch_col<double, 64, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<double, 256, 1, 8, 4, 1, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<double, 256, 1, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<double, 256, 2, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<double, 512, 1, 4, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<double, 512, 1, 8, 4, 2, 1, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<double, 512, 2, 4, 4, 2, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<double, 512, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<double, 1024, 1, 8, 4, 2, 1, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<double, 1024, 2, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<double, 2048, 1, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<double, 4096, 1, 8, 4, 1, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_row<double, 32, 4, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<double, 128, 2, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<double, 128, 4, 8, 1, 3, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<double, 128, 4, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<double, 256, 4, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<double, 256, 4, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<double, 512, 2, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<double, 512, 4, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<double, 1024, 4, 8, 1, 1, 0>(const void*, int, int, const void*, hipStream_t);
}  // namespace af
