// FP32 stage-kernel instantiations, part 4 of 8.
// Generated by gen_kernel_tables.py -- edit that script, not this file.
#include "kernels_inst.hpp"

namespace af {
template hipError_t launch_col<float, 128, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 256, 1, 16, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 256, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 256, 2, 16, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 512, 1, 4, 4, 2, 2, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 512, 1, 4, 4, 2, 2, 3, 10>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 512, 2, 4, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 512, 2, 8, 4, 4, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 512, 4, 8, 4, 4, 2, 3, 8>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 1024, 1, 16, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 1024, 2, 4, 4, 2, 2, 1, 11>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 1024, 2, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 1024, 2, 16, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 2048, 1, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 2048, 2, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 9>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 25>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 2048, 4, 16, 4, 4, 1, 1, 16>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 4096, 2, 16, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 4096, 2, 16, 4, 4, 1, 1, 26>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 128, 4, 16, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 256, 2, 16, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 256, 4, 8, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 256, 4, 16, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 512, 4, 4, 2, 1, 1>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 512, 4, 8, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 1024, 2, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 1024, 4, 8, 4, 1, 4>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 2048, 2, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
}  // namespace af
