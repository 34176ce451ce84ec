// FP32 stage-kernel instantiations, part 0 of 8.
// Generated by gen_kernel_tables.py -- edit that script, not this file.
#include "kernels_inst.hpp"

namespace af {
template hipError_t launch_col<float, 32, 2, 8, 2, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 256, 1, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 256, 2, 4, 4, 2, 2, 3, 3>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 256, 2, 8, 4, 2, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 256, 4, 16, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 512, 1, 4, 4, 2, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 512, 1, 8, 4, 2, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 512, 2, 8, 4, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 512, 4, 8, 2, 4, 2, 3, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 1024, 1, 4, 4, 2, 2, 3, 1>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 1024, 2, 4, 4, 2, 1, 1, 3>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 1024, 2, 4, 4, 2, 2, 3, 9>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 1024, 2, 8, 4, 2, 1, 1, 10>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 1024, 4, 8, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 2048, 2, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 1>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 2048, 2, 16, 4, 2, 1, 1, 17>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 2048, 4, 16, 4, 4, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 4096, 1, 8, 2, 2, 1, 1, 0>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_col<float, 4096, 2, 16, 4, 4, 1, 1, 10>(const void*, int, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 128, 4, 8, 2, 3, 1>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 256, 2, 4, 2, 3, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 256, 4, 4, 2, 3, 1>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 256, 4, 8, 4, 3, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 512, 2, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 512, 4, 4, 2, 3, 5>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 512, 4, 16, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 1024, 4, 8, 2, 1, 0>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 1024, 4, 16, 2, 1, 4>(const void*, int, int, const void*, hipStream_t);
template hipError_t launch_row<float, 2048, 4, 16, 4, 1, 4>(const void*, int, int, const void*, hipStream_t);
}  // namespace af
