// Instantiation helpers shared by kernels_f32.hip and kernels_f64.hip.
#pragma once
#include <hip/hip_ext.h>

#include "registry.hpp"
#include "stage_kernels.hpp"

namespace af {

// Launches through hipExtLaunchKernelGGL, which also records the profiling
// events of g_launch_events when they are set.
template <class K, class... A>
inline void launch(K kernel, dim3 grid, dim3 block, hipStream_t s, A... args) {
  hipExtLaunchKernelGGL(kernel, grid, block, 0, s, g_launch_events.start, g_launch_events.stop, 0, args...);
}

template <class T, int NX2, int RB, int R, int BX, int G, int PP>
hipError_t launch_row(const void* descs, int ninst, int tiles, const void* pool, hipStream_t s) {
  launch(row_pass<T, NX2, RB, R, BX, G, PP>, dim3(ninst * tiles), dim3(RB * (NX2 / R)), s, (const InstDesc<T>*)descs,
         tiles, (const cx<T>*)pool);
  return hipGetLastError();
}

template <class T, int NY, int W, int R, int BY, int BX, int GF, int GI, int PP>
hipError_t launch_col(const void* descs, int ninst, int tiles, int stage, const void* pool, hipStream_t s) {
  const dim3 grid(ninst * tiles), block(W * (NY / R));
  const InstDesc<T>* dd = (const InstDesc<T>*)descs;
  const cx<T>* pp = (const cx<T>*)pool;
  switch (stage) {
    case 0: launch(col_pass<T, NY, W, R, BY, BX, GF, GI, PP, 0>, grid, block, s, dd, tiles, pp); break;
    case 1: launch(col_pass<T, NY, W, R, BY, BX, GF, GI, PP, 1>, grid, block, s, dd, tiles, pp); break;
    case 2: launch(col_pass<T, NY, W, R, BY, BX, GF, GI, PP, 2>, grid, block, s, dd, tiles, pp); break;
    case 3: launch(col_pass<T, NY, W, R, BY, BX, GF, GI, PP, 3>, grid, block, s, dd, tiles, pp); break;
    case 4: launch(col_pass<T, NY, W, R, BY, BX, GF, GI, PP, 4>, grid, block, s, dd, tiles, pp); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

inline int conv_blocks(size_t n) { return (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192); }

template <class T>
hipError_t launch_to_bricks(const void* src, void* dst, int nx, int ny, int by, int bx, hipStream_t s) {
  hipLaunchKernelGGL((to_bricks<T>), dim3(conv_blocks((size_t)ny * (nx / 2))), dim3(256), 0, s,
                     (const double*)src, (cx<T>*)dst, nx, ny, by, bx);
  return hipGetLastError();
}

template <class T>
hipError_t launch_from_bricks(const void* src, void* dst, int nx, int ny, int by, int bx, hipStream_t s) {
  hipLaunchKernelGGL((from_bricks<T>), dim3(conv_blocks((size_t)ny * (nx / 2 + 1))), dim3(256), 0, s,
                     (const cx<T>*)src, (double*)dst, nx, ny, by, bx);
  return hipGetLastError();
}

// Integrating-factor table of an FP32 run (written once, at set-up).
inline hipError_t launch_ifac_init(const void* desc, int nx2, int ny, int w, hipStream_t s) {
  hipLaunchKernelGGL(ifac_init<float>, dim3(conv_blocks((size_t)ny * nx2)), dim3(256), 0, s, (const InstDesc<float>*)desc,
                     nx2, ny, w);
  return hipGetLastError();
}

}  // namespace af
