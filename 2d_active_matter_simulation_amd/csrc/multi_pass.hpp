// One launch per pass over several size classes (the ensemble's single-launch
// mode).  A mixed launch's blocks are the blocks of each class in turn: block
// b < nblocks[0] runs class 0's body for its block b, the next nblocks[1]
// run class 1's, and so on.  Every class of a mixed launch uses a schedule of
// the same block size (the launch has one), so each block runs exactly the
// barrier sequence of its own class's kernel; LDS is the largest of the
// classes' needs.  A block computes what the class's own kernel computes for
// that block, so results are bitwise those of the per-class launches with the
// same schedules.  nblocks[k] is a multiple of 8 for every class but the
// last, so each class's blocks keep their position modulo the 8 XCDs and the
// column pass's XCD-aware tile order keeps its effect.
#pragma once
#include "registry.hpp"
#include "stage_kernels.hpp"

namespace af {

// A column-pass configuration without its stage.
template <class T, int NY, int W, int R, int BY, int BX, int GF, int GI, int PP>
struct ColCfgT {
  template <int S> using At = ColPass<T, NY, W, R, BY, BX, GF, GI, PP, S>;
  static constexpr int NT = At<1>::NT;
};

template <class C, class... Rest>
struct MultiTraits {
  static constexpr int NT = C::NT;
  static constexpr size_t kSmem =
      sizeof(typename C::Smem) > MultiTraits<Rest...>::kSmem ? sizeof(typename C::Smem) : MultiTraits<Rest...>::kSmem;
  static constexpr bool kSameNT = (C::NT == MultiTraits<Rest...>::NT) && MultiTraits<Rest...>::kSameNT;
};
template <class C>
struct MultiTraits<C> {
  static constexpr int NT = C::NT;
  static constexpr size_t kSmem = sizeof(typename C::Smem);
  static constexpr bool kSameNT = true;
};

template <int I, class C, class... Rest>
__device__ __forceinline__ void row_dispatch(const MultiArgs& a, int b, unsigned char* smem) {
  using T = typename C::Real;
  if constexpr (sizeof...(Rest) == 0) {
    C::run((const InstDesc<T>*)a.descs[I], a.tiles[I], (const cx<T>*)a.pool[I], b, a.nblocks[I],
           *reinterpret_cast<typename C::Smem*>(smem));
  } else {
    if (b < a.nblocks[I])
      C::run((const InstDesc<T>*)a.descs[I], a.tiles[I], (const cx<T>*)a.pool[I], b, a.nblocks[I],
             *reinterpret_cast<typename C::Smem*>(smem));
    else
      row_dispatch<I + 1, Rest...>(a, b - a.nblocks[I], smem);
  }
}

template <class... C>
__global__ void __launch_bounds__(MultiTraits<C...>::NT) row_multi(MultiArgs a) {
  static_assert(MultiTraits<C...>::kSameNT, "one block size for every class of a mixed launch");
  __shared__ __align__(16) unsigned char smem[MultiTraits<C...>::kSmem];
  row_dispatch<0, C...>(a, blockIdx.x, smem);
}

template <int I, class C, class... Rest>
__device__ __forceinline__ void col_dispatch(const MultiArgs& a, int b, unsigned char* smem) {
  using T = typename C::Real;
  if constexpr (sizeof...(Rest) == 0) {
    C::run((const InstDesc<T>*)a.descs[I], a.tiles[I], (const cx<T>*)a.pool[I], b, a.nblocks[I],
           *reinterpret_cast<typename C::Smem*>(smem));
  } else {
    if (b < a.nblocks[I])
      C::run((const InstDesc<T>*)a.descs[I], a.tiles[I], (const cx<T>*)a.pool[I], b, a.nblocks[I],
             *reinterpret_cast<typename C::Smem*>(smem));
    else
      col_dispatch<I + 1, Rest...>(a, b - a.nblocks[I], smem);
  }
}

template <int S, class... Cfg>
__global__ void __launch_bounds__(MultiTraits<typename Cfg::template At<S>...>::NT) col_multi(MultiArgs a) {
  using Tr = MultiTraits<typename Cfg::template At<S>...>;
  static_assert(Tr::kSameNT, "one block size for every class of a mixed launch");
  __shared__ __align__(16) unsigned char smem[Tr::kSmem];
  col_dispatch<0, typename Cfg::template At<S>...>(a, blockIdx.x, smem);
}

inline int multi_total(const MultiArgs& a, int n) {
  int t = 0;
  for (int k = 0; k < n; ++k) t += a.nblocks[k];
  return t;
}

template <class... C>
hipError_t launch_row_multi(const MultiArgs& a, hipStream_t s) {
  const int total = multi_total(a, (int)sizeof...(C));
  if (total > 0) hipLaunchKernelGGL((row_multi<C...>), dim3(total), dim3(MultiTraits<C...>::NT), 0, s, a);
  return hipGetLastError();
}

template <class... Cfg>
hipError_t launch_col_multi(const MultiArgs& a, int stage, hipStream_t s) {
  const int total = multi_total(a, (int)sizeof...(Cfg));
  if (total == 0) return hipSuccess;
  const dim3 grid(total), block(MultiTraits<typename Cfg::template At<1>...>::NT);
  switch (stage) {
    case 1: hipLaunchKernelGGL((col_multi<1, Cfg...>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((col_multi<2, Cfg...>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((col_multi<3, Cfg...>), grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL((col_multi<4, Cfg...>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace af
