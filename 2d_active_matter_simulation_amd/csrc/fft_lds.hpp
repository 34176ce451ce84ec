// Device FFT engine for gfx950: Stockham radix passes with the data of one
// transform spread over L/R lanes (R elements per lane, in registers) and
// exchanged through LDS between passes.
//
// Element order contract used by every kernel in this library: a lane with
// index j in [0, L/R) holds, in v[r], the element at position
//     p = j + r * (L/R),      r = 0..R-1
// both before and after fft_core().  The first radix pass reads exactly these
// positions (so it can be fed straight from global memory) and the last pass
// writes exactly these positions (so its output can go straight to global
// memory or into the elementwise stage maths).  Input and output are in
// natural order (Stockham autosort), so no bit reversal is needed anywhere.
//
// Written from the published Stockham formulation (e.g. Govindaraju et al.,
// "High performance discrete Fourier transforms on graphics processors",
// SC'08) and the textbook even/odd split for real transforms.
#pragma once
#include <hip/hip_runtime.h>
#include "layout.hpp"

namespace af {

template <class T> __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T> __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T> __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// a * conj(b)
template <class T> __device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
template <class T> __device__ __forceinline__ cx<T> cscale(T s, cx<T> a) { return {s * a.re, s * a.im}; }
template <class T> __device__ __forceinline__ cx<T> cconj(cx<T> a) { return {a.re, -a.im}; }
template <class T> __device__ __forceinline__ cx<T> cmuli(cx<T> a) { return {-a.im, a.re}; }

// exp(-2 pi i k/16), k = 0..7, correctly rounded doubles.
#define AF_C1 0.92387953251128675613
#define AF_S1 0.38268343236508977173
#define AF_C2 0.70710678118654752440

// Multiply v by W_R^k = exp(DIR * 2 pi i k / R) for a compile-time k < R/2
// (DIR = -1 forward, +1 inverse).
template <int R, int K, int DIR, class T>
__device__ __forceinline__ cx<T> twc(cx<T> v) {
  constexpr int k16 = K * (16 / R);  // same angle on the 16-point circle
  if constexpr (k16 == 0) {
    return v;
  } else if constexpr (k16 == 4) {
    // exp(DIR * i pi/2) = DIR * i
    return DIR < 0 ? cx<T>{v.im, -v.re} : cx<T>{-v.im, v.re};
  } else {
    constexpr double tr[8] = {1.0, AF_C1, AF_C2, AF_S1, 0.0, -AF_S1, -AF_C2, -AF_C1};
    constexpr double ti[8] = {0.0, -AF_S1, -AF_C2, -AF_C1, -1.0, -AF_C1, -AF_C2, -AF_S1};
    const T wr = (T)tr[k16];
    const T wi = (T)(DIR < 0 ? ti[k16] : -ti[k16]);
    return {v.re * wr - v.im * wi, v.re * wi + v.im * wr};
  }
}

// In-register DFT of R = 1, 2, 4, 8 or 16 points (radix-2 decimation in time,
// recursively), natural order in and out.
template <int R, int DIR, class T>
struct Dft {
  template <int K>
  __device__ __forceinline__ static void comb(cx<T>* v, const cx<T>* e, const cx<T>* o) {
    if constexpr (K < R / 2) {
      const cx<T> t = twc<R, K, DIR, T>(o[K]);
      v[K] = e[K] + t;
      v[K + R / 2] = e[K] - t;
      comb<K + 1>(v, e, o);
    }
  }
  __device__ __forceinline__ static void run(cx<T>* v) {
    if constexpr (R == 1) {
    } else if constexpr (R == 2) {
      const cx<T> a = v[0], b = v[1];
      v[0] = a + b;
      v[1] = a - b;
    } else {
      cx<T> e[R / 2], o[R / 2];
#pragma unroll
      for (int i = 0; i < R / 2; ++i) { e[i] = v[2 * i]; o[i] = v[2 * i + 1]; }
      Dft<R / 2, DIR, T>::run(e);
      Dft<R / 2, DIR, T>::run(o);
      comb<0>(v, e, o);
    }
  }
};

constexpr int ilog(int n, int b) { return n < b ? 0 : 1 + ilog(n / b, b); }
constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

// Full length-L transform over one LDS area.  `Lds` provides load(p)/store(p,v)
// for positions p in [0, L) of this lane's transform.  tw points at
// tw[m] = exp(-2 pi i m / L), m in [0, L).  All lanes of the block must call
// this together (it contains block barriers).  On entry the LDS area may still
// be read by other lanes of a previous call only if the caller placed a
// barrier before; this function's first LDS access is a store after pass 1.
template <int L, int R, int DIR, class T, class Lds>
__device__ __forceinline__ void fft_core(cx<T> (&v)[R], int j, Lds& lds, const cx<T>* __restrict__ tw) {
  static_assert((L & (L - 1)) == 0 && (R & (R - 1)) == 0 && R <= L, "power-of-two sizes");
  constexpr int NP = ilog(L, R);           // full radix-R passes
  constexpr int RL = L / ipow(R, NP);      // trailing radix (1 if L is a power of R)
  constexpr int S = L / R;                 // lanes per transform
  int Ns = 1;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    if (p > 0) {
      __syncthreads();
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = lds.load(j + r * S);
      const int jm = j & (Ns - 1);
      const int stride = L / (Ns * R);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const cx<T> w = tw[r * jm * stride];
        v[r] = DIR < 0 ? cmul(v[r], w) : cmulc(v[r], w);
      }
    }
    Dft<R, DIR, T>::run(v);
    const bool last = (p == NP - 1) && (RL == 1);
    if (!last) {
      if (p > 0) __syncthreads();
      const int base = (j / Ns) * Ns * R + (j & (Ns - 1));
#pragma unroll
      for (int r = 0; r < R; ++r) lds.store(base + r * Ns, v[r]);
    }
    Ns *= R;
  }
  if constexpr (RL > 1) {
    __syncthreads();
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = lds.load(j + m * S);
    // R/RL butterflies of radix RL on elements m = q + (R/RL) r'
#pragma unroll
    for (int q = 0; q < R / RL; ++q) {
      const int b = j + q * S;
      cx<T> u[RL];
#pragma unroll
      for (int rr = 0; rr < RL; ++rr) {
        cx<T> x = v[q + (R / RL) * rr];
        if (rr > 0) {
          const cx<T> w = tw[rr * b];
          x = DIR < 0 ? cmul(x, w) : cmulc(x, w);
        }
        u[rr] = x;
      }
      Dft<RL, DIR, T>::run(u);
#pragma unroll
      for (int rr = 0; rr < RL; ++rr) v[q + (R / RL) * rr] = u[rr];
    }
  }
}

}  // namespace af
