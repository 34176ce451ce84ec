// Device FFT engine for gfx950: Stockham radix passes with the data of one
// transform spread over L/R lanes (R elements per lane, in registers) and
// exchanged through LDS between passes.
//
// Element order contract used by every kernel in this library: a lane with
// index j in [0, L/R) holds, in v[r], the element at position
//     p = j + r * (L/R),      r = 0..R-1
// both before and after fft_multi().  The first radix pass reads exactly these
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

// LDS index of element i = p*W + c (position p of interleaved transform c)
// for the exchange written by a radix-R pass of stride Ns.  The rule is chosen
// against gfx950's LDS (64 banks of 4 bytes): the paired 8-byte accesses
// fft_multi compiles to (ds_read2/ds_write2) are served 16 lanes at a time and
// conflict-free when those lanes' 8-byte elements differ mod 16.
//   8-byte elements:  P = min(Ns*W, 16/R) extra elements after every 16, none
//                     once Ns*W >= 16 (stores already hit 16 consecutive
//                     elements);
//   16-byte elements: Ns*W extra per Ns*R*W, none once Ns*W*16 >= 256.
// Radix 16 with W >= 2 is the exception at Ns = 1: its padded load offsets
// are too far apart to pair, so the loads compile to single ds_read_b64,
// which are served 32 lanes at a time (elements must differ mod 32); 2 extra
// elements after every 32 keep those loads and the paired stores
// conflict-free (1 after every 16 left a 2-way conflict on the loads).
// Every store and load of every schedule is conflict-free under these
// rules, and for a given exchange each lane's R stores and R loads sit at one
// lane base plus lane-independent offsets (lds_pad of the offset alone), so
// fft_multi computes one base per lane and uses the offsets as immediates.
template <int W, int R, int ESZ>
__host__ __device__ constexpr int lds_pad(int i, int Ns) {
  if constexpr (ESZ == 8) {
    if (Ns * W >= 16) return i;
    if (R == 16 && Ns == 1 && W >= 2) return i + (i / 32) * 2;
    const int P = (Ns * W < 16 / R) ? Ns * W : 16 / R;
    return i + (i / 16) * P;
  } else {
    return (Ns * W * ESZ >= 256) ? i : i + (i / (Ns * R * W)) * (Ns * W);
  }
}
// Elements of one LDS area for transforms of length L (covers every padding).
template <int L, int W, int R> constexpr int lds_area_elems() { return L * W + (L * W) / R + 1; }

// Twiddle table of fft_multi, in the order the radix passes read it: for
// pass P >= 1 (stride Ns = R^P) entry off(P) + jm*(R-1) + (r-1) holds
// w^(r*jm*L/(Ns*R)), jm < Ns, r in [1, R); for a trailing radix-RL pass entry
// off(NP) + b*(RL-1) + (rr-1) holds w^(rr*b), b < L/RL, rr in [1, RL); with
// w = exp(-2 pi i / L).  Lanes with consecutive jm (or b) read entries an
// odd number of 8- or 16-byte elements apart, which never share a bank, and
// each lane's R-1 twiddles of a pass sit at one base plus immediate offsets.
// L - R entries (none when L == R).
template <int L, int R>
struct TwLayout {
  static constexpr int NP = ilog(L, R);
  static constexpr int RL = L / ipow(R, NP);
  static constexpr int kSize = L - R;
  static constexpr int kAlloc = kSize > 0 ? kSize : 1;
  static constexpr int off(int P) { return ipow(R, P) - R; }
  // exponent m of the entry i (w^m)
  __host__ __device__ static int exponent(int i) {
    int P = 1;
    for (; P < NP; ++P)
      if (i < off(P + 1)) {
        const int k = i - off(P), jm = k / (R - 1), r = k % (R - 1) + 1;
        return r * jm * (L / (ipow(R, P) * R));
      }
    if constexpr (RL > 1) {
      const int k = i - off(NP), b = k / (RL - 1), rr = k % (RL - 1) + 1;
      return rr * b;
    } else {
      return 0;   // no entries past the full passes
    }
  }
};

// NF length-L transforms done together (one LDS exchange per radix pass for
// all of them).  lds points at this block's NF areas (lds_area_elems each) of
// W interleaved transforms; this lane is transform lane j of column c.  tw
// points at the TwLayout<L, R> table (normally an LDS copy).
// All lanes of the block must call this together (it contains block
// barriers).  The first LDS access is a store after radix pass 1, so the
// caller must place a barrier before the call if other lanes may still be
// reading the same LDS areas from an earlier use.
// BUF (elements): offset of a second set of LDS areas, or 0.  With two sets
// the exchanges alternate between them (exchange e writes set e % 2), so the
// barrier that keeps a pass's stores from overwriting data other lanes are
// still loading is not needed: one barrier per exchange instead of two.
template <int L, int R, int DIR, int NF, int W, class T, int BUF = 0>
struct FftMulti {
  static_assert((L & (L - 1)) == 0 && (R & (R - 1)) == 0 && R <= L && R <= 16, "power-of-two sizes");
  static constexpr int NP = ilog(L, R);       // full radix-R passes
  static constexpr int RL = L / ipow(R, NP);  // trailing radix (1 if L is a power of R)
  static constexpr int S = L / R;             // lanes per transform
  static constexpr int ESZ = (int)sizeof(cx<T>);
  // LDS set (0 or 1) the transform's final loads read, -1 if it has no
  // exchange; its first stores (after pass 0) always go to set 0.  With two
  // sets (BUF != 0) a caller may skip the barrier before the next transform
  // when kLastLoadSet == 1: the next one's first stores cannot overwrite
  // data a lagging lane still loads.
  static constexpr int kLastLoadSet = (RL > 1) ? (NP - 1) % 2 : (NP >= 2 ? (NP - 2) % 2 : -1);
  static constexpr int AREA = lds_area_elems<L, W, R>();
  static __device__ __forceinline__ int pad(int i, int Ns) { return lds_pad<W, R, ESZ>(i, Ns); }
  static constexpr bool pads_fit() {
    for (int ns = 1; ns < L; ns *= R)
      if (lds_pad<W, R, ESZ>(L * W - 1, ns) >= AREA) return false;
    return true;
  }
  static_assert(pads_fit(), "every padded exchange fits the LDS area");

  // loads of the positions j + r*S written by the pass of stride Nsp
  template <int Nsp>
  static __device__ __forceinline__ void load(cx<T> (&v)[NF][R], int j, int c, const cx<T>* lds) {
    const int a = pad(j * W + c, Nsp);
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int r = 0; r < R; ++r) v[f][r] = lds[f * AREA + a + lds_pad<W, R, ESZ>(r * S * W, Nsp)];
  }

  template <int P>
  static __device__ __forceinline__ void pass(cx<T> (&v)[NF][R], int j, int c, cx<T>* lds, const cx<T>* tw) {
    if constexpr (P < NP) {
      constexpr int Ns = ipow(R, P);
      if constexpr (P > 0) {
        __syncthreads();
        load<Ns / R>(v, j, c, lds + ((P - 1) % 2) * BUF);
        const int jm = j & (Ns - 1);
        const cx<T>* twp = tw + (TwLayout<L, R>::off(P) - 1) + jm * (R - 1);
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const cx<T> w = twp[r];
#pragma unroll
          for (int f = 0; f < NF; ++f) v[f][r] = DIR < 0 ? cmul(v[f][r], w) : cmulc(v[f][r], w);
        }
      }
#pragma unroll
      for (int f = 0; f < NF; ++f) Dft<R, DIR, T>::run(v[f]);
      constexpr bool last = (P == NP - 1) && (RL == 1);
      if constexpr (!last) {
        if constexpr (P > 0 && BUF == 0) __syncthreads();
        const int base = (j / Ns) * Ns * R + (j & (Ns - 1));
        const int a = pad(base * W + c, Ns);
        cx<T>* const st = lds + (P % 2) * BUF;
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
          for (int r = 0; r < R; ++r) st[f * AREA + a + lds_pad<W, R, ESZ>(r * Ns * W, Ns)] = v[f][r];
      }
      pass<P + 1>(v, j, c, lds, tw);
    } else if constexpr (RL > 1) {
      constexpr int Nsp = ipow(R, NP - 1);
      __syncthreads();
      load<Nsp>(v, j, c, lds + ((NP - 1) % 2) * BUF);
      // R/RL butterflies of radix RL on elements m = q + (R/RL) r'
#pragma unroll
      for (int q = 0; q < R / RL; ++q) {
        const int b = j + q * S;
        const cx<T>* twp = tw + (TwLayout<L, R>::off(NP) - 1) + b * (RL - 1);
        cx<T> w[RL];
#pragma unroll
        for (int rr = 1; rr < RL; ++rr) w[rr] = twp[rr];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          cx<T> u[RL];
#pragma unroll
          for (int rr = 0; rr < RL; ++rr) {
            cx<T> x = v[f][q + (R / RL) * rr];
            if (rr > 0) x = DIR < 0 ? cmul(x, w[rr]) : cmulc(x, w[rr]);
            u[rr] = x;
          }
          Dft<RL, DIR, T>::run(u);
#pragma unroll
          for (int rr = 0; rr < RL; ++rr) v[f][q + (R / RL) * rr] = u[rr];
        }
      }
    }
  }
};

template <int L, int R, int DIR, int NF, class T, class Lds>
__device__ __forceinline__ void fft_multi(cx<T> (&v)[NF][R], int j, Lds& lds, const cx<T>* tw) {
  static_assert(Lds::kL == L && Lds::kR == R, "LDS areas laid out for this transform");
  static_assert(NF <= Lds::kFields, "fields fit the LDS areas of one set");
  FftMulti<L, R, DIR, NF, Lds::kW, T, Lds::kBuf>::template pass<0>(v, j, lds.c, lds.base, tw);
}

}  // namespace af
