// Fused RK-stage kernels for the James-Wilczek vorticity step.
//
// One RK stage of the reference (nonl_func + reality_func + integrate_func_k,
// ActFlow_V3-2.cu:736-821, 1036-1243) is two kernels here:
//
//   row pass   (mixed space -> real space -> mixed space, one block per brick row)
//     reads  U, V, Wm : the y-inverse-transformed v_x, v_y, omega (Ny x Nx/2)
//     row C2R x3, products A=w ux, W=w uy, C=|u|^2 u (nonl_subfunc_1,
//     .cu:1192-1205), F1 = lam A + beta Cy, F2 = lam W - beta Cx, row R2C x2,
//     P = -i kx F1^ ; Q = F2^      (the kx half of nonl_subfunc_2, .cu:1221)
//     FP64 block partials of sum|u|^2 (energy, .cu:1118-1127) and sum C
//     (mean-velocity RHS, .cu:1086-1087)
//     writes 2P, 2Q (the column pass folds the 1/2 into the mask factor)
//
//   column pass (one block per brick column)
//     reads  P, Q, omega_old, acc
//     column FFT x2, N^ = (P^ - i ky Q^) * mask   (rest of nonl_subfunc_2)
//     Hermitian projection of acc's kx=0 column (reality_func, .cu:1237-1242)
//     IF-RK accumulate + next stage input (integrate_func_1..4, .cu:957-1018)
//     mean-velocity ODE on tile 0 (.cu:756-821), NaN watchdog (.cu:884-897)
//     Biot-Savart v^ = (i ky, -i kx) s/k^2 with mean injection
//     (nonl_subfunc_0 .cu:1176-1180, .cu:1053-1054), column inverse FFT x3
//     writes acc (or the new state at stage 4), U, V, Wm
//
// The intermediate "mixed" fields are never materialised in the reference's
// form: the reference's 7 full 2-D transforms and 6 elementwise kernels per
// stage become 2 launches that each touch every field once.
//
// Field layout in HBM (all complex fields, Ny rows x Nx/2 columns; the
// kx = Nx/2 Nyquist column is not stored because every supported dealiasing
// mask zeroes it): bricks of BY rows x BX columns stored contiguously,
//   idx(y, x) = ((y/BY)*(Nx/2/BX) + x/BX) * BY*BX + (y%BY)*BX + x%BX,
// so a row block reads whole bricks and a column block reads BX-wide bricks.
#pragma once
#include <type_traits>

#include "fft_lds.hpp"
#include "layout.hpp"

namespace af {

// Pointers read from an instance descriptor are generic; the small arrays
// (partials, twiddle pool) are read through an address-space-1 view (global
// instructions instead of FLAT).  Fields go through BufField below.
#if defined(__HIP_DEVICE_COMPILE__)
template <class T> using gptr = __attribute__((address_space(1))) T*;
#else   // host pass of hipcc: these functions are never executed on the host
template <class T> using gptr = T*;
#endif
template <class T> __device__ __forceinline__ gptr<T> glob(T* p) { return (gptr<T>)p; }

// Cache-policy bits of gfx950 buffer instructions (their sc0 / nt / sc1
// fields): a device-scope store (sc1) is written through the XCD's L2, so the
// end-of-kernel write-back has nothing left to flush; a streaming access (nt)
// marks its lines for early eviction.  Measured on the stage kernels: the row
// pass's loads (read once) and stores gain from nt loads and sc1 stores at
// every size; for the column pass's stores and its P, Q loads the better
// choice depends on the size and batch, so those are schedule options.
constexpr int kCacheDefault = 0, kCacheNt = 2, kCacheSc1 = 16;

// Buffer-resource view of a complex field: the base lives in a 4-SGPR buffer
// descriptor, a lane's byte offset in a VGPR and a uniform byte offset (e.g. a
// row group's) in an SGPR, so an access costs no per-lane address arithmetic.
// Raw buffer accesses (stride 0, descriptor word 3 of the gfx9 family);
// offsets stay far below num_records (fields hold at most 2^27 bytes).
template <class T>
struct BufField {
#if defined(__HIP_DEVICE_COMPILE__)
  // 16-byte elements (complex double) are stored with global_store_dwordx4:
  // the compiler treats a buffer_store_dwordx4 with an SGPR soffset as free of
  // the store-data hazard, and on gfx950 its data registers were observed to be
  // overwritten before the store read them (corrupt dwords in FP64 fields under
  // register pressure).  8-byte stores and all loads use the buffer path.
  __amdgpu_buffer_rsrc_t rs;
  gptr<char> base;
  __device__ __forceinline__ explicit BufField(const T* p)
      : rs(__builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(p), 0, 0x7fffffff, 0x00020000)),
        base((gptr<char>)glob(const_cast<T*>(p))) {}
  template <int CP = kCacheDefault>
  __device__ __forceinline__ T ld(unsigned vbyte, unsigned sbyte) const {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "float, complex float or double");
    if constexpr (sizeof(T) == 4)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, vbyte, sbyte, CP));
    else if constexpr (sizeof(T) == 8)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, vbyte, sbyte, CP));
    else
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b128(rs, vbyte, sbyte, CP));
  }
  // CP: cache-policy bits of 8-byte stores (16-byte stores take the default)
  template <int CP = kCacheDefault>
  __device__ __forceinline__ void st(unsigned vbyte, unsigned sbyte, T v) const {
    if constexpr (sizeof(T) == 8)
      __builtin_amdgcn_raw_buffer_store_b64(
          __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(rs, 0, 0, 0)), v), rs, vbyte, sbyte, CP);
    else   // 16-byte stores go through a global store (see above)
      *(gptr<T>)(base + sbyte + vbyte) = v;
  }
#else   // host pass of hipcc: never executed on the host
  __device__ explicit BufField(const T*) {}
  template <int CP = kCacheDefault>
  __device__ T ld(unsigned, unsigned) const { return T{}; }
  template <int CP = kCacheDefault>
  __device__ void st(unsigned, unsigned, T) const {}
#endif
};

// 1/x: hardware reciprocal (1 ulp) in FP32, correctly rounded division in FP64.
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }

// Largest |ky index| kept by the mask in column x (-1: none), so the per-mode
// test of the column pass is one integer compare.  Mode 0: the largest y with
// 9 (x^2 Ny^2 + y^2 Nx^2) <= Nx^2 Ny^2; mode 1: 3y <= Ny when 3x <= Nx.
__device__ __forceinline__ int mask_ymax(int x, int nx, int ny, int mode) {
  if (mode != 0) return (3 * x <= nx) ? ny / 3 : -1;
  // all products below are integers < 2^53, exact in double
  const double NX = nx, NY = ny, X = x;
  const double rhs = NY * NY * (NX * NX - 9.0 * X * X);   // >= 9 NX^2 y^2
  if (rhs < 0) return -1;
  const double c = 9.0 * NX * NX;
  double y = floor(sqrt(rhs / c));
  if (c * (y + 1) * (y + 1) <= rhs) y += 1;
  if (c * y * y > rhs) y -= 1;
  return (int)y;
}

// LDS view of NFM field areas, each holding W interleaved transforms of
// length L (the W columns of a column block or the RB rows of a row block);
// with PP a second set of NFM areas follows, used by alternate exchanges.
template <class T, int L, int W, int R, int NFM = 1, int PP = 0>
struct LdsView {
  static constexpr int kL = L, kW = W, kR = R, kFields = NFM;
  static constexpr int kBuf = PP ? NFM * lds_area_elems<L, W, R>() : 0;   // elements
  static constexpr int kElems = (PP ? 2 : 1) * NFM * lds_area_elems<L, W, R>();
  cx<T>* base;
  int c;
  // unpadded access for the even/odd pair exchange of the real transforms,
  // in set `set` (0 or 1; one set only without PP)
  __device__ __forceinline__ cx<T> load_plain(int f, int p, int set = 0) const {
    return base[set * kBuf + f * lds_area_elems<L, W, R>() + p * W + c];
  }
  __device__ __forceinline__ void store_plain(int f, int p, cx<T> v, int set = 0) {
    base[set * kBuf + f * lds_area_elems<L, W, R>() + p * W + c] = v;
  }
};

// Twiddle pool in HBM: tw[L + m] = exp(-2 pi i m / L) for every power of two L.
template <class T>
__device__ __forceinline__ const cx<T>* tw_of(const cx<T>* pool, int L) { return pool + L; }

// Copies n twiddles into LDS (caller places a barrier before their first use).
template <class T, int NT>
__device__ __forceinline__ void stage_twiddles(cx<T>* dst, const cx<T>* __restrict__ src, int n) {
  const gptr<const cx<T>> g = glob(src);
  for (int i = threadIdx.x; i < n; i += NT) dst[i] = g[i];
}
// Builds the pass-ordered twiddle table of a length-L radix-R fft_multi in
// LDS from the natural-order table src[m] = exp(-2 pi i m / L).
template <class T, int NT, int L, int R>
__device__ __forceinline__ void stage_fft_twiddles(cx<T>* dst, const cx<T>* __restrict__ src) {
  const gptr<const cx<T>> g = glob(src);
  for (int i = threadIdx.x; i < TwLayout<L, R>::kSize; i += NT) dst[i] = g[TwLayout<L, R>::exponent(i)];
}

// Workgroups are dispatched to the 8 XCDs (each with its own L2) round-robin
// by workgroup id.  When K consecutive column tiles share 128-byte L2 lines
// (bricks narrower than a line), tiles k*K .. k*K+K-1 are given to workgroup
// ids that land on one XCD, so each line is fetched from HBM once instead of
// once per XCD.  A permutation of [0, n): full groups of 8*K ids are
// remapped, a trailing partial group is left as is.
constexpr int kXcds = 8;
template <int K>
__device__ __forceinline__ int xcd_tile_order(int i, int n) {
  if constexpr (K <= 1) {
    return i;
  } else {
    constexpr int G = kXcds * K;
    const int g = i / G;
    if ((g + 1) * G > n) return i;
    const int r = i - g * G;
    return g * G + (r % kXcds) * K + r / kXcds;
  }
}

// Wave sum of NV doubles, valid in lane 0 of each (possibly partial) wave,
// in a fixed order.  DPP row permutations fold the 16 lanes of each row
// (neighbour, pair of neighbours, half-row mirror, row mirror: every lane
// ends with its row's sum), then lane 0 adds the sums of rows 1-3 read with
// readlane.  A VALU move per 32-bit half and step instead of an LDS-unit
// permute keeps the chain short.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}
template <int NV, int NT>
__device__ __forceinline__ void wave_sum_lane0(double (&v)[NV]) {
  constexpr int WW = NT < 64 ? NT : 64;   // lanes of a (possibly partial) wave
  static_assert((WW & (WW - 1)) == 0, "power-of-two lane count");
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    if constexpr (WW >= 2) v[k] += dpp_f64<0xB1>(v[k]);    // quad_perm [1,0,3,2]
    if constexpr (WW >= 4) v[k] += dpp_f64<0x4E>(v[k]);    // quad_perm [2,3,0,1]
    if constexpr (WW >= 8) v[k] += dpp_f64<0x141>(v[k]);   // row_half_mirror
    if constexpr (WW >= 16) v[k] += dpp_f64<0x140>(v[k]);  // row_mirror
  }
  // every lane adds, so the row sums are read from the values before the adds
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double rowsum = v[k];
#pragma unroll
    for (int row = 1; row < WW / 16; ++row) v[k] += readlane_f64(rowsum, 16 * row);
  }
}

// Deterministic block sum of NV doubles; result valid in thread 0.
template <int NV, int NT>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red /* LDS, NV * waves */) {
  wave_sum_lane0<NV, NT>(v);
  constexpr int NW = (NT + 63) / 64;
  if constexpr (NW > 1) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
      for (int k = 0; k < NV; ++k) red[wv * NV + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        double s = red[k];
        for (int w = 1; w < NW; ++w) s += red[w * NV + k];
        v[k] = s;
      }
  }
}

// First half of a deterministic block sum that needs no barrier of its own:
// each wave sums its lanes and lane 0 parks the wave sums in
// red[wave * NV + k]; after a later barrier, lds_wave_sum adds them in wave
// order.
template <int NV, int NT>
__device__ __forceinline__ void wave_fold_to_lds(double (&v)[NV], double* red /* LDS, NV * waves */) {
  wave_sum_lane0<NV, NT>(v);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) red[(threadIdx.x >> 6) * NV + k] = v[k];
}
template <int NV, int NT>
__device__ __forceinline__ double lds_wave_sum(const double* red, int k) {
  double s = red[k];
#pragma unroll
  for (int w = 1; w < (NT + 63) / 64; ++w) s += red[w * NV + k];
  return s;
}

// ---------------------------------------------------------------------------
// Row pass.  Block = RB rows (one brick row, or half of one) x TR = NX2/R
// lanes per row.
// G = 3: the three C2R transforms share their LDS exchanges (and the two
// R2C transforms theirs); G = 1: one field at a time (less LDS).
// PP, option bits: 1 = two sets of LDS areas, one barrier per transform
// exchange (more LDS); 4 = streaming stores of P, Q (otherwise device-scope
// write-through stores); 32 = paired rows (RowPassPaired below: rows y and
// y + RB/2 of the block as the real and imaginary parts of one full-length
// complex transform); 64 = half-brick row blocks, 128 = P, Q stored in bricks
// twice as tall, 256 = U, V, W read from bricks of the row tile's height
// (RowTiles below).  U, V, W are read with streaming loads.
// ---------------------------------------------------------------------------
// Row tiles of a launch.  Without option 64 a row tile is one brick row (the
// brick height is RB); with it, bricks are 2 RB rows high and a row tile is
// the upper or lower half of a brick row, so a block holds half the rows (half
// the LDS, twice the blocks) while the column pass keeps its taller bricks.
// The two halves share the 128-byte lines of their brick row in every field
// stored in those bricks, so tiles 2b and 2b + 1 go to workgroup ids on one
// XCD (xcd_tile_order) and a line is fetched into (and written back from)
// one L2.
template <int RB, int BX, int NX2, int PP>
struct RowTiles {
  static constexpr bool kHalf = (PP & 64) != 0;
  static constexpr int BYB = kHalf ? 2 * RB : RB;   // brick height
  __device__ __forceinline__ static int order(int block, int nblocks) {
    if constexpr (kHalf) return xcd_tile_order<2>(block, nblocks);
    else return block;
  }
  // element offset of row yy of row tile `tile` at column 0
  __device__ __forceinline__ static unsigned base(int tile, int yy) {
    const int brow = kHalf ? tile >> 1 : tile;
    const int y0 = kHalf ? (tile & 1) * RB : 0;
    return (unsigned)brow * (NX2 / BX) * (BYB * BX) + (unsigned)(y0 + yy) * BX;
  }
  // offset of column k from column 0 of the same row
  __device__ __forceinline__ static unsigned col(int k) { return (unsigned)(k / BX) * (BYB * BX) + (k % BX); }
  // P, Q (written here, read by the column pass) live in bricks twice as tall
  // with option 128, and U, V, W (read here, written by the column pass) with
  // option 256 in bricks of the row tile's height, so that a tile reads one
  // contiguous run: the offsets of row yy of row tile `tile` and of column k
  // for a brick height H
  static constexpr int BYQ = (PP & 128) ? 2 * BYB : BYB;
  static constexpr int BYU = (PP & 256) ? RB : BYB;
  template <int H>
  __device__ __forceinline__ static unsigned base_h(int tile, int yy) {
    const unsigned y = (unsigned)(tile * RB + yy);   // the row (row tiles are RB rows)
    return (y / H) * (NX2 / BX) * (H * BX) + (y % H) * BX;
  }
  template <int H>
  __device__ __forceinline__ static unsigned col_h(int k) { return (unsigned)(k / BX) * (H * BX) + (k % BX); }
};

// RowPass<...>::run is the body of one block (block `block` of the `nblocks`
// of its class in the launch, block size NT, LDS `sm`); row_pass is the kernel
// of one size class, and a mixed launch (multi_pass.hpp) calls the bodies of
// several classes.
template <class T, int NX2, int RB, int R, int BX, int G, int PP>
struct RowPass {
  using Real = T;
  static constexpr int TR = NX2 / R;
  static constexpr int NT = RB * TR;
  using View = LdsView<T, NX2, RB, R, G, PP & 1>;
  // 16-byte aligned: FP64 exchanges then compile to ds_read_b128 /
  // ds_write_b128 (with the struct's natural 8-byte alignment the compiler
  // splits every 16-byte LDS access into two 8-byte ones)
  struct alignas(16) Smem {
    cx<T> lds_mem[View::kElems];
    cx<T> tw_h[TwLayout<NX2, R>::kAlloc];   // the half-length transforms
    cx<T> tw_n[NX2];                        // exp(-2 pi i k / Nx), k < Nx/2, the even/odd split
    double red[4 * ((NT + 63) / 64)];
  };
  __device__ static void run(const InstDesc<T>* __restrict__ descs, int tiles_per_inst,
                             const cx<T>* __restrict__ pool, int block, int nblocks, Smem& sm);
};

template <class T, int NX2, int RB, int R, int BX, int G, int PP>
struct RowPassPaired;

template <class T, int NX2, int RB, int R, int BX, int G, int PP>
__global__ void __launch_bounds__(RB * (NX2 / R))
row_pass(const InstDesc<T>* __restrict__ descs, int tiles_per_inst, const cx<T>* __restrict__ pool) {
  using P = std::conditional_t<(PP & 32) != 0, RowPassPaired<T, NX2, RB, R, BX, G, PP>, RowPass<T, NX2, RB, R, BX, G, PP>>;
  __shared__ typename P::Smem sm;
  P::run(descs, tiles_per_inst, pool, blockIdx.x, gridDim.x, sm);
}

template <class T, int NX2, int RB, int R, int BX, int G, int PP>
__device__ __forceinline__ void RowPass<T, NX2, RB, R, BX, G, PP>::run(const InstDesc<T>* __restrict__ descs,
                                                                      int tiles_per_inst,
                                                                      const cx<T>* __restrict__ pool, int block,
                                                                      int nblocks, Smem& sm) {
  static_assert(NX2 % BX == 0 && TR % BX == 0, "row lanes must cover whole bricks");
  static_assert(G == 1 || G == 3, "group of 1 or 3 fields");
  using Tiles = RowTiles<RB, BX, NX2, PP>;
  constexpr int GR = G == 3 ? 2 : 1;   // R2C group
  constexpr int kCP = (PP & 4) ? kCacheNt : kCacheSc1;
  cx<T>* const lds_mem = sm.lds_mem;
  cx<T>* const tw_h = sm.tw_h;
  cx<T>* const tw_n = sm.tw_n;
  double* const red = sm.red;

  const int bid = Tiles::order(block, nblocks);
  const int inst = bid / tiles_per_inst;
  const int tile = bid - inst * tiles_per_inst;
  const InstDesc<T>& d = descs[inst];
  if (d.nx == 0) return;   // an empty entry of an engine's descriptor array: the whole block exits
  // lane -> (row yy of the block, transform lane j); a wave still covers a
  // contiguous range of whole bricks (or half bricks)
  const int t = threadIdx.x;
  const int yy = t % RB;
  const int j = t / RB;
  View lds{lds_mem, yy};

  // element j + r*TR of this lane's row: lane offset base + col(j) plus
  // r*TR*BYB (TR is a whole number of bricks); fields hold at most 2^23
  // elements, so 32-bit offsets from a uniform base suffice
  constexpr unsigned ESZ = sizeof(cx<T>);
  const unsigned vb = (Tiles::template base_h<Tiles::BYU>(tile, yy) + Tiles::template col_h<Tiles::BYU>(j)) * ESZ;
  auto rb = [](int r) -> unsigned { return (unsigned)r * TR * Tiles::BYU * ESZ; };
  // the same for the P, Q layout
  const unsigned vbq = (Tiles::template base_h<Tiles::BYQ>(tile, yy) + Tiles::template col_h<Tiles::BYQ>(j)) * ESZ;
  auto rbq = [](int r) -> unsigned { return (unsigned)r * TR * Tiles::BYQ * ESZ; };
  // Field by field: loads complete in issue order, so the first field's
  // exchange (and with G = 1 its whole transform) waits for its own loads
  // only while the other two fields are still in flight.
  cx<T> f3[3][R];
  {
    const BufField<cx<T>> gum(d.um), gvm(d.vm), gwm(d.wm);
#pragma unroll
    for (int r = 0; r < R; ++r) f3[0][r] = gum.template ld<kCacheNt>(vb, rb(r));
#pragma unroll
    for (int r = 0; r < R; ++r) f3[1][r] = gvm.template ld<kCacheNt>(vb, rb(r));
#pragma unroll
    for (int r = 0; r < R; ++r) f3[2][r] = gwm.template ld<kCacheNt>(vb, rb(r));
  }
  stage_fft_twiddles<T, NT, NX2, R>(tw_h, tw_of(pool, NX2));
  stage_twiddles<T, NT>(tw_n, tw_of(pool, 2 * NX2), NX2);

  // With two LDS sets (PP) the grouped path puts the pair exchanges in the
  // set the neighbouring transform passes do not touch at that moment, which
  // drops three barriers: the one between the pre-exchange loads and the
  // first C2R stores (those go to set 0, the exchange uses set 1), the one
  // before the R2C transforms when the C2R's last loads read set 1, and the
  // one before the post-exchange stores (into the set the R2C's last loads
  // did not read; every earlier load of it finished before their barrier).
  constexpr bool kDbl = View::kBuf != 0;
  constexpr int kLastC2R = FftMulti<NX2, R, +1, 3, RB, T, View::kBuf>::kLastLoadSet;
  constexpr int kLastR2C = FftMulti<NX2, R, -1, 2, RB, T, View::kBuf>::kLastLoadSet;
  constexpr int kPreSet = (kDbl && G == 3) ? 1 : 0;
  constexpr int kPostSet = (kDbl && G == 3 && kLastR2C >= 0) ? 1 - kLastR2C : 0;

  // Row C2R of a half spectrum X[0..NX2) (X[NX2] = 0, Im X[0] ignored):
  // Z[k] = (X[k] + conj X[M-k]) + i e^{+2 pi i k/N} (X[k] - conj X[M-k]),
  // then inverse length-M transform -> (x[2n], x[2n+1]).
  auto c2r_pre = [&](cx<T>(&q)[R], int f) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = j + r * TR;
      cx<T> xk = q[r], xr;
      if (k == 0) {
        xk.im = (T)0;
        xr = cx<T>{(T)0, (T)0};
      } else {
        xr = cconj(lds.load_plain(f, NX2 - k, kPreSet));
      }
      const cx<T> s = xk + xr, dd = xk - xr;
      q[r] = s + cmuli(cmulc(dd, tw_n[k]));
    }
  };
  if constexpr (G == 3) {
#pragma unroll
    for (int f = 0; f < 3; ++f)
#pragma unroll
      for (int r = 0; r < R; ++r) lds.store_plain(f, j + r * TR, f3[f][r], kPreSet);
    __syncthreads();
#pragma unroll
    for (int f = 0; f < 3; ++f) c2r_pre(f3[f], f);
    if constexpr (!kDbl) __syncthreads();
    fft_multi<NX2, R, +1, 3>(f3, j, lds, tw_h);
  } else {
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      cx<T>(&q)[1][R] = *reinterpret_cast<cx<T>(*)[1][R]>(&f3[f]);
      __syncthreads();
#pragma unroll
      for (int r = 0; r < R; ++r) lds.store_plain(0, j + r * TR, q[0][r]);
      __syncthreads();
      c2r_pre(q[0], 0);
      __syncthreads();
      fft_multi<NX2, R, +1, 1>(q, j, lds, tw_h);
    }
  }

  // Real-space products (nonl_subfunc_1) and the two combined fields.  The
  // lane's 2R points are summed in T (sum |u|^2 reuses Q = |u|^2 of the cubic
  // term), then converted once: the block sum below and the sum over row
  // blocks (column pass, tile 0) are FP64 in a fixed order.  In FP32 a lane
  // sum of 16-32 FP32 products is as accurate as the products themselves.
  const T lam = (T)d.lam, bet = (T)d.beta;
  T esum_l = 0, gxs_l = 0, gys_l = 0;
  cx<T> f2[2][R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    T o1[2], o2[2];
    const T uxa[2] = {f3[0][r].re, f3[0][r].im}, uya[2] = {f3[1][r].re, f3[1][r].im};
    const T wa[2] = {f3[2][r].re, f3[2][r].im};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const T ux = uxa[h], uy = uya[h], wv = wa[h];
      const T A = wv * ux, Wp = wv * uy;
      const T Q = ux * ux + uy * uy;
      const T Cx = Q * ux, Cy = Q * uy;
      o1[h] = lam * A + bet * Cy;
      o2[h] = lam * Wp - bet * Cx;
      esum_l += Q;
      gxs_l += Cx;
      gys_l += Cy;
    }
    f2[0][r] = cx<T>{o1[0], o1[1]};
    f2[1][r] = cx<T>{o2[0], o2[1]};
  }
  const double esum = (double)esum_l, gxs = (double)gxs_l, gys = (double)gys_l;

  // Row R2C of packed real pairs: Z = FFT_M(z), then
  // 2 X[k] = (Z[k] + conj Z[M-k]) - i e^{-2 pi i k/N} (Z[k] - conj Z[M-k]).
  // P and Q are stored as twice the half spectra: the column pass applies
  // the 1/2 together with the dealiasing mask (an exact power-of-two scaling
  // that commutes with the transforms, so the results are unchanged).
  auto r2c_post = [&](cx<T>(&q)[R], int f) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = j + r * TR;
      const cx<T> zk = q[r], zm = cconj(lds.load_plain(f, (NX2 - k) & (NX2 - 1), kPostSet));
      q[r] = (zk + zm) - cmul(cmuli(zk - zm), tw_n[k]);
    }
  };
  // P = -i kx X1 (field 0), Q = X2 (field 1)
  const T tkx = (T)d.tkx;
  const BufField<cx<T>> gpm(d.pm), gqm(d.qm);
  auto store_out = [&](const cx<T>(&q)[R], int f) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (f == 0) {
        const T kx = tkx * (T)(j + r * TR);
        gpm.template st<kCP>(vbq, rbq(r), cx<T>{kx * q[r].im, -kx * q[r].re});
      } else {
        gqm.template st<kCP>(vbq, rbq(r), q[r]);
      }
    }
  };
  if constexpr (GR == 2) {
    if constexpr (!(kDbl && kLastC2R == 1)) __syncthreads();
    fft_multi<NX2, R, -1, 2>(f2, j, lds, tw_h);
    if constexpr (!kDbl || kLastR2C < 0) __syncthreads();
#pragma unroll
    for (int f = 0; f < 2; ++f)
#pragma unroll
      for (int r = 0; r < R; ++r) lds.store_plain(f, j + r * TR, f2[f][r], kPostSet);
    __syncthreads();
#pragma unroll
    for (int f = 0; f < 2; ++f) r2c_post(f2[f], f);
#pragma unroll
    for (int f = 0; f < 2; ++f) store_out(f2[f], f);
  } else {
    // each field is stored as soon as it is done, so P's stores drain while
    // Q is transformed
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      cx<T>(&q)[1][R] = *reinterpret_cast<cx<T>(*)[1][R]>(&f2[f]);
      __syncthreads();
      fft_multi<NX2, R, -1, 1>(q, j, lds, tw_h);
      __syncthreads();
#pragma unroll
      for (int r = 0; r < R; ++r) lds.store_plain(0, j + r * TR, q[0][r]);
      __syncthreads();
      r2c_post(q[0], 0);
      store_out(q[0], f);
    }
  }

  double sums[3] = {esum, gxs, gys};
  block_sum<3, NT>(sums, red);
  if (t == 0) {
    double* p = d.partials + (size_t)tile * 4;
    p[0] = sums[0];
    p[1] = sums[1];
    p[2] = sums[2];
  }
}

// ---------------------------------------------------------------------------
// Row pass with paired rows (option bit 32).  Rows a = yy and b = yy + RB/2
// of the block's brick row travel together as z = a + i b through one
// length-N (N = Nx) complex transform, so the real-FFT even/odd split and its
// twiddles are not needed:
//   C2R: Z[k] = A[k] + i B[k] and Z[N-k] = conj A[k] + i conj B[k] for
//        0 < k < N/2, Z[0] = Re A[0] + i Re B[0], Z[N/2] = 0; z = IFFT_N(Z)
//        gives a in Re z, b in Im z (the same unnormalised inverse as the
//        half-length path);
//   R2C: Z = FFT_N(p_a + i p_b), 2A[k] = Z[k] + conj Z[N-k],
//        2B[k] = -i (Z[k] - conj Z[N-k]) for k < N/2 (twice the half
//        spectra, as the half-length path stores them).
// Lane t -> pair yy = t % (RB/2), transform lane j = t / (RB/2); a lane holds
// positions j + r*S (S = N/R) as in fft_multi: r < R/2 are the half-spectrum
// columns it loads and stores, r >= R/2 their mirror images, exchanged
// through LDS.  Products, partial sums and the outputs are those of RowPass.
template <class T, int NX2, int RB, int R, int BX, int G, int PP>
struct RowPassPaired {
  using Real = T;
  static constexpr int N = 2 * NX2;         // full row length
  static constexpr int W2 = RB / 2;         // row pairs of the block
  static constexpr int S = N / R;           // lanes per pair
  static constexpr int NT = W2 * S;         // = RB * NX2 / R, the block of RowPass
  static_assert(RB % 2 == 0 && R % 2 == 0, "pairs of rows, half of each lane's points mirrored");
  using View = LdsView<T, N, W2, R, G, PP & 1>;
  struct alignas(16) Smem {
    cx<T> lds_mem[View::kElems];
    cx<T> tw[TwLayout<N, R>::kAlloc];
    double red[4 * ((NT + 63) / 64)];
  };
  __device__ static void run(const InstDesc<T>* __restrict__ descs, int tiles_per_inst,
                             const cx<T>* __restrict__ pool, int block, int nblocks, Smem& sm);
};

template <class T, int NX2, int RB, int R, int BX, int G, int PP>
__device__ __forceinline__ void RowPassPaired<T, NX2, RB, R, BX, G, PP>::run(const InstDesc<T>* __restrict__ descs,
                                                                            int tiles_per_inst,
                                                                            const cx<T>* __restrict__ pool, int block,
                                                                            int nblocks, Smem& sm) {
  static_assert(G == 1 || G == 3, "group of 1 or 3 fields");
  static_assert(S % BX == 0, "pair lanes cover whole bricks");
  using Tiles = RowTiles<RB, BX, NX2, PP>;
  static_assert(Tiles::BYQ == Tiles::BYB && Tiles::BYU == Tiles::BYB, "paired rows: one brick shape for all fields");
  constexpr int GR = G == 3 ? 2 : 1;   // R2C group
  constexpr int H = R / 2;             // loaded / stored points per lane and row
  constexpr int kCP = (PP & 4) ? kCacheNt : kCacheSc1;
  const int bid = Tiles::order(block, nblocks);
  const int inst = bid / tiles_per_inst;
  const int tile = bid - inst * tiles_per_inst;
  const InstDesc<T>& d = descs[inst];
  if (d.nx == 0) return;   // an empty entry of an engine's descriptor array: the whole block exits
  const int t = threadIdx.x;
  const int yy = t % W2;   // pair: rows yy and yy + W2
  const int j = t / W2;
  View lds{sm.lds_mem, yy};

  // brick offsets as in RowPass; row b sits W2 rows below row a in every brick
  constexpr unsigned ESZ = sizeof(cx<T>);
  const unsigned va = (Tiles::base(tile, yy) + Tiles::col(j)) * ESZ, vbo = (unsigned)(W2 * BX) * ESZ;
  auto rb = [](int r) -> unsigned { return (unsigned)r * S * Tiles::BYB * ESZ; };   // column j + r*S

  // f[field][r]: r < H: A at column j + r*S; r >= H: B at column j + (r-H)*S
  // field by field, as in RowPass
  cx<T> f3[3][R];
  {
    const BufField<cx<T>> gf[3] = {BufField<cx<T>>(d.um), BufField<cx<T>>(d.vm), BufField<cx<T>>(d.wm)};
#pragma unroll
    for (int f = 0; f < 3; ++f)
#pragma unroll
      for (int r = 0; r < H; ++r) {
        f3[f][r] = gf[f].template ld<kCacheNt>(va, rb(r));
        f3[f][r + H] = gf[f].template ld<kCacheNt>(va + vbo, rb(r));
      }
  }
  stage_fft_twiddles<T, NT, N, R>(sm.tw, tw_of(pool, N));

  // C2R assembly of one field: A, B of column k stored at k and N/2 + k; the
  // lane's mirrored positions p = j + r*S (r >= H, p > N/2) read column N - p.
  auto c2r_store = [&](cx<T>(&q)[R], int f) {
#pragma unroll
    for (int r = 0; r < H; ++r) {
      lds.store_plain(f, j + r * S, q[r]);
      lds.store_plain(f, NX2 + j + r * S, q[r + H]);
    }
  };
  auto c2r_assemble = [&](cx<T>(&q)[R], int f) {
    cx<T> z[R];
#pragma unroll
    for (int r = 0; r < H; ++r) {
      cx<T> a = q[r], b = q[r + H];
      if (r == 0 && j == 0) { a.im = (T)0; b.im = (T)0; }   // column 0: imaginary parts ignored
      z[r] = cx<T>{a.re - b.im, a.im + b.re};
    }
#pragma unroll
    for (int r = H; r < R; ++r) {
      const int p = j + r * S;   // > N/2 except p == N/2 (j == 0, r == H): the Nyquist column, zero
      if (r == H && j == 0) {
        z[r] = cx<T>{(T)0, (T)0};
      } else {
        const cx<T> a = lds.load_plain(f, N - p), b = lds.load_plain(f, NX2 + N - p);
        z[r] = cx<T>{a.re + b.im, b.re - a.im};
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) q[r] = z[r];
  };
  if constexpr (G == 3) {
#pragma unroll
    for (int f = 0; f < 3; ++f) c2r_store(f3[f], f);
    __syncthreads();
#pragma unroll
    for (int f = 0; f < 3; ++f) c2r_assemble(f3[f], f);
    __syncthreads();
    fft_multi<N, R, +1, 3>(f3, j, lds, sm.tw);
  } else {
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      cx<T>(&q)[1][R] = *reinterpret_cast<cx<T>(*)[1][R]>(&f3[f]);
      __syncthreads();
      c2r_store(q[0], 0);
      __syncthreads();
      c2r_assemble(q[0], 0);
      __syncthreads();
      fft_multi<N, R, +1, 1>(q, j, lds, sm.tw);
    }
  }

  // products: (re, im) of each point are (row a, row b) at one x
  const T lam = (T)d.lam, bet = (T)d.beta;
  T esum_l = 0, gxs_l = 0, gys_l = 0;
  cx<T> f2[2][R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    T o1[2], o2[2];
    const T uxa[2] = {f3[0][r].re, f3[0][r].im}, uya[2] = {f3[1][r].re, f3[1][r].im};
    const T wa[2] = {f3[2][r].re, f3[2][r].im};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const T ux = uxa[h], uy = uya[h], wv = wa[h];
      const T A = wv * ux, Wp = wv * uy;
      const T Q = ux * ux + uy * uy;
      const T Cx = Q * ux, Cy = Q * uy;
      o1[h] = lam * A + bet * Cy;
      o2[h] = lam * Wp - bet * Cx;
      esum_l += Q;
      gxs_l += Cx;
      gys_l += Cy;
    }
    f2[0][r] = cx<T>{o1[0], o1[1]};
    f2[1][r] = cx<T>{o2[0], o2[1]};
  }
  const double esum = (double)esum_l, gxs = (double)gxs_l, gys = (double)gys_l;

  // R2C: forward transforms, then the mirror exchange (the lane's r < H
  // columns k read Z[N - k]; column 0 reads itself)
  const T tkx = (T)d.tkx;
  const BufField<cx<T>> gpm(d.pm), gqm(d.qm);
  auto r2c_out = [&](cx<T>(&q)[R], int f, const BufField<cx<T>>& g, bool is_p) {
#pragma unroll
    for (int r = 0; r < H; ++r) {
      const int k = j + r * S;
      const cx<T> zk = q[r], zm = cconj(lds.load_plain(f, (N - k) & (N - 1)));
      cx<T> a2 = zk + zm;                 // 2A[k]
      cx<T> b2 = cmuli(zm - zk);          // 2B[k] = -i (Z[k] - conj Z[N-k])
      if (is_p) {                         // P = -i kx (twice the half spectrum)
        const T kx = tkx * (T)k;
        a2 = cx<T>{kx * a2.im, -kx * a2.re};
        b2 = cx<T>{kx * b2.im, -kx * b2.re};
      }
      g.template st<kCP>(va, rb(r), a2);
      g.template st<kCP>(va + vbo, rb(r), b2);
    }
  };
  if constexpr (GR == 2) {
    __syncthreads();
    fft_multi<N, R, -1, 2>(f2, j, lds, sm.tw);
    __syncthreads();
#pragma unroll
    for (int f = 0; f < 2; ++f)
#pragma unroll
      for (int r = 0; r < R; ++r) lds.store_plain(f, j + r * S, f2[f][r]);
    __syncthreads();
    r2c_out(f2[0], 0, gpm, true);
    r2c_out(f2[1], 1, gqm, false);
  } else {
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      cx<T>(&q)[1][R] = *reinterpret_cast<cx<T>(*)[1][R]>(&f2[f]);
      __syncthreads();
      fft_multi<N, R, -1, 1>(q, j, lds, sm.tw);
      __syncthreads();
#pragma unroll
      for (int r = 0; r < R; ++r) lds.store_plain(0, j + r * S, q[0][r]);
      __syncthreads();
      r2c_out(q[0], 0, f == 0 ? gpm : gqm, f == 0);
    }
  }

  double sums[3] = {esum, gxs, gys};
  block_sum<3, NT>(sums, sm.red);
  if (t == 0) {
    double* p = d.partials + (size_t)tile * 4;
    p[0] = sums[0];
    p[1] = sums[1];
    p[2] = sums[2];
  }
}

// ---------------------------------------------------------------------------
// Column pass.  Block = W adjacent columns (of one BX-wide brick column) x
// TC = NY/R lanes per column.
// stage 0: initial pass only (v^ from the state, inverse transforms).
// stage 1..4: full stage.  The stage is a template argument: each stage's
// kernel keeps only the state it reads live (fewer registers, more waves).  GF = 2 / GI = 3: the forward (P, Q) and inverse
// (v_x, v_y, omega) transforms share LDS exchanges; 1: one at a time.
// PP, option bits: 1 = double-buffered LDS exchanges; 2 = streaming stores;
// 8 = streaming loads of P, Q; 16 = FP32 integrating factors from the table
// at radix 16 too; 128 = P, Q in bricks twice as tall as the other fields';
// 256 = U, V, W in bricks half as tall.
// ---------------------------------------------------------------------------
// ColPass<...>::run is the body of one block (block `block` of `nblocks` of
// the class in the launch); col_pass is the kernel of one size class.
template <class T, int NY, int W, int R, int BY, int BX, int GF, int GI, int PP, int stage>
struct ColPass {
  using Real = T;
  static constexpr int TC = NY / R;
  static constexpr int NT = W * TC;
  static constexpr int GMAX = GF > GI ? GF : GI;
  using View = LdsView<T, NY, W, R, GMAX, PP & 1>;
  struct alignas(16) Smem {   // 16-byte aligned: see RowPass::Smem
    cx<T> lds_mem[View::kElems];
    cx<T> tw_y[TwLayout<NY, R>::kAlloc];
    double mean_in[4];                       // tile 0: M, M_old
    double mean_red[3 * ((NT + 63) / 64)];   // tile 0: wave sums of the row partials
    long long nstep_in;
  };
  __device__ static void run(const InstDesc<T>* __restrict__ descs, int tiles_per_inst,
                             const cx<T>* __restrict__ pool, int block, int nblocks, Smem& sm);
};

template <class T, int NY, int W, int R, int BY, int BX, int GF, int GI, int PP, int stage>
__global__ void __launch_bounds__(W * (NY / R))
col_pass(const InstDesc<T>* __restrict__ descs, int tiles_per_inst, const cx<T>* __restrict__ pool) {
  using P = ColPass<T, NY, W, R, BY, BX, GF, GI, PP, stage>;
  __shared__ typename P::Smem sm;
  P::run(descs, tiles_per_inst, pool, blockIdx.x, gridDim.x, sm);
}

template <class T, int NY, int W, int R, int BY, int BX, int GF, int GI, int PP, int stage>
__device__ __forceinline__ void ColPass<T, NY, W, R, BY, BX, GF, GI, PP, stage>::run(
    const InstDesc<T>* __restrict__ descs, int tiles_per_inst, const cx<T>* __restrict__ pool, int block,
    int nblocks, Smem& sm) {
  static_assert((GF == 1 || GF == 2) && (GI == 1 || GI == 3), "field groups");
  constexpr int kCP = (PP & 2) ? kCacheNt : kCacheDefault;     // field stores
  constexpr int kPqCP = (PP & 8) ? kCacheNt : kCacheDefault;   // P, Q loads
  static_assert(NY > R, "every column transform has a barrier before its first LDS or twiddle read");
  // With two LDS sets (PP), a transform whose predecessor's last loads read
  // set 1 needs no barrier before it: its first stores go to set 0.  Tile 0
  // keeps the barrier after the projection, whose exchange reads set 0.
  constexpr bool kSkipPreFft = View::kBuf != 0 && FftMulti<NY, R, -1, 1, W, T, View::kBuf>::kLastLoadSet == 1;
  cx<T>* const lds_mem = sm.lds_mem;
  cx<T>* const tw_y = sm.tw_y;
  double* const mean_in = sm.mean_in;
  double* const mean_red = sm.mean_red;
  long long& nstep_in = sm.nstep_in;

  // a tile reads BY x W elements of each BY-row band: how many column tiles
  // share a 128-byte line
  constexpr int kLine = 128 / (BY * W * (int)sizeof(cx<T>));
  const int bid = xcd_tile_order<kLine>(block, nblocks);
  const int inst = bid / tiles_per_inst;
  const int tile = bid - inst * tiles_per_inst;
  const InstDesc<T>& d = descs[inst];
  if (d.nx == 0) return;   // an empty entry of an engine's descriptor array: the whole block exits
  const int t = threadIdx.x;
  const int c = t % W, j = t / W;
  const int x = tile * W + c;
  const int nx2 = d.nx >> 1;
  View lds{lds_mem, c};

  // bricks of BY x BX; this block's W columns sit inside one brick column
  static_assert(BX % W == 0, "column block within one brick column");
  const unsigned col_base = (unsigned)(tile * W / BX) * (BY * BX) + (unsigned)(tile * W % BX) + c;
  const unsigned brow = (unsigned)nx2 * BY;   // elements per brick row
  auto addr = [&](int y) -> unsigned { return col_base + (unsigned)(y / BY) * brow + (y % BY) * BX; };
  // Element j + r*TC of this lane's column: a lane offset plus a uniform
  // offset per r (TC is a whole number of bricks for every schedule).
  static_assert(TC % BY == 0, "column lanes span whole bricks");
  constexpr unsigned ESZ = sizeof(cx<T>);
  const unsigned vb = addr(j) * ESZ;
  const unsigned rstride = (unsigned)(TC / BY) * brow * ESZ;   // bytes, uniform
  auto ld = [&](const cx<T>* f, int r) -> cx<T> { return BufField<cx<T>>(f).ld(vb, (unsigned)r * rstride); };
  // P, Q in bricks twice as tall with option 128 (written so by the row pass)
  constexpr int BYQ = (PP & 128) ? 2 * BY : BY;
  static_assert(TC % BYQ == 0, "column lanes span whole P, Q bricks");
  const unsigned brow_q = (unsigned)nx2 * BYQ;
  const unsigned vbq =
      ((unsigned)(tile * W / BX) * (BYQ * BX) + (unsigned)(tile * W % BX) + c + (unsigned)(j / BYQ) * brow_q +
       (unsigned)(j % BYQ) * BX) * ESZ;
  const unsigned rstride_q = (unsigned)(TC / BYQ) * brow_q * ESZ;
  // U, V, W in bricks of half the height with option 256 (the row pass's
  // half-brick tiles then read contiguous brick rows)
  constexpr int BYU = (PP & 256) ? BY / 2 : BY;
  static_assert(TC % BYU == 0 && BYU >= 1, "column lanes span whole U, V, W bricks");
  const unsigned brow_u = (unsigned)nx2 * BYU;
  const unsigned vbu =
      ((unsigned)(tile * W / BX) * (BYU * BX) + (unsigned)(tile * W % BX) + c + (unsigned)(j / BYU) * brow_u +
       (unsigned)(j % BYU) * BX) * ESZ;
  const unsigned rstride_u = (unsigned)(TC / BYU) * brow_u * ESZ;
  auto st_u = [&](cx<T>* f, int r, cx<T> v) { BufField<cx<T>>(f).template st<kCP>(vbu, (unsigned)r * rstride_u, v); };
  auto st = [&](cx<T>* f, int r, cx<T> v) { BufField<cx<T>>(f).template st<kCP>(vb, (unsigned)r * rstride, v); };

  const double dt = d.dt, dt2 = d.dt2, dt3 = d.dt3, dt6 = d.dt6;
  const bool col0 = (x == 0);
  const double kxd = d.tkx * (double)x;
  const double kx2 = kxd * kxd;
  const int ymax = mask_ymax(x, d.nx, d.ny, d.mask_mode);
  cx<T> sv[R];   // spectral input of the next stage
  double minj[2] = {0.0, 0.0};

  if constexpr (stage == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) sv[r] = ld(d.wo, r);
    stage_fft_twiddles<T, NT, NY, R>(tw_y, tw_of(pool, NY));
    if (tile == 0 && t == 0) {
      minj[0] = d.mean->M[0];
      minj[1] = d.mean->M[1];
    }
    __syncthreads();
  } else {
    cx<T> pq[2][R], wv[R], av[R];
    // Tile 0: the row-block partials of the mean ODE, spread over every lane
    // of the block; each wave folds them into LDS, and the ODE adds the wave
    // sums in wave order (a fixed order).  The
    // mean state waits in LDS too, so nothing stays in registers across the
    // transforms and no lane holds the block up before its first barrier.
#pragma unroll
    for (int r = 0; r < R; ++r) {
      pq[0][r] = BufField<cx<T>>(d.pm).template ld<kPqCP>(vbq, (unsigned)r * rstride_q);
      pq[1][r] = BufField<cx<T>>(d.qm).template ld<kPqCP>(vbq, (unsigned)r * rstride_q);
    }
    if (tile == 0) {
      double ps[3] = {0.0, 0.0, 0.0};
      const gptr<double> pp = glob(d.partials);
      for (int i = t; i < d.row_tiles; i += NT) {
        ps[0] += pp[(unsigned)i * 4 + 0];
        ps[1] += pp[(unsigned)i * 4 + 1];
        ps[2] += pp[(unsigned)i * 4 + 2];
      }
      wave_fold_to_lds<3, NT>(ps, mean_red);
      if (t == 0) {
        mean_in[0] = d.mean->M[0];
        mean_in[1] = d.mean->M[1];
        mean_in[2] = d.mean->Mo[0];
        mean_in[3] = d.mean->Mo[1];
        nstep_in = d.mean->nstep;
      }
    }
    // av: the projection source (acc, or omega_old at stage 1, whose
    // integrating factor is even in ky); wv: omega_old for stages 1-3.
    auto load_av = [&]() {
#pragma unroll
      for (int r = 0; r < R; ++r) av[r] = ld((stage == 1) ? d.wo : d.acc, r);
    };
    // FP32 with radix <= 8 (or option 16): integrating factors from the run's
    // table (FP64-evaluated, rounded once), loaded with omega_old; otherwise
    // (FP64, or radix 16 in a latency-bound single run, where the extra loads
    // cost more than the FP64 exp they save) evaluated in the update.
    constexpr bool kTable = sizeof(T) == 4 && (R <= 8 || (PP & 16));
    float ef[2][R];
    auto load_wv = [&]() {
      if (stage == 2 || stage == 3) {
#pragma unroll
        for (int r = 0; r < R; ++r) wv[r] = ld(d.wo, r);
      }
      if constexpr (kTable) {
        if (stage <= 3) {
          const BufField<float> tf(d.ifac);
          const unsigned vbf = ((unsigned)tile * NY * W + (unsigned)t) * 4u, rsf = (unsigned)TC * W * 4u;
          const unsigned plane = (unsigned)d.ny * (unsigned)nx2 * 4u;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            ef[0][r] = tf.ld(vbf, (unsigned)r * rsf);
            if (stage != 2) ef[1][r] = tf.ld(vbf, plane + (unsigned)r * rsf);
          }
        }
      }
    };
    // Hermitian projection of the kx = 0 column (reality_func, .cu:1237-1242):
    // a(y) <- (a(y) + conj a(Ny - y)) / 2 for y not 0 or Ny/2.  The column
    // lives in tile 0's block, so partners are exchanged through LDS (free
    // between transforms; the caller places a barrier before the next one).
    auto project = [&]() {
      if (stage == 1) {
#pragma unroll
        for (int r = 0; r < R; ++r) wv[r] = av[r];
      }
      if (tile == 0) {
        cx<T>* xcol = lds_mem;
        __syncthreads();
        if (c == 0) {
#pragma unroll
          for (int r = 0; r < R; ++r) xcol[j + r * TC] = av[r];
        }
        __syncthreads();
        if (col0) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int y = j + r * TC;
            if ((y & (NY / 2 - 1)) != 0) av[r] = cscale((T)0.5, av[r] + cconj(xcol[(NY - y) & (NY - 1)]));
          }
        }
      }
    };
    // Register pressure: with one forward transform at a time, wv is loaded
    // and the projection done between the two so that P^, av and wv are not
    // all live across a transform.
    load_av();
    if constexpr (GF == 2) load_wv();
    // no barrier here: the transform's first twiddle and LDS reads follow one
    stage_fft_twiddles<T, NT, NY, R>(tw_y, tw_of(pool, NY));
    if constexpr (GF == 2) {
      fft_multi<NY, R, -1, 2>(pq, j, lds, tw_y);
      project();
    } else {
      fft_multi<NY, R, -1, 1>(*reinterpret_cast<cx<T>(*)[1][R]>(&pq[0]), j, lds, tw_y);
      project();
      load_wv();
      if (!kSkipPreFft || tile == 0) __syncthreads();
      fft_multi<NY, R, -1, 1>(*reinterpret_cast<cx<T>(*)[1][R]>(&pq[1]), j, lds, tw_y);
    }
    cx<T>* const dst = (stage == 4) ? d.wo : d.acc;
    cx<T> a00 = {(T)0, (T)0};   // new acc (state at stage 4) of this lane's r = 0 mode
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int y = j + r * TC;
      const int nys = (y <= NY / 2) ? y : y - NY;
      const double kyd = d.tky * (double)nys;
      const double k2 = kx2 + kyd * kyd;
      // the row pass stores 2 P, 2 Q: the mask factor carries the 1/2
      const T keepf = ((nys < 0 ? -nys : nys) <= ymax) ? (T)0.5 : (T)0;
      // N^ = (P^ - i ky Q^) * mask
      const cx<T> nh = cscale(keepf, pq[0][r] + cmuli(cscale((T)(-kyd), pq[1][r])));
      T E = (T)1, E2 = (T)1;
      if (stage <= 3) {
        if constexpr (kTable) {
          E2 = ef[0][r];
          E = ef[1][r];
        } else {
          const double sig = d.c0 + d.c2 * k2 + d.c4 * (k2 * k2);
          const double e2d = exp(sig * dt * 0.5);
          E2 = (T)e2d;
          E = (T)(e2d * e2d);
        }
      }
      cx<T> a = (stage == 1) ? cscale(E, av[r]) : av[r];
      cx<T> s;
      if (stage == 1) {
        a = a + cscale((T)(dt6 * (double)E), nh);
        s = cscale(E2, wv[r] + cscale((T)dt2, nh));
      } else if (stage == 2) {
        a = a + cscale((T)(dt3 * (double)E2), nh);
        s = cscale(E2, wv[r] + cscale((T)dt2, nh));
      } else if (stage == 3) {
        a = a + cscale((T)(dt3 * (double)E2), nh);
        s = cscale(E, wv[r] + cscale((T)dt, nh));
      } else {
        a = a + cscale((T)dt6, nh);
        s = a;
      }
      st(dst, r, a);
      if (r == 0) a00 = a;
      sv[r] = s;
    }

    // Mean-velocity ODE (host updates of .cu:729-821), FP64, tile 0 only.
    if (tile == 0 && t == 0) {
      MeanState* ms = d.mean;
      const double s0 = lds_wave_sum<3, NT>(mean_red, 0);
      const long long nstep = nstep_in;
      const double g[2] = {-d.beta * lds_wave_sum<3, NT>(mean_red, 1), -d.beta * lds_wave_sum<3, NT>(mean_red, 2)};
      double M[2] = {mean_in[0], mean_in[1]}, Mo[2] = {mean_in[2], mean_in[3]}, mm[2];
      if (stage == 1) {
        d.energy[nstep % d.ring_cap] = s0 * d.invN * 0.5;
        Mo[0] = M[0];
        Mo[1] = M[1];
        if (d.mean_scheme == 1) { M[0] = d.e * Mo[0]; M[1] = d.e * Mo[1]; }
      }
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        if (stage == 1) { M[i] += dt6 * d.e * g[i];  mm[i] = d.e2 * (Mo[i] + dt2 * g[i]); }
        else if (stage == 2) { M[i] += dt3 * d.e2 * g[i]; mm[i] = d.e2 * (Mo[i] + dt2 * g[i]); }
        else if (stage == 3) { M[i] += dt3 * d.e2 * g[i]; mm[i] = d.e * (Mo[i] + dt * g[i]); }
        else { M[i] += dt6 * g[i]; mm[i] = M[i]; }
      }
      ms->M[0] = M[0]; ms->M[1] = M[1];
      ms->Mo[0] = Mo[0]; ms->Mo[1] = Mo[1];
      ms->m[0] = mm[0]; ms->m[1] = mm[1];
      if (stage == 4) {
        // check_func (.cu:884-897) on the new Re w^[0], owned by this lane
        if (ms->nan_step < 0 && a00.re != a00.re) ms->nan_step = nstep + 1;
        ms->nstep = nstep + 1;
      }
      minj[0] = mm[0];
      minj[1] = mm[1];
    }
  }

  // Biot-Savart + mean injection + three column inverse transforms:
  // v_x^ = i ky s^/k^2, v_y^ = -i kx s^/k^2 (nonl_subfunc_0, .cu:1176-1180),
  // scaled by 1/N so the row pass sees physical fields.
  T fx[R], fy[R];
  const T kxT = (T)kxd, tkyT = (T)d.tky, invNT = (T)d.invN;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int y = j + r * TC;
    const int nys = (y <= NY / 2) ? y : y - NY;
    const T ky = tkyT * (T)nys;
    const T k2 = kxT * kxT + ky * ky;
    const T ik = (k2 > (T)0) ? invNT * recip(k2) : (T)0;
    fx[r] = ky * ik;
    fy[r] = -kxT * ik;
  }
  const T invN = (T)d.invN;
  auto make = [&](cx<T>(&q)[R], int f) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (f == 2) {
        q[r] = cscale(invN, sv[r]);
      } else if (col0 && j + r * TC == 0) {
        q[r] = cx<T>{(T)(minj[f] * d.invN), (T)0};
      } else {
        q[r] = cscale(f == 0 ? fx[r] : fy[r], cmuli(sv[r]));
      }
    }
  };
  if constexpr (GI == 3) {
    cx<T> q3[3][R];
#pragma unroll
    for (int f = 0; f < 3; ++f) make(q3[f], f);
    if (!kSkipPreFft || stage == 0 || tile == 0) __syncthreads();
    fft_multi<NY, R, +1, 3>(q3, j, lds, tw_y);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      st_u(d.um, r, q3[0][r]);
      st_u(d.vm, r, q3[1][r]);
      st_u(d.wm, r, q3[2][r]);
    }
  } else {
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      cx<T> q[1][R];
      make(q[0], f);
      if (!kSkipPreFft || (f == 0 && (stage == 0 || tile == 0))) __syncthreads();
      fft_multi<NY, R, +1, 1>(q, j, lds, tw_y);
      cx<T>* const dst = (f == 0) ? d.um : (f == 1) ? d.vm : d.wm;
#pragma unroll
      for (int r = 0; r < R; ++r) st_u(dst, r, q[0][r]);
    }
  }
}

// Reference layout (Ny x (Nx/2+1), interleaved doubles) <-> bricks.
__device__ __forceinline__ size_t brick_index(int y, int x, int nx2, int by, int bx) {
  return ((size_t)(y / by) * (nx2 / bx) + x / bx) * (size_t)(by * bx) + (y % by) * bx + x % bx;
}
// Integrating factors of every mode of an FP32 run, evaluated in FP64 exactly
// as the column pass of an FP64 run does and rounded once to FP32.  Stored in
// the order the column pass reads them: plane 0 exp(sigma dt/2), plane 1
// exp(sigma dt), each [column tile][y][column in tile] for tiles of W columns,
// so a column block reads one contiguous run.  Instantiated for T = float.
template <class T>
__global__ void ifac_init(const InstDesc<T>* __restrict__ dp, int nx2, int ny, int w) {
  const InstDesc<T>& d = *dp;
  const size_t n = (size_t)ny * nx2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(i / nx2), x = (int)(i % nx2);
    const int nys = (y <= ny / 2) ? y : y - ny;
    const double kxd = d.tkx * (double)x;
    const double kyd = d.tky * (double)nys;
    const double k2 = kxd * kxd + kyd * kyd;
    const double sig = d.c0 + d.c2 * k2 + d.c4 * (k2 * k2);
    const double e2d = exp(sig * d.dt * 0.5);
    const size_t b = ((size_t)(x / w) * ny + y) * w + x % w;
    d.ifac[b] = (float)e2d;
    d.ifac[n + b] = (float)(e2d * e2d);
  }
}

template <class T>
__global__ void to_bricks(const double* __restrict__ src, cx<T>* __restrict__ dst, int nx, int ny, int by,
                          int bx) {
  const int nx2 = nx / 2, nxh = nx2 + 1;
  const size_t n = (size_t)ny * nx2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(i / nx2), x = (int)(i % nx2);
    const size_t s = (size_t)y * nxh + x;
    dst[brick_index(y, x, nx2, by, bx)] = cx<T>{(T)src[2 * s], (T)src[2 * s + 1]};
  }
}
template <class T>
__global__ void from_bricks(const cx<T>* __restrict__ src, double* __restrict__ dst, int nx, int ny, int by,
                            int bx) {
  const int nx2 = nx / 2, nxh = nx2 + 1;
  const size_t n = (size_t)ny * nxh;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(i / nxh), x = (int)(i % nxh);
    double re = 0, im = 0;
    if (x < nx2) {
      const cx<T> v = src[brick_index(y, x, nx2, by, bx)];
      re = (double)v.re;
      im = (double)v.im;
    }
    dst[2 * i] = re;
    dst[2 * i + 1] = im;
  }
}

}  // namespace af
