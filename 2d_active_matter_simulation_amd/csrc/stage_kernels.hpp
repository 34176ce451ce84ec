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
//     writes P, Q
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
#include "fft_lds.hpp"
#include "layout.hpp"

namespace af {

__device__ __forceinline__ bool keep_mode(int x, int yabs, int nx, int ny, int mode) {
  const long long X = x, Y = yabs, NX = nx, NY = ny;
  if (mode == 0) return 9 * (X * X * NY * NY + Y * Y * NX * NX) <= NX * NX * NY * NY;
  return 3 * X <= NX && 3 * Y <= NY;
}

// LDS view for the column pass: W interleaved columns, one pad element every
// R*W elements to spread the first pass's stride-R stores over the banks.
template <class T, int W, int R>
struct LdsCol {
  cx<T>* base;
  int c;
  __device__ __forceinline__ int a(int p) const {
    const int i = p * W + c;
    return i + i / (R * W);
  }
  __device__ __forceinline__ cx<T> load(int p) const { return base[a(p)]; }
  __device__ __forceinline__ void store(int p, cx<T> v) { base[a(p)] = v; }
};
template <int NY, int W, int R> constexpr int lds_col_elems() { return NY * W + NY * W / (R * W) + 1; }

// LDS view for the row pass: RB interleaved rows.
template <class T, int RB, int R>
struct LdsRow {
  cx<T>* base;
  int yy;
  __device__ __forceinline__ int a(int p) const {
    const int i = p * RB + yy;
    return i + i / (R * RB);
  }
  __device__ __forceinline__ cx<T> load(int p) const { return base[a(p)]; }
  __device__ __forceinline__ void store(int p, cx<T> v) { base[a(p)] = v; }
};
template <int NX2, int RB, int R> constexpr int lds_row_elems() { return NX2 * RB + NX2 * RB / (R * RB) + 1; }

// Twiddle pool: tw[L + m] = exp(-2 pi i m / L) for every power of two L.
template <class T>
__device__ __forceinline__ const cx<T>* tw_of(const cx<T>* pool, int L) { return pool + L; }

template <class T>
__device__ __forceinline__ T sq(T a) { return a * a; }

// Deterministic block sum of NV doubles; result valid in thread 0.
template <int NV, int NT>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red /* LDS, NV*NT/64 */) {
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
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

// ---------------------------------------------------------------------------
// Row pass.  Block = RB rows (one brick row) x TR = NX2/R lanes per row.
// Thread mapping keeps a wave's loads contiguous in the brick layout.
// ---------------------------------------------------------------------------
template <class T, int NX2, int RB, int R, int BX>
__global__ void __launch_bounds__(RB * (NX2 / R))
row_pass(const InstDesc<T>* __restrict__ descs, int tiles_per_inst, const cx<T>* __restrict__ pool) {
  constexpr int TR = NX2 / R;
  constexpr int NT = RB * TR;
  static_assert(TR % BX == 0, "row lanes must cover whole bricks");
  __shared__ cx<T> lds_mem[lds_row_elems<NX2, RB, R>()];
  __shared__ double red[4 * ((NT + 63) / 64)];

  const int inst = blockIdx.x / tiles_per_inst;
  const int tile = blockIdx.x - inst * tiles_per_inst;
  const InstDesc<T>& d = descs[inst];
  const int t = threadIdx.x;
  const int xx = t % BX;
  const int yy = (t / BX) % RB;
  const int jb = t / (BX * RB);
  const int j = jb * BX + xx;
  LdsRow<T, RB, R> lds{lds_mem, yy};
  const cx<T>* twh = tw_of(pool, NX2);       // length-Nx/2 transform
  const cx<T>* twn = tw_of(pool, 2 * NX2);   // exp(-2 pi i k / Nx)

  // brick address of (row tile, position k) for this lane's row
  const size_t tile_base = (size_t)tile * (NX2 / BX) * (RB * BX) + yy * BX + xx;
  auto addr = [&](int k) -> size_t { return tile_base + (size_t)(k / BX) * (RB * BX); };

  cx<T> u[R], v[R], w[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const size_t a = addr(j + r * TR);
    u[r] = d.um[a];
    v[r] = d.vm[a];
    w[r] = d.wm[a];
  }

  // Row C2R of a half spectrum X[0..NX2) (X[NX2] = 0, Im X[0] ignored):
  // Z[k] = (X[k] + conj X[M-k]) + i e^{+2 pi i k/N} (X[k] - conj X[M-k]),
  // then inverse length-M transform -> (x[2n], x[2n+1]).
  auto c2r = [&](cx<T>(&q)[R]) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) lds.store(j + r * TR, q[r]);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = j + r * TR;
      cx<T> xk = q[r], xr;
      if (k == 0) {
        xk.im = (T)0;
        xr = cx<T>{(T)0, (T)0};
      } else {
        xr = cconj(lds.load(NX2 - k));
      }
      const cx<T> s = xk + xr, dd = xk - xr;
      q[r] = s + cmuli(cmulc(dd, twn[k]));
    }
    __syncthreads();
    fft_core<NX2, R, +1>(q, j, lds, twh);
  };
  c2r(u);
  c2r(v);
  c2r(w);

  // Real-space products (nonl_subfunc_1) and the two combined fields.
  const T lam = (T)d.lam, bet = (T)d.beta;
  double esum = 0, gxs = 0, gys = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    T f1[2], f2[2];
    const T uxa[2] = {u[r].re, u[r].im}, uya[2] = {v[r].re, v[r].im}, wa[2] = {w[r].re, w[r].im};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const T ux = uxa[h], uy = uya[h], wv = wa[h];
      const T A = wv * ux, Wp = wv * uy;
      const T Q = ux * ux + uy * uy;
      const T Cx = Q * ux, Cy = Q * uy;
      f1[h] = lam * A + bet * Cy;
      f2[h] = lam * Wp - bet * Cx;
      esum += (double)ux * (double)ux + (double)uy * (double)uy;
      gxs += (double)Cx;
      gys += (double)Cy;
    }
    u[r] = cx<T>{f1[0], f1[1]};
    v[r] = cx<T>{f2[0], f2[1]};
  }

  // Row R2C of packed real pairs: Z = FFT_M(z), then
  // X[k] = 1/2 (Z[k] + conj Z[M-k]) - 1/2 i e^{-2 pi i k/N} (Z[k] - conj Z[M-k]).
  auto r2c = [&](cx<T>(&q)[R]) {
    __syncthreads();
    fft_core<NX2, R, -1>(q, j, lds, twh);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) lds.store(j + r * TR, q[r]);
    __syncthreads();
    const T h = (T)0.5;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = j + r * TR;
      const cx<T> zk = q[r], zm = cconj(lds.load((NX2 - k) & (NX2 - 1)));
      const cx<T> a = cscale(h, zk + zm), dd = zk - zm;
      q[r] = a + cmul(cmuli(cscale(-h, dd)), twn[k]);
    }
  };
  r2c(u);
  r2c(v);

#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k = j + r * TR;
    const T kx = (T)(d.tkx * (double)k);
    const size_t a = addr(k);
    d.pm[a] = cx<T>{kx * u[r].im, -kx * u[r].re};   // -i kx X1
    d.qm[a] = v[r];
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
// Column pass.  Block = W adjacent columns x TC = NY/R lanes per column.
// stage 0: initial pass only (v^ from the state, inverse transforms).
// stage 1..4: full stage.
// ---------------------------------------------------------------------------
template <class T, int NY, int W, int R, int BY>
__global__ void __launch_bounds__(W * (NY / R))
col_pass(const InstDesc<T>* __restrict__ descs, int tiles_per_inst, int stage, const cx<T>* __restrict__ pool) {
  constexpr int TC = NY / R;
  constexpr int NT = W * TC;
  __shared__ cx<T> lds_mem[lds_col_elems<NY, W, R>()];

  const int inst = blockIdx.x / tiles_per_inst;
  const int tile = blockIdx.x - inst * tiles_per_inst;
  const InstDesc<T>& d = descs[inst];
  const int t = threadIdx.x;
  const int c = t % W, j = t / W;
  const int x = tile * W + c;
  const int nx2 = d.nx >> 1;
  LdsCol<T, W, R> lds{lds_mem, c};
  const cx<T>* twy = tw_of(pool, NY);

  const size_t col_base = (size_t)tile * (BY * W) + c;
  const size_t brow = (size_t)(nx2 / W) * (BY * W);  // elements per brick row
  auto addr = [&](int y) -> size_t { return col_base + (size_t)(y / BY) * brow + (y % BY) * W; };

  const double dt = d.dt;
  const bool col0 = (x == 0);
  cx<T> sv[R];   // spectral input of the next stage
  double minj[2] = {0.0, 0.0};

  if (stage == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) sv[r] = d.wo[addr(j + r * TC)];
    if (tile == 0 && t == 0) {
      minj[0] = d.mean->M[0];
      minj[1] = d.mean->M[1];
    }
  } else {
    cx<T> pv[R], qv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const size_t a = addr(j + r * TC);
      pv[r] = d.pm[a];
      qv[r] = d.qm[a];
    }
    fft_core<NY, R, -1>(pv, j, lds, twy);
    __syncthreads();
    fft_core<NY, R, -1>(qv, j, lds, twy);
    // N^ = (P^ - i ky Q^) * mask; pv then holds N^.
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int y = j + r * TC;
      const int nys = (y <= NY / 2) ? y : y - NY;
      const double kyd = d.tky * (double)nys;
      const T keepf = keep_mode(x, nys < 0 ? -nys : nys, d.nx, d.ny, d.mask_mode) ? (T)1 : (T)0;
      pv[r] = cscale(keepf, pv[r] + cmuli(cscale((T)(-kyd), qv[r])));
    }
    // State: omega_old (stages 1-3), acc (stages 2-4), and the projection
    // partner (Ny - y) of the kx = 0 column.
    cx<T> wv[R], av[R], pw[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const size_t a = addr(j + r * TC);
      if (stage <= 3) wv[r] = d.wo[a];
      if (stage >= 2) av[r] = d.acc[a];
      if (col0) {
        const int yp = (NY - (j + r * TC)) & (NY - 1);
        pw[r] = (stage == 1) ? d.wo[addr(yp)] : d.acc[addr(yp)];
      }
    }

    cx<T> an[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int y = j + r * TC;
      const int nys = (y <= NY / 2) ? y : y - NY;
      const double kxd = d.tkx * (double)x, kyd = d.tky * (double)nys;
      const double k2 = kxd * kxd + kyd * kyd;
      const cx<T> nh = pv[r];
      const double sig = d.c0 + d.c2 * k2 + d.c4 * (k2 * k2);
      const T E = (T)exp(sig * dt);
      const T E2 = (T)exp(sig * dt * 0.5);
      const bool proj = col0 && (y & (NY / 2 - 1)) != 0;   // y not 0 or Ny/2
      cx<T> a;
      if (stage == 1) {
        a = cscale(E, wv[r]);
        if (proj) a = cscale((T)0.5, a + cconj(cscale(E, pw[r])));
      } else {
        a = av[r];
        if (proj) a = cscale((T)0.5, a + cconj(pw[r]));
      }
      cx<T> s;
      if (stage == 1) {
        a = a + cscale((T)(dt / 6 * (double)E), nh);
        s = cscale(E2, wv[r] + cscale((T)(dt / 2), nh));
      } else if (stage == 2) {
        a = a + cscale((T)(dt / 3 * (double)E2), nh);
        s = cscale(E2, wv[r] + cscale((T)(dt / 2), nh));
      } else if (stage == 3) {
        a = a + cscale((T)(dt / 3 * (double)E2), nh);
        s = cscale(E, wv[r] + cscale((T)dt, nh));
      } else {
        a = a + cscale((T)(dt / 6), nh);
        s = a;
      }
      an[r] = a;
      sv[r] = s;
    }
    __syncthreads();   // every projection-partner read precedes the writes below
    cx<T>* dst = (stage == 4) ? d.wo : d.acc;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[addr(j + r * TC)] = an[r];

    // Mean-velocity ODE (host updates of .cu:729-821), FP64, tile 0 only.
    if (tile == 0 && t < 64) {
      const int nt = d.row_tiles;
      double s0 = 0, s1 = 0, s2 = 0;
      for (int i = t; i < nt; i += 64) {
        s0 += d.partials[(size_t)i * 4 + 0];
        s1 += d.partials[(size_t)i * 4 + 1];
        s2 += d.partials[(size_t)i * 4 + 2];
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        s0 += __shfl_xor(s0, off, 64);
        s1 += __shfl_xor(s1, off, 64);
        s2 += __shfl_xor(s2, off, 64);
      }
      if (t == 0) {
        MeanState* ms = d.mean;
        const double g[2] = {-d.beta * s1, -d.beta * s2};
        double M[2] = {ms->M[0], ms->M[1]}, Mo[2] = {ms->Mo[0], ms->Mo[1]}, mm[2];
        if (stage == 1) {
          d.energy[ms->nstep % d.ring_cap] = s0 * d.invN * 0.5;
          Mo[0] = M[0];
          Mo[1] = M[1];
          if (d.mean_scheme == 1) { M[0] = d.e * Mo[0]; M[1] = d.e * Mo[1]; }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          if (stage == 1) { M[i] += dt / 6 * d.e * g[i];  mm[i] = d.e2 * (Mo[i] + dt / 2 * g[i]); }
          else if (stage == 2) { M[i] += dt / 3 * d.e2 * g[i]; mm[i] = d.e2 * (Mo[i] + dt / 2 * g[i]); }
          else if (stage == 3) { M[i] += dt / 3 * d.e2 * g[i]; mm[i] = d.e * (Mo[i] + dt * g[i]); }
          else { M[i] += dt / 6 * g[i]; mm[i] = M[i]; }
        }
        ms->M[0] = M[0]; ms->M[1] = M[1];
        ms->Mo[0] = Mo[0]; ms->Mo[1] = Mo[1];
        ms->m[0] = mm[0]; ms->m[1] = mm[1];
        if (stage == 4) {
          // check_func (.cu:884-897) on the new Re w^[0], owned by this lane
          if (ms->nan_step < 0 && an[0].re != an[0].re) ms->nan_step = ms->nstep + 1;
          ms->nstep += 1;
        }
        minj[0] = mm[0];
        minj[1] = mm[1];
      }
    }
  }

  // Biot-Savart + mean injection + three column inverse transforms.
  const T invN = (T)d.invN;
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    cx<T> q[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int y = j + r * TC;
      const int nys = (y <= NY / 2) ? y : y - NY;
      const double kxd = d.tkx * (double)x, kyd = d.tky * (double)nys;
      const double k2 = kxd * kxd + kyd * kyd;
      if (f == 2) {
        q[r] = cscale(invN, sv[r]);
      } else if (col0 && y == 0) {
        q[r] = cx<T>{(T)(minj[f] * d.invN), (T)0};
      } else {
        const double fac = (f == 0 ? kyd : -kxd) / k2 * d.invN;
        q[r] = cscale((T)fac, cmuli(sv[r]));
      }
    }
    __syncthreads();
    fft_core<NY, R, +1>(q, j, lds, twy);
    cx<T>* dst = (f == 0) ? d.um : (f == 1) ? d.vm : d.wm;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[addr(j + r * TC)] = q[r];
  }
}

// Reference layout (Ny x (Nx/2+1), interleaved doubles) <-> bricks.
__device__ __forceinline__ size_t brick_index(int y, int x, int nx2, int by, int bx) {
  return ((size_t)(y / by) * (nx2 / bx) + x / bx) * (size_t)(by * bx) + (y % by) * bx + x % bx;
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
