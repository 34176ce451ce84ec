// CPU oracle for the ActFlow time step -- TEST INFRASTRUCTURE ONLY.
//
// Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg, as the checker.  The product library never links or calls it.
//
// A restatement of the reference step (ActFlow_V3-2.cu) with a CPU FFT in
// place of cuFFT.  Each function cites the reference lines it follows.
// Pinning (tests/test_oracle_*.py):
//   Fft1 / RealFft / Fft2    brute-force long-double DFT, numpy rfft2/irfft2
//                            golden vectors, round trip, Parseval
//   Stepper::rhs             Biot-Savart single-mode closed form, shell
//                            invariance of the advection term, shear-flow
//                            cubic closed form, energy closed form
//   Stepper::step            lambda=beta=0 exact E^n, mean-ODE scalar
//                            recurrence, uniform-advection O(dt^2) closed form,
//                            measured temporal order 2 (variant IF-RK); mean
//                            ODE with sigma(0)=0 and omega=0 converging at
//                            order 4 to |M(t)|^2 = |M0|^2/(1+2b|M0|^2 t/N^4)
//   Stepper::project         lambda=beta=0 one step from a kx=0 column with
//                            independent rows y, Ny-y equals P(E omega0)
//                            exactly (average-and-conjugate, Nyquist row kept)
//   masks                    brute-force lattice counts
#include "oracle.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

constexpr long double kPi = 3.141592653589793238462643383279502884L;

template <class T> struct C { T re, im; };

template <class T> inline C<T> cadd(C<T> a, C<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T> inline C<T> csub(C<T> a, C<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T> inline C<T> cmul(C<T> a, C<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T> inline C<T> cscale(T s, C<T> a) { return {s * a.re, s * a.im}; }
template <class T> inline C<T> cconj(C<T> a) { return {a.re, -a.im}; }
// multiplication by the complex unit i() of cuComplexBinOp.h:13-17
template <class T> inline C<T> cmuli(C<T> a) { return {-a.im, a.re}; }

static bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// exp(-2*pi*i*k/n) in long double, reduced to the first octant so that the
// rounded values are symmetric (cos/sin of the same reduced angle).
static void twiddle_ld(long k, long n, long double* c, long double* s) {
  k %= n;
  if (k < 0) k += n;
  // angle = 2*pi*k/n in [0, 2pi); use symmetry about quadrants
  long q = (4 * k) / n;             // quadrant index 0..3
  long r = 4 * k - q * n;           // remainder in [0, n)
  long double a = kPi * 0.5L * (long double)r / (long double)n;  // in [0, pi/2)
  long double ca = cosl(a), sa = sinl(a);
  long double cc, ss;               // cos/sin of full angle
  switch (q) {
    case 0: cc = ca; ss = sa; break;
    case 1: cc = -sa; ss = ca; break;
    case 2: cc = -ca; ss = -sa; break;
    default: cc = sa; ss = -ca; break;
  }
  *c = cc;
  *s = -ss;                         // exp(-i angle)
}

// Iterative radix-2 Cooley-Tukey (decimation in time), unnormalised.
// Stands in for the column passes of cufftExecR2C/C2R (.cu:504-508).
template <class T>
struct Fft1 {
  int n = 0;
  std::vector<C<T>> w;   // w[k] = exp(-2 pi i k / n), k < n/2
  std::vector<int> rev;
  void init(int n_) {
    n = n_;
    w.resize(n / 2 > 0 ? n / 2 : 1);
    for (int k = 0; k < n / 2; ++k) {
      long double c, s;
      twiddle_ld(k, n, &c, &s);
      w[k] = {(T)c, (T)s};
    }
    rev.assign(n, 0);
    int lg = 0;
    while ((1 << lg) < n) ++lg;
    for (int i = 0; i < n; ++i) {
      int r = 0;
      for (int b = 0; b < lg; ++b)
        if (i & (1 << b)) r |= 1 << (lg - 1 - b);
      rev[i] = r;
    }
  }
  // sign -1: forward (exp(-i...)), +1: inverse (exp(+i...))
  void run(C<T>* a, int sign) const {
    for (int i = 0; i < n; ++i)
      if (i < rev[i]) { C<T> t = a[i]; a[i] = a[rev[i]]; a[rev[i]] = t; }
    for (int len = 2; len <= n; len <<= 1) {
      const int half = len >> 1, step = n / len;
      for (int i = 0; i < n; i += len)
        for (int k = 0; k < half; ++k) {
          C<T> t = w[k * step];
          if (sign > 0) t.im = -t.im;
          C<T> u = a[i + k], v = cmul(a[i + k + half], t);
          a[i + k] = cadd(u, v);
          a[i + k + half] = csub(u, v);
        }
    }
  }
};

// Real transform of length N through a complex transform of length N/2
// (the even/odd split).  r2c: N reals -> N/2+1 complex.  c2r: N/2+1 complex
// -> N reals (times N), using only Re of the k=0 and k=N/2 bins, i.e. the
// real part of the full inverse DFT of the Hermitian extension.
template <class T>
struct RealFft {
  int N = 0, M = 0;
  Fft1<T> half;
  std::vector<C<T>> tw;  // exp(-2 pi i k / N), k = 0..M
  void init(int N_) {
    N = N_;
    M = N / 2;
    half.init(M);
    tw.resize(M + 1);
    for (int k = 0; k <= M; ++k) {
      long double c, s;
      twiddle_ld(k, N, &c, &s);
      tw[k] = {(T)c, (T)s};
    }
  }
  void r2c(const T* x, C<T>* X, C<T>* z) const {
    for (int m = 0; m < M; ++m) z[m] = {x[2 * m], x[2 * m + 1]};
    half.run(z, -1);
    const T h = (T)0.5;
    for (int k = 0; k <= M; ++k) {
      C<T> zk = z[k % M], zm = cconj(z[(M - k) % M]);
      C<T> a = cscale(h, cadd(zk, zm));
      C<T> d = csub(zk, zm);
      C<T> b = cmul(cmuli(cscale(-h, d)), tw[k]);  // -1/2 i w_k d
      X[k] = cadd(a, b);
    }
  }
  void c2r(const C<T>* X, T* x, C<T>* z) const {
    const C<T> x0 = {X[0].re, (T)0}, xm = {X[M].re, (T)0};
    for (int k = 0; k < M; ++k) {
      C<T> xk = (k == 0) ? x0 : X[k];
      C<T> xr = cconj((k == 0) ? xm : X[M - k]);
      C<T> s = cadd(xk, xr), d = csub(xk, xr);
      z[k] = cadd(s, cmuli(cmul(d, cconj(tw[k]))));
    }
    half.run(z, +1);
    for (int m = 0; m < M; ++m) { x[2 * m] = z[m].re; x[2 * m + 1] = z[m].im; }
  }
};

// Unnormalised 2-D R2C / C2R of an ny x nx real field (cufftPlan2d(Ny, Nx),
// .cu:507-508): rows by RealFft, then columns by Fft1.
template <class T>
struct Fft2 {
  int nx = 0, ny = 0, nxh = 0, nthreads = 1;
  RealFft<T> rowf;
  Fft1<T> colf;
  std::vector<std::vector<C<T>>> work;  // per thread
  void init(int nx_, int ny_, int nth) {
    nx = nx_; ny = ny_; nxh = nx / 2 + 1; nthreads = nth;
    rowf.init(nx);
    colf.init(ny);
    work.assign(nthreads, std::vector<C<T>>(ny > nx ? ny : nx));
  }
  void r2c(const T* in, C<T>* out) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int y = 0; y < ny; ++y)
      rowf.r2c(in + (size_t)y * nx, out + (size_t)y * nxh, work[omp_get_thread_num()].data());
    cols(out, -1);
  }
  // in is overwritten (used as the mixed-space buffer)
  void c2r(C<T>* in, T* out) {
    cols(in, +1);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int y = 0; y < ny; ++y)
      rowf.c2r(in + (size_t)y * nxh, out + (size_t)y * nx, work[omp_get_thread_num()].data());
  }
  void cols(C<T>* a, int sign) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int x = 0; x < nxh; ++x) {
      C<T>* buf = work[omp_get_thread_num()].data();
      for (int y = 0; y < ny; ++y) buf[y] = a[(size_t)y * nxh + x];
      colf.run(buf, sign);
      for (int y = 0; y < ny; ++y) a[(size_t)y * nxh + x] = buf[y];
    }
  }
};

// Dealiasing masks.  The reference's cutoff_func (.cu:899-919) intends an
// isotropic cut x^2 + min(y, Ny-y)^2 <= R_sq; mode 3 keeps its literal 32-bit
// arithmetic, modes 0/1 are the 2/3 rules, mode 2 keeps everything.
static bool keep_mode(int x, int y, int nx, int ny, int mode) {
  const int64_t nxm = x, nym = (y <= ny - y) ? y : ny - y;
  const int64_t NX = nx, NY = ny;
  switch (mode) {
    case 0: return 9 * (nxm * nxm * NY * NY + nym * nym * NX * NX) <= NX * NX * NY * NY;
    case 1: return 3 * nxm <= NX && 3 * nym <= NY;
    case 2: return true;
    default: {
      const int nxh = nx / 2 + 1;
      // R_sq{(Nxh-1)*(Nxh-1) * Ny*Ny/4} evaluated left to right in int32
      uint32_t r = (uint32_t)(nxh - 1) * (uint32_t)(nxh - 1);
      r *= (uint32_t)ny;
      r *= (uint32_t)ny;
      int32_t rs = (int32_t)r / 4;
      const int32_t a = x * x + y * y, b = x * x + (ny - y) * (ny - y);
      return a <= rs || b <= rs;
    }
  }
}

template <class T>
struct Stepper {
  int nx, ny, nxh, nc, nthreads;
  double dt, lam, beta, c0, c2, c4, lx, ly;
  int dealias, mean_scheme;
  T lam_s, beta_s;                 // lambda/N^2, beta/N^3 (.cu:432-436)
  std::vector<T> kx, ky, k2, mask, E, E2;
  double e, e2;                    // mean-mode factors (.cu:704-705)
  std::vector<C<T>> w, w_old, s, acc, nonl, vxh, vyh, buf0, buf1, buf2, buf3;
  std::vector<T> ux, uy, wr, ra, rb, rc, rd;
  double M[2] = {0, 0};
  Fft2<T> fft;

  void init(int nx_, int ny_, double lx_, double ly_, double dt_, double lam_, double beta_,
            double c0_, double c2_, double c4_, int dealias_, int mean_scheme_, int nth) {
    nx = nx_; ny = ny_; nxh = nx / 2 + 1; nc = ny * nxh; nthreads = nth;
    lx = lx_; ly = ly_; dt = dt_; lam = lam_; beta = beta_; c0 = c0_; c2 = c2_; c4 = c4_;
    dealias = dealias_; mean_scheme = mean_scheme_;
    const double N = (double)nx * (double)ny;
    lam_s = (T)(lam / (N * N));
    beta_s = (T)(beta / (N * N * N));
    kx.resize(nc); ky.resize(nc); k2.resize(nc); mask.resize(nc); E.resize(nc); E2.resize(nc);
    const double tkx = 2.0 * (double)kPi / lx, tky = 2.0 * (double)kPi / ly;
    for (int c = 0; c < nc; ++c) {           // wavenumber_func, .cu:1329-1349
      const int x = c % nxh, y = c / nxh;
      const double kxd = tkx * x;
      const double kyd = (y < ny / 2 + 1) ? tky * y : tky * (y - ny);
      const double k2d = kxd * kxd + kyd * kyd;
      kx[c] = (T)kxd; ky[c] = (T)kyd; k2[c] = (T)k2d;
      mask[c] = keep_mode(x, y, nx, ny, dealias) ? (T)1 : (T)0;
      const double sig = c0 + c2 * k2d + c4 * (k2d * k2d);  // lin_func, .cu:1030
      E[c] = (T)std::exp(sig * dt);
      E2[c] = (T)std::exp(sig * dt * 0.5);
    }
    e = std::exp(c0 * dt);                  // lin_v, .cu:704
    e2 = std::exp(c0 * dt * 0.5);           // lin2_v, .cu:705
    for (auto* v : {&w, &w_old, &s, &acc, &nonl, &vxh, &vyh, &buf0, &buf1, &buf2, &buf3})
      v->assign(nc, C<T>{0, 0});
    for (auto* v : {&ux, &uy, &wr, &ra, &rb, &rc, &rd}) v->assign((size_t)nx * ny, (T)0);
    fft.init(nx, ny, nthreads);
  }

  // Stage RHS F(s, m) -> nonl, g, energy: nonl_func / modified_nonl_func
  // (.cu:1036-1161) with nonl_subfunc_0/1/2 (.cu:1163-1225).
  void rhs(const C<T>* sin, const double m[2], C<T>* out, double g[2], double* energy) {
    // nonl_subfunc_0 (.cu:1176-1180): v = (s*i)*ky/k^2, (s*i)*(-kx)/k^2, c >= 1
    for (int c = 1; c < nc; ++c) {
      C<T> si = cmuli(sin[c]);
      C<T> a = cscale(ky[c], si), b = cscale((T)-kx[c], si);
      vxh[c] = {a.re / k2[c], a.im / k2[c]};
      vyh[c] = {b.re / k2[c], b.im / k2[c]};
    }
    // mean injection (.cu:1053-1054)
    vxh[0] = {(T)m[0], (T)0};
    vyh[0] = {(T)m[1], (T)0};
    buf0.assign(sin, sin + nc);
    // C2R x3 (.cu:1057-1059)
    fft.c2r(vxh.data(), ux.data());
    fft.c2r(vyh.data(), uy.data());
    fft.c2r(buf0.data(), wr.data());
    const size_t nr = (size_t)nx * ny;
    if (energy) {  // .cu:1118-1127, accumulated in double, row partials in order
      std::vector<double> part(ny, 0.0);
#pragma omp parallel for num_threads(nthreads) schedule(static)
      for (int y = 0; y < ny; ++y) {
        double acc_e = 0;
        for (int x = 0; x < nx; ++x) {
          const double a = ux[(size_t)y * nx + x], b = uy[(size_t)y * nx + x];
          acc_e += a * a + b * b;
        }
        part[y] = acc_e;
      }
      double tot = 0;
      for (int y = 0; y < ny; ++y) tot += part[y];
      const double N = (double)nx * (double)ny;
      *energy = tot / (2.0 * N * N * N);
    }
    // nonl_subfunc_1 (.cu:1192-1205): arr=w*ux, w_appr=w*uy, Q=|u|^2, vx*=Q, vy*=Q
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long r = 0; r < (long)nr; ++r) {
      const T wv = wr[r], a = ux[r], b = uy[r];
      ra[r] = wv * a;                  // A  (arr)
      rb[r] = wv * b;                  // W  (w_appr)
      const T q = a * a + b * b;
      rc[r] = a * q;                   // Cx (vx)
      rd[r] = b * q;                   // Cy (vy)
    }
    // R2C x4 (.cu:1069-1072)
    fft.r2c(rb.data(), buf0.data());   // W-hat  (w_appr_comp)
    fft.r2c(ra.data(), buf1.data());   // A-hat  (arr_comp)
    fft.r2c(rc.data(), buf2.data());   // Cx-hat (vx_comp)
    fft.r2c(rd.data(), buf3.data());   // Cy-hat (vy_comp)
    // nonl_subfunc_2 (.cu:1221-1223)
    for (int c = 0; c < nc; ++c) {
      C<T> t1 = cadd(cscale(lam_s, buf1[c]), cscale(beta_s, buf3[c]));
      C<T> t2 = csub(cscale(lam_s, buf0[c]), cscale(beta_s, buf2[c]));
      C<T> u = csub(cscale((T)-kx[c], t1), cscale(ky[c], t2));
      out[c] = cscale(mask[c], cmuli(u));
    }
    // .cu:1086-1087
    g[0] = -(double)beta_s * (double)buf2[0].re;
    g[1] = -(double)beta_s * (double)buf3[0].re;
  }

  // reality_func (.cu:1228-1243), applied to acc
  void project(C<T>* a) {
    const T h = (T)0.5;
    for (int y = 1; y < ny / 2; ++y) {
      C<T> v = cscale(h, cadd(a[(size_t)y * nxh], cconj(a[(size_t)(ny - y) * nxh])));
      a[(size_t)y * nxh] = v;
      a[(size_t)(ny - y) * nxh] = cconj(v);
    }
  }

  // One time step, .cu:718-821 (Appendix A of the design notes).
  bool step(double* energy) {
    // integrate_func_0 (.cu:949-954); mean copies (.cu:729-732)
    for (int c = 0; c < nc; ++c) {
      w_old[c] = w[c];
      s[c] = w[c];
      acc[c] = cscale(E[c], w[c]);
    }
    const double Mo[2] = {M[0], M[1]};
    double m[2] = {M[0], M[1]};
    if (mean_scheme == 1) { M[0] = e * Mo[0]; M[1] = e * Mo[1]; }
    double g[2];
    // stage 1: modified_nonl_func, reality_func, integrate_func_1, .cu:736-759
    rhs(s.data(), m, nonl.data(), g, energy);
    project(acc.data());
    for (int c = 0; c < nc; ++c) {
      acc[c] = cadd(acc[c], cscale((T)(dt / 6 * (double)E[c]), nonl[c]));
      s[c] = cscale(E2[c], cadd(w_old[c], cscale((T)(dt / 2), nonl[c])));
    }
    for (int i = 0; i < 2; ++i) { M[i] += dt / 6 * e * g[i]; m[i] = e2 * (Mo[i] + dt / 2 * g[i]); }
    // stage 2: .cu:762-781, integrate_func_2
    rhs(s.data(), m, nonl.data(), g, nullptr);
    project(acc.data());
    for (int c = 0; c < nc; ++c) {
      acc[c] = cadd(acc[c], cscale((T)(dt / 3 * (double)E2[c]), nonl[c]));
      s[c] = cscale(E2[c], cadd(w_old[c], cscale((T)(dt / 2), nonl[c])));
    }
    for (int i = 0; i < 2; ++i) { M[i] += dt / 3 * e2 * g[i]; m[i] = e2 * (Mo[i] + dt / 2 * g[i]); }
    // stage 3: .cu:784-802, integrate_func_3
    rhs(s.data(), m, nonl.data(), g, nullptr);
    project(acc.data());
    for (int c = 0; c < nc; ++c) {
      acc[c] = cadd(acc[c], cscale((T)(dt / 3 * (double)E2[c]), nonl[c]));
      s[c] = cscale(E[c], cadd(w_old[c], cscale((T)dt, nonl[c])));
    }
    for (int i = 0; i < 2; ++i) { M[i] += dt / 3 * e2 * g[i]; m[i] = e * (Mo[i] + dt * g[i]); }
    // stage 4: .cu:805-821, integrate_func_4
    rhs(s.data(), m, nonl.data(), g, nullptr);
    project(acc.data());
    for (int c = 0; c < nc; ++c) acc[c] = cadd(acc[c], cscale((T)(dt / 6), nonl[c]));
    for (int i = 0; i < 2; ++i) M[i] += dt / 6 * g[i];
    w.swap(acc);
    return std::isnan((double)w[0].re);     // check_func, .cu:884-897
  }
};

}  // namespace

struct oracle_run {
  int precision;
  Stepper<double> d;
  Stepper<float> f;
};

// C linkage comes from the declarations in oracle.h.

oracle_run* oracle_create(int nx, int ny, double lx, double ly, double dt, double lambda,
                          double beta, double c0, double c2, double c4, int precision,
                          int dealias, int mean_scheme, int nthreads) {
  if (!is_pow2(nx) || !is_pow2(ny) || nx < 4 || ny < 2) return nullptr;
  if (precision != 32 && precision != 64) return nullptr;
  if (nthreads < 1) nthreads = 1;
  oracle_run* r = new oracle_run();
  r->precision = precision;
  if (precision == 64)
    r->d.init(nx, ny, lx, ly, dt, lambda, beta, c0, c2, c4, dealias, mean_scheme, nthreads);
  else
    r->f.init(nx, ny, lx, ly, dt, lambda, beta, c0, c2, c4, dealias, mean_scheme, nthreads);
  return r;
}

void oracle_destroy(oracle_run* r) { delete r; }

template <class T>
static void set_state_t(Stepper<T>& s, const double* w, const double M[2]) {
  for (int c = 0; c < s.nc; ++c) s.w[c] = {(T)w[2 * c], (T)w[2 * c + 1]};
  s.M[0] = M[0];
  s.M[1] = M[1];
}
template <class T>
static void get_state_t(const Stepper<T>& s, double* w, double M[2]) {
  for (int c = 0; c < s.nc; ++c) { w[2 * c] = s.w[c].re; w[2 * c + 1] = s.w[c].im; }
  M[0] = s.M[0];
  M[1] = s.M[1];
}

void oracle_set_state(oracle_run* r, const double* w, const double M[2]) {
  if (r->precision == 64) set_state_t(r->d, w, M); else set_state_t(r->f, w, M);
}
void oracle_get_state(const oracle_run* r, double* w, double M[2]) {
  if (r->precision == 64) get_state_t(r->d, w, M); else get_state_t(r->f, w, M);
}

int oracle_step(oracle_run* r, int nsteps, double* energy) {
  for (int i = 0; i < nsteps; ++i) {
    bool bad = (r->precision == 64) ? r->d.step(energy ? energy + i : nullptr)
                                    : r->f.step(energy ? energy + i : nullptr);
    if (bad) return 1;
  }
  return 0;
}

template <class T>
static void rhs_t(Stepper<T>& s, const double* sh, const double m[2], double* out, double g[2],
                  double* energy) {
  std::vector<C<T>> in(s.nc), o(s.nc);
  for (int c = 0; c < s.nc; ++c) in[c] = {(T)sh[2 * c], (T)sh[2 * c + 1]};
  s.rhs(in.data(), m, o.data(), g, energy);
  for (int c = 0; c < s.nc; ++c) { out[2 * c] = o[c].re; out[2 * c + 1] = o[c].im; }
}

void oracle_rhs(oracle_run* r, const double* s_hat, const double m[2], double* nonl_hat,
                double g[2], double* energy) {
  if (r->precision == 64) rhs_t(r->d, s_hat, m, nonl_hat, g, energy);
  else rhs_t(r->f, s_hat, m, nonl_hat, g, energy);
}

double oracle_mask(const oracle_run* r, int c) {
  return r->precision == 64 ? (double)r->d.mask[c] : (double)r->f.mask[c];
}

template <class T>
static void rfft2_t(int ny, int nx, const double* in, double* out) {
  Fft2<T> f;
  f.init(nx, ny, 1);
  std::vector<T> x((size_t)nx * ny);
  for (size_t i = 0; i < x.size(); ++i) x[i] = (T)in[i];
  std::vector<C<T>> X((size_t)ny * (nx / 2 + 1));
  f.r2c(x.data(), X.data());
  for (size_t i = 0; i < X.size(); ++i) { out[2 * i] = X[i].re; out[2 * i + 1] = X[i].im; }
}
template <class T>
static void irfft2_t(int ny, int nx, const double* in, double* out) {
  Fft2<T> f;
  f.init(nx, ny, 1);
  std::vector<C<T>> X((size_t)ny * (nx / 2 + 1));
  for (size_t i = 0; i < X.size(); ++i) X[i] = {(T)in[2 * i], (T)in[2 * i + 1]};
  std::vector<T> x((size_t)nx * ny);
  f.c2r(X.data(), x.data());
  for (size_t i = 0; i < x.size(); ++i) out[i] = x[i];
}
template <class T>
static void fft1_t(int n, int sign, const double* in, double* out) {
  Fft1<T> f;
  f.init(n);
  std::vector<C<T>> a(n);
  for (int i = 0; i < n; ++i) a[i] = {(T)in[2 * i], (T)in[2 * i + 1]};
  f.run(a.data(), sign);
  for (int i = 0; i < n; ++i) { out[2 * i] = a[i].re; out[2 * i + 1] = a[i].im; }
}

void oracle_rfft2(int ny, int nx, int precision, const double* in, double* out) {
  if (precision == 64) rfft2_t<double>(ny, nx, in, out); else rfft2_t<float>(ny, nx, in, out);
}
void oracle_irfft2(int ny, int nx, int precision, const double* in, double* out) {
  if (precision == 64) irfft2_t<double>(ny, nx, in, out); else irfft2_t<float>(ny, nx, in, out);
}
void oracle_fft1(int n, int sign, int precision, const double* in, double* out) {
  if (precision == 64) fft1_t<double>(n, sign, in, out); else fft1_t<float>(n, sign, in, out);
}

