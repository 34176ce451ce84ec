// Host-side file contract of the reference: specifications.txt parsing
// (ActFlow_V3-2.cu:320-420), the `spec` example writer (.cu:250-266),
// snapshot text files (save_func .cu:1245-1266, reader .cu:615-634),
// energy.txt lines (.cu:1268-1278) and the initial conditions (.cu:535-566,
// 660-679; the band-limited seeded IC has no reference counterpart).
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "status.hpp"

using namespace af;

namespace {

void set_err(char* err, size_t errlen, const std::string& msg) {
  g_last_error = msg;
  if (err && errlen) {
    std::strncpy(err, msg.c_str(), errlen - 1);
    err[errlen - 1] = 0;
  }
}

void copy_str(char* dst, size_t n, const std::string& s) {
  std::strncpy(dst, s.c_str(), n - 1);
  dst[n - 1] = 0;
}

// %g with p significant digits: the default std::ostream formatting at
// precision p, which is what the reference's `<<` produces.
std::string fmt_g(double v, int p) {
  char b[64];
  std::snprintf(b, sizeof b, "%.*g", p, v);
  return b;
}

// ---- a small host FFT used only for the `random` initial condition --------
void fft_inplace(std::vector<std::complex<double>>& a, int sign) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const double ang = sign * 2.0 * M_PI / (double)len;
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; ++k) {
        const std::complex<double> w(std::cos(ang * (double)k), std::sin(ang * (double)k));
        const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
  }
}

bool keep_host(int x, int y, int nx, int ny, int mode) {
  const long long X = x, Y = (y <= ny - y) ? y : ny - y, NX = nx, NY = ny;
  if (mode == 0) return 9 * (X * X * NY * NY + Y * Y * NX * NX) <= NX * NX * NY * NY;
  return 3 * X <= NX && 3 * Y <= NY;
}

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double unit53(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

extern "C" {

void af_spec_defaults(af_spec* s) {
  std::memset(s, 0, sizeof(*s));
  // float globals of .cu:184-203; Ly{Lx} and Ny{Nx} bind at static init
  s->T = 100.0f;
  s->dt = 0.002f;
  s->Lx = 250.0f;
  s->Ly = 250.0f;
  s->Nx = 1024;
  s->Ny = 1024;
  s->lambda = 7.0f;
  s->alpha = -0.8f;
  s->beta = 0.01f;
  s->dt_snap = 0.5f;
  copy_str(s->initial, sizeof s->initial, "random");
  copy_str(s->directory, sizeof s->directory, "./");
  s->precision = 32;
  s->dealias = 0;
  s->mean_scheme = 0;
  s->snapshot_precision = 6;
  s->M = (int)std::round((float)s->T / (float)s->dt);
  s->snap = (int)std::round((float)s->dt_snap / (float)s->dt);
}

af_status af_spec_read(const char* dir, af_spec* s, char* err, size_t errlen) {
  if (!dir || !s) return fail(AF_EINVAL, "NULL argument");
  copy_str(s->directory, sizeof s->directory, dir);
  std::ifstream f(std::string(dir) + "/specifications.txt");
  if (!f.is_open()) {
    set_err(err, errlen, "Error: Unable to open \"specifications.txt\".");
    return AF_EIO;
  }
  std::string line, val;
  auto num = [&](float* out) -> bool {
    std::getline(f, val);
    try { *out = std::stof(val); } catch (...) { return false; }
    return true;
  };
  auto inum = [&](int* out) -> bool {
    std::getline(f, val);
    try { *out = std::stoi(val); } catch (...) { return false; }
    return true;
  };
  while (std::getline(f, line)) {
    bool ok = true;
    float v = 0;
    if (line == "T:") { ok = num(&v); s->T = v; }
    else if (line == "dt:") { ok = num(&v); s->dt = v; }
    else if (line == "Lx:") { ok = num(&v); s->Lx = v; }
    else if (line == "Ly:") { ok = num(&v); s->Ly = v; }
    else if (line == "Nx:") ok = inum(&s->Nx);
    else if (line == "Ny:") ok = inum(&s->Ny);
    else if (line == "lambda:") { ok = num(&v); s->lambda = v; }
    else if (line == "alpha:") { ok = num(&v); s->alpha = v; }
    else if (line == "beta:") { ok = num(&v); s->beta = v; }
    else if (line == "dt_snap:") { ok = num(&v); s->dt_snap = v; }
    else if (line == "initial_condition:") { std::getline(f, val); copy_str(s->initial, sizeof s->initial, val); }
    else if (line == "save_directory:") { std::getline(f, val); copy_str(s->directory, sizeof s->directory, val); }
    // extension keys (no reference counterpart)
    else if (line == "Gamma0:") { ok = num(&v); s->gamma0 = v; s->has_gamma = 1; }
    else if (line == "Gamma2:") { ok = num(&v); s->gamma2 = v; s->has_gamma = 1; }
    else if (line == "precision:") ok = inum(&s->precision);
    else if (line == "dealias:") ok = inum(&s->dealias);
    else if (line == "mean_scheme:") ok = inum(&s->mean_scheme);
    else if (line == "snapshot_precision:") ok = inum(&s->snapshot_precision);
    else {
      set_err(err, errlen,
              "Warning: specification file contains the following unknown specification: " + line +
                  "\nFor additional informatin on the specifications file start the program with the "
                  "argument \"help\".");
      return AF_EINVAL;
    }
    if (!ok) {
      set_err(err, errlen, "Error: invalid value \"" + val + "\" for specification " + line);
      return AF_EINVAL;
    }
  }
  // derived step counts, float arithmetic as in .cu:409, 420
  s->M = (int)std::round((float)s->T / (float)s->dt);
  s->snap = (int)std::round((float)s->dt_snap / (float)s->dt);
  return AF_OK;
}

af_status af_spec_write_example(const char* path, const af_spec* s) {
  std::ofstream f(path, std::ios::out);
  if (!f.is_open()) return fail(AF_EIO, "Error: unable to create example file");
  // float values printed with the default stream format (.cu:259-264)
  f << "T:\n" << (float)s->T << "\ndt:\n" << (float)s->dt << "\nLx:\n" << (float)s->Lx << "\nLy:\n"
    << (float)s->Ly << "\nNx:\n" << s->Nx << "\nNy:\n" << s->Ny << "\nlambda:\n" << (float)s->lambda
    << "\nalpha:\n" << (float)s->alpha << "\nbeta:\n" << (float)s->beta << "\ndt_snap:\n" << (float)s->dt_snap
    << "\ninitial_condition:\n" << s->initial << "\nsave_directory:\n" << s->directory;
  return f.good() ? AF_OK : fail(AF_EIO, "write failed");
}

af_status af_snapshot_write(const char* path, int nx, int ny, const double* w, const double M[2], int precision) {
  if (!path || !w || !M || nx < 2 || ny < 1) return fail(AF_EINVAL, "bad arguments");
  if (precision < 1 || precision > 17) precision = 6;
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(AF_EIO, std::string("cannot open ") + path);
  const size_t nc = (size_t)ny * (nx / 2 + 1);
  // Formatting (a %g per value, ~4 M values at 2048^2) costs far more than
  // the write: rounds of up to 16 blocks of kBlock modes are formatted by one
  // thread each and written in block order, so the bytes are those of one
  // sequential pass.
  constexpr size_t kBlock = 1 << 16;
  const size_t nblocks = (nc + kBlock - 1) / kBlock;
  const unsigned hw = std::thread::hardware_concurrency();
  const size_t nthreads = std::min<size_t>(std::min<size_t>(hw ? hw : 1, 16), nblocks);
  std::vector<std::string> text(nthreads);
  auto format = [&](size_t blk, std::string& out) {
    out.clear();
    char b[64];
    const size_t c1 = std::min(nc, (blk + 1) * kBlock);
    for (size_t c = blk * kBlock; c < c1; ++c) {
      const int n = std::snprintf(b, sizeof b, "%.*g,%.*g,", precision, w[2 * c], precision, w[2 * c + 1]);
      out.append(b, (size_t)n);
    }
  };
  bool ok = true;
  for (size_t b0 = 0; b0 < nblocks && ok; b0 += nthreads) {
    const size_t nb = std::min(nthreads, nblocks - b0);
    std::vector<std::thread> pool;
    size_t started = 1;   // blocks b0 + 1 .. b0 + started - 1 run on threads
    try {
      for (; started < nb; ++started) pool.emplace_back(format, b0 + started, std::ref(text[started]));
    } catch (const std::system_error&) {
      // no more threads: the rest of the round is formatted here
    }
    format(b0, text[0]);
    for (size_t i = started; i < nb; ++i) format(b0 + i, text[i]);
    for (auto& t : pool) t.join();
    for (size_t i = 0; i < nb && ok; ++i) ok = std::fwrite(text[i].data(), 1, text[i].size(), f) == text[i].size();
  }
  char b[64];
  const int n = std::snprintf(b, sizeof b, "%.*g,%.*g", precision, M[0], precision, M[1]);
  ok = ok && std::fwrite(b, 1, (size_t)n, f) == (size_t)n;
  return (std::fclose(f) == 0 && ok) ? AF_OK : fail(AF_EIO, "write");
}

af_status af_snapshot_read(const char* path, int nx, int ny, double* w, double M[2]) {
  if (!path || !w || !M) return fail(AF_EINVAL, "bad arguments");
  std::ifstream f(path, std::ios::in | std::ios::binary);
  if (!f.is_open()) return fail(AF_EIO, "Error: unable to open previous simulation");
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  const char* p = text.c_str();
  const char* end = p + text.size();
  // value, then skip through the next ',' (operator>> then ignore(1000, ','))
  auto next = [&](double* out) -> bool {
    char* q = nullptr;
    errno = 0;
    const double v = std::strtod(p, &q);
    if (q == p) return false;
    *out = v;
    p = q;
    while (p < end && *p != ',') ++p;
    if (p < end) ++p;
    return true;
  };
  const size_t nc = (size_t)ny * (nx / 2 + 1);
  for (size_t i = 0; i < 2 * nc; ++i)
    if (!next(&w[i])) return fail(AF_EIO, "snapshot too short or malformed");
  if (!next(&M[0]) || !next(&M[1])) return fail(AF_EIO, "snapshot lacks the mean velocity");
  return AF_OK;
}

af_status af_energy_append(const char* path, double t, double E) {
  FILE* f = std::fopen(path, "a");
  if (!f) return fail(AF_EIO, "Error: failed to save the mean energy");
  std::fprintf(f, "%s %s\n", fmt_g(t, 6).c_str(), fmt_g(E, 6).c_str());
  return std::fclose(f) == 0 ? AF_OK : fail(AF_EIO, "Error: failed to save the mean energy");
}

af_status af_ic_bandlimited(int nx, int ny, unsigned long long seed, double amp, double kmax, double* w) {
  if (!w || nx < 2 || ny < 2 || !(kmax >= 0)) return fail(AF_EINVAL, "bad arguments");
  const int nxh = nx / 2 + 1;
  const size_t nc = (size_t)ny * nxh;
  std::memset(w, 0, 2 * nc * sizeof(double));
  const int K = (int)std::floor(kmax);
  const double k2max = kmax * kmax;
  if (K >= nx / 2 || K >= ny / 2) return fail(AF_EINVAL, "kmax must be below the Nyquist index");
  const double N = (double)nx * (double)ny;
  uint64_t st = seed;
  // visiting order: nx = 0..K, then ny = -K..K; the independent half is
  // nx > 0, or nx = 0 with ny > 0
  for (int i = 0; i <= K; ++i)
    for (int j = -K; j <= K; ++j) {
      const double r2 = (double)i * i + (double)j * j;
      if (r2 == 0 || r2 > k2max) continue;
      if (i == 0 && j < 0) continue;
      const double u1 = 1.0 - unit53(st);   // (0, 1]
      const double u2 = unit53(st);
      const double u3 = unit53(st);
      const double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
      const double th = 2.0 * M_PI * u3;
      const double re = N * amp * g * std::cos(th), im = N * amp * g * std::sin(th);
      const int y = (j + ny) % ny;
      w[2 * ((size_t)y * nxh + i)] = re;
      w[2 * ((size_t)y * nxh + i) + 1] = im;
      if (i == 0) {
        const int yc = (ny - j) % ny;
        w[2 * ((size_t)yc * nxh)] = re;
        w[2 * ((size_t)yc * nxh) + 1] = -im;
      }
    }
  return AF_OK;
}

af_status af_ic_random(int nx, int ny, unsigned long long seed, int dealias, double* w) {
  if (!w || nx < 2 || ny < 2 || (nx & (nx - 1)) || (ny & (ny - 1))) return fail(AF_EINVAL, "bad arguments");
  const int nxh = nx / 2 + 1;
  // real field in the padded order of .cu:548-557
  std::mt19937_64 gen(seed);
  std::uniform_real_distribution<float> dist(-1.f, 1.f);
  std::vector<double> real((size_t)nx * ny);
  const size_t padded = (size_t)2 * ny * nxh;
  for (size_t r = 0; r < padded; ++r) {
    const size_t col = r % (size_t)(nx + 2);
    if (col == (size_t)nx || col == (size_t)nx + 1) continue;
    real[(r / (nx + 2)) * nx + col] = dist(gen);
  }
  // unnormalised 2-D R2C (.cu:665)
  std::vector<std::complex<double>> spec((size_t)ny * nxh);
  std::vector<std::complex<double>> row(nx), col(ny);
  for (int y = 0; y < ny; ++y) {
    for (int x = 0; x < nx; ++x) row[x] = real[(size_t)y * nx + x];
    fft_inplace(row, -1);
    for (int x = 0; x < nxh; ++x) spec[(size_t)y * nxh + x] = row[x];
  }
  for (int x = 0; x < nxh; ++x) {
    for (int y = 0; y < ny; ++y) col[y] = spec[(size_t)y * nxh + x];
    fft_inplace(col, -1);
    for (int y = 0; y < ny; ++y) spec[(size_t)y * nxh + x] = col[y];
  }
  // dealiasing_func (.cu:921-932) and w[0] = 0 (.cu:676-679)
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nxh; ++x) {
      const size_t c = (size_t)y * nxh + x;
      const bool keep = keep_host(x, y, nx, ny, dealias);
      w[2 * c] = keep ? spec[c].real() : 0.0;
      w[2 * c + 1] = keep ? spec[c].imag() : 0.0;
    }
  w[0] = 0;
  w[1] = 0;
  return AF_OK;
}

}  // extern "C"
