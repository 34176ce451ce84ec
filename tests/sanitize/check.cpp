// Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the file
// formats and initial conditions of the C ABI (csrc/host_io.cpp, HIP-free)
// and the CPU oracle (oracle/oracle.cpp), driven through their C entry points
// on valid and malformed inputs.  Built and run by tests/test_sanitize.py;
// any sanitizer report aborts the program (-fno-sanitize-recover=all), any
// failed check returns 1.
//   check <golden dir holding specifications.txt> <scratch dir>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/actflow.h"
#include "../../oracle/oracle.h"

static int failures = 0;
#define CHECK(c)                                                    \
  do {                                                              \
    if (!(c)) {                                                     \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      ++failures;                                                   \
    }                                                               \
  } while (0)

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  const std::string golden = argv[1], scratch = argv[2];

  // specifications.txt: the reference's example, the example writer, bad input
  af_spec s;
  af_spec_defaults(&s);
  char err[512];
  CHECK(af_spec_read(golden.c_str(), &s, err, sizeof err) == AF_OK);
  CHECK(s.Nx == 1024 && s.Ny == 1024 && s.M == 1000000 && s.snap == 200000);
  CHECK(af_spec_write_example((scratch + "/specifications.txt").c_str(), &s) == AF_OK);
  af_spec s2;
  af_spec_defaults(&s2);
  CHECK(af_spec_read(scratch.c_str(), &s2, err, sizeof err) == AF_OK);
  CHECK(af_spec_read((scratch + "/no_such_dir").c_str(), &s2, err, sizeof err) != AF_OK);
  CHECK(af_spec_read(scratch.c_str(), &s2, err, 1) == AF_OK);   // 1-byte message buffer

  // initial conditions, snapshot round trip at 17 digits, a truncated snapshot
  for (int nx : {32, 64}) {
    const int ny = 2 * nx;
    const size_t n = 2 * (size_t)ny * (nx / 2 + 1);
    std::vector<double> w(n), r(n);
    double M[2] = {1.5, -2.25}, Mr[2];
    CHECK(af_ic_bandlimited(nx, ny, 7, 1e-3, 6.0, w.data()) == AF_OK);
    CHECK(af_ic_random(nx, ny, 42, 0, r.data()) == AF_OK);
    CHECK(af_ic_bandlimited(nx, ny, 7, 1e-3, (double)nx, w.data()) != AF_OK);   // kmax beyond Nyquist
    const std::string path = scratch + "/snap.bin";
    CHECK(af_snapshot_write(path.c_str(), nx, ny, w.data(), M, 17) == AF_OK);
    CHECK(af_snapshot_read(path.c_str(), nx, ny, r.data(), Mr) == AF_OK);
    CHECK(std::memcmp(w.data(), r.data(), n * sizeof(double)) == 0 && Mr[0] == M[0] && Mr[1] == M[1]);
    std::vector<double> big(2 * (size_t)ny * (nx + 1));
    CHECK(af_snapshot_read(path.c_str(), 2 * nx, ny, big.data(), Mr) != AF_OK);   // too short for the grid
    CHECK(af_snapshot_read((scratch + "/missing.bin").c_str(), nx, ny, r.data(), Mr) != AF_OK);
  }
  {   // a state formatted in several blocks on several threads (512 x 512: 131584 modes)
    const int nx = 512, ny = 512;
    const size_t n = 2 * (size_t)ny * (nx / 2 + 1);
    std::vector<double> w(n), r(n);
    double M[2] = {0.125, -3.5e-7}, Mr[2];
    for (size_t i = 0; i < n; ++i) w[i] = std::sin(0.731 * (double)i) * std::pow(10.0, (double)(i % 25) - 12.0);
    const std::string path = scratch + "/snap_big.bin";
    CHECK(af_snapshot_write(path.c_str(), nx, ny, w.data(), M, 17) == AF_OK);
    CHECK(af_snapshot_read(path.c_str(), nx, ny, r.data(), Mr) == AF_OK);
    CHECK(std::memcmp(w.data(), r.data(), n * sizeof(double)) == 0 && Mr[0] == M[0] && Mr[1] == M[1]);
  }
  CHECK(af_energy_append((scratch + "/energy.txt").c_str(), 0.005, 1.25) == AF_OK);

  // oracle: transforms and a few steps of both precisions, square and not
  for (int prec : {64, 32}) {
    const int nx = 32, ny = 64, nxh = nx / 2 + 1;
    std::vector<double> x((size_t)nx * ny), back((size_t)nx * ny), X(2 * (size_t)ny * nxh);
    for (size_t i = 0; i < x.size(); ++i) x[i] = std::sin(0.37 * (double)i) + 0.1 * (double)(i % 7);
    oracle_rfft2(ny, nx, prec, x.data(), X.data());
    oracle_irfft2(ny, nx, prec, X.data(), back.data());
    double worst = 0;
    for (size_t i = 0; i < x.size(); ++i) worst = std::fmax(worst, std::fabs(back[i] / (nx * ny) - x[i]));
    CHECK(worst < (prec == 64 ? 1e-12 : 1e-4));
    std::vector<double> in(2 * 16), out(2 * 16);
    for (int i = 0; i < 32; ++i) in[i] = 0.25 * i;
    oracle_fft1(16, -1, prec, in.data(), out.data());

    oracle_run* o = oracle_create(nx, ny, 2 * M_PI, 4 * M_PI, 2e-4, 3.5, 0.01, 0.8, 0.045, -2.5e-5, prec, 0, 0, 1);
    CHECK(o != nullptr);
    std::vector<double> w(2 * (size_t)ny * nxh), nl(w.size()), en(3);
    CHECK(af_ic_bandlimited(nx, ny, 3, 0.5, 6.0, w.data()) == AF_OK);
    const double M0[2] = {0.1 * nx * ny, -0.05 * nx * ny};
    double g[2], e1;
    oracle_set_state(o, w.data(), M0);
    oracle_rhs(o, w.data(), M0, nl.data(), g, &e1);
    CHECK(oracle_step(o, 3, en.data()) == 0);
    CHECK(std::isfinite(en[2]) && en[2] > 0);
    double Mo[2];
    oracle_get_state(o, w.data(), Mo);
    CHECK(oracle_mask(o, 0) == 1.0);
    oracle_destroy(o);
  }
  CHECK(oracle_create(30, 32, 1, 1, 1e-3, 0, 0, 0, 0, 0, 64, 0, 0, 1) == nullptr);   // not a power of two

  if (failures) std::fprintf(stderr, "%d checks failed\n", failures);
  else std::printf("sanitized host checks ok\n");
  return failures ? 1 : 0;
}
