// Host stepper behind the C ABI: one af_run = one grid resident on one HIP
// device, driven on its own stream.  Replaces the setup and time loop of
// ActFlow_V3-2.cu:409-862 (managed buffers, cuFFT plans, 29 host syncs per
// step) with one device allocation, a captured hipGraph of the 8 launches of a
// step, and one synchronisation per chunk of steps.
#include "run_impl.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace af {

af_status hip_fail(hipError_t e, const char* what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return AF_EDEVICE;
}

static bool pow2_in(int n, int lo, int hi) { return n >= lo && n <= hi && (n & (n - 1)) == 0; }

af_status validate(const af_params* p) {
  if (!p) return fail(AF_EINVAL, "params is NULL");
  if (!pow2_in(p->nx, 32, 4096) || !pow2_in(p->ny, 32, 4096))
    return fail(AF_EUNSUPPORTED, "nx and ny must be powers of two in [32, 4096]");
  if (p->precision != 32 && p->precision != 64) return fail(AF_EINVAL, "precision must be 32 or 64");
  if (p->dealias != 0 && p->dealias != 1)
    return fail(AF_EUNSUPPORTED, "dealias must be 0 (isotropic 2/3) or 1 (square 2/3) on this backend");
  if (p->mean_scheme != 0 && p->mean_scheme != 1) return fail(AF_EINVAL, "mean_scheme must be 0 or 1");
  if (!(p->lx > 0) || !(p->ly > 0) || !std::isfinite(p->lx) || !std::isfinite(p->ly))
    return fail(AF_EINVAL, "box lengths must be positive");
  if (!std::isfinite(p->dt) || !std::isfinite(p->lambda) || !std::isfinite(p->beta) || !std::isfinite(p->c0) ||
      !std::isfinite(p->c2) || !std::isfinite(p->c4))
    return fail(AF_EINVAL, "non-finite parameter");
  return AF_OK;
}

// exp(-2 pi i m / L) in long double with the angle reduced to [0, pi/4] by
// the symmetries of the unit circle, then rounded once.
static void unit_root(long m, long L, long double* c, long double* s) {
  const long double pi = 3.141592653589793238462643383279502884L;
  m %= L;
  // 8 m / L in [0, 8): octant o, offset f in [0, 1)
  const long num = 8 * m;
  const long o = num / L;
  const long double f = (long double)(num - o * L) / (long double)L;  // fraction of an octant
  const long double a = f * pi / 4;                                   // in [0, pi/4)
  const long double ca = cosl(a), sa = sinl(a), cb = cosl(pi / 4 - a), sb = sinl(pi / 4 - a);
  long double cc, ss;  // cos, sin of theta = 2 pi m / L = o*pi/4 + a
  switch (o) {
    case 0: cc = ca; ss = sa; break;
    case 1: cc = sb; ss = cb; break;       // pi/2 - (pi/4 - a)
    case 2: cc = -sa; ss = ca; break;      // pi/2 + a
    case 3: cc = -cb; ss = sb; break;      // pi - (pi/4 - a)
    case 4: cc = -ca; ss = -sa; break;     // pi + a
    case 5: cc = -sb; ss = -cb; break;     // 3pi/2 - (pi/4 - a)
    case 6: cc = sa; ss = -ca; break;      // 3pi/2 + a
    default: cc = cb; ss = -sb; break;     // 2pi - (pi/4 - a)
  }
  *c = cc;
  *s = -ss;
}

template <class T>
static void build_pool(std::vector<unsigned char>& out, int lmax) {
  std::vector<cx<T>> v((size_t)2 * lmax, cx<T>{0, 0});
  for (int L = 1; L <= lmax; L *= 2)
    for (int m = 0; m < L; ++m) {
      long double c, s;
      unit_root(m, L, &c, &s);
      v[(size_t)L + m] = cx<T>{(T)c, (T)s};
    }
  out.resize(v.size() * sizeof(cx<T>));
  std::memcpy(out.data(), v.data(), out.size());
}

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

Geometry geometry(const af_params* p, bool batched) {
  Geometry g;
  g.prec = p->precision;
  g.nx2 = p->nx / 2;
  g.ny = p->ny;
  g.sched = active_schedule(g.prec, g.nx2, g.ny, batched);
  if (g.sched) {
    g.by = g.sched->by;
    g.bx = g.sched->bx;
    g.rc = find_row_cfg(g.prec, g.sched);
    g.cc = find_col_cfg(g.prec, g.sched);
  }
  g.to_bricks = to_bricks_launch(g.prec);
  g.from_bricks = from_bricks_launch(g.prec);
  g.esz = g.prec == 64 ? 2 * sizeof(double) : 2 * sizeof(float);
  g.nfield = (size_t)g.ny * g.nx2;
  g.row_tiles = g.rc ? g.ny / g.rc->rb : 0;
  g.col_tiles = g.cc ? g.nx2 / g.cc->w : 0;
  g.lmax = p->nx > p->ny ? p->nx : p->ny;
  size_t off = 0;
  for (int i = 0; i < 7; ++i) { g.off_field[i] = off; off += align256(g.nfield * g.esz); }
  g.off_partials = off; off += align256((size_t)g.row_tiles * 4 * sizeof(double));
  g.off_mean = off; off += align256(sizeof(MeanState));
  g.off_energy = off; off += align256((size_t)kRingCap * sizeof(double));
  g.off_desc = off; off += align256(sizeof(InstDesc<double>));
  g.off_pool = off; off += align256((size_t)2 * g.lmax * g.esz);
  g.off_ifac = off;
  // the column pass reads the table (kTable): FP32 at radix <= 8, or by option
  g.ifac = g.prec == 32 && g.cc && (g.cc->r <= 8 || (g.cc->pp & kOptIfTable));
  if (g.ifac) off += align256(2 * g.nfield * sizeof(float));
  g.bytes = off;
  return g;
}

template <class T>
static InstDesc<T> make_desc(const af_params* p, const Geometry& g, unsigned char* base) {
  InstDesc<T> d;
  std::memset(&d, 0, sizeof(d));
  cx<T>** f[7] = {&d.wo, &d.acc, &d.um, &d.vm, &d.wm, &d.pm, &d.qm};
  for (int i = 0; i < 7; ++i) *f[i] = reinterpret_cast<cx<T>*>(base + g.off_field[i]);
  d.ifac = g.ifac ? reinterpret_cast<float*>(base + g.off_ifac) : nullptr;
  d.partials = reinterpret_cast<double*>(base + g.off_partials);
  d.mean = reinterpret_cast<MeanState*>(base + g.off_mean);
  d.energy = reinterpret_cast<double*>(base + g.off_energy);
  d.ring_cap = kRingCap;
  d.nx = p->nx;
  d.ny = p->ny;
  d.mask_mode = p->dealias;
  d.mean_scheme = p->mean_scheme;
  d.row_tiles = g.row_tiles;
  const double twopi = 6.283185307179586476925286766559;
  d.dt = p->dt;
  d.tkx = twopi / p->lx;
  d.tky = twopi / p->ly;
  d.c0 = p->c0;
  d.c2 = p->c2;
  d.c4 = p->c4;
  d.lam = p->lambda;
  d.beta = p->beta;
  d.invN = 1.0 / ((double)p->nx * (double)p->ny);
  d.e = std::exp(p->c0 * p->dt);
  d.e2 = std::exp(p->c0 * p->dt * 0.5);
  d.dt2 = p->dt / 2;
  d.dt3 = p->dt / 3;
  d.dt6 = p->dt / 6;
  return d;
}

// Writes the descriptor, twiddle pool and a fresh mean state into the run's
// allocation (synchronously on `stream`).
af_status init_device_block(const af_params* p, const Geometry& g, unsigned char* base, hipStream_t stream) {
  std::vector<unsigned char> pool;
  if (g.prec == 64) build_pool<double>(pool, g.lmax); else build_pool<float>(pool, g.lmax);
  hipError_t e = hipMemcpyAsync(base + g.off_pool, pool.data(), pool.size(), hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) return hip_fail(e, "upload twiddles");
  if (g.prec == 64) {
    InstDesc<double> d = make_desc<double>(p, g, base);
    e = hipMemcpyAsync(base + g.off_desc, &d, sizeof(d), hipMemcpyHostToDevice, stream);
  } else {
    InstDesc<float> d = make_desc<float>(p, g, base);
    e = hipMemcpyAsync(base + g.off_desc, &d, sizeof(d), hipMemcpyHostToDevice, stream);
  }
  if (e != hipSuccess) return hip_fail(e, "upload descriptor");
  if (g.ifac) {
    e = ifac_init_launch()(base + g.off_desc, g.nx2, g.ny, g.cc->w, stream);
    if (e != hipSuccess) return hip_fail(e, "integrating factors");
  }
  MeanState ms;
  std::memset(&ms, 0, sizeof(ms));
  ms.nan_step = -1;
  e = hipMemcpyAsync(base + g.off_mean, &ms, sizeof(ms), hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) return hip_fail(e, "upload mean state");
  e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return hip_fail(e, "init sync");
  return AF_OK;
}

// The kx = Nx/2 column is not stored on this backend (every supported mask
// removes it from the nonlinear term).  A state that carries it would be
// changed by dropping it, so such a state is refused instead.
af_status nyquist_check(const Geometry& g, const double* omega_hat) {
  const size_t nxh = (size_t)g.nx2 + 1;
  for (int y = 0; y < g.ny; ++y) {
    const double* v = omega_hat + 2 * ((size_t)y * nxh + g.nx2);
    if (v[0] != 0.0 || v[1] != 0.0)
      return fail(AF_EUNSUPPORTED, "the state has a nonzero kx = Nx/2 column (row " + std::to_string(y) +
                                       "), which this backend does not represent; zero it to load the state");
  }
  return AF_OK;
}

// Reference-layout state -> bricks + mean, then the stage-0 column pass.
// `keep` (optional): a device buffer of state_staging_doubles(g) doubles to
// use instead of a temporary allocation.
af_status load_state(const Geometry& g, unsigned char* base, const double* omega_hat, const double M[2],
                     hipStream_t stream, double* keep) {
  const size_t n = state_staging_doubles(g);
  double* staging = keep;
  hipError_t e = keep ? hipSuccess : hipMalloc(&staging, n * sizeof(double));
  if (e != hipSuccess) return fail(AF_ENOMEM, std::string("staging: ") + hipGetErrorString(e));
  af_status st = AF_OK;
  MeanState ms;
  std::memset(&ms, 0, sizeof(ms));
  ms.M[0] = M[0];
  ms.M[1] = M[1];
  ms.Mo[0] = M[0];
  ms.Mo[1] = M[1];
  ms.nan_step = -1;
  do {
    if ((e = hipMemcpyAsync(staging, omega_hat, n * sizeof(double), hipMemcpyHostToDevice, stream)) != hipSuccess) {
      st = hip_fail(e, "upload state"); break;
    }
    if ((e = g.to_bricks(staging, base + g.off_field[0], g.nx2 * 2, g.ny, g.by, g.bx, stream)) != hipSuccess) {
      st = hip_fail(e, "to_bricks"); break;
    }
    if ((e = hipMemcpyAsync(base + g.off_mean, &ms, sizeof(ms), hipMemcpyHostToDevice, stream)) != hipSuccess) {
      st = hip_fail(e, "upload mean"); break;
    }
    if ((e = g.cc->launch(base + g.off_desc, 1, g.col_tiles, 0, base + g.off_pool, stream)) != hipSuccess) {
      st = hip_fail(e, "initial column pass"); break;
    }
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) { st = hip_fail(e, "set_state sync"); break; }
  } while (0);
  if (!keep) (void)hipFree(staging);
  return st;
}

af_status store_state(const Geometry& g, unsigned char* base, double* omega_hat, double M[2], hipStream_t stream,
                      double* keep) {
  const size_t n = state_staging_doubles(g);
  double* staging = keep;
  hipError_t e = keep ? hipSuccess : hipMalloc(&staging, n * sizeof(double));
  if (e != hipSuccess) return fail(AF_ENOMEM, std::string("staging: ") + hipGetErrorString(e));
  af_status st = AF_OK;
  MeanState ms;
  do {
    if ((e = g.from_bricks(base + g.off_field[0], staging, g.nx2 * 2, g.ny, g.by, g.bx, stream)) != hipSuccess) {
      st = hip_fail(e, "from_bricks"); break;
    }
    if ((e = hipMemcpyAsync(omega_hat, staging, n * sizeof(double), hipMemcpyDeviceToHost, stream)) != hipSuccess) {
      st = hip_fail(e, "download state"); break;
    }
    if ((e = hipMemcpyAsync(&ms, base + g.off_mean, sizeof(ms), hipMemcpyDeviceToHost, stream)) != hipSuccess) {
      st = hip_fail(e, "download mean"); break;
    }
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) { st = hip_fail(e, "get_state sync"); break; }
    if (M) { M[0] = ms.M[0]; M[1] = ms.M[1]; }
  } while (0);
  if (!keep) (void)hipFree(staging);
  return st;
}

// The 8 launches of one step: row pass and column pass for stages 1..4.
hipError_t launch_step(const Geometry& g, const void* descs, int ninst, const void* pool, hipStream_t s) {
  for (int stage = 1; stage <= 4; ++stage) {
    hipError_t e = g.rc->launch(descs, ninst, g.row_tiles, pool, s);
    if (e != hipSuccess) return e;
    e = g.cc->launch(descs, ninst, g.col_tiles, stage, pool, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace af

using namespace af;

// Step graphs of 1, 2, 4, 8 and 16 steps: a chunk of k steps replays the
// binary decomposition of k (16-step graphs, then at most one of each smaller
// size), so there are few gaps between graph launches at any k.
constexpr int kGraphLevels = 5;
constexpr int kGraphSteps = 1 << (kGraphLevels - 1);

// Host side of the per-chunk hand-back, mirroring the run's device block from
// its mean state (NaN flag, step count) to the end of its energy ring, in
// pinned memory: one direct DMA copy per chunk.
struct HostChunk {
  MeanState ms;
  unsigned char pad[256 - sizeof(MeanState)];
  double energy[kRingCap];
};
static_assert(sizeof(MeanState) <= 256, "the mean state fits its 256-byte slot");

struct af_run {
  af_params p;
  Geometry g;
  hipStream_t stream = nullptr;
  unsigned char* base = nullptr;
  HostChunk* host = nullptr;        // pinned
  double* staging = nullptr;        // reference-layout state for set/get, allocated on first use
  hipGraph_t graph[kGraphLevels] = {};       // graph[l]: 2^l steps
  hipGraphExec_t gexec[kGraphLevels] = {};
  bool use_graph = true;
  bool graph_failed = false;
  bool ready = false;
  long long steps = 0;
  int device = 0;                   // the HIP device current at creation
};

static af_status prepare_graphs(af_run* r);

extern "C" {

void af_params_reference(af_params* p, int nx, int ny, double lx, double ly, double dt, double lambda, double alpha,
                         double beta) {
  std::memset(p, 0, sizeof(*p));
  p->nx = nx; p->ny = ny; p->lx = lx; p->ly = ly; p->dt = dt;
  p->lambda = lambda; p->beta = beta;
  p->c0 = -1.0 - alpha; p->c2 = 2.0; p->c4 = -1.0;
  p->precision = 64; p->dealias = 0; p->mean_scheme = 0;
}

void af_params_jw(af_params* p, int nx, int ny, double lx, double ly, double dt, double lambda, double alpha,
                  double beta, double gamma0, double gamma2) {
  af_params_reference(p, nx, ny, lx, ly, dt, lambda, alpha, beta);
  p->c0 = -alpha; p->c2 = -gamma0; p->c4 = -gamma2;
}

size_t af_run_device_bytes(const af_params* p) {
  if (validate(p) != AF_OK) return 0;
  return geometry(p).bytes;
}

af_status af_run_create(const af_params* p, af_run** out) {
  if (!out) return fail(AF_EINVAL, "out is NULL");
  *out = nullptr;
  af_status st = validate(p);
  if (st != AF_OK) return st;
  af_run* r = new af_run();
  r->p = *p;
  r->g = geometry(p);
  if (!r->g.rc || !r->g.cc) { delete r; return fail(AF_EUNSUPPORTED, "no kernel for this size"); }
  if (r->g.off_energy - r->g.off_mean != offsetof(HostChunk, energy)) {   // layout HostChunk mirrors
    delete r;
    return fail(AF_EDEVICE, "internal: energy ring does not follow the mean state");
  }
  hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete r; return hip_fail(e, "stream create"); }
  e = hipGetDevice(&r->device);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(r->stream);
    delete r;
    return hip_fail(e, "current device");
  }
  e = hipMalloc(&r->base, r->g.bytes);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(r->stream);
    delete r;
    return fail(AF_ENOMEM, std::string("device allocation: ") + hipGetErrorString(e));
  }
  e = hipHostMalloc((void**)&r->host, sizeof(HostChunk), hipHostMallocDefault);
  if (e != hipSuccess) {
    r->host = nullptr;
    af_run_destroy(r);
    return fail(AF_ENOMEM, std::string("pinned host buffer: ") + hipGetErrorString(e));
  }
  st = init_device_block(p, r->g, r->base, r->stream);
  // the step graphs depend only on the run's buffers: captured, instantiated
  // and uploaded once, here, so that nothing of that host work sits between a
  // caller's set_state and its first steps
  if (st == AF_OK) st = prepare_graphs(r);
  if (st != AF_OK) { af_run_destroy(r); return st; }
  *out = r;
  return AF_OK;
}

static af_status ensure_staging(af_run* r) {
  if (r->staging) return AF_OK;
  hipError_t e = hipMalloc(&r->staging, state_staging_doubles(r->g) * sizeof(double));
  if (e != hipSuccess) {
    r->staging = nullptr;
    return fail(AF_ENOMEM, std::string("staging: ") + hipGetErrorString(e));
  }
  return AF_OK;
}

static af_status prepare_graphs(af_run* r);

af_status af_run_set_state(af_run* r, const double* omega_hat, const double M[2]) {
  if (!r || !omega_hat || !M) return fail(AF_EINVAL, "NULL argument");
  r->ready = false;
  r->steps = 0;
  af_status st = nyquist_check(r->g, omega_hat);
  if (st == AF_OK) st = ensure_staging(r);
  if (st == AF_OK) st = load_state(r->g, r->base, omega_hat, M, r->stream, r->staging);
  r->ready = (st == AF_OK);
  return st;
}

af_status af_run_get_state(af_run* r, double* omega_hat, double M[2]) {
  if (!r || !omega_hat || !M) return fail(AF_EINVAL, "NULL argument");
  if (!r->ready) return fail(AF_EINVAL, "state not set");
  af_status st = ensure_staging(r);
  if (st != AF_OK) return st;
  return store_state(r->g, r->base, omega_hat, M, r->stream, r->staging);
}

// Captures `steps` steps into (graph, exec); false (and nothing kept) on failure.
static bool capture_steps(af_run* r, int steps, hipGraph_t* graph, hipGraphExec_t* exec) {
  const void* desc = r->base + r->g.off_desc;
  const void* pool = r->base + r->g.off_pool;
  hipError_t e = hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    hipError_t le = hipSuccess;
    for (int i = 0; i < steps && le == hipSuccess; ++i) le = launch_step(r->g, desc, 1, pool, r->stream);
    e = hipStreamEndCapture(r->stream, graph);
    if (le == hipSuccess && e == hipSuccess) e = hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0);
    else if (e == hipSuccess) e = le;
  }
  if (e != hipSuccess || !*exec) {
    (void)hipGetLastError();
    if (*graph) { (void)hipGraphDestroy(*graph); *graph = nullptr; }
    *exec = nullptr;
    return false;
  }
  return true;
}

// Captures, instantiates and uploads the step graphs of 1, 2, 4, 8 and 16 steps (once per run: they
// depend only on the run's buffers, not on its state), so that the first
// chunk of steps after set_state already replays resident graphs.  A failed
// capture is not an error: the run then launches its kernels directly.
static af_status prepare_graphs(af_run* r) {
  if (!r->use_graph || r->graph_failed) return AF_OK;
  for (int l = 0; l < kGraphLevels && !r->graph_failed; ++l)
    if (!r->gexec[l] && !capture_steps(r, 1 << l, &r->graph[l], &r->gexec[l])) r->graph_failed = true;
  if (r->graph_failed) return AF_OK;
  hipError_t e = hipSuccess;
  for (int l = 0; l < kGraphLevels && e == hipSuccess; ++l) e = hipGraphUpload(r->gexec[l], r->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess) return hip_fail(e, "graph upload");
  return AF_OK;
}

// Enqueues 2^level steps: as a captured graph, or as plain launches when
// graphs are off or could not be captured.
static hipError_t run_steps(af_run* r, int level) {
  const int steps = 1 << level;
  if (r->use_graph && !r->graph_failed) {
    if (r->gexec[level] || capture_steps(r, steps, &r->graph[level], &r->gexec[level]))
      return hipGraphLaunch(r->gexec[level], r->stream);
    r->graph_failed = true;
  }
  const void* desc = r->base + r->g.off_desc;
  const void* pool = r->base + r->g.off_pool;
  hipError_t e = hipSuccess;
  for (int i = 0; i < steps && e == hipSuccess; ++i) e = launch_step(r->g, desc, 1, pool, r->stream);
  return e;
}

af_status af_run_step(af_run* r, int nsteps, double* energy, long long* nan_step) {
  if (!r) return fail(AF_EINVAL, "run is NULL");
  if (nsteps < 0) return fail(AF_EINVAL, "nsteps < 0");
  if (!r->ready) return fail(AF_EINVAL, "state not set");
  int done = 0;
  while (done < nsteps) {
    const int k = (nsteps - done) < kRingCap ? (nsteps - done) : kRingCap;
    hipError_t e = hipSuccess;
    for (int i = 0; i + kGraphSteps <= k && e == hipSuccess; i += kGraphSteps) e = run_steps(r, kGraphLevels - 1);
    for (int l = kGraphLevels - 2; l >= 0 && e == hipSuccess; --l)
      if (k & (1 << l)) e = run_steps(r, l);
    // the chunk's mean state and, if asked for, the energy ring (the k new
    // entries are steps .. steps+k-1 mod the ring size): one copy into pinned
    // memory of the mean state alone or of the block from it to the ring's end
    // (the ring follows the mean state's 256-byte slot), one synchronisation
    if (e == hipSuccess)
      e = hipMemcpyAsync(r->host, r->base + r->g.off_mean, energy ? sizeof(HostChunk) : sizeof(MeanState),
                         hipMemcpyDeviceToHost, r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) {
      // some steps of the chunk may have run: the host step count no longer
      // matches the device's, so the run needs a new set_state
      r->ready = false;
      return hip_fail(e, "step");
    }
    if (energy)
      for (int i = 0; i < k; ++i) energy[done + i] = r->host->energy[(r->steps + i) % kRingCap];
    r->steps += k;
    done += k;
    const MeanState& ms = r->host->ms;
    if (ms.nan_step >= 0) {
      if (nan_step) *nan_step = ms.nan_step;
      return fail(AF_ENAN, "array contains nan values after " + std::to_string(ms.nan_step) + " timesteps");
    }
  }
  return AF_OK;
}

long long af_run_steps_done(const af_run* r) { return r ? r->steps : -1; }

af_status af_run_diagnostics(af_run* r, double* energy, double* enstrophy, double* spectrum, int nshell) {
  if (!r || nshell < 0 || (nshell > 0 && !spectrum)) return fail(AF_EINVAL, "bad arguments");
  if (!r->ready) return fail(AF_EINVAL, "state not set");
  return diagnostics(r->p, r->g, r->base, r->stream, energy, enstrophy, spectrum, nshell);
}

af_status af_run_diagnostics_noalloc(af_run* r, double* energy, double* enstrophy, double* spectrum, int nshell) {
  if (!r || nshell < 0 || (nshell > 0 && !spectrum)) return fail(AF_EINVAL, "bad arguments");
  if (!r->ready) return fail(AF_EINVAL, "state not set");
  return diagnostics_noalloc(r->p, r->g, r->base, r->stream, r->device, energy, enstrophy, spectrum, nshell);
}

af_status af_run_use_graph(af_run* r, int on) {
  if (!r) return fail(AF_EINVAL, "run is NULL");
  r->use_graph = on != 0;
  return AF_OK;
}

af_status af_run_profile(af_run* r, int nsteps, double* row_ms, double* col_ms, double* total_ms) {
  if (!r || nsteps <= 0) return fail(AF_EINVAL, "bad arguments");
  if (!r->ready) return fail(AF_EINVAL, "state not set");
  // events at each kernel's own start and end, plus two around the whole run.
  // Every timed step follows a graph of kLead steps: enqueueing a launch with
  // its event pair takes the host longer than the kernel takes the device, so
  // timed launches alone would leave the device idle between kernels; behind
  // the graph they run back to back, as the graph-launched steps of
  // af_run_step do.
  constexpr int kLeadLevel = 2, kLead = 1 << kLeadLevel;
  const int nl = 8 * nsteps;
  std::vector<hipEvent_t> ev(2 * nl + 2, nullptr);
  for (auto& x : ev)
    if (hipEventCreate(&x) != hipSuccess) {
      for (auto& y : ev)
        if (y) (void)hipEventDestroy(y);
      return fail(AF_EDEVICE, "event create");
    }
  const void* desc = r->base + r->g.off_desc;
  const void* pool = r->base + r->g.off_pool;
  hipError_t e = hipEventRecord(ev[2 * nl], r->stream);
  int k = 0;
  for (int s = 0; s < nsteps && e == hipSuccess; ++s) {
    g_launch_events = {};   // the lead steps record no events (plain launches when graphs are off)
    e = run_steps(r, kLeadLevel);
    for (int stage = 1; stage <= 4 && e == hipSuccess; ++stage) {
      g_launch_events = {ev[2 * k], ev[2 * k + 1]};
      ++k;
      e = r->g.rc->launch(desc, 1, r->g.row_tiles, pool, r->stream);
      g_launch_events = {ev[2 * k], ev[2 * k + 1]};
      ++k;
      if (e == hipSuccess) e = r->g.cc->launch(desc, 1, r->g.col_tiles, stage, pool, r->stream);
    }
  }
  g_launch_events = {};
  if (e == hipSuccess) e = hipEventRecord(ev[2 * nl + 1], r->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  double rs = 0, cs = 0;
  if (e == hipSuccess) {
    for (int i = 0; i < nl && e == hipSuccess; ++i) {
      float ms = 0;
      e = hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]);
      if (i % 2 == 0) rs += ms; else cs += ms;
    }
    float tot = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&tot, ev[2 * nl], ev[2 * nl + 1]);
    if (e == hipSuccess) {
      if (row_ms) *row_ms = rs / (4.0 * nsteps);
      if (col_ms) *col_ms = cs / (4.0 * nsteps);
      if (total_ms) *total_ms = tot;
      r->steps += (long long)nsteps * (kLead + 1);
    }
  }
  // some steps may have run: the host step count no longer matches the device's
  if (e != hipSuccess) r->ready = false;
  for (auto& x : ev) (void)hipEventDestroy(x);
  if (e != hipSuccess) return hip_fail(e, "profile");
  return AF_OK;
}

static void fill_schedule(const Schedule* s, af_schedule* out) {
  out->by = s->by;
  out->bx = s->bx;
  out->row_radix = s->row_r;
  out->row_group = s->row_g;
  out->col_radix = s->col_r;
  out->col_fwd_group = s->col_gf;
  out->col_inv_group = s->col_gi;
  out->col_width = s->col_w;
  out->options = s->pp;
}

int af_schedule_count(int precision, int nx, int ny) { return schedule_count(precision, nx / 2, ny); }

af_status af_schedule_get(int precision, int nx, int ny, int index, af_schedule* out) {
  const Schedule* s = schedule_at(precision, nx / 2, ny, index);
  if (!s || !out) return fail(AF_EINVAL, "no such schedule");
  fill_schedule(s, out);
  return AF_OK;
}

af_status af_schedule_select(int precision, int nx, int ny, int index) {
  return select_schedule(precision, nx / 2, ny, index) ? AF_OK : fail(AF_EINVAL, "no such schedule");
}

af_status af_schedule_active(int precision, int nx, int ny, af_schedule* out) {
  const Schedule* s = active_schedule(precision, nx / 2, ny);
  if (!s || !out) return fail(AF_EINVAL, "no schedule for this size");
  fill_schedule(s, out);
  return AF_OK;
}

void af_run_destroy(af_run* r) {
  if (!r) return;
  if (r->stream) (void)hipStreamSynchronize(r->stream);
  for (int l = 0; l < kGraphLevels; ++l) {
    if (r->gexec[l]) (void)hipGraphExecDestroy(r->gexec[l]);
    if (r->graph[l]) (void)hipGraphDestroy(r->graph[l]);
  }
  if (r->base) (void)hipFree(r->base);
  if (r->host) (void)hipHostFree(r->host);
  if (r->staging) (void)hipFree(r->staging);
  if (r->stream) (void)hipStreamDestroy(r->stream);
  delete r;
}

}  // extern "C"
