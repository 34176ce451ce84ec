// Diagnostics of a run's current state (no reference counterpart; the
// reference only writes the energy, .cu:1118-1127): energy 1/2<|u|^2>,
// enstrophy 1/2<omega^2> and the shell spectrum E(k), unit-width shells in
// |k|/k_unit.  Energy and enstrophy are deterministic FP64 block sums reduced
// in block order on the host; spectrum bins are FP64 atomic sums (order of
// additions not fixed, ~1e-16 relative).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "engine_impl.hpp"

namespace af {

template <class T>
__global__ void __launch_bounds__(256) diag_kernel(const cx<T>* __restrict__ w, int nx2, int ny, int by, int bx,
                                                   double tkx, double tky, double kunit, double invN2,
                                                   double* __restrict__ part, double* __restrict__ spec, int nshell) {
  double e = 0, z = 0;
  const size_t n = (size_t)ny * nx2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(i / nx2), x = (int)(i % nx2);
    const size_t a = ((size_t)(y / by) * (nx2 / bx) + x / bx) * (size_t)(by * bx) + (y % by) * bx + x % bx;
    const double re = (double)w[a].re, im = (double)w[a].im;
    const int nys = (y <= ny / 2) ? y : y - ny;
    const double kx = tkx * x, ky = tky * nys, k2 = kx * kx + ky * ky;
    const double wt = (x == 0) ? 1.0 : 2.0;
    const double a2 = wt * (re * re + im * im) * invN2;
    z += 0.5 * a2;
    if (k2 > 0) {
      const double ek = 0.5 * a2 / k2;
      e += ek;
      const int s = (int)floor(sqrt(k2) / kunit + 0.5);
      if (s < nshell) atomicAdd(&spec[s], ek);
    }
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    e += __shfl_xor(e, off, 64);
    z += __shfl_xor(z, off, 64);
  }
  __shared__ double red[2 * 4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) { red[2 * wv] = e; red[2 * wv + 1] = z; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double se = red[0], sz = red[1];
    for (int k = 1; k < 4; ++k) { se += red[2 * k]; sz += red[2 * k + 1]; }
    part[2 * blockIdx.x] = se;
    part[2 * blockIdx.x + 1] = sz;
  }
}

constexpr int kDiagBlocks = 512;

af_status diagnostics(const af_params& p, const Geometry& g, unsigned char* base, hipStream_t stream, double* energy,
                      double* enstrophy, double* spectrum, int nshell) {
  const int blocks = kDiagBlocks;
  const size_t bytes = (size_t)2 * blocks * sizeof(double) + (size_t)(nshell > 0 ? nshell : 1) * sizeof(double);
  double* buf = nullptr;
  hipError_t e = hipMalloc(&buf, bytes);
  if (e != hipSuccess) return fail(AF_ENOMEM, "diagnostics scratch");
  double* part = buf;
  double* spec = buf + 2 * blocks;
  const double twopi = 6.283185307179586476925286766559;
  const double tkx = twopi / p.lx, tky = twopi / p.ly;
  const double kunit = tkx < tky ? tkx : tky;
  const double N = (double)p.nx * (double)p.ny;
  std::vector<double> hp(2 * blocks), hs(nshell > 0 ? nshell : 1);
  MeanState ms;
  af_status st = AF_OK;
  do {
    if ((e = hipMemsetAsync(spec, 0, (size_t)(nshell > 0 ? nshell : 1) * sizeof(double), stream)) != hipSuccess) {
      st = hip_fail(e, "diag clear"); break;
    }
    if (g.prec == 64)
      hipLaunchKernelGGL((diag_kernel<double>), dim3(blocks), dim3(256), 0, stream,
                         (const cx<double>*)(base + g.off_field[0]), g.nx2, g.ny, g.by, g.bx, tkx, tky, kunit,
                         1.0 / (N * N), part, spec, nshell);
    else
      hipLaunchKernelGGL((diag_kernel<float>), dim3(blocks), dim3(256), 0, stream,
                         (const cx<float>*)(base + g.off_field[0]), g.nx2, g.ny, g.by, g.bx, tkx, tky, kunit,
                         1.0 / (N * N), part, spec, nshell);
    if ((e = hipGetLastError()) != hipSuccess) { st = hip_fail(e, "diag launch"); break; }
    if ((e = hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, stream)) !=
        hipSuccess) { st = hip_fail(e, "diag copy"); break; }
    if ((e = hipMemcpyAsync(hs.data(), spec, hs.size() * sizeof(double), hipMemcpyDeviceToHost, stream)) !=
        hipSuccess) { st = hip_fail(e, "diag copy"); break; }
    if ((e = hipMemcpyAsync(&ms, base + g.off_mean, sizeof(ms), hipMemcpyDeviceToHost, stream)) != hipSuccess) {
      st = hip_fail(e, "diag copy"); break;
    }
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) { st = hip_fail(e, "diag sync"); break; }
    double se = 0, sz = 0;
    for (int b = 0; b < blocks; ++b) { se += hp[2 * b]; sz += hp[2 * b + 1]; }
    const double emean = 0.5 * (ms.M[0] * ms.M[0] + ms.M[1] * ms.M[1]) / (N * N);   // k = 0: mean flow
    if (energy) *energy = se + emean;
    if (enstrophy) *enstrophy = sz;
    if (spectrum && nshell > 0) {
      for (int s = 0; s < nshell; ++s) spectrum[s] = hs[s];
      spectrum[0] += emean;
    }
  } while (0);
  (void)hipFree(buf);
  return st;
}

// ---- allocation-free variant --------------------------------------------
// Device memory of af_run_diagnostics_noalloc: the block partials, the two
// results and up to kNoAllocMaxShells spectrum bins, in one static array per
// device (40 KiB), held until the process ends.  Its address is looked up
// once per device (hipGetSymbolAddress, that device current) and only that
// address is passed to kernels and copies.  Every call on a device shares the
// array: overlapping calls on one device race.
constexpr int kNoAllocMaxShells = 4096;
constexpr int kScratchOut = 2 * kDiagBlocks;     // energy, enstrophy
constexpr int kScratchSpec = kScratchOut + 2;    // spectrum bins
static __device__ __align__(256) double scratch[kScratchSpec + kNoAllocMaxShells];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices];

// The address of `scratch` on device dev (looked up once, with dev current).
static af_status scratch_on(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(AF_EDEVICE, "device index beyond the static scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int prev = -1;
    hipError_t e = hipGetDevice(&prev);
    if (e == hipSuccess && prev != dev) e = hipSetDevice(dev);
    void* ptr = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (prev >= 0 && prev != dev) (void)hipSetDevice(prev);
    if (e != hipSuccess) return hip_fail(e, "scratch address");
    if (!ptr) return fail(AF_EDEVICE, "scratch address is NULL");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return AF_OK;
}

// Zeroes n doubles (a kernel, not a memset, so the call stays a plain kernel
// sequence).
__global__ void __launch_bounds__(256) diag_clear(double* __restrict__ p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0.0;
}

// Sums the block partials in block order (as the host does for
// af_run_diagnostics), adds the mean flow's energy and writes the results.
__global__ void __launch_bounds__(64) diag_finish(const double* __restrict__ part, int nblocks,
                                                  const MeanState* __restrict__ ms, double N2,
                                                  double* __restrict__ out, double* __restrict__ spec, int nshell) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  double se = 0, sz = 0;
  for (int b = 0; b < nblocks; ++b) { se += part[2 * b]; sz += part[2 * b + 1]; }
  const double emean = 0.5 * (ms->M[0] * ms->M[0] + ms->M[1] * ms->M[1]) / N2;
  out[0] = se + emean;
  out[1] = sz;
  if (nshell > 0) spec[0] += emean;
}

af_status diagnostics_noalloc(const af_params& p, const Geometry& g, unsigned char* base, hipStream_t stream, int device,
                              double* energy, double* enstrophy, double* spectrum, int nshell) {
  if (nshell > kNoAllocMaxShells)
    return fail(AF_EINVAL, "nshell " + std::to_string(nshell) + " exceeds the static scratch maximum of " +
                               std::to_string(kNoAllocMaxShells) + " shells");
  double* sp = nullptr;
  af_status st = scratch_on(device, &sp);
  if (st != AF_OK) return st;
  double* part = sp;
  double* out = sp + kScratchOut;
  double* spec = sp + kScratchSpec;
  const double twopi = 6.283185307179586476925286766559;
  const double tkx = twopi / p.lx, tky = twopi / p.ly;
  const double kunit = tkx < tky ? tkx : tky;
  const double N = (double)p.nx * (double)p.ny;
  const int nclear = nshell > 0 ? nshell : 1;
  double res[2];
  hipError_t e = hipSuccess;
  hipLaunchKernelGGL(diag_clear, dim3((nclear + 255) / 256), dim3(256), 0, stream, spec, nclear);
  if (g.prec == 64)
    hipLaunchKernelGGL((diag_kernel<double>), dim3(kDiagBlocks), dim3(256), 0, stream,
                       (const cx<double>*)(base + g.off_field[0]), g.nx2, g.ny, g.by, g.bx, tkx, tky, kunit,
                       1.0 / (N * N), part, spec, nshell);
  else
    hipLaunchKernelGGL((diag_kernel<float>), dim3(kDiagBlocks), dim3(256), 0, stream,
                       (const cx<float>*)(base + g.off_field[0]), g.nx2, g.ny, g.by, g.bx, tkx, tky, kunit,
                       1.0 / (N * N), part, spec, nshell);
  hipLaunchKernelGGL(diag_finish, dim3(1), dim3(64), 0, stream, (const double*)part, kDiagBlocks,
                     (const MeanState*)(base + g.off_mean), N * N, out, spec, nshell);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "diag launch");
  e = hipMemcpyAsync(res, out, sizeof(res), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess && spectrum && nshell > 0)
    e = hipMemcpyAsync(spectrum, spec, (size_t)nshell * sizeof(double), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return hip_fail(e, "diag copy");
  if (energy) *energy = res[0];
  if (enstrophy) *enstrophy = res[1];
  return AF_OK;
}

// ---- engine: mean states of a class's runs ------------------------------
// One thread per descriptor entry copies that run's mean state (NaN flag,
// step count, M) into out[entry], pinned host memory the device writes
// directly, so the host reads every run's state after one event instead of
// one copy per run.  Empty entries (nx == 0) are skipped.
template <class T>
__global__ void __launch_bounds__(64) gather_mean_kernel(const InstDesc<T>* __restrict__ descs, int n,
                                                         MeanState* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const InstDesc<T>& d = descs[i];
  if (d.nx == 0) return;
  out[i] = *d.mean;
}

hipError_t gather_means(int prec, const void* descs, int n, MeanState* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 grid((n + 63) / 64), block(64);
  if (prec == 64)
    hipLaunchKernelGGL((gather_mean_kernel<double>), grid, block, 0, s, (const InstDesc<double>*)descs, n, out);
  else
    hipLaunchKernelGGL((gather_mean_kernel<float>), grid, block, 0, s, (const InstDesc<float>*)descs, n, out);
  return hipGetLastError();
}

}  // namespace af
