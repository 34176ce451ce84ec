// Measurement tool (not part of the library): what a pass boundary costs on
// one gfx950 GPU, for the launch shape of the 1024^2 single-run passes
// (256 workgroups of 512 lanes, one per CU).
//
//   graph:      one kernel per phase, phases chained in a hipGraph (how the
//               library steps: the boundary is the end of one kernel and the
//               start of the next);
//   persistent: one cooperative launch, phases separated by a grid barrier
//               (what a persistent step kernel would need).  The L2s of the 8
//               XCDs are not coherent within a launch, so a barrier that
//               publishes a phase's stores to the other XCDs needs an
//               agent-scope release (L2 write-back) before it arrives and an
//               agent-scope acquire (L2 invalidate) after it leaves.
//               Mode 1: every lane fences; mode 2: the block's lanes meet at
//               a workgroup barrier and lane 0 fences and arrives; mode 3:
//               arrivals are counted per XCD (the XCC_ID hardware register),
//               and only the last workgroup of an XCD to arrive writes that
//               XCD's L2 back and arrives at the global count, after every
//               workgroup of the XCD waited for its own stores to reach L2.
//
// Each phase every block writes NV float4 per lane into its slice of one of
// two buffers; after the boundary each block reads the slice another block
// wrote and counts wrong values (0 expected).  Spin waits are bounded: past
// the limit a block raises an abort flag that every spinning block reads, so
// all waves exit.  Usage: probe_grid_barrier [phases]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                                    \
  do {                                                                                           \
    hipError_t e_ = (x);                                                                         \
    if (e_ != hipSuccess) {                                                                      \
      std::printf("HIP error %s at line %d: %s\n", #x, __LINE__, hipGetErrorString(e_));         \
      std::exit(2);                                                                              \
    }                                                                                            \
  } while (0)

namespace {

constexpr int kLanes = 512;
constexpr int kBlocks = 256;
constexpr int kMaxNv = 8;
constexpr unsigned kSpinLimit = 1u << 22;

__device__ __forceinline__ float4 tag(int p, int b, int t, int v) {
  return make_float4((float)p, (float)b, (float)t, (float)v);
}

template <int NV, bool NT>
__device__ __forceinline__ void write_phase(float4* buf, int p, int b) {
  const int t = threadIdx.x;
  float4* s = buf + (size_t)b * kLanes * NV;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const float4 q = tag(p, b, t, v);
    if constexpr (NT) {
      typedef float f4v __attribute__((ext_vector_type(4)));
      f4v w = {q.x, q.y, q.z, q.w};
      __builtin_nontemporal_store(w, reinterpret_cast<f4v*>(s + v * kLanes + t));
    } else {
      s[v * kLanes + t] = q;
    }
  }
}

template <int NV>
__device__ __forceinline__ void check_phase(const float4* buf, int p, int b, int nb, unsigned* err) {
  const int t = threadIdx.x;
  const int o = (b * 37 + 11 + p) % nb;   // another block, a different one each phase
  const float4* s = buf + (size_t)o * kLanes * NV;
  int bad = 0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const float4 x = s[v * kLanes + t], e = tag(p, o, t, v);
    bad += (x.x != e.x) | (x.y != e.y) | (x.z != e.z) | (x.w != e.w);
  }
  if (bad) atomicAdd(err, (unsigned)bad);
}

template <int MODE>
__device__ __forceinline__ void grid_barrier(unsigned* count, unsigned target, unsigned* abort_flag) {
  if (MODE == 1) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    if (MODE == 2) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    while (__hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      if (++spins > kSpinLimit) {
        __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  if (MODE == 1) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Waits until *c >= target or the abort flag is up (raising it past the spin
// limit); lane 0 only.
__device__ __forceinline__ void spin_until(unsigned* c, unsigned target, unsigned* abort_flag) {
  unsigned spins = 0;
  while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    if (++spins > kSpinLimit) {
      __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

__device__ __forceinline__ unsigned xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 15u; }

// Mode 3 (see the header).  xcd_count[16] counts arrivals per XCD (monotonic),
// pop[x] is the number of workgroups on XCD x and nxcd the number of XCDs used.
__device__ __forceinline__ void grid_barrier_xcd(unsigned* count, unsigned* xcd_count, unsigned pop, unsigned nxcd,
                                                 unsigned p, unsigned* abort_flag) {
  __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): this wave's stores have reached the L2
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned old = __hip_atomic_fetch_add(&xcd_count[xcc_id()], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old + 1 == (p + 1) * pop) {   // last of its XCD: every store of the XCD is in its L2
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    spin_until(count, (p + 1) * nxcd, abort_flag);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

template <int NV, int MODE, bool NT>
__global__ void __launch_bounds__(kLanes) persistent(float4* b0, float4* b1, int phases, unsigned* count,
                                                     unsigned* abort_flag, unsigned* err) {
  const int b = blockIdx.x, nb = gridDim.x;
  // mode 3: count + 1 holds the start count, count + 16.. the per-XCD
  // populations and count + 32.. the per-XCD arrival counts
  __shared__ unsigned pop_nx[2];
  if constexpr (MODE == 3) {
    if (threadIdx.x == 0) {
      __hip_atomic_fetch_add(&count[16 + xcc_id()], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&count[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      spin_until(&count[1], (unsigned)nb, abort_flag);
      unsigned nx = 0;
      for (int x = 0; x < 16; ++x) nx += __hip_atomic_load(&count[16 + x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 0;
      pop_nx[0] = __hip_atomic_load(&count[16 + xcc_id()], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      pop_nx[1] = nx;
    }
    __syncthreads();
  }
  for (int p = 0; p < phases; ++p) {
    float4* buf = (p & 1) ? b1 : b0;
    write_phase<NV, NT>(buf, p, b);
    if constexpr (MODE == 3)
      grid_barrier_xcd(count, count + 32, pop_nx[0], pop_nx[1], (unsigned)p, abort_flag);
    else
      grid_barrier<MODE>(count, (unsigned)(p + 1) * nb, abort_flag);
    check_phase<NV>(buf, p, b, nb, err);
  }
}

// one phase per kernel: check what the previous phase wrote, write this one
template <int NV, bool NT>
__global__ void __launch_bounds__(kLanes) phase_kernel(float4* b0, float4* b1, int p, unsigned* err) {
  const int b = blockIdx.x, nb = gridDim.x;
  if (p > 0) check_phase<NV>(((p - 1) & 1) ? b1 : b0, p - 1, b, nb, err);
  write_phase<NV, NT>((p & 1) ? b1 : b0, p, b);
}

struct Bufs {
  float4 *b0, *b1;
  unsigned *count, *abort_flag, *err;
};

template <int NV, bool NT>
void run_graph(const Bufs& m, int phases, hipStream_t s) {
  hipGraph_t g;
  hipGraphExec_t ge;
  CK(hipStreamBeginCapture(s, hipStreamCaptureModeGlobal));
  for (int p = 0; p < phases; ++p) phase_kernel<NV, NT><<<kBlocks, kLanes, 0, s>>>(m.b0, m.b1, p, m.err);
  CK(hipStreamEndCapture(s, &g));
  CK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
  CK(hipGraphUpload(ge, s));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  CK(hipGraphLaunch(ge, s));   // untimed warm-up
  CK(hipStreamSynchronize(s));
  CK(hipMemset(m.err, 0, 4));
  CK(hipEventRecord(e0, s));
  CK(hipGraphLaunch(ge, s));
  CK(hipEventRecord(e1, s));
  CK(hipStreamSynchronize(s));
  float ms = 0;
  CK(hipEventElapsedTime(&ms, e0, e1));
  unsigned bad = 0;
  CK(hipMemcpy(&bad, m.err, 4, hipMemcpyDeviceToHost));
  std::printf("graph      NV=%d %s: %8.3f us per phase (%d phases, %.1f MB written and read per phase), "
              "wrong values %u\n",
              NV, NT ? "nt stores     " : "default stores", ms * 1e3 / phases, phases,
              kBlocks * kLanes * NV * 16.0 / 1e6, bad);
  std::fflush(stdout);
  CK(hipEventDestroy(e0));
  CK(hipEventDestroy(e1));
  CK(hipGraphExecDestroy(ge));
  CK(hipGraphDestroy(g));
}

template <int NV, int MODE, bool NT>
void run_persistent(const Bufs& m, int phases, hipStream_t s) {
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  for (int rep = 0; rep < 2; ++rep) {
    CK(hipMemset(m.count, 0, 48 * sizeof(unsigned)));
    CK(hipMemset(m.abort_flag, 0, 4));
    CK(hipMemset(m.err, 0, 4));
    CK(hipDeviceSynchronize());
    float4 *b0 = m.b0, *b1 = m.b1;
    unsigned *count = m.count, *abort_flag = m.abort_flag, *err = m.err;
    void* args[] = {&b0, &b1, &phases, &count, &abort_flag, &err};
    CK(hipEventRecord(e0, s));
    CK(hipLaunchCooperativeKernel((const void*)persistent<NV, MODE, NT>, dim3(kBlocks), dim3(kLanes), args, 0, s));
    CK(hipEventRecord(e1, s));
    CK(hipStreamSynchronize(s));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, e0, e1));
    unsigned bad = 0, aborted = 0;
    CK(hipMemcpy(&bad, m.err, 4, hipMemcpyDeviceToHost));
    CK(hipMemcpy(&aborted, m.abort_flag, 4, hipMemcpyDeviceToHost));
    std::printf("persistent NV=%d %s mode %d: %8.3f us per phase (%d phases), wrong values %u, aborted %u\n", NV,
                NT ? "nt stores     " : "default stores", MODE, ms * 1e3 / phases, phases, bad, aborted);
    std::fflush(stdout);
    if (aborted) std::exit(3);
  }
  CK(hipEventDestroy(e0));
  CK(hipEventDestroy(e1));
}

}  // namespace

int main(int argc, char** argv) {
  const int phases = argc > 1 ? std::atoi(argv[1]) : 1000;
  if (phases < 1 || phases > 100000) {
    std::printf("phases must be in [1, 100000]\n");
    return 1;
  }
  int dev = 0, coop = 0, occ = 0;
  CK(hipGetDevice(&dev));
  CK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  if (!coop) {
    std::printf("cooperative launch not supported\n");
    return 1;
  }
  CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)persistent<kMaxNv, 1, false>, kLanes, 0));
  std::printf("%d workgroups of %d lanes; persistent kernel occupancy %d workgroups per CU\n", kBlocks, kLanes, occ);
  Bufs m;
  const size_t n = (size_t)kBlocks * kLanes * kMaxNv;
  CK(hipMalloc(&m.b0, n * sizeof(float4)));
  CK(hipMalloc(&m.b1, n * sizeof(float4)));
  CK(hipMalloc(&m.count, 48 * sizeof(unsigned)));
  CK(hipMalloc(&m.abort_flag, 4));
  CK(hipMalloc(&m.err, 4));
  hipStream_t s;
  CK(hipStreamCreate(&s));
  run_graph<1, false>(m, phases, s);
  run_graph<8, false>(m, phases, s);
  run_graph<8, true>(m, phases, s);
  run_persistent<1, 1, false>(m, phases, s);
  run_persistent<1, 2, false>(m, phases, s);
  run_persistent<8, 1, false>(m, phases, s);
  run_persistent<8, 2, false>(m, phases, s);
  run_persistent<8, 2, true>(m, phases, s);
  run_persistent<1, 3, false>(m, phases, s);
  run_persistent<8, 3, false>(m, phases, s);
  CK(hipStreamDestroy(s));
  CK(hipFree(m.b0));
  CK(hipFree(m.b1));
  CK(hipFree(m.count));
  CK(hipFree(m.abort_flag));
  CK(hipFree(m.err));
  std::printf("done\n");
  return 0;
}
