// Internals shared by the single-run stepper (run.cpp) and the ensemble
// engine (engine.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/actflow.h"
#include "layout.hpp"
#include "registry.hpp"
#include "status.hpp"

namespace af {

constexpr int kRingCap = 1024;   // per-run ring of per-step energies

// Sizes, launch configuration and the offsets of every buffer inside a run's
// single device allocation.
struct Geometry {
  int prec = 0, nx2 = 0, ny = 0, by = 0, bx = 0, lmax = 0;
  const Schedule* sched = nullptr;
  const RowCfg* rc = nullptr;
  const ColCfg* cc = nullptr;
  ConvLaunch to_bricks = nullptr, from_bricks = nullptr;
  size_t esz = 0, nfield = 0;
  int row_tiles = 0, col_tiles = 0;
  size_t off_field[7] = {0, 0, 0, 0, 0, 0, 0};  // wo, acc, um, vm, wm, pm, qm
  size_t off_partials = 0, off_mean = 0, off_energy = 0, off_desc = 0, off_pool = 0;
  size_t off_ifac = 0;   // FP32 integrating-factor table (2 float planes) when ifac
  bool ifac = false;     // FP32 with a column radix <= 8
  size_t bytes = 0;
};

af_status hip_fail(hipError_t e, const char* what);
af_status validate(const af_params* p);
Geometry geometry(const af_params* p, bool batched = false);
af_status init_device_block(const af_params* p, const Geometry& g, unsigned char* base, hipStream_t stream);
af_status nyquist_check(const Geometry& g, const double* omega_hat);
// Doubles of the reference-layout state (the staging buffer of set/get).
inline size_t state_staging_doubles(const Geometry& g) { return (size_t)g.ny * (g.nx2 + 1) * 2; }
af_status load_state(const Geometry& g, unsigned char* base, const double* omega_hat, const double M[2],
                     hipStream_t stream, double* keep = nullptr);
af_status store_state(const Geometry& g, unsigned char* base, double* omega_hat, double M[2], hipStream_t stream,
                      double* keep = nullptr);
hipError_t launch_step(const Geometry& g, const void* descs, int ninst, const void* pool, hipStream_t s);
af_status diagnostics_noalloc(const af_params& p, const Geometry& g, unsigned char* base, hipStream_t stream, int device,
                              double* energy, double* enstrophy, double* spectrum, int nshell);
af_status diagnostics(const af_params& p, const Geometry& g, unsigned char* base, hipStream_t stream, double* energy,
                      double* enstrophy, double* spectrum, int nshell);

}  // namespace af
