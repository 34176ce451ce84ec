// Plain-data structures shared by the host stepper and the device kernels.
#pragma once
#include <cstddef>

namespace af {

// Interleaved complex value as stored in HBM (8 bytes FP32, 16 bytes FP64).
template <class T> struct cx { T re, im; };

// Per-instance device state of the mean-velocity ODE and step bookkeeping.
struct MeanState {
  double M[2];        // unnormalised k=0 coefficient of u (snapshot convention)
  double Mo[2];       // value at the start of the current step
  double m[2];        // last injected stage input
  long long nstep;    // completed steps since set_state
  long long nan_step; // first step after which Re w^[0] was NaN, or -1
};

// Everything one kernel needs for one instance; an array of these drives the
// batched (ensemble) launches, a single element the single-run path.
template <class T>
struct InstDesc {
  cx<T>* wo;    // state omega^ (also omega_old during a step)
  cx<T>* acc;   // RK accumulator
  cx<T>* um;    // mixed v_x
  cx<T>* vm;    // mixed v_y
  cx<T>* wm;    // mixed omega
  cx<T>* pm;    // mixed 2P, P = -i kx (lam A + beta Cy)^x
  cx<T>* qm;    // mixed 2Q, Q = (lam W - beta Cx)^x  (the 1/2 is applied by the column pass)
  float* ifac;  // FP32 runs: integrating factors in column-block order, plane
                // 0 exp(sigma dt/2), plane 1 exp(sigma dt) (FP64-evaluated,
                // rounded once); null for FP64 runs (evaluated in the kernel)
  double* partials;   // [row tiles][4]: sum|u|^2, sum Cx, sum Cy, unused
  MeanState* mean;
  double* energy;     // ring of per-step stage-1 energies
  int ring_cap;
  int nx, ny;
  int mask_mode;      // 0 isotropic 2/3, 1 square 2/3
  int mean_scheme;    // 0 reference, 1 lawson
  int row_tiles;      // number of row blocks (partials entries)
  double dt, tkx, tky, c0, c2, c4, lam, beta, invN, e, e2;
  double dt2, dt3, dt6;   // dt/2, dt/3, dt/6
};

}  // namespace af
