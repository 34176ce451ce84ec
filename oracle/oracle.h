/*
 * CPU oracle for the ActFlow time step -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the correctness checker and the CPU baseline.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product library (libactflow.so) never links, loads or calls it.
 *
 * It restates the reference step of ActFlow_V3-2.cu line by line:
 *   - stage RHS          nonl_func / modified_nonl_func   .cu:1036-1161
 *   - Biot-Savart         nonl_subfunc_0                  .cu:1163-1181
 *   - real-space products nonl_subfunc_1                  .cu:1183-1206
 *   - N-hat assembly      nonl_subfunc_2                  .cu:1208-1225
 *   - Hermitian projection reality_func                   .cu:1228-1243
 *   - IF-RK stages        integrate_func_0..4             .cu:940-1018
 *   - mean-velocity ODE   host updates                    .cu:729-732,756-821
 *   - energy              modified_nonl_func              .cu:1116-1127
 *   - wavenumbers         wavenumber_func                 .cu:1329-1349
 *   - integrating factors lin_func + lin_v/lin2_v         .cu:1020-1033,699-705
 * with its own CPU FFT (iterative radix-2, twiddles in long double) standing
 * in for cuFFT's unnormalised 2-D R2C/C2R (.cu:504-508).
 *
 * Array layout is the reference's: a half spectrum of Ny rows by Nxh=Nx/2+1
 * columns, element c = y*Nxh + x, interleaved (re, im) doubles at the ABI.
 */
#ifndef ACTFLOW_ORACLE_H
#define ACTFLOW_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_run oracle_run;

/* dealias: 0 isotropic 2/3, 1 square 2/3, 2 none, 3 reference literal
 *          (32-bit wrapped radius of .cu:911, i.e. keep all at 256^2, only
 *          k=0 from 512^2 on).
 * mean_scheme: 0 reference (no e*M_old term, .cu:756-821), 1 lawson.
 * precision: 32 or 64 (arithmetic type of fields and FFTs; integrating
 *            factors are evaluated in double and rounded; reductions in double). */
oracle_run* oracle_create(int nx, int ny, double lx, double ly, double dt,
                          double lambda, double beta, double c0, double c2,
                          double c4, int precision, int dealias,
                          int mean_scheme, int nthreads);
void oracle_destroy(oracle_run*);
void oracle_set_state(oracle_run*, const double* omega_hat, const double M[2]);
void oracle_get_state(const oracle_run*, double* omega_hat, double M[2]);
/* Runs nsteps full steps; energy[i] (if non-NULL) receives the stage-1
 * energy of step i (the state before that step, .cu:1118-1129).
 * Returns 0, or 1 if Re omega_hat[0] became NaN (.cu:884-897). */
int oracle_step(oracle_run*, int nsteps, double* energy);
/* One evaluation of the stage RHS F(s, m) -> (N-hat, g, energy), .cu:1036-1161. */
void oracle_rhs(oracle_run*, const double* s_hat, const double m[2],
                double* nonl_hat, double g[2], double* energy);
/* Dealiasing mask value at half-spectrum element c (1.0 or 0.0). */
double oracle_mask(const oracle_run*, int c);

/* Stand-alone transforms (unnormalised, reference layout) for FFT tests. */
void oracle_rfft2(int ny, int nx, int precision, const double* real_in,
                  double* half_out);
void oracle_irfft2(int ny, int nx, int precision, const double* half_in,
                   double* real_out);
void oracle_fft1(int n, int sign, int precision, const double* in, double* out);

#ifdef __cplusplus
}
#endif
#endif
