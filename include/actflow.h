/*
 * actflow.h -- C ABI of libactflow.so, the MI355X (gfx950) implementation of
 * the ActFlow time step (James-Wilczek vorticity equation with the k=0
 * mean-velocity ODE) from DominikSuchla/2D_active_matter_simulation.
 *
 * The reference exposes no in-process API: its step is inline in main()
 * (ActFlow_V3-2.cu:718-862) with state in globals and managed buffers
 * (.cu:183-207, 441-495).  Each entry point below names the reference code it
 * replaces.  No device pointer and no HIP/torch type crosses this boundary;
 * the caller owns every host array, which set/get copy.
 *
 * Conventions shared with the reference:
 *   - half spectrum of Ny rows x Nxh = Nx/2+1 columns, element c = y*Nxh + x
 *     (.cu:416-419), interleaved (re, im) doubles here;
 *   - omega_hat is the UNNORMALISED forward DFT (cuFFT R2C, .cu:507; the
 *     numpy rfft2 convention, image.py:69);
 *   - M[2] is the k=0 DFT coefficient of u, i.e. Nx*Ny*<u> (.cu:1053-1054);
 *   - energy of step m is 1/2<|u|^2> of the state before step m
 *     (.cu:1118-1129), labelled m*dt in energy.txt (.cu:1275).
 */
#ifndef ACTFLOW_H
#define ACTFLOW_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  AF_OK = 0,
  AF_EINVAL = 1,       /* bad argument (sizes, precision, mask mode, ...)      */
  AF_ENOMEM = 2,       /* device or host allocation failed / over HBM budget   */
  AF_EDEVICE = 3,      /* HIP runtime error                                    */
  AF_ENAN = 4,         /* Re omega_hat[0] became NaN (check_func, .cu:884-897)  */
  AF_EIO = 5,          /* file I/O failure                                     */
  AF_EUNSUPPORTED = 6  /* valid for the reference but not on this backend      */
} af_status;

/* Run parameters.  sigma(k) = c0 + c2 k^2 + c4 k^4 is the linear operator.
 * The reference's -(1+lap)^2 - alpha (.cu:1030) is c0 = -1-alpha, c2 = 2,
 * c4 = -1 (af_params_reference); the James-Wilczek form -alpha + Gamma0 lap
 * - Gamma2 lap^2 is c0 = -alpha, c2 = -Gamma0, c4 = -Gamma2 (af_params_jw). */
typedef struct {
  int nx, ny;             /* powers of two, 32..4096 each                    */
  double lx, ly;          /* box lengths (.cu:187-188)                        */
  double dt;              /* time step (.cu:185)                              */
  double lambda, beta;    /* advection, cubic friction (.cu:199-203)          */
  double c0, c2, c4;      /* sigma(k) coefficients                            */
  int precision;          /* 32: FP32 fields/FFTs, FP64 IF factors and ODE;   *
                           * 64: FP64 throughout                              */
  int dealias;            /* 0 isotropic 2/3 (default), 1 square 2/3          */
  int mean_scheme;        /* 0 reference update (.cu:756-821), 1 lawson       */
} af_params;

typedef struct af_run af_run;

const char* af_status_str(af_status s);
/* Last error message of this thread (HIP error text or argument check). */
const char* af_last_error(void);

/* Fill p from the reference's parameters (alpha of .cu:201, c0=-1-alpha,
 * c2=2, c4=-1), precision 64, isotropic 2/3, reference mean scheme. */
void af_params_reference(af_params* p, int nx, int ny, double lx, double ly, double dt,
                         double lambda, double alpha, double beta);
/* Same with the James-Wilczek coefficients (c0=-alpha, c2=-Gamma0, c4=-Gamma2). */
void af_params_jw(af_params* p, int nx, int ny, double lx, double ly, double dt,
                  double lambda, double alpha, double beta, double gamma0, double gamma2);

/* Replaces the setup of .cu:409-521 and 696-705 (derived constants, the 22
 * cudaMallocManaged buffers, cufftPlan2d x2, wavenumber/cutoff/lin tables).
 * Allocates the run's device memory on the current HIP device. */
af_status af_run_create(const af_params* p, af_run** out);
/* Replaces the state load of .cu:615-634 / the IC hand-off of .cu:660-679.
 * omega_hat: 2*Ny*(Nx/2+1) doubles (reference layout).  M: 2 doubles.
 * The kx = Nx/2 column is not represented on this backend (every supported
 * mask removes it from the nonlinear term): a state whose kx = Nx/2 column
 * is nonzero is refused with AF_EUNSUPPORTED (dropping it would change the
 * run); zero that column to load such a state.  (The run's step graphs are
 * captured and uploaded by af_run_create, so no capture happens between
 * set_state and the first steps.) */
af_status af_run_set_state(af_run* r, const double* omega_hat, const double M[2]);
/* Replaces the device->host copy of save_func (.cu:1259-1260). */
af_status af_run_get_state(af_run* r, double* omega_hat, double M[2]);
/* Replaces nsteps iterations of the time loop .cu:718-821 (and check_func,
 * .cu:850-861).  energy: nsteps doubles or NULL.  Returns AF_ENAN if
 * Re omega_hat[0] became NaN; *nan_step (if non-NULL) then receives the
 * 1-based step count since set_state at which it happened. */
af_status af_run_step(af_run* r, int nsteps, double* energy, long long* nan_step);
/* Steps completed since set_state. */
long long af_run_steps_done(const af_run* r);
/* Diagnostics of the current state (no reference counterpart; the reference
 * writes only the energy, .cu:1118-1127): energy 1/2<|u|^2> including the
 * mean flow, enstrophy 1/2<omega^2>, and (if nshell > 0) the shell spectrum
 * spectrum[s] = sum of 1/2|u^(k)|^2/N^2 over |k|/k_unit in [s-1/2, s+1/2),
 * k_unit = min(2pi/Lx, 2pi/Ly); the mean flow is in shell 0. */
af_status af_run_diagnostics(af_run* r, double* energy, double* enstrophy, double* spectrum, int nshell);
/* The same diagnostics without any allocation: no hipMalloc/hipFree and no
 * host allocation.  No counterpart in the reference; for callers that cannot
 * allocate (tight serving loops, harnesses that forbid device allocation).
 * Its device memory is a static per-device array of 40 KiB (block partials,
 * results and up to 4096 spectrum bins); nshell > 4096 returns AF_EINVAL.
 * Results equal af_run_diagnostics's: energy and enstrophy bitwise, the
 * spectrum to the rounding of its atomic sums.  Calls that overlap on one
 * device share that array and race: serialise them. */
af_status af_run_diagnostics_noalloc(af_run* r, double* energy, double* enstrophy, double* spectrum, int nshell);
void af_run_destroy(af_run* r);

/* Per-kernel timing with HIP events on the run's stream: runs 5*nsteps steps,
 * nsteps of them with each launch recording an event pair at its kernel's own
 * start and end (hipExtLaunchKernel), each of those after a 4-step graph so
 * that the device runs the timed kernels back to back as in af_run_step, and
 * returns the mean duration in milliseconds of the row pass and of the column
 * pass (stages 1-4) over the timed steps, and the wall time of the whole
 * region.  No reference counterpart (measurement hook for bench.py). */
af_status af_run_profile(af_run* r, int nsteps, double* row_ms, double* col_ms, double* total_ms);
/* Use (1) or bypass (0) a captured hipGraph for one step (default 1). */
af_status af_run_use_graph(af_run* r, int on);
/* Bytes of device memory held by a run with these parameters (an af_run
 * also keeps, from its first set_state or get_state, a staging buffer of
 * 16*Ny*(Nx/2+1) bytes for the reference-layout state). */
size_t af_run_device_bytes(const af_params* p);

/* ---- schedules (no reference counterpart: cuFFT plans, .cu:504-508) ----
 * A schedule fixes the brick shape of the fields in HBM and, per pass, the
 * radix and how many fields share an LDS exchange.  Index 0 is the built-in
 * default; runs take the schedule active for their size when created
 * (af_schedule_select, else the tuned tables written by autotune.py next to
 * the library -- schedules_batched.txt first for runs admitted to an engine,
 * then schedules.txt -- else the default).  af_schedule_active reports the
 * single-run choice. */
typedef struct {
  int by, bx;               /* brick rows x columns                           */
  int row_radix, row_group; /* row pass radix, fields per LDS exchange        */
  int col_radix, col_fwd_group, col_inv_group;
  int col_width;            /* columns per column-pass block (divides bx)     */
  int options;              /* bits: 1 two sets of LDS areas, one barrier per *
                             * FFT exchange instead of two (more LDS per      *
                             * block); 2 streaming (nt) column-pass stores;   *
                             * 4 streaming row-pass stores instead of device- *
                             * scope write-through (sc1); 8 streaming column- *
                             * pass loads of P, Q; 16 FP32 integrating        *
                             * factors from a per-run table also at radix 16; *
                             * 32 paired rows (two rows of a row block as one *
                             * full-length complex transform); 64 half-brick  *
                             * row blocks (by/2 rows per row-pass block);     *
                             * 128 P, Q in bricks of 2*by rows; 256 (with 64) *
                             * U, V, W in bricks of by/2 rows (64, 128, 256:  *
                             * FP32 schedules).  2, 4 and 16 act on FP32 runs *
                             * only (FP64 stores take the default policy).    */
} af_schedule;
int af_schedule_count(int precision, int nx, int ny);
af_status af_schedule_get(int precision, int nx, int ny, int index, af_schedule* out);
/* Process-wide choice for runs created afterwards; index -1 clears it. */
af_status af_schedule_select(int precision, int nx, int ny, int index);
af_status af_schedule_active(int precision, int nx, int ny, af_schedule* out);

/* ---- ensemble engine: many resident runs, batched launches ------------- */
typedef struct af_engine af_engine;
/* hbm_budget: bytes the engine may hold in run state (admission limit). */
af_status af_engine_create(size_t hbm_budget, af_engine** out);
/* Admit a run; fails with AF_ENOMEM if it would exceed the budget.  The
 * state is copied in; *slot receives its handle. */
af_status af_engine_add(af_engine* e, const af_params* p, const double* omega_hat,
                        const double M[2], int* slot);
af_status af_engine_remove(af_engine* e, int slot);
/* Advance every resident run by nsteps steps, all runs of one size class in
 * the same kernel launches.  Per-run results are bitwise identical to stepping
 * the run alone in an engine (and to af_run_step whenever the engine's tuned
 * schedule for the size is the single-run one; otherwise they agree to the
 * rounding of the precision). */
af_status af_engine_step(af_engine* e, int nsteps);
/* Advance only the listed slots (same batching within the list). */
af_status af_engine_step_slots(af_engine* e, const int* slots, int nslots, int nsteps);
/* Single-launch mode (off by default): when the size classes stepped
 * together are exactly an instantiated mix (the FP32 256^2 + 512^2 + 1024^2
 * ensemble with 256-lane schedules in both passes), every pass of a step is
 * ONE launch over all runs of all those classes instead of one launch per
 * class on concurrent streams; results are bitwise those of the per-class
 * launches with the same schedules.  Otherwise stepping is unchanged.  No
 * reference counterpart (the reference runs one grid per process). */
af_status af_engine_set_single_launch(af_engine* e, int on);
/* 1 if single-launch mode is on and the engine's current classes form a mix. */
int af_engine_single_launch_active(af_engine* e);
af_status af_engine_get_state(af_engine* e, int slot, double* omega_hat, double M[2]);
/* Copies the last n per-step energies of a slot (oldest first). */
af_status af_engine_get_energy(af_engine* e, int slot, int n, double* energy);
long long af_engine_steps_done(const af_engine* e, int slot);
af_status af_engine_diagnostics(af_engine* e, int slot, double* energy, double* enstrophy, double* spectrum,
                                int nshell);
size_t af_engine_bytes_in_use(const af_engine* e);
int af_engine_count(const af_engine* e);
void af_engine_destroy(af_engine* e);

/* ---- serving loop (no reference counterpart) -------------------------- */
typedef struct {
  double arrival_s;          /* arrival, seconds after af_serve starts           */
  af_params params;
  int horizon;               /* steps to run                                     */
  int snapshot_every;        /* steps between snapshots, 0 = none                */
  unsigned long long seed;   /* band-limited IC (af_ic_bandlimited)              */
  double ic_amp, ic_kmax;
} af_request;
typedef struct {
  double admit_s, finish_s;          /* seconds after start; -1 if never       */
  double step_p50_s, step_p99_s;     /* interval between consecutive steps     */
  int snapshots;                     /* snapshots delivered                    */
  int status;                        /* af_status of this request              */
  double M[2];                       /* final mean velocity coefficient        */
} af_request_result;
typedef struct {
  double wall_s, grid_point_steps, throughput;   /* throughput: pt-steps/s      */
  double step_p50_s, step_p99_s;     /* over every step of every request        */
  double ttc_p50_s, ttc_p99_s;       /* time to completion, finish - arrival    */
  double queue_p50_s, queue_p99_s;   /* admission wait, admit - arrival         */
  int max_resident, ticks, completed;
  size_t peak_bytes;                 /* peak run-state bytes held               */
} af_serve_stats;
/* Called on the calling thread for every snapshot, in completion order;
 * omega_hat is in the reference layout (valid during the call only). */
typedef void (*af_snapshot_fn)(void* user, int request, long long step, const double* omega_hat,
                               const double M[2], int nx, int ny);
/* Runs the requests with continuous batching: admission in arrival order by
 * the HBM budget; the resident runs of one size class advance together in
 * chunks of at most max_tick_steps (a chunk ends at the next snapshot or
 * completion step of any of the class's runs), each class on its own as soon
 * as its previous chunk completes; snapshots copied back asynchronously,
 * finished runs retired so waiting requests can enter.  stats->ticks counts
 * the chunks of all classes.  final_states (optional) holds per request NULL or a buffer of
 * 2*Ny*(Nx/2+1) doubles for the final state.  A request's result is bitwise
 * the result of running it alone.  A request whose state goes NaN fails alone
 * (result status AF_ENAN, finish_s -1, the check_func stop of .cu:850-861);
 * the others go on and af_serve returns AF_OK. */
af_status af_serve(const af_request* reqs, int n, size_t hbm_budget, int max_tick_steps, af_snapshot_fn fn,
                   void* user, double** final_states, af_request_result* results, af_serve_stats* stats);

/* ---- host formats (reference file contract) --------------------------- */
typedef struct {
  double T, dt, Lx, Ly;   /* .cu:184-188 */
  int Nx, Ny;             /* .cu:189-190 */
  double lambda, alpha, beta, dt_snap;   /* .cu:199-203, 196 */
  char initial[256];      /* initial_condition line (.cu:192)  */
  char directory[1024];   /* save_directory (.cu:194)          */
  /* extension keys (absent from the reference) */
  int has_gamma;          /* Gamma0:/Gamma2: given -> JW coefficients */
  double gamma0, gamma2;
  int precision, dealias, mean_scheme, snapshot_precision;
  int M, snap;            /* round(T/dt), round(dt_snap/dt) (.cu:409-420) */
} af_spec;

/* Defaults of .cu:183-204 (note Ly and Ny are bound at static init). */
void af_spec_defaults(af_spec* s);
/* Parse <dir>/specifications.txt with the semantics of .cu:320-408.
 * err (optional, errlen bytes) receives the reference's message on failure. */
af_status af_spec_read(const char* path, af_spec* s, char* err, size_t errlen);
/* Write the example file of the `spec` subcommand (.cu:250-266). */
af_status af_spec_write_example(const char* path, const af_spec* s);
/* Snapshot text format of save_func (.cu:1245-1266): "re,im," per element
 * then "Mx,My", at `precision` significant digits (6 = reference). */
af_status af_snapshot_write(const char* path, int nx, int ny, const double* omega_hat,
                            const double M[2], int precision);
/* Reader with the semantics of .cu:615-634 (>> then ignore up to ','). */
af_status af_snapshot_read(const char* path, int nx, int ny, double* omega_hat, double M[2]);
/* One energy line "t E\n" appended (.cu:1268-1278). */
af_status af_energy_append(const char* path, double t, double E);

/* Band-limited seeded initial condition (no reference counterpart; the
 * reference's `random` IC is seeded from the clock, .cu:539-547):
 * c_n = amp * g * exp(i theta) on 0 < |n|^2 <= kmax^2 (lattice indices),
 * g ~ N(0,1) by Box-Muller, theta ~ U[0,2pi), SplitMix64(seed), Hermitian on
 * the nx = 0 column, stored as Nx*Ny*c_n (unnormalised DFT).  M = 0. */
af_status af_ic_bandlimited(int nx, int ny, unsigned long long seed, double amp, double kmax,
                            double* omega_hat);
/* The reference's `random [n]` IC (.cu:535-566): U(-1,1) real field from
 * mt19937_64(seed), R2C on the host, masked, omega_hat[0] = 0.  seed is
 * explicit here (the reference derives it from the clock). */
af_status af_ic_random(int nx, int ny, unsigned long long seed, int dealias, double* omega_hat);

#ifdef __cplusplus
}
#endif
#endif
