"""Python binding of libactflow.so, the MI355X (gfx950) ActFlow time step.

Thin ctypes layer over the C ABI in include/actflow.h; every method calls the
HIP library and raises if it is missing or returns an error.  There is no CPU
fallback: the compute path is the HIP extension or nothing.

The package directory name starts with a digit, so import it with
``importlib.import_module("2d_active_matter_simulation_amd")``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libactflow.so")
CLI_PATH = os.path.join(_HERE, "actflow")

AF_OK, AF_EINVAL, AF_ENOMEM, AF_EDEVICE, AF_ENAN, AF_EIO, AF_EUNSUPPORTED = range(7)


class ActFlowError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[status {status}] {msg}")
        self.status = status


class AfParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("lambda_", ctypes.c_double), ("beta", ctypes.c_double),
        ("c0", ctypes.c_double), ("c2", ctypes.c_double), ("c4", ctypes.c_double),
        ("precision", ctypes.c_int), ("dealias", ctypes.c_int), ("mean_scheme", ctypes.c_int),
    ]


class AfSchedule(ctypes.Structure):
    _fields_ = [("by", ctypes.c_int), ("bx", ctypes.c_int), ("row_radix", ctypes.c_int),
                ("row_group", ctypes.c_int), ("col_radix", ctypes.c_int), ("col_fwd_group", ctypes.c_int),
                ("col_inv_group", ctypes.c_int), ("col_width", ctypes.c_int), ("options", ctypes.c_int)]

    def as_tuple(self):
        return (self.by, self.bx, self.row_radix, self.row_group, self.col_radix, self.col_fwd_group,
                self.col_inv_group, self.col_width, self.options)


class AfRequest(ctypes.Structure):
    _fields_ = [("arrival_s", ctypes.c_double), ("params", AfParams), ("horizon", ctypes.c_int),
                ("snapshot_every", ctypes.c_int), ("seed", ctypes.c_ulonglong), ("ic_amp", ctypes.c_double),
                ("ic_kmax", ctypes.c_double)]


class AfRequestResult(ctypes.Structure):
    _fields_ = [("admit_s", ctypes.c_double), ("finish_s", ctypes.c_double), ("step_p50_s", ctypes.c_double),
                ("step_p99_s", ctypes.c_double), ("snapshots", ctypes.c_int), ("status", ctypes.c_int),
                ("M", ctypes.c_double * 2)]


class AfServeStats(ctypes.Structure):
    _fields_ = [("wall_s", ctypes.c_double), ("grid_point_steps", ctypes.c_double), ("throughput", ctypes.c_double),
                ("step_p50_s", ctypes.c_double), ("step_p99_s", ctypes.c_double), ("ttc_p50_s", ctypes.c_double),
                ("ttc_p99_s", ctypes.c_double), ("queue_p50_s", ctypes.c_double), ("queue_p99_s", ctypes.c_double),
                ("max_resident", ctypes.c_int), ("ticks", ctypes.c_int), ("completed", ctypes.c_int),
                ("peak_bytes", ctypes.c_size_t)]


SNAPSHOT_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong,
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                               ctypes.c_int)


class AfSpec(ctypes.Structure):
    _fields_ = [
        ("T", ctypes.c_double), ("dt", ctypes.c_double), ("Lx", ctypes.c_double), ("Ly", ctypes.c_double),
        ("Nx", ctypes.c_int), ("Ny", ctypes.c_int),
        ("lambda_", ctypes.c_double), ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
        ("dt_snap", ctypes.c_double),
        ("initial", ctypes.c_char * 256), ("directory", ctypes.c_char * 1024),
        ("has_gamma", ctypes.c_int), ("gamma0", ctypes.c_double), ("gamma2", ctypes.c_double),
        ("precision", ctypes.c_int), ("dealias", ctypes.c_int), ("mean_scheme", ctypes.c_int),
        ("snapshot_precision", ctypes.c_int),
        ("M", ctypes.c_int), ("snap", ctypes.c_int),
    ]


_lib = None


def lib() -> ctypes.CDLL:
    """Load libactflow.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.af_status_str.restype = ctypes.c_char_p
        L.af_last_error.restype = ctypes.c_char_p
        L.af_run_create.argtypes = [ctypes.POINTER(AfParams), ctypes.POINTER(vp)]
        L.af_run_set_state.argtypes = [vp, dp, dp]
        L.af_run_get_state.argtypes = [vp, dp, dp]
        L.af_run_step.argtypes = [vp, ctypes.c_int, dp, ctypes.POINTER(ctypes.c_longlong)]
        L.af_run_steps_done.argtypes = [vp]
        L.af_run_steps_done.restype = ctypes.c_longlong
        L.af_run_destroy.argtypes = [vp]
        L.af_run_destroy.restype = None
        L.af_run_profile.argtypes = [vp, ctypes.c_int, dp, dp, dp]
        L.af_run_diagnostics.argtypes = [vp, dp, dp, dp, ctypes.c_int]
        L.af_run_diagnostics_noalloc.argtypes = [vp, dp, dp, dp, ctypes.c_int]
        L.af_engine_diagnostics.argtypes = [vp, ctypes.c_int, dp, dp, dp, ctypes.c_int]
        L.af_run_use_graph.argtypes = [vp, ctypes.c_int]
        L.af_run_device_bytes.argtypes = [ctypes.POINTER(AfParams)]
        L.af_run_device_bytes.restype = ctypes.c_size_t
        L.af_params_reference.argtypes = [ctypes.POINTER(AfParams), ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 6
        L.af_params_reference.restype = None
        L.af_params_jw.argtypes = [ctypes.POINTER(AfParams), ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 8
        L.af_params_jw.restype = None
        L.af_schedule_count.argtypes = [ctypes.c_int] * 3
        L.af_schedule_get.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(AfSchedule)]
        L.af_schedule_select.argtypes = [ctypes.c_int] * 4
        L.af_schedule_active.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(AfSchedule)]
        L.af_serve.argtypes = [ctypes.POINTER(AfRequest), ctypes.c_int, ctypes.c_size_t, ctypes.c_int, SNAPSHOT_FN,
                               vp, ctypes.POINTER(dp), ctypes.POINTER(AfRequestResult), ctypes.POINTER(AfServeStats)]
        L.af_engine_create.argtypes = [ctypes.c_size_t, ctypes.POINTER(vp)]
        L.af_engine_add.argtypes = [vp, ctypes.POINTER(AfParams), dp, dp, ctypes.POINTER(ctypes.c_int)]
        L.af_engine_remove.argtypes = [vp, ctypes.c_int]
        L.af_engine_step.argtypes = [vp, ctypes.c_int]
        L.af_engine_step_slots.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int]
        L.af_engine_set_single_launch.argtypes = [vp, ctypes.c_int]
        L.af_engine_single_launch_active.argtypes = [vp]
        L.af_engine_get_state.argtypes = [vp, ctypes.c_int, dp, dp]
        L.af_engine_get_energy.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.af_engine_steps_done.argtypes = [vp, ctypes.c_int]
        L.af_engine_steps_done.restype = ctypes.c_longlong
        L.af_engine_bytes_in_use.argtypes = [vp]
        L.af_engine_bytes_in_use.restype = ctypes.c_size_t
        L.af_engine_count.argtypes = [vp]
        L.af_engine_destroy.argtypes = [vp]
        L.af_engine_destroy.restype = None
        L.af_spec_defaults.argtypes = [ctypes.POINTER(AfSpec)]
        L.af_spec_defaults.restype = None
        L.af_spec_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(AfSpec), ctypes.c_char_p, ctypes.c_size_t]
        L.af_spec_write_example.argtypes = [ctypes.c_char_p, ctypes.POINTER(AfSpec)]
        L.af_snapshot_write.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int]
        L.af_snapshot_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, dp, dp]
        L.af_energy_append.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_double]
        L.af_ic_bandlimited.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_double,
                                        ctypes.c_double, dp]
        L.af_ic_random.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_int, dp]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(st: int):
    if st != AF_OK:
        L = lib()
        raise ActFlowError(st, f"{L.af_status_str(st).decode()}: {L.af_last_error().decode()}")


def params(nx, ny=None, lx=2 * np.pi, ly=None, dt=2e-4, lam=3.5, alpha=-0.8, beta=0.01, gamma0=None, gamma2=None,
           precision=64, dealias=0, mean_scheme=0) -> AfParams:
    """Run parameters.  With gamma0/gamma2 the James-Wilczek linear operator
    sigma(k) = -alpha - Gamma0 k^2 - Gamma2 k^4 is used, otherwise the
    reference's -(1-k^2)^2 - alpha (ActFlow_V3-2.cu:1030)."""
    ny = nx if ny is None else ny
    ly = lx if ly is None else ly
    p = AfParams()
    if gamma0 is None:
        lib().af_params_reference(ctypes.byref(p), nx, ny, lx, ly, dt, lam, alpha, beta)
    else:
        lib().af_params_jw(ctypes.byref(p), nx, ny, lx, ly, dt, lam, alpha, beta, gamma0, gamma2)
    p.precision, p.dealias, p.mean_scheme = precision, dealias, mean_scheme
    return p


def half_shape(p: AfParams):
    return (p.ny, p.nx // 2 + 1)


class Run:
    """One resident grid on the GPU (af_run)."""

    def __init__(self, p: AfParams):
        self.p = p
        self._h = ctypes.c_void_p()
        _check(lib().af_run_create(ctypes.byref(p), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().af_run_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, omega_hat: np.ndarray, M=(0.0, 0.0)):
        w = np.ascontiguousarray(omega_hat, dtype=np.complex128).reshape(half_shape(self.p))
        m = np.asarray(M, dtype=np.float64)
        _check(lib().af_run_set_state(self._h, _dp(w.view(np.float64)), _dp(m)))

    def get_state(self):
        w = np.zeros(half_shape(self.p), dtype=np.complex128)
        m = np.zeros(2)
        _check(lib().af_run_get_state(self._h, _dp(w.view(np.float64)), _dp(m)))
        return w, m

    def step(self, nsteps: int, energy: bool = False):
        e = np.zeros(max(nsteps, 1))
        nan = ctypes.c_longlong(-1)
        st = lib().af_run_step(self._h, nsteps, _dp(e) if energy else None, ctypes.byref(nan))
        if st == AF_ENAN:
            raise ActFlowError(st, f"NaN after step {nan.value}")
        _check(st)
        return e[:nsteps] if energy else None

    @property
    def steps_done(self) -> int:
        return lib().af_run_steps_done(self._h)

    def diagnostics(self, nshell: int = 0):
        """(energy 1/2<|u|^2>, enstrophy 1/2<w^2>, shell spectrum E(k) or None)."""
        e, z = ctypes.c_double(), ctypes.c_double()
        spec = np.zeros(max(nshell, 1))
        _check(lib().af_run_diagnostics(self._h, ctypes.byref(e), ctypes.byref(z),
                                        _dp(spec) if nshell else None, nshell))
        return e.value, z.value, (spec if nshell else None)

    def diagnostics_noalloc(self, nshell: int = 0):
        """As diagnostics(), through the allocation-free entry point (static
        per-device scratch, nshell <= 4096)."""
        e, z = ctypes.c_double(), ctypes.c_double()
        spec = np.zeros(max(nshell, 1))
        _check(lib().af_run_diagnostics_noalloc(self._h, ctypes.byref(e), ctypes.byref(z),
                                                _dp(spec) if nshell else None, nshell))
        return e.value, z.value, (spec if nshell else None)

    def use_graph(self, on: bool):
        _check(lib().af_run_use_graph(self._h, 1 if on else 0))

    def profile(self, nsteps: int):
        """Mean row-pass / column-pass kernel time (ms) over nsteps steps timed
        with HIP events on the run's stream, each after a 4-step graph (so the
        run advances 5 * nsteps steps), and the total wall of the region."""
        r, c, t = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().af_run_profile(self._h, nsteps, ctypes.byref(r), ctypes.byref(c), ctypes.byref(t)))
        return r.value, c.value, t.value


def schedules(precision, nx, ny=None):
    """The instantiated schedules of a size (index 0 = built-in default)."""
    ny = nx if ny is None else ny
    out = []
    for i in range(lib().af_schedule_count(precision, nx, ny)):
        sc = AfSchedule()
        _check(lib().af_schedule_get(precision, nx, ny, i, ctypes.byref(sc)))
        out.append(sc.as_tuple())
    return out


def select_schedule(precision, nx, ny, index):
    """Process-wide schedule for runs of this size created afterwards."""
    _check(lib().af_schedule_select(precision, nx, ny, index))


def active_schedule(precision, nx, ny=None):
    ny = nx if ny is None else ny
    sc = AfSchedule()
    _check(lib().af_schedule_active(precision, nx, ny, ctypes.byref(sc)))
    return sc.as_tuple()


def device_bytes(p: AfParams) -> int:
    return lib().af_run_device_bytes(ctypes.byref(p))


class Engine:
    """Ensemble engine (af_engine): many resident runs, batched launches."""

    def __init__(self, hbm_budget: int):
        self._h = ctypes.c_void_p()
        _check(lib().af_engine_create(hbm_budget, ctypes.byref(self._h)))
        self.params = {}

    def close(self):
        if self._h:
            lib().af_engine_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, p: AfParams, omega_hat: np.ndarray, M=(0.0, 0.0)) -> int:
        w = np.ascontiguousarray(omega_hat, dtype=np.complex128).reshape(half_shape(p))
        m = np.asarray(M, dtype=np.float64)
        slot = ctypes.c_int(-1)
        _check(lib().af_engine_add(self._h, ctypes.byref(p), _dp(w.view(np.float64)), _dp(m), ctypes.byref(slot)))
        self.params[slot.value] = p
        return slot.value

    def remove(self, slot: int):
        _check(lib().af_engine_remove(self._h, slot))
        self.params.pop(slot, None)

    def step(self, nsteps: int, slots=None):
        if slots is None:
            _check(lib().af_engine_step(self._h, nsteps))
        else:
            arr = (ctypes.c_int * len(slots))(*slots)
            _check(lib().af_engine_step_slots(self._h, arr, len(slots), nsteps))

    def set_single_launch(self, on: bool):
        """One launch per pass over all size classes when they form an
        instantiated mix (af_engine_set_single_launch)."""
        _check(lib().af_engine_set_single_launch(self._h, 1 if on else 0))

    @property
    def single_launch_active(self) -> bool:
        return bool(lib().af_engine_single_launch_active(self._h))

    def get_state(self, slot: int):
        p = self.params[slot]
        w = np.zeros(half_shape(p), dtype=np.complex128)
        m = np.zeros(2)
        _check(lib().af_engine_get_state(self._h, slot, _dp(w.view(np.float64)), _dp(m)))
        return w, m

    def energy(self, slot: int, n: int):
        e = np.zeros(n)
        _check(lib().af_engine_get_energy(self._h, slot, n, _dp(e)))
        return e

    def steps_done(self, slot: int) -> int:
        return lib().af_engine_steps_done(self._h, slot)

    def diagnostics(self, slot: int, nshell: int = 0):
        e, z = ctypes.c_double(), ctypes.c_double()
        spec = np.zeros(max(nshell, 1))
        _check(lib().af_engine_diagnostics(self._h, slot, ctypes.byref(e), ctypes.byref(z),
                                           _dp(spec) if nshell else None, nshell))
        return e.value, z.value, (spec if nshell else None)

    @property
    def bytes_in_use(self) -> int:
        return lib().af_engine_bytes_in_use(self._h)

    @property
    def count(self) -> int:
        return lib().af_engine_count(self._h)


# ---- serving ---------------------------------------------------------------

@dataclass
class Request:
    arrival_s: float
    params: AfParams
    horizon: int
    snapshot_every: int = 0
    seed: int = 1
    ic_amp: float = 1e-3
    ic_kmax: float = 10.0


def served_dt(n, lam, alpha, beta):
    """Time step for a request of the served workload (N^2 grid, box 2pi):
    1e-4 up to 1024^2 and 5e-5 at 2048^2, lowered further by an advective
    CFL estimate dt <= 2 / (lam u_max k_max), k_max = N/3 (2/3 mask) and
    u_max = 3 sqrt(-2 alpha / beta) (three times the saturated rms velocity,
    sqrt(-2 alpha/beta), which the 1024^2 runs with alpha = -0.8, beta = 0.01
    show: <E> ~ 81).  The coefficient 2 keeps the advective term inside the
    explicit RK4 stability interval on the imaginary axis (|z| <= 2.8) with a
    margin.  The 1024^2 run at dt = 1e-4, lam = 3.5 sits at 4.5 by this
    measure and is stable over 300k steps, but a 512^2 request at 2.9
    (lam 4.17, beta 0.0077, dt 1e-4) went NaN after 18.5k steps in FP32 while
    its FP64 run stayed finite: near the limit, whether a chaotic run survives
    depends on its rounding."""
    base = 1e-4 if n <= 1024 else 5e-5
    if lam <= 0 or alpha >= 0 or beta <= 0:
        return base
    umax = 3.0 * np.sqrt(-2.0 * alpha / beta)
    return float(min(base, 2.0 / (lam * umax * (n / 3.0))))


def serve(requests, hbm_budget, max_tick_steps=50, on_snapshot=None, keep_final=False):
    """Run requests through the continuous-batching loop (af_serve).
    on_snapshot(request_index, step, omega_hat, M) receives every snapshot.
    Returns (per-request result dicts, stats dict, final states or None)."""
    n = len(requests)
    arr = (AfRequest * max(n, 1))()
    for i, r in enumerate(requests):
        arr[i].arrival_s = r.arrival_s
        arr[i].params = r.params
        arr[i].horizon = r.horizon
        arr[i].snapshot_every = r.snapshot_every
        arr[i].seed = r.seed
        arr[i].ic_amp = r.ic_amp
        arr[i].ic_kmax = r.ic_kmax
    res = (AfRequestResult * max(n, 1))()
    stats = AfServeStats()
    finals = None
    fptr = None
    if keep_final:
        finals = [np.zeros((r.params.ny, r.params.nx // 2 + 1), dtype=np.complex128) for r in requests]
        fptr = (ctypes.POINTER(ctypes.c_double) * max(n, 1))(*[_dp(f.view(np.float64)) for f in finals])

    def _cb(user, req, step, w, M, nx, ny):
        if on_snapshot is not None:
            a = np.ctypeslib.as_array(w, shape=(ny * (nx // 2 + 1) * 2,)).copy()
            on_snapshot(req, step, a.view(np.complex128).reshape(ny, nx // 2 + 1), (M[0], M[1]))

    cb = SNAPSHOT_FN(_cb)
    _check(lib().af_serve(arr, n, hbm_budget, max_tick_steps, cb, None, fptr, res, ctypes.byref(stats)))
    out = [dict(admit_s=res[i].admit_s, finish_s=res[i].finish_s, step_p50_s=res[i].step_p50_s,
                step_p99_s=res[i].step_p99_s, snapshots=res[i].snapshots, status=res[i].status,
                M=(res[i].M[0], res[i].M[1])) for i in range(n)]
    st = {k: getattr(stats, k) for k, _ in AfServeStats._fields_}
    return out, st, finals


# ---- host formats --------------------------------------------------------

def ic_bandlimited(nx, ny=None, seed=1, amp=1e-3, kmax=10.0) -> np.ndarray:
    ny = nx if ny is None else ny
    w = np.zeros((ny, nx // 2 + 1), dtype=np.complex128)
    _check(lib().af_ic_bandlimited(nx, ny, seed, amp, kmax, _dp(w.view(np.float64))))
    return w


def ic_random(nx, ny=None, seed=42, dealias=0) -> np.ndarray:
    ny = nx if ny is None else ny
    w = np.zeros((ny, nx // 2 + 1), dtype=np.complex128)
    _check(lib().af_ic_random(nx, ny, seed, dealias, _dp(w.view(np.float64))))
    return w


def spec_defaults() -> AfSpec:
    s = AfSpec()
    lib().af_spec_defaults(ctypes.byref(s))
    return s


def spec_read(directory: str) -> AfSpec:
    s = spec_defaults()
    err = ctypes.create_string_buffer(2048)
    st = lib().af_spec_read(directory.encode(), ctypes.byref(s), err, 2048)
    if st != AF_OK:
        raise ActFlowError(st, err.value.decode())
    return s


def snapshot_write(path, omega_hat, M, precision=6):
    w = np.ascontiguousarray(omega_hat, dtype=np.complex128)
    ny, nxh = w.shape
    m = np.asarray(M, dtype=np.float64)
    _check(lib().af_snapshot_write(str(path).encode(), 2 * (nxh - 1), ny, _dp(w.view(np.float64)), _dp(m), precision))


def snapshot_read(path, nx, ny):
    w = np.zeros((ny, nx // 2 + 1), dtype=np.complex128)
    m = np.zeros(2)
    _check(lib().af_snapshot_read(str(path).encode(), nx, ny, _dp(w.view(np.float64)), _dp(m)))
    return w, m


def snapshot_image(path, nx, ny, spectral=False):
    """The array the reference's plotting script draws for one snapshot
    (image.py:49-69), without the plotting.

    spectral=False: the real-space vorticity, the inverse real DFT of the
    stored half spectrum normalised by 1/(Nx Ny) (image.py:67; the stored
    values are unnormalised DFT coefficients, so this is the physical field).

    spectral=True: |w^| on the full (kx, ky) plane, Ny x Nx.  Column c holds
    kx = c - s and row r the half-spectrum row (r - s) mod Ny, with
    s = Nx/2 - 1 for both axes (image.py:64 shifts the rows by Nx//2 - 1 also
    when Ny != Nx; kept for identical images); a column with kx < 0 shows the
    Hermitian partner |w^(-kx, -ky)| (image.py:55-62)."""
    w, _ = snapshot_read(path, nx, ny)
    if not spectral:
        return np.fft.irfft2(w, s=(ny, nx))
    shift = nx // 2 - 1
    kx = np.arange(nx) - shift                       # -(Nx/2 - 1) .. Nx/2
    row = (np.arange(ny) - shift) % ny               # half-spectrum row of each image row
    mag = np.abs(w)
    img = np.empty((ny, nx))
    pos = kx >= 0
    img[:, pos] = mag[row][:, kx[pos]]
    img[:, ~pos] = mag[(-row) % ny][:, -kx[~pos]]
    return img


def energy_append(path, t, E):
    _check(lib().af_energy_append(str(path).encode(), t, E))
