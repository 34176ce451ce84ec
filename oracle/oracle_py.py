"""ctypes binding of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker; the product never imports it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: build with `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 8 + [ctypes.c_int] * 4
        L.oracle_destroy.argtypes = [vp]
        L.oracle_destroy.restype = None
        L.oracle_set_state.argtypes = [vp, dp, dp]
        L.oracle_set_state.restype = None
        L.oracle_get_state.argtypes = [vp, dp, dp]
        L.oracle_get_state.restype = None
        L.oracle_step.argtypes = [vp, ctypes.c_int, dp]
        L.oracle_rhs.argtypes = [vp, dp, dp, dp, dp, dp]
        L.oracle_rhs.restype = None
        L.oracle_mask.argtypes = [vp, ctypes.c_int]
        L.oracle_mask.restype = ctypes.c_double
        for f in (L.oracle_rfft2, L.oracle_irfft2):
            f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
            f.restype = None
        L.oracle_fft1.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        L.oracle_fft1.restype = None
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """CPU restatement of the reference step (see oracle.h)."""

    def __init__(self, nx, ny=None, lx=2 * np.pi, ly=None, dt=2e-4, lam=3.5, beta=0.01, c0=None, c2=2.0, c4=-1.0,
                 alpha=-0.8, precision=64, dealias=0, mean_scheme=0, nthreads=1):
        ny = nx if ny is None else ny
        ly = lx if ly is None else ly
        c0 = -1.0 - alpha if c0 is None else c0
        self.nx, self.ny = nx, ny
        self._h = lib().oracle_create(nx, ny, lx, ly, dt, lam, beta, c0, c2, c4, precision, dealias, mean_scheme,
                                      nthreads)
        if not self._h:
            raise ValueError("oracle_create rejected the parameters")

    @classmethod
    def from_params(cls, p, nthreads=1, dealias=None):
        return cls(p.nx, p.ny, p.lx, p.ly, p.dt, p.lambda_, p.beta, p.c0, p.c2, p.c4, precision=p.precision,
                   dealias=p.dealias if dealias is None else dealias, mean_scheme=p.mean_scheme, nthreads=nthreads)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    @property
    def shape(self):
        return (self.ny, self.nx // 2 + 1)

    def set_state(self, w, M=(0.0, 0.0)):
        w = np.ascontiguousarray(w, dtype=np.complex128).reshape(self.shape)
        m = np.asarray(M, dtype=np.float64)
        lib().oracle_set_state(self._h, _dp(w.view(np.float64)), _dp(m))

    def get_state(self):
        w = np.zeros(self.shape, dtype=np.complex128)
        m = np.zeros(2)
        lib().oracle_get_state(self._h, _dp(w.view(np.float64)), _dp(m))
        return w, m

    def step(self, n, energy=False):
        e = np.zeros(max(n, 1))
        bad = lib().oracle_step(self._h, n, _dp(e) if energy else None)
        if bad:
            raise FloatingPointError("oracle: NaN in w[0]")
        return e[:n] if energy else None

    def rhs(self, s, m=(0.0, 0.0)):
        s = np.ascontiguousarray(s, dtype=np.complex128).reshape(self.shape)
        mm = np.asarray(m, dtype=np.float64)
        out = np.zeros(self.shape, dtype=np.complex128)
        g = np.zeros(2)
        en = np.zeros(1)
        lib().oracle_rhs(self._h, _dp(s.view(np.float64)), _dp(mm), _dp(out.view(np.float64)), _dp(g), _dp(en))
        return out, g, en[0]

    def mask(self):
        n = self.shape[0] * self.shape[1]
        return np.array([lib().oracle_mask(self._h, c) for c in range(n)]).reshape(self.shape)


def rfft2(x, precision=64):
    x = np.ascontiguousarray(x, dtype=np.float64)
    ny, nx = x.shape
    out = np.zeros((ny, nx // 2 + 1), dtype=np.complex128)
    lib().oracle_rfft2(ny, nx, precision, _dp(x), _dp(out.view(np.float64)))
    return out


def irfft2(X, nx, precision=64):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    ny = X.shape[0]
    out = np.zeros((ny, nx))
    lib().oracle_irfft2(ny, nx, precision, _dp(X.view(np.float64)), _dp(out))
    return out


def fft1(a, sign=-1, precision=64):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    out = np.zeros_like(a)
    lib().oracle_fft1(len(a), sign, precision, _dp(a.view(np.float64)), _dp(out.view(np.float64)))
    return out
