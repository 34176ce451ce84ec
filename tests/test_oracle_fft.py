"""Pins the oracle's CPU FFT (the stand-in for cuFFT's unnormalised 2-D
R2C/C2R of ActFlow_V3-2.cu:507-508) against numpy golden vectors, brute-force
long-double DFTs and the transform identities."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def dft_brute(a, sign):
    n = len(a)
    k = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(k, k).astype(np.longdouble) / n)
    return W @ a


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 64])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1_bruteforce(orc, n, sign):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    got = orc.fft1(a, sign)
    ref = dft_brute(a, sign)
    assert np.abs(got - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max())


def test_fft1_float32(orc):
    rng = np.random.default_rng(3)
    a = rng.standard_normal(256) + 1j * rng.standard_normal(256)
    got = orc.fft1(a, -1, precision=32)
    ref = np.fft.fft(a)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-6


def test_rfft2_golden(orc):
    d = np.load(os.path.join(GOLDEN, "rfft2_cases.npz"))
    for key in d.files:
        if not key.startswith("x_"):
            continue
        ny, nx = map(int, key[2:].split("_"))
        x, X = d[key], d[f"X_{ny}_{nx}"]
        got = orc.rfft2(x)
        assert np.abs(got - X).max() <= 1e-13 * np.abs(X).max(), key
        back = orc.irfft2(X, nx)
        assert np.abs(back - nx * ny * x).max() <= 1e-12 * nx * ny, key


def test_irfft2_nonhermitian_golden(orc):
    """C2R of a non-Hermitian half spectrum equals numpy's (Re of the full
    inverse DFT: the imaginary parts of the kx=0 and kx=Nx/2 bins are dropped
    and the kx=0/Nyquist columns are Hermitian-projected)."""
    d = np.load(os.path.join(GOLDEN, "rfft2_cases.npz"))
    for key in d.files:
        if not key.startswith("H_"):
            continue
        ny, nx = map(int, key[2:].split("_"))
        got = orc.irfft2(d[key], nx)
        ref = d[f"h_{ny}_{nx}"]
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), key


def test_rfft2_bruteforce_small(orc):
    rng = np.random.default_rng(11)
    ny, nx = 8, 16
    x = rng.standard_normal((ny, nx)).astype(np.longdouble)
    ky = np.arange(ny)[:, None, None, None]
    kx = np.arange(nx // 2 + 1)[None, :, None, None]
    yy = np.arange(ny)[None, None, :, None]
    xx = np.arange(nx)[None, None, None, :]
    ph = np.exp(-2j * np.pi * (ky * yy / ny + kx * xx / nx))
    ref = np.sum(ph * x[None, None, :, :], axis=(2, 3))
    got = orc.rfft2(x.astype(np.float64))
    assert np.abs(got - ref.astype(np.complex128)).max() < 1e-13 * np.abs(ref).max()


def test_parseval_and_hermitian(orc):
    rng = np.random.default_rng(5)
    x = rng.standard_normal((32, 64))
    X = orc.rfft2(x)
    # Parseval: sum x^2 = (1/N) sum_full |X|^2 with half-spectrum weights
    wt = np.full(X.shape, 2.0)
    wt[:, 0] = 1
    wt[:, -1] = 1
    assert np.isclose(np.sum(x ** 2), np.sum(wt * np.abs(X) ** 2) / x.size, rtol=1e-13)
    # the kx=0 column of a real field's R2C is Hermitian in ky
    col = X[:, 0]
    assert np.allclose(col[1:], np.conj(col[1:][::-1]), rtol=0, atol=1e-12 * np.abs(col).max())


def test_shift_theorem(orc):
    rng = np.random.default_rng(9)
    x = rng.standard_normal((16, 16))
    X = orc.rfft2(x)
    Xs = orc.rfft2(np.roll(x, 3, axis=1))
    kx = np.arange(9)[None, :]
    assert np.allclose(Xs, X * np.exp(-2j * np.pi * 3 * kx / 16), atol=1e-12 * np.abs(X).max())
