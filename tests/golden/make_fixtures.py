"""Writes the golden fixtures under tests/golden/ (run in the build container).

Everything here is computed by numpy or by a plain restatement written in this
file, never by the product or the oracle:
  rfft2_cases.npz      seeded real fields and numpy.fft.rfft2 / irfft2 of them
                       (the unnormalised R2C/C2R convention of cufftPlan2d,
                       ActFlow_V3-2.cu:507-508, as read back by image.py:69)
  bandlimited_ic.npz   the band-limited seeded IC: SplitMix64 (Steele, Lea,
                       Flood 2014) + Box-Muller, visiting nx = 0..K, ny = -K..K
specifications.txt is the reference's example input, copied verbatim as data.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MASK64 = (1 << 64) - 1


def rfft_cases():
    rng = np.random.default_rng(20240611)
    out = {}
    for (ny, nx) in [(8, 8), (16, 32), (32, 16), (64, 64), (4, 8), (128, 32)]:
        x = rng.standard_normal((ny, nx))
        out[f"x_{ny}_{nx}"] = x
        out[f"X_{ny}_{nx}"] = np.fft.rfft2(x)
        # a non-Hermitian half spectrum: C2R semantics take Re of the full IDFT
        H = rng.standard_normal((ny, nx // 2 + 1)) + 1j * rng.standard_normal((ny, nx // 2 + 1))
        out[f"H_{ny}_{nx}"] = H
        out[f"h_{ny}_{nx}"] = np.fft.irfft2(H, s=(ny, nx)) * (nx * ny)
    np.savez(os.path.join(HERE, "rfft2_cases.npz"), **out)


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def unit(self):
        return (self.next() >> 11) * 2.0 ** -53


def bandlimited(seed, amp, kmax):
    g = SplitMix64(seed)
    K = int(np.floor(kmax))
    modes, vals = [], []
    for i in range(0, K + 1):
        for j in range(-K, K + 1):
            r2 = i * i + j * j
            if r2 == 0 or r2 > kmax * kmax or (i == 0 and j < 0):
                continue
            u1 = 1.0 - g.unit()
            u2 = g.unit()
            u3 = g.unit()
            gg = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
            th = 2.0 * np.pi * u3
            modes.append((i, j))
            vals.append(amp * gg * np.exp(1j * th))
    return np.array(modes, dtype=np.int64), np.array(vals)


def ic_cases():
    out = {}
    for seed in (1, 2, 12345):
        m, v = bandlimited(seed, 1e-3, 10.0)
        out[f"modes_{seed}"] = m
        out[f"vals_{seed}"] = v
    m, v = bandlimited(7, 2.5, 4.5)
    out["modes_7_big"] = m
    out["vals_7_big"] = v
    np.savez(os.path.join(HERE, "bandlimited_ic.npz"), **out)


if __name__ == "__main__":
    rfft_cases()
    ic_cases()
    print("fixtures written to", HERE)
