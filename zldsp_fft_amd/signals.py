"""Seeded synthetic inputs for the FFT workloads (tests and bench share them).

SplitMix64 streams (one 64-bit draw per value) so that any language can
reproduce them; uniforms use the top 53 bits, normals use Box-Muller.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Draws offset .. offset+count-1 of the SplitMix64 stream of `seed`."""
    i = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, count: int, lo: float = -1.0, hi: float = 1.0, offset: int = 0) -> np.ndarray:
    """float64 uniform in [lo, hi)."""
    u = (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def normal(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """float64 standard normals (Box-Muller on consecutive uniform pairs)."""
    m = (count + 1) // 2
    u = uniform(seed, 2 * m, 0.0, 1.0, 2 * offset)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2 * np.pi * u2)
    z[1::2] = r * np.sin(2 * np.pi * u2)
    return z[:count]


def complex_uniform(seed: int, batch: int, n: int) -> np.ndarray:
    """complex128 (batch, n), re and im i.i.d. U[-1, 1)."""
    v = uniform(seed, 2 * batch * n).reshape(batch, n, 2)
    return v[..., 0] + 1j * v[..., 1]


def complex_normal(seed: int, batch: int, n: int) -> np.ndarray:
    """complex128 (batch, n), re and im i.i.d. N(0, 1)."""
    v = normal(seed, 2 * batch * n).reshape(batch, n, 2)
    return v[..., 0] + 1j * v[..., 1]


def hann(n: int) -> np.ndarray:
    """Periodic Hann window w[n] = 0.5 - 0.5 cos(2 pi n / N)."""
    return 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(n) / n)


def analyzer_frames(seed: int, frames: int, n: int, sigma: float = 0.1) -> np.ndarray:
    """float64 (frames, n): Hann-windowed N(0, sigma^2) noise plus three
    sinusoids per frame with frequency ~ U[0, N/2) bins, amplitude ~ U[0.1, 1]
    and phase ~ U[0, 2 pi)."""
    g = sigma * normal(seed, frames * n).reshape(frames, n)
    p = uniform(seed + 1000003, frames * 9, 0.0, 1.0).reshape(frames, 3, 3)
    f = p[..., 0] * (n / 2)
    a = 0.1 + 0.9 * p[..., 1]
    ph = 2 * np.pi * p[..., 2]
    t = np.arange(n)
    s = np.zeros((frames, n))
    for j in range(3):
        s += a[:, j, None] * np.sin(2 * np.pi * f[:, j, None] * t[None, :] / n + ph[:, j, None])
    return hann(n)[None, :] * (g + s)


def dynamic_range_set(seed: int, batch: int, n: int) -> np.ndarray:
    """complex128 (batch, n): a unit complex exponential, a -60 dB one and
    -100 dB complex white noise (sigma = 1e-5 per component)."""
    p = uniform(seed, batch * 2, 0.0, 1.0).reshape(batch, 2)
    t = np.arange(n)
    f1 = p[:, 0:1] * n
    f2 = p[:, 1:2] * n
    noise = 1e-5 * complex_normal(seed + 7, batch, n)
    return np.exp(2j * np.pi * f1 * t / n) + 1e-3 * np.exp(2j * np.pi * f2 * t / n) + noise


def rel_l2(y: np.ndarray, ref: np.ndarray, axis=-1) -> np.ndarray:
    """||y - ref|| / ||ref|| along `axis` (per transform)."""
    num = np.linalg.norm(np.asarray(y, dtype=np.complex128) - ref, axis=axis)
    den = np.linalg.norm(ref, axis=axis)
    return num / np.where(den == 0, 1.0, den)
