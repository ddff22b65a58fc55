"""Writes tests/golden/fixtures.npz: small input/output vectors for CFFT and
RFFT, computed by a brute-force DFT in long double (x87 80-bit), independent of
both numpy.fft and the oracle.

The reference (ZL-Audio/zldsp_fft) ships no test vectors, so these fixtures
are the mathematics of its contract (forward sign -, unnormalised; backward
unnormalised inverse; RFFT backward = (N/2) irfft ignoring Im X[0], Im X[N/2]),
on inputs rounded to float32 so that the FP32 GPU path sees the same values.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", ".."))
from zldsp_fft_amd import signals  # noqa: E402

LD = np.longdouble
CLD = np.clongdouble
PI_LD = LD("3.14159265358979323846264338327950288")


def dft_ld(x: np.ndarray, sign: int) -> np.ndarray:
    """sum_n x[n] exp(sign * 2 pi i n k / N), angles reduced exactly mod N."""
    n = x.size
    idx = np.arange(n, dtype=np.int64)
    e = np.outer(idx, idx) % n
    ang = LD(sign) * LD(2) * PI_LD * e.astype(LD) / LD(n)
    w = np.cos(ang) + 1j * np.sin(ang).astype(LD)
    xl = x.astype(CLD)
    return (w.astype(CLD) * xl[None, :]).sum(axis=1)


def main() -> None:
    out = {}
    for order in range(0, 11):
        n = 1 << order
        x = signals.complex_uniform(100 + order, 1, n)[0].astype(np.complex64).astype(np.complex128)
        out[f"cfft_x_{order}"] = x
        out[f"cfft_fwd_{order}"] = dft_ld(x, -1).astype(np.complex128)
        out[f"cfft_bwd_{order}"] = dft_ld(x, +1).astype(np.complex128)
    for order in range(1, 11):
        n = 1 << order
        m = n // 2
        x = signals.uniform(200 + order, n).astype(np.float32).astype(np.float64)
        X = dft_ld(x.astype(np.complex128), -1)[: m + 1]
        out[f"rfft_x_{order}"] = x
        out[f"rfft_fwd_{order}"] = X.astype(np.complex128)
        # backward input: random bins incl. imag at DC / Nyquist, which the contract ignores
        b = signals.complex_uniform(300 + order, 1, m + 1)[0].astype(np.complex64).astype(np.complex128)
        bl = b.astype(CLD)
        t = np.arange(n, dtype=np.int64)
        acc = np.full(n, LD(bl[0].real)) + LD(bl[m].real) * np.where(t % 2 == 0, LD(1), LD(-1))
        if m > 1:
            k = np.arange(1, m, dtype=np.int64)
            e = np.outer(t, k) % n
            ang = LD(2) * PI_LD * e.astype(LD) / LD(n)
            acc = acc + LD(2) * (np.cos(ang) * bl[1:m].real.astype(LD)[None, :]
                                 - np.sin(ang) * bl[1:m].imag.astype(LD)[None, :]).sum(axis=1)
        out[f"rfft_bins_{order}"] = b
        out[f"rfft_bwd_{order}"] = (acc / LD(2)).astype(np.float64)
    path = os.path.join(os.path.dirname(__file__), "fixtures.npz")
    np.savez_compressed(path, **out)
    print(path, sum(v.nbytes for v in out.values()), "bytes")


if __name__ == "__main__":
    main()
