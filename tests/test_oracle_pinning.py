"""Pins the CPU oracle (oracle/zlfft_oracle.c) before anything is checked
against it.

Each oracle function is checked against something independent of its own
formulas: numpy.fft (pocketfft, FP64), a long-double brute-force DFT (the
committed fixtures in tests/golden), closed forms (impulse, constant, single
exponential), invariants (Parseval, linearity, layout invariance, round trips)
and the reference's own scaling rules.  Planted-mistake builds of the same C
source (oracle/_build/liboracle_mut*.so: wrong twiddle sign, wrong twiddle
index, swapped butterfly outputs, a dropped RFFT twiddle term, a misplaced DIF
twiddle) must fail at least one of these checks, so the checks have teeth.
"""
import numpy as np
import pytest

import oracle
from zldsp_fft_amd import signals

rel = lambda a, b: float(np.linalg.norm(np.asarray(a, np.complex128) - b) / max(np.linalg.norm(b), 1e-300))
LAYOUTS = [("aos", "aos"), ("aos", "soa"), ("soa", "aos"), ("soa", "soa")]
# a forced macro split (reference init.hpp:316-338) so small orders take the DIF path
MACRO = dict(l1_order=7, switch_order_=9)


def _x(order, seed=1):
    return signals.complex_uniform(seed, 1, 1 << order)[0]


@pytest.mark.parametrize("order", range(0, 17))
def test_f64_forward_backward_vs_numpy(order):
    x = _x(order, order)
    n = 1 << order
    for lin, lout in LAYOUTS:
        assert rel(oracle.cfft(x, True, "f64", lin, lout), np.fft.fft(x)) < 1e-14
        assert rel(oracle.cfft(x, False, "f64", lin, lout), n * np.fft.ifft(x)) < 1e-14


@pytest.mark.parametrize("order", range(9, 17))
def test_f64_macro_path_vs_numpy(order):
    x = _x(order, 50 + order)
    n = 1 << order
    for lin, lout in LAYOUTS[::3]:
        assert rel(oracle.cfft(x, True, "f64", lin, lout, **MACRO), np.fft.fft(x)) < 1e-14
        assert rel(oracle.cfft(x, False, "f64", lin, lout, **MACRO), n * np.fft.ifft(x)) < 1e-14


def test_golden_cfft(fixtures):
    for order in range(0, 11):
        x = fixtures[f"cfft_x_{order}"]
        assert rel(oracle.cfft(x, True, "f64"), fixtures[f"cfft_fwd_{order}"]) < 1e-14
        assert rel(oracle.cfft(x, False, "f64"), fixtures[f"cfft_bwd_{order}"]) < 1e-14
        assert rel(oracle.cfft(x, True, "f32"), fixtures[f"cfft_fwd_{order}"]) < 5e-7
        if order >= 9:
            assert rel(oracle.cfft(x, True, "f64", **MACRO), fixtures[f"cfft_fwd_{order}"]) < 1e-14


def test_golden_rfft(fixtures):
    for order in range(1, 11):
        n = 1 << order
        x = fixtures[f"rfft_x_{order}"]
        for lay in ("aos", "soa"):
            assert rel(oracle.rfft_forward(x, "f64", lay), fixtures[f"rfft_fwd_{order}"]) < 1e-14
            y = oracle.rfft_backward(fixtures[f"rfft_bins_{order}"], n, "f64", lay)
            assert rel(y, fixtures[f"rfft_bwd_{order}"]) < 1e-14
        sq = oracle.rfft_sqr_mag(x, "f64")
        assert rel(sq, np.abs(fixtures[f"rfft_fwd_{order}"]) ** 2) < 1e-13


@pytest.mark.parametrize("order", [0, 1, 2, 3, 4, 5, 6, 7, 10, 11])
def test_closed_forms(order):
    """impulse -> exp(-2 pi i m k / N); constant -> N delta; exp(+2 pi i f n/N) -> N delta[k - f]."""
    n = 1 << order
    k = np.arange(n)
    for m in {v for v in (0, 1, n // 3, n - 1) if v < n}:
        d = np.zeros(n, complex)
        d[m] = 1.0
        assert np.allclose(oracle.cfft(d, True, "f64"), np.exp(-2j * np.pi * m * k / n), atol=1e-14)
        assert np.allclose(oracle.cfft(d, False, "f64"), np.exp(2j * np.pi * m * k / n), atol=1e-14)
    assert np.allclose(oracle.cfft(np.ones(n, complex), True, "f64"), n * (k == 0), atol=1e-12)
    f = max(n // 4, 0) if n > 1 else 0
    e = np.exp(2j * np.pi * f * np.arange(n) / n)
    assert np.allclose(oracle.cfft(e, True, "f64"), n * (k == f), atol=1e-13 * n)


def test_invariants():
    for order in (5, 8, 12):
        n = 1 << order
        x, y = _x(order, 3), _x(order, 4)
        X = oracle.cfft(x, True, "f64")
        assert abs(np.sum(abs(X) ** 2) - n * np.sum(abs(x) ** 2)) / np.sum(abs(X) ** 2) < 1e-13  # Parseval
        lin = oracle.cfft(2.5 * x - 1j * y, True, "f64")
        assert rel(lin, 2.5 * X - 1j * oracle.cfft(y, True, "f64")) < 1e-14  # linearity
        assert rel(oracle.cfft(X, False, "f64"), n * x) < 1e-14  # round trip N x


@pytest.mark.parametrize("order", range(1, 15))
def test_rfft_vs_numpy(order):
    n = 1 << order
    x = signals.uniform(order, n)
    X = np.fft.rfft(x)
    for lay in ("aos", "soa"):
        Y = oracle.rfft_forward(x, "f64", lay)
        assert rel(Y, X) < 1e-14
        assert Y[0].imag == 0.0 and Y[-1].imag == 0.0
    assert rel(oracle.rfft_sqr_mag(x, "f64"), np.abs(X) ** 2) < 1e-13
    # backward = (N/2) irfft; the imaginary parts of bins 0 and N/2 are ignored
    B = signals.complex_uniform(order + 99, 1, n // 2 + 1)[0]
    want = (n / 2) * np.fft.irfft(B, n)
    for lay in ("aos", "soa"):
        assert rel(oracle.rfft_backward(B, n, "f64", lay), want) < 1e-13
    # round trip (N/2) x (reference H1)
    assert rel(oracle.rfft_backward(X, n, "f64"), (n / 2) * x) < 1e-13


def test_rfft_macro_path():
    for order in range(11, 15):
        n = 1 << order
        x = signals.uniform(order + 7, n)
        assert rel(oracle.rfft_forward(x, "f64", **MACRO), np.fft.rfft(x)) < 1e-14


@pytest.mark.parametrize("order", range(0, 17))
def test_f32_restatement_accuracy(order):
    """The FP32 restatement (reference CFFT<float>) stays within rel-L2 5e-7 of FP64."""
    x = _x(order, 11 + order)
    n = 1 << order
    assert rel(oracle.cfft(x, True, "f32"), np.fft.fft(x)) < 5e-7
    assert rel(oracle.cfft(x, False, "f32", "soa", "soa"), n * np.fft.ifft(x)) < 5e-7
    if order >= 1:
        xr = x.real
        assert rel(oracle.rfft_forward(xr, "f32"), np.fft.rfft(xr)) < 5e-7


def test_cospi_and_twiddle_convention():
    """cospi(x) = cos(x pi) in double (reference math.hpp:19-25)."""
    assert oracle.cospi(0.0) == 1.0 and oracle.sinpi(0.0) == 0.0
    assert oracle.cospi(0.5) == np.cos(0.5 * np.pi)  # not exactly 0: the reference's quirk
    assert abs(oracle.sinpi(-0.25) + np.sqrt(0.5)) <= 2.3e-16  # cos/sin of x*pi: within 1 ulp


def test_switch_order_rule():
    """Host rule of reference system.hpp:131-157 on the cache sizes quoted for
    a Xeon with L1d 48 KiB / L2 2 MiB / L3 105 MiB and an EPYC-like host."""
    assert oracle.switch_order_from("f32", 48 << 10, 2 << 20, 110100480) == (10, 21)
    assert oracle.switch_order_from("f64", 48 << 10, 2 << 20, 110100480) == (9, 20)
    assert oracle.switch_order_from("f32", 48 << 10, 1 << 20, 32 << 20) == (10, 14)
    assert oracle.switch_order_from("f32", 32768, 262144, 0) == (10, 14)


def _pinning_checks(variant: str) -> list[str]:
    """The checks above in compact form; returns the names of failed ones."""
    failed = []
    for order in (4, 5, 6, 7, 10):
        x = _x(order, 5)
        if rel(oracle.cfft(x, True, "f64", variant=variant), np.fft.fft(x)) > 1e-12:
            failed.append(f"cfft{order}")
    for order in (11, 12):
        x = _x(order, 6)
        if rel(oracle.cfft(x, True, "f64", variant=variant, **MACRO), np.fft.fft(x)) > 1e-12:
            failed.append(f"macro{order}")
    for order in (3, 8):
        n = 1 << order
        x = signals.uniform(order, n)
        if rel(oracle.rfft_forward(x, "f64", variant=variant), np.fft.rfft(x)) > 1e-12:
            failed.append(f"rfft{order}")
        B = np.fft.rfft(x)
        if rel(oracle.rfft_backward(B, n, "f64", variant=variant), (n / 2) * x) > 1e-12:
            failed.append(f"irfft{order}")
    return failed


def test_pinning_checks_pass_on_oracle():
    assert _pinning_checks("") == []


@pytest.mark.parametrize("mut", [1, 2, 3, 4, 5])
def test_planted_mistakes_are_caught(mut):
    lib = oracle.load(f"_mut{mut}")
    assert lib.zlo_mutation() == mut
    assert _pinning_checks(f"_mut{mut}"), f"planted mistake {mut} went unnoticed"


def test_batch_threads_match_single():
    order, batch = 9, 13
    x = signals.complex_uniform(21, batch, 1 << order)
    y1 = oracle.cfft_batch(x, True, "f32", nthreads=1)
    y4 = oracle.cfft_batch(x, True, "f32", nthreads=4)
    assert np.array_equal(y1, y4)
    assert np.array_equal(y1[5], oracle.cfft(x[5], True, "f32"))
