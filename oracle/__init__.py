"""CPU oracle for the zldsp::fft hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this package, and only as the checker
(or, for the baseline, as the CPU port being timed).  The product package
``zldsp_fft_amd`` never imports it and has no CPU fallback.

The numerical work happens in ``zlfft_oracle.c`` (a plain-C restatement of the
reference algorithm, built by ``oracle/Makefile`` into ``oracle/_build``); this
module is a thin ctypes wrapper that converts numpy arrays to and from the
AoS / SoA layouts of the reference API (reference
``src/common/zldsp_fft_common_structure.hpp:8-33``).

Every function here is pinned by ``tests/test_oracle_pinning.py`` against
numpy.fft (pocketfft, FP64), a brute-force long-double DFT, closed forms and
the committed fixtures in ``tests/golden``; planted-mistake builds of the same
C source (``liboracle_mut*.so``) must fail those checks.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")

_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)


def build(quiet: bool = True) -> None:
    """Compile the oracle (and its planted-mistake variants) with make."""
    out = subprocess.run(["make", "-C", _HERE, "-j4"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def _bind(lib: ctypes.CDLL) -> ctypes.CDLL:
    for sfx, P in (("f32", _F), ("f64", _D)):
        f = getattr(lib, f"zlo_cfft_{sfx}")
        f.argtypes = [ctypes.c_int] * 5 + [P, P, ctypes.c_int, P, P]
        f.restype = ctypes.c_int
        f = getattr(lib, f"zlo_rfft_{sfx}")
        f.argtypes = [ctypes.c_int] * 5 + [P, P, P, P]
        f.restype = ctypes.c_int
        f = getattr(lib, f"zlo_cfft_batch_{sfx}")
        f.argtypes = [ctypes.c_int, ctypes.c_size_t] + [ctypes.c_int] * 4 + [P, P, ctypes.c_int, P, P, ctypes.c_int]
        f.restype = ctypes.c_int
    lib.zlo_cospi_f64.argtypes = [ctypes.c_double]
    lib.zlo_cospi_f64.restype = ctypes.c_double
    lib.zlo_sinpi_f64.argtypes = [ctypes.c_double]
    lib.zlo_sinpi_f64.restype = ctypes.c_double
    lib.zlo_mutation.restype = ctypes.c_int
    lib.zlo_switch_order.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.zlo_switch_order.restype = None
    lib.zlo_switch_order_from.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.zlo_switch_order_from.restype = None
    return lib


_LIBS: dict[str, ctypes.CDLL] = {}


def load(variant: str = "") -> ctypes.CDLL:
    """Load ``liboracle{variant}.so`` (variant '' or '_mutN'), building if absent."""
    if variant not in _LIBS:
        path = os.path.join(_BUILD, f"liboracle{variant}.so")
        if not os.path.exists(path):
            build()
        _LIBS[variant] = _bind(ctypes.CDLL(path))
    return _LIBS[variant]


def switch_order(precision: str = "f32") -> tuple[int, int]:
    """(max_l1_order, switch_order) by the reference's host rule.

    Reference ``src/common/zldsp_fft_common_system.hpp:131-157`` with cache
    sizes from sysconf, computed in C (``zlo_switch_order``) exactly as there.
    """
    lib = load()
    a, b = ctypes.c_int(), ctypes.c_int()
    lib.zlo_switch_order(4 if precision == "f32" else 8, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def switch_order_from(precision: str, l1: int, l2: int, l3: int) -> tuple[int, int]:
    """The same rule for given cache sizes in bytes."""
    lib = load()
    a, b = ctypes.c_int(), ctypes.c_int()
    lib.zlo_switch_order_from(4 if precision == "f32" else 8, l1, l2, l3, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def _dt(precision: str):
    if precision not in ("f32", "f64"):
        raise ValueError(precision)
    return (np.float32, _F) if precision == "f32" else (np.float64, _D)


def _ptr(a: np.ndarray | None, P):
    return None if a is None else a.ctypes.data_as(P)


def _split(x: np.ndarray, layout: str, dt):
    """complex vector -> (in0, in1) in the requested user layout"""
    if layout == "aos":
        a = np.empty(2 * x.size, dtype=dt)
        a[0::2] = x.real
        a[1::2] = x.imag
        return a, None
    if layout == "soa":
        return np.ascontiguousarray(x.real, dtype=dt), np.ascontiguousarray(x.imag, dtype=dt)
    raise ValueError(layout)


def _alloc(n: int, layout: str, dt):
    if layout == "aos":
        return np.zeros(2 * n, dtype=dt), None
    return np.zeros(n, dtype=dt), np.zeros(n, dtype=dt)


def _join(a0, a1, layout: str) -> np.ndarray:
    if layout == "aos":
        return a0[0::2] + 1j * a0[1::2].astype(a0.dtype)
    return a0 + 1j * a1


def _lay(v: str) -> int:
    return 0 if v == "aos" else 1


def _plan_args(precision, l1_order, sw):
    d1, d2 = switch_order(precision)
    return (d1 if l1_order is None else l1_order), (d2 if sw is None else sw)


def cfft(x, forward: bool = True, precision: str = "f64", in_layout: str = "aos", out_layout: str = "aos",
         l1_order: int | None = None, switch_order_: int | None = None, variant: str = "") -> np.ndarray:
    """CFFT<F>::forward / backward on one transform (reference
    ``src/zldsp_fft_cfft.hpp:38-117``): unnormalised, backward = swap, forward
    DFT, swap.  Returns a complex numpy array (complex64 for f32)."""
    lib = load(variant)
    dt, P = _dt(precision)
    x = np.asarray(x).reshape(-1)
    n = x.size
    order = n.bit_length() - 1
    if n != 1 << order:
        raise ValueError("size must be a power of two")
    l1, sw = _plan_args(precision, l1_order, switch_order_)
    i0, i1 = _split(x, in_layout, dt)
    o0, o1 = _alloc(n, out_layout, dt)
    fn = getattr(lib, f"zlo_cfft_{precision}")
    rc = fn(order, int(forward), l1, sw, _lay(in_layout), _ptr(i0, P), _ptr(i1, P), _lay(out_layout),
            _ptr(o0, P), _ptr(o1, P))
    if rc != 0:
        raise RuntimeError(f"oracle cfft failed rc={rc}")
    return _join(o0, o1, out_layout)


def cfft_batch(x: np.ndarray, forward: bool = True, precision: str = "f32", nthreads: int = 1,
               l1_order: int | None = None, switch_order_: int | None = None) -> np.ndarray:
    """Batched AoS->AoS CFFT over the rows of ``x`` (batch, N), split across
    ``nthreads`` host threads, one workspace per thread (reference instances are
    not shareable, ``src/common/zldsp_fft_common_init.hpp:24-43``)."""
    lib = load()
    dt, P = _dt(precision)
    x = np.asarray(x)
    batch, n = x.shape
    order = n.bit_length() - 1
    l1, sw = _plan_args(precision, l1_order, switch_order_)
    a = np.empty((batch, 2 * n), dtype=dt)
    a[:, 0::2] = x.real
    a[:, 1::2] = x.imag
    o = np.empty_like(a)
    rc = getattr(lib, f"zlo_cfft_batch_{precision}")(order, batch, int(forward), l1, sw, 0, _ptr(a, P), None, 0,
                                                      _ptr(o, P), None, nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle batch failed rc={rc}")
    return o[:, 0::2] + 1j * o[:, 1::2].astype(dt)


def cfft_batch_raw(order: int, aos_in: np.ndarray, aos_out: np.ndarray, forward: bool, precision: str,
                   nthreads: int, l1_order: int | None = None, switch_order_: int | None = None) -> None:
    """Same as cfft_batch on preallocated interleaved buffers (no conversions),
    for timing the CPU baseline."""
    lib = load()
    _, P = _dt(precision)
    n = 1 << order
    batch = aos_in.size // (2 * n)
    l1, sw = _plan_args(precision, l1_order, switch_order_)
    rc = getattr(lib, f"zlo_cfft_batch_{precision}")(order, batch, int(forward), l1, sw, 0, _ptr(aos_in, P), None,
                                                      0, _ptr(aos_out, P), None, nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle batch failed rc={rc}")


def _rfft(mode: int, rorder: int, precision: str, layout: str, a0, a1, real, sq, l1_order, sw, variant):
    lib = load(variant)
    _, P = _dt(precision)
    l1, sw = _plan_args(precision, l1_order, sw)
    rc = getattr(lib, f"zlo_rfft_{precision}")(rorder, mode, l1, sw, _lay(layout), _ptr(a0, P), _ptr(a1, P),
                                                _ptr(real, P), _ptr(sq, P))
    if rc != 0:
        raise RuntimeError(f"oracle rfft failed rc={rc}")


def _rorder(n: int) -> int:
    order = n.bit_length() - 1
    if n != 1 << order or order < 1:
        raise ValueError("RFFT size must be a power of two >= 2")
    return order


def rfft_forward(x, precision: str = "f64", layout: str = "aos", l1_order=None, switch_order_=None,
                 variant: str = "") -> np.ndarray:
    """RFFT<F>::forward (reference ``src/zldsp_fft_rfft.hpp:68-79``): N reals ->
    N/2+1 bins, in AoS or SoA."""
    dt, _ = _dt(precision)
    x = np.ascontiguousarray(np.asarray(x).reshape(-1), dtype=dt)
    rorder = _rorder(x.size)
    nb = x.size // 2 + 1
    o0, o1 = _alloc(nb, layout, dt)
    _rfft(0, rorder, precision, layout, o0, o1, x, None, l1_order, switch_order_, variant)
    return _join(o0, o1, layout)


def rfft_sqr_mag(x, precision: str = "f64", l1_order=None, switch_order_=None, variant: str = "") -> np.ndarray:
    """RFFT<F>::forward_sqr_mag (reference ``src/zldsp_fft_rfft.hpp:104-109``)."""
    dt, _ = _dt(precision)
    x = np.ascontiguousarray(np.asarray(x).reshape(-1), dtype=dt)
    rorder = _rorder(x.size)
    sq = np.zeros(x.size // 2 + 1, dtype=dt)
    _rfft(1, rorder, precision, "aos", None, None, x, sq, l1_order, switch_order_, variant)
    return sq


def rfft_backward(X, n: int, precision: str = "f64", layout: str = "aos", l1_order=None, switch_order_=None,
                  variant: str = "") -> np.ndarray:
    """RFFT<F>::backward (reference ``src/zldsp_fft_rfft.hpp:86-97``): N/2+1 bins
    -> N reals, equal to (N/2) * irfft with Im X[0], Im X[N/2] ignored."""
    dt, _ = _dt(precision)
    rorder = _rorder(n)
    X = np.asarray(X).reshape(-1)
    if X.size != n // 2 + 1:
        raise ValueError("expected N/2+1 bins")
    a0, a1 = _split(X, layout, dt)
    real = np.zeros(n, dtype=dt)
    _rfft(2, rorder, precision, layout, a0, a1, real, None, l1_order, switch_order_, variant)
    return real


def cospi(x: float) -> float:
    return load().zlo_cospi_f64(float(x))


def sinpi(x: float) -> float:
    return load().zlo_sinpi_f64(float(x))
