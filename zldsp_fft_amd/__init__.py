"""zldsp_fft_amd: the zldsp::fft CFFT / RFFT API on MI355X (gfx950).

Python mirror of the reference's two classes (reference
``src/zldsp_fft_cfft.hpp:11-118`` and ``src/zldsp_fft_rfft.hpp:11-139``) over the
C-ABI in ``include/zlfft.h`` (``libzlfft.so``, built in-tree by
``zldsp_fft_amd/build.py``).  The same names, argument meaning and overloads:

* ``CFFT(order)``: ``get_size()``, ``get_order()``, ``forward(in, out)``,
  ``backward(in, out)``, where each buffer is AoS (a complex64 array / tensor,
  the reference's ``std::complex<float>*``) or SoA (a 2-sequence
  ``(real, imag)`` of float32 arrays, the reference's ``std::array<F*, 2>``).
* ``RFFT(order)``: ``forward(real, out)``, ``backward(in, real)``,
  ``forward_sqr_mag(real, out)``, ``get_size()``, ``get_order()``.

Host buffers (numpy arrays or CPU tensors, pinned or not) may hold one
transform or a batch back to back, and the call is synchronous, like the
reference's; batches run through the library's pipelined host path.  CUDA
(HIP) tensors may hold a batch of transforms back to back (``numel`` a
multiple of the per-transform size); the whole batch runs as one launch on
torch's current stream.

There is no CPU fallback: if ``libzlfft.so`` is missing or no GPU is present
the calls raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import Sequence

import numpy as np

__all__ = ["CFFT", "RFFT", "ZlfftError", "lib_path", "load_library", "AOS", "SOA"]

_HERE = os.path.dirname(os.path.abspath(__file__))
# the in-package build; ZLFFT_LIB names another build of the same library
lib_path = os.environ.get("ZLFFT_LIB") or os.path.join(_HERE, "libzlfft.so")

# constants of include/zlfft.h
OK = 0
KIND_CFFT, KIND_RFFT = 0, 1
PREC_F32 = 0
PREC_F16_MFMA, PREC_F16_MFMA_SPLIT = 1, 2
PREC_BF16_MFMA, PREC_BF16_MFMA_SPLIT, PREC_BF16_MFMA_SPLIT3 = 3, 4, 5
PREC_F64 = 6
AOS, SOA = 0, 1
FORWARD, BACKWARD = 0, 1
MAX_CFFT_ORDER = 28
MAX_CFFT_ORDER_F64 = 27
MAX_LDS_ORDER = 14

_ERRNAMES = {1: "INVALID_ORDER", 2: "NULL", 3: "ALLOC", 4: "HIP", 5: "INVALID_ARG", 6: "UNSUPPORTED"}

EXPORTS = (
    "zlfft_plan_create", "zlfft_plan_destroy", "zlfft_plan_get_size", "zlfft_plan_get_order",
    "zlfft_cfft_exec", "zlfft_rfft_forward_exec", "zlfft_rfft_forward_sqr_mag_exec", "zlfft_rfft_backward_exec",
    "zlfft_cfft_execute", "zlfft_rfft_forward", "zlfft_rfft_forward_sqr_mag", "zlfft_rfft_backward",
    "zlfft_plan_copy_rfft_twiddles", "zlfft_last_error", "zlfft_version",
    "zlfft_ragged_create", "zlfft_ragged_create_config", "zlfft_ragged_exec", "zlfft_ragged_destroy",
    "zlfft_rfft_stft_exec",
    "zlfft_cfft_execute_batch", "zlfft_rfft_forward_batch", "zlfft_rfft_forward_sqr_mag_batch",
    "zlfft_rfft_backward_batch", "zlfft_cfft_exec_noalloc",
)
STFT_BINS, STFT_POWER, STFT_DB = 0, 1, 2
FORWARD_SQR_MAG = 2


class RaggedDesc(ctypes.Structure):
    """zlfft_ragged_desc of include/zlfft.h"""
    _fields_ = [("kind", ctypes.c_int), ("order", ctypes.c_int), ("direction", ctypes.c_int),
                ("layout", ctypes.c_int), ("in0", ctypes.c_void_p), ("in1", ctypes.c_void_p),
                ("out0", ctypes.c_void_p), ("out1", ctypes.c_void_p)]


# numpy image of zlfft_ragged_desc (48 bytes), for building large tables without
# one Python object per transform
RAGGED_DESC_DTYPE = np.dtype([("kind", "<i4"), ("order", "<i4"), ("direction", "<i4"), ("layout", "<i4"),
                              ("in0", "<u8"), ("in1", "<u8"), ("out0", "<u8"), ("out1", "<u8")])
assert RAGGED_DESC_DTYPE.itemsize == ctypes.sizeof(RaggedDesc)


class ZlfftError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"zlfft error {code} ({_ERRNAMES.get(code, '?')}): {msg}")
        self.code = code


_LIB: ctypes.CDLL | None = None


def load_library() -> ctypes.CDLL:
    """Load libzlfft.so and declare its signatures (no GPU work happens here)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(lib_path):
        raise ZlfftError(6, f"{lib_path} is not built; run python -m zldsp_fft_amd.build")
    L = ctypes.CDLL(lib_path)
    vp, sz, ci = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
    L.zlfft_plan_create.argtypes = [ctypes.POINTER(vp), ci, ci, sz]
    L.zlfft_plan_destroy.argtypes = [vp]
    L.zlfft_plan_get_size.argtypes = [vp]
    L.zlfft_plan_get_size.restype = sz
    L.zlfft_plan_get_order.argtypes = [vp]
    L.zlfft_plan_get_order.restype = sz
    L.zlfft_cfft_exec.argtypes = [vp, ci, ci, vp, vp, ci, vp, vp, sz, vp]
    L.zlfft_cfft_exec_noalloc.argtypes = [vp, ci, ci, vp, vp, ci, vp, vp, sz, vp]
    L.zlfft_rfft_forward_exec.argtypes = [vp, vp, ci, vp, vp, sz, vp]
    L.zlfft_rfft_forward_sqr_mag_exec.argtypes = [vp, vp, vp, sz, vp]
    L.zlfft_rfft_backward_exec.argtypes = [vp, ci, vp, vp, vp, sz, vp]
    L.zlfft_cfft_execute.argtypes = [vp, ci, ci, vp, vp, ci, vp, vp]
    L.zlfft_rfft_forward.argtypes = [vp, vp, ci, vp, vp]
    L.zlfft_rfft_forward_sqr_mag.argtypes = [vp, vp, vp]
    L.zlfft_rfft_backward.argtypes = [vp, ci, vp, vp, vp]
    L.zlfft_plan_copy_rfft_twiddles.argtypes = [vp, vp, sz]
    L.zlfft_last_error.restype = ctypes.c_char_p
    L.zlfft_ragged_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(RaggedDesc), sz]
    L.zlfft_ragged_create_config.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(RaggedDesc), sz,
                                             ctypes.POINTER(ctypes.c_int)]
    L.zlfft_ragged_exec.argtypes = [vp, vp]
    L.zlfft_ragged_destroy.argtypes = [vp]
    L.zlfft_cfft_execute_batch.argtypes = [vp, ci, ci, vp, vp, ci, vp, vp, sz]
    L.zlfft_rfft_forward_batch.argtypes = [vp, vp, ci, vp, vp, sz]
    L.zlfft_rfft_forward_sqr_mag_batch.argtypes = [vp, vp, vp, sz]
    L.zlfft_rfft_backward_batch.argtypes = [vp, ci, vp, vp, vp, sz]
    L.zlfft_rfft_stft_exec.argtypes = [vp, vp, sz, vp, sz, ci, ci, vp, vp, ctypes.c_float, vp]
    for name in EXPORTS:
        fn = getattr(L, name)
        if name not in ("zlfft_plan_get_size", "zlfft_plan_get_order", "zlfft_last_error"):
            fn.restype = ci
    _LIB = L
    return L


def _check(rc: int) -> None:
    if rc != OK:
        msg = load_library().zlfft_last_error().decode(errors="replace")
        raise ZlfftError(rc, msg)


# ---------------------------------------------------------------- buffers

def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _is_dev(x) -> bool:
    return _is_torch(x) and x.is_cuda


def _contig(x, name: str):
    if _is_torch(x):
        if not x.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
    elif not (isinstance(x, np.ndarray) and x.flags.c_contiguous):
        raise ValueError(f"{name} must be a C-contiguous numpy array or tensor")
    return x


def _ptr(x) -> int:
    return x.data_ptr() if _is_torch(x) else x.ctypes.data


def _nbytes(x) -> int:
    return x.numel() * x.element_size() if _is_torch(x) else x.nbytes


def _is_soa(buf) -> bool:
    return isinstance(buf, (tuple, list))


def _dtype_name(buf) -> str:
    return str(buf.dtype).replace("torch.", "")


def _complex_view(buf, name: str, f64: bool = False):
    """AoS buffer: complex64 or float32 with interleaved re, im (FP64 plans:
    complex128 or interleaved float64)."""
    _contig(buf, name)
    ok = ("complex128", "float64") if f64 else ("complex64", "float32")
    if _dtype_name(buf) not in ok:
        raise TypeError(f"{name}: AoS buffers are {ok[0]} (or interleaved {ok[1]})")
    return buf


def _float_view(buf, name: str, f64: bool = False):
    _contig(buf, name)
    want = "float64" if f64 else "float32"
    if _dtype_name(buf) != want:
        raise TypeError(f"{name}: {want} required")
    return buf


class _Side:
    """One side (input or output) of a call: layout, pointers, element count."""

    def __init__(self, buf, name: str, f64: bool = False):
        ss = 8 if f64 else 4
        if _is_soa(buf):
            if len(buf) != 2:
                raise ValueError(f"{name}: SoA buffers are (real, imag)")
            re, im = _float_view(buf[0], name + "[0]", f64), _float_view(buf[1], name + "[1]", f64)
            if _nbytes(re) != _nbytes(im):
                raise ValueError(f"{name}: real and imag sizes differ")
            self.layout, self.p0, self.p1 = SOA, _ptr(re), _ptr(im)
            self.ncomplex = _nbytes(re) // ss
            self.dev = _is_dev(re) and _is_dev(im)
            self.host = not (_is_dev(re) or _is_dev(im))
            self.ref = (re, im)
        else:
            b = _complex_view(buf, name, f64)
            self.layout, self.p0, self.p1 = AOS, _ptr(b), None
            self.ncomplex = _nbytes(b) // (2 * ss)
            self.dev = _is_dev(b)
            self.host = not self.dev
            self.ref = (b,)


def _stream_of(*tensors) -> int:
    import torch
    dev = next(t for t in tensors if _is_dev(t)).device
    return torch.cuda.current_stream(dev).cuda_stream


def _after_torch(*bufs) -> None:
    """Before a synchronous call that mixes host and device buffers: wait for
    torch's current stream on every device a device buffer lives on, so a
    kernel still producing an input (or still reading an output) on that
    stream has finished.  (The library itself waits for the legacy default
    stream; torch's current stream may be another one.)"""
    import torch
    devs = {b.device for b in bufs if _is_dev(b)}
    for d in devs:
        torch.cuda.current_stream(d).synchronize()


def _batch(n_have: int, n_per: int, name: str) -> int:
    if n_per == 0 or n_have % n_per != 0 or n_have == 0:
        raise ValueError(f"{name}: holds {n_have} elements, not a positive multiple of {n_per}")
    return n_have // n_per


# ---------------------------------------------------------------- plans

class _Plan:
    def __init__(self, kind: int, order: int, precision: int = PREC_F32):
        if isinstance(order, bool) or not isinstance(order, (int, np.integer)) or order < 0:
            raise ValueError("order must be a non-negative integer")
        self._lib = load_library()
        h = ctypes.c_void_p()
        _check(self._lib.zlfft_plan_create(ctypes.byref(h), kind, int(precision), int(order)))
        self._h = h
        self.precision = int(precision)
        self._f64 = self.precision == PREC_F64
        self._ss = 8 if self._f64 else 4

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.zlfft_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def get_size(self) -> int:
        return int(self._lib.zlfft_plan_get_size(self._h))

    def get_order(self) -> int:
        return int(self._lib.zlfft_plan_get_order(self._h))

    @property
    def handle(self) -> int:
        return self._h.value


class CFFT(_Plan):
    """``zldsp::fft::CFFT<float>`` (reference ``src/zldsp_fft_cfft.hpp:11-118``).

    ``precision`` selects the arithmetic: PREC_F32 (the plain path), PREC_F64
    (``CFFT<double>``: complex128 / float64 buffers) or one of the MFMA split
    precisions (4096-point AoS transforms only)."""

    def __init__(self, order: int, precision: int = PREC_F32):
        super().__init__(KIND_CFFT, order, precision)

    def _run(self, direction: int, in_buffer, out_buffer) -> None:
        n = self.get_size()
        si, so = _Side(in_buffer, "in_buffer", self._f64), _Side(out_buffer, "out_buffer", self._f64)
        if si.dev and so.dev:
            b = _batch(si.ncomplex, n, "in_buffer")
            if _batch(so.ncomplex, n, "out_buffer") != b:
                raise ValueError("in_buffer and out_buffer hold different batch sizes")
            _check(self._lib.zlfft_cfft_exec(self._h, direction, si.layout, si.p0, si.p1, so.layout, so.p0, so.p1,
                                             b, _stream_of(*si.ref, *so.ref)))
            return
        b = _batch(si.ncomplex, n, "in_buffer")
        if _batch(so.ncomplex, n, "out_buffer") != b:
            raise ValueError("in_buffer and out_buffer hold different batch sizes")
        _after_torch(*si.ref, *so.ref)
        _check(self._lib.zlfft_cfft_execute_batch(self._h, direction, si.layout, si.p0, si.p1, so.layout, so.p0, so.p1,
                                                  b))

    def forward(self, in_buffer, out_buffer) -> None:
        """AoS/SoA -> AoS/SoA forward (reference ``src/zldsp_fft_cfft.hpp:38-95``)."""
        self._run(FORWARD, in_buffer, out_buffer)

    def backward(self, in_buffer, out_buffer) -> None:
        """AoS/SoA -> AoS/SoA backward (reference ``src/zldsp_fft_cfft.hpp:47-103``)."""
        self._run(BACKWARD, in_buffer, out_buffer)


class RFFT(_Plan):
    """``zldsp::fft::RFFT<float>`` (reference ``src/zldsp_fft_rfft.hpp:11-139``);
    ``precision=PREC_F64`` gives ``RFFT<double>`` (float64 / complex128 buffers)."""

    def __init__(self, order: int, precision: int = PREC_F32):
        super().__init__(KIND_RFFT, order, precision)

    def forward(self, in_buffer, out_buffer) -> None:
        """N reals -> N/2+1 bins, AoS or SoA (reference ``src/zldsp_fft_rfft.hpp:68-79``)."""
        n = self.get_size()
        nb = n // 2 + 1
        x = _float_view(in_buffer, "in_buffer", self._f64)
        so = _Side(out_buffer, "out_buffer", self._f64)
        if _is_dev(x) and so.dev:
            b = _batch(x.numel(), n, "in_buffer")
            if _batch(so.ncomplex, nb, "out_buffer") != b:
                raise ValueError("in_buffer and out_buffer hold different batch sizes")
            _check(self._lib.zlfft_rfft_forward_exec(self._h, _ptr(x), so.layout, so.p0, so.p1, b,
                                                     _stream_of(x, *so.ref)))
            return
        b = _batch(_nbytes(x) // self._ss, n, "in_buffer")
        if _batch(so.ncomplex, nb, "out_buffer") != b:
            raise ValueError("in_buffer and out_buffer hold different batch sizes")
        _after_torch(x, *so.ref)
        _check(self._lib.zlfft_rfft_forward_batch(self._h, _ptr(x), so.layout, so.p0, so.p1, b))

    def backward(self, in_buffer, out_buffer) -> None:
        """N/2+1 bins (AoS or SoA) -> N reals (reference ``src/zldsp_fft_rfft.hpp:86-97``)."""
        n = self.get_size()
        nb = n // 2 + 1
        si = _Side(in_buffer, "in_buffer", self._f64)
        y = _float_view(out_buffer, "out_buffer", self._f64)
        if si.dev and _is_dev(y):
            b = _batch(si.ncomplex, nb, "in_buffer")
            if _batch(y.numel(), n, "out_buffer") != b:
                raise ValueError("in_buffer and out_buffer hold different batch sizes")
            _check(self._lib.zlfft_rfft_backward_exec(self._h, si.layout, si.p0, si.p1, _ptr(y), b,
                                                      _stream_of(*si.ref, y)))
            return
        b = _batch(si.ncomplex, nb, "in_buffer")
        if _batch(_nbytes(y) // self._ss, n, "out_buffer") != b:
            raise ValueError("in_buffer and out_buffer hold different batch sizes")
        _after_torch(*si.ref, y)
        _check(self._lib.zlfft_rfft_backward_batch(self._h, si.layout, si.p0, si.p1, _ptr(y), b))

    def forward_sqr_mag(self, in_buffer, out_buffer) -> None:
        """N reals -> N/2+1 squared magnitudes (reference ``src/zldsp_fft_rfft.hpp:104-109``)."""
        n = self.get_size()
        nb = n // 2 + 1
        x = _float_view(in_buffer, "in_buffer", self._f64)
        y = _float_view(out_buffer, "out_buffer", self._f64)
        if _is_dev(x) and _is_dev(y):
            b = _batch(x.numel(), n, "in_buffer")
            if _batch(y.numel(), nb, "out_buffer") != b:
                raise ValueError("in_buffer and out_buffer hold different batch sizes")
            _check(self._lib.zlfft_rfft_forward_sqr_mag_exec(self._h, _ptr(x), _ptr(y), b, _stream_of(x, y)))
            return
        b = _batch(_nbytes(x) // self._ss, n, "in_buffer")
        if _batch(_nbytes(y) // self._ss, nb, "out_buffer") != b:
            raise ValueError("in_buffer and out_buffer hold different batch sizes")
        _after_torch(x, y)
        _check(self._lib.zlfft_rfft_forward_sqr_mag_batch(self._h, _ptr(x), _ptr(y), b))

    def stft(self, signal, hop: int, out_buffer, window=None, output: int = STFT_BINS, db_floor: float = 1e-20,
             nframes: int | None = None) -> int:
        """STFT front-end (``zlfft_rfft_stft_exec``): frame f is
        ``signal[f * hop : f * hop + N] * window`` (device tensors), transformed
        to N/2+1 bins (AoS tensor or (real, imag) pair), squared magnitudes or
        dB levels per frame, back to back in ``out_buffer``.  ``nframes``
        defaults to every whole frame the signal holds.  Returns nframes."""
        n = self.get_size()
        nb = n // 2 + 1
        x = _float_view(signal, "signal", self._f64)
        if not _is_dev(x):
            raise ValueError("signal must be a device tensor")
        length = x.numel()
        if hop < 0:
            raise ValueError("hop must be >= 0")
        if nframes is None:
            if length < n:
                raise ValueError("signal is shorter than one frame")
            nframes = 1 if hop == 0 else (length - n) // hop + 1
        if nframes > 0 and (nframes - 1) * hop + n > length:
            raise ValueError(f"{nframes} frames at hop {hop} need {(nframes - 1) * hop + n} samples, have {length}")
        w = None
        if window is not None:
            w = _float_view(window, "window", self._f64)
            if w.numel() != n or not _is_dev(w):
                raise ValueError(f"window must be a device tensor of {n} values")
        if output == STFT_BINS:
            so = _Side(out_buffer, "out_buffer", self._f64)
            layout, o0, o1, have = so.layout, so.p0, so.p1, so.ncomplex
            refs = so.ref
        else:
            y = _float_view(out_buffer, "out_buffer", self._f64)
            layout, o0, o1, have, refs = AOS, _ptr(y), None, y.numel(), (y,)
        if have != nframes * nb:
            raise ValueError(f"out_buffer holds {have} values per side, need {nframes * nb}")
        _check(self._lib.zlfft_rfft_stft_exec(self._h, _ptr(x), int(hop), None if w is None else _ptr(w), nframes,
                                              int(output), layout, o0, o1, float(db_floor), _stream_of(x, *refs)))
        return nframes

    def rfft_twiddles(self) -> np.ndarray:
        """The plan's (wc, ws) table for k = 1..N/4, copied to the host."""
        m = self.get_size() // 4
        out = np.zeros(2 * m, dtype=np.float32)
        _check(self._lib.zlfft_plan_copy_rfft_twiddles(self._h, out.ctypes.data, out.size))
        return out.reshape(m, 2)


class RaggedBatch:
    """Many transforms of different orders, kinds, directions and layouts as
    one persistent kernel in a captured hipGraph (``zlfft_ragged_*``).

    ``descs`` is a sequence of ``(kind, order, direction, layout, in, out)``
    where ``in`` / ``out`` are device tensors (AoS: one tensor; SoA: a
    ``(real, imag)`` pair) exactly as the corresponding ``CFFT`` / ``RFFT``
    call takes them.  The tensors must stay alive as long as the batch.
    """

    def __init__(self, descs, logr=None):
        """``logr``: optional points per thread (log2: 3, 4 or 5) for CFFT
        orders 0..14 instead of the tuned table."""
        table = np.zeros(len(descs), dtype=RAGGED_DESC_DTYPE)
        for i, (kind, order, direction, layout, inb, outb) in enumerate(descs):
            def ptrs(b):
                if isinstance(b, (tuple, list)):
                    return _ptr(b[0]), _ptr(b[1])
                return _ptr(b), 0
            i0, i1 = ptrs(inb)
            o0, o1 = ptrs(outb)
            table[i] = (kind, order, direction, layout, i0, i1, o0, o1)
        self._create(table, [(d[4], d[5]) for d in descs], logr)

    @classmethod
    def from_table(cls, table: np.ndarray, keep=(), logr=None) -> "RaggedBatch":
        """A batch from a RAGGED_DESC_DTYPE array of device addresses; `keep`
        holds whatever owns that memory for the lifetime of the batch."""
        self = cls.__new__(cls)
        self._create(np.ascontiguousarray(table, dtype=RAGGED_DESC_DTYPE), keep, logr)
        return self

    def _create(self, table, keep, logr):
        self._lib = load_library()
        self._keep = keep
        h = ctypes.c_void_p()
        arr = table.ctypes.data_as(ctypes.POINTER(RaggedDesc))
        if logr is None:
            _check(self._lib.zlfft_ragged_create(ctypes.byref(h), arr, len(table)))
        else:
            lr = (ctypes.c_int * 15)(*logr)
            _check(self._lib.zlfft_ragged_create_config(ctypes.byref(h), arr, len(table), lr))
        self._h = h

    def run(self, stream: int | None = None) -> None:
        """Runs every transform: one graph launch on `stream` (default: torch's current stream)."""
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        _check(self._lib.zlfft_ragged_exec(self._h, stream))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.zlfft_ragged_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
