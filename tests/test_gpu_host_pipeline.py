"""Calls on host buffers (the reference's calling convention, batched): the
library pipelines chunks through pinned staging on two streams.  Checked
against the FP64 oracle (rel-L2 <= 1e-6) and across memory kinds: pageable,
pinned and device buffers of one batch give identical results (same chunking,
same kernels), in place included."""
import numpy as np
import pytest

import oracle
import zldsp_fft_amd as zf
from zldsp_fft_amd import signals

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
TOL = 1e-6


def test_cfft_host_batch_many_chunks():
    order, batch = 12, 3000  # 64 KiB per transform: 12 chunks of 256, both slots
    n = 1 << order
    x = signals.complex_uniform(11, batch, n).astype(np.complex64)
    plan = zf.CFFT(order)
    y = np.zeros_like(x)
    plan.forward(x, y)
    idx = [0, 255, 256, 1999, 2999]
    ref = np.stack([oracle.cfft(x[i].astype(np.complex128), True, "f64") for i in idx])
    assert signals.rel_l2(y[idx].astype(np.complex128), ref).max() <= TOL
    # pinned host buffers: transferred directly, same chunks -> same bits
    xp = torch.from_numpy(x).pin_memory()
    yp = torch.zeros(x.shape, dtype=torch.complex64).pin_memory()
    plan.forward(xp, yp)
    assert np.array_equal(yp.numpy(), y)
    # device input, pageable output
    yh = np.zeros_like(x)
    xd = torch.from_numpy(x).cuda()
    plan.forward(xd, yh)
    assert signals.rel_l2(yh[idx].astype(np.complex128), ref).max() <= TOL
    # in place on pageable memory, SoA -> SoA backward round trip
    re, im = np.ascontiguousarray(y.real), np.ascontiguousarray(y.imag)
    plan.backward((re, im), (re, im))
    back = re.astype(np.float64) + 1j * im.astype(np.float64)
    assert signals.rel_l2(back[idx], n * x[idx].astype(np.complex128)).max() <= TOL


def test_rfft_host_batches():
    order, frames = 11, 5000
    n, nb = 1 << order, (1 << order) // 2 + 1
    x = signals.uniform(12, frames * n, -1.0, 1.0).astype(np.float32).reshape(frames, n)
    plan = zf.RFFT(order)
    bins = np.zeros((frames, nb), np.complex64)
    plan.forward(x, bins)
    pw = np.zeros((frames, nb), np.float32)
    plan.forward_sqr_mag(x, pw)
    re, im = np.zeros((frames, nb), np.float32), np.zeros((frames, nb), np.float32)
    plan.forward(x, (re, im))
    assert np.array_equal(re + 1j * im, bins)
    idx = [0, 777, 4999]
    ref = np.stack([oracle.rfft_forward(x[i].astype(np.float64), "f64") for i in idx])
    assert signals.rel_l2(bins[idx].astype(np.complex128), ref).max() <= TOL
    assert signals.rel_l2(pw[idx].astype(np.float64), np.abs(ref) ** 2).max() <= TOL
    back = np.zeros_like(x)
    plan.backward(bins, back)
    assert signals.rel_l2(back[idx].astype(np.float64), (n / 2) * x[idx].astype(np.float64)).max() <= TOL


def test_host_batches_multipass_and_f64():
    # multi-pass plan: one slot (shared workspace), several chunks
    order, batch = 16, 40
    n = 1 << order
    x = signals.complex_normal(13, batch, n).astype(np.complex64)
    plan = zf.CFFT(order)
    y = np.zeros_like(x)
    plan.forward(x, y)
    assert signals.rel_l2(y.astype(np.complex128), np.fft.fft(x.astype(np.complex128), axis=1)).max() <= TOL
    # FP64 host batch
    order, batch = 10, 700
    n = 1 << order
    x64 = signals.complex_uniform(14, batch, n)
    p64 = zf.CFFT(order, zf.PREC_F64)
    y64 = np.zeros_like(x64)
    p64.forward(x64, y64)
    assert signals.rel_l2(y64, np.fft.fft(x64, axis=1)).max() <= 1e-13


def test_device_input_still_being_produced():
    """A device input whose producing kernels are still queued on the legacy
    default stream (torch's default stream), passed straight to the C-ABI with
    a host output: the library's streams wait for that work before reading,
    so the result is the transform of the finished input.  The same through
    the Python wrapper on a non-default torch stream, which it synchronises."""
    import ctypes
    order, batch = 12, 2048
    n = 1 << order
    lib = zf.load_library()
    plan = zf.CFFT(order)
    g = torch.Generator(device="cuda").manual_seed(21)
    base = torch.randn((batch, n), dtype=torch.complex64, device="cuda", generator=g)
    torch.cuda.synchronize()
    x = base
    for _ in range(40):  # a few ms of queued work producing x
        x = x * 1.0001 + 0.5
    y = np.zeros((batch, n), np.complex64)
    rc = lib.zlfft_cfft_execute_batch(plan._h, zf.FORWARD, zf.AOS, ctypes.c_void_p(x.data_ptr()), None, zf.AOS,
                                      y.ctypes.data, None, batch)
    assert rc == 0, lib.zlfft_last_error()
    idx = [0, 1000, batch - 1]
    xs = x[idx].cpu().numpy().astype(np.complex128)
    ref = np.stack([oracle.cfft(r, True, "f64") for r in xs])
    assert signals.rel_l2(y[idx].astype(np.complex128), ref).max() <= TOL
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        x2 = base
        for _ in range(40):
            x2 = x2 * 0.9999 - 0.25
        y2 = np.zeros((batch, n), np.complex64)
        plan.forward(x2, y2)
    xs2 = x2[idx].cpu().numpy().astype(np.complex128)
    ref2 = np.stack([oracle.cfft(r, True, "f64") for r in xs2])
    assert signals.rel_l2(y2[idx].astype(np.complex128), ref2).max() <= TOL
