"""zlfft_cfft_exec_noalloc, the allocation-free variant of zlfft_cfft_exec:
multi-pass intermediates in the library's static device scratch (64 MiB per
GPU) instead of the plan workspace.  Through the C-ABI, its outputs equal
zlfft_cfft_exec's bit for bit (same kernels, same plan), on two calls in a row
with the second smaller, for multi-pass and one-kernel plans, in both
directions and several layouts; a call needing more than the scratch holds is
refused with ZLFFT_ERR_INVALID_ARG."""
import ctypes

import numpy as np
import pytest

import oracle
import zldsp_fft_amd as zf
from zldsp_fft_amd import signals

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
SCRATCH = 64 << 20


def _ptrs(buf):
    if isinstance(buf, tuple):
        return buf[0].data_ptr(), buf[1].data_ptr()
    return buf.data_ptr(), None


def _run(lib, fn, plan, direction, lin, x, lout, y, batch):
    i0, i1 = _ptrs(x)
    o0, o1 = _ptrs(y)
    rc = fn(plan._h, direction, lin, i0, i1, lout, o0, o1, batch, torch.cuda.current_stream().cuda_stream)
    assert rc == 0, lib.zlfft_last_error()


def _buf(batch, n, layout, seed=None):
    if layout == zf.AOS:
        if seed is None:
            return torch.full((batch, n), complex(np.nan, np.nan), dtype=torch.complex64, device="cuda")
        x = signals.complex_uniform(seed, batch, n).astype(np.complex64)
        return torch.from_numpy(x).cuda()
    if seed is None:
        return (torch.full((batch, n), np.nan, device="cuda"), torch.full((batch, n), np.nan, device="cuda"))
    x = signals.complex_uniform(seed, batch, n)
    return (torch.from_numpy(x.real.astype(np.float32)).cuda(), torch.from_numpy(x.imag.astype(np.float32)).cuda())


def _same(a, b):
    if isinstance(a, tuple):
        return all(torch.equal(p, q) for p, q in zip(a, b))
    return torch.equal(a, b)


# (order, first batch, second batch): the first call fills the scratch
# (order 22: 2 x 32 MiB of intermediates, the whole 64 MiB), the second is smaller
CASES = [(22, 1, 1), (16, 64, 8), (20, 4, 1), (18, 16, 3), (12, 300, 7)]


@pytest.mark.parametrize("order,b1,b2", CASES)
def test_noalloc_matches_exec(order, b1, b2):
    lib = zf.load_library()
    n = 1 << order
    plan = zf.CFFT(order)
    for direction in (zf.FORWARD, zf.BACKWARD):
        for lin, lout in ((zf.AOS, zf.AOS), (zf.SOA, zf.SOA), (zf.AOS, zf.SOA)):
            for batch, seed in ((b1, 100 + order), (b2, 200 + order)):
                x = _buf(batch, n, lin, seed)
                ya, yb = _buf(batch, n, lout), _buf(batch, n, lout)
                _run(lib, lib.zlfft_cfft_exec, plan, direction, lin, x, lout, ya, batch)
                _run(lib, lib.zlfft_cfft_exec_noalloc, plan, direction, lin, x, lout, yb, batch)
                torch.cuda.synchronize()
                assert _same(ya, yb), (order, batch, direction, lin, lout)
    # and one sampled transform against the FP64 oracle
    x = signals.complex_uniform(300 + order, 1, n)
    xd = torch.from_numpy(x.astype(np.complex64)).cuda()
    y = torch.empty_like(xd)
    _run(lib, lib.zlfft_cfft_exec_noalloc, plan, zf.FORWARD, zf.AOS, xd, zf.AOS, y, 1)
    torch.cuda.synchronize()
    ref = oracle.cfft(x[0].astype(np.complex64).astype(np.complex128), True, "f64")
    assert signals.rel_l2(y.cpu().numpy()[0].astype(np.complex128), ref) <= 1e-6


def test_noalloc_refuses_more_than_the_scratch():
    lib = zf.load_library()
    plan = zf.CFFT(22)
    n = 1 << 22
    x = torch.zeros((2, n), dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    rc = lib.zlfft_cfft_exec_noalloc(plan._h, zf.FORWARD, zf.AOS, x.data_ptr(), None, zf.AOS, y.data_ptr(), None, 2,
                                     torch.cuda.current_stream().cuda_stream)
    assert rc == 5, rc  # ZLFFT_ERR_INVALID_ARG: 2 x 64 MiB of intermediates
    assert b"64" in lib.zlfft_last_error() or b"67108864" in lib.zlfft_last_error()
