"""Error paths of the C-ABI, forced from the host with the library's fault
injection (ZLFFT_INJECT, read when a plan or ragged batch is created): device
allocations that fail at creation or at a later call, and kernel launches that
fail.  Each call returns the code a real failure would (ZLFFT_ERR_ALLOC = 3,
ZLFFT_ERR_HIP = 4) with a message, leaves nothing half-built, and plans created
afterwards without the injection work normally.  No kernel runs in a failing
state: the injected launch failures replace the launch itself."""
import os

import numpy as np
import pytest

import oracle
import zldsp_fft_amd as zf
from zldsp_fft_amd import signals, workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
ALLOC, HIP = 3, 4


def injected(mode, make):
    os.environ["ZLFFT_INJECT"] = mode
    try:
        return make()
    finally:
        del os.environ["ZLFFT_INJECT"]


def expect(code, fn):
    with pytest.raises(zf.ZlfftError) as ei:
        fn()
    assert ei.value.code == code, ei.value
    assert str(ei.value)


@pytest.mark.parametrize("kind,order,prec", [
    (zf.CFFT, 10, zf.PREC_F32), (zf.CFFT, 20, zf.PREC_F32), (zf.RFFT, 11, zf.PREC_F32),
    (zf.CFFT, 12, zf.PREC_F16_MFMA_SPLIT), (zf.CFFT, 12, zf.PREC_F64), (zf.RFFT, 18, zf.PREC_F64)])
def test_allocation_failure_at_creation(kind, order, prec):
    expect(ALLOC, lambda: injected("alloc", lambda: kind(order, prec)))
    with kind(order, prec) as plan:  # afterwards, without the injection
        assert plan.get_order() == order


def test_workspace_and_staging_allocation_failures():
    n = 1 << 18
    plan = injected("work_alloc", lambda: zf.CFFT(18))
    x = torch.zeros(n, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    expect(ALLOC, lambda: plan.forward(x, y))          # multi-pass workspace
    expect(ALLOC, lambda: plan.forward(x, y))          # and again: nothing half-built
    small = injected("work_alloc", lambda: zf.CFFT(10))
    xh = np.zeros(1024, np.complex64)
    expect(ALLOC, lambda: small.forward(xh, np.empty_like(xh)))  # host-pipeline staging
    plan.close()
    small.close()
    ok = zf.CFFT(18)
    xr = signals.complex_uniform(5, 1, n).astype(np.complex64)
    xd = torch.from_numpy(xr).cuda().view(-1)
    ok.forward(xd, y)
    torch.cuda.synchronize()
    ref = oracle.cfft(xr[0].astype(np.complex128), True, "f64")
    assert signals.rel_l2(y.cpu().numpy().astype(np.complex128), ref) <= 1e-6


def test_launch_failures():
    x = torch.zeros(4 * 1024, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    for make, call in (
            (lambda: zf.CFFT(10), lambda p: p.forward(x, y)),                       # LDS kernel
            (lambda: zf.CFFT(12, zf.PREC_F16_MFMA_SPLIT), lambda p: p.forward(x, y)),  # MFMA kernel
            (lambda: zf.CFFT(16), lambda p: p.forward(torch.zeros(1 << 16, dtype=torch.complex64, device="cuda"),
                                                      torch.empty(1 << 16, dtype=torch.complex64, device="cuda"))),
            (lambda: zf.RFFT(17), lambda p: p.forward(torch.zeros(1 << 17, device="cuda"),
                                                      torch.empty((1 << 16) + 1, dtype=torch.complex64,
                                                                  device="cuda"))),
            (lambda: zf.RFFT(17), lambda p: p.backward(torch.zeros((1 << 16) + 1, dtype=torch.complex64,
                                                                   device="cuda"),
                                                       torch.empty(1 << 17, device="cuda"))),
            (lambda: zf.RFFT(11), lambda p: p.stft(torch.zeros(4096, device="cuda"), 512,
                                                   torch.empty(5 * 1025, device="cuda"), output=zf.STFT_POWER)),
            (lambda: zf.CFFT(10), lambda p: p.forward(np.zeros(1024, np.complex64), np.empty(1024, np.complex64))),
    ):
        plan = injected("launch", make)
        expect(HIP, lambda: call(plan))
        plan.close()
    torch.cuda.synchronize()


def test_ragged_failures():
    descs = workloads.ragged_mix(50, 3)
    expect(ALLOC, lambda: injected("alloc", lambda: zf.RaggedBatch(descs)))
    rb = injected("launch", lambda: zf.RaggedBatch(descs))
    expect(HIP, rb.run)
    rb.close()
    rb = zf.RaggedBatch(descs)  # without the injection
    rb.run()
    torch.cuda.synchronize()
    rb.close()
