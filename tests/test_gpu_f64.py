"""CFFT<double> / RFFT<double> on the GPU (ZLFFT_PREC_F64; the reference's
classes are templates on F, src/zldsp_fft_cfft.hpp:11, src/zldsp_fft_rfft.hpp:11).

Oracles: the FP64 restatement of the reference (oracle, the reference's own
CFFT<double> algorithm), the long-double brute-force DFT fixtures, and
numpy.fft (pocketfft) for the multi-pass sizes.  Tolerance: rel-L2 per
transform <= 1e-13 (FP64 rounding, 2^-53 = 1.1e-16, times the log2 N growth of
an FFT's error, with margin); the measured values are printed."""
import numpy as np
import pytest

import oracle
import zldsp_fft_amd as zf
from zldsp_fft_amd import signals

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-13
LAYOUTS = [(zf.AOS, zf.AOS), (zf.AOS, zf.SOA), (zf.SOA, zf.AOS), (zf.SOA, zf.SOA)]


def make_in(x, layout):
    if layout == zf.AOS:
        return torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex128)).cuda()
    return (torch.from_numpy(np.ascontiguousarray(x.real)).cuda(), torch.from_numpy(np.ascontiguousarray(x.imag)).cuda())


def make_out(shape, layout):
    if layout == zf.AOS:
        return torch.full(shape, complex(np.nan, np.nan), dtype=torch.complex128, device="cuda")
    return (torch.full(shape, np.nan, dtype=torch.float64, device="cuda"),
            torch.full(shape, np.nan, dtype=torch.float64, device="cuda"))


def fetch(buf):
    torch.cuda.synchronize()
    if isinstance(buf, tuple):
        return buf[0].cpu().numpy() + 1j * buf[1].cpu().numpy()
    return buf.cpu().numpy()


@pytest.mark.parametrize("order", range(0, 14))
def test_cfft64_lds_orders(order):
    """Orders 0-13 (one LDS-resident launch), every layout pair, both
    directions, against the FP64 oracle."""
    n = 1 << order
    x = signals.complex_uniform(100 + order, 3, n)
    plan = zf.CFFT(order, zf.PREC_F64)
    worst = 0.0
    for forward in (True, False):
        ref = np.stack([oracle.cfft(r, forward, "f64") for r in x])
        for lin, lout in LAYOUTS:
            y = make_out((3, n), lout)
            (plan.forward if forward else plan.backward)(make_in(x, lin), y)
            e = signals.rel_l2(fetch(y), ref).max()
            worst = max(worst, e)
            assert e <= TOL, (order, forward, lin, lout, e)
    print(f"f64 cfft order {order}: max relL2 vs FP64 oracle {worst:.2e}", flush=True)


@pytest.mark.parametrize("order", [5, 8, 10, 12, 13])
def test_cfft64_streaming_batches(order):
    """Batches larger than the resident workgroups (the persistent FP64
    streaming kernel, 8 points per thread, where two workgroups fit a CU),
    every layout pair, both directions, a part-filled last group: sampled
    transforms against the FP64 oracle."""
    n = 1 << order
    batch = max(2048, (1 << 21) // n) + 3
    g = torch.Generator(device="cuda").manual_seed(200 + order)
    x = torch.randn((batch, n), dtype=torch.complex128, device="cuda", generator=g)
    plan = zf.CFFT(order, zf.PREC_F64)
    idx = np.unique(np.linspace(0, batch - 1, 10).astype(int))
    xs = x[idx].cpu().numpy()
    worst = 0.0
    for forward in (True, False):
        ref = np.stack([oracle.cfft(r, forward, "f64") for r in xs])
        for lin, lout in LAYOUTS:
            xin = x if lin == zf.AOS else (x.real.contiguous(), x.imag.contiguous())
            y = make_out((batch, n), lout)
            (plan.forward if forward else plan.backward)(xin, y)
            got = fetch(tuple(t[idx] for t in y) if isinstance(y, tuple) else y[idx])
            e = signals.rel_l2(got, ref).max()
            worst = max(worst, e)
            assert e <= TOL, (order, forward, lin, lout, e)
    print(f"f64 streaming order {order} batch {batch}: max relL2 vs FP64 oracle {worst:.2e}", flush=True)


def test_cfft64_golden_fixtures(fixtures):
    """Against the long-double brute-force DFT fixtures (orders 0-10)."""
    for order in range(0, 11):
        x = fixtures[f"cfft_x_{order}"]
        plan = zf.CFFT(order, zf.PREC_F64)
        for key, run in (("fwd", plan.forward), ("bwd", plan.backward)):
            y = make_out((1 << order,), zf.AOS)
            run(make_in(x, zf.AOS), y)
            ref = fixtures[f"cfft_{key}_{order}"]
            assert signals.rel_l2(fetch(y)[None], ref[None])[0] <= 1e-14, (order, key)


@pytest.mark.parametrize("order,batch", [(14, 4), (15, 2), (16, 2), (18, 1), (20, 1), (22, 1)])
def test_cfft64_multipass_orders(order, batch):
    """Orders 14-22 run as HBM passes in FP64: forward against numpy.fft
    (FP64), backward as N ifft, AoS and SoA, and the round trip N x."""
    n = 1 << order
    x = signals.complex_normal(200 + order, batch, n)
    plan = zf.CFFT(order, zf.PREC_F64)
    for lin, lout in ((zf.AOS, zf.AOS), (zf.SOA, zf.SOA), (zf.AOS, zf.SOA)):
        y = make_out((batch, n), lout)
        plan.forward(make_in(x, lin), y)
        e = signals.rel_l2(fetch(y), np.fft.fft(x, axis=1)).max()
        assert e <= TOL, (order, lin, lout, e)
        z = make_out((batch, n), lin)
        plan.backward(y, z)
        e2 = signals.rel_l2(fetch(z), n * x).max()
        assert e2 <= TOL, (order, lin, lout, e2)
    print(f"f64 cfft order {order}: fwd {e:.2e} round trip {e2:.2e}", flush=True)


@pytest.mark.parametrize("order", [1, 2, 3, 5, 8, 11, 14, 15, 16, 17, 18, 19, 20, 21])
def test_rfft64(order):
    """RFFT<double>: forward AoS / SoA, forward_sqr_mag and backward (scaled by
    N/2, imaginary parts of bins 0 and N/2 ignored), against the FP64 oracle;
    exact zero imaginary parts at DC and Nyquist.  Orders 15-21 run as HBM
    passes with the post-/pre-twiddle in the last / first pass, whose
    workgroups hold 32, 16 or 8 column pairs (radix 128-256, 512, 1024)."""
    n = 1 << order
    nb = n // 2 + 1
    batch = 3
    x = signals.uniform(300 + order, batch * n, -1.0, 1.0).reshape(batch, n)
    plan = zf.RFFT(order, zf.PREC_F64)
    ref = np.stack([oracle.rfft_forward(r, "f64") for r in x])
    xd = torch.from_numpy(x).cuda()
    for lout in (zf.AOS, zf.SOA):
        y = make_out((batch, nb), lout)
        plan.forward(xd, y)
        got = fetch(y)
        assert signals.rel_l2(got, ref).max() <= TOL, (order, lout)
        assert np.all(got[:, 0].imag == 0.0) and np.all(got[:, -1].imag == 0.0)
    p = torch.full((batch, nb), np.nan, dtype=torch.float64, device="cuda")
    plan.forward_sqr_mag(xd, p)
    assert signals.rel_l2(fetch(p), np.abs(ref) ** 2).max() <= TOL, order
    bins = signals.complex_uniform(400 + order, batch, nb)  # random imag at DC / Nyquist: ignored
    refb = np.stack([oracle.rfft_backward(b, n, "f64") for b in bins])
    for lin in (zf.AOS, zf.SOA):
        o = torch.full((batch, n), np.nan, dtype=torch.float64, device="cuda")
        plan.backward(make_in(bins, lin), o)
        assert signals.rel_l2(fetch(o), refb).max() <= TOL, (order, lin)


@pytest.mark.parametrize("order", [5, 8, 10, 11, 12, 13])
def test_rfft64_streaming_batches(order):
    """RFFT<double> batches larger than the resident workgroups: the persistent
    streaming kernel (real orders up to 12) and the one-launch kernel above,
    AoS / SoA bins and squared magnitudes on sampled frames against the FP64
    oracle, the part-filled last group included; DC and Nyquist exactly real."""
    n = 1 << order
    nb = n // 2 + 1
    batch = max(4096, (1 << 22) // n) + 3
    g = torch.Generator(device="cuda").manual_seed(600 + order)
    xd = torch.rand((batch, n), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    plan = zf.RFFT(order, zf.PREC_F64)
    idx = np.unique(np.linspace(0, batch - 1, 10).astype(int))
    x = xd[idx].cpu().numpy()
    ref = np.stack([oracle.rfft_forward(r, "f64") for r in x])
    for lout in (zf.AOS, zf.SOA):
        y = make_out((batch, nb), lout)
        plan.forward(xd, y)
        torch.cuda.synchronize()
        got = fetch(tuple(t[idx] for t in y) if isinstance(y, tuple) else y[idx])
        e = signals.rel_l2(got, ref).max()
        print(f"f64 rfft stream order={order} batch={batch} lout={lout} relL2={e:.2e}", flush=True)
        assert e <= TOL, (order, lout)
        assert np.all(got[:, 0].imag == 0.0) and np.all(got[:, -1].imag == 0.0)
    p = torch.full((batch, nb), np.nan, dtype=torch.float64, device="cuda")
    plan.forward_sqr_mag(xd, p)
    torch.cuda.synchronize()
    assert signals.rel_l2(p[idx].cpu().numpy(), np.abs(ref) ** 2).max() <= TOL, order
    assert not torch.isnan(p).any()
    # backward over the same batch from random bins (AoS and SoA): the
    # streaming C2R kernel up to real order 12; imaginary parts of DC and
    # Nyquist ignored
    Xb = torch.randn((batch, nb), dtype=torch.complex128, device="cuda", generator=g)
    refb = np.stack([oracle.rfft_backward(b, n, "f64") for b in Xb[idx].cpu().numpy()])
    outs = []
    for lin in (zf.AOS, zf.SOA):
        o = torch.full((batch, n), np.nan, dtype=torch.float64, device="cuda")
        plan.backward(Xb if lin == zf.AOS else (Xb.real.contiguous(), Xb.imag.contiguous()), o)
        torch.cuda.synchronize()
        e = signals.rel_l2(o[idx].cpu().numpy(), refb).max()
        print(f"f64 rfft stream backward order={order} batch={batch} lin={lin} relL2={e:.2e}", flush=True)
        assert e <= TOL, (order, lin)
        assert not torch.isnan(o).any()
        outs.append(o)
    if order >= 8:  # both layouts stream (or both take the one-launch kernel): the same bits
        assert torch.equal(outs[0], outs[1]), order


def test_f64_host_pointers_and_dtype_checks():
    """Synchronous host-pointer calls (numpy complex128 / float64) equal the
    device batch bit for bit; FP32 buffers are refused by an FP64 plan."""
    order = 9
    n = 1 << order
    x = signals.complex_uniform(7, 1, n)[0]
    plan = zf.CFFT(order, zf.PREC_F64)
    yh = np.empty(n, dtype=np.complex128)
    plan.forward(np.ascontiguousarray(x), yh)
    yd = make_out((n,), zf.AOS)
    plan.forward(make_in(x, zf.AOS), yd)
    assert np.array_equal(yh, fetch(yd))
    r = zf.RFFT(order, zf.PREC_F64)
    xr = signals.uniform(8, n, -1.0, 1.0)
    bh = np.empty(n // 2 + 1, dtype=np.complex128)
    r.forward(xr, bh)
    assert signals.rel_l2(bh[None], oracle.rfft_forward(xr, "f64")[None])[0] <= TOL
    with pytest.raises(TypeError):
        plan.forward(torch.zeros(n, dtype=torch.complex64, device="cuda"),
                     torch.zeros(n, dtype=torch.complex64, device="cuda"))
    with pytest.raises(TypeError):
        r.forward(torch.zeros(n, dtype=torch.float32, device="cuda"),
                  torch.zeros(n // 2 + 1, dtype=torch.complex128, device="cuda"))
