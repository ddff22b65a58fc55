"""Multi-pass (HBM) transforms, CFFT orders above 14 and RFFT orders above 15,
against the oracle.  Tolerance rel-L2 <= 1e-6 per transform vs FP64."""
import numpy as np
import pytest

import oracle
import zldsp_fft_amd as zf
from zldsp_fft_amd import signals
from test_gpu_parity import LAYOUTS, fetch, make_in, make_out, rel, to_dev_complex

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
TOL = 1e-6


@pytest.mark.parametrize("order", range(15, 23))
def test_cfft_large_orders(order):
    n = 1 << order
    batch = 2 if order <= 19 else 1
    x = signals.complex_uniform(4000 + order, batch, n).astype(np.complex64).astype(np.complex128)
    plan = zf.CFFT(order)
    for forward in (True, False):
        ref = np.stack([oracle.cfft(r, forward, "f64") for r in x])
        first = None
        for lin, lout in (LAYOUTS if order in (15, 16, 21, 22) else LAYOUTS[::3]):
            yo = make_out((batch, n), lout)
            (plan.forward if forward else plan.backward)(make_in(x, lin), yo)
            y = fetch(yo)
            e = rel(y, ref).max()
            print(f"order={order} fwd={forward} lin={lin} lout={lout} relL2={e:.3e}", flush=True)
            assert e <= TOL
            if first is None:
                first = y
            else:
                assert np.array_equal(y, first)


@pytest.mark.parametrize("order,batch", [(16, 64), (18, 16), (20, 4), (22, 1)])
def test_large_soa_config(order, batch):
    """SoA CFFTs, re/im ~ N(0, 1), 4,194,304 points per call: forward vs the FP64
    oracle on the first and last transform, every transform against numpy's
    FP64 FFT (pocketfft, pinned to the oracle at < 1e-14 in the CPU tests), and
    backward(forward(x)) = N x on all."""
    n = 1 << order
    g = torch.Generator(device="cuda").manual_seed(3)
    re = torch.randn((batch, n), device="cuda", generator=g)
    im = torch.randn((batch, n), device="cuda", generator=g)
    plan = zf.CFFT(order)
    yr, yi = torch.empty_like(re), torch.empty_like(im)
    zr, zi = torch.empty_like(re), torch.empty_like(im)
    plan.forward((re, im), (yr, yi))
    plan.backward((yr, yi), (zr, zi))
    torch.cuda.synchronize()
    num = torch.sqrt(torch.sum((zr - n * re) ** 2 + (zi - n * im) ** 2, dim=1))
    den = torch.sqrt(torch.sum((n * re) ** 2 + (n * im) ** 2, dim=1))
    rt = (num / den).max().item()
    idx = sorted({0, batch - 1})
    x = (re[idx].cpu().numpy().astype(np.float64) + 1j * im[idx].cpu().numpy().astype(np.float64))
    y = (yr[idx].cpu().numpy().astype(np.float64) + 1j * yi[idx].cpu().numpy().astype(np.float64))
    ref = np.stack([oracle.cfft(r, True, "f64") for r in x])
    e = rel(y, ref).max()
    xa = re.cpu().numpy().astype(np.float64) + 1j * im.cpu().numpy().astype(np.float64)
    ya = yr.cpu().numpy().astype(np.float64) + 1j * yi.cpu().numpy().astype(np.float64)
    e_all = rel(ya, np.fft.fft(xa, axis=1)).max()
    print(f"order={order} batch={batch}: fwd relL2={e:.3e} (all {batch} vs numpy {e_all:.3e}) "
          f"round trip relL2={rt:.3e}", flush=True)
    assert e <= TOL and e_all <= TOL and rt <= TOL


@pytest.mark.parametrize("order", [16, 17, 19, 20, 21, 22])
def test_rfft_large_orders(order):
    """RFFT forward (bins, both layouts; squared magnitudes) and backward
    (bins from both layouts, random imaginary parts at DC and Nyquist) on 3
    frames: the post-twiddle runs in the last HBM pass and the pre-twiddle in
    the first, which pair column i with column cols - i, and every frame but
    the first starts off a cache line.  A guard past the last frame stays
    untouched."""
    n, nb, batch, guard = 1 << order, (1 << order) // 2 + 1, 3, 64
    x = signals.uniform(5000 + order, batch * n).reshape(batch, n).astype(np.float32).astype(np.float64)
    plan = zf.RFFT(order)
    xd = torch.from_numpy(x.astype(np.float32)).cuda()
    ref = np.stack([oracle.rfft_forward(r, "f64") for r in x])
    first = None
    for lay in (zf.AOS, zf.SOA):
        y = make_out((batch * nb + guard,), lay)
        yv = y[:batch * nb] if lay == zf.AOS else (y[0][:batch * nb], y[1][:batch * nb])
        plan.forward(xd, yv)
        Y = fetch(y)
        assert np.all(np.isnan(Y[batch * nb:].real)), "write past the last frame"
        Y = Y[:batch * nb].reshape(batch, nb)
        e = rel(Y, ref).max()
        print(f"rfft order={order} layout={lay} bins relL2={e:.3e}", flush=True)
        assert e <= TOL
        assert np.all(Y[:, 0].imag == 0.0) and np.all(Y[:, -1].imag == 0.0)
        if first is None:
            first = Y
        else:
            assert np.array_equal(Y, first)
    sq = torch.full((batch * nb + guard,), np.nan, dtype=torch.float32, device="cuda")
    plan.forward_sqr_mag(xd, sq[:batch * nb])
    torch.cuda.synchronize()
    s = sq.cpu().numpy().astype(np.float64)
    assert np.all(np.isnan(s[batch * nb:]))
    s = s[:batch * nb].reshape(batch, nb)
    p = np.abs(ref) ** 2
    assert (np.linalg.norm(s - p, axis=1) / np.linalg.norm(p, axis=1)).max() <= TOL
    bins = signals.complex_uniform(6000 + order, batch, nb).astype(np.complex64).astype(np.complex128)
    want = np.stack([oracle.rfft_backward(b, n, "f64") for b in bins])
    got = None
    for lay in (zf.AOS, zf.SOA):
        r = torch.full((batch * n + guard,), np.nan, dtype=torch.float32, device="cuda")
        src = make_in(bins.reshape(1, -1), lay)
        src = src[0] if lay == zf.AOS else (src[0][0], src[1][0])
        plan.backward(src, r[:batch * n])
        torch.cuda.synchronize()
        rr = r.cpu().numpy().astype(np.float64)
        assert np.all(np.isnan(rr[batch * n:]))
        rr = rr[:batch * n].reshape(batch, n)
        e = rel(rr, want).max()
        print(f"rfft order={order} layout={lay} backward relL2={e:.3e}", flush=True)
        assert e <= TOL
        if got is None:
            got = rr
        else:
            assert np.array_equal(rr, got)
