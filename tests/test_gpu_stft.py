"""STFT front-end (zlfft_rfft_stft_exec): frames at a hop, windowed at load
time, transformed to bins, squared magnitudes or dB.  Checked against the
FP64 oracle on frames cut and windowed on the host: rel-L2 <= 1e-6 (FP32
plans) / 1e-13 (FP64 plans) for bins and powers; dB within 1e-3 dB where the
power is above 1e-6 of the frame's peak (the dB of an FP32 power with
relative error e is off by 4.34 e dB)."""
import numpy as np
import pytest

import oracle
import zldsp_fft_amd as zf
from zldsp_fft_amd import signals

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def host_frames(x, n, hop, nframes, w):
    fr = np.stack([x[f * hop:f * hop + n] for f in range(nframes)])
    return fr * (w if w is not None else 1.0)


@pytest.mark.parametrize("order,prec,hop,windowed", [
    (5, zf.PREC_F32, 8, True), (11, zf.PREC_F32, 512, True), (11, zf.PREC_F32, 2051, False),
    (14, zf.PREC_F32, 1, True), (15, zf.PREC_F32, 8192, True), (17, zf.PREC_F32, 40001, True),
    (11, zf.PREC_F64, 512, True), (16, zf.PREC_F64, 16384, True),
])
def test_stft_outputs(order, prec, hop, windowed):
    n = 1 << order
    nb = n // 2 + 1
    nframes = 5
    f64 = prec == zf.PREC_F64
    dt, tol = (torch.float64, 1e-13) if f64 else (torch.float32, 1e-6)
    length = (nframes - 1) * hop + n + 3
    x = signals.uniform(500 + order, length, -1.0, 1.0)
    w = signals.hann(n) if windowed else None
    xd = torch.from_numpy(x).to(dt).cuda()
    wd = None if w is None else torch.from_numpy(w).to(dt).cuda()
    xs = xd.double().cpu().numpy()  # the values the GPU reads
    ws = None if wd is None else wd.double().cpu().numpy()
    ref = np.stack([oracle.rfft_forward(fr, "f64") for fr in host_frames(xs, n, hop, nframes, ws)])
    plan = zf.RFFT(order, prec)
    ctype = torch.complex128 if f64 else torch.complex64
    # bins, AoS and SoA
    yb = torch.full((nframes, nb), complex(np.nan, np.nan), dtype=ctype, device="cuda")
    assert plan.stft(xd, hop, yb, window=wd, nframes=nframes) == nframes
    got = yb.cpu().numpy().astype(np.complex128)
    e = signals.rel_l2(got, ref).max()
    assert e <= tol, (order, hop, e)
    assert np.all(got[:, 0].imag == 0) and np.all(got[:, -1].imag == 0)
    re = torch.full((nframes, nb), np.nan, dtype=dt, device="cuda")
    im = torch.full((nframes, nb), np.nan, dtype=dt, device="cuda")
    plan.stft(xd, hop, (re, im), window=wd, nframes=nframes)
    assert np.array_equal(re.cpu().numpy() + 1j * im.cpu().numpy(), yb.cpu().numpy())
    # power and dB
    pw = torch.full((nframes, nb), np.nan, dtype=dt, device="cuda")
    plan.stft(xd, hop, pw, window=wd, output=zf.STFT_POWER, nframes=nframes)
    pref = np.abs(ref) ** 2
    assert signals.rel_l2(pw.cpu().numpy(), pref).max() <= tol
    floor = 1e-12
    db = torch.full((nframes, nb), np.nan, dtype=dt, device="cuda")
    plan.stft(xd, hop, db, window=wd, output=zf.STFT_DB, db_floor=floor, nframes=nframes)
    dref = 10 * np.log10(np.maximum(pref, floor))
    big = pref > 1e-6 * pref.max(axis=1, keepdims=True)
    assert np.abs(db.cpu().numpy() - dref)[big].max() <= 1e-3
    assert np.all(db.cpu().numpy() >= 10 * np.log10(floor) - 1e-3)
    print(f"stft order {order} {'f64' if f64 else 'f32'} hop {hop}: bins relL2 {e:.2e}", flush=True)


def test_stft_defaults_and_errors():
    order, hop = 11, 256
    n = 1 << order
    plan = zf.RFFT(order)
    x = torch.rand(10 * n, device="cuda")
    nf = (x.numel() - n) // hop + 1
    out = torch.empty((nf, n // 2 + 1), dtype=torch.complex64, device="cuda")
    assert plan.stft(x, hop, out) == nf
    # the plain RFFT of the frames cut on the device gives the same bins
    frames = x.unfold(0, n, hop).contiguous()
    ref = torch.empty_like(out)
    plan.forward(frames, ref)
    torch.cuda.synchronize()
    a, b = out.cpu().numpy(), ref.cpu().numpy()
    assert signals.rel_l2(a, b).max() <= 1e-6
    with pytest.raises(ValueError):
        plan.stft(x, hop, out, nframes=nf + 1)  # past the end of the signal
    with pytest.raises(zf.ZlfftError) as e:
        plan.stft(x, hop, torch.empty((nf, n // 2 + 1), device="cuda"), output=zf.STFT_DB, db_floor=0.0)
    assert e.value.code == 5


@pytest.mark.parametrize("order,hop,nframes,windowed", [
    (9, 77, 33001, True), (11, 333, 8193, True), (11, 512, 8195, False), (12, 1000, 4099, True),
    (16, 4097, 21, False), (17, 12345, 9, True), (20, 300001, 3, True),
])
def test_stft_large_batches(order, hop, nframes, windowed):
    """Frame counts above any resident grid, so the persistent streaming form
    runs (including a part-filled last group and odd hops, i.e. frames that
    start off an 8-byte boundary); orders above 15 run as HBM passes whose
    first pass reads the frames from the signal.  Bins and powers are bit-identical to the
    plain RFFT of the same frames cut and windowed on the device (one FP32
    product per sample either way); dB levels match 10 log10 of those powers
    to 1e-4 dB."""
    n = 1 << order
    nb = n // 2 + 1
    plan = zf.RFFT(order)
    g = torch.Generator(device="cuda").manual_seed(order)
    x = torch.rand((nframes - 1) * hop + n, device="cuda", generator=g) * 2 - 1
    w = torch.from_numpy(signals.hann(n).astype(np.float32)).cuda() if windowed else None
    frames = x.unfold(0, n, hop)
    frames = (frames * w if windowed else frames).contiguous()
    ref = torch.empty((nframes, nb), dtype=torch.complex64, device="cuda")
    plan.forward(frames, ref)
    out = torch.full((nframes, nb), complex(np.nan, np.nan), dtype=torch.complex64, device="cuda")
    assert plan.stft(x, hop, out, window=w, nframes=nframes) == nframes
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    re = torch.full((nframes, nb), np.nan, device="cuda")
    im = torch.full((nframes, nb), np.nan, device="cuda")
    plan.stft(x, hop, (re, im), window=w, nframes=nframes)
    assert torch.equal(re, ref.real) and torch.equal(im, ref.imag)
    pref = torch.empty((nframes, nb), device="cuda")
    plan.forward_sqr_mag(frames, pref)
    pw = torch.full((nframes, nb), np.nan, device="cuda")
    plan.stft(x, hop, pw, window=w, output=zf.STFT_POWER, nframes=nframes)
    assert torch.equal(pw, pref)
    floor = 1e-9
    db = torch.full((nframes, nb), np.nan, device="cuda")
    plan.stft(x, hop, db, window=w, output=zf.STFT_DB, db_floor=floor, nframes=nframes)
    dref = 10 * np.log10(np.maximum(pref.cpu().numpy().astype(np.float64), floor))
    assert np.abs(db.cpu().numpy() - dref).max() <= 1e-4
    print(f"stft streaming order {order} hop {hop} frames {nframes}: bit-identical bins/powers", flush=True)
