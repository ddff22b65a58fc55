"""The largest orders the library accepts (include/zlfft.h: CFFT orders up to
28 in FP32 and 27 in FP64, RFFT up to 29 and 28), checked through properties
that do not need a host FFT of that size, computed on the GPU in FP64:

* impulse delta[n - m] -> e^{-2 pi i m k / N} (exact phases from m k mod N):
  fixes the sign convention, the index order and every twiddle;
* complex exponential e^{2 pi i f n / N} -> N delta[k - f];
* backward(forward(x)) = N x (CFFT), (N/2) x (RFFT), and Parseval's
  sum |X|^2 = N sum |x|^2, on random inputs;
* at order 23 (RFFT 24) a random input against numpy's FP64 FFT.

Tolerance: rel-L2 1e-6 for FP32 plans, 1e-12 for FP64 plans."""
import numpy as np
import pytest

import zldsp_fft_amd as zf

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
TOL32, TOL64 = 1e-6, 1e-12


def rel(a, b):
    a = a.to(torch.complex128) if a.is_complex() else a.to(torch.float64)
    b = b.to(torch.complex128) if b.is_complex() else b.to(torch.float64)
    return (torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)).item()


def phases(n, m, dev, sign=-1.0):
    """e^{sign 2 pi i m k / N}, k < N, from the exact integer m k mod N."""
    k = torch.arange(n, dtype=torch.int64, device=dev)
    e = (k * m) % n
    ang = (sign * 2.0 * np.pi / n) * e.to(torch.float64)
    return torch.complex(torch.cos(ang), torch.sin(ang))


def free():
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


CASES = [(o, zf.PREC_F32) for o in range(23, zf.MAX_CFFT_ORDER + 1)] + \
        [(o, zf.PREC_F64) for o in range(23, zf.MAX_CFFT_ORDER_F64 + 1)]


@pytest.mark.parametrize("order,prec", CASES)
def test_cfft_top_orders(order, prec):
    n = 1 << order
    dev = torch.device("cuda")
    f64 = prec == zf.PREC_F64
    ct = torch.complex128 if f64 else torch.complex64
    tol = TOL64 if f64 else TOL32
    plan = zf.CFFT(order, prec)
    # impulse at m (odd, so m k mod N runs through every residue class)
    m = (0x9E3779B1 * order) % n | 1
    x = torch.zeros(n, dtype=ct, device=dev)
    x[m] = 1
    y = torch.empty_like(x)
    plan.forward(x, y)
    e_imp = rel(y, phases(n, m, dev))
    plan.backward(x, y)
    e_imp_b = rel(y, phases(n, m, dev, +1.0))
    del x
    # exponential at bin f
    f = n // 3
    x = phases(n, f, dev, +1.0).to(ct)
    plan.forward(x, y)
    want = torch.zeros(n, dtype=torch.complex128, device=dev)
    want[f] = n
    e_exp = rel(y, want)
    del x, want
    # random: round trip and Parseval
    g = torch.Generator(device=dev).manual_seed(order)
    x = torch.randn(n, dtype=ct, device=dev, generator=g)
    plan.forward(x, y)
    z = torch.empty_like(x)
    plan.backward(y, z)
    e_rt = rel(z, n * x.to(torch.complex128))
    px = torch.sum(torch.abs(x.to(torch.complex128)) ** 2).item()
    py = torch.sum(torch.abs(y.to(torch.complex128)) ** 2).item()
    e_pars = abs(py - n * px) / (n * px)
    e_np = 0.0
    if order == 23:
        ref = np.fft.fft(x.cpu().numpy().astype(np.complex128))
        e_np = float(np.linalg.norm(y.cpu().numpy() - ref) / np.linalg.norm(ref))
    print(f"order={order} f64={f64}: impulse {e_imp:.2e}/{e_imp_b:.2e} exp {e_exp:.2e} round trip {e_rt:.2e} "
          f"Parseval {e_pars:.2e} numpy {e_np:.2e}", flush=True)
    del x, y, z, plan
    free()
    assert max(e_imp, e_imp_b, e_exp, e_rt, e_np) <= tol and e_pars <= tol


RCASES = [(o, zf.PREC_F32) for o in (18, 20, 22, 24, 26, zf.MAX_CFFT_ORDER + 1)] + \
         [(o, zf.PREC_F64) for o in (18, 24, zf.MAX_CFFT_ORDER_F64 + 1)]


@pytest.mark.parametrize("order,prec", RCASES)
def test_rfft_top_orders(order, prec):
    n, nb = 1 << order, (1 << order) // 2 + 1
    dev = torch.device("cuda")
    f64 = prec == zf.PREC_F64
    rt = torch.float64 if f64 else torch.float32
    ct = torch.complex128 if f64 else torch.complex64
    tol = TOL64 if f64 else TOL32
    plan = zf.RFFT(order, prec)
    # real impulse at m: bins e^{-2 pi i m k / N}, k <= N/2; DC / Nyquist exactly real
    m = (0x9E3779B1 * order) % n | 1
    x = torch.zeros(n, dtype=rt, device=dev)
    x[m] = 1
    y = torch.empty(nb, dtype=ct, device=dev)
    plan.forward(x, y)
    e_imp = rel(y, phases(n, m, dev)[:nb])
    assert y[0].imag.item() == 0.0 and y[-1].imag.item() == 0.0
    sq = torch.empty(nb, dtype=rt, device=dev)
    plan.forward_sqr_mag(x, sq)
    e_sq = rel(sq, torch.ones(nb, dtype=torch.float64, device=dev))
    del x
    # random: round trip (N/2) x and Parseval over the half spectrum
    g = torch.Generator(device=dev).manual_seed(100 + order)
    x = torch.rand(n, dtype=rt, device=dev, generator=g) * 2 - 1
    plan.forward(x, y)
    z = torch.empty_like(x)
    plan.backward(y, z)
    e_rt = rel(z, (n / 2) * x.to(torch.float64))
    a = torch.abs(y.to(torch.complex128)) ** 2
    energy = (2 * torch.sum(a) - a[0] - a[-1]).item()
    px = torch.sum(x.to(torch.float64) ** 2).item()
    e_pars = abs(energy - n * px) / (n * px)
    e_np = 0.0
    if order == 24:
        ref = np.fft.rfft(x.cpu().numpy().astype(np.float64))
        e_np = float(np.linalg.norm(y.cpu().numpy() - ref) / np.linalg.norm(ref))
    print(f"rfft order={order} f64={f64}: impulse {e_imp:.2e} sqr_mag {e_sq:.2e} round trip {e_rt:.2e} "
          f"Parseval {e_pars:.2e} numpy {e_np:.2e}", flush=True)
    del x, y, z, sq, plan
    free()
    assert max(e_imp, e_sq, e_rt, e_np) <= tol and e_pars <= tol


def test_orders_beyond_the_maximum_are_refused():
    for kind, order, prec in ((zf.CFFT, zf.MAX_CFFT_ORDER + 1, zf.PREC_F32),
                              (zf.CFFT, zf.MAX_CFFT_ORDER_F64 + 1, zf.PREC_F64),
                              (zf.RFFT, zf.MAX_CFFT_ORDER + 2, zf.PREC_F32),
                              (zf.RFFT, zf.MAX_CFFT_ORDER_F64 + 2, zf.PREC_F64)):
        with pytest.raises(zf.ZlfftError) as ei:
            kind(order, prec)
        assert ei.value.code == 1  # ZLFFT_ERR_INVALID_ORDER
