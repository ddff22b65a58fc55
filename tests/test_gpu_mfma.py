"""MFMA split-precision 4096-point CFFTs against the FP64 oracle.

Radix-16 passes run as 16x16 DFT-matrix products on FP16 / BF16 matrix cores
with FP32 accumulate.  Two input sets: re/im i.i.d. N(0, 1), and a
dynamic-range set (unit complex exponential + -60 dB exponential + -100 dB
complex white noise).  Reported per precision: rel-L2 (max over the batch)
and max-bin error max_k |y_k - y*_k| / max_k |y*_k|.  The split FP16 and
three-way split BF16 variants must meet rel-L2 <= 1e-6; the unsplit and
two-way BF16 variants are measured and bounded loosely (their mantissas cannot
reach 1e-6).
"""
import numpy as np
import pytest

import oracle
import zldsp_fft_amd as zf
from zldsp_fft_amd import signals

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

PRECS = {
    "f16": (zf.PREC_F16_MFMA, 2e-3),
    "f16_split": (zf.PREC_F16_MFMA_SPLIT, 1e-6),
    "bf16": (zf.PREC_BF16_MFMA, 2e-2),
    "bf16_split": (zf.PREC_BF16_MFMA_SPLIT, 1e-4),
    "bf16_split3": (zf.PREC_BF16_MFMA_SPLIT3, 1e-6),
}


def _sets(batch):
    a = signals.complex_normal(4, batch, 4096)
    b = signals.dynamic_range_set(5, batch, 4096)
    return {"normal": a.astype(np.complex64).astype(np.complex128),
            "dynamic": b.astype(np.complex64).astype(np.complex128)}


@pytest.fixture(scope="module")
def inputs():
    batch = 24
    sets = _sets(batch)
    refs = {k: (np.stack([oracle.cfft(r, True, "f64") for r in v]),
                np.stack([oracle.cfft(r, False, "f64") for r in v])) for k, v in sets.items()}
    return sets, refs


@pytest.mark.parametrize("name", list(PRECS))
def test_mfma_precision(name, inputs):
    prec, tol = PRECS[name]
    sets, refs = inputs
    plan = zf.CFFT(12, prec)
    for sname, x in sets.items():
        xd = torch.from_numpy(x.astype(np.complex64)).cuda()
        for forward, ref in ((True, refs[sname][0]), (False, refs[sname][1])):
            y = torch.empty_like(xd)
            (plan.forward if forward else plan.backward)(xd, y)
            torch.cuda.synchronize()
            yh = y.cpu().numpy().astype(np.complex128)
            rl2 = signals.rel_l2(yh, ref).max()
            mbe = (np.abs(yh - ref).max(axis=1) / np.abs(ref).max(axis=1)).max()
            print(f"mfma {name:12s} set={sname:8s} fwd={forward}: relL2={rl2:.3e} max-bin={mbe:.3e}", flush=True)
            assert np.all(np.isfinite(yh))
            assert rl2 <= tol


def test_mfma_unsupported_shapes():
    with pytest.raises(zf.ZlfftError) as e:
        zf.CFFT(10, zf.PREC_F16_MFMA_SPLIT)
    assert e.value.code == 6
    plan = zf.CFFT(12, zf.PREC_F16_MFMA_SPLIT)
    x = torch.zeros(4096, dtype=torch.float32, device="cuda")
    y = torch.zeros(4096, dtype=torch.float32, device="cuda")
    with pytest.raises(zf.ZlfftError) as e:
        plan.forward((x, x), (y, y))
    assert e.value.code == 6


@pytest.mark.parametrize("name", ["f16_split", "bf16_split3", "f16"])
def test_mfma_persistent_form_matches(name):
    """A batch larger than one wave of resident workgroups runs the persistent
    form (register prefetch of the next transform, sink loads past the end);
    its outputs equal those of the one-transform-per-workgroup form (the same
    inputs in chunks of 100) bit for bit, in both directions, and a sample
    meets the precision's tolerance against the FP64 oracle."""
    prec, tol = PRECS[name]
    plan = zf.CFFT(12, prec)
    batch = 1500
    x = signals.complex_normal(9, batch, 4096).astype(np.complex64)
    xd = torch.from_numpy(x).cuda()
    for forward in (True, False):
        run = plan.forward if forward else plan.backward
        y = torch.empty_like(xd)
        run(xd, y)
        yc = torch.empty_like(xd)
        for s in range(0, batch, 100):
            run(xd[s:s + 100], yc[s:s + 100])
        torch.cuda.synchronize()
        assert torch.equal(y, yc), (name, forward)
        idx = [0, 511, 512, 1023, 1499]
        ref = np.stack([oracle.cfft(x[i].astype(np.complex128), forward, "f64") for i in idx])
        got = y[idx].cpu().numpy().astype(np.complex128)
        assert signals.rel_l2(got, ref).max() <= tol, (name, forward)


def _dynamic_range_gpu(seed, batch, n):
    """The dynamic-range set, generated on the device: per transform a unit
    complex exponential at frequency f1, a -60 dB one at f2 (f1, f2 ~ U[0, N)
    bins, SplitMix64) and complex white noise of sigma 1e-5 per component."""
    p = torch.from_numpy(signals.uniform(seed, batch * 2, 0.0, 1.0).reshape(batch, 2)).cuda()
    t = torch.arange(n, dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(seed)
    noise = 1e-5 * torch.randn((batch, n), dtype=torch.complex128, device="cuda", generator=g) * np.sqrt(2.0)
    ph1 = 2 * np.pi * (p[:, 0:1] * n) * t / n
    ph2 = 2 * np.pi * (p[:, 1:2] * n) * t / n
    x = torch.polar(torch.ones_like(ph1), ph1) + 1e-3 * torch.polar(torch.ones_like(ph2), ph2) + noise
    return x.to(torch.complex64)


@pytest.mark.parametrize("name", ["f16_split", "bf16_split3"])
def test_mfma_full_batch_both_sets(name):
    """The full 16384 x 4096-point batch of the headline workload on the split
    MFMA kernels (the persistent form), both input sets, both directions:
    every transform within rel-L2 1e-6 of the FP64 GPU path (CFFT<double>,
    itself within 5e-16 of the FP64 oracle), and sampled transforms against
    the oracle directly."""
    prec, tol = PRECS[name]
    batch, n = 16384, 4096
    plan, p64 = zf.CFFT(12, prec), zf.CFFT(12, zf.PREC_F64)
    g = torch.Generator(device="cuda").manual_seed(4)
    sets = {"normal": torch.randn((batch, n), dtype=torch.complex64, device="cuda", generator=g),
            "dynamic": _dynamic_range_gpu(5, batch, n)}
    for sname, x in sets.items():
        x64 = x.to(torch.complex128)
        for forward in (True, False):
            y = torch.empty_like(x)
            (plan.forward if forward else plan.backward)(x, y)
            y64 = torch.empty_like(x64)
            (p64.forward if forward else p64.backward)(x64, y64)
            d = y.to(torch.complex128) - y64
            rl2 = (torch.linalg.vector_norm(d, dim=1) / torch.linalg.vector_norm(y64, dim=1)).max().item()
            mbe = (d.abs().amax(dim=1) / y64.abs().amax(dim=1)).max().item()
            print(f"mfma {name} full batch set={sname} fwd={forward}: relL2 max={rl2:.3e} max-bin={mbe:.3e}",
                  flush=True)
            assert torch.isfinite(y).all().item()
            assert rl2 <= tol
            idx = [0, 8191, 16383]
            ref = np.stack([oracle.cfft(r, forward, "f64") for r in x[idx].cpu().numpy().astype(np.complex128)])
            assert signals.rel_l2(y[idx].cpu().numpy().astype(np.complex128), ref).max() <= tol
            del y, y64, d
        del x64
    torch.cuda.empty_cache()
