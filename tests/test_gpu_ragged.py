"""Ragged batches: 10,000 transforms with orders uniform in 6..14, half CFFT
and half RFFT, half forward and half backward, half AoS and half SoA, inputs
uniform in [-1, 1), in one persistent graph-captured launch.  Every output
must match the per-transform plain path within rel-L2 1e-6 (and a sample the
FP64 oracle); replays of the graph give identical results."""
import numpy as np
import pytest

import oracle
import zldsp_fft_amd as zf
from zldsp_fft_amd import autotune, workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
TOL = 1e-6


def as_np(b):
    if isinstance(b, tuple):
        return b[0].cpu().numpy().astype(np.float64) + 1j * b[1].cpu().numpy().astype(np.float64)
    return b.cpu().numpy()


def plain(desc, plans):
    kind, order, direction, layout, inb, outb = desc
    key = (kind, order)
    if key not in plans:
        plans[key] = zf.CFFT(order) if kind == zf.KIND_CFFT else zf.RFFT(order)
    p = plans[key]
    out = (tuple(torch.empty_like(t) for t in outb) if isinstance(outb, tuple) else torch.empty_like(outb))
    if kind == zf.KIND_CFFT:
        (p.forward if direction == zf.FORWARD else p.backward)(inb, out)
    else:
        (p.forward if direction == zf.FORWARD else p.backward)(inb, out)
    return out


def test_ragged_mixed_10000():
    descs = workloads.ragged_mix(10000, 6)
    rb = zf.RaggedBatch(descs)
    rb.run()
    torch.cuda.synchronize()
    first = [as_np(d[5]) for d in descs]
    # replays: identical results (graph + queue counter reset by a kernel node)
    for _ in range(2):
        for d in descs:
            o = d[5]
            for t in (o if isinstance(o, tuple) else (o,)):
                t.fill_(float("nan"))
        rb.run()
        torch.cuda.synchronize()
        for d, f in zip(descs, first):
            assert np.array_equal(as_np(d[5]), f, equal_nan=False)
    plans = {}
    worst = 0.0
    for i, d in enumerate(descs):
        ref = as_np(plain(d, plans))
        got = first[i]
        e = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        worst = max(worst, e)
        assert e <= TOL, (i, d[:4], e)
    print(f"ragged 10000 vs plain path: max relL2 {worst:.3e}", flush=True)
    # a sample against the FP64 oracle
    for i in range(0, 10000, 997):
        kind, order, direction, layout, inb, outb = descs[i]
        x = as_np(inb)
        n = 1 << order
        if kind == zf.KIND_CFFT:
            ref = oracle.cfft(x.astype(np.complex128), direction == zf.FORWARD, "f64")
        elif direction == zf.FORWARD:
            ref = oracle.rfft_forward(x.astype(np.float64), "f64")
        else:
            ref = oracle.rfft_backward(x.astype(np.complex128), n, "f64")
        e = float(np.linalg.norm(first[i] - ref) / np.linalg.norm(ref))
        assert e <= TOL, (i, descs[i][:4], e)


def test_ragged_sqr_mag_and_small_orders():
    descs = []
    g = torch.Generator(device="cuda").manual_seed(1)
    for order in range(4, 16):
        n = 1 << order
        x = torch.rand(n, generator=g, device="cuda") * 2 - 1
        descs.append((zf.KIND_RFFT, order, zf.FORWARD_SQR_MAG, zf.AOS, x, torch.empty(n // 2 + 1, device="cuda")))
    for order in range(3, 15):
        n = 1 << order
        x = torch.view_as_complex((torch.rand((n, 2), generator=g, device="cuda") * 2 - 1).contiguous())
        descs.append((zf.KIND_CFFT, order, zf.BACKWARD, zf.AOS, x, torch.empty(n, dtype=torch.complex64, device="cuda")))
    rb = zf.RaggedBatch(descs)
    rb.run()
    torch.cuda.synchronize()
    for kind, order, direction, layout, inb, outb in descs:
        n = 1 << order
        x = inb.cpu().numpy()
        if kind == zf.KIND_RFFT:
            ref = np.abs(oracle.rfft_forward(x.astype(np.float64), "f64")) ** 2
        else:
            ref = oracle.cfft(x.astype(np.complex128), False, "f64")
        got = outb.cpu().numpy()
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= TOL, (kind, order)


def test_ragged_single_empty_and_lopsided():
    """Edge cases of the per-workgroup runs: one transform (one workgroup),
    an empty set (a no-op), and one order-14 CFFT among 3000 order-3 ones
    (most workgroups run only tiny items, one runs the large one), all
    layouts of the large one; each output against the FP64 oracle, and the
    large one also against the plain path."""
    g = torch.Generator(device="cuda").manual_seed(7)

    def cplx(n):
        return torch.view_as_complex((torch.rand((n, 2), generator=g, device="cuda") * 2 - 1).contiguous())

    x = cplx(1 << 14)
    y = torch.empty_like(x)
    rb = zf.RaggedBatch([(zf.KIND_CFFT, 14, zf.FORWARD, zf.AOS, x, y)])
    rb.run()
    torch.cuda.synchronize()
    ref = oracle.cfft(x.cpu().numpy().astype(np.complex128), True, "f64")
    assert np.linalg.norm(y.cpu().numpy() - ref) / np.linalg.norm(ref) <= TOL
    rb.close()

    empty = zf.RaggedBatch([])
    empty.run()
    torch.cuda.synchronize()
    empty.close()

    for layout in (zf.AOS, zf.SOA):
        big = cplx(1 << 14)
        if layout == zf.SOA:
            bin_ = (big.real.contiguous(), big.imag.contiguous())
            bout = (torch.empty(1 << 14, device="cuda"), torch.empty(1 << 14, device="cuda"))
        else:
            bin_, bout = big, torch.empty_like(big)
        descs = [(zf.KIND_CFFT, 14, zf.BACKWARD, layout, bin_, bout)]
        for _ in range(3000):
            t = cplx(8)
            descs.append((zf.KIND_CFFT, 3, zf.FORWARD, zf.AOS, t, torch.empty_like(t)))
        rb = zf.RaggedBatch(descs)
        rb.run()
        torch.cuda.synchronize()
        got = as_np(descs[0][5])
        ref = oracle.cfft(big.cpu().numpy().astype(np.complex128), False, "f64")
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= TOL, layout
        pl = as_np(plain(descs[0], {}))
        assert np.linalg.norm(got - pl) / np.linalg.norm(pl) <= TOL, layout
        xs = np.stack([d[4].cpu().numpy() for d in descs[1:]]).astype(np.complex128)
        ys = np.stack([d[5].cpu().numpy() for d in descs[1:]])
        refs = np.fft.fft(xs, axis=1)  # order 3 forward: the 8-point DFT
        e = np.linalg.norm(ys - refs, axis=1) / np.linalg.norm(refs, axis=1)
        assert e.max() <= TOL, layout
        rb.close()


def test_ragged_rejects_bad_descriptors():
    x = torch.zeros(64, dtype=torch.complex64, device="cuda")
    with pytest.raises(zf.ZlfftError) as e:
        zf.RaggedBatch([(zf.KIND_CFFT, 2, zf.FORWARD, zf.AOS, x, x)])
    assert e.value.code == 1
    with pytest.raises(zf.ZlfftError) as e:
        zf.RaggedBatch([(zf.KIND_CFFT, 15, zf.FORWARD, zf.AOS, x, x)])
    assert e.value.code == 1
    with pytest.raises(zf.ZlfftError) as e:  # 8 points per thread: 2^14 / 8 threads exceed one workgroup (512)
        zf.RaggedBatch([(zf.KIND_CFFT, 14, zf.FORWARD, zf.AOS, x, x)], logr=[3] * 15)
    assert e.value.code == 5
    with pytest.raises(zf.ZlfftError) as e:
        zf.RaggedBatch([(zf.KIND_CFFT, 6, zf.FORWARD_SQR_MAG, zf.AOS, x, x)])
    assert e.value.code == 5


def _oracle(kind, direction, order, x):
    if kind == zf.KIND_CFFT:
        return oracle.cfft(x.astype(np.complex128), direction == zf.FORWARD, "f64")
    if direction == zf.FORWARD:
        return oracle.rfft_forward(x.astype(np.float64), "f64")
    if direction == zf.FORWARD_SQR_MAG:
        return np.abs(oracle.rfft_forward(x.astype(np.float64), "f64")) ** 2
    return oracle.rfft_backward(x.astype(np.complex128), 1 << order, "f64")


def test_ragged_points_per_thread_choices_agree():
    """Every points-per-thread configuration the autotuner chooses between
    (8, 16, 32) gives the same results (per transform, rel-L2 1e-6) on every
    order it supports, and matches the FP64 oracle; the descriptor table is
    built in bulk (RaggedBatch.from_table) over one buffer per variant."""
    for order in range(3, 15):
        ut = workloads.UniformTable(order, 4 * len(workloads.VARIANTS), seed=order)
        res = {}
        for logr in autotune.LOGRS:
            if not autotune.legal(order, logr):
                continue
            table = [autotune.default_logr(n) for n in range(15)]
            table[order] = logr  # also the CFFT order of the RFFT entries (real size 2^(order + 1))
            rb = zf.RaggedBatch.from_table(ut.table, keep=ut, logr=table)
            rb.run()
            torch.cuda.synchronize()
            res[logr] = [o.cpu().numpy().astype(np.float64) for o in ut.outputs()]
            rb.close()
        first = min(res)
        for lg in res:
            for a, b in zip(res[lg], res[first]):
                e = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
                assert e.max() <= TOL, (order, lg)
        for v in range(len(ut.bufs)):
            (kind, direction, layout, o), x, got = ut.transform(v, 3)
            ref = _oracle(kind, direction, o, x)
            assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= TOL, (order, kind, direction, layout)
