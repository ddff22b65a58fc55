"""Device-side synthetic workloads shared by the tests, bench.py and the
autotuner (torch supplies the device buffers)."""
from __future__ import annotations

import numpy as np

import zldsp_fft_amd as zf
from zldsp_fft_amd import signals


_ALIGN = 64  # floats: every buffer of a slab starts on a 256-byte boundary


def _slab_views(torch, slab, sizes):
    """Consecutive 256-byte-aligned float views of a 1-D tensor, one per size."""
    out, off = [], 0
    for n in sizes:
        out.append(slab[off:off + n])
        off += (n + _ALIGN - 1) // _ALIGN * _ALIGN
    return out


def _slab_len(sizes) -> int:
    return sum((n + _ALIGN - 1) // _ALIGN * _ALIGN for n in sizes)


def ragged_mix(count: int, seed: int, orders=(6, 14)):
    """The mixed ragged set: orders uniform in [lo, hi], CFFT / RFFT, forward /
    backward and AoS / SoA each 50/50 (SplitMix64 choices), inputs U[-1, 1).

    Returns descriptors (kind, order, direction, layout, in, out) shaped as the
    CFFT / RFFT calls take them, whose buffers are views into one input slab
    (filled by a single device RNG call) and one output slab, so building the
    set costs three device allocations, not one per transform."""
    import torch
    u = signals.uniform(seed, 4 * count, 0.0, 1.0).reshape(count, 4)
    lo, hi = orders
    shapes = []  # per descriptor: (kind, order, direction, layout, in float counts, out float counts)
    for i in range(count):
        order = lo + min(hi - lo, int(u[i, 0] * (hi - lo + 1)))
        kind = zf.KIND_CFFT if u[i, 1] < 0.5 else zf.KIND_RFFT
        direction = zf.FORWARD if u[i, 2] < 0.5 else zf.BACKWARD
        layout = zf.AOS if u[i, 3] < 0.5 else zf.SOA
        fi, ki = _side_floats(kind, direction, layout, order, 0)
        fo, ko = _side_floats(kind, direction, layout, order, 1)
        shapes.append((kind, order, direction, layout, [fi] * ki, [fo] * ko))
    in_sizes = [n for s in shapes for n in s[4]]
    out_sizes = [n for s in shapes for n in s[5]]
    g = torch.Generator(device="cuda").manual_seed(seed)
    in_slab = torch.rand(_slab_len(in_sizes), generator=g, device="cuda")
    in_slab.mul_(2).sub_(1)
    out_slab = torch.empty(_slab_len(out_sizes), device="cuda")
    ins = iter(_slab_views(torch, in_slab, in_sizes))
    outs = iter(_slab_views(torch, out_slab, out_sizes))
    descs = []
    for kind, order, direction, layout, fin, fout in shapes:
        complex_in = kind == zf.KIND_CFFT or direction == zf.BACKWARD
        complex_out = kind == zf.KIND_CFFT or direction == zf.FORWARD

        def side(views, cplx):
            if len(views) == 2:
                return tuple(views)
            v = views[0]
            return torch.view_as_complex(v.view(-1, 2)) if cplx else v

        inb = side([next(ins) for _ in fin], complex_in)
        outb = side([next(outs) for _ in fout], complex_out)
        descs.append((kind, order, direction, layout, inb, outb))
    return descs


VARIANTS = ([(zf.KIND_CFFT, d, l) for d in (zf.FORWARD, zf.BACKWARD) for l in (zf.AOS, zf.SOA)]
            + [(zf.KIND_RFFT, d, l) for d in (zf.FORWARD, zf.BACKWARD) for l in (zf.AOS, zf.SOA)]
            + [(zf.KIND_RFFT, zf.FORWARD_SQR_MAG, zf.AOS)])


def _side_floats(kind, direction, layout, order, side):
    """Floats per transform of one buffer of one side (`side` 0: in, 1: out)
    and how many such buffers (2 for SoA complex sides)."""
    n = 1 << order
    nb = n // 2 + 1
    if kind == zf.KIND_CFFT:
        return 2 * n if layout == zf.AOS else n, 1 if layout == zf.AOS else 2
    real_side = (side == 0) == (direction != zf.BACKWARD)
    if real_side:
        return n, 1
    if direction == zf.FORWARD_SQR_MAG:
        return nb, 1
    return (2 * nb, 1) if layout == zf.AOS else (nb, 2)


class UniformTable:
    """count transforms of one CFFT order (RFFT entries have real order + 1)
    cycling through every variant, as a descriptor table over one input and
    one output buffer per variant (U[-1, 1) inputs): no per-transform objects."""

    def __init__(self, order: int, count: int, seed: int):
        import torch
        g = torch.Generator(device="cuda").manual_seed(seed)
        self.table = np.zeros(count, dtype=zf.RAGGED_DESC_DTYPE)
        self.bufs = []
        idx = np.arange(count)
        for v, (kind, d, l) in enumerate(VARIANTS):
            rows = idx[idx % len(VARIANTS) == v]
            m = len(rows)
            if m == 0:
                continue
            o = order + (1 if kind == zf.KIND_RFFT else 0)
            fi, ki = _side_floats(kind, d, l, o, 0)
            fo, ko = _side_floats(kind, d, l, o, 1)
            ins = [torch.rand(m * fi, generator=g, device="cuda") * 2 - 1 for _ in range(ki)]
            outs = [torch.empty(m * fo, device="cuda") for _ in range(ko)]
            self.bufs.append(((kind, d, l, o, fo), ins, outs))
            k = np.arange(m, dtype=np.uint64)
            t = self.table[rows]
            t["kind"], t["order"], t["direction"], t["layout"] = kind, o, d, l
            t["in0"] = ins[0].data_ptr() + 4 * fi * k
            t["in1"] = ins[1].data_ptr() + 4 * fi * k if ki == 2 else 0
            t["out0"] = outs[0].data_ptr() + 4 * fo * k
            t["out1"] = outs[1].data_ptr() + 4 * fo * k if ko == 2 else 0
            self.table[rows] = t

    def outputs(self):
        """Each output buffer as a (transforms, floats per transform) view."""
        return [o.view(-1, meta[4]) for meta, _, outs in self.bufs for o in outs]

    def transform(self, v: int, i: int):
        """(kind, direction, layout, order), input and output of transform i
        of variant v, as numpy (complex sides as complex128)."""
        (kind, d, l, o, fo), ins, outs = self.bufs[v]

        def side(bufs, per, complex_side):
            rows = [b.view(-1, per)[i].cpu().numpy().astype(np.float64) for b in bufs]
            if not complex_side:
                return rows[0]
            if len(rows) == 2:
                return rows[0] + 1j * rows[1]
            return rows[0][0::2] + 1j * rows[0][1::2]

        m = outs[0].numel() // fo
        fi = ins[0].numel() // m
        cin = kind == zf.KIND_CFFT or d == zf.BACKWARD
        cout = kind == zf.KIND_CFFT or d == zf.FORWARD
        return (kind, d, l, o), side(ins, fi, cin), side(outs, fo, cout)

    def nbytes(self) -> int:
        return sum(4 * t.numel() for _, ins, outs in self.bufs for t in ins + outs)


def algorithmic_bytes(desc) -> int:
    """One read of the input and one write of the output of one descriptor."""
    kind, order, direction, layout = desc[:4]
    n = 1 << order
    if kind == zf.KIND_CFFT:
        return 16 * n
    nb = n // 2 + 1
    if direction == zf.FORWARD:
        return 4 * n + 8 * nb
    if direction == zf.FORWARD_SQR_MAG:
        return 4 * n + 4 * nb
    return 8 * nb + 4 * n


def time_ms(torch, fn, reps: int = 20, warmup: int = 5) -> float:
    """Mean wall time of fn() on torch's current stream, by HIP events."""
    for _ in range(warmup):
        fn()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps
