"""Device-side synthetic workloads shared by the tests, bench.py and the
autotuner (torch supplies the device buffers)."""
from __future__ import annotations

import numpy as np

import zldsp_fft_amd as zf
from zldsp_fft_amd import signals


def _rand(torch, g, shape):
    return torch.rand(shape, generator=g, device="cuda", dtype=torch.float32) * 2 - 1


def ragged_desc(torch, g, kind, order, direction, layout):
    """One descriptor (kind, order, direction, layout, in, out) with random
    U[-1, 1) input and an empty output, shaped as the CFFT / RFFT call takes them."""
    n = 1 << order
    nb = n // 2 + 1
    if kind == zf.KIND_CFFT:
        if layout == zf.AOS:
            inb = torch.view_as_complex(_rand(torch, g, (n, 2)).contiguous())
            outb = torch.empty(n, dtype=torch.complex64, device="cuda")
        else:
            inb = (_rand(torch, g, n), _rand(torch, g, n))
            outb = (torch.empty(n, device="cuda"), torch.empty(n, device="cuda"))
    elif direction in (zf.FORWARD, zf.FORWARD_SQR_MAG):
        inb = _rand(torch, g, n)
        if direction == zf.FORWARD_SQR_MAG:
            outb = torch.empty(nb, device="cuda")
        elif layout == zf.AOS:
            outb = torch.empty(nb, dtype=torch.complex64, device="cuda")
        else:
            outb = (torch.empty(nb, device="cuda"), torch.empty(nb, device="cuda"))
    else:
        inb = (torch.view_as_complex(_rand(torch, g, (nb, 2)).contiguous()) if layout == zf.AOS
               else (_rand(torch, g, nb), _rand(torch, g, nb)))
        outb = torch.empty(n, device="cuda")
    return (kind, order, direction, layout, inb, outb)


def ragged_mix(count: int, seed: int, orders=(6, 14)):
    """The mixed ragged set: orders uniform in [lo, hi], CFFT / RFFT, forward /
    backward and AoS / SoA each 50/50 (SplitMix64 choices), inputs U[-1, 1)."""
    import torch
    u = signals.uniform(seed, 4 * count, 0.0, 1.0).reshape(count, 4)
    lo, hi = orders
    g = torch.Generator(device="cuda").manual_seed(seed)
    out = []
    for i in range(count):
        order = lo + min(hi - lo, int(u[i, 0] * (hi - lo + 1)))
        kind = zf.KIND_CFFT if u[i, 1] < 0.5 else zf.KIND_RFFT
        direction = zf.FORWARD if u[i, 2] < 0.5 else zf.BACKWARD
        layout = zf.AOS if u[i, 3] < 0.5 else zf.SOA
        out.append(ragged_desc(torch, g, kind, order, direction, layout))
    return out


def ragged_uniform_order(order: int, count: int, seed: int):
    """count transforms of one size, cycling through every variant (CFFT
    forward/backward x AoS/SoA, RFFT forward/backward x AoS/SoA, sqr-mag);
    `order` is the real size for the RFFT entries, so all share one CFFT order
    only for the CFFT ones."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    variants = [(zf.KIND_CFFT, d, l) for d in (zf.FORWARD, zf.BACKWARD) for l in (zf.AOS, zf.SOA)]
    variants += [(zf.KIND_RFFT, d, l) for d in (zf.FORWARD, zf.BACKWARD) for l in (zf.AOS, zf.SOA)]
    variants += [(zf.KIND_RFFT, zf.FORWARD_SQR_MAG, zf.AOS)]
    out = []
    for i in range(count):
        kind, d, l = variants[i % len(variants)]
        out.append(ragged_desc(torch, g, kind, order + (1 if kind == zf.KIND_RFFT else 0), d, l))
    return out


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
