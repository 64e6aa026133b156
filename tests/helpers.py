"""Shared test helpers: seeded cases and source traces."""
import numpy as np
from scipy.integrate import cumulative_trapezoid

from swit_amd import workloads


def integrated_sources(w, amp=1.0):
    """(ns, nt) float32: cumtrapz of the Ricker with unit spacing (toolbox/solver.py:47)."""
    wav = workloads.ricker(w.nt, w.dt, w.f0, amp)
    src = cumulative_trapezoid(wav, initial=0).astype(np.float32)
    return np.broadcast_to(src, (w.ns, w.nt)).copy()


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))


def quantize_strips(boundary, mode):
    """The 16-bit strip encoding applied to FP32 strips (ns, nt, S), returned decoded:
    per shot and step, v * 2^k with k = min(126, 141 - e(m)), e(m) the biased FP32
    exponent of the row's max |v| (so the max lands in [2^14, 2^15)), rounded to nearest
    even in FP16 or BF16, times 2^-k."""
    b = np.ascontiguousarray(boundary, dtype=np.float32)
    m = np.abs(b).max(axis=-1, keepdims=True).astype(np.float32)
    e = ((m.view(np.uint32) >> 23) & 0xFF).astype(np.int64)
    k = np.minimum(126, 141 - e)
    inv = ((k + 127).astype(np.uint32) << 23).view(np.float32)
    scale = ((127 - k).astype(np.uint32) << 23).view(np.float32)
    x = (b * inv).astype(np.float32)
    if mode == "f16":
        q = x.astype(np.float16).astype(np.float32)
    elif mode == "bf16":
        bits = x.view(np.uint32).astype(np.uint64)
        bits = (bits + 0x7FFF + ((bits >> 16) & 1)) >> 16
        q = (bits.astype(np.uint32) << 16).view(np.float32)
    else:
        raise ValueError(mode)
    return (q * scale).astype(np.float32)
