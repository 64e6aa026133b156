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
