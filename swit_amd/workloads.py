"""Seeded synthetic workloads: velocity models, density, wavelet and acquisition for
the solver's benchmark and parity configurations.

Density follows Gardner, rho = 310 vp^0.25, as in every reference example
(examples/case-00-marmousi2/SWIT_workflow_LBGFS.py:49-50).  Coordinates are
(ix dx, iz dx) metres; receivers sit at every surface cell (z = 0).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.ndimage import gaussian_filter


@dataclass
class Workload:
    name: str
    nx: int
    nz: int
    dx: float
    dt: float
    nt: int
    npml: int
    f0: float
    vp_true: np.ndarray
    vp_init: np.ndarray
    src_xz: np.ndarray            # (ns, 2) metres
    rec_xz: np.ndarray            # (ns, ng, 2) metres
    extra: dict = field(default_factory=dict)

    @property
    def ns(self):
        return self.src_xz.shape[0]

    @property
    def ng(self):
        return self.rec_xz.shape[1]

    def rho(self, vp):
        return gardner(vp)


def gardner(vp):
    return (310.0 * np.power(np.asarray(vp, dtype=np.float64), 0.25)).astype(np.float32)


def ricker(nt, dt, f0, amp=1.0):
    """Ricker with t0 = 1.2/f0 on t = k dt (toolbox/solver.py:183-197)."""
    t = np.linspace(0, dt * nt, num=nt, endpoint=False)
    t0 = 1.2 / f0
    temp = (np.pi * f0) ** 2
    return amp * (1 - 2 * temp * (t - t0) ** 2) * np.exp(-temp * (t - t0) ** 2)


def surface_receivers(nx, dx, ns):
    rec = np.zeros((nx, 2), np.float32)
    rec[:, 0] = np.arange(nx) * dx
    return np.broadcast_to(rec, (ns, nx, 2)).copy()


def shots_at(ix_list, iz, dx):
    ix_list = np.asarray(ix_list)
    s = np.zeros((len(ix_list), 2), np.float32)
    s[:, 0] = ix_list * dx
    s[:, 1] = iz * dx
    return s


def layered_faulted(nx, nz, vmin, vmax, nlayers, rng, water_rows=0, dip_deg=8.0, faults=()):
    """Dipping layers (velocity increasing with depth) cut by 60-degree normal faults.
    faults: sequence of (x0_cells, throw_cells)."""
    x = np.arange(nx)[:, None].astype(np.float64)
    z = np.arange(nz)[None, :].astype(np.float64)
    top = water_rows + 2
    base = np.sort(rng.uniform(top, nz - 3, nlayers - 1))
    dips = np.deg2rad(rng.uniform(-dip_deg, dip_deg, nlayers - 1))
    zz = np.broadcast_to(z, (nx, nz)).copy()
    for x0, throw in faults:
        hanging = x > x0 + z / np.tan(np.deg2rad(60.0))
        zz = np.where(hanging, zz - throw, zz)
    idx = np.zeros((nx, nz), np.int64)
    for k in range(nlayers - 1):
        depth = base[k] + np.tan(dips[k]) * (x - nx / 2.0)
        idx += (zz > depth).astype(np.int64)
    vel = np.linspace(vmin, vmax, nlayers)
    vp = vel[np.clip(idx, 0, nlayers - 1)]
    if water_rows:
        vp[:, :water_rows] = 1500.0
    return vp.astype(np.float32)


def config1(seed=1):
    """201x81, dx 25 m, 40-cell PML, nt 2000, dt 2 ms; 6 flat layers 1500-4000 m/s plus a
    Gaussian anomaly (sigma 8 cells, +300 m/s); source at cell (100, 2); 201 receivers."""
    rng = np.random.default_rng(seed)
    nx, nz, dx = 201, 81, 25.0
    iface = np.sort(rng.choice(np.arange(8, 75), 5, replace=False))
    vel = np.linspace(1500, 4000, 6)
    z = np.arange(nz)
    vp = np.empty((nx, nz))
    vp[:] = vel[np.searchsorted(iface, z, side="right")][None, :]
    X, Z = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    vp += 300.0 * np.exp(-((X - 100) ** 2 + (Z - 45) ** 2) / (2 * 8.0 ** 2))
    vp = vp.astype(np.float32)
    init = gaussian_filter(vp.astype(np.float64), 10, mode="nearest").astype(np.float32)
    return Workload("201x81 flat layers + anomaly, 1 shot", nx, nz, dx, 0.002, 2000, 40, 5.0,
                    vp, init, shots_at([100], 2, dx), surface_receivers(nx, dx, 1))


def config2(seed=2, nt=4000):
    """481x141, dx 25 m, 40-cell PML, nt 4000, dt 1.9 ms; 12 dipping layers 1500-5500 m/s,
    two normal faults, 10-cell water layer; 49 shots every 10 cells at depth 2; 481
    receivers."""
    rng = np.random.default_rng(seed)
    nx, nz, dx = 481, 141, 25.0
    faults = [(rng.uniform(120, 200), rng.uniform(5, 15)), (rng.uniform(280, 380), rng.uniform(5, 15))]
    vp = layered_faulted(nx, nz, 1500.0, 5500.0, 12, rng, water_rows=10, faults=faults)
    init = gaussian_filter(vp.astype(np.float64), 15, mode="nearest").astype(np.float32)
    src = shots_at(np.arange(0, 481, 10), 2, dx)
    return Workload("481x141 dipping layers + faults, 49 shots", nx, nz, dx, 0.0019, nt, 40, 5.0,
                    vp, init, src, surface_receivers(nx, dx, len(src)))


def config3(seed=3):
    """161x61, dx 20 m, nt 1500, dt 1.5 ms, 8 shots; smooth 1500->3000 m/s background and a
    smooth random perturbation scaled to 50 m/s."""
    rng = np.random.default_rng(seed)
    nx, nz, dx = 161, 61, 20.0
    vp = np.broadcast_to(np.linspace(1500, 3000, nz)[None, :], (nx, nz)).astype(np.float32)
    dm = gaussian_filter(rng.standard_normal((nx, nz)), 6)
    dm = (dm / np.abs(dm).max() * 50.0).astype(np.float32)
    src = shots_at(10 + 20 * np.arange(8), 2, dx)
    w = Workload("161x61 smooth gradient, 8 shots", nx, nz, dx, 0.0015, 1500, 40, 5.0,
                 (vp + dm).astype(np.float32), vp.copy(), src, surface_receivers(nx, dx, 8))
    w.extra["dm"] = dm
    return w


def config4(seed=4):
    """Three grid families (uneven sizes and record lengths), 22/21/21 shots,
    layered-plus-fault models 1500-6000 m/s."""
    rng = np.random.default_rng(seed)
    fams = [(501, 171, 20.0, 5000, 0.0015, 22), (481, 121, 25.0, 4000, 0.0020, 21),
            (301, 101, 15.0, 3000, 0.0012, 21)]
    out = []
    for nx, nz, dx, nt, dt, ns in fams:
        faults = [(rng.uniform(0.3 * nx, 0.45 * nx), rng.uniform(4, 10))]
        vp = layered_faulted(nx, nz, 1500.0, 6000.0, 10, rng, faults=faults)
        init = gaussian_filter(vp.astype(np.float64), 10, mode="nearest").astype(np.float32)
        ix = np.linspace(0, nx - 1, ns).astype(int)
        out.append(Workload(f"{nx}x{nz}", nx, nz, dx, dt, nt, 40, 5.0, vp, init,
                            shots_at(ix, 2, dx), surface_receivers(nx, dx, ns)))
    return out


def config5(seed=5):
    """801x267, dx 12.5 m, 50-cell PML, nt 6000, dt 1 ms, 81 shots, 801 receivers; folded
    layers 2000-6000 m/s from a random-phase fold field."""
    rng = np.random.default_rng(seed)
    nx, nz, dx = 801, 267, 12.5
    x = np.arange(nx)[:, None]
    z = np.arange(nz)[None, :].astype(np.float64)
    fold = np.zeros((nx, 1))
    for k in range(1, 5):
        fold += rng.uniform(2, 10) / k * np.sin(2 * np.pi * k * x / nx + rng.uniform(0, 2 * np.pi))
    zz = z + fold * (z / nz)
    nl = 14
    edges = np.linspace(0, nz, nl + 1)[1:-1]
    idx = np.searchsorted(edges, zz)
    vp = np.linspace(2000, 6000, nl)[np.clip(idx, 0, nl - 1)].astype(np.float32)
    init = gaussian_filter(vp.astype(np.float64), 15, mode="nearest").astype(np.float32)
    src = shots_at(np.arange(0, 801, 10), 2, dx)
    return Workload("801x267 folded layers, 81 shots", nx, nz, dx, 0.001, 6000, 50, 5.0, vp, init,
                    src, surface_receivers(nx, dx, len(src)))


def small(seed=0, nx=64, nz=40, npml=10, nt=400, ns=3, dx=10.0, dt=0.001):
    """Small seeded heterogeneous case for parity tests: layers, a fault and random
    smooth texture; surface receivers plus two buried ones; sources at varied depth."""
    rng = np.random.default_rng(seed)
    vp = layered_faulted(nx, nz, 1500.0, 3200.0, 5, rng, water_rows=3,
                         faults=[(nx * 0.4, 3.0)])
    vp = vp + (gaussian_filter(rng.standard_normal((nx, nz)), 2) * 200).astype(np.float32)
    vp = np.clip(vp, 1400, 3400).astype(np.float32)
    init = gaussian_filter(vp.astype(np.float64), 4, mode="nearest").astype(np.float32)
    sx = np.linspace(3, nx - 4, ns)
    sz = np.array([0, 2, 7, 1, 3, 12][:ns] + [2] * max(0, ns - 6))[:ns]
    sz = np.minimum(sz, nz - 4)
    src = np.stack([sx * dx, sz * dx], 1).astype(np.float32)
    rec = np.zeros((nx + 2, 2), np.float32)
    rec[:nx, 0] = np.arange(nx) * dx
    rec[nx] = [(nx // 3) * dx + 3.0, min(15, nz - 3) * dx + 4.0]
    rec[nx + 1] = [(2 * nx // 3) * dx, 1 * dx]
    return Workload("small", nx, nz, dx, dt, nt, npml, 12.0, vp, init, src,
                    np.broadcast_to(rec, (ns,) + rec.shape).copy())
