"""ORACLE -- test infrastructure, not product code.

ctypes front end of the C restatement in ``fd2d_oracle.c`` (FP32 and FP64 builds of
the reference scheme).  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` may import this module, and only as the checker.

Conventions follow the reference's file formats (fd2dmpi/io.f90, toolbox/solver.py):
  * models ``vp``, ``rho``: numpy (nx, nz), iz fastest;
  * ``src_xz``: (ns, 2) metres; ``rec_xz``: (ns, ng, 2) metres (float32 after the
    reference's ``%10.1f`` coordinate file, toolbox/solver.py:178-179);
  * ``src``: (ns, nt) the *integrated* source trace the solver reads from
    ``src%d.bin`` (toolbox/solver.py:47,112-113);
  * seismograms (ns, ng, nt); strips (ns, nt, 9nx+6nz) with row it-1 = step it;
  * per-shot kernels (ns, nx, nz).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle_fd2d.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, no reference sources involved)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _vp(dtype):
    return ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for suf in ("f32", "f64"):
            rt = ctypes.c_float if suf == "f32" else ctypes.c_double
            getattr(L, f"orc_pad_{suf}").argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 2
            getattr(L, f"orc_pad_{suf}").restype = None
            getattr(L, f"orc_vmin_{suf}").argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p]
            getattr(L, f"orc_vmin_{suf}").restype = rt
            getattr(L, f"orc_setup_pml_{suf}").argtypes = [ctypes.c_int] * 3 + [rt, rt, ctypes.c_void_p]
            getattr(L, f"orc_setup_pml_{suf}").restype = None
            getattr(L, f"orc_forward_shot_{suf}").argtypes = (
                [ctypes.c_int] * 5 + [ctypes.c_float] * 2 + [ctypes.c_void_p] * 3
                + [ctypes.c_float] * 2 + [ctypes.c_int] + [ctypes.c_void_p] * 2
                + [ctypes.c_void_p] * 7)
            getattr(L, f"orc_forward_shot_{suf}").restype = ctypes.c_int
            getattr(L, f"orc_adjoint_shot_{suf}").argtypes = (
                [ctypes.c_int] * 5 + [ctypes.c_float] * 2 + [ctypes.c_void_p] * 3
                + [ctypes.c_float] * 2 + [ctypes.c_int] + [ctypes.c_void_p] * 2
                + [ctypes.c_void_p] * 11)
            getattr(L, f"orc_adjoint_shot_{suf}").restype = ctypes.c_int
            getattr(L, f"orc_scale_gradient_{suf}").argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 2
            getattr(L, f"orc_scale_gradient_{suf}").restype = None
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _dt(precision):
    if precision not in ("f32", "f64"):
        raise ValueError(precision)
    return np.float32 if precision == "f32" else np.float64


class Model:
    """Padded model + damping, as set up by solver.f90:53-76."""

    def __init__(self, vp, rho, dx, npml, precision="f32", damp_override=None, exact_input=False):
        """exact_input: with precision 'f64', keep vp / rho as given instead of rounding
        them through float32 (the reference reads float32 files); used to difference
        models that differ by less than a float32 step."""
        L = lib()
        self.precision = precision
        T = _dt(precision)
        keep = exact_input and precision == "f64"
        vp = np.ascontiguousarray(np.asarray(vp, dtype=np.float64 if keep else np.float32).astype(T))
        rho = np.ascontiguousarray(np.asarray(rho, dtype=np.float64 if keep else np.float32).astype(T))
        self.nx, self.nz = vp.shape
        self.npml = int(npml)
        self.dx = float(np.float32(dx))
        nxp, nzp = self.nx + 2 * npml, self.nz + 2 * npml
        self.c = np.zeros((nxp, nzp), T)
        self.den = np.zeros((nxp, nzp), T)
        getattr(L, f"orc_pad_{precision}")(self.nx, self.nz, npml, _ptr(vp), _ptr(self.c))
        getattr(L, f"orc_pad_{precision}")(self.nx, self.nz, npml, _ptr(rho), _ptr(self.den))
        self.vmin = getattr(L, f"orc_vmin_{precision}")(self.nx, self.nz, npml, _ptr(self.c))
        self.damp = np.zeros((nxp, nzp), T)
        getattr(L, f"orc_setup_pml_{precision}")(self.nx, self.nz, npml, self.dx, self.vmin,
                                                 _ptr(self.damp))
        if damp_override is not None:
            self.damp[...] = damp_override


def _geom(src_xz, rec_xz, ns):
    src_xz = np.asarray(src_xz, dtype=np.float32).reshape(ns, 2)
    if isinstance(rec_xz, (list, tuple)):
        recs = [np.ascontiguousarray(np.asarray(r, dtype=np.float32).reshape(-1, 2)) for r in rec_xz]
    else:
        r = np.asarray(rec_xz, dtype=np.float32)
        recs = [np.ascontiguousarray(r[i].reshape(-1, 2)) for i in range(ns)]
    return src_xz, recs


def forward(model: Model, nt, dt, src_xz, rec_xz, src, free_surface=1, save_strips=True,
            psnap=False, threads=None):
    """Forward job for every shot (fdcore.f90:30-198, one shot per call)."""
    L = lib()
    prec = model.precision
    T = _dt(prec)
    src = np.ascontiguousarray(np.asarray(src, dtype=np.float32).astype(T))
    ns = src.shape[0]
    src_xz, recs = _geom(src_xz, rec_xz, ns)
    nx, nz = model.nx, model.nz
    S = 9 * nx + 6 * nz
    ngmax = max(r.shape[0] for r in recs)
    out = {
        "seis": np.zeros((ns, ngmax, nt), T),
        "boundary": np.zeros((ns, nt, S), T) if save_strips else None,
        "p_end": np.zeros((ns, nx, nz), T) if save_strips else None,
        "u_end": np.zeros((ns, nx, nz), T) if save_strips else None,
        "w_end": np.zeros((ns, nx, nz), T) if save_strips else None,
        "psnap": np.zeros((ns, nt, nx, nz), T) if psnap else None,
    }
    fn = getattr(L, f"orc_forward_shot_{prec}")

    def run(i):
        ng = recs[i].shape[0]
        xg = np.ascontiguousarray(recs[i][:, 0])
        zg = np.ascontiguousarray(recs[i][:, 1])
        csg = np.zeros((ng, nt), T)
        b = out["boundary"][i] if save_strips else None
        pe = out["p_end"][i] if save_strips else None
        ue = out["u_end"][i] if save_strips else None
        we = out["w_end"][i] if save_strips else None
        ps = out["psnap"][i] if psnap else None
        rc = fn(nx, nz, model.npml, nt, free_surface, model.dx, float(np.float32(dt)),
                _ptr(model.c), _ptr(model.den), _ptr(model.damp),
                float(src_xz[i, 0]), float(src_xz[i, 1]), ng, _ptr(xg), _ptr(zg),
                _ptr(src[i]), _ptr(csg), _ptr(b), _ptr(pe), _ptr(ue), _ptr(we), _ptr(ps))
        if rc:
            raise MemoryError("oracle forward failed")
        out["seis"][i, :ng] = csg

    _map(run, ns, threads)
    return out


def adjoint(model: Model, nt, dt, src_xz, rec_xz, src, adj_src, fwd, free_surface=1,
            record=False, prec_snap=False, threads=None):
    """Adjoint job (solver.f90:120-281) for every shot; per-shot kernels plus the
    FP64 cross-shot sums and the -2/c scaling of solver.f90:270."""
    L = lib()
    prec = model.precision
    T = _dt(prec)
    src = np.ascontiguousarray(np.asarray(src, dtype=np.float32).astype(T))
    adj_src = np.asarray(adj_src, dtype=np.float32).astype(T)
    ns = src.shape[0]
    src_xz, recs = _geom(src_xz, rec_xz, ns)
    nx, nz = model.nx, model.nz
    dg = np.zeros((ns, nx, nz), T)
    dfw = np.zeros((ns, nx, nz), T)
    dbk = np.zeros((ns, nx, nz), T)
    rec = np.zeros(adj_src.shape, T) if record else None
    psn = np.zeros((ns, nt, nx, nz), T) if prec_snap else None
    fn = getattr(L, f"orc_adjoint_shot_{prec}")

    def run(i):
        ng = recs[i].shape[0]
        xg = np.ascontiguousarray(recs[i][:, 0])
        zg = np.ascontiguousarray(recs[i][:, 1])
        sa = np.ascontiguousarray(adj_src[i, :ng])
        b = np.ascontiguousarray(fwd["boundary"][i])
        pe = np.array(fwd["p_end"][i], copy=True)
        ue = np.array(fwd["u_end"][i], copy=True)
        we = np.array(fwd["w_end"][i], copy=True)
        r = np.zeros((ng, nt), T) if record else None
        q = psn[i] if prec_snap else None
        g = np.zeros((nx, nz), T)
        f = np.zeros((nx, nz), T)
        k = np.zeros((nx, nz), T)
        rc = fn(nx, nz, model.npml, nt, free_surface, model.dx, float(np.float32(dt)),
                _ptr(model.c), _ptr(model.den), _ptr(model.damp),
                float(src_xz[i, 0]), float(src_xz[i, 1]), ng, _ptr(xg), _ptr(zg),
                _ptr(src[i]), _ptr(sa), _ptr(b), _ptr(pe), _ptr(ue), _ptr(we),
                _ptr(g), _ptr(f), _ptr(k), _ptr(r), _ptr(q))
        if rc:
            raise MemoryError("oracle adjoint failed")
        dg[i], dfw[i], dbk[i] = g, f, k
        if record:
            rec[i, :ng] = r

    _map(run, ns, threads)
    npml = model.npml
    c_int = model.c[npml:npml + nx, npml:npml + nz].astype(np.float64)
    g = dg.astype(np.float64).sum(0)
    out = {
        "dg": dg, "dfw": dfw, "dbk": dbk,
        "grad": -2.0 * g / c_int,
        "illum_forw": dfw.astype(np.float64).sum(0),
        "illum_back": dbk.astype(np.float64).sum(0),
    }
    if record:
        out["adj_rec"] = rec
    if prec_snap:
        out["prec"] = psn
    return out


def _map(fn, n, threads):
    if threads is None:
        threads = min(n, os.cpu_count() or 1)
    if threads <= 1 or n <= 1:
        for i in range(n):
            fn(i)
        return
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(fn, range(n)))


def ricker(nt, dt, f0, amp=1.0):
    """toolbox/solver.py:183-197 (Ricker, t0 = 1.2/f0, t = k dt)."""
    t = np.linspace(0, dt * nt, num=nt, endpoint=False)
    t0 = 1.2 / f0
    temp = (np.pi * f0) ** 2
    return amp * (1 - 2 * temp * (t - t0) ** 2) * np.exp(-temp * (t - t0) ** 2)


def integrate_source(wavelet):
    """toolbox/solver.py:47: cumulative trapezoid with unit spacing, initial 0."""
    from scipy.integrate import cumulative_trapezoid
    return cumulative_trapezoid(wavelet, axis=-1, initial=0)
