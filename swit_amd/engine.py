"""Device plan: owns one libswit_fd plan and moves torch device buffers through it.

PyTorch provides device memory and the stream; all solver arithmetic runs in the
HIP kernels of libswit_fd.so.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native


def _require_gpu():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("swit_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def pack_receivers(rec_xz, ns):
    """(ns, ng, 2) array or list of (ng_i, 2) arrays -> (ng[ns], rec[ns, ngmax, 2])."""
    if isinstance(rec_xz, (list, tuple)):
        recs = [np.asarray(r, dtype=np.float32).reshape(-1, 2) for r in rec_xz]
    else:
        a = np.asarray(rec_xz, dtype=np.float32)
        if a.ndim == 2:
            a = np.broadcast_to(a, (ns,) + a.shape)
        recs = [a[i].reshape(-1, 2) for i in range(ns)]
    if len(recs) != ns:
        raise ValueError(f"receiver list has {len(recs)} shots, expected {ns}")
    ng = np.array([r.shape[0] for r in recs], dtype=np.int32)
    ngmax = int(ng.max()) if ns else 0
    out = np.zeros((ns, max(ngmax, 1), 2), dtype=np.float32)
    for i, r in enumerate(recs):
        out[i, : r.shape[0]] = r
    return ng, out


class Plan:
    """One solver plan: grid, acquisition and device scratch for every shot.

    path: 'plain' (per-shot, per-step launches) or 'batched' (all shots per launch).
    strips: 'f32' | 'f16' | 'bf16' storage of the boundary strips, or 'none' for a plan
    that only models (observed data, line-search forwards, the exact adjoint): it holds
    no strips, and strip-saving forwards and adjoint() are refused.
    """

    def __init__(self, nx, nz, dx, dt, nt, npml, src_xz, rec_xz, free_surface=1,
                 path="batched", strips="f32", exact=False):
        torch = _require_gpu()
        self._torch = torch
        self._lib = _native.load()
        src_xz = np.ascontiguousarray(np.asarray(src_xz, dtype=np.float32).reshape(-1, 2))
        self.ns = src_xz.shape[0]
        ng, rec = pack_receivers(rec_xz, self.ns)
        self.ng = ng
        self.ngmax = rec.shape[1]
        self.nx, self.nz, self.nt, self.npml = int(nx), int(nz), int(nt), int(npml)
        self.dx, self.dt = float(np.float32(dx)), float(np.float32(dt))
        flags = {"plain": _native.SWIT_PATH_PLAIN, "batched": _native.SWIT_PATH_BATCHED}[path]
        flags |= {"f32": 0, "f16": _native.SWIT_STRIPS_F16, "bf16": _native.SWIT_STRIPS_BF16,
                  "none": _native.SWIT_PLAN_NO_STRIPS}[strips]
        if exact:
            # exact=True keeps the whole divergence history; exact='checkpoint' keeps
            # states every ~sqrt(nt) steps and recomputes segments in the adjoint
            flags |= _native.SWIT_PLAN_EXACT_CKPT if exact == "checkpoint" else _native.SWIT_PLAN_EXACT
        self.path, self.strips = path, strips
        g = _native.SwitGrid(self.nx, self.nz, self.npml, self.nt, int(free_surface),
                             self.dx, self.dt)
        self._keep = (np.ascontiguousarray(ng), src_xz, np.ascontiguousarray(rec))
        sh = _native.SwitShots(
            self.ns, self.ngmax,
            self._keep[0].ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            self._keep[1].ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            self._keep[2].ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        h = ctypes.c_void_p()
        torch.cuda.init()
        _native.check(self._lib.swit_plan_create(ctypes.byref(g), ctypes.byref(sh), flags,
                                                 ctypes.byref(h)))
        self._h = h
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.bwd_steps = 2          # backward time steps per launch (batched path)

    # -- helpers ----------------------------------------------------------------
    def _stream(self):
        return ctypes.c_void_p(self._torch.cuda.current_stream(self.device).cuda_stream)

    def _dev(self, x, dtype, shape):
        torch = self._torch
        t = torch.as_tensor(x, dtype=dtype, device=self.device).contiguous()
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"expected shape {tuple(shape)}, got {tuple(t.shape)}")
        return t

    def _out(self, out):
        """A seismogram output: a new (ns, ngmax, nt) float32 tensor, or `out` checked to
        be exactly that (contiguous, on the plan's device), since the kernels write
        ns * ngmax * nt floats through its pointer."""
        torch = self._torch
        shape = (self.ns, self.ngmax, self.nt)
        if out is None:
            return torch.empty(shape, dtype=torch.float32, device=self.device)
        if (not torch.is_tensor(out) or out.dtype != torch.float32 or tuple(out.shape) != shape
                or out.device != self.device or not out.is_contiguous()):
            raise ValueError(f"out must be a contiguous float32 tensor of shape {shape} on {self.device}")
        return out

    @property
    def scratch_bytes(self):
        return int(self._lib.swit_plan_scratch_bytes(self._h))

    def exact_generation(self):
        """Generation of the divergence history the plan holds (raised by every
        forward_exact; 0 when none is valid)."""
        return int(self._lib.swit_plan_exact_generation(self._h))

    def set_option(self, option, value):
        """option: 'graph' | 'profile' | 'fwd_group' | 'fwd_colgroup' | 'bwd_colgroup' |
        'bwd_steps' | 'bwd_tile_x' | 'bwd_balance' | 'exact_fused' | 'streams' (batched
        path).  'streams' 2 runs the two halves of the shots on two streams; 'bwd_balance'
        splits the one-step backward kernel's work over the XCDs by estimated cost (1, the
        default) or by item count (0).  None of them changes results."""
        code = {"graph": _native.SWIT_OPT_GRAPH, "profile": _native.SWIT_OPT_PROFILE,
                "fwd_group": _native.SWIT_OPT_FWD_GROUP, "fwd_colgroup": _native.SWIT_OPT_FWD_COLGROUP,
                "bwd_colgroup": _native.SWIT_OPT_BWD_COLGROUP, "bwd_steps": _native.SWIT_OPT_BWD_STEPS,
                "bwd_tile_x": _native.SWIT_OPT_BWD_TILE_X, "exact_fused": _native.SWIT_OPT_EXACT_FUSED,
                "streams": _native.SWIT_OPT_STREAMS, "bwd_balance": _native.SWIT_OPT_BWD_BALANCE}[option]
        _native.check(self._lib.swit_plan_set_option(self._h, code, int(value)))
        if option == "bwd_steps":
            self.bwd_steps = int(value)

    def autotune(self, steps=20):
        """Time the batched step kernels for every candidate forward shots-per-workgroup
        and column-group order (backward: steps per launch, the one-step kernel's tile
        width and column-group order) and keep the fastest.  Invalidates saved strips.
        Returns {'fwd_group', 'fwd_colgroup', 'bwd_colgroup', 'bwd_steps', 'bwd_tile_x',
        'bwd_balance', 'fwd_us', 'bwd_us'} (times per time step)."""
        choice = (ctypes.c_int * 6)()
        us = (ctypes.c_double * 2)()
        _native.check(self._lib.swit_plan_autotune(self._h, int(steps), choice, us, self._stream()))
        self.bwd_steps = choice[3]
        return {"fwd_group": choice[0], "fwd_colgroup": choice[1], "bwd_colgroup": choice[2],
                "bwd_steps": choice[3], "bwd_tile_x": choice[4], "bwd_balance": choice[5],
                "fwd_us": round(us[0], 2),
                "bwd_us": round(us[1], 2)}

    def kernel_time(self, which):
        """(total_ms, launches) of the forward (which='forward') or backward step kernel
        accumulated while profiling was on."""
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        _native.check(self._lib.swit_plan_kernel_time(self._h, {"forward": 0, "backward": 1}[which],
                                                      ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    # -- jobs ---------------------------------------------------------------------
    def set_model(self, vp, rho):
        torch = self._torch
        self._vp = self._dev(vp, torch.float32, (self.nx, self.nz))
        self._rho = self._dev(rho, torch.float32, (self.nx, self.nz))
        _native.check(self._lib.swit_plan_set_model(self._h, ctypes.c_void_p(self._vp.data_ptr()),
                                                    ctypes.c_void_p(self._rho.data_ptr()),
                                                    self._stream()))

    def forward(self, src, save_strips=True, out=None):
        """src: (ns, nt) integrated source traces -> seismograms (ns, ngmax, nt).
        save_strips: 1/True save strips for adjoint(), 0/False discard saved strips,
        2 save nothing but keep earlier saved strips."""
        torch = self._torch
        save_strips = int(save_strips)
        src = self._dev(src, torch.float32, (self.ns, self.nt))
        out = self._out(out)
        _native.check(self._lib.swit_forward(self._h, ctypes.c_void_p(src.data_ptr()),
                                             ctypes.c_void_p(out.data_ptr()), save_strips,
                                             self._stream()))
        self._src = src
        return out

    def adjoint(self, src, adjsrc):
        """-> (grad, illum_forw, illum_back), each (nx, nz) float64 on the device."""
        torch = self._torch
        src = self._dev(src, torch.float32, (self.ns, self.nt))
        adjsrc = self._dev(adjsrc, torch.float32, (self.ns, self.ngmax, self.nt))
        outs = [torch.empty((self.nx, self.nz), dtype=torch.float64, device=self.device) for _ in range(3)]
        _native.check(self._lib.swit_adjoint(self._h, ctypes.c_void_p(src.data_ptr()),
                                             ctypes.c_void_p(adjsrc.data_ptr()),
                                             *[ctypes.c_void_p(o.data_ptr()) for o in outs],
                                             self._stream()))
        return tuple(outs)

    def forward_exact(self, src, out=None):
        """Forward (same seismograms as forward()) that also stores the divergence
        history for adjoint_exact().  Needs a plan created with exact=True."""
        torch = self._torch
        src = self._dev(src, torch.float32, (self.ns, self.nt))
        out = self._out(out)
        _native.check(self._lib.swit_forward_exact(self._h, ctypes.c_void_p(src.data_ptr()),
                                                   ctypes.c_void_p(out.data_ptr()), self._stream()))
        self._src = src
        return out

    def adjoint_exact(self, gseis):
        """dJ/dseis (ns, ngmax, nt) -> dJ/dvp (nx, nz) float64: the exact discrete adjoint
        of forward_exact's scheme (density and damping held fixed)."""
        torch = self._torch
        g = self._dev(gseis, torch.float32, (self.ns, self.ngmax, self.nt))
        out = torch.empty((self.nx, self.nz), dtype=torch.float64, device=self.device)
        _native.check(self._lib.swit_adjoint_exact(self._h, ctypes.c_void_p(g.data_ptr()),
                                                   ctypes.c_void_p(out.data_ptr()), self._stream()))
        return out

    def saved_strips(self):
        """Saved boundary strips decoded to FP32: (ns, nt, 9nx+6nz) on the device."""
        torch = self._torch
        out = torch.empty((self.ns, self.nt, 9 * self.nx + 6 * self.nz), dtype=torch.float32, device=self.device)
        _native.check(self._lib.swit_plan_strips(self._h, ctypes.c_void_p(out.data_ptr()), self._stream()))
        return out

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.swit_plan_destroy(h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PlanGroup:
    """Several batched plans (uneven grids, shot counts, record lengths) run as one job:
    each forward / adjoint is a single hipGraph launch with every plan's time loop on
    its own branch.  Results equal those of each plan run alone."""

    def __init__(self, plans):
        self.plans = list(plans)
        if not self.plans:
            raise ValueError("need at least one plan")
        self._lib = self.plans[0]._lib
        self._torch = self.plans[0]._torch
        arr = (ctypes.c_void_p * len(self.plans))(*[p._h.value for p in self.plans])
        h = ctypes.c_void_p()
        _native.check(self._lib.swit_group_create(arr, len(self.plans), ctypes.byref(h)))
        self._h = h

    @staticmethod
    def _ptrs(tensors):
        return (ctypes.c_void_p * len(tensors))(*[t.data_ptr() for t in tensors])

    def forward(self, srcs, save_strips=True):
        """srcs: one (ns_i, nt_i) source array per plan -> list of (ns_i, ngmax_i, nt_i)."""
        torch = self._torch
        ins = [p._dev(s, torch.float32, (p.ns, p.nt)) for p, s in zip(self.plans, srcs, strict=True)]
        outs = [torch.empty((p.ns, p.ngmax, p.nt), dtype=torch.float32, device=p.device) for p in self.plans]
        _native.check(self._lib.swit_group_forward(self._h, self._ptrs(ins), self._ptrs(outs), int(save_strips),
                                                   self.plans[0]._stream()))
        for p, s in zip(self.plans, ins):
            p._src = s
        return outs

    def adjoint(self, srcs, adjsrcs):
        """-> list of (grad, illum_forw, illum_back) per plan, each (nx_i, nz_i) float64."""
        torch = self._torch
        ins = [p._dev(s, torch.float32, (p.ns, p.nt)) for p, s in zip(self.plans, srcs, strict=True)]
        adj = [p._dev(a, torch.float32, (p.ns, p.ngmax, p.nt)) for p, a in zip(self.plans, adjsrcs, strict=True)]
        outs = [[torch.empty((p.nx, p.nz), dtype=torch.float64, device=p.device) for _ in range(3)]
                for p in self.plans]
        _native.check(self._lib.swit_group_adjoint(
            self._h, self._ptrs(ins), self._ptrs(adj), self._ptrs([o[0] for o in outs]),
            self._ptrs([o[1] for o in outs]), self._ptrs([o[2] for o in outs]), self.plans[0]._stream()))
        return [tuple(o) for o in outs]

    def gradient(self, srcs, obs, kind):
        """Every plan's whole gradient in one graph launch (swit_group_gradient): forward
        with strips, misfit `kind` and adjoint source against obs[i], adjoint.  Returns
        (misfits, seismograms, (grad, illum_forw, illum_back) per plan); each plan's
        results equal forward, engine.misfit and adjoint run one after the other.  The
        returned tensors are the group's own and are overwritten by the next call."""
        torch = self._torch
        code = _MISFIT_KIND[kind.lower()]
        ins = [p._dev(s, torch.float32, (p.ns, p.nt)) for p, s in zip(self.plans, srcs, strict=True)]
        ob = [p._dev(o, torch.float32, (p.ns, p.ngmax, p.nt)) for p, o in zip(self.plans, obs, strict=True)]
        # the graph holds the arrays' addresses: the outputs and work arrays are kept and
        # reused from call to call (so the graph is captured once), and the observed
        # gathers are held until the next call
        if getattr(self, "_bufs", None) is None:
            seis = [torch.empty((p.ns, p.ngmax, p.nt), dtype=torch.float32, device=p.device) for p in self.plans]
            self._bufs = (seis, [torch.empty_like(x) for x in seis],
                          [torch.empty(p.ns * p.ngmax, dtype=torch.float64, device=p.device) for p in self.plans],
                          [torch.empty(1, dtype=torch.float64, device=p.device) for p in self.plans],
                          [[torch.empty((p.nx, p.nz), dtype=torch.float64, device=p.device) for _ in range(3)]
                           for p in self.plans])
        seis, adj, work, J, outs = self._bufs
        self._obs_held = ob
        _native.check(self._lib.swit_group_gradient(
            self._h, code, self._ptrs(ins), self._ptrs(ob), self._ptrs(seis), self._ptrs(adj), self._ptrs(work),
            self._ptrs(J), self._ptrs([o[0] for o in outs]), self._ptrs([o[1] for o in outs]),
            self._ptrs([o[2] for o in outs]), self.plans[0]._stream()))
        for p, s in zip(self.plans, ins):
            p._src = s
        return [j[0] for j in J], seis, [tuple(o) for o in outs]

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.swit_group_destroy(h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Lanes:
    """The shots of one job split into `lanes` batched plans (contiguous shot ranges),
    each running its whole job on a stream of its own.  The lanes' step kernels run side
    by side and fill each other's idle workgroup slots: at 481x141x4000 with 49 shots a
    gradient on two lanes runs in 473-480 ms against 511-514 ms for one plan (three or
    four lanes, or a second lane started after the first's forward, were slower).  Gradients
    and illuminations are the lanes' FP64 sums added in lane order, so they equal the
    one-plan result to FP64 rounding.  Plan keyword arguments (strips, ...) apply to
    every lane."""

    def __init__(self, nx, nz, dx, dt, nt, npml, src_xz, rec_xz, lanes=2, **plan_kw):
        torch = _require_gpu()
        self._torch = torch
        src_xz = np.asarray(src_xz, dtype=np.float32).reshape(-1, 2)
        ns = src_xz.shape[0]
        ng, rec = pack_receivers(rec_xz, ns)
        lanes = max(1, min(int(lanes), ns))
        b = np.linspace(0, ns, lanes + 1).round().astype(int)
        self.ranges = [(int(b[i]), int(b[i + 1])) for i in range(lanes)]
        self.plans = [Plan(nx, nz, dx, dt, nt, npml, src_xz[a:e], [rec[i, :ng[i]] for i in range(a, e)], **plan_kw)
                      for a, e in self.ranges]
        self.streams = [torch.cuda.Stream() for _ in self.ranges]
        self.ns, self.nx, self.nz, self.nt = ns, int(nx), int(nz), int(nt)
        self.ngmax = max(p.ngmax for p in self.plans)
        self.device = self.plans[0].device

    def _run(self, fn):
        """fn(i, plan, shot range) on lane i's stream, every lane; returns the results
        after the caller's stream has waited for all lanes.  Device tensors in the
        results are marked as used by the caller's stream, so their memory is not
        handed out again on a lane's stream before the caller is done with them."""
        torch = self._torch
        cur = torch.cuda.current_stream(self.device)
        out = []
        for i, (p, r, st) in enumerate(zip(self.plans, self.ranges, self.streams)):
            st.wait_stream(cur)
            with torch.cuda.stream(st):
                out.append(fn(i, p, r))
        for st in self.streams:
            cur.wait_stream(st)
        for res in out:
            for t in (res if isinstance(res, (tuple, list)) else (res,)):
                if torch.is_tensor(t) and t.is_cuda:
                    t.record_stream(cur)
        return out

    def set_model(self, vp, rho):
        torch = self._torch
        vp = self.plans[0]._dev(vp, torch.float32, (self.nx, self.nz))
        rho = self.plans[0]._dev(rho, torch.float32, (self.nx, self.nz))
        self._run(lambda i, p, r: p.set_model(vp, rho))

    def autotune(self, steps=20):
        """Each lane's plan tuned alone, one after the other (the timings must not share
        the GPU)."""
        return [p.autotune(steps) for p in self.plans]

    def forward(self, src, save_strips=True):
        """src (ns, nt) -> seismograms (ns, ngmax, nt) of every shot."""
        torch = self._torch
        src = self.plans[0]._dev(src, torch.float32, (self.ns, self.nt))
        outs = self._run(lambda i, p, r: p.forward(src[r[0]:r[1]], save_strips=save_strips))
        seis = torch.zeros((self.ns, self.ngmax, self.nt), dtype=torch.float32, device=self.device)
        for (a, e), o in zip(self.ranges, outs):
            seis[a:e, :o.shape[1]] = o
        return seis

    def gradient(self, src, obs, kind, dt, process=None):
        """Per lane: forward with strips, process(lane, shot range, gathers) if given (in
        place), misfit `kind` and adjoint source against obs[range] (engine.misfit), and
        the adjoint.  Returns (misfit, grad, illum_forw, illum_back), FP64 sums over the
        lanes in lane order."""
        torch = self._torch
        src = self.plans[0]._dev(src, torch.float32, (self.ns, self.nt))

        def lane(i, p, r):
            syn = p.forward(src[r[0]:r[1]], save_strips=True)
            if process is not None:
                process(i, r, syn)
            J, adj, _ = misfit(kind, obs[r[0]:r[1], :p.ngmax], syn, dt)
            return (J,) + p.adjoint(src[r[0]:r[1]], adj)
        outs = self._run(lane)
        tot = [o for o in outs[0]]
        for o in outs[1:]:
            tot = [a + b for a, b in zip(tot, o)]
        return tuple(tot)

    def misfit(self, src, obs, kind, dt, process=None):
        """Forward without strips, optional processing, misfit; the lanes' sum."""
        torch = self._torch
        src = self.plans[0]._dev(src, torch.float32, (self.ns, self.nt))

        def lane(i, p, r):
            syn = p.forward(src[r[0]:r[1]], save_strips=False)
            if process is not None:
                process(i, r, syn)
            return misfit(kind, obs[r[0]:r[1], :p.ngmax], syn, dt)[0]
        outs = self._run(lane)
        J = outs[0]
        for x in outs[1:]:
            J = J + x
        return J

    def close(self):
        for p in self.plans:
            p.close()


_MISFIT_KIND = {"waveform": _native.SWIT_MISFIT_WAVEFORM,
                "globalcorrelation": _native.SWIT_MISFIT_GLOBAL_CORRELATION,
                "envelope": _native.SWIT_MISFIT_ENVELOPE,
                "traveltime": _native.SWIT_MISFIT_TRAVELTIME}


def misfit(kind, obs, syn, dt):
    """Misfit and adjoint source of device gathers (toolbox/misfit.py rules).

    kind: 'Waveform', 'Globalcorrelation', 'Envelope' or 'Traveltime'; obs, syn:
    float32 device tensors of one shape (..., nt).  Returns (misfit as a float64
    device scalar, adjoint source of syn's shape, per-trace misfits)."""
    torch = _require_gpu()
    lib = _native.load()
    code = _MISFIT_KIND[kind.lower()]
    if obs.shape != syn.shape:
        raise ValueError("obs and syn are not consistent.")
    o = obs.to(torch.float32).contiguous()
    s = syn.to(torch.float32).contiguous()
    nt = int(o.shape[-1])
    ntr = int(o.numel() // nt)
    adj = torch.empty_like(s)
    work = torch.empty(ntr, dtype=torch.float64, device=s.device)
    out = torch.empty(1, dtype=torch.float64, device=s.device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(s.device).cuda_stream)
    _native.check(lib.swit_misfit(code, ntr, nt, float(np.float32(dt)), ctypes.c_void_p(o.data_ptr()),
                                  ctypes.c_void_p(s.data_ptr()), ctypes.c_void_p(adj.data_ptr()),
                                  ctypes.c_void_p(work.data_ptr()), ctypes.c_void_p(out.data_ptr()), stream))
    return out[0], adj, work.view(o.shape[:-1])


def grad_precond(grad, forw, back, vpmax, marine, grad_mute=0, grad_thred=0.0, grad_smooth=0, mask=None):
    """grad_precond (toolbox/postprocess.py:20-86) on the device: grad, forw, back (nx, nz)
    float32 (as the solver's .bin files hold them; host or device), mask (nx, nz) or None.
    Returns the preconditioned gradient as a float64 (nx, nz) device tensor."""
    torch = _require_gpu()
    lib = _native.load()
    dev = torch.device("cuda", torch.cuda.current_device())

    def f32(a):
        return torch.as_tensor(np.asarray(a) if not torch.is_tensor(a) else a).to(dev, torch.float32).contiguous()
    g, fw, bk = f32(grad), f32(forw), f32(back)
    if g.dim() != 2 or fw.shape != g.shape or bk.shape != g.shape:
        raise ValueError("grad, forw and back must be (nx, nz) arrays of one shape")
    nx, nz = int(g.shape[0]), int(g.shape[1])
    m = None
    if mask is not None:
        m = torch.as_tensor(np.asarray(mask) if not torch.is_tensor(mask) else mask).to(dev, torch.float64)
        if tuple(m.shape) != (nx, nz):
            raise ValueError('Wrong size of grad mask: the size of the mask should be identical to the size of vp model')
        m = m.contiguous()
    prm = _native.SwitPrecond(nx, nz, int(bool(marine)), int(grad_mute), float(grad_thred), int(grad_smooth),
                              float(vpmax))
    nbytes = ctypes.c_size_t()
    _native.check(lib.swit_grad_precond_work(ctypes.byref(prm), ctypes.byref(nbytes)))
    work = torch.empty(int(nbytes.value), dtype=torch.uint8, device=dev)
    out = torch.empty((nx, nz), dtype=torch.float64, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _native.check(lib.swit_grad_precond(ctypes.byref(prm), ctypes.c_void_p(g.data_ptr()),
                                        ctypes.c_void_p(fw.data_ptr()), ctypes.c_void_p(bk.data_ptr()),
                                        ctypes.c_void_p(m.data_ptr()) if m is not None else None,
                                        ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(work.data_ptr()),
                                        nbytes.value, stream))
    return out


def process_params(optim, ngathers, ntr, nt, dt):
    """swit_process parameters for `optim`'s processing (toolbox/preprocess.py:47-200)
    and the tables it takes: (params, mute window [100] or None, taper [nt] or None)."""
    from scipy.signal import lfilter_zi
    from . import preprocess
    p = _native.SwitProcessParams()
    p.ngathers, p.ntr, p.nt, p.dt = int(ngathers), int(ntr), int(nt), float(dt)
    p.mute_offset_short = int(bool(optim.mute_offset_short))
    p.mute_offset_long = int(bool(optim.mute_offset_long))
    p.mute_offset_short_dis = float(optim.mute_offset_short_dis)
    p.mute_offset_long_dis = float(optim.mute_offset_long_dis)
    p.mute_late_arrival = int(bool(optim.mute_late_arrival))
    p.mute_late_window = float(optim.mute_late_window)
    win = taper = None
    if p.mute_late_arrival:
        L = preprocess.MUTE_LENGTH
        win = np.sin(np.linspace(0, np.pi, 2 * L))[0:L]
    ba = preprocess.filter_design(optim, dt)
    if ba is not None:
        b, a = ba
        if len(b) != len(a) or len(b) > _native.SWIT_PROCESS_MAX_TAPS or a[0] != 1.0:
            raise ValueError("filter coefficients must have len(b) = len(a) <= %d and a[0] = 1"
                             % _native.SWIT_PROCESS_MAX_TAPS)
        zi = lfilter_zi(b, a)
        p.filter_taps = len(b)
        for k in range(len(b)):
            p.b[k], p.a[k] = float(b[k]), float(a[k])
        for k in range(len(zi)):
            p.zi[k] = float(zi[k])
        p.detrend_taper = 1
        taper = preprocess.obspy_taper(nt)
    nz = optim.normalize
    p.norm_trace = (_native.SWIT_NORM["l1"] if 'L1-Trace' in nz else _native.SWIT_NORM["l2"] if 'L2-Trace' in nz
                    else _native.SWIT_NORM["max"] if 'Max-Trace' in nz else 0)
    p.norm_event = (_native.SWIT_NORM["l1"] if 'L1-Event' in nz else _native.SWIT_NORM["l2"] if 'L2-Event' in nz
                    else 0)
    return p, win, taper


def process_gathers(optim, data, dt, offsets=None, params=None, first_break_ms=None):
    """process_workflow's steps (toolbox/preprocess.py:47-200) on device gathers, in
    place: data float32 device tensor [ngathers][ntr][nt] (or [ntr][nt]); offsets
    [ngathers][ntr] group x - source x (required with an offset mute).  params: the
    (params, mute window, taper) of process_params, to override the derived ones.
    first_break_ms: [ngathers][ntr] first-break times in ms from the traces' SU headers
    (the reference's use_first_break_in_su_header) to centre the window mute on, or None
    to pick the first break from the data.  Returns data."""
    torch = _require_gpu()
    lib = _native.load()
    if data.dtype != torch.float32 or not data.is_cuda or not data.is_contiguous():
        raise ValueError("data must be a contiguous float32 device tensor")
    shp = data.shape if data.dim() == 3 else (1,) + tuple(data.shape)
    ng, ntr, nt = (int(x) for x in shp)
    p, win, taper = params if params is not None else process_params(optim, ng, ntr, nt, dt)
    dev = data.device
    off = None
    if p.mute_offset_short or p.mute_offset_long:
        if offsets is None:
            raise ValueError("offset mutes need the trace offsets")
        off = torch.as_tensor(np.asarray(offsets, np.float32).reshape(-1), device=dev)
        if off.numel() != ng * ntr:
            raise ValueError("offsets must hold one value per trace")
    w = torch.as_tensor(win, dtype=torch.float64, device=dev) if win is not None else None
    pk = None
    if first_break_ms is not None and p.mute_late_arrival:
        from .preprocess import header_picks
        fb = np.asarray(first_break_ms).reshape(-1)
        if fb.size != ng * ntr:
            raise ValueError("first_break_ms must hold one value per trace")
        pk = torch.as_tensor(header_picks(fb, optim.mute_late_window, dt).astype(np.int32), device=dev)
    tp = torch.as_tensor(taper, dtype=torch.float64, device=dev) if taper is not None else None
    nbytes = ctypes.c_size_t()
    _native.check(lib.swit_process_work(ctypes.byref(p), ctypes.byref(nbytes)))
    work = torch.empty(int(nbytes.value), dtype=torch.uint8, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)

    def ptr(t):
        return ctypes.c_void_p(t.data_ptr()) if t is not None else None
    _native.check(lib.swit_process(ctypes.byref(p), ptr(off), ptr(w), ptr(tp), ptr(pk), ptr(data), ptr(work),
                                   nbytes.value, stream))
    return data
