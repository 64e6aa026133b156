"""The inversion loop of the reference toolbox (toolbox/inversion.py:30-482) with every
vector resident on the device: model, gradient, search direction, the L-BFGS S / Y
history and the seismograms stay in HBM, and one iteration runs forward (boundary
strips kept on the device), processing (swit_process), misfit and adjoint source
(swit_misfit), adjoint and gradient preconditioning (swit_grad_precond) without SU
or .bin files in between.

The loop keeps the file-based loop's call order and decisions (swit_amd.inversion,
itself inversion.py's): forward with saved strips, processing, misfit, adjoint,
NLCG (Polak-Ribiere, restart after 6, descent check) or L-BFGS (memory 5, Liu-Nocedal
scaling, descent check, restart on failure) with float32 S / Y, backtracking with
forward-only runs and the same step safeguards and clipping, then the decision.  It
writes the same outputs under homepath/outputs (gradient / direction / velocity per
iteration, misfit and line-search logs).  Inner products accumulate in FP64 on the
device; the reference takes float32 ones of its float32 history, so directions agree
to float32 rounding.
"""
from __future__ import annotations

import numpy as np

from . import inversion as _host
from . import preprocess, su


def _torch():
    import torch
    return torch


def _dot(a, b):
    """FP64 inner product of two vectors (a float)."""
    torch = _torch()
    return float(torch.dot(a.reshape(-1).to(torch.float64), b.reshape(-1).to(torch.float64)))


class DeviceProblem:
    """The FWI problem of a SWIT simulation on one GPU: batched plans of simu's grid and
    acquisition (`lanes` shot ranges on streams of their own, engine.Lanes), the
    integrated source (toolbox/solver.py:41, cumulative trapezoid), the processed
    observed gathers and optim's processing, misfit and preconditioning.

    obs: observed gathers (ns, ng, nt) as the solver writes them (host or device), or
    None to model them in simu_true."""

    def __init__(self, simu, optim, obs=None, simu_true=None, lanes=2):
        from .engine import _require_gpu
        torch = _require_gpu()
        from scipy.integrate import cumulative_trapezoid
        from . import solver
        from .engine import Lanes
        inp = solver.solver_inputs(simu)
        self.nx, self.nz, self.nt, self.dt = inp["nx"], inp["nz"], inp["nt"], inp["dt"]
        self.ns, self.ng = simu.source.n, simu.receiver.n
        self.optim = optim
        self.kind = optim.misfit_type
        if self.kind.lower() == 'rtm':
            raise ValueError("the device loop takes a misfit, not 'RTM'")
        # two lanes (shot halves on streams of their own): their step kernels run side by
        # side and fill each other's idle workgroup slots
        self.lanes = Lanes(inp["nx"], inp["nz"], inp["dx"], inp["dt"], inp["nt"], inp["npml"], inp["src_xz"],
                           inp["rec_xz"], lanes=lanes, free_surface=inp["free_surface"], path="batched",
                           strips=solver.STRIPS)
        self.device = self.lanes.device
        src = cumulative_trapezoid(simu.source.wavelet, axis=-1, initial=0)
        src = np.array(np.broadcast_to(np.asarray(src, dtype=np.float32), (self.ns, self.nt)))
        self.src = torch.from_numpy(src).to(self.device)
        self.rho = torch.from_numpy(inp["rho"]).to(self.device)
        # the SU header's offsets, nint(gx) - nint(sx) (io.f90:687-774), as the file-based
        # loop's processing reads them
        self.offsets = (su.nint(inp["rec_xz"][:, :, 0]) - su.nint(inp["src_xz"][:, None, 0])).astype(np.float64)
        self.identity = preprocess._identity(optim)
        self.forwards = 0
        self.adjoints = 0
        if obs is None:
            if simu_true is None:
                raise ValueError("need observed gathers or simu_true")
            vt = torch.from_numpy(solver.solver_inputs(simu_true)["vp"]).to(self.device)
            rt = torch.from_numpy(solver.solver_inputs(simu_true)["rho"]).to(self.device)
            self.lanes.set_model(vt, rt)
            obs = self.lanes.forward(self.src, save_strips=0)
        self.obs = self._process(torch.as_tensor(np.asarray(obs) if not torch.is_tensor(obs) else obs,
                                                 dtype=torch.float32).to(self.device).contiguous().clone())
        # pick the step kernels' launch configuration once, in the starting model (the
        # results do not depend on it)
        self.lanes.set_model(torch.from_numpy(inp["vp"]).to(self.device), self.rho)
        self.tuned = self.lanes.autotune(steps=20)

    def _process(self, gathers, offsets=None):
        if not self.identity:
            preprocess.process_gathers_device(self.optim, gathers, self.dt,
                                              self.offsets if offsets is None else offsets)
        return gathers

    def _lane_process(self, i, r, syn):
        """Processing of lane i's gathers (shots r[0]..r[1]-1), in place."""
        if not self.identity:
            self._process(syn, self.offsets[r[0]:r[1]])

    def _set(self, v):
        torch = _torch()
        self.lanes.set_model(v.reshape(self.nx, self.nz).to(torch.float32), self.rho)

    def misfit(self, v):
        """Misfit of model vector v: forward without strips, processing, misfit."""
        self._set(v)
        self.forwards += 1
        return float(self.lanes.misfit(self.src, self.obs, self.kind, self.dt, process=self._lane_process))

    def gradient(self, v):
        """(misfit, preconditioned gradient vector float64) of model vector v, as
        forward(savesnap=1), process_workflow, misfit, adjoint (with grad_precond)."""
        torch = _torch()
        from .engine import grad_precond
        self._set(v)
        self.forwards += 1
        J, g, fw, bk = self.lanes.gradient(self.src, self.obs, self.kind, self.dt, process=self._lane_process)
        self.adjoints += 1
        o = self.optim
        # the solver's kernel and illuminations are float32 files (toolbox/solver.py:86-92)
        gp = grad_precond(g.to(torch.float32), fw.to(torch.float32), bk.to(torch.float32), o.vpmax,
                          o.marine_or_land in ['Marine', 'Offshore'], o.grad_mute, o.grad_thred, o.grad_smooth,
                          o.grad_mask)
        return float(J), gp.reshape(-1)

    def close(self):
        self.lanes.close()


# -- search directions (inversion.py:112-263) -----------------------------------------------
def NLCG(inv_scheme):
    """inversion.py:112-171 on device vectors."""
    cg_itmax = 6
    g_now, g_old, d_old = inv_scheme['g_now'], inv_scheme['g_old'], inv_scheme['d_old']
    cg_it = inv_scheme['cg_it'] + 1
    if cg_it == 1:
        d_now = -g_now
        status = 0
    elif cg_it > cg_itmax:
        d_now = -g_now
        status = 1
        print('restarting NLCG... [periodic restart]')
    else:
        beta = _dot(g_now, g_now - g_old) / _dot(g_old, g_old)
        if beta <= 0.:
            beta = 0.
            status = 1
            print('restarting NLCG... [negative beta]')
        else:
            status = 0
        d_now = -g_now + beta * d_old
    if _dot(d_now, g_now) / _dot(g_now, g_now) > 0.:
        d_now = -g_now
        status = 1
        print('restarting NLCG... [not a descent direction]')
    else:
        print('find NLCG direction')
    if status:
        cg_it = 0
    inv_scheme['d_now'] = d_now
    inv_scheme['cg_it'] = cg_it
    return inv_scheme


def lbfgs_direction(S, Y, k, g_now):
    """Two-loop recursion over the k newest pairs (row 0 newest) with the Liu-Nocedal
    initial scaling (inversion.py:222-239); S, Y: (memory, n) float32 arrays.  The
    updates scale the float32 rows in FP64, as numpy does for a float64 scalar times a
    float32 column."""
    f64 = _torch().float64
    rho = [0.0] * k
    alpha = [0.0] * k
    q = g_now.to(f64)
    for i in range(k):
        rho[i] = 1. / _dot(Y[i], S[i])
        alpha[i] = rho[i] * _dot(S[i], q)
        q = q - alpha[i] * Y[i].to(f64)
    z = (_dot(Y[0], S[0]) / _dot(Y[0], Y[0])) * q
    for i in range(k - 1, -1, -1):
        beta = rho[i] * _dot(Y[i], z)
        z = z + (alpha[i] - beta) * S[i].to(f64)
    return z


def LBFGS(inv_scheme):
    """inversion.py:174-263 on device vectors; the float32 history rows shift down by
    one (row 0 newest) as the reference's memmap columns do."""
    memory = inv_scheme['memory']
    memory_used = inv_scheme['memory_used']
    lbfgs_it = inv_scheme['lbfgs_it'] + 1
    g_now, g_old = inv_scheme['g_now'], inv_scheme['g_old']
    v_now, v_old = inv_scheme['v_now'], inv_scheme['v_old']
    S, Y = inv_scheme['S'], inv_scheme['Y']
    if lbfgs_it == 1:
        d_now = -g_now
        status = 0
    else:
        s = v_now - v_old
        y = g_now - g_old
        if memory_used > 0:
            S[1:] = S[:-1].clone()
            Y[1:] = Y[:-1].clone()
        S[0] = s
        Y[0] = y
        memory_used = min(memory_used + 1, memory)
        z = lbfgs_direction(S, Y, memory_used, g_now)
        if _dot(-z, g_now) / _dot(g_now, g_now) > 0:
            d_now = -g_now
            status = 1
            print('restarting LBFGS... [not a descent direction]')
        else:
            d_now = -z
            status = 0
    if status:
        lbfgs_it = 1
        memory_used = 0
        S.zero_()
        Y.zero_()
    else:
        print('find LBFGS direction')
    inv_scheme['memory_used'] = memory_used
    inv_scheme['lbfgs_it'] = lbfgs_it
    inv_scheme['d_now'] = d_now
    return inv_scheme


# -- line search, decisions, state (inversion.py:266-474) ------------------------------------
def model_update(vp, update, vpmax, vpmin):
    """inversion.py:465-474: step, then the reference's two masked assignments in its
    order (values above vpmax become vpmax, then values below vpmin become vpmin)."""
    torch = _torch()
    v = vp + update
    v = torch.where(v > vpmax, torch.full_like(v, vpmax), v)
    return torch.where(v < vpmin, torch.full_like(v, vpmin), v)


def backtrack(problem, optim, inv_scheme):
    """inversion.py:266-338 (see swit_amd.inversion.backtrack), with forward-only misfit
    evaluations of the device problem."""
    dec, search_max, vmax_thresh, vmin_thresh = 0.5, 6, 300, 20
    f_now, v_now, d_now = inv_scheme['f_now'], inv_scheme['v_now'], inv_scheme['d_now']
    if optim.scheme in ['LBFGS']:
        dmax = float(d_now.abs().max())
        if inv_scheme['lbfgs_it'] == 1:
            step = optim.step_length
        elif dmax > vmax_thresh:
            step = 1.0 * vmax_thresh / dmax
        elif dmax < vmin_thresh:
            step = 1.0 * vmin_thresh / dmax
        else:
            step = 1.0
    else:
        step = optim.step_length
    ls_iter = 0
    while True:
        ls_iter += 1
        v_try = model_update(v_now, d_now * step, optim.vpmax, optim.vpmin)
        f_try = problem.misfit(v_try)
        if f_try >= f_now:
            print("line search: step = %.4f, f_now = %.4e, f_try = %.4e. Not decrease." % (step, f_now, f_try))
            width = dec
        else:
            inv_scheme = _host.updata_inv_scheme(inv_scheme, 0, ls_iter, step, f_try, v_try)
            print("line search: step = %.4f, f_now = %.4e, f_try = %.4e. Succeed." % (step, f_now, f_try))
            return inv_scheme
        if ls_iter >= search_max:
            inv_scheme = _host.updata_inv_scheme(inv_scheme, -1, ls_iter, step, f_try, v_try)
            print("line search fails: exceed the iteration of linesearch.")
            return inv_scheme
        step = step * width


def optimize_init(n, optim, device):
    """inversion.py:341-400 on the device (memory 5 for L-BFGS)."""
    torch = _torch()
    temp = torch.zeros(n, dtype=torch.float64, device=device)
    common = {'g_now': temp, 'g_old': temp, 'd_now': temp, 'v_old': temp, 'v_now': temp,
              'v_try': temp, 'f_old': 0., 'f_now': 0., 'f_try': 0., 'ls_iter': 0, 'ls_status': 0,
              'ls_step': 0.}
    if optim.scheme in ['NLCG']:
        return dict(common, d_old=temp, cg_it=0)
    memory = 5
    return dict(common, memory=memory, memory_used=0, lbfgs_it=0,
                S=torch.zeros((memory, n), dtype=torch.float32, device=device),
                Y=torch.zeros((memory, n), dtype=torch.float32, device=device))


def desicion_make(n, optim, inv_scheme, device):
    """inversion.py:403-449."""
    if optim.scheme in ['NLCG']:
        if inv_scheme['ls_status'] == -1 and inv_scheme['cg_it'] > 1:
            inv_scheme['cg_it'] = 0
            inv_scheme['g_old'] = inv_scheme['g_old'] * 0.
            inv_scheme['d_old'] = inv_scheme['d_old'] * 0.
            print('line search fails: restarting the line search ...')
        else:
            inv_scheme['g_old'] = inv_scheme['g_now']
            inv_scheme['d_old'] = inv_scheme['d_now']
            inv_scheme['f_old'] = inv_scheme['f_now']
            inv_scheme['f_now'] = inv_scheme['f_try']
            inv_scheme['v_now'] = inv_scheme['v_try']
            print('accept line search result')
    elif optim.scheme in ['LBFGS']:
        if inv_scheme['ls_status'] == -1:
            old = inv_scheme
            inv_scheme = optimize_init(n, optim, device)
            for key in ('v_now', 'f_now', 'g_now', 'd_now'):
                inv_scheme[key] = old[key]
            inv_scheme['v_old'] = old['v_now']
            print('line search fails: restarting the line search ...')
        else:
            inv_scheme['g_old'] = inv_scheme['g_now']
            inv_scheme['v_old'] = inv_scheme['v_now']
            inv_scheme['v_now'] = inv_scheme['v_try']
            inv_scheme['f_old'] = inv_scheme['f_now']
            inv_scheme['f_now'] = inv_scheme['f_try']
            print('accept line search result')
    return inv_scheme


def save_inv_scheme(simu, optim, inv_scheme):
    """toolbox/tools.py:340-359 from device vectors (copied to the host once each)."""
    host = dict(inv_scheme)
    for key in ('g_now', 'd_now', 'v_now'):
        host[key] = inv_scheme[key].cpu().numpy()
    _host.save_inv_scheme(simu, optim, host)


def inversion(simu, optim, inv_model, problem=None, obs=None, simu_true=None):
    """inversion.py:30-79 with device-resident vectors.  problem: a DeviceProblem (one is
    built from simu, optim and obs / simu_true when None).  Returns the final state
    (device vectors); simu.model.vp holds the final model on the host."""
    torch = _torch()
    simu.model.vp = inv_model['vp']
    simu.model.rho = inv_model['rho']
    own = problem is None
    if own:
        problem = DeviceProblem(simu, optim, obs=obs, simu_true=simu_true)
    nx, nz = problem.nx, problem.nz
    n = nx * nz
    dev = problem.device
    inv_scheme = optimize_init(n, optim, dev)
    v0 = torch.from_numpy(np.ascontiguousarray(np.asarray(simu.model.vp, dtype=np.float32)).reshape(-1)).to(dev)
    inv_scheme['v_now'] = v0.to(torch.float64)
    inv_scheme['v_old'] = v0.to(torch.float64)
    try:
        while optim.iter < optim.maxiter:
            optim.iter += 1
            print('\n-----------  iteration %d  -----------\n' % optim.iter)
            inv_scheme['f_now'], inv_scheme['g_now'] = problem.gradient(inv_scheme['v_now'])
            if optim.scheme in ['NLCG']:
                inv_scheme = NLCG(inv_scheme)
            elif optim.scheme in ['LBFGS']:
                inv_scheme = LBFGS(inv_scheme)
            inv_scheme = backtrack(problem, optim, inv_scheme)
            inv_scheme = desicion_make(n, optim, inv_scheme, dev)
            save_inv_scheme(simu, optim, inv_scheme)
        print('\n-----------  iteration end  -----------\n')
    finally:
        simu.model.vp = inv_scheme['v_now'].cpu().numpy().reshape(nx, nz)
        if own:
            problem.close()
    return inv_scheme
