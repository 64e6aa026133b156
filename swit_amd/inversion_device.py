"""FWI on one GPU with every vector resident in HBM: the inversion loop of the
reference toolbox (toolbox/inversion.py:30-504) as its callers run it, without SU or
.bin files in between.  The model, gradient, search direction, the L-BFGS history and
the seismograms stay on the device; one iteration runs the forward with boundary
strips kept on the device, processing (swit_process), misfit and adjoint source
(swit_misfit), the adjoint and gradient preconditioning (swit_grad_precond).

The decisions are the reference loop's, so that the misfit history and the models
match it:
  * NLCG, Polak-Ribiere+: steepest descent on the first step, after 6 steps, on
    beta <= 0 and on a non-descent direction (inversion.py:111-171);
  * L-BFGS with 5 pairs, newest first, float32 history, Liu-Nocedal scaling; steepest
    descent, and a cleared history, on the first step or a non-descent direction
    (inversion.py:174-271);
  * backtracking (inversion.py:274-338): at most 6 trials from optim.step_length (the
    first L-BFGS step and every NLCG step) or from a step that bounds the largest
    model change to [20, 300] m/s, halving; trial models clipped to [vpmin, vpmax];
  * accept or restart (inversion.py:403-449), and the same outputs under
    homepath/outputs (tools.py:340-359): gradient, direction and model per iteration,
    the misfit and line-search logs, and the reference's log lines.
Inner products accumulate in FP64 on the device, where the reference forms float32
ones of its float32 history, so the directions agree to float32 rounding.
"""
from __future__ import annotations

import numpy as np

import os

from . import preprocess, su
from .tools import savebinfloat32


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


# -- the loop ---------------------------------------------------------------------------
_MEMORY = 5                 # L-BFGS pairs
_CG_RESTART = 6             # NLCG steps before a periodic restart
_TRIALS = 6                 # line-search trials
_DV_MAX, _DV_MIN = 300.0, 20.0   # bounds on the largest model change of a later L-BFGS step


def model_update(vp, update, vpmax, vpmin):
    """vp + update bounded as the reference bounds it (inversion.py:465-474): entries
    above vpmax become vpmax, then entries below vpmin become vpmin."""
    torch = _torch()
    v = vp + update
    v = torch.where(v > vpmax, torch.full_like(v, vpmax), v)
    return torch.where(v < vpmin, torch.full_like(v, vpmin), v)


def lbfgs_direction(S, Y, k, g):
    """H g by the two-loop recursion over the k newest pairs (row 0 newest) of the
    float32 histories S, Y (memory, n), with H0 = (y0.s0 / y0.y0) I; the float32 rows
    are scaled in FP64."""
    f64 = _torch().float64
    rho = [1.0 / _dot(Y[i], S[i]) for i in range(k)]
    alpha = [0.0] * k
    q = g.to(f64)
    for i in range(k):
        alpha[i] = rho[i] * _dot(S[i], q)
        q = q - alpha[i] * Y[i].to(f64)
    z = (_dot(Y[0], S[0]) / _dot(Y[0], Y[0])) * q
    for i in reversed(range(k)):
        z = z + (alpha[i] - rho[i] * _dot(Y[i], z)) * S[i].to(f64)
    return z


class _State:
    """The loop's state on the device: current and previous model, gradient and
    direction, misfits, the method's counters and the L-BFGS history."""

    def __init__(self, v0, scheme):
        torch = _torch()
        self.scheme = scheme
        self.v = v0.to(torch.float64)
        n, dev = self.v.numel(), self.v.device
        self.v_prev = self.v.clone()
        self.v_try = self.v
        self.f = self.f_prev = self.f_try = 0.0
        self.g = torch.zeros(n, dtype=torch.float64, device=dev)
        self.d = self.g
        self.step, self.trials, self.failed = 0.0, 0, False
        self.S = self.Y = None
        self.restart()

    def restart(self):
        """Forget the method's memory (a new start of NLCG / L-BFGS)."""
        torch = _torch()
        n, dev = self.v.numel(), self.v.device
        self.g_prev = torch.zeros(n, dtype=torch.float64, device=dev)
        self.d_prev = self.g_prev
        self.count = 0              # NLCG steps since a restart / L-BFGS iteration number
        self.used = 0               # L-BFGS pairs held
        if self.scheme == 'LBFGS':
            self.S = torch.zeros((_MEMORY, n), dtype=torch.float32, device=dev)
            self.Y = torch.zeros((_MEMORY, n), dtype=torch.float32, device=dev)

    def direction(self):
        if self.scheme == 'NLCG':
            self._nlcg()
        else:
            self._lbfgs()

    def _nlcg(self):
        g = self.g
        self.count += 1
        restart = False
        if self.count == 1:
            d = -g
        elif self.count > _CG_RESTART:
            d, restart = -g, True
            print('restarting NLCG... [periodic restart]')
        else:
            beta = _dot(g, g - self.g_prev) / _dot(self.g_prev, self.g_prev)
            if beta <= 0.0:
                beta, restart = 0.0, True
                print('restarting NLCG... [negative beta]')
            d = -g + beta * self.d_prev
        if _dot(d, g) / _dot(g, g) > 0.0:
            d, restart = -g, True
            print('restarting NLCG... [not a descent direction]')
        else:
            print('find NLCG direction')
        if restart:
            self.count = 0
        self.d = d

    def _lbfgs(self):
        g = self.g
        self.count += 1
        if self.count == 1:
            self.d = -g
            print('find LBFGS direction')
            return
        if self.used:
            self.S[1:] = self.S[:-1].clone()
            self.Y[1:] = self.Y[:-1].clone()
        self.S[0] = self.v - self.v_prev
        self.Y[0] = g - self.g_prev
        self.used = min(self.used + 1, _MEMORY)
        z = lbfgs_direction(self.S, self.Y, self.used, g)
        if _dot(-z, g) / _dot(g, g) > 0:
            self.d = -g
            self.count, self.used = 1, 0
            self.S.zero_()
            self.Y.zero_()
            print('restarting LBFGS... [not a descent direction]')
        else:
            self.d = -z
            print('find LBFGS direction')

    def line_search(self, problem, optim):
        if self.scheme == 'LBFGS' and self.count != 1:
            dmax = float(self.d.abs().max())
            step = _DV_MAX / dmax if dmax > _DV_MAX else _DV_MIN / dmax if dmax < _DV_MIN else 1.0
        else:
            step = optim.step_length
        for trial in range(1, _TRIALS + 1):
            v_try = model_update(self.v, self.d * step, optim.vpmax, optim.vpmin)
            f_try = problem.misfit(v_try)
            self.trials, self.step, self.f_try, self.v_try = trial, step, f_try, v_try
            if f_try < self.f:
                self.failed = False
                print("line search: step = %.4f, f_now = %.4e, f_try = %.4e. Succeed." % (step, self.f, f_try))
                return
            print("line search: step = %.4f, f_now = %.4e, f_try = %.4e. Not decrease." % (step, self.f, f_try))
            step *= 0.5
        self.failed = True
        print("line search fails: exceed the iteration of linesearch.")

    def decide(self):
        """Take the trial model, or keep the model and restart the method."""
        if self.scheme == 'NLCG' and self.failed and self.count > 1:
            self.count = 0
            self.g_prev = self.g_prev * 0.0
            self.d_prev = self.d_prev * 0.0
            print('line search fails: restarting the line search ...')
        elif self.scheme == 'LBFGS' and self.failed:
            # a fresh start of the whole state (inversion.py:425-434): the logs then show
            # no line search and a previous misfit of 0
            self.restart()
            self.v_prev = self.v
            self.f_prev = 0.0
            self.trials, self.step = 0, 0.0
            print('line search fails: restarting the line search ...')
        else:
            self.g_prev, self.d_prev, self.v_prev = self.g, self.d, self.v
            self.f_prev, self.f = self.f, self.f_try
            self.v = self.v_try
            print('accept line search result')


def _log(path, value, integer=False):
    """One value per line appended to a text log, '%d' or '%.4e' (tools.py:307-319)."""
    with open(path, 'a') as f:
        f.write(('%d\n' if integer else '%.4e\n') % value)


def save_outputs(simu, optim, st):
    """Per-iteration outputs under homepath/outputs, as tools.py:340-359 writes them:
    gradient, direction and model as float32 .bin files; the misfit log (the starting
    misfit first, at the first iteration), the line-search step and trial count."""
    out = simu.system.homepath + 'outputs'
    it = optim.iter
    for sub, name, vec in (('gradient', 'grad', st.g), ('direction', 'dirc', st.d), ('velocity', 'vp', st.v)):
        os.makedirs(out + '/' + sub, exist_ok=True)
        savebinfloat32(out + '/%s/%s-%d.bin' % (sub, name, it), vec.cpu().numpy())
    if it == 1:
        _log(out + '/misfit_data.dat', st.f_prev)
    _log(out + '/misfit_data.dat', st.f)
    _log(out + '/line_search_step.dat', st.step)
    _log(out + '/line_search_iteration.dat', st.trials, integer=True)


def inversion(simu, optim, inv_model, problem=None, obs=None, simu_true=None):
    """The loop of inversion.py:30-79 with device-resident vectors.  problem: a
    DeviceProblem (or any object with misfit(v) and gradient(v) -> (f, g) on device
    vectors); one is built from simu, optim and obs / simu_true when None.  Returns the
    final state; simu.model.vp holds the final model on the host."""
    torch = _torch()
    simu.model.vp = inv_model['vp']
    simu.model.rho = inv_model['rho']
    own = problem is None
    if own:
        problem = DeviceProblem(simu, optim, obs=obs, simu_true=simu_true)
    nx, nz = problem.nx, problem.nz
    dev = problem.device
    v0 = torch.from_numpy(np.ascontiguousarray(np.asarray(simu.model.vp, dtype=np.float32)).reshape(-1)).to(dev)
    st = _State(v0, optim.scheme)
    try:
        while optim.iter < optim.maxiter:
            optim.iter += 1
            print('\n-----------  iteration %d  -----------\n' % optim.iter)
            st.f, st.g = problem.gradient(st.v)
            st.direction()
            st.line_search(problem, optim)
            st.decide()
            save_outputs(simu, optim, st)
        print('\n-----------  iteration end  -----------\n')
    finally:
        simu.model.vp = st.v.cpu().numpy().reshape(nx, nz)
        if own:
            problem.close()
    return st
