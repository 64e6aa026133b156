"""ORACLE -- test infrastructure, not product code.

A NumPy restatement of the reference inversion loop's decisions
(toolbox/inversion.py:30-504), the checker of the device-resident loop
(swit_amd.inversion_device).  Only tests/ may import this module.

The loop is driven by two callbacks, so a test can put any problem under it (a toy
quadratic, the drop-in solver through files, the device problem):

    evaluate(v)  -> (f, g)   misfit and preconditioned gradient at model vector v
                             (the reference: forward with snapshots, processing,
                             misfit, adjoint; inversion.py:48-60)
    trial(v)     -> f        misfit only (a line-search forward; inversion.py:319-325)

One iteration: evaluate at the current model, a search direction, a backtracking line
search, and the accept / restart decision.  The precision flow is the reference's: the
L-BFGS history lives in float32 columns (its memmaps, inversion.py:212-223), so its
inner products over two history columns are float32 ones; everything else is float64.

  NLCG      Polak-Ribiere+ (inversion.py:111-171): restart to steepest descent on the
            first step, after 6 steps, on beta <= 0 and on a non-descent direction
  L-BFGS    memory 5, newest pair first, Liu-Nocedal scaling of the initial inverse
            Hessian (inversion.py:174-271); steepest descent on the first step or a
            non-descent direction, which also clears the history
  search    at most 6 trials (inversion.py:274-338): the first L-BFGS step and every
            NLCG step start at optim.step_length, later L-BFGS steps at 1 unless the
            largest direction entry exceeds 300 (step 300/max) or is below 20 (step
            20/max); every trial model is clipped to [vpmin, vpmax] (values above vpmax
            first); a trial is accepted when its misfit is below the current one,
            otherwise the step halves; the sixth failure ends the search as failed
  decision  (inversion.py:403-449) NLCG: a failed search after more than one CG step
            clears the CG memory and keeps the model; otherwise (including a failed
            search on the first CG step) the trial model is taken.  L-BFGS: a failed
            search keeps the model and re-initialises the whole state, so that the
            iteration logs 0 trials, step 0 and a previous misfit of 0; a successful one
            takes the trial model

How it is pinned (tests/test_oracle_inversion.py): the two-loop recursion equals the
dense BFGS inverse-Hessian update built from the same pairs (Nocedal and Wright,
algorithm 7.4 against eq. 6.17); Polak-Ribiere by its formula; the clipping order; on a
separable quadratic the first step is steepest descent, every accepted model lowers the
misfit and L-BFGS converges; later L-BFGS steps bound the largest model change to
[20, 300]; forced line-search failures take each scheme's restart path.
"""
from __future__ import annotations

import numpy as np

MEMORY = 5
CG_RESTART = 6
TRIALS = 6
STEP_MAX_ENTRY = 300.0
STEP_MIN_ENTRY = 20.0


def clip(v, vmax, vmin):
    """Model update bound: entries above vmax become vmax, then entries below vmin
    become vmin (the reference's two masked assignments, in that order)."""
    out = np.array(v, dtype=np.float64, copy=True)
    out[out > vmax] = vmax
    out[out < vmin] = vmin
    return out


def two_loop(S, Y, k, g):
    """L-BFGS direction H g from the k newest pairs (column 0 newest) of the float32
    histories S, Y (n, MEMORY), with H0 = (y0.s0 / y0.y0) I."""
    rho = np.zeros(k)
    alpha = np.zeros(k)
    q = g
    for i in range(k):
        rho[i] = 1.0 / np.dot(Y[:, i], S[:, i])
        alpha[i] = rho[i] * np.dot(S[:, i], q)
        q = q - alpha[i] * Y[:, i]
    z = (np.dot(Y[:, 0], S[:, 0]) / np.dot(Y[:, 0], Y[:, 0])) * q
    for i in reversed(range(k)):
        z = z + (alpha[i] - rho[i] * np.dot(Y[:, i], z)) * S[:, i]
    return z


def polak_ribiere(g, g_prev):
    return np.dot(g, g - g_prev) / np.dot(g_prev, g_prev)


def descent_measure(d, g):
    """d.g / g.g: positive when d points uphill."""
    return np.dot(d, g) / np.dot(g, g)


class ReferenceLoop:
    """The loop's state and one iteration at a time.

    scheme: 'NLCG' or 'LBFGS'; step_length, vpmax, vpmin: optim's.  After each
    iteration, `history` holds (misfit after the decision, line-search trials, step,
    search status) and `models` the model vector after the decision."""

    def __init__(self, v0, scheme, step_length, vpmax, vpmin):
        self.scheme = scheme
        self.step_length, self.vpmax, self.vpmin = step_length, vpmax, vpmin
        n = np.size(v0)
        self.v = np.asarray(v0, dtype=np.float64).reshape(-1).copy()
        self.v_prev = self.v.copy()
        self.f = 0.0                        # misfit of the accepted model (0 before the first)
        self.f_first = None                 # misfit of the starting model
        self.history, self.models, self.directions, self.gradients = [], [], [], []
        self._reset(n)

    def _reset(self, n):
        self.g_prev = np.zeros(n)
        self.d_prev = np.zeros(n)
        self.steps = 0                      # NLCG steps since a restart / L-BFGS iteration count
        self.used = 0                       # L-BFGS pairs held
        self.S = np.zeros((n, MEMORY), np.float32)
        self.Y = np.zeros((n, MEMORY), np.float32)

    # -- directions -------------------------------------------------------------------
    def _nlcg(self, g):
        self.steps += 1
        restart = False
        if self.steps == 1:
            d = -g
        elif self.steps > CG_RESTART:
            d, restart = -g, True
        else:
            beta = polak_ribiere(g, self.g_prev)
            if beta <= 0.0:
                beta, restart = 0.0, True
            d = -g + beta * self.d_prev
        if descent_measure(d, g) > 0.0:
            d, restart = -g, True
        if restart:
            self.steps = 0
        return d

    def _lbfgs(self, g):
        self.steps += 1
        if self.steps == 1:
            return -g
        s, y = self.v - self.v_prev, g - self.g_prev
        if self.used:
            self.S[:, 1:] = self.S[:, :-1]
            self.Y[:, 1:] = self.Y[:, :-1]
        self.S[:, 0], self.Y[:, 0] = s, y
        self.used = min(self.used + 1, MEMORY)
        z = two_loop(self.S, self.Y, self.used, g)
        if descent_measure(-z, g) > 0:
            self.steps, self.used = 1, 0
            self.S[:] = 0.0
            self.Y[:] = 0.0
            return -g
        return -z

    # -- one iteration --------------------------------------------------------------------
    def iterate(self, evaluate, trial):
        f_now, g = evaluate(self.v)
        f_now = float(f_now)
        if self.f_first is None:
            self.f_first = f_now
        g = np.asarray(g, dtype=np.float64).reshape(-1)
        d = self._nlcg(g) if self.scheme == 'NLCG' else self._lbfgs(g)
        # line search
        if self.scheme == 'LBFGS' and self.steps != 1:
            dmax = np.abs(d).max()
            step = (STEP_MAX_ENTRY / dmax if dmax > STEP_MAX_ENTRY else
                    STEP_MIN_ENTRY / dmax if dmax < STEP_MIN_ENTRY else 1.0)
        else:
            step = self.step_length
        for tries in range(1, TRIALS + 1):
            v_try = clip(self.v + d * step, self.vpmax, self.vpmin)
            f_try = float(trial(v_try))
            if f_try < f_now:
                failed = False
                break
            if tries == TRIALS:
                failed = True
                break
            step *= 0.5
        # decision
        if self.scheme == 'NLCG':
            if failed and self.steps > 1:
                self.steps = 0
                self.g_prev = self.g_prev * 0.0
                self.d_prev = self.d_prev * 0.0
                self.f = f_now
            else:
                self.g_prev, self.d_prev = g, d
                self.f, self.v = f_try, v_try
        else:
            if failed:
                self._reset(self.v.size)
                self.v_prev = self.v.copy()
                self.f = f_now
                tries, step = 0, 0.0          # the re-initialised state logs no search
            else:
                self.g_prev, self.v_prev = g, self.v
                self.f, self.v = f_try, v_try
        self.history.append((self.f, tries, step, -1 if failed else 0))
        self.models.append(self.v.copy())
        self.directions.append(d)
        self.gradients.append(g)
        return self.f
