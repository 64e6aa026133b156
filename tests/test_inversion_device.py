"""The product's inversion loop (swit_amd.inversion_device) against the oracle loop
(oracle/inversion_loop.py), on CPU tensors with the same problem: the same line-search
trial counts and steps, misfits and models to float32 rounding (the device loop's
inner products are FP64 where the reference's L-BFGS history products are float32),
through convergence on a quadratic and through each scheme's failure / restart path."""
import types

import numpy as np
import pytest

from oracle import inversion_loop as L

NX, NZ = 6, 5


class _Problem:
    """evaluate / trial callbacks as a device problem (torch CPU tensors) and as the
    oracle's (numpy)."""

    def __init__(self, evaluate, trial):
        import torch
        self.torch = torch
        self.device = torch.device('cpu')
        self.nx, self.nz = NX, NZ
        self._e, self._t = evaluate, trial

    def misfit(self, v):
        return float(self._t(v.numpy()))

    def gradient(self, v):
        f, g = self._e(v.numpy())
        return float(f), self.torch.from_numpy(np.asarray(g, np.float64))


def _run_device(tmp_path, scheme, evaluate, trial, maxiter):
    from swit_amd import inversion_device as D
    home = str(tmp_path) + '/'
    simu = types.SimpleNamespace(model=types.SimpleNamespace(nx=NX, nz=NZ, vp=None, rho=None),
                                 system=types.SimpleNamespace(homepath=home))
    optim = types.SimpleNamespace(scheme=scheme, iter=0, maxiter=maxiter, misfit_type='Waveform',
                                  step_length=0.05, vpmax=3500.0, vpmin=1400.0)
    vp0 = np.full((NX, NZ), 2500.0)
    st = D.inversion(simu, optim, {'vp': vp0, 'rho': np.ones_like(vp0)}, problem=_Problem(evaluate, trial))
    return home, st, simu


def _compare(tmp_path, scheme, make, maxiter):
    """make() -> fresh (evaluate, trial) callbacks for each of the two runs."""
    from swit_amd.tools import loadbinfloat32
    home, st, simu = _run_device(tmp_path, scheme, *make(), maxiter)
    ref = L.ReferenceLoop(np.full(NX * NZ, 2500.0), scheme, 0.05, 3500.0, 1400.0)
    evaluate, trial = make()
    for _ in range(maxiter):
        ref.iterate(evaluate, trial)
    out = home + 'outputs/'
    np.testing.assert_array_equal(np.loadtxt(out + 'line_search_iteration.dat'), [h[1] for h in ref.history])
    np.testing.assert_allclose(np.loadtxt(out + 'line_search_step.dat'), [h[2] for h in ref.history], rtol=1e-4)
    f_ref = [ref.f_first if ref.history[0][3] == 0 or scheme == 'NLCG' else 0.0] + [h[0] for h in ref.history]
    np.testing.assert_allclose(np.loadtxt(out + 'misfit_data.dat'), f_ref, rtol=2e-4)
    for i in range(maxiter):
        np.testing.assert_allclose(loadbinfloat32(out + 'velocity/vp-%d.bin' % (i + 1)), ref.models[i], rtol=1e-5)
        np.testing.assert_allclose(loadbinfloat32(out + 'direction/dirc-%d.bin' % (i + 1)), ref.directions[i],
                                   rtol=1e-4, atol=1e-4 * np.abs(ref.directions[i]).max())
    np.testing.assert_allclose(simu.model.vp.ravel(), ref.models[-1], rtol=1e-5)
    return ref


def _quadratic():
    rng = np.random.default_rng(3)
    target = rng.uniform(1800, 3200, NX * NZ)
    w = rng.uniform(0.5, 4.0, NX * NZ)

    def f(v):
        r = v - target
        return 0.5 * float(np.sum(w * r * r))
    return f, (lambda v: (f(v), w * (v - target)))


@pytest.mark.parametrize("scheme", ["LBFGS", "NLCG"])
def test_device_loop_matches_oracle_on_a_quadratic(tmp_path, scheme):
    f, fg = _quadratic()
    ref = _compare(tmp_path, scheme, lambda: (fg, f), 10)
    assert ref.history[-1][0] < 0.5 * ref.f_first


def test_device_loop_lbfgs_failure_restart(tmp_path):
    """Iteration 2's line search never succeeds: the model is kept, the history is
    cleared and iteration 3 starts again from steepest descent."""
    f, fg = _quadratic()

    def make():
        calls = {'n': 0}

        def trial(v):
            calls['n'] += 1
            return 1e30 if 2 <= calls['n'] <= 7 else f(v)
        return fg, trial
    ref = _compare(tmp_path, "LBFGS", make, 4)
    assert [h[3] for h in ref.history] == [0, -1, 0, 0]


def test_device_loop_nlcg_failure_paths(tmp_path):
    """A failed search on the first CG step takes the trial model; a failed search on a
    conjugate step (beta > 0) keeps the model and clears the CG memory."""
    gv = np.linspace(-1.0, 2.0, NX * NZ)

    def make():
        seq = iter([(10.0, gv), (6.0, 1.1 * gv), (5.0, 1.2 * gv), (5.0, 1.3 * gv)])
        trials = iter([1e30] * 6 + [5.0] + [1e30] * 6 + [4.0] * 6)
        return (lambda v: next(seq)), (lambda v: next(trials))
    ref = _compare(tmp_path, "NLCG", make, 4)
    assert [h[3] for h in ref.history] == [-1, 0, -1, 0]


def test_device_lbfgs_direction_matches_oracle():
    import torch
    from swit_amd import inversion_device as D
    rng = np.random.default_rng(5)
    n, k = 40, 5
    S = rng.standard_normal((n, L.MEMORY)).astype(np.float32)
    Y = (S * rng.uniform(0.5, 2.0, (n, 1))).astype(np.float32)
    g = rng.standard_normal(n)
    zh = L.two_loop(S, Y, k, g)
    zd = D.lbfgs_direction(torch.from_numpy(np.ascontiguousarray(S.T)), torch.from_numpy(np.ascontiguousarray(Y.T)),
                           k, torch.from_numpy(g))
    np.testing.assert_allclose(zd.numpy(), zh, rtol=1e-5, atol=1e-6 * np.abs(zh).max())
