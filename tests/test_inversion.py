"""The inversion loop restatement (swit_amd.inversion, toolbox/inversion.py) on the
CPU: the L-BFGS two-loop recursion equals the dense BFGS inverse-Hessian update built
from the same pairs (Nocedal & Wright, algorithm 7.4 vs eq. 6.17); Polak-Ribiere and
the descent test by their formulas; and the whole loop (NLCG and L-BFGS, backtracking,
decisions, outputs) on a toy separable quadratic standing in for the solver."""
import os
import types

import numpy as np
import pytest

from swit_amd import inversion as inv


def test_two_loop_equals_dense_bfgs():
    rng = np.random.default_rng(0)
    n, k = 7, 4
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    S = rng.standard_normal((n, k))
    Y = A @ S
    g = rng.standard_normal(n)
    z = inv.lbfgs_direction(S, Y, k, g)
    gamma = Y[:, 0] @ S[:, 0] / (Y[:, 0] @ Y[:, 0])
    H = gamma * np.eye(n)
    for i in range(k - 1, -1, -1):          # oldest pair first, newest (column 0) last
        s, y = S[:, i], Y[:, i]
        rho = 1.0 / (y @ s)
        V = np.eye(n) - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    np.testing.assert_allclose(z, H @ g, rtol=1e-10)
    # the newest secant condition holds: H y0 = s0
    np.testing.assert_allclose(H @ Y[:, 0], S[:, 0], rtol=1e-10)


def test_polak_ribiere_and_descent():
    g_now = np.array([1.0, 2.0, -1.0])
    g_old = np.array([0.5, -1.0, 2.0])
    assert inv.pollak_ribere(g_now, g_old) == pytest.approx(g_now @ (g_now - g_old) / (g_old @ g_old))
    assert inv.check_descent(-g_now, g_now) == pytest.approx(-1.0)
    v = inv.model_update(np.array([1500., 3000., 2000.]), np.array([-200., 800., 5.]), 3500., 1400.)
    np.testing.assert_array_equal(v, [1400., 3500., 2005.])


def _toy(tmp_path, scheme, monkeypatch, maxiter):
    """misfit f(v) = 1/2 sum w (v - v*)^2 through stand-ins for forward / misfit /
    adjoint / process_workflow that read simu.model.vp."""
    nx, nz = 6, 5
    rng = np.random.default_rng(3)
    target = rng.uniform(1800, 3200, (nx, nz))
    w = rng.uniform(0.5, 4.0, (nx, nz)).ravel()
    calls = []

    class Solver:
        @staticmethod
        def forward(simu, simu_type='obs', savesnap=0):
            calls.append(('forward', savesnap))

        @staticmethod
        def adjoint(simu, optim):
            calls.append(('adjoint',))
            return w * (simu.model.vp.ravel() - target.ravel())

    def misfit(simu, kind):
        r = simu.model.vp.ravel() - target.ravel()
        return 0.5 * float(np.sum(w * r * r))

    monkeypatch.setattr(inv, '_solver', lambda: Solver)
    monkeypatch.setattr(inv, '_misfit', misfit)
    monkeypatch.setattr(inv.preprocess, 'process_workflow', lambda *a, **k: calls.append(('process',)))
    home = str(tmp_path) + '/'
    simu = types.SimpleNamespace(model=types.SimpleNamespace(nx=nx, nz=nz, vp=None, rho=None),
                                 system=types.SimpleNamespace(homepath=home))
    optim = types.SimpleNamespace(scheme=scheme, iter=0, maxiter=maxiter, misfit_type='Waveform',
                                  step_length=0.05, vpmax=3500.0, vpmin=1400.0)
    vp0 = np.full((nx, nz), 2500.0)
    out = inv.inversion(simu, optim, {'vp': vp0, 'rho': np.ones_like(vp0)})
    return out, calls, home, misfit(types.SimpleNamespace(model=types.SimpleNamespace(vp=vp0)), None)


@pytest.mark.parametrize("scheme", ["LBFGS", "NLCG"])
def test_inversion_loop_on_quadratic(tmp_path, monkeypatch, scheme):
    out, calls, home, f0 = _toy(tmp_path, scheme, monkeypatch, maxiter=12)
    f = np.loadtxt(home + 'outputs/misfit_data.dat')
    # iteration 1 logs f_old (the starting misfit, '%.4e') and then f_now
    assert f[0] == pytest.approx(f0, rel=1e-4) and f[1] < f0
    assert np.all(np.diff(f[1:]) <= 0)                        # accepted steps never increase it
    # L-BFGS converges on a quadratic; NLCG keeps the fixed initial step of the
    # reference (inversion.py:305-306) and only has to make steady progress
    assert f[-1] < (1e-3 if scheme == "LBFGS" else 0.1) * f0
    # per iteration: forward with snapshots, processing, adjoint, then line-search forwards
    assert calls[0] == ('forward', 1) and calls[1] == ('process',) and calls[2] == ('adjoint',)
    assert ('forward', 0) in calls
    for i in range(1, 13):
        for sub, stem in (('gradient', 'grad'), ('direction', 'dirc'), ('velocity', 'vp')):
            assert os.path.exists(home + 'outputs/%s/%s-%d.bin' % (sub, stem, i))


@pytest.mark.parametrize("scheme", ["LBFGS", "NLCG"])
def test_device_loop_matches_file_loop(tmp_path, monkeypatch, scheme):
    """The device-resident loop (swit_amd.inversion_device), run on CPU tensors with the
    same toy problem, takes the same decisions as the file-based loop: identical
    line-search iteration counts and steps, misfits and models to float32 rounding
    (its inner products are FP64, the file loop's L-BFGS history products float32)."""
    import torch
    from swit_amd import inversion_device as dinv
    from swit_amd.tools import loadbinfloat32
    maxiter = 10
    _, _, home_a, _ = _toy(tmp_path / 'a', scheme, monkeypatch, maxiter)
    nx, nz = 6, 5
    rng = np.random.default_rng(3)
    target = torch.from_numpy(rng.uniform(1800, 3200, (nx, nz)).ravel())
    w = torch.from_numpy(rng.uniform(0.5, 4.0, (nx, nz)).ravel())

    class Problem:
        device = torch.device('cpu')

        def __init__(self):
            self.nx, self.nz = nx, nz

        def misfit(self, v):
            r = v - target
            return 0.5 * float(torch.sum(w * r * r))

        def gradient(self, v):
            return self.misfit(v), w * (v - target)

    home_b = str(tmp_path / 'b') + '/'
    simu = types.SimpleNamespace(model=types.SimpleNamespace(nx=nx, nz=nz, vp=None, rho=None),
                                 system=types.SimpleNamespace(homepath=home_b))
    optim = types.SimpleNamespace(scheme=scheme, iter=0, maxiter=maxiter, misfit_type='Waveform',
                                  step_length=0.05, vpmax=3500.0, vpmin=1400.0)
    vp0 = np.full((nx, nz), 2500.0)
    dinv.inversion(simu, optim, {'vp': vp0, 'rho': np.ones_like(vp0)}, problem=Problem())
    for name in ('line_search_iteration.dat',):
        np.testing.assert_array_equal(np.loadtxt(home_a + 'outputs/' + name), np.loadtxt(home_b + 'outputs/' + name))
    np.testing.assert_allclose(np.loadtxt(home_a + 'outputs/line_search_step.dat'),
                               np.loadtxt(home_b + 'outputs/line_search_step.dat'), rtol=1e-4)
    np.testing.assert_allclose(np.loadtxt(home_a + 'outputs/misfit_data.dat'),
                               np.loadtxt(home_b + 'outputs/misfit_data.dat'), rtol=2e-4)
    for i in range(1, maxiter + 1):
        va = loadbinfloat32(home_a + 'outputs/velocity/vp-%d.bin' % i)
        vb = loadbinfloat32(home_b + 'outputs/velocity/vp-%d.bin' % i)
        np.testing.assert_allclose(vb, va, rtol=1e-5)
    np.testing.assert_allclose(simu.model.vp.ravel(), va, rtol=1e-5)


def test_device_lbfgs_direction_matches_host():
    import torch
    from swit_amd import inversion_device as dinv
    rng = np.random.default_rng(5)
    n, k = 40, 5
    S = rng.standard_normal((n, k)).astype(np.float32)
    Y = (S * rng.uniform(0.5, 2.0, (n, 1))).astype(np.float32)
    g = rng.standard_normal(n)
    zh = inv.lbfgs_direction(S, Y, k, g)
    zd = dinv.lbfgs_direction(torch.from_numpy(np.ascontiguousarray(S.T)), torch.from_numpy(np.ascontiguousarray(Y.T)),
                              k, torch.from_numpy(g))
    np.testing.assert_allclose(zd.numpy(), zh, rtol=1e-5, atol=1e-6 * np.abs(zh).max())
