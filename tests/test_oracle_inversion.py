"""Pins of the oracle's inversion loop (oracle/inversion_loop.py, the restatement of
toolbox/inversion.py's decisions) by what the mathematics and the reference's rules fix."""
import numpy as np
import pytest

from oracle import inversion_loop as L


def test_two_loop_equals_dense_bfgs():
    """The two-loop recursion over k pairs (newest in column 0) applies the inverse
    Hessian of k BFGS updates of H0 = gamma I, gamma = y0.s0 / y0.y0, applied oldest
    pair first (Nocedal and Wright, algorithm 7.4 against eq. 6.17); the newest secant
    condition H y0 = s0 holds."""
    rng = np.random.default_rng(0)
    n, k = 7, 4
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    S = np.zeros((n, L.MEMORY), np.float32)
    Y = np.zeros((n, L.MEMORY), np.float32)
    S[:, :k] = rng.standard_normal((n, k))
    Y[:, :k] = (A @ S[:, :k].astype(np.float64)).astype(np.float32)
    g = rng.standard_normal(n)
    z = L.two_loop(S, Y, k, g)
    S64, Y64 = S.astype(np.float64), Y.astype(np.float64)
    H = (Y64[:, 0] @ S64[:, 0]) / (Y64[:, 0] @ Y64[:, 0]) * np.eye(n)
    for i in reversed(range(k)):
        s, y = S64[:, i], Y64[:, i]
        r = 1.0 / (y @ s)
        V = np.eye(n) - r * np.outer(y, s)
        H = V.T @ H @ V + r * np.outer(s, s)
    np.testing.assert_allclose(z, H @ g, rtol=1e-5)
    np.testing.assert_allclose(H @ Y64[:, 0], S64[:, 0], rtol=1e-10)


def test_formulas_and_clipping_order():
    g, gp = np.array([1.0, 2.0, -1.0]), np.array([0.5, -1.0, 2.0])
    assert L.polak_ribiere(g, gp) == pytest.approx(g @ (g - gp) / (gp @ gp))
    assert L.descent_measure(-g, g) == pytest.approx(-1.0)
    # above vpmax first, then below vpmin: with vmin > vmax everything ends at vmin
    np.testing.assert_array_equal(L.clip([1500., 3000., 2000.], 3500., 1400.), [1500., 3000., 2000.])
    np.testing.assert_array_equal(L.clip([1300., 3600., 2005.], 3500., 1400.), [1400., 3500., 2005.])
    np.testing.assert_array_equal(L.clip([1000., 5000.], 1500., 2000.), [2000., 2000.])


def _quadratic(n=30, seed=3):
    rng = np.random.default_rng(seed)
    target = rng.uniform(1800, 3200, n)
    w = rng.uniform(0.5, 4.0, n)

    def f(v):
        r = v - target
        return 0.5 * float(np.sum(w * r * r))
    return f, (lambda v: (f(v), w * (v - target))), target


@pytest.mark.parametrize("scheme", ["LBFGS", "NLCG"])
def test_loop_descends_on_a_quadratic(scheme):
    """Every accepted model lowers the misfit; L-BFGS converges on a separable quadratic
    (its secant pairs recover the diagonal Hessian); NLCG keeps the fixed initial step
    of the reference and only has to make steady progress."""
    f, fg, target = _quadratic()
    lp = L.ReferenceLoop(np.full(30, 2500.0), scheme, 0.05, 3500.0, 1400.0)
    for _ in range(15):
        lp.iterate(fg, f)
    fs = [lp.f_first] + [h[0] for h in lp.history]
    assert np.all(np.diff(fs) <= 0) and fs[1] < fs[0]
    assert fs[-1] < (1e-6 if scheme == "LBFGS" else 0.2) * fs[0]
    # the first step of both schemes is steepest descent with optim.step_length
    assert lp.history[0][2] == pytest.approx(0.05) or lp.history[0][1] > 1
    np.testing.assert_allclose(lp.directions[0], -lp.gradients[0])


def test_lbfgs_later_step_bounds_the_largest_change():
    """After the first iteration the L-BFGS trial step makes the largest model change
    300 m/s when the direction's largest entry exceeds 300, 20 when it is below 20,
    and the full step otherwise."""
    f, fg, _ = _quadratic()
    lp = L.ReferenceLoop(np.full(30, 2500.0), "LBFGS", 0.05, 3500.0, 1400.0)
    for _ in range(6):
        lp.iterate(fg, f)
    for (fv, tries, step, status), d in zip(lp.history[1:], lp.directions[1:]):
        if tries == 1:
            dmax = np.abs(d).max()
            expect = 300.0 / dmax if dmax > 300 else 20.0 / dmax if dmax < 20 else 1.0
            assert step == pytest.approx(expect)


def test_failed_line_search_restart_paths():
    """A trial misfit that never decreases: L-BFGS keeps the model, logs 0 as the
    previous misfit and starts afresh (steepest descent next time); NLCG on its first
    step takes the last trial model anyway (the reference's rule), and on a later step
    keeps the model and clears its memory."""
    f, fg, _ = _quadratic()
    up = lambda v: 1e30
    v0 = np.full(30, 2500.0)
    lb = L.ReferenceLoop(v0, "LBFGS", 0.05, 3500.0, 1400.0)
    lb.iterate(fg, f)
    lb.iterate(fg, up)
    assert lb.history[-1][1:] == (0, 0.0, -1)         # the re-initialised state logs no search
    np.testing.assert_array_equal(lb.models[-1], lb.models[-2])
    assert lb.steps == 0 and lb.used == 0 and not lb.S.any()
    lb.iterate(fg, f)
    np.testing.assert_allclose(lb.directions[-1], -lb.gradients[-1])
    cg = L.ReferenceLoop(v0, "NLCG", 0.05, 3500.0, 1400.0)
    cg.iterate(fg, up)
    assert cg.history[-1][3] == -1 and not np.array_equal(cg.models[-1], v0)
    # a conjugate step (beta > 0: the gradient grows along itself) whose search fails
    gv = np.linspace(-1.0, 2.0, 30)
    cg = L.ReferenceLoop(v0, "NLCG", 0.05, 3500.0, 1400.0)
    cg.iterate(lambda v: (10.0, gv), lambda v: 5.0)
    m = cg.models[-1].copy()
    cg.iterate(lambda v: (5.0, 1.1 * gv), up)
    np.testing.assert_allclose(cg.directions[-1], -1.1 * gv - 0.11 * gv, rtol=1e-12)
    np.testing.assert_array_equal(cg.models[-1], m)
    assert cg.steps == 0 and not cg.d_prev.any() and cg.f == 5.0
