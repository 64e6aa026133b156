"""Pins of the oracle's gradient preconditioning (oracle/precond.py, the restatement of
toolbox/postprocess.py:20-111 and toolbox/tools.py:239-275) by what the mathematics
fixes, so that a wrong kernel width, a missing normalisation, a transposed axis or a
wrong taper fails one of them."""
import numpy as np
import pytest

from oracle import precond as P


@pytest.mark.parametrize("span", [1, 3, 8])
def test_smoothing_of_an_impulse_is_the_kernel(span):
    """Far from the edges the normalised convolution of a unit impulse is the kernel
    w(i) w(k) / (sum w)^2 with w(k) = exp(-0.5 (2k)^2 / span^2): a Gaussian of standard
    deviation `span` sampled at x = 2k (linspace(-2 span, 2 span, 2 span + 1))."""
    n = 6 * span + 9
    f = np.zeros((n, n + 3))
    c = (n // 2, (n + 3) // 2)
    f[c] = 1.0
    out = P.smooth(f, span)
    k = np.arange(-span, span + 1)
    w = np.exp(-0.5 * (2.0 * k) ** 2 / span ** 2)
    expect = np.outer(w, w) / w.sum() ** 2
    win = out[c[0] - span:c[0] + span + 1, c[1] - span:c[1] + span + 1]
    np.testing.assert_allclose(win, expect, rtol=1e-12, atol=1e-15)
    assert out.sum() == pytest.approx(1.0, rel=1e-12)
    assert np.count_nonzero(np.abs(out) > 1e-300) == (2 * span + 1) ** 2


def test_smoothing_keeps_constants_and_acts_along_each_axis():
    """Normalised convolution leaves a constant field unchanged, edges included; a field
    varying along x only stays constant along z, and its smoothing along x equals the
    1-D normalised convolution with w."""
    rng = np.random.default_rng(0)
    np.testing.assert_allclose(P.smooth(np.full((30, 17), 2.5), 4), 2.5, rtol=1e-13)
    prof = rng.standard_normal(40)
    out = P.smooth(np.repeat(prof[:, None], 23, axis=1), 3)
    np.testing.assert_allclose(out, np.repeat(out[:, :1], 23, axis=1), rtol=1e-12)
    k = np.arange(-3, 4)
    w = np.exp(-2.0 * k * k / 9.0)
    num = np.convolve(prof, w, mode="same")
    den = np.convolve(np.ones(40), w, mode="same")
    np.testing.assert_allclose(out[:, 0], num / den, rtol=1e-12)


def test_marine_taper():
    t = P.taper(9, 20, 6, 0.0, True)
    assert np.all(t[:, :6] == 0.0) and np.all(t[:, 6:] == 1.0)


def test_land_taper():
    """(1 - (1 - thred) t)^2 with t the smoothed, normalised Hamming half-window: the
    weight is thred^2 where t = 1 (its maximum), rises monotonically with depth and is
    1 below the smoothed window's reach (size + size // 2 rows)."""
    size, thred = 10, 0.05
    t = P.taper(40, 60, size, thred, False)
    col = t[20]
    np.testing.assert_allclose(t, np.repeat(col[None], 40, axis=0), rtol=1e-12)
    assert np.all(np.diff(col[:size + size // 2 + 1]) >= -1e-15)
    np.testing.assert_allclose(col[size + size // 2 + 1:], 1.0, rtol=0, atol=1e-15)
    assert col.min() == pytest.approx(thred ** 2, rel=1e-9)
    assert 0.0 < col[0] < 1.0


def test_precondition_scales_and_divides():
    """With uniform illuminations the preconditioner is 1 everywhere, so without taper,
    smoothing or mask the result is the gradient scaled to max |g| = vpmax; with
    illumination that decays with depth the deep gradient is boosted (divided by a
    smaller preconditioner) and the floor 1e-4 bounds the boost."""
    rng = np.random.default_rng(1)
    g = rng.standard_normal((50, 45)).astype(np.float32)
    ones = np.ones_like(g)
    out = P.precondition(g, ones, ones, 3000.0, True)
    g64 = g.astype(np.float64)
    np.testing.assert_allclose(out, g64 * (3000.0 / np.abs(g64).max()), rtol=1e-13)
    depth = np.exp(-np.arange(45) / 5.0)[None, :].repeat(50, axis=0).astype(np.float32)
    out = P.precondition(np.ones_like(g), depth, depth, 1.0, True)
    assert np.all(np.diff(out[25]) >= 0) and out[25, -1] > 5 * out[25, 0]
    assert out.max() == pytest.approx(1.0)
    tiny = np.full_like(g, 1e-12)
    tiny[0, 0] = 1.0
    out = P.precondition(np.ones_like(g), tiny, tiny, 1.0, True)
    assert out.min() >= out.max() * P.FLOOR * (1 - 1e-12)
    mask = np.zeros((50, 45))
    mask[10:20, 5:9] = 1.0
    out = P.precondition(g, ones, ones, 2.0, True, mask=mask)
    assert np.all(out[mask == 0] == 0) and np.abs(out).max() == pytest.approx(2.0)
