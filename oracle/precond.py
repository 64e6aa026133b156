"""ORACLE -- test infrastructure, not product code.

NumPy restatement of the reference's gradient preconditioning, the checker of the
device preconditioner (swit_grad_precond).  Only tests/ may import this module.

  smooth(field, span)        toolbox/tools.py:239-275 (smooth2d with gauss2)
  taper(nx, nz, ...)         toolbox/postprocess.py:88-111 (grad_taper)
  precondition(...)          toolbox/postprocess.py:20-86 (grad_precond)

Arrays are numpy (nx, nz) with iz fastest, the layout of the solver's .bin files.

How each function is pinned (tests/test_oracle_precond.py): the smoothing of a unit
impulse far from the edges is the normalised kernel exp(-2 k^2 / span^2) in each axis
(the reference samples a Gaussian of standard deviation `span` at spacing 2); a
constant field is left unchanged up to the edges (normalised convolution), and a field
that varies along one axis only is smoothed along that axis only; the marine taper is
zero above the mute row and one below; the land taper is 1 - (1 - thred)^2 at the
surface and 1 far below the mute, and monotone in between; the preconditioned gradient
has maximum magnitude vpmax, is unchanged in sign, and equals the raw gradient scaled
when the illuminations are uniform and no smoothing is asked for.
"""
from __future__ import annotations

import numpy as np
from scipy.signal import convolve2d
from scipy.signal.windows import hamming

# floor of the normalised preconditioner (postprocess.py:55)
FLOOR = 1e-4


def smooth(field, span):
    """Normalised Gaussian smoothing: 'same'-size zero-padded convolution with the
    (2 span + 1)^2 kernel w(i) w(k), w(k) = exp(-2 k^2 / span^2) for k = -span..span
    (the reference's kernel: a Gaussian of standard deviation `span` evaluated at
    x = 2k), divided by the same convolution of a field of ones so that the edges are
    not darkened."""
    span = int(span)
    k = np.arange(-span, span + 1, dtype=np.float64)
    w = np.exp(-2.0 * k * k / float(span * span))
    kern = np.outer(w, w)
    kern /= kern.sum()
    num = convolve2d(np.asarray(field, dtype=np.float64), kern, mode="same")
    den = convolve2d(np.ones(np.shape(field)), kern, mode="same")
    return num / den


def taper(nx, nz, size, thred, marine):
    """Depth taper of the gradient: marine data zero the top `size` rows; land data
    damp them smoothly: the descending half of a 2*size Hamming window in the top rows,
    smoothed with span size // 2 and normalised, gives t, and the taper is
    (1 - (1 - thred) t)^2."""
    if marine:
        out = np.ones((nx, nz))
        out[:, :size] = 0.0
        return out
    ramp = np.zeros((nx, nz))
    ramp[:, :size] = hamming(2 * size)[size:]
    ramp = smooth(ramp, size // 2)
    ramp = ramp / ramp.max() * (1.0 - thred)
    return (1.0 - ramp) ** 2


def precondition(grad, forw, back, vpmax, marine, grad_mute=0, grad_thred=0.0, grad_smooth=0, mask=None):
    """The preconditioned gradient (nx, nz) float64 from the solver's float32 kernel and
    illuminations:
      1. the taper product, kept in the gradient's float32 (the reference multiplies in
         place into its float32 array);
      2. each illumination smoothed with span 40 (half the shorter side if that is 40
         or less) and scaled to a maximum of 1; their sum scaled to a maximum of 1 and
         floored at 1e-4; the gradient divided by it;
      3. smoothing with span grad_smooth, below the mute rows for marine data;
      4. the mask product, then the scale that makes max |g| = vpmax."""
    g = np.array(grad, dtype=np.float32, copy=True)
    nx, nz = g.shape
    if grad_mute > 0:
        g *= taper(nx, nz, grad_mute, grad_thred, marine)     # float32 result, as in place
    span = 40 if min(nx, nz) > 40 else min(nx, nz) // 2
    f = smooth(forw, span)
    b = smooth(back, span)
    p = f / f.max() + b / b.max()
    p = p / p.max()
    p = np.where(p < FLOOR, FLOOR, p)
    g = g / p
    if grad_smooth > 0:
        if marine:
            g[:, grad_mute:] = smooth(g[:, grad_mute:], grad_smooth)
        else:
            g = smooth(g, grad_smooth)
    if mask is not None:
        if np.shape(mask) != g.shape:
            raise ValueError("mask shape %s differs from the model's %s" % (np.shape(mask), g.shape))
        g = g * mask
    return g * (vpmax / np.abs(g).max())
