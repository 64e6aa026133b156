"""Gradient preconditioning of the reference toolbox (toolbox/postprocess.py:20-111):
water/land taper, illumination preconditioner (F^ + B^)/max with floor 1e-4 after
Gaussian smoothing (span 40), gradient smoothing, optional mask, and scaling to
vpmax / max|g|.  The reference's illumination figures (plot_model2D) are not drawn.
"""
from __future__ import annotations

import numpy as np
from scipy.signal.windows import hamming

from .tools import array2vector, smooth2d, vector2array


def grad_precond(simu, optim, grad, forw, back):
    nx, nz = simu.model.nx, simu.model.nz
    vpmax = optim.vpmax
    marine_or_land = optim.marine_or_land
    grad_mute = optim.grad_mute
    grad_thred = optim.grad_thred
    grad_smooth = optim.grad_smooth
    grad_mask = optim.grad_mask
    # the gradient keeps its dtype (float32 from the solver's files) through the in-place
    # taper product, as in the reference; the division by the preconditioner promotes it
    grad = np.array(vector2array(np.asarray(grad), nx, nz))
    forw = vector2array(np.asarray(forw), nx, nz)
    back = vector2array(np.asarray(back), nx, nz)
    if grad_mute > 0:
        grad *= grad_taper(nx, nz, tapersize=grad_mute, thred=grad_thred, marine_or_land=marine_or_land)
    span = 40 if min(nx, nz) > 40 else int(min(nx, nz) / 2)
    forw = smooth2d(forw, span)
    back = smooth2d(back, span)
    epsilon = 0.0001
    forw = forw / np.max(forw)
    back = back / np.max(back)
    precond = forw + back
    precond = precond / np.max(precond)
    precond[precond < epsilon] = epsilon
    grad = grad / np.power(precond, 1)
    if grad_smooth > 0:
        if marine_or_land in ['Marine', 'Offshore']:
            grad[:, grad_mute:] = smooth2d(grad[:, grad_mute:], span=grad_smooth)
        else:
            grad = smooth2d(grad, span=grad_smooth)
    if grad_mask is not None:
        if np.shape(grad_mask) != np.shape(grad):
            raise ValueError('Wrong size of grad mask: the size of the mask should be identical to the size of vp model')
        grad = grad * grad_mask
    grad *= vpmax / abs(grad).max()
    return array2vector(grad)


def grad_precond_device(simu, optim, grad, forw, back):
    """grad_precond on the GPU (swit_amd.engine.grad_precond, the same steps in FP64, the
    smoothing as two 1-D passes): equal to grad_precond to FP64 rounding."""
    from .engine import grad_precond as _device
    nx, nz = simu.model.nx, simu.model.nz
    g = _device(vector2array(np.asarray(grad, np.float32), nx, nz), vector2array(np.asarray(forw, np.float32), nx, nz),
                vector2array(np.asarray(back, np.float32), nx, nz), optim.vpmax,
                optim.marine_or_land in ['Marine', 'Offshore'], optim.grad_mute, optim.grad_thred,
                optim.grad_smooth, optim.grad_mask)
    return array2vector(g.cpu().numpy())


def grad_taper(nx, nz, tapersize=20, thred=0.05, marine_or_land='Marine'):
    """postprocess.py:88-111 (Hamming window from scipy.signal.windows)."""
    if marine_or_land in ['Marine', 'Offshore']:
        taper = np.ones((nx, nz))
        taper[:, :tapersize] = 0.0
    else:
        H = hamming(tapersize * 2)[tapersize:]
        taper = np.zeros((nx, nz))
        taper[:, :tapersize] = H
        taper = smooth2d(taper, span=tapersize // 2)
        taper /= taper.max()
        taper *= (1 - thred)
        taper = -taper + 1
        taper = taper * taper
    return taper
