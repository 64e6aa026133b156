"""Gradient preconditioning for the drop-in's adjoint() (toolbox/postprocess.py:20-111):
the same call as the reference's grad_precond(simu, optim, grad, forw, back), run on
the device by swit_grad_precond (FP64: taper, illumination preconditioner with its
1e-4 floor, separable Gaussian smoothing, mask, scaling to vpmax / max|g|)."""
from __future__ import annotations

import numpy as np

from .tools import array2vector, vector2array


def grad_precond(simu, optim, grad, forw, back):
    """grad, forw, back: the solver's kernel and illuminations as 1-D (nx*nz) float32
    vectors (the .bin files' contents).  Returns the preconditioned gradient as a 1-D
    float64 vector."""
    from .engine import grad_precond as _device
    nx, nz = simu.model.nx, simu.model.nz

    def field(v):
        return vector2array(np.asarray(v, np.float32), nx, nz)
    g = _device(field(grad), field(forw), field(back), optim.vpmax, optim.marine_or_land in ['Marine', 'Offshore'],
                optim.grad_mute, optim.grad_thred, optim.grad_smooth, optim.grad_mask)
    return array2vector(g.cpu().numpy())
