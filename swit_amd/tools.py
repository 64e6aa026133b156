"""File helpers of the FWI host side, following the reference's formats
(toolbox/tools.py): raw little-endian float32 binaries with numpy (nx, nz) order,
(nx, nz) <-> vector reshapes, Gaussian smoothing, and clearing the synthetic-data
folder between forward runs."""
from __future__ import annotations

import os
import shutil

import numpy as np
from scipy import signal as _signal


def savebinfloat32(filename, obj):
    """toolbox/tools.py:72-78."""
    with open(filename, "wb") as fp:
        np.asarray(obj, dtype=np.float32).tofile(fp)


def loadbinfloat32(filename):
    """toolbox/tools.py:81-87."""
    with open(filename, "rb") as fp:
        return np.array(np.fromfile(fp, dtype=np.float32))


def array2vector(array):
    """toolbox/tools.py:225-229 ((nx, nz) -> nx*nz, iz fastest)."""
    nx, nz = array.shape[0:2]
    return array.reshape(nx * nz)


def vector2array(vector, nx, nz):
    """toolbox/tools.py:232-236."""
    return vector.reshape((nx, nz))


def cleandata(datapath):
    """toolbox/tools.py:322-327: empty a non-empty data folder."""
    if os.path.isdir(datapath) and os.listdir(datapath):
        shutil.rmtree(datapath)
        os.makedirs(datapath)


def gauss2(X, Y, mu, sigma, normalize=True):
    """Bivariate Gaussian on a grid (toolbox/tools.py:239-252)."""
    D = sigma[0, 0] * sigma[1, 1] - sigma[0, 1] * sigma[1, 0]
    B = np.linalg.inv(sigma)
    X = X - mu[0]
    Y = Y - mu[1]
    Z = B[0, 0] * X ** 2.0 + B[0, 1] * X * Y + B[1, 0] * X * Y + B[1, 1] * Y ** 2.0
    Z = np.exp(-0.5 * Z)
    if normalize:
        Z *= (2.0 * np.pi * np.sqrt(D)) ** (-1.0)
    return Z


def smooth2d(Z, span=10):
    """Normalised Gaussian smoothing with a (4 span + 1)^2 kernel of standard deviation
    span, edge-corrected by the smoothed ones field (toolbox/tools.py:255-275)."""
    Z = np.copy(Z)
    x = np.linspace(-2.0 * span, 2.0 * span, 2 * span + 1)
    y = np.linspace(-2.0 * span, 2.0 * span, 2 * span + 1)
    X, Y = np.meshgrid(x, y)
    mu = np.array([0.0, 0.0])
    sigma = np.diag([span, span]) ** 2.0
    F = gauss2(X, Y, mu, sigma)
    F = F / np.sum(F)
    W = np.ones(Z.shape)
    Z = _signal.convolve2d(Z, F, "same")
    W = _signal.convolve2d(W, F, "same")
    return Z / W
