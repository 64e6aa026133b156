"""File helpers of the FWI host side, following the reference's formats
(toolbox/tools.py): raw little-endian float32 binaries with numpy (nx, nz) order,
(nx, nz) <-> vector reshapes, and clearing the synthetic-data folder between forward
runs."""
from __future__ import annotations

import os
import shutil

import numpy as np


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
