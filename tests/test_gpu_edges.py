"""Empty and degenerate inputs are refused with SWIT_EINVAL (and a message) before any
work is queued: a plan with no shots, no receivers or a one-sample record; a misfit
over no traces; processing of no gathers; a plan group of no plans.  A plan is never
returned for a refused geometry."""
import ctypes

import numpy as np
import pytest


def _create(ns, ngmax, nt, nrec_per_shot=None):
    from swit_amd import _native
    lib = _native.load()
    g = _native.SwitGrid(20, 16, 5, nt, 1, 10.0, 0.001)
    ng = np.full(max(ns, 1), ngmax if nrec_per_shot is None else nrec_per_shot, np.int32)
    src = np.zeros((max(ns, 1), 2), np.float32)
    rec = np.zeros((max(ns, 1), max(ngmax, 1), 2), np.float32)
    sh = _native.SwitShots(ns, ngmax, ng.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                           src.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                           rec.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    h = ctypes.c_void_p()
    rc = lib.swit_plan_create(ctypes.byref(g), ctypes.byref(sh), _native.SWIT_PATH_BATCHED, ctypes.byref(h))
    return rc, h, lib


@pytest.mark.gpu
@pytest.mark.parametrize("ns,ngmax,nt", [(0, 3, 50), (2, 0, 50), (2, 3, 1), (2, 3, 0)])
def test_empty_geometry_is_refused(ns, ngmax, nt):
    from swit_amd import _native
    rc, h, lib = _create(ns, ngmax, nt)
    assert rc == _native.SWIT_EINVAL and not h.value
    assert lib.swit_last_error()


@pytest.mark.gpu
def test_smallest_geometry_is_accepted():
    from swit_amd import _native
    rc, h, lib = _create(1, 1, 2)
    assert rc == _native.SWIT_OK and h.value
    assert lib.swit_plan_destroy(h) == _native.SWIT_OK


@pytest.mark.gpu
def test_empty_misfit_processing_and_group_are_refused():
    import torch
    from swit_amd import _native, base, engine
    empty = torch.zeros((0, 100), dtype=torch.float32, device="cuda")
    with pytest.raises(_native.SwitError) as e:
        engine.misfit("Waveform", empty, empty, 0.001)
    assert e.value.code == _native.SWIT_EINVAL
    o = base.optimize('Waveform', 'LBFGS', 1, 0.05, 6000.0, 1000.0, 'Marine', 0, 0, 'None', 0.0, 0.0, 0, 0.0,
                      ['Max-Trace'], 0, 0, 0.0, 0.0)
    with pytest.raises(_native.SwitError) as e:
        engine.process_gathers(o, torch.zeros((0, 4, 100), dtype=torch.float32, device="cuda"), 0.001)
    assert e.value.code == _native.SWIT_EINVAL
    lib = _native.load()
    out = ctypes.c_void_p()
    assert lib.swit_group_create(None, 0, ctypes.byref(out)) == _native.SWIT_EINVAL and not out.value
