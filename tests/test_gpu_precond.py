"""grad_precond on the GPU (swit_amd.engine.grad_precond, and the drop-in's
swit_amd.postprocess.grad_precond over it) against the oracle's numpy restatement of
toolbox/postprocess.py:20-111 (oracle/precond.py, pinned by tests/test_oracle_precond.py):
marine and land tapers, with and without gradient smoothing
and mask, small and large spans, grids under 40 cells (span min(nx, nz)/2).
Tolerance: relative L2 1e-12 (FP64 sums in another order).  With the land taper the
reference rounds grad * taper to float32 (in-place product on the float32 gradient); a
taper that differs from numpy's in the last FP64 bit can move that product by one
float32 ulp in a few cells, so those cases are held to relative L2 1e-8 (about 1e-7 on
each such cell)."""
import numpy as np
import pytest

from oracle import precond as P
from swit_amd import postprocess


def _case(nx, nz, seed):
    rng = np.random.default_rng(seed)
    grad = rng.standard_normal((nx, nz)).astype(np.float32)
    forw = (rng.random((nx, nz)) ** 2 * 3.0 + 1e-3).astype(np.float32)
    back = (rng.random((nx, nz)) ** 3 * 7.0).astype(np.float32)
    return grad, forw, back


@pytest.mark.gpu
@pytest.mark.parametrize("nx,nz,marine,mute,thred,smooth,use_mask", [
    (120, 70, True, 8, 0.0, 3, False),
    (120, 70, True, 0, 0.0, 0, False),
    (120, 70, False, 12, 0.05, 5, True),
    (481, 141, True, 10, 0.0, 2, False),     # the bench grid: 81-tap smoothing of the illuminations
    (30, 24, False, 6, 0.1, 2, False),       # under 40 cells: span min(nx, nz)/2
    (60, 50, False, 0, 0.0, 4, True),
])
def test_device_grad_precond_matches_numpy(nx, nz, marine, mute, thred, smooth, use_mask):
    from swit_amd.engine import grad_precond
    grad, forw, back = _case(nx, nz, nx + nz + mute)
    mask = np.random.default_rng(3).random((nx, nz)) if use_mask else None

    class M:
        pass
    simu, optim = M(), M()
    simu.model = M()
    simu.model.nx, simu.model.nz = nx, nz
    optim.vpmax, optim.marine_or_land = 3500.0, ('Marine' if marine else 'Land')
    optim.grad_mute, optim.grad_thred, optim.grad_smooth, optim.grad_mask = mute, thred, smooth, mask
    ref = P.precondition(grad, forw, back, 3500.0, marine, mute, thred, smooth, mask).ravel()
    got = grad_precond(grad, forw, back, 3500.0, marine, mute, thred, smooth, mask).cpu().numpy().ravel()
    via = postprocess.grad_precond(simu, optim, grad.ravel(), forw.ravel(), back.ravel())
    np.testing.assert_array_equal(via, got)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"{nx}x{nz} marine={marine} mute={mute} smooth={smooth}: rel L2 {err:.2e}")
    land_taper = not marine and mute > 0
    assert err <= (1e-8 if land_taper else 1e-12)
    assert np.abs(got).max() == pytest.approx(3500.0, rel=1e-15)
