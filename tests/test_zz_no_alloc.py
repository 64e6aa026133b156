"""swit_grad_precond_noalloc, the preconditioner entry point that allocates nothing (its
workspace is a device-resident static array), against swit_grad_precond with a
workspace the caller allocates.  Both run the same kernels in the same order, so their
results must be identical bit for bit.

Calls in a row share the static workspace and find the previous call's contents in it:
a large call, then a smaller one, then the large one again, each with other parameters,
must each equal a fresh workspaced call.  The same holds on a stream other than the
default one.  A model whose workspace exceeds the static array is refused with
SWIT_EINVAL before anything runs."""
import ctypes

import numpy as np
import pytest

from swit_amd import _native


def _inputs(torch, nx, nz, seed, mask):
    rng = np.random.default_rng(seed)
    grad = rng.standard_normal((nx, nz)).astype(np.float32)
    forw = (rng.random((nx, nz)) ** 2 * 3.0 + 1e-3).astype(np.float32)
    back = (rng.random((nx, nz)) ** 3 * 7.0).astype(np.float32)
    dev = torch.device("cuda", 0)
    t = [torch.from_numpy(a).to(dev) for a in (grad, forw, back)]
    m = torch.from_numpy(rng.random((nx, nz))).to(dev) if mask else None
    return t, m


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _run(torch, L, prm, t, m, stream, noalloc):
    nx, nz = prm.nx, prm.nz
    out = torch.full((nx, nz), float("nan"), dtype=torch.float64, device=t[0].device)
    sp = ctypes.c_void_p(stream.cuda_stream)
    if noalloc:
        rc = L.swit_grad_precond_noalloc(ctypes.byref(prm), _ptr(t[0]), _ptr(t[1]), _ptr(t[2]), _ptr(m),
                                         _ptr(out), sp)
    else:
        nb = ctypes.c_size_t()
        _native.check(L.swit_grad_precond_work(ctypes.byref(prm), ctypes.byref(nb)))
        work = torch.full((int(nb.value),), 0xA5, dtype=torch.uint8, device=t[0].device)
        rc = L.swit_grad_precond(ctypes.byref(prm), _ptr(t[0]), _ptr(t[1]), _ptr(t[2]), _ptr(m), _ptr(out),
                                 _ptr(work), nb.value, sp)
    _native.check(rc)
    stream.synchronize()
    return out.cpu().numpy()


# (nx, nz, marine, grad_mute, grad_thred, grad_smooth, mask): a large land case, a
# smaller marine one, and the large one again
_CALLS = [
    (801, 267, False, 12, 0.001, 4, True),
    (120, 70, True, 8, 0.0, 3, False),
    (801, 267, False, 12, 0.001, 4, True),
]


@pytest.mark.gpu
@pytest.mark.parametrize("side_stream", [False, True])
def test_noalloc_equals_workspace_entry_on_calls_in_a_row(side_stream):
    import torch
    L = _native.load()
    stream = torch.cuda.Stream(device=0) if side_stream else torch.cuda.current_stream(0)
    for k, (nx, nz, marine, mute, thred, smooth, mask) in enumerate(_CALLS):
        print(f"call {k}: {nx}x{nz} marine={marine} mute={mute} smooth={smooth} mask={mask} "
              f"side_stream={side_stream}", flush=True)
        prm = _native.SwitPrecond(nx, nz, int(marine), mute, thred, smooth, 3500.0)
        t, m = _inputs(torch, nx, nz, 100 + k, mask)
        want = _run(torch, L, prm, t, m, stream, noalloc=False)
        got = _run(torch, L, prm, t, m, stream, noalloc=True)
        assert np.isfinite(want).all()
        np.testing.assert_array_equal(got, want)
        assert np.abs(got).max() == pytest.approx(3500.0, rel=1e-15)


@pytest.mark.gpu
def test_noalloc_refuses_models_beyond_its_workspace():
    import torch
    L = _native.load()
    t, _ = _inputs(torch, 4, 4, 1, False)
    out = torch.empty((4, 4), dtype=torch.float64, device="cuda:0")
    # 1500 x 1500 cells need 72 MB of workspace, above the 64 MiB array; the check comes
    # before any launch, so the small buffers are never read
    prm = _native.SwitPrecond(1500, 1500, 1, 0, 0.0, 0, 3500.0)
    nb = ctypes.c_size_t()
    _native.check(L.swit_grad_precond_work(ctypes.byref(prm), ctypes.byref(nb)))
    assert nb.value > _native.SWIT_PRECOND_SCRATCH_BYTES
    rc = L.swit_grad_precond_noalloc(ctypes.byref(prm), _ptr(t[0]), _ptr(t[1]), _ptr(t[2]), None, _ptr(out),
                                     ctypes.c_void_p(torch.cuda.current_stream(0).cuda_stream))
    assert rc == _native.SWIT_EINVAL
    assert b"workspace" in L.swit_last_error()
