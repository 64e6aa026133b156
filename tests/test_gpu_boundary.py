"""The C ABI's refusals and state rules beyond argument checks:

  * swit_plan_set_model refuses a model with a vp or rho value that is not finite and
    positive, or whose largest vp breaks the stability limit dt < sqrt(3/8) dx / max(vp)
    (toolbox/base.py:57-59), and leaves the plan without a model;
  * a plan created with SWIT_PLAN_NO_STRIPS models with the same seismograms as a
    strip-holding plan, holds less memory, and refuses strip-saving forwards and
    adjoints; its autotuner tunes the forward only;
  * the exact adjoint keeps its own copy of the forward's sources, so a reference
    forward with other sources between swit_forward_exact and swit_adjoint_exact leaves
    the (checkpointed) gradient unchanged; the history's generation names the forward
    it belongs to, and torch.autograd refuses a backward whose history was replaced;
  * Plan.forward refuses an `out` tensor of the wrong shape, dtype or layout."""
import numpy as np
import pytest

from helpers import integrated_sources
from swit_amd import workloads


def _plan(w, **kw):
    import torch
    from swit_amd.engine import Plan
    p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, **kw)
    p.set_model(torch.from_numpy(w.vp_init), torch.from_numpy(workloads.gardner(w.vp_init)))
    return p


@pytest.mark.gpu
@pytest.mark.parametrize("path", ["batched", "plain"])
def test_set_model_refuses_unstable_or_invalid_models(path):
    import torch
    from swit_amd._native import SWIT_EINVAL, SWIT_ESTATE, SwitError
    w = workloads.small(seed=61, nx=40, nz=30, npml=8, nt=100, ns=2)
    p = _plan(w, path=path)
    src = torch.from_numpy(integrated_sources(w))
    ref = p.forward(src, save_strips=False).cpu().numpy()
    rho = torch.from_numpy(workloads.gardner(w.vp_init))
    limit = np.sqrt(3.0 / 8.0) * np.float32(w.dx) / np.float32(w.dt)      # max vp allowed, m/s
    bad = []
    v = w.vp_init.copy()
    v[5, 7] = np.float32(limit * 1.001)                 # one cell over the stability limit
    bad.append((v, rho))
    v = w.vp_init.copy()
    v[3, 3] = np.nan
    bad.append((v, rho))
    v = w.vp_init.copy()
    v[0, 0] = 0.0
    bad.append((v, rho))
    r = rho.clone()
    r[2, 9] = -1.0
    bad.append((w.vp_init, r))
    for vp, rr in bad:
        with pytest.raises(SwitError) as e:
            p.set_model(torch.from_numpy(np.ascontiguousarray(vp)), rr)
        assert e.value.code == SWIT_EINVAL
        # no model is loaded after a refusal
        with pytest.raises(SwitError) as e:
            p.forward(src, save_strips=False)
        assert e.value.code == SWIT_ESTATE
    v = w.vp_init.copy()
    v[5, 7] = np.float32(limit * 0.999)                 # just inside the limit: accepted
    p.set_model(torch.from_numpy(v), rho)
    p.set_model(torch.from_numpy(w.vp_init), rho)
    np.testing.assert_array_equal(p.forward(src, save_strips=False).cpu().numpy(), ref)
    p.close()


@pytest.mark.gpu
def test_plan_without_strips():
    import torch
    from swit_amd._native import SWIT_EINVAL, SWIT_ESTATE, SwitError
    w = workloads.small(seed=62, nx=50, nz=30, npml=8, nt=160, ns=3)
    src = torch.from_numpy(integrated_sources(w))
    full = _plan(w)
    bare = _plan(w, strips="none")
    assert bare.scratch_bytes < full.scratch_bytes
    a = full.forward(src, save_strips=True).cpu().numpy()
    b = bare.forward(src, save_strips=False).cpu().numpy()
    c = bare.forward(src, save_strips=2).cpu().numpy()
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a, c)
    with pytest.raises(SwitError) as e:
        bare.forward(src, save_strips=True)
    assert e.value.code == SWIT_EINVAL
    with pytest.raises(SwitError) as e:
        bare.adjoint(src, torch.from_numpy(a))
    assert e.value.code == SWIT_ESTATE
    ch = bare.autotune(steps=3)
    assert ch["fwd_us"] > 0 and ch["bwd_us"] == 0
    np.testing.assert_array_equal(bare.forward(src, save_strips=False).cpu().numpy(), a)
    plain = _plan(w, strips="none", path="plain")
    np.testing.assert_array_equal(plain.forward(src, save_strips=False).cpu().numpy(), a)
    with pytest.raises(SwitError) as e:
        plain.forward(src, save_strips=True)
    assert e.value.code == SWIT_EINVAL
    for p in (full, bare, plain):
        p.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [True, "checkpoint"])
def test_exact_adjoint_keeps_its_forward(mode):
    import torch
    from swit_amd.autograd import l2_misfit, seismograms
    w = workloads.small(seed=63, nx=44, nz=30, npml=8, nt=150, ns=2)
    src = torch.from_numpy(integrated_sources(w)).cuda()
    p = _plan(w, exact=mode)
    g0 = p.forward_exact(src) * 0.5 + 1e-3
    ref = p.adjoint_exact(g0).cpu().numpy()
    gen = p.exact_generation()
    assert gen > 0
    p.forward_exact(src)
    assert p.exact_generation() == gen + 1
    # a reference forward and adjoint with other sources in between
    p.forward(src * 3.0, save_strips=True)
    p.adjoint(src * 3.0, g0)
    assert p.exact_generation() == gen + 1
    np.testing.assert_array_equal(p.adjoint_exact(g0).cpu().numpy(), ref)
    # a new model invalidates the history
    p.set_model(torch.from_numpy(w.vp_init), torch.from_numpy(workloads.gardner(w.vp_init)))
    assert p.exact_generation() == 0
    # autograd: a backward whose history was replaced raises
    rho = torch.from_numpy(workloads.gardner(w.vp_init)).cuda()
    vp = torch.tensor(w.vp_init, device="cuda", requires_grad=True)
    s1 = seismograms(vp, p, src, rho)
    s2 = seismograms(vp, p, src * 2.0, rho)
    with pytest.raises(RuntimeError):
        l2_misfit(s1, torch.zeros_like(s1)).backward()
    l2_misfit(s2, torch.zeros_like(s2)).backward()
    assert torch.isfinite(vp.grad).all()
    p.close()


@pytest.mark.gpu
def test_forward_out_is_checked():
    import torch
    w = workloads.small(seed=64, nx=36, nz=24, npml=6, nt=80, ns=2)
    p = _plan(w)
    src = torch.from_numpy(integrated_sources(w))
    shape = (p.ns, p.ngmax, p.nt)
    ok = torch.empty(shape, dtype=torch.float32, device="cuda")
    assert p.forward(src, out=ok, save_strips=False) is ok
    for bad in (torch.empty((p.ns, p.ngmax, p.nt - 1), dtype=torch.float32, device="cuda"),
                torch.empty(shape, dtype=torch.float64, device="cuda"),
                torch.empty(shape, dtype=torch.float32),
                torch.empty((p.ns, p.nt, p.ngmax), dtype=torch.float32, device="cuda").transpose(1, 2)):
        with pytest.raises(ValueError):
            p.forward(src, out=bad, save_strips=False)
    p.close()
