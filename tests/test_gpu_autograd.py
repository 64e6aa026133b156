"""Exact adjoint behind torch.autograd (swit_amd.autograd): the seismograms equal the
ordinary forward's bit for bit; the gradient equals torch autograd through an FP64
restatement of the same scheme (tolerance: relative L2 1e-4, FP32 against FP64); and
the directional derivative <g, dm> matches FP64 central finite differences of the FP64
oracle at eps = 1e-2 .. 1e-5 (tolerance 1e-3 relative) for the L2 and the
global-correlation misfits, on the 161x61, 8-shot, 1500-step case (whole history and
checkpoints) and on two shots of the 481x141, 4000-step benchmark grid (checkpoints)."""
import numpy as np
import pytest

from helpers import integrated_sources, rel_l2, torch_forward
from oracle import oracle as O
from swit_amd import workloads


def _plan(w, exact=True):
    from swit_amd.engine import Plan
    return Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, path="batched", exact=exact)


@pytest.mark.gpu
def test_exact_forward_equals_forward():
    import torch
    w = workloads.small(seed=51, ns=3, nt=200)
    src = torch.from_numpy(integrated_sources(w))
    p = _plan(w)
    p.set_model(torch.from_numpy(w.vp_init), torch.from_numpy(workloads.gardner(w.vp_init)))
    a = p.forward(src, save_strips=True).cpu().numpy()
    b = p.forward_exact(src).cpu().numpy()
    p.close()
    np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("free_surface", [1, 0])
def test_exact_gradient_matches_autograd_restatement(free_surface):
    import torch
    from swit_amd.autograd import l2_misfit, seismograms
    from swit_amd.engine import Plan
    w = workloads.small(seed=52, nx=40, nz=30, npml=8, nt=220, ns=2)
    src = integrated_sources(w)
    rho = workloads.gardner(w.vp_init)
    p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, free_surface=free_surface,
             path="batched", exact=True)
    p.set_model(torch.from_numpy(w.vp_true), torch.from_numpy(workloads.gardner(w.vp_true)))
    obs = p.forward_exact(torch.from_numpy(src)).clone()
    vp = torch.tensor(w.vp_init, device="cuda", requires_grad=True)
    seis = seismograms(vp, p, torch.from_numpy(src).cuda(), torch.from_numpy(rho).cuda())
    l2_misfit(seis, obs).backward()
    g_gpu = vp.grad.double().cpu().numpy()
    # FP64 autograd through the restatement, same frozen damping
    m = O.Model(w.vp_init, rho, w.dx, w.npml, "f64")
    vt = torch.tensor(w.vp_init.astype(np.float64), requires_grad=True)
    st = torch_forward(vt, rho, w.dx, w.dt, w.npml, w.src_xz, w.rec_xz, src, m.damp, free_surface)
    obs_np = obs.cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(st.detach().numpy(), seis.detach().cpu().numpy(), rtol=0,
                               atol=1e-5 * np.abs(obs_np).max())
    (0.5 * ((st - torch.from_numpy(obs_np)) ** 2).sum()).backward()
    g_ref = vt.grad.numpy()
    err = rel_l2(g_gpu, g_ref)
    print(f"free_surface={free_surface}: exact gradient relL2 vs FP64 autograd {err:.3e}")
    assert err <= 1e-4
    p.close()


def _finite_difference_check(w, dm, mode, sel=None):
    """<g, dm> of the exact adjoint (FP32 kernels, plan mode `mode`) against FP64 central
    finite differences of the FP64 oracle along dm at eps = 1e-2 .. 1e-5, for the L2 and
    global-correlation misfits; relative error <= 1e-3.  The damping profile is frozen
    at the unperturbed model's (the exact adjoint holds it fixed)."""
    import torch
    from helpers import oracle_threads
    from swit_amd.autograd import global_correlation_misfit, l2_misfit, seismograms
    from swit_amd.engine import Plan
    sel = slice(None) if sel is None else sel
    sxz, rxz = w.src_xz[sel], w.rec_xz[sel]
    src = integrated_sources(w)[sel]
    rho = workloads.gardner(w.vp_init)
    p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, sxz, rxz, path="batched", exact=mode)
    p.set_model(torch.from_numpy(w.vp_true), torch.from_numpy(workloads.gardner(w.vp_true)))
    obs = p.forward_exact(torch.from_numpy(src)).clone()
    obs_np = obs.cpu().numpy().astype(np.float64)
    vp0 = w.vp_init.astype(np.float32)
    m0 = O.Model(vp0, rho, w.dx, w.npml, "f64")
    th = oracle_threads()
    fns = {"l2": l2_misfit, "gc": global_correlation_misfit}
    gd = {}
    for kind, fn in fns.items():
        vp = torch.tensor(vp0, device="cuda", requires_grad=True)
        seis = seismograms(vp, p, torch.from_numpy(src).cuda(), torch.from_numpy(rho).cuda())
        fn(seis, obs).backward()
        gd[kind] = float(np.sum(vp.grad.double().cpu().numpy() * dm))
    p.close()

    def J_oracle(vp):
        m = O.Model(vp, rho, w.dx, w.npml, "f64", damp_override=m0.damp, exact_input=True)
        s = torch.from_numpy(O.forward(m, w.nt, w.dt, sxz, rxz, src, save_strips=False, threads=th)["seis"])
        ot = torch.from_numpy(obs_np)
        return {k: float(fn(s, ot)) for k, fn in fns.items()}

    worst = 0.0
    for eps in (1e-2, 1e-3, 1e-4, 1e-5):
        jp, jm = J_oracle(vp0 + eps * dm), J_oracle(vp0 - eps * dm)
        for kind in fns:
            fd = (jp[kind] - jm[kind]) / (2 * eps)
            rel = abs(gd[kind] - fd) / abs(fd)
            worst = max(worst, rel)
            print(f"{w.nx}x{w.nz}x{w.nt} {mode} {kind} eps={eps:g}: <g,dm> {gd[kind]:.9e}  FD {fd:.9e}  "
                  f"rel {rel:.2e}", flush=True)
            assert rel <= 1e-3, (kind, eps, rel)
    return worst


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [True, "checkpoint"])
def test_exact_gradient_finite_differences_161x61(mode):
    """The autograd check configuration: 161x61, nt 1500, 8 shots, smooth perturbation
    scaled to 50 m/s, with the whole divergence history and with checkpoints."""
    w = workloads.config3()
    _finite_difference_check(w, w.extra["dm"].astype(np.float64), mode)


@pytest.mark.gpu
def test_exact_gradient_finite_differences_481x141x4000():
    """The benchmark's grid and record length (481x141, npml 40, nt 4000), shots 0 and
    24, a smooth random perturbation (N(0,1) noise Gaussian-filtered with sigma 6 cells,
    scaled to 50 m/s); checkpointed sweep, as the benchmark's autograd leg runs it when
    the whole history does not fit."""
    from scipy.ndimage import gaussian_filter
    w = workloads.config2()
    rng = np.random.default_rng(22)
    dm = gaussian_filter(rng.standard_normal((w.nx, w.nz)), 6)
    dm = dm / np.abs(dm).max() * 50.0
    worst = _finite_difference_check(w, dm, "checkpoint", sel=[0, 24])
    print(f"481x141x4000: worst relative error of <g,dm> against FP64 finite differences {worst:.2e}")


@pytest.mark.gpu
@pytest.mark.parametrize("nt", [220, 2, 17])
def test_checkpointed_exact_adjoint_equals_full_history(nt):
    """exact='checkpoint' recomputes each ~sqrt(nt)-step segment from a stored state:
    the gradient is bit-identical to the full-history one (a partial last segment,
    a one-step job and a square step count included)."""
    import torch
    from swit_amd.engine import Plan
    w = workloads.small(seed=53, nx=44, nz=28, npml=8, nt=nt, ns=3)
    src = torch.from_numpy(integrated_sources(w))
    out = {}
    for mode in (True, "checkpoint"):
        p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, path="batched", exact=mode)
        p.set_model(torch.from_numpy(w.vp_init), torch.from_numpy(workloads.gardner(w.vp_init)))
        seis = p.forward_exact(src)
        g = p.adjoint_exact(seis * 0.5 + 1e-3)
        out[str(mode)] = (seis.cpu().numpy(), g.cpu().numpy(), p.scratch_bytes)
        p.close()
    np.testing.assert_array_equal(out["checkpoint"][0], out["True"][0])
    np.testing.assert_array_equal(out["checkpoint"][1], out["True"][1])
    # with one step the divergence acts on the zero initial state: no K dependence
    assert (np.abs(out["True"][1]).max() > 0) == (nt > 2)
    print(nt, "scratch bytes full / checkpoint:", out["True"][2], out["checkpoint"][2])


@pytest.mark.gpu
@pytest.mark.parametrize("nx,nz,npml,nt,ns,mode,fs", [
    (44, 28, 8, 160, 3, True, 1),
    (40, 30, 59, 120, 2, True, 1),        # free surface deep in a wide absorbing layer
    (70, 150, 20, 130, 2, "checkpoint", 1),
    (33, 9, 3, 101, 2, True, 1),           # thinnest absorbing layer and interior
    (100, 40, 12, 90, 3, "checkpoint", 0),
    (50, 60, 45, 80, 2, True, 0),
])
def test_fused_exact_step_equals_separate_passes(nx, nz, npml, nt, ns, mode, fs):
    """The fused exact-adjoint step kernel (one launch per reverse step) gives the seven
    separate passes' gradient bit for bit, with the full history and with checkpoints,
    with and without the free surface, for tilings that put the mirror rows at many
    positions in the tiles' windows.  Both add the receivers that share a recording cell
    (the case's last receiver shares a surface receiver's) in receiver order."""
    import torch
    from swit_amd.engine import Plan
    w = workloads.small(seed=nx + nz + npml, nx=nx, nz=nz, npml=npml, nt=nt, ns=ns)
    src = torch.from_numpy(integrated_sources(w))
    out = {}
    for fused in (1, 0):
        for graph in (1, 0):
            p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, free_surface=fs,
                     path="batched", exact=mode)
            p.set_option("exact_fused", fused)
            p.set_option("graph", graph)
            p.set_model(torch.from_numpy(w.vp_init), torch.from_numpy(workloads.gardner(w.vp_init)))
            seis = p.forward_exact(src)
            out[fused, graph] = p.adjoint_exact(seis * 0.5 + 1e-3).cpu().numpy()
            p.close()
    assert np.abs(out[0, 1]).max() > 0
    for key, g in out.items():
        np.testing.assert_array_equal(g, out[0, 1], err_msg=str(key))
