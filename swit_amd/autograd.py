"""Differentiable forward modelling: velocity model -> seismograms as a
torch.autograd.Function whose backward is the hand-written exact adjoint
(swit_forward_exact / swit_adjoint_exact of libswit_fd).

    from swit_amd.autograd import seismograms
    seis = seismograms(vp, plan, src, rho)          # (ns, ngmax, nt), differentiable in vp
    loss = l2_misfit(seis, obs); loss.backward()    # vp.grad = exact dJ/dvp

The gradient is that of the discrete scheme the forward runs (4th-order staggered
velocity-pressure, fdcore.f90:75-128), through K = den c^2 dt/dx and the PML padding
of the model (io.f90:231-300); density and the damping profile (including its vmin)
are held fixed.  The plan keeps one divergence history, so a backward belongs to the
last forward through that plan: a backward whose forward's history has been replaced
(another forward_exact, a new model, the autotuner) raises instead of returning the
gradient of another forward.  Create the plan with exact=True (whole history,
ns * nt * (nx+2npml) * ld floats) or exact='checkpoint' (states every ~sqrt(nt) steps,
segments recomputed in the backward; the same gradient bit for bit).
plan.set_option("streams", 2) runs each half of the shots on a stream of its own
(forward and backward; the same gradient bit for bit, about 9 % faster at 49 shots).
"""
from __future__ import annotations

import torch


class _Seismograms(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, plan, src, rho):
        plan.set_model(vp.detach().to(torch.float32), rho)
        seis = plan.forward_exact(src)
        ctx.plan = plan
        ctx.generation = plan.exact_generation()
        ctx.vp_dtype = vp.dtype
        return seis

    @staticmethod
    def backward(ctx, gseis):
        if ctx.plan.exact_generation() != ctx.generation:
            raise RuntimeError("the plan's divergence history no longer belongs to this forward (another "
                               "forward_exact, a new model or the autotuner ran on the plan since)")
        grad = ctx.plan.adjoint_exact(gseis.contiguous())
        return grad.to(ctx.vp_dtype), None, None, None


def seismograms(vp, plan, src, rho):
    """Differentiable seismograms (ns, ngmax, nt) of model vp (nx, nz) on `plan`
    (a batched plan created with exact=True), integrated sources src (ns, nt) and the
    fixed density rho (nx, nz)."""
    return _Seismograms.apply(vp, plan, src, rho)


def l2_misfit(seis, obs):
    """1/2 sum (seis - obs)^2, accumulated in FP64."""
    r = seis.to(torch.float64) - obs.to(torch.float64)
    return 0.5 * (r * r).sum()


def global_correlation_misfit(seis, obs):
    """-sum over live traces of the Pearson coefficient of the normalised traces
    (toolbox/misfit.py:338-355), FP64; traces with a zero norm are skipped."""
    s = seis.to(torch.float64)
    o = obs.to(torch.float64)
    live = (s.norm(dim=-1) > 0) & (o.norm(dim=-1) > 0)
    s, o = s[live], o[live]
    s = s / s.norm(dim=-1, keepdim=True)
    o = o / o.norm(dim=-1, keepdim=True)
    sc = s - s.mean(-1, keepdim=True)
    oc = o - o.mean(-1, keepdim=True)
    c = (sc * oc).sum(-1) / torch.sqrt((sc * sc).sum(-1) * (oc * oc).sum(-1))
    return -c.sum()
