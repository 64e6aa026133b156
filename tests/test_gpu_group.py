"""Plan groups: uneven grids, shot counts, record lengths and strip precisions in one
graph launch give exactly what each plan gives alone (bit-identical seismograms,
gradients and illuminations), across repeated calls and a forward-group change."""
import numpy as np
import pytest

from helpers import integrated_sources
from swit_amd import workloads


def _cases():
    return [dict(seed=41, nx=60, nz=30, npml=10, nt=180, ns=3),
            dict(seed=42, nx=45, nz=52, npml=12, nt=120, ns=2, dx=15.0),
            dict(seed=43, nx=33, nz=20, npml=8, nt=240, ns=4)]


def _plan(w, strips):
    import torch
    from swit_amd.engine import Plan
    p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, path="batched", strips=strips)
    p.set_model(torch.from_numpy(w.vp_init), torch.from_numpy(workloads.gardner(w.vp_init)))
    return p


@pytest.mark.gpu
def test_group_matches_single_plans():
    import torch
    from swit_amd.engine import PlanGroup
    ws = [workloads.small(**c) for c in _cases()]
    strips = ["f32", "bf16", "f32"]
    srcs = [integrated_sources(w) for w in ws]
    # alone
    ref = []
    for w, st, s in zip(ws, strips, srcs):
        p = _plan(w, st)
        seis = p.forward(torch.from_numpy(s), save_strips=True)
        adj = (seis * 1e-2).contiguous()
        g = [t.cpu().numpy() for t in p.adjoint(torch.from_numpy(s), adj)]
        ref.append((seis.cpu().numpy(), adj.cpu().numpy(), g))
        p.close()
    # grouped
    plans = [_plan(w, st) for w, st in zip(ws, strips)]
    grp = PlanGroup(plans)
    for rep in range(2):
        if rep == 1:
            plans[2].set_option("fwd_group", 2)       # captured arguments change: recapture
        seis = grp.forward(srcs, save_strips=True)
        outs = grp.adjoint(srcs, [r[1] for r in ref])
        torch.cuda.synchronize()
        for i, (r_seis, _, r_g) in enumerate(ref):
            np.testing.assert_array_equal(seis[i].cpu().numpy(), r_seis, err_msg=f"plan {i} rep {rep}")
            for a, b in zip(outs[i], r_g):
                np.testing.assert_array_equal(a.cpu().numpy(), b, err_msg=f"plan {i} rep {rep}")
    # an adjoint after a forward without strips is refused
    grp.forward(srcs, save_strips=False)
    from swit_amd._native import SwitError
    with pytest.raises(SwitError):
        grp.adjoint(srcs, [r[1] for r in ref])
    grp.close()
    for p in plans:
        p.close()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["Waveform", "Globalcorrelation"])
def test_group_gradient_equals_the_steps_one_by_one(kind):
    """swit_group_gradient (one graph: forward, misfit and adjoint source, adjoint per
    plan, plans on concurrent branches) gives each plan the bits of swit_forward,
    swit_misfit and swit_adjoint run one after the other; a second call (the graph
    replayed) gives the same bits again."""
    import torch
    from helpers import integrated_sources
    from swit_amd import workloads
    from swit_amd.engine import Plan, PlanGroup, misfit
    ws = [workloads.small(seed=51, nx=50, nz=34, npml=10, nt=210, ns=3),
          workloads.small(seed=52, nx=64, nz=28, npml=8, nt=160, ns=2, dx=12.0, dt=0.0012)]
    plans, srcs, obs, ref = [], [], [], []
    for w in ws:
        src = torch.from_numpy(integrated_sources(w)).cuda()
        p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz)
        p.set_model(torch.from_numpy(w.vp_true).cuda(), torch.from_numpy(workloads.gardner(w.vp_true)).cuda())
        o = p.forward(src, save_strips=False).clone()
        p.set_model(torch.from_numpy(w.vp_init).cuda(), torch.from_numpy(workloads.gardner(w.vp_init)).cuda())
        syn = p.forward(src, save_strips=True)
        J, adj, _ = misfit(kind, o, syn, w.dt)
        g = p.adjoint(src, adj)
        ref.append((float(J), syn.cpu().numpy(), [t.cpu().numpy() for t in g]))
        plans.append(p)
        srcs.append(src)
        obs.append(o)
    grp = PlanGroup(plans)
    for _ in range(2):
        Js, seis, outs = grp.gradient(srcs, obs, kind)
        torch.cuda.synchronize()
        for (J0, s0, g0), J, s, g in zip(ref, Js, seis, outs):
            assert float(J) == J0
            np.testing.assert_array_equal(s.cpu().numpy(), s0)
            for a, b in zip(g, g0):
                np.testing.assert_array_equal(a.cpu().numpy(), b)
    grp.close()
    for p in plans:
        p.close()
