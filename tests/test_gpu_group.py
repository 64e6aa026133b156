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
