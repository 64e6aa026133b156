"""engine.Lanes: the shots of one job split into plans on streams of their own.  Its
seismograms equal one plan's bit for bit; its gradient and illuminations, FP64 sums of
the lanes' sums, equal one plan's to FP64 rounding; its misfit equals the one-plan
misfit of the same gathers."""
import numpy as np
import pytest

from helpers import integrated_sources
from swit_amd import workloads


@pytest.mark.gpu
@pytest.mark.parametrize("lanes,strips,ragged", [(2, "f32", False), (3, "f16", False), (1, "f32", False),
                                                 (2, "f32", True)])
def test_lanes_match_one_plan(lanes, strips, ragged):
    import torch
    from swit_amd.engine import Lanes, Plan, misfit
    w = workloads.small(seed=41, nx=70, nz=40, npml=10, nt=300, ns=5)
    if ragged:
        # fewer receivers on the last shots: the second lane's plan has a smaller ngmax
        w.rec_xz = [w.rec_xz[i, :w.rec_xz.shape[1] - 9 * i] for i in range(w.ns)]
    src = torch.from_numpy(integrated_sources(w)).cuda()
    vt, rt = torch.from_numpy(w.vp_true).cuda(), torch.from_numpy(workloads.gardner(w.vp_true)).cuda()
    vi, ri = torch.from_numpy(w.vp_init).cuda(), torch.from_numpy(workloads.gardner(w.vp_init)).cuda()
    p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, strips=strips)
    p.set_model(vt, rt)
    obs = p.forward(src, save_strips=False).clone()
    if ragged:                                         # padded receiver traces are dead (zero) traces
        for i in range(w.ns):
            assert not torch.any(obs[i, len(w.rec_xz[i]):])
    p.set_model(vi, ri)
    syn = p.forward(src, save_strips=True)
    J1, adj, _ = misfit("Waveform", obs, syn, w.dt)
    g1 = p.adjoint(src, adj)
    L = Lanes(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, lanes=lanes, strips=strips)
    L.set_model(vt, rt)
    np.testing.assert_array_equal(L.forward(src, save_strips=False).cpu().numpy(), obs.cpu().numpy())
    L.set_model(vi, ri)
    for _ in range(2):                                 # a second call reuses the lanes' buffers
        J, g, fw, bk = L.gradient(src, obs, "Waveform", w.dt)
        torch.cuda.synchronize()
        assert float(J) == pytest.approx(float(J1), rel=1e-13)
        for a, b in zip((g, fw, bk), g1):
            a, b = a.cpu().numpy(), b.cpu().numpy()
            assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(b)
    assert float(L.misfit(src, obs, "Waveform", w.dt)) == pytest.approx(float(J1), rel=1e-13)
    L.close()
    p.close()
