"""Two streams (SWIT_OPT_STREAMS 2): the first and second half of the shots run their
time loops on two streams, forked from and joined back into the caller's stream.  Every
result must equal the one-stream run bit for bit: seismograms, saved strips (FP32 and
16-bit), gradient and illuminations, with graphs and with direct launches, with the
one-step and the two-step backward kernel, for odd and even shot counts (one shot runs
on one stream), and after option changes that drop the captured graphs."""
import numpy as np
import pytest

from helpers import integrated_sources
from swit_amd import workloads


def _run(w, strips, steps, streams, graph, src, obs):
    import torch
    from swit_amd.engine import Plan
    p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, strips=strips)
    p.set_model(torch.from_numpy(w.vp_init), torch.from_numpy(workloads.gardner(w.vp_init)))
    p.set_option("bwd_steps", steps)
    p.set_option("bwd_balance", 1 if steps == 1 else 0)
    p.set_option("streams", streams)
    p.set_option("graph", graph)
    out = []
    for rep in range(2):                     # the second pass replays the captured graphs
        syn = p.forward(torch.from_numpy(src), save_strips=True)
        strip = p.saved_strips().cpu().numpy()
        adj = (syn.cpu().numpy() - obs).astype(np.float32)
        g = [t.cpu().numpy() for t in p.adjoint(torch.from_numpy(src), torch.from_numpy(adj))]
        out.append([syn.cpu().numpy(), strip] + g)
    p.close()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("nx,nz,npml,nt,ns,strips,steps", [
    (60, 40, 10, 200, 2, "f32", 1),
    (60, 40, 10, 201, 5, "f32", 2),
    (48, 32, 12, 180, 3, "f16", 1),
    (48, 32, 12, 181, 4, "bf16", 2),
    (40, 30, 10, 120, 1, "f32", 1),          # one shot: one stream
])
@pytest.mark.parametrize("graph", [1, 0])
def test_two_streams_match_one_stream(nx, nz, npml, nt, ns, strips, steps, graph):
    from oracle import oracle as O
    w = workloads.small(seed=nx + ns + nt, nx=nx, nz=nz, npml=npml, nt=nt, ns=ns)
    src = integrated_sources(w)
    m = O.Model(w.vp_true, workloads.gardner(w.vp_true), w.dx, w.npml, "f32")
    obs = O.forward(m, w.nt, w.dt, w.src_xz, w.rec_xz, src, save_strips=False)["seis"]
    one = _run(w, strips, steps, 1, graph, src, obs)
    two = _run(w, strips, steps, 2, graph, src, obs)
    names = ("seismograms", "strips", "gradient", "forward illumination", "backward illumination")
    for rep in range(2):
        for name, a, b in zip(names, one[rep], two[rep]):
            np.testing.assert_array_equal(b, a, err_msg=f"{name}, pass {rep}")
    assert np.abs(one[0][2]).max() > 0


@pytest.mark.gpu
def test_two_streams_after_option_changes_and_profile():
    """Changing the backward kernel or tile width drops the captured graphs of both
    streams; the kernel-time counters still count one launch per time step."""
    import torch
    from swit_amd import _native
    from swit_amd.engine import Plan
    w = workloads.small(seed=9, nx=56, nz=36, npml=10, nt=150, ns=4)
    src = integrated_sources(w)
    p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz)
    p.set_model(torch.from_numpy(w.vp_init), torch.from_numpy(workloads.gardner(w.vp_init)))
    with pytest.raises(_native.SwitError):
        p.set_option("streams", 3)
    res = {}
    for streams in (1, 2):
        p.set_option("streams", streams)
        for steps, tx in ((2, 24), (1, 24), (1, 16)):
            p.set_option("bwd_steps", steps)
            p.set_option("bwd_tile_x", tx)
            syn = p.forward(torch.from_numpy(src), save_strips=True)
            g = p.adjoint(torch.from_numpy(src), syn * 1e-3)[0]
            res[streams, steps, tx] = (syn.cpu().numpy(), g.cpu().numpy())
    for key in ((2, 24), (1, 24), (1, 16)):
        np.testing.assert_array_equal(res[(2,) + key][0], res[(1,) + key][0])
        np.testing.assert_array_equal(res[(2,) + key][1], res[(1,) + key][1])
    p.set_option("profile", 1)
    syn = p.forward(torch.from_numpy(src), save_strips=True)
    p.adjoint(torch.from_numpy(src), syn * 1e-3)
    _, nf = p.kernel_time("forward")
    _, nb = p.kernel_time("backward")
    assert nf == w.nt - 1 and nb == w.nt - 1
    p.close()


def _run_exact(w, mode, streams, graph, fused, src, obs):
    import torch
    from swit_amd.engine import Plan
    p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, path="batched", exact=mode)
    p.set_model(torch.from_numpy(w.vp_init), torch.from_numpy(workloads.gardner(w.vp_init)))
    p.set_option("streams", streams)
    p.set_option("graph", graph)
    p.set_option("exact_fused", fused)
    out = []
    for rep in range(2):
        syn = p.forward_exact(torch.from_numpy(src))
        gseis = (syn - torch.from_numpy(obs).to(syn.device)).contiguous()
        out.append((syn.cpu().numpy(), p.adjoint_exact(gseis).cpu().numpy()))
    p.close()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [True, "checkpoint"])
@pytest.mark.parametrize("ns,nt,graph,fused", [
    (2, 160, 1, 1),
    (5, 161, 1, 1),
    (4, 150, 0, 1),
    (3, 151, 1, 0),                          # separate adjoint passes: the adjoint on one stream
])
def test_two_streams_exact_match_one_stream(mode, ns, nt, graph, fused):
    """The exact forward and adjoint (full divergence history, or checkpoints with
    segment recomputation) on two streams equal the one-stream run bit for bit."""
    from oracle import oracle as O
    w = workloads.small(seed=ns + nt, nx=52, nz=36, npml=10, nt=nt, ns=ns)
    src = integrated_sources(w)
    m = O.Model(w.vp_true, workloads.gardner(w.vp_true), w.dx, w.npml, "f32")
    obs = O.forward(m, w.nt, w.dt, w.src_xz, w.rec_xz, src, save_strips=False)["seis"].astype(np.float32)
    one = _run_exact(w, mode, 1, graph, fused, src, obs)
    two = _run_exact(w, mode, 2, graph, fused, src, obs)
    for rep in range(2):
        np.testing.assert_array_equal(two[rep][0], one[rep][0], err_msg=f"seismograms, pass {rep}")
        np.testing.assert_array_equal(two[rep][1], one[rep][1], err_msg=f"gradient, pass {rep}")
    assert np.abs(one[0][1]).max() > 0
