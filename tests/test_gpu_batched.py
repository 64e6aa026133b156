"""Batched path: every mode (graph replay, direct launches, group sizes, repeated
calls) against the FP32 oracle and against itself."""
import numpy as np
import pytest

from helpers import integrated_sources, rel_l2
from oracle import oracle as O
from swit_amd import workloads


def _case(seed=11, **kw):
    w = workloads.small(seed=seed, **kw)
    src = integrated_sources(w)
    m = O.Model(w.vp_true, workloads.gardner(w.vp_true), w.dx, w.npml, "f32")
    obs = O.forward(m, w.nt, w.dt, w.src_xz, w.rec_xz, src, save_strips=False)["seis"]
    m = O.Model(w.vp_init, workloads.gardner(w.vp_init), w.dx, w.npml, "f32")
    f = O.forward(m, w.nt, w.dt, w.src_xz, w.rec_xz, src)
    adj = (f["seis"] - obs).astype(np.float32)
    a = O.adjoint(m, w.nt, w.dt, w.src_xz, w.rec_xz, src, adj, f)
    return w, src, adj, f, a


def _plan(w, path="batched"):
    import torch
    from swit_amd.engine import Plan
    p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, path=path)
    p.set_model(torch.from_numpy(w.vp_init), torch.from_numpy(workloads.gardner(w.vp_init)))
    return p


def _run(p, src, adj):
    import torch
    seis = p.forward(torch.from_numpy(src), save_strips=True).cpu().numpy()
    g = [t.cpu().numpy() for t in p.adjoint(torch.from_numpy(src), torch.from_numpy(adj))]
    return seis, g


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [1, 0])
def test_batched_matches_oracle(graph):
    w, src, adj, f, a = _case()
    p = _plan(w)
    p.set_option("graph", graph)
    seis, (g, fw, bk) = _run(p, src, adj)
    strips = p.saved_strips().cpu().numpy()
    p.close()
    print(f"graph={graph}: seis max|d| {np.abs(seis - f['seis']).max():.3e}, grad relL2 {rel_l2(g, a['grad']):.3e}")
    np.testing.assert_array_equal(seis, f["seis"])
    np.testing.assert_array_equal(strips, f["boundary"])
    # per-shot FP32 imaging in the reference's order: bit-identical per-shot kernels,
    # FP64 cross-shot sums in the same shot order as the oracle
    assert rel_l2(g, a["grad"]) <= 1e-12
    assert rel_l2(fw, a["illum_forw"]) <= 1e-12
    assert rel_l2(bk, a["illum_back"]) <= 1e-12


@pytest.mark.gpu
def test_group_sizes_and_replays_agree():
    w, src, adj, f, a = _case(seed=12, ns=5)
    p = _plan(w)
    ref = None
    # shots per forward workgroup, with the two-step backward kernel or the one-step one
    # and its XCD work split by estimated cost (1) or by item count (0)
    for G, steps, bal in ((1, 2, 1), (3, 1, 1), (5, 1, 0), (2, 2, 0), (3, 1, 1)):
        p.set_option("fwd_group", G)
        p.set_option("bwd_steps", steps)
        p.set_option("bwd_balance", bal)
        for rep in range(2):
            seis, g = _run(p, src, adj)
            np.testing.assert_array_equal(seis, f["seis"])
            if ref is None:
                ref = g
            for x, y in zip(g, ref):
                assert rel_l2(x, y) == 0.0, (G, steps, bal, rep)
    assert rel_l2(ref[0], a["grad"]) <= 1e-12
    p.close()


@pytest.mark.gpu
def test_adjoint_needs_saved_strips():
    import torch
    from swit_amd._native import SwitError
    w, src, adj, f, a = _case(seed=13, ns=2, nt=120)
    p = _plan(w)
    p.forward(torch.from_numpy(src), save_strips=False)
    with pytest.raises(SwitError):
        p.adjoint(torch.from_numpy(src), torch.from_numpy(adj))
    p.close()


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [1, 0])
def test_profile_counts_every_step(graph):
    """Profiling counts nt-1 launches per time loop, accumulates over calls, restarts
    from zero when switched on again, and leaves the results unchanged."""
    w, src, adj, f, a = _case(seed=14, ns=2, nt=150)
    p = _plan(w)
    p.set_option("graph", graph)
    seis0, g0 = _run(p, src, adj)
    p.set_option("profile", 1)
    seis1, g1 = _run(p, src, adj)
    _run(p, src, adj)
    fm, fn = p.kernel_time("forward")
    bm, bn = p.kernel_time("backward")
    p.set_option("profile", 0)
    _run(p, src, adj)
    assert p.kernel_time("forward")[1] == fn
    p.set_option("profile", 1)
    assert p.kernel_time("backward") == (0.0, 0)
    p.close()
    assert fn == 2 * (w.nt - 1) and bn == 2 * (w.nt - 1)
    assert fm > 0 and bm > 0
    np.testing.assert_array_equal(seis1, seis0)
    for x, y in zip(g1, g0):
        np.testing.assert_array_equal(x, y)


@pytest.mark.gpu
@pytest.mark.parametrize("nx,nz,npml,nt,ns", [
    (40, 30, 59, 160, 2),     # free-surface rows straddle the default backward tile boundary
    (70, 150, 20, 220, 3),    # several tile rows in both step kernels
    (33, 9, 3, 150, 2),       # thinnest absorbing layer and interior the plan accepts
    (100, 40, 12, 200, 4),    # several tile columns, uneven last column
])
def test_batched_tilings(nx, nz, npml, nt, ns):
    w, src, adj, f, a = _case(seed=nx + nz, nx=nx, nz=nz, npml=npml, nt=nt, ns=ns)
    for graph in (1, 0):
        p = _plan(w)
        p.set_option("graph", graph)
        seis, (g, fw, bk) = _run(p, src, adj)
        p.close()
        np.testing.assert_array_equal(seis, f["seis"])
        assert rel_l2(g, a["grad"]) <= 1e-12, (graph, rel_l2(g, a["grad"]))
        assert rel_l2(fw, a["illum_forw"]) <= 1e-12
        assert rel_l2(bk, a["illum_back"]) <= 1e-12


@pytest.mark.gpu
def test_autotune_keeps_results():
    """The autotuner only changes the work order and shots per workgroup: results are
    bit-identical before and after; strips saved before tuning are invalidated."""
    import torch
    from swit_amd._native import SwitError
    w, src, adj, f, a = _case(seed=15, ns=5, nt=160)
    p = _plan(w)
    seis0, g0 = _run(p, src, adj)
    p.forward(torch.from_numpy(src), save_strips=True)
    ch = p.autotune(steps=5)
    print("autotune", ch)
    assert 1 <= ch["fwd_group"] <= w.ns and ch["fwd_colgroup"] >= 1 and ch["bwd_colgroup"] >= 1
    assert ch["bwd_balance"] in (0, 1)
    assert ch["fwd_us"] > 0 and ch["bwd_us"] > 0
    with pytest.raises(SwitError):
        p.adjoint(torch.from_numpy(src), torch.from_numpy(adj))
    seis1, g1 = _run(p, src, adj)
    g_after = g1
    for opt, v in (("fwd_colgroup", 1), ("bwd_colgroup", 3), ("fwd_group", 2), ("bwd_balance", 0)):
        p.set_option(opt, v)
    seis2, g2 = _run(p, src, adj)
    p.close()
    np.testing.assert_array_equal(seis1, seis0)
    np.testing.assert_array_equal(seis2, seis0)
    for x, y, z, q in zip(g_after, g1, g2, g0):
        np.testing.assert_array_equal(x, q)
        np.testing.assert_array_equal(y, q)
        np.testing.assert_array_equal(z, q)
