"""Two-step backward kernel (two reverse time steps per launch) against the one-step
kernel at both its tile widths and XCD work splits, and the oracle: identical bits for
odd and even step counts, every strip mode,
free-surface rows at many offsets inside the two-step windows, and receivers on every
row of columns at and around tile edges (adjoint sources injected in both steps'
regions, receivers sharing cells added in receiver order)."""
import numpy as np
import pytest

from helpers import integrated_sources, rel_l2
from oracle import oracle as O
from swit_amd import workloads


def _receivers(w, tiles=(16, 24)):
    """Surface row, plus full-depth lines on interior columns whose padded column lies
    at or next to a tile edge (of each tile width), plus duplicates of a few cells."""
    nxp = w.nx + 2 * w.npml
    cols = []
    for tile in tiles:
        found = []
        for edge in range(tile, nxp, tile):              # last padded column of a tile
            for ix in (edge - 2, edge, edge + 1, edge + 3):
                jx = ix - w.npml                         # interior column (1-based)
                if 1 <= jx <= w.nx and jx - 1 not in cols and jx - 1 not in found:
                    found.append(jx - 1)
        cols += found[:4]
    rec = [[x * w.dx, 0.0] for x in range(w.nx)]
    for c in cols:
        rec += [[c * w.dx, z * w.dx] for z in range(1, w.nz - 1)]
    rec += [[cols[0] * w.dx, 2 * w.dx], [cols[0] * w.dx + 1.0, 2 * w.dx + 2.0], [3 * w.dx, 0.0]]
    rec = np.asarray(rec, np.float32)
    return np.broadcast_to(rec, (w.ns,) + rec.shape).copy()


def _plan(w, strips, steps):
    import torch
    from swit_amd.engine import Plan
    p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, strips=strips)
    p.set_model(torch.from_numpy(w.vp_init), torch.from_numpy(workloads.gardner(w.vp_init)))
    p.set_option("bwd_steps", steps)
    return p


@pytest.mark.gpu
@pytest.mark.parametrize("nx,nz,npml,nt,ns,strips", [
    (60, 50, 10, 161, 2, "f32"),    # odd number of reverse steps: pairs only
    (60, 50, 10, 160, 2, "f32"),    # even: the pass ends with one one-step launch
    (40, 30, 59, 120, 2, "f16"),    # wide absorbing layer, free surface deep in the grid
    (90, 140, 20, 151, 3, "bf16"),  # several two-step tile rows and columns
    (33, 9, 3, 101, 2, "f32"),      # thinnest absorbing layer and interior
    (50, 60, 45, 90, 2, "f32"),
    (50, 60, 50, 91, 2, "f16"),
    (40, 30, 10, 3, 2, "f32"),      # one two-step launch is the whole reverse pass
    (40, 30, 10, 2, 2, "bf16"),     # one one-step launch is the whole reverse pass
])
def test_two_step_backward_matches_one_step(nx, nz, npml, nt, ns, strips):
    import torch
    w = workloads.small(seed=nx + nz + npml, nx=nx, nz=nz, npml=npml, nt=nt, ns=ns)
    w.rec_xz = _receivers(w)
    src = integrated_sources(w)
    m = O.Model(w.vp_true, workloads.gardner(w.vp_true), w.dx, w.npml, "f32")
    obs = O.forward(m, w.nt, w.dt, w.src_xz, w.rec_xz, src, save_strips=False)["seis"]
    out = {}
    # the one-step kernel at both its tile widths and every XCD work split, and the
    # two-step kernel
    cases = [(1, 16, 0), (1, 24, 0), (2, 24, 0), (1, 24, 1), (1, 16, 1)]
    for steps, tx, bal in cases:
        for graph in (1, 0):
            p = _plan(w, strips, steps)
            p.set_option("bwd_tile_x", tx)
            p.set_option("bwd_balance", bal)
            p.set_option("graph", graph)
            syn = p.forward(torch.from_numpy(src), save_strips=True)
            adj = (syn.cpu().numpy() - obs).astype(np.float32)
            out[steps, tx, bal, graph] = [t.cpu().numpy() for t in p.adjoint(torch.from_numpy(src),
                                                                             torch.from_numpy(adj))]
            p.close()
    ref = out[1, 24, 0, 1]
    for key, g in out.items():
        for x, y in zip(g, ref):
            np.testing.assert_array_equal(x, y, err_msg=str(key))
    if strips == "f32":
        mi = O.Model(w.vp_init, workloads.gardner(w.vp_init), w.dx, w.npml, "f32")
        f = O.forward(mi, w.nt, w.dt, w.src_xz, w.rec_xz, src)
        a = O.adjoint(mi, w.nt, w.dt, w.src_xz, w.rec_xz, src, (f["seis"] - obs).astype(np.float32), f)
        assert rel_l2(ref[0], a["grad"]) <= 1e-12
        assert rel_l2(ref[1], a["illum_forw"]) <= 1e-12
        assert rel_l2(ref[2], a["illum_back"]) <= 1e-12


@pytest.mark.gpu
def test_two_step_is_the_default_and_profiles_per_step():
    """Plans start with the two-step kernel; profiling counts time steps (nt-1 per
    reverse pass) whichever kernel runs."""
    import torch
    w = workloads.small(seed=5, nx=48, nz=36, npml=10, nt=140, ns=2)
    src = integrated_sources(w)
    res = {}
    for steps in (2, 1):
        p = _plan(w, "f32", steps)
        p.set_option("profile", 1)
        syn = p.forward(torch.from_numpy(src), save_strips=True)
        g = p.adjoint(torch.from_numpy(src), syn * 0.5)[0].cpu().numpy()
        bm, bn = p.kernel_time("backward")
        p.close()
        assert bn == w.nt - 1 and bm > 0
        res[steps] = g
    np.testing.assert_array_equal(res[1], res[2])
    from swit_amd import _native
    from swit_amd.engine import Plan
    p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz)
    assert p.bwd_steps == 2
    with pytest.raises(_native.SwitError):
        p.set_option("bwd_tile_x", 20)
    with pytest.raises(_native.SwitError):
        p.set_option("bwd_balance", 2)
    p.set_model(torch.from_numpy(w.vp_init), torch.from_numpy(workloads.gardner(w.vp_init)))
    tuned = p.autotune(steps=4)
    assert tuned["bwd_tile_x"] in (16, 24) and tuned["bwd_steps"] in (1, 2) and tuned["bwd_balance"] in (0, 1)
    p.close()


@pytest.mark.gpu
def test_interior_tiles_with_buried_sources_and_receivers():
    """Deep interior: on a 120 x 150 interior with a 10-cell absorbing layer, 4 of the
    18 backward tiles at width 24 and 6 of the 27 at width 16 have their whole window
    inside the interior (interior rows 50-107).  Sources buried in those tiles exercise
    the reconstruction's source subtraction away from every strip, and receivers
    scattered through them (two sharing a cell) the adjoint injection.  Both widths, the
    two-step kernel and the oracle agree."""
    import torch
    w = workloads.small(seed=77, nx=120, nz=150, npml=10, nt=180, ns=3)
    w.src_xz = np.array([[40, 60], [63, 80], [92, 101]], np.float32) * w.dx
    rec = [[x * w.dx, 0.0] for x in range(w.nx)]
    rec += [[x * w.dx, z * w.dx] for x in range(14, 112, 7) for z in (49, 55, 71, 106)]
    rec += [[50 * w.dx, 70 * w.dx], [50 * w.dx + 1.0, 70 * w.dx + 1.0]]
    w.rec_xz = np.broadcast_to(np.asarray(rec, np.float32), (w.ns, len(rec), 2)).copy()
    src = integrated_sources(w)
    m = O.Model(w.vp_true, workloads.gardner(w.vp_true), w.dx, w.npml, "f32")
    obs = O.forward(m, w.nt, w.dt, w.src_xz, w.rec_xz, src, save_strips=False)["seis"]
    out = {}
    for steps, tx, bal in ((1, 16, 0), (1, 24, 0), (2, 24, 0), (1, 24, 1), (1, 16, 1)):
        p = _plan(w, "f32", steps)
        p.set_option("bwd_tile_x", tx)
        p.set_option("bwd_balance", bal)
        syn = p.forward(torch.from_numpy(src), save_strips=True)
        adj = (syn.cpu().numpy() - obs).astype(np.float32)
        out[steps, tx, bal] = [t.cpu().numpy() for t in p.adjoint(torch.from_numpy(src), torch.from_numpy(adj))]
        p.close()
    ref = out[1, 24, 0]
    for key, g in out.items():
        for x, y in zip(g, ref):
            np.testing.assert_array_equal(x, y, err_msg=str(key))
    mi = O.Model(w.vp_init, workloads.gardner(w.vp_init), w.dx, w.npml, "f32")
    f = O.forward(mi, w.nt, w.dt, w.src_xz, w.rec_xz, src)
    a = O.adjoint(mi, w.nt, w.dt, w.src_xz, w.rec_xz, src, (f["seis"] - obs).astype(np.float32), f)
    assert rel_l2(ref[0], a["grad"]) <= 1e-12
    assert rel_l2(ref[1], a["illum_forw"]) <= 1e-12
    assert rel_l2(ref[2], a["illum_back"]) <= 1e-12
