"""16-bit boundary strips (FP16 / BF16 with a per-step scale) on the batched path.

Encoding per shot and step: the values times the power of two 2^k that puts the
row's max |v| into [2^14, 2^15), rounded once to nearest even.  The forward is unchanged (seismograms bit-exact with the FP32 oracle).  The adjoint
must equal the FP32 oracle's adjoint run on the oracle's strips passed through the
same encoding (helpers.quantize_strips), shot by shot in the same arithmetic, so the
FP64 cross-shot sums agree to rounding of the sum (relative L2 <= 1e-12).  The error
the encoding itself adds is reported against the FP32-strip gradient and the FP64
oracle, and bounded: relative L2 <= 5e-3 (FP16) and <= 3e-2 (BF16)."""
import numpy as np
import pytest

from helpers import integrated_sources, quantize_strips, rel_l2
from oracle import oracle as O
from swit_amd import workloads

BOUND = {"f16": 5e-3, "bf16": 3e-2}


def test_quantize_strips_closed_forms():
    """Power-of-two scaling is exact; zero rows stay zero; the error relative to the
    row max is at most 2^-11 (FP16) / 2^-8 (BF16)."""
    row = np.array([[[0.0, 1.0, -0.5, 0.25 + 2.0 ** -14, 4.0]]], np.float32)
    zero = np.zeros((1, 1, 5), np.float32)
    for mode, bits in (("f16", 11), ("bf16", 8)):
        q = quantize_strips(row, mode)[0, 0]
        # max 4 -> 2^12: 4, 1, -0.5 exact; 0.25 + 2^-14 -> 1024.25 rounds to 1024
        np.testing.assert_array_equal(q, np.array([0.0, 1.0, -0.5, 0.25, 4.0], np.float32))
        np.testing.assert_array_equal(quantize_strips(zero, mode), zero)
        r = np.random.default_rng(1).standard_normal((2, 3, 1000)).astype(np.float32)
        err = np.abs(quantize_strips(r, mode) - r) / np.abs(r).max(axis=-1, keepdims=True)
        assert err.max() <= 2.0 ** -bits + 1e-9
        assert err.max() > 2.0 ** -(bits + 2)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["f16", "bf16"])
@pytest.mark.parametrize("nt", [150, 2, 3, 33, 65])
def test_16bit_strips_match_encoded_oracle(mode, nt):
    import torch
    from swit_amd.engine import Plan
    w = workloads.small(seed=31, ns=3, nt=nt)
    src = integrated_sources(w)
    m = O.Model(w.vp_true, workloads.gardner(w.vp_true), w.dx, w.npml, "f32")
    obs = O.forward(m, w.nt, w.dt, w.src_xz, w.rec_xz, src, save_strips=False)["seis"]
    m = O.Model(w.vp_init, workloads.gardner(w.vp_init), w.dx, w.npml, "f32")
    f = O.forward(m, w.nt, w.dt, w.src_xz, w.rec_xz, src)
    adj = (f["seis"] - obs).astype(np.float32)
    if nt <= 3:
        adj = (f["seis"] + 1e-3).astype(np.float32)         # a nonzero source for tiny nt
    a32 = O.adjoint(m, w.nt, w.dt, w.src_xz, w.rec_xz, src, adj, f)
    fq = dict(f)
    fq["boundary"] = quantize_strips(f["boundary"], mode)
    aq = O.adjoint(m, w.nt, w.dt, w.src_xz, w.rec_xz, src, adj, fq)
    m64 = O.Model(w.vp_init, workloads.gardner(w.vp_init), w.dx, w.npml, "f64")
    f64 = O.forward(m64, w.nt, w.dt, w.src_xz, w.rec_xz, src)
    a64 = O.adjoint(m64, w.nt, w.dt, w.src_xz, w.rec_xz, src, adj, f64)

    p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, path="batched", strips=mode)
    p.set_model(torch.from_numpy(w.vp_init), torch.from_numpy(workloads.gardner(w.vp_init)))
    seis = p.forward(torch.from_numpy(src), save_strips=True).cpu().numpy()
    strips = p.saved_strips().cpu().numpy()
    eq = quantize_strips(f["boundary"], mode)
    bad = np.argwhere(strips.view(np.uint32) != eq.view(np.uint32))
    if len(bad):
        print(f"[{mode} nt={nt}] {len(bad)} strip values differ from the encoded oracle, first rows "
              f"{sorted(set(map(tuple, bad[:, :2].tolist())))[:8]}")
        s0, r0, k0 = bad[0]
        print(f"  e.g. shot {s0} row {r0} slot {k0}: gpu {strips[s0, r0, k0]!r} expected {eq[s0, r0, k0]!r} "
              f"fp32 {f['boundary'][s0, r0, k0]!r} row max {np.abs(f['boundary'][s0, r0]).max()!r}")
    np.testing.assert_array_equal(strips, eq)
    g, fw, bk = [t.cpu().numpy() for t in p.adjoint(torch.from_numpy(src), torch.from_numpy(adj))]
    # a second adjoint from the same strips gives the same result
    g2 = p.adjoint(torch.from_numpy(src), torch.from_numpy(adj))[0].cpu().numpy()
    p.close()
    print(f"[{mode} nt={nt}] grad relL2 vs encoded oracle {rel_l2(g, aq['grad']):.3e}, "
          f"vs FP32 strips {rel_l2(g, a32['grad']):.3e}, vs FP64 oracle {rel_l2(g, a64['grad']):.3e}; "
          f"illum_forw vs FP32 {rel_l2(fw, a32['illum_forw']):.3e}")
    np.testing.assert_array_equal(seis, f["seis"])
    np.testing.assert_array_equal(g2, g)
    assert rel_l2(g, aq["grad"]) <= 1e-12
    assert rel_l2(fw, aq["illum_forw"]) <= 1e-12
    assert rel_l2(bk, aq["illum_back"]) <= 1e-12
    assert rel_l2(g, a64["grad"]) <= BOUND[mode]
