"""Parity at the benchmark configurations' full sizes (grid, absorbing layer, record
length, shot count), through the C ABI, against the oracle:

  * 201x81, npml 40, nt 2000, 1 shot: plain and batched paths;
  * 481x141, npml 40, nt 4000, all 49 shots: the batched path as the benchmark runs it
    (autotuned), and the plain path;
  * 501x171x5000 / 481x121x4000 / 301x101x3000, 22 + 21 + 21 shots: one plan group, the
    whole gradient in one graph launch with the global-correlation misfit on the device;
  * 801x267, npml 50, nt 6000, shots 0 and 40: FP32 strips against the FP32 oracle.

Against the FP32 oracle (the reference's arithmetic in its order): seismograms and
strips bit-exact; per-shot FP32 kernels and illuminations in the reference's order, so
the FP64 shot sums agree to relative L2 1e-12.  Against the FP64 oracle (its own
observed data, residual and adjoint): the FP32 results' relative L2 errors are printed,
and the gradient's is held to 1e-3.  The 801x267 precision check against the FP64
oracle runs that oracle in a CPU-only subprocess (tests/precision_c5.py)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import integrated_sources, oracle_threads, rel_l2
from oracle import oracle as O
from swit_amd import workloads

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _oracle_gradient(w, prec, sel=None, adj_fn=None, strips=False):
    """Observed data (true model), synthetics with strips and the adjoint (initial model)
    of the oracle in precision `prec`; adj_fn(obs, syn) -> adjoint source (default
    syn - obs in the oracle's precision)."""
    sel = slice(None) if sel is None else sel
    th = oracle_threads()
    src = integrated_sources(w)[sel]
    sxz, rxz = w.src_xz[sel], w.rec_xz[sel]
    mt = O.Model(w.vp_true, w.rho(w.vp_true), w.dx, w.npml, prec)
    obs = O.forward(mt, w.nt, w.dt, sxz, rxz, src, save_strips=False, threads=th)["seis"]
    m = O.Model(w.vp_init, w.rho(w.vp_init), w.dx, w.npml, prec)
    f = O.forward(m, w.nt, w.dt, sxz, rxz, src, threads=th)
    adj = adj_fn(obs, f["seis"]) if adj_fn is not None else (f["seis"] - obs)
    a = O.adjoint(m, w.nt, w.dt, sxz, rxz, src, adj, f, threads=th)
    out = {"obs": obs, "syn": f["seis"], "adj": adj, "grad": a["grad"], "illum_forw": a["illum_forw"],
           "illum_back": a["illum_back"]}
    if strips:
        out["boundary"] = f["boundary"]
    return out


def _gpu_gradient(w, path="batched", sel=None, strips="f32", autotune=False, keep_strips=False):
    """The same sequence on the GPU: observed data, synthetics with strips, adjoint source
    syn - obs (FP32 on the device, as the oracle's FP32 run forms it), adjoint."""
    import torch
    from swit_amd.engine import Plan
    sel = slice(None) if sel is None else sel
    src = torch.from_numpy(integrated_sources(w)[sel]).cuda()
    p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz[sel], w.rec_xz[sel], path=path, strips=strips)
    p.set_model(torch.from_numpy(w.vp_true).cuda(), torch.from_numpy(w.rho(w.vp_true)).cuda())
    tuned = p.autotune(steps=10) if autotune else None
    obs = p.forward(src, save_strips=False).clone()
    p.set_model(torch.from_numpy(w.vp_init).cuda(), torch.from_numpy(w.rho(w.vp_init)).cuda())
    syn = p.forward(src, save_strips=True)
    out = {"obs": obs.cpu().numpy(), "syn": syn.cpu().numpy()}
    if keep_strips:
        out["boundary"] = p.saved_strips().cpu().numpy()
    g, fw, bk = p.adjoint(src, syn - obs)
    out.update(grad=g.cpu().numpy(), illum_forw=fw.cpu().numpy(), illum_back=bk.cpu().numpy(), tuned=tuned)
    p.close()
    return out


def _compare(tag, gpu, o32, o64):
    np.testing.assert_array_equal(gpu["obs"], o32["obs"], err_msg=tag + " observed data")
    np.testing.assert_array_equal(gpu["syn"], o32["syn"], err_msg=tag + " synthetics")
    if "boundary" in gpu:
        np.testing.assert_array_equal(gpu["boundary"], o32["boundary"], err_msg=tag + " strips")
    for k in ("grad", "illum_forw", "illum_back"):
        e = rel_l2(gpu[k], o32[k])
        assert e <= 1e-12, (tag, k, e)
    e64 = {k: rel_l2(gpu[k], o64[k]) for k in ("syn", "grad", "illum_forw", "illum_back")}
    print(f"{tag}: vs FP64 oracle " + ", ".join(f"{k} {v:.2e}" for k, v in e64.items()), flush=True)
    assert e64["grad"] <= 1e-3


@pytest.mark.gpu
def test_config_201x81x2000_one_shot_plain_and_batched():
    w = workloads.config1()
    o32 = _oracle_gradient(w, "f32", strips=True)
    o64 = _oracle_gradient(w, "f64")
    for path in ("plain", "batched"):
        gpu = _gpu_gradient(w, path, keep_strips=True)
        _compare(f"201x81x2000 {path}", gpu, o32, o64)


@pytest.mark.gpu
def test_config_481x141x4000_49_shots_batched_autotuned():
    """The benchmark's workload, all 49 shots, with the launch configuration the
    autotuner picks (as bench.py runs it)."""
    w = workloads.config2()
    o32 = _oracle_gradient(w, "f32", strips=True)
    gpu = _gpu_gradient(w, "batched", autotune=True, keep_strips=True)
    print("autotuned:", gpu["tuned"], flush=True)
    o64 = _oracle_gradient(w, "f64")
    _compare("481x141x4000 batched", gpu, o32, o64)


@pytest.mark.gpu
def test_config_481x141x4000_49_shots_plain():
    w = workloads.config2()
    o32 = _oracle_gradient(w, "f32")
    gpu = _gpu_gradient(w, "plain")
    o64 = _oracle_gradient(w, "f64")
    _compare("481x141x4000 plain", gpu, o32, o64)


@pytest.mark.gpu
def test_config_uneven_64_shots_one_plan_group():
    """All three grid families in one plan group; the whole gradient in one graph launch
    (swit_group_gradient) with the global-correlation misfit and adjoint source on the
    device.  Each family's seismograms against the FP32 oracle bit for bit; its adjoint
    source against the host restatement of misfit.py's (float32 output rounding); its
    gradient against the FP32 oracle's adjoint fed that adjoint source (1e-12), and
    against the FP64 oracle with its own host adjoint source (1e-3)."""
    import torch
    from swit_amd import misfit as host
    from swit_amd.engine import Plan, PlanGroup
    ws = workloads.config4()
    plans, srcs, obs = [], [], []
    for w in ws:
        src = torch.from_numpy(integrated_sources(w)).cuda()
        p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz)
        p.set_model(torch.from_numpy(w.vp_true).cuda(), torch.from_numpy(w.rho(w.vp_true)).cuda())
        p.autotune(steps=10)
        plans.append(p)
        srcs.append(src)
    grp = PlanGroup(plans)
    obs = [o.clone() for o in grp.forward(srcs, save_strips=False)]
    for p, w in zip(plans, ws):
        p.set_model(torch.from_numpy(w.vp_init).cuda(), torch.from_numpy(w.rho(w.vp_init)).cuda())
    Js, syn, outs = grp.gradient(srcs, obs, "Globalcorrelation")
    torch.cuda.synchronize()
    adj_dev = [a.cpu().numpy() for a in grp._bufs[1]]
    res = [(float(J), s.cpu().numpy(), [t.cpu().numpy() for t in o]) for J, s, o in zip(Js, syn, outs)]
    obs_np = [o.cpu().numpy() for o in obs]
    grp.close()
    for p in plans:
        p.close()
    for i, w in enumerate(ws):
        tag = f"{w.nx}x{w.nz}x{w.nt} ({w.ns} shots)"
        J, syn_np, (g, fw, bk) = res[i]
        o32 = _oracle_gradient(w, "f32", adj_fn=lambda o, s, a=adj_dev[i]: a)
        np.testing.assert_array_equal(obs_np[i], o32["obs"], err_msg=tag)
        np.testing.assert_array_equal(syn_np, o32["syn"], err_msg=tag)
        dt = float(np.float32(w.dt))
        adj_host = np.stack([host.adjoint_source_global_correlation(o32["obs"][k], o32["syn"][k], dt)
                             for k in range(w.ns)])
        J_host = sum(host.misfit_global_correlation(o32["obs"][k], o32["syn"][k], dt) for k in range(w.ns))
        scale = np.abs(adj_host).max()
        assert np.abs(adj_dev[i] - adj_host).max() <= 1e-6 * scale, tag
        assert abs(J - J_host) <= 1e-6 * abs(J_host), (tag, J, J_host)
        for k, v in (("grad", g), ("illum_forw", fw), ("illum_back", bk)):
            e = rel_l2(v, o32[k])
            assert e <= 1e-12, (tag, k, e)
        o64 = _oracle_gradient(w, "f64", adj_fn=lambda o, s: np.stack(
            [host.adjoint_source_global_correlation(o[k], s[k], dt) for k in range(w.ns)]))
        e64 = rel_l2(g, o64["grad"])
        print(f"{tag}: misfit {J:.9e} (host {J_host:.9e}); gradient vs FP64 oracle {e64:.2e}", flush=True)
        assert e64 <= 1e-3


@pytest.mark.gpu
def test_config_801x267x6000_two_shots_fp32():
    w = workloads.config5()
    sel = [0, w.ns // 2]
    o32 = _oracle_gradient(w, "f32", sel=sel, strips=True)
    gpu = _gpu_gradient(w, "batched", sel=sel, keep_strips=True)
    np.testing.assert_array_equal(gpu["obs"], o32["obs"])
    np.testing.assert_array_equal(gpu["syn"], o32["syn"])
    np.testing.assert_array_equal(gpu["boundary"], o32["boundary"])
    for k in ("grad", "illum_forw", "illum_back"):
        assert rel_l2(gpu[k], o32[k]) <= 1e-12, k


@pytest.mark.gpu
def test_config_801x267x6000_precision_against_fp64_oracle(tmp_path):
    """Two shots with FP32, FP16 and BF16 strips (bench.py --workload precision) against
    the FP64 oracle run by tests/precision_c5.py in a process with no GPU runtime:
    gradient relative L2 and waveform misfit relative error <= 1e-3."""
    import json
    out = str(tmp_path / "p.npz")
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--workload", "precision", "--out", out],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([sys.executable, os.path.join(REPO, "tests", "precision_c5.py"), out,
                        str(oracle_threads())], capture_output=True, text=True, timeout=900,
                       env=dict(os.environ, HIP_VISIBLE_DEVICES="", CUDA_VISIBLE_DEVICES=""))
    assert r.returncode == 0, r.stdout + r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print(res, flush=True)
    for st in ("f32", "f16", "bf16"):
        assert res["grad_rel_l2_" + st] <= 1e-3, st
        assert res["misfit_rel_" + st] <= 1e-3, st
    assert res["grad_rel_l2_f32"] <= 1e-5
