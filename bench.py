"""Benchmark: shot-gradients/s of the 2-D acoustic FWI gradient on one MI355X.

Workload (default): 481 x 141 grid, dx 25 m, 40-cell PML, nt 4000, dt 1.9 ms, 49 shots,
481 surface receivers, seeded dipping-layer + fault model (swit_amd.workloads.config2).
One step = one full L2 gradient over every shot: model upload, forward with strip
saving and receiver sampling, adjoint source (syn - obs), reverse-time reconstruction,
adjoint propagation, imaging and the cross-shot sum.  Inputs (models, sources,
observed data) are resident in HBM before the timed region.

Prints one JSON line with the metric, a roofline object for the dominant kernel (the
backward step kernel, which does two reverse time steps per launch unless the autotuner
picks the one-step kernel; its average duration comes from HIP events on the launch
stream that bracket each time loop of the timed steps) and a CPU baseline (the FP32
oracle on the host cores, bounded sample).  Side legs, outside the timed steps: the
plain per-shot path on a sample of shots, the same gradient on two concurrent lanes
(engine.Lanes), the same gradient through torch.autograd with the exact adjoint, and
the four device misfits and the device shot-gather processing on the step's gathers.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(nx, nz, npml, ng):
    """Per shot and time step: (forward, backward) algorithmic bytes.
    forward  = 36 N_pad (p,u,w read+write, coefficients) + 4 S (strip write) + 4 ng
    backward = 36 N_pad (adjoint) + 32 N_int (reconstruction) + 24 N_int (imaging)
               + 4 S (strip read) + 4 ng (adjoint source read)."""
    n_pad = (nx + 2 * npml) * (nz + 2 * npml)
    n_int = nx * nz
    S = 9 * nx + 6 * nz
    fwd = 36 * n_pad + 4 * S + 4 * ng
    bwd = 36 * n_pad + 32 * n_int + 24 * n_int + 4 * S + 4 * ng
    return fwd, bwd


def load_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary
    (profiles/traffic_<kernel>.json: FETCH_SIZE and WRITE_SIZE in KiB per dispatch,
    scaled by the copy-kernel calibration factor), or None."""
    p = os.path.join(REPO, "profiles", f"traffic_{kernel}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        t = json.load(f)
    return t.get("bytes_per_launch")


def autograd_leg(w, src, obs, vp_init, rho_init, args):
    """The same gradient through torch.autograd with the exact adjoint: one step is
    seismograms(vp) (forward with the divergence history), 1/2 |d - d_obs|^2 and
    backward (exact adjoint of all shots), timed on one stream and then with the
    shots' halves on two streams (the value; the gradients must be identical).  The
    whole divergence history takes
    ns * nt * (nx+2npml) * ld * 4 bytes; when the device cannot hold it the plan keeps
    checkpoints and recomputes segments instead."""
    import torch
    from swit_amd.autograd import l2_misfit, seismograms
    from swit_amd.engine import Plan
    need = w.ns * w.nt * (w.nx + 2 * w.npml) * ((w.nz + 2 * w.npml + 31) // 32 * 32) * 4
    free, _ = torch.cuda.mem_get_info()
    mode = True if need <= 0.8 * free else "checkpoint"
    pe = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, path="batched", exact=mode)

    def step():
        vp = vp_init.clone().requires_grad_(True)
        loss = l2_misfit(seismograms(vp, pe, src, rho_init), obs)
        loss.backward()
        return vp.grad

    def timed(streams):
        pe.set_option("streams", streams)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g = step()
        torch.cuda.synchronize()
        return g, time.perf_counter() - t0

    # one stream, then the shots' halves on two streams (bit-identical gradients)
    g1, el1 = timed(1)
    g, el = timed(2)
    same = bool(torch.equal(g, g1))
    fd = fd_self_check(w, pe, src, obs, vp_init, rho_init, g)
    pe.close()
    # algorithmic bytes of one autograd gradient: the exact forward (36 N_pad + the
    # divergence history 4 N_pad + 4 ng per shot and step) and the fused exact adjoint
    # (adjoint state 24 N_pad, history read 4 N_pad, K-gradient read+write 8 N_pad,
    # coefficients 12 N_pad, adjoint-source read 4 ng); a checkpointed sweep recomputes
    # the forward once more
    n_pad = (w.nx + 2 * w.npml) * (w.nz + 2 * w.npml)
    fwd_b = 40 * n_pad + 4 * w.ng
    per_grad = w.ns * (w.nt - 1) * ((2 if mode == "checkpoint" else 1) * fwd_b + 48 * n_pad + 4 * w.ng)
    gbs = per_grad * args.steps / el / 1e9
    return {"value": round(w.ns * args.steps / el, 3), "unit": "shot-gradients/s",
            "ms_per_step": round(1000 * el / args.steps, 2), "streams": 2,
            "one_stream_value": round(w.ns * args.steps / el1, 3), "identical_to_one_stream": same,
            "history": "full" if mode is True else "checkpoints", "history_gib": round(need / 2**30, 1),
            "algorithmic_bytes_per_gradient": int(per_grad), "achieved_gbs": round(gbs, 1),
            "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
            "finite": bool(torch.isfinite(g).all()), "fd_self_check": fd}


def fd_self_check(w, plan, src, obs, vp, rho, g):
    """<g, dm> of the exact adjoint against central differences of the same GPU forward
    (FP32 seismograms, FP64 misfit) along a smooth perturbation dm (N(0,1) noise
    Gaussian-filtered with sigma 6 cells, max 50 m/s), at eps = 1 and 0.5.  A check that
    the gradient is the derivative of the forward the GPU runs, at the benchmark's size.
    Rounding in the FP32 forward limits it: at 481x141x4000 with 49 shots the relative
    error is about 4e-3 at eps 1 and 1e-2 at eps 0.5, growing as eps shrinks, which is
    cancellation in J(+) - J(-) rather than a gradient error.  The error against FP64
    finite differences of the FP64 oracle is a test (two of these shots)."""
    import torch
    from scipy.ndimage import gaussian_filter
    from swit_amd.autograd import l2_misfit, seismograms
    rng = np.random.default_rng(22)
    dm = gaussian_filter(rng.standard_normal((w.nx, w.nz)), 6)
    dm = torch.from_numpy(dm / np.abs(dm).max() * 50.0).to(vp.device)
    gd = float((g.double() * dm).sum())
    out = {"dm_max_m_per_s": 50.0, "g_dot_dm": gd}
    with torch.no_grad():
        for eps in (1.0, 0.5):
            jp = float(l2_misfit(seismograms((vp + eps * dm).float(), plan, src, rho), obs))
            jm = float(l2_misfit(seismograms((vp - eps * dm).float(), plan, src, rho), obs))
            fdv = (jp - jm) / (2 * eps)
            out[f"rel_err_eps_{eps:g}"] = abs(gd - fdv) / abs(fdv)
    return out


def lanes_leg(w, src, obs, vp, rho, g_ref, args, lanes=2):
    """The same gradient on engine.Lanes: the shots split into `lanes` plans, each
    running forward, adjoint source and adjoint on a stream of its own, so that the
    lanes' step kernels run side by side.  Timed as the main steps (warm-up, then
    `steps` gradients between synchronisations); the gradient is checked against the
    main step's (FP64 sums in another order)."""
    import torch
    from swit_amd.engine import Lanes
    L = Lanes(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, lanes=lanes, strips=args.strips)
    L.set_model(torch.from_numpy(w.vp_true).to(vp.device), torch.from_numpy(w.rho(w.vp_true)).to(vp.device))
    tuned = L.autotune(steps=50) if not args.no_autotune else None
    L.set_model(vp, rho)
    for _ in range(args.warmup):
        L.gradient(src, obs, "Waveform", w.dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        J, g, fw, bk = L.gradient(src, obs, "Waveform", w.dt)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    rel = float((g - g_ref).norm() / g_ref.norm())
    L.close()
    return {"lanes": lanes, "value": round(w.ns * args.steps / el, 3), "unit": "shot-gradients/s",
            "ms_per_step": round(1000 * el / args.steps, 2), "grad_rel_vs_main": rel,
            "autotuned": [t.get("bwd_steps") for t in tuned] if tuned else None}


def two_streams_leg(w, plan, step, g_ref, args):
    """The main plan with SWIT_OPT_STREAMS 2 (each half of the shots runs its time loops
    on a stream of its own): timed as the main steps; the gradient and illuminations
    must equal the one-stream step's bit for bit."""
    import torch
    plan.set_option("streams", 2)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g2 = step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    plan.set_option("streams", 1)
    same = all(bool(torch.equal(a, b)) for a, b in zip(g2, g_ref))
    return {"value": round(w.ns * args.steps / el, 3), "unit": "shot-gradients/s",
            "ms_per_step": round(1000 * el / args.steps, 2), "identical_to_one_stream": same}


def plain_leg(w, src, obs, vp_true, rho_true, vp_init, rho_init, batched_value, args):
    """The plain per-shot GPU path (one thread per cell, one launch per sweep, shot by
    shot), two ways, outside the timed steps:
      * `dropin`: one L2 gradient of all shots through the drop-in solver as its callers
        run it (swit_amd.solver with PATH='plain': forward with snapshots writing the SU
        files, processing, misfit, adjoint with the adjoint sources read from the files,
        the .bin outputs and the preconditioner), observed SU files written first;
      * `resident`: the plain plan's forward and adjoint with inputs already on the
        device, as the batched step is timed.
    batched_over_plain compares the batched step with the drop-in gradient."""
    import shutil
    import tempfile
    import torch
    from swit_amd import base, misfit as host_misfit, preprocess, solver
    from swit_amd.engine import Plan
    k = w.ns if args.plain_shots < 0 else min(args.plain_shots, w.ns)
    pp = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz[:k], w.rec_xz[:k], path="plain")
    pp.set_model(vp_init, rho_init)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    syn_k = pp.forward(src[:k], save_strips=True)
    pp.adjoint(src[:k], syn_k - obs[:k])
    torch.cuda.synchronize()
    resident = time.perf_counter() - t1
    pp.close()
    del syn_k
    # the drop-in, files on a RAM-backed folder when there is one
    root = "/dev/shm" if os.path.isdir("/dev/shm") else None
    home = tempfile.mkdtemp(prefix="swit_plain_", dir=root) + "/"
    wav = np.broadcast_to(workloads_ricker(w), (k, w.nt)).copy()

    def simulate(vp, rho):
        model = base.model(w.nx, w.nz, w.dx, w.dt, w.nt, True, w.npml, vp.astype(np.float64),
                           rho.astype(np.float64))
        return base.simulate(model, base.source(w.f0, k, w.src_xz[:k].astype(np.float64), wav),
                             base.receiver(w.ng, w.rec_xz[:k].astype(np.float64)), base.system(home, 1))
    optim = base.optimize('Waveform', 'LBFGS', 1, 0.05, 6000.0, 1000.0, 'Marine', 10, 0, 'None', 1.0, 10.0,
                          False, 0.0, 'None', False, False, 0.0, 1e9)
    path0 = solver.PATH
    solver.PATH = "plain"
    try:
        st = simulate(w.vp_true, w.rho(w.vp_true))
        solver.forward(st, 'obs', savesnap=0)
        preprocess.process_workflow(st, optim, 'obs')
        simu = simulate(w.vp_init, w.rho(w.vp_init))
        t1 = time.perf_counter()
        solver.forward(simu, 'syn', savesnap=1)
        preprocess.process_workflow(simu, optim, 'syn')
        J = host_misfit.misfit(simu, 'Waveform')
        g = solver.adjoint(simu, optim)
        dropin = time.perf_counter() - t1
    finally:
        solver.PATH = path0
        solver.release()
        shutil.rmtree(home, ignore_errors=True)
    return {"value": round(k / dropin, 3), "unit": "shot-gradients/s", "shots": k,
            "path": "drop-in solver (toolbox/solver.py surface), PATH='plain', SU files",
            "seconds": round(dropin, 2), "misfit": float(J), "finite": bool(np.isfinite(g).all()),
            "resident": {"value": round(k / resident, 3), "seconds": round(resident, 2),
                         "note": "plain plan forward + adjoint, inputs on the device"},
            "batched_over_plain": round(batched_value / (k / dropin), 2),
            "batched_over_plain_resident": round(batched_value / (k / resident), 2)}


def workloads_ricker(w):
    from swit_amd import workloads
    return workloads.ricker(w.nt, w.dt, w.f0)


def misfit_leg(obs, syn, dt, reps=5):
    """Device misfits and adjoint sources (swit_misfit) of every kind on the step's
    gathers [ns][ng][nt]: average time per call from HIP events on torch's current
    stream (the stream the misfit kernels are launched on), the HBM rate of the
    algorithmic bytes (read obs and syn, write the adjoint source: 12 bytes a sample),
    and the host restatement (numpy/scipy, as the reference computes them) timed on
    one shot and scaled to all of them."""
    import torch
    from swit_amd import misfit as host
    from swit_amd.engine import misfit
    ns, nt = obs.shape[0], obs.shape[-1]
    nbytes = 12 * obs.numel()
    o1, s1 = obs[0].cpu().numpy(), syn[0].cpu().numpy()
    out = {}
    for kind in ("Waveform", "Globalcorrelation", "Envelope", "Traveltime"):
        misfit(kind, obs, syn, dt)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            J, _, _ = misfit(kind, obs, syn, dt)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        t0 = time.perf_counter()
        host._MISFIT[kind.lower()](o1, s1, dt)
        host._ADJ[kind.lower()](o1, s1, dt)
        host_ms = 1000 * (time.perf_counter() - t0) * ns
        out[kind.lower()] = {"ms": round(ms, 3), "achieved_gbs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                             "host_ms_all_shots": round(host_ms, 1), "misfit": float(J)}
    out["traces"] = int(obs.numel() // nt)
    out["nt"] = int(nt)
    return out


def process_leg(syn, dt, offsets, reps=3):
    """Device shot-gather processing (swit_process) of the step's gathers with the
    reference's usual settings (first-break late mute, 2-12 Hz Butterworth bandpass
    with detrend and taper, L2 trace normalisation): average time per call from HIP
    events on torch's current stream, against the host restatement (numpy/scipy, the
    reference's float32 flow) timed on one shot and scaled to all of them."""
    import torch
    from swit_amd import base, preprocess
    from swit_amd.engine import process_gathers
    o = base.optimize('Waveform', 'LBFGS', 1, 0.05, 6000.0, 1000.0, 'Marine', 0, 0, 'Bandpass', 2.0, 12.0, 1, 0.5,
                      ['L2-Trace'], 0, 0, 0.0, 0.0)
    x = syn.clone()
    process_gathers(o, x, dt, offsets)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms = 0.0
    for _ in range(reps):
        x.copy_(syn)
        e0.record()
        process_gathers(o, x, dt, offsets)
        e1.record()
        torch.cuda.synchronize()
        ms += e0.elapsed_time(e1) / reps
    s1 = syn[0].cpu().numpy()
    t0 = time.perf_counter()
    h = preprocess.apply_mute(o, s1, dt, np.asarray(offsets)[0])
    h = preprocess.apply_filter(o, h, dt)
    preprocess.apply_normalize(o, h)
    host_ms = 1000 * (time.perf_counter() - t0) * syn.shape[0]
    return {"ms": round(ms, 3), "host_ms_all_shots": round(host_ms, 1),
            "steps": "late first-break mute, bandpass 2-12 Hz filtfilt + detrend + taper, L2-Trace",
            "traces": int(syn.numel() // syn.shape[-1]), "nt": int(syn.shape[-1])}


def host_cores():
    """CPUs this process may run on (its affinity mask)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def cpu_baseline(w, src, obs_syn_adj, nshots):
    """FP32 oracle (C restatement of the reference scheme) on the host: nshots
    shot-gradients of the same workload, one thread per shot on every CPU the process
    may run on."""
    from oracle import oracle as O
    threads = max(1, min(nshots, host_cores()))
    rho = w.rho(w.vp_init)
    t0 = time.perf_counter()
    m = O.Model(w.vp_init, rho, w.dx, w.npml, "f32")
    sel = slice(0, nshots)
    f = O.forward(m, w.nt, w.dt, w.src_xz[sel], w.rec_xz[sel], src[sel], threads=threads)
    adj = obs_syn_adj(f["seis"])
    O.adjoint(m, w.nt, w.dt, w.src_xz[sel], w.rec_xz[sel], src[sel], adj, f, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": nshots / dt, "unit": "shot-gradients/s", "cores": threads, "kind": "port",
            "sample": f"{nshots} of the workload's {w.ns} shots at full nt (observed data given), FP32 oracle "
                      f"(forward with strips, reconstruction, adjoint, imaging), one thread per shot on the "
                      f"{threads} CPUs of the process's affinity, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nt", type=int, default=4000)
    ap.add_argument("--path", default="batched", choices=["batched", "plain"])
    ap.add_argument("--strips", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch configuration")
    ap.add_argument("--no-autograd", action="store_true",
                    help="skip the exact-adjoint (torch.autograd) gradient leg")
    ap.add_argument("--cpu-shots", type=int, default=0, help="CPU baseline sample size (0: every shot)")
    ap.add_argument("--no-lanes", action="store_true",
                    help="skip the two-lane and two-stream legs (their half-size launches of the same kernels "
                         "would mix into a profiler's per-kernel averages)")
    ap.add_argument("--plain-shots", type=int, default=-1,
                    help="plain-path comparison: shots (-1: all of them; 0: skip)")
    ap.add_argument("--iters", type=int, default=10, help="L-BFGS iterations of --workload lbfgs")
    ap.add_argument("--out", default="precision_c5.npz",
                    help="with --workload precision: where the two-shot gradients go")
    ap.add_argument("--workload", default="c2", choices=["c2", "uneven", "lbfgs", "precision"],
                    help="c2: the 481x141x4000, 49-shot gradient (default); uneven: three grid "
                         "families (501x171x5000, 481x121x4000, 301x101x3000; 64 shots) as one plan "
                         "group with the global-correlation misfit on the device; lbfgs: --iters "
                         "L-BFGS iterations of the device-resident inversion loop "
                         "on the 801x267x6000, 81-shot model with 16-bit strips (--strips, default f16); "
                         "precision: two shots of that model with FP32 and 16-bit strips, gradients to --out "
                         "for tests/precision_c5.py")
    ap.add_argument("--bwd-steps", type=int, default=0, choices=[0, 1, 2],
                    help="reverse steps per backward launch (1 or 2) after autotuning; 0: the autotuner's "
                         "choice")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (5 times): the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}))
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(shots would shard across GPUs with one all-reduce per gradient; not implemented)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "uneven":
        return bench_uneven(args)
    if args.workload == "lbfgs":
        return bench_lbfgs(args)
    if args.workload == "precision":
        return bench_precision(args)

    import torch
    from swit_amd import workloads
    from swit_amd.engine import Plan
    from scipy.integrate import cumulative_trapezoid

    w = workloads.config2(nt=args.nt)
    wav = workloads.ricker(w.nt, w.dt, w.f0)
    src_np = np.broadcast_to(cumulative_trapezoid(wav, initial=0).astype(np.float32), (w.ns, w.nt)).copy()
    dev = torch.device("cuda", 0)
    src = torch.from_numpy(src_np).to(dev)
    vp_true = torch.from_numpy(w.vp_true).to(dev)
    vp_init = torch.from_numpy(w.vp_init).to(dev)
    rho_true = torch.from_numpy(w.rho(w.vp_true)).to(dev)
    rho_init = torch.from_numpy(w.rho(w.vp_init)).to(dev)

    plan = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, path=args.path, strips=args.strips)
    plan.set_model(vp_true, rho_true)
    tuned = plan.autotune(steps=50) if args.path == "batched" and not args.no_autotune else None
    if args.bwd_steps and args.path == "batched":
        plan.set_option("bwd_steps", args.bwd_steps)
    obs = plan.forward(src, save_strips=False).clone()
    torch.cuda.synchronize()

    def step():
        plan.set_model(vp_init, rho_init)
        syn = plan.forward(src, save_strips=True)
        adj = syn - obs
        return plan.adjoint(src, adj)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # kernel timing inside the timed region: one pair of HIP events, recorded on the
    # plan's launch stream, brackets each time loop (nt-1 step kernels back to back)
    if args.path == "batched":
        plan.set_option("profile", 1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    value = w.ns * args.steps / elapsed

    roofline = None
    if args.path == "batched":
        f_ms, f_n = plan.kernel_time("forward")
        b_ms, b_n = plan.kernel_time("backward")
        plan.set_option("profile", 0)
        fwd_b, bwd_b = algorithmic_bytes(w.nx, w.nz, w.npml, w.ng)
        # time steps per launch: the backward kernel may do two reverse steps per launch;
        # the profiled loops count time steps, so a launch lasts spl times the per-step
        # time and moves spl times the per-step algorithmic bytes
        spl = plan.bwd_steps
        kern = {"forward": dict(name="k_fwd_fused", steps=1, total_ms=f_ms, n=f_n, step_bytes=fwd_b * w.ns),
                "backward": dict(name="k_bwd_fused2" if spl == 2 else "k_bwd_fused", steps=spl, total_ms=b_ms,
                                 n=b_n, step_bytes=bwd_b * w.ns)}
        for k in kern.values():
            k["avg_launch_ms"] = k["total_ms"] / max(k["n"], 1) * k["steps"]
            k["bytes_per_launch"] = k["step_bytes"] * k["steps"]
            k["gbs"] = k["bytes_per_launch"] / (k["avg_launch_ms"] * 1e-3) / 1e9
        dom = max(kern, key=lambda k: kern[k]["total_ms"])
        oth = kern["forward" if dom == "backward" else "backward"]
        kd = kern[dom]
        roofline = {"bound": "hbm", "achieved": round(kd["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(kd["gbs"] / HBM_PEAK_GBS, 4), "traffic": load_traffic(kd["name"]),
                    "kernel": kd["name"], "avg_launch_us": round(kd["avg_launch_ms"] * 1e3, 2),
                    "time_steps_per_launch": kd["steps"],
                    "algorithmic_bytes_per_launch": int(kd["bytes_per_launch"]),
                    "other_kernel": {"name": oth["name"], "avg_launch_us": round(oth["avg_launch_ms"] * 1e3, 2),
                                     "time_steps_per_launch": oth["steps"], "achieved_gbs": round(oth["gbs"], 1)},
                    "time_loops_gbs": round((kern["forward"]["step_bytes"] * f_n + kern["backward"]["step_bytes"]
                                             * b_n) / ((f_ms + b_ms) * 1e-3) / 1e9, 1)}

    lanes = (lanes_leg(w, src, obs, vp_init, rho_init, g[0], args)
             if args.path == "batched" and not args.no_lanes else None)
    two_streams = (two_streams_leg(w, plan, step, g, args)
                   if args.path == "batched" and not args.no_lanes else None)
    misfits = processing = None
    if args.path == "batched":
        syn_now = plan.forward(src, save_strips=False)
        misfits = misfit_leg(obs, syn_now, w.dt)
        from swit_amd.su import nint
        offsets = (nint(w.rec_xz[:, :, 0]) - nint(w.src_xz[:, None, 0])).astype(np.float64)
        processing = process_leg(syn_now, w.dt, offsets)

    autograd = None
    if args.path == "batched" and not args.no_autograd:
        plan.close()
        autograd = autograd_leg(w, src, obs, vp_init, rho_init, args)

    plain = None
    if args.plain_shots and args.path == "batched":
        plain = plain_leg(w, src, obs, vp_true, rho_true, vp_init, rho_init, value, args)

    cpu = None
    if not args.no_cpu:
        obs_np = obs.cpu().numpy()
        n = args.cpu_shots or w.ns
        cpu = cpu_baseline(w, src_np, lambda syn: (syn - obs_np[: syn.shape[0]]).astype(np.float32), n)

    line = {
        "metric": "shot-gradients/s at 481\u00d7141\u00d74000 steps, batched vs plain; GB/s vs HBM peak",
        "value": round(value, 3),
        "unit": "shot-gradients/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 12-layer dipping model with two faults, Ricker 5 Hz)",
        "config": {"workload": f"{w.nx}x{w.nz} grid, npml {w.npml}, nt {w.nt}, dt {w.dt}, {w.ns} shots, "
                               f"{w.ng} receivers, L2 gradient", "path": args.path, "strips": args.strips,
                   "shots_per_step": w.ns, "autotuned": tuned, "bwd_steps": plan.bwd_steps},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain_baseline": plain,
        "autograd": autograd,
        "concurrent_lanes": lanes,
        "two_streams": two_streams,
        "device_misfits": misfits,
        "device_processing": processing,
    }
    if plain and autograd and "value" in autograd:
        autograd["over_plain"] = round(autograd["value"] / plain["value"], 2)
    print(json.dumps(line), flush=True)
    plan.close()


def bench_lbfgs(args):
    """L-BFGS iterations of the device-resident inversion loop (swit_amd.inversion_device,
    the reference loop's decisions with every vector on the device) on the 801x267x6000,
    81-shot model, 16-bit strips (--strips, default f16), FP64 shot sum.  The observed
    data are modelled on the device in the true model first (timed apart)."""
    import shutil
    import tempfile
    from swit_amd import base, inversion_device, solver, workloads
    strips = args.strips if args.strips != "f32" else "f16"
    w = workloads.config5()
    root = "/dev/shm" if os.path.isdir("/dev/shm") else None
    home = tempfile.mkdtemp(prefix="swit_lbfgs_", dir=root) + "/"
    wav = np.broadcast_to(workloads.ricker(w.nt, w.dt, w.f0), (w.ns, w.nt)).copy()

    def simulate(vp):
        model = base.model(w.nx, w.nz, w.dx, w.dt, w.nt, True, w.npml, vp.astype(np.float64),
                           w.rho(vp).astype(np.float64))
        return base.simulate(model, base.source(w.f0, w.ns, w.src_xz.astype(np.float64), wav),
                             base.receiver(w.ng, w.rec_xz.astype(np.float64)), base.system(home, 1))

    optim = base.optimize('Waveform', 'LBFGS', args.iters, 0.05, 6000.0, 2000.0, 'Marine', 5, 0, 'None',
                          1.0, 10.0, False, 0.0, 'None', False, False, 0.0, 1e9)
    solver.STRIPS = strips
    init = {'vp': w.vp_init.astype(np.float64), 'rho': w.rho(w.vp_init).astype(np.float64)}
    t0 = time.perf_counter()
    problem = inversion_device.DeviceProblem(simulate(w.vp_init), optim, simu_true=simulate(w.vp_true))
    t_obs = time.perf_counter() - t0
    simu = simulate(w.vp_init)
    t0 = time.perf_counter()
    inversion_device.inversion(simu, optim, init, problem=problem)
    el = time.perf_counter() - t0
    problem.close()
    f = np.loadtxt(home + 'outputs/misfit_data.dat')
    steps = np.loadtxt(home + 'outputs/line_search_iteration.dat').reshape(-1)
    shutil.rmtree(home, ignore_errors=True)

    line = {
        "metric": "L-BFGS iterations of the device-resident inversion loop, 801x267x6000, 81 shots",
        "value": round(args.iters / el * 60.0, 3), "unit": "iterations/min", "n_gpus": 1,
        "steps": args.iters, "warmup": 0, "ms_per_step": round(1000 * el / args.iters, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded folded layers 2000-6000 m/s, Ricker 5 Hz)",
        "config": {"workload": f"{w.nx}x{w.nz} dx {w.dx} npml {w.npml} nt {w.nt} dt {w.dt}, {w.ns} shots, "
                               f"{w.ng} receivers, waveform misfit, device-resident loop (no SU files)",
                   "strips": strips, "cross_shot_sum": "f64"},
        "misfit": [float(x) for x in f], "line_search_forwards": [int(x) for x in steps],
        "observed_data_s": round(t_obs, 2),
    }
    print(json.dumps(line), flush=True)


def bench_precision(args):
    """GPU half of the 801x267x6000 precision sample: two shots (the first and the
    middle one) of the 81-shot workload, L2 gradient with FP32 and 16-bit strips, the
    waveform misfit; written to --out (npz) for tests/precision_c5.py, which runs the
    FP64 oracle on the same two shots in a process of its own (no GPU runtime there)."""
    import torch
    from scipy.integrate import cumulative_trapezoid
    from swit_amd import workloads
    from swit_amd.engine import Plan, misfit
    w = workloads.config5()
    sel = [0, w.ns // 2]
    dev = torch.device("cuda", 0)
    src = cumulative_trapezoid(workloads.ricker(w.nt, w.dt, w.f0), initial=0).astype(np.float32)
    src = torch.from_numpy(np.broadcast_to(src, (len(sel), w.nt)).copy()).to(dev)
    out = {"sel": np.array(sel)}
    for st in ("f32", "f16", "bf16"):
        p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz[sel], w.rec_xz[sel], strips=st)
        p.set_model(torch.from_numpy(w.vp_true).to(dev), torch.from_numpy(w.rho(w.vp_true)).to(dev))
        obs = p.forward(src, save_strips=False).clone()
        p.set_model(torch.from_numpy(w.vp_init).to(dev), torch.from_numpy(w.rho(w.vp_init)).to(dev))
        syn = p.forward(src, save_strips=True)
        J, adj, _ = misfit("Waveform", obs, syn, w.dt)
        g, fw, bk = p.adjoint(src, adj)
        out["grad_" + st] = g.cpu().numpy()
        out["misfit_" + st] = float(J)
        torch.cuda.synchronize()
        p.close()
        print(f"{st}: misfit {out['misfit_' + st]:.9e}", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.savez(args.out, **out)
    print(json.dumps({"workload": "801x267x6000 precision sample, shots " + str(sel), "out": args.out}))


def bench_uneven(args):
    """Three grid families in one plan group: each step is one gradient of all 64 shots,
    as a group forward with strips (every plan's time loop on a branch of one graph
    launch), the global-correlation misfit and adjoint source on the device, and a group
    adjoint.  Timed beside it: the whole gradient in one graph launch
    (swit_group_gradient, where one plan's adjoint runs beside another's forward), and
    the plans one after the other."""
    import torch
    from scipy.integrate import cumulative_trapezoid
    from swit_amd import workloads
    from swit_amd.engine import Plan, PlanGroup, misfit

    ws = workloads.config4()
    dev = torch.device("cuda", 0)
    srcs, plans = [], []
    for w in ws:
        wav = workloads.ricker(w.nt, w.dt, w.f0)
        srcs.append(torch.from_numpy(np.broadcast_to(cumulative_trapezoid(wav, initial=0).astype(np.float32),
                                                     (w.ns, w.nt)).copy()).to(dev))
        p = Plan(w.nx, w.nz, w.dx, w.dt, w.nt, w.npml, w.src_xz, w.rec_xz, path="batched", strips=args.strips)
        p.set_model(torch.from_numpy(w.vp_true).to(dev), torch.from_numpy(w.rho(w.vp_true)).to(dev))
        if not args.no_autotune:
            p.autotune(steps=30)
        if args.bwd_steps:
            p.set_option("bwd_steps", args.bwd_steps)
        plans.append(p)
    group = PlanGroup(plans)
    obs = [o.clone() for o in group.forward(srcs, save_strips=False)]
    models = [(torch.from_numpy(w.vp_init).to(dev), torch.from_numpy(w.rho(w.vp_init)).to(dev)) for w in ws]
    torch.cuda.synchronize()

    def step(mode):
        """one gradient of all 64 shots: 'gradient' = one graph launch per gradient
        (swit_group_gradient), 'phases' = a group forward, the misfits, a group
        adjoint, 'plans' = the plans one after the other"""
        for p, (vp, rho) in zip(plans, models):
            p.set_model(vp, rho)
        if mode == "gradient":
            Js, _, out = group.gradient(srcs, obs, "Globalcorrelation")
            J = Js[0]
            for x in Js[1:]:
                J = J + x
            return J, out
        if mode == "phases":
            syn = group.forward(srcs, save_strips=True)
        else:
            syn = [p.forward(s, save_strips=True) for p, s in zip(plans, srcs)]
        res = [misfit("Globalcorrelation", o, s_, w.dt) for o, s_, w in zip(obs, syn, ws)]
        adj = [r[1] for r in res]
        if mode == "phases":
            out = group.adjoint(srcs, adj)
        else:
            out = [p.adjoint(s, a) for p, s, a in zip(plans, srcs, adj)]
        J = res[0][0]
        for r in res[1:]:
            J = J + r[0]
        return J, out

    def timed(mode):
        for _ in range(args.warmup):
            step(mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            J, _ = step(mode)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, float(J)

    el, J = timed("phases")
    el_one, J_one = timed("gradient")
    el_seq, J_seq = timed("plans")
    nshots = sum(w.ns for w in ws)
    line = {
        "metric": "shot-gradients/s, uneven 3-grid batch in one plan group (global-correlation misfit)",
        "value": round(nshots * args.steps / el, 3), "unit": "shot-gradients/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1000 * el / args.steps, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded layered-plus-fault models 1500-6000 m/s, Ricker 5 Hz)",
        "config": {"workload": "; ".join(f"{w.nx}x{w.nz} dx {w.dx} nt {w.nt} dt {w.dt}, {w.ns} shots" for w in ws),
                   "path": "batched group: group forward, misfits, group adjoint", "strips": args.strips,
                   "shots_per_step": nshots},
        "misfit": J,
        "one_launch": {"value": round(nshots * args.steps / el_one, 3), "ms_per_step": round(1000 * el_one / args.steps, 2),
                       "misfit": J_one, "note": "whole gradient in one graph launch (swit_group_gradient)"},
        "plan_by_plan": {"value": round(nshots * args.steps / el_seq, 3), "ms_per_step": round(1000 * el_seq / args.steps, 2),
                         "misfit": J_seq, "group_over_plan_by_plan": round(el_seq / el, 3)},
    }
    print(json.dumps(line), flush=True)
    group.close()
    for p in plans:
        p.close()


if __name__ == "__main__":
    main()
