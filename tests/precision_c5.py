"""CPU half of the 801x267x6000 precision sample (not collected by pytest: it takes a
few minutes).  `python bench.py --workload precision --out P.npz` writes the GPU's
two-shot gradients (FP32, FP16 and BF16 strips) and waveform misfit; this script runs
the FP64 oracle on the same two shots of the same workload, with the same FP32 source
traces and the same time step, and prints the relative L2 errors as one JSON line.

    python tests/precision_c5.py P.npz [threads]

The oracle runs in this process only; no GPU runtime is loaded here."""
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

from scipy.integrate import cumulative_trapezoid  # noqa: E402

from oracle import oracle as O  # noqa: E402
from swit_amd import misfit, workloads  # noqa: E402


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def main():
    gpu = np.load(sys.argv[1])
    threads = int(sys.argv[2]) if len(sys.argv) > 2 else 2
    w = workloads.config5()
    sel = [int(i) for i in gpu["sel"]]
    dt = float(np.float32(w.dt))
    src = cumulative_trapezoid(workloads.ricker(w.nt, w.dt, w.f0), initial=0).astype(np.float32)
    src = np.broadcast_to(src, (len(sel), w.nt)).copy()
    sxz, rxz = w.src_xz[sel], w.rec_xz[sel]
    t0 = time.perf_counter()
    mt = O.Model(w.vp_true, w.rho(w.vp_true), w.dx, w.npml, "f64")
    obs = O.forward(mt, w.nt, dt, sxz, rxz, src, save_strips=False, threads=threads)["seis"]
    m = O.Model(w.vp_init, w.rho(w.vp_init), w.dx, w.npml, "f64")
    f = O.forward(m, w.nt, dt, sxz, rxz, src, threads=threads)
    adj = f["seis"] - obs
    g64 = O.adjoint(m, w.nt, dt, sxz, rxz, src, adj, f, threads=threads)["grad"]
    J64 = sum(misfit.misfit_waveform(obs[i], f["seis"][i], dt) for i in range(len(sel)))
    out = {"workload": "801x267x6000, shots %s of 81, L2 gradient" % sel, "oracle_s": round(time.perf_counter() - t0, 1),
           "misfit_fp64": J64}
    for st in ("f32", "f16", "bf16"):
        out["grad_rel_l2_" + st] = rel_l2(gpu["grad_" + st], g64)
        out["misfit_rel_" + st] = abs(float(gpu["misfit_" + st]) - J64) / J64
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
