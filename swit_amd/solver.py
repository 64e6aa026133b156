"""Drop-in replacement of the reference's solver adaptor (toolbox/solver.py).

Same call surface and side effects for its callers (toolbox/inversion.py:48,60,325,
example workflows):

    forward(simu, simu_type='obs' | 'syn', savesnap=0 | 1) -> None
        writes data/<simu_type>/src<i>_sg.su (1-based, unpadded) and chdirs to
        homepath; 'syn' first clears data/syn/ (toolbox/solver.py:24-58)
    adjoint(simu, optim) -> 1-D (nx*nz) preconditioned gradient
        adjoint sources from the misfit module, gradient / illuminations written to
        data/syn/src0_kernel_vp.bin, src0_illum_forw.bin, src0_illum_back.bin, then
        grad_precond (toolbox/solver.py:61-92)
    write_parfile(simu, simu_type, src, savesnap=0), source_wavelet(nt, dt, f0, srctype)

Instead of `mpirun fd2dmpi` with files in between, both jobs run on the GPU through
libswit_fd.so (swit_amd.engine.Plan, batched path): the boundary strips and final
state of the last strip-saving 'syn' forward stay in device memory for the adjoint.
Observed-data forwards never disturb that state: they run on a plan of their own that
holds no strips and is kept for the next call.  Seismograms go to the host and into
their SU files one shot at a time.  Inputs are rounded exactly as the reference's file
round trip rounds them (float32 models and sources, '%f' dx/dt, '%10.1f' coordinates).
Any solver failure raises ValueError('Forward solver crash' / 'Adjoint solver crash')
as the reference does.
"""
from __future__ import annotations

import os

import numpy as np
from scipy.integrate import cumulative_trapezoid

from . import su
from .tools import cleandata, loadbinfloat32, savebinfloat32

# Write the reference's parfile / model / source / coord.txt files as well (only
# needed to run the Fortran solver on the same inputs; no caller reads them).
WRITE_PARFILES = False
# Device path of the plan: 'batched' (default) or 'plain' (per-shot launches).
PATH = "batched"
# Boundary-strip storage: 'f32' (reference precision), 'f16' or 'bf16'.
STRIPS = "f32"
# Time the batched step kernels' launch configurations once per synthetic-data plan,
# at its first model, and keep the fastest (the results do not depend on the choice).
AUTOTUNE = True
# Batched plans with two or more shots run the two halves of the shots on two streams,
# so that each time step's two launches overlap (bit-identical results; 1 turns it off).
STREAMS = 2

_JOBS = {}         # synthetic-data plan per homepath
_OBS_JOBS = {}     # observed-data plan (no strips) per homepath


def _misfit_module():
    from . import misfit
    return misfit


def _postprocess_module():
    from . import postprocess
    return postprocess


class _Job:
    """Device plan and saved-state bookkeeping for one homepath and geometry."""

    def __init__(self, key, inputs, strips=None):
        from .engine import Plan
        self.key = key
        self.plan = Plan(inputs["nx"], inputs["nz"], inputs["dx"], inputs["dt"], inputs["nt"],
                         inputs["npml"], inputs["src_xz"], inputs["rec_xz"],
                         free_surface=inputs["free_surface"], path=PATH, strips=strips or STRIPS)
        if PATH == "batched":
            self.plan.set_option("streams", STREAMS)
        self.fwd_src = None          # source traces of the last strip-saving forward
        self.model = None            # (vp, rho) the plan currently holds
        self.tuned = None            # the autotuner's choice, once made

    def set_model(self, vp, rho, tune=False):
        import torch
        if self.model is not None and np.array_equal(self.model[0], vp) and np.array_equal(self.model[1], rho):
            return
        self.plan.set_model(torch.from_numpy(vp), torch.from_numpy(rho))
        self.model = (vp.copy(), rho.copy())
        if tune and self.tuned is None and PATH == "batched":
            # before any strip-saving forward: tuning runs over the plan's buffers
            self.tuned = self.plan.autotune(steps=20)


def solver_inputs(simu):
    """The values the reference solver reads back from the files write_parfile writes:
    float32 models and sources, dx/dt through '%f', coordinates through '%10.1f'
    (toolbox/solver.py:104-180, fd2dmpi/io.f90:72-96,126-228)."""
    m = simu.model
    srcn, recn = simu.source.n, simu.receiver.n
    sxz = np.asarray(simu.source.xz, dtype=np.float64).reshape(srcn, 2)
    rxz = np.asarray(simu.receiver.xz, dtype=np.float64).reshape(srcn, recn, 2)
    rnd = np.vectorize(lambda v: float('%10.1f' % v))
    return {
        "nx": int(m.nx), "nz": int(m.nz), "nt": int(m.nt), "npml": int(m.pml),
        "dx": float(np.float32(float('%f' % m.dx))),
        "dt": float(np.float32(float('%f' % m.dt))),
        "free_surface": 1 if m.fs else 0,
        "vp": np.ascontiguousarray(np.asarray(m.vp, dtype=np.float32)),
        "rho": np.ascontiguousarray(np.asarray(m.rho, dtype=np.float32)),
        "src_xz": rnd(sxz).astype(np.float32),
        "rec_xz": rnd(rxz).astype(np.float32),
    }


def _key(inputs):
    return (inputs["nx"], inputs["nz"], inputs["nt"], inputs["npml"], inputs["dx"], inputs["dt"],
            inputs["free_surface"], inputs["src_xz"].tobytes(), inputs["rec_xz"].tobytes(), PATH, STRIPS)


def _job(simu, inputs):
    """The synthetic-data plan of simu's homepath (created or rebuilt on a new geometry)."""
    key = _key(inputs)
    home = simu.system.homepath
    job = _JOBS.get(home)
    if job is None or job.key != key:
        if job is not None:
            job.plan.close()
        job = _JOBS[home] = _Job(key, inputs)
    return job


def _obs_job(simu, inputs):
    """The observed-data plan of simu's homepath: no strips, kept across calls (created
    or rebuilt on a new geometry)."""
    key = ("obs",) + _key(inputs)
    home = simu.system.homepath
    job = _OBS_JOBS.get(home)
    if job is None or job.key != key:
        if job is not None:
            job.plan.close()
        job = _OBS_JOBS[home] = _Job(key, inputs, strips="none")
    return job


def release(homepath=None):
    """Free the device plans (all, or those of homepath)."""
    for jobs in (_JOBS, _OBS_JOBS):
        for h in ([homepath] if homepath else list(jobs)):
            job = jobs.pop(h, None)
            if job is not None:
                job.plan.close()


def forward(simu, simu_type='obs', savesnap=0):
    """forward solver (toolbox/solver.py:24-58)."""
    if simu_type not in ['obs', 'syn']:
        raise ValueError('Forward: unsupport simulation tyep.')
    homepath = simu.system.homepath
    srcn = simu.source.n
    if simu_type in ['syn']:
        cleandata(homepath + 'data/syn/')
    for isrc in range(srcn):
        ifolder = homepath + 'data/%s/src%d_snapshot' % (simu_type, isrc + 1)
        if not os.path.exists(ifolder) and savesnap == 1:
            os.makedirs(ifolder)
    src = cumulative_trapezoid(simu.source.wavelet, axis=-1, initial=0)
    if WRITE_PARFILES:
        write_parfile(simu, simu_type, src, savesnap=savesnap)
    os.chdir(homepath)
    try:
        import torch
        inp = solver_inputs(simu)
        src32 = np.array(np.broadcast_to(np.asarray(src, dtype=np.float32), (srcn, inp["nt"])))
        if simu_type == 'syn':
            # the synthetic plan keeps the boundary strips of a snapshot-saving run for
            # adjoint(); a run without snapshots drops them, as cleandata drops the files
            job = _job(simu, inp)
            job.set_model(inp["vp"], inp["rho"], tune=AUTOTUNE)
            seis = job.plan.forward(torch.from_numpy(src32), save_strips=1 if savesnap == 1 else 0)
            job.fwd_src = src32 if savesnap == 1 else None
        else:
            # observed data run on a plan of their own, holding no strips, so the
            # synthetic plan's model and saved state (data/syn/ in the reference) are
            # left as they are
            obs = _obs_job(simu, inp)
            obs.set_model(inp["vp"], inp["rho"])
            seis = obs.plan.forward(torch.from_numpy(src32), save_strips=0)
        # one shot's gather at a time to the host and into its SU file
        ng = simu.receiver.n
        for isrc in range(srcn):
            su.write(homepath + 'data/%s/src%d_sg.su' % (simu_type, isrc + 1), seis[isrc, :ng].cpu().numpy(),
                     inp["dt"], inp["src_xz"][isrc, 0], inp["rec_xz"][isrc, :, 0])
        del seis
    except Exception as exc:  # the reference reports a failed solver run this way
        print(exc)
        raise ValueError('Forward solver crash') from exc


def adjoint(simu, optim):
    """adjoint solver (toolbox/solver.py:61-92)."""
    homepath = simu.system.homepath
    misfit_type = optim.misfit_type
    src = _misfit_module().adjoint_source(simu, misfit_type)
    if WRITE_PARFILES:
        write_parfile(simu, 'adj', src, savesnap=0)
    os.chdir(homepath)
    try:
        import torch
        inp = solver_inputs(simu)
        job = _job(simu, inp)
        if job.fwd_src is None:
            raise RuntimeError('no saved forward state: run forward(simu, "syn", savesnap=1) first')
        if not (np.array_equal(job.model[0], inp["vp"]) and np.array_equal(job.model[1], inp["rho"])):
            # the model changed since the snapshot-saving forward: the reference would
            # reconstruct the old wavefield inside the new model; here the forward is
            # rerun in the new model so the gradient belongs to it
            job.set_model(inp["vp"], inp["rho"])
            job.plan.forward(torch.from_numpy(job.fwd_src), save_strips=1)
        adj = np.asarray(src, dtype=np.float32).reshape(simu.source.n, simu.receiver.n, inp["nt"])
        g, fw, bk = job.plan.adjoint(torch.from_numpy(job.fwd_src), torch.from_numpy(np.ascontiguousarray(adj)))
        grad = g.cpu().numpy().astype(np.float32)
        forw = fw.cpu().numpy().astype(np.float32)
        back = bk.cpu().numpy().astype(np.float32)
    except Exception as exc:
        print(exc)
        raise ValueError('Adjoint solver crash') from exc
    savebinfloat32(homepath + 'data/syn/src0_kernel_vp.bin', grad)
    savebinfloat32(homepath + 'data/syn/src0_illum_forw.bin', forw)
    savebinfloat32(homepath + 'data/syn/src0_illum_back.bin', back)
    grad = loadbinfloat32(homepath + 'data/syn/src0_kernel_vp.bin')
    forw = loadbinfloat32(homepath + 'data/syn/src0_illum_forw.bin')
    back = loadbinfloat32(homepath + 'data/syn/src0_illum_back.bin')
    return _postprocess_module().grad_precond(simu, optim, grad, forw, back)


def write_parfile(simu, simu_type, src, savesnap=0):
    """Parameter, source, model and geometry files of the reference solver
    (toolbox/solver.py:95-180)."""
    savestep = simu.model.savestep
    path = simu.system.homepath
    if simu_type in ['obs', 'syn']:
        parpath = path + 'parfile/forward_parfile/parfile'
        srcpath = path + 'parfile/forward_source/'
    elif simu_type in ['adj']:
        parpath = path + 'parfile/adjoint_parfile/parfile'
        srcpath = path + 'parfile/adjoint_source/'
    else:
        raise ValueError(simu_type)
    for d in (os.path.dirname(parpath), srcpath, path + 'parfile/model'):
        os.makedirs(d, exist_ok=True)
    for isrc in range(simu.source.n):
        savebinfloat32(srcpath + 'src%d.bin' % (isrc + 1), src[isrc, :])
    savebinfloat32(path + 'parfile/model/vel.bin', simu.model.vp)
    savebinfloat32(path + 'parfile/model/rho.bin', simu.model.rho)
    lines = ['######################################### ', '#                                         ',
             '#     fd2dmpi input parameter file        ', '#                                         ',
             '######################################### ', '                                          ']
    lines.append('jobtype=forward_modeling' if simu_type in ['obs', 'syn'] else 'jobtype=adjoint_modeling')
    lines.append('COORD_FILE=%s' % (path + 'parfile/model/coord.txt'))
    lines.append('DATA_OUT=%s' % (path + ('data/obs/src' if simu_type == 'obs' else 'data/syn/src')))
    lines.append('VEL_IN=%s' % (path + 'parfile/model/vel.bin'))
    lines.append('DENSITYFILE=%s' % (path + 'parfile/model/rho.bin'))
    lines.append('FILEFORMAT=su')
    lines.append('NX=%d' % simu.model.nx)
    lines.append('NZ=%d' % simu.model.nz)
    lines.append('DX=%f' % simu.model.dx)
    lines.append('NPML=%d' % simu.model.pml)
    lines.append('NT_WORK=%d' % simu.model.nt)
    lines.append('DT_WORK=%f' % simu.model.dt)
    lines.append('FREESURFACE=1' if simu.model.fs else 'FREESURFACE=0')
    lines.append('STORE_SNAP=%d' % savesnap)
    lines.append('STORE_STEP=%d' % savestep)
    with open(parpath, 'w') as fp:
        fp.write('\n'.join(lines) + '\n')
    srcn, recn = simu.source.n, simu.receiver.n
    srcxz, recxz = simu.source.xz, simu.receiver.xz
    with open(path + 'parfile/model/coord.txt', 'w') as fp:
        for isrc in range(srcn):
            for irec in range(recn):
                fp.write('%6i %6i %10.1f %10.1f %10.1f %10.1f %10.1f\n' % (
                    isrc + 1, irec + 1, srcxz[isrc, 0], srcxz[isrc, 1],
                    recxz[isrc, irec, 0], recxz[isrc, irec, 1], 1))


def source_wavelet(nt, dt, f0, srctype):
    """Ricker with t0 = 1.2/f0 (toolbox/solver.py:183-197)."""
    t = np.linspace(0, dt * nt, num=nt, endpoint=False)
    if srctype.lower() in ['ricker']:
        t0 = 1.2 / f0
        temp = (np.pi * f0) ** 2
        return (1 - 2 * temp * (t - t0) ** 2) * np.exp(- temp * (t - t0) ** 2)
    raise ValueError('Other wavelets can be implemented here.')
