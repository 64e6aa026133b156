"""The drop-in solver (swit_amd.solver, toolbox/solver.py's surface) on the GPU against
the oracle: SU seismograms and the raw kernel / illumination files bit-exact, the
preconditioned gradient identical, for the reference call sequence
(toolbox/inversion.py:48,54,57,60)."""
import os

import numpy as np
import pytest

from helpers import integrated_sources
from oracle import oracle as O
from swit_amd import workloads


def _simu(home, w, vp):
    from swit_amd import base
    model = base.model(w.nx, w.nz, w.dx, w.dt, w.nt, True, w.npml, vp.astype(np.float64),
                       workloads.gardner(vp).astype(np.float64))
    wav = np.broadcast_to(workloads.ricker(w.nt, w.dt, w.f0), (w.ns, w.nt)).copy()
    source = base.source(w.f0, w.ns, w.src_xz.astype(np.float64), wav)
    receiver = base.receiver(w.rec_xz.shape[1], w.rec_xz.astype(np.float64))
    return base.simulate(model, source, receiver, base.system(home, 4))


def _optim():
    from swit_amd import base
    return base.optimize('Waveform', 'LBFGS', 1, 0.05, 3500.0, 1400.0, 'Marine', 3, 2, 'None', 1.0, 10.0,
                         False, 0.0, 'None', False, False, 0.0, 1e9)


@pytest.mark.gpu
def test_dropin_sequence_matches_oracle(tmp_path):
    from oracle import precond as P
    from swit_amd import misfit, preprocess, solver, su, tools
    w = workloads.small(seed=21, nx=56, nz=36, npml=10, nt=300, ns=3)
    w.rec_xz = w.rec_xz[:, :w.nx]                     # surface receivers, all inside x range
    home = str(tmp_path) + "/"
    simu_true = _simu(home, w, w.vp_true)
    simu = _simu(home, w, w.vp_init)
    optim = _optim()
    solver.forward(simu_true, 'obs', savesnap=0)
    solver.forward(simu, 'syn', savesnap=1)
    assert os.getcwd() == os.path.realpath(home.rstrip("/"))
    preprocess.process_workflow(simu, optim, 'obs')
    preprocess.process_workflow(simu, optim, 'syn')
    J = misfit.misfit(simu, 'Waveform')
    g = solver.adjoint(simu, optim)

    src = integrated_sources(w)
    inp = solver.solver_inputs(simu)
    mt = O.Model(w.vp_true, workloads.gardner(w.vp_true), w.dx, w.npml, "f32")
    m = O.Model(w.vp_init, workloads.gardner(w.vp_init), w.dx, w.npml, "f32")
    obs = O.forward(mt, w.nt, inp["dt"], inp["src_xz"], inp["rec_xz"], src, save_strips=False)["seis"]
    f = O.forward(m, w.nt, inp["dt"], inp["src_xz"], inp["rec_xz"], src)
    for i in range(w.ns):
        d_obs, _ = su.read(home + 'data/obs/src%d_sg.su' % (i + 1))
        d_syn, _ = su.read(home + 'data/syn/src%d_sg.su' % (i + 1))
        np.testing.assert_array_equal(d_obs, obs[i])
        np.testing.assert_array_equal(d_syn, f["seis"][i])
    adj = (f["seis"].astype(np.float64) - obs.astype(np.float64))
    a = O.adjoint(m, w.nt, inp["dt"], inp["src_xz"], inp["rec_xz"], src, adj.astype(np.float32), f)
    kern = tools.loadbinfloat32(home + 'data/syn/src0_kernel_vp.bin')
    np.testing.assert_array_equal(kern, a["grad"].astype(np.float32).ravel())
    np.testing.assert_array_equal(tools.loadbinfloat32(home + 'data/syn/src0_illum_forw.bin'),
                                  a["illum_forw"].astype(np.float32).ravel())
    f32 = [a[k].astype(np.float32) for k in ("grad", "illum_forw", "illum_back")]
    ref = P.precondition(*f32, optim.vpmax, True, optim.grad_mute, optim.grad_thred, optim.grad_smooth,
                         optim.grad_mask).ravel()
    # adjoint() preconditions on the device: the oracle's numpy steps to FP64 rounding
    assert np.linalg.norm(g - ref) <= 1e-12 * np.linalg.norm(ref)
    r = f["seis"].astype(np.float64) - obs
    assert J == pytest.approx(sum(misfit.misfit_waveform(obs[i], f["seis"][i], inp["dt"]) for i in range(w.ns)))
    assert J > 0
    solver.release()


@pytest.mark.gpu
def test_obs_forward_keeps_saved_state_and_syn_without_snapshots_clears_it(tmp_path):
    from swit_amd import preprocess, solver
    w = workloads.small(seed=22, nx=40, nz=30, npml=8, nt=200, ns=2)
    w.rec_xz = w.rec_xz[:, :w.nx]
    home = str(tmp_path) + "/"
    simu = _simu(home, w, w.vp_init)
    simu_true = _simu(home, w, w.vp_true)
    optim = _optim()
    solver.forward(simu, 'syn', savesnap=1)
    solver.forward(simu_true, 'obs', savesnap=0)       # must not discard the saved state
    preprocess.process_workflow(simu, optim, 'obs')
    preprocess.process_workflow(simu, optim, 'syn')
    g = solver.adjoint(simu, optim)
    assert np.all(np.isfinite(g)) and np.abs(g).max() > 0
    solver.forward(simu, 'syn', savesnap=0)            # line-search forward: state gone
    preprocess.process_workflow(simu, optim, 'syn')
    with pytest.raises(ValueError, match='Adjoint solver crash'):
        solver.adjoint(simu, optim)
    solver.release()


def _file_loop_callbacks(simu, optim):
    """The reference loop's solver calls through files (toolbox/inversion.py:48-60 for a
    gradient, :319-325 for a line-search trial) as callbacks of the oracle loop."""
    from swit_amd import misfit, preprocess, solver
    nx, nz = simu.model.nx, simu.model.nz

    def evaluate(v):
        simu.model.vp = np.asarray(v, np.float64).reshape(nx, nz)
        solver.forward(simu, 'syn', savesnap=1)
        preprocess.process_workflow(simu, optim, 'syn')
        f = misfit.misfit(simu, optim.misfit_type)
        return f, solver.adjoint(simu, optim)

    def trial(v):
        simu.model.vp = np.asarray(v, np.float64).reshape(nx, nz)
        solver.forward(simu, 'syn', savesnap=0)
        preprocess.process_workflow(simu, optim, 'syn')
        return misfit.misfit(simu, optim.misfit_type)
    return evaluate, trial


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["LBFGS", "NLCG"])
def test_reference_loop_on_the_dropin_reduces_misfit(tmp_path, scheme):
    """Three iterations of the reference inversion loop's call sequence (the oracle's
    restatement) on the drop-in solver through its SU and .bin files: every accepted
    model lowers the waveform misfit."""
    from oracle import inversion_loop as L
    from swit_amd import base, preprocess, solver
    w = workloads.small(seed=23, nx=60, nz=36, npml=10, nt=300, ns=3)
    w.rec_xz = w.rec_xz[:, :w.nx]
    home = str(tmp_path) + "/"
    simu_true = _simu(home, w, w.vp_true)
    optim = base.optimize('Waveform', scheme, 3, 0.05, 3500.0, 1400.0, 'Marine', 3, 2, 'None', 1.0, 10.0,
                          False, 0.0, 'None', False, False, 0.0, 1e9)
    solver.forward(simu_true, 'obs', savesnap=0)
    preprocess.process_workflow(simu_true, optim, 'obs')
    simu = _simu(home, w, w.vp_init)
    loop = L.ReferenceLoop(w.vp_init.astype(np.float64).ravel(), scheme, optim.step_length, optim.vpmax, optim.vpmin)
    evaluate, trial = _file_loop_callbacks(simu, optim)
    for _ in range(3):
        loop.iterate(evaluate, trial)
    f = [loop.f_first] + [h[0] for h in loop.history]
    print(scheme, f)
    assert np.all(np.diff(f) < 0)
    solver.release()


@pytest.mark.gpu
@pytest.mark.parametrize("kind,filt,norm,mute", [("Waveform", "None", "None", False),
                                                 ("Globalcorrelation", "Bandpass", ["L2-Trace"], False),
                                                 ("Envelope", "Bandpass", ["Max-Trace"], True)])
def test_device_loop_matches_reference_loop(tmp_path, kind, filt, norm, mute):
    """The device-resident loop (swit_amd.inversion_device: device processing, misfit,
    adjoint source, preconditioning and L-BFGS vectors, no SU files) against the
    reference loop's call sequence (the oracle's restatement) on the drop-in solver
    through its files, over three L-BFGS iterations: the same line-search decisions,
    misfits to 1e-4 (the logs keep 5 digits) and models to float32 rounding of the
    different inner products and processing."""
    from oracle import inversion_loop as L
    from swit_amd import base, inversion_device, preprocess, solver
    from swit_amd.tools import loadbinfloat32
    w = workloads.small(seed=24, nx=60, nz=36, npml=10, nt=300, ns=3)
    w.rec_xz = w.rec_xz[:, :w.nx]

    def optim():
        # with `mute`: traces closer than 45 m to the source are zeroed
        return base.optimize(kind, 'LBFGS', 3, 0.05, 3500.0, 1400.0, 'Marine', 3, 2, filt, 3.0, 25.0,
                             False, 0.0, norm, mute, False, 45.0, 1e9)
    init = {'vp': w.vp_init.astype(np.float64), 'rho': workloads.gardner(w.vp_init).astype(np.float64)}
    home_a = str(tmp_path / 'a') + "/"
    os.makedirs(home_a)
    simu_true = _simu(home_a, w, w.vp_true)
    o = optim()
    solver.forward(simu_true, 'obs', savesnap=0)
    preprocess.process_workflow(simu_true, o, 'obs')
    simu = _simu(home_a, w, w.vp_init)
    simu.model.rho = init['rho']
    ref = L.ReferenceLoop(init['vp'].ravel(), 'LBFGS', o.step_length, o.vpmax, o.vpmin)
    evaluate, trial = _file_loop_callbacks(simu, o)
    for _ in range(3):
        ref.iterate(evaluate, trial)
    solver.release()

    home_b = str(tmp_path / 'b') + "/"
    os.makedirs(home_b)
    simu = _simu(home_b, w, w.vp_init)
    o = optim()
    inversion_device.inversion(simu, o, dict(init), simu_true=_simu(home_b, w, w.vp_true))
    fa = np.array([ref.f_first] + [h[0] for h in ref.history])
    fb = np.loadtxt(home_b + 'outputs/misfit_data.dat')
    print(kind, fa, fb)
    np.testing.assert_array_equal(np.loadtxt(home_b + 'outputs/line_search_iteration.dat'),
                                  [h[1] for h in ref.history])
    np.testing.assert_allclose(fb, fa, rtol=1e-4)
    vb = loadbinfloat32(home_b + 'outputs/velocity/vp-3.bin')
    np.testing.assert_allclose(vb, ref.models[-1], rtol=1e-4)
    # a failed line search (an L-BFGS restart) keeps the misfit; none increases it
    assert np.all(np.diff(fb) <= 0) and fb[-1] < fb[0]