"""Host-side formats and caller modules (CPU): SU files, solver input rounding,
parfile text, Ricker / source integration, misfits and adjoint sources."""
import os
import struct

import numpy as np
import pytest
from scipy.integrate import cumulative_trapezoid

from swit_amd import misfit, su
from swit_amd import solver as S


def test_su_header_byte_layout(tmp_path):
    """Standard SU/SEG-Y trace header offsets: tracl 0, tracr 4, fldr 8, ep 16, cdp 20,
    cdpt 24, offset 36, sx 72, gx 80, ns 114, dt 116, d1 180, f1 184, d2 188, f2 192;
    values as io.f90:522-571 sets them."""
    p = tmp_path / "g.su"
    data = np.arange(3 * 5, dtype=np.float32).reshape(3, 5)
    xg = np.array([0.0, 25.0, 37.5], np.float32)
    su.write(p, data, np.float32(0.0019), np.float32(12.5), xg)
    raw = p.read_bytes()
    assert len(raw) == 3 * (240 + 5 * 4)
    for i in range(3):
        h = raw[i * 260:(i + 1) * 260]
        assert struct.unpack_from("<i", h, 0)[0] == i + 1       # tracl
        assert struct.unpack_from("<i", h, 4)[0] == i + 1       # tracr
        assert struct.unpack_from("<i", h, 8)[0] == 1           # fldr
        assert struct.unpack_from("<i", h, 16)[0] == 1          # ep
        sx = 13                                                  # nint(12.5) rounds away from 0
        gx = [0, 25, 38][i]
        assert struct.unpack_from("<i", h, 72)[0] == sx
        assert struct.unpack_from("<i", h, 80)[0] == gx
        assert struct.unpack_from("<i", h, 36)[0] == gx - sx    # offset
        cdp = [6, 19, 25][i]                                     # nint((12.5 + xg) / 2)
        assert struct.unpack_from("<i", h, 20)[0] == cdp
        assert struct.unpack_from("<i", h, 24)[0] == cdp
        assert struct.unpack_from("<h", h, 114)[0] == 5          # ns
        assert struct.unpack_from("<h", h, 116)[0] == 1900       # dt in microseconds
        assert struct.unpack_from("<f", h, 180)[0] == np.float32(0.0019)
        assert struct.unpack_from("<f", h, 184)[0] == 0.0
        assert struct.unpack_from("<f", h, 188)[0] == 1.0
        assert struct.unpack_from("<f", h, 192)[0] == 1.0
        np.testing.assert_array_equal(np.frombuffer(h[240:], "<f4"), data[i])
    d, hdr = su.read(p)
    np.testing.assert_array_equal(d, data)
    assert su.sampling(p) == (3, 5, pytest.approx(0.0019))


def test_nint_matches_fortran():
    np.testing.assert_array_equal(su.nint([0.5, 1.5, 2.5, -0.5, -1.5, 2.49]), [1, 2, 3, -1, -2, 2])


def test_source_wavelet_and_integration():
    nt, dt, f0 = 500, 0.002, 5.0
    w = S.source_wavelet(nt, dt, f0, "ricker")
    t = np.arange(nt) * dt
    a = (np.pi * f0 * (t - 1.2 / f0)) ** 2
    np.testing.assert_allclose(w, (1 - 2 * a) * np.exp(-a), rtol=1e-12, atol=1e-15)
    c = cumulative_trapezoid(w, initial=0)
    np.testing.assert_allclose(c[1:], np.cumsum(0.5 * (w[1:] + w[:-1])), rtol=1e-12, atol=1e-12)
    with pytest.raises(ValueError):
        S.source_wavelet(nt, dt, f0, "gaussian")


class _Obj:
    pass


def _simu(tmp_path, nx=20, nz=10, ns=2):
    simu = _Obj()
    simu.model, simu.source, simu.receiver, simu.system = _Obj(), _Obj(), _Obj(), _Obj()
    m = simu.model
    m.nx, m.nz, m.dx, m.dt, m.nt, m.pml, m.fs, m.savestep = nx, nz, 12.5, 0.0013, 100, 10, True, 1
    m.vp = np.full((nx, nz), 2000.0)
    m.rho = np.full((nx, nz), 2100.0)
    simu.source.n = ns
    simu.source.xz = np.array([[10.04, 0.0], [120.06, 25.0]])[:ns]
    simu.source.wavelet = np.zeros((ns, m.nt))
    simu.receiver.n = nx
    simu.receiver.xz = np.stack([np.stack([np.arange(nx) * 12.5 + 0.04, np.zeros(nx)], 1)] * ns)
    simu.system.homepath = str(tmp_path) + "/"
    return simu


def test_solver_inputs_round_like_the_files(tmp_path):
    inp = S.solver_inputs(_simu(tmp_path))
    assert inp["dt"] == float(np.float32(0.0013))
    assert inp["dx"] == float(np.float32(12.5))
    assert inp["src_xz"][0, 0] == np.float32(10.0)       # '%10.1f' of 10.04
    assert inp["src_xz"][1, 0] == np.float32(120.1)
    assert inp["rec_xz"][0, 3, 0] == np.float32(37.5)
    assert inp["vp"].dtype == np.float32 and inp["free_surface"] == 1


def test_write_parfile_text(tmp_path):
    simu = _simu(tmp_path)
    S.write_parfile(simu, "syn", np.zeros((2, 100)), savesnap=1)
    txt = open(tmp_path / "parfile/forward_parfile/parfile").read().splitlines()
    assert "jobtype=forward_modeling" in txt
    assert "DX=12.500000" in txt and "DT_WORK=0.001300" in txt
    assert "STORE_SNAP=1" in txt and "FREESURFACE=1" in txt and "NPML=10" in txt
    coord = open(tmp_path / "parfile/model/coord.txt").read().splitlines()
    assert coord[0] == "%6i %6i %10.1f %10.1f %10.1f %10.1f %10.1f" % (1, 1, 10.04, 0.0, 0.04, 0.0, 1)
    assert len(coord) == 2 * 20
    assert np.fromfile(tmp_path / "parfile/model/vel.bin", np.float32).size == 200


# -- misfits -------------------------------------------------------------------------------
def _ricker_gather(nt=400, dt=0.002, shifts=(0, 3, -5)):
    t = np.arange(nt) * dt
    out = []
    for k in shifts:
        a = (np.pi * 8 * (t - 0.3 - k * dt)) ** 2
        out.append((1 - 2 * a) * np.exp(-a))
    return np.array(out)


def test_waveform_misfit_and_adjoint():
    obs = _ricker_gather(shifts=(0, 0, 0))
    syn = _ricker_gather(shifts=(0, 3, -5))
    dt = 0.002
    r = syn - obs
    assert misfit.misfit_waveform(obs, syn, dt) == pytest.approx(np.sum(np.sqrt(np.sum(r * r * dt, -1))))
    np.testing.assert_allclose(misfit.adjoint_source_waveform(obs, syn, dt), r)
    z = obs.copy()
    z[1] = 0.0                                           # dead trace contributes nothing
    assert misfit.adjoint_source_waveform(z, syn, dt)[1].max() == 0.0


def test_traveltime_lag_is_the_shift():
    obs = _ricker_gather(shifts=(0, 0, 0))
    syn = _ricker_gather(shifts=(0, 3, -5))
    nt = obs.shape[1]
    lag = misfit.cross_correlate_max(obs, syn) - nt + 1
    np.testing.assert_array_equal(np.abs(lag), [0, 3, 5])
    assert misfit.misfit_traveltime(obs, syn, 0.002) == pytest.approx(0.5 * (9 + 25) * 0.002)


def test_global_correlation_misfit_and_adjoint():
    obs = _ricker_gather(shifts=(0, 0, 0))
    syn = 2.5 * _ricker_gather(shifts=(0, 3, -5))
    dt = 0.002
    assert misfit.misfit_global_correlation(obs, obs * 3, dt) == pytest.approx(-3.0)
    # for zero-mean traces the coefficient is the normalised dot product, and the
    # adjoint source is its exact derivative
    o = obs - obs.mean(-1, keepdims=True)
    s = syn - syn.mean(-1, keepdims=True)
    adj = misfit.adjoint_source_global_correlation(o, s, dt)
    rng = np.random.default_rng(0)
    ds = rng.standard_normal(s.shape)
    ds -= ds.mean(-1, keepdims=True)
    eps = 1e-6

    def J(x):
        return -np.sum(np.sum((x / np.linalg.norm(x, axis=-1, keepdims=True)) *
                              (o / np.linalg.norm(o, axis=-1, keepdims=True)), -1))
    fd = (J(s + eps * ds) - J(s - eps * ds)) / (2 * eps)
    assert np.sum(adj * ds) == pytest.approx(fd, rel=1e-6)


def test_envelope_zero_for_identical_traces():
    obs = _ricker_gather()
    assert misfit.misfit_envelope(obs, obs, 0.002) == pytest.approx(0.0, abs=1e-20)
    np.testing.assert_allclose(misfit.adjoint_source_envelope(obs, obs, 0.002), 0.0, atol=1e-12)


def test_envelope_adjoint_is_the_misfit_gradient():
    """The envelope misfit J = dt sum r^2 (r = E_syn - E_obs) has dJ/dsyn = 2 dt adj:
    adj = 2 s r - H(2 H(s) r) uses that the discrete Hilbert transform is antisymmetric
    (H^T = -H), so a dropped sign or an untransposed H fails this directional check."""
    rng = np.random.default_rng(3)
    obs = rng.standard_normal((3, 96))
    syn = rng.standard_normal((3, 96))
    dt = 0.004
    adj = misfit.adjoint_source_envelope(obs, syn, dt)
    ds = rng.standard_normal(syn.shape)
    eps = 1e-6
    fd = (misfit.misfit_envelope(obs, syn + eps * ds, dt) - misfit.misfit_envelope(obs, syn - eps * ds, dt)) / (2 * eps)
    assert 2 * dt * np.sum(adj * ds) == pytest.approx(fd, rel=1e-7)


def test_hilbert_transform_of_cosine_is_sine():
    """scipy.signal.hilbert on a whole number of periods: H cos = sin, so the envelope is
    flat (closed form used by the device envelope test)."""
    from scipy.signal import hilbert
    n = 240
    t = np.arange(n)
    for k in (1, 7, 119):
        np.testing.assert_allclose(hilbert(np.cos(2 * np.pi * k * t / n)).imag, np.sin(2 * np.pi * k * t / n),
                                   atol=1e-12)


class _RecordingPlan:
    """Stands in for swit_amd.engine.Plan to check the drop-in's plan bookkeeping on the
    CPU: every call is logged, seismograms are zeros of the right shape."""
    log = []

    def __init__(self, nx, nz, dx, dt, nt, npml, src_xz, rec_xz, **kw):
        import torch
        self.ns, self.ng, self.nt, self.nx, self.nz = src_xz.shape[0], rec_xz.shape[1], nt, nx, nz
        self.strips = kw.get("strips", "f32")
        self.id = len([e for e in self.log if e[0] == "new"])
        self.log.append(("new", self.id))
        self.torch = torch
        self.options = {}

    def set_option(self, option, value):
        self.options[option] = value

    def set_model(self, vp, rho):
        self.log.append(("model", self.id, float(vp.double().sum())))

    def autotune(self, steps=20):
        self.log.append(("tune", self.id))
        return {}

    def forward(self, src, save_strips):
        self.log.append(("forward", self.id, int(save_strips)))
        return self.torch.zeros(self.ns, self.ng, self.nt)

    def adjoint(self, src, adj):
        self.log.append(("adjoint", self.id))
        z = self.torch.ones(self.nx, self.nz, dtype=self.torch.float64)
        return z, z, z

    def close(self):
        self.log.append(("close", self.id))


def test_dropin_plan_bookkeeping(tmp_path, monkeypatch):
    from swit_amd import engine, preprocess, solver, workloads
    from test_gpu_dropin import _optim, _simu
    import types
    monkeypatch.setattr(engine, "Plan", _RecordingPlan)
    # no GPU here for the preconditioner: the raw kernel stands in for it
    monkeypatch.setattr(solver, "_postprocess_module", lambda: types.SimpleNamespace(
        grad_precond=lambda simu, optim, grad, forw, back: np.asarray(grad, np.float64)))
    _RecordingPlan.log = []
    log = _RecordingPlan.log
    w = workloads.small(seed=22, nx=40, nz=30, npml=8, nt=60, ns=2)
    w.rec_xz = w.rec_xz[:, :w.nx]
    home = str(tmp_path) + "/"
    simu = _simu(home, w, w.vp_init)
    simu_true = _simu(home, w, w.vp_true)
    optim = _optim()
    solver.forward(simu, 'syn', savesnap=1)
    solver.forward(simu_true, 'obs', savesnap=0)
    # the observed run used a plan of its own (kept for the next observed run); the
    # synthetic plan (0) saw one model, the autotuner (once, before any strip-saving
    # forward) and one strip-saving forward
    assert [e for e in log if e[1] == 0] == [("new", 0), ("model", 0, float(np.float32(w.vp_init).astype(np.float64).sum())),
                                             ("tune", 0), ("forward", 0, 1)]
    assert [e[0] for e in log if e[1] == 1] == ["new", "model", "forward"]
    assert solver._OBS_JOBS[home].plan.strips == "none"
    # both batched plans run the halves of the shots on two streams
    assert solver._OBS_JOBS[home].plan.options == {"streams": 2}
    assert solver._JOBS[home].plan.options == {"streams": 2}
    solver.forward(simu_true, 'obs', savesnap=0)           # same model: no new plan, no upload
    assert [e[0] for e in log if e[1] == 1] == ["new", "model", "forward", "forward"]
    preprocess.process_workflow(simu, optim, 'obs')
    preprocess.process_workflow(simu, optim, 'syn')
    solver.adjoint(simu, optim)
    assert log[-1] == ("adjoint", 0)
    # a model change after the saving forward: the forward is rerun in the new model
    n = len(log)
    simu.model.vp = simu.model.vp * 1.01
    solver.adjoint(simu, optim)
    assert [e[0] for e in log[n:]] == ["model", "forward", "adjoint"] and log[n + 1][2] == 1
    # a forward without snapshots drops the saved state
    solver.forward(simu, 'syn', savesnap=0)
    preprocess.process_workflow(simu, optim, 'syn')
    with pytest.raises(ValueError, match='Adjoint solver crash'):
        solver.adjoint(simu, optim)
    solver.release()
    assert sorted(log[-2:]) == [("close", 0), ("close", 1)]
