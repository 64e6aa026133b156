"""Device misfits and adjoint sources against the host restatement of
toolbox/misfit.py (swit_amd.misfit): waveform adjoint source bit-exact (FP32
difference), misfits and global-correlation adjoint sources to FP64-vs-FP32 rounding
(relative 1e-6); dead traces give zero; the cross-trace sum is reproducible."""
import numpy as np
import pytest

from swit_amd import misfit as host


def _gathers(seed=0, ntr=37, nt=1500):
    rng = np.random.default_rng(seed)
    t = np.arange(nt) * 0.002
    obs = np.zeros((ntr, nt), np.float32)
    syn = np.zeros((ntr, nt), np.float32)
    for i in range(ntr):
        t0 = 0.3 + 0.01 * i
        a = (np.pi * 8.0) ** 2
        obs[i] = (1 - 2 * a * (t - t0) ** 2) * np.exp(-a * (t - t0) ** 2)
        syn[i] = 1.7 * (1 - 2 * a * (t - t0 - 0.006) ** 2) * np.exp(-a * (t - t0 - 0.006) ** 2)
    syn += (0.01 * rng.standard_normal(syn.shape)).astype(np.float32)
    obs[3] = 0.0                       # dead observed trace
    syn[5] = 0.0                       # dead synthetic trace
    return obs, syn


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["Waveform", "Globalcorrelation"])
def test_device_misfit_matches_host(kind):
    import torch
    from swit_amd.engine import misfit
    obs, syn = _gathers()
    dt = 0.002
    J, adj, per = misfit(kind, torch.from_numpy(obs).cuda(), torch.from_numpy(syn).cuda(), dt)
    J2, adj2, _ = misfit(kind, torch.from_numpy(obs).cuda(), torch.from_numpy(syn).cuda(), dt)
    J, J2 = float(J), float(J2)
    adj = adj.cpu().numpy()
    if kind == "Waveform":
        jh = host.misfit_waveform(obs.astype(np.float64), syn.astype(np.float64), dt)
        ah = host.adjoint_source_waveform(obs, syn, dt)
        np.testing.assert_array_equal(adj, ah.astype(np.float32))
    else:
        jh = host.misfit_global_correlation(obs.astype(np.float64), syn.astype(np.float64), dt)
        ah = host.adjoint_source_global_correlation(obs.astype(np.float64), syn.astype(np.float64), dt)
        np.testing.assert_allclose(adj, ah, rtol=0, atol=1e-6 * np.abs(ah).max())
    assert J == J2                                    # fixed-order sum
    assert J == pytest.approx(jh, rel=1e-6)
    assert np.all(adj[3] == 0) and np.all(adj[5] == 0)
    assert float(per[3]) == 0.0 and float(per[5]) == 0.0


def _fft_gathers(nt, ntr=19, seed=1, dt=0.002):
    """Ricker pairs with per-trace shifts of whole samples (traveltime lags known up to
    the noise), a dead observed and a dead synthetic trace."""
    rng = np.random.default_rng(seed)
    t = np.arange(nt) * dt
    a = (np.pi * 9.0) ** 2
    obs = np.zeros((ntr, nt), np.float32)
    syn = np.zeros((ntr, nt), np.float32)
    for i in range(ntr):
        t0 = (0.25 + 0.5 * rng.random()) * nt * dt
        sh = (i % 7 - 3) * dt
        obs[i] = (1 - 2 * a * (t - t0) ** 2) * np.exp(-a * (t - t0) ** 2)
        syn[i] = 1.3 * (1 - 2 * a * (t - t0 - sh) ** 2) * np.exp(-a * (t - t0 - sh) ** 2)
    syn += (0.002 * rng.standard_normal(syn.shape)).astype(np.float32)
    obs[2] = 0.0
    syn[4] = 0.0
    return obs, syn


# 1500 = 4 3 5 5 5, 2048 = 4^5 2, 4000 = 4 4 2 5 5 5 (6000 and 4097 = 17 241 use the
# large-size kernel; 1009 is prime: one direct stage)
FFT_LENGTHS = [7, 1009, 1500, 2048, 4000, 4097, 6000]


@pytest.mark.gpu
@pytest.mark.parametrize("nt", FFT_LENGTHS)
def test_device_envelope_matches_host(nt):
    """Envelope (power 2): misfit and adjoint source against the host restatement.  In
    FP64 (scipy's transforms of float64 traces) to 1e-9 in the misfit and float32
    rounding in the adjoint source; against the reference's own dtype flow (scipy keeps
    float32 traces in complex64) to single-precision transform error (1e-4)."""
    import torch
    from swit_amd.engine import misfit
    dt = 0.002
    obs, syn = _fft_gathers(nt)
    J, adj, per = misfit("Envelope", torch.from_numpy(obs).cuda(), torch.from_numpy(syn).cuda(), dt)
    J2, _, _ = misfit("Envelope", torch.from_numpy(obs).cuda(), torch.from_numpy(syn).cuda(), dt)
    J, J2, adj, per = float(J), float(J2), adj.cpu().numpy(), per.cpu().numpy()
    assert J == J2
    o64, s64 = obs.astype(np.float64), syn.astype(np.float64)
    jh = host.misfit_envelope(o64, s64, float(np.float32(dt)))
    ah = host.adjoint_source_envelope(o64, s64, dt)
    assert J == pytest.approx(jh, rel=1e-9)
    np.testing.assert_allclose(adj, ah, rtol=0, atol=2e-7 * np.abs(ah).max())
    jr = host.misfit_envelope(obs, syn, dt)
    ar = host.adjoint_source_envelope(obs, syn, dt)
    assert J == pytest.approx(jr, rel=1e-4)
    np.testing.assert_allclose(adj, ar, rtol=0, atol=1e-4 * np.abs(ar).max())
    assert np.all(adj[2] == 0) and np.all(adj[4] == 0) and per[2] == 0 and per[4] == 0


@pytest.mark.gpu
def test_device_envelope_of_cosines():
    """Closed form: the discrete Hilbert transform of cos(2 pi k t / n) (0 < k < n/2) is
    sin(2 pi k t / n), so a cosine of amplitude b has the flat envelope b^2 and the
    misfit of amplitude b against amplitude c is n (b^2 - c^2)^2 dt per trace; the
    adjoint source is 2 s r - H(2 H(s) r) = 2 r (s - H(H s)) = 4 r s."""
    import torch
    from swit_amd.engine import misfit
    n, dt = 1500, 0.004
    t = np.arange(n)
    ks = [1, 17, 300, 749]
    syn = np.array([1.5 * np.cos(2 * np.pi * k * t / n) for k in ks], np.float32)
    obs = np.array([0.5 * np.cos(2 * np.pi * k * t / n + 0.3) for k in ks], np.float32)
    J, adj, per = misfit("Envelope", torch.from_numpy(obs).cuda(), torch.from_numpy(syn).cuda(), dt)
    r = 1.5 ** 2 - 0.5 ** 2
    assert float(J) == pytest.approx(len(ks) * n * r * r * float(np.float32(dt)), rel=1e-5)
    np.testing.assert_allclose(adj.cpu().numpy(), 4 * r * syn.astype(np.float64), rtol=0, atol=1e-4)


@pytest.mark.gpu
@pytest.mark.parametrize("nt", FFT_LENGTHS)
def test_device_traveltime_matches_host(nt):
    """Traveltime: the lag of every trace equals the host's (scipy FFT correlation, in
    both dtype flows), the misfit matches to FP64 rounding and the adjoint source to
    float32 output rounding; a trace whose central difference is zero is dead."""
    import torch
    from swit_amd.engine import misfit
    dt = 0.002
    obs, syn = _fft_gathers(nt, seed=nt)
    if nt > 2:
        syn[6] = 0.25                       # constant: central difference all zero
    J, adj, per = misfit("Traveltime", torch.from_numpy(obs).cuda(), torch.from_numpy(syn).cuda(), dt)
    J2, _, _ = misfit("Traveltime", torch.from_numpy(obs).cuda(), torch.from_numpy(syn).cuda(), dt)
    J, J2, adj, per = float(J), float(J2), adj.cpu().numpy(), per.cpu().numpy()
    assert J == J2
    dtf = float(np.float32(dt))
    lag32 = host.cross_correlate_max(obs, syn) - nt + 1
    lag64 = host.cross_correlate_max(obs.astype(np.float64), syn.astype(np.float64)) - nt + 1
    der = host._traveltime_derivative(syn, dt)
    ok = host._live(obs, syn) & (np.sum(np.abs(der), axis=-1) > 0)
    np.testing.assert_array_equal(lag32[ok], lag64[ok])
    np.testing.assert_allclose(per, np.where(ok, 0.5 * lag64.astype(np.float64) ** 2 * dtf, 0.0), rtol=1e-12)
    assert J == pytest.approx(host.misfit_traveltime(obs, syn, dtf), rel=1e-12, abs=0)
    ah = host.adjoint_source_traveltime(obs, syn, dtf)
    np.testing.assert_allclose(adj, ah, rtol=0, atol=1e-6 * max(np.abs(ah).max(), 1e-30))
    assert np.all(adj[2] == 0) and np.all(adj[4] == 0)
    if nt > 2:
        assert np.all(adj[6] == 0) and per[6] == 0
