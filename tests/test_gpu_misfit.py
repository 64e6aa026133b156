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
