"""Misfit functions and adjoint sources of the reference toolbox (toolbox/misfit.py),
vectorised over the traces of a gather.  Gathers are read from the processed SU files
data/{obs,syn}/src<i>_sg_proc.su as the reference does (misfit.py:43-44,96-100).

Trace rules follow misfit.py line by line: traces whose observed or synthetic L2 norm
is zero contribute nothing; 'Waveform' sums sqrt(sum r^2 dt) per trace (:263-282) with
adjoint source syn - obs (:128-148); 'Globalcorrelation' uses the Pearson coefficient of
the normalised traces (np.corrcoef, :338-355) and the adjoint of :219-242;
'Traveltime' the FFT cross-correlation lag (:359-373, :312-334, :188-215);
'Envelope' power 2 with the Hilbert transform (:286-308, :151-185); 'RTM' adjoint
source = observed data.
"""
from __future__ import annotations

import numpy as np
from scipy import fftpack
from scipy.signal import hilbert

from . import su


def _gathers(homepath, isrc, need_syn=True):
    obs, ho = su.read(homepath + 'data/obs/src%d_sg_proc.su' % (isrc + 1))
    syn = None
    if need_syn:
        syn, hs = su.read(homepath + 'data/syn/src%d_sg_proc.su' % (isrc + 1))
        if obs.shape != syn.shape or ho["dt"][0] != hs["dt"][0]:
            raise ValueError('obs and syn are not consistent.')
    dt = float(ho["dt"][0]) * 1e-6
    return obs, syn, dt


def _live(obs, syn):
    return (np.linalg.norm(obs, axis=-1) > 0) & (np.linalg.norm(syn, axis=-1) > 0)


# -- per-gather misfits -----------------------------------------------------------------
def misfit_waveform(obs, syn, dt):
    r = (syn - obs)[_live(obs, syn)]
    return float(np.sum(np.sqrt(np.sum(r * r * dt, axis=-1))))


def misfit_envelope(obs, syn, dt, p=2.0):
    live = _live(obs, syn)
    se = syn[live] ** 2.0 + hilbert(syn[live], axis=-1).imag ** 2.0
    oe = obs[live] ** 2.0 + hilbert(obs[live], axis=-1).imag ** 2.0
    rsd = se ** (p / 2.0) - oe ** (p / 2.0)
    return float(np.sum(np.linalg.norm(rsd, axis=-1) ** 2) * dt)


def cross_correlate_max(obs, syn):
    """Lag of the cross-correlation maximum in samples (misfit.py:359-373)."""
    nt = obs.shape[-1]
    a = fftpack.fft(obs)
    b = -fftpack.fft(syn).conjugate()
    cc = np.argmax(np.abs(fftpack.ifft(a * b)), -1) - 1
    cc[np.where(cc < nt // 2)] = cc[np.where(cc < nt // 2)] + nt
    return cc


def _traveltime_derivative(syn, dt):
    adj = np.zeros_like(syn, dtype=np.float64)
    adj[:, 1:-1] = (syn[:, 2:] - syn[:, 0:-2]) / (2.0 * dt)
    return adj


def misfit_traveltime(obs, syn, dt):
    nt = obs.shape[-1]
    cc = cross_correlate_max(obs, syn)
    der = _traveltime_derivative(syn, dt)
    ok = _live(obs, syn) & (np.sum(np.abs(der), axis=-1) > 0)
    return float(np.sum(0.5 * np.power(cc[ok] - nt + 1, 2) * dt))


def misfit_global_correlation(obs, syn, dt):
    live = _live(obs, syn)
    o = obs[live] / np.linalg.norm(obs[live], axis=-1, keepdims=True)
    s = syn[live] / np.linalg.norm(syn[live], axis=-1, keepdims=True)
    return float(-np.sum([np.corrcoef(s[i], o[i])[0, 1] for i in range(o.shape[0])]))


# -- per-gather adjoint sources ------------------------------------------------------------
def adjoint_source_waveform(obs, syn, dt):
    adj = np.zeros(obs.shape)
    live = _live(obs, syn)
    adj[live] = syn[live] - obs[live]
    return adj


def adjoint_source_envelope(obs, syn, dt, p=2.0):
    adj = np.zeros(obs.shape)
    live = _live(obs, syn)
    s, o = syn[live], obs[live]
    sh = hilbert(s, axis=-1).imag
    oh = hilbert(o, axis=-1).imag
    se = s ** 2.0 + sh ** 2.0
    oe = o ** 2.0 + oh ** 2.0
    rsd = se ** (p / 2.0) - oe ** (p / 2.0)
    a = p * se ** (p / 2.0 - 1.0) * s * rsd
    dh = p * se ** (p / 2.0 - 1.0) * sh
    a += (-hilbert(dh * rsd, axis=-1)).imag
    adj[live] = a
    return adj


def adjoint_source_traveltime(obs, syn, dt):
    nt = obs.shape[-1]
    cc = cross_correlate_max(obs, syn)
    adj = _traveltime_derivative(syn, dt)
    ok = _live(obs, syn) & (np.sum(np.abs(adj), axis=-1) > 0)
    norm = np.sum(adj * adj, axis=-1) * dt
    scale = np.where(ok, (cc - nt + 1) * dt / np.where(ok, norm, 1.0), 0.0)
    return adj * scale[:, None]


def adjoint_source_global_correlation(obs, syn, dt):
    adj = np.zeros(obs.shape)
    live = np.where(_live(obs, syn))[0]
    for i in live:
        on = np.linalg.norm(obs[i])
        sn = np.linalg.norm(syn[i])
        o, s = obs[i] / on, syn[i] / sn
        adj[i] = 1.0 / sn * (s * np.corrcoef(s, o)[0, 1] - o)
    return adj


_MISFIT = {'waveform': misfit_waveform, 'envelope': misfit_envelope, 'traveltime': misfit_traveltime,
           'globalcorrelation': misfit_global_correlation}
_ADJ = {'waveform': adjoint_source_waveform, 'envelope': adjoint_source_envelope,
        'traveltime': adjoint_source_traveltime, 'globalcorrelation': adjoint_source_global_correlation}


def misfit(simu, misfit_type):
    """Sum of the per-shot misfits (misfit.py:23-36)."""
    kind = misfit_type.lower()
    total = 0.0
    for isrc in range(simu.source.n):
        if kind == 'rtm':
            continue
        obs, syn, dt = _gathers(simu.system.homepath, isrc)
        total += _MISFIT[kind](obs, syn, dt)
    return np.float64(total)


def adjoint_source(simu, misfit_type):
    """Adjoint sources of every shot, float32 array (srcn, recn, nt) (misfit.py:74-88).
    float32 is the precision the solver reads them in (the reference writes them to
    src%d.bin as float32, toolbox/solver.py:112-113); each shot's result is stored as it
    is computed, so the host holds one gather set and one shot's temporaries (at
    801x267x6000 with 81 shots: 1.6 GB, where a list of float64 gathers stacked into
    one array would hold 6.2 GB at its peak)."""
    kind = misfit_type.lower()
    out = None
    for isrc in range(simu.source.n):
        if kind == 'rtm':
            obs, _, _ = _gathers(simu.system.homepath, isrc, need_syn=False)
            a = obs
        else:
            obs, syn, dt = _gathers(simu.system.homepath, isrc)
            a = _ADJ[kind](obs, syn, dt)
        if out is None:
            out = np.empty((simu.source.n,) + a.shape, np.float32)
        out[isrc] = a
    return out
