"""Device shot-gather processing (swit_process) against the host restatement of
toolbox/preprocess.py (swit_amd.preprocess, the reference's float32 flow):
- mutes (offset and first-break window, early and late; first breaks picked from the
  data or taken from the SU headers): bit-exact;
- the filter stage alone: bit-exact against scipy.signal.filtfilt of the float32
  gather, stored as float32, for bandpass, lowpass and highpass designs;
- filter + demean + linear detrend + taper: to 2e-6 of the trace maximum (scipy fits
  the line in float32 LAPACK, the device in FP64 rounded to float32);
- normalisations: max bit-exact, L1 / L2 to float32 rounding of the norms;
- the whole workflow over several gathers, and argument errors."""
import ctypes

import numpy as np
import pytest

from swit_amd import base, preprocess


def _optim(**kw):
    o = base.optimize('Waveform', 'LBFGS', 1, 0.05, 6000.0, 1000.0, 'Marine', 0, 0, 'None', 0.0, 0.0, 0, 0.0,
                      ['None'], 0, 0, 0.0, 0.0)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _gathers(ng=3, ntr=23, nt=1200, dt=0.002, seed=0):
    rng = np.random.default_rng(seed)
    t = np.arange(nt) * dt
    d = np.zeros((ng, ntr, nt), np.float32)
    off = np.zeros((ng, ntr))
    for g in range(ng):
        for i in range(ntr):
            off[g, i] = (i - ntr // 2 + g) * 25.0
            t0 = 0.1 + abs(off[g, i]) / 2000.0
            a = (np.pi * 12.0) ** 2
            d[g, i] = (1 - 2 * a * (t - t0) ** 2) * np.exp(-a * (t - t0) ** 2)
            d[g, i] += 0.3 * np.sin(2 * np.pi * 3.0 * t + i) + 0.01 * rng.standard_normal(nt)
    d[1, 4] = 0.0
    return d, off


def _device(optim, d, dt, off, params=None):
    import torch
    from swit_amd import engine
    x = torch.from_numpy(d.copy()).cuda()
    engine.process_gathers(optim, x, dt, off, params=params)
    return x.cpu().numpy()


def _host(optim, d, dt, off):
    out = []
    for g in range(d.shape[0]):
        x = preprocess.apply_mute(optim, d[g], dt, off[g])
        x = preprocess.apply_filter(optim, x, dt)
        out.append(preprocess.apply_normalize(optim, x))
    return np.array(out)


@pytest.mark.gpu
@pytest.mark.parametrize("window", [0.08, -0.08])
def test_mutes_bit_exact(window):
    dt = 0.002
    d, off = _gathers()
    o = _optim(mute_offset_short=1, mute_offset_short_dis=60.0, mute_offset_long=1, mute_offset_long_dis=240.0,
               mute_late_arrival=1, mute_late_window=window)
    got = _device(o, d, dt, off)
    want = _host(o, d, dt, off)
    assert np.any(got == 0) and np.any(got != 0)
    np.testing.assert_array_equal(got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("window", [0.08, -0.05])
def test_header_first_break_mute_bit_exact(window):
    """First breaks from the SU headers (per-trace ms): the device centres the window
    mute on int((t / 1000 + |window|) / dt) as the host restatement does."""
    import torch
    from swit_amd import engine
    dt = 0.002
    d, off = _gathers(seed=4)
    fb = np.random.default_rng(7).integers(0, 1500, size=d.shape[:2])
    o = _optim(mute_late_arrival=1, mute_late_window=window, mute_offset_short=1, mute_offset_short_dis=30.0)
    x = torch.from_numpy(d.copy()).cuda()
    engine.process_gathers(o, x, dt, off, first_break_ms=fb)
    got = x.cpu().numpy()
    want = np.array([preprocess.apply_mute(o, d[g], dt, off[g], fb[g]) for g in range(d.shape[0])])
    np.testing.assert_array_equal(got, want)
    assert not np.array_equal(got, _device(o, d, dt, off))      # the data picker gives other windows


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["bandpass", "lowpass", "highpass"])
def test_filter_stage_bit_exact(kind):
    from scipy.signal import filtfilt
    from swit_amd import engine
    dt = 0.002
    d, off = _gathers(nt=1000)
    o = _optim(fre_filter=kind, fre_low=4.0 if kind == 'bandpass' else 15.0, fre_high=18.0 if kind == 'bandpass'
               else 6.0)
    p, win, taper = engine.process_params(o, *d.shape, dt)
    p.detrend_taper = 0
    got = _device(o, d, dt, off, params=(p, win, None))
    b, a = preprocess.filter_design(o, dt)
    want = filtfilt(b, a, d).astype(np.float32)
    np.testing.assert_array_equal(got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["bandpass", "lowpass", "highpass"])
def test_filter_detrend_taper(kind):
    dt = 0.002
    d, off = _gathers(nt=1500, seed=1)
    o = _optim(fre_filter=kind, fre_low=3.0 if kind == 'bandpass' else 20.0, fre_high=25.0 if kind == 'bandpass'
               else 5.0)
    got = _device(o, d, dt, off)
    want = _host(o, d, dt, off)
    scale = np.abs(want).max(axis=-1, keepdims=True) + 1e-30
    assert np.abs(got - want).max() / scale.max() < 2e-6
    assert np.all(np.abs(got - want) <= 2e-6 * scale)


@pytest.mark.gpu
@pytest.mark.parametrize("norm", [['Max-Trace'], ['L1-Trace'], ['L2-Trace'], ['L1-Event'], ['L2-Event'],
                                  ['L2-Trace', 'L1-Event'], ['Max-Trace', 'L2-Event']])
def test_normalisations(norm):
    dt = 0.002
    d, off = _gathers(seed=2)
    o = _optim(normalize=norm)
    got = _device(o, d, dt, off)
    want = _host(o, d, dt, off)
    if norm == ['Max-Trace']:
        np.testing.assert_array_equal(got, want)
    else:
        np.testing.assert_allclose(got, want, rtol=2e-6, atol=0)


@pytest.mark.gpu
def test_whole_workflow():
    dt = 0.0015
    d, off = _gathers(ng=4, ntr=31, nt=2001, dt=dt, seed=3)
    o = _optim(mute_offset_short=1, mute_offset_short_dis=30.0, mute_late_arrival=1, mute_late_window=0.15,
               fre_filter='bandpass', fre_low=2.0, fre_high=30.0, normalize=['L2-Trace', 'L2-Event'])
    got = _device(o, d, dt, off)
    want = _host(o, d, dt, off)
    np.testing.assert_allclose(got, want, rtol=0, atol=5e-6 * np.abs(want).max())


@pytest.mark.gpu
def test_argument_errors():
    import torch
    from swit_amd import _native, engine
    o = _optim(fre_filter='lowpass', fre_low=10.0)
    x = torch.zeros((2, 20), dtype=torch.float32, device="cuda")        # nt = 20 <= 3 * 5 taps? 20 > 15: ok
    engine.process_gathers(o, x, 0.002)
    x = torch.zeros((2, 15), dtype=torch.float32, device="cuda")
    with pytest.raises(_native.SwitError):
        engine.process_gathers(o, x, 0.002)
    o = _optim(mute_offset_short=1, mute_offset_short_dis=10.0)
    with pytest.raises(ValueError):
        engine.process_gathers(o, torch.zeros((2, 50), dtype=torch.float32, device="cuda"), 0.002)
    p = _native.SwitProcessParams()
    p.ngathers, p.ntr, p.nt, p.norm_trace = 1, 1, 10, 7
    nb = ctypes.c_size_t()
    assert _native.load().swit_process_work(ctypes.byref(p), ctypes.byref(nb)) == _native.SWIT_EINVAL
