"""Shot-gather processing between the forward run and the misfit (toolbox/preprocess.py):
data/<type>/src<i>_sg.su -> data/<type>/src<i>_sg_proc.su with offset/arrival mutes,
Butterworth filtering (scipy butter/filtfilt, order 4) and normalisation, applied in
the reference's order (mute, filter, normalize; preprocess.py:47-61).

With no filter, mute or normalisation (the benchmark configurations) the processed
file is a byte copy of the solver's SU file.  The reference's detrend/taper after
filtering and first-break picking come from obspy, which is not available here: they
are restated with scipy (demean, linear detrend, 5 % Hann taper) and their parity is
unpinned.
"""
from __future__ import annotations

import shutil

import numpy as np
from scipy.signal import butter, detrend, filtfilt

from . import su


def butter_bandpass_filter(data, lowcut, highcut, dt, order=4):
    nyq = 0.5 / dt
    b, a = butter(order, [lowcut / nyq, highcut / nyq], btype='band')
    return filtfilt(b, a, data)


def butter_lowpass_filter(data, highcut, dt, order=4):
    nyq = 0.5 / dt
    b, a = butter(order, highcut / nyq, btype='low')
    return filtfilt(b, a, data)


def butter_highpass_filter(data, lowcut, dt, order=4):
    nyq = 0.5 / dt
    b, a = butter(order, lowcut / nyq, btype='hp')
    return filtfilt(b, a, data)


def _hann_taper(n, pct=0.05):
    w = np.ones(n)
    k = int(pct * n)
    if k > 0:
        ramp = 0.5 * (1 - np.cos(np.pi * np.arange(k) / k))
        w[:k] = ramp
        w[n - k:] = ramp[::-1]
    return w


def apply_filter(optim, data, dt):
    """preprocess.py:64-95."""
    kind = optim.fre_filter.lower()
    if kind == 'none':
        return data
    if kind == 'bandpass':
        out = butter_bandpass_filter(data, optim.fre_low, optim.fre_high, dt)
    elif kind == 'lowpass':
        out = butter_lowpass_filter(data, optim.fre_low, dt)
    elif kind == 'highpass':
        out = butter_highpass_filter(data, optim.fre_high, dt)
    else:
        raise ValueError(optim.fre_filter)
    out = detrend(out - out.mean(axis=-1, keepdims=True), axis=-1, type='linear')
    return out * _hann_taper(out.shape[-1])


def mask(itmin, itmax, nt, length):
    """Tapered mute mask (preprocess.py:238-262)."""
    m = np.ones(nt)
    win = np.sin(np.linspace(0, np.pi, 2 * length))[0:length]
    if 1 < itmin < itmax < nt:
        m[0:itmin] = 0.0
        m[itmin:itmax] = win * m[itmin:itmax]
    elif itmin < 1 <= itmax:
        m[0:itmax] = win[length - itmax:length] * m[0:itmax]
    elif itmin < nt < itmax:
        m[0:itmin] = 0.0
        m[itmin:nt] = win[0:nt - itmin] * m[itmin:nt]
    elif itmin > nt:
        m[:] = 0.0
    return m


def brutal_picker(data):
    """First sample above 0.1 % of the trace maximum (preprocess.py:203-209)."""
    thr = 0.001 * np.max(np.abs(data), axis=-1)
    return np.array([(np.abs(data[i]) > thr[i]).argmax() for i in range(data.shape[0])])


def apply_mute(optim, data, dt, offsets):
    """Offset mutes and first-break window mute (preprocess.py:98-151)."""
    data = data.copy()
    length = 100
    nt = data.shape[-1]
    off = np.abs(offsets)
    if optim.mute_offset_short:
        data[off < optim.mute_offset_short_dis] = 0.0
    if optim.mute_offset_long:
        data[off > optim.mute_offset_long_dis] = 0.0
    if optim.mute_late_arrival:
        mutetype = 'late' if optim.mute_late_window > 0.0 else 'early'
        pick = brutal_picker(data) + np.ceil(abs(optim.mute_late_window) / dt)
        for i in range(data.shape[0]):
            itmin = int(pick[i] - length / 2)
            itmax = int(itmin + length)
            mk = mask(itmin, itmax, nt, length)
            data[i] *= mk if mutetype == 'early' else (1.0 - mk)
    return data


def apply_normalize(optim, data):
    """preprocess.py:154-200."""
    data = data.copy()
    nz = optim.normalize
    if 'L1-Trace' in nz:
        w = np.linalg.norm(data, ord=1, axis=-1)
    elif 'L2-Trace' in nz:
        w = np.linalg.norm(data, ord=2, axis=-1)
    elif 'Max-Trace' in nz:
        w = np.max(np.abs(data), axis=-1)
    else:
        w = None
    if w is not None:
        data[w > 0] /= w[w > 0, None]
    if 'L1-Event' in nz:
        data /= np.sum(np.linalg.norm(data, ord=1, axis=-1))
    elif 'L2-Event' in nz:
        data /= np.sum(np.linalg.norm(data, ord=2, axis=-1))
    return data


def _identity(optim):
    return (optim.fre_filter.lower() == 'none' and not optim.mute_late_arrival and
            not optim.mute_offset_short and not optim.mute_offset_long and
            not any(k in optim.normalize for k in ('L1-Trace', 'L2-Trace', 'Max-Trace', 'L1-Event', 'L2-Event')))


def process_workflow(simu, optim, simu_type='syn', use_first_break_in_su_header=False):
    """Process every shot gather of data/<simu_type>/ (preprocess.py:23-44)."""
    if use_first_break_in_su_header:
        raise NotImplementedError('first breaks from SU headers are not supported')
    datapath = simu.system.homepath + 'data/' + simu_type + '/'
    for isrc in range(simu.source.n):
        load = datapath + 'src%d_sg.su' % (isrc + 1)
        save = datapath + 'src%d_sg_proc.su' % (isrc + 1)
        if _identity(optim):
            shutil.copyfile(load, save)
            continue
        data, hdr = su.read(load)
        dt = float(hdr["dt"][0]) * 1e-6
        data = data.astype(np.float64)
        data = apply_mute(optim, data, dt, hdr["offset"].astype(np.float64))
        data = apply_filter(optim, data, dt)
        data = apply_normalize(optim, data)
        su.write(save, data.astype(np.float32), np.float32(hdr["d1"][0]), hdr["sx"][0], hdr["gx"])
