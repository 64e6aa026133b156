"""Shot-gather processing between the forward run and the misfit (toolbox/preprocess.py):
data/<type>/src<i>_sg.su -> data/<type>/src<i>_sg_proc.su with offset/arrival mutes,
Butterworth filtering (scipy butter/filtfilt, order 4) and normalisation, applied in
the reference's order (mute, filter, normalize; preprocess.py:47-61).

The reference works on obspy traces, whose data stay float32 throughout: the mute
multiplies in place, the filtered float64 result is stored back as float32, and the
detrend ('demean', then 'linear', both scipy.signal.detrend on float32 data) and the
5 % Hann taper of obspy's Trace.taper follow the filter (preprocess.py:64-95).  This
restatement keeps that float32 flow, trace by trace, with the same scipy calls.
obspy itself is not available here: its detrend and taper are restated from their
documented behaviour and their parity is unpinned.

With no filter, mute or normalisation (the benchmark configurations) the processed
file is a byte copy of the solver's SU file.

`process_gathers_device` runs the same steps on device gathers through swit_process.
"""
from __future__ import annotations

import shutil

import numpy as np
from scipy.signal import butter, detrend, filtfilt
from scipy.signal.windows import hann

from . import su

MUTE_LENGTH = 100        # samples of the arrival-mute taper (preprocess.py:101)


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


def filter_design(optim, dt, order=4):
    """(b, a) of the Butterworth filter apply_filter uses, or None for 'none'
    (preprocess.py:75-81, :273-296): 'bandpass' fre_low..fre_high, 'lowpass' at
    fre_low, 'highpass' at fre_high."""
    kind = optim.fre_filter.lower()
    nyq = 0.5 / dt
    if kind == 'none':
        return None
    if kind == 'bandpass':
        return butter(order, [optim.fre_low / nyq, optim.fre_high / nyq], btype='band')
    if kind == 'lowpass':
        return butter(order, optim.fre_low / nyq, btype='low')
    if kind == 'highpass':
        return butter(order, optim.fre_high / nyq, btype='hp')
    raise ValueError(optim.fre_filter)


def obspy_taper(npts, max_percentage=0.05):
    """Weights of obspy's Trace.taper(max_percentage, type='hann'): half-length
    min(int(p npts), int(npts / 2)), sides from scipy's symmetric Hann window of
    2 half-length + 1 points (2 half-length when that equals npts), ones between."""
    wlen = min(int(max_percentage * npts), int(npts / 2))
    sides = hann(2 * wlen) if 2 * wlen == npts else hann(2 * wlen + 1)
    return np.hstack((sides[:wlen], np.ones(npts - 2 * wlen), sides[len(sides) - wlen:]))


def apply_filter(optim, data, dt):
    """preprocess.py:64-95 on a float32 gather: filtfilt (float64) stored back as
    float32, then per trace demean, linear detrend and the 5 % Hann taper."""
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
    out = out.astype(np.float32)
    taper = obspy_taper(out.shape[-1])
    for i in range(out.shape[0]):
        tr = detrend(out[i], type='constant')
        tr = detrend(tr, type='linear')
        tr *= taper
        out[i] = tr
    return out


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


def header_picks(first_break_ms, mute_late_window, dt):
    """Window-mute centres (samples) from the SU headers' first-break times in ms:
    int((t_ms / 1000 + |mute_late_window|) / dt) per trace (preprocess.py:134-139)."""
    t = np.asarray(first_break_ms, dtype=np.float64)
    return np.trunc((t / 1000.0 + abs(mute_late_window)) / dt).astype(np.int64)


def apply_mute(optim, data, dt, offsets, first_break_ms=None):
    """Offset mutes, then the first-break window mute (preprocess.py:98-151), in place
    on a copy of the float32 gather as the reference's in-place trace products.
    first_break_ms: per-trace first breaks from the SU headers (the reference's
    use_first_break_in_su_header), or None to pick them from the data."""
    data = data.copy()
    length = MUTE_LENGTH
    nt = data.shape[-1]
    off = np.abs(offsets)
    if optim.mute_offset_short:
        data[off < optim.mute_offset_short_dis] = 0.0
    if optim.mute_offset_long:
        data[off > optim.mute_offset_long_dis] = 0.0
    if optim.mute_late_arrival:
        mutetype = 'late' if optim.mute_late_window > 0.0 else 'early'
        if first_break_ms is not None:
            pick = header_picks(first_break_ms, optim.mute_late_window, dt)
        else:
            pick = brutal_picker(data) + np.ceil(abs(optim.mute_late_window) / dt)
        for i in range(data.shape[0]):
            itmin = int(pick[i] - length / 2)
            itmax = int(itmin + length)
            mk = mask(itmin, itmax, nt, length)
            data[i] *= mk if mutetype == 'early' else (1.0 - mk)
    return data


def apply_normalize(optim, data):
    """preprocess.py:154-200 on a float32 gather, trace by trace as the reference (its
    norms are float32 reductions; the event sum starts from 0. and adds them in trace
    order)."""
    data = data.copy()
    nz = optim.normalize
    if 'L1-Trace' in nz or 'L2-Trace' in nz or 'Max-Trace' in nz:
        for i in range(data.shape[0]):
            if 'L1-Trace' in nz:
                w = np.linalg.norm(data[i], ord=1)
            elif 'L2-Trace' in nz:
                w = np.linalg.norm(data[i], ord=2)
            else:
                w = np.max(abs(data[i]))
            if w > 0:
                data[i] /= w
    if 'L1-Event' in nz or 'L2-Event' in nz:
        w = 0.
        for i in range(data.shape[0]):
            w += np.linalg.norm(data[i], ord=1 if 'L1-Event' in nz else 2)
        for i in range(data.shape[0]):
            data[i] /= w
    return data


def _identity(optim):
    return (optim.fre_filter.lower() == 'none' and not optim.mute_late_arrival and
            not optim.mute_offset_short and not optim.mute_offset_long and
            not any(k in optim.normalize for k in ('L1-Trace', 'L2-Trace', 'Max-Trace', 'L1-Event', 'L2-Event')))


def process_workflow(simu, optim, simu_type='syn', use_first_break_in_su_header=False):
    """Process every shot gather of data/<simu_type>/ (preprocess.py:23-44).  With
    use_first_break_in_su_header the window mute is centred on each trace's first break
    from its SU header (mute-start time, ms) instead of a pick from the data.  The
    processed file keeps the input traces' headers."""
    datapath = simu.system.homepath + 'data/' + simu_type + '/'
    for isrc in range(simu.source.n):
        load = datapath + 'src%d_sg.su' % (isrc + 1)
        save = datapath + 'src%d_sg_proc.su' % (isrc + 1)
        if _identity(optim):
            shutil.copyfile(load, save)
            continue
        data, hdr = su.read(load)
        dt = float(hdr["dt"][0]) * 1e-6
        fb = hdr["muts"].astype(np.float64) if use_first_break_in_su_header else None
        data = apply_mute(optim, data.astype(np.float32), dt, hdr["offset"].astype(np.float64), fb)
        data = apply_filter(optim, data, dt)
        data = apply_normalize(optim, data)
        su.write_with_headers(save, data.astype(np.float32), hdr)


def process_gathers_device(optim, data, dt, offsets=None, first_break_ms=None):
    """apply_mute, apply_filter and apply_normalize on device gathers (float32 device
    tensor [ngathers][ntr][nt] or [ntr][nt], processed in place) through swit_process;
    offsets: group x - source x per trace; first_break_ms: per-trace first breaks from
    the SU headers, or None to pick them from the data."""
    from . import engine
    return engine.process_gathers(optim, data, dt, offsets, first_break_ms=first_break_ms)
