"""Seismic Unix (SU) trace files, byte-compatible with the reference solver's writer.

Layout (fd2dmpi/su.f90, fd2dmpi/io.f90:522-571 write_sufile_2d_coord): every trace is
a 240-byte little-endian header followed by ns float32 samples.  The writer fills the
fields the reference fills and leaves the rest zero (io.f90:687-774 clean_suheader):

    tracl = tracr = trace number (1-based)   fldr = ep = 1
    sx = nint(xs)   gx = nint(xg)   offset = gx - sx   cdp = cdpt = nint(0.5 (xs + xg))
    ns = nt (int16)   dt = int(dt * 1e6) (int16, float32 product)
    d1 = dt   f1 = 0   d2 = 1   f2 = 1   (float32)

The mute-start time (muts, int16 ms at byte 110) is zero in the solver's files; field
data carry first breaks there, which processing can use (toolbox/preprocess.py:134-139).
"""
from __future__ import annotations

import numpy as np

HEADER_BYTES = 240

# byte offsets and types of the SEG-Y/SU trace header fields the solver writes/reads
_FIELDS = {
    "tracl": (0, "<i4"), "tracr": (4, "<i4"), "fldr": (8, "<i4"), "tracf": (12, "<i4"),
    "ep": (16, "<i4"), "cdp": (20, "<i4"), "cdpt": (24, "<i4"), "trid": (28, "<i2"),
    "offset": (36, "<i4"), "sx": (72, "<i4"), "sy": (76, "<i4"), "gx": (80, "<i4"),
    "gy": (84, "<i4"), "muts": (110, "<i2"), "mute": (112, "<i2"), "ns": (114, "<i2"), "dt": (116, "<i2"),
    "d1": (180, "<f4"), "f1": (184, "<f4"), "d2": (188, "<f4"), "f2": (192, "<f4"),
}

_HDR = np.dtype({"names": list(_FIELDS), "formats": [v[1] for v in _FIELDS.values()],
                 "offsets": [v[0] for v in _FIELDS.values()], "itemsize": HEADER_BYTES})


def nint(x):
    """Fortran NINT: nearest integer, halves away from zero."""
    x = np.asarray(x, dtype=np.float64)
    return np.where(x >= 0, np.floor(x + 0.5), np.ceil(x - 0.5)).astype(np.int64)


def headers(nt, dt, xs, xg):
    """Trace headers of one shot gather, as io.f90:522-571 fills them.
    dt, xs, xg are the solver's float32 values (parfile / coord.txt precision)."""
    xg = np.asarray(xg, dtype=np.float32).ravel()
    xs = np.float32(xs)
    n = xg.size
    h = np.zeros(n, dtype=_HDR)
    h["tracl"] = h["tracr"] = np.arange(1, n + 1)
    h["fldr"] = 1
    h["ep"] = 1
    sx = nint(xs)
    gx = nint(xg)
    h["sx"] = sx
    h["gx"] = gx
    h["offset"] = gx - sx
    cdp = nint(np.float32(0.5) * (xs + xg))        # single-precision sum, as in the reference
    h["cdp"] = cdp
    h["cdpt"] = cdp
    h["ns"] = nt
    h["dt"] = int(np.float32(dt) * np.float32(1.0e6))
    h["d1"] = np.float32(dt)
    h["d2"] = 1.0
    h["f1"] = 0.0
    h["f2"] = 1.0
    return h


def write(path, data, dt, xs, xg, first_break_ms=None):
    """Write a shot gather data[ng, nt] (float32) with reference headers.  first_break_ms:
    per-trace first-break times in ms for the mute-start field (muts, int16), which the
    solver leaves zero; processing can centre its window mute on them."""
    data = np.ascontiguousarray(np.asarray(data, dtype="<f4"))
    ng, nt = data.shape
    h = headers(nt, dt, xs, xg)
    if first_break_ms is not None:
        h["muts"] = np.asarray(first_break_ms).reshape(ng)
    write_with_headers(path, data, h)


def write_with_headers(path, data, hdr):
    """Write data[ng, nt] (float32) under the given 240-byte trace headers (as read),
    with the sample count set to nt: every other field is kept, as processing that
    rewrites a gather keeps its traces' headers."""
    data = np.ascontiguousarray(np.asarray(data, dtype="<f4"))
    ng, nt = data.shape
    h = np.array(hdr, dtype=_HDR).reshape(ng)
    h["ns"] = nt
    rec = np.zeros(ng, dtype=np.dtype([("h", "V240"), ("d", "<f4", (nt,))]))
    rec["h"] = h.view("V240")
    rec["d"] = data
    with open(path, "wb") as f:
        rec.tofile(f)


def read(path):
    """Read an SU file -> (data[ng, nt] float32, headers structured array).  ns is taken
    from the first header, as obspy does for fixed-length files."""
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size < HEADER_BYTES:
        raise ValueError(f"{path}: not an SU file")
    ns = int(np.frombuffer(raw[114:116].tobytes(), "<i2")[0]) & 0xFFFF
    rec_bytes = HEADER_BYTES + 4 * ns
    if raw.size % rec_bytes:
        raise ValueError(f"{path}: size {raw.size} is not a multiple of {rec_bytes}")
    ng = raw.size // rec_bytes
    rec = raw.reshape(ng, rec_bytes)
    hdr = np.frombuffer(rec[:, :HEADER_BYTES].copy().tobytes(), dtype=_HDR)
    data = np.frombuffer(rec[:, HEADER_BYTES:].copy().tobytes(), dtype="<f4").reshape(ng, ns)
    return data.astype(np.float32), hdr


def sampling(path):
    """(ng, nt, dt) of an SU file: dt from the microsecond header field."""
    data, hdr = read(path)
    return data.shape[0], data.shape[1], float(hdr["dt"][0]) * 1e-6
