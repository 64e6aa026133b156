"""CPU checks of the C ABI: the library is built, loads, and exports every entry point
include/swit_fd.h declares (no compute call is made)."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    hdr = open(os.path.join(REPO, "include", "swit_fd.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(swit_[a-z_]+)\s*\(", hdr)))


def test_header_declares_the_python_binding_exports():
    from swit_amd import _native
    assert sorted(_native.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol():
    from swit_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libswit_fd.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_native.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_binding_loads_and_reports_no_error():
    from swit_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libswit_fd.so not built")
    L = _native.load()
    assert L.swit_last_error() == b""
    assert L.swit_plan_scratch_bytes(None) == 0
    assert L.swit_plan_destroy(None) == 0


def test_missing_library_fails_loudly(tmp_path):
    from swit_amd import _native
    with pytest.raises(_native.NativeLibraryMissing):
        _native.load(str(tmp_path / "absent.so"))


def test_workspace_queries_validate_without_a_gpu():
    """swit_process_work and swit_grad_precond_work are host-only: they check the
    parameters and size the workspace (the filter's time-major forward output for up
    to 16384 traces, the event norms; the preconditioner's four fields)."""
    from swit_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libswit_fd.so not built")
    L = _native.load()
    nb = ctypes.c_size_t()
    p = _native.SwitProcessParams()
    p.ngathers, p.ntr, p.nt, p.dt = 3, 100, 1000, 0.002
    assert L.swit_process_work(ctypes.byref(p), ctypes.byref(nb)) == 0 and nb.value == 256
    p.filter_taps = 9
    p.a[0] = 1.0
    assert L.swit_process_work(ctypes.byref(p), ctypes.byref(nb)) == 0
    assert nb.value == 8 * (1000 + 6 * 9) * 300
    p.ngathers, p.ntr = 200, 801                      # more traces than one filter chunk
    assert L.swit_process_work(ctypes.byref(p), ctypes.byref(nb)) == 0
    assert nb.value == 8 * (1000 + 6 * 9) * 16384
    bad = []
    for field, value in (("filter_taps", 17), ("filter_taps", 1), ("nt", 27), ("norm_trace", 4),
                         ("norm_event", 3), ("ngathers", 0)):
        q = _native.SwitProcessParams.from_buffer_copy(p)
        setattr(q, field, value)
        bad.append(L.swit_process_work(ctypes.byref(q), ctypes.byref(nb)))
    q = _native.SwitProcessParams.from_buffer_copy(p)
    q.a[0] = 2.0                                      # coefficients must be normalised
    bad.append(L.swit_process_work(ctypes.byref(q), ctypes.byref(nb)))
    q = _native.SwitProcessParams.from_buffer_copy(p)
    q.mute_late_arrival, q.dt = 1, 0.0                # the window mute needs dt
    bad.append(L.swit_process_work(ctypes.byref(q), ctypes.byref(nb)))
    assert bad == [_native.SWIT_EINVAL] * len(bad)
    assert L.swit_last_error() != b""
    pc = _native.SwitPrecond(481, 141, 1, 5, 0.0, 40, 5500.0)
    assert L.swit_grad_precond_work(ctypes.byref(pc), ctypes.byref(nb)) == 0
    assert nb.value >= 8 * 4 * 481 * 141
    pc.nx = 0
    assert L.swit_grad_precond_work(ctypes.byref(pc), ctypes.byref(nb)) == _native.SWIT_EINVAL
