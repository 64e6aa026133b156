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
