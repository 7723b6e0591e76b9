"""The C ABI library (CPU only): it loads, exports every function declared in include/*.h,
and its request struct matches the ctypes mirror.  No device call is made."""
import ctypes
import glob
import os
import re

import pytest

from thewalrus_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(wh_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_built():
    assert os.path.exists(_native.LIB_PATH), "run python -m thewalrus_amd._build"


def test_exports_every_declared_symbol():
    L = _native.load()
    names = declared_functions()
    assert names >= set(_native.EXPORTS)
    for name in sorted(names):
        assert hasattr(L, name), name


def test_request_layout():
    """wh_request: 3 int64 offsets, 2 doubles, 4 int32, 32 int32 multiplicities = 184 bytes."""
    R = _native.WhRequest
    assert ctypes.sizeof(R) == 184
    assert R.oddloop.offset == 24 and R.E.offset == 40 and R.edge_reps.offset == 56


def test_error_string_without_device():
    L = _native.load()
    assert isinstance(L.wh_last_error(), bytes)
    assert L.wh_version() >= 100


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError):
        _native.load(str(tmp_path / "nope.so"))
