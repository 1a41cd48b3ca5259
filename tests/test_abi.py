"""The libbmx C ABI on the CPU: the library loads, exports every entry point
include/bmx.h declares, and its struct sizes agree with the ctypes mirror.
No compute calls (there is no GPU here)."""
import ctypes
import os
import re

import pytest

from bmx import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "bmx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bmx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    lib = abi.load_library()
    names = _header_functions()
    assert names, "no declarations parsed from include/bmx.h"
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(abi.EXPORTS) == names


def test_struct_layout_matches_header():
    lib = abi.load_library()
    assert lib.bmx_abi_version() == 4
    assert abi.check_layout(lib)
    assert ctypes.sizeof(abi.EleRec) == lib.bmx_sizeof(0)


def test_ctx_create_without_device_is_an_error_not_a_crash():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    lib = abi.load_library()
    h = ctypes.c_void_p()
    rc = lib.bmx_ctx_create(ctypes.byref(h), 0, 0)
    assert rc < 0 and not h.value
    assert lib.bmx_last_error(None) == b"null context"
