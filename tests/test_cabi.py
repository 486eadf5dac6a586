"""The C ABI (include/dsx.h): libdsx.so loads, exports every declared entry
point, and rejects bad arguments with an error code and message before any
GPU work (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import pytest

from dsx import _lib
from dsx._build import LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dsx.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|int64_t|const char \*)\s*(dsx_[a-z0-9_]+)\s*\(", text)))


def test_library_built():
    assert os.path.exists(LIB_PATH), "run __graft_entry__.build() first"


def test_every_declared_symbol_is_exported():
    names = _declared()
    assert len(names) >= 12
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True)
    exported = set(re.findall(r"\sT\s(dsx_[a-z0-9_]+)", out.stdout))
    missing = [n for n in names if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    assert set(names) == set(_lib.PROTOTYPES), "ctypes prototypes out of sync with include/dsx.h"


def test_version_and_error_paths():
    L = _lib.lib()
    assert L.dsx_version() == 0x000100
    rc = L.dsx_corr_i8(None, None, None, 1, 2, 1, 8, 1, 1, 0, 1, None)
    assert rc == -1 and b"null" in L.dsx_last_error()
    buf = ctypes.create_string_buffer(64)
    p = ctypes.cast(buf, ctypes.c_void_p)
    rc = L.dsx_corr_i8(p, None, p, 4, 2, 1, 3, 1, 1, 0, 1, None)     # odd t_int
    assert rc == -1 and b"t_int" in L.dsx_last_error()
    rc = L.dsx_dedisp_detect(p, 1, 1, 1, 0, 10, 0, p, None, 0, p, 1, 1, p, p, p, p, p, 1, 0, 65, 1, 64, 10, 0, 0,
                             ctypes.c_float(8.0), ctypes.c_double(1.0), p, 1, p, None)
    assert rc == -1 and b"n_centres" in L.dsx_last_error()
    rc = L.dsx_dedisp_detect(p, 1, 1, 1, 0, 10, 0, p, None, 0, p, 1, 1, p, p, p, None, p, 1, 0, 64, 1, 64, 10, 0, 0,
                             ctypes.c_float(8.0), ctypes.c_double(1.0), p, 1, p, None)
    assert rc == -1 and b"null" in L.dsx_last_error()
    # the prime kernel's centres must be held by the ring
    rc = L.dsx_dedisp_prime(p, 0, 1, 1, 10, 20, 0, p, 1, p, p, -64, 64, p, None)
    assert rc == -1 and b"not held" in L.dsx_last_error()
    rc = L.dsx_image_subbands(p, p, p, p, p, 1, 1, 1, 1, 100, 1, 2, ctypes.c_float(1.0), 1, None)
    assert rc == -1 and b"power of two" in L.dsx_last_error()
    rc = L.dsx_ring_advance(p, p, p, 1, 1, 1, 1, 0, 1, 1, None)
    assert rc == -1 and b"differ" in L.dsx_last_error()
    # two window buffers loaded LDS-direct (dma, q_total == r_total on an unpaired plan) need a float32
    # history (ring_dtype 1 = fp16); an unpaired plan with a second buffer but no dma is not a plan
    rc = L.dsx_dedisp_detect(p, 1, 1, 1, 0, 10, 0, p, None, 0, p, 1, 1, p, p, p, p, p, 1, 0, 64, 1, 64, 10, 10, 1,
                             ctypes.c_float(8.0), ctypes.c_double(1.0), p, 1, p, None)
    assert rc == -1 and b"float32 history" in L.dsx_last_error()
    rc = L.dsx_dedisp_detect(p, 0, 1, 1, 0, 10, 0, p, None, 0, p, 1, 1, p, p, p, p, p, 1, 0, 64, 1, 64, 10, 10, 0,
                             ctypes.c_float(8.0), ctypes.c_double(1.0), p, 1, p, None)
    assert rc == -1 and b"describe a plan" in L.dsx_last_error()
    buf2 = ctypes.create_string_buffer(64)
    p2 = ctypes.cast(buf2, ctypes.c_void_p)
    rc = L.dsx_downsample_time(p, p2, 1, 10, 4, 3, 0, None)
    assert rc == -1 and b"must divide" in L.dsx_last_error()
    rc = L.dsx_combine(p, p, p2, 0, 4, 4, 0, 4, None)
    assert rc == -1 and b"bad shape" in L.dsx_last_error()


def test_call_raises_runtime_error():
    with pytest.raises(RuntimeError, match="dsx_boxcar"):
        _lib.call("dsx_boxcar", None, None, 1, 1, 1, None)
