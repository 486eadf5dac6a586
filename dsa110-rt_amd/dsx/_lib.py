"""ctypes binding of ``libdsx.so`` (the C ABI declared in ``include/dsx.h``).

The fast path has no CPU fallback: if the library is missing or no GPU is
visible, every stage call raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import torch

from ._build import LIB_PATH

DSX_F32, DSX_F16, DSX_I8, DSX_I32 = 0, 1, 2, 3

_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_F = ctypes.c_float
_D = ctypes.c_double

#: name -> argtypes, exactly the prototypes of include/dsx.h
PROTOTYPES = {
    "dsx_version": [],
    "dsx_last_error": [],
    "dsx_unpack_i4": [_P, _P, _P, _I, _I, _I, _I, _F, _P],
    "dsx_corr_i8": [_P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _P],
    "dsx_corr_i8_stage1": [_P, _P, _P, _I, _P, _I, _I, _I, _I, _I, _I, _I, _P],
    "dsx_static_sky": [_P, _I, _I, _I, _I64, _P, _P, _I, _I, _I, _I, _P],
    "dsx_corr_cal": [_P, _P, _P, _P, _P, _I, _P, _I, _I, _I, _I, _I, _I, _I, _I, _F, _P],
    "dsx_apply_cal": [_P, _P, _P, _P, _I, _P, _P, _I, _I, _I, _P],
    "dsx_grid_gather": [_P, _P, _P, _P, _I, _I, _I, _I, _P],
    "dsx_image_subbands": [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _F, _I, _P],
    "dsx_image_subbands_f32": [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _F, _I, _P],
    "dsx_dirty_image": [_P, _P, _I, _I, _F, _P],
    "dsx_image_cube": [_P, _P, _I, _I, _I, _I, _I, _I, _F, _P, _P, _P, _I, _P, _I, _P, _I, _P, _P],
    "dsx_ring_advance": [_P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _P],
    "dsx_downsample_time": [_P, _P, _I, _I, _I64, _I, _I, _P],
    "dsx_dedisp_plan": [_P, _I, _I, _I, _I, _I, _I, _I, _P, _P, _I, _P, _P],
    "dsx_dedisp_detect": [_P, _I, _I, _I, _I, _I, _I64, _P, _P, _I, _P, _I, _I, _P, _P, _P, _P, _P,
                          _I, _I64, _I, _I, _I, _I, _I, _I, _F, _D, _P, _I, _P, _P],
    "dsx_dedisp_prime": [_P, _I, _I, _I, _I, _I, _I64, _P, _I, _P, _P, _I64, _I, _P, _P],
    "dsx_stats_decay": [_P, _I64, _P, _I, _D, _P],
    "dsx_combine": [_P, _P, _P, _I, _I, _I64, _I, _I, _P],
    "dsx_cube_cumsum": [_P, _I, _P, _I, _I64, _I, _P],
    "dsx_cube_candidates": [_P, _I, _I64, _I, _P, _P, _I, _F, _I, _I, _I, _P, _P, _P, _P, _I64, _P],
    "dsx_cube_box_samples": [_P, _I, _I64, _I, _P, _I, _I, _I, _P, _I, _P, _P],
    "dsx_cube_candidates_noalloc": [_P, _I, _I64, _I, _P, _P, _I, _F, _I, _I, _I, _P, _P, _P, _P],
    "dsx_cube_candidates_noalloc_capacity": [],
    "dsx_synth_block": [_P, _U64, _I64, _I, _I, _I, _P, _I, _P, _P, _P, _P],
    "dsx_boxcar": [_P, _P, _I, _I, _I, _P],
    "dsx_multi_boxcar_argmax": [_P, _I, _P, _P, _P, _I, _P, _I, _P, _I, _I, _I, _I, _P],
}

_lib: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    """Load ``libdsx.so`` once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it first (python -c "
                "'import __graft_entry__ as g; g.build()')")
        handle = ctypes.CDLL(LIB_PATH)
        for name, argtypes in PROTOTYPES.items():
            fn = getattr(handle, name)
            fn.argtypes = argtypes
            fn.restype = (ctypes.c_char_p if name == "dsx_last_error" else
                          ctypes.c_int64 if name.endswith("_capacity") else ctypes.c_int)
        _lib = handle
    return _lib


def call(name: str, *args) -> int:
    """Invoke ``name``; raise ``RuntimeError`` with ``dsx_last_error()`` on failure."""
    fn = getattr(lib(), name)
    rc = fn(*args)
    if rc < 0:
        msg = lib().dsx_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed ({rc}): {msg}")
    return rc


def version() -> int:
    return lib().dsx_version()


def require_gpu(*tensors: torch.Tensor) -> None:
    """The fast path runs on the GPU only (no CPU fallback)."""
    if not torch.cuda.is_available():
        raise RuntimeError("dsx needs a ROCm GPU; no device is visible")
    for t in tensors:
        if t is None:
            continue
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
            raise ValueError(f"expected a CUDA (ROCm) tensor, got {type(t).__name__} on "
                             f"{getattr(t, 'device', '?')}")
        if not t.is_contiguous():
            raise ValueError("tensors passed to dsx must be contiguous")


def ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream(s: Optional[torch.cuda.Stream] = None):
    s = torch.cuda.current_stream() if s is None else s
    return ctypes.c_void_p(s.cuda_stream)
