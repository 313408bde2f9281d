"""ctypes binding of libgsevt_hip.so (the C ABI declared in include/gsevt.h).

There is no fallback: if the library is missing or a call fails, an exception
says so.  Every entry point takes raw device pointers, sizes and an explicit
HIP stream.
"""
from __future__ import annotations

import ctypes
import os

import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgsevt_hip.so")

_vp = ctypes.c_void_p
_i = ctypes.c_int
_ll = ctypes.c_longlong
_f = ctypes.c_float
_d = ctypes.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    "gsevt_last_error": (ctypes.c_char_p, []),
    "gsevt_clear_error": (None, []),
    "gsevt_abi_version": (_i, []),
    "gsevt_preprocess": (_i, [_i, _i, _i, _vp, _vp, _f, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _f, _f,
                              _i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "gsevt_duplicate_with_keys": (_i, [_i, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp, _vp, _ll, _vp]),
    "gsevt_tile_ranges": (_i, [_vp, _ll, _vp, _i, _vp, _vp]),
    "gsevt_blend_forward": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gsevt_blend_backward": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gsevt_pose_partials_count": (_i, [_i]),
    "gsevt_preprocess_backward": (_i, [_i, _i, _i, _vp, _vp, _f, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i,
                                       _f, _f, _i, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                       _vp, _vp]),
    "gsevt_pose_finalize": (_i, [_vp, _i, _vp, _vp, _d, _vp, _vp, _vp, _i, _vp]),
    "gsevt_mark_visible": (_i, [_i, _vp, _vp, _vp, _vp, _vp]),
    "gsevt_fill_u32": (_i, [_vp, _ll, ctypes.c_uint32, _vp]),
    "gsevt_scan_scratch_elems": (_ll, [_ll]),
    "gsevt_scan_inclusive_u32": (_i, [_vp, _vp, _ll, _vp, _vp, _vp]),
    "gsevt_sort_scratch_elems": (_ll, [_ll]),
    "gsevt_sort_passes": (_i, [_i]),
    "gsevt_sort_pairs": (_i, [_vp, _vp, _vp, _vp, _vp, _ll, _i, _vp, _vp]),
    "gsevt_event_scatter": (_i, [_i, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp]),
    "gsevt_undistort_map": (_i, [_i, _i, _vp, _vp, _vp, _vp]),
    "gsevt_event_frame_scratch_bytes": (_ll, [_i, _i]),
    "gsevt_event_frame": (_i, [_i, _i, _vp, _vp, _i, _d, _vp, _vp, _vp, _vp]),
    "gsevt_event_frame_noalloc": (_i, [_i, _i, _vp, _vp, _i, _d, _vp, _vp, _vp]),
    "gsevt_resize_nearest": (_i, [_i, _i, _vp, _i, _i, _vp, _vp]),
}

_LIB = None


def load():
    """Load the library (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                           "g.build()'` (or `make -C gs-evt_amd/csrc`); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


class GsevtError(RuntimeError):
    pass


def call(name, *args):
    L = load()
    rc = getattr(L, name)(*args)
    if rc != 0:
        msg = L.gsevt_last_error().decode()
        raise GsevtError(f"{name} failed ({rc}): {msg}")
    return rc


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_of(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
