"""On-disk inputs of the tracker, read by libgsevt_io.so (C++, OpenMP).

* ``load_ply``: the 3DGS PLY map as GaussianModel.load_ply builds it
  (gaussian_splatting/scene/gaussian_model.py:238-327): raw pre-activation
  values; f_rest_* in index order reshaped (P, 3, K-1) and transposed to
  (P, K-1, 3); the reference's assertion on the number of f_rest_*
  properties.
* ``load_events_from_txt``: the events.txt reader of
  utils/event_camera/event.py:11-39 with the same arguments and result:
  "ts x y p" lines, events before ``start_time`` skipped, windows of exactly
  ``max_events_per_frame`` events, leftovers dropped, at most ``array_nums``
  windows.  A line that is not four integers raises when the reference's loop
  would reach it: ValueError for a field that is not an integer, IndexError
  for fewer than four fields (as ``int()`` and ``line_data[k]`` raise there).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_LIB = None


class GsevtIOError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgsevt_io.so")
        if not os.path.exists(path):
            raise GsevtIOError(f"{path} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        vp, ll, i = ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int
        L.gsevt_io_last_error.restype = ctypes.c_char_p
        L.gsevt_ply_info.argtypes = [ctypes.c_char_p, ctypes.POINTER(ll), ctypes.POINTER(i), ctypes.POINTER(i),
                                     ctypes.POINTER(i)]
        L.gsevt_ply_info.restype = i
        L.gsevt_ply_read.argtypes = [ctypes.c_char_p] + [vp] * 7
        L.gsevt_ply_read.restype = i
        L.gsevt_events_txt_lines.argtypes = [ctypes.c_char_p]
        L.gsevt_events_txt_lines.restype = ll
        L.gsevt_events_txt_read.argtypes = [ctypes.c_char_p, ll, vp, vp, vp, vp, ctypes.POINTER(ll),
                                            ctypes.POINTER(ll), ctypes.POINTER(i)]
        L.gsevt_events_txt_read.restype = i
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise GsevtIOError(lib().gsevt_io_last_error().decode(errors="replace"))


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def read_ply(path):
    """Vertex properties of a 3DGS PLY file as float32 arrays: xyz [P,3],
    normals [P,3], f_dc [P,3], f_rest [P,n_rest] (index order), opacity
    [P,1], scales [P,n_scale], rots [P,n_rot] (index order)."""
    L = lib()
    bp = os.fsencode(path)
    n = ctypes.c_longlong()
    nr, ns, nq = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(L.gsevt_ply_info(bp, ctypes.byref(n), ctypes.byref(nr), ctypes.byref(ns), ctypes.byref(nq)))
    P = n.value
    out = dict(xyz=np.empty((P, 3), np.float32), normals=np.zeros((P, 3), np.float32),
               f_dc=np.empty((P, 3), np.float32), f_rest=np.empty((P, nr.value), np.float32),
               opacity=np.empty((P, 1), np.float32), scales=np.empty((P, ns.value), np.float32),
               rots=np.empty((P, nq.value), np.float32))
    _check(L.gsevt_ply_read(bp, _p(out["xyz"]), _p(out["normals"]), _p(out["f_dc"]), _p(out["f_rest"]),
                            _p(out["opacity"]), _p(out["scales"]), _p(out["rots"])))
    return out


def ply_to_raw(r, max_sh_degree=3):
    """The raw-parameter layout of gaussian_model.py:298-323 from read_ply's
    arrays (features_dc [P,1,3], features_rest [P,K-1,3])."""
    from .synthetic import RawScene
    K = (max_sh_degree + 1) ** 2
    P = r["xyz"].shape[0]
    if r["f_rest"].shape[1] != 3 * K - 3:
        raise AssertionError(f"{r['f_rest'].shape[1]} f_rest_* properties, {3 * K - 3} expected for SH degree "
                             f"{max_sh_degree}")
    fdc = r["f_dc"].reshape(P, 3, 1).transpose(0, 2, 1)
    frest = r["f_rest"].reshape(P, 3, K - 1).transpose(0, 2, 1)
    return RawScene(xyz=r["xyz"], features_dc=np.ascontiguousarray(fdc), features_rest=np.ascontiguousarray(frest),
                    scaling=r["scales"], rotation=r["rots"], opacity=r["opacity"], sh_degree=max_sh_degree)


def load_ply(path, max_sh_degree=3, device="cuda"):
    """GaussianModel of a 3DGS PLY file (active SH degree = max_sh_degree)."""
    from .model import GaussianModel
    raw = ply_to_raw(read_ply(path), max_sh_degree)
    return GaussianModel.from_raw(raw.xyz, raw.features_dc, raw.features_rest, raw.scaling, raw.rotation,
                                  raw.opacity, raw.sh_degree, device=device)


def read_events_txt(path):
    """All lines of an events.txt file: (ts int64, x, y, p int32, index of the
    first malformed line or -1, its kind: 1 not an integer, 2 fewer than four
    fields, 3 outside int32)."""
    L = lib()
    bp = os.fsencode(path)
    n = L.gsevt_events_txt_lines(bp)
    if n < 0:
        _check(-1)
    cap = max(int(n), 1)
    ts = np.empty(cap, np.int64)
    x = np.empty(cap, np.int32)
    y = np.empty(cap, np.int32)
    p = np.empty(cap, np.int32)
    got, bad, kind = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_int()
    _check(L.gsevt_events_txt_read(bp, cap, _p(ts), _p(x), _p(y), _p(p), ctypes.byref(got), ctypes.byref(bad),
                                   ctypes.byref(kind)))
    m = got.value
    return ts[:m], x[:m], y[:m], p[:m], int(bad.value), int(kind.value)


def load_events_from_txt(data_path, max_events_per_frame, array_nums=None, start_time=None):
    """List of EventArray windows, as utils/event_camera/event.py:11-39."""
    from .frame import EventArray
    ts, x, y, p, bad, kind = read_events_txt(data_path)
    M = int(max_events_per_frame)
    keep = np.arange(ts.shape[0]) if start_time is None else np.nonzero(ts >= start_time)[0]
    n_win = keep.shape[0] // M if M > 0 else 0
    if array_nums is not None:
        n_win = min(n_win, int(array_nums))
    # the reference parses line by line and stops after the last full window
    # when array_nums is reached; a malformed line before that point raises
    if array_nums is not None and n_win == int(array_nums) and n_win > 0:
        stop_line = int(keep[n_win * M - 1])
    else:
        stop_line = ts.shape[0] - 1
    if bad >= 0 and bad <= stop_line:
        if kind == 2:   # line_data[k] of a short line
            raise IndexError(f"{data_path}: line {bad + 1} has fewer than 4 fields")
        if kind == 3:
            raise OverflowError(f"{data_path}: line {bad + 1}: x, y or p outside int32")
        raise ValueError(f"{data_path}: line {bad + 1}: a field is not an integer")
    sel = keep[:n_win * M]
    return [EventArray(ts[sel[w * M:(w + 1) * M]], x[sel[w * M:(w + 1) * M]], y[sel[w * M:(w + 1) * M]],
                       p[sel[w * M:(w + 1) * M]]) for w in range(n_win)]
