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
* ``load_events_from_bag``: the ROS1 bag reader of
  utils/event_camera/event.py:42-69 with the same arguments and result (the
  topic is an extra keyword with the reference's value as default).
"""
from __future__ import annotations

import ctypes
import math
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
        L.gsevt_bag_open.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i, ctypes.c_double, ctypes.POINTER(ll),
                                     ctypes.POINTER(ll)]
        L.gsevt_bag_open.restype = vp
        L.gsevt_bag_events.argtypes = [vp, i, vp, vp, vp, vp, vp]
        L.gsevt_bag_events.restype = i
        L.gsevt_bag_close.argtypes = [vp]
        L.gsevt_bag_close.restype = None
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


BAG_TOPIC = "/prophesee/left/events"   # utils/event_camera/event.py:46


def read_events_bag(path, topic=BAG_TOPIC, start_time=None, event_stamps=False):
    """Events of the EventArray messages of ``topic``: (ts us int64, x, y, p
    int32, events per message int64), messages in bag time order.  ``ts`` is
    the message's header.stamp (the event's own ts with ``event_stamps``) in
    integer microseconds; messages with bag time below ``start_time`` (seconds)
    are left out."""
    L = lib()
    bp, tp = os.fsencode(path), topic.encode()
    has = 0 if start_time is None else 1
    st = 0.0 if start_time is None else float(start_time)
    nm, ne = ctypes.c_longlong(), ctypes.c_longlong()
    h = L.gsevt_bag_open(bp, tp, has, st, ctypes.byref(nm), ctypes.byref(ne))
    if not h:
        _check(-1)
    try:
        m, n = nm.value, ne.value
        counts = np.empty(max(m, 1), np.int64)
        ts = np.empty(max(n, 1), np.int64)
        x = np.empty(ts.shape[0], np.int32)
        y = np.empty(ts.shape[0], np.int32)
        p = np.empty(ts.shape[0], np.int32)
        _check(L.gsevt_bag_events(h, int(bool(event_stamps)), _p(counts), _p(ts), _p(x), _p(y), _p(p)))
    finally:
        L.gsevt_bag_close(h)
    return ts[:n], x[:n], y[:n], p[:n], counts[:m]


def bag_windows(counts, max_events_per_frame, array_nums=None):
    """[start, end) event ranges of the windows load_events_from_bag makes
    from messages of ``counts`` events (utils/event_camera/event.py:54-67):
    a window closes when it holds max_events_per_frame events or more (so
    windows of one event when that is 1 or less) and may span messages; the
    open window at the end is dropped.  Once ``array_nums`` windows exist, the
    reference's ``break`` leaves only the loop over one message's events: the
    rest of that message is dropped, reading goes on with the next message,
    and each later window that closes drops the rest of its message too."""
    thr = max(1, math.ceil(max_events_per_frame))
    out = []
    start = pos = 0           # open window's first event; current message's first event
    for k in np.asarray(counts, np.int64).tolist():
        end = pos + k
        close = start + thr   # exclusive end of the open window once full
        while close <= end:
            out.append((start, close))
            start = close
            if array_nums is not None and len(out) >= array_nums:
                start = end   # rest of this message dropped
                break
            close = start + thr
        pos = end
    return out


def load_events_from_bag(bag_file, max_events_per_frame, array_nums=None, start_time=None, topic=BAG_TOPIC):
    """List of EventArray windows, as utils/event_camera/event.py:42-69.
    Every event carries its message's header.stamp, as there, here in
    integer microseconds so that EventArray.duration()/time() read it as they
    read the events.txt timestamps."""
    from .frame import EventArray
    ts, x, y, p, counts = read_events_bag(bag_file, topic, start_time)
    return [EventArray(ts[s:e], x[s:e], y[s:e], p[s:e])
            for s, e in bag_windows(counts, max_events_per_frame, array_nums)]
