"""On-disk readers (libgsevt_io.so) against independent parses.

* PLY: files written by a numpy structured-array writer in the 3DGS attribute
  order of gaussian_model.py:173-218, in ascii and both binary byte orders,
  with shuffled property order, mixed scalar types and an element with list
  properties before the vertex element; the reader's arrays equal the
  written values rounded to float32 and the raw layout equals the scene the
  file was written from.  Missing properties, truncation and a wrong SH
  degree raise.
* events.txt: windows equal those of a line-by-line restatement of the
  reference loop (utils/event_camera/event.py:11-39) for start_time /
  array_nums combinations; a 1M-line file parses equal to pandas; malformed
  lines raise only when the reference's loop reaches them.
"""
import os
import pathlib
import sys
import time

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gs-evt_amd"))

from gsevt import io as gio          # noqa: E402
from gsevt import synthetic          # noqa: E402
from ply_writer import write_ply, gaussian_vertex   # noqa: E402


@pytest.fixture(scope="module")
def raw():
    return synthetic.patch_scene(P=3000, sh_degree=3, seed=1234, n_patches=8)


@pytest.mark.parametrize("fmt", ["binary_little_endian", "binary_big_endian", "ascii"])
def test_ply_roundtrip_3dgs_layout(tmp_path, raw, fmt):
    path = str(tmp_path / f"m_{fmt}.ply")
    write_ply(path, gaussian_vertex(raw), fmt=fmt)
    r = gio.read_ply(path)
    back = gio.ply_to_raw(r, 3)
    np.testing.assert_array_equal(back.xyz, raw.xyz.astype(np.float32))
    np.testing.assert_array_equal(back.features_dc, raw.features_dc.astype(np.float32))
    np.testing.assert_array_equal(back.features_rest, raw.features_rest.astype(np.float32))
    np.testing.assert_array_equal(back.opacity, raw.opacity.astype(np.float32))
    np.testing.assert_array_equal(back.scaling, raw.scaling.astype(np.float32))
    np.testing.assert_array_equal(back.rotation, raw.rotation.astype(np.float32))


def test_ply_any_order_types_and_leading_list_element(tmp_path, raw):
    rng = np.random.default_rng(3)
    vert = gaussian_vertex(raw)
    order = rng.permutation(len(vert))
    dts = ["f8" if k % 3 == 0 else "f4" for k in range(len(vert))]
    shuffled = [(vert[k][0], dts[k], vert[k][2].astype(np.float64) * 1.0000001 if dts[k] == "f8" else vert[k][2])
                for k in order]
    shuffled.append(("red", "u1", rng.integers(0, 255, raw.xyz.shape[0])))
    # an element with a list property before the vertex element
    counts = rng.integers(0, 5, 4).astype(np.uint8)
    body = b"".join(bytes([c]) + rng.standard_normal(c).astype("<f4").tobytes() + np.int32(7).tobytes()
                    for c in counts)
    pre = [("camera", 4, ["property list uchar float intr", "property int id"], body)]
    path = str(tmp_path / "mixed.ply")
    write_ply(path, shuffled, pre_elements=pre)
    r = gio.read_ply(path)
    want = {name: np.asarray(c, np.float64 if dt == "f8" else np.float32).astype(np.float32)
            for name, dt, c in shuffled}
    np.testing.assert_array_equal(r["xyz"][:, 1], want["y"])
    np.testing.assert_array_equal(r["opacity"][:, 0], want["opacity"])
    for i in range(45):
        np.testing.assert_array_equal(r["f_rest"][:, i], want[f"f_rest_{i}"])
    for i in range(4):
        np.testing.assert_array_equal(r["rots"][:, i], want[f"rot_{i}"])


def test_ply_errors(tmp_path, raw):
    vert = gaussian_vertex(raw)
    p1 = str(tmp_path / "missing.ply")
    write_ply(p1, [v for v in vert if v[0] != "opacity"])
    with pytest.raises(gio.GsevtIOError, match="opacity"):
        gio.read_ply(p1)
    p2 = str(tmp_path / "trunc.ply")
    write_ply(p2, vert)
    data = pathlib.Path(p2).read_bytes()
    pathlib.Path(p2).write_bytes(data[:-100])
    with pytest.raises(gio.GsevtIOError, match="truncated"):
        gio.read_ply(p2)
    p3 = str(tmp_path / "sh.ply")
    write_ply(p3, vert)
    with pytest.raises(AssertionError):
        gio.ply_to_raw(gio.read_ply(p3), max_sh_degree=2)
    with pytest.raises(gio.GsevtIOError):
        gio.read_ply(str(tmp_path / "nope.ply"))


def _reference_windows(lines, M, array_nums=None, start_time=None):
    """Line-by-line restatement of the reference's loop."""
    out, cur = [], []
    for line in lines:
        d = [int(item) for item in line.strip().split(" ")]
        if start_time is not None and d[0] < start_time:
            continue
        cur.append((d[0], d[1], d[2], d[3]))
        if len(cur) == M:
            out.append(cur)
            cur = []
            if array_nums is not None and len(out) >= array_nums:
                break
    return out


def _write_events(path, n, seed):
    rng = np.random.default_rng(seed)
    ts = np.sort(rng.integers(1_600_000_000_000, 1_600_000_100_000, n))
    x = rng.integers(0, 640, n)
    y = rng.integers(0, 480, n)
    p = rng.integers(0, 2, n)
    with open(path, "w") as f:
        for a, b, c, d in zip(ts, x, y, p):
            f.write(f"{a} {b} {c} {d}\n")
    return ts, x, y, p


@pytest.mark.parametrize("array_nums,start", [(None, None), (3, None), (None, "mid"), (2, "mid")])
def test_events_windows_match_reference_loop(tmp_path, array_nums, start):
    path = str(tmp_path / "ev.txt")
    ts, *_ = _write_events(path, 5300, 5)
    start_time = int(ts[1234]) if start == "mid" else None
    lines = pathlib.Path(path).read_text().splitlines(keepends=True)
    ref = _reference_windows(lines, 1000, array_nums, start_time)
    got = gio.load_events_from_txt(path, 1000, array_nums, start_time)
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        r = np.array(r)
        np.testing.assert_array_equal(g.ts, r[:, 0])
        np.testing.assert_array_equal(g.x, r[:, 1])
        np.testing.assert_array_equal(g.y, r[:, 2])
        np.testing.assert_array_equal(g.p, r[:, 3])


def test_events_large_file_matches_pandas_and_reports_rate(tmp_path):
    import pandas as pd
    path = str(tmp_path / "big.txt")
    ts, x, y, p = _write_events(path, 1_000_000, 6)
    t0 = time.perf_counter()
    a = gio.read_events_txt(path)
    dt = time.perf_counter() - t0
    df = pd.read_csv(path, sep=" ", header=None, dtype=np.int64)
    np.testing.assert_array_equal(a[0], df[0].to_numpy())
    np.testing.assert_array_equal(a[1], df[1].to_numpy())
    np.testing.assert_array_equal(a[3], df[3].to_numpy())
    assert a[4] == -1 and a[5] == 0
    mb = os.path.getsize(path) / 1e6
    print(f"[io] 1M events ({mb:.1f} MB) parsed in {dt * 1e3:.1f} ms = {1e6 / dt / 1e6:.1f} M events/s")


def test_events_malformed_line_raises_only_when_reached(tmp_path):
    path = str(tmp_path / "bad.txt")
    _write_events(path, 3000, 7)
    lines = pathlib.Path(path).read_text().splitlines(keepends=True)
    lines[2500] = "12 34\n"
    pathlib.Path(path).write_text("".join(lines))
    # two windows of 1000 end at line 1999: the bad line is never read
    assert len(gio.load_events_from_txt(path, 1000, array_nums=2)) == 2
    with pytest.raises(IndexError, match="line 2501"):
        gio.load_events_from_txt(path, 1000)
    with pytest.raises(IndexError):
        _reference_windows(lines, 1000)
    assert len(_reference_windows(lines, 1000, array_nums=2)) == 2
    lines[2500] = "12 34 x 1\n"
    pathlib.Path(path).write_text("".join(lines))
    with pytest.raises(ValueError, match="line 2501"):
        gio.load_events_from_txt(path, 1000)
    with pytest.raises(ValueError):
        _reference_windows(lines, 1000)
    lines[2500] = "12 34 56 1 99\n"          # extra integer fields are ignored there
    pathlib.Path(path).write_text("".join(lines))
    w = gio.load_events_from_txt(path, 1000)
    r = _reference_windows(lines, 1000)
    assert len(w) == len(r) == 3 and w[2].ts[500] == 12 and w[2].p[500] == 1


def test_events_edge_cases(tmp_path):
    path = str(tmp_path / "few.txt")
    _write_events(path, 999, 8)
    assert gio.load_events_from_txt(path, 1000) == []          # leftovers dropped
    e = str(tmp_path / "empty.txt")
    open(e, "w").close()
    assert gio.load_events_from_txt(e, 10) == []
    nonl = str(tmp_path / "nonl.txt")
    pathlib.Path(nonl).write_text("1 2 3 1\n4 5 6 0")                 # last line without newline
    w = gio.load_events_from_txt(nonl, 2)
    assert len(w) == 1 and list(w[0].ts) == [1, 4] and list(w[0].p) == [1, 0]


def test_io_library_exports_every_declared_symbol():
    import ctypes
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    body = re.sub(r"/\*.*?\*/", "", pathlib.Path(root, "include", "gsevt_io.h").read_text(), flags=re.S)
    names = sorted(set(re.findall(r"\b(gsevt_[a-z0-9_]+)\s*\(", body)))
    assert len(names) == 8
    L = ctypes.CDLL(os.path.join(root, "gs-evt_amd", "gsevt", "libgsevt_io.so"))
    assert not [n for n in names if not hasattr(L, n)]


def test_events_parse_rate_vs_line_loop(tmp_path):
    """Reported figures: the C++ parser against the reference's per-line
    loop (restated above, without its progress bar) on the same file."""
    path = str(tmp_path / "rate.txt")
    _write_events(path, 200_000, 9)
    lines = pathlib.Path(path).read_text().splitlines(keepends=True)
    t0 = time.perf_counter()
    _reference_windows(lines, 10_000)
    t_loop = time.perf_counter() - t0
    t0 = time.perf_counter()
    gio.load_events_from_txt(path, 10_000)
    t_cpp = time.perf_counter() - t0
    print(f"[io] 200k events: line loop {200_000 / t_loop / 1e6:.2f} M events/s, "
          f"libgsevt_io {200_000 / t_cpp / 1e6:.1f} M events/s ({os.cpu_count()} CPUs)")
    assert t_cpp < t_loop


@pytest.mark.parametrize("count", [float("nan"), float("inf"), 1.5, 1e30, 2.0 ** 40])
def test_ply_malformed_list_count_raises(tmp_path, raw, count):
    """A list count that is not a finite non-negative integer, or whose items
    cannot fit in the rest of the file, is an error (not a huge or undefined
    record size)."""
    vert = gaussian_vertex(raw)
    body = np.float32(count).tobytes() + np.zeros(2, "<f4").tobytes() + np.int32(7).tobytes()
    pre = [("camera", 1, ["property list float float intr", "property int id"], body)]
    path = str(tmp_path / "badlist.ply")
    write_ply(path, vert, pre_elements=pre)
    with pytest.raises(gio.GsevtIOError):
        gio.read_ply(path)
    # the same element with a valid count of 2 reads
    ok = str(tmp_path / "oklist.ply")
    write_ply(ok, vert, pre_elements=[("camera", 1, pre[0][2],
                                       np.float32(2).tobytes() + np.zeros(2, "<f4").tobytes()
                                       + np.int32(7).tobytes())])
    np.testing.assert_array_equal(gio.read_ply(ok)["xyz"][:, 0], raw.xyz[:, 0].astype(np.float32))


_GUARD_SHIM = r"""
#define _GNU_SOURCE
#include <dlfcn.h>
#include <sys/mman.h>
#include <sys/types.h>
/* file mappings get an inaccessible page right below them, so a read one
   byte before a mapping faults */
static void* guarded(void* addr, size_t len, int prot, int flags, int fd, off_t off) {
  static void* (*real)(void*, size_t, int, int, int, off_t);
  if (!real) real = (void* (*)(void*, size_t, int, int, int, off_t))dlsym(RTLD_NEXT, "mmap");
  if (fd < 0 || addr) return real(addr, len, prot, flags, fd, off);
  char* base = (char*)real(0, len + 8192, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
  if (base == MAP_FAILED) return base;
  return real(base + 4096, len, prot, flags | MAP_FIXED, fd, off);
}
void* mmap(void* a, size_t l, int p, int f, int fd, off_t o) { return guarded(a, l, p, f, fd, o); }
void* mmap64(void* a, size_t l, int p, int f, int fd, off_t o) { return guarded(a, l, p, f, fd, o); }
"""


def test_events_file_shorter_than_thread_count(tmp_path):
    """A 7-byte events file read by 64 OpenMP threads, with the page below
    the file mapping made inaccessible: the newline-aligned chunk split must
    not read before the mapping (it did for chunk starts at 0)."""
    import subprocess
    import sys
    shim_c = tmp_path / "guard.c"
    shim_c.write_text(_GUARD_SHIM)
    shim = str(tmp_path / "libguard.so")
    subprocess.run(["gcc", "-O1", "-shared", "-fPIC", "-o", shim, str(shim_c), "-ldl"], check=True)
    path = tmp_path / "tiny.txt"
    path.write_text("1 2 3 1")     # 7 bytes, no newline
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import sys; sys.path[:0] = [%r, %r]; from gsevt import io as gio; "
            "w = gio.load_events_from_txt(%r, 1); "
            "assert len(w) == 1 and list(w[0].ts) == [1] and list(w[0].x) == [2], w; print('ok')"
            % (root, os.path.join(root, "gs-evt_amd"), str(path)))
    env = dict(os.environ, OMP_NUM_THREADS="64")
    env["LD_PRELOAD"] = shim + (":" + os.environ["LD_PRELOAD"] if os.environ.get("LD_PRELOAD") else "")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), (r.returncode, r.stdout[-500:], r.stderr[-2000:])
