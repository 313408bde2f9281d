"""ROS1 bag event reader (libgsevt_io.so) against independently written bags.

rosbag itself is not installed, so the reader is checked against bags that
tests/bagfile.py writes from the published format 2.0 record layout, with
chunks compressed by Python's bz2 and the system liblz4 (parity with rosbag's
own reader: unpinned).  What is pinned:

* the decoded events, per-message counts and stamps equal what was written,
  for none / bz2 / lz4 chunks, several connections on the topic whose
  messages interleave in bag time (rosbag yields them in time order; ties in
  file order here), another topic of another type in the same chunks, empty
  messages, a nonzero polarity byte other than 1, and x, y up to 65535;
* the windows equal those of a literal restatement of the reference loop
  (utils/event_camera/event.py:42-69) over the decoded messages, for
  max_events_per_frame / array_nums / start_time combinations, including the
  array_nums quirk (its break leaves only the inner loop) and a message whose
  bag time equals start_time (kept: the test is `<`);
* chunks are selected through the chunk info records: chunks of another
  topic only are skipped, mixed chunks (the topic listed second) are read;
* an unindexed bag, a file that is not a bag, a truncated chunk, a topic of
  another layout and a missing file raise; an absent topic gives no windows.
"""
import pathlib
import random
import struct

import numpy as np
import pytest

from bagfile import BagWriter, EVENT_ARRAY_DEF, encode_event_array, lz4_available
from gsevt import io as gio

TOPIC = gio.BAG_TOPIC


def _make_bag(path, compression, seed=0, n_msgs=60, chunk_bytes=1500):
    """Two connections on TOPIC (prophesee and dvs types) whose bag times
    interleave with ties, plus an IMU-like topic; returns the messages of
    TOPIC in file order as (bag secs, bag nsecs, stamp secs, stamp nsecs,
    [(x, y, ts_secs, ts_nsecs, pol)])."""
    rng = random.Random(seed)
    w = BagWriter(path, compression, chunk_bytes)
    c0 = w.add_connection(TOPIC, "prophesee_event_msgs/EventArray", EVENT_ARRAY_DEF.format(pkg="prophesee_event_msgs"))
    c1 = w.add_connection(TOPIC, "dvs_msgs/EventArray", EVENT_ARRAY_DEF.format(pkg="dvs_msgs"))
    imu = w.add_connection("/imu", "sensor_msgs/Imu", "float64 a\nfloat64 b\n")
    msgs = []
    for i in range(n_msgs):
        conn = c0 if i % 3 else c1
        # bag times: mostly increasing, but connection c1 runs 2 steps late
        # and every 7th message ties with its predecessor
        step = i - 2 if conn == c1 else i
        bs, bn = 100 + step // 10, (step % 10) * 100_000_000 + (0 if i % 7 else 5)
        ss, sn = bs, bn + 123_457                 # header stamp differs from the bag time
        k = rng.choice([0, 1, 5, 17, 40])
        ev = [(rng.randrange(65536), rng.randrange(65536), ss, sn + j * 1000, rng.choice([0, 1, 2]))
              for j in range(k)]
        w.write(conn, bs, bn, encode_event_array(i, ss, sn, ev))
        msgs.append((bs, bn, ss, sn, ev))
        if i % 4 == 0:
            w.write(imu, bs, bn, struct.pack("<dd", 1.0, 2.0))
    w.close()
    return msgs


def _time_order(msgs):
    return sorted(msgs, key=lambda m: (m[0], m[1]))     # stable: ties keep file order


def _reference_windows(msgs, M, array_nums=None, start_time=None):
    """utils/event_camera/event.py:42-69 over decoded messages in bag time
    order; events as (header stamp us, x, y, p)."""
    event_arrays, cur = [], []
    for bs, bn, ss, sn, events in msgs:
        if start_time is not None and float(bs) + float(bn) / 1e9 < start_time:
            continue
        for x, y, _, _, pol in events:
            cur.append((ss * 1_000_000 + sn // 1000, x, y, 1 if pol else 0))
            if len(cur) >= M:
                event_arrays.append(cur)
                cur = []
                if array_nums is not None and len(event_arrays) >= array_nums:
                    break
    return event_arrays


COMPRESSIONS = ["none", "bz2"] + (["lz4"] if lz4_available() else [])


@pytest.mark.parametrize("compression", COMPRESSIONS)
def test_bag_events_roundtrip(tmp_path, compression):
    path = str(tmp_path / f"ev_{compression}.bag")
    msgs = _time_order(_make_bag(path, compression))
    ts, x, y, p, counts = gio.read_events_bag(path)
    np.testing.assert_array_equal(counts, [len(m[4]) for m in msgs])
    ev = [(m[2] * 1_000_000 + m[3] // 1000, e[0], e[1], 1 if e[4] else 0) for m in msgs for e in m[4]]
    assert len(ev) == ts.shape[0] and len(ev) > 0
    np.testing.assert_array_equal(ts, [e[0] for e in ev])
    np.testing.assert_array_equal(x, [e[1] for e in ev])
    np.testing.assert_array_equal(y, [e[2] for e in ev])
    np.testing.assert_array_equal(p, [e[3] for e in ev])
    assert x.max() > 60000 and set(np.unique(p).tolist()) <= {0, 1}
    ts2, *_ = gio.read_events_bag(path, event_stamps=True)
    np.testing.assert_array_equal(ts2, [e[2] * 1_000_000 + e[3] // 1000 for m in msgs for e in m[4]])


@pytest.mark.parametrize("M,array_nums,start", [(16, None, None), (16, 3, None), (1, 5, None), (0, None, None),
                                                 (25, None, "tie"), (9, 4, "tie"), (7.5, None, None)])
def test_bag_windows_match_reference_loop(tmp_path, M, array_nums, start):
    path = str(tmp_path / "w.bag")
    msgs = _time_order(_make_bag(path, "none", seed=3))
    start_time = None
    if start == "tie":    # the bag time of a message in the middle, exactly
        bs, bn = msgs[len(msgs) // 2][:2]
        start_time = float(bs) + float(bn) / 1e9
    want = _reference_windows(msgs, M, array_nums, start_time)
    got = gio.load_events_from_bag(path, M, array_nums=array_nums, start_time=start_time)
    assert len(got) == len(want) > 0
    for g, wv in zip(got, want):
        np.testing.assert_array_equal(np.stack([g.ts, g.x, g.y, g.p], 1), np.array(wv, np.int64))
    if array_nums is not None:      # the reference's break does not stop the outer loop
        assert len(got) >= array_nums


def test_bag_start_time_is_bag_time(tmp_path):
    path = str(tmp_path / "s.bag")
    w = BagWriter(path)
    c = w.add_connection(TOPIC, "dvs_msgs/EventArray", "")      # no definition: accepted by type name
    w.write(c, 10, 0, encode_event_array(0, 50, 0, [(1, 2, 50, 0, 1)]))
    w.write(c, 10, 1, encode_event_array(1, 5, 0, [(3, 4, 5, 0, 0)]))
    w.close()
    ts, x, y, p, counts = gio.read_events_bag(path, start_time=10.000000001)
    assert counts.tolist() == [1] and x.tolist() == [3] and ts.tolist() == [5_000_000]
    assert gio.read_events_bag(path, start_time=10.0)[4].tolist() == [1, 1]
    assert gio.load_events_from_bag(path, 5) == []         # the open window at the end is dropped


def test_bag_absent_topic_and_empty_bag(tmp_path):
    path = str(tmp_path / "e.bag")
    _make_bag(path, "none", n_msgs=5)
    assert gio.load_events_from_bag(path, 4, topic="/other") == []
    empty = str(tmp_path / "empty.bag")
    BagWriter(empty).close()
    assert gio.read_events_bag(empty)[0].shape == (0,)


def test_bag_errors(tmp_path):
    good = str(tmp_path / "g.bag")
    _make_bag(good, "bz2", n_msgs=30)
    with pytest.raises(gio.GsevtIOError, match="cannot open"):
        gio.read_events_bag(str(tmp_path / "missing.bag"))
    txt = tmp_path / "t.bag"
    txt.write_text("1 2 3 1\n")
    with pytest.raises(gio.GsevtIOError, match="not a ROS bag"):
        gio.read_events_bag(str(txt))
    unidx = str(tmp_path / "u.bag")
    w = BagWriter(unidx)
    c = w.add_connection(TOPIC, "dvs_msgs/EventArray", EVENT_ARRAY_DEF.format(pkg="dvs_msgs"))
    w.write(c, 1, 0, encode_event_array(0, 1, 0, [(1, 1, 1, 0, 1)]))
    w.close(index=False)
    with pytest.raises(gio.GsevtIOError, match="unindexed"):
        gio.read_events_bag(unidx)
    # a byte of the first chunk's compressed data flipped: bz2 refuses it
    data = bytearray(pathlib.Path(good).read_bytes())
    pos = 13 + 4096 + 200
    data[pos] ^= 0xFF
    bad = tmp_path / "c.bag"
    bad.write_bytes(bytes(data))
    with pytest.raises(gio.GsevtIOError, match="bz2|truncated|shorter"):
        gio.read_events_bag(str(bad))
    # the index section cut short
    cut = tmp_path / "cut.bag"
    cut.write_bytes(pathlib.Path(good).read_bytes()[:-10])
    with pytest.raises(gio.GsevtIOError, match="truncated"):
        gio.read_events_bag(str(cut))
    # an event message one byte short of its declared events
    short = str(tmp_path / "short.bag")
    w = BagWriter(short)
    c = w.add_connection(TOPIC, "dvs_msgs/EventArray", EVENT_ARRAY_DEF.format(pkg="dvs_msgs"))
    w.write(c, 1, 0, encode_event_array(0, 1, 0, [(1, 1, 1, 0, 1), (2, 2, 1, 0, 0)])[:-1])
    w.close()
    with pytest.raises(gio.GsevtIOError, match="shorter than its EventArray"):
        gio.read_events_bag(short)
    # a topic whose Event is laid out otherwise
    other = str(tmp_path / "o.bag")
    w = BagWriter(other)
    d = EVENT_ARRAY_DEF.format(pkg="my_msgs").replace("uint16 x\nuint16 y", "uint32 x\nuint32 y")
    c = w.add_connection(TOPIC, "my_msgs/EventArray", d)
    w.write(c, 1, 0, b"\0" * 40)
    w.close()
    with pytest.raises(gio.GsevtIOError, match="uint16 x, uint16 y"):
        gio.read_events_bag(other)
    # an unknown type without a definition
    nodef = str(tmp_path / "n.bag")
    w = BagWriter(nodef)
    c = w.add_connection(TOPIC, "my_msgs/EventArray", "")
    w.write(c, 1, 0, b"\0" * 40)
    w.close()
    with pytest.raises(gio.GsevtIOError, match="no message definition"):
        gio.read_events_bag(nodef)


def test_bag_layout_accepts_qualified_names(tmp_path):
    path = str(tmp_path / "q.bag")
    w = BagWriter(path)
    d = EVENT_ARRAY_DEF.format(pkg="prophesee_event_msgs").replace(
        "Header header", "std_msgs/Header header  # with a comment").replace(
        "Event[] events", "prophesee_event_msgs/Event[] events")
    c = w.add_connection(TOPIC, "prophesee_event_msgs/EventArray", d)
    w.write(c, 1, 0, encode_event_array(0, 1, 0, [(7, 8, 1, 0, 1)]))
    w.close()
    ts, x, y, p, counts = gio.read_events_bag(path)
    assert (x.tolist(), y.tolist(), p.tolist()) == ([7], [8], [1])


@pytest.mark.parametrize("compression", COMPRESSIONS)
def test_bag_chunks_without_the_topic_are_skipped(tmp_path, compression):
    """Chunks are chosen through the chunk info records: a chunk of another
    topic only is not read, mixed chunks are, and every event arrives."""
    path = str(tmp_path / "mixed.bag")
    w = BagWriter(path, compression, chunk_bytes=10 ** 9)        # chunks closed by hand below
    imu = w.add_connection("/imu", "sensor_msgs/Imu", "float64 a\n")
    c = w.add_connection(TOPIC, "dvs_msgs/EventArray", EVENT_ARRAY_DEF.format(pkg="dvs_msgs"))
    want = []
    for k in range(4):
        w.write(imu, 1 + k, 0, struct.pack("<d", float(k)))
        w._flush()                                                  # a chunk of /imu only
        ev = [(k, 10 + j, 1 + k, 5, j % 2) for j in range(3 + k)]
        w.write(imu, 1 + k, 1, struct.pack("<d", 0.5))              # listed first in the chunk info
        w.write(c, 1 + k, 2, encode_event_array(k, 1 + k, 2, ev))
        w._flush()                                                  # a mixed chunk
        want += [(e[0], e[1], 1 if e[4] else 0) for e in ev]
    w.write(imu, 9, 0, struct.pack("<d", 9.0))
    w.close()
    assert len(w.chunk_infos) == 9
    ts, x, y, p, counts = gio.read_events_bag(path)
    assert counts.tolist() == [3, 4, 5, 6]
    assert list(zip(x.tolist(), y.tolist(), p.tolist())) == want
