"""Writer of ROS1 bags (format 2.0) for the bag reader's tests.

Record layout of the published rosbag format 2.0: "#ROSBAG V2.0\\n", a bag
header record padded to 4096 bytes, chunk records each followed by one index
data record per connection in it, then the index section (connection records,
chunk info records).  Every record is u32 header length, header fields (u32
length, "name=" + value), u32 data length, data.  bz2 chunks are compressed by
Python's bz2 module and lz4 chunks by the system liblz4's frame compressor, so
the reader's decoding is checked against independent encoders.
"""
from __future__ import annotations

import bz2
import ctypes
import struct

OP_MSG, OP_BAG_HEADER, OP_INDEX, OP_CHUNK, OP_CHUNK_INFO, OP_CONNECTION = 2, 3, 4, 5, 6, 7

EVENT_ARRAY_DEF = """# This message contains an array of events
Header header
uint32 height
uint32 width
Event[] events

================================================================================
MSG: std_msgs/Header
# sequence ID
uint32 seq
time stamp
string frame_id

================================================================================
MSG: {pkg}/Event
# A DVS event
uint16 x
uint16 y
time ts
bool polarity
"""


def _field(name, value):
    b = name.encode() + b"=" + value
    return struct.pack("<I", len(b)) + b


def _fields(d):
    return b"".join(_field(k, v) for k, v in d.items())


def _record(header, data):
    h = _fields(header)
    return struct.pack("<I", len(h)) + h + struct.pack("<I", len(data)) + data


def _time(secs, nsecs):
    return struct.pack("<II", secs, nsecs)


def encode_event_array(seq, secs, nsecs, events, frame_id=b"cam", height=480, width=640):
    """Serialised EventArray: events as (x, y, ts_secs, ts_nsecs, polarity byte)."""
    out = [struct.pack("<III", seq, secs, nsecs), struct.pack("<I", len(frame_id)), frame_id,
           struct.pack("<III", height, width, len(events))]
    out += [struct.pack("<HHIIB", *e) for e in events]
    return b"".join(out)


def lz4_frame(data):
    L = ctypes.CDLL("liblz4.so.1")
    L.LZ4F_compressFrameBound.restype = ctypes.c_size_t
    L.LZ4F_compressFrameBound.argtypes = [ctypes.c_size_t, ctypes.c_void_p]
    L.LZ4F_compressFrame.restype = ctypes.c_size_t
    L.LZ4F_compressFrame.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_void_p]
    L.LZ4F_isError.restype = ctypes.c_uint
    L.LZ4F_isError.argtypes = [ctypes.c_size_t]
    cap = L.LZ4F_compressFrameBound(len(data), None)
    dst = ctypes.create_string_buffer(cap)
    n = L.LZ4F_compressFrame(dst, cap, data, len(data), None)
    assert not L.LZ4F_isError(n)
    return dst.raw[:n]


def lz4_available():
    try:
        ctypes.CDLL("liblz4.so.1")
        return True
    except OSError:
        return False


class BagWriter:
    def __init__(self, path, compression="none", chunk_bytes=2048):
        self.path, self.compression, self.chunk_bytes = path, compression, chunk_bytes
        self.conns = []            # (id, topic, connection header bytes)
        self.chunk = bytearray()
        self.chunk_msgs = {}       # conn -> [(secs, nsecs, offset)]
        self.chunk_infos = []      # (pos, start, end, {conn: count})
        self.body = bytearray()    # everything after the bag header record
        self.written = set()

    def add_connection(self, topic, type_, definition, md5="0" * 32):
        cid = len(self.conns)
        hdr = _fields({"topic": topic.encode(), "type": type_.encode(), "md5sum": md5.encode(),
                       "message_definition": definition.encode()})
        self.conns.append((cid, topic, hdr))
        return cid

    def _conn_record(self, cid):
        _, topic, hdr = self.conns[cid]
        return _record({"op": bytes([OP_CONNECTION]), "conn": struct.pack("<I", cid), "topic": topic.encode()}, hdr)

    def write(self, conn, secs, nsecs, payload):
        if conn not in self.written:           # a connection record precedes its first message
            self.chunk += self._conn_record(conn)
            self.written.add(conn)
        self.chunk_msgs.setdefault(conn, []).append((secs, nsecs, len(self.chunk)))
        self.chunk += _record({"op": bytes([OP_MSG]), "conn": struct.pack("<I", conn),
                               "time": _time(secs, nsecs)}, payload)
        if len(self.chunk) >= self.chunk_bytes:
            self._flush()

    def _flush(self):
        if not self.chunk_msgs:
            return
        raw = bytes(self.chunk)
        data = {"none": lambda b: b, "bz2": bz2.compress, "lz4": lz4_frame}[self.compression](raw)
        pos = 4096 + 13 + len(self.body)        # file offset of the chunk record
        self.body += _record({"op": bytes([OP_CHUNK]), "compression": self.compression.encode(),
                              "size": struct.pack("<I", len(raw))}, data)
        times = []
        for cid, entries in self.chunk_msgs.items():
            self.body += _record({"op": bytes([OP_INDEX]), "ver": struct.pack("<I", 1),
                                  "conn": struct.pack("<I", cid), "count": struct.pack("<I", len(entries))},
                                 b"".join(_time(s, n) + struct.pack("<I", o) for s, n, o in entries))
            times += [(s, n) for s, n, _ in entries]
        self.chunk_infos.append((pos, min(times), max(times), {c: len(e) for c, e in self.chunk_msgs.items()}))
        self.chunk = bytearray()
        self.chunk_msgs = {}

    def close(self, index=True):
        self._flush()
        index_pos = 4096 + 13 + len(self.body)
        for cid in range(len(self.conns)):
            self.body += self._conn_record(cid)
        for pos, t0, t1, counts in self.chunk_infos:
            self.body += _record({"op": bytes([OP_CHUNK_INFO]), "ver": struct.pack("<I", 1),
                                  "chunk_pos": struct.pack("<Q", pos), "start_time": _time(*t0),
                                  "end_time": _time(*t1), "count": struct.pack("<I", len(counts))},
                                 b"".join(struct.pack("<II", c, n) for c, n in counts.items()))
        h = _fields({"op": bytes([OP_BAG_HEADER]), "index_pos": struct.pack("<Q", index_pos if index else 0),
                     "conn_count": struct.pack("<I", len(self.conns)),
                     "chunk_count": struct.pack("<I", len(self.chunk_infos))})
        pad = 4096 - 8 - len(h)
        header = struct.pack("<I", len(h)) + h + struct.pack("<I", pad) + b" " * pad
        assert len(header) == 4096
        with open(self.path, "wb") as f:
            f.write(b"#ROSBAG V2.0\n" + header + bytes(self.body))
