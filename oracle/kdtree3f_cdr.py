"""ORACLE -- test infrastructure only (tests/ use it as the checker).

An independent restatement of the KDTree3f wire format and of MCAP framing,
written from the specifications rather than from the product code:

- the ROS 2 message kd_slam2_msgs/msg/KDTree3f (KDTree3f.msg, KDNode3f.msg),
  with std_msgs/Header {builtin_interfaces/Time stamp {int32 sec,
  uint32 nanosec}, string frame_id}, serialized as OMG CDR little endian
  (encapsulation 0x0001, options 0; every primitive aligned to its size
  counted from the end of the 4-byte encapsulation; strings as uint32 length
  including the NUL, the bytes, the NUL; sequences as uint32 count +
  elements), which is how rclcpp::Serialization stores it (reference call
  sites: kd_slam2/src/lib/kd_io/rosbag2_message_writer.cpp:257-286,
  rosbag2_message_reader.cpp:150-190);
- MCAP records (https://mcap.dev/spec): opcode byte, uint64 length, fields;
  this writer puts the messages inside one uncompressed Chunk, as rosbag2's
  MCAP storage does, to exercise the product reader's chunk path.

The message is described as a field list and walked generically; no code is
shared with kd_slam2_amd/csrc/kdh_io.cpp.
"""
import struct

# (name, kind) with kind one of the primitive codes, "string", ("array", code, n),
# ("sequence", fields)
_PRIM = {"int32": "<i", "uint32": "<I", "float32": "<f", "float64": "<d"}

NODE_FIELDS = [
    ("mean", ("array", "float32", 3)),
    ("direction", ("array", "float32", 3)),
    ("idx_start", "uint32"),
    ("idx_end", "uint32"),
    ("idx_left", "int32"),
    ("idx_right", "int32"),
    ("dts_mean", "float64"),
    ("dts_covariance", "float64"),
]

TREE_FIELDS = [
    ("stamp_sec", "int32"),
    ("stamp_nanosec", "uint32"),
    ("frame_id", "string"),
    ("src_topic", "string"),
    ("ts_start", "float64"),
    ("root_mean", ("array", "float32", 3)),
    ("root_eigenvalues", ("array", "float32", 3)),
    ("root_eigenvectors", ("array", "float32", 9)),
    ("nodes", ("sequence", NODE_FIELDS)),
]


class _Writer:
    def __init__(self):
        self.body = bytearray()

    def _pad(self, size):
        while len(self.body) % size:
            self.body.append(0)

    def prim(self, code, value):
        fmt = _PRIM[code]
        self._pad(struct.calcsize(fmt))
        self.body += struct.pack(fmt, value)

    def fields(self, spec, values):
        for name, kind in spec:
            v = values[name]
            if kind == "string":
                raw = v.encode() + b"\0"
                self.prim("uint32", len(raw))
                self.body += raw
            elif isinstance(kind, tuple) and kind[0] == "array":
                for x in v:
                    self.prim(kind[1], x)
            elif isinstance(kind, tuple) and kind[0] == "sequence":
                self.prim("uint32", len(v))
                for item in v:
                    self.fields(kind[1], item)
            else:
                self.prim(kind, v)


class _Reader:
    def __init__(self, body):
        self.body = body
        self.pos = 0

    def prim(self, code):
        fmt = _PRIM[code]
        size = struct.calcsize(fmt)
        self.pos += (-self.pos) % size
        (v,) = struct.unpack_from(fmt, self.body, self.pos)
        self.pos += size
        return v

    def fields(self, spec):
        out = {}
        for name, kind in spec:
            if kind == "string":
                n = self.prim("uint32")
                out[name] = self.body[self.pos:self.pos + n].rstrip(b"\0").decode()
                self.pos += n
            elif isinstance(kind, tuple) and kind[0] == "array":
                out[name] = [self.prim(kind[1]) for _ in range(kind[2])]
            elif isinstance(kind, tuple) and kind[0] == "sequence":
                out[name] = [self.fields(kind[1]) for _ in range(self.prim("uint32"))]
            else:
                out[name] = self.prim(kind)
        return out


def encode(msg):
    """msg: dict with the TREE_FIELDS names (nodes: list of dicts)."""
    w = _Writer()
    w.fields(TREE_FIELDS, msg)
    return bytes([0x00, 0x01, 0x00, 0x00]) + bytes(w.body)


def decode(data):
    assert data[0] == 0x00 and data[1] == 0x01, "CDR little endian"
    return _Reader(bytes(data[4:])).fields(TREE_FIELDS)


# ---- MCAP ---------------------------------------------------------------------

MAGIC = b"\x89MCAP0\r\n"


def _s(text):
    raw = text.encode()
    return struct.pack("<I", len(raw)) + raw


def _record(op, body):
    return struct.pack("<BQ", op, len(body)) + body


def mcap_with_chunk(messages, topic, schema_name, schema_text, compression=""):
    """An MCAP file whose schema, channel and messages (list of (log_time,
    bytes)) sit in one Chunk record; the chunk records are stored as is, so
    only compression "" gives a readable file."""
    inner = _record(0x03, struct.pack("<H", 1) + _s(schema_name) + _s("ros2msg") +
                    struct.pack("<I", len(schema_text.encode())) + schema_text.encode())
    inner += _record(0x04, struct.pack("<HH", 7, 1) + _s(topic) + _s("cdr") + struct.pack("<I", 0))
    for seq, (t, data) in enumerate(messages):
        inner += _record(0x05, struct.pack("<HIQQ", 7, seq, t, t) + data)
    t0 = min(t for t, _ in messages) if messages else 0
    t1 = max(t for t, _ in messages) if messages else 0
    chunk = struct.pack("<QQQI", t0, t1, len(inner), 0) + _s(compression) + struct.pack("<Q", len(inner)) + inner
    out = MAGIC + _record(0x01, _s("ros2") + _s("oracle"))
    out += _record(0x06, chunk)
    out += _record(0x0F, struct.pack("<I", 0))
    out += _record(0x02, struct.pack("<QQI", 0, 0, 0))
    return out + MAGIC


def mcap_records(data):
    """The top-level records of an MCAP file: (offset, opcode, body)."""
    assert data[:8] == MAGIC and data[-8:] == MAGIC
    out, pos = [], 8
    while pos < len(data) - 8:
        op, n = struct.unpack_from("<BQ", data, pos)
        out.append((pos, op, data[pos + 9:pos + 9 + n]))
        pos += 9 + n
    assert pos == len(data) - 8
    return out
