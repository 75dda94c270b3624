"""The KDTree3f wire and bag format (include/kdh_io.h) from Python.

A tree travels as the ROS 2 message kd_slam2_msgs/msg/KDTree3f in CDR and is
stored in rosbag2 bags (MCAP, or the older sqlite3 .db3 storage).  The
reference converts with rclcpp::Serialization inside rosbag2_cpp
(kd_slam2/src/lib/kd_io/rosbag2_message_writer.cpp:257-286 writes,
rosbag2_message_reader.cpp:150-190 reads); here the codec and the MCAP
framing are libkdhip.so's, and a tree read from a bag uploads with
DeviceTree like any other.
"""
import ctypes as C
import sqlite3

import numpy as np

from . import _lib as L
from .api import HostTree

SCHEMA_NAME = "kd_slam2_msgs/msg/KDTree3f"


class KDTree3fMessage:
    """The fields of KDTree3f; nodes is a Node_ array (NODE_DTYPE) whose
    idx_start / idx_end travel as uint32."""

    def __init__(self, nodes, frame_id="", src_topic="", stamp_sec=0, stamp_nanosec=0, ts_start=0.0,
                 root_mean=(0.0, 0.0, 0.0), root_eigenvalues=(0.0, 0.0, 0.0), root_eigenvectors=(0.0,) * 9):
        self.nodes = np.ascontiguousarray(nodes, dtype=L.NODE_DTYPE)
        self.frame_id = frame_id
        self.src_topic = src_topic
        self.stamp_sec = int(stamp_sec)
        self.stamp_nanosec = int(stamp_nanosec)
        self.ts_start = float(ts_start)
        self.root_mean = np.asarray(root_mean, dtype=np.float32).reshape(3)
        self.root_eigenvalues = np.asarray(root_eigenvalues, dtype=np.float32).reshape(3)
        self.root_eigenvectors = np.asarray(root_eigenvectors, dtype=np.float32).reshape(9)  # column-major

    def _info(self):
        info = L.KDTree3fInfo()
        info.stamp_sec = self.stamp_sec
        info.stamp_nanosec = self.stamp_nanosec
        info.frame_id = self.frame_id.encode()
        info.src_topic = self.src_topic.encode()
        info.ts_start = self.ts_start
        info.root_mean[:] = self.root_mean.tolist()
        info.root_eigenvalues[:] = self.root_eigenvalues.tolist()
        info.root_eigenvectors[:] = self.root_eigenvectors.tolist()
        return info


def encode(msg):
    """Serialize a KDTree3fMessage (CDR, little endian): bytes."""
    info = msg._info()
    size = C.c_size_t()
    L.check(L.lib().kdh_kdtree3f_encoded_size(C.byref(info), len(msg.nodes), C.byref(size)), what="encoded size")
    buf = (C.c_uint8 * size.value)()
    n = C.c_size_t()
    L.check(L.lib().kdh_kdtree3f_encode(C.byref(info), msg.nodes.ctypes.data_as(C.c_void_p), len(msg.nodes), buf,
                                        size.value, C.byref(n)), what="kdh_kdtree3f_encode")
    return C.string_at(buf, n.value)


def decode(data):
    """Parse a serialized KDTree3f: a KDTree3fMessage."""
    # bytes are read in place (the library only reads the buffer); other
    # buffer types are copied once
    raw = C.c_char_p(data) if isinstance(data, bytes) else (C.c_uint8 * len(data)).from_buffer_copy(data)
    info = L.KDTree3fInfo()
    nn = C.c_int64()
    cap = len(data) + 1
    fid, top = C.create_string_buffer(cap), C.create_string_buffer(cap)
    L.check(L.lib().kdh_kdtree3f_decode(raw, len(data), C.byref(info), fid, cap, top, cap, None, 0, C.byref(nn)),
            what="kdh_kdtree3f_decode")
    nodes = np.zeros(nn.value, dtype=L.NODE_DTYPE)
    L.check(L.lib().kdh_kdtree3f_decode(raw, len(data), C.byref(info), fid, cap, top, cap,
                                        nodes.ctypes.data_as(C.c_void_p), len(nodes), C.byref(nn)),
            what="kdh_kdtree3f_decode")
    return KDTree3fMessage(nodes, fid.value.decode(), top.value.decode(), info.stamp_sec, info.stamp_nanosec,
                           info.ts_start, list(info.root_mean), list(info.root_eigenvalues),
                           list(info.root_eigenvectors))


def message_from_tree(tree, frame_id="", src_topic="", stamp_sec=0, stamp_nanosec=0):
    """A KDTree3fMessage of a HostTree.  The root eigen-decomposition
    (Tree_::root_eigenvalues / root_eigenvectors, set from the root's
    labelling in TreeCPU_, tree/tree_cpu_.h:72-88) is recomputed in FP64
    from the tree's points when it holds them, else left at zero."""
    mean = tree.nodes["mean"][0] if len(tree.nodes) else np.zeros(3, np.float32)
    evals, evecs = np.zeros(3), np.zeros((3, 3))
    if tree.points is not None and len(tree.points) >= 3:
        P = np.asarray(tree.points, dtype=np.float64)
        mu = P.mean(axis=0)
        w, V = np.linalg.eigh((P - mu).T @ (P - mu) / len(P))
        m32 = mu.astype(np.float32)
        for k in (0, 2):  # reoriented against the mean, as the labelling does
            if float(np.dot(m32, V[:, k].astype(np.float32))) > 0:
                V[:, k] = -V[:, k]
        if np.linalg.det(V) < 0:
            V[:, 0] = -V[:, 0]
        evals, evecs = w, V
    return KDTree3fMessage(tree.nodes, frame_id, src_topic, stamp_sec, stamp_nanosec, tree.ts_start, mean, evals,
                           evecs.T.reshape(9))  # column-major


def tree_from_message(msg):
    """A HostTree (Node_ array and the recomputed leaf list) of a message."""
    nl = C.c_int64()
    L.check(L.lib().kdh_tree_leaves(msg.nodes.ctypes.data_as(C.c_void_p), len(msg.nodes), None, 0, C.byref(nl)),
            what="kdh_tree_leaves")
    leaves = np.zeros(nl.value, dtype=np.uint64)
    L.check(L.lib().kdh_tree_leaves(msg.nodes.ctypes.data_as(C.c_void_p), len(msg.nodes),
                                    leaves.ctypes.data_as(C.c_void_p), len(leaves), C.byref(nl)),
            what="kdh_tree_leaves")
    return HostTree(msg.nodes, leaves, None, msg.ts_start)


def read_mcap(path):
    """Every message of an MCAP file: dicts with topic, schema_name,
    message_encoding, log_time, publish_time, sequence, data (bytes)."""
    r = C.c_void_p()
    rc = L.lib().kdh_mcap_reader_open(path.encode(), C.byref(r))
    if rc != L.KDH_OK:
        raise RuntimeError(f"kdh_mcap_reader_open({path}): {L.ERRORS.get(rc, rc)}")
    out = []
    try:
        m = L.McapMessage()
        while True:
            rc = L.lib().kdh_mcap_reader_next(r, C.byref(m))
            if rc == L.KDH_ERR_NOT_FOUND:
                break
            if rc != L.KDH_OK:
                raise RuntimeError(f"MCAP {path}: {L.lib().kdh_mcap_reader_error(r).decode()}")
            out.append(dict(topic=m.topic.decode(), schema_name=m.schema_name.decode(),
                            message_encoding=m.message_encoding.decode(), log_time=m.log_time,
                            publish_time=m.publish_time, sequence=m.sequence,
                            data=C.string_at(m.data, m.size)))
    finally:
        L.lib().kdh_mcap_reader_close(r)
    return out


def write_mcap(path, records):
    """Write KDTree3f messages: records is a list of (topic, log_time_ns,
    KDTree3fMessage or serialized bytes)."""
    w = C.c_void_p()
    L.check(L.lib().kdh_mcap_writer_open(path.encode(), C.byref(w)), what="kdh_mcap_writer_open")
    channels = {}
    try:
        schema = L.lib().kdh_kdtree3f_schema_text()
        for topic, t, msg in records:
            if topic not in channels:
                cid = C.c_uint16()
                L.check(L.lib().kdh_mcap_writer_add_channel(w, topic.encode(), SCHEMA_NAME.encode(), schema,
                                                            C.byref(cid)), what="kdh_mcap_writer_add_channel")
                channels[topic] = cid.value
            data = msg if isinstance(msg, (bytes, bytearray)) else encode(msg)
            # bytes are passed in place (the writer only reads them)
            buf = C.c_char_p(data) if isinstance(data, bytes) else (C.c_uint8 * len(data)).from_buffer_copy(data)
            L.check(L.lib().kdh_mcap_writer_write(w, channels[topic], int(t), int(t), buf, len(data)),
                    what="kdh_mcap_writer_write")
    finally:
        L.check(L.lib().kdh_mcap_writer_close(w), what="kdh_mcap_writer_close")


def read_db3(path):
    """The messages of a rosbag2 sqlite3 storage file (.db3): dicts with
    topic, type, serialization_format, timestamp, data, in timestamp order."""
    con = sqlite3.connect(f"file:{path}?mode=ro", uri=True)
    try:
        rows = con.execute("SELECT topics.name, topics.type, topics.serialization_format, messages.timestamp, "
                           "messages.data FROM messages JOIN topics ON messages.topic_id = topics.id "
                           "ORDER BY messages.timestamp, messages.id").fetchall()
    finally:
        con.close()
    return [dict(topic=a, type=b, serialization_format=c, timestamp=d, data=bytes(e)) for a, b, c, d, e in rows]


def trees_in_bag(path, topic=None):
    """(log_time, KDTree3fMessage) for every KDTree3f message of an MCAP or
    .db3 file (optionally one topic)."""
    if str(path).endswith(".db3"):
        items = [(m["timestamp"], m["topic"], m["type"], m["data"]) for m in read_db3(path)]
    else:
        items = [(m["log_time"], m["topic"], m["schema_name"], m["data"]) for m in read_mcap(str(path))]
    return [(t, decode(d)) for t, tp, ty, d in items if ty == SCHEMA_NAME and (topic is None or tp == topic)]
