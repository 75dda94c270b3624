"""The KDTree3f wire and bag format (include/kdh_io.h, kd_slam2_amd/kdtree_msg.py)
against the CDR and MCAP specifications: known-answer bytes derived by hand,
the oracle's independent restatement (oracle/kdtree3f_cdr.py) byte for byte,
round trips, the reader's chunk path and its errors, and trees that survive a
bag round trip unchanged.  CPU only."""
import struct

import numpy as np
import pytest

from kd_slam2_amd import TreeGenerator3D, kdtree_msg as KM, synth
from kd_slam2_amd._lib import NODE_DTYPE
from oracle import kdtree3f_cdr as OC
from oracle import pyoracle as O


def _node(mean, direction, start, end, left, right, dts, cov):
    n = np.zeros(1, dtype=NODE_DTYPE)
    n["mean"], n["direction"] = mean, direction
    n["idx_start"], n["idx_end"], n["idx_left"], n["idx_right"] = start, end, left, right
    n["dts_mean"], n["dts_covariance"] = dts, cov
    return n


def test_known_answer_bytes_one_node():
    """A one-node message laid out by hand from the CDR rules.  Offsets below
    count from the end of the 4-byte encapsulation header."""
    nodes = _node([1.0, 2.0, 3.0], [0.0, 0.0, -1.0], 5, 9, -1, -1, 0.25, 0.5)
    msg = KM.KDTree3fMessage(nodes, frame_id="ab", src_topic="xyz", stamp_sec=7, stamp_nanosec=11, ts_start=1.5,
                             root_mean=[1, 2, 3], root_eigenvalues=[4, 5, 6], root_eigenvectors=list(range(9)))
    body = b""
    body += struct.pack("<iI", 7, 11)                    # 0: stamp
    body += struct.pack("<I", 3) + b"ab\0"               # 8: frame_id, length 3 with the NUL -> ends at 15
    body += b"\0"                                        # 15: pad to 16
    body += struct.pack("<I", 4) + b"xyz\0"              # 16: src_topic -> ends at 24 (8-aligned)
    body += struct.pack("<d", 1.5)                       # 24: ts_start
    body += struct.pack("<3f", 1, 2, 3) + struct.pack("<3f", 4, 5, 6) + struct.pack("<9f", *range(9))  # 32..92
    body += struct.pack("<I", 1)                         # 92: node count -> 96
    body += struct.pack("<3f3f", 1, 2, 3, 0, 0, -1)      # 96: mean, direction -> 120
    body += struct.pack("<IIii", 5, 9, -1, -1)           # 120: indices -> 136 (8-aligned)
    body += struct.pack("<dd", 0.25, 0.5)                # 136: dts -> 152
    want = bytes([0, 1, 0, 0]) + body
    got = KM.encode(msg)
    assert got == want
    assert OC.encode(OC.decode(want)) == want


def _random_message(rng, n_nodes, frame, topic):
    nodes = np.zeros(n_nodes, dtype=NODE_DTYPE)
    nodes["mean"] = rng.normal(size=(n_nodes, 3)).astype(np.float32)
    nodes["direction"] = rng.normal(size=(n_nodes, 3)).astype(np.float32)
    nodes["idx_start"] = rng.integers(0, 2**32, n_nodes, dtype=np.uint64)
    nodes["idx_end"] = rng.integers(0, 2**32, n_nodes, dtype=np.uint64)
    nodes["idx_left"] = rng.integers(-1, 1000, n_nodes)
    nodes["idx_right"] = rng.integers(-1, 1000, n_nodes)
    nodes["dts_mean"] = rng.normal(size=n_nodes)
    nodes["dts_covariance"] = rng.normal(size=n_nodes)
    return KM.KDTree3fMessage(nodes, frame, topic, int(rng.integers(-100, 100)), int(rng.integers(0, 10**9)),
                              float(rng.normal()), rng.normal(size=3), rng.normal(size=3), rng.normal(size=9))


def _as_oracle(msg):
    return dict(stamp_sec=msg.stamp_sec, stamp_nanosec=msg.stamp_nanosec, frame_id=msg.frame_id,
                src_topic=msg.src_topic, ts_start=msg.ts_start, root_mean=msg.root_mean.tolist(),
                root_eigenvalues=msg.root_eigenvalues.tolist(), root_eigenvectors=msg.root_eigenvectors.tolist(),
                nodes=[dict(mean=n["mean"].tolist(), direction=n["direction"].tolist(), idx_start=int(n["idx_start"]),
                            idx_end=int(n["idx_end"]), idx_left=int(n["idx_left"]), idx_right=int(n["idx_right"]),
                            dts_mean=float(n["dts_mean"]), dts_covariance=float(n["dts_covariance"]))
                       for n in msg.nodes])


def test_encoding_equals_the_oracle_for_every_padding():
    """Strings of every length mod 8 put the doubles and the nodes at every
    alignment; node counts 0, 1, 2, 7."""
    rng = np.random.default_rng(3)
    for n_nodes in (0, 1, 2, 7):
        for lf in range(9):
            for lt in range(0, 9, 3):
                msg = _random_message(rng, n_nodes, "f" * lf, "t" * lt)
                data = KM.encode(msg)
                assert data == OC.encode(_as_oracle(msg)), (n_nodes, lf, lt)
                back = KM.decode(data)
                assert back.frame_id == msg.frame_id and back.src_topic == msg.src_topic
                assert back.nodes.tobytes() == msg.nodes.tobytes()
                for k in ("stamp_sec", "stamp_nanosec", "ts_start"):
                    assert getattr(back, k) == getattr(msg, k)
                for k in ("root_mean", "root_eigenvalues", "root_eigenvectors"):
                    assert np.array_equal(getattr(back, k), getattr(msg, k))


def test_indices_travel_as_uint32():
    """idx_start / idx_end narrow like static_cast<uint32_t> (the reference
    writer); the reader widens them back."""
    nodes = _node([0, 0, 0], [1, 0, 0], 2**32 + 5, 2**33 + 7, 1, 2, 0.0, 0.0)
    back = KM.decode(KM.encode(KM.KDTree3fMessage(nodes)))
    assert int(back.nodes["idx_start"][0]) == 5 and int(back.nodes["idx_end"][0]) == 7


def test_malformed_buffers_are_errors():
    msg = KM.KDTree3fMessage(_node([1, 2, 3], [0, 0, 1], 0, 4, -1, -1, 0, 0), frame_id="map")
    data = KM.encode(msg)
    for cut in (3, 10, 40, len(data) - 1):
        with pytest.raises(RuntimeError):
            KM.decode(data[:cut])
    with pytest.raises(RuntimeError):
        KM.decode(bytes([0, 0, 0, 0]) + data[4:])  # big endian representation
    big = bytearray(data)
    big[92 + 4:96 + 4] = struct.pack("<I", 10**8)  # node count beyond the buffer
    with pytest.raises(RuntimeError):
        KM.decode(bytes(big))


def test_mcap_round_trip_and_tree_survives(tmp_path):
    """A built tree through a message and an MCAP file comes back with the
    same nodes and leaf list, and the oracle aligns against it to the same
    pose."""
    w = synth.config1(n_map=30_000, n_scan=4_096)
    g = TreeGenerator3D(leaf_max_points=16)
    fixed, moving = g.makeTree(w["map"]), g.makeTree(w["scan"], w["stamps"])
    path = str(tmp_path / "trees.mcap")
    msgs = [("/kd/map", 100, KM.message_from_tree(fixed, frame_id="map", src_topic="/points")),
            ("/kd/scan", 200, KM.message_from_tree(moving, frame_id="base", src_topic="/points")),
            ("/kd/scan", 300, KM.message_from_tree(moving, frame_id="base"))]
    KM.write_mcap(path, msgs)
    raw = open(path, "rb").read()
    assert raw[:8] == OC.MAGIC and raw[-8:] == OC.MAGIC
    got = KM.read_mcap(path)
    assert [(m["topic"], m["log_time"], m["sequence"]) for m in got] == [("/kd/map", 100, 0), ("/kd/scan", 200, 0),
                                                                          ("/kd/scan", 300, 1)]
    assert all(m["schema_name"] == KM.SCHEMA_NAME and m["message_encoding"] == "cdr" for m in got)
    trees = KM.trees_in_bag(path)
    assert len(trees) == 3
    f2, m2 = KM.tree_from_message(trees[0][1]), KM.tree_from_message(trees[1][1])
    assert f2.nodes.tobytes() == fixed.nodes.tobytes() and np.array_equal(f2.leaves, fixed.leaves)
    assert m2.nodes.tobytes() == moving.nodes.tobytes() and np.array_equal(m2.leaves, moving.leaves)
    p = O.from_kdh_params(__import__("kd_slam2_amd").default_params(), 8)
    a = O.compute(fixed, moving, w["X0"], w["Xf"], p)
    b = O.compute(f2, m2, w["X0"], w["Xf"], p)
    assert np.array_equal(a[0], b[0])
    # the root eigenvectors form a rotation
    E = trees[0][1].root_eigenvectors.reshape(3, 3).T
    assert abs(np.linalg.det(E) - 1) < 1e-5


def test_mcap_reader_follows_uncompressed_chunks_and_rejects_compressed(tmp_path):
    rng = np.random.default_rng(9)
    msgs = [_random_message(rng, k, "m", "s") for k in (1, 3, 0)]
    datas = [KM.encode(m) for m in msgs]
    schema = __import__("kd_slam2_amd")._lib.lib().kdh_kdtree3f_schema_text().decode()
    path = tmp_path / "chunked.mcap"
    path.write_bytes(OC.mcap_with_chunk([(10 * i, d) for i, d in enumerate(datas)], "/t", KM.SCHEMA_NAME, schema))
    got = KM.read_mcap(str(path))
    assert [m["data"] for m in got] == datas and [m["topic"] for m in got] == ["/t"] * 3
    bad = tmp_path / "zstd.mcap"
    bad.write_bytes(OC.mcap_with_chunk([(0, datas[0])], "/t", KM.SCHEMA_NAME, schema, compression="zstd"))
    with pytest.raises(RuntimeError, match="compressed chunk"):
        KM.read_mcap(str(bad))


def test_db3_storage(tmp_path):
    """rosbag2's sqlite3 storage: topics and messages tables."""
    import sqlite3

    msg = KM.message_from_tree(TreeGenerator3D().makeTree(synth.config1(n_map=5_000, n_scan=2_048)["map"]))
    db = tmp_path / "bag_0.db3"
    con = sqlite3.connect(db)
    con.execute("CREATE TABLE topics(id INTEGER PRIMARY KEY, name TEXT NOT NULL, type TEXT NOT NULL, "
                "serialization_format TEXT NOT NULL, offered_qos_profiles TEXT NOT NULL)")
    con.execute("CREATE TABLE messages(id INTEGER PRIMARY KEY, topic_id INTEGER NOT NULL, "
                "timestamp INTEGER NOT NULL, data BLOB NOT NULL)")
    con.execute("INSERT INTO topics VALUES (1, '/kd/map', ?, 'cdr', '')", (KM.SCHEMA_NAME,))
    con.execute("INSERT INTO messages VALUES (1, 1, 42, ?)", (KM.encode(msg),))
    con.commit()
    con.close()
    (t, back), = KM.trees_in_bag(str(db))
    assert t == 42 and back.nodes.tobytes() == msg.nodes.tobytes()


def test_mcap_writer_layout(tmp_path):
    """The writer's file, walked by the oracle's record reader: header
    profile, data section, DataEnd, summary groups that the SummaryOffset
    records and the Footer point at, and statistics that count the messages."""
    rng = np.random.default_rng(1)
    path = tmp_path / "w.mcap"
    KM.write_mcap(str(path), [("/a", 5, _random_message(rng, 2, "f", "s")), ("/b", 9, _random_message(rng, 0, "", "")),
                              ("/a", 7, _random_message(rng, 1, "g", ""))])
    data = path.read_bytes()
    recs = OC.mcap_records(data)
    ops = [op for _, op, _ in recs]
    assert ops == [0x01, 0x03, 0x04, 0x05, 0x04, 0x05, 0x05, 0x0F, 0x03, 0x04, 0x04, 0x0B, 0x0E, 0x0E, 0x0E, 0x02]
    assert recs[0][2] == OC._s("ros2") + OC._s("kd_slam2_amd")
    summary_start, offsets_start, crc = struct.unpack("<QQI", recs[-1][2])
    assert summary_start == recs[8][0] and offsets_start == recs[12][0] and crc == 0
    for off, op, body in recs[12:15]:
        gop, start, length = struct.unpack("<BQQ", body)
        members = [r for r in recs if start <= r[0] < start + length]
        assert members and all(r[1] == gop for r in members)
        assert sum(9 + len(r[2]) for r in members) == length
    st = recs[11][2]
    n_msg, n_schema, n_chan, n_att, n_meta, n_chunk, t0, t1, mlen = struct.unpack_from("<QHIIIIQQI", st)
    assert (n_msg, n_schema, n_chan, n_att, n_meta, n_chunk, t0, t1) == (3, 1, 2, 0, 0, 0, 5, 9)
    counts = dict(struct.unpack_from("<HQ", st, 46 + 10 * k) for k in range(mlen // 10))
    assert counts == {0: 2, 1: 1}


def test_count_only_decode_agrees_with_the_full_parse_on_every_cut():
    """kdh_kdtree3f_decode asked for the count only (nodes NULL) finds the end
    of the node records from their fixed layout; it must accept and reject
    exactly the buffers the node-by-node parse does.  Frame ids of several
    lengths put the first node at both 8-byte phases."""
    import ctypes as C

    from kd_slam2_amd import _lib as L

    nodes = np.concatenate([_node([k, 2, 3], [0, 0, 1], k, k + 4, -1, -1, 0.5 * k, 0.25) for k in range(5)])
    for fid in ("", "m", "map", "odom", "map_frame_x"):
        data = KM.encode(KM.KDTree3fMessage(nodes, frame_id=fid, src_topic="/t"))
        for cut in range(4, len(data) + 1):
            raw = (C.c_uint8 * cut).from_buffer_copy(data[:cut])
            out = []
            for fill in (False, True):
                info, nn = L.KDTree3fInfo(), C.c_int64()
                fb, tb = C.create_string_buffer(64), C.create_string_buffer(64)
                buf = np.zeros(len(nodes), dtype=NODE_DTYPE)
                rc = L.lib().kdh_kdtree3f_decode(raw, cut, C.byref(info), fb, 64, tb, 64,
                                                 buf.ctypes.data_as(C.c_void_p) if fill else None,
                                                 len(buf) if fill else 0, C.byref(nn))
                out.append((rc, nn.value if rc == L.KDH_OK else None))
            assert out[0] == out[1], (fid, cut, out)
            assert (out[1][0] == L.KDH_OK) == (cut == len(data)), (fid, cut)
