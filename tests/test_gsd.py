"""GSD file layer and hoomd schema (host code, no GPU).

The ``gsd`` package is not installed, so files are checked against the
published layout (file layer 2.0): a file assembled here byte by byte with
``struct`` must read back, and a written file must hold its header, index
entries, names and data at the offsets the layout fixes.
"""

import struct

import numpy as np
import pytest

from azplugins_amd.harness import gsd
from azplugins_amd.harness.state import Snapshot


def _header(raw):
    return struct.unpack_from("<QQQQQII64s64s", raw, 0)


def test_header_index_and_names_at_their_offsets(tmp_path):
    fn = tmp_path / "a.gsd"
    pos = np.arange(12, dtype=np.float32).reshape(4, 3)
    with gsd.GSDFile(fn, "w", application="app", schema="hoomd", schema_version=(1, 4)) as f:
        f.write_chunk("particles/position", pos)
        f.write_chunk("configuration/step", np.array([7], dtype=np.uint64))
        f.end_frame()
    raw = fn.read_bytes()
    magic, iloc, ialloc, nloc, nalloc, sver, fver, app, schema = _header(raw)
    assert magic == 0x65DF65DF65DF65DF
    assert fver == 2 << 16 and sver == (1 << 16) | 4
    assert app.rstrip(b"\0") == b"app" and schema.rstrip(b"\0") == b"hoomd"
    assert len(app) == 64 and raw[176:256] == b"\0" * 80
    # names in order of first use, 64 bytes each, then an empty name
    assert raw[nloc:nloc + 64].rstrip(b"\0") == b"particles/position"
    assert raw[nloc + 64:nloc + 128].rstrip(b"\0") == b"configuration/step"
    assert raw[nloc + 128:nloc + 192] == b"\0" * 64
    # index entries: frame, N, location, M, id, type, flags; sorted by id
    e0 = struct.unpack_from("<QQqIHBB", raw, iloc)
    e1 = struct.unpack_from("<QQqIHBB", raw, iloc + 32)
    assert e0[0] == 0 and e0[1] == 4 and e0[3] == 3 and e0[4] == 0 and e0[5] == 9 and e0[6] == 0
    assert e1[0] == 0 and e1[1] == 1 and e1[3] == 1 and e1[4] == 1 and e1[5] == 4
    assert raw[e0[2]:e0[2] + 48] == pos.tobytes()
    assert struct.unpack_from("<Q", raw, e1[2])[0] == 7
    # the rest of the index is unused (location 0)
    assert struct.unpack_from("<q", raw, iloc + 64 + 16)[0] == 0
    assert ialloc >= 2 and nalloc >= 3


def _hand_built(fn):
    """Two frames written with struct only: frame 0 holds N, types, typeid,
    position and the step; frame 1 holds the step and new positions."""
    names = [b"configuration/step", b"particles/N", b"particles/types", b"particles/typeid",
             b"particles/position"]
    p0 = np.array([[0.5, -1.0, 2.0], [1.5, 0.25, -3.0]], dtype="<f4")
    p1 = p0 + np.float32(0.125)
    types = np.array([[65, 0, 0], [66, 98, 0]], dtype="<i1")  # "A", "Bb"
    blobs = [
        (0, 0, struct.pack("<Q", 10), 1, 1, 4),
        (0, 1, struct.pack("<I", 2), 1, 1, 3),
        (0, 2, types.tobytes(), 2, 3, 5),
        (0, 3, struct.pack("<II", 1, 0), 2, 1, 3),
        (0, 4, p0.tobytes(), 2, 3, 9),
        (1, 0, struct.pack("<Q", 20), 1, 1, 4),
        (1, 4, p1.tobytes(), 2, 3, 9),
    ]
    n_index, n_names = 16, 8
    iloc = 256
    nloc = iloc + 32 * n_index
    data = nloc + 64 * n_names
    out = bytearray(struct.pack("<QQQQQII64s64s80s", 0x65DF65DF65DF65DF, iloc, n_index, nloc, n_names,
                                (1 << 16) | 4, 2 << 16, b"hand", b"hoomd", b""))
    out += b"\0" * (32 * n_index + 64 * n_names)
    for k, n in enumerate(names):
        out[nloc + 64 * k:nloc + 64 * k + len(n)] = n
    loc = data
    for k, (frame, nid, b, N, M, t) in enumerate(blobs):
        out[iloc + 32 * k:iloc + 32 * k + 32] = struct.pack("<QQqIHBB", frame, N, loc, M, nid, t, 0)
        out += b
        loc += len(b)
    fn.write_bytes(bytes(out))
    return p0, p1


def test_reads_a_file_assembled_from_the_layout(tmp_path):
    fn = tmp_path / "hand.gsd"
    p0, p1 = _hand_built(fn)
    with gsd.open(fn, "rb") as t:
        assert len(t) == 2
        f0, f1 = t[0], t[1]
        assert t.file.application == "hand"
    assert f0.configuration.step == 10 and f1.configuration.step == 20
    assert f0.particles.types == ["A", "Bb"] and f1.particles.types == ["A", "Bb"]
    assert np.array_equal(f0.particles.typeid, [1, 0]) and np.array_equal(f1.particles.typeid, [1, 0])
    assert np.array_equal(f0.particles.position, p0) and np.array_equal(f1.particles.position, p1)
    # chunks absent from both frames read as the schema's defaults
    assert np.array_equal(f1.particles.mass, [1.0, 1.0])
    assert np.array_equal(f1.particles.orientation, [[1, 0, 0, 0]] * 2)
    assert np.array_equal(f1.particles.body, [-1, -1])
    assert np.array_equal(f1.configuration.box, [1, 1, 1, 0, 0, 0])
    assert f1.configuration.dimensions == 3


@pytest.mark.parametrize("dtype", ["u1", "u2", "u4", "u8", "i1", "i2", "i4", "i8", "f4", "f8"])
def test_every_type_round_trips(tmp_path, dtype):
    fn = tmp_path / "t.gsd"
    rng = np.random.default_rng(1)
    a = (rng.normal(size=(5, 3)) * 100).astype(dtype)
    b = (rng.normal(size=7) * 100).astype(dtype)
    with gsd.GSDFile(fn, "w") as f:
        f.write_chunk("a", a)
        f.write_chunk("b", b)
        f.end_frame()
    with gsd.GSDFile(fn, "r") as f:
        ra, rb = f.read_chunk(0, "a"), f.read_chunk(0, "b")
    assert ra.dtype == np.dtype(dtype) and ra.shape == (5, 3) and np.array_equal(ra, a)
    assert rb.shape == (7,) and np.array_equal(rb, b)


def test_index_and_names_grow_and_append_reopens(tmp_path):
    """300 names (the name list starts at 64) over 3 frames (900 index
    entries; the index starts at 128), then two more frames appended after
    the file was closed."""
    fn = tmp_path / "g.gsd"
    names = [f"log/q{k:03d}" for k in range(300)]
    with gsd.GSDFile(fn, "w") as f:
        for fr in range(3):
            for k, n in enumerate(names):
                f.write_chunk(n, np.array([fr * 1000 + k], dtype=np.int64))
            f.end_frame()
    with gsd.GSDFile(fn, "a") as f:
        assert f.nframes == 3
        for fr in range(3, 5):
            f.write_chunk(names[5], np.array([fr], dtype=np.int64))
            f.end_frame()
    with gsd.GSDFile(fn, "r") as f:
        assert f.nframes == 5
        for fr in range(3):
            got = [int(f.read_chunk(fr, n)[0]) for n in names]
            assert got == [fr * 1000 + k for k in range(300)]
        assert int(f.read_chunk(4, names[5])[0]) == 4
        assert not f.chunk_exists(4, names[6])
        assert f.find_matching_chunk_names("log/q29") == [f"log/q29{k}" for k in range(10)]


def test_uncommitted_frame_is_not_read(tmp_path):
    fn = tmp_path / "u.gsd"
    f = gsd.GSDFile(fn, "w")
    f.write_chunk("x", np.array([1.0]))
    f.end_frame()
    f.write_chunk("x", np.array([2.0]))
    f.flush()  # data of frame 1 is in the file, its index entry is not
    with gsd.GSDFile(fn, "r") as r:
        assert r.nframes == 1 and float(r.read_chunk(0, "x")[0]) == 1.0
    f.close()


def test_errors(tmp_path):
    fn = tmp_path / "e.gsd"
    fn.write_bytes(b"\0" * 300)
    with pytest.raises(ValueError, match="magic"):
        gsd.GSDFile(fn, "r")
    with pytest.raises(ValueError, match="mode"):
        gsd.GSDFile(fn, "q")
    with gsd.GSDFile(tmp_path / "ok.gsd", "w") as f:
        with pytest.raises(ValueError, match="empty"):
            f.write_chunk("x", np.zeros(0))
        with pytest.raises(TypeError):
            f.write_chunk("x", np.zeros(2, dtype=np.complex64))
        f.write_chunk("x", np.zeros(2))
        with pytest.raises(ValueError, match="twice"):
            f.write_chunk("x", np.zeros(2))
        f.end_frame()
    with pytest.raises(FileExistsError):
        gsd.GSDFile(tmp_path / "ok.gsd", "x")
    with gsd.GSDFile(tmp_path / "ok.gsd", "r") as f:
        with pytest.raises(PermissionError):
            f.write_chunk("y", np.zeros(1))
        with pytest.raises(KeyError):
            f.read_chunk(0, "nope")


def _frame(N, seed, step):
    rng = np.random.default_rng(seed)
    fr = gsd.Frame()
    fr.configuration.step = step
    fr.configuration.box = [10, 11, 12, 0, 0, 0]
    fr.particles.N = N
    fr.particles.types = ["A", "B", "long_name"]
    fr.particles.typeid = rng.integers(0, 3, N)
    fr.particles.position = rng.uniform(-5, 5, (N, 3))
    fr.particles.velocity = rng.normal(size=(N, 3))
    fr.particles.image = rng.integers(-2, 3, (N, 3))
    return fr.validate()


def test_schema_round_trip_and_fallback(tmp_path):
    fn = tmp_path / "s.gsd"
    a, b = _frame(50, 1, 0), _frame(50, 2, 100)
    with gsd.open(fn, "wb") as t:
        t.append(a)
        # later frames: only position (and the step) are written
        t.append(b, only={"particles/position"})
    with gsd.open(fn, "rb") as t:
        f = t.file
        # defaults are not stored; the step always is
        assert not f.chunk_exists(0, "particles/mass") and not f.chunk_exists(0, "particles/orientation")
        assert f.chunk_exists(0, "particles/velocity") and f.chunk_exists(1, "configuration/step")
        assert not f.chunk_exists(1, "particles/velocity") and f.chunk_exists(1, "particles/position")
        r0, r1 = t[0], t[-1]
        assert len(t[:]) == 2
    for name in ("typeid", "position", "velocity", "image", "mass", "orientation"):
        assert np.array_equal(getattr(r0.particles, name), getattr(a.particles, name)), name
    assert r0.particles.types == ["A", "B", "long_name"]
    assert r1.configuration.step == 100
    assert np.array_equal(r1.particles.position, b.particles.position)
    # velocity was not written at frame 1: it reads as frame 0's
    assert np.array_equal(r1.particles.velocity, a.particles.velocity)
    assert np.array_equal(r1.configuration.box, np.float32([10, 11, 12, 0, 0, 0]))


def test_snapshot_frame_conversion():
    fr = _frame(20, 3, 55)
    fr.particles.mass = np.linspace(1, 2, 20)
    fr.particles.body = np.arange(20) - 5
    snap = Snapshot.from_gsd_frame(fr)
    assert snap.configuration.step == 55 and snap.particles.N == 20
    assert snap.particles.position.dtype == np.float64
    assert np.array_equal(snap.particles.position, fr.particles.position.astype(np.float64))
    back = snap._gsd_frame(step=55)
    ref = dict(fr._chunks())
    for name, arr in back._chunks():
        assert np.array_equal(arr, ref[name]), name
