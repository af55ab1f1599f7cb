"""GSD files: the trajectory and restart format HOOMD-blue reads and writes.

Two layers, as in the published GSD specification (file layer 2.0) and its
``hoomd`` schema:

* ``GSDFile`` — the file layer.  A 256-byte header, an index of 32-byte
  entries (frame, N, location, M, id, type, flags) that says where each
  chunk's data lies, a name list of 64-byte NUL-padded names that the
  index's ``id`` points into, and the chunk data.  All integers are
  little-endian.  Data is appended as chunks are written; a frame's index
  entries go into the index, sorted by name id, when the frame ends, so a
  reader never sees a half-written frame.  A full index or name list is
  moved to the end of the file at twice the size and the header re-pointed.
  Unused index entries have location 0; the name list ends at the first
  empty name.
* ``HOOMDTrajectory`` / ``Frame`` — the ``hoomd`` schema (version 1.4):
  ``configuration/{step,dimensions,box}`` and ``particles/{N,types,typeid,
  mass,charge,diameter,body,moment_inertia,position,orientation,velocity,
  angmom,image}`` with the schema's types and defaults.  A chunk absent from
  frame i reads as frame 0's chunk, else as the default; ``append`` writes
  at frame 0 every chunk that differs from its default and at later frames
  every chunk that differs from what frame 0 resolves to (the step always).

The reference plugs into HOOMD-blue, whose GSD support comes from the
``gsd`` package (gsd 5.0.1 in the reference's lock file, ``pixi.lock:32``);
that package is not present here, so this module is written from the
format's specification and is checked by round trips and byte-level layout
tests rather than against files the ``gsd`` package wrote.
"""

import io
import os
import struct

import numpy as np

MAGIC = 0x65DF65DF65DF65DF
FILE_VERSION = (2 << 16) | 0
NAME_SIZE = 64
_HEADER = struct.Struct("<QQQQQII64s64s80s")
assert _HEADER.size == 256
INDEX_DTYPE = np.dtype([("frame", "<u8"), ("N", "<u8"), ("location", "<i8"), ("M", "<u4"),
                        ("id", "<u2"), ("type", "u1"), ("flags", "u1")])
assert INDEX_DTYPE.itemsize == 32
# type ids of the file layer
TYPES = {1: np.dtype("<u1"), 2: np.dtype("<u2"), 3: np.dtype("<u4"), 4: np.dtype("<u8"),
         5: np.dtype("<i1"), 6: np.dtype("<i2"), 7: np.dtype("<i4"), 8: np.dtype("<i8"),
         9: np.dtype("<f4"), 10: np.dtype("<f8")}
_TYPE_IDS = {(dt.kind, dt.itemsize): k for k, dt in TYPES.items()}
_INITIAL_INDEX = 128
_INITIAL_NAMES = 64


def make_version(major, minor):
    return (int(major) << 16) | int(minor)


def _pad(s, n):
    b = s.encode("ascii")
    if len(b) >= n:
        raise ValueError(f"name longer than {n - 1} bytes: {s!r}")
    return b + b"\0" * (n - len(b))


class GSDFile:
    """A GSD file opened for reading ('r'), reading and appending frames
    ('r+', 'a' — 'a' creates a missing file), or created empty ('w'
    truncates, 'x' refuses an existing file)."""

    def __init__(self, name, mode="r", application="azplugins_amd", schema="hoomd",
                 schema_version=(1, 4)):
        if mode not in ("r", "r+", "w", "x", "a"):
            raise ValueError(f"mode must be r, r+, w, x or a, not {mode!r}")
        self.name = os.fspath(name)
        self.mode = mode
        if mode in ("w", "x") or (mode == "a" and not os.path.exists(self.name)):
            if mode == "x" and os.path.exists(self.name):
                raise FileExistsError(self.name)
            self._create(application, schema, make_version(*schema_version))
        self._fh = io.open(self.name, "rb" if mode == "r" else "r+b")
        self._load()
        self._pending = []

    # ---- layout ----
    def _create(self, application, schema, schema_version):
        index_loc = _HEADER.size
        names_loc = index_loc + _INITIAL_INDEX * INDEX_DTYPE.itemsize
        head = _HEADER.pack(MAGIC, index_loc, _INITIAL_INDEX, names_loc, _INITIAL_NAMES, schema_version,
                            FILE_VERSION, _pad(application, NAME_SIZE), _pad(schema, NAME_SIZE), b"\0" * 80)
        with io.open(self.name, "wb") as fh:
            fh.write(head)
            fh.write(b"\0" * (_INITIAL_INDEX * INDEX_DTYPE.itemsize + _INITIAL_NAMES * NAME_SIZE))

    def _load(self):
        fh = self._fh
        fh.seek(0)
        raw = fh.read(_HEADER.size)
        if len(raw) < _HEADER.size:
            raise ValueError(f"{self.name}: not a GSD file (short header)")
        (magic, self._index_loc, self._index_alloc, self._names_loc, self._names_alloc, self.schema_version,
         self.file_version, app, schema, _) = _HEADER.unpack(raw)
        if magic != MAGIC:
            raise ValueError(f"{self.name}: not a GSD file (bad magic)")
        if self.file_version >> 16 != 2:
            raise ValueError(f"{self.name}: GSD file layer version {self.file_version >> 16}.x is not supported")
        self.application = app.rstrip(b"\0").decode("ascii")
        self.schema = schema.rstrip(b"\0").decode("ascii")
        fh.seek(self._names_loc)
        names = np.frombuffer(fh.read(self._names_alloc * NAME_SIZE), dtype=f"S{NAME_SIZE}")
        self._names = []
        for n in names:
            if not n:
                break
            self._names.append(n.decode("ascii"))
        self._name_id = {n: k for k, n in enumerate(self._names)}
        fh.seek(self._index_loc)
        index = np.frombuffer(fh.read(self._index_alloc * INDEX_DTYPE.itemsize), dtype=INDEX_DTYPE)
        used = np.flatnonzero(index["location"] == 0)
        self._n_index = int(used[0]) if used.size else self._index_alloc
        self._index = index[:self._n_index].copy()
        self.nframes = int(self._index["frame"][-1]) + 1 if self._n_index else 0
        self._lookup = {(int(e["frame"]), int(e["id"])): k for k, e in enumerate(self._index)}

    def _write_header(self):
        fh = self._fh
        fh.seek(8)
        fh.write(struct.pack("<QQQQ", self._index_loc, self._index_alloc, self._names_loc, self._names_alloc))

    def _grow(self, which):
        """Move the index or the name list to the end of the file at twice its size."""
        fh = self._fh
        fh.seek(0, os.SEEK_END)
        end = fh.tell()
        if which == "index":
            fh.seek(self._index_loc)
            old = fh.read(self._index_alloc * INDEX_DTYPE.itemsize)
            fh.seek(end)
            fh.write(old + b"\0" * len(old))
            self._index_loc, self._index_alloc = end, 2 * self._index_alloc
        else:
            fh.seek(self._names_loc)
            old = fh.read(self._names_alloc * NAME_SIZE)
            fh.seek(end)
            fh.write(old + b"\0" * len(old))
            self._names_loc, self._names_alloc = end, 2 * self._names_alloc
        self._write_header()

    # ---- writing ----
    def _writable(self):
        if self.mode == "r":
            raise PermissionError(f"{self.name} is open read-only")

    def write_chunk(self, name, data):
        """Appends chunk `name` of the current frame: a 1-D array (M = 1) or
        a 2-D array (N rows of M columns) of a file-layer type."""
        self._writable()
        a = np.ascontiguousarray(data)
        if a.ndim == 0:
            a = a.reshape(1)
        if a.ndim > 2:
            raise ValueError("chunks are 1-D or 2-D")
        tid = _TYPE_IDS.get((a.dtype.kind, a.dtype.itemsize))
        if tid is None:
            raise TypeError(f"chunk {name!r}: type {a.dtype} has no GSD type id")
        if a.size == 0:
            raise ValueError(f"chunk {name!r} is empty (a GSD chunk holds at least one value)")
        a = a.astype(TYPES[tid], copy=False)
        if name not in self._name_id:
            if len(self._names) == self._names_alloc:
                self._grow("names")
            self._fh.seek(self._names_loc + len(self._names) * NAME_SIZE)
            self._fh.write(_pad(name, NAME_SIZE))
            self._name_id[name] = len(self._names)
            self._names.append(name)
        N = a.shape[0]
        M = a.shape[1] if a.ndim == 2 else 1
        if any(p[4] == self._name_id[name] for p in self._pending):
            raise ValueError(f"chunk {name!r} written twice in one frame")
        self._fh.seek(0, os.SEEK_END)
        loc = self._fh.tell()
        self._fh.write(a.tobytes())
        self._pending.append((self.nframes, N, loc, M, self._name_id[name], tid, 0))

    def end_frame(self):
        """Commits the current frame's index entries (sorted by name id)."""
        self._writable()
        entries = np.array(sorted(self._pending, key=lambda p: p[4]), dtype=INDEX_DTYPE)
        while self._n_index + len(entries) > self._index_alloc:
            self._grow("index")
        self._fh.seek(self._index_loc + self._n_index * INDEX_DTYPE.itemsize)
        self._fh.write(entries.tobytes())
        for e in entries:
            self._lookup[(int(e["frame"]), int(e["id"]))] = self._n_index
            self._n_index += 1
        self._index = np.concatenate([self._index, entries])
        self._pending = []
        self.nframes += 1
        self._fh.flush()

    def flush(self):
        if self.mode != "r":
            self._fh.flush()

    # ---- reading ----
    def chunk_exists(self, frame, name):
        k = self._name_id.get(name)
        return k is not None and (int(frame), k) in self._lookup

    def read_chunk(self, frame, name):
        """The chunk's data: shape (N,) when M = 1, else (N, M)."""
        k = self._name_id.get(name)
        e = None if k is None else self._lookup.get((int(frame), k))
        if e is None:
            raise KeyError(f"frame {frame} has no chunk {name!r}")
        ent = self._index[e]
        dt = TYPES.get(int(ent["type"]))
        if dt is None:
            raise ValueError(f"chunk {name!r}: unknown GSD type id {int(ent['type'])}")
        N, M = int(ent["N"]), int(ent["M"])
        self._fh.seek(int(ent["location"]))
        raw = self._fh.read(N * M * dt.itemsize)
        if len(raw) != N * M * dt.itemsize:
            raise ValueError(f"chunk {name!r} of frame {frame} is truncated")
        a = np.frombuffer(raw, dtype=dt).copy()
        return a if M == 1 else a.reshape(N, M)

    def find_matching_chunk_names(self, prefix):
        return [n for n in self._names if n.startswith(prefix)]

    def close(self):
        if self._fh is not None:
            if self.mode != "r":
                self._fh.flush()
            self._fh.close()
            self._fh = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---- the hoomd schema ----
# name -> (dtype, columns or None, default value of one row)
SCHEMA = {
    "configuration/step": (np.uint64, None, 0),
    "configuration/dimensions": (np.uint8, None, 3),
    "configuration/box": (np.float32, None, None),
    "particles/N": (np.uint32, None, 0),
    "particles/typeid": (np.uint32, None, 0),
    "particles/mass": (np.float32, None, 1.0),
    "particles/charge": (np.float32, None, 0.0),
    "particles/diameter": (np.float32, None, 1.0),
    "particles/body": (np.int32, None, -1),
    "particles/moment_inertia": (np.float32, 3, (0.0, 0.0, 0.0)),
    "particles/position": (np.float32, 3, (0.0, 0.0, 0.0)),
    "particles/orientation": (np.float32, 4, (1.0, 0.0, 0.0, 0.0)),
    "particles/velocity": (np.float32, 3, (0.0, 0.0, 0.0)),
    "particles/angmom": (np.float32, 4, (0.0, 0.0, 0.0, 0.0)),
    "particles/image": (np.int32, 3, (0, 0, 0)),
}
PER_PARTICLE = [k for k in SCHEMA if k.startswith("particles/") and k != "particles/N"]
_DEFAULT_BOX = np.array([1, 1, 1, 0, 0, 0], dtype=np.float32)


class _Group:
    pass


class Frame:
    """One frame of the hoomd schema (the fields of ``gsd.hoomd.Frame``)."""

    def __init__(self):
        self.configuration = _Group()
        self.configuration.step = 0
        self.configuration.dimensions = 3
        self.configuration.box = _DEFAULT_BOX.copy()
        self.particles = _Group()
        self.particles.N = 0
        self.particles.types = ["A"]
        for key in PER_PARTICLE:
            setattr(self.particles, key.split("/")[1], None)

    def _default(self, key):
        dt, w, d = SCHEMA[key]
        N = int(self.particles.N)
        return np.full((N,) if w is None else (N, w), d, dtype=dt)

    def validate(self):
        """Converts every field to the schema's type and shape; an unset
        per-particle field becomes its default."""
        c = self.configuration
        c.step = int(c.step)
        c.dimensions = int(c.dimensions)
        c.box = np.asarray(c.box, dtype=np.float32).reshape(6)
        p = self.particles
        p.N = int(p.N)
        p.types = [str(t) for t in p.types]
        for key in PER_PARTICLE:
            attr = key.split("/")[1]
            v = getattr(p, attr)
            dt, w, _ = SCHEMA[key]
            shape = (p.N,) if w is None else (p.N, w)
            v = self._default(key) if v is None else np.ascontiguousarray(v, dtype=dt).reshape(shape)
            setattr(p, attr, v)
        return self

    def _chunks(self):
        """(name, array) for every chunk of the frame, in schema order."""
        self.validate()
        c, p = self.configuration, self.particles
        out = [("configuration/step", np.array([c.step], dtype=np.uint64)),
               ("configuration/dimensions", np.array([c.dimensions], dtype=np.uint8)),
               ("configuration/box", c.box),
               ("particles/N", np.array([p.N], dtype=np.uint32)),
               ("particles/types", _types_array(p.types))]
        out += [(key, getattr(p, key.split("/")[1])) for key in PER_PARTICLE]
        return out


def _types_array(types):
    width = max(len(t) for t in types) + 1 if types else 1
    a = np.zeros((len(types), width), dtype=np.int8)
    for k, t in enumerate(types):
        b = np.frombuffer(t.encode("ascii"), dtype=np.int8)
        a[k, :b.size] = b
    return a


def _types_list(a):
    a = np.asarray(a, dtype=np.int8).reshape(a.shape[0], -1)
    return [bytes(r.tobytes()).split(b"\0", 1)[0].decode("ascii") for r in a]


class HOOMDTrajectory:
    """A sequence of ``Frame`` over a ``GSDFile`` with the hoomd schema
    (``gsd.hoomd.open``'s object): ``len``, indexing, iteration, ``append``."""

    def __init__(self, file):
        if file.schema != "hoomd":
            raise ValueError(f"{file.name}: schema {file.schema!r}, expected 'hoomd'")
        if file.schema_version >> 16 >= 2:
            raise ValueError(f"{file.name}: hoomd schema {file.schema_version >> 16}.x is not supported")
        self.file = file
        self._initial = None

    def __len__(self):
        return self.file.nframes

    def _initial_chunks(self):
        if self._initial is None:
            self._initial = dict(self._read_chunks(0, fallback=False)) if len(self) else {}
        return self._initial

    def _read_chunks(self, i, fallback=True):
        f = self.file
        keys = ["configuration/step", "configuration/dimensions", "configuration/box", "particles/N",
                "particles/types"] + PER_PARTICLE
        out = []
        for key in keys:
            if f.chunk_exists(i, key):
                out.append((key, f.read_chunk(i, key)))
            elif fallback and key in self._initial_chunks():
                out.append((key, self._initial_chunks()[key]))
        return out

    def _read_frame(self, i):
        chunks = dict(self._read_chunks(i))
        fr = Frame()
        c, p = fr.configuration, fr.particles
        if "configuration/step" in chunks:
            c.step = int(chunks["configuration/step"][0])
        if "configuration/dimensions" in chunks:
            c.dimensions = int(chunks["configuration/dimensions"][0])
        if "configuration/box" in chunks:
            c.box = chunks["configuration/box"].astype(np.float32)
        if "particles/N" in chunks:
            p.N = int(chunks["particles/N"][0])
        if "particles/types" in chunks:
            p.types = _types_list(chunks["particles/types"])
        for key in PER_PARTICLE:
            if key in chunks:
                dt, w, _ = SCHEMA[key]
                v = chunks[key].astype(dt)
                shape = (p.N,) if w is None else (p.N, w)
                if v.size != int(np.prod(shape)):
                    raise ValueError(f"frame {i}: {key} holds {v.size} values for N = {p.N}")
                setattr(p, key.split("/")[1], v.reshape(shape))
        return fr.validate()

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self._read_frame(k) for k in range(*i.indices(len(self)))]
        i = int(i)
        if i < 0:
            i += len(self)
        if not 0 <= i < len(self):
            raise IndexError(f"frame {i} out of range ({len(self)} frames)")
        return self._read_frame(i)

    def __iter__(self):
        for i in range(len(self)):
            yield self._read_frame(i)

    def append(self, frame, only=None):
        """Writes `frame` as the next frame.  `only` (chunk names) limits a
        later frame to those chunks; the step is always written."""
        first = len(self) == 0
        init = {} if first else self._initial_chunks()
        for name, data in frame._chunks():
            if name == "configuration/step":
                self.file.write_chunk(name, data)
                continue
            if not first:
                if only is not None and name not in only:
                    continue
                if name in init:
                    if not _equal(data, init[name]):
                        self.file.write_chunk(name, data)
                    continue
            if data.size and not _is_default(name, data):
                self.file.write_chunk(name, data)
        self.file.end_frame()
        if first:
            self._initial = None

    def flush(self):
        self.file.flush()

    def close(self):
        self.file.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _is_default(name, data):
    """Whether a chunk holds the schema's default (per row, for per-particle
    chunks), compared without building a default array."""
    if name == "particles/types":
        return _types_list(data) == ["A"]
    if name == "configuration/box":
        return np.array_equal(data, _DEFAULT_BOX)
    if name == "particles/N":
        return int(data[0]) == 0
    dt, w, d = SCHEMA[name]
    row = np.asarray(d, dtype=dt)
    return bool(np.all(data == row))


def _equal(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b)


def open(name, mode="r"):  # noqa: A001 - gsd.hoomd.open's name
    """``gsd.hoomd.open``: modes 'r', 'r+', 'w', 'x', 'a' (a 'b' suffix, as
    in 'rb' or 'wb', is accepted and ignored)."""
    m = mode.replace("b", "")
    return HOOMDTrajectory(GSDFile(name, m, schema="hoomd", schema_version=(1, 4)))
