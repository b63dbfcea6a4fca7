"""Minimal reader for GSD trajectory files (file layer v1 and v2, HOOMD schema).

Trajectory ingestion for the batched path and for the reference's own fixtures
(``tests/integration/files/lj/lj.gsd``,
``tests/validation/files/steinhardt_average/Test_Configuration.gsd`` in
freud 3.5.0).  The ``gsd`` package is not installed in this environment, so this
implements the published on-disk format directly: a 256-byte header, an index
of 32-byte chunk entries ``(frame u64, N u64, location i64, M u32, id u16,
type u8, flags u8)`` and a name list (64-byte fixed entries in v1,
NUL-separated strings in v2).  Frames inherit a chunk from frame 0 when they do
not store it, as the HOOMD schema specifies.
"""

import struct

import numpy as np

_MAGIC = 0x65DF65DF65DF65DF
_TYPES = {
    1: np.uint8,
    2: np.uint16,
    3: np.uint32,
    4: np.uint64,
    5: np.int8,
    6: np.int16,
    7: np.int32,
    8: np.int64,
    9: np.float32,
    10: np.float64,
    11: np.uint8,
}


class Frame:
    """One HOOMD frame: ``box`` (Lx, Ly, Lz, xy, xz, yz), ``dimensions``, ``positions``."""

    def __init__(self, box, dimensions, positions, step):
        self.box = box
        self.dimensions = dimensions
        self.positions = positions
        self.step = step

    @property
    def points(self):
        return self.positions

    def freud_box(self):
        """Box parameters with HOOMD's 2D convention (Lz, xz, yz zeroed)."""
        b = np.array(self.box, dtype=np.float32).copy()
        if self.dimensions == 2:
            b[[2, 4, 5]] = 0
        return b


class GSDFile:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self._buf = fh.read()
        b = self._buf
        magic, idx_loc, idx_alloc, nl_loc, nl_alloc, _schema_ver, gsd_ver = struct.unpack("<QQQQQII", b[:48])
        if magic != _MAGIC:
            raise ValueError(f"{path} is not a GSD file")
        self.gsd_version = gsd_ver
        major = gsd_ver >> 16
        if major == 1 or gsd_ver < 0x10000:
            names = []
            for e in range(nl_alloc):
                raw = b[nl_loc + 64 * e: nl_loc + 64 * (e + 1)]
                name = raw.split(b"\x00", 1)[0].decode()
                if not name:
                    break
                names.append(name)
        else:
            names = []
            pos = nl_loc
            end = len(b)
            while pos < end:
                z = b.index(b"\x00", pos)
                if z == pos:
                    break
                names.append(b[pos:z].decode())
                pos = z + 1
        self._names = names
        self._index = {}
        n_frames = 0
        for e in range(idx_alloc):
            off = idx_loc + 32 * e
            if off + 32 > len(b):
                break
            frame, n, loc, m, cid, ty, _fl = struct.unpack("<QQqIHBB", b[off: off + 32])
            if loc == 0:
                continue
            self._index[(frame, names[cid])] = (n, m, loc, ty)
            n_frames = max(n_frames, frame + 1)
        self.n_frames = n_frames

    def read_chunk(self, frame, name):
        key = (frame, name)
        if key not in self._index:
            key = (0, name)
            if key not in self._index:
                return None
        n, m, loc, ty = self._index[key]
        dt = _TYPES[ty]
        count = n * m
        arr = np.frombuffer(self._buf, dtype=dt, count=count, offset=loc)
        return arr.reshape(n, m) if m > 1 else arr.copy()

    def __len__(self):
        return self.n_frames

    def __getitem__(self, f):
        if f < 0:
            f += self.n_frames
        box = self.read_chunk(f, "configuration/box")
        box = np.array([1, 1, 1, 0, 0, 0], np.float32) if box is None else np.asarray(box, np.float32).ravel()
        dims = self.read_chunk(f, "configuration/dimensions")
        dims = 3 if dims is None else int(np.asarray(dims).ravel()[0])
        step = self.read_chunk(f, "configuration/step")
        step = 0 if step is None else int(np.asarray(step).ravel()[0])
        pos = self.read_chunk(f, "particles/position")
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        return Frame(box, dims, pos, step)

    def __iter__(self):
        for f in range(self.n_frames):
            yield self[f]


def open_gsd(path):
    return GSDFile(path)
