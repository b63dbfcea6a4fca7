"""Reader for DCD trajectory files (CHARMM / NAMD / HOOMD layout).

Trajectory ingestion beside ``freud_amd.gsd``: the reference's reader
integration test analyses ``tests/integration/files/lj/lj.dcd`` (freud 3.5.0,
``tests/integration/test_reader_integrations.py:28,50``) through MDAnalysis,
which is not installed here, so this implements the published file layout
directly.  A DCD file is a sequence of Fortran unformatted records, each framed
by its byte length (a 4-byte integer before and after; a file whose first
marker is not 84 in either byte order is rejected):

* header: ``"CORD"`` and 20 integer control words -- [0] frames, [1] first
  step, [2] steps between frames, [8] fixed atoms, [9] time step (a float),
  [10] unit cell present, [11] fourth dimension present, [19] CHARMM version;
* title: a count and that many 80-character lines;
* atom count;
* per frame: the unit cell record when present (6 float64, in the order
  A, cos-or-angle gamma, B, cos-or-angle beta, cos-or-angle alpha, C), then
  the x, y and z records (float32 per atom), and a fourth one when present.

Angles are stored as cosines by CHARMM, NAMD >= 2.5 and HOOMD (all three
values in [-1, 1]); an angle is then 90 degrees minus arcsin of the stored
value, which is exactly 90 degrees for an orthogonal cell.  Files with fixed
atoms (per-frame coordinates of the free atoms only) are not supported.

Each frame's box is what freud makes of an MDAnalysis frame of the same file:
``NeighborQuery.from_system`` passes MDAnalysis' ``triclinic_dimensions`` (the
cell vectors as rows: a = (A, 0, 0), b = (B cos gamma, B sin gamma, 0),
c = (C cos beta, C (cos alpha - cos beta cos gamma) / sin gamma, c_z), float32)
to ``Box.from_matrix`` as it stands (``freud/locality.py:316-324``,
``freud/box.py:846-882``).  For orthogonal cells, such as HOOMD's, this is the
diagonal box.
"""

import struct

import numpy as np


class DCDFrame:
    """One frame: ``box`` (a freud_amd Box), ``positions`` (N, 3) float32, ``step``."""

    def __init__(self, box, positions, step, cell):
        self.box = box
        self.positions = positions
        self.step = step
        self.cell = cell  # (A, B, C, alpha, beta, gamma), angles in degrees

    @property
    def points(self):
        return self.positions


def cell_vectors(a, b, c, alpha, beta, gamma):
    """Rows a, b, c of the cell from lengths and angles in degrees (float64)."""
    if alpha == 90.0 and beta == 90.0 and gamma == 90.0:
        return np.diag([a, b, c]).astype(np.float64)
    ca, cb, cg = (np.cos(np.radians(t)) for t in (alpha, beta, gamma))
    sg = np.sin(np.radians(gamma))
    m = np.zeros((3, 3))
    m[0] = (a, 0.0, 0.0)
    m[1] = (b * cg, b * sg, 0.0)
    cy = c * (ca - cb * cg) / sg
    m[2] = (c * cb, cy, np.sqrt(max(c * c - (c * cb) ** 2 - cy * cy, 0.0)))
    return m


class DCDFile:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self._buf = fh.read()
        b = self._buf
        for end in ("<", ">"):
            if len(b) >= 4 and struct.unpack(end + "i", b[:4])[0] == 84:
                self._end = end
                break
        else:
            raise ValueError(f"{path} is not a DCD file (first record marker is not 84)")
        e = self._end
        pos = 0
        rec, pos = self._record(pos)
        if rec[:4] != b"CORD":
            raise ValueError(f"{path} is not a coordinate DCD file")
        ctl = struct.unpack(e + "20i", rec[4:84])
        self.n_frames_header = ctl[0]
        self.first_step = ctl[1]
        self.step_interval = ctl[2]
        self.fixed_atoms = ctl[8]
        self.charmm = ctl[19] != 0
        self.has_cell = bool(ctl[10]) if self.charmm else False
        self.has_4d = bool(ctl[11]) if self.charmm else False
        if self.fixed_atoms:
            raise ValueError("DCD files with fixed atoms are not supported")
        rec, pos = self._record(pos)
        ntitle = struct.unpack(e + "i", rec[:4])[0]
        self.title = [rec[4 + 80 * k: 4 + 80 * (k + 1)].rstrip(b"\x00 ").decode(errors="replace")
                      for k in range(ntitle)]
        rec, pos = self._record(pos)
        self.n_atoms = struct.unpack(e + "i", rec[:4])[0]
        self._first = pos
        per = 3 * (8 + 4 * self.n_atoms) + (8 + 48 if self.has_cell else 0) + (8 + 4 * self.n_atoms
                                                                                   if self.has_4d else 0)
        self._frame_bytes = per
        self.n_frames = (len(b) - pos) // per

    def _record(self, pos):
        e = self._end
        (n,) = struct.unpack(e + "i", self._buf[pos:pos + 4])
        body = self._buf[pos + 4:pos + 4 + n]
        (n2,) = struct.unpack(e + "i", self._buf[pos + 4 + n:pos + 8 + n])
        if n2 != n or len(body) != n:
            raise ValueError("corrupt DCD record")
        return body, pos + 8 + n

    def __len__(self):
        return self.n_frames

    def __getitem__(self, f):
        from freud_amd.box import Box

        if f < 0:
            f += self.n_frames
        if not 0 <= f < self.n_frames:
            raise IndexError("frame index out of range")
        pos = self._first + f * self._frame_bytes
        e = self._end
        cell = None
        if self.has_cell:
            rec, pos = self._record(pos)
            A, g, B, bt, al, C = struct.unpack(e + "6d", rec)
            angles = [g, bt, al]
            if all(-1.0 <= x <= 1.0 for x in angles):
                angles = [90.0 - np.degrees(np.arcsin(x)) for x in angles]
            g, bt, al = angles
            cell = (A, B, C, al, bt, g)
        xyz = []
        for _ in range(3):
            rec, pos = self._record(pos)
            xyz.append(np.frombuffer(rec, dtype=e + "f4", count=self.n_atoms))
        positions = np.ascontiguousarray(np.stack(xyz, axis=1), dtype=np.float32)
        if cell is None:
            raise ValueError("DCD frame without a unit cell: no box")
        box = Box.from_matrix(cell_vectors(*cell).astype(np.float32))
        step = self.first_step + f * self.step_interval
        return DCDFrame(box, positions, step, cell)

    def __iter__(self):
        for f in range(self.n_frames):
            yield self[f]


def open_dcd(path):
    return DCDFile(path)
