"""Periodic simulation box (the API of ``freud.box.Box``, freud/box.py:28-992).

The box stores float32 lengths and tilt factors like the reference's C++ Box
(freud/box/Box.h:44-606) and every coordinate transform below repeats that
file's float32 operation order, so wrapped vectors are bit-identical to the
reference's.  These are host-side utilities; the GPU kernels apply the same
arithmetic inside the pair loops.
"""

import warnings

import numpy as np

_f = np.float32


class Box:
    """Triclinic box with lengths ``Lx, Ly, Lz`` and tilts ``xy, xz, yz``."""

    def __init__(self, Lx, Ly, Lz=0, xy=0, xz=0, yz=0, is2D=None):
        if is2D is None:
            is2D = Lz == 0
        if is2D:
            if not (Lx and Ly):
                msg = "Lx and Ly must be nonzero for 2D boxes."
                raise ValueError(msg)
            if Lz != 0 or xz != 0 or yz != 0:
                warnings.warn("Specifying z-dimensions in a 2-dimensional box has no effect!", stacklevel=2)
        elif not (Lx and Ly and Lz):
            msg = "Lx, Ly, and Lz must be nonzero for 3D boxes."
            raise ValueError(msg)
        self._is2D = bool(is2D)
        self._set_L(Lx, Ly, Lz)
        self._xy = _f(xy)
        self._xz = _f(xz)
        self._yz = _f(yz)
        self._periodic = np.array([True, True, True])

    # Box.h:97-115
    def _set_L(self, Lx, Ly, Lz):
        self._L = np.array([Lx, Ly, 0 if self._is2D else Lz], dtype=_f)

    @property
    def L(self):
        return self._L.astype(np.float64)

    @L.setter
    def L(self, value):
        try:
            if len(value) != 3:
                msg = "setL must be called with a scalar or a list of length 3."
                raise ValueError(msg)
        except TypeError:
            value = (value, value, value)
        if self.is2D and value[2] != 0:
            warnings.warn("Specifying z-dimensions in a 2-dimensional box has no effect!", stacklevel=2)
        self._set_L(*value)

    @property
    def Lx(self):
        return float(self._L[0])

    @Lx.setter
    def Lx(self, value):
        self._L[0] = value

    @property
    def Ly(self):
        return float(self._L[1])

    @Ly.setter
    def Ly(self, value):
        self._L[1] = value

    @property
    def Lz(self):
        return float(self._L[2])

    @Lz.setter
    def Lz(self, value):
        if not self._is2D:
            self._L[2] = value

    @property
    def xy(self):
        return float(self._xy)

    @xy.setter
    def xy(self, value):
        self._xy = _f(value)

    @property
    def xz(self):
        return float(self._xz)

    @xz.setter
    def xz(self, value):
        self._xz = _f(value)

    @property
    def yz(self):
        return float(self._yz)

    @yz.setter
    def yz(self, value):
        self._yz = _f(value)

    def __eq__(self, other):
        if not isinstance(other, Box):
            return False
        return (
            np.array_equal(self._L, other._L)
            and self._xy == other._xy
            and self._xz == other._xz
            and self._yz == other._yz
            and self._is2D == other._is2D
            and np.array_equal(self._periodic, other._periodic)
        )

    @property
    def dimensions(self):
        return 2 if self._is2D else 3

    @dimensions.setter
    def dimensions(self, value):
        if value not in (2, 3):
            msg = "dimensions must be 2 or 3"
            raise ValueError(msg)
        self._is2D = value == 2
        if self._is2D:
            self._L[2] = 0

    @property
    def is2D(self):
        return self._is2D

    @property
    def L_inv(self):
        with np.errstate(divide="ignore"):
            inv = (_f(1) / self._L).astype(_f)
        if self._is2D:
            inv[2] = 0
        return inv.astype(np.float64)

    @property
    def volume(self):
        # Box.h:198-205
        if self._is2D:
            return float(self._L[0] * self._L[1])
        return float(self._L[0] * self._L[1] * self._L[2])

    @property
    def _lo(self):
        return -(self._L / _f(2))

    def make_absolute(self, fractional_coordinates, out=None):
        """Box.h:212-222."""
        f = np.asarray(fractional_coordinates, dtype=_f)
        single = f.ndim == 1
        f = np.atleast_2d(f)
        lo = self._lo
        v = lo + f * self._L
        v[:, 0] = v[:, 0] + (self._xy * v[:, 1] + self._xz * v[:, 2])
        v[:, 1] = v[:, 1] + self._yz * v[:, 2]
        if self._is2D:
            v[:, 2] = 0
        v = v.astype(_f)
        if out is not None:
            out[...] = v.reshape(out.shape)
            return out
        return v[0] if single else v

    def make_fractional(self, absolute_coordinates, out=None):
        """Box.h:243-255."""
        v = np.atleast_2d(np.asarray(absolute_coordinates, dtype=_f))
        single = np.asarray(absolute_coordinates).ndim == 1
        d = v - self._lo
        t = (self._xz - self._yz * self._xy) * v[:, 2] + self._xy * v[:, 1]
        d[:, 0] = d[:, 0] - t
        d[:, 1] = d[:, 1] - self._yz * v[:, 2]
        with np.errstate(divide="ignore", invalid="ignore"):
            d = (d / self._L).astype(_f)
        if self._is2D:
            d[:, 2] = 0
        if out is not None:
            out[...] = d.reshape(out.shape)
            return out
        return d[0] if single else d

    def get_images(self, vecs):
        """Box.h:262-276."""
        v = np.atleast_2d(np.asarray(vecs, dtype=_f))
        single = np.asarray(vecs).ndim == 1
        f = self.make_fractional(v) - _f(0.5)
        if self._is2D:
            f[:, 2] = 0
        img = np.where(f >= 0, f + _f(0.5), f - _f(0.5)).astype(np.int32)
        return img[0] if single else img

    def get_box_vector(self, i):
        """Box.h:503-518 lattice vector i."""
        Lx, Ly, Lz = self._L
        if i == 0:
            return np.array([Lx, 0, 0], dtype=np.float64)
        if i == 1:
            return np.array([Ly * self._xy, Ly, 0], dtype=_f).astype(np.float64)
        if i == 2 and not self._is2D:
            return np.array([Lz * self._xz, Lz * self._yz, Lz], dtype=_f).astype(np.float64)
        msg = "Box lattice vector index requested does not exist."
        raise IndexError(msg)

    @property
    def v1(self):
        return self.get_box_vector(0)

    @property
    def v2(self):
        return self.get_box_vector(1)

    @property
    def v3(self):
        return self.get_box_vector(2)

    def wrap(self, vecs, out=None):
        """Box.h:307-329 (fractional round trip with modulusPositive)."""
        v = np.atleast_2d(np.asarray(vecs, dtype=_f))
        single = np.asarray(vecs).ndim == 1
        f = self.make_fractional(v)
        one = _f(1)
        f = np.fmod(np.fmod(f, one) + one, one).astype(_f)
        w = self.make_absolute(f)
        if out is not None:
            out[...] = w.reshape(out.shape)
            return out
        return w[0] if single else w

    def unwrap(self, vecs, imgs, out=None):
        """Box.h:341-357."""
        v = np.atleast_2d(np.asarray(vecs, dtype=_f))
        single = np.asarray(vecs).ndim == 1
        im = np.atleast_2d(np.asarray(imgs)).astype(_f)
        im = np.broadcast_to(im, v.shape)
        a = self.get_box_vector(0).astype(_f)
        b = self.get_box_vector(1).astype(_f)
        r = v + a * im[:, 0:1] + b * im[:, 1:2]
        if not self._is2D:
            r = r + self.get_box_vector(2).astype(_f) * im[:, 2:3]
        r = r.astype(_f)
        if out is not None:
            out[...] = r.reshape(out.shape)
            return out
        return r[0] if single else r

    def compute_distances(self, query_points, points):
        """Box.h:395-401 |wrap(r_j - r_i)| per pair."""
        q = np.atleast_2d(np.asarray(query_points, dtype=_f))
        p = np.atleast_2d(np.asarray(points, dtype=_f))
        if len(q) != len(p):
            msg = "The number of query points and points must match."
            raise ValueError(msg)
        r = self.wrap(p - q)
        return np.sqrt((r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]) + r[:, 2] * r[:, 2]).astype(_f)

    def compute_all_distances(self, query_points, points):
        q = np.atleast_2d(np.asarray(query_points, dtype=_f))
        p = np.atleast_2d(np.asarray(points, dtype=_f))
        r = self.wrap((p[None, :, :] - q[:, None, :]).reshape(-1, 3))
        d = np.sqrt((r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]) + r[:, 2] * r[:, 2]).astype(_f)
        return d.reshape(len(q), len(p))

    def contains(self, points):
        img = np.atleast_2d(self.get_images(np.atleast_2d(points)))
        return np.all(img == 0, axis=1)

    @property
    def cubic(self):
        return (
            not self._is2D
            and self._L[0] == self._L[1] == self._L[2]
            and self._xy == self._xz == self._yz == 0
        )

    @property
    def periodic(self):
        return self._periodic.copy()

    @periodic.setter
    def periodic(self, periodic):
        p = np.broadcast_to(np.asarray(periodic, dtype=bool), (3,))
        self._periodic = p.copy()

    @property
    def periodic_x(self):
        return bool(self._periodic[0])

    @property
    def periodic_y(self):
        return bool(self._periodic[1])

    @property
    def periodic_z(self):
        return bool(self._periodic[2])

    def to_dict(self):
        return {
            "Lx": self.Lx,
            "Ly": self.Ly,
            "Lz": self.Lz,
            "xy": self.xy,
            "xz": self.xz,
            "yz": self.yz,
            "dimensions": self.dimensions,
        }

    def to_matrix(self):
        return np.asarray([[self.Lx, self.xy * self.Ly, self.xz * self.Lz],
                           [0, self.Ly, self.yz * self.Lz],
                           [0, 0, self.Lz]])

    def to_box_params(self):
        return (self.Lx, self.Ly, self.Lz, self.xy, self.xz, self.yz)

    def __repr__(self):
        return (
            f"freud_amd.box.Box(Lx={self.Lx}, Ly={self.Ly}, Lz={self.Lz}, xy={self.xy}, "
            f"xz={self.xz}, yz={self.yz}, is2D={self.is2D})"
        )

    def __str__(self):
        return repr(self)

    def __mul__(self, scale):
        if scale > 0:
            return Box(self.Lx * scale, self.Ly * scale, self.Lz * scale, self.xy, self.xz, self.yz, self.is2D)
        msg = "Box can only be multiplied by positive values."
        raise ValueError(msg)

    def __rmul__(self, scale):
        return self * scale

    @classmethod
    def from_box(cls, box, dimensions=None):
        """freud/box.py:750-843."""
        if isinstance(box, Box):
            b = Box(box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, box.is2D)
            if dimensions is not None and dimensions != b.dimensions:
                b = Box(box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, dimensions == 2)
            return b
        if np.asarray(box).shape == (3, 3):
            return cls.from_matrix(box, dimensions=dimensions)
        try:
            Lx = box.Lx
            Ly = box.Ly
            Lz = getattr(box, "Lz", 0)
            xy = getattr(box, "xy", 0)
            xz = getattr(box, "xz", 0)
            yz = getattr(box, "yz", 0)
            if dimensions is None:
                dimensions = getattr(box, "dimensions", None)
        except AttributeError:
            try:
                Lx = box["Lx"]
                Ly = box["Ly"]
                Lz = box.get("Lz", 0)
                xy = box.get("xy", 0)
                xz = box.get("xz", 0)
                yz = box.get("yz", 0)
                if dimensions is None:
                    dimensions = box.get("dimensions", None)
            except (IndexError, KeyError, TypeError) as exc:
                if len(box) not in {2, 3, 6}:
                    msg = "List-like objects must have length 2, 3, or 6 to be converted to freud.box.Box."
                    raise ValueError(msg) from exc
                Lx = box[0]
                Ly = box[1]
                Lz = box[2] if len(box) > 2 else 0
                xy, xz, yz = box[3:6] if len(box) == 6 else (0, 0, 0)
        if dimensions is None:
            dimensions = 2 if Lz == 0 else 3
        is2D = dimensions == 2
        return cls(Lx=Lx, Ly=Ly, Lz=Lz, xy=xy, xz=xz, yz=yz, is2D=is2D)

    @classmethod
    def from_matrix(cls, box_matrix, dimensions=None):
        m = np.asarray(box_matrix, dtype=np.float64)
        v0, v1, v2 = m[:, 0], m[:, 1], m[:, 2]
        Lx = np.sqrt(np.dot(v0, v0))
        a2x = np.dot(v0, v1) / Lx
        Ly = np.sqrt(np.dot(v1, v1) - a2x * a2x)
        xy = a2x / Ly
        v0xv1 = np.cross(v0, v1)
        v0xv1mag = np.sqrt(np.dot(v0xv1, v0xv1))
        Lz = np.dot(v2, v0xv1) / v0xv1mag
        if Lz != 0:
            a3x = np.dot(v0, v2) / Lx
            xz = a3x / Lz
            yz = (np.dot(v1, v2) - a2x * a3x) / (Ly * Lz)
        else:
            xz = yz = 0
        if dimensions is None:
            dimensions = 2 if Lz == 0 else 3
        return cls(Lx=Lx, Ly=Ly, Lz=Lz, xy=xy, xz=xz, yz=yz, is2D=dimensions == 2)

    @classmethod
    def cube(cls, L=None):
        if L is None:
            msg = "cube() missing 1 required positional argument: L"
            raise TypeError(msg)
        return cls(Lx=L, Ly=L, Lz=L, xy=0, xz=0, yz=0, is2D=False)

    @classmethod
    def square(cls, L=None):
        if L is None:
            msg = "square() missing 1 required positional argument: L"
            raise TypeError(msg)
        return cls(Lx=L, Ly=L, Lz=0, xy=0, xz=0, yz=0, is2D=True)

    def nearest_plane_distance(self):
        """Box.h:489-497."""
        Lx, Ly, Lz = self._L
        xy, xz, yz = self._xy, self._xz, self._yz
        t = xy * yz - xz
        dx = Lx / np.sqrt(_f(1) + xy * xy + t * t)
        dy = Ly / np.sqrt(_f(1) + yz * yz)
        return np.array([dx, dy, Lz], dtype=_f)
