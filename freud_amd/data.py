"""Seeded system generators (the API of freud/data.py:21-376 plus the workload
generators of the benchmark configurations).

Positions are produced on the host with numpy; they need not be bit-identical
to the reference's generators (only the analyses' outputs on a given input are
compared).
"""

import numpy as np

import freud_amd.box


class UnitCell:
    """A lattice unit cell: box of lattice vectors + fractional basis positions."""

    def __init__(self, box, basis_positions=None):
        if basis_positions is None:
            basis_positions = [[0, 0, 0]]
        self._box = freud_amd.box.Box.from_box(box)
        self._basis_positions = np.asarray(basis_positions, dtype=np.float64)

    @property
    def box(self):
        return self._box

    @property
    def basis_positions(self):
        return self._basis_positions

    @property
    def dimensions(self):
        return self._box.dimensions

    @property
    def lattice_vectors(self):
        return np.stack([self._box.v1, self._box.v2, self._box.v3 if not self._box.is2D else np.zeros(3)])

    def generate_system(self, num_replicas=1, scale=1.0, sigma_noise=0.0, seed=None):
        """Replicate the cell ``num_replicas`` times per axis (positions grouped by
        basis site, like numpy.repeat), scale, add Gaussian noise, wrap."""
        try:
            nx, ny, nz = num_replicas
        except TypeError:
            nx = ny = num_replicas
            nz = 1 if self.box.is2D else num_replicas
        if not all(int(n) == n and n > 0 for n in (nx, ny, nz)):
            msg = "The number of replicas must be a positive integer in each dimension."
            raise ValueError(msg)
        if self.box.is2D and nz != 1:
            msg = "The number of replicas in z must be 1 for a 2D unit cell."
            raise ValueError(msg)
        b = self.box
        box = freud_amd.box.Box(b.Lx * nx * scale, b.Ly * ny * scale, b.Lz * nz * scale, b.xy, b.xz, b.yz, b.is2D)
        cells = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
        frac = (self._basis_positions[:, None, :] + cells[None, :, :]) / np.array([nx, ny, nz], dtype=np.float64)
        frac = frac.reshape(-1, 3)
        if b.is2D:
            frac[:, 2] = 0
        positions = box.make_absolute(frac.astype(np.float32)).astype(np.float64)
        if sigma_noise != 0:
            rs = np.random.RandomState(seed)
            noise = rs.normal(0.0, sigma_noise, size=positions.shape)
            if b.is2D:
                noise[:, 2] = 0
            positions = positions + noise
        positions = box.wrap(positions.astype(np.float32))
        return box, positions

    @classmethod
    def fcc(cls):
        return cls([1, 1, 1], [[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])

    @classmethod
    def bcc(cls):
        return cls([1, 1, 1], [[0, 0, 0], [0.5, 0.5, 0.5]])

    @classmethod
    def sc(cls):
        return cls([1, 1, 1])

    @classmethod
    def square(cls):
        return cls([1, 1], [[0, 0, 0]])

    @classmethod
    def hex(cls):
        return cls([1, np.sqrt(3)], [[0, 0, 0], [0.5, 0.5, 0]])


def make_random_system(box_size, num_points, is2D=False, seed=None):
    """Uniform random points in a cubic (square) box (freud/data.py:350-376)."""
    box = freud_amd.box.Box.square(box_size) if is2D else freud_amd.box.Box.cube(box_size)
    rs = np.random.RandomState(seed)
    frac = rs.random_sample((num_points, 3))
    if is2D:
        frac[:, 2] = 0
    return box, box.make_absolute(frac.astype(np.float32))


# ---------------------------------------------------------------------------
# Workload generators of the benchmark configurations (seeded, float32)


def uniform_cube(n, L, seed=0):
    """N uniform points ``((rng.random((N,3)) - 0.5) * L)`` in a cube of side L."""
    rng = np.random.default_rng(seed)
    pts = ((rng.random((n, 3)) - 0.5) * L).astype(np.float32)
    return freud_amd.box.Box.cube(L), pts


def fcc_jittered(cells, a, sigma, seed=1):
    """FCC lattice of cells^3 x 4 points (lattice constant a), Gaussian jitter
    sigma per coordinate, wrapped into the periodic cube."""
    box, pts = UnitCell.fcc().generate_system(cells, scale=a, sigma_noise=sigma, seed=seed)
    return box, np.ascontiguousarray(pts, dtype=np.float32)


def bcc_jittered(cells, a, sigma, seed=1):
    box, pts = UnitCell.bcc().generate_system(cells, scale=a, sigma_noise=sigma, seed=seed)
    return box, np.ascontiguousarray(pts, dtype=np.float32)


def triclinic_uniform(n, density, seed=2, tilt=0.3):
    """Uniform fractional coordinates in a triclinic box with tilts ~ U[-tilt, tilt]."""
    rng = np.random.default_rng(seed)
    L = (n / density) ** (1.0 / 3.0)
    xy, xz, yz = rng.uniform(-tilt, tilt, size=3)
    box = freud_amd.box.Box(L, L, L, xy, xz, yz)
    frac = rng.random((n, 3)).astype(np.float32)
    return box, np.ascontiguousarray(box.make_absolute(frac), dtype=np.float32)


def hex2d_jittered(n_target, spacing=1.0, jitter=0.05, seed=3):
    """2D hexagonal lattice of about n_target points, jitter in units of the spacing."""
    ny = int(np.sqrt(n_target / (2 * np.sqrt(3)))) or 1
    nx = int(round(n_target / (2 * ny))) or 1
    box, pts = UnitCell.hex().generate_system((nx, ny, 1), scale=spacing, sigma_noise=jitter * spacing, seed=seed)
    return box, np.ascontiguousarray(pts, dtype=np.float32)


def two_phase(n_solid_cells, rho_solid=1.05, rho_liquid=0.95, sigma=0.04, seed=5):
    """Slab A: FCC (n_solid_cells^3 x 4 points at rho_solid, jitter sigma*a);
    slab B: as many uniform-liquid points at rho_liquid with the same y-z
    cross-section, stacked along x."""
    rng = np.random.default_rng(seed)
    a = (4.0 / rho_solid) ** (1.0 / 3.0)
    n_s = 4 * n_solid_cells ** 3
    Ls = n_solid_cells * a
    Lx_liq = n_s / (rho_liquid * Ls * Ls)
    Lx = Ls + Lx_liq
    box = freud_amd.box.Box(Lx, Ls, Ls)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.stack(np.meshgrid(*[np.arange(n_solid_cells)] * 3, indexing="ij"), -1).reshape(-1, 3)
    sol = ((cells[:, None, :] + basis[None]) * a).reshape(-1, 3)
    sol += rng.normal(0.0, sigma * a, size=sol.shape)
    liq = rng.random((n_s, 3)) * np.array([Lx_liq, Ls, Ls])
    liq[:, 0] += Ls
    pts = np.concatenate([sol, liq]) - np.array([Lx, Ls, Ls]) / 2
    pts = box.wrap(pts.astype(np.float32))
    labels = np.concatenate([np.ones(n_s, np.int8), np.zeros(n_s, np.int8)])
    return box, np.ascontiguousarray(pts, dtype=np.float32), labels


_BASIS = {
    "fcc": np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]),
    "bcc": np.array([[0, 0, 0], [0.5, 0.5, 0.5]]),
}


def mixed_frames(n_frames=256, n_min=10_000, n_max=200_000, tilt=0.3, seed=4):
    """A trajectory of frames of uneven size: N per frame log-uniform in
    [n_min, n_max]; each frame is a uniform liquid (density 0.8-1.0) or a
    jittered FCC/BCC crystal (density 1.0, sigma 0.03-0.1 a), in a cubic box or
    a triclinic box with tilts uniform in [-tilt, tilt].  Returns a list of
    (box, float32 points) and a list of frame kinds."""
    rng = np.random.default_rng(seed)
    frames, kinds = [], []
    for _ in range(n_frames):
        n_target = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        kind = ("liquid", "fcc", "bcc")[int(rng.integers(3))]
        tilts = rng.uniform(-tilt, tilt, 3) if rng.random() < 0.5 else np.zeros(3)
        if kind == "liquid":
            rho = rng.uniform(0.8, 1.0)
            n = n_target
            L = (n / rho) ** (1.0 / 3.0)
            frac = rng.random((n, 3))
            noise = None
        else:
            basis = _BASIS[kind]
            m = len(basis)
            cells = max(2, int(round((n_target / m) ** (1.0 / 3.0))))
            a = (m / 1.0) ** (1.0 / 3.0)
            L = cells * a
            grid = np.stack(np.meshgrid(*[np.arange(cells)] * 3, indexing="ij"), -1).reshape(-1, 3)
            frac = ((grid[:, None, :] + basis[None]) / cells).reshape(-1, 3)
            noise = rng.normal(0.0, rng.uniform(0.03, 0.1) * a, size=frac.shape)
        box = freud_amd.box.Box(L, L, L, *tilts)
        pts = box.make_absolute(frac.astype(np.float32)).astype(np.float64)
        if noise is not None:
            pts = pts + noise
        pts = box.wrap(pts.astype(np.float32))
        frames.append((box, np.ascontiguousarray(pts, dtype=np.float32)))
        kinds.append(kind)
    return frames, kinds
