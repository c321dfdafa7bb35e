"""Structured benchmark meshes and problem builders.

``grid_mesh`` restates ValoMC's ``createGridMesh`` (createGridMesh.m:67-287)
with its exact node, element and boundary ordering, returning 0-based int64
index arrays (ValoMC.m:128-130 subtracts 1 before calling the engine).  The
problem builders instantiate the benchmark configurations: homogeneous 2D
squares, the 384k-tet heterogeneous box with a pencil source, the jittered
frequency-domain box with a Gaussian-weighted source, and the batched
source/parameter grid.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np


@dataclass
class Problem:
    """Engine inputs in the reference's conventions (row = element)."""

    dim: int
    H: np.ndarray            # (Ne, dim+1) int64, 0-based
    BH: np.ndarray           # (Nb, dim) int64
    r: np.ndarray            # (Nr, dim) float64
    mua: np.ndarray
    mus: np.ndarray
    g: np.ndarray
    n: np.ndarray
    BCType: np.ndarray       # (Nb,) uint8 characters
    BCn: np.ndarray
    BCLNormal: Optional[np.ndarray] = None
    BCLDirType: Optional[np.ndarray] = None
    BCIntensity: Optional[np.ndarray] = None
    GaussianSigma: Optional[np.ndarray] = None
    HN: Optional[np.ndarray] = None
    f: float = 0.0
    phase0: float = 0.0
    Nphoton: int = 1_000_000
    seed: int = 5489
    # model constants, the reference's members MC3D::weight0, chance and c0
    # (MC3D.hpp:146,160, defaults at :178-183): roulette threshold, roulette
    # survival chance, speed of light in mm/s
    weight0: float = 0.001
    chance: float = 0.1
    c0: float = 2.99792458e11
    notes: dict = field(default_factory=dict)

    @property
    def Ne(self) -> int:
        return int(self.H.shape[0])

    @property
    def Nb(self) -> int:
        return int(self.BH.shape[0])

    @property
    def Nr(self) -> int:
        return int(self.r.shape[0])


def _nodes_vec(v: np.ndarray, h: Optional[float] = None) -> np.ndarray:
    v = np.asarray(v, dtype=np.float64).ravel()
    d = abs(v[1] - v[0]) if v.size > 1 else float(h)
    g = v - d / 2.0
    return np.concatenate([g, [g[-1] + d]])


def grid_mesh(xvec, yvec, zvec=None, h: Optional[float] = None):
    """createGridMesh(xvec, yvec[, zvec]) with 0-based indices.

    Returns (r, H, BH).  ``h`` gives the cell size for axes with a single
    centre (createGridMesh itself needs two).  Node numbering follows MATLAB's meshgrid + (:)
    column-major flattening: y fastest, then x, then z
    (createGridMesh.m:101-102, 195-196, 201).
    """
    gx = _nodes_vec(xvec, h)
    gy = _nodes_vec(yvec, h)
    if zvec is None:
        nxn, nyn = gx.size, gy.size
        X, Y = np.meshgrid(gx, gy)                    # (nyn, nxn)
        r = np.stack([X.ravel(order="F"), Y.ravel(order="F")], axis=1)
        nxv, nyv = nxn - 1, nyn - 1
        ysize, xsize = nyv + 1, nxv + 1
        i = np.repeat(np.arange(1, nxv + 1), nyv)     # x voxel, outer loop
        j = np.tile(np.arange(1, nyv + 1), nxv)       # y voxel, inner loop
        k = (i - 1) * ysize + j                       # 1-based lower-left node
        t1 = np.stack([k, k + 1, k + ysize], axis=1)               # createGridMesh.m:122-130
        t2 = np.stack([k + 1, k + ysize + 1, k + ysize], axis=1)   # createGridMesh.m:131-139
        H = np.concatenate([t1, t2]) - 1
        jy = np.arange(1, nyv + 1)
        jx = np.arange(1, nxv + 1)
        b1 = np.stack([jy, jy + 1], axis=1)                                    # BL -> TL
        b2 = np.stack([ysize * jx, ysize * (jx + 1)], axis=1)                  # top row
        b3 = np.stack([ysize * xsize - (jy - 1), ysize * xsize - jy], axis=1)  # right column, down
        b4 = np.stack([ysize * (xsize - 1) + 1 - ysize * (jx - 1),
                       ysize * (xsize - 1) + 1 - ysize * jx], axis=1)          # bottom row, leftwards
        BH = np.concatenate([b1, b2, b3, b4]) - 1
        return r, H.astype(np.int64), BH.astype(np.int64)

    gz = _nodes_vec(zvec, h)
    nxn, nyn, nzn = gx.size, gy.size, gz.size
    # meshgrid(gx, gy, gz) has shape (nyn, nxn, nzn); (:) is column-major
    X, Y, Z = np.meshgrid(gx, gy, gz)
    r = np.stack([X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")], axis=1)

    def ijk(i, j, k):  # createGridMesh.m:206, 1-based
        return (k - 1) * (nxn * nyn) + (i - 1) * nyn + j

    points = np.array([[0, 0, 0], [0, 1, 0], [1, 0, 0], [1, 1, 0],
                       [0, 0, 1], [0, 1, 1], [1, 0, 1], [1, 1, 1]])
    conn = np.array([[5, 1, 2, 3], [6, 5, 2, 3], [6, 7, 5, 3],
                     [6, 4, 7, 3], [6, 2, 4, 3], [6, 8, 7, 4]]) - 1
    nxv, nyv, nzv = nxn - 1, nyn - 1, nzn - 1
    # loop order t, k, i, j (createGridMesh.m:213-227): j fastest
    kk, ii, jj = np.meshgrid(np.arange(1, nzv + 1), np.arange(1, nxv + 1),
                             np.arange(1, nyv + 1), indexing="ij")
    kk, ii, jj = kk.ravel(), ii.ravel(), jj.ravel()
    blocks = []
    for t in range(6):
        off = points[conn[t]]
        cols = [ijk(ii + off[c, 0], jj + off[c, 1], kk + off[c, 2]) for c in range(4)]
        blocks.append(np.stack(cols, axis=1))
    H = np.concatenate(blocks) - 1

    rows = []
    # XY planes (createGridMesh.m:233-247): for i, for j -> 4 rows each
    I, J = np.meshgrid(np.arange(1, nxv + 1), np.arange(1, nyv + 1), indexing="ij")
    I, J = I.ravel(), J.ravel()
    k0, k1 = np.ones_like(I), np.full_like(I, nzv + 1)
    xy = np.stack([
        np.stack([ijk(I, J, k0), ijk(I + 1, J, k0), ijk(I, J + 1, k0)], 1),
        np.stack([ijk(I + 1, J + 1, k0), ijk(I, J + 1, k0), ijk(I + 1, J, k0)], 1),
        np.stack([ijk(I, J, k1), ijk(I + 1, J, k1), ijk(I, J + 1, k1)], 1),
        np.stack([ijk(I + 1, J + 1, k1), ijk(I, J + 1, k1), ijk(I + 1, J, k1)], 1),
    ], axis=1).reshape(-1, 3)
    rows.append(xy)
    # YZ planes (:250-264): for j, for k
    J, K = np.meshgrid(np.arange(1, nyv + 1), np.arange(1, nzv + 1), indexing="ij")
    J, K = J.ravel(), K.ravel()
    i0, i1 = np.ones_like(J), np.full_like(J, nxv + 1)
    yz = np.stack([
        np.stack([ijk(i0, J, K), ijk(i0, J + 1, K), ijk(i0, J, K + 1)], 1),
        np.stack([ijk(i0, J + 1, K + 1), ijk(i0, J, K + 1), ijk(i0, J + 1, K)], 1),
        np.stack([ijk(i1, J, K), ijk(i1, J + 1, K), ijk(i1, J, K + 1)], 1),
        np.stack([ijk(i1, J + 1, K + 1), ijk(i1, J, K + 1), ijk(i1, J + 1, K)], 1),
    ], axis=1).reshape(-1, 3)
    rows.append(yz)
    # ZX planes (:267-281): for k, for i
    K, I = np.meshgrid(np.arange(1, nzv + 1), np.arange(1, nxv + 1), indexing="ij")
    K, I = K.ravel(), I.ravel()
    j0, j1 = np.ones_like(K), np.full_like(K, nyv + 1)
    zx = np.stack([
        np.stack([ijk(I, j0, K), ijk(I, j0, K + 1), ijk(I + 1, j0, K)], 1),
        np.stack([ijk(I + 1, j0, K + 1), ijk(I + 1, j0, K), ijk(I, j0, K + 1)], 1),
        np.stack([ijk(I, j1, K), ijk(I, j1, K + 1), ijk(I + 1, j1, K)], 1),
        np.stack([ijk(I + 1, j1, K + 1), ijk(I + 1, j1, K), ijk(I, j1, K + 1)], 1),
    ], axis=1).reshape(-1, 3)
    rows.append(zx)
    BH = np.concatenate(rows) - 1
    return r, H.astype(np.int64), BH.astype(np.int64)


def rectangular_mesh(xsize: float, ysize: float, dh: float):
    """createRectangularMesh (createRectangularMesh.m:28)."""
    xv = np.arange(-xsize / 2 + dh / 2, xsize / 2 - dh / 2 + dh / 4, dh)
    yv = np.arange(-ysize / 2 + dh / 2, ysize / 2 - dh / 2 + dh / 4, dh)
    return grid_mesh(xv, yv)


def _chars(s: str, n: int) -> np.ndarray:
    return np.full(n, ord(s), dtype=np.uint8)


def homogeneous(dim, r, H, BH, mua, mus, g, n, n_out=1.0) -> Problem:
    Ne, Nb = H.shape[0], BH.shape[0]
    return Problem(dim=dim, H=H, BH=BH, r=r,
                   mua=np.full(Ne, float(mua)), mus=np.full(Ne, float(mus)),
                   g=np.full(Ne, float(g)), n=np.full(Ne, float(n)),
                   BCType=_chars("a", Nb), BCn=np.full(Nb, float(n_out)))


# ---------------------------------------------------------------------------
# Benchmark configurations
# ---------------------------------------------------------------------------

def square_2d(nphoton: int = 1_000_000, seed: int = 1) -> Problem:
    """10x10 mm square, 1 mm grid split in 2 triangles (200 elements, 40
    edges); mua=0.01, mus=1, g=0.9, n=1.3, n=1.0 outside; cosinic source on
    boundary edges 4-7 (1-based), CW."""
    r, H, BH = rectangular_mesh(10, 10, 1)
    p = homogeneous(2, r, H, BH, 0.01, 1.0, 0.9, 1.3, 1.0)
    p.BCType[3:7] = ord("c")
    p.Nphoton, p.seed = nphoton, seed
    return p


def sphere_box_3d(nphoton: int = 100_000_000, seed: int = 1234, side_mm: int = 40,
                  h: float = 1.0, pencil=(0.3, 0.2)) -> Problem:
    """40x40x40 mm box on a 1 mm grid, 6 tets per cube (384,000 tets).
    Background mua=0.01, mus=1, g=0.9, n=1.4, n=1.0 outside.  A 5 mm sphere
    centred at default_rng(seed).uniform(-12, 12, 3) has mua=0.05, mus=2
    (membership by tet centroid).  Pencil source at (x, y) on the top face
    z=+side/2, offset off the central vertex, CW."""
    half = side_mm / 2.0
    v = np.arange(-half + h / 2, half - h / 2 + h / 4, h)
    r, H, BH = grid_mesh(v, v, v)
    p = homogeneous(3, r, H, BH, 0.01, 1.0, 0.9, 1.4, 1.0)
    rng = np.random.default_rng(seed)
    centre = rng.uniform(-12.0, 12.0, 3)
    cen = r[H].mean(axis=1)
    inside = np.linalg.norm(cen - centre, axis=1) <= 5.0
    p.mua[inside] = 0.05
    p.mus[inside] = 2.0
    px, py = pencil
    ib = locate_top_triangle(r, BH, px, py, half)
    p.BCType[ib] = ord("p")
    p.BCLDirType = _chars("n", p.Nb)
    p.BCLDirType[ib] = ord("p")
    p.BCLNormal = np.zeros((p.Nb, 3))
    p.BCLNormal[ib] = (px, py, half)
    p.Nphoton, p.seed = nphoton, seed
    p.notes.update(sphere_centre=centre.tolist(), sphere_elements=int(inside.sum()),
                   pencil_face=int(ib), pencil=(px, py, half))
    return p


def batched_box_instances(seed: int = 1234, photon_scale: float = 1.0):
    """Batched forward solves on the 384k-tet box of ``sphere_box_3d``:
    64 pencil positions on an 8x8 lattice of top-face points (5 mm cell
    centres offset by (0.3, 0.2) mm off the vertices) x 4 optical-parameter
    sets (mua and mus of every element times 1 + U(-0.1, 0.1), seeded) = 256
    instances, photon counts log-uniform in [1e6, 1e8] (seeded) times
    ``photon_scale``.

    Returns (base problem, sources, media, photons): ``sources`` is a list of
    64 dicts of BCType / BCLDirType / BCLNormal overrides, ``media`` a list
    of 4 (mua, mus) pairs and ``photons`` an int array [4, 64]."""
    base = sphere_box_3d(nphoton=1, seed=seed)
    half = 20.0
    rng = np.random.default_rng(seed + 1)
    sources = []
    for i in range(8):
        for j in range(8):
            x, y = -17.5 + 5.0 * i + 0.3, -17.5 + 5.0 * j + 0.2
            r = np.asarray(base.r)
            BH = np.asarray(base.BH)
            ib = locate_top_triangle(r, BH, x, y, half)
            bct = np.full(base.Nb, ord("a"), dtype=np.uint8)
            bct[ib] = ord("p")
            ldt = np.full(base.Nb, ord("n"), dtype=np.uint8)
            ldt[ib] = ord("p")
            bcl = np.zeros((base.Nb, 3))
            bcl[ib] = (x, y, half)
            sources.append(dict(BCType=bct, BCLDirType=ldt, BCLNormal=bcl, pencil=(x, y, half), face=int(ib)))
    media = []
    for _ in range(4):
        mua = np.asarray(base.mua) * (1.0 + rng.uniform(-0.1, 0.1, base.Ne))
        mus = np.asarray(base.mus) * (1.0 + rng.uniform(-0.1, 0.1, base.Ne))
        media.append((mua, mus))
    photons = np.round(10.0 ** rng.uniform(6.0, 8.0, (4, 64)) * photon_scale).astype(np.int64)
    return base, sources, media, photons


def locate_top_triangle(r, BH, x, y, ztop) -> int:
    """Boundary triangle on the plane z=ztop that contains (x, y)."""
    tri = r[BH]                                     # (Nb, 3, 3)
    on_top = np.all(np.abs(tri[:, :, 2] - ztop) < 1e-9, axis=1)
    a, b, c = tri[:, 0, :2], tri[:, 1, :2], tri[:, 2, :2]
    p = np.array([x, y])

    def cross(u, v):
        return u[:, 0] * v[:, 1] - u[:, 1] * v[:, 0]
    d1 = cross(b - a, p - a)
    d2 = cross(c - b, p - b)
    d3 = cross(a - c, p - c)
    inside = ((d1 >= 0) & (d2 >= 0) & (d3 >= 0)) | ((d1 <= 0) & (d2 <= 0) & (d3 <= 0))
    idx = np.nonzero(on_top & inside)[0]
    if idx.size == 0:
        raise ValueError("point not on the top face")
    return int(idx[0])


def signed_measure(r: np.ndarray, H: np.ndarray) -> np.ndarray:
    """Signed volume (tetrahedra) or signed area (triangles) of every
    element, in the node order of its row of H."""
    H = np.asarray(H)
    a = r[H[:, 0]]
    if H.shape[1] == 4:
        b, c, d = r[H[:, 1]] - a, r[H[:, 2]] - a, r[H[:, 3]] - a
        return np.einsum("ij,ij->i", np.cross(b, c), d) / 6.0
    b, c = r[H[:, 1]] - a, r[H[:, 2]] - a
    return (b[:, 0] * c[:, 1] - b[:, 1] * c[:, 0]) / 2.0


# an element keeps at least this fraction of its unjittered signed measure
JITTER_MIN_QUALITY = 0.05


def jitter_interior(r: np.ndarray, h: float, amp: float, seed: int, H: Optional[np.ndarray] = None,
                    min_quality: float = JITTER_MIN_QUALITY, max_rounds: int = 200) -> np.ndarray:
    """Move interior nodes by U(-amp*h, amp*h) per axis; boundary nodes stay
    so the box faces remain planar.

    With the element table ``H``, offsets that would turn an element inside
    out are redrawn: after the first draw, every interior node of an element
    whose signed volume (2D: area) fell below ``min_quality`` times its
    unjittered value gets a fresh offset from the same seeded stream, until
    no such element is left.  The bound stays +-amp*h and the result is a
    conforming mesh with every element of its original orientation.  (At
    +-0.2h an unconditional draw inverts 57 of the 2,035,200 tets of the
    60x60x40 mm box at h = 0.75 mm.)"""
    rng = np.random.default_rng(seed)
    lo, hi = r.min(axis=0), r.max(axis=0)
    interior = np.all((r > lo + 1e-9) & (r < hi - 1e-9), axis=1)
    out = r.copy()
    out[interior] += rng.uniform(-amp * h, amp * h, size=(int(interior.sum()), r.shape[1]))
    if H is None:
        return out
    H = np.asarray(H)
    v0 = signed_measure(r, H)
    bad = np.nonzero(signed_measure(out, H) / v0 < min_quality)[0]
    for _ in range(max_rounds):
        if bad.size == 0:
            return out
        nodes = np.unique(H[bad].ravel())
        nodes = nodes[interior[nodes]]
        out[nodes] = r[nodes] + rng.uniform(-amp * h, amp * h, size=(nodes.size, r.shape[1]))
        # only elements that share a moved node can have changed
        moved = np.zeros(r.shape[0], dtype=bool)
        moved[nodes] = True
        touch = np.nonzero(moved[H].any(axis=1))[0]
        q = signed_measure(out, H[touch]) / v0[touch]
        bad = touch[q < min_quality]
    raise RuntimeError(f"jitter_interior: {bad.size} elements still below quality after {max_rounds} redraws")


def smooth_field(cen: np.ndarray, corr_mm: float, seed: int, grid_h: float = 1.0) -> np.ndarray:
    """Unit-variance Gaussian random field at the points `cen`: white noise
    on a regular grid, Gaussian-smoothed with sigma=corr_mm (FFT), sampled at
    the nearest grid node."""
    rng = np.random.default_rng(seed)
    lo = cen.min(axis=0) - 3 * corr_mm
    hi = cen.max(axis=0) + 3 * corr_mm
    shape = tuple(int(math.ceil((hi[d] - lo[d]) / grid_h)) + 1 for d in range(cen.shape[1]))
    noise = rng.standard_normal(shape)
    F = np.fft.rfftn(noise)
    freqs = [np.fft.fftfreq(s, d=grid_h) for s in shape[:-1]] + [np.fft.rfftfreq(shape[-1], d=grid_h)]
    K2 = sum(np.meshgrid(*[f ** 2 for f in freqs], indexing="ij"))
    field = np.fft.irfftn(F * np.exp(-2 * (math.pi ** 2) * (corr_mm ** 2) * K2), s=shape,
                          axes=tuple(range(len(shape))))
    field = (field - field.mean()) / field.std()
    idx = np.clip(np.rint((cen - lo) / grid_h).astype(np.int64), 0, np.array(shape) - 1)
    return field[tuple(idx[:, d] for d in range(cen.shape[1]))]


def lognormal_from_field(z: np.ndarray, mean: float, cv: float) -> np.ndarray:
    s = math.sqrt(math.log(1.0 + cv * cv))
    return mean * np.exp(s * z - 0.5 * s * s)


def fd_box_3d(nphoton: int = 300_000_000, seed: int = 7, h: float = 0.75,
              size=(60.0, 60.0, 40.0), corr_mm: float = 5.0, sigma_mm: float = 2.0,
              freq_hz: float = 100e6) -> Problem:
    """60x60x40 mm box at h=0.75 mm (80x80x53 cells, 2,035,200 tets),
    interior nodes jittered by U(+-0.2h), with offsets that would invert a
    tet redrawn (``jitter_interior``); per-element mua/mus lognormal from
    a smoothed Gaussian field (5 mm correlation); g=0.9, n=1.37, n=1.0
    outside; 100 MHz.  The Gaussian-profile source is direct ('l') light on
    the top face with BCIntensity = exp(-rho^2/2sigma^2), truncated at 4 sigma."""
    nx, ny = int(round(size[0] / h)), int(round(size[1] / h))
    nz = int(size[2] // h)
    xv = (np.arange(nx) + 0.5) * h - size[0] / 2
    yv = (np.arange(ny) + 0.5) * h - size[1] / 2
    zv = (np.arange(nz) + 0.5) * h - nz * h / 2
    r, H, BH = grid_mesh(xv, yv, zv)
    r = jitter_interior(r, h, 0.2, seed, H)
    cen = r[H].mean(axis=1)
    mua = lognormal_from_field(smooth_field(cen, corr_mm, seed + 1), 0.02, 0.30)
    mus = lognormal_from_field(smooth_field(cen, corr_mm, seed + 2), 1.5, 0.20)
    Ne, Nb = H.shape[0], BH.shape[0]
    p = Problem(dim=3, H=H, BH=BH, r=r, mua=mua, mus=mus, g=np.full(Ne, 0.9),
                n=np.full(Ne, 1.37), BCType=_chars("a", Nb), BCn=np.full(Nb, 1.0))
    ztop = r[:, 2].max()
    tri = r[BH]
    top = np.all(np.abs(tri[:, :, 2] - ztop) < 1e-9, axis=1)
    rho = np.linalg.norm(tri[:, :, :2].mean(axis=1), axis=1)
    src = top & (rho <= 4 * sigma_mm)
    p.BCType[src] = ord("l")
    p.BCIntensity = np.where(src, np.exp(-rho ** 2 / (2 * sigma_mm ** 2)), 0.0)
    p.f, p.Nphoton, p.seed = freq_hz, nphoton, seed
    p.notes.update(cells=(nx, ny, nz), h=h, corr_mm=corr_mm, sigma_mm=sigma_mm,
                   source_faces=int(src.sum()))
    return p


def serving_pool(seed: int = 2024, scale: float = 1.0):
    """Seeded pool of structured-then-jittered meshes for the serving mix:
    four 2D squares of about 1k-200k triangles and four 3D boxes of about
    50k-2M tetrahedra (interior nodes moved by U(+-0.2h), offsets that
    would invert an element redrawn), each with a
    lognormal smooth-field medium and a boundary source (2D: cosinic edges
    on the left side; 3D: Gaussian-profile direct light on the top face,
    sigma = 1/10 of the side).  ``scale`` < 1 shrinks every mesh (cell counts
    times scale^(1/dim)) for quick runs.  Returns a list of Problems with
    ``notes['kind']`` = '2d' or '3d'."""
    pool = []
    for k, n in enumerate((23, 50, 110, 316)):
        n = max(4, int(round(n * scale ** 0.5)))
        side = 10.0
        h = side / n
        v = (np.arange(n) + 0.5) * h - side / 2
        r, H, BH = grid_mesh(v, v, h=h)
        r = jitter_interior(r, h, 0.2, seed + k, H)
        cen = r[H].mean(axis=1)
        mua = lognormal_from_field(smooth_field(cen, 1.0, seed + 10 + k, grid_h=min(0.5, h)), 0.02, 0.3)
        mus = lognormal_from_field(smooth_field(cen, 1.0, seed + 20 + k, grid_h=min(0.5, h)), 1.0, 0.2)
        Ne, Nb = H.shape[0], BH.shape[0]
        p = Problem(dim=2, H=H, BH=BH, r=r, mua=mua, mus=mus, g=np.full(Ne, 0.9), n=np.full(Ne, 1.33),
                    BCType=_chars("a", Nb), BCn=np.full(Nb, 1.0))
        seg = r[BH]
        left = np.all(np.abs(seg[:, :, 0] - r[:, 0].min()) < 1e-9, axis=1) & (np.abs(seg[:, :, 1].mean(axis=1)) < 2.0)
        p.BCType[left] = ord("c")
        p.notes.update(kind="2d", cells=n, h=h)
        pool.append(p)
    for k, n in enumerate((21, 33, 52, 69)):
        n = max(3, int(round(n * scale ** (1.0 / 3.0))))
        side = 30.0
        h = side / n
        v = (np.arange(n) + 0.5) * h - side / 2
        r, H, BH = grid_mesh(v, v, v, h=h)
        r = jitter_interior(r, h, 0.2, seed + 100 + k, H)
        cen = r[H].mean(axis=1)
        gh = max(h, 0.5)
        mua = lognormal_from_field(smooth_field(cen, 3.0, seed + 110 + k, grid_h=gh), 0.02, 0.3)
        mus = lognormal_from_field(smooth_field(cen, 3.0, seed + 120 + k, grid_h=gh), 1.5, 0.2)
        Ne, Nb = H.shape[0], BH.shape[0]
        p = Problem(dim=3, H=H, BH=BH, r=r, mua=mua, mus=mus, g=np.full(Ne, 0.9), n=np.full(Ne, 1.37),
                    BCType=_chars("a", Nb), BCn=np.full(Nb, 1.0))
        ztop = r[:, 2].max()
        tri = r[BH]
        top = np.all(np.abs(tri[:, :, 2] - ztop) < 1e-9, axis=1)
        rho = np.linalg.norm(tri[:, :, :2].mean(axis=1), axis=1)
        sig = side / 10.0
        src = top & (rho <= 4 * sig)
        p.BCType[src] = ord("l")
        p.BCIntensity = np.where(src, np.exp(-rho ** 2 / (2 * sig ** 2)), 0.0)
        p.notes.update(kind="3d", cells=n, h=h)
        pool.append(p)
    return pool


def serving_requests(n: int, seed: int = 2025, median: float = 1e7, sigma: float = 1.0,
                     p2d: float = 0.4, p_fd: float = 0.5, freq_hz: float = 100e6):
    """Seeded request mix for the serving loop: mesh kind 2D with probability
    p2d (else 3D), a mesh of that kind drawn uniformly from the pool,
    photons lognormal (median, sigma) clipped to [1e4, 1e9], CW or ``freq_hz``
    with probability p_fd.  Returns a list of dicts (kind, index, photons, f,
    seed); arrival times are drawn separately for a chosen rate."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        kind = "2d" if rng.random() < p2d else "3d"
        idx = int(rng.integers(0, 4))
        ph = int(np.clip(np.exp(rng.normal(math.log(median), sigma)), 1e4, 1e9))
        f = freq_hz if rng.random() < p_fd else 0.0
        out.append(dict(kind=kind, index=idx, photons=ph, f=f, seed=1 + i))
    return out
