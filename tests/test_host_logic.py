"""Host-side logic that needs no GPU: the benchmark input generators
(createGridMesh numbering against the reference's pixel/voxel examples,
batched instances, serving pool and request mix), the serving cache key,
and bench.py's refusal of more than one GPU."""
import math
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from valomc_amd import meshgen  # noqa: E402


def test_batched_instances_follow_the_lattice():
    base, sources, media, photons = meshgen.batched_box_instances(seed=1234, photon_scale=1.0)
    assert base.Ne == 384000 and len(sources) == 64 and len(media) == 4
    assert photons.shape == (4, 64)
    assert photons.min() >= 0.99e6 and photons.max() <= 1.01e8
    r, BH = np.asarray(base.r), np.asarray(base.BH)
    for s in sources:
        faces = np.nonzero(s["BCType"] == ord("p"))[0]
        assert faces.tolist() == [s["face"]]
        x, y, z = s["pencil"]
        tri = r[BH[s["face"]]]
        assert np.allclose(tri[:, 2], 20.0) and z == 20.0
        # the pencil point lies inside its face's triangle (barycentric test)
        a, b, c = tri[:, :2]
        m = np.array([b - a, c - a]).T
        u, v = np.linalg.solve(m, np.array([x, y]) - a)
        assert u >= 0 and v >= 0 and u + v <= 1
        assert np.allclose(s["BCLNormal"][s["face"]], (x, y, z))
    for mua, mus in media:
        rel = mua / np.asarray(base.mua)
        assert rel.min() >= 0.9 - 1e-12 and rel.max() <= 1.1 + 1e-12
        assert not np.allclose(rel, 1.0)
    # seeded
    _, s2, m2, p2 = meshgen.batched_box_instances(seed=1234, photon_scale=1.0)
    assert np.array_equal(p2, photons) and np.array_equal(m2[0][0], media[0][0])


def test_serving_pool_sizes_and_sources():
    pool = meshgen.serving_pool(seed=2024, scale=0.02)   # shrunk for the CPU suite
    kinds = [p.notes["kind"] for p in pool]
    assert kinds == ["2d"] * 4 + ["3d"] * 4
    for p in pool:
        assert (p.BCType != ord("a")).sum() > 0
        r = np.asarray(p.r)
        # jitter moved interior nodes only: the box faces stay planar
        lo, hi = r.min(axis=0), r.max(axis=0)
        on_face = np.any(np.isclose(r, lo) | np.isclose(r, hi), axis=1)
        assert on_face.sum() > 0
    full_sizes = [(23, 2 * 23 ** 2), (316, 2 * 316 ** 2), (21, 6 * 21 ** 3), (69, 6 * 69 ** 3)]
    assert full_sizes[0][1] >= 1000 and full_sizes[1][1] <= 200000
    assert full_sizes[2][1] >= 50000 and full_sizes[3][1] <= 2000000


def test_serving_request_mix():
    reqs = meshgen.serving_requests(4000, seed=2025)
    f2d = np.mean([q["kind"] == "2d" for q in reqs])
    ffd = np.mean([q["f"] > 0 for q in reqs])
    ph = np.array([q["photons"] for q in reqs])
    assert abs(f2d - 0.4) < 0.03 and abs(ffd - 0.5) < 0.03
    assert 0.8e7 < np.median(ph) < 1.25e7
    sig = np.std(np.log(ph))
    assert 0.9 < sig < 1.1
    assert reqs == meshgen.serving_requests(4000, seed=2025)


def test_serving_cache_key_tracks_content():
    from valomc_amd.serving import content_key
    p = meshgen.square_2d(nphoton=10, seed=1)
    q = meshgen.square_2d(nphoton=10, seed=1)
    assert content_key(p) == content_key(q)
    q.mua = np.asarray(q.mua) * 1.01
    assert content_key(p) != content_key(q)


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_grid_mesh_2d_follows_the_documented_numbering():
    """createGridMesh on a 2x2 pixel grid, against the numbering drawn in
    examples/pixeltest.m (docs/pixeltest.html): vertices 1..9 with y
    ascending first, lower triangles t1..t4 of the pixels in meshgrid order,
    then the upper triangles t5..t8, every pixel split along the diagonal
    from its top-left to its bottom-right corner."""
    r, H, BH = meshgen.grid_mesh([-0.5, 0.5], [-0.5, 0.5])
    assert np.array_equal(r, [[x, y] for x in (-1, 0, 1) for y in (-1, 0, 1)])
    tris = [{1, 2, 4}, {2, 3, 5}, {4, 5, 7}, {5, 6, 8}, {2, 4, 5}, {3, 5, 6}, {5, 7, 8}, {6, 8, 9}]
    assert [set(t) for t in (H + 1).tolist()] == tris
    # the boundary runs once round the square, clockwise from the lower
    # left edge (createGridMesh.m:141-176)
    assert (BH + 1).tolist() == [[1, 2], [2, 3], [3, 6], [6, 9], [9, 8], [8, 7], [7, 4], [4, 1]]


def test_grid_mesh_2d_config_square():
    """createRectangularMesh(10, 10, 1): 121 nodes, 200 triangles, 40 edges;
    boundary rows 4-7 (1-based) are the edges of x = -5 with y in [-2, 2],
    the cosinic source of the 2D benchmark square."""
    r, H, BH = meshgen.rectangular_mesh(10, 10, 1)
    assert (r.shape[0], H.shape[0], BH.shape[0]) == (121, 200, 40)
    e = r[BH[3:7]]
    assert np.all(e[:, :, 0] == -5.0)
    assert e[:, :, 1].min() == -2.0 and e[:, :, 1].max() == 2.0
    # each element: positive area, the mesh tiles the square
    a = r[H]
    area = 0.5 * np.abs((a[:, 1, 0] - a[:, 0, 0]) * (a[:, 2, 1] - a[:, 0, 1])
                        - (a[:, 2, 0] - a[:, 0, 0]) * (a[:, 1, 1] - a[:, 0, 1]))
    assert np.allclose(area, 0.5) and math.isclose(area.sum(), 100.0)


def test_grid_mesh_3d_voxel_order():
    """createGridMesh in 3D, against examples/voxeltest.m: six tetrahedra per
    voxel; element e lies in voxel e mod nvox, voxels in meshgrid order (y
    ascending first, then x, then z); the six tetrahedra of a voxel fill
    it; the boundary triangles are the voxel faces on the outer surface,
    each a face of exactly one tetrahedron."""
    xv, yv, zv = np.arange(3) + 0.5, np.arange(4) + 0.5, np.arange(2) + 0.5
    r, H, BH = meshgen.grid_mesh(xv, yv, zv)
    nx, ny, nz = len(xv), len(yv), len(zv)
    nvox = nx * ny * nz
    assert H.shape == (6 * nvox, 4) and r.shape == ((nx + 1) * (ny + 1) * (nz + 1), 3)
    cen = r[H].mean(axis=1)
    ix, iy, iz = (np.floor(cen[:, d]).astype(int) for d in range(3))
    v = np.arange(6 * nvox) % nvox
    assert np.array_equal(iy + ny * ix + ny * nx * iz, v)
    t = r[H]
    vol = np.abs(np.einsum("ij,ij->i", t[:, 1] - t[:, 0], np.cross(t[:, 2] - t[:, 0], t[:, 3] - t[:, 0]))) / 6
    assert np.allclose(np.bincount(v, weights=vol), 1.0)
    assert BH.shape[0] == 4 * (nx * ny + ny * nz + nz * nx)
    b = r[BH]
    on_surface = np.zeros(BH.shape[0], dtype=bool)
    for d, hi in ((0, nx), (1, ny), (2, nz)):
        on_surface |= np.all(b[:, :, d] == 0, axis=1) | np.all(b[:, :, d] == hi, axis=1)
    assert on_surface.all()
    faces = {}
    for e, row in enumerate(H.tolist()):
        for f in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
            faces.setdefault(tuple(sorted(row[i] for i in f)), []).append(e)
    assert all(len(faces.get(tuple(sorted(row)), [])) == 1 for row in BH.tolist())


def test_pencil_face_of_the_box():
    """The 384k-tet box: 68,921 nodes, 19,200 boundary triangles; the pencil
    at (0.3, 0.2) on the top face lands on boundary row 3282 (0-based)."""
    p = meshgen.sphere_box_3d(nphoton=1)
    assert (p.Ne, p.Nb, p.Nr) == (384000, 19200, 68921)
    assert p.notes["pencil_face"] == 3282
    tri = np.asarray(p.r)[np.asarray(p.BH)[3282]]
    assert np.all(tri[:, 2] == 20.0)


def _unjittered(p):
    """The structured grid a jittered benchmark mesh was drawn from: the
    same element table, nodes moved back to the lattice (interior nodes
    move by at most 0.2 h, so rounding to the lattice recovers them)."""
    r = np.asarray(p.r)
    lo = r.min(axis=0)
    h = p.notes["h"]
    return lo + np.rint((r - lo) / h) * h, h


def _check_jittered(p):
    r0, h = _unjittered(p)
    r = np.asarray(p.r)
    q = meshgen.signed_measure(r, p.H) / meshgen.signed_measure(r0, p.H)
    # every element keeps its orientation (and a minimum quality)
    assert q.min() >= meshgen.JITTER_MIN_QUALITY, (p.notes, q.min(), int(np.sum(q <= 0)))
    # the offsets stay within +-0.2 h, boundary nodes do not move
    off = np.abs(r - r0)
    assert off.max() <= 0.2 * h + 1e-12
    lo, hi = r0.min(axis=0), r0.max(axis=0)
    on_box = np.any((np.abs(r0 - lo) < 1e-9) | (np.abs(r0 - hi) < 1e-9), axis=1)
    assert off[on_box].max() < 1e-9
    assert off[~on_box].max() > 0.15 * h


def test_jitter_redraws_offsets_that_invert_elements():
    """An unconditional +-0.2h draw turns a few elements inside out (57 of
    the config-3 box's tets); with the element table the offsets that would
    are redrawn, and the draw stays seeded."""
    r, H, BH = meshgen.grid_mesh(np.arange(12) + 0.5, np.arange(12) + 0.5, np.arange(12) + 0.5)
    raw = meshgen.jitter_interior(r, 1.0, 0.45, seed=3)
    assert np.sum(meshgen.signed_measure(raw, H) / meshgen.signed_measure(r, H) <= 0) > 0
    fixed = meshgen.jitter_interior(r, 1.0, 0.45, seed=3, H=H)
    q = meshgen.signed_measure(fixed, H) / meshgen.signed_measure(r, H)
    assert q.min() >= meshgen.JITTER_MIN_QUALITY
    assert np.array_equal(fixed, meshgen.jitter_interior(r, 1.0, 0.45, seed=3, H=H))
    assert np.abs(fixed - r).max() <= 0.45


def test_fd_box_and_serving_meshes_keep_their_orientation():
    p = meshgen.fd_box_3d(nphoton=1, seed=7)
    assert p.Ne == 2035200
    _check_jittered(p)
    for q in meshgen.serving_pool(seed=2024):
        _check_jittered(q)


def test_oracle_loses_no_photon_on_the_fd_box():
    """On the repaired 2M-tet box no photon leaves the mesh topology (the
    inverted tets of an unconditional jitter lost 279 of 20,000)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    p = meshgen.fd_box_3d(nphoton=20000, seed=7)
    r = pyoracle.run(p, 8, pyoracle.RNG_MT19937)
    assert r.loss == 0
    led = r.ledger
    assert abs(led.sum() - 2 * led[3] - p.Nphoton) <= 1e-9 * p.Nphoton
