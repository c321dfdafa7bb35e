"""Host-side members of the drop-in MC3D/MC2D classes (no GPU needed).

tests/host_classes.cpp fills a class from a golden case's arrays, calls
ErrorChecks() and Init() (the host pre-pass through the C ABI) and prints
the members the reference's C++ users read: the light-source list
(LightSources, LightSourcesMother, LightSourcesCDF, MC3D.hpp:155-157,
941-1002), the random stream (InitRand and the four uniform maps,
MC3D.hpp:276-303), the sharing assignment (MC3D.hpp:200-273) and the
geometry queries (MC3D.hpp:306-457, MC2D.hpp:303-390) and
BuildNeighbourhoods() on a topology alone (MC3D.hpp:835-871, as
createBH3mex calls it).  They are checked against an element-major scan
of the reference-written neighbour table, numpy cross products and
std::mt19937.
"""
import os
import subprocess

import numpy as np
import pytest

import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
SOURCES = {3: set(b"lLiIcCp"), 2: set(b"licg")}   # MC3D.hpp:957-959, MC2D.hpp:866-869
FACES = {3: [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)], 2: [(0, 1), (1, 2), (2, 0)]}


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    import valomc_amd
    valomc_amd.load_library()
    out = str(tmp_path_factory.mktemp("hc") / "host_classes")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I",
                           os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "valomc_amd", "host"),
                           os.path.join(ROOT, "tests", "host_classes.cpp"), "-o", out,
                           "-L" + os.path.join(ROOT, "valomc_amd"), "-lvalomc",
                           "-Wl,-rpath," + os.path.join(ROOT, "valomc_amd")])
    return out


def _run(exe, d, tmp):
    p = cases.from_arrays(d)
    cm = lambda a, t: np.asfortranarray(np.asarray(a, dtype=t)).ravel(order="F")
    for k, t in (("H", np.int64), ("BH", np.int64), ("r", np.float64), ("mua", np.float64), ("mus", np.float64),
                 ("g", np.float64), ("n", np.float64), ("BCn", np.float64)):
        cm(getattr(p, k), t).tofile(os.path.join(tmp, k))
    np.asarray(p.BCType, dtype=np.uint8).tofile(os.path.join(tmp, "BCType"))
    inten = np.ones(p.Nb) if p.BCIntensity is None else np.asarray(p.BCIntensity, dtype=np.float64)
    inten.tofile(os.path.join(tmp, "BCIntensity"))
    r = subprocess.run([exe, tmp, str(p.dim), str(p.Ne), str(p.Nb), str(p.Nr)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    out = {}
    for ln in r.stdout.splitlines():
        k, *v = ln.split()
        out[k] = v
    return p, inten, out


@pytest.mark.parametrize("name", ["p3_relative_cos_intensity", "p3_cosinic_hetero_n", "p3_single_cube",
                                  "p2_square_cosinic", "p2_mirror_hetero", "p2_doc_quad"])
def test_class_members(exe, tmp_path, name):
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    p, inten, out = _run(exe, d, str(tmp_path))
    dim, nv = p.dim, p.dim + 1
    r, H, BH = np.asarray(p.r), np.asarray(p.H), np.asarray(p.BH)
    hn = d["out_T1_HN"].reshape(-1, nv) if d["out_T1_HN"].ndim == 1 else d["out_T1_HN"]

    # light sources: element-major scan of the reference's neighbour table
    faces, mothers = [], []
    for el in range(p.Ne):
        for f in range(nv):
            if hn[el, f] <= -1 and int(p.BCType[-1 - hn[el, f]]) in SOURCES[dim]:
                faces.append(-1 - hn[el, f])
                mothers.append(el)
    assert int(out["nsources"][0]) == len(faces) > 0
    assert [int(x) for x in out["sources"]] == faces
    assert [int(x) for x in out["mothers"]] == mothers
    tri = r[BH[faces]]
    if dim == 3:
        size = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)
    else:
        size = np.linalg.norm(tri[:, 1] - tri[:, 0], axis=1)
    w = np.cumsum(size * inten[faces])
    assert np.allclose([float(x) for x in out["cdf"]], w / w[-1], rtol=1e-12, atol=0)

    # BuildNeighbourhoods() on topology alone (createBH3mex's use): the
    # reference's table; without BH the boundary faces stay INT_FAST64_MAX
    hn_cm = np.asfortranarray(hn).ravel(order="F")
    assert [int(x) for x in out["hn_with_bh"]] == hn_cm.tolist()
    if dim == 3:
        expect = np.where(hn_cm < 0, np.iinfo(np.int64).max, hn_cm)
        assert [int(x) for x in out["hn_topology"]] == expect.tolist()

    # host random stream, assignment
    assert out["rng_mismatch"] == ["0"]
    assert out["shared"] == ["1", "1", "1"]
    assert out["own_outputs"] == ["1", "0", str(p.Ne)]
    assert out["copied_seed"] == ["77", "1"]

    # geometry
    v = r[H]
    if dim == 3:
        vol = np.abs(np.einsum("ij,ij->i", np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), v[:, 3] - v[:, 0])) / 6
    else:
        a, b2 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
        vol = 0.5 * np.abs(a[:, 0] * b2[:, 1] - a[:, 1] * b2[:, 0])
    assert np.allclose([float(x) for x in out["volume"]], vol, rtol=1e-13)
    b = r[BH]

    def normal(pts):   # unit normal in the reference's orientation
        if dim == 3:
            c = np.cross(pts[:, 1] - pts[:, 0], pts[:, 2] - pts[:, 0])
        else:
            e = pts[:, 1] - pts[:, 0]
            c = np.stack([e[:, 1], -e[:, 0]], axis=1)
        return c / np.linalg.norm(c, axis=1, keepdims=True)

    def measure(pts):
        if dim == 3:
            return 0.5 * np.linalg.norm(np.cross(pts[:, 1] - pts[:, 0], pts[:, 2] - pts[:, 0]), axis=1)
        return np.linalg.norm(pts[:, 1] - pts[:, 0], axis=1)

    assert np.allclose([float(x) for x in out["barea"]], measure(b), rtol=1e-13)
    assert np.allclose(np.array(out["bnormal"], float).reshape(-1, dim), normal(b), atol=1e-13)
    nf = nv + 1   # the program asks for one face index past the last: zero
    fa = np.array(out["farea"], float).reshape(p.Ne, nf)
    fn = np.array(out["fnormal"], float).reshape(p.Ne, nf, dim)
    for f, idx in enumerate(FACES[dim]):
        pts = v[:, list(idx)]
        assert np.allclose(fa[:, f], measure(pts), rtol=1e-13), f
        assert np.allclose(fn[:, f], normal(pts), atol=1e-13), f
    assert np.all(fa[:, nv] == 0) and np.all(fn[:, nv] == 0)
