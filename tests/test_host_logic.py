"""Host-side logic that needs no GPU: the benchmark input generators
(batched instances, serving pool and request mix), the serving cache key,
and bench.py's refusal of more than one GPU."""
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
