"""The benchmark configurations on the GPU, against the oracle.

Each test runs a configuration's own workload (its mesh at full size, its
medium and sources), with photon counts the CPU oracle finishes in seconds:

* the 384k-tet heterogeneous box (pencil source, CW): FP64 on the device
  photon for photon against the Philox-mode oracle across the whole
  Morton-reordered mesh; FP32 against FP64 on the same streams;
* the repaired 2,035,200-tet jittered box (smooth-field medium, Gaussian
  source, 100 MHz): the same on ER and EI, and no photon lost;
* the 256-instance batch (64 pencil positions x 4 parameter sets): every
  item bit-equal to its solve alone (fixed point), and batch items against
  the oracle run with the item's seed and stream tag;
* the serving mix (2D and 3D jittered meshes, CW and 100 MHz): every
  request served in the mix bit-equal to its solve alone, and served
  requests against the oracle;
* the box and the FD box against the reference's own MT19937 streams:
  z-scores of spatially binned tallies over independent seeds.
"""
import numpy as np
import pytest

import pyoracle
from valomc_amd import meshgen

pytestmark = pytest.mark.gpu

OUT = ("ER", "EI", "EBR", "EBI", "DEBR", "DEBI")


@pytest.fixture(scope="module")
def vmc():
    import valomc_amd
    print("device:", valomc_amd.device_info(), flush=True)
    return valomc_amd


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a - b)


def _match_oracle(res, ref, tol=1e-9, keys=OUT):
    for k in keys:
        a, b = getattr(res, k), getattr(ref, k)
        if np.linalg.norm(b) == 0:
            assert np.max(np.abs(a)) < 1e-12, k
        else:
            assert rel_l2(a, b) < tol, (k, rel_l2(a, b))


def _same_bits(a, b):
    for k in OUT:
        assert np.array_equal(getattr(a, k), getattr(b, k), equal_nan=True), k


@pytest.fixture(scope="module")
def box(vmc):
    p = meshgen.sphere_box_3d(nphoton=20000, seed=1234)
    with vmc.Problem(p) as prob:
        yield p, prob


def test_box_fp64_matches_oracle_photon_for_photon(vmc, box):
    """The 384k-tet box at 2e4 photons: the device follows every photon
    through the Morton-ordered records the way the oracle follows it
    through the reference's element numbering."""
    p, prob = box
    print(f"box Ne={p.Ne} N={p.Nphoton}", flush=True)
    res = prob.run(p.Nphoton, seed=p.seed, precision=vmc.FP64)
    ref = pyoracle.run(p, 16, pyoracle.RNG_PHILOX)
    _match_oracle(res, ref)
    assert res.loss == ref.loss == 0
    led = res.ledger
    assert abs(led[0] + led[1] + led[2] - led[3] + led[4] - p.Nphoton) <= 1e-9 * p.Nphoton


def test_box_device_slots(vmc, box):
    """vmc_problem_slots: whole cells per tally sector, so the 40^3-cube box
    takes 8 slots per cube of six tetrahedra (4 per 3); the serving cache
    sizes resident meshes from it."""
    from valomc_amd import serving
    p, prob = box
    assert prob.slots == 64000 * 8 == p.Ne * 4 // 3
    assert serving.device_bytes(p) == serving.device_bytes(p, prob.slots)


def test_box_fp32_within_one_percent_of_fp64(vmc, box):
    p, prob = box
    a = prob.run(2_000_000, seed=5, precision=vmc.FP64)
    b = prob.run(2_000_000, seed=5, precision=vmc.FP32)
    for k in ("ER", "EBR"):
        assert rel_l2(getattr(b, k), getattr(a, k)) < 1e-2, (k, rel_l2(getattr(b, k), getattr(a, k)))


def test_fd_box_fp64_matches_oracle_and_loses_nothing(vmc):
    """The repaired 2M-tet frequency-domain box: ER and EI photon for photon
    against the oracle, FP32 within 1 % of FP64, and no photon leaves the
    mesh topology."""
    p = meshgen.fd_box_3d(nphoton=5000, seed=7)
    print(f"fd box Ne={p.Ne} N={p.Nphoton}", flush=True)
    with vmc.Problem(p) as prob:
        res = prob.run(p.Nphoton, seed=p.seed, f=p.f, precision=vmc.FP64)
        ref = pyoracle.run(p, 16, pyoracle.RNG_PHILOX)
        _match_oracle(res, ref)
        assert res.loss == ref.loss == 0
        a = prob.run(1_000_000, seed=3, f=p.f, precision=vmc.FP64)
        b = prob.run(1_000_000, seed=3, f=p.f, precision=vmc.FP32)
    assert a.loss == 0 and b.loss == 0
    for k in ("ER", "EI", "EBR", "EBI"):
        assert rel_l2(getattr(b, k), getattr(a, k)) < 1e-2, (k, rel_l2(getattr(b, k), getattr(a, k)))


def _spatial_bins(p, size):
    """Element -> index of the cube of side `size` (mm) holding its centroid."""
    r, H = np.asarray(p.r), np.asarray(p.H)
    q = np.floor((r[H].mean(axis=1) - r.min(axis=0)) / size).astype(np.int64)
    dims = q.max(axis=0) + 1
    return (q[:, 0] * dims[1] + q[:, 1]) * dims[2] + q[:, 2], int(np.prod(dims))


def _binned_z(g, r, pilot, bins, nbins, K):
    """z of the difference of the means of K runs a side, per bin of summed
    tallies, over the bins a separate pilot run finds above 1e-3 of its
    largest (choosing bins by the tested runs' own spread would favour bins
    whose samples happened to agree).  Returns z and the Bonferroni bound
    at 1e-3 for the largest |z| under a t distribution with 2(K-1) degrees
    of freedom (the standard errors come from K runs a side)."""
    from scipy import stats
    G = np.array([np.bincount(bins, x, nbins) for x in g])
    Rr = np.array([np.bincount(bins, x, nbins) for x in r])
    pil = np.abs(np.bincount(bins, pilot, nbins))
    ok = pil > 1e-3 * pil.max()
    se = np.sqrt(G.var(0, ddof=1) / K + Rr.var(0, ddof=1) / K)
    z = (G.mean(0)[ok] - Rr.mean(0)[ok]) / se[ok]
    nu = 2 * (K - 1)
    return z, stats.t.ppf(1.0 - 1e-3 / (2 * max(ok.sum(), 1)), nu), nu


def _check_z(tag, z, zmax, nu, nmin):
    print(f"{tag}: {z.size} bins, mean z^2 {np.mean(z ** 2):.3f} (t: {nu / (nu - 2):.3f}), "
          f"max |z| {np.abs(z).max():.2f} (bound {zmax:.2f})", flush=True)
    assert z.size >= nmin, (tag, z.size)
    assert 0.6 < np.mean(z ** 2) < 1.8, (tag, np.mean(z ** 2))
    assert np.abs(z).max() < zmax, (tag, np.abs(z).max())


def test_box_z_scores_against_reference_streams(vmc, box):
    """The full 384k-tet box against the reference's own MT19937 streams
    (the oracle's MT mode reproduces the reference bit for bit): K seeds a
    side; element fluence summed over 2 mm cubes (48 tets each) and exitance
    per boundary face must differ by standard-normal-looking z.  A bias of
    0.5 % in the fluence gives |z| above 20 here."""
    import copy
    p, prob = box
    K, n = 8, 200000
    g = [prob.run(n, seed=2000 + k, precision=vmc.FP64) for k in range(K)]
    pilot = prob.run(n, seed=999, precision=vmc.FP64)
    r = []
    for k in range(K):
        q = copy.copy(p)
        q.Nphoton, q.seed = n, 7000 + 32 * k   # thread t of an MT run uses seed + t: no two runs share a stream
        r.append(pyoracle.run(q, 16, pyoracle.RNG_MT19937))
    bins, nb = _spatial_bins(p, 2.0)
    _check_z("box ER", *_binned_z([x.ER for x in g], [x.ER for x in r], pilot.ER, bins, nb, K), 2000)
    _check_z("box EBR", *_binned_z([x.EBR for x in g], [x.EBR for x in r], pilot.EBR, np.arange(p.Nb), p.Nb, K), 2000)


def test_fd_box_z_scores_against_reference_streams(vmc):
    """The 2M-tet frequency-domain box: ER and EI summed over 3 mm cubes
    against the reference's MT19937 streams, as for the box."""
    p = meshgen.fd_box_3d(nphoton=1, seed=7)
    K, n = 6, 20000
    with vmc.Problem(p) as prob:
        g = [prob.run(n, seed=3000 + k, f=p.f, precision=vmc.FP64) for k in range(K)]
        pilot = prob.run(n, seed=999, f=p.f, precision=vmc.FP64)
    r = []
    for k in range(K):
        p.Nphoton, p.seed = n, 9000 + 32 * k
        r.append(pyoracle.run(p, 16, pyoracle.RNG_MT19937))
    bins, nb = _spatial_bins(p, 3.0)
    for key in ("ER", "EI"):
        _check_z(f"fd box {key}", *_binned_z([getattr(x, key) for x in g], [getattr(x, key) for x in r],
                                              getattr(pilot, key), bins, nb, K), 1000)


def _instance_problem(base, source, medium):
    """The oracle's input for one batch instance: the base box with the
    instance's source and optical parameters."""
    import copy
    q = copy.copy(base)
    q.mua, q.mus = medium
    q.BCType, q.BCLDirType, q.BCLNormal = source["BCType"], source["BCLDirType"], source["BCLNormal"]
    return q


@pytest.fixture(scope="module")
def batch(vmc):
    """All 256 instances of the batch workload, photon counts scaled to
    1e3..1e5 (log-uniform, the workload's draw times 1e-3)."""
    base_p, sources, media, photons = meshgen.batched_box_instances(seed=1234, photon_scale=1e-3)
    base = vmc.Problem(base_p)
    meds = [base.variant(mua=mua, mus=mus) for mua, mus in media]
    probs = [[m.variant(BCType=s["BCType"], BCLDirType=s["BCLDirType"], BCLNormal=s["BCLNormal"])
              for s in sources] for m in meds]
    items = [(probs[i][j], dict(nphoton=int(photons[i, j]), seed=1000 + 64 * i + j, stream_tag=64 * i + j))
             for i in range(4) for j in range(64)]
    yield base_p, sources, media, items
    for row in probs:
        for pr in row:
            pr.close()
    for m in meds:
        m.close()
    base.close()


def test_batch_of_256_items_equals_items_alone(vmc, batch):
    _, _, _, items = batch
    total = sum(kw["nphoton"] for _, kw in items)
    print(f"batch: {len(items)} items, {total} photons", flush=True)
    res, stats = vmc.batch_run(items, precision=vmc.FP32, tally=vmc.TALLY_FIXED64)
    assert stats.photons == total
    for (prob, kw), r in zip(items, res):
        kw = dict(kw)
        n = kw.pop("nphoton")
        alone = prob.run(n, precision=vmc.FP32, tally=vmc.TALLY_FIXED64, **kw)
        _same_bits(r, alone)
        assert r.loss == alone.loss == 0
        # each instance's own weight ledger closes and equals its solve alone
        led = np.asarray(r.ledger)
        assert abs(led[0] + led[1] + led[2] - led[3] + led[4] - n) <= 1e-6 * n, led
        assert np.allclose(led, alone.ledger, rtol=1e-9, atol=1e-9 * n), (led, alone.ledger)


@pytest.mark.parametrize("tally", ["fp64", "fixed"])
def test_batch_items_match_oracle(vmc, batch, tally):
    """Two instances of different source and medium, run inside the FP64
    batch, against the oracle on the instance's inputs with the item's seed
    and stream tag.  Fixed-point tallies round every flushed run of
    deposits to 2^-32 of a weight, hence their 1e-7 bound (1e-9 for FP64
    tallies)."""
    base_p, sources, media, items = batch
    t = vmc.TALLY_FIXED64 if tally == "fixed" else vmc.TALLY_FP64
    res, _ = vmc.batch_run(items, precision=vmc.FP64, tally=t)
    for i, j in ((0, 5), (3, 60)):
        k = 64 * i + j
        _, kw = items[k]
        q = _instance_problem(base_p, sources[j], media[i])
        q.Nphoton, q.seed = kw["nphoton"], kw["seed"]
        ref = pyoracle.run(q, 16, pyoracle.RNG_PHILOX, stream_tag=kw["stream_tag"])
        print(f"item {k}: {q.Nphoton} photons, tag {kw['stream_tag']}", flush=True)
        _match_oracle(res[k], ref, tol=1e-7 if tally == "fixed" else 1e-9)
        assert res[k].loss == ref.loss


@pytest.fixture(scope="module")
def pool(vmc):
    """The serving pool at reduced mesh size (every mesh kind and both
    dimensions) and a request stream drawn from the serving mix with its
    photon budgets scaled down."""
    pool = meshgen.serving_pool(seed=2024, scale=0.05)
    probs = [vmc.Problem(p) for p in pool]
    reqs = meshgen.serving_requests(24, seed=2025)
    yield pool, probs, reqs
    for pr in probs:
        pr.close()


def _request(pool, q):
    k = q["index"] + (0 if q["kind"] == "2d" else 4)
    return k, max(1, q["photons"] // 200)


def test_served_requests_equal_solves_alone(vmc, pool):
    """Requests arriving while others run share launches; each result is the
    same as the solve alone, bit for bit with fixed-point tallies."""
    meshes, probs, reqs = pool
    with vmc.Server(precision=vmc.FP32, tally=vmc.TALLY_FIXED64, launch_photons=200000, min_slice=4096) as srv:
        tickets = []
        for i, q in enumerate(reqs):
            k, n = _request(meshes, q)
            tickets.append((k, n, q, srv.submit(probs[k], n, seed=q["seed"], f=q["f"], stream_tag=i)))
        served = [(k, n, q, i, srv.wait(t)) for i, (k, n, q, t) in enumerate(tickets)]
        launches = srv.counters()["launches"]
    print(f"{len(served)} requests in {launches} launches", flush=True)
    assert launches < sum(-(-n // 200000) for _, n, _, _ in tickets)   # requests did share launches
    for k, n, q, i, r in served:
        alone = probs[k].run(n, seed=q["seed"], f=q["f"], stream_tag=i, precision=vmc.FP32, tally=vmc.TALLY_FIXED64)
        _same_bits(r, alone)
        assert r.loss == alone.loss


def test_served_requests_match_oracle(vmc, pool):
    """A 2D and a 3D frequency-domain request served in a mix on an FP64
    server, against the oracle with the request's seed and stream tag."""
    meshes, probs, reqs = pool
    chosen = {}
    for i, q in enumerate(reqs):
        if q["f"] > 0 and q["kind"] not in chosen:
            chosen[q["kind"]] = i
    assert set(chosen) == {"2d", "3d"}
    with vmc.Server(precision=vmc.FP64, tally=vmc.TALLY_FP64, launch_photons=100000, min_slice=4096) as srv:
        tickets = {}
        for i, q in enumerate(reqs):
            k, n = _request(meshes, q)
            tickets[i] = (k, n, q, srv.submit(probs[k], min(n, 20000), seed=q["seed"], f=q["f"], stream_tag=i))
        done = {i: srv.wait(t[3]) for i, t in tickets.items()}
    for kind, i in chosen.items():
        k, n, q, _ = tickets[i]
        p = meshes[k]
        p.Nphoton, p.seed, p.f = min(n, 20000), q["seed"], q["f"]
        print(f"request {i}: {kind} mesh {k} Ne={p.Ne} N={p.Nphoton}", flush=True)
        ref = pyoracle.run(p, 16, pyoracle.RNG_PHILOX, stream_tag=i)
        _match_oracle(done[i], ref)
        assert done[i].loss == ref.loss == 0


def test_batch_groups_items_by_roulette_constants(vmc):
    """Items whose weight0/chance differ run in separate launches of one
    batch (the constants are launch-wide); each still equals its solve
    alone, and c0 may differ inside one launch."""
    import cases
    p = cases.p3_roulette_constants()
    with vmc.Problem(p) as prob:
        items = [(prob, dict(nphoton=20000, seed=3, f=p.f, weight0=0.3, chance=0.2, c0=2.2e11)),
                 (prob, dict(nphoton=15000, seed=4, f=p.f, weight0=0.3, chance=0.2, c0=2.9e11)),
                 (prob, dict(nphoton=10000, seed=5, f=p.f)),
                 (prob, dict(nphoton=12000, seed=6, f=p.f, weight0=0.05, chance=0.5))]
        res, _ = vmc.batch_run(items, precision=vmc.FP32, tally=vmc.TALLY_FIXED64)
        for (_, kw), r in zip(items, res):
            kw = dict(kw)
            alone = prob.run(kw.pop("nphoton"), precision=vmc.FP32, tally=vmc.TALLY_FIXED64, **kw)
            _same_bits(r, alone)
        assert not np.array_equal(res[0].ER, res[2].ER)


def test_model_constants_are_validated(vmc):
    import cases
    p = cases.p3_single_cube()
    with vmc.Problem(p) as prob:
        for kw in (dict(chance=0.0), dict(chance=1.5), dict(weight0=-1.0), dict(c0=0.0),
                   dict(weight0=float("nan"))):
            with pytest.raises(vmc.VmcError) as ei:
                prob.run(1000, **kw)
            assert ei.value.status == vmc.STATUS_INVALID_ARGUMENT, kw
        with pytest.raises(vmc.VmcError):
            prob.run(2 ** 31, tally=vmc.TALLY_FIXED64)


@pytest.mark.parametrize("kw", [dict(lds_slots=-1), dict(cache_epoch=-1), dict(lds_slots=-1, chunk=4096)])
def test_many_tiny_items_without_cache_epochs(vmc, kw):
    """Items of a few photons each pad their last dispenser chunk, so a wave
    can draw only padding while other items' ids remain; with cache epochs
    off it must refill rather than leave.  Every item equals its solve
    alone."""
    import cases
    p = cases.p3_pencil_sphere()
    rng = np.random.default_rng(4)
    with vmc.Problem(p) as prob:
        items = [(prob, dict(nphoton=int(n), seed=100 + i)) for i, n in enumerate(rng.integers(1, 90, 300))]
        res, stats = vmc.batch_run(items, precision=vmc.FP32, tally=vmc.TALLY_FIXED64, **kw)
        assert stats.photons == sum(k["nphoton"] for _, k in items)
        for (_, k), r in zip(items, res):
            alone = prob.run(k["nphoton"], seed=k["seed"], precision=vmc.FP32, tally=vmc.TALLY_FIXED64)
            _same_bits(r, alone)
