"""Batched forward solves (vmc_batch_run) against single runs.

Every item keeps its own random streams (seed, stream tag, photon index)
and its own tallies, so in fixed-point mode each item's result in a mixed
batch equals its result alone bit for bit; FP64 tallies agree to rounding.
The batch mixes problems of both dimensions, CW and frequency-domain items,
and uneven photon counts (including one not a multiple of the dispenser's
chunk and an empty item).
"""
import os

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OUT = ("ER", "EI", "EBR", "EBI", "DEBR", "DEBI")


@pytest.fixture(scope="module")
def vmc():
    import valomc_amd
    return valomc_amd


def _golden(name):
    return cases.from_arrays(dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)))


MIX = [("p3_pencil_sphere", dict(nphoton=30000, seed=3)),
       ("p3_pencil_sphere", dict(nphoton=12345, seed=3, stream_tag=7)),
       ("p3_cosinic_hetero_n", dict(nphoton=20000, seed=11)),
       ("p3_direct_abs_fd", dict(nphoton=15000, seed=5)),
       ("p2_square_cosinic", dict(nphoton=25000, seed=2)),
       ("p2_iso_gauss_fd", dict(nphoton=10000, seed=9)),
       ("p3_iso_mirror", dict(nphoton=0, seed=1))]


def _check_ledger(r, alone, n, tol):
    """An item's own weight ledger (absorbed, exited, killed, gained, lost)
    closes on its photon count and equals the ledger of its run alone: the
    same photons, summed in another order."""
    led = np.asarray(r.ledger)
    assert abs(led[0] + led[1] + led[2] - led[3] + led[4] - n) <= tol * max(n, 1), led
    assert np.allclose(led, alone.ledger, rtol=1e-9, atol=1e-9 * max(n, 1)), (led, alone.ledger)


def _mix(vmc):
    probs = {}
    items = []
    for name, kw in MIX:
        p = _golden(name)
        if name not in probs:
            probs[name] = (p, vmc.Problem(p))
        pp, prob = probs[name]
        items.append((prob, dict(kw, f=pp.f, phase0=pp.phase0)))
    return probs, items


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_batch_items_equal_single_runs(vmc, precision):
    prec = vmc.FP32 if precision == "fp32" else vmc.FP64
    probs, items = _mix(vmc)
    try:
        print(f"batch of {len(items)} items, {precision}, fixed-point tallies", flush=True)
        res, stats = vmc.batch_run(items, precision=prec, tally=vmc.TALLY_FIXED64)
        assert stats.photons == sum(kw["nphoton"] for _, kw in items)
        # the batch's ledger is the sum of its items'
        assert np.allclose(stats.ledger, np.sum([r.ledger for r in res], axis=0), rtol=1e-12, atol=1e-9)
        for (prob, kw), r in zip(items, res):
            kw = dict(kw)
            n = kw.pop("nphoton")
            alone = prob.run(n, precision=prec, tally=vmc.TALLY_FIXED64, **kw)
            for k in OUT:
                assert np.array_equal(getattr(r, k), getattr(alone, k), equal_nan=True), (kw, k)
            assert r.loss == alone.loss
            _check_ledger(r, alone, n, 1e-6 if precision == "fp32" else 1e-9)
    finally:
        for _, prob in probs.values():
            prob.close()


def test_batch_fp64_tallies_match_single_runs(vmc):
    probs, items = _mix(vmc)
    try:
        res, _ = vmc.batch_run(items, precision=vmc.FP64)
        for (prob, kw), r in zip(items, res):
            kw = dict(kw)
            n = kw.pop("nphoton")
            alone = prob.run(n, precision=vmc.FP64, **kw)
            _check_ledger(r, alone, n, 1e-9)
            for k in OUT:
                a, b = getattr(r, k), getattr(alone, k)
                if n == 0:
                    assert np.array_equal(a, b, equal_nan=True), k   # 0/0 like the reference
                    continue
                nb = np.linalg.norm(b)
                assert np.linalg.norm(a - b) <= 1e-12 * max(nb, 1e-300), (kw, k)
    finally:
        for _, prob in probs.values():
            prob.close()


def test_batch_rejects_mixed_precision(vmc):
    p = _golden("p3_pencil_sphere")
    with vmc.Problem(p) as prob:
        items = [(prob, dict(nphoton=100, precision=vmc.FP32)), (prob, dict(nphoton=100, precision=vmc.FP64))]
        with pytest.raises(vmc.VmcError):
            vmc.batch_run(items)


def _fresh_equal(vmc, p_mod, variant, **kw):
    with vmc.Problem(p_mod) as fresh:
        a = fresh.run(20000, seed=4, tally=vmc.TALLY_FIXED64, **kw)
    b = variant.run(20000, seed=4, tally=vmc.TALLY_FIXED64, **kw)
    for k in OUT:
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_variants_equal_fresh_problems(vmc, precision):
    """A variant (shared mesh and records, other medium or sources) gives the
    bits of a problem created from scratch with the same inputs."""
    prec = vmc.FP32 if precision == "fp32" else vmc.FP64
    p = _golden("p3_cosinic_hetero_n")
    with vmc.Problem(p) as base:
        # medium override
        p2 = _golden("p3_cosinic_hetero_n")
        p2.mua = np.asarray(p2.mua) * 1.1
        p2.mus = np.asarray(p2.mus) * 0.9
        v = base.variant(mua=p2.mua, mus=p2.mus)
        _fresh_equal(vmc, p2, v, precision=prec)
        v.close()
        # light-source override: move the source to other boundary faces
        p3 = _golden("p3_cosinic_hetero_n")
        bct = np.asarray(p3.BCType).copy()
        src = np.nonzero(bct != ord("a"))[0]
        bct[src] = ord("a")
        bct[np.arange(0, p3.Nb, 7)] = ord("l")
        p3.BCType = bct
        v = base.variant(BCType=bct)
        _fresh_equal(vmc, p3, v, precision=prec)
        # the base is untouched by its variants
        _fresh_equal(vmc, _golden("p3_cosinic_hetero_n"), base, precision=prec)
        v.close()


def test_variant_without_light_source_is_refused(vmc):
    p = _golden("p3_pencil_sphere")
    with vmc.Problem(p) as base:
        with pytest.raises(vmc.VmcError) as ei:
            base.variant(BCType=np.full(p.Nb, ord("a"), dtype=np.uint8))
        assert ei.value.status == 1   # 1 + NO_LIGHTSOURCE


def test_many_items_equal_single_runs(vmc):
    """Two dozen items of one problem, with uneven photon counts (one large
    item, an empty one, a single photon, one just past a dispenser chunk):
    chunks of many items are in flight at once and items end at different
    times, yet each item equals its solo run bit for bit (fixed point) and
    keeps its own weight ledger (24 items share a wave's four ledger slots)."""
    p = _golden("p3_pencil_sphere")
    q = _golden("p3_direct_abs_fd")
    counts = [200000, 0, 1, 129] + [3000 + 1500 * k for k in range(20)]
    try:
        prob, probq = vmc.Problem(p), vmc.Problem(q)
        items = [(prob, dict(nphoton=n, seed=21, stream_tag=k)) for k, n in enumerate(counts)]
        print(f"batch of {len(items)} CW items", flush=True)
        res, stats = vmc.batch_run(items, precision=vmc.FP32, tally=vmc.TALLY_FIXED64)
        assert stats.photons == sum(counts)
        assert np.allclose(stats.ledger, np.sum([r.ledger for r in res], axis=0), rtol=1e-12, atol=1e-9)
        for (pr, kw), r in zip(items, res):
            kw = dict(kw)
            n = kw.pop("nphoton")
            alone = pr.run(n, precision=vmc.FP32, tally=vmc.TALLY_FIXED64, **kw)
            for k in OUT:
                assert np.array_equal(getattr(r, k), getattr(alone, k), equal_nan=True), (kw, k)
            assert r.loss == alone.loss
            _check_ledger(r, alone, n, 1e-6)
        itemsq = [(probq, dict(nphoton=n, seed=5, stream_tag=k, f=q.f, phase0=q.phase0))
                  for k, n in enumerate(counts[4:])]
        print(f"batch of {len(itemsq)} FD items", flush=True)
        resq, _ = vmc.batch_run(itemsq, precision=vmc.FP32, tally=vmc.TALLY_FIXED64)
        for (pr, kw), r in zip(itemsq, resq):
            kw = dict(kw)
            n = kw.pop("nphoton")
            alone = pr.run(n, precision=vmc.FP32, tally=vmc.TALLY_FIXED64, **kw)
            for k in OUT:
                assert np.array_equal(getattr(r, k), getattr(alone, k), equal_nan=True), (kw, k)
            _check_ledger(r, alone, n, 1e-6)
    finally:
        prob.close()
        probq.close()


@pytest.mark.parametrize("epoch", [1, 3, 64])
def test_batch_cache_epochs_keep_results(vmc, epoch):
    """A batch's cache epochs (all waves meet at a barrier, move the slots'
    sums to the replica and free the slots) at short intervals, so they run
    many times while items start and end: every item still equals its solve
    alone bit for bit (fixed point), CW and frequency domain, and its ledger
    closes."""
    p = _golden("p3_pencil_sphere")
    q = _golden("p3_direct_abs_fd")
    counts = [40000, 0, 1, 129] + [2000 + 900 * k for k in range(12)]
    try:
        prob, probq = vmc.Problem(p), vmc.Problem(q)
        items = [(prob, dict(nphoton=n, seed=31, stream_tag=k)) for k, n in enumerate(counts)]
        items += [(probq, dict(nphoton=n, seed=7, stream_tag=k, f=q.f, phase0=q.phase0))
                  for k, n in enumerate(counts[4:])]
        print(f"batch of {len(items)} items, cache epoch every {epoch} passes", flush=True)
        res, stats = vmc.batch_run(items, precision=vmc.FP32, tally=vmc.TALLY_FIXED64, cache_epoch=epoch)
        assert stats.photons == sum(kw["nphoton"] for _, kw in items)
        for (pr, kw), r in zip(items, res):
            kw = dict(kw)
            n = kw.pop("nphoton")
            alone = pr.run(n, precision=vmc.FP32, tally=vmc.TALLY_FIXED64, **kw)
            for k in OUT:
                assert np.array_equal(getattr(r, k), getattr(alone, k), equal_nan=True), (kw, k)
            assert r.loss == alone.loss
            _check_ledger(r, alone, n, 1e-6)
    finally:
        prob.close()
        probq.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_batch_plan_replays_equal_eager_batches(vmc, precision):
    """A batch recorded once as a HIP graph (2D and 3D, CW and FD groups, an
    empty item) and replayed with new seeds, with eager launches between the
    replays: every replay equals the eager batch of the same seeds bit for
    bit, a medium changed between replays is used by the next one, and
    items that do not match the plan are refused."""
    probs, items = _mix(vmc)
    kw_common = dict(precision=vmc.FP32 if precision == "fp32" else vmc.FP64, tally=vmc.TALLY_FIXED64)
    try:
        with vmc.BatchPlan(items, **kw_common) as plan:
            for rep in range(3):
                seeded = [(prob, dict(kw, seed=kw["seed"] + 100 * rep)) for prob, kw in items]
                print(f"replay {rep}", flush=True)
                got, _ = plan.run(seeded)
                want, _ = vmc.batch_run(seeded, **kw_common)
                for (_, kw), g, w in zip(seeded, got, want):
                    for k in OUT:
                        assert np.array_equal(getattr(g, k), getattr(w, k), equal_nan=True), (rep, kw, k)
                    assert g.loss == w.loss
                prob0, kw0 = items[0]
                prob0.run(5000, seed=rep, precision=vmc.FP32)   # eager work between replays
            pp, prob = probs["p3_pencil_sphere"]
            prob.set_medium(mua=np.asarray(pp.mua) * 1.5)
            print("replay after set_medium", flush=True)
            got, _ = plan.run()
            want, _ = vmc.batch_run(items, **kw_common)
            for (_, kw), g, w in zip(items, got, want):
                for k in OUT:
                    assert np.array_equal(getattr(g, k), getattr(w, k), equal_nan=True), ("medium", kw, k)
            bad = [(prob, dict(kw, nphoton=kw["nphoton"] + 1)) for prob, kw in items]
            with pytest.raises(vmc.VmcError):
                plan.run(bad)
            with pytest.raises(vmc.VmcError):
                plan.run(items[:-1])
    finally:
        for _, prob in probs.values():
            prob.close()


def test_batch_autotune_keeps_results(vmc):
    """vmc_batch_autotune picks valid knobs for a mixed batch, and the batch
    run with them equals the batch run with the defaults (fixed point)."""
    probs, items = _mix(vmc)
    common = dict(precision=vmc.FP32, tally=vmc.TALLY_FIXED64)
    try:
        t = vmc.batch_autotune(items, trial_fraction=0.5, **common)
        assert t.pop("trials") >= 10
        assert t["block_size"] in (1024, 512, 256) and t["chunk"] in (128, 64, 256, 512)
        assert t["refill_min"] in (6, 10, 16, 24, 32) and t["replicas"] in (0, 8, 1) and t["lds_slots"] in (0, -1)
        assert t["cache_epoch"] in (0, -1, 256, 4096)
        base, _ = vmc.batch_run(items, **common)
        tuned, _ = vmc.batch_run(items, **common, **t)
        for (_, kw), r0, r1 in zip(items, base, tuned):
            for k in OUT:
                assert np.array_equal(getattr(r0, k), getattr(r1, k), equal_nan=True), (t, kw, k)
    finally:
        for _, prob in probs.values():
            prob.close()
