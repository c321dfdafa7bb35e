"""GPU parity: the HIP transport path against the oracle.

* FP64 on the device vs the CPU restatement driven by the same per-photon
  Philox streams: the same photons take the same paths, so the tallies
  agree to rounding (relative L2 <= 1e-9) on every golden case.
* Against the reference's own MT19937 streams the agreement is
  statistical: relative L2 within a few two-seed spreads of the reference.
* FP32 transport vs FP64 with the same streams: <= 1% relative L2.
* Energy ledger, fixed-point reproducibility across launch shapes, abort,
  error codes and the MC3D/MC2D class mirror.
"""
import os

import numpy as np
import pytest

import cases
import pyoracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OUT = ("ER", "EI", "EBR", "EBI", "DEBR", "DEBI")


@pytest.fixture(scope="module")
def vmc():
    import valomc_amd
    print("device:", valomc_amd.device_info(), flush=True)
    return valomc_amd


def _golden(name):
    return cases.from_arrays(dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)))


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a - b)


def _run(vmc, p, **kw):
    with vmc.Problem(p) as prob:
        return prob.run(kw.pop("nphoton", p.Nphoton), seed=kw.pop("seed", p.seed), f=p.f, phase0=p.phase0,
                        weight0=p.weight0, chance=p.chance, c0=p.c0, **kw)


@pytest.mark.parametrize("tally", ["fp64", "fixed"])
@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_fp64_matches_oracle_on_same_streams(vmc, name, tally):
    """FP64 transport against the Philox-mode oracle, photon for photon,
    with FP64 tallies (relative L2 <= 1e-9) and with the bit-reproducible
    fixed-point tallies, whose every flushed run of deposits is rounded to a
    multiple of 2^-32 of a photon weight: a CW visit deposits ~1e-3 of a
    weight, so an element's sum carries ~1e-8 / sqrt(runs) relative
    rounding and the bound is 1e-7.  The roulette cases run the reference's
    weight0/chance/c0 members at non-default values."""
    p = _golden(name)
    print(f"case {name} Ne={p.Ne} N={p.Nphoton} tally={tally}", flush=True)
    ref = pyoracle.run(p, 4, pyoracle.RNG_PHILOX)
    res = _run(vmc, p, precision=vmc.FP64, tally=vmc.TALLY_FIXED64 if tally == "fixed" else vmc.TALLY_FP64)
    tol = 1e-7 if tally == "fixed" else 1e-9
    for k in OUT:
        a, b = getattr(res, k), getattr(ref, k)
        if np.linalg.norm(b) == 0:
            assert np.max(np.abs(a)) < 1e-12, k
        else:
            assert rel_l2(a, b) < tol, (k, rel_l2(a, b))
    assert abs(res.loss - ref.loss) <= max(2, p.Nphoton // 1000), (res.loss, ref.loss)


@pytest.mark.parametrize("nphot", [0, 1, 2, 65])
@pytest.mark.parametrize("name", ["p3_pencil_sphere", "p3_direct_abs_fd", "p2_square_cosinic", "p2_iso_gauss_fd"])
def test_tiny_photon_counts_match_oracle(vmc, name, nphot):
    """Ragged minimal runs: no photon (every divided tally is 0/0 = NaN,
    as in the reference), one, two, and one more than a wave.  FP64 output
    equals the oracle's on the same streams, NaNs in the same words, and
    loses the same photons; the FP32 path gives the same NaN pattern."""
    p = _golden(name)
    p.Nphoton = nphot
    print(f"case {name} N={nphot}", flush=True)
    ref = pyoracle.run(p, 1, pyoracle.RNG_PHILOX)
    res = _run(vmc, p, precision=vmc.FP64)
    res32 = _run(vmc, p, precision=vmc.FP32)
    for k in OUT:
        a, a32, b = getattr(res, k), getattr(res32, k), getattr(ref, k)
        assert np.array_equal(np.isnan(a), np.isnan(b)), k
        assert np.array_equal(np.isnan(a32), np.isnan(b)), k
        fin = ~np.isnan(b)
        if nphot == 0:
            assert np.all(a[fin] == 0) and np.all(a32[fin] == 0), k
        elif np.linalg.norm(b[fin]) > 0:
            assert rel_l2(a[fin], b[fin]) < 1e-9, (k, rel_l2(a[fin], b[fin]))
    assert res.loss == ref.loss


def test_fixed_point_range_is_reported(vmc):
    """Fixed-point words are exact while a flushed sum stays below 2^30
    weight units.  A cube 1e10 mm across, non-scattering (mus = 0), index
    matched (n = 1) and non-absorbing (mua = 0), makes each tet crossing
    deposit w * ds ~ 1e10 (the reference's non-absorbing tally), so the run
    ends with VMC_ERR_TALLY_RANGE, alone and in a batch; the same inputs
    with FP64 tallies are fine."""
    p = _golden("p3_single_cube")
    p.r = np.asarray(p.r, dtype=np.float64) * 1e10
    p.mua = np.zeros_like(p.mua)
    p.mus = np.zeros_like(p.mus)
    p.n = np.ones_like(p.n)
    with vmc.Problem(p) as prob:
        ok = prob.run(2000, seed=3, precision=vmc.FP64)
        assert np.all(np.isfinite(ok.ER)) and ok.ER.max() > 0 and ok.loss == 0
        for prec in (vmc.FP64, vmc.FP32):
            with pytest.raises(vmc.VmcError) as e:
                prob.run(2000, seed=3, precision=prec, tally=vmc.TALLY_FIXED64)
            assert e.value.status == vmc.STATUS_TALLY_RANGE
        q = _golden("p3_pencil_sphere")
        with vmc.Problem(q) as fine:
            items = [(fine, dict(nphoton=2000, seed=1)), (prob, dict(nphoton=2000, seed=3))]
            with pytest.raises(vmc.VmcError) as e:
                vmc.batch_run(items, precision=vmc.FP32, tally=vmc.TALLY_FIXED64)
            assert e.value.status == vmc.STATUS_TALLY_RANGE


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_ledger_closes(vmc, name):
    p = _golden(name)
    for prec, tol in ((vmc.FP64, 1e-9), (vmc.FP32, 1e-6)):
        res = _run(vmc, p, precision=prec)
        absorbed, exited, killed, gained, lost = res.ledger
        total = absorbed + exited + killed - gained + lost
        assert abs(total - p.Nphoton) <= tol * p.Nphoton, (prec, total)


@pytest.mark.parametrize("name", ["p3_pencil_sphere", "p3_cosinic_hetero_n", "p3_direct_abs_fd",
                                  "p2_square_cosinic", "p2_iso_gauss_fd", "p2_mirror_hetero"])
def test_fp32_within_one_percent_of_fp64(vmc, name):
    p = _golden(name)
    p.Nphoton = 200000
    a = _run(vmc, p, precision=vmc.FP64)
    b = _run(vmc, p, precision=vmc.FP32)
    for k in ("ER", "EBR"):
        assert rel_l2(getattr(b, k), getattr(a, k)) < 1e-2, (k, rel_l2(getattr(b, k), getattr(a, k)))
    if p.f > 0:
        assert rel_l2(b.EI, a.EI) < 1e-2


@pytest.mark.parametrize("name,nphot", [("p2_square_cosinic", 400000), ("p3_pencil_sphere", 200000),
                                        ("p3_iso_mirror", 200000), ("p2_given_directions", 200000)])
def test_statistical_parity_with_reference_streams(vmc, name, nphot):
    """GPU (Philox) vs oracle (the reference's MT19937 streams): the
    difference stays within the spread of two independent reference runs."""
    p = _golden(name)
    p.Nphoton = nphot
    r1 = pyoracle.run(p, 8, pyoracle.RNG_MT19937)
    p2 = _golden(name)
    p2.Nphoton, p2.seed = nphot, p.seed + 1000
    r2 = pyoracle.run(p2, 8, pyoracle.RNG_MT19937)
    g = _run(vmc, p, precision=vmc.FP64, seed=777)
    for k in ("ER", "EBR"):
        spread = rel_l2(getattr(r2, k), getattr(r1, k))
        diff = rel_l2(getattr(g, k), getattr(r1, k))
        assert diff < 2.5 * spread, (k, diff, spread)
    # loss rate agrees with the reference's
    assert abs(g.loss - r1.loss) <= 5 * np.sqrt(max(r1.loss, 1)) + 5


def test_fixed_point_tallies_are_bit_reproducible(vmc):
    p = _golden("p3_pencil_sphere")
    p.Nphoton = 100000
    with vmc.Problem(p) as prob:
        a = prob.run(p.Nphoton, seed=5, precision=vmc.FP32, tally=vmc.TALLY_FIXED64)
        b = prob.run(p.Nphoton, seed=5, precision=vmc.FP32, tally=vmc.TALLY_FIXED64,
                     block_size=128, replicas=3, refill_min=16, blocks_per_cu=2)
        c = prob.run(p.Nphoton, seed=5, precision=vmc.FP32, tally=vmc.TALLY_FP64)
    for k in OUT:
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert rel_l2(c.ER, a.ER) < 1e-8


@pytest.mark.parametrize("name", ["p3_pencil_sphere", "p3_direct_abs_fd", "p2_iso_gauss_fd"])
def test_photon_ranges_split_a_run(vmc, name):
    """A run walks stream ids first_photon .. first_photon + N - 1, so
    disjoint ranges (the shares of several launches or GPUs) add up to the
    whole run: with fixed-point tallies the raw sums are exact integers, so
    the photon-weighted outputs of the ranges equal the whole run's to the
    rounding of the normalisation, and losses and element steps add up
    exactly.  Ragged shares, one of a single photon."""
    p = _golden(name)
    n_all = 120000
    cuts = [0, 17000, 60000, 60001, n_all]
    print(f"case {name}", flush=True)
    with vmc.Problem(p) as prob:
        def run(n, first):
            return prob.run(n, seed=21, f=p.f, phase0=p.phase0, precision=vmc.FP32, tally=vmc.TALLY_FIXED64,
                            first_photon=first)
        whole = run(n_all, 0)
        parts = [(b - a, run(b - a, a)) for a, b in zip(cuts[:-1], cuts[1:])]
    assert sum(r.loss for _, r in parts) == whole.loss
    assert sum(r.steps for _, r in parts) == whole.steps
    for k in OUT:
        want = getattr(whole, k)
        got = sum(getattr(r, k) * n for n, r in parts) / n_all
        assert np.max(np.abs(got - want)) <= 1e-13 * max(np.max(np.abs(want)), 1e-300), k
    # a range is the same photons whichever run they are in
    with vmc.Problem(p) as prob:
        a = prob.run(1000, seed=21, f=p.f, phase0=p.phase0, precision=vmc.FP32, tally=vmc.TALLY_FIXED64,
                     first_photon=5000)
        b = prob.run(1000, seed=21, f=p.f, phase0=p.phase0, precision=vmc.FP32, tally=vmc.TALLY_FIXED64,
                     first_photon=5000, block_size=128)
        c = prob.run(1000, seed=21, f=p.f, phase0=p.phase0, precision=vmc.FP32, tally=vmc.TALLY_FIXED64)
    assert all(np.array_equal(getattr(a, k), getattr(b, k)) for k in OUT)
    assert not np.array_equal(a.ER, c.ER)


@pytest.mark.parametrize("name", ["p3_pencil_sphere", "p3_direct_abs_fd", "p2_iso_gauss_fd",
                                  "p2_mirror_hetero"])
def test_tally_paths_agree_bit_for_bit(vmc, name):
    """Fixed-point adds are exact and order-independent, so the LDS cache
    (kept for the whole launch or handed back every few loop passes), the
    coalesced flush and per-lane atomics must give the same bits, for
    element words (CW and FD) and every boundary word (3D and 2D)."""
    p = _golden(name)
    p.Nphoton = 50000
    print(f"case {name}", flush=True)
    with vmc.Problem(p) as prob:
        run = lambda **kw: prob.run(p.Nphoton, seed=11, f=p.f, phase0=p.phase0, precision=vmc.FP32,
                                    tally=vmc.TALLY_FIXED64, **kw)
        base = run()
        for kw in (dict(lds_slots=-1, direct_flush=1), dict(lds_slots=-1), dict(direct_flush=1),
                   dict(lds_slots=7), dict(block_size=64), dict(block_size=64, lds_slots=-1),
                   dict(cache_epoch=1), dict(cache_epoch=5, lds_slots=7), dict(cache_epoch=3, block_size=128)):
            r = run(**kw)
            for k in OUT:
                assert np.array_equal(getattr(r, k), getattr(base, k)), (kw, k)


@pytest.mark.parametrize("name", ["p3_single_cube", "p3_single_cube_fd", "p3_direct_abs_fd", "p2_square_cosinic",
                                  "p2_iso_gauss_fd", "p2_mirror_hetero"])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_small_mesh_in_lds_agrees_bit_for_bit(vmc, name, precision):
    """A mesh whose element records fit a workgroup's LDS share is walked
    from LDS with a cache slot per sector (reported as lds_mesh); explicit
    cache knobs keep the records in device memory.  With fixed-point tallies
    both give the same bits, the same steps and the same losses."""
    p = _golden(name.replace("_fd", "") if name == "p3_single_cube_fd" else name)
    if name == "p3_single_cube_fd":
        p.f, p.phase0 = 100e6, 0.3   # the 3D frequency-domain tallies of a mesh that fits in FP64 too
    p.Nphoton = 30000
    prec = vmc.FP32 if precision == "fp32" else vmc.FP64
    print(f"case {name} {precision}", flush=True)
    with vmc.Problem(p) as prob:
        run = lambda **kw: prob.run(p.Nphoton, seed=17, f=p.f, phase0=p.phase0, precision=prec,
                                    tally=vmc.TALLY_FIXED64, **kw)
        small = run()
        # 384 FP64 tetrahedra (256-byte records) exceed the 80 KB of a
        # 256-lane FP64 workgroup; every other case here fits
        assert small.lds_mesh == (not (precision == "fp64" and p.dim == 3 and p.Ne > 100)), p.Ne
        for kw in (dict(lds_slots=-1), dict(lds_slots=5), dict(cache_epoch=2)):
            r = run(**kw)
            assert not r.lds_mesh
            for k in OUT:
                assert np.array_equal(getattr(r, k), getattr(small, k)), (kw, k)
            assert (r.loss, r.steps) == (small.loss, small.steps), kw


def test_mesh_residency_follows_size(vmc):
    """The 200-triangle square of the reference's 2D example is held in LDS
    in FP32 and FP64; the 1296-tetrahedron golden mesh (166 KB of FP32
    records) is not."""
    from valomc_amd import meshgen
    sq = meshgen.square_2d(nphoton=20000, seed=1)
    big = _golden("p3_pencil_sphere")
    with vmc.Problem(sq) as a, vmc.Problem(big) as b:
        assert a.run(20000, precision=vmc.FP32).lds_mesh
        assert a.run(20000, precision=vmc.FP64).lds_mesh
        assert not b.run(20000, precision=vmc.FP32).lds_mesh
        assert not b.run(20000, precision=vmc.FP64).lds_mesh


def test_mesh_residency_around_the_lds_limit(vmc):
    """Square grids from 288 to 1352 triangles cross the LDS limits of the
    small-mesh path (about 950 element slots in FP32, 320 in FP64 CW and
    224 in FP64 frequency domain): every size runs, the largest ones from
    device memory, and each equals the general path bit for bit in fixed
    point."""
    from valomc_amd import meshgen
    seen = set()
    for cells in (12, 16, 20, 22, 24, 26):
        r, H, BH = meshgen.rectangular_mesh(10, 10, 10.0 / cells)
        p = meshgen.homogeneous(2, r, H, BH, 0.01, 1.0, 0.9, 1.3, 1.0)
        p.BCType[:cells // 2] = ord("c")
        with vmc.Problem(p) as prob:
            for prec in (vmc.FP32, vmc.FP64):
                for f in (0.0, 100e6):
                    run = lambda **kw: prob.run(20000, seed=5, f=f, precision=prec, tally=vmc.TALLY_FIXED64, **kw)
                    a, b = run(), run(lds_slots=-1)
                    print(f"{2 * cells * cells} triangles, prec {prec}, f {f}: lds_mesh {a.lds_mesh}", flush=True)
                    seen.add((prec, a.lds_mesh))
                    for k in OUT:
                        assert np.array_equal(getattr(a, k), getattr(b, k), equal_nan=True), (cells, prec, f, k)
    assert seen == {(vmc.FP32, True), (vmc.FP32, False), (vmc.FP64, True), (vmc.FP64, False)}


def test_tally_flush_mode_is_validated(vmc):
    """direct_flush takes 0 or 1 only."""
    p = _golden("p3_pencil_sphere")
    with vmc.Problem(p) as prob:
        for q in (-1, 2, 1024):
            with pytest.raises(vmc.VmcError):
                prob.run(1000, direct_flush=q)


def test_results_do_not_depend_on_launch_shape(vmc):
    p = _golden("p2_square_cosinic")
    p.Nphoton = 50000
    with vmc.Problem(p) as prob:
        a = prob.run(p.Nphoton, seed=9, precision=vmc.FP64)
        b = prob.run(p.Nphoton, seed=9, precision=vmc.FP64, block_size=64, replicas=1, refill_min=32)
        c = prob.run(p.Nphoton, seed=9, precision=vmc.FP64, blocks_per_cu=1)
        assert c.grid_blocks == a.grid_blocks // 2 and c.grid_blocks > 0   # FP64: two 256-lane groups per CU
        for k in OUT:
            assert rel_l2(getattr(a, k), getattr(c, k)) < 1e-12, k
    for k in OUT:
        assert rel_l2(getattr(a, k), getattr(b, k)) < 1e-12, k


def test_abort_keeps_planned_normalisation(vmc):
    p = _golden("p2_square_cosinic")
    p.Nphoton = 200000
    calls = []

    def progress(pct):
        calls.append(pct)
        return len(calls) < 5

    with vmc.Problem(p) as prob:
        full = prob.run(p.Nphoton, seed=3)
        part = prob.run(p.Nphoton, seed=3, progress=progress, check=False)
    assert part.status == vmc.STATUS_ABORTED
    assert len(calls) == 5
    # 5 of 20 chunks ran and the tallies are still divided by the planned N
    # (MC3D.hpp:1714-1718)
    ratio = part.ER.sum() / full.ER.sum()
    assert 0.2 < ratio < 0.3


@pytest.mark.parametrize("name", sorted(cases.ERROR_CASES))
def test_error_status_on_device_path(vmc, name):
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    with pytest.raises(vmc.VmcError) as ei:
        vmc.Problem(cases.from_arrays(d))
    assert ei.value.status == int(d["out_status"])


def test_class_mirror_mc2d(vmc):
    p = _golden("p2_doc_quad")
    mc = vmc.MC2D()
    mc.H, mc.BH, mc.r = p.H, p.BH, p.r
    mc.mua, mc.mus, mc.g, mc.n = p.mua, p.mus, p.g, p.n
    mc.BCType, mc.BCn = p.BCType, p.BCn
    mc.Nphoton, mc.seed = 50000, 4
    mc.ErrorChecks()
    mc.Init()
    assert mc.HN.shape == (2, 3)
    mc.MonteCarlo()
    ref = pyoracle.run(p, 1, pyoracle.RNG_MT19937)
    assert mc.ER.shape == (2,) and np.all(mc.ER > 0)
    assert rel_l2(mc.ER, ref.ER) < 0.05


def test_one_shot_mex_style_entry(vmc):
    import ctypes as C
    lib = vmc.load_library()
    p = _golden("p3_single_cube")
    cm = lambda a, t: np.ascontiguousarray(np.asarray(a, dtype=t).ravel(order="F"))
    H, BH, r = cm(p.H, np.int64), cm(p.BH, np.int64), cm(p.r, np.float64)
    arr = {k: cm(getattr(p, k), np.float64) for k in ("mua", "mus", "g", "n", "BCn")}
    bct = cm(p.BCType, np.uint8)
    out = {k: np.zeros(p.Ne if k in ("ER", "EI") else p.Nb) for k in OUT}
    hn = np.zeros(p.Ne * 4, dtype=np.int64)
    loss = C.c_int64(0)
    V = C.c_void_p
    fn = lib.vmc_monte_carlo
    fn.argtypes = [C.c_int32, C.c_int64, C.c_int64, C.c_int64, V, V, V, V, V, V, V, V, V, V, V, V, V,
                   C.c_double, C.c_double, C.c_int64, V, C.c_uint64, C.c_int32, V, V, V, V, V, V, V,
                   C.POINTER(C.c_int64)]
    st = fn(3, p.Ne, p.Nb, p.Nr, H.ctypes.data, None, BH.ctypes.data, r.ctypes.data, bct.ctypes.data, None,
            None, None, arr["BCn"].ctypes.data, arr["mua"].ctypes.data, arr["mus"].ctypes.data,
            arr["g"].ctypes.data, arr["n"].ctypes.data, 0.0, 0.0, 20000, None, 11, 0,
            *[out[k].ctypes.data for k in OUT], hn.ctypes.data, C.byref(loss))
    assert st == 0
    p.Nphoton, p.seed = 20000, 11
    ref = pyoracle.run(p, 2, pyoracle.RNG_PHILOX)
    assert rel_l2(out["ER"], ref.ER) < 1e-9
    assert np.array_equal(hn.reshape(4, p.Ne).T, ref.HN)


@pytest.mark.parametrize("name,nphot", [("p3_pencil_sphere", 100000), ("p2_square_cosinic", 200000),
                                        ("p3_cosinic_hetero_n", 100000)])
def test_per_element_z_scores_against_reference_streams(vmc, name, nphot):
    """Element by element against the reference's own MT19937 streams: K
    independent seeds on each side give a mean and a standard error per
    element; over the elements whose mean is known to 5 % on both sides,
    z = (gpu - ref) / sqrt(se_gpu^2 + se_ref^2) must look standard normal
    (mean z^2 near 1, no |z| beyond 4.5).  A bias of a few percent anywhere
    (a dropped term, a wrong branch) shows up as large z on the well-sampled
    elements even when the global relative L2 test passes."""
    K = 10
    p = _golden(name)
    g, r = [], []
    with vmc.Problem(p) as prob:
        for k in range(K):
            g.append(prob.run(nphot, seed=1000 + k, f=p.f, phase0=p.phase0, precision=vmc.FP64).ER)
    for k in range(K):
        q = _golden(name)
        # thread t of an MT run uses seed + t: space the seeds so no two
        # runs share a stream
        q.Nphoton, q.seed = nphot, 5000 + 16 * k
        r.append(pyoracle.run(q, 8, pyoracle.RNG_MT19937).ER)
    g, r = np.array(g), np.array(r)
    mg, mr = g.mean(0), r.mean(0)
    sg, sr = g.std(0, ddof=1) / np.sqrt(K), r.std(0, ddof=1) / np.sqrt(K)
    ok = (mg > 0) & (mr > 0) & (sg < 0.05 * mg) & (sr < 0.05 * mr)
    assert ok.sum() >= 50, ok.sum()
    z = (mg[ok] - mr[ok]) / np.sqrt(sg[ok] ** 2 + sr[ok] ** 2)
    # the standard errors are estimated from K runs a side, so z follows a
    # t distribution with about 2(K-1) degrees of freedom; the largest |z|
    # over n elements is bounded by the Bonferroni quantile at 1e-3
    from scipy import stats
    nu = 2 * (K - 1)
    zmax = stats.t.ppf(1.0 - 1e-3 / (2 * ok.sum()), nu)
    print(f"{name}: {ok.sum()} elements, mean z^2 {np.mean(z ** 2):.3f} (t: {nu / (nu - 2):.3f}), "
          f"max |z| {np.abs(z).max():.2f} (bound {zmax:.2f})", flush=True)
    assert 0.5 < np.mean(z ** 2) < 2.0, np.mean(z ** 2)
    assert np.abs(z).max() < zmax, np.abs(z).max()
