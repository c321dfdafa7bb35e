"""vmc_problem_autotune: the launch shape and tally strategy it picks are
valid run settings, and, like every launch knob, they leave the results
unchanged (bit for bit with fixed-point tallies)."""
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


@pytest.mark.parametrize("name,precision", [("p3_pencil_sphere", "fp32"), ("p3_direct_abs_fd", "fp32"),
                                            ("p2_square_cosinic", "fp64")])
def test_autotuned_settings_are_valid_and_keep_results(vmc, name, precision):
    prec = vmc.FP32 if precision == "fp32" else vmc.FP64
    p = _golden(name)
    kw = dict(seed=17, f=p.f, phase0=p.phase0, precision=prec, tally=vmc.TALLY_FIXED64)
    print(f"case {name} {precision}", flush=True)
    with vmc.Problem(p) as prob:
        t = prob.autotune(40000, trial_photons=20000, **kw)
        assert t["trials"] >= 10
        blocks = (256, 128, 64) if prec == vmc.FP64 else (1024, 512, 256)
        assert t["block_size"] in blocks
        assert t["refill_min"] in (6, 10, 16, 24, 32)
        assert t["chunk"] in (128, 64, 256, 512)
        assert t["replicas"] in (0, 8, 1)
        assert t["lds_slots"] in (0, -1)
        assert t["cache_epoch"] in (0, -1, 256, 4096)
        trials = t.pop("trials")
        assert trials > 0
        base = prob.run(40000, **kw)
        tuned = prob.run(40000, **kw, **t)
        for k in OUT:
            assert np.array_equal(getattr(base, k), getattr(tuned, k), equal_nan=True), (t, k)


def test_chunk_size_leaves_results_unchanged(vmc):
    p = _golden("p3_cosinic_hetero_n")
    with vmc.Problem(p) as prob:
        run = lambda **kw: prob.run(30000, seed=4, precision=vmc.FP32, tally=vmc.TALLY_FIXED64, **kw)
        base = run()
        for c in (64, 256, 4096):
            r = run(chunk=c)
            for k in OUT:
                assert np.array_equal(getattr(base, k), getattr(r, k)), (c, k)
        for bad in (16, 32, 100, 8192):   # below a wave, not a power of 2, too large
            with pytest.raises(vmc.VmcError):
                run(chunk=bad)
