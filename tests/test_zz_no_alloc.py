"""vmc_problem_run_noalloc against vmc_problem_run.

The allocation-free entry point keeps the tallies of a run in the
library's static device scratch; its results must be those of
vmc_problem_run: bit for bit with fixed-point tallies; to rounding with FP64
tallies (relative L2 <= 1e-9), whose sums depend on the order of the
atomics and, in FP32 transport, on which visits a lane sums in FP32 before
they reach the FP64 words -- two vmc_problem_run calls differ the same
way.  Two calls in a row, the second on a
smaller problem, so a run never sees what the previous one left in the
shared scratch.
"""
import os

import numpy as np
import pytest

import cases
from valomc_amd import meshgen

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


def _compare(vmc, prob, n, tally, **kw):
    ref = prob.run(n, tally=tally, **kw)
    got = prob.run(n, tally=tally, allocate=False, **kw)
    assert got.steps == ref.steps and got.loss == ref.loss
    for k in OUT:
        a, b = getattr(got, k), getattr(ref, k)
        if tally == vmc.TALLY_FIXED64:
            assert np.array_equal(a, b), k
        else:
            assert rel_l2(a, b) < 1e-9, (k, rel_l2(a, b))
    return got


@pytest.mark.parametrize("tally", ["fixed", "fp64"])
def test_noalloc_equals_run_large_then_small(vmc, tally):
    t = vmc.TALLY_FIXED64 if tally == "fixed" else vmc.TALLY_FP64
    big = meshgen.sphere_box_3d(nphoton=1, seed=1234)
    print(f"noalloc: 384k-tet box, then smaller problems, tally {tally}", flush=True)
    with vmc.Problem(big) as prob:
        _compare(vmc, prob, 2_000_000, t, seed=3, precision=vmc.FP32)
        # FP64 records are allocated by that precision's first run
        prob.run(1000, seed=1, precision=vmc.FP64)
        _compare(vmc, prob, 300_000, t, seed=4, precision=vmc.FP64)
    for name in ("p3_direct_abs_fd", "p2_iso_gauss_fd"):
        p = _golden(name)
        with vmc.Problem(p) as prob:
            prob.run(1000, seed=1, f=p.f, phase0=p.phase0, precision=vmc.FP32)
            _compare(vmc, prob, 50_000, t, seed=5, f=p.f, phase0=p.phase0, precision=vmc.FP32)


def test_noalloc_follows_a_new_medium(vmc):
    """A medium changed after the first run is uploaded into the records'
    existing allocation."""
    p = _golden("p3_pencil_sphere")
    with vmc.Problem(p) as prob:
        prob.run(1000, seed=1, precision=vmc.FP32)
        prob.set_medium(mua=np.asarray(p.mua) * 1.5)
        _compare(vmc, prob, 100_000, vmc.TALLY_FIXED64, seed=6, precision=vmc.FP32)


def test_noalloc_refuses_what_it_cannot_do_without_allocating(vmc):
    p = _golden("p3_pencil_sphere")
    with vmc.Problem(p) as prob:
        # no run of this precision yet: the records are not on the device
        with pytest.raises(vmc.VmcError) as ei:
            prob.run(1000, precision=vmc.FP32, allocate=False)
        assert ei.value.status == vmc.STATUS_INVALID_ARGUMENT
    big = meshgen.sphere_box_3d(nphoton=1, seed=1234)
    with vmc.Problem(big) as prob:
        prob.run(1000, precision=vmc.FP32)
        # 64 explicit replicas of the 384k-element tallies exceed the scratch
        with pytest.raises(vmc.VmcError) as ei:
            prob.run(1000, precision=vmc.FP32, replicas=64, allocate=False)
        assert ei.value.status == 101
        # and the scratch is still good afterwards
        _compare(vmc, prob, 200_000, vmc.TALLY_FIXED64, seed=8, precision=vmc.FP32)
