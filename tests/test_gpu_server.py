"""Serving loop (vmc_server_*): continuous batching of forward solves.

A request's photons keep their stream ids however the server slices them
over launches and whatever shares the launches, so in fixed-point mode each
result equals vmc_problem_run's for the same inputs bit for bit: alone, in a
mix submitted together, and when submitted while others are running, under
both scheduling policies (fair shares, shortest remaining time first).  A
small arena forces requests to queue for admission.
"""
import os
import time

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


REQS = [("p3_pencil_sphere", dict(nphoton=600000, seed=3)),
        ("p3_direct_abs_fd", dict(nphoton=300000, seed=5)),
        ("p2_square_cosinic", dict(nphoton=500000, seed=2)),
        ("p3_cosinic_hetero_n", dict(nphoton=250000, seed=11, stream_tag=4)),
        ("p2_iso_gauss_fd", dict(nphoton=200000, seed=9)),
        ("p3_pencil_sphere", dict(nphoton=77777, seed=8))]


def _problems(vmc):
    probs = {}
    for name, _ in REQS:
        if name not in probs:
            p = _golden(name)
            probs[name] = (p, vmc.Problem(p))
    return probs


@pytest.mark.parametrize("arena,policy", [(0, "SERVE_SHORTEST"), (24 << 20, "SERVE_SHORTEST"),
                                          (0, "SERVE_FAIR"), (24 << 20, "SERVE_FAIR")])
def test_server_results_equal_single_runs(vmc, arena, policy):
    probs = _problems(vmc)
    try:
        singles = []
        for name, kw in REQS:
            p, prob = probs[name]
            kw = dict(kw, f=p.f, phase0=p.phase0)
            n = kw.pop("nphoton")
            singles.append(prob.run(n, precision=vmc.FP32, tally=vmc.TALLY_FIXED64, **kw))
        # small launches so requests are sliced; the second half arrives while
        # the first half runs
        with vmc.Server(precision=vmc.FP32, tally=vmc.TALLY_FIXED64, arena_bytes=arena,
                        launch_photons=200000, min_slice=50000, policy=getattr(vmc, policy)) as srv:
            tickets = []
            for k, (name, kw) in enumerate(REQS):
                p, prob = probs[name]
                kw = dict(kw, f=p.f, phase0=p.phase0)
                n = kw.pop("nphoton")
                tickets.append(srv.submit(prob, n, **kw))
                if k == 2:
                    time.sleep(0.05)
            results = [srv.wait(t) for t in tickets]
            c = srv.counters()
            print("server counters", c, flush=True)
            assert c["launches"] > len(REQS)   # requests were sliced over launches
        for (name, kw), a, b in zip(REQS, results, singles):
            for k in OUT:
                assert np.array_equal(getattr(a, k), getattr(b, k)), (name, kw, k)
            assert a.loss == b.loss
            # the request's own weight ledger, summed over its launches
            n = kw["nphoton"]
            led = np.asarray(a.ledger)
            assert abs(led[0] + led[1] + led[2] - led[3] + led[4] - n) <= 1e-6 * n, (name, led)
            assert np.allclose(led, b.ledger, rtol=1e-9, atol=1e-9 * n), (name, led, b.ledger)
    finally:
        for _, prob in probs.values():
            prob.close()


def test_server_refuses_request_larger_than_arena(vmc):
    p = _golden("p3_pencil_sphere")
    with vmc.Problem(p) as prob, vmc.Server(arena_bytes=4096) as srv:
        with pytest.raises(vmc.VmcError) as ei:
            srv.submit(prob, 1000)
        assert ei.value.status == 101


def test_server_policy_is_validated(vmc):
    with pytest.raises(vmc.VmcError):
        vmc.Server(policy=2)
