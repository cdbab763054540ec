"""Serving loop: requests admitted while others run give the same result as
when run alone (independent contexts, deterministic kernels: bit for bit in
FP64 and in mixed precision)."""
import copy

import numpy as np
import pytest

from flurrypp_amd.serve import Request, Server, poisson_mix

pytestmark = pytest.mark.gpu


def test_results_same_alone_and_in_the_mix():
    reqs = poisson_mix(10, rate=50.0, seed=7, steady_tol=1e-4, unsteady_steps=30, n3d=(3, 4), p3d=(2, 3),
                       nx2d=(8, 16), ny2d=(4, 6), p2d=(1, 3), max_steps=600)
    mix = Server(max_active=6, steps_per_tick=5, check_every=10, keep_results=True, pmg=True)
    stats = mix.run(copy.deepcopy(reqs))
    assert stats["requests"] == len(reqs)
    assert stats["p99_s"] >= stats["p50_s"] > 0
    got = {r.rid: r for r in mix.done}
    # overlap happened: some request was admitted before another finished
    spans = sorted((r.admitted, r.finished) for r in mix.done)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    for r in reqs[:4]:
        alone = Server(max_active=1, steps_per_tick=5, check_every=10, keep_results=True, pmg=True)
        r1 = copy.deepcopy(r)
        r1.arrival = 0.0
        alone.run([r1])
        a = alone.done[0]
        b = got[r.rid]
        assert a.steps == b.steps and a.status == b.status
        assert np.array_equal(a.result, b.result, equal_nan=True), f"request {r.rid} ({r.kind}, {r.precision}) differs"


def test_admission_respects_budget():
    reqs = [Request(i, "bump2d", (12, 6), 2, 0.0, steady_tol=None, n_steps=20) for i in range(6)]
    srv = Server(hbm_budget=1.0, max_active=8, steps_per_tick=5)  # budget admits one at a time
    srv.run(reqs)
    spans = sorted((r.admitted, r.finished) for r in srv.done)
    assert all(spans[i + 1][0] >= spans[i][1] - 1e-9 for i in range(len(spans) - 1))
