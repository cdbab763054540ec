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


def test_steps_exact_when_ticks_do_not_divide_them():
    # 13 steps at 5 per tick: the last tick takes 3, and the result equals 13
    # updates of the request alone; a steady request is checked at multiples of
    # check_every although ticks are 7 steps
    from flurrypp_amd._native import PATH_FUSED
    from flurrypp_amd.serve import build_case
    from flurrypp_amd.solver import Solver
    un = Request(0, "vortex3d", (3,), 2, 0.0, n_steps=13)
    st = Request(1, "bump2d", (8, 4), 1, 0.0, steady_tol=1e-30, max_steps=25)
    srv = Server(max_active=2, steps_per_tick=7, check_every=10, keep_results=True)
    srv.run([copy.deepcopy(un), copy.deepcopy(st)])
    got = {r.rid: r for r in srv.done}
    assert got[0].steps == 13 and got[0].status == "done"
    assert got[1].steps == 25 and got[1].status == "max_steps"
    setup, U = build_case(un)
    s = Solver(setup.params, setup=setup)
    s.initializeSolution(U)
    s.update(13, path=PATH_FUSED)
    ref = s.get_U()
    s.close()
    assert np.array_equal(got[0].result, ref)


@pytest.mark.parametrize("kind,size,order", [("vortex3d", (12,), 3), ("bump2d", (48, 16), 4)])
def test_footprint_covers_what_a_context_allocates(kind, size, order):
    # admission by HBM budget relies on footprint_bytes bounding a context's
    # device memory once both precisions have run (graphs included)
    import ctypes as C
    from flurrypp_amd._native import PATH_FUSED, PATH_MIXED, hip_lib
    from flurrypp_amd.serve import build_case, footprint_bytes
    from flurrypp_amd.solver import Solver
    hip_lib()
    rt = C.CDLL("libamdhip64.so")

    def free():
        f, t = C.c_size_t(), C.c_size_t()
        assert rt.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    def run(setup, U):
        s = Solver(setup.params, setup=setup)
        s.initializeSolution(U)
        s.update(2, path=PATH_FUSED)
        s.update(2, path=PATH_MIXED)
        return s

    r = Request(0, kind, size, order, 0.0, n_steps=2)
    setup, U = build_case(r)
    run(setup, U).close()   # the runtime's one-time allocations (code objects, scratch) happen here
    f0 = free()
    s = run(setup, U)
    used = f0 - free()
    s.close()
    est = footprint_bytes(setup.nDims, setup.order, setup.nEles, bool(setup.params.viscous))
    print(f"footprint {kind} {size} P{order}: used {used / 2**20:.1f} MiB, estimate {est / 2**20:.1f} MiB", flush=True)
    assert used <= est, (used, est)
