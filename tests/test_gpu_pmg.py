"""P-multigrid on the device (fhip_pmg_cycle) against the oracle's cycle:
FP64 paths within 1e-12 relative L2 after several update + cycle
iterations, the mixed path within 1e-6."""
import numpy as np
import pytest

from flurrypp_amd import cases
from flurrypp_amd._native import PATH_AUTO, PATH_FUSED, PATH_MIXED, PATH_STAGED
from flurrypp_amd.mesh import pmg_operators
from flurrypp_amd.solver import MultiGrid, Solver
from oracle.oracle import Oracle, pmg_cycle

pytestmark = pytest.mark.gpu

CASES = {
    "bump2d_p3_levels1-3": (lambda: cases.bump2d(nx=10, ny=4, order=3, mach=0.3), 1, 1),
    "bump2d_p2_levels0-2_smooth2": (lambda: cases.bump2d(nx=8, ny=4, order=2, mach=0.4), 0, 2),
    "vortex3d_p2_levels1-2": (lambda: cases.vortex3d(n=3, order=2), 1, 1),
    "tgv3d_ns_p2_levels1-2": (lambda: cases.tgv3d(n=3, order=2, viscous=1), 1, 1),
    "bump2d_p5_levels0-5": (lambda: cases.bump2d(nx=6, ny=5, order=5, mach=0.39), 0, 1),
}


def _run(name, path, iters=4):
    make, low, smooth = CASES[name]
    setup, U = make()
    print(f"pmg case {name} path {path}", flush=True)
    fine = Solver(setup.params, setup=setup)
    fine.initializeSolution(U)
    mg = MultiGrid(fine, low_order=low, smooth_steps=smooth)
    ors = []
    for s in mg.levels:
        o = Oracle.from_setup(s.host)
        o.initializeSolution(s.host.initial_condition() if s is not fine else U, s.params.dtType)
        ors.append(o)
    D = setup.nDims
    R = [pmg_operators(D, s.order)[1] for s in mg.levels[1:]]
    Pr = [pmg_operators(D, s.order)[0] for s in mg.levels[:-1]]
    for _ in range(iters):
        fine.update(1, path=path)
        mg.cycle(path)
        ors[-1].update(1)
        pmg_cycle(ors, R, Pr, smooth)
    out = []
    for s, o in zip(mg.levels, ors):
        a, b = s.get_U(), o.get_U()
        out.append(np.linalg.norm(a - b) / np.linalg.norm(b))
    mg.close()
    fine.close()
    return out


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("path", [PATH_STAGED, PATH_FUSED, PATH_AUTO])
def test_pmg_cycle_matches_oracle(name, path):
    errs = _run(name, path)
    assert max(errs) <= 1e-12, errs


@pytest.mark.parametrize("name", ["bump2d_p3_levels1-3", "tgv3d_ns_p2_levels1-2"])
def test_pmg_cycle_mixed_precision(name):
    errs = _run(name, PATH_MIXED)
    print(f"mixed pmg {name}: {errs}")
    assert max(errs) <= 1e-6, errs


@pytest.mark.parametrize("path", [PATH_FUSED, PATH_MIXED, PATH_STAGED])
def test_pmg_cycle_graph_replay_matches_eager(path):
    """From the third cycle on, fhip_pmg_cycle replays one graph captured over
    all levels' streams; every replay equals the launch-by-launch cycle bit for
    bit, with eager work and synchronisation between the replays."""
    from flurrypp_amd import _native as nat
    runs = []
    for graphs in (1, 0):
        setup, U = cases.bump2d(nx=10, ny=5, order=3, mach=0.35, time_type=6, cfl=0.9)
        fine = Solver(setup.params, setup=setup)
        fine.initializeSolution(U)
        mg = MultiGrid(fine, low_order=0 if path != PATH_FUSED else 1, smooth_steps=2)
        for s in mg.levels:
            nat.hip_check(nat.hip_lib().fhip_set_option(s._ctx, 1, graphs))
        states = []
        for _ in range(6):
            fine.update(1, path=path)
            mg.cycle(path)
            states.append((fine.get_U(), fine.residual_norm(2), mg.coarse[0].get_U()))
        runs.append(states)
        mg.close()
        fine.close()
    for i, (a, b) in enumerate(zip(*runs)):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), f"cycle {i}"
