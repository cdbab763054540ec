"""Batched solves (flurrypp_amd.batch over fhip_update_many): every mesh in a
batch evolves exactly as it does alone, and steady runs stop per mesh."""
import numpy as np
import pytest

from flurrypp_amd import cases
from flurrypp_amd._native import PATH_FUSED, PATH_MIXED
from flurrypp_amd.batch import BatchedSolver, Member
from flurrypp_amd.solver import Solver

pytestmark = pytest.mark.gpu


def members():
    out = []
    for nx, ny, P, M in ((12, 5, 2, 0.3), (16, 6, 2, 0.45), (10, 4, 3, 0.2), (14, 5, 3, 0.35), (8, 4, 1, 0.5)):
        s, U = cases.bump2d(nx=nx, ny=ny, order=P, mach=M, seed=nx * 31 + ny)
        out.append(Member(s, U, ("bump", nx, ny, P, M)))
    for n in (6, 9):  # global CFL dt per mesh
        s, U = cases.vortex2d(n=n, order=3)
        out.append(Member(s, U, ("vortex", n)))
    return out


@pytest.mark.parametrize("path", [PATH_FUSED, PATH_MIXED])
def test_batch_equals_alone(path):
    ms = members()
    b = BatchedSolver(ms, path=path)
    b.step(12)
    got = b.states()
    res = b.residuals(2)
    for i, m in enumerate(ms):
        s = Solver(m.setup.params, setup=m.setup)
        s.initializeSolution(m.U)
        s.update(12, path=path)
        ref = s.get_U()
        assert np.array_equal(got[i], ref), f"member {m.tag} differs (max {np.abs(got[i] - ref).max():.3e})"
        r_alone = s.residual_norm(2)
        assert np.allclose(res[i], r_alone, rtol=1e-10, atol=0), (m.tag, res[i], r_alone)
        s.close()
    b.close()


def test_steady_stop_per_mesh():
    ms = []
    for nx, ny, P, M in ((8, 4, 1, 0.3), (8, 4, 2, 0.4), (10, 4, 2, 0.25)):
        s, U = cases.bump2d(nx=nx, ny=ny, order=P, mach=M, seed=3 + nx)
        ms.append(Member(s, U))
    b = BatchedSolver(ms)
    out = b.solve_steady(tol=1e-4, max_steps=20000, check_every=100)
    assert out["converged"].all(), out
    states = b.states()
    for i in range(len(ms)):
        assert np.isfinite(states[i]).all()
        assert out["residual"][i][0] <= 1e-4
    # frozen members keep their state
    before = b.states()
    b.step(50)
    after = b.states()
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    b.close()


@pytest.mark.parametrize("merged", [True, False])
def test_batch_pmg_equals_alone(merged):
    """Batched P-multigrid gives each mesh the same state as its own Solver +
    MultiGrid, bit for bit: with one cycle over coarse levels merged across
    the orders (fhip_pmg_create_batch), and with a cycle per order group."""
    from flurrypp_amd.solver import MultiGrid
    ms = []
    for nx, ny, P, M in ((8, 4, 2, 0.3), (10, 4, 2, 0.4), (8, 5, 3, 0.25), (6, 4, 1, 0.35)):
        s, U = cases.bump2d(nx=nx, ny=ny, order=P, mach=M, seed=nx * 7 + ny)
        ms.append(Member(s, U, ("bump", nx, ny, P, M)))
    b = BatchedSolver(ms, path=PATH_FUSED, pmg=True, low_order=0, merged_cycle=merged)
    assert bool(b._merged) == merged
    b.step(6)
    got = b.states()
    for i, m in enumerate(ms):
        s = Solver(m.setup.params, setup=m.setup)
        s.initializeSolution(m.U)
        mg = MultiGrid(s, low_order=0, smooth_steps=1)
        for _ in range(6):
            s.update(1, path=PATH_FUSED)
            mg.cycle(PATH_FUSED)
        ref = s.get_U()
        assert np.array_equal(got[i], ref), f"member {m.tag} differs (max {np.abs(got[i] - ref).max():.3e})"
        mg.close()
        s.close()
    b.close()


def test_batch_pmg_steady_converges():
    ms = []
    for nx, ny, P, M in ((8, 4, 1, 0.3), (8, 4, 2, 0.4), (10, 4, 3, 0.25)):
        s, U = cases.bump2d(nx=nx, ny=ny, order=P, mach=M, seed=5 + nx)
        ms.append(Member(s, U))
    b = BatchedSolver(ms, pmg=True, low_order=0)
    out = b.solve_steady(tol=1e-8, max_steps=6000, check_every=100)
    print("pmg steady steps", out["steps"], [r[0] for r in out["residual"]])
    assert out["converged"].all(), out
    b.close()


@pytest.mark.parametrize("path,low", [(PATH_FUSED, 0), (PATH_MIXED, 0), (PATH_FUSED, 1)])
def test_batch_merged_cycle_equals_group_cycles(path, low):
    """The merged cycles and one cycle per order group give the same states
    bit for bit, over orders 1, 2 and 4 with two smoothing steps, a gap in the
    orders (no P3 group: the P4 cycle's order-2 level holds P4 meshes only)
    and members frozen by a steady stop."""
    ms = []
    for nx, ny, P, M in ((8, 4, 1, 0.3), (10, 5, 2, 0.4), (9, 4, 2, 0.25), (7, 4, 4, 0.35), (6, 5, 4, 0.2)):
        s, U = cases.bump2d(nx=nx, ny=ny, order=P, mach=M, seed=nx * 13 + ny, time_type=6, cfl=0.9)
        ms.append(Member(s, U, ("bump", nx, ny, P, M)))
    out = []
    for merged in (True, False):
        b = BatchedSolver(ms, path=path, pmg=True, low_order=low, smooth_steps=2, merged_cycle=merged)
        # timeType 6: orders <= 2 and > 2 have their own RK coefficients, so
        # two merged cycles (P1 and P2; P4); with lowOrder 1 the P1 group has
        # no cycle
        fams = [[1, 2], [4]] if low == 0 else [[2], [4]]
        assert [[b.groups[gi].m.order for gi in mc["fine"]] for mc in b._merged] == (fams if merged else [])
        r = b.solve_steady(tol=1e-6, max_steps=300, check_every=20)
        out.append((b.states(), r["steps"], r["converged"]))
        b.close()
    (s1, k1, c1), (s2, k2, c2) = out
    assert list(k1) == list(k2) and list(c1) == list(c2)
    print("steps", list(k1), "converged", list(c1))
    for i, (x, y) in enumerate(zip(s1, s2)):
        assert np.array_equal(x, y), f"member {ms[i].tag} differs (max {np.abs(x - y).max():.3e})"


def test_batch_multigrid_rejects_mismatched_levels():
    """fhip_pmg_create_batch checks that every coarse level holds the meshes of
    the fine level one order up and of the coarse level above it, and that the
    fine orders ascend; a refused hierarchy leaves no object behind."""
    import ctypes as C
    from flurrypp_amd import _native as nat
    from flurrypp_amd.batch import BatchGroup, MergedSetup
    from flurrypp_amd.mesh import pmg_operators
    from flurrypp_amd.solver import HostSetup
    import copy
    ms = []
    for nx, ny, P in ((8, 4, 1), (6, 4, 2)):
        s, U = cases.bump2d(nx=nx, ny=ny, order=P, mach=0.3, seed=nx)
        ms.append(Member(s, U))
    def level(members, q):
        cm = []
        for m in members:
            pc = copy.copy(m.setup.params)
            pc.order = q
            hs = HostSetup(pc, m.setup.mesh)
            cm.append(Member(hs, hs.initial_condition()))
        g = BatchGroup(MergedSetup(cm))
        g.initialize()
        return g
    f1, f2 = BatchGroup(MergedSetup([ms[0]])), BatchGroup(MergedSetup([ms[1]]))
    f1.initialize()
    f2.initialize()
    good = [level(ms, 0), level(ms[1:], 1)]     # order 0: both meshes; order 1: the P2 mesh
    bad = [level(ms[1:], 0), level(ms[1:], 1)]  # order 0 misses the P1 mesh
    R = [pmg_operators(2, q + 1, ms[0].setup.params.sptsTypeQuad)[1] for q in (0, 1)]
    Pr = [pmg_operators(2, q, ms[0].setup.params.sptsTypeQuad)[0] for q in (0, 1)]
    Ra = (nat._dp * 2)(*[nat.dptr(a) for a in R])
    Pa = (nat._dp * 2)(*[nat.dptr(a) for a in Pr])
    lib = nat.hip_lib()
    def create(fine, coarse):
        h = C.c_void_p()
        fa = (C.c_void_p * len(fine))(*[g._ctx.value for g in fine])
        ca = (C.c_void_p * len(coarse))(*[g._ctx.value for g in coarse])
        rc = lib.fhip_pmg_create_batch(fa, len(fine), ca, len(coarse), Ra, Pa, 1, C.byref(h))
        return rc, h
    rc, h = create([f1, f2], bad)
    assert rc != 0 and not h.value and b"must hold the meshes" in lib.fhip_last_error()
    rc, h = create([f2, f1], good)
    assert rc != 0 and not h.value and b"ascend" in lib.fhip_last_error()
    rc, h = create([f1, f2], good)
    assert rc == 0 and h.value
    nat.hip_check(lib.fhip_pmg_cycle(h, PATH_FUSED))
    lib.fhip_pmg_destroy(h)
    for g in good + bad + [f1, f2]:
        g.close()


def test_steady_stops_at_max_steps():
    # max_steps not a multiple of check_every: the last chunk is shortened
    ms = members()[:3]
    b = BatchedSolver(ms)
    out = b.solve_steady(tol=1e-30, max_steps=130, check_every=50)
    assert list(out["steps"]) == [130] * len(ms), out["steps"]
    got = b.states()
    b.close()
    s = Solver(ms[0].setup.params, setup=ms[0].setup)
    s.initializeSolution(ms[0].U)
    s.update(130, path=PATH_FUSED)
    assert np.array_equal(got[0], s.get_U())
    s.close()
