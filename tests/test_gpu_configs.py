"""The configured workloads at their own sizes.

* Taylor-Green Navier-Stokes at 64^3 P3 (16.8 M DOF): the mixed-precision
  path against the FP64 path from the same state (density relative L2,
  kinetic-energy dissipation rate), and conservation on both.
* Batched Gaussian-bump channels drawn from the configured ranges (16..128 x
  8..48 quads, P 1..5, Mach 0.2-0.5, local CFL): every member of one batch
  against the oracle run alone; 512 cases in one batch: a sample equal to its
  runs alone, and the free stream held to round-off by all 512 in straight
  channels.
* The served mix at 8..40^3 vortex columns and the bump range: every result
  the same alone and in the mix, at two arrival rates.
"""
import copy

import numpy as np
import pytest

from flurrypp_amd import cases
from flurrypp_amd._native import PATH_FUSED, PATH_MIXED
from flurrypp_amd.batch import BatchedSolver, Member
from flurrypp_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _moments(setup, U):
    """Mass, momentum and energy integrals sum w detJ U (FP64)."""
    wj = np.einsum("s,se->se", setup.opers.w_spts, setup.geo.detjac_spts)
    return np.einsum("se,sek->k", wj, U)


def test_tgv_64_mixed_against_fp64():
    print("TGV 64^3 P3 NS: FP64 and mixed, 10 RK44 steps each", flush=True)
    setup, U = cases.tgv3d(n=64, order=3, viscous=1)
    s = Solver(setup.params, setup=setup)
    out = {}
    for name, path in (("fp64", PATH_FUSED), ("mixed", PATH_MIXED)):
        s.set_U(U)
        s.update(10, path=path)
        s.check_state()
        out[name] = s.get_U()
    s.close()
    U64, U32 = out["fp64"], out["mixed"]
    rho = _rel(U32[..., 0], U64[..., 0])
    print(f"density rel L2 mixed vs FP64: {rho:.3e}", flush=True)
    assert rho <= 1e-6
    assert _rel(U32, U64) <= 1e-6
    # kinetic-energy dissipation rate over the 10 steps
    from bench import kinetic_energy
    T = 10 * setup.params.dt
    ek0 = kinetic_energy(setup, U)
    eps64 = -(kinetic_energy(setup, U64) - ek0) / T
    eps32 = -(kinetic_energy(setup, U32) - ek0) / T
    print(f"KE dissipation rate: FP64 {eps64:.6e}, mixed {eps32:.6e}", flush=True)
    assert eps64 > 0 and abs(eps32 - eps64) <= 1e-3 * eps64
    # conservation: the periodic box keeps its mass and energy and its zero
    # total momentum, up to the drift of each path's rounding: per element,
    # sum_s w r_s is a sum of face fluxes of size |F| ~ p ~ 71 rounded at 1e-16
    # (FP64) or 6e-8 (FP32 residuals), and the roundings do not cancel across
    # elements (FP64 at 48^3: mass 6e-11, energy 1e-12 relative over 5 steps)
    m0 = _moments(setup, U)
    wj = np.einsum("s,se->se", setup.opers.w_spts, setup.geo.detjac_spts)
    mom_scale = float(np.einsum("se,se->", wj, np.abs(U[..., 1])))  # integral of |rho u|
    for name, V in out.items():
        m = _moments(setup, V)
        drift = np.abs(m[[0, 4]] - m0[[0, 4]]) / np.abs(m0[[0, 4]])
        mom = np.abs(m[1:4]) / mom_scale
        print(f"{name}: mass/energy drift {drift}, momentum {mom} of its scale", flush=True)
        assert np.all(drift <= (1e-9 if name == "fp64" else 1e-8)), (name, drift)
        assert np.all(mom <= (1e-12 if name == "fp64" else 1e-9)), (name, mom)


def _bump_members(n, seed, height=0.0625, pmin=1):
    """n channels drawn from the configured ranges; orders cycle through
    pmin..5 so every order is present."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nx, ny = int(rng.integers(16, 129)), int(rng.integers(8, 49))
        M = float(rng.uniform(0.2, 0.5))
        P = pmin + i % (6 - pmin)
        s, U = cases.bump2d(nx=nx, ny=ny, order=P, mach=M, seed=seed * 1009 + i, time_type=6, cfl=0.9,
                            height=height)
        out.append(Member(s, U, (nx, ny, P, M)))
    return out


def test_batched_members_match_oracle():
    from oracle.oracle import Oracle
    ms = _bump_members(5, seed=41)
    print("batch of 5 channels vs oracle:", [m.tag[:3] for m in ms], flush=True)
    b = BatchedSolver(ms, path=PATH_FUSED)
    b.step(10)
    got = b.states()
    b.close()
    for i, m in enumerate(ms):
        o = Oracle.from_setup(m.setup)
        o.initializeSolution(m.U, m.setup.params.dtType)
        o.update(10)
        ref = o.get_U()
        err = _rel(got[i], ref)
        print(f"member {m.tag[:3]}: rel L2 vs oracle {err:.3e}", flush=True)
        assert err <= 1e-12, (m.tag, err)


def test_512_cases_in_one_batch():
    ms = _bump_members(512, seed=43)
    print(f"512 channels, {sum(m.setup.nEles * m.setup.nSpts for m in ms):,} DOF", flush=True)
    b = BatchedSolver(ms, path=PATH_FUSED)
    b.step(10)
    got = b.states()
    b.close()
    assert all(np.isfinite(x).all() for x in got)
    for i in (0, 129, 258, 511):
        m = ms[i]
        s = Solver(m.setup.params, setup=m.setup)
        s.initializeSolution(m.U)
        s.update(10, path=PATH_FUSED)
        assert np.array_equal(got[i], s.get_U()), f"member {i} {m.tag} differs from its run alone"
        s.close()
    # straight channels: the free stream is steady; every member keeps it
    flat = _bump_members(512, seed=44, height=0.0)
    b = BatchedSolver(flat, path=PATH_FUSED)
    b.step(10)
    res = b.residuals(2)
    st = b.states()
    b.close()
    worst = max(float(np.abs(x - m.U).max() / np.abs(m.U).max()) for x, m in zip(st, flat))
    print(f"free stream over 512 straight channels: worst state change {worst:.3e}, "
          f"worst residual {max(float(r.max()) for r in res):.3e}", flush=True)
    assert worst <= 1e-12
    assert max(float(r.max()) for r in res) <= 1e-11


@pytest.mark.parametrize("rate", [5.0, 200.0])
def test_served_mix_same_alone(rate):
    from flurrypp_amd.serve import Server, poisson_mix
    reqs = poisson_mix(8, rate=rate, seed=17, steady_tol=1e-6, unsteady_steps=20, n3d=(8, 40), p3d=(2, 4),
                       nx2d=(16, 128), ny2d=(8, 48), p2d=(1, 5), max_steps=300)
    # the smallest and the largest vortex columns of the range are in the mix
    for r, n, P in ((reqs[1], 8, 2), (reqs[6], 40, 4)):
        r.kind, r.size, r.order, r.steady_tol, r.n_steps, r.mach = "vortex3d", (n,), P, None, 20, 0.5
    print(f"served mix at {rate:g}/s:", [(r.kind, r.size, r.order, r.precision) for r in reqs], flush=True)
    mix = Server(max_active=6, steps_per_tick=5, check_every=50, keep_results=True, pmg=True)
    stats = mix.run(copy.deepcopy(reqs))
    assert stats["requests"] == len(reqs)
    got = {r.rid: r for r in mix.done}
    for r in reqs:
        alone = Server(max_active=1, steps_per_tick=5, check_every=50, keep_results=True, pmg=True)
        r1 = copy.deepcopy(r)
        r1.arrival = 0.0
        alone.run([r1])
        a, m = alone.done[0], got[r.rid]
        assert a.steps == m.steps and a.status == m.status, (r.rid, a.steps, m.steps, a.status, m.status)
        assert np.array_equal(a.result, m.result, equal_nan=True), f"request {r.rid} ({r.kind} {r.size})"
