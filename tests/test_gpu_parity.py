"""Parity of the HIP path (libflurry_hip, through its C ABI) with the CPU
restatement in oracle/, on the same host set-up and initial state.

Bars: index maps bit-exact; FP64 arrays within 1e-13 relative L2 for one
residual evaluation and 1e-12 relative L2 for the state after N full RK steps
(the reference's own -O2/-Ofast builds differ by ~1e-15 after 100 steps).
"""
import numpy as np
import pytest

from flurrypp_amd import cases
from flurrypp_amd._native import PATH_FUSED, PATH_STAGED
from flurrypp_amd.solver import Solver
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL_STAGE = 1e-13
TOL_TRAJ = 1e-12


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


CASES = {
    "vortex2d_p3_jitter_rusanov": lambda: cases.vortex2d(n=8, order=3),
    "vortex2d_p2_roe": lambda: cases.vortex2d(n=6, order=2, riemann=1),
    "vortex2d_p1_dt_fixed": lambda: cases.vortex2d(n=5, order=1, dt_type=0),
    "vortex2d_p4_jameson": lambda: cases.vortex2d(n=4, order=4, time_type=6),
    "tgv3d_p3_ns_jitter": lambda: cases.tgv3d(n=3, order=3, viscous=1, jitter=0.1),
    "tgv3d_p2_euler_jitter": lambda: cases.tgv3d(n=4, order=2, viscous=0, jitter=0.1),
    "bump2d_p2_char_slip_local_dt": lambda: cases.bump2d(nx=12, ny=5, order=2),
    "bump2d_p3_ns_adiabatic": lambda: _bump_ns(),
    "bump2d_p0_char_slip_local_dt": lambda: cases.bump2d(nx=12, ny=5, order=0),
    "tgv3d_p0_ns_jitter": lambda: cases.tgv3d(n=4, order=0, viscous=1, jitter=0.1),
    "vortex2d_p5_jitter": lambda: cases.vortex2d(n=4, order=5),
    "bump2d_p5_char_slip_local_dt": lambda: cases.bump2d(nx=6, ny=4, order=5),
    "tgv3d_p5_ns_jitter": lambda: cases.tgv3d(n=2, order=5, viscous=1, jitter=0.1),
    "box3d_p2_ns_walls": lambda: _box_ns(),
    "bump2d_p2_roe_subin_subout_sym": lambda: _channel_subsonic(),
    "box3d_p2_euler_supin_supout_sym": lambda: _box_supersonic(),
}


def _perturb(U):
    return U * (1 + 0.01 * np.sin(np.arange(U.size)).reshape(U.shape))


def _channel_subsonic():
    """2-D Euler bump channel with subsonic inflow and outflow ends (sub_in,
    sub_out), a symmetry floor and a slip-wall lid; Roe flux at interior
    points (boundary points take Rusanov, src/face.cpp:174-177)."""
    from flurrypp_amd.input import InputParams
    from flurrypp_amd.mesh import Mesh, bc_number
    from flurrypp_amd.solver import HostSetup
    g = 1.4
    p = InputParams.from_dict(dict(equation=1, nDims=2, order=2, meshType=0, meshFileName="-", timeType=4, dtType=1,
                                   CFL=0.5, maxTime=1e9, iterMax=10, icType=0, riemannType=1, gamma=g, rhoBound=1.0,
                                   uBound=0.4, vBound=0.0, pBound=1.0 / g, mesh_bound={"fluid": "fluid"}))
    nx, ny = 10, 4
    xv, c2v = cases.bump_mesh(nx, ny, 0.1).nodes()
    vid = lambda i, j: j * (nx + 1) + i  # noqa: E731
    sides = [("sub_in", [vid(0, j) for j in range(ny + 1)]), ("sub_out", [vid(nx, j) for j in range(ny + 1)]),
             ("symmetry", [vid(i, 0) for i in range(nx + 1)]), ("slip_wall", [vid(i, ny) for i in range(nx + 1)])]
    m = Mesh.from_arrays(xv, c2v, [(bc_number(b), np.array(v)) for b, v in sides])
    s = HostSetup(p, m)
    return s, _perturb(s.initial_condition())


def _box_supersonic():
    """3-D Euler box at Mach 1.5 with supersonic inflow and outflow ends
    (sup_in, sup_out), symmetry floor and lid, slip side walls."""
    from flurrypp_amd.input import InputParams
    from flurrypp_amd.mesh import Mesh
    from flurrypp_amd.solver import HostSetup
    g = 1.4
    p = InputParams.from_dict(dict(equation=1, nDims=3, order=2, meshType=1, nx=3, ny=3, nz=3, xmin=0, xmax=1, ymin=0,
                                   ymax=1, zmin=0, zmax=1, timeType=4, dtType=1, CFL=0.5, maxTime=1e9, iterMax=10,
                                   icType=0, riemannType=0, gamma=g, rhoBound=1.0, uBound=1.5 * np.sqrt(g),
                                   vBound=0.1, wBound=-0.05, pBound=1.0))
    m = Mesh.cartesian(3, (3, 3, 3), (0, 0, 0), (1, 1, 1),
                       bcs={"left": "sup_in", "right": "sup_out", "bottom": "symmetry", "top": "symmetry",
                            "front": "slip_wall", "back": "slip_wall"})
    s = HostSetup(p, m)
    return s, _perturb(s.initial_condition())


def _bump_ns():
    """Viscous channel with adiabatic no-slip walls and characteristic ends."""
    from flurrypp_amd.input import InputParams
    from flurrypp_amd.mesh import bc_number
    from flurrypp_amd.solver import HostSetup
    p = InputParams.from_dict(dict(equation=1, nDims=2, order=3, meshType=0, meshFileName="-", timeType=4, dtType=1,
                                   CFL=0.5, maxTime=1e9, iterMax=10, icType=0, viscous=1, Re=200, MachBound=0.3,
                                   MachWall=0.0, fixVis=0, LDG_penFact=0.5, LDG_tau=1.0,
                                   mesh_bound={"fluid": "fluid"}))
    m = cases.bump_mesh(10, 4, 0.1)
    xv, c2v = m.nodes()
    # same channel, walls adiabatic no-slip
    from flurrypp_amd.mesh import Mesh
    nx, ny = 10, 4
    vid = lambda i, j: j * (nx + 1) + i  # noqa: E731
    walls = [vid(i, 0) for i in range(nx + 1)] + [vid(i, ny) for i in range(nx + 1)]
    ends = [vid(0, j) for j in range(ny + 1)] + [vid(nx, j) for j in range(ny + 1)]
    m2 = Mesh.from_arrays(xv, c2v, [(bc_number("char"), np.array(ends)),
                                    (bc_number("adiabatic_noslip"), np.array(walls))])
    s = HostSetup(p, m2)
    return s, s.initial_condition()


def _box_ns():
    """3-D viscous box with every viscous boundary kind: isothermal no-slip
    floor, moving isothermal lid, adiabatic no-slip and slip side walls,
    characteristic ends; Sutherland viscosity, LDG penalty 0.5."""
    from flurrypp_amd.input import InputParams
    from flurrypp_amd.mesh import Mesh
    from flurrypp_amd.solver import HostSetup
    p = InputParams.from_dict(dict(equation=1, nDims=3, order=2, meshType=1, nx=3, ny=3, nz=3, xmin=0, xmax=1, ymin=0,
                                   ymax=1, zmin=0, zmax=1, timeType=4, dtType=1, CFL=0.5, maxTime=1e9, iterMax=10,
                                   icType=0, viscous=1, Re=100, MachBound=0.3, MachWall=0.1, fixVis=0,
                                   LDG_penFact=0.5, LDG_tau=1.0, TWall=320.0))
    m = Mesh.cartesian(3, (3, 3, 3), (0, 0, 0), (1, 1, 1),
                       bcs={"bottom": "isothermal_noslip", "top": "isothermal_noslip_moving",
                            "front": "adiabatic_noslip", "back": "slip_wall", "left": "char", "right": "char"})
    s = HostSetup(p, m)
    U = s.initial_condition()
    return s, U * (1 + 0.01 * np.sin(np.arange(U.size)).reshape(U.shape))


@pytest.fixture(scope="module", params=list(CASES))
def case(request):
    setup, U = CASES[request.param]()
    return request.param, setup, U


def test_face_maps_bit_exact(case):
    name, setup, U = case
    s = Solver(setup.params, setup=setup)
    o = Oracle.from_setup(setup)
    for m in ("face_fpt_L", "face_fpt_R"):
        assert np.array_equal(s.get(m), o.get(m)), f"{name}: {m} differs"
    s.close()


def test_one_residual_every_array(case):
    name, setup, U = case
    p = setup.params
    s = Solver(p, setup=setup)
    o = Oracle.from_setup(setup)
    s.initializeSolution(U)
    o.initializeSolution(U, p.dtType)
    if p.dtType != 0:
        assert abs(s.calcDt(True) - o.calcDt(True)) <= 1e-14 * abs(o.calcDt(True))
    s.calcResidual(0)
    o.calcResidual(0)
    names = ["U_fpts", "F_spts", "Fn_fpts", "disFn_fpts", "divF_spts", "waveSp"]
    if p.viscous:
        names += ["dU_spts", "dU_fpts", "dUc_fpts"]
    # Arrays formed by cancellation are measured against the size of the
    # terms that cancel: disFn = Fn - (extrapolated F.n), divF = div(F).
    # Gradients of fields with a large mean (the energy) likewise.
    scale = {"disFn_fpts": "Fn_fpts", "divF_spts": "F_spts", "dUc_fpts": "U_fpts", "dU_spts": "U_spts",
             "dU_fpts": "U_fpts"}
    for a in names:
        ref = o.get(a)
        den = max(np.linalg.norm(ref), np.linalg.norm(o.get(scale[a])) if a in scale else 0.0)
        err = float(np.linalg.norm(s.get(a) - ref) / max(den, 1e-300))
        print(f"{name} {a} rel_l2={err:.3e}", flush=True)
        assert err < TOL_STAGE, f"{name}: {a} rel L2 {err:.3e}"


@pytest.mark.parametrize("path", [PATH_STAGED, PATH_FUSED])
def test_trajectory(case, path):
    name, setup, U = case
    p = setup.params
    s = Solver(p, setup=setup)
    if path == PATH_FUSED and not s._fused_ok():
        pytest.skip("fused path does not cover this configuration")
    o = Oracle.from_setup(setup)
    s.initializeSolution(U)
    o.initializeSolution(U, p.dtType)
    nsteps = 10
    s.update(nsteps, path=path)
    o.update(nsteps)
    err = rel_l2(s.get_U(), o.get_U())
    print(f"{name} path={path} U rel_l2 after {nsteps} steps = {err:.3e}", flush=True)
    assert err < TOL_TRAJ
    if p.dtType == 1:
        # global CFL dt of the last step (src/solver.cpp:447-468)
        assert abs(s.dt - o.dt) <= 1e-12 * abs(o.dt), (s.dt, o.dt)
    elif p.dtType == 2:
        e_s, e_o = s.get("ele_dt"), o.get("ele_dt")
        assert rel_l2(e_s, e_o) <= 1e-12
    r_s, r_o = s.residual_norm(2), o.residual_norm(2)
    assert np.allclose(r_s, r_o, rtol=1e-10, atol=0)


MIXED_TOL = 1e-6  # relative L2 of the state against the FP64 oracle


@pytest.mark.parametrize("name", ["vortex2d_p3_jitter_rusanov", "tgv3d_p3_ns_jitter", "bump2d_p3_ns_adiabatic",
                                  "vortex2d_p2_roe", "box3d_p2_ns_walls"])
def test_mixed_precision_error(name):
    """Mixed path (FP32 fluxes/intermediates, FP64 state and RK sums) stays
    within 1e-6 relative L2 of the FP64 oracle after 10 RK steps."""
    from flurrypp_amd._native import PATH_MIXED
    setup, U = CASES[name]()
    p = setup.params
    s = Solver(p, setup=setup)
    o = Oracle.from_setup(setup)
    s.initializeSolution(U)
    o.initializeSolution(U, p.dtType)
    s.update(10, path=PATH_MIXED)
    o.update(10)
    a, b = s.get_U(), o.get_U()
    err = rel_l2(a, b)
    err_rho = rel_l2(a[..., 0], b[..., 0])
    print(f"{name} mixed: U rel_l2={err:.3e} rho rel_l2={err_rho:.3e}", flush=True)
    assert err < MIXED_TOL and err_rho < MIXED_TOL


def test_graph_replay_matches_eager():
    """The fused step replayed from a captured HIP graph equals the eager
    launches bit for bit, over several replays with host work between."""
    from flurrypp_amd import _native as nat
    setup, U = cases.vortex2d(n=8, order=3)
    a = Solver(setup.params, setup=setup)
    b = Solver(setup.params, setup=setup)
    nat.hip_lib().fhip_set_option(b._ctx, 1, 0)  # graphs off
    for x in (a, b):
        x.initializeSolution(U)
    for rep in range(4):
        a.update(3, path=PATH_FUSED)
        b.update(3, path=PATH_FUSED)
        assert np.array_equal(a.get_U(), b.get_U()), f"replay {rep}"


def test_granular_sequence_matches_update():
    """calcResidual/timeStepA/timeStepB/copy* called one by one, as
    solver::update does, equal fhip_update."""
    setup, U = cases.vortex2d(n=6, order=3)
    p = setup.params
    a = Solver(p, setup=setup)
    b = Solver(p, setup=setup)
    a.initializeSolution(U)
    b.initializeSolution(U)
    a.update(1, path=PATH_STAGED)
    # by hand
    nRK = p.nRKSteps
    b.calcDt(True)
    b.copyUspts_U0()
    for step in range(nRK - 1):
        b.calcResidual(step)
        b.timeStepA(step, p.RKa[step + 1])
    b.calcResidual(nRK - 1)
    b.copyU0_Uspts()
    for step in range(nRK):
        b.timeStepB(step)
    assert np.array_equal(a.get_U(), b.get_U())


def test_multigrid_source_term():
    """timeStepA/B with a PMG source term (src/solver.cpp:472-486, 506-520)."""
    setup, U = cases.vortex2d(n=5, order=2)
    p = setup.params
    s = Solver(p, setup=setup)
    o = Oracle.from_setup(setup)
    s.initializeSolution(U)
    o.initializeSolution(U, p.dtType)
    src = np.random.default_rng(1).standard_normal(U.shape) * 1e-3
    for x in (s, o):
        x.set_dt(1e-3)
        x.copyUspts_U0()
        x.calcResidual(0)
        x.timeStepA(0, 0.5, src)
        x.calcResidual(1)
        x.timeStepB(1, src)
    assert rel_l2(s.get_U(), o.get_U()) < 1e-14


@pytest.mark.parametrize("path", [PATH_STAGED, PATH_FUSED])
def test_vortex2d_full_size_100_steps(path):
    """The 2-D Euler vortex at its full configuration (32x32 jittered periodic
    quads, P3, Rusanov, RK44, global CFL 0.5) for 100 steps against the
    oracle: state within 1e-12 relative L2, residual norms within round-off."""
    setup, U = cases.vortex2d(n=32, order=3)
    p = setup.params
    print(f"vortex2d 32x32 P3 path {path}", flush=True)
    s = Solver(p, setup=setup)
    o = Oracle.from_setup(setup)
    s.initializeSolution(U)
    o.initializeSolution(U, p.dtType)
    s.update(100, path=path)
    o.update(100)
    err = rel_l2(s.get_U(), o.get_U())
    print(f"vortex2d 32x32 P3 path={path} U rel_l2 after 100 steps = {err:.3e}", flush=True)
    assert err < TOL_TRAJ  # a state this close after 100 steps also means the same dt at every step
    np.testing.assert_allclose(s.residual_norm(2), o.residual_norm(2), rtol=1e-10, atol=0)
    s.close()


# the BF16 matrix-core forms of the mixed path (z contractions of P3 hexes as
# products of BF16 pieces, FP32 accumulation) against the FP64 oracle after
# 10 RK steps; hi+lo keeps about 16 bits of each factor, hi+mid+lo 24
BF16_TOL = {4: 2e-5, 5: 1e-6}


@pytest.mark.parametrize("path", [4, 5])
@pytest.mark.parametrize("name", ["tgv3d_p3_ns_jitter", "vortex3d_p3"])
def test_mixed_bf16_error(name, path):
    if name == "vortex3d_p3":
        setup, U = cases.vortex3d(n=4, order=3)
    else:
        setup, U = CASES[name]()
    p = setup.params
    s = Solver(p, setup=setup)
    o = Oracle.from_setup(setup)
    s.initializeSolution(U)
    o.initializeSolution(U, p.dtType)
    s.update(10, path=path)
    o.update(10)
    a, b = s.get_U(), o.get_U()
    err, err_rho = rel_l2(a, b), rel_l2(a[..., 0], b[..., 0])
    print(f"{name} BF16 path {path}: U rel_l2={err:.3e} rho rel_l2={err_rho:.3e}", flush=True)
    assert err < BF16_TOL[path] and err_rho < BF16_TOL[path]


def test_bf16_paths_refuse_other_shapes():
    from flurrypp_amd._native import FhipError
    setup, U = cases.vortex2d(n=6, order=3)
    s = Solver(setup.params, setup=setup)
    s.initializeSolution(U)
    with pytest.raises(FhipError, match="P3 hexes"):
        s.update(1, path=4)
    s.close()
