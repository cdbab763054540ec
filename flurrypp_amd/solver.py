"""Host mirror of FlurryPP's `solver` class over libflurry_hip.

``Solver`` keeps the reference's member names (setup, initializeSolution,
update, calcResidual, timeStepA, timeStepB, calcDt, copyUspts_U0,
copyU0_Uspts, extrapolateU) and the same array names, but every step runs in
libflurry_hip on the GPU.  Host set-up (mesh, operators, metrics, faces)
comes from libflurry_host.
"""
from __future__ import annotations

import copy
import ctypes as C
import math
from pathlib import Path

import numpy as np

from . import _native as nat
from .input import InputParams
from .mesh import Mesh, operators, pmg_operators, sizes

PI = 4.0 * math.atan(1.0)  # include/global.hpp:72


def build_mesh(p: InputParams) -> Mesh:
    """geo::setup for meshType 0 (Gmsh) or 1 (Cartesian), src/geo.cpp:63-117."""
    if p.meshType == 1:
        lo = (p.xmin, p.ymin, p.zmin)
        hi = (p.xmax, p.ymax, p.zmax)
        bcs = {"bottom": p.create_bcBottom, "right": p.create_bcRight, "top": p.create_bcTop,
               "left": p.create_bcLeft, "front": p.create_bcFront, "back": p.create_bcBack}
        m = Mesh.cartesian(p.nDims, (p.nx, p.ny, p.nz), lo, hi, bcs)
        # createMesh overrides the periodic lengths (src/geo.cpp:1502-1504)
        p.periodicDX, p.periodicDY = p.xmax - p.xmin, p.ymax - p.ymin
        p.periodicDZ = (p.zmax - p.zmin) if p.nDims == 3 else 0.0
    elif p.meshType == 0:
        m = Mesh.read_gmsh(p.meshFileName, p.meshBounds)
    else:
        raise ValueError("meshType must be 0 or 1")
    m.set_periodic(p.periodicDX, p.periodicDY, p.periodicDZ, p.periodicTol)
    return m


def fhip_params(p: InputParams, nEles: int) -> nat.FhipParams:
    S, F, Ff = sizes(p.nDims, p.order)
    fp = nat.FhipParams()
    fp.n_dims, fp.n_fields, fp.order = p.nDims, p.nFields, p.order
    fp.n_eles, fp.n_spts, fp.n_fpts, fp.n_fpts_face = nEles, S, F, Ff
    fp.equation, fp.viscous, fp.riemann_type = p.equation, p.viscous, p.riemannType
    fp.time_type, fp.dt_type, fp.n_rk_steps = p.timeType, p.dtType, p.nRKSteps
    fp.fix_vis, fp.init_iter = int(getattr(p, "fixVis", 1)), p.initIter
    for i, a in enumerate(p.RKa):
        fp.rk_a[i] = a
    for i, b in enumerate(p.RKb):
        fp.rk_b[i] = b
    fp.cfl = getattr(p, "CFL", 0.0)
    fp.dt = p.dt
    fp.gamma, fp.prandtl = p.gamma, getattr(p, "prandtl", 0.72)
    fp.mu_inf, fp.rt_inf, fp.c_sth = getattr(p, "mu_inf", 0.0), getattr(p, "rt_inf", 0.0), getattr(p, "c_sth", 0.0)
    fp.pen_fact, fp.tau = p.penFact, p.tau
    fp.rho_bound, fp.u_bound, fp.v_bound, fp.w_bound, fp.p_bound = p.rhoBound, p.uBound, p.vBound, p.wBound, p.pBound
    fp.t_wall, fp.r_gas = getattr(p, "TWall", 300.0), getattr(p, "RGas", 286.9)
    fp.u_wall, fp.v_wall, fp.w_wall = getattr(p, "uWall", 0.0), getattr(p, "vWall", 0.0), getattr(p, "wWall", 0.0)
    return fp


def initial_condition(p: InputParams, pos_spts: np.ndarray) -> np.ndarray:
    """ele::setInitialCondition (src/ele.cpp:733-832) on pos_spts [S][E][D]."""
    S, E, D = pos_spts.shape
    N = D + 2
    U = np.zeros((S, E, N))
    g = p.gamma
    x, y = pos_spts[..., 0], pos_spts[..., 1]
    if p.icType == 0:
        rho, vx, vy = p.rhoIC, p.vxIC, p.vyIC
        vz = p.vzIC if D == 3 else 0.0
        U[..., 0] = rho
        U[..., 1] = rho * vx
        U[..., 2] = rho * vy
        if D == 3:
            U[..., 3] = rho * vz
        U[..., D + 1] = p.pIC / (g - 1) + (0.5 * rho * (vx * vx + vy * vy + vz * vz))
    elif p.icType == 1:
        eps = 5.0
        f = 1.0 - (x * x + y * y)
        rho = np.maximum(np.power(1.0 - eps * eps * (g - 1.0) / (8.0 * g * PI * PI) * np.exp(f),
                                  1.0 / (g - 1.0) + 1e-5), 1e-3)
        vx = 1.0 - eps * y / (2.0 * PI) * np.exp(f / 2.0)
        vy = 1.0 + eps * x / (2.0 * PI) * np.exp(f / 2.0)
        pr = np.power(rho, g)
        U[..., 0], U[..., 1], U[..., 2] = rho, rho * vx, rho * vy
        U[..., D + 1] = pr / (g - 1) + (0.5 * rho * (vx * vx + vy * vy))
    elif p.icType == 2:
        eps, rc, Minf, Uinf, rhoInf = 1.0, 1.0, 0.3, 1.0, 1.0
        th = p.vortexAngle
        Pinf = Minf ** -2 / g
        eM = (eps * Minf) * (eps * Minf)
        f = -(x * x + y * y) / (rc * rc)
        vx = Uinf * (math.cos(th) - y * eps / rc * np.exp(f / 2.0))
        vy = Uinf * (math.sin(th) + x * eps / rc * np.exp(f / 2.0))
        rho = rhoInf * np.power(1.0 - (g - 1.0) / 2.0 * eM * np.exp(f), g / (g - 1.0))
        pr = Pinf * np.power(1.0 - (g - 1.0) / 2.0 * eM * np.exp(f), g / (g - 1.0))
        U[..., 0], U[..., 1], U[..., 2] = rho, rho * vx, rho * vy
        U[..., D + 1] = pr / (g - 1) + (0.5 * rho * (vx * vx + vy * vy))
    else:
        raise ValueError(f"icType {p.icType} not available")
    return U


class HostSetup:
    """Host half of solver::setup (src/solver.cpp:60-115): mesh connectivity,
    operators, metric terms, face list and the fhip_params block."""

    def __init__(self, params: InputParams, mesh: Mesh | None = None):
        p = params
        self.params = p
        self.mesh = mesh if mesh is not None else build_mesh(p)
        if not getattr(self.mesh, "processed", False):
            self.mesh.process()
        self.nDims, self.nFields, self.order = p.nDims, p.nFields, p.order
        self.nEles = self.mesh.nEles
        self.nSpts, self.nFpts, self.nFptsFace = sizes(self.nDims, self.order)
        self.opers = operators(self.nDims, self.order, p.sptsTypeQuad, p.vcjhSchemeQuad)
        self.geo = self.mesh.geometry(self.order, p.sptsTypeQuad)
        self.faces = self.mesh.faces()
        self.fp = fhip_params(p, self.nEles)

    def initial_condition(self) -> np.ndarray:
        return initial_condition(self.params, self.geo.pos_spts)


class Solver:
    """FR solver whose hot path runs in libflurry_hip."""

    def __init__(self, params: InputParams, mesh: Mesh | None = None, setup: HostSetup | None = None):
        self.host = setup if setup is not None else HostSetup(params, mesh)
        self.params = self.host.params
        self.setup()

    # solver::setup (src/solver.cpp:60-115): device half
    def setup(self) -> None:
        h = self.host
        self.mesh = h.mesh
        self.nDims, self.nFields, self.order, self.nEles = h.nDims, h.nFields, h.order, h.nEles
        self.nSpts, self.nFpts, self.nFptsFace = h.nSpts, h.nFpts, h.nFptsFace
        self.opers, self.geo, self.faces, self._fp = h.opers, h.geo, h.faces, h.fp
        f = self.faces
        o, g = self.opers, self.geo
        self._keep = [o.spts_to_fpts, o.grad_spts, o.extrap_fn, o.correction, o.correct_u, o.w_spts, o.w_face,
                      g.jginv_spts, g.detjac_spts, g.norm_fpts, g.da_fpts,
                      f.eleL, f.locL, f.eleR, f.locR, f.relRot, f.bc]
        d = nat.FhipDesc()
        d.params = C.pointer(self._fp)
        (d.opp_spts_to_fpts, d.opp_grad_spts, d.opp_extrapolate_fn, d.opp_correction, d.opp_correct_u,
         d.w_spts, d.w_face) = (nat.dptr(a) for a in self._keep[:7])
        d.jginv_spts, d.detjac_spts, d.norm_fpts, d.da_fpts = (nat.dptr(a) for a in self._keep[7:11])
        d.n_faces = len(f.eleL)
        (d.face_ele_l, d.face_loc_l, d.face_ele_r, d.face_loc_r, d.face_rel_rot,
         d.face_bc) = (nat.iptr(a) for a in self._keep[11:])
        self._desc = d
        lib = nat.hip_lib()
        hh = C.c_void_p()
        nat.hip_check(lib.fhip_create(C.byref(d), C.byref(hh)))
        self._ctx = hh

    def close(self) -> None:
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            nat.hip_lib().fhip_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -----------------------------------------------------------------
    @property
    def shape_spts(self):
        return (self.nSpts, self.nEles, self.nFields)

    def set_U(self, U: np.ndarray) -> None:
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.shape_spts)
        nat.hip_check(nat.hip_lib().fhip_set_state(self._ctx, nat.dptr(U)))

    def get_U(self) -> np.ndarray:
        U = np.zeros(self.shape_spts)
        nat.hip_check(nat.hip_lib().fhip_get_state(self._ctx, nat.dptr(U)))
        return U

    U_spts = property(get_U, set_U)

    def get(self, name: str, step: int = 0) -> np.ndarray:
        lib = nat.hip_lib()
        which = nat.ARR[name]
        n = lib.fhip_array_size(self._ctx, which)
        if n < 0:
            raise nat.FhipError(1, lib.fhip_last_error().decode())
        dtype = np.int32 if name.startswith("face_fpt") else np.float64
        out = np.zeros(n, dtype=dtype)
        nat.hip_check(lib.fhip_get_array(self._ctx, which, step, out.ctypes.data_as(C.c_void_p)))
        return out

    def put(self, name: str, a: np.ndarray, step: int = 0) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        nat.hip_check(nat.hip_lib().fhip_set_array(self._ctx, nat.ARR[name], step, a.ctypes.data_as(C.c_void_p)))

    # -- reference members ---------------------------------------------------------
    def initializeSolution(self, U: np.ndarray | None = None, directory: str = ".") -> None:
        """solver::initializeSolution (src/solver.cpp:1892-1943); U overrides the
        IC; with `restart 1` the state (and time) come from the plot file of
        restartIter (solver::readRestartFile, src/solver.cpp:1789-1890)."""
        p = self.params
        if U is None and getattr(p, "restart", 0):
            from .output import read_restart, vtu_name
            U, t = read_restart(Path(directory) / vtu_name(p.dataFileName, p.restartIter), self.host)
            if t is not None:
                p.time = t
        if U is None:
            U = initial_condition(self.params, self.geo.pos_spts)
        self.set_U(U)
        if self.params.dtType != 0:
            self.extrapolateU()
            nat.hip_check(nat.hip_lib().fhip_calc_wave_speed(self._ctx))

    def extrapolateU(self) -> None:
        nat.hip_check(nat.hip_lib().fhip_extrapolate_u(self._ctx))

    def calcResidual(self, step: int) -> None:
        nat.hip_check(nat.hip_lib().fhip_calc_residual(self._ctx, step))

    def timeStepA(self, step: int, rkval: float, src: np.ndarray | None = None) -> None:
        nat.hip_check(nat.hip_lib().fhip_time_step_a(self._ctx, step, rkval, nat.dptr(src)))

    def timeStepB(self, step: int, src: np.ndarray | None = None) -> None:
        nat.hip_check(nat.hip_lib().fhip_time_step_b(self._ctx, step, nat.dptr(src)))

    def copyUspts_U0(self) -> None:
        nat.hip_check(nat.hip_lib().fhip_copy_u_to_u0(self._ctx))

    def copyU0_Uspts(self) -> None:
        nat.hip_check(nat.hip_lib().fhip_copy_u0_to_u(self._ctx))

    def calcDt(self, refresh: bool | None = None) -> float:
        if refresh is None:
            refresh = self.iter == self.params.initIter + 1
        dt = C.c_double()
        nat.hip_check(nat.hip_lib().fhip_calc_dt(self._ctx, int(refresh), C.byref(dt)))
        self.params.dt = dt.value
        return dt.value

    def set_dt(self, dt: float) -> None:
        nat.hip_check(nat.hip_lib().fhip_set_dt(self._ctx, dt))
        self.params.dt = dt

    @property
    def dt(self) -> float:
        v = C.c_double()
        nat.hip_check(nat.hip_lib().fhip_get_dt(self._ctx, C.byref(v)))
        return v.value

    @property
    def iter(self) -> int:
        v = C.c_int()
        nat.hip_lib().fhip_get_iter(self._ctx, C.byref(v))
        return v.value

    def update(self, n_steps: int = 1, path: int = nat.PATH_AUTO, sync: bool = True) -> None:
        """n_steps of the main-loop body (iter++, solver::update)."""
        lib = nat.hip_lib()
        nat.hip_check(lib.fhip_update(self._ctx, n_steps, path))
        if sync:
            nat.hip_check(lib.fhip_sync(self._ctx))

    def _fused_ok(self) -> bool:
        return bool(nat.hip_lib().fhip_path_supported(self._ctx, nat.PATH_FUSED))

    def sync(self) -> None:
        nat.hip_check(nat.hip_lib().fhip_sync(self._ctx))

    def residual_norm(self, res_type: int | None = None) -> np.ndarray:
        out = np.zeros(self.nFields)
        nat.hip_check(nat.hip_lib().fhip_residual_norm(self._ctx, res_type or self.params.resType, nat.dptr(out)))
        return out

    def check_state(self) -> None:
        nat.hip_check(nat.hip_lib().fhip_check_state(self._ctx))

    # -- timing ------------------------------------------------------------------------
    def writeData(self, directory: str = ".") -> Path:
        """writeData (src/output.cpp:43-58): CSV (plotType 0) or ParaView (1)."""
        from .output import vtu_name, write_csv, write_vtu
        p = self.params
        U = self.get_U()
        if getattr(p, "plotType", 1) == 0:
            return write_csv(Path(directory) / f"{p.dataFileName}.csv.{self.iter:09d}", self.host, U)
        return write_vtu(Path(directory) / vtu_name(p.dataFileName, self.iter), self.host, U, p.time, self.iter)

    def autotune(self, path: int = nat.PATH_FUSED) -> list[int]:
        """Pick the block size of every fused kernel by timing them on the current state."""
        out = np.zeros(5, dtype=np.int32)
        nat.hip_check(nat.hip_lib().fhip_autotune(self._ctx, path, nat.iptr(out)))
        return out.tolist()

    def block_sizes(self, path: int = nat.PATH_FUSED) -> list[int]:
        out = np.zeros(5, dtype=np.int32)
        nat.hip_check(nat.hip_lib().fhip_get_block_sizes(self._ctx, path, nat.iptr(out)))
        return out.tolist()

    def set_block_sizes(self, sizes, path: int = nat.PATH_FUSED) -> None:
        a = np.ascontiguousarray(sizes, dtype=np.int32)
        nat.hip_check(nat.hip_lib().fhip_set_block_sizes(self._ctx, path, nat.iptr(a)))

    def enable_timing(self, on: bool = True) -> None:
        nat.hip_lib().fhip_enable_timing(self._ctx, int(on))

    def kernel_times(self) -> dict[str, tuple[float, int]]:
        n, L = 64, 64
        names = C.create_string_buffer(n * L)
        avg = np.zeros(n)
        cnt = (C.c_longlong * n)()
        k = nat.hip_lib().fhip_kernel_times(self._ctx, n, names, L, nat.dptr(avg), cnt)
        out = {}
        for i in range(k):
            nm = names.raw[i * L:(i + 1) * L].split(b"\0", 1)[0].decode()
            out[nm] = (float(avg[i]), int(cnt[i]))
        return out

    def desc_arrays(self) -> dict:
        """The host arrays handed to fhip_create (for checkers)."""
        o, g, f = self.opers, self.geo, self.faces
        return dict(opers=o, geo=g, faces=f, params=self._fp)


class MultiGrid:
    """P-multigrid (class multiGrid, src/multigrid.cpp:46-143,169-313): coarse
    solvers of orders lowOrder..P-1 on the fine solver's mesh, one cycle after
    every fine update.  As in multiGrid::setup, each coarse level gets a copy of
    the fine input (its RK coefficients included) with only the order changed,
    and its own initial condition and wave speeds."""

    def __init__(self, fine: Solver, low_order: int | None = None, smooth_steps: int | None = None):
        p = fine.params
        low = getattr(p, "lowOrder", 0) if low_order is None else low_order
        self.smooth = getattr(p, "smoothSteps", 1) if smooth_steps is None else smooth_steps
        if not 0 <= low < fine.order:
            raise ValueError("P-multigrid needs 0 <= lowOrder < order")
        self.fine = fine
        self.coarse: list[Solver] = []
        for P in range(low, fine.order):
            pc = copy.copy(p)
            pc.order = P
            s = Solver(pc, setup=HostSetup(pc, fine.mesh))
            s.initializeSolution()
            self.coarse.append(s)
        levels = self.coarse + [fine]
        D = fine.nDims
        sp = p.sptsTypeQuad
        self.restrict_ops = [pmg_operators(D, lv.order, sp)[1] for lv in levels[1:]]
        self.prolong_ops = [pmg_operators(D, lv.order, sp)[0] for lv in levels[:-1]]
        n = len(levels)
        hs = (C.c_void_p * n)(*[lv._ctx.value for lv in levels])
        R = (nat._dp * (n - 1))(*[nat.dptr(a) for a in self.restrict_ops])
        Pr = (nat._dp * (n - 1))(*[nat.dptr(a) for a in self.prolong_ops])
        h = C.c_void_p()
        nat.hip_check(nat.hip_lib().fhip_pmg_create(hs, n, R, Pr, self.smooth, C.byref(h)))
        self._mg = h

    @property
    def levels(self) -> list[Solver]:
        return self.coarse + [self.fine]

    def cycle(self, path: int = nat.PATH_AUTO, sync: bool = True) -> None:
        nat.hip_check(nat.hip_lib().fhip_pmg_cycle(self._mg, path))
        if sync:
            self.fine.sync()

    def close(self) -> None:
        if getattr(self, "_mg", None) is not None and self._mg.value:
            nat.hip_lib().fhip_pmg_destroy(self._mg)
            self._mg = None
        for s in self.coarse:
            s.close()
        self.coarse = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run(params: InputParams, log=print, output_dir: str | None = ".") -> Solver:
    """The main loop of src/flurry.cpp:120-155 (residual monitor and plot
    files included; output_dir=None writes no files).  The error monitor
    (writeError / writeAllError against a test case's exact solution) is
    output outside the residual + RK path and is not reproduced."""
    s = Solver(params)
    s.initializeSolution()
    p = params
    mg = MultiGrid(s) if getattr(p, "PMG", 0) else None
    if output_dir is not None:
        s.writeData(output_dir)
    while s.iter < p.iterMax and p.time < p.maxTime:
        s.update(1)
        if mg is not None:
            mg.cycle()
        p.time += s.dt
        it = s.iter
        if output_dir is not None and (it % p.plotFreq == 0 or it == p.iterMax or p.time >= p.maxTime):
            s.writeData(output_dir)
        if it % p.monitorResFreq == 0 or it == p.initIter + 1 or p.time >= p.maxTime:
            dt = s.dt
            if dt < 1e-13:
                raise RuntimeError("Instability detected - dt approaching zero!")
            res = s.residual_norm(p.resType)
            if np.any(np.isnan(res)):
                raise RuntimeError("NaN Encountered in Solution Residual!")
            log(f"{it:8d}  " + "  ".join(f"{r:.6e}" for r in res))
    return s
