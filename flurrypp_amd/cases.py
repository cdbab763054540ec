"""Workloads: input options, meshes and initial states of the test and
benchmark cases (the options are FlurryPP input keys).

* ``vortex2d``  2-D Euler isentropic vortex on a jittered periodic quad mesh
* ``tgv3d``     3-D Taylor-Green vortex (Navier-Stokes or Euler) on a periodic hex mesh
* ``bump2d``    2-D Euler Gaussian-bump channel (characteristic inlet/outlet, slip walls)
* ``vortex3d``  3-D Euler vortex column on a periodic hex mesh

Jitter: every vertex not on a boundary moves by u*amp*h per coordinate,
u uniform in [-1, 1] from splitmix64(seed ^ (vertex*nDims + d)).
"""
from __future__ import annotations

import math

import numpy as np

from .input import InputParams
from .mesh import Mesh, bc_number
from .solver import PI, HostSetup

SEED = 20261202


def _params(**kw) -> InputParams:
    return InputParams.from_dict(kw)


# --------------------------------------------------------------------------
def vortex2d(n: int = 32, order: int = 3, jitter: float = 0.1, mach: float = 0.5, beta: float = 5.0,
             cfl: float = 0.5, riemann: int = 0, dt_type: int = 1, iters: int = 100, seed: int = SEED,
             time_type: int = 4):
    """2-D Euler isentropic vortex on [-5,5]^2, freestream (M*sqrt(gamma), 0),
    rho_inf = p_inf = 1; vortex of the reference form (src/ele.cpp:765-777)."""
    p = _params(equation=1, nDims=2, order=order, meshType=1, nx=n, ny=n, xmin=-5, xmax=5, ymin=-5, ymax=5,
                timeType=time_type, dtType=dt_type, CFL=cfl, dt=1e-3, maxTime=1e9, iterMax=iters, icType=0,
                riemannType=riemann, gamma=1.4, rhoBound=1.0, uBound=mach * math.sqrt(1.4), vBound=0.0, pBound=1.0)
    h = 10.0 / n
    m = Mesh.cartesian(2, (n, n), (-5, -5), (5, 5))
    if jitter:
        m.jitter((h, h), jitter, seed)
    m.set_periodic(10.0, 10.0, 0.0, 1e-6)
    setup = HostSetup(p, m)
    U = vortex_state(setup.geo.pos_spts, p.gamma, beta, p.uBound, 0.0)
    return setup, U


def vortex_state(pos: np.ndarray, g: float, eps: float, uinf: float, vinf: float) -> np.ndarray:
    S, E, D = pos.shape
    x, y = pos[..., 0], pos[..., 1]
    f = 1.0 - (x * x + y * y)
    rho = np.maximum(np.power(1.0 - eps * eps * (g - 1.0) / (8.0 * g * PI * PI) * np.exp(f), 1.0 / (g - 1.0) + 1e-5),
                     1e-3)
    vx = uinf - eps * y / (2.0 * PI) * np.exp(f / 2.0)
    vy = vinf + eps * x / (2.0 * PI) * np.exp(f / 2.0)
    pr = np.power(rho, g)
    U = np.zeros((S, E, D + 2))
    U[..., 0], U[..., 1], U[..., 2] = rho, rho * vx, rho * vy
    U[..., D + 1] = pr / (g - 1) + 0.5 * rho * (vx * vx + vy * vy)
    return U


# --------------------------------------------------------------------------
def tgv3d(n: int = 48, order: int = 3, viscous: int = 1, riemann: int = 0, mach: float = 0.1, re: float = 1600.0,
          dt: float | None = None, cfl: float = 0.5, iters: int = 200, jitter: float = 0.0, seed: int = SEED):
    """Taylor-Green vortex on [0,2pi]^3 (periodic), rho=1, V0=1,
    p = 1/(gamma M^2) + (cos2x+cos2y)(cos2z+2)/16.  Viscous: Re=1600,
    Pr=0.71, constant viscosity mu = 1/Re (TBound = TGas, fixVis 1), LDG
    penFact 0, tau 1.  dt: fixed; None -> one CFL evaluation at t=0."""
    L = 2.0 * math.pi
    kw = dict(equation=1, nDims=3, order=order, meshType=1, nx=n, ny=n, nz=n, xmin=0, xmax=L, ymin=0, ymax=L,
              zmin=0, zmax=L, timeType=4, dtType=0, dt=dt if dt else 1e-3, iterMax=iters, icType=0,
              riemannType=riemann, gamma=1.4, rhoBound=1.0, uBound=1.0, vBound=0.0, wBound=0.0,
              pBound=1.0 / (1.4 * mach * mach), viscous=viscous)
    if viscous:
        kw.update(Re=re, MachBound=mach, TBound=291.15, TGas=291.15, fixVis=1, prandtl=0.71, MachWall=0.0,
                  LDG_penFact=0.0, LDG_tau=1.0)
    p = _params(**kw)
    m = Mesh.cartesian(3, (n, n, n), (0, 0, 0), (L, L, L))
    if jitter:
        m.jitter((L / n,) * 3, jitter, seed)
    m.set_periodic(L, L, L, 1e-6)
    setup = HostSetup(p, m)
    U = tgv_state(setup.geo.pos_spts, p.gamma, mach)
    if dt is None:
        p.dt = cfl_dt(setup, U, cfl)
        setup.fp.dt = p.dt
    return setup, U


def tgv_state(pos: np.ndarray, g: float, mach: float) -> np.ndarray:
    x, y, z = pos[..., 0], pos[..., 1], pos[..., 2]
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u)
    pr = 1.0 / (g * mach * mach) + (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0) / 16.0
    rho = np.ones_like(u)
    U = np.zeros(pos.shape[:2] + (5,))
    U[..., 0] = rho
    U[..., 1], U[..., 2], U[..., 3] = rho * u, rho * v, rho * w
    U[..., 4] = pr / (g - 1) + 0.5 * rho * (u * u + v * v + w * w)
    return U


def cfl_dt(setup: HostSetup, U: np.ndarray, cfl: float) -> float:
    """Global CFL time step of ele::calcDt (src/ele.cpp:1127-1146) for state U,
    with wave speeds from the extrapolated state (ele::calcWaveSpFpts)."""
    g = setup.params.gamma
    D = setup.nDims
    Uf = np.einsum("fs,sek->fek", setup.opers.spts_to_fpts, U)
    rho = Uf[..., 0]
    vel = Uf[..., 1:D + 1] / rho[..., None]
    pr = (g - 1) * (Uf[..., D + 1] - 0.5 * rho * (vel ** 2).sum(-1))
    vn = (vel * setup.geo.norm_fpts).sum(-1)
    ws = np.abs(vn) + np.sqrt(np.maximum(g * pr / rho, 0.0))
    Ff = setup.nFptsFace
    wf = np.tile(setup.opers.w_face, setup.nFpts // Ff)
    vol = np.einsum("s,se->e", setup.opers.w_spts, setup.geo.detjac_spts)
    iw = np.einsum("fe,f,fe->e", ws, wf, setup.geo.da_fpts)
    lim = {0: 1.393, 1: 0.464, 2: 0.235, 3: 0.1453, 4: 0.100, 5: 0.0736}[setup.order]
    return float(np.min(cfl * lim * 2 * vol / iw))


# --------------------------------------------------------------------------
def bump_mesh(nx: int, ny: int, jitter: float = 0.1, seed: int = SEED, height: float = 0.0625) -> Mesh:
    """Channel [-1.5,1.5] x [y_b(x), 0.8], y_b = height exp(-25 x^2);
    bottom/top slip walls, inlet/outlet characteristic."""
    xs = np.linspace(-1.5, 1.5, nx + 1)
    yb = height * np.exp(-25.0 * xs * xs)
    j = np.arange(ny + 1)
    X = np.repeat(xs[None, :], ny + 1, 0)
    Y = yb[None, :] + (0.8 - yb[None, :]) * (j[:, None] / ny)
    xv = np.stack([X.ravel(), Y.ravel()], 1)
    vid = lambda i, jj: jj * (nx + 1) + i  # noqa: E731
    c2v = np.array([[vid(i, jj), vid(i + 1, jj), vid(i + 1, jj + 1), vid(i, jj + 1)]
                    for i in range(nx) for jj in range(ny)], dtype=np.int32)
    bottom = [vid(i, 0) for i in range(nx + 1)]
    top = [vid(i, ny) for i in range(nx + 1)]
    inlet = [vid(0, jj) for jj in range(ny + 1)]
    outlet = [vid(nx, jj) for jj in range(ny + 1)]
    hx, hy = 3.0 / nx, 0.8 / ny
    if jitter:
        rng_off = np.zeros_like(xv)
        for v in range(xv.shape[0]):
            i, jj = v % (nx + 1), v // (nx + 1)
            if 0 < i < nx and 0 < jj < ny:
                for d, hd in enumerate((hx, hy)):
                    rng_off[v, d] = (2.0 * _u01(seed ^ (v * 2 + d)) - 1.0) * jitter * hd
        xv = xv + rng_off
    return Mesh.from_arrays(xv, c2v, [(bc_number("char"), np.array(inlet + outlet)),
                                      (bc_number("slip_wall"), np.array(bottom + top))])


def _u01(x: int) -> float:
    M = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & M
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M
    x ^= x >> 31
    return (x >> 11) * (1.0 / 9007199254740992.0)


def bump2d(nx: int = 32, ny: int = 12, order: int = 2, mach: float = 0.3, cfl: float = 0.5, dt_type: int = 2,
           time_type: int = 4, jitter: float = 0.1, seed: int = SEED, iters: int = 100, dt: float = 1e-3,
           height: float = 0.0625):
    """Subsonic Euler flow over a Gaussian bump: rho=1, p=1/gamma (c=1),
    inlet velocity (M, 0); local-CFL time stepping by default (dt is used
    with dt_type 0 only).  height 0 gives a straight channel, in which the
    free stream is an exact steady state."""
    g = 1.4
    p = _params(equation=1, nDims=2, order=order, meshType=0, meshFileName="-", timeType=time_type, dtType=dt_type,
                CFL=cfl, dt=dt, maxTime=1e9, iterMax=iters, icType=0, riemannType=0, gamma=g, rhoBound=1.0, uBound=mach,
                vBound=0.0, pBound=1.0 / g, mesh_bound={"fluid": "fluid"})
    m = bump_mesh(nx, ny, jitter, seed, height)
    setup = HostSetup(p, m)
    return setup, setup.initial_condition()


# --------------------------------------------------------------------------
def vortex3d(n: int = 8, order: int = 3, cfl: float = 0.5, beta: float = 5.0, mach: float = 0.5, iters: int = 100):
    """Euler vortex column (axis z) on the periodic box [-5,5]^2 x [0,10/n*n_z]."""
    p = _params(equation=1, nDims=3, order=order, meshType=1, nx=n, ny=n, nz=n, xmin=-5, xmax=5, ymin=-5,
                ymax=5, zmin=-5, zmax=5, timeType=4, dtType=1, CFL=cfl, maxTime=1e9, iterMax=iters, icType=0,
                riemannType=0, gamma=1.4, rhoBound=1.0, uBound=mach * math.sqrt(1.4), vBound=0.0, wBound=0.0,
                pBound=1.0)
    m = Mesh.cartesian(3, (n, n, n), (-5, -5, -5), (5, 5, 5))
    m.set_periodic(10.0, 10.0, 10.0, 1e-6)
    setup = HostSetup(p, m)
    U2 = vortex_state(setup.geo.pos_spts[..., :2], 1.4, beta, p.uBound, 0.0)
    U = np.zeros(U2.shape[:2] + (5,))
    U[..., 0:3] = U2[..., 0:3]
    U[..., 4] = U2[..., 3]
    return setup, U
