"""Serving loop: solve requests that arrive over time, with continuous batching.

Every admitted request gets its own libflurry_hip context (own stream, own
device memory, state resident on the GPU until it finishes).  Each tick of the
loop advances every active request by `steps_per_tick` RK steps through one
fhip_update_many call, so the kernels of all active requests are in flight on
the device together; finished requests are evicted (their memory released)
and waiting requests are admitted as slots and HBM budget free up.  Requests
are either steady (stop when the residual falls below a tolerance) or
unsteady (a fixed number of steps), in FP64 or mixed precision.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native as nat
from . import cases
from .solver import MultiGrid, Solver


@dataclass
class Request:
    rid: int
    kind: str                      # "bump2d" or "vortex3d"
    size: tuple                    # (nx, ny) or (n,)
    order: int
    arrival: float                 # seconds after the start of the run
    mach: float = 0.3
    precision: str = "fp64"        # "fp64" or "mixed"
    steady_tol: float | None = None
    n_steps: int | None = None
    max_steps: int = 200000
    # filled in by the server
    admitted: float | None = None
    finished: float | None = None
    steps: int = 0
    residual: float | None = None
    dof: int = 0
    result: np.ndarray | None = None
    status: str = ""               # "converged", "max_steps", "diverged" or "done" (unsteady)
    # mixed-precision steady requests finish in FP64 (see Server.promote_below):
    # the step at which this one switched, or None
    promoted_at: int | None = None
    best: float = float("inf")     # lowest residual seen so far, and the check at which it was seen
    best_check: int = 0
    checks: int = 0

    @property
    def path(self) -> int:
        """The path the request runs on now."""
        return nat.PATH_MIXED if self.precision == "mixed" and self.promoted_at is None else nat.PATH_FUSED


def build_case(r: Request):
    if r.kind == "bump2d":
        # settings of FlurryPP's P-multigrid channel case: tuned steady RK, local CFL 0.9
        return cases.bump2d(nx=r.size[0], ny=r.size[1], order=r.order, mach=r.mach, seed=1000 + r.rid,
                            time_type=6, cfl=0.9)
    if r.kind == "vortex3d":
        return cases.vortex3d(n=r.size[0], order=r.order, mach=r.mach)
    raise ValueError(r.kind)


def footprint_bytes(nDims: int, order: int, nEles: int, viscous: bool = False, nRK: int = 4,
                    pmg_source: bool = True) -> int:
    """Device bytes one context holds once both precisions have run: the
    staged arrays and geometry, the fused FP64 state, metric terms and face
    tables, and the FP64 and FP32 work arrays (counted per element from
    fhip_create, fused_setup and ensure_work), plus 25% and 16 MiB for
    allocation granularity (measured: a 3-D P3 context of 169 MiB of arrays
    holds 206 MiB).  The runtime's one-time cost per process (code
    objects, scratch) is not included."""
    P1 = order + 1
    D = nDims
    S, N = P1 ** D, D + 2
    F = 2 * D * P1 ** (D - 1)
    SN, FN = S * N, F * N
    # staged: U, U0, F_spts, divF[nRK]; U_fpts, disFn, Fn, face Fn; wave speeds, metrics, face tables
    d = SN * (2 + D + nRK) + 4 * FN + F * (3.5 + D) + D * D * S + S + 1
    # fused FP64: U, its stage copy, RK accumulator, divF0; wave speeds, metrics (+1/detJ), face
    # tables, the points' side codes, element volumes
    d += 4 * SN + F * (2 + D) + D * D * S + 2 * S + 1.5 * F + 0.5 * F + 1
    # FP64 work: divF, U/Fn/disFn at the flux points
    d += SN + 3 * FN
    # FP32 work (half a double each): the same, stage residuals, FP32 U, FP32 geometry
    d += 0.5 * (SN + 3 * FN + (nRK - 1) * SN + SN + D * D * S + S + D * F + F)
    if viscous:
        d += D * SN + D * FN + FN            # staged gradients and dUc
        d += D * F                           # fused face normals
        d += 2 * FN + 0.5 * (2 * FN + D * F)  # dUc and LDG side parts, FP64 and FP32
    if pmg_source:
        d += SN                              # P-multigrid source, element-major
    return int(8 * d * nEles * 1.25) + (16 << 20)


@dataclass
class Server:
    hbm_budget: float = 64e9       # bytes this server may hold on the device
    max_active: int = 64           # concurrent requests (slots)
    steps_per_tick: int = 10
    check_every: int = 50          # steady requests: residual check interval (steps)
    keep_results: bool = False
    pmg: bool = False              # steady requests: P-multigrid down to order 0 after every step
    # A mixed-precision steady request continues in FP64 once its residual is
    # below promote_below, or once it has not halved over stall_checks residual
    # checks: its fluxes and residual are FP32 sums, whose rounding floors
    # the residual near 1e-7 (measured on the bump channels), above the 1e-8
    # to 1e-10 tolerances steady requests ask for.  The FP64 state is kept
    # by both paths, so the switch only changes the kernels.
    promote_below: float = 1e-6
    stall_checks: int = 8
    log: object = None
    active: list = field(default_factory=list)
    done: list = field(default_factory=list)

    def _room(self, r: Request) -> int:
        """Steps request r may take before its next residual check (steady) or
        its last step (unsteady)."""
        if r.steady_tol is None:
            return max(0, r.n_steps - r.steps)
        return max(0, min(r.max_steps - r.steps, self.check_every - r.steps % self.check_every))

    def run(self, requests: list[Request]) -> dict:
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting: list[Request] = []
        used = 0.0
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        ctx = {}
        total_dof_stages = 0
        while pending or waiting or self.active:
            t = now()
            while pending and pending[0].arrival <= t:
                waiting.append(pending.pop(0))
            # admission: first come, first served, within slots and HBM budget
            while waiting and len(self.active) < self.max_active:
                r = waiting[0]
                setup, U = build_case(r)
                visc = bool(setup.params.viscous)
                need = footprint_bytes(setup.nDims, setup.order, setup.nEles, visc)
                if self.pmg and r.steady_tol is not None:
                    need += sum(footprint_bytes(setup.nDims, q, setup.nEles, visc) for q in range(setup.order))
                if used + need > self.hbm_budget and self.active:
                    break
                waiting.pop(0)
                s = Solver(setup.params, setup=setup)
                s.initializeSolution(U)
                mg = MultiGrid(s, low_order=0, smooth_steps=2) if (self.pmg and r.steady_tol is not None
                                                                    and setup.order > 0) else None
                r.admitted, r.dof = now(), setup.nEles * setup.nSpts
                ctx[r.rid] = (s, need, setup.params.nRKSteps, mg)
                used += need
                self.active.append(r)
            if not self.active:
                if pending:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - now())))
                continue
            # one tick: the active requests of each precision advance together,
            # by at most steps_per_tick and never past a member's next residual
            # check or its last step
            lib = nat.hip_lib()
            taken = {}
            for path in (nat.PATH_FUSED, nat.PATH_MIXED):
                group = [r for r in self.active if r.path == path]
                if not group:
                    continue
                k = min([self.steps_per_tick] + [self._room(r) for r in group])
                arr = (C.c_void_p * len(group))(*[ctx[r.rid][0]._ctx.value for r in group])
                cyc = [ctx[r.rid][3] for r in group if ctx[r.rid][3] is not None]
                if k > 0 and not cyc:
                    nat.hip_check(lib.fhip_update_many(arr, len(group), k, path))
                elif k > 0:
                    for _ in range(k):
                        nat.hip_check(lib.fhip_update_many(arr, len(group), 1, path))
                        for mg in cyc:
                            mg.cycle(path, sync=False)
                for r in group:
                    taken[r.rid] = k
            for r in self.active:
                ctx[r.rid][0].sync()
                r.steps += taken[r.rid]
                total_dof_stages += r.dof * ctx[r.rid][2] * taken[r.rid]
            # completion and eviction
            still = []
            for r in self.active:
                s, need, _, mg = ctx[r.rid]
                fin = False
                if r.steady_tol is not None:
                    if r.steps % self.check_every == 0 or r.steps >= r.max_steps:
                        res = s.residual_norm(2)
                        r.residual = float(res[0])
                        r.checks += 1
                        if r.residual < 0.5 * r.best:
                            r.best, r.best_check = r.residual, r.checks
                        if not np.isfinite(res).all():
                            fin, r.status = True, "diverged"
                        elif r.residual <= r.steady_tol:
                            fin, r.status = True, "converged"
                        elif r.steps >= r.max_steps:
                            fin, r.status = True, "max_steps"
                        elif r.path == nat.PATH_MIXED and (r.residual <= self.promote_below or
                                                           r.checks - r.best_check >= self.stall_checks):
                            r.promoted_at = r.steps
                            if self.log is not None:
                                self.log(f"request {r.rid} ({r.kind} P{r.order}) continues in FP64 after "
                                         f"{r.steps} mixed steps, residual {r.residual:.3e}")
                else:
                    fin = r.steps >= r.n_steps
                    if fin:
                        r.status = "done"
                if fin:
                    r.finished = now()
                    if self.log is not None:
                        self.log(f"request {r.rid} ({r.kind} P{r.order}, {r.precision}) {r.status} after "
                                 f"{r.steps} steps, {r.finished - r.arrival:.3f} s")
                    if self.keep_results:
                        r.result = s.get_U()
                    if mg is not None:
                        mg.close()
                    s.close()
                    used -= need
                    del ctx[r.rid]
                    self.done.append(r)
                else:
                    still.append(r)
            self.active = still
        wall = now()
        out = {"requests": len(self.done), "wall_s": wall}
        out.update(latency_stats(self.done))
        out["by_precision"] = {p: latency_stats([r for r in self.done if r.precision == p]) for p in ("fp64", "mixed")}
        out["promoted_to_fp64"] = sum(1 for r in self.done if r.promoted_at is not None)
        out["throughput_req_per_s"] = len(self.done) / wall if wall > 0 else None
        out["gdof_stage_per_s"] = total_dof_stages / wall / 1e9 if wall > 0 else None
        return out


def _pct(x: list[float]) -> dict:
    return {"p50_s": float(np.percentile(x, 50)) if len(x) else None,
            "p99_s": float(np.percentile(x, 99)) if len(x) else None}


def latency_stats(done: list[Request]) -> dict:
    """Status counts and latencies (arrival to finish): p50/p99 over the
    requests that reached a solution (converged steady, finished unsteady),
    and separately over those that stopped without one (max_steps,
    diverged), so that no request is left out of the report."""
    count = {k: sum(1 for r in done if r.status == k) for k in ("converged", "max_steps", "diverged", "done")}
    ok = [r.finished - r.arrival for r in done if r.status in ("converged", "done")]
    bad = [r.finished - r.arrival for r in done if r.status in ("max_steps", "diverged")]
    out = {"status": count}
    out.update(_pct(ok))
    out["unsolved"] = _pct(bad)
    out["all"] = _pct(ok + bad)
    return out


def poisson_mix(n: int, rate: float, seed: int = 1, frac_2d: float = 0.7, steady_tol: float = 1e-6,
                unsteady_steps: int = 100, n3d=(8, 16), p3d=(2, 4), nx2d=(16, 48), ny2d=(8, 16),
                p2d=(1, 3), mixed_frac: float = 0.5, max_steps: int = 3000) -> list[Request]:
    """Poisson arrivals at `rate` per second of a 2-D steady / 3-D unsteady mix;
    steady requests give up after max_steps."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        prec = "mixed" if rng.random() < mixed_frac else "fp64"
        if rng.random() < frac_2d:
            out.append(Request(i, "bump2d", (int(rng.integers(nx2d[0], nx2d[1] + 1)),
                                             int(rng.integers(ny2d[0], ny2d[1] + 1))),
                               int(rng.integers(p2d[0], p2d[1] + 1)), t, mach=float(rng.uniform(0.2, 0.5)),
                               precision=prec, steady_tol=steady_tol, max_steps=max_steps))
        else:
            out.append(Request(i, "vortex3d", (int(rng.integers(n3d[0], n3d[1] + 1)),),
                               int(rng.integers(p3d[0], p3d[1] + 1)), t, mach=0.5, precision=prec,
                               n_steps=unsteady_steps))
    return out
