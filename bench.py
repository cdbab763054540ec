#!/usr/bin/env python3
"""Benchmark of the FR residual + RK update on one MI355X.

Workload (default): 3-D compressible Navier-Stokes Taylor-Green vortex,
48^3 periodic hex mesh on [0,2pi]^3, P=3 (64 solution points per element,
7,077,888 DOF x 5 fields), Rusanov inviscid + LDG viscous flux, Re=1600,
Mach 0.1, Pr=0.71, RK44, fixed dt from a CFL-0.5 evaluation at t=0, FP64.
One benchmark step is one RK44 time step (4 stages).  The metric counts
DOF x stages per second (DOF = elements x solution points).

Prints one JSON line.  The timed region runs `--steps` RK steps through
fhip_update on the device in its production mode (one captured HIP graph
replayed per RK step, no synchronisation inside the region).  A further
`--kernel-steps` pass launches kernel by kernel with HIP events on the
context's stream around every launch; its per-kernel averages give the
roofline fraction of the dominant kernel.  The CPU
baseline times the C restatement in oracle/ (serial) on a smaller Taylor-Green
mesh on this host, on one core and in --cpu-procs concurrent single-thread
processes.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "GDOF·stage/s at P3 hex TGV, FP64 vs mixed; p50/p99 solve latency batched"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n: int, steps: int) -> dict:
    from flurrypp_amd import cases
    from oracle.oracle import Oracle
    setup, U = cases.tgv3d(n=n, order=3, viscous=1)
    o = Oracle.from_setup(setup)
    o.initializeSolution(U, 0)
    t0 = time.perf_counter()
    o.update(steps)
    el = time.perf_counter() - t0
    dof = setup.nEles * setup.nSpts
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), model)
    except OSError:
        pass
    return {"value": dof * 4 * steps / el / 1e9, "unit": "GDOF·stage/s", "cores": 1, "kind": "port",
            "host_cpu": model, "nproc": os.cpu_count(),
            "sample": f"C restatement (oracle/, serial, gcc -O3) on the same Taylor-Green NS case at {n}^3 hex P3, "
                      f"{steps} RK44 steps ({dof} DOF), {el:.1f} s"}


_CPU_CHILD = """
import sys, time
sys.path.insert(0, {root!r})
from flurrypp_amd import cases
from oracle.oracle import Oracle
setup, U = cases.tgv3d(n={n}, order=3, viscous=1)
o = Oracle.from_setup(setup)
o.initializeSolution(U, 0)
t0 = time.time()
o.update({steps})
print(t0, time.time(), setup.nEles * setup.nSpts)
"""


def cpu_baseline_processes(n: int, steps: int, procs: int) -> dict:
    """The same sample in `procs` independent single-thread processes at once
    (the reference's viscous path runs one thread per process): aggregate
    DOF-stages over the span from the first timed start to the last end."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS="1")
    code = _CPU_CHILD.format(root=str(ROOT), n=n, steps=steps)
    ps = [subprocess.Popen([sys.executable, "-c", code], stdout=subprocess.PIPE, env=env, text=True)
          for _ in range(procs)]
    rows = []
    for p in ps:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"CPU baseline process failed ({p.returncode})")
        t0, t1, dof = out.split()
        rows.append((float(t0), float(t1), int(dof)))
    span = max(r[1] for r in rows) - min(r[0] for r in rows)
    work = sum(r[2] * 4 * steps for r in rows)
    return {"value": work / span / 1e9, "unit": "GDOF·stage/s", "processes": procs, "threads_per_process": 1,
            "sample": f"{procs} concurrent copies of the {n}^3 sample, {span:.1f} s from the first timed start "
                      f"to the last end"}


def batched_form(n_cases: int, seed: int, tol: float, max_steps: int, nx_hi: int, ny_hi: int,
                 cfl: list[float] | None = None) -> dict:
    """Independent 2-D Euler Gaussian-bump channels (sizes, orders 1..5 and
    inlet Mach 0.2-0.5 drawn seeded), merged per order, with the settings of
    FlurryPP's own P-multigrid channel case (tests/euler/channel/input_channelPMG:
    tuned steady RK (timeType 6), local CFL 0.9, lowOrder 0, smoothSteps 2),
    run until each case's L2 density residual is <= tol; converged and
    diverged cases are frozen, the rest stop at max_steps.  cfl[P-1] is the
    local CFL number of order P (default 0.9 for every order)."""
    from flurrypp_amd import cases
    from flurrypp_amd.batch import BatchedSolver, Member
    rng = np.random.default_rng(seed)
    ms = []
    for i in range(n_cases):
        nx, ny = int(rng.integers(16, nx_hi + 1)), int(rng.integers(8, ny_hi + 1))
        P, M = int(rng.integers(1, 6)), float(rng.uniform(0.2, 0.5))
        s, U = cases.bump2d(nx=nx, ny=ny, order=P, mach=M, seed=seed * 100003 + i, time_type=6,
                            cfl=cfl[P - 1] if cfl else 0.9)
        ms.append(Member(s, U, (nx, ny, P, M)))
    t = time.perf_counter()
    b = BatchedSolver(ms, pmg=True, low_order=0, smooth_steps=2, autotune=True)
    setup_s = time.perf_counter() - t
    tuned = {(f"P{g}/level P{q}" if g else f"merged level P{q}"): v for (g, q), v in sorted(b.block_sizes.items())}
    def progress(step, nconv, ndiv):
        if step % 500 == 0:
            log(f"bench: batched step {step}: {nconv} converged, {ndiv} diverged")
    out = b.solve_steady(tol=tol, max_steps=max_steps, check_every=50, stop_on_nan=False, progress=progress)
    b.close()
    conv, div = out["converged"], out["diverged"]
    by_order = {P: {"cases": int(sum(1 for m in ms if m.tag[2] == P)),
                    "converged": int(sum(1 for i, m in enumerate(ms) if m.tag[2] == P and conv[i])),
                    "diverged": int(sum(1 for i, m in enumerate(ms) if m.tag[2] == P and div[i]))}
                for P in range(1, 6)}
    tc = out["time_s"][conv]
    dof = sum(m.setup.nEles * m.setup.nSpts for m in ms)
    cfls = ", ".join(f"P{P} {cfl[P - 1] if cfl else 0.9:g}" for P in range(1, 6))
    return {
        "workload": f"{n_cases} 2-D Euler Gaussian-bump channels, quads 16..{nx_hi} x 8..{ny_hi}, P 1..5, "
                    f"Mach 0.2-0.5, timeType 6, local CFL ({cfls}), P-multigrid (orders 0..P-1, 2 smoothing steps), "
                    f"steady to L2 residual {tol:g} within {max_steps} steps",
        "cases": n_cases, "dof": dof, "converged": int(conv.sum()), "diverged": int(div.sum()),
        "unconverged": int(n_cases - conv.sum() - div.sum()),
        "fractions": {"converged": round(float(conv.mean()), 4), "diverged": round(float(div.mean()), 4),
                      "unconverged": round(float(1 - conv.mean() - div.mean()), 4)},
        "by_order": by_order, "wall_s": round(out["wall_s"], 3),
        # time to converged over the converged cases only; the others are
        # counted above (diverged cases are frozen when their residual turns
        # NaN, unconverged ones run to max_steps)
        "p50_s": float(np.percentile(tc, 50)) if len(tc) else None,
        "p99_s": float(np.percentile(tc, 99)) if len(tc) else None,
        "steps_p50": int(np.percentile(out["steps"], 50)), "steps_max": int(out["steps"].max()),
        "value": round(out["dof_stages"] / out["wall_s"] / 1e9, 4), "unit": "GDOF·stage/s (fine-level RK stages)",
        "host_setup_s": round(setup_s, 2), "block_sizes": tuned,
    }


def served_form(n_req: int, rates: list[float], seed: int) -> dict:
    """Poisson arrivals of the same request mix at each rate in `rates`: 70%
    steady 2-D bump channels (16..128 x 8..48 quads, P 1..5, Mach 0.2-0.5,
    P-multigrid, residual 1e-8 within 6000 steps), 30% unsteady 3-D periodic
    Euler vortex columns (8..40^3 hexes, P 2..4, 100 steps); half FP64, half
    mixed precision (mixed steady requests finish in FP64)."""
    from flurrypp_amd.serve import Server, poisson_mix
    out = {"workload": f"{n_req} requests per rate: 70% steady 2-D bump (16..128 x 8..48, P 1..5, residual 1e-8 "
                       f"within 6000 steps, P-multigrid), 30% 3-D Euler vortex (8..40^3, P 2..4, 100 steps); "
                       f"FP64/mixed 50/50", "rates": {}}
    for rate in rates:
        reqs = poisson_mix(n_req, rate=rate, seed=seed, steady_tol=1e-8, unsteady_steps=100, n3d=(8, 40),
                           p3d=(2, 4), nx2d=(16, 128), ny2d=(8, 48), p2d=(1, 5), max_steps=6000)
        srv = Server(hbm_budget=200e9, max_active=32, steps_per_tick=10, check_every=50, pmg=True, log=log)
        out["rates"][f"{rate:g}/s"] = srv.run(reqs)
    return out


def tgv_mixed_form(n: int, order: int, steps: int, warmup: int, tune: bool) -> dict:
    """The Taylor-Green NS case at n^3 on the mixed path beside the FP64 path,
    both from the same initial state: throughput of each and the mixed
    path's error against the FP64 one (density L2, kinetic-energy
    dissipation rate)."""
    from flurrypp_amd import cases
    from flurrypp_amd._native import PATH_FUSED, PATH_MIXED, PATH_MIXED_BF16, PATH_MIXED_BF16X3
    from flurrypp_amd.solver import Solver
    t = time.perf_counter()
    setup, U = cases.tgv3d(n=n, order=order, viscous=1)
    s = Solver(setup.params, setup=setup)
    s.initializeSolution(U)
    host = time.perf_counter() - t
    dof, nRK = setup.nEles * setup.nSpts, setup.params.nRKSteps
    runs = {}
    paths = [("fp64", PATH_FUSED), ("mixed", PATH_MIXED)]
    if order == 3:
        paths += [("bf16_hi_lo", PATH_MIXED_BF16), ("bf16_hi_mid_lo", PATH_MIXED_BF16X3)]
    for name, path in paths:
        s.set_U(U)
        if tune:
            s.autotune(path)
        s.update(warmup, path=path)
        s.sync()
        t0 = time.perf_counter()
        s.update(steps, path=path, sync=False)
        s.sync()
        runs[name] = (time.perf_counter() - t0, s.get_U())
        log(f"bench: TGV {n}^3 {name}: {dof * nRK * steps / runs[name][0] / 1e9:.3f} GDOF-stage/s")
    s.close()
    el64, U64 = runs["fp64"]
    T = (warmup + steps) * setup.params.dt
    ek0, ek64 = kinetic_energy(setup, U), kinetic_energy(setup, U64)
    eps64 = -(ek64 - ek0) / T

    def against_fp64(name):
        el, V = runs[name]
        eps = -(kinetic_energy(setup, V) - ek0) / T
        return {"value": round(dof * nRK * steps / el / 1e9, 4), "speedup_vs_fp64": round(el64 / el, 3),
                "rel_l2_density_vs_fp64": float(np.linalg.norm(V[..., 0] - U64[..., 0]) / np.linalg.norm(U64[..., 0])),
                "rel_l2_state_vs_fp64": float(np.linalg.norm(V - U64) / np.linalg.norm(U64)),
                "ke_dissipation_rel_err": abs(eps - eps64) / abs(eps64) if eps64 else None}
    out = {"workload": f"3D Navier-Stokes Taylor-Green vortex, {n}^3 periodic hex, P{order}, Rusanov+LDG, RK44",
           "elements": setup.nEles, "dof": dof, "host_setup_s": round(host, 2), "unit": "GDOF·stage/s",
           "fp64_value": round(dof * nRK * steps / el64 / 1e9, 4), "after_steps": warmup + steps}
    out.update(against_fp64("mixed"))
    # the operators' z contractions as BF16 pieces on the matrix cores (FP32
    # accumulation), beside the FP32 vector-unit form above
    for name in ("bf16_hi_lo", "bf16_hi_mid_lo"):
        if name in runs:
            out[name] = against_fp64(name)
    return out


def vortex2d_form(steps: int, tune: bool) -> dict:
    """The 2-D Euler isentropic vortex at its full size (32x32 jittered
    periodic quads, P3, Rusanov, RK44, CFL 0.5) on the FP64 fused path: a
    launch-bound case (16,384 DOF)."""
    from flurrypp_amd import cases
    from flurrypp_amd._native import PATH_FUSED
    from flurrypp_amd.solver import Solver
    setup, U = cases.vortex2d(n=32, order=3)
    s = Solver(setup.params, setup=setup)
    s.initializeSolution(U)
    if tune:
        s.autotune(PATH_FUSED)
    s.update(10, path=PATH_FUSED)
    s.sync()
    t0 = time.perf_counter()
    s.update(steps, path=PATH_FUSED, sync=False)
    s.sync()
    el = time.perf_counter() - t0
    s.close()
    dof, nRK = setup.nEles * setup.nSpts, setup.params.nRKSteps
    return {"workload": "2D Euler isentropic vortex, 32x32 jittered periodic quads, P3, Rusanov, RK44, CFL 0.5",
            "dof": dof, "steps": steps, "value": round(dof * nRK * steps / el / 1e9, 4), "unit": "GDOF·stage/s",
            "ms_per_step": round(el / steps * 1e3, 4)}


def dense_operator_form(n: int, order: int, reps: int = 5) -> dict:
    """The operator applications of one Navier-Stokes RK stage as the reference
    forms them: dense GEMMs over every element (src/solver.cpp:556-1158,
    extrapolateU, calcGradU_spts, correctGradU, extrapolateGradU,
    extrapolateNormalFlux, calcDivF_spts, correctDivFlux), on the hex P3
    Taylor-Green shapes, through rocBLAS on the matrix cores: in FP64, and in
    BF16 with FP32 accumulation as hi+lo split pairs.  GEMMs only (no flux,
    face or RK work), to set beside the fused path's whole stage."""
    import ctypes as C
    from flurrypp_amd.mesh import operators
    hip = C.CDLL("libamdhip64.so")
    rb = C.CDLL("/opt/rocm/lib/librocblas.so")

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc})")

    op = operators(3, order)
    F, S = op.spts_to_fpts.shape
    N = 5
    cols = n ** 3 * N
    # (operator, count): A is rows x k in row-major, applied to a k x cols block
    ops = [(op.spts_to_fpts, 1), (op.grad_spts[0], 3), (op.correct_u[0], 3), (op.spts_to_fpts, 3),
           (op.extrap_fn[0], 3), (op.grad_spts[0], 3), (op.correction, 1)]
    flop = sum(2.0 * a.shape[0] * a.shape[1] * cols * c for a, c in ops)
    handle = C.c_void_p()
    ok(rb.rocblas_create_handle(C.byref(handle)), "rocblas_create_handle")
    bufs = []

    def dmalloc(nbytes, fill):
        p = C.c_void_p()
        ok(hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)), "hipMalloc")
        ok(hip.hipMemset(p, C.c_int(fill), C.c_size_t(nbytes)), "hipMemset")
        bufs.append(p)
        return p

    def upload(arr):
        arr = np.ascontiguousarray(arr)
        p = dmalloc(arr.nbytes, 0)
        ok(hip.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), C.c_int(1)), "hipMemcpy")
        return p

    def timed(run):
        run()
        ok(hip.hipDeviceSynchronize(), "sync")
        ev0, ev1 = C.c_void_p(), C.c_void_p()
        hip.hipEventCreate(C.byref(ev0))
        hip.hipEventCreate(C.byref(ev1))
        hip.hipEventRecord(ev0, None)
        for _ in range(reps):
            run()
        hip.hipEventRecord(ev1, None)
        hip.hipEventSynchronize(ev1)
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), ev0, ev1)
        return ms.value / reps

    out = {"workload": f"operator GEMMs of one NS RK stage, {n}^3 hex P{order} ({cols:,} columns), rocBLAS",
           "gemms_per_stage": sum(c for _, c in ops), "gflop_per_stage": round(flop / 1e9, 2)}
    try:
        # FP64: row-major C[m x cols] = A[m x k] B[k x cols] is column-major C^T = B^T A^T
        big = max(F, S) * cols * 8
        B64, C64 = dmalloc(big, 0x3F), dmalloc(big, 0)
        A64 = [(upload(a.astype(np.float64)), a.shape, c) for a, c in ops]
        one, zero = C.c_double(1.0), C.c_double(0.0)

        def run64():
            for pa, (m, k), c in A64:
                for _ in range(c):
                    ok(rb.rocblas_dgemm(handle, 111, 111, cols, m, k, C.byref(one), B64, cols, pa, k,
                                        C.byref(zero), C64, cols), "rocblas_dgemm")
        t64 = timed(run64)
        out["fp64"] = {"ms_per_stage": round(t64, 4), "TFLOP/s": round(flop / t64 / 1e9, 2),
                       "frac_of_fp64_matrix_peak": round(flop / t64 / 1e9 / FP64_PEAK_TFS, 4)}
        # BF16 inputs (hi and lo parts), FP32 accumulate and output
        Bb, Cf = dmalloc(max(F, S) * cols * 2, 0x3F), dmalloc(max(F, S) * cols * 4, 0)

        def split(a):
            # hi = a rounded down to BF16's 8 mantissa bits, lo = the rest (BF16 bit patterns)
            a32 = np.ascontiguousarray(a, dtype=np.float32)
            hi = (a32.view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32)
            lo = (a32 - hi).astype(np.float32)
            return [(x.view(np.uint32) >> 16).astype(np.uint16) for x in (hi, lo)]
        Abf = [(upload(h), upload(lo), a.shape, c) for a, c in ops for h, lo in [split(a)]]
        f1, f0 = C.c_float(1.0), C.c_float(0.0)

        def runbf():
            for ph, pl, (m, k), c in Abf:
                for _ in range(c):
                    for pa, beta in ((ph, f0), (pl, f1)):
                        ok(rb.rocblas_gemm_ex(handle, 111, 111, cols, m, k, C.byref(f1), Bb, 168, cols, pa, 168, k,
                                              C.byref(beta), Cf, 151, cols, Cf, 151, cols, 151, 0, 0, 0),
                           "rocblas_gemm_ex")
        tbf = timed(runbf)
        out["bf16_hi_lo"] = {"ms_per_stage": round(tbf, 4), "gemms": 2 * out["gemms_per_stage"],
                             "TFLOP/s": round(2 * flop / tbf / 1e9, 2)}
    finally:
        for p in bufs:
            hip.hipFree(p)
        rb.rocblas_destroy_handle(handle)
    return out


def kinetic_energy(setup, U) -> float:
    """E_k = (1/|Omega|) sum w detJ rho |u|^2 / 2 over solution points (FP64)."""
    D = setup.nDims
    wj = np.einsum("s,se->se", setup.opers.w_spts, setup.geo.detjac_spts)
    rho = U[..., 0]
    ke = 0.5 * np.sum(U[..., 1:D + 1] ** 2, axis=-1) / rho
    return float(np.sum(wj * ke) / np.sum(wj))


def kernel_table(kt: dict, work: dict) -> dict:
    """Per-kernel average time and launches, with the algorithmic bytes per
    launch, achieved GB/s and fraction of HBM peak where the model has them."""
    kernels = {}
    for name, (avg_ms, cnt) in kt.items():
        ent = {"avg_ms": round(avg_ms, 5), "launches": cnt}
        w = work.get(name)
        if w is not None and avg_ms > 0:
            ent["GB/s"] = round(w.bytes / (avg_ms * 1e-3) / 1e9, 1)
            ent["frac_of_hbm_peak"] = round(w.bytes / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
            ent["bytes_per_launch"] = w.bytes
        kernels[name] = ent
    return kernels


def traffic_for(kernel: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary."""
    p = ROOT / "profiles" / "traffic.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
        ent = d.get(workload, {}).get(kernel)
        return None if ent is None else float(ent["bytes_per_launch"])
    except Exception:
        return None


def kernel_pass(s, path: int, steps: int) -> dict:
    """Per-kernel average device times: `steps` more RK steps launched kernel
    by kernel with HIP events recorded on the context's stream around every
    launch (the captured graph of the timed region has no events inside),
    read after the pass."""
    s.enable_timing(True)
    s.sync()
    s.update(steps, path=path, sync=False)
    s.sync()
    kt = s.kernel_times()
    s.enable_timing(False)
    return kt


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--kernel-steps", type=int, default=20,
                    help="RK steps of the per-kernel timing pass after the timed region")
    ap.add_argument("--n", type=int, default=48, help="elements per direction")
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--inviscid", action="store_true")
    ap.add_argument("--path", choices=["fused", "staged"], default="fused")
    ap.add_argument("--cpu-n", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-procs", type=int, default=min(16, os.cpu_count() or 1),
                    help="concurrent single-thread copies for the all-cores CPU figure (<= 1 skips it)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default block sizes (256)")
    ap.add_argument("--no-dense", action="store_true", help="skip timing the dense operator GEMMs")
    ap.add_argument("--mixed-n", type=int, default=64,
                    help="elements per direction of the large mixed-precision Taylor-Green form (0 skips it)")
    ap.add_argument("--batch-cases", type=int, default=64, help="0 skips the batched form")
    ap.add_argument("--batch-tol", type=float, default=1e-10)
    ap.add_argument("--batch-max-steps", type=int, default=3000)
    ap.add_argument("--batch-nx", type=int, default=128)
    ap.add_argument("--batch-ny", type=int, default=48)
    ap.add_argument("--batch-cfl", type=str, default="0.9,0.9,0.9,0.9,0.9",
                    help="local CFL number of orders 1..5, comma-separated")
    ap.add_argument("--serve-requests", type=int, default=24, help="0 skips the served form")
    ap.add_argument("--serve-rates", type=str, default="5,20,80", help="Poisson arrival rates (per s), comma-separated")
    a = ap.parse_args()

    if a.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {a.gpus} is not supported: it measures one GPU's share "
              "of the path (a mesh shards by element partition, one context per GPU).", file=sys.stderr)
        return 2

    from flurrypp_amd.build import ensure_built
    ensure_built()
    from flurrypp_amd import cases, perfmodel
    from flurrypp_amd._native import PATH_FUSED, PATH_STAGED
    from flurrypp_amd.solver import Solver

    visc = 0 if a.inviscid else 1
    log(f"bench: setting up TGV {a.n}^3 P{a.order} {'NS' if visc else 'Euler'}")
    t = time.perf_counter()
    setup, U = cases.tgv3d(n=a.n, order=a.order, viscous=visc)
    log(f"bench: host set-up {time.perf_counter() - t:.1f} s, dt={setup.params.dt:.6e}")
    s = Solver(setup.params, setup=setup)
    s.initializeSolution(U)
    path = PATH_FUSED if a.path == "fused" else PATH_STAGED
    tuned = {}
    if a.path == "fused" and not a.no_tune:
        from flurrypp_amd._native import PATH_MIXED
        t = time.perf_counter()
        tuned["fp64"] = s.autotune(PATH_FUSED)
        if not a.no_mixed:
            tuned["mixed"] = s.autotune(PATH_MIXED)
        log(f"bench: autotuned block sizes {tuned} in {time.perf_counter() - t:.1f} s")
    s.update(a.warmup, path=path)
    s.check_state()
    log("bench: warmup done")

    # timed region: the production mode, one captured HIP graph replayed per
    # RK step (fused path), no events inside
    s.sync()
    t0 = time.perf_counter()
    s.update(a.steps, path=path, sync=False)
    s.sync()
    el = time.perf_counter() - t0
    s.check_state()
    U64 = s.get_U()
    kt = kernel_pass(s, path, a.kernel_steps)

    nRK = setup.params.nRKSteps
    dof = setup.nEles * setup.nSpts
    value = dof * nRK * a.steps / el / 1e9
    workload = f"tgv3d_{a.n}_p{a.order}_{'ns' if visc else 'euler'}_{a.path}"

    # roofline of the dominant kernel
    nFF = len(setup.faces.eleL) * setup.nFptsFace
    nFFb = setup.faces.nBnd * setup.nFptsFace
    work = perfmodel.fused_work(setup.nDims, setup.order, setup.nEles, nFF, nFFb, bool(visc), nRK,
                                setup.params.timeType)
    kernels = kernel_table(kt, work)
    dom = max((k for k in kernels if k in work), key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches"],
              default=None)
    roof = None
    if dom:
        w = work[dom]
        avg_s = kernels[dom]["avg_ms"] * 1e-3
        if perfmodel.hbm_bound(w):
            ach = w.bytes / avg_s / 1e9
            roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic_for(dom, workload), "kernel": dom}
        else:
            ach = w.flops / avg_s / 1e12
            roof = {"bound": "mfma", "achieved": round(ach, 2), "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                    "frac": round(ach / FP64_PEAK_TFS, 4), "traffic": traffic_for(dom, workload), "kernel": dom}

    forms = {}
    if a.path == "fused" and not a.no_mixed:
        # mixed-precision form on the same case, from the same initial state
        from flurrypp_amd._native import PATH_MIXED
        s.set_U(U)
        s.update(a.warmup, path=PATH_MIXED)
        s.sync()
        t0 = time.perf_counter()
        s.update(a.steps, path=PATH_MIXED, sync=False)
        s.sync()
        elm = time.perf_counter() - t0
        U32 = s.get_U()
        ktm = kernel_pass(s, PATH_MIXED, a.kernel_steps)
        rho_err = float(np.linalg.norm(U32[..., 0] - U64[..., 0]) / np.linalg.norm(U64[..., 0]))
        u_err = float(np.linalg.norm(U32 - U64) / np.linalg.norm(U64))
        T = (a.warmup + a.steps) * setup.params.dt
        ek0, ek64, ek32 = (kinetic_energy(setup, x) for x in (U, U64, U32))
        eps64, eps32 = -(ek64 - ek0) / T, -(ek32 - ek0) / T
        forms["mixed"] = {
            "value": round(dof * nRK * a.steps / elm / 1e9, 4), "unit": "GDOF·stage/s",
            "ms_per_step": round(elm / a.steps * 1e3, 4), "speedup_vs_fp64": round(el / elm, 3),
            "rel_l2_density_vs_fp64": rho_err, "rel_l2_state_vs_fp64": u_err,
            "ke_dissipation_rate_fp64": eps64, "ke_dissipation_rate_mixed": eps32,
            "ke_dissipation_rel_err": abs(eps32 - eps64) / abs(eps64) if eps64 else None,
            "after_steps": a.warmup + a.steps,
            "kernels": kernel_table(ktm, perfmodel.mixed_work(setup.nDims, setup.order, setup.nEles, nFF, nFFb,
                                                              bool(visc), nRK, setup.params.timeType)),
        }

    # one plain copy of the FP64 state (E*S*N doubles read and written) after
    # the timed regions: the known-size kernel that scales rocprofv3's
    # FETCH_SIZE / WRITE_SIZE counters to bytes
    s.copyUspts_U0()
    s.sync()
    s.close()
    def optional(name, fn):
        # a side form that fails (a missing library, say) must not cost the headline line
        try:
            return fn()
        except Exception as ex:
            log(f"bench: {name} form failed: {ex!r}")
            return {"error": repr(ex)}

    if a.path == "fused":
        forms["vortex2d_32"] = optional("vortex2d_32", lambda: vortex2d_form(100, not a.no_tune))
        if visc and not a.no_dense:
            log("bench: dense operator GEMMs (rocBLAS)")
            dense = optional("dense_operators", lambda: dense_operator_form(a.n, a.order))
            if "error" not in dense:
                dense["fused_fp64_stage_ms"] = round(sum(kernels[k]["avg_ms"] for k in kernels
                                                         if k in work and k != "fused_extrapolate"), 4)
            forms["dense_operators"] = dense
    if a.path == "fused" and not a.no_mixed and a.mixed_n > 0:
        log(f"bench: mixed-precision form at {a.mixed_n}^3")
        forms[f"mixed_{a.mixed_n}"] = optional(f"mixed_{a.mixed_n}", lambda: tgv_mixed_form(
            a.mixed_n, a.order, a.steps, a.warmup, not a.no_tune))
    if a.batch_cases > 0:
        log(f"bench: batched form ({a.batch_cases} cases)")
        forms["batched"] = optional("batched", lambda: batched_form(
            a.batch_cases, 11, a.batch_tol, a.batch_max_steps, a.batch_nx, a.batch_ny,
            [float(x) for x in a.batch_cfl.split(",")]))
    if a.serve_requests > 0:
        log(f"bench: served form ({a.serve_requests} requests)")
        forms["served"] = optional("served", lambda: served_form(
            a.serve_requests, [float(x) for x in a.serve_rates.split(",")], 13))

    cpu = None
    if not a.no_cpu:
        log("bench: CPU baseline (oracle restatement)")
        cpu = cpu_baseline(a.cpu_n, a.cpu_steps)
        if a.cpu_procs > 1:
            log(f"bench: CPU baseline in {a.cpu_procs} processes")
            cpu["processes"] = optional("cpu processes",
                                        lambda: cpu_baseline_processes(a.cpu_n, a.cpu_steps, a.cpu_procs))

    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "GDOF·stage/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(el / a.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D {'Navier-Stokes' if visc else 'Euler'} Taylor-Green vortex, {a.n}^3 periodic "
                               f"hex, P{a.order}, Rusanov{'+LDG' if visc else ''}, RK44, FP64",
                   "elements": setup.nEles, "dof": dof, "stages_per_step": nRK, "path": a.path,
                   "timing": ("timed region replays one captured HIP graph per RK step; per-kernel averages "
                              f"from a further {a.kernel_steps}-step pass with HIP events around every launch"),
                   "dt": setup.params.dt,
                   "block_sizes": tuned or "default 256 (extrapolate, face_inviscid, volume, face_boundary, rk)"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "forms": forms,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
