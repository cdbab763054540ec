"""Batch members on the CPU oracle: member i of the 64-case batched bump
form (the same seeded draws as bench.py's batched form) stepped on the oracle
as the device steps it alone (fine update, then a P-multigrid cycle over
orders 0..P-1 with 2 smoothing steps, per step), its density residual checked
every 50 steps, until the residual is NaN, reaches 1e-10 or max_steps.

    python tests/batched_divergence.py MEMBER MAX_STEPS

prints one JSON line (status, the step of the last check, residual history).
The device record of the same members is tests/batched64_members.json."""
from __future__ import annotations

import copy
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from flurrypp_amd import cases  # noqa: E402
from flurrypp_amd.mesh import pmg_operators  # noqa: E402
from flurrypp_amd.solver import HostSetup  # noqa: E402
from oracle.oracle import Oracle, pmg_cycle  # noqa: E402

SEED, CHECK = 11, 50


def member(i: int):
    """(setup, U, (nx, ny, P, M)) of member i: bench.py's batched_form draws."""
    rng = np.random.default_rng(SEED)
    for _ in range(i + 1):
        nx, ny = int(rng.integers(16, 129)), int(rng.integers(8, 49))
        P, M = int(rng.integers(1, 6)), float(rng.uniform(0.2, 0.5))
    s, U = cases.bump2d(nx=nx, ny=ny, order=P, mach=M, seed=SEED * 100003 + i, time_type=6, cfl=0.9)
    return s, U, (nx, ny, P, M)


def run_member(i: int, max_steps: int, tol: float = 1e-10) -> dict:
    s, U, tag = member(i)
    lv = []
    for q in range(0, s.order):
        pc = copy.copy(s.params)
        pc.order = q
        lv.append(HostSetup(pc, s.mesh))
    lv.append(s)
    ors = [Oracle.from_setup(h) for h in lv]
    for h, o in zip(lv[:-1], ors[:-1]):
        o.initializeSolution(h.initial_condition(), h.params.dtType)
    ors[-1].initializeSolution(U, s.params.dtType)
    R = [pmg_operators(2, h.order)[1] for h in lv[1:]]
    Pr = [pmg_operators(2, h.order)[0] for h in lv[:-1]]
    t0 = time.perf_counter()
    hist, status, step = [], "max_steps", 0
    for step in range(1, max_steps + 1):
        ors[-1].update(1)
        pmg_cycle(ors, R, Pr, 2)
        if step % CHECK == 0:
            r = float(ors[-1].residual_norm(2)[0])
            hist.append((step, r))
            if not np.isfinite(r):
                status = "diverged"
                break
            if r <= tol:
                status = "converged"
                break
    return {"i": i, "nx": tag[0], "ny": tag[1], "P": tag[2], "M": tag[3], "status": status, "steps": step,
            "cpu_s": round(time.perf_counter() - t0, 1), "hist": hist}


if __name__ == "__main__":
    print(json.dumps(run_member(int(sys.argv[1]), int(sys.argv[2]))), flush=True)
