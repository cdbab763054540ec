"""Writes the golden fixtures of the FR residual + RK update under tests/golden/.

Each fixture is the oracle's output (the C restatement under oracle/, test
infrastructure) on one small seeded case of the scope's configurations:

* ``u0_sha256``  SHA-256 of the initial state's FP64 bytes (the case is rebuilt
  from its seeded generator; the hash proves the inputs are the same),
* ``face_fpt_L``/``face_fpt_R``  the face flux-point maps (int32),
* ``divF0``  divF_spts[0] of calcResidual(0) at the initial state, and
  ``F_scale`` the L2 norm of F_spts there (divF is formed by cancellation and
  is measured against the size of the fluxes that cancel),
* ``U10``, ``res10``, ``dt10``  state, L2 residual norm and dt after 10 update()
  calls (solver::update, src/solver.cpp:319-363).

Run from the repository root: ``python tests/golden/make_golden.py``.
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from flurrypp_amd import cases  # noqa: E402

HERE = Path(__file__).resolve().parent
STEPS = 10

# name -> seeded case; shapes follow the scope's configurations at oracle size
CASES = {
    # C1 shape: 2-D isentropic vortex, jittered periodic quads, P3, Rusanov, RK44, global CFL
    "vortex2d_p3_8x8_rusanov": lambda: cases.vortex2d(n=8, order=3),
    # 2-D Roe flux, P2
    "vortex2d_p2_6x6_roe": lambda: cases.vortex2d(n=6, order=2, riemann=1),
    # C4 shape: Gaussian-bump channel, characteristic ends, slip walls, local dt
    "bump2d_p2_12x5_char_slip_localdt": lambda: cases.bump2d(nx=12, ny=5, order=2),
    # C2 shape: Navier-Stokes Taylor-Green, jittered periodic hexes, P3, Rusanov + LDG
    "tgv3d_p3_3cube_ns": lambda: cases.tgv3d(n=3, order=3, viscous=1, jitter=0.1),
    # C5 shape: 3-D periodic Euler vortex, P2
    "vortex3d_p2_3cube": lambda: cases.vortex3d(n=3, order=2),
}


def u0_hash(U: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(U, dtype=np.float64).tobytes()).hexdigest()


def oracle_outputs(setup, U0) -> dict:
    from oracle.oracle import Oracle
    dt_type = setup.params.dtType
    o = Oracle.from_setup(setup)
    o.initializeSolution(U0, dt_type)
    out = {"face_fpt_L": o.get("face_fpt_L").astype(np.int32), "face_fpt_R": o.get("face_fpt_R").astype(np.int32)}
    o.calcResidual(0)
    out["divF0"] = o.get("divF_spts", 0).copy()
    out["F_scale"] = np.float64(np.linalg.norm(o.get("F_spts")))
    t = Oracle.from_setup(setup)
    t.initializeSolution(U0, dt_type)
    t.update(STEPS)
    out["U10"] = t.get_U().copy()
    out["res10"] = np.asarray(t.residual_norm(2), dtype=np.float64)
    out["dt10"] = np.float64(t.calcDt(False) if dt_type != 0 else setup.params.dt)
    return out


def main() -> int:
    for name, make in CASES.items():
        setup, U0 = make()
        d = oracle_outputs(setup, U0)
        d["u0_sha256"] = np.array(u0_hash(U0))
        np.savez_compressed(HERE / f"{name}.npz", **d)
        print(f"{name}: {sum(v.nbytes for v in d.values()) / 1024:.0f} KiB", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
