"""The HIP paths against the committed golden fixtures (tests/golden/*.npz,
the oracle's outputs on five small seeded cases): face maps bit-exact, the
staged path's stage-0 residual within 1e-13 relative L2 (measured, like the
other residual parity tests, against the larger of its own norm and that of
the fluxes whose divergence it is), the state after 10
RK steps within 1e-12 on the staged and fused FP64 paths and within 1e-6 on
the mixed path, the residual norm within 1e-10 (relative to the largest
field's)."""
from pathlib import Path

import numpy as np
import pytest

from flurrypp_amd._native import PATH_FUSED, PATH_MIXED, PATH_STAGED
from flurrypp_amd.solver import Solver
from tests.golden.make_golden import CASES, STEPS, u0_hash

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", list(CASES))
def test_maps_and_residual_match_golden(name):
    g = _load(name)
    setup, U0 = CASES[name]()
    assert u0_hash(U0) == str(g["u0_sha256"])
    print(f"golden {name}: maps and residual", flush=True)
    s = Solver(setup.params, setup=setup)
    s.initializeSolution(U0)
    assert np.array_equal(s.get("face_fpt_L"), g["face_fpt_L"])
    assert np.array_equal(s.get("face_fpt_R"), g["face_fpt_R"])
    s.calcResidual(0)
    d = s.get("divF_spts", 0)
    err = float(np.linalg.norm(d - g["divF0"]) / max(np.linalg.norm(g["divF0"]), float(g["F_scale"])))
    s.close()
    assert err <= 1e-13, err


@pytest.mark.parametrize("path,tol", [(PATH_STAGED, 1e-12), (PATH_FUSED, 1e-12), (PATH_MIXED, 1e-6)])
@pytest.mark.parametrize("name", list(CASES))
def test_trajectory_matches_golden(name, path, tol):
    g = _load(name)
    setup, U0 = CASES[name]()
    print(f"golden {name}: path {path}", flush=True)
    s = Solver(setup.params, setup=setup)
    if path != PATH_STAGED and not s._fused_ok():
        s.close()
        pytest.skip("fused path does not cover this configuration")
    s.initializeSolution(U0)
    s.update(STEPS, path=path)
    err = rel(s.get_U(), g["U10"])
    res = s.residual_norm(2)
    s.close()
    print(f"golden {name}: path {path} U rel L2 {err:.3e}", flush=True)
    assert err <= tol, err
    if path != PATH_MIXED:
        # fields at round-off level (w of a vortex column) compare absolutely
        np.testing.assert_allclose(res, g["res10"], rtol=1e-10, atol=1e-10 * float(np.max(np.abs(g["res10"]))))
