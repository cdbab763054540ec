"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle) against the oracle and the host set-up, on the CPU.

The seeded inputs must hash to the fixture's value; the face maps the host
set-up hands the oracle must equal the fixture's bit for bit; the oracle must
reproduce the stage-0 residual and the 10-step state, residual norm and dt.
The oracle is deterministic C (no FMA contraction), so the bars are
round-off, not tolerances for a different algorithm.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle.oracle import Oracle
from tests.golden.make_golden import CASES, STEPS, u0_hash

GOLDEN = Path(__file__).resolve().parent / "golden"


def _load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reproduces_golden(name):
    g = _load(name)
    setup, U0 = CASES[name]()
    assert u0_hash(U0) == str(g["u0_sha256"]), "seeded inputs changed"
    o = Oracle.from_setup(setup)
    o.initializeSolution(U0, setup.params.dtType)
    assert np.array_equal(o.get("face_fpt_L"), g["face_fpt_L"])
    assert np.array_equal(o.get("face_fpt_R"), g["face_fpt_R"])
    o.calcResidual(0)
    assert rel(o.get("divF_spts", 0), g["divF0"]) <= 1e-14
    t = Oracle.from_setup(setup)
    t.initializeSolution(U0, setup.params.dtType)
    t.update(STEPS)
    assert rel(t.get_U(), g["U10"]) <= 1e-14
    np.testing.assert_allclose(t.residual_norm(2), g["res10"], rtol=1e-12, atol=0)


def test_golden_fixtures_present():
    for name in CASES:
        assert (GOLDEN / f"{name}.npz").exists(), name
