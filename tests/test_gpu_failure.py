"""Failure detection (the reference's NaN / instability checks,
src/output.cpp:1016-1045, which exit): the library reports a non-physical
state as FHIP_ESTATE instead, and in a batch one failing mesh is reported
diverged while every other mesh evolves exactly as it does alone."""
import numpy as np
import pytest

from flurrypp_amd import cases
from flurrypp_amd._native import PATH_FUSED, FhipError

FHIP_ESTATE = 3  # include/flurry_hip.h
from flurrypp_amd.batch import BatchedSolver, Member
from flurrypp_amd.solver import Solver

pytestmark = pytest.mark.gpu


def test_check_state_flags_nan_and_negative_pressure():
    setup, U = cases.vortex2d(n=6, order=2)
    s = Solver(setup.params, setup=setup)
    s.initializeSolution(U)
    s.check_state()                       # a physical state passes
    bad = U.copy()
    bad[3, 5, 0] = np.nan
    s.set_U(bad)
    with pytest.raises(FhipError, match="non-finite or non-physical") as ei:
        s.check_state()
    assert ei.value.code == FHIP_ESTATE
    bad = U.copy()
    bad[0, 2, 3] = 0.0                    # energy 0: negative pressure
    s.set_U(bad)
    with pytest.raises(FhipError, match="non-finite or non-physical"):
        s.check_state()
    s.set_U(U)
    s.check_state()
    s.close()


def test_nan_in_one_batch_member_leaves_the_others_exact():
    ms = []
    for nx, ny, P, M in ((8, 4, 2, 0.3), (10, 4, 2, 0.4), (9, 5, 2, 0.25)):
        s, U = cases.bump2d(nx=nx, ny=ny, order=P, mach=M, seed=5 + nx)
        ms.append(Member(s, U))
    poisoned = [Member(m.setup, m.U.copy()) for m in ms]
    poisoned[1].U[0, 0, 0] = np.nan
    b = BatchedSolver(poisoned)
    out = b.solve_steady(tol=1e-30, max_steps=60, check_every=20, stop_on_nan=False)
    assert list(out["diverged"]) == [False, True, False], out["diverged"]
    got = b.states()
    b.close()
    for i in (0, 2):
        s = Solver(ms[i].setup.params, setup=ms[i].setup)
        s.initializeSolution(ms[i].U)
        s.update(60, path=PATH_FUSED)
        assert np.array_equal(got[i], s.get_U()), f"member {i} differs"
        s.close()
