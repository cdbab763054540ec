"""Checkpoint / resume through FlurryPP's plot file: N steps, write, restart
from the file in a new solver, N more steps == 2N steps straight (to the
16-digit text precision of the file).  Fixed dt: with a CFL dt the restarted
run refreshes its wave speeds on its first step (iter == initIter+1), as
FlurryPP does, so it is not expected to equal the straight run."""
import copy

import numpy as np
import pytest

from flurrypp_amd import cases
from flurrypp_amd.solver import HostSetup, Solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("make", [lambda: cases.vortex2d(n=6, order=3, dt_type=0),
                                  lambda: cases.tgv3d(n=3, order=2, viscous=1, jitter=0.1)])
def test_restart_continues_the_run(tmp_path, make):
    setup, U = make()
    p = setup.params
    p.dataFileName = "chk"
    a = Solver(p, setup=setup)
    a.initializeSolution(U)
    a.update(10)
    p.time = 10 * a.dt
    a.writeData(tmp_path)
    it = a.iter
    a.update(10)
    ref = a.get_U()
    a.close()

    q = copy.copy(p)
    q.restart, q.restartIter, q.initIter = 1, it, it
    q.time = 0.0
    b = Solver(q, setup=HostSetup(q, setup.mesh))
    b.initializeSolution(directory=tmp_path)
    assert q.time == pytest.approx(p.time, rel=1e-15)
    b.update(10)
    got = b.get_U()
    b.close()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-13
