"""The main loop (flurrypp_amd.solver.run, src/flurry.cpp:120-155): residual
monitor lines at iteration initIter+1, every monitorResFreq iterations and
when maxTime is reached; the state after the loop equals the same number of
update() calls."""
import numpy as np
import pytest

from flurrypp_amd.input import InputParams
from flurrypp_amd.solver import Solver, run

pytestmark = pytest.mark.gpu


def _params(**kw):
    d = dict(equation=1, nDims=2, order=2, meshType=1, nx=6, ny=6, xmin=-5, xmax=5, ymin=-5, ymax=5,
             timeType=4, dtType=1, CFL=0.5, maxTime=1e9, iterMax=12, icType=2, monitorResFreq=5, plotFreq=1000,
             resType=2)
    d.update(kw)
    return InputParams.from_dict(d)


def test_run_monitor_and_state():
    lines = []
    s = run(_params(), log=lines.append, output_dir=None)
    its = [int(x.split()[0]) for x in lines]
    assert its == [1, 5, 10], lines
    assert s.iter == 12
    got = s.get_U()
    s.close()
    ref = Solver(_params())
    ref.initializeSolution()
    ref.update(12)
    assert np.array_equal(got, ref.get_U())
    ref.close()


def test_run_stops_at_max_time():
    probe = Solver(_params())
    probe.initializeSolution()
    probe.update(1)
    dt = probe.dt
    probe.close()
    lines = []
    # maxTime between the 3rd and 4th step (dt is refreshed every step; at CFL 0.5
    # on this smooth vortex it changes by well under 1% over a few steps)
    s = run(_params(maxTime=3.5 * dt, monitorResFreq=100), log=lines.append, output_dir=None)
    assert s.iter == 4, s.iter
    its = [int(x.split()[0]) for x in lines]
    assert its == [1, 4], lines   # initIter+1, then the step that reached maxTime
    s.close()
