"""FlurryPP input files (include/input.hpp, src/input.cpp:257-596): keys,
defaults, RK tables and the viscous non-dimensionalisation.

The non-dimensional state is pinned by what the physics fixes, not by the
reference's own formulas: at the free stream, Sutherland's law evaluated the
way the viscous flux evaluates it gives mu = 1/Re; the temperature recovered
from the state gives T/TGas = TBound/TGas; p = 1/(gamma M^2) with rho = |u| = 1;
the moving wall's speed is MachWall sqrt(TWall) / (MachBound sqrt(TBound)).
When the reference's test inputs are present they are parsed too (else those
cases skip)."""
import math
from pathlib import Path

import pytest

from flurrypp_amd.input import InputParams

REF_TESTS = Path("/root/reference/tests")

COUETTE = dict(equation=1, order=2, timeType=6, dtType=2, CFL=0.2, dt=1e-6, iterMax=10000, maxTime=5000,
               viscous=1, riemannType=0, nDims=2, PMG=1, smoothSteps=1, icType=0, meshType=0,
               meshFileName="couette.msh", periodicDX=2, rhoBound=1, uBound=0.1, vBound=0.0, wBound=0.0,
               pBound=0.7142857143, MachBound=0.2, Re=200, Lref=1.0, TBound=300, nxBound=1, nyBound=0, nzBound=0,
               TWall=300, MachWall=0.2, nxWall=1, nyWall=0, nzWall=0, LDG_penFact=0, LDG_tau=0.1,
               mesh_bound={"bottom": "isothermal_noslip", "top": "isothermal_noslip_moving", "left": "periodic",
                           "right": "periodic", "fluid": "fluid"})


def _check_nondimensional(p, re, mach, t_bound, t_wall, mach_wall):
    g = p.gamma
    assert p.rhoBound == 1.0
    assert math.isclose(math.hypot(p.uBound, p.vBound, getattr(p, "wBound", 0.0)), 1.0, rel_tol=1e-14)
    assert math.isclose(p.pBound, 1.0 / (g * mach * mach), rel_tol=1e-13)
    # Sutherland's law as the viscous flux applies it, at the free-stream temperature
    e = p.pBound / ((g - 1) * p.rhoBound)
    rt = (g - 1) * e / p.rt_inf
    mu = p.mu_inf * rt ** 1.5 * (1 + p.c_sth) / (rt + p.c_sth)
    assert math.isclose(mu, 1.0 / re, rel_tol=1e-13)
    assert math.isclose(rt, t_bound / p.TGas, rel_tol=1e-13)
    assert math.isclose(p.TWall, t_wall / t_bound, rel_tol=1e-14)
    assert math.isclose(p.uWall, mach_wall * math.sqrt(t_wall) / (mach * math.sqrt(t_bound)), rel_tol=1e-13)


def test_couette_options_and_nondimensional_state():
    p = InputParams.from_dict(COUETTE)
    assert (p.order, p.timeType, p.dtType, p.PMG, p.smoothSteps, p.lowOrder) == (2, 6, 2, 1, 1, 0)
    assert p.nRKSteps == 4 and p.RKa == [0.0, 0.153, 0.442, 0.931, 1.0]   # tuned steady RK, P <= 2
    assert (p.penFact, p.tau) == (0.0, 0.1)
    _check_nondimensional(p, re=200, mach=0.2, t_bound=300, t_wall=300, mach_wall=0.2)


def test_rk_tables():
    base = dict(equation=1, nDims=2, order=3, meshType=1, nx=2, ny=2, viscous=0, iterMax=10, dtType=0, dt=1e-3)
    for tt, n in ((0, 1), (2, 2), (4, 4), (5, 4), (6, 4)):
        p = InputParams.from_dict(dict(base, timeType=tt))
        assert p.nRKSteps == n
    p4 = InputParams.from_dict(dict(base, timeType=4))
    assert math.isclose(sum(p4.RKb), 1.0, rel_tol=1e-15)          # consistent RK weights
    p6 = InputParams.from_dict(dict(base, timeType=6))
    assert p6.RKa == [0.0, 0.158, 0.477, 1.0, 1.0]                  # tuned steady RK, P > 2


@pytest.mark.skipif(not REF_TESTS.exists(), reason="reference test inputs not present")
def test_reference_input_files():
    ch = InputParams.from_file(REF_TESTS / "euler/channel/input_channelPMG")
    assert (ch.order, ch.timeType, ch.dtType, ch.CFL, ch.PMG, ch.lowOrder, ch.smoothSteps) == (3, 6, 2, 0.9, 1, 0, 2)
    assert (ch.monitorResFreq, ch.plotFreq, ch.iterMax, ch.resType) == (200, 2000, 4000, 2)
    assert ch.meshFileName == "channelBump.msh" and ch.viscous == 0
    co = InputParams.from_file(REF_TESTS / "navier-stokes/couette/input_couette")
    ref = InputParams.from_dict(COUETTE)
    for k in ("order", "timeType", "dtType", "CFL", "viscous", "PMG", "smoothSteps", "penFact", "tau", "mu_inf",
              "rt_inf", "c_sth", "pBound", "uBound", "TWall", "uWall", "periodicDX", "RKa"):
        assert getattr(co, k) == getattr(ref, k), k
