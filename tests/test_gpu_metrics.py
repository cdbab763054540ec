"""Metric terms rebuilt from the element vertices inside the fused kernels:
fhip_create verifies them against the stored arrays (bit for bit), the
option can then be turned on, and runs with and without it give the same
state bit for bit."""
import numpy as np
import pytest

from flurrypp_amd import _native as nat
from flurrypp_amd import cases
from flurrypp_amd._native import PATH_FUSED, PATH_MIXED
from flurrypp_amd.solver import Solver

pytestmark = pytest.mark.gpu

CASES = {
    "tgv3d_p3_ns_jitter": lambda: cases.tgv3d(n=3, order=3, viscous=1, jitter=0.1),
    "bump2d_p2": lambda: cases.bump2d(nx=10, ny=4, order=2),
    "vortex2d_p4_jitter": lambda: cases.vortex2d(n=4, order=4),
    "tgv3d_p1_euler_jitter": lambda: cases.tgv3d(n=4, order=1, viscous=0, jitter=0.1),
}


def _opt(s):
    v = np.zeros(1, dtype=np.int32)
    nat.hip_check(nat.hip_lib().fhip_get_option(s._ctx, nat.OPT_METRICS_FROM_NODES, nat.iptr(v)))
    return int(v[0])


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("path", [PATH_FUSED, PATH_MIXED])
def test_metrics_from_vertices_match_stored(name, path):
    out = []
    for on in (1, 0):
        setup, U = CASES[name]()
        s = Solver(setup.params, setup=setup)
        assert _opt(s) == 0, "off by default"
        nat.hip_check(nat.hip_lib().fhip_set_option(s._ctx, nat.OPT_METRICS_FROM_NODES, on))
        assert _opt(s) == on, "vertex metrics not verified at fhip_create"
        s.initializeSolution(U)
        s.update(6, path=path)
        out.append(s.get_U())
        s.close()
    assert np.array_equal(out[0], out[1])
