"""Block shapes of the fused kernels: every valid block size gives the same
state bit for bit, and fhip_autotune leaves the state unchanged and picks
valid sizes."""
import numpy as np
import pytest

from flurrypp_amd import cases
from flurrypp_amd._native import PATH_FUSED, PATH_MIXED
from flurrypp_amd.solver import Solver

pytestmark = pytest.mark.gpu

CASES = {
    "tgv3d_p3_ns": lambda: cases.tgv3d(n=3, order=3, viscous=1, jitter=0.1),
    "bump2d_p2": lambda: cases.bump2d(nx=10, ny=4, order=2),
    "vortex2d_p5": lambda: cases.vortex2d(n=4, order=5),
}


def _run(make, path, sizes=None, tune=False, steps=5):
    setup, U = make()
    s = Solver(setup.params, setup=setup)
    s.initializeSolution(U)
    got = None
    if tune:
        before = s.get_U()
        got = s.autotune(path)
        assert np.array_equal(before, s.get_U())
    if sizes is not None:
        s.set_block_sizes(sizes, path)
        assert s.block_sizes(path) == list(sizes)
    s.update(steps, path=path)
    out = s.get_U()
    s.close()
    return out, got


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("path", [PATH_FUSED, PATH_MIXED])
def test_block_sizes_do_not_change_results(name, path):
    make = CASES[name]
    ref, _ = _run(make, path)
    setup, U = make()
    probe = Solver(setup.params, setup=setup)
    tried = 0
    for k in range(5):
        for nt in (64, 128, 192, 320, 512):
            sizes = [256] * 5
            sizes[k] = nt
            try:
                probe.set_block_sizes(sizes, path)   # rejected: fewer threads than points, or LDS > 64 KiB
            except Exception:
                continue
            got, _ = _run(make, path, sizes=sizes)
            assert np.array_equal(got, ref), (name, sizes)
            tried += 1
    probe.close()
    assert tried >= 10


@pytest.mark.parametrize("path", [PATH_FUSED, PATH_MIXED])
def test_autotune_keeps_state_and_results(path):
    make = CASES["tgv3d_p3_ns"]
    ref, _ = _run(make, path)
    got, sizes = _run(make, path, tune=True)
    assert all(64 <= x <= 512 for x in sizes), sizes
    assert np.array_equal(got, ref)


def test_invalid_block_size_rejected():
    setup, U = CASES["vortex2d_p5"]()
    s = Solver(setup.params, setup=setup)
    with pytest.raises(Exception):
        s.set_block_sizes([256, 256, 32, 256, 256])   # 32 threads < 36 points per element
    s.close()


@pytest.mark.parametrize("path", [PATH_FUSED, PATH_MIXED])
def test_default_sizes_fit_p0_hexes(path):
    # 256 threads of P0 hexes (one point per element) would need 70 KiB of LDS
    # in the FP64 RK and viscous volume kernels: the defaults come down to fit
    make = lambda: cases.tgv3d(n=4, order=0, viscous=1, jitter=0.1)
    setup, U = make()
    s = Solver(setup.params, setup=setup)
    sizes = s.block_sizes(path)
    s.set_block_sizes(sizes, path)  # raises unless every size is valid
    s.close()
    if path == PATH_FUSED:
        assert sizes[2] <= 192 and sizes[4] <= 192, sizes
    ref, _ = _run(make, path)
    got, _ = _run(make, path, sizes=[128] * 5)
    assert np.array_equal(got, ref)
