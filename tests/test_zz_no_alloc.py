"""fhip_residual_norm_noalloc, called through the C ABI, against
fhip_residual_norm on the same context: the same bits for every norm type,
on two calls in a row whose second context is smaller (the scratch array
keeps the first call's contents), and FHIP_ESIZE one element count past its
documented maximum."""
import ctypes as C

import numpy as np
import pytest

from flurrypp_amd import _native as nat
from flurrypp_amd import cases
from flurrypp_amd.solver import Solver

pytestmark = pytest.mark.gpu

FHIP_ESIZE = 4


def _lib():
    lib = nat.hip_lib()
    lib.fhip_residual_norm_noalloc.argtypes = [C.c_void_p, C.c_int, nat._dp]
    return lib


def _norms(s: Solver, nfields: int, fn) -> list[np.ndarray]:
    out = []
    for t in (1, 2, 3):
        r = np.full(nfields, np.nan)
        nat.hip_check(fn(s._ctx, t, nat.dptr(r)))
        out.append(r)
    return out


@pytest.mark.parametrize("path", [nat.PATH_FUSED, nat.PATH_STAGED])
def test_noalloc_matches_residual_norm_twice(path):
    lib = _lib()
    print(f"no-alloc residual norm, path {path}: TGV 8^3 P3 then vortex 6x6 P2", flush=True)
    big = cases.tgv3d(n=8, order=3, viscous=1)
    small = cases.vortex2d(n=6, order=2)
    for setup, U in (big, small):  # second call: fewer elements, fewer fields
        s = Solver(setup.params, setup=setup)
        s.initializeSolution(U)
        s.update(3, path=path)
        ref = _norms(s, setup.nFields, lib.fhip_residual_norm)
        got = _norms(s, setup.nFields, lib.fhip_residual_norm_noalloc)
        for t, (a, b) in enumerate(zip(got, ref), start=1):
            assert np.all(np.isfinite(b)) and np.all(b > 0), (t, b)
            assert np.array_equal(a, b), (t, a, b)
        s.close()


def test_noalloc_size_limit():
    """The largest accepted 2-D context (1023^2 elements) works; 1025^2,
    past (16384/4 - 1) * 256 = 1,048,320 elements, returns FHIP_ESIZE."""
    lib = _lib()
    print("no-alloc residual norm: 1023^2 and 1025^2 P0 quads", flush=True)
    for n, ok in ((1023, True), (1025, False)):
        setup, U = cases.vortex2d(n=n, order=0, jitter=0.0)
        s = Solver(setup.params, setup=setup)
        s.initializeSolution(U)
        s.calcResidual(0)
        r = np.zeros(4)
        rc = lib.fhip_residual_norm_noalloc(s._ctx, 2, nat.dptr(r))
        if ok:
            assert rc == 0, nat.hip_lib().fhip_last_error()
            ref = np.zeros(4)
            nat.hip_check(lib.fhip_residual_norm(s._ctx, 2, nat.dptr(ref)))
            assert np.array_equal(r, ref)
        else:
            assert rc == FHIP_ESIZE
            assert b"maximum" in nat.hip_lib().fhip_last_error()
        s.close()
