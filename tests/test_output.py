"""Plot / restart files (no GPU): FlurryPP's .vtu layout written and read back."""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from flurrypp_amd import cases
from flurrypp_amd.mesh import ref_points
from flurrypp_amd.output import plot_positions, read_restart, write_csv, write_vtu
from flurrypp_amd.solver import HostSetup


def _perturbed(make):
    s, U = make()
    U = U * (1 + 0.01 * np.sin(np.arange(U.size)).reshape(U.shape))
    return s, U


@pytest.mark.parametrize("make", [lambda: cases.bump2d(nx=5, ny=3, order=2),
                                  lambda: cases.tgv3d(n=2, order=3, viscous=1, jitter=0.1)])
def test_vtu_round_trip(tmp_path, make):
    s, U = _perturbed(make)
    f = write_vtu(tmp_path / "run_000000007.vtu", s, U, time=1.25, iteration=7)
    root = ET.parse(f).getroot()
    pieces = root.findall("./UnstructuredGrid/Piece")
    assert len(pieces) == s.nEles
    n1 = s.order + 3
    assert pieces[0].get("NumberOfPoints") == str(n1 ** s.nDims)
    assert pieces[0].get("NumberOfCells") == str((s.order + 2) ** s.nDims)
    U2, t = read_restart(f, s)
    assert t == 1.25
    assert np.abs(U2 - U).max() <= 1e-14 * np.abs(U).max()


def test_plot_points_contain_solution_points_and_vertices():
    s, _ = cases.bump2d(nx=4, ny=3, order=3)
    X = plot_positions(s)
    n1 = s.order + 3
    P1 = s.order + 1
    for spt in range(P1 * P1):
        i, j = spt % P1, spt // P1
        np.testing.assert_allclose(X[(i + 1) + n1 * (j + 1)], s.geo.pos_spts[spt], rtol=0, atol=1e-13)
    xv, c2v = s.mesh.nodes()
    for corner, node in ((0, 0), (n1 - 1, 1), (n1 * n1 - 1, 2), (n1 * (n1 - 1), 3)):
        np.testing.assert_allclose(X[corner], xv[c2v[:, node]], rtol=0, atol=1e-13)


def test_restart_at_another_order_interpolates_exactly():
    """A file written at P=2 read by a P=3 solver: conservative fields that are
    degree-2 polynomials in the reference coordinates come back exact."""
    s2, _ = cases.bump2d(nx=4, ny=3, order=2)
    p3 = __import__("copy").copy(s2.params)
    p3.order = 3
    s3 = HostSetup(p3, s2.mesh)

    def field(xr, e):
        x, y = xr[:, 0], xr[:, 1]
        rho = 1.0 + 0.1 * x + 0.05 * y * y + 0.01 * e
        mx = 0.3 + 0.02 * x * y
        my = -0.01 * y
        E = 2.5 + 0.1 * x * x - 0.03 * y
        return np.stack([rho, mx, my, E], axis=-1)

    x2, _, _ = ref_points(2, 2)
    x3, _, _ = ref_points(2, 3)
    U2 = np.stack([field(x2, e) for e in range(s2.nEles)], axis=1)
    U3 = np.stack([field(x3, e) for e in range(s3.nEles)], axis=1)
    import tempfile, pathlib
    with tempfile.TemporaryDirectory() as d:
        f = write_vtu(pathlib.Path(d) / "r.vtu", s2, U2)
        got, _ = read_restart(f, s3)
    np.testing.assert_allclose(got, U3, rtol=0, atol=1e-12)


def test_csv(tmp_path):
    s, U = _perturbed(lambda: cases.bump2d(nx=3, ny=2, order=1))
    f = write_csv(tmp_path / "x.csv", s, U)
    rows = f.read_text().splitlines()
    assert rows[0] == "x,y,z,rho,u,v,p"
    assert len(rows) == 1 + s.nEles * (s.nSpts + s.nFpts)
    first = [float(v) for v in rows[1].split(",")]
    np.testing.assert_allclose(first[:2], s.geo.pos_spts[0, 0], atol=1e-14)
    assert first[3] == pytest.approx(U[0, 0, 0], abs=1e-14)
