"""FlurryPP's plot and restart files.

* ParaView `.vtu` (`writeParaview`, src/output.cpp:137-513, serial branch): one
  `<Piece>` per element at the plot points -- the solution points bordered by
  the element's edges/faces, (P+3)^D points forming (P+2)^D linear sub-cells
  (`getLocPpts`, src/points.cpp:164-196) -- holding Density, Velocity and
  Pressure extrapolated from the solution points with `opp_spts_to_ppts`
  (src/operators.cpp:143-169, src/solver.cpp:630-670), and the simulation time
  and iteration in XML comments.
* Restart (`solver::readRestartFile` / `ele::restart`, src/solver.cpp:1789-1890,
  src/ele.cpp:1497-1920): the same file read back; the interior plot points are
  the solution points, their primitives become conservative variables, and a
  file written at another order is interpolated to this one.
* CSV (`writeCSV`, src/output.cpp:60-135): x, y, z, then rho, u, v, p at the
  solution points and the flux points of every element.

Euler/Navier-Stokes fields only (the reference's advection-diffusion branch
is not written).
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np

from .mesh import pts1d


def _lag(x: np.ndarray, y: float, mode: int) -> float:
    """Lagrange basis function `mode` through nodes x at y (src/polynomials.cpp:39-50)."""
    v = 1.0
    for i in range(len(x)):
        if i != mode:
            v = v * ((y - x[i]) / (x[mode] - x[i]))
    return v


def ppts_1d(order: int, spts_type: str = "Legendre") -> np.ndarray:
    return np.concatenate([[-1.0], pts1d(order, spts_type), [1.0]])


def _grid_index(D: int, n1: int, p: int):
    return (p % n1, (p // n1) % n1, p // (n1 * n1)) if D == 3 else (p % n1, p // n1, 0)


def interp_matrix(D: int, order_from: int, at_1d: np.ndarray, spts_type: str = "Legendre") -> np.ndarray:
    """Lagrange interpolation from the solution points of `order_from` to the
    tensor grid of `at_1d`, factors multiplied x, y, z in that order."""
    s1 = pts1d(order_from, spts_type)
    P1, n1 = order_from + 1, len(at_1d)
    S, T = P1 ** D, n1 ** D
    L = np.array([[_lag(s1, y, i) for i in range(P1)] for y in at_1d])
    out = np.zeros((T, S))
    for t in range(T):
        it, jt, kt = _grid_index(D, n1, t)
        for s in range(S):
            i, j, k = _grid_index(D, P1, s)
            v = L[it, i] * L[jt, j]
            if D == 3:
                v = v * L[kt, k]
            out[t, s] = v
    return out


def primitives(U: np.ndarray, D: int, gamma: float) -> np.ndarray:
    """rho, u, v, [w], p per point (src/solver.cpp:642-657); U [..., N]."""
    rho = U[..., 0]
    u = U[..., 1] / rho
    v = U[..., 2] / rho
    V = np.empty_like(U)
    V[..., 0], V[..., 1], V[..., 2] = rho, u, v
    vsq = u * u + v * v
    if D == 3:
        w = U[..., 3] / rho
        vsq = vsq + w * w
        V[..., 3] = w
    V[..., D + 1] = (gamma - 1) * (U[..., D + 1] - 0.5 * rho * vsq)
    return V


def conservative(V: np.ndarray, D: int, gamma: float) -> np.ndarray:
    """Inverse of `primitives` as ele::restart writes it (src/ele.cpp:1812-1822)."""
    U = np.empty_like(V)
    U[..., 0] = V[..., 0]
    U[..., 1] = V[..., 0] * V[..., 1]
    U[..., 2] = V[..., 0] * V[..., 2]
    vsq = V[..., 1] * V[..., 1] + V[..., 2] * V[..., 2]
    if D == 3:
        U[..., 3] = V[..., 0] * V[..., 3]
        vsq = vsq + V[..., 3] * V[..., 3]
    U[..., D + 1] = V[..., D + 1] / (gamma - 1) + 0.5 * V[..., 0] * vsq
    return U


def _shape_linear(D: int, x: float, y: float, z: float) -> np.ndarray:
    """Bilinear / trilinear shape functions in FlurryPP's corner order."""
    c = np.array([-1.0, 1.0])
    lx = [_lag(c, x, 0), _lag(c, x, 1)]
    ly = [_lag(c, y, 0), _lag(c, y, 1)]
    q = [lx[0] * ly[0], lx[1] * ly[0], lx[1] * ly[1], lx[0] * ly[1]]
    if D == 2:
        return np.array(q)
    lz = [_lag(c, z, 0), _lag(c, z, 1)]
    return np.array([v * lz[0] for v in q] + [v * lz[1] for v in q])


def plot_positions(setup) -> np.ndarray:
    """Physical coordinates of every element's plot points, [nPpts][nEles][D]."""
    D, P = setup.nDims, setup.order
    xv, c2v = setup.mesh.nodes()
    t1 = ppts_1d(P, setup.params.sptsTypeQuad)
    n1 = len(t1)
    sh = np.array([_shape_linear(D, *[t1[a] for a in _grid_index(D, n1, p)][:D], *([0.0] * (3 - D)))
                   for p in range(n1 ** D)])
    X = xv[c2v[:, : 2 ** D]]                 # [E][nodes][D]
    return np.einsum("pn,end->ped", sh, X)


def _fmt(a) -> str:
    return " ".join(format(float(x), ".16g") for x in np.ravel(a))


def write_vtu(path, setup, U: np.ndarray, time: float = 0.0, iteration: int = 0) -> Path:
    """FlurryPP's serial ParaView file of state U [nSpts][nEles][nFields]."""
    p = setup.params
    D, P, E = setup.nDims, setup.order, setup.nEles
    if p.equation != 1:
        raise ValueError("only the Euler / Navier-Stokes fields are written")
    n1 = P + 3
    nP, nC = n1 ** D, (P + 2) ** D
    A = interp_matrix(D, P, ppts_1d(P, p.sptsTypeQuad), p.sptsTypeQuad)
    Vs = primitives(np.asarray(U, dtype=np.float64), D, p.gamma)
    Vp = np.einsum("ps,sek->pek", A, Vs)
    X = plot_positions(setup)
    if D == 2:
        conn = [[j * n1 + i, j * n1 + i + 1, (j + 1) * n1 + i + 1, (j + 1) * n1 + i]
                for j in range(n1 - 1) for i in range(n1 - 1)]
    else:
        def g(i, j, k):
            return i + n1 * (j + n1 * k)
        conn = [[g(i, j, k), g(i + 1, j, k), g(i + 1, j + 1, k), g(i, j + 1, k),
                 g(i, j, k + 1), g(i + 1, j, k + 1), g(i + 1, j + 1, k + 1), g(i, j + 1, k + 1)]
                for k in range(n1 - 1) for j in range(n1 - 1) for i in range(n1 - 1)]
    conn_txt = "".join(" ".join(str(v) for v in c) + " \n" for c in conn)
    nv = 4 if D == 2 else 8
    offsets = " ".join(str((k + 1) * nv) for k in range(nC)) + " "
    types = " ".join(str(9 if D == 2 else 12) for _ in range(nC)) + " "
    out = ['<?xml version="1.0" ?>',
           '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" '
           'compressor="vtkZLibDataCompressor">',
           f"<!-- TIME {format(float(time), '.16g')} -->", f"<!-- ITER {int(iteration)} -->", "\t<UnstructuredGrid>"]
    for e in range(E):
        v = Vp[:, e, :]
        vel = np.zeros((nP, 3))
        vel[:, :D] = v[:, 1:D + 1]
        pts = np.zeros((nP, 3))
        pts[:, :D] = X[:, e, :]
        out += [f'\t\t<Piece NumberOfPoints="{nP}" NumberOfCells="{nC}">', "\t\t\t<PointData>",
                '\t\t\t\t<DataArray type="Float32" Name="Density" format="ascii">', _fmt(v[:, 0]) + " ",
                "\t\t\t\t</DataArray>",
                '\t\t\t\t<DataArray type="Float32" NumberOfComponents="3" Name="Velocity" format="ascii">',
                _fmt(vel) + " ", "\t\t\t\t</DataArray>",
                '\t\t\t\t<DataArray type="Float32" Name="Pressure" format="ascii">', _fmt(v[:, D + 1]) + " ",
                "\t\t\t\t</DataArray>", "\t\t\t</PointData>", "\t\t\t<Points>",
                '\t\t\t\t<DataArray type="Float32" NumberOfComponents="3" format="ascii">', _fmt(pts) + " ",
                "\t\t\t\t</DataArray>", "\t\t\t</Points>", "\t\t\t<Cells>",
                '\t\t\t\t<DataArray type="Int32" Name="connectivity" format="ascii">', conn_txt.rstrip("\n"),
                "\t\t\t\t</DataArray>", '\t\t\t\t<DataArray type="Int32" Name="offsets" format="ascii">', offsets,
                "\t\t\t\t</DataArray>", '\t\t\t\t<DataArray type="UInt8" Name="types" format="ascii">', types,
                "\t\t\t\t</DataArray>", "\t\t\t</Cells>", "\t\t</Piece>"]
    out += ["\t</UnstructuredGrid>", "</VTKFile>"]
    path = Path(path)
    path.write_text("\n".join(out) + "\n")
    return path


def vtu_name(data_file_name: str, iteration: int) -> str:
    return f"{data_file_name}_{iteration:09d}.vtu"


def read_restart(path, setup) -> tuple[np.ndarray, float | None]:
    """State at the solution points of `setup` (and the time) from a plot file."""
    p = setup.params
    D, order_now = setup.nDims, setup.order
    text = Path(path).read_text().splitlines()
    time = None
    li = 0
    while li < len(text):
        w = text[li].split()
        if w and w[0] == "<!--" and len(w) > 2 and w[1] == "TIME":
            time = float(w[2])
        if w and w[0] == "<UnstructuredGrid>":
            break
        li += 1
    else:
        raise ValueError("Cannot find UnstructuredData tag in restart file.")
    li += 1
    U = np.zeros((setup.nSpts, setup.nEles, setup.nFields))
    interp = {}
    for e in range(setup.nEles):
        m = re.search(r'NumberOfPoints="(\d+)"\s+NumberOfCells="(\d+)"', text[li])
        if not m:
            raise ValueError(f"restart file: no piece for element {e}")
        n_cells = int(m.group(2))
        order = round(n_cells ** (1.0 / D)) - 2
        P1, n1 = order + 1, order + 3
        fields = {}
        li += 1
        while "</Piece>" not in text[li]:
            m2 = re.search(r'<DataArray[^>]*Name="(\w+)"', text[li])
            if m2 and m2.group(1) in ("Density", "Velocity", "Pressure"):
                vals = np.array(text[li + 1].split(), dtype=np.float64)
                fields[m2.group(1)] = vals.reshape(n1 ** D, -1)
                li += 2
                continue
            li += 1
        li += 1
        # interior plot points are the solution points (x fastest)
        idx = []
        for s in range(P1 ** D):
            i, j, k = _grid_index(D, P1, s)
            idx.append((i + 1) + n1 * ((j + 1) + (n1 * (k + 1) if D == 3 else 0)))
        idx = np.array(idx)
        V = np.zeros((P1 ** D, setup.nFields))
        V[:, 0] = fields["Density"][idx, 0]
        V[:, 1:D + 1] = fields["Velocity"][idx, :D]
        V[:, D + 1] = fields["Pressure"][idx, 0]
        Ue = conservative(V, D, p.gamma)
        if order != order_now:
            if order not in interp:
                interp[order] = interp_matrix(D, order, pts1d(order_now, p.sptsTypeQuad), p.sptsTypeQuad)
            Ue = interp[order] @ Ue
        U[:, e, :] = Ue
    return U, time


def write_csv(path, setup, U: np.ndarray) -> Path:
    """x,y,z,rho,u,v,p at every element's solution points then flux points."""
    p = setup.params
    D = setup.nDims
    Vs = primitives(np.asarray(U, dtype=np.float64), D, p.gamma)
    Uf = np.einsum("fs,sek->fek", setup.opers.spts_to_fpts, U)
    Vf = primitives(Uf, D, p.gamma)
    ps, pf = setup.geo.pos_spts, setup.geo.pos_fpts
    lines = ["x,y,z,rho,u,v,p"]  # the reference's header, also in 3-D
    for e in range(setup.nEles):
        for V, X in ((Vs, ps), (Vf, pf)):
            for q in range(V.shape[0]):
                xyz = [format(X[q, e, d], ".15f") for d in range(D)] + (["0.0"] if D == 2 else [])
                lines.append(",".join(xyz + [format(V[q, e, k], ".15f") for k in range(setup.nFields)]))
    path = Path(path)
    path.write_text("\n".join(lines) + "\n")
    return path
