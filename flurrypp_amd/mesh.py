"""Meshes and their FR set-up, through libflurry_host (include/flurry_host.h).

``Mesh`` owns a host mesh (vertices, linear quads/hexes, boundaries) and
produces FlurryPP's face list and metric terms; ``operators`` returns the FR
operator matrices of one element type and order.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._native import dptr, host_check, host_lib, iptr

PTS = {"legendre": 0, "lobatto": 1, "equidistant": 2}


def bc_number(name: str) -> int:
    v = host_lib().fhost_bc_from_name(name.encode())
    if v < -1:
        raise ValueError(f"Unrecognized boundary condition: {name!r}")
    return v


def sizes(nDims: int, order: int) -> tuple[int, int, int]:
    s, f, ff = C.c_int(), C.c_int(), C.c_int()
    host_check(host_lib().fhost_sizes(nDims, order, C.byref(s), C.byref(f), C.byref(ff)))
    return s.value, f.value, ff.value


@dataclass
class Operators:
    nDims: int
    order: int
    spts_to_fpts: np.ndarray
    grad_spts: np.ndarray
    extrap_fn: np.ndarray
    correction: np.ndarray
    correct_u: np.ndarray
    w_spts: np.ndarray
    w_face: np.ndarray


def operators(nDims: int, order: int, spts_type: str = "Legendre", vcjh: int = 0) -> Operators:
    S, F, Ff = sizes(nDims, order)
    D = nDims
    a = dict(spts_to_fpts=np.zeros((F, S)), grad_spts=np.zeros((D, S, S)), extrap_fn=np.zeros((D, F, S)),
             correction=np.zeros((S, F)), correct_u=np.zeros((D, S, F)), w_spts=np.zeros(S), w_face=np.zeros(Ff))
    host_check(host_lib().fhost_operators(nDims, order, PTS[spts_type.lower()], vcjh, *(dptr(a[k]) for k in
               ("spts_to_fpts", "grad_spts", "extrap_fn", "correction", "correct_u", "w_spts", "w_face"))))
    return Operators(nDims, order, **a)


def pmg_operators(nDims: int, order: int, spts_type: str = "Legendre"):
    """P-multigrid transfer operators of one order (oper::setupPMG,
    src/operators.cpp:1553-1634): prolongation [(P+2)^D][S] to order P+1 and
    restriction [P^D][S] to order P-1 (None at order 0)."""
    S = (order + 1) ** nDims
    pro = np.zeros(((order + 2) ** nDims, S))
    res = np.zeros((order ** nDims, S)) if order > 0 else None
    host_check(host_lib().fhost_pmg_operators(nDims, order, PTS[spts_type.lower()], dptr(pro),
                                              dptr(res) if res is not None else None))
    return pro, res


def pts1d(order: int, spts_type: str = "Legendre") -> np.ndarray:
    out = np.zeros(order + 1)
    host_check(host_lib().fhost_pts1d(PTS[spts_type.lower()], order, dptr(out)))
    return out


def ref_points(nDims: int, order: int, spts_type: str = "Legendre"):
    S, F, _ = sizes(nDims, order)
    ls, lf, tn = np.zeros((S, nDims)), np.zeros((F, nDims)), np.zeros((F, nDims))
    host_check(host_lib().fhost_ref_points(nDims, order, PTS[spts_type.lower()], dptr(ls), dptr(lf), dptr(tn)))
    return ls, lf, tn


@dataclass
class Faces:
    eleL: np.ndarray
    locL: np.ndarray
    eleR: np.ndarray
    locR: np.ndarray
    relRot: np.ndarray
    bc: np.ndarray
    gid: np.ndarray
    periodic: np.ndarray
    nInt: int
    nBnd: int


@dataclass
class Geometry:
    pos_spts: np.ndarray
    pos_fpts: np.ndarray
    jginv_spts: np.ndarray
    detjac_spts: np.ndarray
    norm_fpts: np.ndarray
    da_fpts: np.ndarray


class Mesh:
    def __init__(self, handle: C.c_void_p):
        self._h = handle
        self.processed = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            host_lib().fhost_mesh_destroy(h)
            self._h = None

    # -- constructors --------------------------------------------------------
    @classmethod
    def cartesian(cls, nDims: int, n: tuple, lo: tuple, hi: tuple, bcs: dict | None = None) -> "Mesh":
        """createMesh (src/geo.cpp:1477-1762).  bcs: side -> BC name for
        bottom/right/top/left (+front/back); default periodic."""
        sides = ["bottom", "right", "top", "left"] + (["front", "back"] if nDims == 3 else [])
        bcs = {k.lower(): v for k, v in (bcs or {}).items()}
        names = (C.c_char_p * len(sides))(*[bcs.get(s, "periodic").encode() for s in sides])
        nx, ny = n[0], n[1]
        nz = n[2] if nDims == 3 else 1
        lo_a = np.array(lo[:nDims], dtype=np.float64)
        hi_a = np.array(hi[:nDims], dtype=np.float64)
        h = C.c_void_p()
        host_check(host_lib().fhost_mesh_cartesian(nDims, nx, ny, nz, dptr(lo_a), dptr(hi_a), names, C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, xv: np.ndarray, c2v: np.ndarray, bounds: list[tuple[int, np.ndarray]],
                    shape_nodes: np.ndarray | None = None) -> "Mesh":
        """Elements by vertex (c2v, FlurryPP corner order); shape_nodes, if
        given, are the geometry nodes of curved elements [nEles][nShape] in
        Gmsh's order (the first 4 or 8 being the vertices)."""
        xv = np.ascontiguousarray(xv, dtype=np.float64)
        c2v = np.ascontiguousarray(c2v, dtype=np.int32)
        nDims = xv.shape[1]
        bc_list = np.array([b for b, _ in bounds], dtype=np.int32)
        ptr = np.zeros(len(bounds) + 1, dtype=np.int32)
        pts = []
        for i, (_, p) in enumerate(bounds):
            pts.append(np.asarray(p, dtype=np.int32))
            ptr[i + 1] = ptr[i] + len(p)
        pts_a = np.ascontiguousarray(np.concatenate(pts) if pts else np.zeros(1, np.int32), dtype=np.int32)
        h = C.c_void_p()
        host_check(host_lib().fhost_mesh_from_arrays(nDims, xv.shape[0], dptr(xv), c2v.shape[0], iptr(c2v),
                                                      len(bounds), iptr(bc_list), iptr(ptr), iptr(pts_a), C.byref(h)))
        m = cls(h)
        if shape_nodes is not None:
            c2n = np.ascontiguousarray(shape_nodes, dtype=np.int32)
            host_check(host_lib().fhost_mesh_set_shape_nodes(h, c2n.shape[1], iptr(c2n)))
        return m

    @classmethod
    def read_gmsh(cls, path: str, mesh_bounds: dict[str, str]) -> "Mesh":
        names = (C.c_char_p * len(mesh_bounds))(*[k.encode() for k in mesh_bounds])
        bcs = (C.c_char_p * len(mesh_bounds))(*[v.encode() for v in mesh_bounds.values()])
        h = C.c_void_p()
        host_check(host_lib().fhost_mesh_read_gmsh(str(path).encode(), len(mesh_bounds), names, bcs, C.byref(h)))
        return cls(h)

    # -- edits ----------------------------------------------------------------
    def jitter(self, h: tuple, amp: float, seed: int) -> "Mesh":
        ha = np.array(h, dtype=np.float64)
        host_check(host_lib().fhost_mesh_jitter(self._h, dptr(ha), amp, seed))
        return self

    def set_periodic(self, dx: float, dy: float, dz: float = np.inf, tol: float = 1e-6) -> "Mesh":
        host_check(host_lib().fhost_mesh_set_periodic(self._h, dx, dy, dz, tol))
        return self

    def process(self) -> "Mesh":
        host_check(host_lib().fhost_mesh_process(self._h))
        self.processed = True
        return self

    # -- queries ---------------------------------------------------------------
    def counts(self) -> dict:
        v = [C.c_int() for _ in range(5)]
        host_lib().fhost_mesh_counts(self._h, *(C.byref(x) for x in v))
        return dict(zip(("nDims", "nVerts", "nEles", "nIntFaces", "nBndFaces"), (x.value for x in v)))

    @property
    def nDims(self) -> int:
        return self.counts()["nDims"]

    @property
    def nEles(self) -> int:
        return self.counts()["nEles"]

    def nodes(self) -> tuple[np.ndarray, np.ndarray]:
        c = self.counts()
        xv = np.zeros((c["nVerts"], c["nDims"]))
        c2v = np.zeros((c["nEles"], 4 if c["nDims"] == 2 else 8), dtype=np.int32)
        host_lib().fhost_mesh_get_nodes(self._h, dptr(xv), iptr(c2v))
        return xv, c2v

    def shape_nodes(self) -> np.ndarray:
        """Geometry nodes of every element [nEles][nShape] (Gmsh order; the
        vertices for linear elements)."""
        ns = C.c_int()
        host_check(host_lib().fhost_mesh_get_shape_nodes(self._h, C.byref(ns), None))
        out = np.zeros((self.nEles, ns.value), dtype=np.int32)
        host_check(host_lib().fhost_mesh_get_shape_nodes(self._h, None, iptr(out)))
        return out

    def faces(self) -> Faces:
        if not self.processed:
            self.process()
        c = self.counts()
        n = c["nIntFaces"] + c["nBndFaces"]
        arr = [np.zeros(n, dtype=np.int32) for _ in range(8)]
        host_check(host_lib().fhost_mesh_get_faces(self._h, *(iptr(a) for a in arr)))
        return Faces(*arr, nInt=c["nIntFaces"], nBnd=c["nBndFaces"])

    def geometry(self, order: int, spts_type: str = "Legendre") -> Geometry:
        c = self.counts()
        D, E = c["nDims"], c["nEles"]
        S, F, _ = sizes(D, order)
        g = Geometry(pos_spts=np.zeros((S, E, D)), pos_fpts=np.zeros((F, E, D)), jginv_spts=np.zeros((D, S, E, D)),
                     detjac_spts=np.zeros((S, E)), norm_fpts=np.zeros((F, E, D)), da_fpts=np.zeros((F, E)))
        host_check(host_lib().fhost_mesh_geometry(self._h, order, PTS[spts_type.lower()], dptr(g.pos_spts),
                                                  dptr(g.pos_fpts), dptr(g.jginv_spts), dptr(g.detjac_spts),
                                                  dptr(g.norm_fpts), dptr(g.da_fpts)))
        return g
