"""Literal restatement of FlurryPP's face construction, for small meshes.

TEST INFRASTRUCTURE ONLY (the checker of libflurry_host's O(n) set-up).

It follows the reference step by step, including its quadratic searches:
  processConn2D / processConn3D   src/geo.cpp:221-570 (matrix::unique, src/matrix.cpp:668-695)
  processPeriodicBoundaries       src/geo.cpp:1764-1852 (checkPeriodicFaces/3D, :1870-1960)
  setupElesFaces (face order)     src/geo.cpp:854-930
  compareOrientation              src/geo.cpp:2078-2225
and returns the per-face records in solver::faces order.
"""
from __future__ import annotations

import math

PERIODIC = 0
HEX_FACE_VERTS = [[0, 1, 2, 3], [4, 5, 6, 7], [3, 0, 4, 7], [2, 1, 5, 6], [1, 0, 4, 5], [3, 2, 6, 7]]
HEX_ORIENT = [[0, 1, 2, 3], [5, 4, 7, 6], [0, 3, 7, 4], [2, 1, 5, 6], [1, 0, 4, 5], [3, 2, 6, 7]]


def _unique(rows):
    """matrix::unique: first-occurrence order; iRow maps rows to unique ids."""
    out, irow = [], [-1] * len(rows)
    for i, r in enumerate(rows):
        for j in range(i):
            if rows[j] == r:
                irow[i] = irow[j]
                break
        if irow[i] == -1:
            out.append(list(r))
            irow[i] = len(out) - 1
    return out, irow


def _periodic2d(xv, e1, e2, dx, dy, tol):
    x11, y11 = xv[e1[0]][0], xv[e1[0]][1]
    x12, y12 = xv[e1[1]][0], xv[e1[1]][1]
    x21, y21 = xv[e2[0]][0], xv[e2[0]][1]
    x22, y22 = xv[e2[1]][0], xv[e2[1]][1]
    a = abs
    if a(a(x21 - x11) - dx) < tol and a(a(x22 - x12) - dx) < tol and a(y21 - y11) < tol and a(y22 - y12) < tol:
        return True
    if a(a(x22 - x11) - dx) < tol and a(a(x21 - x12) - dx) < tol and a(y22 - y11) < tol and a(y21 - y12) < tol:
        return True
    if a(a(y21 - y11) - dy) < tol and a(a(y22 - y12) - dy) < tol and a(x21 - x11) < tol and a(x22 - x12) < tol:
        return True
    if a(a(y22 - y11) - dy) < tol and a(a(y21 - y12) - dy) < tol and a(x22 - x11) < tol and a(x21 - x12) < tol:
        return True
    return False


def _nc(xv, f):
    v1 = [xv[f[1]][d] - xv[f[0]][d] for d in range(3)]
    v2 = [xv[f[2]][d] - xv[f[0]][d] for d in range(3)]
    c = [0.0, 0.0, 0.0]
    for j in range(len(f)):
        for d in range(3):
            c[d] += xv[f[j]][d]
    c = [x / len(f) for x in c]
    n = [v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2], v1[0] * v2[1] - v1[1] * v2[0]]
    m = math.sqrt(sum(x * x for x in n))
    return [x / m for x in n], c


def _mul(a, b):
    try:
        return a * b
    except Exception:
        return float("nan")


def _periodic3d(xv, f1, f2, dx, dy, dz, tol):
    n1, c1 = _nc(xv, f1)
    n2, c2 = _nc(xv, f2)
    dot = sum(n1[i] * n2[i] for i in range(3))
    if abs(1 - abs(dot)) > tol:
        return False
    nd = [abs(_mul(n1[0], dx)), abs(_mul(n1[1], dy)), abs(_mul(n1[2], dz))]
    for i in range(3):
        diff = abs(abs(c2[i] - c1[i]) - nd[i])
        if diff > tol:  # NaN passes
            return False
    return True


def build_faces(D, xv, c2v, bc_list, bnd_pts, dx, dy, dz, tol):
    """Returns a list of (eleL, locL, eleR, locR, relRot, bc, gid) in solver::faces order."""
    E = len(c2v)
    nf = 4 if D == 2 else 6
    rows = []
    for e in range(E):
        for f in range(nf):
            if D == 2:
                a, b = c2v[e][f], c2v[e][(f + 1) % 4]
                rows.append((min(a, b), max(a, b)))
            else:
                rows.append(tuple(sorted(c2v[e][i] for i in HEX_FACE_VERTS[f])))
    f2v, iF = _unique(rows)
    nFaces = len(f2v)
    int_faces, bnd_faces = [], []
    seen = list(iF)
    for i in range(len(seen)):
        if seen[i] != -1:
            ff = [j for j in range(len(seen)) if seen[j] == seen[i]]
            if len(ff) == 2:
                int_faces.append(seen[i])
            elif len(ff) == 1:
                bnd_faces.append(seen[i])
            for j in ff:
                seen[j] = -1
    bc_type = [-1] * len(bnd_faces)
    for i, fb in enumerate(bnd_faces):
        for b in range(len(bc_list)):
            if all(v in bnd_pts[b] for v in f2v[fb]):
                bc_type[i] = bc_list[b]
                break
    c2f = [[-1] * nf for _ in range(E)]
    f2c = [[-1, -1] for _ in range(nFaces)]
    for ic in range(E):
        for j in range(nf):
            ff = iF[ic * nf + j]
            c2f[ic][j] = ff
            if f2c[ff][0] == -1:
                f2c[ff][0] = ic
            else:
                f2c[ff][1] = ic
    iper = [i for i in range(len(bnd_faces)) if bc_type[i] == PERIODIC]
    for i in iper:
        if bnd_faces[i] == -10:
            continue
        for j in iper:
            if i == j or bnd_faces[i] == -10 or bnd_faces[j] == -10:
                continue
            if D == 2:
                m = _periodic2d(xv, f2v[bnd_faces[i]], f2v[bnd_faces[j]], dx, dy, tol)
            else:
                m = _periodic3d(xv, f2v[bnd_faces[i]], f2v[bnd_faces[j]], dx, dy, dz, tol)
            if m:
                assert i < j
                bi, bj = bnd_faces[i], bnd_faces[j]
                int_faces.append(bi)
                f2c[bi][1] = f2c[bj][0]
                f2c[bj][0] = -1
                ic = f2c[bi][1]
                fid = c2f[ic].index(bj)
                c2f[ic][fid] = bi
                bnd_faces[i] = bnd_faces[j] = -10
                bc_type[i] = bc_type[j] = -10
                break
    keep = [k for k in range(len(bnd_faces)) if bnd_faces[k] != -10]
    bnd_faces = [bnd_faces[k] for k in keep]
    bc_type = [bc_type[k] for k in keep]

    def orient(ic1, f1, ic2, f2):
        if D == 2:
            return 0
        t1 = [c2v[ic1][k] for k in HEX_ORIENT[f1]]
        t2 = [c2v[ic2][k] for k in HEX_ORIENT[f2]]
        for i in range(4):
            if t1[i] == t2[0]:
                return i
        if _periodic3d(xv, t1, t2, dx, dy, dz, tol):
            c1 = [sum(xv[v][d] for v in t1) / 4 for d in range(3)]
            c2 = [sum(xv[v][d] for v in t2) / 4 for d in range(3)]
            fd = [c2[d] - c1[d] for d in range(3)]
            m = math.sqrt(sum(x * x for x in fd))
            fd = [x / m for x in fd]
            for i in range(4):
                pd = [xv[t2[0]][d] - xv[t1[i]][d] for d in range(3)]
                pm = math.sqrt(sum(x * x for x in pd))
                dot = sum(fd[d] * pd[d] / pm for d in range(3))
                if abs(1 - abs(dot)) < tol:
                    return i
            raise RuntimeError("Unable to orient periodic faces")
        raise RuntimeError("Internal faces improperly matched")

    faces = []
    for ff in int_faces:
        ic1, ic2 = f2c[ff]
        fid1, fid2 = c2f[ic1].index(ff), c2f[ic2].index(ff)
        faces.append((ic1, fid1, ic2, fid2, orient(ic1, fid1, ic2, fid2), -1, ff))
    for i, ff in enumerate(bnd_faces):
        ic = f2c[ff][0]
        faces.append((ic, c2f[ic].index(ff), -1, -1, 0, bc_type[i], ff))
    return faces


def cartesian(D, n, lo, hi, bcs):
    """createMesh vertex/element/boundary numbering (src/geo.cpp:1477-1762);
    bcs: list of BC numbers for bottom, right, top, left[, front, back]."""
    nx, ny = n[0], n[1]
    nz = n[2] if D == 3 else 1
    dx = (hi[0] - lo[0]) / nx
    dy = (hi[1] - lo[1]) / ny
    xv, c2v = [], []
    if D == 2:
        for i in range(ny + 1):
            for j in range(nx + 1):
                xv.append([lo[0] + j * dx, lo[1] + i * dy])
        for i in range(nx):
            for j in range(ny):
                c2v.append([j * (nx + 1) + i, j * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i])
    else:
        dz = (hi[2] - lo[2]) / nz
        for k in range(nz + 1):
            for j in range(ny + 1):
                for i in range(nx + 1):
                    xv.append([lo[0] + i * dx, lo[1] + j * dy, lo[2] + k * dz])
        X = nx + 1
        Y = ny + 1
        for k in range(nz):
            for i in range(nx):
                for j in range(ny):
                    c2v.append([i + X * (j + Y * k), i + X * (j + Y * k) + 1, i + X * (j + 1 + Y * k) + 1,
                                i + X * (j + 1 + Y * k), i + X * (j + Y * (k + 1)), i + X * (j + Y * (k + 1)) + 1,
                                i + X * (j + 1 + Y * (k + 1)) + 1, i + X * (j + 1 + Y * (k + 1))])
    bc_list = sorted(set(bcs))
    pts = {b: {0} for b in bc_list}
    if D == 2:
        for ix in range(nx):
            pts[bcs[0]].update([ix, ix + 1])
            pts[bcs[2]].update([(nx + 1) * ny + ix + 1, (nx + 1) * ny + ix])
        for iy in range(ny):
            pts[bcs[3]].update([(iy + 1) * (nx + 1), iy * (nx + 1)])
            pts[bcs[1]].update([iy * (nx + 1) + nx, (iy + 1) * (nx + 1) + nx])
    else:
        X, XY = nx + 1, (nx + 1) * (ny + 1)
        for ix in range(nx):
            for iy in range(ny):
                pts[bcs[0]].update([iy * X + ix, iy * X + ix + 1, (iy + 1) * X + ix + 1, (iy + 1) * X + ix])
                pts[bcs[2]].update([XY * nz + iy * X + ix, XY * nz + iy * X + ix + 1, XY * nz + (iy + 1) * X + ix + 1,
                                    XY * nz + (iy + 1) * X + ix])
        for iz in range(nz):
            for iy in range(ny):
                pts[bcs[3]].update([iz * XY + iy * X, (iz + 1) * XY + iy * X, (iz + 1) * XY + (iy + 1) * X,
                                    iz * XY + (iy + 1) * X])
                pts[bcs[1]].update([iz * XY + iy * X + nx, (iz + 1) * XY + iy * X + nx,
                                    (iz + 1) * XY + (iy + 1) * X + nx, iz * XY + (iy + 1) * X + nx])
            for ix in range(nx):
                pts[bcs[5]].update([iz * XY + ix, (iz + 1) * XY + ix, (iz + 1) * XY + ix + 1, iz * XY + ix + 1])
                pts[bcs[4]].update([iz * XY + ny * X + ix, (iz + 1) * XY + ny * X + ix,
                                    (iz + 1) * XY + ny * X + ix + 1, iz * XY + ny * X + ix + 1])
    return xv, c2v, bc_list, [pts[b] for b in bc_list]
