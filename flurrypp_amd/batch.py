"""Batched solves: many independent meshes advanced together.

Meshes of the same order (and options) are merged into one libflurry_hip
context: elements and faces are concatenated, every mesh keeps its own
boundary (free-stream) state through the boundary-state table, its own dt
(per element for local CFL, per mesh for global CFL) and its own residual.
Groups of different orders run as separate contexts whose RK steps are
interleaved by fhip_update_many, each on its own stream, so their kernels
overlap on the device.  Converged meshes are frozen with per-mesh activity
flags.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as nat
from .mesh import Faces, Geometry, Operators
from .solver import HostSetup


@dataclass
class Member:
    setup: HostSetup
    U: np.ndarray
    tag: object = None


class MergedSetup:
    """HostSetup-compatible union of several meshes of one order."""

    def __init__(self, members: list[Member]):
        s0 = members[0].setup
        for m in members:
            s = m.setup
            if (s.nDims, s.order, s.nFields) != (s0.nDims, s0.order, s0.nFields):
                raise ValueError("a merged batch needs one element type and order")
            for k in ("equation", "viscous", "riemannType", "timeType", "dtType", "gamma"):
                if getattr(s.params, k) != getattr(s0.params, k):
                    raise ValueError(f"batch members differ in {k}")
        self.params = s0.params
        self.nDims, self.nFields, self.order = s0.nDims, s0.nFields, s0.order
        self.nSpts, self.nFpts, self.nFptsFace = s0.nSpts, s0.nFpts, s0.nFptsFace
        self.opers: Operators = s0.opers
        self.members = members
        self.ele_start = np.zeros(len(members) + 1, dtype=np.int32)
        for i, m in enumerate(members):
            self.ele_start[i + 1] = self.ele_start[i] + m.setup.nEles
        self.nEles = int(self.ele_start[-1])
        cat1 = lambda name: np.ascontiguousarray(np.concatenate([getattr(m.setup.geo, name) for m in members], axis=1))  # noqa: E731
        self.geo = Geometry(pos_spts=cat1("pos_spts"), pos_fpts=cat1("pos_fpts"), jginv_spts=cat1("jginv_spts"),
                            detjac_spts=cat1("detjac_spts"), norm_fpts=cat1("norm_fpts"), da_fpts=cat1("da_fpts"))
        eL, lL, eR, lR, rr, bc, gid, per, bst = [], [], [], [], [], [], [], [], []
        for i, m in enumerate(members):
            f = m.setup.faces
            off = int(self.ele_start[i])
            eL.append(f.eleL + off)
            eR.append(np.where(f.eleR >= 0, f.eleR + off, -1))
            lL.append(f.locL)
            lR.append(f.locR)
            rr.append(f.relRot)
            bc.append(f.bc)
            gid.append(f.gid)
            per.append(f.periodic)
            bst.append(np.where(f.eleR >= 0, -1, i))
        c32 = lambda xs: np.ascontiguousarray(np.concatenate(xs), dtype=np.int32)  # noqa: E731
        self.faces = Faces(c32(eL), c32(lL), c32(eR), c32(lR), c32(rr), c32(bc), c32(gid), c32(per),
                           nInt=sum(m.setup.faces.nInt for m in members),
                           nBnd=sum(m.setup.faces.nBnd for m in members))
        self.face_bnd_state = c32(bst)
        self.bnd_states = np.ascontiguousarray(
            [[m.setup.params.rhoBound, m.setup.params.uBound, m.setup.params.vBound,
              getattr(m.setup.params, "wBound", 0.0), m.setup.params.pBound] for m in members], dtype=np.float64)
        fp = nat.FhipParams.from_buffer_copy(s0.fp)
        fp.n_eles = self.nEles
        self.fp = fp
        self.U = np.ascontiguousarray(np.concatenate([m.U for m in members], axis=1))


class BatchGroup:
    """One libflurry_hip context holding several meshes of one order."""

    def __init__(self, merged: MergedSetup):
        self.m = merged
        o, g, f = merged.opers, merged.geo, merged.faces
        self._keep = [o.spts_to_fpts, o.grad_spts, o.extrap_fn, o.correction, o.correct_u, o.w_spts, o.w_face,
                      g.jginv_spts, g.detjac_spts, g.norm_fpts, g.da_fpts,
                      f.eleL, f.locL, f.eleR, f.locR, f.relRot, f.bc, merged.ele_start, merged.bnd_states,
                      merged.face_bnd_state]
        d = nat.FhipDesc()
        d.params = C.pointer(merged.fp)
        (d.opp_spts_to_fpts, d.opp_grad_spts, d.opp_extrapolate_fn, d.opp_correction, d.opp_correct_u,
         d.w_spts, d.w_face) = (nat.dptr(a) for a in self._keep[:7])
        d.jginv_spts, d.detjac_spts, d.norm_fpts, d.da_fpts = (nat.dptr(a) for a in self._keep[7:11])
        d.n_faces = len(f.eleL)
        (d.face_ele_l, d.face_loc_l, d.face_ele_r, d.face_loc_r, d.face_rel_rot,
         d.face_bc) = (nat.iptr(a) for a in self._keep[11:17])
        d.n_meshes = len(merged.members)
        d.mesh_ele_start = nat.iptr(merged.ele_start)
        d.n_bnd_states = len(merged.members)
        d.bnd_states = nat.dptr(merged.bnd_states)
        d.face_bnd_state = nat.iptr(merged.face_bnd_state)
        self._desc = d
        h = C.c_void_p()
        nat.hip_check(nat.hip_lib().fhip_create(C.byref(d), C.byref(h)))
        self._ctx = h
        self.nMesh = len(merged.members)
        self.active = np.ones(self.nMesh, dtype=np.int32)

    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            nat.hip_lib().fhip_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize(self):
        lib = nat.hip_lib()
        nat.hip_check(lib.fhip_set_state(self._ctx, nat.dptr(self.m.U)))
        if self.m.params.dtType != 0:
            nat.hip_check(lib.fhip_extrapolate_u(self._ctx))
            nat.hip_check(lib.fhip_calc_wave_speed(self._ctx))

    def residuals(self, res_type: int = 2) -> np.ndarray:
        out = np.zeros((self.nMesh, self.m.nFields))
        nat.hip_check(nat.hip_lib().fhip_mesh_residual_norms(self._ctx, res_type, nat.dptr(out)))
        return out

    def set_active(self, active: np.ndarray) -> None:
        self.active = np.ascontiguousarray(active, dtype=np.int32)
        nat.hip_check(nat.hip_lib().fhip_set_mesh_active(self._ctx, nat.iptr(self.active)))

    def member_states(self) -> list[np.ndarray]:
        U = np.zeros((self.m.nSpts, self.m.nEles, self.m.nFields))
        nat.hip_check(nat.hip_lib().fhip_get_state(self._ctx, nat.dptr(U)))
        s = self.m.ele_start
        return [U[:, s[i]:s[i + 1]] for i in range(self.nMesh)]


class BatchedSolver:
    """Independent cases grouped by order; one context per order group."""

    def __init__(self, members: list[Member], path: int = nat.PATH_FUSED):
        self.members = members
        self.path = path
        groups: dict[tuple, list[int]] = {}
        for i, m in enumerate(members):
            groups.setdefault((m.setup.nDims, m.setup.order), []).append(i)
        self.index = []  # (group, position) of each member
        self.groups: list[BatchGroup] = []
        pos = [None] * len(members)
        for key in sorted(groups):
            idx = groups[key]
            g = BatchGroup(MergedSetup([members[i] for i in idx]))
            for j, i in enumerate(idx):
                pos[i] = (len(self.groups), j)
            self.groups.append(g)
        self.index = pos
        for g in self.groups:
            g.initialize()
        self.converged = np.zeros(len(members), dtype=bool)

    def close(self):
        for g in self.groups:
            g.close()

    def step(self, n: int = 1) -> None:
        arr = (C.c_void_p * len(self.groups))(*[g._ctx.value for g in self.groups])
        nat.hip_check(nat.hip_lib().fhip_update_many(arr, len(self.groups), n, self.path))
        for g in self.groups:
            nat.hip_check(nat.hip_lib().fhip_sync(g._ctx))

    def residuals(self, res_type: int = 2) -> list[np.ndarray]:
        per = [g.residuals(res_type) for g in self.groups]
        return [per[gi][j] for gi, j in self.index]

    def states(self) -> list[np.ndarray]:
        per = [g.member_states() for g in self.groups]
        return [per[gi][j] for gi, j in self.index]

    def solve_steady(self, tol: float = 1e-10, max_steps: int = 100000, check_every: int = 50,
                     res_type: int = 2, field: int = 0) -> dict:
        """Step until every member's residual (field `field`) is <= tol; freeze
        converged members.  Returns steps and residual per member."""
        steps = np.zeros(len(self.members), dtype=np.int64)
        done = 0
        last = [None] * len(self.members)
        while done < max_steps and not self.converged.all():
            self.step(check_every)
            done += check_every
            res = self.residuals(res_type)
            for i, r in enumerate(res):
                if not self.converged[i]:
                    last[i] = r
                    steps[i] = done
                    if np.isnan(r).any():
                        raise RuntimeError(f"NaN in the residual of batch member {i}")
                    if r[field] <= tol:
                        self.converged[i] = True
            for gi, g in enumerate(self.groups):
                act = np.array([0 if self.converged[i] else 1 for i, (gg, _) in enumerate(self.index) if gg == gi],
                               dtype=np.int32)
                g.set_active(act)
        return {"steps": steps, "residual": last, "converged": self.converged.copy()}
