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

import copy
import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import _native as nat
from .mesh import Faces, Geometry, Operators, pmg_operators
from .solver import HostSetup


@dataclass
class Member:
    setup: HostSetup
    U: np.ndarray
    tag: object = None


# fhip_params fields a merged context keeps per mesh (the free stream, through
# its boundary-state table) or sums (the element count)
_PER_MESH = ("n_eles", "rho_bound", "u_bound", "v_bound", "w_bound", "p_bound")


# fhip_params fields that follow from the order
_PER_ORDER = ("order", "n_spts", "n_fpts", "n_fpts_face")


def context_key(setup, with_order: bool = True) -> tuple:
    """What one context takes from a case besides its mesh and free stream:
    element type and order, the operators' point set and correction scheme,
    and every other fhip_params field (dt only where it is fixed).  Cases
    merge into one context only when these are equal.  with_order=False
    leaves out the order and what follows from it."""
    p, fp = setup.params, setup.fp
    vals = [setup.nDims, setup.order if with_order else None, str(p.sptsTypeQuad).lower(), p.vcjhSchemeQuad]
    for name, _ in nat.FhipParams._fields_:
        if name in _PER_MESH or (name == "dt" and fp.dt_type != 0) or (not with_order and name in _PER_ORDER):
            continue
        v = getattr(fp, name)
        vals.append(tuple(v) if isinstance(v, C.Array) else v)
    return tuple(vals)


class MergedSetup:
    """HostSetup-compatible union of several meshes of one order."""

    def __init__(self, members: list[Member]):
        s0 = members[0].setup
        k0 = context_key(s0)
        for m in members:
            if context_key(m.setup) != k0:
                raise ValueError("batch members of one context must share element type, order, operators and "
                                 "every solver option but the free stream")
        self.params = s0.params
        self.nDims, self.nFields, self.order = s0.nDims, s0.nFields, s0.order
        self.nSpts, self.nFpts, self.nFptsFace = s0.nSpts, s0.nFpts, s0.nFptsFace
        self.opers: Operators = s0.opers
        self.members = members
        self.ele_start = np.zeros(len(members) + 1, dtype=np.int32)
        for i, m in enumerate(members):
            self.ele_start[i + 1] = self.ele_start[i] + m.setup.nEles
        self.nEles = int(self.ele_start[-1])
        # element axis: 1 for [pt][ele](...), 2 for JGinv_spts [dim][spt][ele][dim]
        def cat(name, axis=1):
            return np.ascontiguousarray(np.concatenate([getattr(m.setup.geo, name) for m in members], axis=axis))
        self.geo = Geometry(pos_spts=cat("pos_spts"), pos_fpts=cat("pos_fpts"), jginv_spts=cat("jginv_spts", 2),
                            detjac_spts=cat("detjac_spts"), norm_fpts=cat("norm_fpts"), da_fpts=cat("da_fpts"))
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
    """Independent cases grouped by order; one context per order group.  With
    pmg=True every group of order P > low_order also gets coarse levels of
    orders low_order..P-1 (the same meshes) and a P-multigrid cycle after each
    of its updates, on its own stream (the groups' cycles overlap on the
    device).  merged_cycle=True (fused paths) instead merges the coarse levels
    of one order across the groups that differ in order alone and runs one
    batched cycle (fhip_pmg_create_batch) for them: a third of the launches,
    but one stream per family of groups.  Both give the same results bit for
    bit; on a 64-case bump-channel batch (P 1..5) the default measured 7.07
    ms per step against 7.31 ms merged."""

    def __init__(self, members: list[Member], path: int = nat.PATH_FUSED, pmg: bool = False, low_order: int = 0,
                 smooth_steps: int = 1, autotune: bool = False, merged_cycle: bool = False):
        self.members = members
        self.path = path
        self.pmg = pmg
        self.smooth = smooth_steps
        # one merged context per element type, order and solver options
        groups: dict[tuple, list[int]] = {}
        for i, m in enumerate(members):
            groups.setdefault(context_key(m.setup), []).append(i)
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
        # P-multigrid: coarse groups per fine group, and one fhip_pmg each
        self.coarse: list[list[BatchGroup]] = [[] for _ in self.groups]
        self._mg: list = [None] * len(self.groups)
        self._ops = []
        # batched cycles: the groups above low_order that differ in order alone
        # (a family; e.g. the tuned steady RK coefficients of timeType 6 differ
        # between orders <= 2 and > 2, and every coarse level keeps its fine
        # level's) share one cycle over coarse levels merged across them
        self._merged: list[dict] = []
        if pmg and merged_cycle and path in (nat.PATH_FUSED, nat.PATH_MIXED):
            fams: dict[tuple, list[int]] = {}
            for gi, g in enumerate(self.groups):
                if g.m.order > low_order:
                    fams.setdefault(context_key(g.m.members[0].setup, with_order=False), []).append(gi)
            for fine in fams.values():
                self._build_merged_cycle(sorted(fine, key=lambda gi: self.groups[gi].m.order), low_order,
                                         smooth_steps)
        elif pmg:
            for gi, g in enumerate(self.groups):
                P = g.m.order
                if P <= low_order:
                    continue
                for q in range(low_order, P):
                    cm = []
                    for m in g.m.members:
                        pc = copy.copy(m.setup.params)
                        pc.order = q
                        hs = HostSetup(pc, m.setup.mesh)
                        cm.append(Member(hs, hs.initial_condition(), m.tag))
                    cg = BatchGroup(MergedSetup(cm))
                    cg.initialize()
                    self.coarse[gi].append(cg)
                levels = self.coarse[gi] + [g]
                D = g.m.nDims
                sp = g.m.params.sptsTypeQuad
                R = [pmg_operators(D, lv.m.order, sp)[1] for lv in levels[1:]]
                Pr = [pmg_operators(D, lv.m.order, sp)[0] for lv in levels[:-1]]
                self._ops.append((R, Pr))
                n = len(levels)
                hs_ = (C.c_void_p * n)(*[lv._ctx.value for lv in levels])
                Ra = (nat._dp * (n - 1))(*[nat.dptr(a) for a in R])
                Pa = (nat._dp * (n - 1))(*[nat.dptr(a) for a in Pr])
                h = C.c_void_p()
                nat.hip_check(nat.hip_lib().fhip_pmg_create(hs_, n, Ra, Pa, smooth_steps, C.byref(h)))
                self._mg[gi] = h
        # block sizes of every context's kernels, timed on its own data;
        # merged coarse levels under the key (0, order)
        self.block_sizes = {}
        if autotune and path in (nat.PATH_FUSED, nat.PATH_MIXED):
            tuned = [(g.m.order, lv) for gi, g in enumerate(self.groups) for lv in self.coarse[gi] + [g]]
            tuned += [(0, lv) for mc in self._merged for lv in mc["coarse"]]
            for key, lv in tuned:
                out = np.zeros(5, dtype=np.int32)
                nat.hip_check(nat.hip_lib().fhip_autotune(lv._ctx, path, nat.iptr(out)))
                self.block_sizes[(key, lv.m.order)] = out.tolist()

    def _build_merged_cycle(self, fine: list[int], low_order: int, smooth_steps: int) -> None:
        """One batched cycle for the fine groups `fine` (ascending order)."""
        mc = {"fine": fine, "coarse": [], "h": None}
        self._merged.append(mc)
        top = self.groups[fine[-1]].m.order
        for q in range(low_order, top):
            cm = []
            for gi in fine:
                if self.groups[gi].m.order <= q:
                    continue
                for m in self.groups[gi].m.members:
                    pc = copy.copy(m.setup.params)
                    pc.order = q
                    hs = HostSetup(pc, m.setup.mesh)
                    cm.append(Member(hs, hs.initial_condition(), m.tag))
            cg = BatchGroup(MergedSetup(cm))
            cg.initialize()
            mc["coarse"].append(cg)
        g0 = self.groups[fine[0]].m
        D, sp = g0.nDims, g0.params.sptsTypeQuad
        R = [pmg_operators(D, q + 1, sp)[1] for q in range(low_order, top)]
        Pr = [pmg_operators(D, q, sp)[0] for q in range(low_order, top)]
        self._ops.append((R, Pr))
        nf, nc = len(fine), len(mc["coarse"])
        fa = (C.c_void_p * nf)(*[self.groups[gi]._ctx.value for gi in fine])
        ca = (C.c_void_p * nc)(*[cg._ctx.value for cg in mc["coarse"]])
        Ra = (nat._dp * nc)(*[nat.dptr(a) for a in R])
        Pa = (nat._dp * nc)(*[nat.dptr(a) for a in Pr])
        h = C.c_void_p()
        nat.hip_check(nat.hip_lib().fhip_pmg_create_batch(fa, nf, ca, nc, Ra, Pa, smooth_steps, C.byref(h)))
        mc["h"] = h

    def close(self):
        for mc in self._merged:
            if mc["h"] is not None and mc["h"].value:
                nat.hip_lib().fhip_pmg_destroy(mc["h"])
            mc["h"] = None
        for h in self._mg:
            if h is not None and h.value:
                nat.hip_lib().fhip_pmg_destroy(h)
        self._mg = [None] * len(self.groups)
        for cs in self.coarse:
            for g in cs:
                g.close()
        for mc in self._merged:
            for g in mc["coarse"]:
                g.close()
        for g in self.groups:
            g.close()

    def step(self, n: int = 1) -> None:
        arr = (C.c_void_p * len(self.groups))(*[g._ctx.value for g in self.groups])
        lib = nat.hip_lib()
        if self._merged:
            for _ in range(n):
                nat.hip_check(lib.fhip_update_many(arr, len(self.groups), 1, self.path))
                for mc in self._merged:
                    nat.hip_check(lib.fhip_pmg_cycle(mc["h"], self.path))
        elif not any(h is not None for h in self._mg):
            nat.hip_check(lib.fhip_update_many(arr, len(self.groups), n, self.path))
        else:
            for _ in range(n):
                nat.hip_check(lib.fhip_update_many(arr, len(self.groups), 1, self.path))
                for h in self._mg:
                    if h is not None:
                        nat.hip_check(lib.fhip_pmg_cycle(h, self.path))
        for g in self.groups:
            nat.hip_check(lib.fhip_sync(g._ctx))

    def residuals(self, res_type: int = 2) -> list[np.ndarray]:
        per = [g.residuals(res_type) for g in self.groups]
        return [per[gi][j] for gi, j in self.index]

    def states(self) -> list[np.ndarray]:
        per = [g.member_states() for g in self.groups]
        return [per[gi][j] for gi, j in self.index]

    def solve_steady(self, tol: float = 1e-10, max_steps: int = 100000, check_every: int = 50,
                     res_type: int = 2, field: int = 0, stop_on_nan: bool = True, progress=None) -> dict:
        """Step until every member's residual (field `field`) is <= tol; freeze
        converged members.  A member whose residual turns NaN raises, or with
        stop_on_nan=False is frozen and reported in "diverged".  Returns steps,
        residual and convergence time per member."""
        steps = np.zeros(len(self.members), dtype=np.int64)
        t_conv = np.full(len(self.members), np.nan)
        diverged = np.zeros(len(self.members), dtype=bool)
        done = 0
        last = [None] * len(self.members)
        dof = np.array([m.setup.nEles * m.setup.nSpts * m.setup.params.nRKSteps for m in self.members])
        work = 0
        t0 = time.perf_counter()
        while done < max_steps and not (self.converged | diverged).all():
            k = min(check_every, max_steps - done)  # never past max_steps
            self.step(k)
            done += k
            work += int(dof[~(self.converged | diverged)].sum()) * k
            res = self.residuals(res_type)
            now = time.perf_counter() - t0
            for i, r in enumerate(res):
                if not (self.converged[i] or diverged[i]):
                    last[i] = r
                    steps[i] = done
                    if not np.isfinite(r).all():
                        if stop_on_nan:
                            raise RuntimeError(f"NaN in the residual of batch member {i}")
                        diverged[i] = True
                    elif r[field] <= tol:
                        self.converged[i] = True
                        t_conv[i] = now
            stopped = self.converged | diverged
            if progress is not None:
                progress(done, int(self.converged.sum()), int(diverged.sum()))
            acts = []
            for gi, g in enumerate(self.groups):
                act = np.array([0 if stopped[i] else 1 for i, (gg, _) in enumerate(self.index) if gg == gi],
                               dtype=np.int32)
                acts.append(act)
                g.set_active(act)
                for cg in self.coarse[gi]:
                    cg.set_active(act)
            # a merged coarse level holds the groups above its order, ascending
            for mc in self._merged:
                for cg in mc["coarse"]:
                    cg.set_active(np.concatenate([acts[gi] for gi in mc["fine"]
                                                  if self.groups[gi].m.order > cg.m.order]))
        return {"steps": steps, "residual": last, "converged": self.converged.copy(), "diverged": diverged,
                "time_s": t_conv,
                "wall_s": time.perf_counter() - t0, "dof_stages": work}
