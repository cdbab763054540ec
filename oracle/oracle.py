"""ctypes wrapper of the CPU restatement (liboracle_fr.so, fr_oracle.h).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker or the CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

ARR = {"U_spts": 0, "U_fpts": 1, "F_spts": 2, "divF_spts": 3, "disFn_fpts": 4, "Fn_fpts": 5, "dU_spts": 6,
       "dU_fpts": 7, "dUc_fpts": 8, "U0": 9, "waveSp": 10, "ele_dt": 11, "face_fpt_L": 12, "face_fpt_R": 13}


class OrParams(C.Structure):
    _fields_ = [
        ("n_dims", C.c_int), ("n_fields", C.c_int), ("order", C.c_int), ("n_eles", C.c_int), ("n_spts", C.c_int),
        ("n_fpts", C.c_int), ("n_fpts_face", C.c_int), ("equation", C.c_int), ("viscous", C.c_int),
        ("riemann_type", C.c_int), ("time_type", C.c_int), ("dt_type", C.c_int), ("n_rk_steps", C.c_int),
        ("fix_vis", C.c_int), ("init_iter", C.c_int), ("rk_a", C.c_double * 8), ("rk_b", C.c_double * 8),
        ("cfl", C.c_double), ("dt", C.c_double), ("gamma", C.c_double), ("prandtl", C.c_double),
        ("mu_inf", C.c_double), ("rt_inf", C.c_double), ("c_sth", C.c_double), ("pen_fact", C.c_double),
        ("tau", C.c_double), ("rho_bound", C.c_double), ("u_bound", C.c_double), ("v_bound", C.c_double),
        ("w_bound", C.c_double), ("p_bound", C.c_double), ("t_wall", C.c_double), ("r_gas", C.c_double),
        ("u_wall", C.c_double), ("v_wall", C.c_double), ("w_wall", C.c_double),
    ]


class OrDesc(C.Structure):
    _fields_ = [("params", C.POINTER(OrParams))] + [(n, _dp) for n in (
        "opp_spts_to_fpts", "opp_grad_spts", "opp_extrapolate_fn", "opp_correction", "opp_correct_u", "w_spts",
        "w_face", "jginv_spts", "detjac_spts", "norm_fpts", "da_fpts")] + [("n_faces", C.c_int)] + [
        (n, _ip) for n in ("face_ele_l", "face_loc_l", "face_ele_r", "face_loc_r", "face_rel_rot", "face_bc")]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return HERE / "liboracle_fr.so"


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = HERE / "liboracle_fr.so"
        if not path.exists():
            build()
        L = C.CDLL(str(path))
        L.or_last_error.restype = C.c_char_p
        L.or_create.argtypes = [C.POINTER(OrDesc), C.POINTER(C.c_void_p)]
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_destroy.restype = None
        L.or_array_size.argtypes = [C.c_void_p, C.c_int]
        L.or_array_size.restype = C.c_longlong
        L.or_get_array.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.or_set_array.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        for n in ("or_extrapolate_u", "or_calc_wave_speed", "or_copy_u_to_u0", "or_copy_u0_to_u", "or_get_iter"):
            getattr(L, n).argtypes = [C.c_void_p]
        L.or_calc_residual.argtypes = [C.c_void_p, C.c_int]
        L.or_time_step_a.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp]
        L.or_time_step_b.argtypes = [C.c_void_p, C.c_int, _dp]
        L.or_calc_dt.argtypes = [C.c_void_p, C.c_int, _dp]
        L.or_set_dt.argtypes = [C.c_void_p, C.c_double]
        L.or_update.argtypes = [C.c_void_p, C.c_int, _dp]
        L.or_residual_norm.argtypes = [C.c_void_p, C.c_int, _dp]
        L.or_get_dt.argtypes = [C.c_void_p]
        L.or_get_dt.restype = C.c_double
        L.or_pmg_cycle.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(_dp), C.POINTER(_dp), C.c_int]
        _lib = L
    return _lib


def _chk(rc):
    if rc != 0:
        raise RuntimeError(lib().or_last_error().decode())


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """One oracle context built from the same host arrays libflurry_hip gets."""

    def __init__(self, fhip_params, opers, geo, faces):
        p = OrParams()
        for name, _ in OrParams._fields_:
            v = getattr(fhip_params, name)
            if name in ("rk_a", "rk_b"):
                for i in range(8):
                    getattr(p, name)[i] = v[i]
            else:
                setattr(p, name, v)
        self._p = p
        self.N, self.S, self.F, self.E, self.D = p.n_fields, p.n_spts, p.n_fpts, p.n_eles, p.n_dims
        keep = [_d(opers.spts_to_fpts), _d(opers.grad_spts), _d(opers.extrap_fn), _d(opers.correction),
                _d(opers.correct_u), _d(opers.w_spts), _d(opers.w_face), _d(geo.jginv_spts), _d(geo.detjac_spts),
                _d(geo.norm_fpts), _d(geo.da_fpts)]
        ik = [np.ascontiguousarray(a, dtype=np.int32) for a in
              (faces.eleL, faces.locL, faces.eleR, faces.locR, faces.relRot, faces.bc)]
        self._keep = keep + ik
        d = OrDesc()
        d.params = C.pointer(p)
        for (name, _), a in zip(OrDesc._fields_[1:12], keep):
            setattr(d, name, a.ctypes.data_as(_dp))
        d.n_faces = len(ik[0])
        for (name, _), a in zip(OrDesc._fields_[13:], ik):
            setattr(d, name, a.ctypes.data_as(_ip))
        self._d = d
        h = C.c_void_p()
        _chk(lib().or_create(C.byref(d), C.byref(h)))
        self._h = h

    @classmethod
    def from_setup(cls, h) -> "Oracle":
        """From a flurrypp_amd.solver.HostSetup (or a Solver)."""
        fp = getattr(h, "fp", None) or getattr(h, "_fp")
        return cls(fp, h.opers, h.geo, h.faces)

    def __del__(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().or_destroy(self._h)
            self._h = None

    def get(self, name: str, step: int = 0) -> np.ndarray:
        w = ARR[name]
        n = lib().or_array_size(self._h, w)
        out = np.zeros(n, dtype=np.int32 if name.startswith("face_fpt") else np.float64)
        _chk(lib().or_get_array(self._h, w, step, out.ctypes.data_as(C.c_void_p)))
        return out

    def put(self, name: str, a, step: int = 0) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        _chk(lib().or_set_array(self._h, ARR[name], step, a.ctypes.data_as(C.c_void_p)))

    def set_U(self, U):
        self.put("U_spts", U)

    def get_U(self):
        return self.get("U_spts").reshape(self.S, self.E, self.N)

    def extrapolateU(self):
        _chk(lib().or_extrapolate_u(self._h))

    def calcWaveSp(self):
        _chk(lib().or_calc_wave_speed(self._h))

    def initializeSolution(self, U, dt_type: int):
        self.set_U(U)
        if dt_type != 0:
            self.extrapolateU()
            self.calcWaveSp()

    def calcResidual(self, step):
        _chk(lib().or_calc_residual(self._h, step))

    def timeStepA(self, step, rk, src=None):
        _chk(lib().or_time_step_a(self._h, step, rk, None if src is None else _d(src).ctypes.data_as(_dp)))

    def timeStepB(self, step, src=None):
        _chk(lib().or_time_step_b(self._h, step, None if src is None else _d(src).ctypes.data_as(_dp)))

    def copyUspts_U0(self):
        _chk(lib().or_copy_u_to_u0(self._h))

    def copyU0_Uspts(self):
        _chk(lib().or_copy_u0_to_u(self._h))

    def calcDt(self, refresh: bool) -> float:
        v = C.c_double()
        _chk(lib().or_calc_dt(self._h, int(refresh), C.byref(v)))
        return v.value

    def set_dt(self, dt):
        _chk(lib().or_set_dt(self._h, dt))

    @property
    def dt(self) -> float:
        """The time step of the last update (or calcDt)."""
        return lib().or_get_dt(self._h)

    def update(self, n=1, src=None):
        _chk(lib().or_update(self._h, n, None if src is None else _d(src).ctypes.data_as(_dp)))

    def residual_norm(self, res_type=2):
        out = np.zeros(self.N)
        _chk(lib().or_residual_norm(self._h, res_type, out.ctypes.data_as(_dp)))
        return out

    @property
    def iter(self):
        return lib().or_get_iter(self._h)


def pmg_cycle(levels: list, restrict_ops: list, prolong_ops: list, smooth_steps: int) -> None:
    """One P-multigrid cycle (multiGrid::cycle, src/multigrid.cpp:169-313)
    over oracle contexts of ascending order (last = fine).  restrict_ops[i]
    maps level i+1 to level i, prolong_ops[i] level i to level i+1."""
    n = len(levels)
    hs = (C.c_void_p * n)(*[lv._h.value for lv in levels])
    R = [_d(r) for r in restrict_ops]
    Pr = [_d(p) for p in prolong_ops]
    Ra = (_dp * max(1, n - 1))(*[r.ctypes.data_as(_dp) for r in R])
    Pa = (_dp * max(1, n - 1))(*[p.ctypes.data_as(_dp) for p in Pr])
    _chk(lib().or_pmg_cycle(hs, n, Ra, Pa, smooth_steps))
