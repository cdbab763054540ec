"""ctypes bindings of libflurry_hip (include/flurry_hip.h) and
libflurry_host (include/flurry_host.h).

The libraries are loaded from ``flurrypp_amd/lib``; a missing library raises
immediately (there is no fallback implementation).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class FhipParams(C.Structure):
    """fhip_params (include/flurry_hip.h)."""
    _fields_ = [
        ("n_dims", C.c_int), ("n_fields", C.c_int), ("order", C.c_int),
        ("n_eles", C.c_int), ("n_spts", C.c_int), ("n_fpts", C.c_int), ("n_fpts_face", C.c_int),
        ("equation", C.c_int), ("viscous", C.c_int), ("riemann_type", C.c_int), ("time_type", C.c_int),
        ("dt_type", C.c_int), ("n_rk_steps", C.c_int), ("fix_vis", C.c_int), ("init_iter", C.c_int),
        ("rk_a", C.c_double * 8), ("rk_b", C.c_double * 8),
        ("cfl", C.c_double), ("dt", C.c_double),
        ("gamma", C.c_double), ("prandtl", C.c_double), ("mu_inf", C.c_double), ("rt_inf", C.c_double),
        ("c_sth", C.c_double), ("pen_fact", C.c_double), ("tau", C.c_double),
        ("rho_bound", C.c_double), ("u_bound", C.c_double), ("v_bound", C.c_double), ("w_bound", C.c_double),
        ("p_bound", C.c_double), ("t_wall", C.c_double), ("r_gas", C.c_double), ("u_wall", C.c_double),
        ("v_wall", C.c_double), ("w_wall", C.c_double),
    ]


class FhipDesc(C.Structure):
    """fhip_desc (include/flurry_hip.h)."""
    _fields_ = [
        ("params", C.POINTER(FhipParams)),
        ("opp_spts_to_fpts", _dp), ("opp_grad_spts", _dp), ("opp_extrapolate_fn", _dp),
        ("opp_correction", _dp), ("opp_correct_u", _dp), ("w_spts", _dp), ("w_face", _dp),
        ("jginv_spts", _dp), ("detjac_spts", _dp), ("norm_fpts", _dp), ("da_fpts", _dp),
        ("n_faces", C.c_int),
        ("face_ele_l", _ip), ("face_loc_l", _ip), ("face_ele_r", _ip), ("face_loc_r", _ip),
        ("face_rel_rot", _ip), ("face_bc", _ip),
        ("n_meshes", C.c_int), ("mesh_ele_start", _ip),
        ("n_bnd_states", C.c_int), ("bnd_states", _dp), ("face_bnd_state", _ip),
    ]


# Array ids (FHIP_ARR_*)
ARR = {
    "U_spts": 0, "U_fpts": 1, "F_spts": 2, "divF_spts": 3, "disFn_fpts": 4, "Fn_fpts": 5,
    "dU_spts": 6, "dU_fpts": 7, "dUc_fpts": 8, "U0": 9, "waveSp": 10, "ele_dt": 11,
    "face_fpt_L": 12, "face_fpt_R": 13,
}
PATH_AUTO, PATH_STAGED, PATH_FUSED, PATH_MIXED = 0, 1, 2, 3
# mixed precision with the z contractions of P3 hexes as BF16 hi+lo (hi+mid+lo) matrix-core products
PATH_MIXED_BF16, PATH_MIXED_BF16X3 = 4, 5


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


_hip = None
_host = None


def hip_lib() -> C.CDLL:
    """libflurry_hip; raises if it has not been built."""
    global _hip
    if _hip is None:
        path = LIB_DIR / "libflurry_hip.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: build it with `python -m flurrypp_amd.build` "
                               "(there is no CPU fallback)")
        lib = C.CDLL(str(path))
        lib.fhip_last_error.restype = C.c_char_p
        lib.fhip_create.argtypes = [C.POINTER(FhipDesc), C.POINTER(C.c_void_p)]
        lib.fhip_destroy.argtypes = [C.c_void_p]
        lib.fhip_destroy.restype = None
        for name in ("fhip_set_state", "fhip_get_state"):
            getattr(lib, name).argtypes = [C.c_void_p, _dp]
        for name in ("fhip_extrapolate_u", "fhip_calc_wave_speed", "fhip_copy_u_to_u0", "fhip_copy_u0_to_u",
                     "fhip_sync", "fhip_check_state"):
            getattr(lib, name).argtypes = [C.c_void_p]
        lib.fhip_calc_residual.argtypes = [C.c_void_p, C.c_int]
        lib.fhip_time_step_a.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp]
        lib.fhip_time_step_b.argtypes = [C.c_void_p, C.c_int, _dp]
        lib.fhip_calc_dt.argtypes = [C.c_void_p, C.c_int, _dp]
        lib.fhip_set_dt.argtypes = [C.c_void_p, C.c_double]
        lib.fhip_get_dt.argtypes = [C.c_void_p, _dp]
        lib.fhip_get_iter.argtypes = [C.c_void_p, _ip]
        lib.fhip_update.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.fhip_path_supported.argtypes = [C.c_void_p, C.c_int]
        lib.fhip_set_option.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.fhip_residual_norm.argtypes = [C.c_void_p, C.c_int, _dp]
        lib.fhip_mesh_residual_norms.argtypes = [C.c_void_p, C.c_int, _dp]
        lib.fhip_set_mesh_active.argtypes = [C.c_void_p, _ip]
        lib.fhip_update_many.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int]
        for name in ("fhip_autotune", "fhip_get_block_sizes"):
            getattr(lib, name).argtypes = [C.c_void_p, C.c_int, _ip]
        lib.fhip_set_block_sizes.argtypes = [C.c_void_p, C.c_int, _ip]
        lib.fhip_pmg_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(_dp), C.POINTER(_dp), C.c_int,
                                        C.POINTER(C.c_void_p)]
        lib.fhip_pmg_create_batch.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p), C.c_int,
                                              C.POINTER(_dp), C.POINTER(_dp), C.c_int, C.POINTER(C.c_void_p)]
        lib.fhip_pmg_cycle.argtypes = [C.c_void_p, C.c_int]
        lib.fhip_pmg_destroy.argtypes = [C.c_void_p]
        lib.fhip_pmg_destroy.restype = None
        lib.fhip_array_size.argtypes = [C.c_void_p, C.c_int]
        lib.fhip_array_size.restype = C.c_longlong
        lib.fhip_get_array.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        lib.fhip_set_array.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        lib.fhip_device_ptr.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        lib.fhip_stream.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        lib.fhip_enable_timing.argtypes = [C.c_void_p, C.c_int]
        lib.fhip_kernel_times.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_int, _dp,
                                          C.POINTER(C.c_longlong)]
        lib.fhip_set_device.argtypes = [C.c_int]
        _hip = lib
    return _hip


def host_lib() -> C.CDLL:
    global _host
    if _host is None:
        path = LIB_DIR / "libflurry_host.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: build it with `python -m flurrypp_amd.build`")
        lib = C.CDLL(str(path))
        lib.fhost_last_error.restype = C.c_char_p
        lib.fhost_bc_from_name.argtypes = [C.c_char_p]
        lib.fhost_sizes.argtypes = [C.c_int, C.c_int, _ip, _ip, _ip]
        lib.fhost_pts1d.argtypes = [C.c_int, C.c_int, _dp]
        lib.fhost_weights1d.argtypes = [C.c_int, _dp]
        lib.fhost_ref_points.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        lib.fhost_operators.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int] + [_dp] * 7
        lib.fhost_mesh_cartesian.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                             C.POINTER(C.c_char_p), C.POINTER(C.c_void_p)]
        lib.fhost_mesh_from_arrays.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _ip, C.c_int, _ip, _ip, _ip,
                                               C.POINTER(C.c_void_p)]
        lib.fhost_mesh_read_gmsh.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p),
                                             C.POINTER(C.c_void_p)]
        lib.fhost_mesh_jitter.argtypes = [C.c_void_p, _dp, C.c_double, C.c_ulonglong]
        lib.fhost_mesh_set_periodic.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double]
        lib.fhost_mesh_process.argtypes = [C.c_void_p]
        lib.fhost_mesh_counts.argtypes = [C.c_void_p, _ip, _ip, _ip, _ip, _ip]
        lib.fhost_mesh_get_nodes.argtypes = [C.c_void_p, _dp, _ip]
        lib.fhost_mesh_get_faces.argtypes = [C.c_void_p] + [_ip] * 8
        lib.fhost_mesh_geometry.argtypes = [C.c_void_p, C.c_int, C.c_int] + [_dp] * 6
        lib.fhost_mesh_set_shape_nodes.argtypes = [C.c_void_p, C.c_int, _ip]
        lib.fhost_mesh_get_shape_nodes.argtypes = [C.c_void_p, _ip, _ip]
        lib.fhost_mesh_destroy.argtypes = [C.c_void_p]
        lib.fhost_mesh_destroy.restype = None
        _host = lib
    return _host


class FhipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[fhip {code}] {msg}")
        self.code = code


def hip_check(rc: int) -> None:
    if rc != 0:
        raise FhipError(rc, hip_lib().fhip_last_error().decode())


def host_check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(host_lib().fhost_last_error().decode())
