"""ctypes binding of libopencloth.so (the C ABI in include/opencloth.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libopencloth.so")

OC_OK = 0
OC_ERR_INVALID_ARGUMENT = -1
OC_ERR_UNSUPPORTED_TOPOLOGY = -2
OC_ERR_DEVICE = -3
OC_ERR_OUT_OF_MEMORY = -4
OC_ERR_NO_DEVICE = -5

SOLVE_NOT_CONVERGED, SOLVE_BREAKDOWN, SOLVE_NONFINITE = 1, 2, 4
OC_FLAG_GROUND_CLAMP, OC_FLAG_CG_PERSISTENT, OC_FLAG_JACOBI, OC_FLAG_LIST_ORDER = 1, 2, 4, 8
OC_FLAG_CG_SINGLE_REDUCTION = 16

# Every symbol include/opencloth.h declares, with (restype, argtypes).
_P = C.c_void_p
_I = C.c_int32
_F = C.c_float
SIGNATURES = {
    "oc_version": (C.c_char_p, []),
    "oc_last_error": (C.c_char_p, []),
    "oc_device_count": (C.c_int, []),
    "oc_grid_triangles": (C.c_int, [_I, _I, _P]),
    "oc_cloth_tile_shape": (C.c_int, [_P, _P]),
    "oc_cloth_create": (C.c_int, [_I, _I, _P, _P, _P, _I, _P, _P, _P]),
    "oc_cloth_destroy": (C.c_int, [_P]),
    "oc_cloth_set_stream": (C.c_int, [_P, _P]),
    "oc_cloth_set_params": (C.c_int, [_P, _P]),
    "oc_cloth_get_params": (C.c_int, [_P, _P]),
    "oc_cloth_set_state": (C.c_int, [_P, _P, _P]),
    "oc_cloth_get_state": (C.c_int, [_P, _P, _P]),
    "oc_cloth_get_dv": (C.c_int, [_P, _P]),
    "oc_cloth_set_dv": (C.c_int, [_P, _P]),
    "oc_cloth_compute_forces": (C.c_int, [_P, _P]),
    "oc_cloth_assemble": (C.c_int, [_P, _F]),
    "oc_cloth_solve": (C.c_int, [_P, _P, _P]),
    "oc_cloth_update": (C.c_int, [_P, _F]),
    "oc_cloth_integrate_implicit": (C.c_int, [_P, _F]),
    "oc_cloth_step": (C.c_int, [_P, _I]),
    "oc_cloth_step_noalloc": (C.c_int, [_P, _I]),
    "oc_cloth_synchronize": (C.c_int, [_P]),
    "oc_cloth_last_status": (C.c_int, [_P, _P, _P]),
    "oc_cloth_sticky_status": (C.c_int, [_P, _P]),
    "oc_cloth_profile_steps": (C.c_int, [_P, _I, _P, _P]),
    "oc_cloth_time_cg_spmv": (C.c_int, [_P, _I, _P]),
    "oc_cloth_time_kernel": (C.c_int, [_P, _I, _I, _P]),
    "oc_cloth_set_spring_stiffness": (C.c_int, [_P, _P]),
    "oc_cloth_get_spring_stiffness": (C.c_int, [_P, _P]),
    "oc_cloth_rollout_forward": (C.c_int, [_P, _I, _I]),
    "oc_cloth_rollout_backward": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "oc_cloth_num_pairs": (C.c_int, [_P, _P]),
    "oc_cloth_get_pairs": (C.c_int, [_P, _P, _P]),
    "oc_cloth_get_system": (C.c_int, [_P, _P, _P]),
    "oc_cloth_ellipsoid_collision": (C.c_int, [_P, _P]),
    "oc_cloth_provot": (C.c_int, [_P, _F]),
    "oc_cloth_step_physics": (C.c_int, [_P, _I, _P, _F]),
    "oc_verlet_create": (C.c_int, [_I, _I, _P, _P, _P, _I, _P, _P, _P]),
    "oc_verlet_destroy": (C.c_int, [_P]),
    "oc_verlet_set_stream": (C.c_int, [_P, _P]),
    "oc_verlet_set_state": (C.c_int, [_P, _P, _P]),
    "oc_verlet_get_state": (C.c_int, [_P, _P, _P]),
    "oc_verlet_compute_forces": (C.c_int, [_P, _P]),
    "oc_verlet_step": (C.c_int, [_P, _I, _P]),
    "oc_vgrid_create": (C.c_int, [_I, _I, _P, _P, _P, _P, _P]),
    "oc_vgrid_destroy": (C.c_int, [_P]),
    "oc_vgrid_set_stream": (C.c_int, [_P, _P]),
    "oc_vgrid_set_params": (C.c_int, [_P, _P]),
    "oc_vgrid_set_state": (C.c_int, [_P, _P, _P]),
    "oc_vgrid_get_state": (C.c_int, [_P, _P, _P]),
    "oc_vgrid_step": (C.c_int, [_P, _I, _P]),
    "oc_vgrid_synchronize": (C.c_int, [_P]),
    "oc_pbd_create": (C.c_int, [_I, _P, _P, _P, _I, _P, _I, _P, _P]),
    "oc_pbd_destroy": (C.c_int, [_P]),
    "oc_pbd_set_stream": (C.c_int, [_P, _P]),
    "oc_pbd_set_state": (C.c_int, [_P, _P, _P]),
    "oc_pbd_get_state": (C.c_int, [_P, _P, _P]),
    "oc_pbd_num_colors": (C.c_int, [_P, _P, _P]),
    "oc_pbd_step": (C.c_int, [_P, _I, _P]),
    "oc_pbd_synchronize": (C.c_int, [_P]),
    "oc_batch_create": (C.c_int, [_I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "oc_batch_destroy": (C.c_int, [_P]),
    "oc_batch_set_stream": (C.c_int, [_P, _P]),
    "oc_batch_set_params": (C.c_int, [_P, _P]),
    "oc_batch_set_schedule": (C.c_int, [_P, _P]),
    "oc_batch_get_schedule": (C.c_int, [_P, _P]),
    "oc_batch_autotune": (C.c_int, [_P, _P]),
    "oc_batch_threads_per_cloth": (C.c_int, [_P, _P]),
    "oc_batch_set_state": (C.c_int, [_P, _P, _P]),
    "oc_batch_get_state": (C.c_int, [_P, _P, _P]),
    "oc_batch_step": (C.c_int, [_P, _I]),
    "oc_batch_step_eager": (C.c_int, [_P, _I]),
    "oc_batch_get_status": (C.c_int, [_P, _P, _P]),
    "oc_batch_sticky_status": (C.c_int, [_P, _P]),
    "oc_batch_synchronize": (C.c_int, [_P]),
    "oc_batch_set_spring_stiffness": (C.c_int, [_P, _P]),
    "oc_batch_get_spring_stiffness": (C.c_int, [_P, _P]),
    "oc_batch_rollout_forward": (C.c_int, [_P, _I, _I]),
    "oc_batch_rollout_backward": (C.c_int, [_P, _P, _P, _P, _P, _P]),
}


class OcParams(C.Structure):
    """oc_params (include/opencloth.h)."""

    _fields_ = [
        ("mass", C.c_float), ("gravity", C.c_float * 3), ("dt", C.c_float),
        ("cg_max_iters", C.c_int32), ("cg_tol", C.c_float), ("cg_tol_mode", C.c_int32),
        ("warm_start", C.c_int32), ("flags", C.c_int32), ("damping", C.c_float),
    ]


class OcEllipsoid(C.Structure):
    """oc_ellipsoid (include/opencloth.h)."""

    _fields_ = [("transform", C.c_float * 16), ("inverse", C.c_float * 16),
                ("center", C.c_float * 3), ("radius", C.c_float)]

    @classmethod
    def from_dict(cls, e: dict) -> "OcEllipsoid":
        o = cls()
        for k in range(16):
            o.transform[k] = float(e["M"][k])
            o.inverse[k] = float(e["Minv"][k])
        for k in range(3):
            o.center[k] = float(e["center"][k])
        o.radius = float(e["radius"])
        return o


class OcVerletParams(C.Structure):
    """oc_verlet_params (include/opencloth.h)."""

    _fields_ = [("mass", C.c_float), ("gravity", C.c_float * 3), ("damping", C.c_float),
                ("dt", C.c_float), ("flags", C.c_int32)]


class OcVgridParams(C.Structure):
    """oc_vgrid_params (include/opencloth.h)."""

    _fields_ = [("mass", C.c_float), ("gravity", C.c_float * 3), ("damping", C.c_float),
                ("dt", C.c_float), ("spacing", C.c_float * 2), ("ks", C.c_float * 3),
                ("kd", C.c_float * 3)]

    @classmethod
    def from_dict(cls, p: dict) -> "OcVgridParams":
        o = cls()
        o.mass, o.damping, o.dt = p["mass"], p["damping"], p["dt"]
        for k in range(3):
            o.gravity[k] = p["gravity"][k]
            o.ks[k] = p["ks"][k]
            o.kd[k] = p["kd"][k]
        o.spacing[0], o.spacing[1] = p["spacing"]
        return o


class OcPbdParams(C.Structure):
    """oc_pbd_params (include/opencloth.h)."""

    _fields_ = [("mass", C.c_float), ("gravity", C.c_float * 3), ("dt", C.c_float),
                ("global_damping", C.c_float), ("k_damp", C.c_float),
                ("solver_iterations", C.c_int32), ("flags", C.c_int32)]


class OcBatchSchedule(C.Structure):
    """oc_batch_schedule (include/opencloth.h)."""

    _fields_ = [("block_threads", C.c_int32), ("vertices_per_thread", C.c_int32)]


class OpenClothError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"opencloth error {code}: {msg}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OpenClothError(-100, f"{LIB_PATH} is missing: run `python -m opencloth_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != OC_OK:
        raise OpenClothError(rc, lib().oc_last_error().decode(errors="replace"))


def tile_shape():
    """(tile_x, tile_y) of the single-cloth kernels (oc_cloth_tile_shape)."""
    tx, ty = C.c_int32(), C.c_int32()
    check(lib().oc_cloth_tile_shape(C.byref(tx), C.byref(ty)))
    return tx.value, ty.value


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
