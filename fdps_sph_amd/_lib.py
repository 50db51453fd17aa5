"""ctypes binding of libfdps_sph_gpu.so (include/fdps_sph_gpu.h).

The library is the product: there is no CPU fallback.  Loading fails with an
ImportError naming the missing file when the HIP build is absent.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfdps_sph_gpu.so")

# error codes (include/fdps_sph_gpu.h)
OK = 0
ERR_ARG = -1
ERR_CAPACITY = -2
ERR_TAG = -3
ERR_HIP = -4
ERR_EOS = -5
ERR_STATE = -6
NOALLOC_MAX_N = 1048576  # SPHGPU_NOALLOC_MAX_N

PREC_FP64 = 0
PREC_MIXED = 1

FIELDS = ["POS", "VEL", "ACC", "GRAV", "POT", "MASS", "SMTH", "DENS", "ENG", "PRES", "SNDS", "DIVV", "ROTV",
          "GRADH", "BAL", "ENGDOT", "DT", "NGB", "MAT", "SYS", "ID", "VELHALF", "ENGHALF"]
FIELD = {name: k for k, name in enumerate(FIELDS)}
VEC_FIELDS = {"POS", "VEL", "ACC", "GRAV", "ROTV", "VELHALF"}
I32_FIELDS = {"NGB", "MAT", "SYS"}
I64_FIELDS = {"ID"}


class SimParams(C.Structure):
    _fields_ = [("n_leaf", C.c_int32), ("n_group", C.c_int32), ("theta", C.c_double), ("damping", C.c_double),
                ("ext_force_until", C.c_double), ("nsys", C.c_int32), ("list_cap_sph", C.c_int32),
                ("list_cap_epj", C.c_int32), ("list_cap_spj", C.c_int32), ("eos_density_fp64", C.c_int32),
                ("list_cap_front", C.c_int32), ("list_front_lds", C.c_int32), ("persistent", C.c_int32),
                ("graph", C.c_int32), ("walk_waves", C.c_int32), ("vsig_split", C.c_int32),
                ("hydro_compact", C.c_int32)]


class KernelStat(C.Structure):
    _fields_ = [("ms", C.c_double), ("launches", C.c_int64), ("n_i", C.c_int64), ("n_j", C.c_int64),
                ("n_s", C.c_int64), ("pairs", C.c_int64)]


class SphGpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sphgpu error {code}: {msg}")
        self.code = code


_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m fdps_sph_amd.build` "
                          "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    i32 = C.c_int32
    i64 = C.c_int64
    PP = C.POINTER(C.c_void_p)
    PI = C.POINTER(C.c_int32)
    L.sphgpu_last_error.restype = C.c_char_p
    L.sphgpu_version.restype = i32
    L.sphgpu_create.argtypes = [C.POINTER(C.c_void_p), i32, i32]
    L.sphgpu_destroy.argtypes = [P]
    L.sphgpu_set_precision.argtypes = [P, i32]
    for k in ("dens", "drvt", "hydro"):
        getattr(L, f"sphgpu_{k}_dispatch").argtypes = [P, i32, i32, PP, PI, PP, PI]
        getattr(L, f"sphgpu_{k}_retrieve").argtypes = [P, i32, i32, PI, PP]
    L.sphgpu_grav_dispatch.argtypes = [P, i32, i32, PP, PI, PP, PI, PP, PI]
    L.sphgpu_grav_retrieve.argtypes = [P, i32, i32, PI, PP]
    L.sphgpu_eos_load_table.argtypes = [P, i32, C.c_char_p]
    L.sphgpu_eos_set_table.argtypes = [P, i32, i32, i32, P]
    L.sphgpu_eos_set_tillotson.argtypes = [P, i32, P]
    L.sphgpu_eos_pressure.argtypes = [P, i64, P, P, P, P, P]
    L.sphgpu_eos_pressure_noalloc.argtypes = [P, i64, P, P, P, P, P]
    L.sphgpu_eos_property.argtypes = [P, i64, P, P, P, i32, P]
    L.sphgpu_eos_internal_energy.argtypes = [P, i64, P, P, P, i32, P]
    L.sphgpu_sim_default_params.argtypes = [C.POINTER(SimParams)]
    L.sphgpu_sim_default_params.restype = None
    L.sphgpu_sim_init.argtypes = [P, C.POINTER(SimParams), i64, P, P, P, P, P, P, P, P]
    L.sphgpu_sim_step.argtypes = [P, i32]
    L.sphgpu_sim_size.argtypes = [P, C.POINTER(i64)]
    L.sphgpu_sim_get.argtypes = [P, i32, P]
    L.sphgpu_sim_get_time.argtypes = [P, P, P]
    L.sphgpu_sim_stage_ms.argtypes = [P, P, i32]
    L.sphgpu_sim_list_builds.argtypes = [P, P, P]
    L.sphgpu_sim_graph_stats.argtypes = [P, P, P]
    L.sphgpu_sim_set_capacities.argtypes = [P, i32, i32, i32]
    L.sphgpu_sim_lists.argtypes = [P, i32, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64), P, P, P, P, P]
    L.sphgpu_sim_nodes.argtypes = [P, C.POINTER(i64), P]
    L.sphgpu_sim_set_stats.argtypes = [P, i32]
    L.sphgpu_sim_kernel_stats.argtypes = [P, C.POINTER(KernelStat), i32]
    _LIB = L
    return L


def check(rc: int) -> None:
    if rc != OK:
        msg = lib().sphgpu_last_error()
        raise SphGpuError(rc, msg.decode() if msg else "")


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
