"""ctypes wrapper of the CPU oracle (oracle/sph_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The shipped package
(fdps_sph_amd) never imports this module.

Two builds of the same C source are available:
  * strict  (-O2 -ffp-contract=off, serial)  -- parity checker
  * fast    (-O3 -ffast-math -funroll-loops -fopenmp, the reference's flags)
            -- CPU baseline timing
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# Byte layouts of the reference boundary types (src/class.h:36-86, 233-453;
# PS::SPJMonopole = {mass, pos}).
EPI_DENS = np.dtype([("pos", "<f8", 3), ("mass", "<f8"), ("smth", "<f8"), ("id", "<i8")])
EPJ_DENS = np.dtype([("mass", "<f8"), ("pos", "<f8", 3), ("smth", "<f8")])
RES_DENS = np.dtype([("dens", "<f8"), ("smth", "<f8")])
EPI_DRVT = np.dtype([("pos", "<f8", 3), ("vel", "<f8", 3), ("smth", "<f8"), ("dens", "<f8")])
EPJ_DRVT = np.dtype([("mass", "<f8"), ("pos", "<f8", 3), ("vel", "<f8", 3), ("smth", "<f8")])
RES_DRVT = np.dtype([("div_v", "<f8"), ("rot_v", "<f8", 3), ("grad_smth", "<f8")])
EPI_HYDRO = np.dtype([("pos", "<f8", 3), ("vel", "<f8", 3), ("smth", "<f8"), ("dens", "<f8"), ("pres", "<f8"),
                      ("snds", "<f8"), ("grad_smth", "<f8"), ("Bal", "<f8"), ("AVa", "<f8"), ("id", "<i8")])
EPJ_HYDRO = np.dtype([("pos", "<f8", 3), ("vel", "<f8", 3), ("dens", "<f8"), ("mass", "<f8"), ("smth", "<f8"),
                      ("pres", "<f8"), ("grad_smth", "<f8"), ("snds", "<f8"), ("Bal", "<f8"), ("AVa", "<f8"),
                      ("id", "<i8")])
RES_HYDRO = np.dtype([("acc", "<f8", 3), ("eng_dot", "<f8"), ("dt", "<f8")])
EPI_GRAV = np.dtype([("pos", "<f8", 3), ("eps2", "<f8"), ("id", "<i8")])
EPJ_GRAV = np.dtype([("pos", "<f8", 3), ("mass", "<f8"), ("id", "<i8")])
SPJ_MONO = np.dtype([("mass", "<f8"), ("pos", "<f8", 3)])
RES_GRAV = np.dtype([("acc", "<f8", 3), ("pot", "<f8"), ("dt", "<f8")])

# SoA particle state, field order of orc_state
STATE_F64 = ["x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "gx", "gy", "gz", "pot", "m", "h", "rho", "u",
             "p", "cs", "divv", "rotx", "roty", "rotz", "gs", "bal", "du", "dti", "vhx", "vhy", "vhz", "uh"]
STATE_I32 = ["mat", "sys", "ngb"]


class _State(C.Structure):
    _fields_ = ([("n", C.c_int64)] + [(f, C.POINTER(C.c_double)) for f in STATE_F64]
                + [(f, C.POINTER(C.c_int32)) for f in STATE_I32] + [("id", C.POINTER(C.c_int64))])


class _Params(C.Structure):
    _fields_ = [("n_leaf", C.c_int), ("n_group", C.c_int), ("theta", C.c_double), ("damping", C.c_double),
                ("ext_force_until", C.c_double), ("nsys", C.c_int), ("fast_lists", C.c_int)]


class _EosSet(C.Structure):
    _fields_ = [("nmat", C.c_int), ("tab", C.POINTER(C.c_void_p))]


def build() -> None:
    """Compile both oracle variants (make in oracle/)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_LIBS: dict = {}


def lib(kind: str = "strict") -> C.CDLL:
    if kind in _LIBS:
        return _LIBS[kind]
    path = os.path.join(HERE, f"liboracle_{kind}.so")
    # make rebuilds only what is older than sphgpu_oracle.c/.h
    try:
        build()
    except (OSError, subprocess.CalledProcessError):
        if not os.path.exists(path):
            raise
    L = C.CDLL(path)
    P = C.c_void_p
    d = C.c_double
    i = C.c_int
    i64 = C.c_int64
    L.orc_W.restype = d
    L.orc_W.argtypes = [P, d]
    L.orc_gradW.argtypes = [P, d, P]
    for name in ["orc_calc_density", "orc_calc_derivative", "orc_calc_hydro", "orc_calc_grav_ep",
                 "orc_calc_grav_sp", "orc_count_ngb"]:
        getattr(L, name).argtypes = [P, i, P, i, P]
    for name in ["orc_clear_dens", "orc_clear_drvt", "orc_clear_hydro", "orc_clear_grav"]:
        getattr(L, name).argtypes = [P, i]
    L.orc_balsara.restype = d
    L.orc_balsara.argtypes = [d, P, d, d]
    L.orc_table_load.restype = P
    L.orc_table_load.argtypes = [C.c_char_p]
    L.orc_table_from_cells.restype = P
    L.orc_table_from_cells.argtypes = [i, i, P]
    L.orc_table_free.argtypes = [P]
    L.orc_table_tillotson.restype = P
    L.orc_table_tillotson.argtypes = [P]
    L.orc_tillotson.argtypes = [P, d, d, P, P]
    L.orc_table_dims.argtypes = [P, C.POINTER(i), C.POINTER(i)]
    L.orc_table_cells.restype = C.POINTER(d)
    L.orc_table_cells.argtypes = [P]
    L.orc_eos_bilinear.restype = d
    L.orc_eos_bilinear.argtypes = [P, d, d, i]
    L.orc_eos_internal_energy.restype = d
    L.orc_eos_internal_energy.argtypes = [P, d, d, i]
    L.orc_sort_state.argtypes = [C.POINTER(_State), i]
    L.orc_sim_init.argtypes = [C.POINTER(_State), C.POINTER(_EosSet), C.POINTER(_Params), P, P]
    L.orc_sim_step.argtypes = [C.POINTER(_State), C.POINTER(_EosSet), C.POINTER(_Params), P, P]
    L.orc_tree_build.restype = P
    L.orc_tree_build.argtypes = [C.POINTER(_State), C.POINTER(_Params)]
    L.orc_tree_free.argtypes = [P]
    L.orc_tree_ngroups.restype = i64
    L.orc_tree_ngroups.argtypes = [P]
    L.orc_tree_group_start.restype = C.POINTER(i64)
    L.orc_tree_group_start.argtypes = [P]
    L.orc_tree_nnodes.restype = i64
    L.orc_tree_nnodes.argtypes = [P]
    L.orc_tree_nlevels.restype = i
    L.orc_tree_nlevels.argtypes = [P]
    L.orc_tree_level_offset.restype = C.POINTER(i64)
    L.orc_tree_level_offset.argtypes = [P]
    L.orc_tree_node.argtypes = [P, i64, P, P, P]
    L.orc_tree_update_maxH.argtypes = [P, C.POINTER(_State)]
    L.orc_list_sph.restype = i64
    L.orc_list_sph.argtypes = [P, C.POINTER(_State), i64, i, i, P, i64]
    L.orc_list_grav.restype = i
    L.orc_list_grav.argtypes = [P, C.POINTER(_State), i64, d, P, i64, P, P, i64, P]
    L.orc_morton.restype = C.c_uint64
    L.orc_morton.argtypes = [d, d, d, P, d, i]
    L.orc_hydro_term_scale.argtypes = [P, i, P, i, P, P]
    L.orc_drvt_term_scale.argtypes = [P, i, P, i, P, P]
    L.orc_state_term_scales.argtypes = [C.POINTER(_State), C.POINTER(_Params), P, P, P, P, P]
    L.orc_wtime.restype = d
    L.orc_num_threads.restype = i
    _LIBS[kind] = L
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- kernels
def W(dr, h, kind="strict") -> float:
    dr = np.ascontiguousarray(dr, dtype=np.float64)
    return lib(kind).orc_W(_p(dr), float(h))


def gradW(dr, h, kind="strict") -> np.ndarray:
    dr = np.ascontiguousarray(dr, dtype=np.float64)
    out = np.zeros(3)
    lib(kind).orc_gradW(_p(dr), float(h), _p(out))
    return out


# ---------------------------------------------------------------- functors
def _functor(name, clear, res_dtype, epi, epj, kind):
    L = lib(kind)
    epi = np.ascontiguousarray(epi)
    epj = np.ascontiguousarray(epj)
    res = np.zeros(len(epi), dtype=res_dtype)
    if clear:
        getattr(L, clear)(_p(res), len(res))
    getattr(L, name)(_p(epi), len(epi), _p(epj), len(epj), _p(res))
    return res


def calc_density(epi, epj, kind="strict"):
    return _functor("orc_calc_density", "orc_clear_dens", RES_DENS, epi, epj, kind)


def count_ngb(epi, epj, kind="strict"):
    out = np.zeros(len(epi), dtype=np.int32)
    lib(kind).orc_count_ngb(_p(np.ascontiguousarray(epi)), len(epi), _p(np.ascontiguousarray(epj)), len(epj),
                            _p(out))
    return out


def calc_derivative(epi, epj, kind="strict"):
    return _functor("orc_calc_derivative", "orc_clear_drvt", RES_DRVT, epi, epj, kind)


def calc_hydro(epi, epj, kind="strict"):
    return _functor("orc_calc_hydro", "orc_clear_hydro", RES_HYDRO, epi, epj, kind)


def calc_grav(epi, epj, spj=None, kind="strict"):
    L = lib(kind)
    epi = np.ascontiguousarray(epi)
    res = np.zeros(len(epi), dtype=RES_GRAV)
    L.orc_clear_grav(_p(res), len(res))
    epj = np.ascontiguousarray(epj)
    L.orc_calc_grav_ep(_p(epi), len(epi), _p(epj), len(epj), _p(res))
    if spj is not None:
        spj = np.ascontiguousarray(spj)
        L.orc_calc_grav_sp(_p(epi), len(epi), _p(spj), len(spj), _p(res))
    return res


def hydro_term_scale(epi, epj, kind="strict"):
    epi = np.ascontiguousarray(epi)
    epj = np.ascontiguousarray(epj)
    a = np.zeros(len(epi))
    u = np.zeros(len(epi))
    lib(kind).orc_hydro_term_scale(_p(epi), len(epi), _p(epj), len(epj), _p(a), _p(u))
    return a, u


def drvt_term_scale(epi, epj, kind="strict"):
    epi = np.ascontiguousarray(epi)
    epj = np.ascontiguousarray(epj)
    d = np.zeros(len(epi))
    r = np.zeros(len(epi))
    lib(kind).orc_drvt_term_scale(_p(epi), len(epi), _p(epj), len(epj), _p(d), _p(r))
    return d, r


def state_term_scales(state, prm, kind="strict"):
    """Per-particle sizes of the summed terms: acc, du, div, rot, grav."""
    n = state.n
    out = [np.zeros(n) for _ in range(5)]
    st = state.cstruct()
    lib(kind).orc_state_term_scales(C.byref(st), C.byref(prm), *[_p(a) for a in out])
    return dict(zip(["acc", "du", "div", "rot", "grav"], out))


def balsara(div_v, rot_v, snds, smth, kind="strict") -> float:
    r = np.ascontiguousarray(rot_v, dtype=np.float64)
    return lib(kind).orc_balsara(float(div_v), _p(r), float(snds), float(smth))


# ---------------------------------------------------------------- EoS
class Table:
    """An EoS table in the reference's rho-u format (src/EoS.h:290-385)."""

    def __init__(self, path=None, cells=None, kind="strict", tillotson=None):
        self._L = lib(kind)
        if tillotson is not None:  # analytic Tillotson parameters (10)
            self._prm_keep = np.ascontiguousarray(tillotson, dtype=np.float64)
            assert self._prm_keep.shape == (10,)
            self._t = self._L.orc_table_tillotson(_p(self._prm_keep))
        elif path is not None:
            self._t = self._L.orc_table_load(path.encode())
            if not self._t:
                raise IOError(f"cannot read EoS table {path}")
        else:
            cells = np.ascontiguousarray(cells, dtype=np.float64)
            self._cells_keep = cells
            self._t = self._L.orc_table_from_cells(cells.shape[0], cells.shape[1], _p(cells))
        nr, nu = C.c_int(), C.c_int()
        self._L.orc_table_dims(self._t, C.byref(nr), C.byref(nu))
        self.nrho, self.nu = nr.value, nu.value

    def __del__(self):
        try:
            self._L.orc_table_free(self._t)
        except Exception:
            pass

    @property
    def handle(self):
        return self._t

    def cells(self) -> np.ndarray:
        p = self._L.orc_table_cells(self._t)
        return np.ctypeslib.as_array(p, shape=(self.nrho, self.nu, 6)).copy()

    def bilinear(self, rho, u, prop) -> float:
        return self._L.orc_eos_bilinear(self._t, float(rho), float(u), int(prop))

    def pressure(self, rho, u):
        return self.bilinear(rho, u, 3)

    def sound_speed(self, rho, u):
        return self.bilinear(rho, u, 4)

    def temperature(self, rho, u):
        return self.bilinear(rho, u, 2)

    def entropy(self, rho, u):
        return self.bilinear(rho, u, 5)

    def internal_energy(self, rho, S, abs_int=True):
        return self._L.orc_eos_internal_energy(self._t, float(rho), float(S), int(bool(abs_int)))


# ---------------------------------------------------------------- state / sim
class State:
    """Numpy SoA particle state mirrored into an orc_state."""

    def __init__(self, n: int):
        self.n = int(n)
        for f in STATE_F64:
            setattr(self, f, np.zeros(self.n))
        for f in STATE_I32:
            setattr(self, f, np.zeros(self.n, dtype=np.int32))
        self.id = np.arange(self.n, dtype=np.int64)
        self.bal[:] = 1.0
        self.gs[:] = 1.0

    def copy(self) -> "State":
        s = State(self.n)
        for f in STATE_F64 + STATE_I32 + ["id"]:
            setattr(s, f, getattr(self, f).copy())
        return s

    def cstruct(self) -> _State:
        st = _State()
        st.n = self.n
        for f in STATE_F64:
            a = getattr(self, f)
            assert a.dtype == np.float64 and a.flags.c_contiguous and len(a) == self.n, f
            setattr(st, f, a.ctypes.data_as(C.POINTER(C.c_double)))
        for f in STATE_I32:
            a = getattr(self, f)
            assert a.dtype == np.int32 and a.flags.c_contiguous and len(a) == self.n, f
            setattr(st, f, a.ctypes.data_as(C.POINTER(C.c_int32)))
        st.id = self.id.ctypes.data_as(C.POINTER(C.c_int64))
        return st


def params(n_leaf=16, n_group=64, theta=0.5, damping=1.0, ext_force_until=5000.0, nsys=1, fast_lists=1):
    return _Params(n_leaf, n_group, theta, damping, ext_force_until, nsys, fast_lists)


class Sim:
    """Full-step oracle: src/main.cpp:69-104 (init) and :125-173 (step)."""

    def __init__(self, state: State, tables, prm: _Params, kind="strict"):
        self.L = lib(kind)
        self.s = state
        self.tables = tables
        self.prm = prm
        nsys = max(1, prm.nsys)
        self.dt = np.zeros(nsys)
        self.time = np.zeros(nsys)
        arr = (C.c_void_p * len(tables))(*[t.handle for t in tables])
        self._arr = arr
        self.eos = _EosSet(len(tables), C.cast(arr, C.POINTER(C.c_void_p)))

    def init(self):
        st = self.s.cstruct()
        self.L.orc_sim_init(C.byref(st), C.byref(self.eos), C.byref(self.prm), _p(self.dt), _p(self.time))

    def step(self):
        st = self.s.cstruct()
        self.L.orc_sim_step(C.byref(st), C.byref(self.eos), C.byref(self.prm), _p(self.dt), _p(self.time))


def sort_state(state: State, nsys=1, kind="strict"):
    st = state.cstruct()
    lib(kind).orc_sort_state(C.byref(st), nsys)


class Tree:
    def __init__(self, state: State, prm: _Params, kind="strict"):
        self.L = lib(kind)
        self.s = state
        self._st = state.cstruct()
        self.prm = prm
        self.t = self.L.orc_tree_build(C.byref(self._st), C.byref(prm))

    def __del__(self):
        try:
            self.L.orc_tree_free(self.t)
        except Exception:
            pass

    @property
    def ngroups(self):
        return self.L.orc_tree_ngroups(self.t)

    def group_start(self):
        g = self.ngroups
        return np.ctypeslib.as_array(self.L.orc_tree_group_start(self.t), shape=(g + 1,)).copy()

    @property
    def nnodes(self):
        return self.L.orc_tree_nnodes(self.t)

    @property
    def nlevels(self):
        return self.L.orc_tree_nlevels(self.t)

    def level_offset(self):
        return np.ctypeslib.as_array(self.L.orc_tree_level_offset(self.t), shape=(self.nlevels + 1,)).copy()

    def node(self, k):
        r = np.zeros(2, dtype=np.int64)
        c = np.zeros(2, dtype=np.int64)
        m = np.zeros(11)
        self.L.orc_tree_node(self.t, int(k), _p(r), _p(c), _p(m))
        return r, c, m

    def update_maxH(self):
        self.L.orc_tree_update_maxH(self.t, C.byref(self._st))

    def list_sph(self, g, kind=0, brute=False, cap=1 << 20):
        out = np.zeros(cap, dtype=np.int32)
        n = self.L.orc_list_sph(self.t, C.byref(self._st), int(g), int(kind), int(brute), _p(out), cap)
        if n < 0:
            raise RuntimeError("list capacity exceeded")
        return out[:n].copy()

    def list_grav(self, g, theta, cap=1 << 20):
        ej = np.zeros(cap, dtype=np.int32)
        es = np.zeros(cap, dtype=np.int32)
        nj = C.c_int64()
        ns = C.c_int64()
        rc = self.L.orc_list_grav(self.t, C.byref(self._st), int(g), float(theta), _p(ej), cap, C.byref(nj),
                                  _p(es), cap, C.byref(ns))
        if rc:
            raise RuntimeError("list capacity exceeded")
        return ej[:nj.value].copy(), es[:ns.value].copy()


def morton(x, y, z, lo, side, bits=21) -> int:
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    return lib().orc_morton(float(x), float(y), float(z), _p(lo), float(side), int(bits))


def wtime(kind="fast") -> float:
    return lib(kind).orc_wtime()


def num_threads(kind="fast") -> int:
    return lib(kind).orc_num_threads()
