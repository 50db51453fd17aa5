"""Host-side mirror of the reference's per-step interfaces over the C-ABI.

* ``Context.calc_density / calc_derivative / calc_hydro / calc_gravity`` take
  lists of (epi, epj[, spj]) walks with the reference's EPI/EPJ/RESULT layouts
  and return RESULT arrays -- the functor calls of src/force.h, batched the way
  FDPS's multiwalk dispatch/retrieve batches them.
* ``Context.pressure / property / internal_energy`` -- EoS_t evaluation
  (src/EoS.h).
* ``Simulation`` -- the device-resident step of src/main.cpp:69-173.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import structs as S
from ._lib import (FIELD, I32_FIELDS, I64_FIELDS, PREC_FP64, PREC_MIXED, VEC_FIELDS, KernelStat, SimParams, check,
                   lib, ptr)


def _walk_arrays(walks: Sequence[np.ndarray], dtype):
    arrs = [np.ascontiguousarray(w, dtype=dtype) for w in walks]
    n = np.array([len(a) for a in arrs], dtype=np.int32)
    p = (C.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])
    return arrs, n, p


class Context:
    """A GPU context (one HIP stream, EoS tables, multiwalk batches)."""

    def __init__(self, device: int = 0, precision: str = "fp64"):
        self._L = lib()
        prec = {"fp64": PREC_FP64, "mixed": PREC_MIXED}[precision]
        h = C.c_void_p()
        check(self._L.sphgpu_create(C.byref(h), int(device), prec))
        self._h = h
        self.precision = precision

    def close(self):
        if getattr(self, "_h", None):
            self._L.sphgpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_precision(self, precision: str):
        """Switch the form of the work launched from now on ("fp64" or "mixed")."""
        check(self._L.sphgpu_set_precision(self._h, {"fp64": PREC_FP64, "mixed": PREC_MIXED}[precision]))
        self.precision = precision

    # ------------------------------------------------------------ EoS
    def load_table(self, material: int, path: str):
        check(self._L.sphgpu_eos_load_table(self._h, int(material), path.encode()))

    def set_table(self, material: int, cells: np.ndarray):
        cells = np.ascontiguousarray(cells, dtype=np.float64)
        assert cells.ndim == 3 and cells.shape[2] == 6
        check(self._L.sphgpu_eos_set_table(self._h, int(material), cells.shape[0], cells.shape[1], ptr(cells)))

    def set_tillotson(self, material: int, params):
        """Analytic Tillotson material; params = (rho0, u0, u_iv, u_cv, A, B, a,
        b, alpha, beta) as src/EoS.h:446 (see eos_tables.GRANITE_TILLOTSON)."""
        prm = np.ascontiguousarray(params, dtype=np.float64)
        assert prm.shape == (10,)
        check(self._L.sphgpu_eos_set_tillotson(self._h, int(material), ptr(prm)))

    def pressure(self, material, dens, eng):
        mat = np.ascontiguousarray(material, dtype=np.int32)
        d = np.ascontiguousarray(dens, dtype=np.float64)
        e = np.ascontiguousarray(eng, dtype=np.float64)
        p = np.empty_like(d)
        c = np.empty_like(d)
        check(self._L.sphgpu_eos_pressure(self._h, len(d), ptr(mat), ptr(d), ptr(e), ptr(p), ptr(c)))
        return p, c

    def property(self, material, dens, eng, prop: int):
        mat = np.ascontiguousarray(material, dtype=np.int32)
        d = np.ascontiguousarray(dens, dtype=np.float64)
        e = np.ascontiguousarray(eng, dtype=np.float64)
        out = np.empty_like(d)
        check(self._L.sphgpu_eos_property(self._h, len(d), ptr(mat), ptr(d), ptr(e), int(prop), ptr(out)))
        return out

    def internal_energy(self, material, dens, ent, abs_int: bool = True):
        mat = np.ascontiguousarray(material, dtype=np.int32)
        d = np.ascontiguousarray(dens, dtype=np.float64)
        s = np.ascontiguousarray(ent, dtype=np.float64)
        out = np.empty_like(d)
        check(self._L.sphgpu_eos_internal_energy(self._h, len(d), ptr(mat), ptr(d), ptr(s), int(bool(abs_int)),
                                                 ptr(out)))
        return out

    # ------------------------------------------------------------ multiwalk functors
    _TYPES = {"dens": (S.EPI_DENS, S.EPJ_DENS, S.RES_DENS), "drvt": (S.EPI_DRVT, S.EPJ_DRVT, S.RES_DRVT),
              "hydro": (S.EPI_HYDRO, S.EPJ_HYDRO, S.RES_HYDRO), "grav": (S.EPI_GRAV, S.EPJ_GRAV, S.RES_GRAV)}

    def dispatch(self, kind: str, walks_i, walks_j, walks_s=None, tag: int = 0) -> np.ndarray:
        """FDPS multiwalk dispatch of one batch of walks on `tag`; returns the
        n_epi array that the matching retrieve takes."""
        di, dj, _ = self._TYPES[kind]
        ai, ni, pi = _walk_arrays(walks_i, di)
        aj, nj, pj = _walk_arrays(walks_j, dj)
        nw = len(ai)
        pn = ni.ctypes.data_as(C.POINTER(C.c_int32))
        pnj = nj.ctypes.data_as(C.POINTER(C.c_int32))
        disp = getattr(self._L, f"sphgpu_{kind}_dispatch")
        if kind == "grav":
            if walks_s is None:
                check(disp(self._h, tag, nw, pi, pn, pj, pnj, None, None))
            else:
                as_, ns, ps = _walk_arrays(walks_s, S.SPJ_MONOPOLE)
                check(disp(self._h, tag, nw, pi, pn, pj, pnj, ps, ns.ctypes.data_as(C.POINTER(C.c_int32))))
        else:
            check(disp(self._h, tag, nw, pi, pn, pj, pnj))
        return ni

    def retrieve(self, kind: str, n_epi: np.ndarray, tag: int = 0):
        """FDPS multiwalk retrieve of the batch dispatched on `tag`."""
        dres = self._TYPES[kind][2]
        n_epi = np.ascontiguousarray(n_epi, dtype=np.int32)
        nw = len(n_epi)
        res = [np.zeros(int(k), dtype=dres) for k in n_epi]
        pr = (C.c_void_p * max(1, nw))(*[r.ctypes.data for r in res])
        check(getattr(self._L, f"sphgpu_{kind}_retrieve")(self._h, tag, nw, n_epi.ctypes.data_as(C.POINTER(C.c_int32)),
                                                          pr))
        return res

    def _run(self, kind, walks_i, walks_j, tag, walks_s=None):
        return self.retrieve(kind, self.dispatch(kind, walks_i, walks_j, walks_s, tag), tag)

    def calc_density(self, walks_i, walks_j, tag=0):
        return self._run("dens", walks_i, walks_j, tag)

    def calc_derivative(self, walks_i, walks_j, tag=0):
        return self._run("drvt", walks_i, walks_j, tag)

    def calc_hydro(self, walks_i, walks_j, tag=0):
        return self._run("hydro", walks_i, walks_j, tag)

    def calc_gravity(self, walks_i, walks_j, walks_s=None, tag=0):
        return self._run("grav", walks_i, walks_j, tag, walks_s)


def default_params(**kw) -> SimParams:
    p = SimParams()
    lib().sphgpu_sim_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class Simulation:
    """Particles resident in HBM; each step runs src/main.cpp:125-173 on the GPU.

    A context holds one simulation: creating another on the same context
    replaces it, and the replaced object then raises on every call."""

    def __init__(self, ctx: Context, ic: dict, **params):
        self.ctx = ctx
        self._L = lib()
        ctx._sim_serial = getattr(ctx, "_sim_serial", 0) + 1
        self._serial = ctx._sim_serial
        self.params = default_params(**params)
        pos = np.ascontiguousarray(ic["pos"], dtype=np.float64)
        vel = np.ascontiguousarray(ic["vel"], dtype=np.float64)
        mass = np.ascontiguousarray(ic["mass"], dtype=np.float64)
        dens = np.ascontiguousarray(ic["dens"], dtype=np.float64)
        eng = np.ascontiguousarray(ic["eng"], dtype=np.float64)
        mat = np.ascontiguousarray(ic["mat"], dtype=np.int32)
        sys_ = ic.get("sys")
        sys_ = None if sys_ is None else np.ascontiguousarray(sys_, dtype=np.int32)
        ids = ic.get("id")
        ids = None if ids is None else np.ascontiguousarray(ids, dtype=np.int64)
        self.n = len(mass)
        check(self._L.sphgpu_sim_init(ctx.handle, C.byref(self.params), self.n, ptr(pos), ptr(vel), ptr(mass),
                                      ptr(dens), ptr(eng), ptr(mat), ptr(sys_), ptr(ids)))
        self.nsys = max(1, self.params.nsys)

    @property
    def _h(self):
        if getattr(self.ctx, "_sim_serial", None) != self._serial:
            raise RuntimeError("this Simulation was replaced by a newer one on the same context")
        return self.ctx.handle

    def step(self, n: int = 1):
        check(self._L.sphgpu_sim_step(self._h, int(n)))

    def get(self, field: str) -> np.ndarray:
        f = field.upper()
        if f in VEC_FIELDS:
            out = np.empty((self.n, 3))
        elif f in I32_FIELDS:
            out = np.empty(self.n, dtype=np.int32)
        elif f in I64_FIELDS:
            out = np.empty(self.n, dtype=np.int64)
        else:
            out = np.empty(self.n)
        check(self._L.sphgpu_sim_get(self._h, FIELD[f], ptr(out)))
        return out

    def time(self):
        dt = np.empty(self.nsys)
        t = np.empty(self.nsys)
        check(self._L.sphgpu_sim_get_time(self._h, ptr(dt), ptr(t)))
        return dt, t

    def stage_ms(self) -> np.ndarray:
        ms = np.zeros(8)
        check(self._L.sphgpu_sim_stage_ms(self._h, ptr(ms), 8))
        return ms

    def list_builds(self) -> tuple[int, int]:
        """(gather-list constructions, force evaluations) since init."""
        g, f = C.c_int64(), C.c_int64()
        check(self._L.sphgpu_sim_list_builds(self._h, C.byref(g), C.byref(f)))
        return g.value, f.value

    def set_capacities(self, sph: int, epj: int, spj: int):
        """Interaction-list capacities of the next force evaluations."""
        check(self._L.sphgpu_sim_set_capacities(self._h, int(sph), int(epj), int(spj)))

    def graph_stats(self) -> tuple[int, int]:
        """(steps run as graph replays, graphs recorded) since init."""
        r, c = C.c_int64(), C.c_int64()
        check(self._L.sphgpu_sim_graph_stats(self._h, C.byref(r), C.byref(c)))
        return r.value, c.value

    def set_stats(self, on: bool = True):
        check(self._L.sphgpu_sim_set_stats(self._h, int(bool(on))))

    def kernel_stats(self) -> dict:
        names = ["dens", "drvt", "hydro", "grav", "dens_fp64", "vsig"]
        arr = (KernelStat * len(names))()
        check(self._L.sphgpu_sim_kernel_stats(self._h, arr, len(names)))
        return {names[k]: {f: getattr(arr[k], f) for f, _ in KernelStat._fields_} for k in range(len(names))}

    def lists(self, kind: int):
        ng, nj, ns = C.c_int64(), C.c_int64(), C.c_int64()
        h = self._h
        check(self._L.sphgpu_sim_lists(h, kind, C.byref(ng), C.byref(nj), C.byref(ns), None, None, None, None, None))
        G = ng.value
        gstart = np.empty(G + 1, dtype=np.int32)
        joff = np.empty(G + 1, dtype=np.int64)
        jidx = np.empty(max(1, nj.value), dtype=np.int32)
        soff = np.empty(G + 1, dtype=np.int64)
        sidx = np.empty(max(1, ns.value), dtype=np.int32)
        check(self._L.sphgpu_sim_lists(h, kind, C.byref(ng), C.byref(nj), C.byref(ns), ptr(gstart), ptr(joff),
                                       ptr(jidx), ptr(soff), ptr(sidx)))
        out = {"group_start": gstart, "j_off": joff, "j_idx": jidx[:nj.value]}
        if kind == 2:
            out["s_off"] = soff
            out["s_idx"] = sidx[:ns.value]
        return out

    def nodes(self) -> np.ndarray:
        nn = C.c_int64()
        check(self._L.sphgpu_sim_nodes(self._h, C.byref(nn), None))
        m = np.empty((nn.value, 10))
        check(self._L.sphgpu_sim_nodes(self._h, C.byref(nn), ptr(m)))
        return m

    def state(self) -> dict:
        names = ["POS", "VEL", "ACC", "GRAV", "POT", "MASS", "SMTH", "DENS", "ENG", "PRES", "SNDS", "DIVV", "ROTV",
                 "GRADH", "BAL", "ENGDOT", "DT", "NGB", "MAT", "SYS", "ID", "VELHALF", "ENGHALF"]
        return {k.lower(): self.get(k) for k in names}
