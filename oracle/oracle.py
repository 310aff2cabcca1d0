"""ctypes binding of the CPU restatement (oracle/cumc_oracle.c).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libcumc_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("coordinate", ctypes.c_int), ("boundary_flag", ctypes.c_int * 6),
        ("recon", ctypes.c_int), ("model", ctypes.c_int), ("geom_source_repeat", ctypes.c_int),
        ("ggas", ctypes.c_double), ("cfl", ctypes.c_double),
        ("rho_floor", ctypes.c_double), ("eps_floor", ctypes.c_double), ("r_sh", ctypes.c_double),
        ("solver", ctypes.c_int),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcConfig), _D, _D, _D]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_ncell.restype = ctypes.c_long
        L.orc_set_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_ncell.argtypes = [ctypes.c_void_p]
        L.orc_set_state.argtypes = [ctypes.c_void_p, _D, _D]
        L.orc_get_state.argtypes = [ctypes.c_void_p, _D, _D, _D, _D]
        L.orc_get_lrk.argtypes = [ctypes.c_void_p, _D]
        L.orc_initial_update.argtypes = [ctypes.c_void_p]
        L.orc_finddt.restype = ctypes.c_double
        L.orc_finddt.argtypes = [ctypes.c_void_p]
        L.orc_rungekutta.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_spatial.argtypes = [ctypes.c_void_p]
        L.orc_boundary.argtypes = [ctypes.c_void_p]
        L.orc_find_divb.argtypes = [ctypes.c_void_p, _D, _D]
        L.orc_floor_counts.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]
        L.orc_sweep_flux.argtypes = [ctypes.c_void_p, ctypes.c_int, _D]
        L.orc_run.restype = ctypes.c_int
        L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _D, _D]
        L.orc_ppm.argtypes = [_D, _D, _D, _D]
        L.orc_tvdmm.argtypes = [_D, _D, _D, _D]
        L.orc_signalspeed.restype = ctypes.c_double
        L.orc_signalspeed.argtypes = [ctypes.c_double] * 5
        L.orc_tvd.argtypes = [ctypes.c_int, _D, _D, _D, _D]
        L.orc_weno5.argtypes = [_D, _D, _D]
        for name in ("orc_hll", "orc_lf", "orc_hllc", "orc_hlld"):
            getattr(L, name).restype = ctypes.c_int
            getattr(L, name).argtypes = [ctypes.c_int, ctypes.c_double, _D, _D, _D]
        L.orc_ptof.argtypes = [ctypes.c_int, _D, _D, _D]
        L.orc_ptou_face.argtypes = [_D, ctypes.c_double, _D]
        L.orc_ptou_cell.argtypes = [_D, ctypes.c_double, _D, _D]
        L.orc_utop_cell.argtypes = [_D, _D, _D, _D]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def _cfg(cfg) -> OrcConfig:
    c = OrcConfig()
    c.nx, c.ny, c.nz = cfg.nx, cfg.ny, cfg.nz
    c.coordinate = cfg.coordinate
    for i in range(6):
        c.boundary_flag[i] = int(cfg.boundary_flag[i])
    c.recon, c.model, c.geom_source_repeat = cfg.recon, cfg.model, cfg.geom_source_repeat
    c.ggas, c.cfl = cfg.ggas, cfg.cfl
    c.rho_floor, c.eps_floor, c.r_sh = cfg.rho_floor, cfg.eps_floor, cfg.r_sh
    c.solver = getattr(cfg, "solver", 2)
    return c


class Oracle:
    """The reference time step on the CPU, one instance per problem."""

    def __init__(self, cfg, xf, yf, zf):
        self.cfg = cfg
        self._xf = np.ascontiguousarray(xf, dtype=np.float64)
        self._yf = np.ascontiguousarray(yf, dtype=np.float64)
        self._zf = np.ascontiguousarray(zf, dtype=np.float64)
        c = _cfg(cfg)
        self.h = lib().orc_create(ctypes.byref(c), _p(self._xf), _p(self._yf), _p(self._zf))
        if not self.h:
            raise MemoryError("orc_create failed")

    @classmethod
    def from_problem(cls, prob):
        o = cls(prob.cfg, prob.xf, prob.yf, prob.zf)
        o.set_state(prob.prim, prob.eps)
        return o

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, prim, eps):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        assert prim.shape == self.cfg.shape + (8,) and eps.shape == self.cfg.shape
        lib().orc_set_state(self.h, _p(prim), _p(eps))

    def get_state(self):
        prim = np.empty(self.cfg.shape + (8,))
        eps = np.empty(self.cfg.shape)
        bcell = np.empty(self.cfg.shape + (3,))
        cons = np.empty(self.cfg.shape + (8,))
        lib().orc_get_state(self.h, _p(prim), _p(eps), _p(bcell), _p(cons))
        return {"prim": prim, "eps": eps, "bcell": bcell, "cons": cons}

    def l_rk(self):
        a = np.empty(self.cfg.shape + (8,))
        lib().orc_get_lrk(self.h, _p(a))
        return a

    def initial_update(self):
        lib().orc_initial_update(self.h)

    def finddt(self) -> float:
        return float(lib().orc_finddt(self.h))

    def rungekutta(self, dt: float):
        lib().orc_rungekutta(self.h, float(dt))

    def spatial(self):
        lib().orc_spatial(self.h)

    def boundary(self):
        lib().orc_boundary(self.h)

    def sweep_flux(self, direction: int):
        out = np.zeros(self.cfg.shape + (7,))
        lib().orc_sweep_flux(self.h, int(direction), _p(out))
        return out

    def find_divb(self):
        a = ctypes.c_double()
        b = ctypes.c_double()
        lib().orc_find_divb(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def floor_counts(self):
        """(density, eps) floor applications by CUSTOM_CHECKRHO since creation."""
        a = ctypes.c_long()
        b = ctypes.c_long()
        lib().orc_floor_counts(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def run(self, max_steps: int, total_time: float = 1.0e300, t0: float = 0.0):
        dts = np.zeros(max(1, min(max_steps, 10 ** 6)))
        t = ctypes.c_double(t0)
        n = lib().orc_run(self.h, int(min(max_steps, dts.size)), float(total_time), _p(dts), ctypes.byref(t))
        return dts[:n].copy(), t.value


# ---- point-level functions --------------------------------------------------------------------------

def set_threads(n: int) -> int:
    """OpenMP threads of the restatement's loops (n <= 0 leaves them); returns the count in force."""
    return int(lib().orc_set_threads(int(n)))


def ppm(d, v):
    d = np.ascontiguousarray(d, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    vm, vp = ctypes.c_double(), ctypes.c_double()
    lib().orc_ppm(_p(d), _p(v), ctypes.byref(vm), ctypes.byref(vp))
    return vm.value, vp.value


def tvdmm(d, v):
    d = np.ascontiguousarray(d, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    vm, vp = ctypes.c_double(), ctypes.c_double()
    lib().orc_tvdmm(_p(d), _p(v), ctypes.byref(vm), ctypes.byref(vp))
    return vm.value, vp.value


def tvd(lim, d, v):
    """TVD edge values with limiter lim (1 minmod, 2 MC, 3 van Leer)."""
    d = np.ascontiguousarray(d, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    vm, vp = ctypes.c_double(), ctypes.c_double()
    lib().orc_tvd(int(lim), _p(d), _p(v), ctypes.byref(vm), ctypes.byref(vp))
    return vm.value, vp.value


def weno5(v):
    """WENO5 edge values (vm, vp) of the middle cell of the 5-cell stencil v."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    vm, vp = ctypes.c_double(), ctypes.c_double()
    lib().orc_weno5(_p(v), ctypes.byref(vm), ctypes.byref(vp))
    return vm.value, vp.value


def signalspeed(cs, bn, bt1, bt2, rho):
    return lib().orc_signalspeed(cs, bn, bt1, bt2, rho)


def hll(direction, gamma, pl, pr):
    pl = np.ascontiguousarray(pl, dtype=np.float64)
    pr = np.ascontiguousarray(pr, dtype=np.float64)
    f = np.zeros(8)
    br = lib().orc_hll(int(direction), float(gamma), _p(pl), _p(pr), _p(f))
    return f, br


def lf(direction, gamma, pl, pr):
    pl = np.ascontiguousarray(pl, dtype=np.float64)
    pr = np.ascontiguousarray(pr, dtype=np.float64)
    f = np.zeros(8)
    br = lib().orc_lf(int(direction), float(gamma), _p(pl), _p(pr), _p(f))
    return f, br


def hllc(direction, gamma, pl, pr):
    """HLLC flux (8 unrotated components) and the wave region: 0 F_L, 1 F*_L, 2 F*_R, 3 F_R."""
    pl = np.ascontiguousarray(pl, dtype=np.float64)
    pr = np.ascontiguousarray(pr, dtype=np.float64)
    f = np.zeros(8)
    br = lib().orc_hllc(int(direction), float(gamma), _p(pl), _p(pr), _p(f))
    return f, br


def hlld(direction, gamma, pl, pr):
    """HLLD flux (8 unrotated components) and the wave region: 0 F_L, 1 F*_L, 2 F**_L,
    3 F**_R, 4 F*_R, 5 F_R."""
    pl = np.ascontiguousarray(pl, dtype=np.float64)
    pr = np.ascontiguousarray(pr, dtype=np.float64)
    f = np.zeros(8)
    br = lib().orc_hlld(int(direction), float(gamma), _p(pl), _p(pr), _p(f))
    return f, br


def ptof(direction, p, u):
    p = np.ascontiguousarray(p, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    f = np.zeros(8)
    lib().orc_ptof(int(direction), _p(p), _p(u), _p(f))
    return f


def ptou_face(p, eps):
    p = np.ascontiguousarray(p, dtype=np.float64)
    u = np.zeros(8)
    lib().orc_ptou_face(_p(p), float(eps), _p(u))
    return u


def ptou_cell(p, eps, bc):
    p = np.ascontiguousarray(p, dtype=np.float64)
    bc = np.ascontiguousarray(bc, dtype=np.float64)
    u = np.zeros(8)
    lib().orc_ptou_cell(_p(p), float(eps), _p(bc), _p(u))
    return u


def utop_cell(u, bc):
    u = np.ascontiguousarray(u, dtype=np.float64)
    bc = np.ascontiguousarray(bc, dtype=np.float64)
    p = np.zeros(8)
    e = ctypes.c_double()
    lib().orc_utop_cell(_p(u), _p(bc), _p(p), ctypes.byref(e))
    return p, e.value
