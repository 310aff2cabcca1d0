"""ctypes binding of libcumc.so, with a host-side mirror of the reference driver's calls.

The methods carry the reference names: ``initial_update`` (solver/initial.f90:83),
``finddt`` (solver/rungekutta.f90:161), ``rungekutta`` (solver/rungekutta.f90:12),
``find_divb`` (solver/mhd.f90:15) and ``run`` for the time loop of solver/main.f90:127-163.
Arrays are the reference's ghost-inclusive Fortran arrays (C shape (nz+6, ny+6, nx+6, nv)).
There is no CPU fallback: without the compiled library every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .config import Config, CConfig, to_cconfig

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcumc.so")
# diagnostic build of the same sources with branch signatures (cumc_batch_branch_sig)
LIB_SIG_PATH = os.path.join(_HERE, "libcumc_sig.so")
_libs = {}
_D = ctypes.POINTER(ctypes.c_double)
_CTX = ctypes.c_void_p


class CumcError(RuntimeError):
    pass


class CConfigP(ctypes.Structure):
    _fields_ = CConfig._fields_ + [("precision", ctypes.c_int)]


PHASES = ("flux_x", "flux_y", "flux_z", "ecorn", "tail", "ghost", "finddt")


class Tuning(ctypes.Structure):
    """cumc_tuning: launch shapes (segment length, plane chunk, persistent workgroups per phase)."""
    _fields_ = [("seg", ctypes.c_int), ("zchunk", ctypes.c_int), ("blocks", ctypes.c_int * len(PHASES)),
                ("tile_w", ctypes.c_int)]

    def as_dict(self):
        return {"seg": self.seg, "zchunk": self.zchunk, "blocks": dict(zip(PHASES, list(self.blocks))),
                "tile_w": self.tile_w}

    @classmethod
    def from_dict(cls, d):
        t = cls()
        t.seg = int(d["seg"])
        t.zchunk = int(d["zchunk"])
        b = d.get("blocks", {})
        for i, name in enumerate(PHASES):
            t.blocks[i] = int(b.get(name, 0))
        t.tile_w = int(d.get("tile_w", 0))
        return t


def lib(variant: str = ""):
    """The loaded library: libcumc.so, or with variant="sig" the diagnostic libcumc_sig.so."""
    path = LIB_SIG_PATH if variant == "sig" else LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            raise CumcError(f"{path} is not built; run __graft_entry__.build() (hipcc, gfx950)")
        L = ctypes.CDLL(path)
        sig = {
            "cumc_create": [ctypes.POINTER(CConfigP), ctypes.POINTER(_CTX)],
            "cumc_destroy": [_CTX],
            "cumc_set_grid": [_CTX, _D, _D, _D],
            "cumc_upload": [_CTX, _D, _D],
            "cumc_download": [_CTX, _D, _D, _D],
            "cumc_download_noalloc": [_CTX, _D, _D, _D],
            "cumc_initial_update": [_CTX],
            "cumc_finddt": [_CTX, _D],
            "cumc_rungekutta": [_CTX, ctypes.c_double],
            "cumc_run": [_CTX, ctypes.c_int, ctypes.c_double, _D, ctypes.POINTER(ctypes.c_int), _D],
            "cumc_find_divb": [_CTX, _D, _D],
            "cumc_boundary": [_CTX],
            "cumc_sweep_flux": [_CTX, ctypes.c_int, _D],
            "cumc_last_step_ms": [_CTX, _D],
            "cumc_profile": [_CTX, ctypes.c_int],
            "cumc_kernel_times": [_CTX, ctypes.c_int, _D, ctypes.POINTER(ctypes.c_long),
                                  ctypes.POINTER(ctypes.c_char_p)],
            "cumc_pattern_probe": [_CTX, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D],
            "cumc_set_zcollapse": [_CTX, ctypes.c_int],
            "cumc_zcollapse_state": [_CTX, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
        }
        T = ctypes.POINTER(Tuning)
        sig.update({
            "cumc_autotune": [_CTX, T],
            "cumc_get_tuning": [_CTX, T],
            "cumc_set_tuning": [_CTX, T],
            "cumc_batch_create": [ctypes.POINTER(_CTX), ctypes.c_int, ctypes.POINTER(_CTX)],
            "cumc_batch_destroy": [_CTX],
            "cumc_batch_run": [_CTX, ctypes.c_int, ctypes.c_int, _D, _D],
            "cumc_batch_autotune": [_CTX, T],
            "cumc_batch_get_tuning": [_CTX, T],
            "cumc_batch_set_tuning": [_CTX, T],
            "cumc_batch_last_ms": [_CTX, _D],
            "cumc_tape_run": [_CTX, ctypes.c_int, ctypes.c_int, _D, _D],
            "cumc_tape_backward": [_CTX, _D, _D, _D, _D, _D],
            "cumc_tape_free": [_CTX],
            "cumc_batch_tape_run": [_CTX, ctypes.c_int, ctypes.c_int, _D, _D],
            "cumc_batch_tape_backward": [_CTX, ctypes.POINTER(_D), ctypes.POINTER(_D)],
            "cumc_batch_tape_free": [_CTX],
            "cumc_batch_tape_recompute_ms": [_CTX, _D],
        })
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.cumc_batch_last_error.restype = ctypes.c_char_p
        L.cumc_batch_last_error.argtypes = [_CTX]
        L.cumc_copy_probe.argtypes = [ctypes.c_long, ctypes.c_int, _D]
        L.cumc_copy_probe.restype = ctypes.c_int
        L.cumc_error_string.restype = ctypes.c_char_p
        L.cumc_error_string.argtypes = [ctypes.c_int]
        L.cumc_last_error.restype = ctypes.c_char_p
        L.cumc_last_error.argtypes = [_CTX]
        L.cumc_batch_branch_sig.argtypes = [_CTX, ctypes.POINTER(ctypes.c_uint64)]
        L.cumc_batch_branch_sig.restype = ctypes.c_int
        _libs[path] = L
    return _libs[path]


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous float64"
    return a.ctypes.data_as(_D)


def copy_probe(nbytes: int, reps: int = 5) -> float:
    """Bandwidth (GB/s, read + write) of a plain device copy of nbytes, run reps times."""
    g = ctypes.c_double()
    r = lib().cumc_copy_probe(int(nbytes), int(reps), ctypes.byref(g))
    if r:
        raise CumcError(f"cumc_copy_probe: {lib().cumc_error_string(r).decode()}")
    return g.value


class Solver:
    """One subdomain of the CUMC3D step resident on the current GPU.  variant="sig" creates it
    in the diagnostic build libcumc_sig.so (same arithmetic, branch signatures recorded)."""

    _L = None
    h = None

    def __init__(self, cfg: Config, xf, yf, zf, variant: str = ""):
        self._L = lib(variant)
        self.variant = variant
        self.cfg = cfg
        c = CConfigP()
        base = to_cconfig(cfg)
        for name, _ in CConfig._fields_:
            setattr(c, name, getattr(base, name))
        c.precision = int(cfg.precision)
        h = _CTX()
        r = self._L.cumc_create(ctypes.byref(c), ctypes.byref(h))
        if r:
            raise CumcError(f"cumc_create: {self._L.cumc_error_string(r).decode()}")
        self.h = h
        self._check(self._L.cumc_set_grid(self.h, *(_p(np.ascontiguousarray(a, dtype=np.float64)) for a in (xf, yf, zf))),
                    "cumc_set_grid")

    @classmethod
    def from_problem(cls, prob, precision: int | None = None, variant: str = ""):
        cfg = prob.cfg
        if precision is not None:
            import dataclasses
            cfg = dataclasses.replace(cfg, precision=precision)
        s = cls(cfg, prob.xf, prob.yf, prob.zf, variant=variant)
        s.upload(prob.prim, prob.eps)
        return s

    def _check(self, r, what):
        if r:
            msg = self._L.cumc_last_error(self.h).decode() if self.h else ""
            raise CumcError(f"{what}: {self._L.cumc_error_string(r).decode()} {msg}")

    def close(self):
        if getattr(self, "h", None):
            self._L.cumc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, prim, eps):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        if prim.shape != self.cfg.shape + (8,) or eps.shape != self.cfg.shape:
            raise ValueError("prim/eps shapes do not match the configuration")
        self._check(self._L.cumc_upload(self.h, _p(prim), _p(eps)), "cumc_upload")

    def download(self):
        prim = np.empty(self.cfg.shape + (8,))
        eps = np.empty(self.cfg.shape)
        bcell = np.empty(self.cfg.shape + (3,))
        self._check(self._L.cumc_download(self.h, _p(prim), _p(eps), _p(bcell)), "cumc_download")
        return {"prim": prim, "eps": eps, "bcell": bcell}

    def initial_update(self):
        self._check(self._L.cumc_initial_update(self.h), "cumc_initial_update")

    def boundary(self):
        self._check(self._L.cumc_boundary(self.h), "cumc_boundary")

    def finddt(self) -> float:
        d = ctypes.c_double()
        self._check(self._L.cumc_finddt(self.h, ctypes.byref(d)), "cumc_finddt")
        return d.value

    def rungekutta(self, dt: float):
        self._check(self._L.cumc_rungekutta(self.h, float(dt)), "cumc_rungekutta")

    def run(self, max_steps: int, total_time: float = 1.0e300, t0: float = 0.0):
        t = ctypes.c_double(t0)
        n = ctypes.c_int(0)
        dts = np.zeros(max(1, max_steps))
        self._check(self._L.cumc_run(self.h, int(max_steps), float(total_time), ctypes.byref(t), ctypes.byref(n),
                                   _p(dts)), "cumc_run")
        return dts[: n.value].copy(), t.value

    def find_divb(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self._L.cumc_find_divb(self.h, ctypes.byref(a), ctypes.byref(b)), "cumc_find_divb")
        return a.value, b.value

    def sweep_flux(self, direction: int):
        out = np.empty(self.cfg.shape + (7,))
        self._check(self._L.cumc_sweep_flux(self.h, int(direction), _p(out)), "cumc_sweep_flux")
        return out

    def profile(self, on: bool = True):
        self._check(self._L.cumc_profile(self.h, 1 if on else 0), "cumc_profile")

    def kernel_times(self):
        """{class: (total_ms, launches)} of the HIP-event timing since profile(True)."""
        n = 16
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_long * n)()
        names = (ctypes.c_char_p * n)()
        k = self._L.cumc_kernel_times(self.h, n, ms, cnt, names)
        if k < 0:
            raise CumcError("cumc_kernel_times failed")
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(k)}

    def last_ms(self) -> float:
        d = ctypes.c_double()
        self._check(self._L.cumc_last_step_ms(self.h, ctypes.byref(d)), "cumc_last_step_ms")
        return d.value

    def autotune(self) -> dict:
        t = Tuning()
        self._check(self._L.cumc_autotune(self.h, ctypes.byref(t)), "cumc_autotune")
        return t.as_dict()

    def pattern_probe(self, nread: int, nwrite: int, reps: int = 5) -> float:
        """GB/s of the stage tail's access pattern (nread FP64 planes read, nwrite written per
        cell, no arithmetic) on this context's layout and column tiles."""
        g = ctypes.c_double()
        self._check(self._L.cumc_pattern_probe(self.h, int(nread), int(nwrite), int(reps), ctypes.byref(g)),
                    "cumc_pattern_probe")
        return g.value

    def tape_run(self, nsteps: int, every: int = 8, dts=None):
        """initial_update + nsteps steps from the uploaded state, checkpointed every `every` steps
        for tape_backward.  dts freezes the time steps.  Returns the steps used."""
        out = np.zeros(max(1, nsteps))
        din = None
        if dts is not None:
            din = np.ascontiguousarray(dts, dtype=np.float64)
            if din.shape != (nsteps,):
                raise ValueError(f"dts must have shape ({nsteps},)")
        self._check(self._L.cumc_tape_run(self.h, int(nsteps), int(every), _p(din) if din is not None else None,
                                        _p(out)), "cumc_tape_run")
        return out[:nsteps].copy()

    def tape_backward(self, seed_prim=None, seed_eps=None, seed_bcell=None):
        """Gradients of <seed, final state> with respect to the uploaded prim and eps."""
        shp = self.cfg.shape

        def arr(a, nv):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            want = shp + ((nv,) if nv > 1 else ())
            if a.shape != want:
                raise ValueError(f"seed shape {a.shape} != {want}")
            return a

        sp, se, sb = arr(seed_prim, 8), arr(seed_eps, 1), arr(seed_bcell, 3)
        gp = np.zeros(shp + (8,))
        ge = np.zeros(shp)
        self._check(self._L.cumc_tape_backward(self.h, _p(sp) if sp is not None else None,
                                             _p(se) if se is not None else None,
                                             _p(sb) if sb is not None else None, _p(gp), _p(ge)),
                    "cumc_tape_backward")
        return gp, ge

    def tape_free(self):
        self._check(self._L.cumc_tape_free(self.h), "cumc_tape_free")

    def tuning(self) -> dict:
        t = Tuning()
        self._check(self._L.cumc_get_tuning(self.h, ctypes.byref(t)), "cumc_get_tuning")
        return t.as_dict()

    def set_tuning(self, d: dict):
        t = Tuning.from_dict(d)
        self._check(self._L.cumc_set_tuning(self.h, ctypes.byref(t)), "cumc_set_tuning")

    def set_zcollapse(self, on: bool):
        """z-collapsed stepping of a periodic nz = 1 context (on by default; off keeps the full
        layout).  Results are bit-identical either way."""
        self._check(self._L.cumc_set_zcollapse(self.h, int(bool(on))), "cumc_set_zcollapse")

    def zcollapse_state(self) -> tuple[bool, bool]:
        """(allowed, active): whether the context may step collapsed, and whether it is now."""
        a, b = ctypes.c_int(0), ctypes.c_int(0)
        self._check(self._L.cumc_zcollapse_state(self.h, ctypes.byref(a), ctypes.byref(b)), "cumc_zcollapse_state")
        return bool(a.value), bool(b.value)


class Batch:
    """Independent boxes advanced together: one persistent launch per phase over all boxes,
    optionally replayed from a hipGraph (cumc_batch_*).  The boxes stay ordinary Solvers."""

    _L = None
    h = None

    def __init__(self, solvers):
        self.solvers = list(solvers)
        if len({s.variant for s in self.solvers}) != 1:
            raise ValueError("the boxes of a batch must come from one library build")
        self._L = self.solvers[0]._L
        arr = (_CTX * len(self.solvers))(*[s.h for s in self.solvers])
        h = _CTX()
        r = self._L.cumc_batch_create(arr, len(self.solvers), ctypes.byref(h))
        if r:
            raise CumcError(f"cumc_batch_create: {self._L.cumc_error_string(r).decode()}")
        self.h = h

    def _check(self, r, what):
        if r:
            msg = self._L.cumc_batch_last_error(self.h).decode() if self.h else ""
            raise CumcError(f"{what}: {self._L.cumc_error_string(r).decode()} {msg}")

    def close(self):
        if getattr(self, "h", None):
            self._L.cumc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps: int, graph_steps: int = 0, dts=None):
        """nsteps steps of every box; dts (nsteps, nbox) freezes the time steps.  Returns the
        (nsteps, nbox) array of steps used."""
        nb = len(self.solvers)
        out = np.zeros((max(1, nsteps), nb))
        din = None
        if dts is not None:
            din = np.ascontiguousarray(dts, dtype=np.float64)
            if din.shape != (nsteps, nb):
                raise ValueError(f"dts must have shape ({nsteps}, {nb})")
        self._check(self._L.cumc_batch_run(self.h, int(nsteps), int(graph_steps), _p(din) if din is not None else None,
                                         _p(out)), "cumc_batch_run")
        return out[:nsteps].copy()

    def autotune(self) -> dict:
        t = Tuning()
        self._check(self._L.cumc_batch_autotune(self.h, ctypes.byref(t)), "cumc_batch_autotune")
        return t.as_dict()

    def tuning(self) -> dict:
        t = Tuning()
        self._check(self._L.cumc_batch_get_tuning(self.h, ctypes.byref(t)), "cumc_batch_get_tuning")
        return t.as_dict()

    def set_tuning(self, d: dict):
        t = Tuning.from_dict(d)
        self._check(self._L.cumc_batch_set_tuning(self.h, ctypes.byref(t)), "cumc_batch_set_tuning")

    def last_ms(self) -> float:
        d = ctypes.c_double()
        self._check(self._L.cumc_batch_last_ms(self.h, ctypes.byref(d)), "cumc_batch_last_ms")
        return d.value

    def tape_run(self, nsteps: int, every: int = 8, dts=None):
        """initial_update + nsteps steps of every box from its uploaded state, checkpointed every
        `every` steps.  dts (nsteps, nbox) freezes the steps.  Returns the (nsteps, nbox) steps."""
        nb = len(self.solvers)
        out = np.zeros((max(1, nsteps), nb))
        din = None
        if dts is not None:
            din = np.ascontiguousarray(dts, dtype=np.float64)
            if din.shape != (nsteps, nb):
                raise ValueError(f"dts must have shape ({nsteps}, {nb})")
        self._check(self._L.cumc_batch_tape_run(self.h, int(nsteps), int(every), _p(din) if din is not None else None,
                                              _p(out)), "cumc_batch_tape_run")
        return out[:nsteps].copy()

    def tape_backward(self, seeds_bcell):
        """Per-box gradients (uploaded prim layout) of sum_b <seed_b, final bcell_b>."""
        nb = len(self.solvers)
        seeds = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in seeds_bcell]
        for s, a in zip(self.solvers, seeds):
            if a is not None and a.shape != s.cfg.shape + (3,):
                raise ValueError("seed shape does not match its box")
        grads = [np.zeros(s.cfg.shape + (8,)) for s in self.solvers]
        sp = (_D * nb)(*[_p(a) if a is not None else None for a in seeds])
        gp = (_D * nb)(*[_p(g) for g in grads])
        self._check(self._L.cumc_batch_tape_backward(self.h, sp, gp), "cumc_batch_tape_backward")
        return grads

    def tape_free(self):
        self._check(self._L.cumc_batch_tape_free(self.h), "cumc_batch_tape_free")

    def branch_sig(self):
        """Per-box branch signatures (uint64) accumulated by the forward kernels since the last
        call, which clears them; only the diagnostic build records them."""
        out = (ctypes.c_uint64 * len(self.solvers))()
        self._check(self._L.cumc_batch_branch_sig(self.h, out), "cumc_batch_branch_sig")
        return np.array(list(out), dtype=np.uint64)

    def tape_recompute_ms(self) -> float:
        """Device ms of the last backward pass spent re-running checkpoint segments forward."""
        d = ctypes.c_double()
        self._check(self._L.cumc_batch_tape_recompute_ms(self.h, ctypes.byref(d)), "cumc_batch_tape_recompute_ms")
        return d.value
