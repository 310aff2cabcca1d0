"""Run configuration and grid faces for the CUMC3D time step.

The reference fixes these as compile-time Fortran PARAMETERs
(solver/definition.f90:29-185, model/<m>/src/defs.f90); here they are a runtime
record that the C ABI receives as ``cumc_config`` (include/cumc.h).
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

import numpy as np

NGHOST = 3  # solver/definition.f90:54
NV = 8  # rho, vx, vy, vz, p, bx, by, bz  (solver/setup.f90:49-116)
IRHO, IVX, IVY, IVZ, ITAU, IBX, IBY, IBZ = range(8)

CARTESIAN, CYLINDRICAL, SPHERICAL = 1, 2, 3  # definition.f90:65-67
TVDMM, PPMC = 1, 4  # definition.f90:158,167
TVDMC, TVDVL, WENO = 2, 3, 5  # definition.f90:161,164,170: declared there, implemented here as extensions
LF, HLL, HLLC, HLLD = 1, 2, 3, 4  # definition.f90:140-149: HLL is the reference's; the others extensions
MODEL_RIEMANN, MODEL_TORUS_CYN, MODEL_TORUS_SPH = 0, 1, 2
FP64, MIXED = 0, 1


@dataclasses.dataclass
class Config:
    nx: int
    ny: int
    nz: int
    coordinate: int = CARTESIAN
    boundary_flag: tuple = (0, 0, 0, 0, 0, 0)
    recon: int = PPMC
    model: int = MODEL_RIEMANN
    geom_source_repeat: int = 8
    ggas: float = 5.0 / 3.0
    cfl: float = 0.7
    rho_floor: float = 0.0
    eps_floor: float = 0.0
    r_sh: float = 1.0
    precision: int = FP64
    solver: int = HLL

    @property
    def shape(self):
        """C-order shape of one ghost-inclusive cell array: (nz+6, ny+6, nx+6)."""
        return (self.nz + 2 * NGHOST, self.ny + 2 * NGHOST, self.nx + 2 * NGHOST)

    @property
    def ncell(self):
        s = self.shape
        return s[0] * s[1] * s[2]

    @property
    def zones(self):
        return self.nx * self.ny * self.nz


class CConfig(ctypes.Structure):
    """Mirror of ``cumc_config`` in include/cumc.h (and ``orc_config`` of the oracle)."""

    _fields_ = [
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("coordinate", ctypes.c_int),
        ("boundary_flag", ctypes.c_int * 6),
        ("recon", ctypes.c_int),
        ("model", ctypes.c_int),
        ("geom_source_repeat", ctypes.c_int),
        ("ggas", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("rho_floor", ctypes.c_double),
        ("eps_floor", ctypes.c_double),
        ("r_sh", ctypes.c_double),
        ("solver", ctypes.c_int),
    ]


def to_cconfig(cfg: Config) -> CConfig:
    c = CConfig()
    c.nx, c.ny, c.nz = cfg.nx, cfg.ny, cfg.nz
    c.coordinate = cfg.coordinate
    for i in range(6):
        c.boundary_flag[i] = int(cfg.boundary_flag[i])
    c.recon = cfg.recon
    c.model = cfg.model
    c.geom_source_repeat = cfg.geom_source_repeat
    c.ggas = cfg.ggas
    c.cfl = cfg.cfl
    c.rho_floor = cfg.rho_floor
    c.eps_floor = cfg.eps_floor
    c.r_sh = cfg.r_sh
    c.solver = cfg.solver
    return c


def uniform_faces(n: int, start: float, end: float, offset: int = 0) -> np.ndarray:
    """Face coordinates xF(-3:n+3) exactly as solver/coords.f90:24-32 builds them:
    xF(j) = start + (offset + j) * dx with dx = (end - start) / n."""
    dx = (end - start) / float(n)
    j = np.arange(-NGHOST, n + NGHOST + 1, dtype=np.float64)
    return start + (float(offset) + j) * dx


def log_faces(n: int, r0: float, r1: float) -> np.ndarray:
    """Log-spaced radial faces xF(j) = r0 * (r1/r0)**(j/n), j = -3..n+3."""
    j = np.arange(-NGHOST, n + NGHOST + 1, dtype=np.float64)
    return r0 * np.power(r1 / r0, j / float(n))


def periodic_angle_faces(n: int) -> np.ndarray:
    return uniform_faces(n, 0.0, 2.0 * math.pi)
