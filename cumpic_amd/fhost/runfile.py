"""Writer/reader of the Fortran host's run and output files (format in model_file.f90)."""
from __future__ import annotations

import numpy as np


def write_runfile(path, prob, max_steps: int, total_time: float, output_time: float = 100.0):
    """output_time: simulated time between HDF5 profiles (the reference's output_profiletime,
    definition.f90:132, 100.0 there); profiles are written only when the host gets a directory."""
    c = prob.cfg
    hdr = np.array([c.nx, c.ny, c.nz, c.coordinate, *c.boundary_flag, c.recon, c.model, c.geom_source_repeat,
                    c.precision, max_steps, c.solver], dtype="<i4")
    rh = np.array([c.ggas, c.cfl, c.rho_floor, c.eps_floor, c.r_sh, total_time, output_time], dtype="<f8")
    with open(path, "wb") as f:
        for a in (hdr, rh, prob.xf, prob.yf, prob.zf, prob.prim, prob.eps):
            f.write(np.ascontiguousarray(a, dtype=a.dtype if a is hdr else "<f8").tobytes())


def read_output(path, cfg):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(4), "<i4")[0])
        t, ds, da = np.frombuffer(f.read(24), "<f8")
        prim = np.frombuffer(f.read(8 * 8 * cfg.ncell), "<f8").reshape(cfg.shape + (8,)).copy()
        eps = np.frombuffer(f.read(8 * cfg.ncell), "<f8").reshape(cfg.shape).copy()
        dts = np.frombuffer(f.read(8 * n), "<f8").copy()
    return {"steps": n, "time": float(t), "divb": (float(ds), float(da)), "prim": prim, "eps": eps, "dts": dts}
