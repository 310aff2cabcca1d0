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
        out = {"steps": n, "time": float(t), "divb": (float(ds), float(da)), "prim": prim, "eps": eps, "dts": dts}
        rest = f.read(12)
        if len(rest) == 12:  # the state the model plugin's GET_MODEL produced
            out["prim0"] = np.frombuffer(f.read(8 * 8 * cfg.ncell), "<f8").reshape(cfg.shape + (8,)).copy()
            out["eps0"] = np.frombuffer(f.read(8 * cfg.ncell), "<f8").reshape(cfg.shape).copy()
    return out


def write_riemann_params(path, test: int, n: int, **kw):
    """&riemann namelist of the Fortran riemann plugin (model/riemann/riemann.f90)."""
    items = {"riemann_test": test, "n": n, **kw}

    def fmt(v):
        return repr(float(v)).replace("e", "D") if isinstance(v, float) else str(int(v))
    with open(path, "w") as f:
        f.write("&riemann\n" + ",\n".join(f"  {k} = {fmt(v)}" for k, v in items.items()) + "\n/\n")


def read_model_only(path):
    """Grid and initial state written by a host run with --model-only."""
    with open(path, "rb") as f:
        nx, ny, nz = (int(v) for v in np.frombuffer(f.read(12), "<i4"))
        xf = np.frombuffer(f.read(8 * (nx + 7)), "<f8").copy()
        yf = np.frombuffer(f.read(8 * (ny + 7)), "<f8").copy()
        zf = np.frombuffer(f.read(8 * (nz + 7)), "<f8").copy()
        shape = (nz + 6, ny + 6, nx + 6)
        ncell = shape[0] * shape[1] * shape[2]
        prim = np.frombuffer(f.read(8 * 8 * ncell), "<f8").reshape(shape + (8,)).copy()
        eps = np.frombuffer(f.read(8 * ncell), "<f8").reshape(shape).copy()
    return {"nx": nx, "ny": ny, "nz": nz, "xf": xf, "yf": yf, "zf": zf, "prim": prim, "eps": eps}
