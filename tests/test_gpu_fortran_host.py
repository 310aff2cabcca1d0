"""The Fortran host driver (cumpic_amd/fhost/cumc3d) runs the 1D Riemann workload through the C
ABI on the GPU; its final state and dt sequence must equal the CPU restatement bit for bit."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from cumpic_amd import build, ics
from cumpic_amd.fhost.runfile import read_output, write_runfile
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def test_fortran_host_riemann_1d(tmp_path):
    print("[case] Fortran host, 1D Riemann 800 cells to t=0.1", flush=True)
    exe = build.FHOST_EXE
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built (no flang)")
    p = ics.riemann_1d()
    run = tmp_path / "c0.run"
    out = tmp_path / "c0.out"
    write_runfile(str(run), p, max_steps=100000, total_time=p.total_time)
    r = subprocess.run([exe, str(run), str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    res = read_output(str(out), p.cfg)
    o = Oracle.from_problem(p)
    o.initial_update()
    dts, t = o.run(100000, p.total_time)
    so = o.get_state()
    assert res["steps"] == len(dts) and res["time"] == t
    assert np.array_equal(res["dts"], dts)
    assert np.array_equal(res["prim"], so["prim"])
    assert np.array_equal(res["eps"], so["eps"])
