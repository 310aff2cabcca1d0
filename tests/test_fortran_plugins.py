"""The Fortran host's model plugins, checked without a GPU (--model-only stops before any device
call): the riemann plugin's GET_MODEL computes Toro 1, Brio-Wu and the MHD rotor in Fortran
(model/riemann/riemann.f90, after the reference's riemann1d.f90:15-89 and riemann2d.f90:15-63),
GETGRID's faces and the filled state must equal the Python generators (ics.riemann_test, pinned in
tests/test_riemann_plugin.py against exact solutions and budgets) bit for bit; the run-file plugin
hands its file's grid and state through unchanged; a bad selector is refused."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from cumpic_amd import build, ics
from cumpic_amd.fhost.runfile import read_model_only, write_riemann_params, write_runfile


def _exe(model):
    exe = build.fhost_exe(model)
    if not os.path.exists(exe):
        if build.flang() is None:
            pytest.skip("no Fortran compiler")
        build.build_fortran_host()
    return exe


@pytest.mark.parametrize("test,n", [(ics.TORO_1, 64), (ics.BRIO_WU, 50), (ics.MHD_ROTOR, 40)])
def test_riemann_plugin_get_model_matches_generators(tmp_path, test, n):
    par, out = tmp_path / "p.nml", tmp_path / "m.out"
    write_riemann_params(str(par), test, n)
    r = subprocess.run([_exe("riemann"), "--model-only", str(par), str(out)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    m = read_model_only(str(out))
    p = ics.riemann_test(test, n=n)
    assert (m["nx"], m["ny"], m["nz"]) == (p.cfg.nx, p.cfg.ny, p.cfg.nz)
    for k in ("xf", "yf", "zf", "prim", "eps"):
        assert np.array_equal(m[k], getattr(p, k)), k


def test_riemann_plugin_refuses_an_unknown_test(tmp_path):
    par, out = tmp_path / "p.nml", tmp_path / "m.out"
    write_riemann_params(str(par), 7, 16)
    r = subprocess.run([_exe("riemann"), "--model-only", str(par), str(out)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode != 0 and "riemann_test" in r.stdout


def test_file_plugin_passes_the_run_file_through(tmp_path):
    p = ics.torus_cyn(12, 6, 12)
    run, out = tmp_path / "t.run", tmp_path / "m.out"
    write_runfile(str(run), p, max_steps=3, total_time=1.0)
    r = subprocess.run([_exe("file"), "--model-only", str(run), str(out)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    m = read_model_only(str(out))
    for k in ("xf", "yf", "zf", "prim", "eps"):
        assert np.array_equal(m[k], getattr(p, k)), k
