"""The Fortran host driver (cumpic_amd/fhost/cumc3d) runs the 1D Riemann workload and the riemann
plugin's named problems (Toro 1, Brio-Wu, the 2D MHD rotor) through the C ABI on the GPU; its
final state and dt sequence must equal the CPU restatement bit for bit."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from cumpic_amd import build, ics
from cumpic_amd.fhost.runfile import read_output, write_runfile
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _named(name):
    if name == "riemann_1d":
        return ics.riemann_1d()
    n = {"toro1": 200, "briowu": 200, "mhdrotor": 48}[name]
    return ics.riemann_test({"toro1": ics.TORO_1, "briowu": ics.BRIO_WU, "mhdrotor": ics.MHD_ROTOR}[name], n=n)


@pytest.mark.parametrize("name", ["riemann_1d", "toro1", "briowu", "mhdrotor"])
def test_fortran_host_riemann_1d(tmp_path, name):
    print(f"[case] Fortran host, {name} to its stop time", flush=True)
    exe = build.FHOST_EXE
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built (no flang)")
    p = _named(name)
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


def test_fortran_host_hdf5_profiles(tmp_path):
    """HDF5 profiles written by the host at the reference's points (main.f90:107-110, 150-160,
    165-168): after initial_update as profile 0, whenever |t - t_last| >= output_profiletime
    (n_iter advancing), and at the end under the next n_iter; each equal bit for bit to the CPU
    restatement's state at that step, in the dataio.f90 layout."""
    from cumpic_amd import dataio
    print("[case] Fortran host HDF5 profiles, 1D Riemann, output every 0.03", flush=True)
    exe = build.FHOST_EXE
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built (no flang)")
    try:
        dataio._hdf5()
    except RuntimeError:
        pytest.skip("the HDF5 C library is not available")
    p = ics.riemann_1d()
    out_time = 0.03
    run, out, prof = tmp_path / "c0.run", tmp_path / "c0.out", tmp_path / "outfile"
    prof.mkdir()
    write_runfile(str(run), p, max_steps=100000, total_time=p.total_time, output_time=out_time)
    r = subprocess.run([exe, str(run), str(out), str(prof)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    res = read_output(str(out), p.cfg)
    o = Oracle.from_problem(p)
    o.initial_update()
    expect = {0: o.get_state()}
    n_iter, t, last = 1, 0.0, 0.0
    for dt in res["dts"]:
        o.rungekutta(float(dt))
        t += float(dt)
        if abs(t - last) >= out_time:
            last = t
            expect[n_iter] = o.get_state()
            n_iter += 1
    expect[n_iter] = o.get_state()
    assert len(expect) >= 4
    assert sorted(os.listdir(prof)) == sorted(dataio.profile_name(k) for k in expect)
    c = p.cfg
    inner = (slice(3, 3 + c.nz), slice(3, 3 + c.ny), slice(3, 3 + c.nx))
    for k, st in expect.items():
        got = dataio.read_profile(str(prof / dataio.profile_name(k)))
        assert set(got) == {"xf", "prim", "eps"}
        assert np.array_equal(got["prim"], st["prim"][inner]), k
        assert np.array_equal(got["eps"], st["eps"][inner]), k
        assert np.array_equal(got["xf"], p.xf[3:3 + c.nx + 1])


@pytest.mark.parametrize("name,test,n", [("toro1", ics.TORO_1, 200), ("briowu", ics.BRIO_WU, 200),
                                         ("mhdrotor", ics.MHD_ROTOR, 48)])
def test_fortran_riemann_plugin_named_problems(tmp_path, name, test, n):
    """The host built with the Fortran riemann plugin (model/riemann/riemann.f90: MODEL_SETUP from a
    &riemann namelist, GETGRID + CUSTOM_GRID, GET_MODEL computing the problem, OPENFILE_CUSTOM) runs
    the problem to its stop time with no Python initial condition; the CPU restatement, started from
    the state that host's GET_MODEL produced, must reach the same final state and dt sequence bit for
    bit (and that state is the Python generator's)."""
    from cumpic_amd.fhost.runfile import write_riemann_params
    print(f"[case] Fortran riemann plugin, {name} n={n} to its stop time", flush=True)
    exe = build.fhost_exe("riemann")
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built (no flang)")
    par, out = tmp_path / "r.nml", tmp_path / "r.out"
    write_riemann_params(str(par), test, n)
    r = subprocess.run([exe, str(par), str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    ref = ics.riemann_test(test, n=n)
    res = read_output(str(out), ref.cfg)
    assert np.array_equal(res["prim0"], ref.prim) and np.array_equal(res["eps0"], ref.eps)
    ref.prim, ref.eps = res["prim0"], res["eps0"]
    o = Oracle.from_problem(ref)
    o.initial_update()
    dts, t = o.run(100000, ref.total_time)
    so = o.get_state()
    print(f"  {res['steps']} steps to t = {res['time']}", flush=True)
    assert res["steps"] == len(dts) and res["time"] == t
    assert np.array_equal(res["dts"], dts)
    assert np.array_equal(res["prim"], so["prim"])
    assert np.array_equal(res["eps"], so["eps"])
