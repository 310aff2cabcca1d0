"""CPU-side checks of the boundary: libcumc.so loads and exports every entry point that
include/cumc.h declares, the ctypes config mirrors the C struct, and the Fortran host links."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

from cumpic_amd import build
from cumpic_amd.config import CConfig

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"\b(cumc_[a-z_0-9]+)\s*\(", txt))
    return sorted(names)


@pytest.mark.parametrize("variant", ["", "sig"])
def test_library_exports_every_declared_symbol(variant):
    path = build.LIB_SIG if variant == "sig" else build.LIB
    if not os.path.exists(path):
        build.build_library(variant=variant)
    lib = ctypes.CDLL(path)
    names = declared_symbols()
    assert "cumc_rungekutta" in names and "cumc_finddt" in names and "cumc_batch_branch_sig" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_config_struct_layout_matches_header():
    from cumpic_amd.cumc import CConfigP
    # 3+1+6+3 ints (52 B) + pad + 5 doubles + solver + precision
    assert ctypes.sizeof(CConfigP) == 104
    assert CConfigP.ggas.offset == 56 and CConfigP.solver.offset == 96 and CConfigP.precision.offset == 100
    assert ctypes.sizeof(CConfig) == 104  # padded to 8 after solver


def test_error_strings_without_gpu():
    lib = ctypes.CDLL(build.LIB)
    lib.cumc_error_string.restype = ctypes.c_char_p
    assert lib.cumc_error_string(0) == b"ok"
    assert lib.cumc_error_string(3) == b"non-finite time step"


def test_fortran_host_links():
    if build.flang() is None:
        pytest.skip("no Fortran compiler in this image")
    exe = build.build_fortran_host()
    assert os.path.exists(exe)
    out = subprocess.run(["nm", "-D", "--undefined-only", exe], stdout=subprocess.PIPE, text=True).stdout
    for sym in ("cumc_create", "cumc_finddt", "cumc_rungekutta", "cumc_initial_update", "cumc_download"):
        assert sym in out


def test_tuning_struct_layout_matches_header():
    from cumpic_amd.cumc import PHASES, Tuning
    txt = open(os.path.join(ROOT, "include", "cumc.h")).read()
    enum = re.search(r"enum \{ CUMC_PH_FLUX_X = 0,([^}]*)\}", txt).group(0)
    assert enum.count("CUMC_PH_") - 1 == len(PHASES)  # CUMC_PH_N closes the list
    assert ctypes.sizeof(Tuning) == 4 * (3 + len(PHASES))
    t = Tuning.from_dict({"seg": 5, "zchunk": 6, "blocks": {"tail": 9}})
    assert t.as_dict()["blocks"]["tail"] == 9 and t.as_dict()["seg"] == 5


def test_batch_boxes_sizes_are_seeded():
    from cumpic_amd import ics
    a = [p.cfg.nx for p in ics.batch_boxes(nbox=6, sizes=(8, 12))]
    b = [p.cfg.nx for p in ics.batch_boxes(nbox=6, sizes=(8, 12))]
    assert a == b and set(a) <= {8, 12}
    p = ics.batch_boxes(nbox=1, sizes=(8,))[0]
    assert p.cfg.nz == 1 and p.cfg.ny == 8 and p.prim.shape == (7, 14, 14, 8)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """The product path has no CPU fallback: without the built HIP library every entry point
    raises instead of computing anything."""
    from cumpic_amd import cumc, ics
    monkeypatch.setattr(cumc, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(cumc, "_libs", {})
    with pytest.raises(cumc.CumcError, match="not built"):
        cumc.Solver.from_problem(ics.small_periodic(8))


def test_create_rejects_bad_configurations_before_any_hip_call():
    """cumc_create validates the whole configuration on the host before it touches a device, so
    every rejection is CUMC_EINVAL (1) with *out left NULL, here without a GPU: sizes (including
    a plane of 2^29 cells or more and dimensions whose padded product overflows 32 bits), the
    option enums, gamma and the CFL number, the boundary flags and the symmetry walls the
    reference stops on (setup.f90:147-151, :206-210)."""
    from cumpic_amd import ics
    from cumpic_amd.config import to_cconfig
    from cumpic_amd.cumc import CConfigP

    lib = ctypes.CDLL(build.LIB)
    lib.cumc_create.argtypes = [ctypes.POINTER(CConfigP), ctypes.POINTER(ctypes.c_void_p)]
    base = to_cconfig(ics.small_periodic(8).cfg)

    def create(**changes):
        c = CConfigP()
        for name, _ in CConfig._fields_:
            setattr(c, name, getattr(base, name))
        c.precision = 0
        for k, v in changes.items():
            if k.startswith("bf"):
                c.boundary_flag[int(k[2:])] = v
            else:
                setattr(c, k, v)
        out = ctypes.c_void_p(12345)
        r = lib.cumc_create(ctypes.byref(c), ctypes.byref(out))
        return r, out.value

    bad = [dict(nx=0), dict(ny=-1), dict(nz=0), dict(coordinate=7), dict(recon=99), dict(solver=99),
           dict(model=99), dict(precision=9), dict(ggas=1.0), dict(ggas=float("nan")), dict(cfl=0.0),
           dict(cfl=float("nan")), dict(bf0=5), dict(bf3=-1), dict(nz=2, bf4=2), dict(bf0=4), dict(bf1=4),
           dict(bf5=3), dict(nx=812, ny=812, nz=812), dict(nx=2**29), dict(nx=2**31 - 1, ny=2**31 - 1),
           dict(nx=60000, ny=60000, nz=1)]
    for changes in bad:
        r, out = create(**changes)
        assert r == 1, (changes, r)
        assert not out, changes
    assert lib.cumc_create(None, None) == 1
