"""cumc_download_noalloc (no device allocation: staging through the library's fixed per-GPU
buffer) against cumc_download, through the C ABI: bit-identical host arrays on two calls in a
row, the first on a state larger than one staging slab (several slabs of whole planes), the
second on a smaller mixed-form state (whose FP64 planes are rebuilt first); and a grid whose
single plane exceeds the buffer is refused."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from cumpic_amd import ics

pytestmark = pytest.mark.gpu


def _both(s):
    from cumpic_amd.cumc import _p, lib
    shp = s.cfg.shape
    ref = s.download()
    got = {"prim": np.full(shp + (8,), -7.0), "eps": np.full(shp, -7.0), "bcell": np.full(shp + (3,), -7.0)}
    r = lib().cumc_download_noalloc(s.h, _p(got["prim"]), _p(got["eps"]), _p(got["bcell"]))
    assert r == 0, lib().cumc_last_error(s.h).decode()
    return ref, got


def test_download_noalloc_matches_download():
    from cumpic_amd.config import MIXED
    from cumpic_amd.cumc import Solver
    # 200^3: one prim plane is 206*206*8 doubles (2.6 MiB), 12 planes per 32 MiB slab, 18 slabs
    big = ics.randomize(ics.spectral_box(n=200, seed=1301, v_rms=0.3, db_rms=0.3, steps=2), seed=1)
    print(f"[case] noalloc download, first call {big.cfg.nx}x{big.cfg.ny}x{big.cfg.nz} FP64 after one step", flush=True)
    s = Solver.from_problem(big)
    s.initial_update()
    s.run(1)
    ref, got = _both(s)
    for k in ref:
        assert np.array_equal(ref[k], got[k]), k
    s.close()
    small = ics.randomize(ics.small_periodic(10, seed=1302), seed=2)
    print("[case] noalloc download, second call 10^3 mixed form after two steps", flush=True)
    m = Solver.from_problem(small, precision=MIXED)
    m.initial_update()
    m.run(2)
    ref, got = _both(m)
    for k in ref:
        assert np.array_equal(ref[k], got[k]), k
    m.close()


def test_download_noalloc_refuses_a_plane_larger_than_the_buffer():
    from cumpic_amd.cumc import Solver, _p, lib
    p = ics.small_periodic(8, seed=1303)
    p.cfg.nx, p.cfg.ny, p.cfg.nz = 730, 730, 1
    from cumpic_amd.config import uniform_faces
    p.xf, p.yf, p.zf = uniform_faces(730, 0, 1), uniform_faces(730, 0, 1), uniform_faces(1, 0, 1)
    print("[case] noalloc download refuses a 730x730 plane with prim", flush=True)
    s = Solver(p.cfg, p.xf, p.yf, p.zf)
    shp = p.cfg.shape
    prim = np.zeros(shp + (8,))
    eps = np.zeros(shp)
    assert lib().cumc_download_noalloc(s.h, _p(prim), None, None) == 1
    assert "staging" in lib().cumc_last_error(s.h).decode()
    # eps alone (one double per cell) fits
    assert lib().cumc_download_noalloc(s.h, None, _p(eps), None) == 0
    s.close()
