"""Mixed-precision form (FP32 reconstruction, HLL and EMFs; FP64 state, RK3, U->P and face B)
against the FP64 path on the same inputs and the same dt sequence.

Tolerance: relative L1 of rho, p, |v| and |B| <= 1e-5 after the run, and the area-weighted
|div B| stays at FP64 round-off (<= 1e-12), because the face-B update is FP64 CT."""
from __future__ import annotations

import numpy as np
import pytest

from cumpic_amd import ics
from cumpic_amd.config import FP64, MIXED
from tests.helpers import interior, rel_l1

pytestmark = pytest.mark.gpu


def fields(state, cfg):
    p = interior(state["prim"], cfg)
    b = interior(state["bcell"], cfg)
    return {"rho": p[..., 0], "p": p[..., 4], "|v|": np.linalg.norm(p[..., 1:4], axis=-1),
            "|B|": np.linalg.norm(b, axis=-1)}


@pytest.mark.parametrize("n,steps", [(32, 20)])
def test_mixed_vs_fp64(n, steps):
    from cumpic_amd.cumc import Solver
    print(f"[case] mixed vs fp64 spectral box n={n} steps={steps}", flush=True)
    p = ics.spectral_box(n=n, seed=1002, steps=steps)
    a = Solver.from_problem(p, precision=FP64)
    b = Solver.from_problem(p, precision=MIXED)
    a.initial_update()
    b.initial_update()
    dts, _ = a.run(steps)
    for dt in dts:
        b.rungekutta(dt)
    fa, fb = fields(a.download(), p.cfg), fields(b.download(), p.cfg)
    errs = {k: rel_l1(fb[k], fa[k]) for k in fa}
    print("rel L1 mixed vs fp64:", errs, flush=True)
    for k, e in errs.items():
        assert e <= 1e-5, (k, e)
    assert max(errs.values()) > 0.0  # the FP32 stages really ran
    _, dabs = b.find_divb()
    assert dabs <= 1e-12, dabs


def test_mixed_torus_runs_and_stays_close():
    from cumpic_amd.cumc import Solver
    print("[case] mixed vs fp64 torus 24x12x24, 5 steps", flush=True)
    p = ics.torus_cyn(24, 12, 24)
    a = Solver.from_problem(p, precision=FP64)
    b = Solver.from_problem(p, precision=MIXED)
    a.initial_update()
    b.initial_update()
    dts, _ = a.run(5)
    for dt in dts:
        b.rungekutta(dt)
    fa, fb = fields(a.download(), p.cfg), fields(b.download(), p.cfg)
    errs = {k: rel_l1(fb[k], fa[k]) for k in fa}
    print("torus rel L1 mixed vs fp64:", errs, flush=True)
    assert errs["rho"] <= 1e-5 and errs["p"] <= 1e-4
    _, dabs = b.find_divb()
    assert dabs <= 1e-12
