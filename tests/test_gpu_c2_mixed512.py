"""The mixed-precision workload at its stated size: the 512^3 periodic spectral box (the 256^3
box's seeded spectrum, seed 1001), 100 FP64 steps, then the mixed form (FP32 reconstruction, HLL
and EMFs; FP64 RK3 state, pressure recovery and face-B update) from the same initial state on the
FP64 run's dt sequence.  Tolerance: relative L1 <= 1e-5 of rho, p, |v| and |B| against FP64, and
max |div B| <= 1e-12 for both forms (area-weighted face sum per cell, solver/mhd.f90:41-50).

Memory: one context at a time (about 67 GB of HBM each), host copies of the two final states."""
from __future__ import annotations

import time

import numpy as np
import pytest

from cumpic_amd import ics

pytestmark = pytest.mark.gpu

N, STEPS = 512, 100


def _fields(st, c):
    sl = (slice(3, 3 + c.nz), slice(3, 3 + c.ny), slice(3, 3 + c.nx))
    p, b = st["prim"][sl], st["bcell"][sl]
    return {"rho": p[..., 0].copy(), "p": p[..., 4].copy(), "|v|": np.linalg.norm(p[..., 1:4], axis=-1),
            "|B|": np.linalg.norm(b, axis=-1)}


@pytest.mark.timeout(1500)
def test_c2_mixed_512_against_fp64():
    from cumpic_amd.config import FP64, MIXED
    from cumpic_amd.cumc import Solver
    print(f"[case] C2: {N}^3 spectral box seed 1001, {STEPS} steps, FP64 then mixed on its dt sequence", flush=True)
    t0 = time.perf_counter()
    prob = ics.spectral_box(n=N, seed=1001, steps=STEPS)
    print(f"  initial condition {time.perf_counter() - t0:.1f} s", flush=True)
    c = prob.cfg
    s = Solver.from_problem(prob, precision=FP64)
    s.initial_update()
    dts, _ = s.run(STEPS)
    _, abs64 = s.find_divb()
    ref = _fields(s.download(), c)
    s.close()
    print(f"  FP64: {STEPS} steps, dt {dts[0]:.4e} .. {dts[-1]:.4e}, max |div B| {abs64:.2e}", flush=True)
    assert np.all(np.isfinite(dts)) and np.all(dts > 0)
    m = Solver.from_problem(prob, precision=MIXED)
    m.initial_update()
    for i, dt in enumerate(dts):
        m.rungekutta(float(dt))
        if i % 25 == 24:
            print(f"  mixed: step {i + 1}", flush=True)
    _, absm = m.find_divb()
    got = _fields(m.download(), c)
    m.close()
    errs = {k: float(np.sum(np.abs(got[k] - ref[k])) / np.sum(np.abs(ref[k]))) for k in ref}
    print(f"  mixed rel L1 vs FP64: {errs}; max |div B| mixed {absm:.2e}", flush=True)
    for k, e in errs.items():
        assert e <= 1e-5, (k, e)
    assert abs64 <= 1e-12 and absm <= 1e-12
