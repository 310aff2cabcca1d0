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


def test_mixed_spherical_torus_runs_and_stays_close():
    """The mixed form in spherical geometry with the torus_sph hooks (its stage tails take the
    geometric sources and the model's gravity and floors) against FP64 on the same dt sequence."""
    from cumpic_amd.cumc import Solver
    print("[case] mixed vs fp64 spherical torus 24x12x16, 5 steps", flush=True)
    p = ics.torus_sph(24, 12, 16)
    a = Solver.from_problem(p, precision=FP64)
    b = Solver.from_problem(p, precision=MIXED)
    a.initial_update()
    b.initial_update()
    dts, _ = a.run(5)
    for dt in dts:
        b.rungekutta(dt)
    fa, fb = fields(a.download(), p.cfg), fields(b.download(), p.cfg)
    errs = {k: rel_l1(fb[k], fa[k]) for k in fa}
    print("spherical torus rel L1 mixed vs fp64:", errs, flush=True)
    assert errs["rho"] <= 1e-5 and errs["p"] <= 1e-4
    assert max(errs.values()) > 0.0
    _, dabs = b.find_divb()
    assert dabs <= 1e-12


def test_mixed_state_handoff():
    """In the mixed form the stage tail writes only the FP32 primitive planes the sweeps read;
    the FP64 primitives are rebuilt from the conserved state when the host asks (download,
    divB, finddt of a changed state, boundary).  A download between steps, an autotune mid-run
    (which saves and restores both sets of planes) and a boundary call must leave the run
    bit-identical to an untouched twin, and the rebuilt state stays within the mixed tolerance
    of the FP64 run."""
    from cumpic_amd.cumc import Solver
    from tests.helpers import assert_bitwise
    print("[case] mixed state hand-off, randomized 24^3 periodic box", flush=True)
    p = ics.randomize(ics.small_periodic(24, seed=2301), seed=5, amp=0.1)
    a = Solver.from_problem(p, precision=MIXED)
    c = Solver.from_problem(p, precision=MIXED)
    a.initial_update()
    c.initial_update()

    def same(what):
        sa, sc = a.download(), c.download()
        for k in ("prim", "eps", "bcell"):
            assert_bitwise(sa[k], sc[k], f"{what} {k}")

    da, _ = a.run(3)
    s1, s2 = a.download(), a.download()
    for k in ("prim", "eps", "bcell"):
        assert_bitwise(s1[k], s2[k], f"repeated download {k}")
    dc, _ = c.run(3)
    assert_bitwise(da, dc, "dt sequence")
    a.autotune()
    da, _ = a.run(3)
    dc, _ = c.run(3)
    assert_bitwise(da, dc, "dt after autotune")
    same("after autotune")
    a.boundary()
    c.boundary()
    da, _ = a.run(2)
    dc, _ = c.run(2)
    assert_bitwise(da, dc, "dt after boundary")
    same("after boundary")
    _, dabs = a.find_divb()
    assert dabs <= 1e-12, dabs
    # against the FP64 run on the same dt sequence
    ff = Solver.from_problem(p, precision=FP64)
    ff.initial_update()
    m = Solver.from_problem(p, precision=MIXED)
    m.initial_update()
    dts, _ = ff.run(6)
    for dt in dts:
        m.rungekutta(dt)
    fa, fb = fields(ff.download(), p.cfg), fields(m.download(), p.cfg)
    errs = {k: rel_l1(fb[k], fa[k]) for k in fa}
    print("rel L1 mixed vs fp64:", errs, flush=True)
    for k, e in errs.items():
        assert e <= 1e-5, (k, e)
    assert max(errs.values()) > 0.0
