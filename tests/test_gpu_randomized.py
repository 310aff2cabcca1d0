"""A seeded sweep over the option space: geometry and model hooks (Cartesian riemann box,
cylindrical torus_cyn, spherical shell, spherical torus_sph), reconstruction (the reference's
TVD_MM and PPM and the extensions), Riemann solver, boundary flags and the geometric-source
repetition, each run a few steps on the GPU (FP64) and on the CPU restatement from the same
seeded state: every configuration must agree bit for bit, dt sequence included."""
from __future__ import annotations

import numpy as np
import pytest

from cumpic_amd import ics
from cumpic_amd.config import HLL, HLLC, HLLD, LF, PPMC, TVDMC, TVDMM, TVDVL, WENO
from oracle.oracle import Oracle
from tests.helpers import assert_bitwise

pytestmark = pytest.mark.gpu

RECONS = [TVDMM, TVDMC, TVDVL, PPMC, WENO]
SOLVERS = [HLL, LF, HLLC, HLLD]


def _config(i):
    rng = np.random.default_rng(4000 + i)
    kind = ["box", "torus_cyn", "sphere", "torus_sph"][i % 4]
    if kind == "box":
        p = ics.randomize(ics.small_periodic(int(rng.integers(6, 11)), seed=4100 + i), seed=i, amp=0.15)
        # x flags 0..3, y 0..4, z 0, 1, 2, 4 (the combinations the reference accepts)
        fx = [int(v) for v in rng.integers(0, 4, 2)]
        fy = [int(v) for v in rng.integers(0, 5, 2)]
        fz = [int(rng.choice([0, 1, 2, 4])) for _ in range(2)]
        p.cfg.boundary_flag = tuple(fx + fy + fz)
    elif kind == "torus_cyn":
        p = ics.torus_cyn(16, 6, 16)
    elif kind == "sphere":
        p = ics.spherical_shell(12, 8, 6, seed=4200 + i, noise=0.2,
                                boundary_flag=tuple(int(v) for v in rng.choice([1, 2], 4)) + (0, 0))
    else:
        p = ics.torus_sph(16, 8, 6)
    p.cfg.recon = int(rng.choice(RECONS))
    p.cfg.solver = int(rng.choice(SOLVERS))
    if kind != "box":
        p.cfg.geom_source_repeat = int(rng.choice([1, 8]))
    return kind, p


@pytest.mark.parametrize("i", range(24))
def test_randomized_configuration_bitwise(i):
    from cumpic_amd.cumc import Solver
    kind, p = _config(i)
    c = p.cfg
    print(f"[case] {i}: {kind} {c.nx}x{c.ny}x{c.nz} recon={c.recon} solver={c.solver} "
          f"flags={c.boundary_flag} repeat={c.geom_source_repeat}, 3 steps", flush=True)
    g, o = Solver.from_problem(p), Oracle.from_problem(p)
    g.initial_update()
    o.initial_update()
    dts_g, _ = g.run(3)
    dts_o, _ = o.run(3)
    assert_bitwise(dts_g, dts_o, "dt sequence")
    sg, so = g.download(), o.get_state()
    for k in ("prim", "eps", "bcell"):
        assert_bitwise(sg[k], so[k], f"config {i} {k}")
    assert g.find_divb() == o.find_divb()


@pytest.mark.parametrize("i", range(0, 24, 3))
def test_randomized_configuration_mixed_close_to_fp64(i):
    """The mixed form on the same draws (its FP32 sweeps, FP32 primitive planes and FP32 ghost
    pass in every geometry): finite, density and pressure within 1e-5 relative L1 of the FP64
    run on the same dt sequence, div B at round-off."""
    from cumpic_amd.config import FP64, MIXED
    from cumpic_amd.cumc import Solver
    from tests.helpers import interior, rel_l1
    kind, p = _config(i)
    print(f"[case] mixed {i}: {kind} recon={p.cfg.recon} solver={p.cfg.solver}, 3 steps", flush=True)
    a = Solver.from_problem(p, precision=FP64)
    b = Solver.from_problem(p, precision=MIXED)
    a.initial_update()
    b.initial_update()
    dts, _ = a.run(3)
    for dt in dts:
        b.rungekutta(float(dt))
    sa, sb = a.download(), b.download()
    assert np.isfinite(sb["prim"]).all()
    for k in (0, 4):
        e = rel_l1(interior(sb["prim"], p.cfg)[..., k], interior(sa["prim"], p.cfg)[..., k])
        assert e <= 1e-5, (i, k, e)
    assert b.find_divb()[1] <= 1e-12
