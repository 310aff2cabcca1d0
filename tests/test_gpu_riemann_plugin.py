"""The riemann plugin's three named problems (Toro's test 1, Brio-Wu, the MHD rotor) run to
their stop times on the GPU and on the CPU restatement: bit for bit, dt sequence included,
with the reference's options (PPM, TVD minmod) and with the extension solvers; and the mixed
form against FP64 on the rotor."""
from __future__ import annotations

import numpy as np
import pytest

from cumpic_amd import ics
from cumpic_amd.config import FP64, HLL, HLLC, HLLD, MIXED, PPMC, TVDMM, WENO
from cumpic_amd.ics import BRIO_WU, MHD_ROTOR, TORO_1
from oracle.oracle import Oracle
from tests.helpers import assert_bitwise, interior, rel_l1

pytestmark = pytest.mark.gpu

CASES = [
    (TORO_1, 200, PPMC, HLL),
    (TORO_1, 200, TVDMM, HLL),
    (BRIO_WU, 200, PPMC, HLL),
    (BRIO_WU, 200, PPMC, HLLD),
    (BRIO_WU, 160, WENO, HLLC),
    (MHD_ROTOR, 64, PPMC, HLL),
    (MHD_ROTOR, 48, TVDMM, HLLD),
]


@pytest.mark.parametrize("test,n,recon,solver", CASES)
def test_named_riemann_problem_bitwise(test, n, recon, solver):
    from cumpic_amd.cumc import Solver
    p = ics.riemann_test(test, n=n, recon=recon)
    p.cfg.solver = solver
    print(f"[case] {p.name} n={n} recon={recon} solver={solver} to t={p.total_time}", flush=True)
    g, o = Solver.from_problem(p), Oracle.from_problem(p)
    g.initial_update()
    o.initial_update()
    dts_g, tg = g.run(10 ** 4, p.total_time)
    dts_o, to = o.run(10 ** 4, p.total_time)
    assert tg == to == p.total_time
    assert_bitwise(dts_g, dts_o, "dt sequence")
    sg, so = g.download(), o.get_state()
    for k in ("prim", "eps", "bcell"):
        assert_bitwise(sg[k], so[k], f"{p.name} {k}")


def test_mhd_rotor_mixed_close_to_fp64():
    """Mixed form on the rotor (strong shear and a compressed shell): density and pressure
    within 1e-5 relative L1 of FP64 on the same dt sequence, div B at round-off."""
    from cumpic_amd.cumc import Solver
    p = ics.riemann_test(MHD_ROTOR, n=96)
    print(f"[case] mixed rotor n=96 to t={p.total_time}", flush=True)
    a = Solver.from_problem(p, precision=FP64)
    b = Solver.from_problem(p, precision=MIXED)
    a.initial_update()
    b.initial_update()
    dts, _ = a.run(10 ** 4, p.total_time)
    for dt in dts:
        b.rungekutta(float(dt))
    sa, sb = a.download(), b.download()
    assert np.isfinite(sb["prim"]).all()
    for k in (0, 4):
        e = rel_l1(interior(sb["prim"], p.cfg)[..., k], interior(sa["prim"], p.cfg)[..., k])
        assert 0.0 < e <= 1e-5, (k, e)
    assert b.find_divb()[1] <= 1e-12
