"""HIP path vs the CPU restatement for the extension options (TVD with the MC / van Leer
limiters, WENO5, the local Lax-Friedrichs and HLLD solvers): bit for bit on the same seeded
inputs, as for the reference's own options, in the three sweep directions, whole steps with
several boundary sets and geometries, and the mixed form against FP64 within its tolerance."""
from __future__ import annotations

import pytest

from cumpic_amd import ics
from cumpic_amd.config import FP64, HLL, HLLC, HLLD, LF, MIXED, PPMC, TVDMC, TVDMM, TVDVL, WENO
from oracle.oracle import Oracle
from tests.helpers import assert_bitwise, interior, rel_l1

pytestmark = pytest.mark.gpu


def _problem(recon, solver, flags=(0, 0, 0, 0, 0, 0), seed=11):
    p = ics.randomize(ics.small_periodic(12, seed=1100 + seed), seed=seed, amp=0.15)
    p.cfg.recon, p.cfg.solver, p.cfg.boundary_flag = recon, solver, flags
    return p


def _states_equal(g, o, what):
    sg, so = g.download(), o.get_state()
    for k in ("prim", "eps", "bcell"):
        assert_bitwise(sg[k], so[k], f"{what} {k}")


@pytest.mark.parametrize("direction", [1, 2, 3])
@pytest.mark.parametrize("recon,solver", [(PPMC, HLLD), (TVDVL, LF), (TVDMC, HLL), (WENO, HLL), (PPMC, HLLC)])
def test_extension_sweep_flux_bitwise(direction, recon, solver):
    from cumpic_amd.cumc import Solver
    print(f"[case] sweep_flux dir={direction} recon={recon} solver={solver} n=12", flush=True)
    p = _problem(recon, solver)
    g, o = Solver.from_problem(p), Oracle.from_problem(p)
    g.initial_update()
    o.initial_update()
    fg, fo = g.sweep_flux(direction), o.sweep_flux(direction)
    c = p.cfg
    n = [c.nx, c.ny, c.nz]
    sl = []
    for ax in (2, 1, 0):
        d = 3 - ax
        hi = n[d - 1] if d == direction else n[d - 1] + 1
        sl.append(slice(2, 2 + hi + 1))
    assert_bitwise(fg[tuple(sl)], fo[tuple(sl)], f"flux dir {direction}")


@pytest.mark.parametrize("recon,solver,flags", [
    (TVDMC, HLL, (0, 0, 0, 0, 0, 0)),
    (TVDVL, HLL, (1, 1, 2, 2, 1, 1)),
    (PPMC, HLLD, (0, 0, 0, 0, 0, 0)),
    (TVDMM, HLLD, (1, 1, 2, 2, 1, 1)),
    (TVDMC, HLLD, (2, 1, 3, 3, 0, 0)),
    (PPMC, LF, (0, 0, 0, 0, 0, 0)),
    (TVDVL, LF, (1, 0, 0, 1, 2, 0)),
    (WENO, HLL, (0, 0, 0, 0, 0, 0)),
    (WENO, HLLD, (1, 1, 2, 2, 1, 1)),
    (PPMC, HLLC, (0, 0, 0, 0, 0, 0)),
    (TVDMC, HLLC, (1, 1, 2, 2, 1, 1)),
])
def test_extension_steps_bitwise(recon, solver, flags):
    from cumpic_amd.cumc import Solver
    print(f"[case] 12^3 steps recon={recon} solver={solver} flags={flags}, 4 steps", flush=True)
    p = _problem(recon, solver, flags)
    g, o = Solver.from_problem(p), Oracle.from_problem(p)
    g.initial_update()
    o.initial_update()
    dts_g, _ = g.run(4)
    dts_o, _ = o.run(4)
    assert_bitwise(dts_g, dts_o, "dt sequence")
    _states_equal(g, o, "4 steps")


@pytest.mark.parametrize("geometry", ["torus_cyn", "spherical"])
def test_extension_curvilinear_steps_bitwise(geometry):
    from cumpic_amd.cumc import Solver
    print(f"[case] {geometry} with HLLD + MC, 3 steps", flush=True)
    if geometry == "torus_cyn":
        p = ics.torus_cyn(20, 8, 20)
    else:
        p = ics.spherical_shell(14, 10, 8, seed=1501, recon=PPMC, boundary_flag=(1, 1, 2, 2, 0, 0), noise=0.2)
    p.cfg.recon, p.cfg.solver = TVDMC, HLLD
    g, o = Solver.from_problem(p), Oracle.from_problem(p)
    g.initial_update()
    o.initial_update()
    dts_g, _ = g.run(3)
    dts_o, _ = o.run(3)
    assert_bitwise(dts_g, dts_o, "dt sequence")
    _states_equal(g, o, f"{geometry} 3 steps")


def test_extension_mixed_form_close_to_fp64():
    """The mixed form instantiates the same templates in FP32: with HLLD and PPM it stays
    within the mixed tolerance of the FP64 run on the same dt sequence."""
    from cumpic_amd.cumc import Solver
    print("[case] mixed vs fp64, HLLD + PPM, spectral box n=24, 10 steps", flush=True)
    p = ics.spectral_box(n=24, seed=1102, steps=10)
    p.cfg.solver = HLLD
    a = Solver.from_problem(p, precision=FP64)
    b = Solver.from_problem(p, precision=MIXED)
    a.initial_update()
    b.initial_update()
    dts, _ = a.run(10)
    for dt in dts:
        b.rungekutta(dt)
    sa, sb = a.download(), b.download()
    for k in (0, 4):
        e = rel_l1(interior(sb["prim"], p.cfg)[..., k], interior(sa["prim"], p.cfg)[..., k])
        assert 0.0 < e <= 1e-5, (k, e)


def test_batch_and_reverse_mode_reject_the_extensions():
    """The batched form and the reverse mode carry the reference's options only (HLL with PPM
    or TVD minmod); other options are refused with an error, not run."""
    from cumpic_amd.cumc import Batch, CumcError, Solver
    print("[case] batch / tape refuse HLLD", flush=True)
    p = ics.small_periodic(8, seed=1103, ndim=2)
    p.cfg.solver = HLLD
    s = Solver.from_problem(p)
    with pytest.raises(CumcError):
        Batch([s])
    s.upload(p.prim, p.eps)
    with pytest.raises(CumcError):
        s.tape_run(2, every=1)
