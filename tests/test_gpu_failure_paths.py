"""Floor and non-finite paths of the HIP step against the CPU restatement, bit for bit.

The riemann plugin's eps floor (model/riemann/src/custom.f90:66-68) guards the cancellation
E - rho v^2/2 - B^2/2 of solver/utop.f90:63-64; the torus plugin floors density and eps
(model/torus_cyn/src/custom.f90:86-116).  Each case first shows on the restatement that the branch
fires (its CUSTOM_CHECKRHO counts), and the GPU state must equal the restatement's after every
step: a floor the GPU skipped or applied elsewhere changes p = rho eps (gamma - 1) and every flux
after it.  One NaN cell must spread through the same cells on both sides, and the CFL reduction
must report it (CUMC_ENONFINITE) instead of dropping it.
"""
from __future__ import annotations

import numpy as np
import pytest

from cumpic_amd import ics
from oracle.oracle import Oracle
from tests.failure_cases import double_rarefaction, with_nan_cell
from tests.helpers import assert_bitwise

pytestmark = pytest.mark.gpu


def _pair(prob):
    from cumpic_amd.cumc import Solver
    return Solver.from_problem(prob), Oracle.from_problem(prob)


def _same(g, o, what):
    sg, so = g.download(), o.get_state()
    for k in ("prim", "eps", "bcell"):
        assert_bitwise(sg[k], so[k], f"{what} {k}")


def test_riemann_eps_floor_fires_bitwise():
    p = double_rarefaction()
    print(f"[case] double rarefaction 1D n={p.cfg.nx}, {p.steps} steps", flush=True)
    g, o = _pair(p)
    g.initial_update()
    o.initial_update()
    for m in range(p.steps):
        dg, do = g.finddt(), o.finddt()
        assert dg == do, (m, dg, do)
        g.rungekutta(dg)
        o.rungekutta(do)
        _same(g, o, f"step {m}")
    nr, ne = o.floor_counts()
    print(f"  eps floored {ne} times (density floor {nr})", flush=True)
    assert ne > 0


def test_torus_floors_fire_bitwise():
    print("[case] cylindrical torus 24x12x24, floors firing, 3 steps", flush=True)
    p = ics.torus_cyn(24, 12, 24)
    g, o = _pair(p)
    g.initial_update()
    o.initial_update()
    for m in range(3):
        dg, do = g.finddt(), o.finddt()
        assert dg == do
        g.rungekutta(dg)
        o.rungekutta(do)
        _same(g, o, f"step {m}")
    nr, ne = o.floor_counts()
    print(f"  density floored {nr} times, eps {ne} times", flush=True)
    assert nr > 0 and ne > 0


def test_nan_cell_propagates_identically():
    from cumpic_amd.cumc import CumcError
    base = ics.randomize(ics.small_periodic(12, seed=1107), seed=5, amp=0.1)
    p = with_nan_cell(base, (6, 6, 6))
    print("[case] one NaN cell, 12^3 periodic, one step at a fixed dt", flush=True)
    g, o = _pair(p)
    g.initial_update()
    o.initial_update()
    _same(g, o, "initial_update")
    g.rungekutta(2.0e-3)
    o.rungekutta(2.0e-3)
    sg, so = g.download(), o.get_state()
    for k in ("prim", "eps", "bcell"):
        assert np.array_equal(np.isnan(sg[k]), np.isnan(so[k])), k
        assert_bitwise(sg[k], so[k], f"NaN step {k}")
    nbad = int(np.isnan(sg["prim"][..., 0]).sum())
    print(f"  NaN density in {nbad} cells on both sides", flush=True)
    assert nbad > 1
    assert o.finddt() == 0.0
    with pytest.raises(CumcError, match="non-finite"):
        g.finddt()


def test_finddt_reports_one_nan_cell():
    from cumpic_amd.cumc import CumcError
    base = ics.randomize(ics.small_periodic(16, seed=1108), seed=6, amp=0.1)
    print("[case] finddt on a 16^3 state with one NaN cell (upload) and after a step", flush=True)
    g, _ = _pair(base)
    g.initial_update()
    assert g.finddt() > 0.0
    g.close()
    g, _ = _pair(with_nan_cell(base, (3, 9, 14)))
    g.initial_update()
    with pytest.raises(CumcError, match="non-finite"):
        g.finddt()
    with pytest.raises(CumcError, match="non-finite"):
        g.run(2)
