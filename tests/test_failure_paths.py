"""Floors and non-finite states on the CPU restatement: the inputs of the GPU failure-path tests
do drive the step into the branches those tests compare (riemann eps floor, torus density and eps
floors), and the CFL reduction reports a state with one NaN cell instead of dropping the cell."""
from __future__ import annotations

import math

import numpy as np

from cumpic_amd import ics
from oracle.oracle import Oracle
from tests.failure_cases import double_rarefaction, with_nan_cell


def test_riemann_eps_floor_fires_in_a_double_rarefaction():
    p = double_rarefaction()
    o = Oracle.from_problem(p)
    o.initial_update()
    assert o.floor_counts() == (0, 0)
    for _ in range(p.steps):
        o.rungekutta(o.finddt())
    nr, ne = o.floor_counts()
    o.close()
    assert nr == 0 and ne > 0
    print(f"eps floored {ne} times in {p.steps} steps")


def test_torus_floors_fire():
    p = ics.torus_cyn(24, 12, 24)
    o = Oracle.from_problem(p)
    o.initial_update()
    o.run(3)
    nr, ne = o.floor_counts()
    o.close()
    assert nr > 0 and ne > 0
    print(f"torus: density floored {nr} times, eps {ne} times in 3 steps")


def test_finddt_reports_one_nan_cell():
    base = ics.randomize(ics.small_periodic(8, seed=1105), seed=3, amp=0.1)
    o = Oracle.from_problem(base)
    o.initial_update()
    dt = o.finddt()
    o.close()
    assert math.isfinite(dt) and dt > 0.0
    o = Oracle.from_problem(with_nan_cell(base))
    o.initial_update()
    assert o.finddt() == 0.0  # reported whatever the reduction order
    o.close()


def test_nan_stays_local_for_one_stage_width():
    """One NaN cell spreads through the stencils, but cells beyond the reach of one RK3 step
    (three stages x 3-cell stencils) stay finite."""
    base = ics.randomize(ics.small_periodic(24, seed=1106), seed=4, amp=0.1)
    o = Oracle.from_problem(with_nan_cell(base, (12, 12, 12)))
    o.initial_update()
    o.rungekutta(1.0e-3)
    st = o.get_state()
    o.close()
    bad = np.argwhere(np.isnan(st["prim"][..., 0]))
    assert len(bad) > 1
    # x column j = 1 is 11 cells from j = 12 one way round the periodic box and 13 the other
    far = st["prim"][ics.ix(1):ics.ix(24) + 1, ics.ix(1):ics.ix(24) + 1, ics.ix(1), 0]
    assert np.isfinite(far).all()
