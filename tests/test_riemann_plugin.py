"""The riemann plugin's three named problems (model/riemann/src/riemann1d.f90, riemann2d.f90)
as built by ``ics.riemann_test``, run on the CPU restatement to their stop times and checked
against what the mathematics fixes: the exact Euler solution for Toro's test 1, the global
momentum budgets that the outflow boundaries' constant boundary fluxes fix in closed form,
exact conservation of mass and energy while the waves stay inside the box, planar symmetry
(v_z, B_z stay exactly zero, B_x stays exactly 0.75 in 1D) and the rotor's point symmetry."""
from __future__ import annotations

import math

import numpy as np
import pytest

from cumpic_amd import ics
from cumpic_amd.ics import BRIO_WU, MHD_ROTOR, TORO_1
from oracle.oracle import Oracle
from tests.helpers import exact_riemann_euler, interior


def _totals(o, cfg):
    s = o.get_state()
    return interior(s["cons"], cfg)[..., :5].sum(axis=(0, 1, 2))


def _run(test, n):
    p = ics.riemann_test(test, n=n)
    o = Oracle.from_problem(p)
    o.initial_update()
    t0 = _totals(o, p.cfg)
    dts, t = o.run(10 ** 6, p.total_time)
    assert t == p.total_time and len(dts) > 10
    return p, o, t0, _totals(o, p.cfg), t


def test_initial_states_follow_the_plugin():
    """Cell values on [0,n]^3 and the derived internal energy: the two 1D problems switch at
    x = 0.5 (cell centres), the rotor's disc is r <= 0.1 about the box centre, spinning with
    |v| = r/0.1 so that its rim moves at unit speed."""
    g = ics.RIEMANN_GGAS
    for test in (TORO_1, BRIO_WU):
        p = ics.riemann_test(test, n=64)
        row = p.prim[2, 2, 2:2 + 65]  # cells j = 0..64 of the k = l = 0 row
        xc = 0.5 * (p.xf[3:3 + 65] + p.xf[2:2 + 65])
        left = xc <= 0.5
        assert left.sum() == 33  # cells 0..32
        np.testing.assert_array_equal(row[:, 0], np.where(left, 1.0, 0.125))
        np.testing.assert_array_equal(row[:, 4], np.where(left, 1.0, 0.1))
        np.testing.assert_array_equal(row[:, 5], 0.75 if test == BRIO_WU else 0.0)
        np.testing.assert_array_equal(row[:, 6], np.where(left, 1.0, -1.0) if test == BRIO_WU else 0.0)
        assert np.all(row[:, 1:4] == 0) and np.all(row[:, 7] == 0)
        np.testing.assert_allclose(p.eps[2, 2, 2:2 + 65], row[:, 4] / row[:, 0] / (g - 1), rtol=1e-15)
    p = ics.riemann_test(MHD_ROTOR, n=100)
    pr = p.prim[2, 2:2 + 101, 2:2 + 101]
    x = 0.5 * (p.xf[3:3 + 101] + p.xf[2:2 + 101])[None, :]
    y = x.T
    r = np.hypot(x - 0.5, y - 0.5)
    v = pr[..., 1:3]
    assert np.all(pr[r <= 0.1][:, 0] == 10.0) and np.all(pr[r >= 0.115][:, 0] == 1.0)
    np.testing.assert_allclose(np.hypot(v[..., 0], v[..., 1])[r <= 0.1], r[r <= 0.1] / 0.1, rtol=1e-14)
    # rigid rotation: v is perpendicular to the radius everywhere, counter-clockwise
    assert np.abs(v[..., 0] * (x - 0.5) + v[..., 1] * (y - 0.5)).max() < 1e-15
    assert np.all(v[..., 1] * (x - 0.5) - v[..., 0] * (y - 0.5) >= 0)
    assert np.all(pr[..., 4] == 0.5) and np.all(pr[..., 5] == 2.5 / math.sqrt(4 * math.pi))
    assert 1.0 <= pr[..., 0].min() and np.hypot(v[..., 0], v[..., 1]).max() <= 1.0


def test_toro1_against_exact_solution():
    """Toro's test 1 at the plugin's gamma = 5/3: small L1 density error against the exact
    Euler solution that shrinks under refinement."""
    errs = []
    for n in (100, 200):
        p, o, *_ = _run(TORO_1, n)
        rho = o.get_state()["prim"][3, 3, 3:3 + n, 0]
        xc = 0.5 * (p.xf[3:3 + n] + p.xf[4:4 + n])
        ex = exact_riemann_euler((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), ics.RIEMANN_GGAS, (xc - 0.5) / 0.2)
        errs.append(float(np.mean(np.abs(rho - ex[:, 0]))))
    assert errs[1] < 0.02 and errs[1] < 0.75 * errs[0], errs


@pytest.mark.parametrize("test", [TORO_1, BRIO_WU])
def test_shock_tube_budgets(test):
    """With outflow ghosts and untouched end cells, each boundary face carries the initial
    state's physical flux: no mass or energy flux, x momentum flux p + (B_t^2 - B_x^2)/2 and
    y momentum flux -B_x B_y.  So dx * sum(rho v_x) = t (P_L - P_R), dx * sum(rho v_y) =
    t B_x (B_yR - B_yL), mass and energy are conserved; a dropped magnetic pressure, a wrong
    stress sign or a leaking boundary flux breaks one of them."""
    p, o, t0, t1, t = _run(test, 200)
    dx = 1.0 / 200
    bx, byl, byr = (0.75, 1.0, -1.0) if test == BRIO_WU else (0.0, 0.0, 0.0)
    pl = 1.0 + 0.5 * (byl ** 2 - bx ** 2)
    pr = 0.1 + 0.5 * (byr ** 2 - bx ** 2)
    assert abs(dx * (t1[1] - t0[1]) - t * (pl - pr)) < 1e-13
    assert abs(dx * (t1[2] - t0[2]) - t * bx * (byr - byl)) < 1e-13
    assert t1[3] == 0.0
    for k in (0, 4):
        assert abs(t1[k] - t0[k]) < 1e-14 * abs(t0[k]), (k, t1[k] - t0[k])
    s = o.get_state()
    assert np.all(s["prim"][..., 3] == 0.0) and np.all(s["prim"][..., 7] == 0.0)
    bxf = s["prim"][3, 3, 2:3 + 200, 5]  # faces 0..n
    np.testing.assert_array_equal(bxf, bx)
    row = interior(s["prim"], p.cfg)[0, 0]
    assert row[0, 0] == 1.0 and row[-1, 0] == 0.125  # the waves have not reached the ends


def test_mhd_rotor_symmetry_and_conservation():
    """The rotor is invariant under the point reflection through the box centre (rho, p, B even,
    v odd); the scheme keeps that to round-off, keeps v_z = B_z = 0, conserves mass and energy
    while the fast front stays inside the box, keeps the total momentum at zero and div B at
    round-off."""
    p, o, t0, t1, t = _run(MHD_ROTOR, 64)
    s = o.get_state()
    q = interior(s["prim"], p.cfg)[0]
    refl = q[::-1, ::-1]
    for k, sign in ((0, 1), (1, -1), (2, -1), (4, 1)):
        assert np.abs(q[..., k] - sign * refl[..., k]).max() < 1e-11 * np.abs(q[..., k]).max(), k
    assert np.all(s["prim"][..., 3] == 0.0) and np.all(s["prim"][..., 7] == 0.0)
    for k in (0, 4):
        assert abs(t1[k] - t0[k]) < 1e-13 * abs(t0[k]), (k, t1[k] - t0[k])
    mom = np.abs(interior(s["cons"], p.cfg)[..., 1:3]).sum()
    assert abs(t1[1]) < 1e-12 * mom and abs(t1[2]) < 1e-12 * mom
    assert o.find_divb()[1] < 1e-12
    edge = np.concatenate([q[0], q[-1], q[:, 0], q[:, -1]])
    # the ambient gas at the walls is untouched but for the scheme's smooth precursor (~1e-11)
    assert np.abs(edge[:, 0] - 1.0).max() < 1e-9 and np.abs(edge[:, 1:3]).max() < 1e-9
    assert q[..., 0].min() > 0.0 and q[..., 4].min() > 0.0
