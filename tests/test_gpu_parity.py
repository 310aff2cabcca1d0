"""HIP path vs the CPU restatement: bit-for-bit on the same seeded inputs.

All arithmetic on both sides is IEEE double with no contracted multiply-adds and correctly
rounded division and square root, in the same operation order, so every entry of the
ghost-inclusive state must agree exactly (FP64 path).  Run with -m gpu on an MI355X.
"""
from __future__ import annotations

import numpy as np
import pytest

from cumpic_amd import ics
from cumpic_amd.config import PPMC, TVDMM
from oracle.oracle import Oracle
from tests.helpers import assert_bitwise

pytestmark = pytest.mark.gpu


def _pair(prob):
    from cumpic_amd.cumc import Solver
    g = Solver.from_problem(prob)
    o = Oracle.from_problem(prob)
    return g, o


def _states_equal(g, o, what):
    sg = g.download()
    so = o.get_state()
    assert_bitwise(sg["prim"], so["prim"], what + " prim")
    assert_bitwise(sg["eps"], so["eps"], what + " eps")
    assert_bitwise(sg["bcell"], so["bcell"], what + " bcell")


def _periodic(n=12, seed=7, recon=PPMC):
    p = ics.randomize(ics.small_periodic(n, seed=1005 + seed), seed=seed, amp=0.15)
    p.cfg.recon = recon
    return p


@pytest.mark.parametrize("direction", [1, 2, 3])
@pytest.mark.parametrize("recon", [PPMC, TVDMM])
def test_sweep_flux_bitwise(direction, recon):
    print(f"[case] sweep_flux dir={direction} recon={recon} n=12", flush=True)
    p = _periodic(recon=recon)
    g, o = _pair(p)
    g.initial_update()
    o.initial_update()
    fg = g.sweep_flux(direction)
    fo = o.sweep_flux(direction)
    c = p.cfg
    n = [c.nx, c.ny, c.nz]
    sl = []
    for ax in (2, 1, 0):  # array axes are (l, k, j)
        d = 3 - ax
        hi = n[d - 1] if d == direction else n[d - 1] + 1
        sl.append(slice(2, 2 + hi + 1))
    sl = tuple(sl)
    assert_bitwise(fg[sl], fo[sl], f"flux dir {direction}")


def test_initial_update_bitwise():
    print("[case] initial_update n=12", flush=True)
    p = _periodic()
    g, o = _pair(p)
    g.initial_update()
    o.initial_update()
    _states_equal(g, o, "initial_update")


@pytest.mark.parametrize("recon", [PPMC, TVDMM])
def test_one_step_bitwise(recon):
    print(f"[case] one RK3 step recon={recon} n=12", flush=True)
    p = _periodic(recon=recon)
    g, o = _pair(p)
    g.initial_update()
    o.initial_update()
    dg, do = g.finddt(), o.finddt()
    assert dg == do, (dg, do)
    g.rungekutta(do)
    o.rungekutta(do)
    _states_equal(g, o, "after one step")


def test_riemann_1d_full_run_bitwise():
    print("[case] 1D Riemann 800 cells to t=0.1", flush=True)
    p = ics.riemann_1d()
    g, o = _pair(p)
    g.initial_update()
    o.initial_update()
    dts_g, tg = g.run(10 ** 5, p.total_time)
    dts_o, to = o.run(10 ** 5, p.total_time)
    assert len(dts_g) == len(dts_o) and tg == to
    assert_bitwise(dts_g, dts_o, "dt sequence")
    _states_equal(g, o, "1D Riemann final")


def test_periodic_3d_steps_bitwise():
    print("[case] 3D periodic 16^3, 5 steps", flush=True)
    p = ics.small_periodic(16)
    g, o = _pair(p)
    g.initial_update()
    o.initial_update()
    dts_g, _ = g.run(5)
    dts_o, _ = o.run(5)
    assert_bitwise(dts_g, dts_o, "dt sequence")
    _states_equal(g, o, "periodic 5 steps")
    sg, ag = g.find_divb()
    so, ao = o.find_divb()
    assert sg == so and ag == ao
    assert ag < 1e-12


def test_torus_cylindrical_steps_bitwise():
    print("[case] cylindrical torus 24x12x24, 3 steps", flush=True)
    p = ics.torus_cyn(24, 12, 24)
    g, o = _pair(p)
    g.initial_update()
    o.initial_update()
    dts_g, _ = g.run(3)
    dts_o, _ = o.run(3)
    assert_bitwise(dts_g, dts_o, "dt sequence")
    _states_equal(g, o, "torus 3 steps")


@pytest.mark.parametrize("flags", [(1, 1, 2, 2, 1, 1), (2, 1, 3, 3, 4, 4), (1, 0, 0, 1, 2, 0)])
def test_boundary_variants_bitwise(flags):
    print(f"[case] boundary flags {flags} n=12, 2 steps", flush=True)
    p = _periodic(seed=3)
    p.cfg.boundary_flag = flags
    g, o = _pair(p)
    g.initial_update()
    o.initial_update()
    _states_equal(g, o, f"initial {flags}")
    dts_g, _ = g.run(2)
    dts_o, _ = o.run(2)
    assert_bitwise(dts_g, dts_o, "dt sequence")
    _states_equal(g, o, f"flags {flags}")


@pytest.mark.parametrize("recon", [PPMC, TVDMM])
@pytest.mark.parametrize("flags", [(1, 1, 1, 1, 0, 0), (2, 2, 2, 2, 0, 0)])
def test_spherical_steps_bitwise(recon, flags):
    """Spherical (r, theta, phi) shell: GEOM_FLUX / GEOM_CT / GEOM_SOURCE / finddt / find_divb of
    coords.f90's spherical branches, with the trigonometric factors taken from host tables."""
    print(f"[case] spherical shell 14x10x8 recon={recon} flags={flags}, 4 steps", flush=True)
    p = ics.spherical_shell(14, 10, 8, seed=1500, recon=recon, boundary_flag=flags, noise=0.2)
    g, o = _pair(p)
    g.initial_update()
    o.initial_update()
    _states_equal(g, o, "spherical initial")
    dts_g, _ = g.run(4)
    dts_o, _ = o.run(4)
    assert_bitwise(dts_g, dts_o, "dt sequence")
    _states_equal(g, o, "spherical 4 steps")
    sg, ag = g.find_divb()
    so, ao = o.find_divb()
    assert sg == so and ag == ao


def test_spherical_mixed_and_batch():
    """The mixed form runs the spherical tail (FP32 fluxes, FP64 state) close to FP64, and a batch of
    spherical boxes equals the boxes run one at a time bit for bit."""
    from cumpic_amd.config import MIXED
    from cumpic_amd.cumc import Batch, Solver
    print("[case] spherical mixed vs FP64, and batched", flush=True)
    p = ics.spherical_shell(16, 12, 8, seed=1501, noise=0.1)
    f64 = Solver.from_problem(p)
    mx = Solver.from_problem(p, precision=MIXED)
    f64.initial_update()
    mx.initial_update()
    dts, _ = f64.run(5)
    for dt in dts:
        mx.rungekutta(float(dt))
    a, b = f64.download()["prim"][3:-3, 3:-3, 3:-3], mx.download()["prim"][3:-3, 3:-3, 3:-3]
    rel = np.sum(np.abs(a - b), axis=(0, 1, 2)) / np.sum(np.abs(a), axis=(0, 1, 2))
    assert np.all(rel[:5] < 1e-5), rel
    _, divb = mx.find_divb()
    assert divb < 1e-15
    probs = [ics.spherical_shell(n, 8, 6, seed=1510 + n, noise=0.1) for n in (10, 14)]
    singles = [Solver.from_problem(q) for q in probs]
    boxes = [Solver.from_problem(q) for q in probs]
    for s in singles + boxes:
        s.initial_update()
    ref = np.array([s.run(3)[0] for s in singles]).T
    got = Batch(boxes).run(3, graph_steps=3)
    assert_bitwise(got, ref, "batched spherical dt")
    for x, y in zip(boxes, singles):
        sx, sy = x.download(), y.download()
        for k in ("prim", "eps", "bcell"):
            assert_bitwise(sx[k], sy[k], f"batched spherical {k}")


def test_torus_spherical_steps_bitwise():
    """torus_sph plugin hooks on the spherical grid: floors, radial Paczynski-Wiita gravity and the
    radial inflow clamp, against the restatement bit for bit."""
    print("[case] spherical torus 24x12x8, 3 steps", flush=True)
    p = ics.torus_sph(24, 12, 8)
    g, o = _pair(p)
    g.initial_update()
    o.initial_update()
    dts_g, _ = g.run(3)
    dts_o, _ = o.run(3)
    assert_bitwise(dts_g, dts_o, "dt sequence")
    _states_equal(g, o, "spherical torus 3 steps")


@pytest.mark.parametrize("shape", ["2d", "3d"])
def test_run_step_graphs_match_direct_steps(shape):
    """cumc_run replays runs of 8 or more steps as hipGraphs (per buffer parity and view); the same
    steps launched one by one (finddt + rungekutta) give the same dt sequence and state bit for
    bit: a z-collapsed 2D box (whose first step runs in the full layout, the later ones collapsed)
    and a 24^3 box, over two runs with a download between them."""
    from cumpic_amd.cumc import Solver
    print(f"[case] run() step graphs vs direct steps, {shape}", flush=True)
    if shape == "2d":
        p = ics.randomize(ics.spectral_box(n=40, seed=5100, ndim=2, v_rms=0.3, db_rms=0.3, ny=48), seed=5101,
                          amp=0.1)
    else:
        p = ics.randomize(ics.small_periodic(24, seed=5200), seed=5201, amp=0.1)
    a, b = Solver.from_problem(p), Solver.from_problem(p)
    a.initial_update()
    b.initial_update()
    for nsteps in (9, 12):
        da, _ = a.run(nsteps)
        db = []
        for _ in range(nsteps):
            dt = b.finddt()
            b.rungekutta(dt)
            db.append(dt)
        assert_bitwise(da, np.array(db), f"dt sequence of {nsteps} steps")
        if shape == "2d":
            assert a.zcollapse_state() == (True, True), "the 2D box should step collapsed"
        sa, sb = a.download(), b.download()
        for k in ("prim", "eps", "bcell"):
            assert_bitwise(sa[k], sb[k], f"{k} after {nsteps} more steps")
