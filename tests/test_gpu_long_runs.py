"""Longer runs of the workload initial conditions, HIP path vs the CPU restatement bit for bit.

The parity tests elsewhere take a few steps of small randomized boxes; these take the seeded
workload builders themselves at reduced size over tens of steps, each step with its own CFL dt:
the 3D periodic spectral box (the |k|^-2 spectrum of the 256^3 workload, seed 1001) at 64^3 for
40 steps, the cylindrical torus (floor atmosphere, gravity, inflow clamp; seed 1003) at
64x32x64 for 20 steps, and the 2D spectral box of the batched workload at 96^2 for 40 steps.  The
dt sequences, the final ghost-inclusive states and max |div B| must agree exactly.
"""
from __future__ import annotations

import pytest

from cumpic_amd import ics
from oracle.oracle import Oracle
from tests.helpers import assert_bitwise

pytestmark = pytest.mark.gpu


def _run_both(prob, nsteps, what):
    from cumpic_amd.cumc import Solver
    g = Solver.from_problem(prob)
    o = Oracle.from_problem(prob)
    g.initial_update()
    o.initial_update()
    dts_g, _ = g.run(nsteps)
    dts_o, _ = o.run(nsteps)
    assert len(dts_g) == nsteps
    assert_bitwise(dts_g, dts_o, what + " dt sequence")
    sg, so = g.download(), o.get_state()
    for k in ("prim", "eps", "bcell"):
        assert_bitwise(sg[k], so[k], f"{what} {k}")
    dg, do = g.find_divb(), o.find_divb()
    assert dg == do, (dg, do)
    g.close()
    o.close()
    return dg


def test_spectral_box_64_cubed_40_steps_bitwise():
    print("[case] 3D periodic spectral box 64^3 (seed 1001), 40 steps", flush=True)
    _, dabs = _run_both(ics.spectral_box(n=64, seed=1001, steps=40), 40, "spectral 64^3")
    assert dabs < 1e-12


def test_torus_64x32x64_20_steps_bitwise():
    print("[case] cylindrical torus 64x32x64 (seed 1003), 20 steps", flush=True)
    _run_both(ics.torus_cyn(64, 32, 64, steps=20), 20, "torus 64x32x64")


def test_spectral_2d_96_squared_40_steps_bitwise():
    print("[case] 2D periodic spectral box 96^2 (seed 1004), 40 steps", flush=True)
    _, dabs = _run_both(ics.spectral_box(n=96, seed=1004, ndim=2, steps=40), 40, "spectral 96^2")
    assert dabs < 1e-12
