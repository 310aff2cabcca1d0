"""Batched form: boxes of uneven sizes advanced by persistent work-list launches (direct and
hipGraph-replayed) must be bit-identical to the same boxes advanced one at a time, and the one
at a time path is bit-identical to the CPU restatement (test_gpu_parity); one box is also
checked against the restatement directly here.
"""
from __future__ import annotations

import numpy as np
import pytest

from cumpic_amd import ics
from oracle.oracle import Oracle
from tests.helpers import assert_bitwise

pytestmark = pytest.mark.gpu


def _boxes():
    """Uneven sizes, 2D (nz = 1) and one 3D box, periodic, PPM."""
    probs = [ics.randomize(ics.spectral_box(n=n, seed=2000 + i, ndim=2, v_rms=0.3, db_rms=0.3), seed=i, amp=0.1)
             for i, n in enumerate((16, 24, 40, 9))]
    probs.append(ics.randomize(ics.small_periodic(10, seed=2100), seed=9, amp=0.1))
    probs.append(ics.spectral_box(n=20, seed=2200, ndim=2, ny=12))
    return probs


def _solvers(probs):
    from cumpic_amd.cumc import Solver
    out = []
    for p in probs:
        s = Solver.from_problem(p)
        s.initial_update()
        out.append(s)
    return out


def _same_state(a, b, what):
    sa, sb = a.download(), b.download()
    for k in ("prim", "eps", "bcell"):
        assert_bitwise(sa[k], sb[k], f"{what} {k}")


@pytest.mark.parametrize("graph_steps", [0, 2, 3])
def test_batch_matches_single_boxes(graph_steps):
    from cumpic_amd.cumc import Batch
    print(f"[case] batch vs single, graph_steps={graph_steps}", flush=True)
    probs = _boxes()
    singles = _solvers(probs)
    boxes = _solvers(probs)
    nsteps = 5
    ref_dts = np.array([s.run(nsteps)[0] for s in singles]).T  # (nsteps, nbox)
    b = Batch(boxes)
    dts = b.run(nsteps, graph_steps=graph_steps)
    assert_bitwise(dts, ref_dts, "per-box dt sequence")
    for i, (x, y) in enumerate(zip(boxes, singles)):
        _same_state(x, y, f"box {i}")
    # a second run continues from there (odd step count flips the buffer roles)
    dts2 = b.run(3, graph_steps=graph_steps)
    ref2 = np.array([s.run(3)[0] for s in singles]).T
    assert_bitwise(dts2, ref2, "second run dt")
    for i, (x, y) in enumerate(zip(boxes, singles)):
        _same_state(x, y, f"box {i} after second run")
        # the boxes stay usable one at a time
        assert x.finddt() == y.finddt()


def test_batch_frozen_dt_matches_rungekutta():
    from cumpic_amd.cumc import Batch
    print("[case] batch frozen dt", flush=True)
    probs = _boxes()[:3]
    singles = _solvers(probs)
    boxes = _solvers(probs)
    rng = np.random.default_rng(5)
    base = np.array([s.finddt() for s in singles])
    dts = base[None, :] * rng.uniform(0.5, 0.9, size=(4, len(probs)))
    for i, s in enumerate(singles):
        for n in range(4):
            s.rungekutta(dts[n, i])
    b = Batch(boxes)
    used = b.run(4, graph_steps=2, dts=dts)
    assert_bitwise(used, dts, "frozen dt echo")
    for i, (x, y) in enumerate(zip(boxes, singles)):
        _same_state(x, y, f"frozen box {i}")


def test_batch_box_vs_restatement():
    from cumpic_amd.cumc import Batch
    print("[case] batch box vs oracle", flush=True)
    probs = _boxes()[:2]
    boxes = _solvers(probs)
    o = Oracle.from_problem(probs[1])
    o.initial_update()
    b = Batch(boxes)
    dts = b.run(4, graph_steps=4)
    odts, _ = o.run(4, 1.0e300)
    assert_bitwise(dts[:, 1], np.asarray(odts), "dt vs oracle")
    sg, so = boxes[1].download(), o.get_state()
    for k in ("prim", "eps", "bcell"):
        assert_bitwise(sg[k], so[k], f"batch box vs oracle {k}")


def test_batch_autotune_keeps_state_and_results():
    from cumpic_amd.cumc import Batch
    print("[case] batch autotune", flush=True)
    probs = _boxes()
    boxes = _solvers(probs)
    singles = _solvers(probs)
    b = Batch(boxes)
    before = [x.download() for x in boxes]
    t = b.autotune()
    assert t["seg"] in (8, 12, 16, 24, 32, 64, 128, 256) and t["zchunk"] in (1, 2, 4, 8, 32)
    for i, x in enumerate(boxes):
        after = x.download()
        for k in ("prim", "eps", "bcell"):
            assert_bitwise(after[k], before[i][k], f"autotune changed box {i} {k}")
    dts = b.run(4, graph_steps=4)
    ref = np.array([s.run(4)[0] for s in singles]).T
    assert_bitwise(dts, ref, "tuned dt")
    for i, (x, y) in enumerate(zip(boxes, singles)):
        _same_state(x, y, f"tuned box {i}")
    # explicit shapes round-trip and still give the same results
    b.set_tuning({"seg": 7, "zchunk": 3, "tile_w": 32, "blocks": {"flux_x": 5, "flux_y": 3, "flux_z": 2,
                                                                  "ecorn": 4, "tail": 6, "ghost": 1, "finddt": 2}})
    assert b.tuning()["seg"] == 7
    dts = b.run(2, graph_steps=0)
    ref = np.array([s.run(2)[0] for s in singles]).T
    assert_bitwise(dts, ref, "odd-shape dt")
    for i, (x, y) in enumerate(zip(boxes, singles)):
        _same_state(x, y, f"odd-shape box {i}")


def test_single_autotune_keeps_state_and_results():
    from cumpic_amd.cumc import Solver
    print("[case] single-box autotune", flush=True)
    p = ics.randomize(ics.small_periodic(24, seed=2300), seed=3, amp=0.1)
    a = Solver.from_problem(p)
    c = Solver.from_problem(p)
    a.initial_update()
    c.initial_update()
    before = a.download()
    t = a.autotune()
    after = a.download()
    for k in ("prim", "eps", "bcell"):
        assert_bitwise(after[k], before[k], f"autotune changed {k}")
    assert t["seg"] >= 1 and t["zchunk"] >= 1 and t["tile_w"] in (0, 32, 64, 128)
    da, _ = a.run(3)
    dc, _ = c.run(3)
    assert_bitwise(da, dc, "dt after autotune")
    _same_state(a, c, "after autotune")
    # every column tile shape gives the same bits
    for w in (32, 64, 128):
        a.set_tuning({"seg": 16, "zchunk": 5, "tile_w": w})
        a.run(1)
        c.run(1)
        _same_state(a, c, f"tile_w {w}")
    da, _ = a.run(3)
    dc, _ = c.run(3)
    assert_bitwise(da, dc, "dt after autotune")
    _same_state(a, c, "after autotune")


def test_batch_rejects_unsupported_boxes():
    from cumpic_amd.cumc import Batch, CumcError, Solver
    from cumpic_amd.config import MIXED
    print("[case] batch argument checks", flush=True)
    p = ics.small_periodic(8, seed=2400)
    s64 = Solver.from_problem(p)
    smx = Solver.from_problem(p, precision=MIXED)
    with pytest.raises(CumcError):
        Batch([s64, smx])
    with pytest.raises(CumcError):
        Batch([s64, s64])
    b = Batch([s64])
    with pytest.raises(CumcError):
        b.run(2, dts=np.array([[1e-3], [np.nan]]))
