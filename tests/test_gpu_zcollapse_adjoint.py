"""Batched reverse mode of z-collapsed boxes (periodic nz = 1 steps on their single plane).

The collapsed forward is bit-identical to the full layout (test_gpu_zcollapse), so both reverse
passes differentiate the same map; they sum the adjoint contributions of the z-neighbour reads
in a different order (the collapsed reverse adds a z neighbour's share straight into the plane,
the full-layout one first into a ghost plane and then through the ghost-fill reverse), so the
gradients agree to rounding, not bit for bit.  One exception is the split between the two
periodic copies of a z face: the collapsed map evolves face B_z of planes 0 and 1 as one value,
so the derivative with respect to the uploaded B_z of both planes is checked as their sum.
Within the collapsed reverse, gradients are reproducible bit for bit (checkpoint intervals, a
second backward pass on one tape), and they match central differences of the forward map.
"""
from __future__ import annotations

import numpy as np
import pytest

from cumpic_amd import ics
from cumpic_amd.config import IBZ
from cumpic_amd.ics import ix

pytestmark = pytest.mark.gpu


def _inner(cfg):
    return (slice(ix(1), ix(cfg.nz) + 1), slice(ix(1), ix(cfg.ny) + 1), slice(ix(1), ix(cfg.nx) + 1))


def _dv(p):
    return (p.xf[4] - p.xf[3]) * (p.yf[4] - p.yf[3]) * (p.zf[4] - p.zf[3])


def _probs():
    probs = [ics.randomize(ics.spectral_box(n=n, seed=4100 + i, ndim=2, v_rms=0.3, db_rms=0.3), seed=i, amp=0.05)
             for i, n in enumerate((12, 20, 9))]
    # short y pencils (per-pencil y reverse) and a 3D box (full layout) in the same batch
    probs.append(ics.randomize(ics.spectral_box(n=14, seed=4200, ndim=2, v_rms=0.3, db_rms=0.3, ny=5), seed=8,
                               amp=0.05))
    probs.append(ics.randomize(ics.small_periodic(8, seed=4300), seed=7, amp=0.05))
    # outflow on every side (z included): the MHD rotor
    probs.append(ics.riemann_test(ics.MHD_ROTOR, n=24))
    return probs


def _box_like(p, seed):
    return ics.randomize(ics.spectral_box(n=p.cfg.nx, seed=seed, ndim=2, v_rms=0.3, db_rms=0.3, ny=p.cfg.ny),
                         seed=seed, amp=0.05)


def _batch(probs, collapse):
    from cumpic_amd.cumc import Batch, Solver
    boxes = [Solver.from_problem(p) for p in probs]
    for bx in boxes:
        bx.set_zcollapse(collapse)
    return Batch(boxes), boxes


def _seeds(boxes, probs):
    out = []
    for s, p in zip(boxes, probs):
        st = s.download()
        seed = np.zeros(p.cfg.shape + (3,))
        inner = _inner(p.cfg)
        seed[inner] = st["bcell"][inner] * _dv(p)
        out.append(seed)
    return out


def _grads(probs, collapse, nsteps, every, dts=None):
    b, boxes = _batch(probs, collapse)
    used = b.tape_run(nsteps, every, dts=dts)
    active = [bx.zcollapse_state()[1] for bx in boxes]
    seeds = _seeds(boxes, probs)
    g = b.tape_backward(seeds)
    return g, used, active, (b, boxes, seeds)


def _merge_bz(g, cfg):
    """The gradient with the uploaded B_z of planes l = 0 and 1 (array indices 2, 3) as one."""
    h = g.copy()
    if cfg.nz == 1:
        h[3, ..., IBZ] = g[2, ..., IBZ] + g[3, ..., IBZ]
        h[2, ..., IBZ] = 0.0
    return h


def test_collapsed_batch_gradient_matches_full_layout():
    print("[case] collapsed batched reverse vs full layout", flush=True)
    probs = _probs()
    gf, dts, act_f, _ = _grads(probs, False, 5, 2)
    gc, dts_c, act_c, _ = _grads(probs, True, 5, 2, dts=dts)
    assert act_f == [False] * 6
    assert act_c == [True, True, True, True, False, True], act_c
    for i, (a, b, p) in enumerate(zip(gc, gf, probs)):
        a, b = _merge_bz(a, p.cfg), _merge_bz(b, p.cfg)
        scale = np.abs(b).max()
        err = np.abs(a - b).max() / scale
        print(f"  box {i}: max |g_collapsed - g_full| / max |g| = {err:.2e}", flush=True)
        assert scale > 0 and err <= 1e-11, (i, err)
        if i == 4:
            assert np.array_equal(a, b)  # the 3D box is reversed in the full layout either way


def test_collapsed_batch_gradient_reproducible():
    print("[case] collapsed batched reverse: checkpoint intervals and a second backward", flush=True)
    probs = _probs()[:3]
    g1, dts, act, (b, boxes, seeds) = _grads(probs, True, 5, 1)
    assert all(act)
    g1b = b.tape_backward(seeds)  # second pass on the same tape
    for x, y in zip(g1, g1b):
        assert np.array_equal(x, y)
    # the tape is reversed in the views it was recorded in, whatever the live states did since:
    # other states uploaded and stepped, one box switched to the full layout
    others = [_box_like(p, seed=4700 + i) for i, p in enumerate(probs)]
    for bx, q in zip(boxes, others):
        bx.upload(q.prim, q.eps)
        bx.initial_update()
    boxes[1].set_zcollapse(False)
    b.run(2)
    g1c = b.tape_backward(seeds)
    for i, (x, y) in enumerate(zip(g1, g1c)):
        assert np.array_equal(x, y), f"box {i} after the live states changed"
    assert boxes[1].zcollapse_state() == (False, False)
    boxes[1].set_zcollapse(True)
    for every in (2, 3, 5):
        g, _, _, _ = _grads(probs, True, 5, every, dts=dts)
        for i, (x, y) in enumerate(zip(g, g1)):
            assert np.array_equal(x, y), f"box {i}, checkpoints every {every}"


def test_collapsed_batch_gradient_central_differences():
    """g . d against (J(v + h d) - J(v - h d)) / 2h on the batch, J = sum over boxes of
    |B|^2/2 dV after 4 steps with frozen time steps, the forward in the collapsed view (bit-
    identical to the full layout); h = 5e-8, retaken at h/4 when a switching surface lies
    within h (as in test_gpu_adjoint)."""
    print("[case] collapsed batched reverse vs central differences", flush=True)
    probs = _probs()[:3]
    g, dts, _, (b, boxes, _) = _grads(probs, True, 4, 2)
    rng = np.random.default_rng(4400)

    def bfield(prims):
        for s, p, pr in zip(boxes, probs, prims):
            s.upload(pr, p.eps)
            s.initial_update()
        b.run(4, dts=dts)
        return [s.download()["bcell"][_inner(p.cfg)] for s, p in zip(boxes, probs)]

    for d in range(4):
        dirs = [rng.uniform(-1.0, 1.0, size=p.prim[_inner(p.cfg) + (slice(1, 4),)].shape) for p in probs]
        ad = sum(float(np.sum(gi[_inner(p.cfg) + (slice(1, 4),)] * di)) for gi, p, di in zip(g, probs, dirs))
        mag = sum(float(np.sum(np.abs(gi[_inner(p.cfg) + (slice(1, 4),)] * di))) for gi, p, di in zip(g, probs, dirs))
        h, tries, rel = 5e-8, [], 1.0
        for _ in range(4):
            pp, pm = [], []
            for p, di in zip(probs, dirs):
                a, c = p.prim.copy(), p.prim.copy()
                a[_inner(p.cfg) + (slice(1, 4),)] += h * di
                c[_inner(p.cfg) + (slice(1, 4),)] -= h * di
                pp.append(a)
                pm.append(c)
            bp, bm = bfield(pp), bfield(pm)
            dj = sum(0.5 * float(np.sum((x - y) * (x + y))) * _dv(p) for x, y, p in zip(bp, bm, probs))
            fd = dj / (2 * h)
            rel = min(abs(fd - ad) / abs(ad), abs(fd - ad) / mag)
            tries.append(f"h={h:.1e}: {abs(fd - ad) / abs(ad):.2e}")
            if rel <= 1e-6:
                break
            h /= 4
        print(f"  dir {d}: adjoint {ad:.12e} | " + ", ".join(tries), flush=True)
        assert rel <= 1e-6, (d, tries)
