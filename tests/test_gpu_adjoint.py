"""Reverse mode of initial_update + N RK3 steps against central finite differences.

Objective: J = sum over interior cells of |bcell|^2/2 dV after N steps with the time steps
frozen.  The gradient with respect to the uploaded velocity field from cumc_tape_backward must
match (J(v + h d) - J(v - h d)) / 2h along seeded random directions d (entries uniform in
[-1, 1]) to 1e-6 relative, with J evaluated by the plain forward path on the GPU and, for some
directions, by the CPU restatement (whose forward is bit-identical, so its differences are the
oracle's own).

Step size: the limiters, HLL region choice and floors make the step piecewise smooth, and a
central difference that straddles a switching surface is off by O(1) in that cell's
contribution; the number of surfaces crossed grows with h.  The difference J(v+) - J(v-) is
formed as sum (B+ - B-)(B+ + B-)/2 dV, which keeps the rounding floor near 1e-17; h starts at
5e-8 and a direction that misses is retaken with h/4 (a switching surface within h), see
_fd_check.  A direction along which the contributions g_i d_i largely cancel has a small
derivative and the same absolute difference noise, so the error is also measured against the
cancellation-free magnitude sum |g_i d_i|; either measure within 1e-6 passes.
"""
from __future__ import annotations

import numpy as np
import pytest

from cumpic_amd import ics
from cumpic_amd.config import PPMC, TVDMM
from cumpic_amd.ics import ix
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _inner(cfg):
    return (slice(ix(1), ix(cfg.nz) + 1), slice(ix(1), ix(cfg.ny) + 1), slice(ix(1), ix(cfg.nx) + 1))


def _dv(p):
    return (p.xf[4] - p.xf[3]) * (p.yf[4] - p.yf[3]) * (p.zf[4] - p.zf[3])


def _B_gpu(s, prob, prim, dts):
    s.upload(prim, prob.eps)
    s.initial_update()
    for dt in dts:
        s.rungekutta(float(dt))
    return s.download()["bcell"][_inner(prob.cfg)]


def _B_oracle(prob, prim, dts):
    o = Oracle.from_problem(prob)
    o.set_state(prim, prob.eps)
    o.initial_update()
    for dt in dts:
        o.rungekutta(float(dt))
    b = o.get_state()["bcell"][_inner(prob.cfg)]
    o.close()
    return b


def _dJ(bp, bm, prob):
    """J(+) - J(-) for J = sum |B|^2/2 dV, formed from the field differences."""
    return 0.5 * float(np.sum((bp - bm) * (bp + bm))) * _dv(prob)


def _gradient(s, prob, nsteps, every, dts=None):
    s.upload(prob.prim, prob.eps)
    used = s.tape_run(nsteps, every, dts)
    st = s.download()
    seed = np.zeros(prob.cfg.shape + (3,))
    inner = _inner(prob.cfg)
    seed[inner] = st["bcell"][inner] * _dv(prob)
    gp, ge = s.tape_backward(seed_bcell=seed)
    return gp, ge, used


def _fd_check(prob, nsteps, every, ndir, seed, h=5e-8, oracle_dirs=0, refine=3):
    """Directional derivatives of the adjoint against central differences of the forward map.

    The adjoint is the derivative on the branch taken at v.  A central difference with step h
    matches it only if no switching surface lies within h of v along d; when a direction misses
    1e-6 at h, the difference is retaken at h/4 (up to `refine` times): a surface at distance
    delta stops mattering once h < delta, while a wrong adjoint matches at no h."""
    from cumpic_amd.cumc import Solver
    s = Solver.from_problem(prob)
    gp, ge, dts = _gradient(s, prob, nsteps, every)
    inner = _inner(prob.cfg)
    gv = gp[inner + (slice(1, 4),)]
    assert np.isfinite(gp).all() and np.abs(gv).max() > 0
    rng = np.random.default_rng(seed)
    worst, refined = 0.0, 0
    for d in range(ndir):
        dirn = rng.uniform(-1.0, 1.0, size=gv.shape)
        ad = float(np.sum(gv * dirn))
        mag = float(np.sum(np.abs(gv * dirn)))
        hh, tries = h, []
        for r in range(refine + 1):
            pp, pm = prob.prim.copy(), prob.prim.copy()
            pp[inner + (slice(1, 4),)] += hh * dirn
            pm[inner + (slice(1, 4),)] -= hh * dirn
            fd = _dJ(_B_gpu(s, prob, pp, dts), _B_gpu(s, prob, pm, dts), prob) / (2 * hh)
            rel = min(abs(fd - ad) / abs(ad), abs(fd - ad) / mag)
            tries.append(f"h={hh:.1e}: {abs(fd - ad) / abs(ad):.2e} ({abs(fd - ad) / mag:.2e} of sum|g d|)")
            if d < oracle_dirs and r == 0:
                bpo, bmo = _B_oracle(prob, pp, dts), _B_oracle(prob, pm, dts)
                assert _dJ(bpo, bmo, prob) / (2 * hh) == fd  # the CPU restatement's forward is bit-identical
            if rel <= TOL:
                break
            hh /= 4
        print(f"  dir {d}: adjoint {ad:.12e} central {fd:.12e} | " + ", ".join(tries), flush=True)
        assert rel <= TOL, (d, ad, fd, tries)
        worst = max(worst, rel)
        refined += len(tries) > 1
    print(f"  directions needing a smaller h: {refined} of {ndir}", flush=True)
    s.close()
    return worst


@pytest.mark.parametrize("ny", [16, 6])
def test_adjoint_fd_2d_periodic_ppm(ny):
    # ny = 16 reverses the y sweep wave-parallel, ny = 6 by per-pencil walks
    print(f"[case] adjoint vs FD, 2D 16x{ny} periodic PPM, 6 steps, checkpoints every 4", flush=True)
    prob = ics.randomize(ics.spectral_box(n=16, seed=3000, ndim=2, v_rms=0.3, db_rms=0.3, ny=ny), seed=1, amp=0.05)
    w = _fd_check(prob, nsteps=6, every=4, ndir=8, seed=31, oracle_dirs=2)
    print(f"  worst rel {w:.2e}")


@pytest.mark.parametrize("recon", [PPMC, TVDMM])
def test_adjoint_fd_3d(recon):
    print(f"[case] adjoint vs FD, 3D 8^3 periodic recon={recon}, 4 steps, checkpoints every 3", flush=True)
    prob = ics.randomize(ics.small_periodic(8, seed=3100), seed=2, amp=0.05)
    prob.cfg.recon = recon
    _fd_check(prob, nsteps=4, every=3, ndir=6, seed=32)


def test_adjoint_fd_3d_smooth_ppm():
    print("[case] adjoint vs FD, 3D 12^3 smooth spectral PPM, 3 steps, checkpoints every 2", flush=True)
    prob = ics.spectral_box(n=12, seed=3150, v_rms=0.1, db_rms=0.1)
    _fd_check(prob, nsteps=3, every=2, ndir=6, seed=34)


def test_adjoint_fd_outflow_and_reflecting():
    print("[case] adjoint vs FD, 3D 8^3 outflow x, periodic y, reflecting z", flush=True)
    prob = ics.randomize(ics.small_periodic(8, seed=3200), seed=3, amp=0.05)
    # boundary flags: x outflow, y periodic, z reflecting (parity copies with bfac)
    prob.cfg.boundary_flag = (1, 1, 0, 0, 2, 2)
    _fd_check(prob, nsteps=3, every=2, ndir=6, seed=33)


def test_checkpoint_interval_does_not_change_the_gradient():
    from cumpic_amd.cumc import Solver
    print("[case] checkpoint interval invariance", flush=True)
    prob = ics.randomize(ics.spectral_box(n=12, seed=3300, ndim=2, v_rms=0.3, db_rms=0.3), seed=4, amp=0.05)
    s = Solver.from_problem(prob)
    g1, e1, d1 = _gradient(s, prob, 7, 1)
    g7, e7, d7 = _gradient(s, prob, 7, 7)
    g3, e3, d3 = _gradient(s, prob, 7, 3, dts=d1)
    assert np.array_equal(d1, d7) and np.array_equal(d1, d3)
    assert np.array_equal(g1, g7) and np.array_equal(g1, g3)
    assert np.array_equal(e1, e3)


def test_tape_forward_matches_run_and_restores_state():
    from cumpic_amd.cumc import Solver
    print("[case] tape forward == plain run", flush=True)
    prob = ics.randomize(ics.small_periodic(8, seed=3400), seed=5, amp=0.05)
    a = Solver.from_problem(prob)
    b = Solver.from_problem(prob)
    dts = a.tape_run(5, 2)
    b.initial_update()
    db, _ = b.run(5)
    assert np.array_equal(dts, db)
    sa, sb = a.download(), b.download()
    for k in ("prim", "eps", "bcell"):
        assert np.array_equal(sa[k], sb[k]), k
    seed = np.ones(prob.cfg.shape + (3,))
    a.tape_backward(seed_bcell=seed)
    sa = a.download()
    for k in ("prim", "eps", "bcell"):
        assert np.array_equal(sa[k], sb[k]), k


def test_autograd_function_matches_tape():
    import torch
    from cumpic_amd.autograd import Rollout, initial_velocity, magnetic_energy
    from cumpic_amd.cumc import Solver
    print("[case] torch.autograd.Function", flush=True)
    prob = ics.randomize(ics.spectral_box(n=12, seed=3500, ndim=2, v_rms=0.3, db_rms=0.3), seed=6, amp=0.05)
    s = Solver.from_problem(prob)
    gp, _, dts = _gradient(s, prob, 5, 2)
    v0 = initial_velocity(prob).clone().requires_grad_(True)
    _, bc = Rollout.apply(v0, s, prob, 5, 2, dts)
    J = magnetic_energy(bc, prob)
    J.backward()
    want = gp[_inner(prob.cfg) + (slice(1, 4),)]
    # same seed up to how torch rounds d(|B|^2/2 dV)/dB
    np.testing.assert_allclose(v0.grad.numpy(), want, rtol=1e-12, atol=1e-14 * np.abs(want).max())


@pytest.mark.parametrize("every", [2, 3, 5])
def test_batched_reverse_matches_single_box_tapes(every):
    """The batched reverse against per-box tapes (checkpoint every 2), with the batch
    checkpointing every 2, 3 (last segment starts on an odd step) or 5 steps (one segment: its
    stage states come from the taping forward and nothing is re-run), for frozen and for found
    time steps."""
    from cumpic_amd.cumc import Batch, Solver
    print(f"[case] batched reverse vs per-box tapes, batch checkpoints every {every}", flush=True)
    probs = [ics.randomize(ics.spectral_box(n=n, seed=3600 + i, ndim=2, v_rms=0.3, db_rms=0.3), seed=i, amp=0.05)
             for i, n in enumerate((12, 20, 9))]
    probs.append(ics.randomize(ics.small_periodic(8, seed=3700), seed=7, amp=0.05))
    # short y pencils: this box's y sweep takes the per-pencil reverse in the same batch
    probs.append(ics.randomize(ics.spectral_box(n=14, seed=3800, ndim=2, v_rms=0.3, db_rms=0.3, ny=5), seed=8,
                               amp=0.05))
    singles = [Solver.from_problem(p) for p in probs]
    ref = []
    for s, p in zip(singles, probs):
        gp, _, dts = _gradient(s, p, 5, 2)
        ref.append((gp, dts))
    # per-box tapes run in the full layout; so does the batch here (its z-collapsed reverse is
    # checked against this one in test_gpu_zcollapse_adjoint)
    boxes = [Solver.from_problem(p) for p in probs]
    for bx in boxes:
        bx.set_zcollapse(False)
    b = Batch(boxes)
    dts = b.tape_run(5, every, dts=np.array([r[1] for r in ref]).T)
    seeds = []
    for s, p in zip(boxes, probs):
        st = s.download()
        seed = np.zeros(p.cfg.shape + (3,))
        inner = _inner(p.cfg)
        seed[inner] = st["bcell"][inner] * _dv(p)
        seeds.append(seed)
    grads = b.tape_backward(seeds)
    for i, (g, r) in enumerate(zip(grads, ref)):
        assert np.array_equal(g, r[0]), f"box {i}"
    if every >= 5:
        assert b.tape_recompute_ms() < 0.05 * b.last_ms()
    # non-frozen run reproduces the per-box CFL sequences, and the same gradients
    for s, p in zip(boxes, probs):
        s.upload(p.prim, p.eps)
    d2 = b.tape_run(5, every)
    assert np.array_equal(d2, np.array([r[1] for r in ref]).T)
    grads = b.tape_backward(seeds)
    for i, (g, r) in enumerate(zip(grads, ref)):
        assert np.array_equal(g, r[0]), f"box {i}, found time steps"
    assert b.last_ms() >= 0.0


def test_second_backward_on_one_tape_is_identical():
    """A tape stays valid until it is freed: a second backward pass on the same tape (as torch's
    retain_graph or several vector-Jacobian products make) gives the same gradients bit for bit,
    also when the first pass re-ran earlier checkpoint segments into the segment buffers
    (nsteps 5, checkpoints every 2: segments [4], [2,3], [0,1])."""
    from cumpic_amd.cumc import Batch, Solver
    print("[case] two backward passes on one tape, single context and batch", flush=True)
    probs = [ics.randomize(ics.spectral_box(n=n, seed=3900 + i, ndim=2, v_rms=0.3, db_rms=0.3), seed=i, amp=0.05)
             for i, n in enumerate((12, 10))]
    p = probs[0]
    s = Solver.from_problem(p)
    s.upload(p.prim, p.eps)
    s.tape_run(5, 2)
    seed = np.zeros(p.cfg.shape + (3,))
    inner = _inner(p.cfg)
    seed[inner] = s.download()["bcell"][inner] * _dv(p)
    g1, e1 = s.tape_backward(seed_bcell=seed)
    g2, e2 = s.tape_backward(seed_bcell=seed)
    assert np.abs(g1).max() > 0
    assert np.array_equal(g1, g2) and np.array_equal(e1, e2)
    s.close()
    boxes = [Solver.from_problem(q) for q in probs]
    for bx in boxes:
        bx.set_zcollapse(False)  # box 0 is compared with the full-layout single-context tape
    b = Batch(boxes)
    b.tape_run(5, 2)
    seeds = []
    for bx, q in zip(boxes, probs):
        sd = np.zeros(q.cfg.shape + (3,))
        inn = _inner(q.cfg)
        sd[inn] = bx.download()["bcell"][inn] * _dv(q)
        seeds.append(sd)
    ga = b.tape_backward(seeds)
    gb = b.tape_backward(seeds)
    for i, (x, y) in enumerate(zip(ga, gb)):
        assert np.abs(x).max() > 0 and np.array_equal(x, y), f"box {i}"
    assert np.array_equal(ga[0], g1)
    b.close()
    for bx in boxes:
        bx.close()
