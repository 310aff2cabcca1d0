"""The diagnostic build libcumc_sig.so and its per-box branch signatures.

libcumc_sig.so is the library compiled with CUMC_BRANCH_SIG: every data-dependent decision of the
reference path (PPM / TVD-minmod limiter switches, HLL speed selections and region, the
signal-speed clamp, the floors) is hashed per box by the batched forward kernels.  Checked here:
- its arithmetic is the product's: a batched run is bit-identical to libcumc.so's (final state and
  time steps);
- signatures are deterministic, and a perturbation of one box moves that box's signature only;
- libcumc.so refuses cumc_batch_branch_sig;
- central differences of small 2D boxes agree with the batched reverse mode to 1e-6 relative to
  |g_b . d_b| on every box-direction pair whose +-h runs reproduce the unperturbed signature
  (no kink crossed), while the pairs with a flip miss by orders of magnitude more.
"""
from __future__ import annotations

import numpy as np
import pytest

from cumpic_amd import ics
from cumpic_amd.ics import ix

pytestmark = pytest.mark.gpu


def _inner(cfg):
    return (slice(ix(1), ix(cfg.nz) + 1), slice(ix(1), ix(cfg.ny) + 1), slice(ix(1), ix(cfg.nx) + 1))


def _dv(p):
    return (p.xf[4] - p.xf[3]) * (p.yf[4] - p.yf[3]) * (p.zf[4] - p.zf[3])


def _boxes(probs, variant, vs=None):
    from cumpic_amd.cumc import Solver
    out = []
    for i, p in enumerate(probs):
        s = Solver(p.cfg, p.xf, p.yf, p.zf, variant=variant)
        prim = p.prim.copy()
        if vs is not None:
            prim[_inner(p.cfg) + (slice(1, 4),)] = vs[i]
        s.upload(prim, p.eps)
        s.initial_update()
        out.append(s)
    return out


def _reset(boxes, probs, vs=None):
    for i, (s, p) in enumerate(zip(boxes, probs)):
        prim = p.prim.copy()
        if vs is not None:
            prim[_inner(p.cfg) + (slice(1, 4),)] = vs[i]
        s.upload(prim, p.eps)
        s.initial_update()


def test_sig_build_is_bitwise_the_product_and_deterministic():
    from cumpic_amd.cumc import Batch, CumcError
    probs = ics.batch_boxes(nbox=3, sizes=(16, 24), steps=6)
    a, b = _boxes(probs, ""), _boxes(probs, "sig")
    A, B = Batch(a), Batch(b)
    da = A.run(6, graph_steps=3)
    B.branch_sig()
    db = B.run(6, graph_steps=3)
    s1 = B.branch_sig()
    assert np.array_equal(da, db)
    for x, y in zip(a, b):
        u, v = x.download(), y.download()
        for k in ("prim", "eps", "bcell"):
            assert np.array_equal(u[k], v[k]), k
    assert (s1 != 0).all()
    print(f"  signatures {[hex(int(x)) for x in s1]}", flush=True)
    # the same run again (time steps frozen to the ones found): the same decisions
    _reset(b, probs)
    B.run(6, graph_steps=3, dts=db)
    assert np.array_equal(B.branch_sig(), s1)
    # direct launches instead of the graph: the same work items, the same signature
    _reset(b, probs)
    B.run(6, graph_steps=0, dts=db)
    assert np.array_equal(B.branch_sig(), s1)
    with pytest.raises(CumcError, match="diagnostic build"):
        A.branch_sig()
    for x in (A, B, *a, *b):
        x.close()


def test_signature_moves_only_in_the_perturbed_box():
    from cumpic_amd.cumc import Batch
    probs = ics.batch_boxes(nbox=3, sizes=(16,), steps=4)
    v0 = [p.prim[_inner(p.cfg) + (slice(1, 4),)].copy() for p in probs]
    b = _boxes(probs, "sig")
    B = Batch(b)
    dts = B.run(4, graph_steps=2)
    base = B.branch_sig()
    rng = np.random.default_rng(7)
    vs = [v.copy() for v in v0]
    vs[1] = vs[1] + 1e-2 * rng.normal(size=vs[1].shape)
    _reset(b, probs, vs)
    B.run(4, graph_steps=2, dts=dts)
    moved = B.branch_sig()
    assert moved[0] == base[0] and moved[2] == base[2]
    assert moved[1] != base[1]
    for x in (B, *b):
        x.close()


def test_fd_agrees_on_flip_free_pairs():
    from cumpic_amd.cumc import Batch
    # h = 1e-5: the rounding of the forward runs (a few units in the last place of each B) enters
    # the central difference as about 1e-6 of |g_b . d_b| at h = 1e-6 on these boxes, and ten
    # times less here
    nsteps, every, ndir, h = 6, 3, 6, 1e-5
    probs = ics.batch_boxes(nbox=4, sizes=(16,), steps=nsteps)
    inner = [_inner(p.cfg) for p in probs]
    # gradient of J_b = sum |B|^2/2 dV per box from the product build's batched tape
    a = _boxes(probs, "")
    for s, p in zip(a, probs):
        s.upload(p.prim, p.eps)
    A = Batch(a)
    dts = A.tape_run(nsteps, every)
    seeds = []
    for s, p, ii in zip(a, probs, inner):
        seed = np.zeros(p.cfg.shape + (3,))
        seed[ii] = s.download()["bcell"][ii] * _dv(p)
        seeds.append(seed)
    grads = [g[ii + (slice(1, 4),)].copy() for g, ii in zip(A.tape_backward(seeds), inner)]
    A.tape_free()
    # forward runs of the diagnostic build
    v0 = [p.prim[ii + (slice(1, 4),)].copy() for p, ii in zip(probs, inner)]
    b = _boxes(probs, "sig")
    B = Batch(b)

    def run(vs):
        _reset(b, probs, vs)
        B.branch_sig()
        B.run(nsteps, graph_steps=every, dts=dts)
        return [s.download()["bcell"][ii] for s, ii in zip(b, inner)], B.branch_sig()

    def half(x, y, w):
        return 0.5 * float(np.sum((x - y) * (x + y))) * w

    _, sig0 = run(v0)
    rng = np.random.default_rng(11)
    free_ok, free_n, flipped, misses = 0, 0, [], []
    for d in range(ndir):
        dirs = [rng.normal(size=g.shape) for g in grads]
        nrm = np.sqrt(sum(float(np.sum(x * x)) for x in dirs))
        dirs = [x / nrm for x in dirs]
        bp, sp = run([v + h * x for v, x in zip(v0, dirs)])
        bm, sm = run([v - h * x for v, x in zip(v0, dirs)])
        for i, p in enumerate(probs):
            w = _dv(p)
            ad = float(np.sum(grads[i] * dirs[i]))
            fd = half(bp[i], bm[i], w) / (2 * h)
            err = abs(fd - ad) / abs(ad)
            free = sp[i] == sig0[i] and sm[i] == sig0[i]
            print(f"  dir {d} box {i}: adjoint {ad:.12e} central {fd:.12e} rel {err:.2e} "
                  f"{'flip-free' if free else 'branch flip'}", flush=True)
            if free:
                free_n += 1
                free_ok += err <= 1e-6
                if err > 1e-6:
                    misses.append((d, i, ad, fd, err))
            else:
                flipped.append(err)
    print(f"  flip-free pairs within 1e-6: {free_ok} of {free_n}; flipped pairs {len(flipped)}"
          + (f", their median error {np.median(flipped):.2e}" if flipped else ""), flush=True)
    assert free_n >= 1 and not misses, misses
    for x in (A, B, *a, *b):
        x.close()
