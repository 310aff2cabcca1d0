"""Pins of reference-specific choices by checks that do not restate the oracle's own formula.

* HLL outer speeds (solver/rsolver.f90:110-115): s_L = min(v_L - c_fL, u_roe - c_roe) with c_roe
  the sqrt(rho)-weighted mean of the two FAST magnetosonic speeds.  Low-beta states with B_n != 0,
  where the Roe-average bound is the active one and the fast speed is far above the sound speed:
  the flux matches only when c_roe averages the closed-form fast speeds, and the same bound built
  from sound speeds is shown to miss.
* finddt's width rule (rungekutta.f90:230-247): dt = cfl min_i(dx_i) / max_i(|v_i| + c_f,i).  The
  rule is invariant under permuting the cell widths among the axes (a per-axis min_i dx_i/lambda_i
  is not), and its dt obeys every one-dimensional Courant condition.
* Ghost passes (boundary.f90): a periodic box must step exactly as the same data tiled twice per
  axis (every stencil sees the same values), and a reflecting x wall as the mirror-extended
  periodic box (v_x and B_x odd), to round-off.  A wrong pass range, pass order or parity factor
  leaves a stale or mis-signed ghost inside some stencil and fails one of them.
"""
from __future__ import annotations

import math

import numpy as np

from cumpic_amd import ics
from cumpic_amd.config import Config, uniform_faces
from oracle import oracle as orc
from oracle.oracle import Oracle

RNG = np.random.default_rng(20261203)


def _fast(rho, p, b, n, g):
    """Fast magnetosonic speed along axis n from its closed form (units with B^2/2 pressure)."""
    a2 = g * p / rho
    va2 = float(b @ b) / rho
    return math.sqrt(0.5 * (a2 + va2 + math.sqrt((a2 + va2) ** 2 - 4.0 * a2 * b[n] ** 2 / rho)))


def _flux(s, g):
    rho, v, p, B = s[0], s[1:4], s[4], s[5:8]
    pt = p + 0.5 * B @ B
    E = p / (g - 1) + 0.5 * rho * v @ v + 0.5 * B @ B
    f = np.zeros(8)
    f[0] = rho * v[0]
    f[1:4] = rho * v[0] * v - B[0] * B
    f[1] += pt
    f[4] = (E + pt) * v[0] - B[0] * (v @ B)
    f[5] = 0.0
    f[6] = v[0] * B[1] - v[1] * B[0]
    f[7] = v[0] * B[2] - v[2] * B[0]
    return f


def _cons(s, g):
    rho, v, p, B = s[0], s[1:4], s[4], s[5:8]
    return np.array([rho, *(rho * v), p / (g - 1) + 0.5 * rho * v @ v + 0.5 * B @ B, *B])


def _hll(sl, sr, g, cbar_from):
    wl, wr = math.sqrt(sl[0]), math.sqrt(sr[0])
    cl, cr = cbar_from(sl), cbar_from(sr)
    ub = (wl * sl[1] + wr * sr[1]) / (wl + wr)
    cb = (wl * cl + wr * cr) / (wl + wr)
    fl_, fr_ = _fast(sl[0], sl[4], sl[5:8], 0, g), _fast(sr[0], sr[4], sr[5:8], 0, g)
    SL, SR = min(sl[1] - fl_, ub - cb), max(sr[1] + fr_, ub + cb)
    roe_active = (ub - cb < sl[1] - fl_) or (ub + cb > sr[1] + fr_)
    FL, FR = _flux(sl, g), _flux(sr, g)
    if SL >= 0:
        return FL, roe_active
    if SR <= 0:
        return FR, roe_active
    return (SR * FL - SL * FR + SL * SR * (_cons(sr, g) - _cons(sl, g))) / (SR - SL), roe_active


def test_hll_outer_speeds_average_fast_speeds():
    g = 5.0 / 3.0
    active, missed = 0, 0
    for _ in range(600):
        bn = RNG.uniform(0.5, 1.5) * RNG.choice([-1, 1])
        sl = np.array([RNG.uniform(0.05, 2), *RNG.uniform(-0.5, 0.5, 3), RNG.uniform(0.001, 0.05), bn,
                       *RNG.uniform(-1.5, 1.5, 2)])
        sr = np.array([RNG.uniform(0.05, 2), *RNG.uniform(-0.5, 0.5, 3), RNG.uniform(0.001, 0.05), bn,
                       *RNG.uniform(-1.5, 1.5, 2)])
        f, _ = orc.hll(1, g, sl, sr)
        want, act = _hll(sl, sr, g, lambda s: _fast(s[0], s[4], s[5:8], 0, g))
        assert np.allclose(f[[0, 1, 2, 3, 4, 6, 7]], want[[0, 1, 2, 3, 4, 6, 7]], rtol=1e-11, atol=1e-12)
        if act:
            active += 1
            alt, _ = _hll(sl, sr, g, lambda s: math.sqrt(g * s[4] / s[0]))
            missed += not np.allclose(f[[0, 1, 2, 3, 4, 6, 7]], alt[[0, 1, 2, 3, 4, 6, 7]], rtol=1e-6, atol=1e-9)
    # the Roe-average bound decides many of these faces, and a sound-speed average misses there
    assert active >= 60, active
    assert missed >= active // 2, (missed, active)


def _uniform_dt(widths, state, cfl=0.6, g=5.0 / 3.0):
    n = 4
    cfg = Config(nx=n, ny=n, nz=n, boundary_flag=(0,) * 6, ggas=g, cfl=cfl)
    o = Oracle(cfg, *(uniform_faces(n, 0, w * n) for w in widths))
    prim = np.zeros(cfg.shape + (8,))
    prim[...] = state
    o.set_state(prim, np.full(cfg.shape, state[4] / state[0] / (g - 1)))
    o.initial_update()
    dt = o.finddt()
    o.close()
    return dt


def test_finddt_width_rule_properties():
    g, cfl = 5.0 / 3.0, 0.6
    for _ in range(6):
        state = np.array([RNG.uniform(0.5, 2), *RNG.uniform(-1, 1, 3), RNG.uniform(0.1, 2), *RNG.uniform(-1, 1, 3)])
        w = RNG.uniform(0.3, 3, 3)
        dts = [_uniform_dt(w[list(p)], state, cfl, g) for p in ((0, 1, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))]
        assert dts[0] == dts[1] == dts[2] == dts[3]
        lam = [abs(state[1 + i]) + _fast(state[0], state[4], state[5:8], i, g) for i in range(3)]
        for i in range(3):
            assert dts[0] * lam[i] / w[i] <= cfl * (1 + 1e-12)
        # the per-axis rule would give a different step whenever the ratios differ
        per_axis = cfl * min(w[i] / lam[i] for i in range(3))
        if abs(per_axis - dts[0]) > 1e-9 * per_axis:
            assert dts[0] < per_axis


def _step(prob, nsteps):
    o = Oracle.from_problem(prob)
    o.initial_update()
    for _ in range(nsteps):
        o.rungekutta(o.finddt())
    st = o.get_state()
    o.close()
    return st


def _inner(st, c):
    return {k: st[k][3:3 + c.nz, 3:3 + c.ny, 3:3 + c.nx] for k in ("prim", "eps")}


def test_periodic_box_equals_its_tiling():
    """8^3 box of width 1/8 cells against the same data tiled 2x2x2 on [0,2]^3: every face
    coordinate is a multiple of 1/8, so all cell widths are exact and the two runs must agree
    bit for bit in every copy of the box."""
    n = 8
    base = ics.randomize(ics.small_periodic(n, seed=1201), seed=9, amp=0.2)
    c = base.cfg
    big = ics.small_periodic(2 * n, seed=1201)
    big.xf = big.yf = big.zf = uniform_faces(2 * n, 0.0, 2.0)
    big.cfg.ggas, big.cfg.cfl = c.ggas, c.cfl
    inn, innb = slice(3, 3 + n), slice(3, 3 + 2 * n)
    for arr_b, arr_s in ((big.prim, base.prim), (big.eps, base.eps)):
        core = arr_s[inn, inn, inn]
        arr_b[...] = 0.0
        arr_b[innb, innb, innb] = np.tile(core, (2, 2, 2) + (1,) * (core.ndim - 3))
    # the face-B planes of the low faces (index 2: j, k or l = 0) are the periodic images of
    # the faces n (2n in the tiled box), as the uploaded state of a periodic run has them
    for prob, m in ((base, n), (big, 2 * n)):
        r = slice(3, 3 + m)
        for v, ax in ((5, 2), (6, 1), (7, 0)):
            dst, src = [r, r, r], [r, r, r]
            dst[ax], src[ax] = 2, 2 + m
            prob.prim[tuple(dst) + (v,)] = prob.prim[tuple(src) + (v,)]
    a = _inner(_step(base, 3), c)
    b = _inner(_step(big, 3), big.cfg)
    for k in a:
        for t in ((0, 0, 0), (1, 0, 1), (0, 1, 1), (1, 1, 0), (1, 1, 1)):
            sub = b[k][n * t[0]:n * t[0] + n, n * t[1]:n * t[1] + n, n * t[2]:n * t[2] + n]
            assert np.array_equal(sub, a[k]), (k, t)


def test_reflecting_wall_equals_mirrored_periodic_box():
    n = 8
    base = ics.randomize(ics.small_periodic(n, seed=1202), seed=11, amp=0.2)
    base.prim[..., 5] = 0.0  # B_x = 0 on every x face (odd about both walls)
    base.cfg.boundary_flag = (2, 2, 0, 0, 0, 0)
    big = ics.small_periodic(2 * n, seed=1202)
    big.cfg.ny, big.cfg.nz = n, n
    big.xf = uniform_faces(2 * n, 0.0, 2.0)
    big.yf, big.zf = base.yf, base.zf
    big.cfg.ggas, big.cfg.cfl = base.cfg.ggas, base.cfg.cfl
    big.prim = np.zeros(big.cfg.shape + (8,))
    big.eps = np.zeros(big.cfg.shape)
    core_p = base.prim[3:3 + n, 3:3 + n, 3:3 + n]
    core_e = base.eps[3:3 + n, 3:3 + n, 3:3 + n]
    mir_p = core_p[:, :, ::-1].copy()
    mir_p[..., 1] *= -1.0
    big.prim[3:3 + n, 3:3 + n, 3:3 + 2 * n] = np.concatenate([core_p, mir_p], axis=2)
    big.eps[3:3 + n, 3:3 + n, 3:3 + 2 * n] = np.concatenate([core_e, core_e[:, :, ::-1]], axis=2)
    # y / z face B: the low faces of the periodic y and z axes are images of faces n
    for b_, m in ((base, n), (big, 2 * n)):
        b_.prim[2, 3:3 + n, 3:3 + m, 7] = b_.prim[2 + n, 3:3 + n, 3:3 + m, 7]
        b_.prim[3:3 + n, 2, 3:3 + m, 6] = b_.prim[3:3 + n, 2 + n, 3:3 + m, 6]
    a = _inner(_step(base, 2), base.cfg)
    b = _inner(_step(big, 2), big.cfg)
    for k in a:
        left = b[k][:, :, :n]
        right = b[k][:, :, n:][:, :, ::-1]
        scale = float(np.max(np.abs(a[k])))
        assert np.max(np.abs(left - a[k])) <= 1e-12 * scale, k
        if k == "prim":
            right = right.copy()
            right[..., 1] *= -1.0
            right[..., 5] = b[k][:, :, n - 1:2 * n - 1, 5][:, :, ::-1] * -1.0
            assert np.max(np.abs(right[..., :5] - a[k][..., :5])) <= 1e-12 * scale
