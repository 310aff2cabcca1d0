"""Adjoints of the point numerics (HLL with signal speeds, PPM slope and edge values, TVD
minmod) against central differences of their primals, on the CPU.

The primals and adjoints are the same source the HIP kernels compile (physics.hpp,
adjoint.hpp), built for the host by tests/native/point_adjoint.hip.  A sample is skipped only
where x - h and x + h take different branches (HLL region, min/max picks, clamps, limiter
cases): there the primal has a kink and no derivative.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "point_adjoint.hip")
OUT = os.path.join(HERE, "native", "_build", "libpoint_adjoint.so")
D = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def lib():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    if not os.path.exists(OUT) or os.path.getmtime(OUT) < max(
            os.path.getmtime(p) for p in (SRC, os.path.join(HERE, "..", "cumpic_amd", "csrc", "adjoint.hpp"),
                                          os.path.join(HERE, "..", "cumpic_amd", "csrc", "physics.hpp"))):
        cmd = ["/opt/rocm/bin/hipcc", "-O1", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
               "--offload-arch=gfx950", "-o", OUT + ".tmp", SRC]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        assert r.returncode == 0, r.stdout
        os.replace(OUT + ".tmp", OUT)
    L = ctypes.CDLL(OUT)
    L.pt_hll.argtypes = [ctypes.c_double, D, D, D]
    L.pt_hll_sig.argtypes = [ctypes.c_double, D, D]
    L.pt_hll_adj.argtypes = [ctypes.c_double, D, D, D, D, D]
    L.pt_signalspeed.argtypes = [D]
    L.pt_signalspeed.restype = ctypes.c_double
    L.pt_signalspeed_adj.argtypes = [D, ctypes.c_double, D]
    L.pt_ppm_edges.argtypes = [D, D, D, D]
    L.pt_ppm_edges_sig.argtypes = [D, D, D]
    L.pt_ppm_edges_adj.argtypes = [D, D, D, D, D, D]
    L.pt_ppm_slope.argtypes = [ctypes.c_double, ctypes.c_double, D]
    L.pt_ppm_slope.restype = ctypes.c_double
    L.pt_ppm_slope_sig.argtypes = [ctypes.c_double, ctypes.c_double, D]
    L.pt_ppm_slope_adj.argtypes = [ctypes.c_double, ctypes.c_double, D, ctypes.c_double, D]
    L.pt_tvd.argtypes = [ctypes.c_double, ctypes.c_double, D, D]
    L.pt_tvd_adj.argtypes = [ctypes.c_double, ctypes.c_double, D, D, D]
    return L


def p(a):
    return a.ctypes.data_as(D)


def fd_check(f, sig, x, grad, h_rel=1e-6, rtol=2e-7):
    """grad[i] vs (f(x+h e_i) - f(x-h e_i)) / 2h for every i whose +-h branch signature matches."""
    s0 = sig(x)
    scale = max(1e-300, float(np.max(np.abs(grad))))
    checked = 0
    for i in range(x.size):
        h = h_rel * max(1.0, abs(x[i]))
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        if sig(xp) != s0 or sig(xm) != s0:
            continue
        d = (f(xp) - f(xm)) / (2 * h)
        assert abs(d - grad[i]) <= rtol * max(scale, abs(d)) + 1e-12, (i, d, grad[i], x)
        checked += 1
    return checked


def hll_state(rng):
    rho = rng.uniform(0.3, 2.0)
    vn = rng.uniform(-3.0, 3.0)
    vt = rng.uniform(-1.0, 1.0, 2)
    pr = rng.uniform(0.1, 2.0)
    b = rng.uniform(-1.5, 1.5, 3)
    return np.array([rho, vn, vt[0], vt[1], pr, b[0], b[1], b[2]])


def test_hll_adjoint_vs_central_differences(lib):
    rng = np.random.default_rng(11)
    g = 5.0 / 3.0
    regions = set()
    total = 0
    for _ in range(300):
        L, R = hll_state(rng), hll_state(rng)
        R[5] = L[5]  # shared normal field
        lam = rng.normal(size=7)

        def split(x):
            a, b = x[:8].copy(), x[8:].copy()
            b[5] = a[5]
            return a, b

        def f(x):
            a, b = split(x)
            out = np.zeros(7)
            lib.pt_hll(g, p(a), p(b), p(out))
            return float(lam @ out)

        def sig(x):
            a, b = split(x)
            return lib.pt_hll_sig(g, p(a), p(b))

        lL, lR = np.zeros(8), np.zeros(8)
        lib.pt_hll_adj(g, p(L), p(R), p(lam), p(lL), p(lR))
        grad = np.concatenate([lL, lR])
        grad[5] = lL[5] + lR[5]  # the shared bn is one input
        grad[13] = 0.0           # R's copy of bn is not independent
        x = np.concatenate([L, R])
        regions.add(sig(x) & 3)

        def f2(xx):
            return f(xx)

        # perturb every input except R's bn (index 13), which split() ties to L's
        idx = [i for i in range(16) if i != 13]
        s0 = sig(x)
        scale = float(np.max(np.abs(grad)))
        for i in idx:
            h = 1e-6 * max(1.0, abs(x[i]))
            xp, xm = x.copy(), x.copy()
            xp[i] += h
            xm[i] -= h
            if sig(xp) != s0 or sig(xm) != s0:
                continue
            d = (f2(xp) - f2(xm)) / (2 * h)
            assert abs(d - grad[i]) <= 2e-7 * max(scale, abs(d)) + 1e-12, (i, d, grad[i])
            total += 1
    assert regions == {0, 1, 2}, regions  # left-supersonic, HLL fan and right-supersonic all exercised
    assert total > 4000


def test_signalspeed_adjoint(lib):
    rng = np.random.default_rng(12)
    n = 0
    for _ in range(200):
        x = np.array([rng.uniform(0.2, 2), *rng.uniform(-2, 2, 3), rng.uniform(0.3, 3)])
        lam = rng.normal()
        gr = np.zeros(5)
        lib.pt_signalspeed_adj(p(x), lam, p(gr))
        n += fd_check(lambda y: lam * lib.pt_signalspeed(p(y)), lambda y: 0, x, gr)
    assert n == 1000



def _ppm_coef_table(rng):
    # coefficients of a non-uniform cell from five random widths (same formulas as the tables)
    d = rng.uniform(0.5, 1.5, 5)
    dm2, dm1, dc, dp1, dp2 = d
    a1R = dc / (dc + dp1)
    a1L = dm1 / (dm1 + dc)
    aR = 2.0 * dp1 * dc / (dp1 + dc)
    aL = 2.0 * dc * dm1 / (dc + dm1)
    dXR = dm1 + dc + dp1 + dp2
    dXL = dm2 + dm1 + dc + dp1
    z1R = (dm1 + dc) / (2.0 * dc + dp1)
    z1L = (dm2 + dm1) / (2.0 * dm1 + dc)
    zR = (dp2 + dp1) / (2.0 * dp1 + dc)
    zL = (dp1 + dc) / (2.0 * dc + dm1)
    return np.array([a1L, aL * (z1L - zL) / dXL, (-dm1 * z1L) / dXL, dc * zL / dXL,
                     a1R, aR * (z1R - zR) / dXR, (-dc * z1R) / dXR, dp1 * zR / dXR])


def test_ppm_edges_adjoint(lib):
    rng = np.random.default_rng(13)
    n = 0
    sigs = set()
    for _ in range(400):
        c = _ppm_coef_table(rng)
        x = rng.normal(size=6)
        x[3:] *= 0.5
        lam = rng.normal(size=2)

        def f(y):
            o = np.zeros(2)
            lib.pt_ppm_edges(p(c), p(y[:3].copy()), p(y[3:].copy()), p(o))
            return float(lam @ o)

        def sig(y):
            return lib.pt_ppm_edges_sig(p(c), p(y[:3].copy()), p(y[3:].copy()))

        lv, ld = np.zeros(3), np.zeros(3)
        lib.pt_ppm_edges_adj(p(c), p(x[:3].copy()), p(x[3:].copy()), p(lam), p(lv), p(ld))
        sigs.add(sig(x))
        n += fd_check(f, sig, x, np.concatenate([lv, ld]))
    assert {1, 2} <= sigs and any(s & 4 for s in sigs) and any(s & 8 for s in sigs), sigs
    assert n > 2000


def test_ppm_slope_adjoint(lib):
    rng = np.random.default_rng(14)
    n = 0
    picks = set()
    for _ in range(400):
        b0, b1 = rng.uniform(0.2, 0.8, 2)
        x = rng.normal(size=3)
        lam = rng.normal()
        gr = np.zeros(3)
        lib.pt_ppm_slope_adj(b0, b1, p(x), lam, p(gr))
        sig = lambda y: lib.pt_ppm_slope_sig(b0, b1, p(y))  # noqa: E731
        picks.add(sig(x) & 7)
        n += fd_check(lambda y: lam * lib.pt_ppm_slope(b0, b1, p(y)), sig, x, gr)
    assert {0, 1, 3, 5} <= picks, picks  # flat, and each of the three limiter picks
    assert n > 1000


def test_tvd_adjoint(lib):
    rng = np.random.default_rng(15)
    n = 0
    for _ in range(300):
        cf, cb = rng.uniform(1.5, 2.5, 2)
        x = rng.normal(size=3)
        lam = rng.normal(size=2)
        gr = np.zeros(3)
        lib.pt_tvd_adj(cf, cb, p(x), p(lam), p(gr))

        def f(y):
            o = np.zeros(2)
            lib.pt_tvd(cf, cb, p(y), p(o))
            return float(lam @ o)

        def sig(y):
            dqf = 2 * (y[2] - y[1]) / cf
            dqb = 2 * (y[1] - y[0]) / cb
            return (np.sign(dqf), np.sign(dqb), abs(dqb) < abs(dqf))

        n += fd_check(f, sig, x, gr)
    assert n > 600
