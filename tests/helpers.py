"""Shared comparison helpers for the parity tests."""
from __future__ import annotations

import math

import numpy as np


def mismatch_report(a: np.ndarray, b: np.ndarray, what: str, max_show: int = 5) -> str:
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    n = int(bad.sum())
    if n == 0:
        return ""
    idx = np.argwhere(bad)[:max_show]
    rows = [f"{tuple(int(x) for x in i)}: got {a[tuple(i)]!r} want {b[tuple(i)]!r}" for i in idx]
    return f"{what}: {n} of {a.size} entries differ; first: " + "; ".join(rows)


def assert_bitwise(a: np.ndarray, b: np.ndarray, what: str):
    msg = mismatch_report(a, b, what)
    assert not msg, msg


def rel_l1(a: np.ndarray, b: np.ndarray) -> float:
    den = float(np.sum(np.abs(b)))
    return float(np.sum(np.abs(a - b))) / den if den > 0 else float(np.sum(np.abs(a - b)))


def interior(arr: np.ndarray, cfg) -> np.ndarray:
    """Cells [1,n]^3 of a ghost-inclusive array (index j+2)."""
    return arr[3:3 + cfg.nz, 3:3 + cfg.ny, 3:3 + cfg.nx]


def exact_riemann_euler(wl, wr, g, xi):
    """Exact solution of the Euler Riemann problem (Toro 2009, ch. 4) sampled at xi = x/t."""
    rl, ul, pl = wl
    rr, ur, pr = wr
    al, ar = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)

    def fk(p, rk, pk, ak):
        if p > pk:
            A, B = 2 / ((g + 1) * rk), (g - 1) / (g + 1) * pk
            return (p - pk) * math.sqrt(A / (p + B))
        return 2 * ak / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)

    lo, hi = 1e-12, 100.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if fk(mid, rl, pl, al) + fk(mid, rr, pr, ar) + (ur - ul) > 0:
            hi = mid
        else:
            lo = mid
    ps = 0.5 * (lo + hi)
    us = 0.5 * (ul + ur) + 0.5 * (fk(ps, rr, pr, ar) - fk(ps, rl, pl, al))
    out = np.zeros((len(xi), 3))
    for i, s in enumerate(xi):
        if s <= us:
            if ps > pl:
                sh = ul - al * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                w = (rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1)), us, ps) if s >= sh else (rl, ul, pl)
            else:
                shl, stl = ul - al, us - al * (ps / pl) ** ((g - 1) / (2 * g))
                if s <= shl:
                    w = (rl, ul, pl)
                elif s >= stl:
                    w = (rl * (ps / pl) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) + (g - 1) / ((g + 1) * al) * (ul - s)
                    w = (rl * c ** (2 / (g - 1)), 2 / (g + 1) * (al + (g - 1) / 2 * ul + s), pl * c ** (2 * g / (g - 1)))
        else:
            if ps > pr:
                sh = ur + ar * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                w = (rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1)), us, ps) if s <= sh else (rr, ur, pr)
            else:
                shr, str_ = ur + ar, us + ar * (ps / pr) ** ((g - 1) / (2 * g))
                if s >= shr:
                    w = (rr, ur, pr)
                elif s <= str_:
                    w = (rr * (ps / pr) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) - (g - 1) / ((g + 1) * ar) * (ur - s)
                    w = (rr * c ** (2 / (g - 1)), 2 / (g + 1) * (-ar + (g - 1) / 2 * ur + s), pr * c ** (2 * g / (g - 1)))
        out[i] = w
    return out
