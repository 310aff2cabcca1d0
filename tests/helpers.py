"""Shared comparison helpers for the parity tests."""
from __future__ import annotations

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
