"""Frozen digests of seven canonical runs (tests/golden/oracle_runs.json, written by
tests/golden/make_golden.py): the CPU restatement must reproduce them bit for bit (so a change of
its arithmetic is deliberate), and so must the HIP FP64 path (gpu-marked), independently of a
side-by-side oracle run."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

from tests.golden.make_golden import OUT, cases, digest

GOLDEN = json.load(open(OUT))


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_oracle_reproduces_golden_digest(name):
    prob, steps = cases()[name]
    d = digest(prob, steps)
    g = GOLDEN[name]
    assert d["steps"] == g["steps"] and d["time"] == g["time"]
    assert d["sha256"] == g["sha256"], (name, d, g)


@pytest.mark.parametrize("threads", [1, 3])
def test_oracle_digest_independent_of_thread_count(threads):
    """The restatement's OpenMP loops write disjoint cells and reduce dt by min, so its result
    does not depend on the thread count (the CPU baseline times it at several)."""
    from oracle import oracle as orc
    name = "periodic_12_ppm_4steps"
    prob, steps = cases()[name]
    before = orc.set_threads(0)
    try:
        orc.set_threads(threads)
        d = digest(prob, steps)
    finally:
        orc.set_threads(before)
    assert d["sha256"] == GOLDEN[name]["sha256"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_gpu_reproduces_golden_digest(name):
    from cumpic_amd.cumc import Solver
    print(f"[case] golden digest {name} on the GPU", flush=True)
    prob, steps = cases()[name]
    s = Solver.from_problem(prob)
    s.initial_update()
    if steps is None:
        dts, t = s.run(100000, prob.total_time)
    else:
        dts, t = s.run(steps)
    st = s.download()
    h = hashlib.sha256()
    for k in ("prim", "eps", "bcell"):
        h.update(np.ascontiguousarray(st[k], dtype="<f8").tobytes())
    h.update(np.ascontiguousarray(dts, dtype="<f8").tobytes())
    g = GOLDEN[name]
    assert len(dts) == g["steps"] and float(t) == g["time"]
    assert h.hexdigest() == g["sha256"], name
