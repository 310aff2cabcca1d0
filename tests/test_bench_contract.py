"""bench.py's contract, checked without a GPU: one GPU only (a clear refusal for more), the
roofline object computed from a kernel timing with the algorithmic bytes per launch that the
design states (411 B per cell of [0,n]^3 for the FP64 stage tail), the calibrated traffic read
from the committed rocprofv3 summary, and the CPU baseline's fields on a tiny box."""
from __future__ import annotations

import importlib.util
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0 and "one GPU" in r.stdout


def test_roofline_object_from_a_kernel_timing():
    b = _bench()
    from cumpic_amd import ics
    cfg = ics.small_periodic(8).cfg
    cfg.nx = cfg.ny = cfg.nz = 256
    assert b.KERNEL_BYTES["fp64"]["tail"] == pytest.approx(410.67, abs=0.01)
    units = b.kernel_units("tail", cfg)
    assert units == 257 ** 3
    # three tail launches of 1.75 ms in total 5.25 ms, the fluxes cheaper: the tail dominates
    kt = {"tail": (5.25, 3), "flux_x": (2.6, 3), "ecorn": (1.5, 3)}
    r = b.roofline(kt, cfg, "fp64")
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert key in r
    assert r["kernel"] == "tail" and r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    expect = 410.6667 * 257 ** 3 / 1.75e-3 / 1e9
    assert r["achieved"] == pytest.approx(expect, rel=1e-4)
    assert r["frac"] == pytest.approx(expect / 8000.0, rel=1e-4)
    # calibrated HBM bytes per launch of the FP64 tail from profiles/pmc_traffic.txt
    assert r["traffic"] == b.read_traffic("tail", "fp64") and 7e9 < r["traffic"] < 9e9


def test_cpu_baseline_fields():
    b = _bench()
    from cumpic_amd import ics
    p = ics.small_periodic(8, seed=1005)
    out, state = b.cpu_baseline(p, steps=1, one_core_n=8)
    for key in ("value", "unit", "cores", "kind", "sample", "value_1core", "sample_1core"):
        assert key in out
    assert out["kind"] == "port" and out["unit"] == "zone-cycles/s" and out["value"] > 0 and out["value_1core"] > 0
    assert state["prim"].shape == p.prim.shape
