"""bench.py's contract, checked without a GPU: one GPU only (a clear refusal for more), the
roofline object computed from a kernel timing with the algorithmic bytes per launch of the stage
tail (512/3 B per interior zone: one stage of the FP64 state traffic), the counter traffic taken
only from this run's own rocprofv3 passes (null without them), the parsing of a rocprofv3 counter
CSV, and the CPU baseline's fields on a tiny box."""
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
    # (26 + 20, 26 + 20, 34 + 20) FP64 words over the three stages of a plain step
    assert b.KERNEL_IO_BYTES["fp64"]["tail"] == pytest.approx((46 + 46 + 54) / 3 * 8, abs=1e-9)
    assert b.kernel_units("tail", cfg) == 257 ** 3
    # three tail launches of 1.75 ms in total 5.25 ms, the fluxes cheaper: the tail dominates
    kt = {"tail": (5.25, 3), "flux_x": (2.6, 3), "ecorn": (1.5, 3)}
    r = b.roofline(kt, cfg, "fp64")
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert key in r
    assert r["kernel"] == "tail" and r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    expect = 512.0 / 3.0 * 256 ** 3 / 1.75e-3 / 1e9
    assert r["achieved"] == pytest.approx(expect, rel=1e-9)
    assert r["frac"] == pytest.approx(expect / 8000.0, rel=1e-9)
    assert r["traffic"] is None and r["traffic_over_algorithmic"] is None
    assert r["kernel_io"]["achieved_GBps"] == pytest.approx(389.3333 * 257 ** 3 / 1.75e-3 / 1e9, rel=1e-4)
    t = {"tail_bytes_per_launch": 8.2e9}
    r = b.roofline(kt, cfg, "fp64", traffic=t)
    assert r["traffic"] == 8.2e9
    assert r["traffic_over_algorithmic"] == pytest.approx(8.2e9 / (512.0 / 3.0 * 256 ** 3))


def test_counter_csv_parsing(tmp_path):
    b = _bench()
    d = tmp_path / "host" / "1234"
    d.mkdir(parents=True)
    rows = [("3", "cumc::k_tail<double, 1, 1>(cumc::Lay)", "100.5"), ("1", "k_copy_probe(double2 const*)", "524288"),
            ("2", "cumc::k_flux_x<double, 4, 2>(cumc::Lay)", "7.0")]
    with open(d / "pmc_counter_collection.csv", "w") as f:
        f.write('"Correlation_Id","Dispatch_Id","Kernel_Name","Counter_Name","Counter_Value"\n')
        for i, n, v in rows:
            f.write(f'"{i}","{i}","{n}","FETCH_SIZE","{v}"\n')
    got = b._counter_rows(str(tmp_path))
    assert [r[0] for r in got] == [1, 2, 3]
    assert got[0][1].startswith("k_copy_probe") and got[2][2] == 100.5
    assert b.STEP_KERNELS.search(got[2][1]) and not b.STEP_KERNELS.search(got[0][1])
    names = {"cumc::k_flux_x<double, 4, 2>(cumc::Lay)": "flux_x", "cumc::k_flux_walk<double, 2, 4, 2>(x)": "flux_y",
             "cumc::k_flux_walk<float, 3, 1, 2>(x)": "flux_z", "cumc::k_ecorn<double>(x)": "ecorn",
             "cumc::k_tail<float, 3, 1>(x)": "tail", "cumc::k_ghost32<2>(x)": "ghost", "k_copy_probe(x)": None}
    for n, c in names.items():
        assert b.kernel_class(n) == c, n


def test_cpu_baseline_fields():
    b = _bench()
    from cumpic_amd import ics
    p = ics.small_periodic(8, seed=1005)
    out, state = b.cpu_baseline(p, steps=1, one_core_n=8)
    for key in ("value", "unit", "cores", "kind", "sample", "value_1core", "sample_1core"):
        assert key in out
    assert out["kind"] == "port" and out["unit"] == "zone-cycles/s" and out["value"] > 0 and out["value_1core"] > 0
    assert state["prim"].shape == p.prim.shape
