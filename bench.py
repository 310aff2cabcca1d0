#!/usr/bin/env python3
"""Benchmark of the MI355X CUMC3D MHD step on one GPU.

Default workload: the 3D periodic spectral MHD box of 256^3 zones, gamma = 5/3, beta = 1,
mean B_z = 1, PPM + HLL + upwind CT + SSP-RK3, seeded synthetic initial condition
(cumpic_amd.ics.spectral_box, seed 1001).  One step = finddt + one RUNGEKUTTA (three stages).

`value` is the FP64 rate.  The same run also times the mixed form (FP32 reconstruction, HLL and
EMFs; FP64 state) on the same workload and reports its rate, its speed-up and its error against
FP64 on the same dt sequence.  Prints ONE JSON line; progress goes to stderr.

It also runs the batched form: 48 independent 2D periodic boxes of sizes drawn by seed from
{64,96,128,192,256}^2 (cumpic_amd.ics.batch_boxes, seed 1004), 40 FP64 steps, advanced one box
at a time, by persistent work-list launches, and by a replayed hipGraph of 8 steps after the
launch-shape tuner, with a bit-for-bit check of the batched result against the one-at-a-time one.

The default run adds the mixed-precision workload at its own size (`c2`: the same seeded
spectrum on the 512^3 box, 100 FP64 steps, then the mixed form on the FP64 run's dt sequence,
relative L1 error of rho, p, |v|, |B| and max |div B| of both forms) and measures the stage
tail's HBM traffic per launch in the same run: two child processes run two steps of the same
state under `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE`, calibrated by a 1 GiB device
copy in the same child.  It also runs the CPU-runnable configuration (`c0`: the seeded 1D MHD
Riemann problem, 800 cells, to t = 0.1 on the GPU and on the CPU restatement, compared bit for
bit) and the curvilinear one (`c3`: the cylindrical torus 256x128x256, 50 steps in both forms,
the mixed form's error, the CPU restatement's rate and a full-size one-step parity check).

  --workload torus      cylindrical torus 256x128x256, 50 steps (torus_cyn hooks)
  --n 512 --steps 100   the 512^3 box of the mixed-precision workload
  --no-batch            skip the batched workload
  --no-c2               skip the 512^3 mixed-precision leg
  --no-c0, --no-c3      skip the 1D Riemann / cylindrical torus legs
  --no-traffic          skip the counter children
  --workload batch      only the batched workload (value = its hipGraph rate)
"""
from __future__ import annotations

import argparse
import csv
import dataclasses
import glob
import json
import os
import re
import shutil
import signal
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s
BYTES_PER_ZONE_CYCLE = 512.0  # 64 FP64 state words per RK3 step
# Algorithmic bytes of one stage-tail launch: the RK3 state traffic of one stage, 512/3 B per
# interior zone (stage 1 reads U0 and writes U1, stages 2 and 3 read U0 and the previous stage and
# write the next: 16 + 24 + 24 FP64 words per zone and step).  The mixed form keeps the FP64 state.
TAIL_ALG_BYTES_PER_ZONE = BYTES_PER_ZONE_CYCLE / 3.0
# bytes each kernel class must load and store per unit as written (its inputs and outputs,
# including the flux, EMF and primitive planes that pass between kernels): reported beside the
# algorithmic figure, not as it
KERNEL_IO_BYTES = {
    "fp64": {
        "flux_x": 15 * 8, "flux_y": 15 * 8, "flux_z": 15 * 8,  # 7 stencil vars + normal B in, 7 fluxes out
        "ecorn": 15 * 8,                                        # 6 B-fluxes + v + cell B in, 3 EMFs out
        # 15 fluxes + 3 EMFs + U0 (stage 1), the combined input (stage 2) or U0 + Uin (stage 3) in;
        # prim 8, eps, bcell 3, Uout 8 out: (46 + 46 + 54) / 3 words averaged over the three stage
        # launches of a plain step
        "tail": (46 + 46 + 54) / 3 * 8,
    },
    # mixed form: fluxes, EMFs and the primitive planes the sweeps read are FP32; the conserved
    # state is FP64.  Tail: 18 FP32 flux/EMF words + U0 (stage 1), the combined input (stage 2) or
    # U0 + Uin (stage 3) in; Uout 8 FP64 + P32 8, EPS32 1, BC32 3 FP32 out
    "mixed": {
        "flux_x": 8 * 4 + 7 * 4, "flux_y": 8 * 4 + 7 * 4, "flux_z": 8 * 4 + 7 * 4,
        "ecorn": 6 * 4 + 6 * 4 + 3 * 4,
        "tail": (2 * (18 * 4 + 8 * 8 + 8 * 8 + 12 * 4) + (18 * 4 + 16 * 8 + 8 * 8 + 12 * 4)) / 3,
    },
}
# kernels of the time step (per form) whose counter bytes make up the step traffic
STEP_KERNELS = re.compile(r"k_(flux_x|flux_walk|ecorn|tail|stage_z|ghost32|ghost|fill_u64|custom_boundary)")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_units(name, cfg):
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    return {
        "flux_x": (nx + 1) * (ny + 2) * (nz + 2),
        "flux_y": (nx + 2) * (ny + 1) * (nz + 2),
        "flux_z": (nx + 2) * (ny + 2) * (nz + 1),
        "ecorn": (nx + 1) * (ny + 1) * (nz + 1),
        "tail": (nx + 1) * (ny + 1) * (nz + 1),
    }.get(name)


def _counter_rows(outdir):
    """(dispatch id, kernel name, counter value) of every dispatch in a rocprofv3 CSV counter
    collection, in dispatch order."""
    rows = []
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                name = r.get("Kernel_Name") or ""
                val = r.get("Counter_Value")
                if name and val not in (None, ""):
                    rows.append((int(r.get("Dispatch_Id") or 0), name, float(val)))
    rows.sort()
    return rows


def pmc_traffic(ic_dir, form, steps=2, timeout=300):
    """Counter bytes measured in this run for one form: a child process runs `steps` RK3 steps of
    the saved state under rocprofv3 --pmc FETCH_SIZE, another under --pmc WRITE_SIZE; each child
    first copies 1 GiB with cumc_copy_probe, whose counter readings give the KiB-to-byte factors.
    Returns the stage tail's bytes per launch and the whole step's bytes per zone-cycle (every
    dispatch from the first sweep on), or None."""
    rp = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(rp):
        log("[bench] traffic: rocprofv3 not found, traffic left null")
        return None
    res = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="cumc_pmc_")
        cmd = [rp, "--pmc", ctr, "--output-format", "csv", "-d", d, "-o", "pmc", "--", sys.executable,
               os.path.abspath(__file__), "--traffic-child", ic_dir, "--forms", form, "--steps", str(steps)]
        log(f"[bench] traffic: {form} {ctr} pass under rocprofv3")
        p = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, start_new_session=True)
        try:
            _, err = p.communicate(timeout=timeout)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.communicate()
            log(f"[bench] traffic: {ctr} pass timed out, traffic left null")
            shutil.rmtree(d, ignore_errors=True)
            return None
        rows = _counter_rows(d)
        shutil.rmtree(d, ignore_errors=True)
        cp = [v for _, n, v in rows if "k_copy_probe" in n]
        first = next((i for i, (_, n, _) in enumerate(rows) if "k_flux_x" in n), None)
        if not cp or first is None:
            log(f"[bench] traffic: {ctr} pass left no usable counters (exit {p.returncode}): "
                f"{err.decode(errors='replace')[-400:]}")
            return None
        tail = [v for _, n, v in rows if "k_tail<" in n]
        step = [v for _, n, v in rows[first:] if STEP_KERNELS.search(n)]
        classes = {}
        for _, n, v in rows[first:]:
            c = kernel_class(n)
            if c:
                classes.setdefault(c, []).append(v)
        res[ctr] = (float(1 << 30) / 1024.0 / cp[0], sum(tail) / len(tail), len(tail), sum(step), classes)
    (ff, tf, nt, sf, cf), (fw, tw, _, sw, cw) = res["FETCH_SIZE"], res["WRITE_SIZE"]
    by_class = {c: (ff * sum(cf[c]) / len(cf[c]) + fw * sum(cw.get(c, [0.0])) / max(1, len(cw.get(c, [])))) * 1024.0
                for c in cf}
    return {"tail_bytes_per_launch": (ff * tf + fw * tw) * 1024.0, "tail_launches": nt,
            "step_bytes": (ff * sf + fw * sw) * 1024.0 / steps, "bytes_per_launch_by_class": by_class,
            "fetch_factor": ff, "write_factor": fw,
            "counters": "rocprofv3 --pmc FETCH_SIZE, then --pmc WRITE_SIZE, KiB per dispatch x factor x 1024"}


def kernel_class(name):
    """Kernel class of a dispatch name (the classes of the library's HIP-event timing)."""
    if "k_flux_x<" in name:
        return "flux_x"
    if "k_flux_walk<" in name:
        m = re.search(r"k_flux_walk<(?:double|float), (\d)", name)
        return {"2": "flux_y", "3": "flux_z"}.get(m.group(1)) if m else None
    for key, c in (("k_ecorn<", "ecorn"), ("k_tail<", "tail"), ("k_ghost", "ghost")):
        if key in name:
            return c
    return None


def traffic_child(args):
    """Body of a counter pass: calibration copy, then `steps` RK3 steps of the saved state."""
    from cumpic_amd.config import FP64, MIXED, Config
    from cumpic_amd.cumc import Solver, copy_probe
    from cumpic_amd.ics import Problem
    meta = json.load(open(os.path.join(args.traffic_child, "meta.json")))
    cfg = Config(**{k: tuple(v) if isinstance(v, list) else v for k, v in meta["cfg"].items()})
    ld = lambda n: np.load(os.path.join(args.traffic_child, n + ".npy"))  # noqa: E731
    prob = Problem(cfg, ld("xf"), ld("yf"), ld("zf"), ld("prim"), ld("eps"))
    copy_probe(1 << 30, reps=1)
    s = Solver.from_problem(prob, precision=FP64 if args.forms == "fp64" else MIXED)
    s.initial_update()
    s.run(args.steps)
    s.close()


def save_ic(prob):
    d = tempfile.mkdtemp(prefix="cumc_ic_")
    for n in ("xf", "yf", "zf", "prim", "eps"):
        np.save(os.path.join(d, n + ".npy"), getattr(prob, n))
    json.dump({"cfg": dataclasses.asdict(prob.cfg)}, open(os.path.join(d, "meta.json"), "w"))
    return d


def cpu_baseline(prob, steps=1, one_core_n=128, total_time=1.0e300, threads=None):
    """The CPU restatement on the host cores: same initial state, at most `steps` RK3 steps (each
    with its finddt) up to total_time; beside it a one-thread figure on a bounded sample (the same
    workload builder at one_core_n^3, one step).  The returned state carries the steps' dt."""
    from cumpic_amd import ics
    from oracle import oracle as orc
    from oracle.oracle import Oracle
    all_cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cores = threads or all_cores
    orc.set_threads(cores)
    o = Oracle.from_problem(prob)
    o.initial_update()
    t0 = time.perf_counter()
    dts, _ = o.run(steps, total_time)
    wall = time.perf_counter() - t0
    state = o.get_state()
    state["dts"] = dts
    o.close()
    orc.set_threads(all_cores)
    c = prob.cfg
    out = {
        "value": c.zones * len(dts) / wall,
        "unit": "zone-cycles/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{c.nx}x{c.ny}x{c.nz} {prob.name}, {len(dts)} RK3 step(s) after initial_update, "
                  f"{wall:.2f} s wall on {cores} OpenMP thread(s)",
    }
    if one_core_n:
        small = ics.spectral_box(n=one_core_n, seed=1001, steps=1)
        orc.set_threads(1)
        try:
            o1 = Oracle.from_problem(small)
            o1.initial_update()
            t0 = time.perf_counter()
            o1.run(1)
            w1 = time.perf_counter() - t0
            o1.close()
        finally:
            orc.set_threads(cores)
        out["value_1core"] = small.cfg.zones / w1
        out["sample_1core"] = f"{one_core_n}^3 periodic spectral box (seed 1001), 1 RK3 step, {w1:.1f} s on 1 thread"
    return out, state


def timed_run(Solver, prob, prec, steps, warmup, tune=True):
    """Upload, initial_update, warm-up steps, then (tune) the library's launch-shape tuner, which
    leaves the state as it found it, then the timed steps."""
    s = Solver.from_problem(prob, precision=prec)
    s.initial_update()
    if warmup:
        s.run(warmup)
    tuning = s.autotune() if tune else s.tuning()
    s.profile(True)
    t0 = time.perf_counter()
    dts, _ = s.run(steps)  # synchronous on return (hipStreamSynchronize)
    wall = time.perf_counter() - t0
    kt = s.kernel_times()
    s.profile(False)
    return s, dts, wall, kt, tuning


def roofline(kt, cfg, tag, traffic=None):
    """Roofline of the stage tail, the kernel that takes the most time: algorithmic bytes per
    launch = 512/3 B per interior zone (one stage's share of the RK3 state traffic) over its
    average HIP-event launch time, against 8 TB/s.  `traffic` is the HBM bytes per launch that
    the counters of this run measured (None: not measured); kernel_io gives the same launch
    measured against the bytes the kernel loads and stores as written."""
    tail_ms, launches = kt["tail"]
    avg_s = tail_ms / launches * 1e-3
    alg = TAIL_ALG_BYTES_PER_ZONE * cfg.zones
    io = KERNEL_IO_BYTES[tag]["tail"] * kernel_units("tail", cfg)
    dom = max(kt, key=lambda k: kt[k][0])
    t = traffic["tail_bytes_per_launch"] if traffic else None
    return {"bound": "hbm", "achieved": alg / avg_s / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": alg / avg_s / HBM_PEAK, "traffic": t, "kernel": "tail", "dominant_by_time": dom,
            "algorithmic_bytes_per_launch": alg,
            "algorithmic_basis": "512/3 B per interior zone: one RK3 stage of the FP64 state traffic",
            "traffic_over_algorithmic": (t / alg) if t else None, "avg_launch_ms": tail_ms / launches,
            "kernel_io": {"bytes_per_launch": io, "achieved_GBps": io / avg_s / 1e9, "frac": io / avg_s / HBM_PEAK,
                          "basis": "the tail's own loads and stores (flux, EMF, state and primitive planes)"}}


def fields(state, cfg):
    sl = (slice(3, 3 + cfg.nz), slice(3, 3 + cfg.ny), slice(3, 3 + cfg.nx))
    p, b = state["prim"][sl], state["bcell"][sl]
    return {"rho": p[..., 0], "p": p[..., 4], "|v|": np.linalg.norm(p[..., 1:4], axis=-1),
            "|B|": np.linalg.norm(b, axis=-1)}


def bench_c0(t_end=0.1):
    """The CPU-runnable configuration: the seeded 1D MHD Riemann problem (800 cells, gamma 2,
    CFL 0.8, outflow, seed 1000) run to t = 0.1 with each step's finddt, in FP64 on the GPU (a
    separate warm-up context first) and on the CPU restatement on one thread; final state and dt
    sequence compared bit for bit."""
    from cumpic_amd import ics
    from cumpic_amd.config import FP64
    from cumpic_amd.cumc import Solver
    prob = ics.riemann_1d(t_end=t_end)
    w = Solver.from_problem(prob, precision=FP64)
    w.initial_update()
    w.run(5)
    w.close()
    s = Solver.from_problem(prob, precision=FP64)
    s.initial_update()
    t0 = time.perf_counter()
    dts, t = s.run(10 ** 6, total_time=t_end)
    wall = time.perf_counter() - t0
    gs = s.download()
    s.close()
    cpu, ostate = cpu_baseline(prob, steps=10 ** 6, one_core_n=0, total_time=t_end, threads=1)
    same = bool(np.array_equal(np.asarray(dts), ostate["dts"]) and
                all(np.array_equal(gs[k], ostate[k]) for k in ("prim", "eps", "bcell")))
    zc = prob.cfg.zones * len(dts)
    out = {"workload": f"1D MHD Riemann problem, {prob.cfg.nx} cells, seeded states (seed 1000), gamma 2, "
                       f"CFL 0.8, outflow, to t = {t_end}",
           "steps": len(dts), "t_end": t, "gpu": {"value": zc / wall, "ms": wall * 1e3},
           "cpu": cpu, "bitwise_vs_cpu": same}
    log(f"[bench] c0: {len(dts)} steps to t={t:.6f}; GPU {wall * 1e3:.1f} ms ({zc / wall:.3e} zone-cycles/s, "
        f"launch-bound at 800 cells), CPU {cpu['value']:.3e} on 1 thread; bitwise: {same}")
    return out


def bench_c3(steps=50, warmup=2):
    """The curvilinear configuration: the cylindrical torus (256x128x256, log R in [1,40], seed
    1003, torus_cyn hooks), `steps` steps in FP64 and in the mixed form after `warmup` steps and
    the tuner, the mixed form's error against FP64 on the FP64 run's dt sequence, the CPU
    restatement's rate on one step, and that step's full-size parity with the GPU."""
    from cumpic_amd import ics
    from cumpic_amd.config import FP64, MIXED
    from cumpic_amd.cumc import Solver
    prob = ics.torus_cyn(256, 128, 256, steps=steps)
    cfg = prob.cfg
    res = {}
    ref_dts = ref = None
    for form, prec in (("fp64", FP64), ("mixed", MIXED)):
        s, dts, wall, kt, tuning = timed_run(Solver, prob, prec, steps, warmup)
        _, dabs = s.find_divb()
        st = s.download()
        s.close()
        res[form] = {"value": cfg.zones * steps / wall, "ms_per_step": wall / steps * 1e3, "divb_max_abs": dabs,
                     "state_finite": bool(np.isfinite(st["prim"]).all()), "tuning": tuning}
        if form == "fp64":
            ref_dts, ref = dts, fields(st, cfg)
        del st
    f64 = Solver.from_problem(prob, precision=FP64)
    f64.initial_update()
    warm_dts = f64.run(warmup)[0] if warmup else np.zeros(0)
    f64.close()
    m = Solver.from_problem(prob, precision=MIXED)
    m.initial_update()
    for dt in list(warm_dts) + list(ref_dts):
        m.rungekutta(float(dt))
    mf = fields(m.download(), cfg)
    m.close()
    res["mixed"]["rel_l1_vs_fp64"] = {k: float(np.sum(np.abs(mf[k] - ref[k])) / np.sum(np.abs(ref[k]))) for k in mf}
    res["mixed"]["speedup_vs_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
    cpu, ostate = cpu_baseline(prob, steps=1, one_core_n=0)
    g = Solver.from_problem(prob, precision=FP64)
    g.initial_update()
    g.run(1)
    gs = g.download()
    g.close()
    same = all(np.array_equal(gs[k], ostate[k], equal_nan=True) for k in ("prim", "eps", "bcell"))
    log(f"[bench] c3: FP64 {res['fp64']['value']:.4e}, mixed {res['mixed']['value']:.4e} zone-cycles/s "
        f"({res['mixed']['speedup_vs_fp64']:.2f}x), mixed rel L1 {res['mixed']['rel_l1_vs_fp64']}; "
        f"CPU {cpu['value']:.3e}; one-step bitwise {same}")
    return {"workload": "cylindrical torus 256x128x256 (log R in [1,40], phi periodic, z in [-20,20]), beta=100, "
                        "PW gravity + floors + inflow clamp, PPM+HLL+upwind CT+SSP-RK3 (seed 1003)",
            "steps": steps, "zones": cfg.zones, **res, "cpu": cpu, "one_step_bitwise_vs_cpu": bool(same)}


def bench_c2(n=512, steps=100, warmup=2):
    """The mixed-precision workload at its stated size: the 256^3 box's seeded spectrum on the
    512^3 box, `steps` FP64 steps (after `warmup` and the tuner), then the mixed form from the same
    initial state on the FP64 run's dt sequence (warm-up steps included).  Reports both rates, the
    speed-up, the relative L1 error of rho, p, |v|, |B| against FP64 and max |div B| of both."""
    from cumpic_amd import ics
    from cumpic_amd.config import FP64, MIXED
    from cumpic_amd.cumc import Solver
    log(f"[bench] c2: building the {n}^3 spectral initial condition")
    t0 = time.perf_counter()
    prob = ics.spectral_box(n=n, seed=1001, steps=steps)
    log(f"[bench] c2: initial condition in {time.perf_counter() - t0:.1f} s")
    cfg = prob.cfg
    s = Solver.from_problem(prob, precision=FP64)
    s.initial_update()
    warm = s.run(warmup)[0] if warmup else np.zeros(0)
    tune64 = s.autotune()
    t0 = time.perf_counter()
    dts, _ = s.run(steps)
    w64 = time.perf_counter() - t0
    _, abs64 = s.find_divb()
    ref = fields(s.download(), cfg)
    s.close()
    log(f"[bench] c2: FP64 {cfg.zones * steps / w64:.4e} zone-cycles/s ({w64 / steps * 1e3:.2f} ms/step)")
    m = Solver.from_problem(prob, precision=MIXED)
    m.initial_update()
    for dt in warm:
        m.rungekutta(float(dt))
    tunem = m.autotune()
    t0 = time.perf_counter()
    for dt in dts:
        m.rungekutta(float(dt))
    wm = time.perf_counter() - t0
    _, absm = m.find_divb()
    mf = fields(m.download(), cfg)
    m.close()
    errs = {k: float(np.sum(np.abs(mf[k] - ref[k])) / np.sum(np.abs(ref[k]))) for k in mf}
    del mf, ref
    log(f"[bench] c2: mixed {cfg.zones * steps / wm:.4e} zone-cycles/s ({wm / steps * 1e3:.2f} ms/step), "
        f"{w64 / wm:.2f}x; rel L1 vs FP64 {errs}; max |div B| FP64 {abs64:.2e}, mixed {absm:.2e}")
    return {"workload": f"3D periodic spectral MHD box {n}^3 (seed 1001, the 256^3 box's spectrum), "
                        f"{steps} steps after {warmup} warm-up, mixed form on the FP64 dt sequence",
            "zones": cfg.zones, "steps": steps,
            "fp64": {"value": cfg.zones * steps / w64, "ms_per_step": w64 / steps * 1e3, "tuning": tune64},
            "mixed": {"value": cfg.zones * steps / wm, "ms_per_step": wm / steps * 1e3, "tuning": tunem},
            "speedup_mixed_vs_fp64": w64 / wm, "rel_l1_vs_fp64": errs,
            "divb_max_abs": {"fp64": abs64, "mixed": absm},
            "targets_met": bool(w64 / wm >= 1.6 and max(errs.values()) <= 1e-5 and max(abs64, absm) <= 1e-12)}


def bench_batch(steps=40, graph_steps=8, fd_dirs=32):
    """Batched 2D boxes: one at a time vs work-list launches vs hipGraph replay."""
    from cumpic_amd import ics
    from cumpic_amd.cumc import Batch, Solver
    log("[bench] batch: building 48 2D boxes")
    probs = ics.batch_boxes(steps=steps)
    zones = sum(p.cfg.zones for p in probs)

    def fresh():
        out = []
        for p in probs:
            s = Solver.from_problem(p)
            s.initial_update()
            out.append(s)
        return out

    seq = fresh()
    seq[0].run(1)  # module load / first-launch cost out of the timed region
    seq[0].close()
    seq[0] = Solver.from_problem(probs[0])
    seq[0].initial_update()
    t0 = time.perf_counter()
    seq_dts = [s.run(steps)[0] for s in seq]
    t_seq = time.perf_counter() - t0
    log(f"[bench] batch: one box at a time {t_seq * 1e3:.1f} ms")

    boxes = fresh()
    b = Batch(boxes)
    t0 = time.perf_counter()
    tune = b.autotune()
    t_tune = time.perf_counter() - t0
    log(f"[bench] batch: tuner {t_tune:.2f} s -> {tune}")
    b.run(graph_steps, graph_steps=graph_steps)  # graph capture + warm-up
    for s, p in zip(boxes, probs):
        s.upload(p.prim, p.eps)
        s.initial_update()
    t0 = time.perf_counter()
    dts = b.run(steps, graph_steps=graph_steps)
    t_graph = time.perf_counter() - t0
    same = bool(np.array_equal(dts, np.array(seq_dts).T))
    for s, q in zip(boxes, seq):
        a, c = s.download(), q.download()
        same = same and all(np.array_equal(a[k], c[k]) for k in ("prim", "eps", "bcell"))
    log(f"[bench] batch: graph {t_graph * 1e3:.1f} ms, bitwise vs one-at-a-time: {same}")
    for s, p in zip(boxes, probs):
        s.upload(p.prim, p.eps)
        s.initial_update()
    t0 = time.perf_counter()
    b.run(steps, graph_steps=0)
    t_direct = time.perf_counter() - t0
    log(f"[bench] batch: direct launches {t_direct * 1e3:.1f} ms")
    adj = bench_adjoint(probs, boxes, b, np.array(seq_dts).T, steps, every=graph_steps, ndir=fd_dirs)
    b.close()
    for s in boxes + seq:
        s.close()
    zc = zones * steps
    return {
        "workload": f"48 2D periodic spectral boxes, sizes drawn by seed 1004 from {{64,96,128,192,256}}^2, "
                    f"{zones} zones, {steps} FP64 steps, per-box CFL dt on the device",
        "zone_cycles": zc,
        "one_at_a_time": {"value": zc / t_seq, "ms": t_seq * 1e3},
        "batched_direct": {"value": zc / t_direct, "ms": t_direct * 1e3},
        "batched_graph": {"value": zc / t_graph, "ms": t_graph * 1e3, "graph_steps": graph_steps},
        "speedup_graph_vs_one_at_a_time": t_seq / t_graph,
        "tuning": tune,
        "tuner_s": t_tune,
        "bitwise_vs_one_at_a_time": same,
        "unit": "zone-cycles/s",
        "adjoint": adj,
    }


def _interior(cfg):
    return (slice(3, 3 + cfg.nz), slice(3, 3 + cfg.ny), slice(3, 3 + cfg.nx))


def bench_adjoint(probs, boxes, batch, dts, steps, every=8, ndir=32, hs=(1e-4, 1e-5, 1e-6), store_all=True):
    """Reverse mode on the batched boxes: J_b = sum |B|^2/2 dV per box after `steps` frozen-dt
    steps; per-box gradients with respect to the initial velocity from the batched tape
    (cumc_batch_tape_run / cumc_batch_tape_backward, checkpoints every `every` steps).  Cost:
    device time of the backward pass over the device time of the same batched forward run.
    Check: central differences of sum_b J_b along ndir seeded random unit directions in the
    concatenated velocity space, each J(+-) from one batched frozen-dt run."""
    log(f"[bench] adjoint: {len(boxes)} boxes, {steps} steps, checkpoints every {every}")
    for s, p in zip(boxes, probs):
        s.upload(p.prim, p.eps)
        s.initial_update()
    batch.run(steps, graph_steps=every, dts=dts)
    fwd_ms = batch.last_ms()
    ref_bcell = [s.download()["bcell"][_interior(p.cfg)] for s, p in zip(boxes, probs)]
    for s, p in zip(boxes, probs):
        s.upload(p.prim, p.eps)
    t0 = time.perf_counter()
    batch.tape_run(steps, every, dts)
    t_tape = time.perf_counter() - t0
    seeds, dvs = [], []
    for s, p in zip(boxes, probs):
        inner = _interior(p.cfg)
        dv = (p.xf[4] - p.xf[3]) * (p.yf[4] - p.yf[3]) * (p.zf[4] - p.zf[3])
        st = s.download()
        seed = np.zeros(p.cfg.shape + (3,))
        seed[inner] = st["bcell"][inner] * dv
        seeds.append(seed)
        dvs.append(dv)
    t0 = time.perf_counter()
    gp = batch.tape_backward(seeds)
    t_back = time.perf_counter() - t0
    back_ms = batch.last_ms()
    recompute_ms = batch.tape_recompute_ms()
    batch.tape_free()
    grads = [g[_interior(p.cfg) + (slice(1, 4),)].copy() for g, p in zip(gp, probs)]
    log(f"[bench] adjoint: forward {fwd_ms:.1f} ms (batched graph), backward {back_ms:.1f} ms device "
        f"({t_back * 1e3:.1f} ms wall incl. seed/gradient transfers), ratio {back_ms / fwd_ms:.2f}; "
        f"the checkpoint segments' forward re-runs took {recompute_ms:.1f} ms on their own stream, "
        f"beside the reverse of the segment after them")
    # one checkpoint segment over all steps: every stage state is kept from the taping forward
    # (about 2570 state planes per box, ~194 GB for the 48 boxes), so the backward re-runs nothing
    stored = None
    if store_all:
        from cumpic_amd.cumc import CumcError
        for s, p in zip(boxes, probs):
            s.upload(p.prim, p.eps)
        try:
            t0 = time.perf_counter()
            batch.tape_run(steps, steps, dts)
            t_tape_all = time.perf_counter() - t0
            gp_all = batch.tape_backward(seeds)
            back_all = batch.last_ms()
            rec_all = batch.tape_recompute_ms()
            same = all(np.array_equal(a, c) for a, c in zip(gp_all, gp))
            stored = {"checkpoint_every": steps, "backward_ms": back_all, "recompute_ms": rec_all,
                      "adjoint_over_forward": back_all / fwd_ms, "tape_forward_wall_ms": t_tape_all * 1e3,
                      "gradient_bitwise_equal_to_checkpointed": bool(same)}
            log(f"[bench] adjoint, every stage state kept (one segment): backward {back_all:.1f} ms device, "
                f"ratio {back_all / fwd_ms:.2f}, re-run {rec_all:.1f} ms, gradients bit-identical to the "
                f"checkpointed run: {same}")
        except CumcError as e:
            log(f"[bench] adjoint with every stage state kept: not run ({e})")
        finally:
            batch.tape_free()

    fd = fd_check(probs, grads, dvs, dts, steps, every, ndir, hs, ref_bcell, tuning=batch.tuning())
    return {"objective": "per box sum |B|^2/2 dV after the steps; gradient w.r.t. initial velocity",
            "steps": steps, "checkpoint_every": every, "forward_ms": fwd_ms, "backward_ms": back_ms,
            "backward_wall_ms": t_back * 1e3, "tape_forward_wall_ms": t_tape * 1e3,
            "adjoint_over_forward": back_ms / fwd_ms, "recompute_ms": recompute_ms,
            "rerun_stream": "re-runs overlap the reverse of the next segment; recompute_ms is their stream time",
            "checkpointing_cost_ms": (back_ms - stored["backward_ms"]) if stored else None,
            "all_states_kept": stored, **fd}


def fd_check(probs, grads, dvs, dts, steps, every, ndir, hs, ref_bcell, tuning=None):
    """Central differences of J = sum_b J_b along ndir seeded unit directions of the concatenated
    initial-velocity space, against the reverse-mode g.d, with every forward run's per-box branch
    signature.  The runs use the diagnostic build libcumc_sig.so (the same arithmetic; its base
    run is checked bit for bit against the product build's final state), which hashes every
    limiter / Riemann-region / floor decision per box.  A box-direction pair is flip-free at h
    when the +h and -h runs both reproduce the unperturbed run's signature: then the difference
    crosses no kink of the map inside that box, and what remains of the error is truncation and
    the rounding of the forward runs.  At the smallest h the rounding part is measured: the
    difference is taken again at h' = h (1 + 2^-6), whose runs round independently while the
    smooth part of the difference moves by O(h^2 / 64), so |FD(h) - FD(h')| / sqrt(2) is the size
    of the rounding error of one central difference.  Errors are relative to |g.d| (per pair,
    g_b.d_b) and, for the directions, also to sum |g_i d_i|."""
    from cumpic_amd.cumc import Batch, Solver
    if ndir == 0:
        return {"fd_rel_error_by_h": {}, "fd_error_over_term_sum_by_h": {}, "directions": 0}
    sboxes = [Solver(p.cfg, p.xf, p.yf, p.zf, variant="sig") for p in probs]
    sb = Batch(sboxes)
    if tuning:
        sb.set_tuning(tuning)

    def run(vs):
        for s, p, v in zip(sboxes, probs, vs):
            prim = p.prim.copy()
            prim[_interior(p.cfg) + (slice(1, 4),)] = v
            s.upload(prim, p.eps)
            s.initial_update()
        sb.branch_sig()
        sb.run(steps, graph_steps=every, dts=dts)
        sig = sb.branch_sig()
        return [s.download()["bcell"][_interior(p.cfg)] for s, p in zip(sboxes, probs)], sig

    v0 = [p.prim[_interior(p.cfg) + (slice(1, 4),)].copy() for p in probs]
    b0, sig0 = run(v0)
    same_build = all(np.array_equal(a, c) for a, c in zip(b0, ref_bcell))
    log(f"[bench] FD check: diagnostic build bit-identical to the product build on the base run: {same_build}")
    rng = np.random.default_rng(1004)
    nb = len(probs)
    rel = {h: [] for h in hs}
    rel_terms = {h: [] for h in hs}
    dir_flip_free = {h: [] for h in hs}
    pair_rel = {h: [] for h in hs}
    pair_free = {h: [] for h in hs}
    h_min = min(hs)
    spread_free, spread_err, spread_noise = [], [], []

    def half_diff(a, c, w):
        return 0.5 * float(np.sum((a - c) * (a + c))) * w

    for d in range(ndir):
        dirs = [rng.normal(size=g.shape) for g in grads]
        nrm = np.sqrt(sum(float(np.sum(x * x)) for x in dirs))
        dirs = [x / nrm for x in dirs]
        adb = np.array([float(np.sum(g * x)) for g, x in zip(grads, dirs)])
        ad = float(adb.sum())
        mag = sum(float(np.sum(np.abs(g * x))) for g, x in zip(grads, dirs))
        def central(h):
            bp, sp = run([v + h * x for v, x in zip(v0, dirs)])
            bm, sm = run([v - h * x for v, x in zip(v0, dirs)])
            djb = np.array([half_diff(a, c, w) for a, c, w in zip(bp, bm, dvs)])
            return djb / (2 * h), float(djb.sum()) / (2 * h), (sp == sig0) & (sm == sig0)

        for h in hs:
            fdb, fd, free = central(h)
            rel[h].append(abs(fd - ad) / max(abs(fd), abs(ad)))
            rel_terms[h].append(abs(fd - ad) / mag)
            dir_flip_free[h].append(bool(free.all()))
            pair_rel[h].extend((np.abs(fdb - adb) / np.abs(adb)).tolist())
            pair_free[h].extend(free.tolist())
            if h == h_min:
                fdb2, _, free2 = central(h * (1.0 + 2.0 ** -6))
                spread_free.extend((free & free2).tolist())
                spread_err.extend((np.abs(fdb - adb) / np.abs(adb)).tolist())
                spread_noise.extend((np.abs(fdb - fdb2) / np.sqrt(2.0) / np.abs(adb)).tolist())
    sb.close()
    for s in sboxes:
        s.close()

    def stats(v):
        v = np.asarray(v, dtype=float)
        if v.size == 0:
            return {"n": 0}
        return {"n": int(v.size), "median": float(np.median(v)), "max": float(np.max(v)),
                "within_1e-6": int(np.sum(v <= 1e-6))}

    def summarise(r):
        return {f"{h:.0e}": stats(v) for h, v in r.items() if v}

    summary, summary_terms = summarise(rel), summarise(rel_terms)
    by_h = {}
    for h in hs:
        e = np.array(pair_rel[h])
        f = np.array(pair_free[h], dtype=bool)
        de = np.array(rel[h])
        df = np.array(dir_flip_free[h], dtype=bool)
        by_h[f"{h:.0e}"] = {
            "directions_flip_free": int(df.sum()),
            "directions_flip_free_error": stats(de[df]),
            "box_pairs": int(e.size),
            "box_pairs_flip_free": int(f.sum()),
            "flip_free_error": stats(e[f]),
            "flipped_error": stats(e[~f]),
        }
    sf = np.array(spread_free, dtype=bool)
    se, sn = np.array(spread_err), np.array(spread_noise)
    rounding = {"h": h_min, "h2": h_min * (1.0 + 2.0 ** -6), "box_pairs_flip_free_at_both": int(sf.sum()),
                "error": stats(se[sf]), "rounding": stats(sn[sf]),
                "error_over_rounding_median": float(np.median(se[sf] / sn[sf])) if sf.any() else None}
    log(f"[bench] adjoint vs central differences ({ndir} unit directions), relative to |g.d|: {summary}")
    log(f"[bench] the same relative to sum |g_i d_i|: {summary_terms}")
    for h, r in by_h.items():
        log(f"[bench] FD h={h}: {r['box_pairs_flip_free']}/{r['box_pairs']} box-direction pairs flip-free; "
            f"flip-free error {r['flip_free_error']}; flipped error {r['flipped_error']}; "
            f"flip-free directions {r['directions_flip_free']}/{ndir}")
    log(f"[bench] FD rounding at h={h_min:.0e} on pairs flip-free at h and h': {rounding}")
    return {"fd_rel_error_by_h": summary, "fd_error_over_term_sum_by_h": summary_terms, "directions": ndir,
            "fd_branch_signature_by_h": by_h, "fd_rounding_at_smallest_h": rounding,
            "fd_diagnostic_build_bitwise": bool(same_build)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["box", "torus", "sphere", "batch"], default="box",
                    help="batch: only the batched 2D boxes (forward and reverse mode), value = batched rate")
    ap.add_argument("--n", type=int, default=256, help="zones per axis of the periodic box")
    ap.add_argument("--forms", default="fp64,mixed", help="comma list of fp64, mixed")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and full-size parity check")
    ap.add_argument("--no-error", action="store_true", help="skip the mixed-vs-FP64 error replay")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 2D-box workload")
    ap.add_argument("--fd-dirs", type=int, default=32, help="directions of the reverse-mode difference check")
    ap.add_argument("--copy-probe-mib", type=int, default=1024,
                    help="plain device copy of this many MiB: the achievable HBM rate, and the counter "
                         "calibration (0 skips it)")
    ap.add_argument("--no-tune", action="store_true", help="time the default launch shapes (no autotune)")
    ap.add_argument("--recon", choices=["default", "ppm", "tvdmm", "tvdmc", "tvdvl", "weno"], default="default",
                    help="reconstruction (ppm and tvdmm are the reference's; the others are extensions)")
    ap.add_argument("--solver", choices=["hll", "lf", "hllc", "hlld"], default="hll",
                    help="Riemann solver (hll is the reference's; lf, hllc and hlld are extensions)")
    ap.add_argument("--no-c2", action="store_true", help="skip the 512^3 mixed-precision leg")
    ap.add_argument("--no-c0", action="store_true", help="skip the 1D Riemann leg (GPU and CPU to t = 0.1)")
    ap.add_argument("--no-c3", action="store_true", help="skip the cylindrical torus leg")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter children")
    ap.add_argument("--traffic-child", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.traffic_child:
        traffic_child(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU halo exchange is a design note, not code)", file=sys.stderr)
        sys.exit(2)

    from cumpic_amd import ics
    from cumpic_amd.config import FP64, MIXED
    from cumpic_amd.cumc import Solver, copy_probe

    copy_gbps = None

    def run_copy_probe():
        if args.copy_probe_mib:
            g = copy_probe(args.copy_probe_mib << 20, reps=3)
            log(f"[bench] copy probe {args.copy_probe_mib} MiB: {g:.1f} GB/s")
            return g
        return None

    if args.workload == "batch":
        copy_gbps = run_copy_probe()
        steps = args.steps or 40
        bt = bench_batch(steps=steps, graph_steps=min(8, steps), fd_dirs=args.fd_dirs)
        g = bt["batched_graph"]
        print(json.dumps({
            "metric": "zone-cycles/s FP64 vs mixed, % HBM roofline, adjoint/forward step cost",
            "value": g["value"], "unit": "zone-cycles/s", "n_gpus": 1, "steps": steps, "warmup": 1,
            "ms_per_step": g["ms"] / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": bt["workload"], "form_of_value": "batched hipGraph"},
            "roofline": None, "cpu_baseline": None, "batch": bt}), flush=True)
        return

    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    if args.workload == "sphere":
        steps = args.steps or 50
        log("[bench] building 256x128x128 spherical torus")
        prob = ics.torus_sph(256, 128, 128, steps=steps)
        workload = ("spherical torus 256x128x128 (log r in [10,110], theta in [0.2,pi-0.2], phi periodic), "
                    "torus_sph hooks: PW gravity + floors + radial inflow clamp, beta=100, "
                    "PPM+HLL+upwind CT+SSP-RK3")
    elif args.workload == "torus":
        steps = args.steps or 50
        log("[bench] building 256x128x256 cylindrical torus")
        prob = ics.torus_cyn(256, 128, 256, steps=steps)
        workload = ("cylindrical torus 256x128x256 (log R in [1,40], phi periodic, z in [-20,20]), beta=100, "
                    "PW gravity + floors + inflow clamp, PPM+HLL+upwind CT+SSP-RK3")
    else:
        n = args.n
        steps = args.steps or (100 if n >= 512 else 200)
        log(f"[bench] building {n}^3 spectral initial condition")
        prob = ics.spectral_box(n=n, seed=1001, steps=steps)
        workload = f"3D periodic spectral MHD box {n}^3, PPM+HLL+upwind CT+SSP-RK3"
    cfg = prob.cfg
    from cumpic_amd import config as C
    if args.recon != "default":
        cfg.recon = {"ppm": C.PPMC, "tvdmm": C.TVDMM, "tvdmc": C.TVDMC, "tvdvl": C.TVDVL, "weno": C.WENO}[args.recon]
    cfg.solver = {"hll": C.HLL, "lf": C.LF, "hllc": C.HLLC, "hlld": C.HLLD}[args.solver]
    if args.recon != "default" or args.solver != "hll":
        workload += f" (options: recon {args.recon}, solver {args.solver})"

    # counter children first, before this process touches the GPU
    traffic = {}
    if not args.no_traffic and args.workload == "box" and steps > 0:
        ic_dir = save_ic(prob)
        try:
            for form in forms:
                traffic[form] = pmc_traffic(ic_dir, form)
                if traffic[form]:
                    t = traffic[form]
                    log(f"[bench] traffic {form}: tail {t['tail_bytes_per_launch'] / 1e9:.3f} GB per launch, "
                        f"step {t['step_bytes'] / cfg.zones:.0f} B per zone-cycle")
        finally:
            shutil.rmtree(ic_dir, ignore_errors=True)
    copy_gbps = run_copy_probe()

    cpu, parity = None, None
    if not args.no_cpu and "fp64" in forms:
        log("[bench] CPU restatement: 1 step on the host cores")
        cpu, ostate = cpu_baseline(prob, steps=1, one_core_n=128 if args.workload == "box" else 0)
        log(f"[bench] cpu {cpu['value']:.3e} zone-cycles/s")
        g = Solver.from_problem(prob, precision=FP64)
        g.initial_update()
        g.run(1)
        gs = g.download()
        g.close()
        same = all(np.array_equal(gs[k], ostate[k], equal_nan=True) for k in ("prim", "eps", "bcell"))
        parity = {"one_step_bitwise": bool(same)}
        del ostate, gs
        log(f"[bench] full-size parity after one step: {parity}")

    results = {}
    ref_state, ref_dts, warm_dts = None, None, None
    for form in forms:
        prec = FP64 if form == "fp64" else MIXED
        log(f"[bench] {form}: upload, initial_update, {args.warmup} warm-up, {steps} timed steps")
        s, dts, wall, kt, tuning = timed_run(Solver, prob, prec, steps, args.warmup, tune=not args.no_tune)
        dsig, dabs = s.find_divb()
        st = s.download()
        finite = bool(np.isfinite(st["prim"]).all())
        tr = traffic.get(form)
        r = {"value": cfg.zones * steps / wall, "ms_per_step": wall / steps * 1e3,
             "roofline": roofline(kt, cfg, form, traffic=tr),
             "step_traffic_bytes_per_zone_cycle": (tr["step_bytes"] / cfg.zones) if tr else None,
             "traffic_bytes_per_launch_by_kernel": tr["bytes_per_launch_by_class"] if tr else None,
             "kernel_ms_per_step": {k: v[0] / steps for k, v in kt.items()}, "divb_max_abs": dabs,
             "state_finite": finite, "tuning": tuning}
        kio = r["roofline"]["kernel_io"]
        if copy_gbps:
            kio["frac_of_copy_probe"] = kio["achieved_GBps"] / copy_gbps
        if form == "fp64":
            # the tail's access pattern alone (stages 1-2: 26 planes read, stage 3: 34; 20 written),
            # same layout and column tiles, no arithmetic: the bandwidth this kernel could reach
            cells = (cfg.nx + 1) * (cfg.ny + 1) * (cfg.nz + 1)
            g1, g23 = s.pattern_probe(26, 20), s.pattern_probe(34, 20)
            t_stage = (2 * 46 * 8 * cells / (g1 * 1e9) + 54 * 8 * cells / (g23 * 1e9)) / 3
            ceiling = KERNEL_IO_BYTES["fp64"]["tail"] * cells / t_stage / 1e9
            kio["pattern_ceiling_GBps"] = ceiling
            kio["frac_of_pattern_ceiling"] = kio["achieved_GBps"] / ceiling
            log(f"[bench] tail access pattern without arithmetic: {g1:.0f} GB/s (26 read + 20 written planes), "
                f"{g23:.0f} GB/s (34 + 20); the tail reaches {kio['achieved_GBps'] / ceiling:.2f} of it")
        s.close()
        if form == "fp64":
            ref_dts = dts
            ref_state = fields(st, cfg)
        results[form] = r
        del st
        log(f"[bench] {form}: {r['value']:.4e} zone-cycles/s ({r['ms_per_step']:.2f} ms/step)")

    if "mixed" in results and ref_state is not None and not args.no_error:
        # error of the mixed form against FP64 on the FP64 run's dt sequence (warm-up + timed)
        log("[bench] mixed vs FP64 error replay")
        f64 = Solver.from_problem(prob, precision=FP64)
        f64.initial_update()
        warm_dts = f64.run(args.warmup)[0] if args.warmup else np.zeros(0)
        f64.close()
        s = Solver.from_problem(prob, precision=MIXED)
        s.initial_update()
        for dt in list(warm_dts) + list(ref_dts):
            s.rungekutta(float(dt))
        mf = fields(s.download(), cfg)
        s.close()
        errs = {k: float(np.sum(np.abs(mf[k] - ref_state[k])) / np.sum(np.abs(ref_state[k]))) for k in mf}
        results["mixed"]["rel_l1_vs_fp64"] = errs
        log(f"[bench] mixed rel L1 vs FP64: {errs}")

    c2 = c0 = c3 = None
    if not args.no_c2 and args.workload == "box" and args.n == 256 and set(forms) == {"fp64", "mixed"}:
        c2 = bench_c2()
    if not args.no_c0 and args.workload == "box" and not args.no_cpu:
        c0 = bench_c0()
    if not args.no_c3 and args.workload == "box" and set(forms) == {"fp64", "mixed"} and not args.no_cpu:
        c3 = bench_c3()

    batch = None
    if not args.no_batch and args.workload == "box":
        batch = bench_batch(fd_dirs=args.fd_dirs)

    main_form = "fp64" if "fp64" in results else forms[0]
    m = results[main_form]
    step_alg = m["value"] * BYTES_PER_ZONE_CYCLE
    out = {
        "metric": "zone-cycles/s FP64 vs mixed, % HBM roofline, adjoint/forward step cost",
        "value": m["value"],
        "unit": "zone-cycles/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_form == "fp64" else "f64-state/f32-flux",
        "data": "synthetic",
        "config": {"workload": workload, "zones": cfg.zones, "grid": [cfg.nx, cfg.ny, cfg.nz],
                   "gamma": cfg.ggas, "cfl": cfg.cfl,
                   "seed": {"box": 1001, "torus": 1003, "sphere": 1007}[args.workload],
                   "form_of_value": main_form},
        "roofline": m["roofline"],
        "cpu_baseline": cpu,
        "step_roofline": {"bytes_per_zone_cycle": BYTES_PER_ZONE_CYCLE, "achieved_GBps": step_alg / 1e9,
                          "frac": step_alg / HBM_PEAK,
                          "traffic_bytes_per_zone_cycle": m["step_traffic_bytes_per_zone_cycle"],
                          "traffic_bytes_per_launch_by_kernel": m["traffic_bytes_per_launch_by_kernel"]},
        "kernel_ms_per_step": m["kernel_ms_per_step"],
        "tuning": m["tuning"],
        "divb_max_abs": m["divb_max_abs"],
        "state_finite": m["state_finite"],
        "parity": parity,
        "copy_probe_GBps": copy_gbps,
    }
    if "mixed" in results and main_form != "mixed":
        mx = results["mixed"]
        out["mixed"] = {"value": mx["value"], "ms_per_step": mx["ms_per_step"],
                        "speedup_vs_fp64": mx["value"] / m["value"], "rel_l1_vs_fp64": mx.get("rel_l1_vs_fp64"),
                        "divb_max_abs": mx["divb_max_abs"], "state_finite": mx["state_finite"],
                        "roofline": mx["roofline"], "kernel_ms_per_step": mx["kernel_ms_per_step"],
                        "tuning": mx["tuning"]}
    out["c0"] = c0
    out["c2"] = c2
    out["c3"] = c3
    out["batch"] = batch
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
