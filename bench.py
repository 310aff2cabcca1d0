#!/usr/bin/env python3
"""Benchmark of the MI355X CUMC3D MHD step on one GPU.

Workload (default): the 3D periodic spectral MHD box of 256^3 zones, gamma = 5/3, beta = 1,
mean B_z = 1, PPM + HLL + upwind CT + SSP-RK3 in FP64, seeded synthetic initial condition
(cumpic_amd.ics.spectral_box, seed 1001).  One step = finddt + one RUNGEKUTTA (three RK stages).

Prints ONE JSON line with the zone-cycles/s of the timed steps, the roofline of the dominant
kernel (HIP events on the library stream during the timed region) and the CPU restatement's
rate on the host cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s
BYTES_PER_ZONE_CYCLE = 512.0  # 64 FP64 state words per RK3 step (read U0 x3, stage states, writes)
# algorithmic bytes per unit of each kernel class: loads + stores of its compulsory FP64 words
KERNEL_BYTES = {
    "flux_x": 15 * 8, "flux_y": 15 * 8, "flux_z": 15 * 8,  # 7 stencil vars + normal B in, 7 fluxes out
    "ecorn": 15 * 8,                                        # 6 B-fluxes + v + cell B in, 3 EMFs out
    # 15 fluxes + 3 EMFs + U0 (+ Uin in stages 2, 3) in; prim 8, eps, bcell 3, Uout 8 out:
    # (46 + 54 + 54) / 3 words averaged over the three stage launches
    "tail": (46 + 54 + 54) / 3 * 8,
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_units(name, n):
    nx = ny = nz = n
    return {
        "flux_x": (nx + 1) * (ny + 2) * (nz + 2),
        "flux_y": (nx + 2) * (ny + 1) * (nz + 2),
        "flux_z": (nx + 2) * (ny + 2) * (nz + 1),
        "ecorn": (nx + 1) * (ny + 1) * (nz + 1),
        "tail": (nx + 1) * (ny + 1) * (nz + 1),
    }.get(name)


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2 and parts[0] == kernel:
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


def cpu_baseline(prob, steps=1):
    """The CPU restatement on the host cores: same initial state, `steps` RK3 steps."""
    from oracle.oracle import Oracle
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = Oracle.from_problem(prob)
    o.initial_update()
    t0 = time.perf_counter()
    o.run(steps)
    wall = time.perf_counter() - t0
    state = o.get_state()
    o.close()
    return {
        "value": prob.cfg.zones * steps / wall,
        "unit": "zone-cycles/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{prob.cfg.nx}x{prob.cfg.ny}x{prob.cfg.nz} spectral box, {steps} RK3 step(s) after initial_update, "
                  f"{wall:.1f} s wall on {cores} OpenMP threads",
    }, state


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="zones per axis of the periodic box")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and full-size parity check")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the multi-GPU halo exchange is a design note, not code)", file=sys.stderr)
        sys.exit(2)

    from cumpic_amd import ics
    from cumpic_amd.config import FP64, MIXED
    from cumpic_amd.cumc import Solver, copy_probe

    copy_gbps = None
    if args.copy_probe_mib:
        copy_gbps = copy_probe(args.copy_probe_mib << 20, reps=3)
        log(f"[bench] copy probe {args.copy_probe_mib} MiB: {copy_gbps:.1f} GB/s")

    n = args.n
    log(f"[bench] building {n}^3 spectral initial condition")
    prob = ics.spectral_box(n=n, seed=1001, steps=args.steps)
    prec = MIXED if args.precision == "mixed" else FP64

    parity = None
    cpu = None
    if not args.no_cpu:
        log("[bench] CPU restatement: 1 step on the host cores")
        cpu, ostate = cpu_baseline(prob, steps=1)
        log(f"[bench] cpu {cpu['value']:.3e} zone-cycles/s")
        g = Solver.from_problem(prob, precision=prec)
        g.initial_update()
        g.run(1)
        gs = g.download()
        g.close()
        if prec == FP64:
            same = all(np.array_equal(gs[k], ostate[k], equal_nan=True) for k in ("prim", "eps", "bcell"))
            parity = {"one_step_bitwise": bool(same)}
        else:
            a = gs["prim"][3:3 + n, 3:3 + n, 3:3 + n, 0]
            b = ostate["prim"][3:3 + n, 3:3 + n, 3:3 + n, 0]
            parity = {"one_step_rel_l1_rho": float(np.sum(np.abs(a - b)) / np.sum(np.abs(b)))}
        del ostate, gs
        log(f"[bench] full-size parity after one step: {parity}")

    log("[bench] GPU: upload, initial_update, warmup")
    s = Solver.from_problem(prob, precision=prec)
    del prob
    s.initial_update()
    if args.warmup:
        s.run(args.warmup)
    s.profile(True)
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    dts, _ = s.run(args.steps)  # synchronous on return (hipStreamSynchronize)
    wall = time.perf_counter() - t0
    kt = s.kernel_times()
    s.profile(False)
    dsig, dabs = s.find_divb()
    final = s.download()["prim"][3:3 + n, 3:3 + n, 3:3 + n]
    finite = bool(np.isfinite(final).all())
    del final
    zones = n ** 3
    value = zones * args.steps / wall

    # dominant kernel of the timed region
    dom = max((k for k in kt if k in KERNEL_BYTES), key=lambda k: kt[k][0])
    tot_ms, launches = kt[dom]
    avg_s = tot_ms / launches * 1e-3
    alg_bytes = KERNEL_BYTES[dom] * kernel_units(dom, n)
    achieved = alg_bytes / avg_s
    roof = {
        "bound": "hbm",
        "achieved": achieved / 1e9,
        "peak": HBM_PEAK / 1e9,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK,
        "traffic": read_traffic(dom),
        "kernel": dom,
        "avg_launch_ms": tot_ms / launches,
    }
    step_alg = value * BYTES_PER_ZONE_CYCLE
    out = {
        "metric": "zone-cycles/s FP64 vs mixed, % HBM roofline, adjoint/forward step cost",
        "value": value,
        "unit": "zone-cycles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == FP64 else "f64-state/f32-flux",
        "data": "synthetic",
        "config": {
            "workload": f"3D periodic spectral MHD box {n}^3, PPM+HLL+upwind CT+SSP-RK3, "
                        f"{'FP64' if prec == FP64 else 'mixed FP32 reconstruction/HLL/EMF, FP64 state'}",
            "zones": zones,
            "grid": [n, n, n],
            "gamma": 5.0 / 3.0,
            "cfl": 0.4,
            "seed": 1001,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "step_roofline": {"bytes_per_zone_cycle": BYTES_PER_ZONE_CYCLE, "achieved_GBps": step_alg / 1e9,
                          "frac": step_alg / HBM_PEAK},
        "kernel_ms_per_step": {k: v[0] / args.steps for k, v in kt.items()},
        "divb_max_abs": dabs,
        "state_finite": finite,
        "copy_probe_GBps": copy_gbps,
        "parity": parity,
        "dt_first_last": [float(dts[0]), float(dts[-1])] if len(dts) else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
