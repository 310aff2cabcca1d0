"""Regenerate tests/golden/oracle_runs.json: SHA-256 digests (and a few summary values) of the
CPU restatement's final state for seven canonical runs.  The reference ships no tests, fixtures or
golden data and its Fortran build needs an HDF5 module this image lacks, so these are not
reference outputs: they freeze the restatement (pinned by tests/test_oracle_pins.py) so that any
later change of its arithmetic is deliberate.  Run from the repository root:

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cumpic_amd import ics  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_runs.json")


def cases():
    """name -> (problem, steps or None for the problem's own end time)"""
    return {
        "riemann_1d_800_t0.1": (ics.riemann_1d(), None),
        "periodic_12_ppm_4steps": (ics.randomize(ics.small_periodic(12, seed=1005), seed=7, amp=0.15), 4),
        "torus_cyn_24x12x24_3steps": (ics.torus_cyn(24, 12, 24), 3),
        "spherical_shell_14x10x8_3steps": (ics.spherical_shell(14, 10, 8, seed=1500, noise=0.2,
                                                               boundary_flag=(1, 1, 1, 1, 0, 0)), 3),
        "toro1_200_t0.2": (ics.riemann_test(ics.TORO_1, n=200), None),
        "briowu_200_t0.1": (ics.riemann_test(ics.BRIO_WU, n=200), None),
        "mhdrotor_64sq_t0.15": (ics.riemann_test(ics.MHD_ROTOR, n=64), None),
    }


def digest(prob, steps):
    o = Oracle.from_problem(prob)
    o.initial_update()
    if steps is None:
        dts, t = o.run(100000, prob.total_time)
    else:
        dts, t = o.run(steps)
    s = o.get_state()
    h = hashlib.sha256()
    for k in ("prim", "eps", "bcell"):
        h.update(np.ascontiguousarray(s[k], dtype="<f8").tobytes())
    h.update(np.ascontiguousarray(dts, dtype="<f8").tobytes())
    return {"sha256": h.hexdigest(), "steps": int(len(dts)), "time": float(t),
            "sum_rho": float(np.sum(s["prim"][..., 0])), "max_abs_divb": float(o.find_divb()[1])}


def main():
    out = {name: digest(p, n) for name, (p, n) in cases().items()}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
