"""Compile the HIP library (and the Fortran host) in-tree for gfx950."""
from __future__ import annotations

import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libcumc.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no fused multiply-add, every operation rounds as in the reference source.
# -packed-fp32-ops off: the compiler must not form v_pk_{add,mul,fma}_f32.
HIP_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Wno-unused-result", "-Wno-unused-value",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build_library(verbose: bool = False) -> str:
    src = os.path.join(PKG, "csrc", "cumc.hip")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, src]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout)
    if verbose:
        print(r.stdout)
    os.replace(tmp, LIB)
    return LIB


def flang() -> str | None:
    for c in ("/opt/rocm/lib/llvm/bin/flang", shutil.which("flang") or "", shutil.which("amdflang") or ""):
        if c and os.path.exists(c):
            return c
    return None


FHOST_DIR = os.path.join(PKG, "fhost")
FHOST_EXE = os.path.join(FHOST_DIR, "cumc3d")


def build_fortran_host() -> str | None:
    """Fortran driver over the C ABI (flang); returns the executable path or None without flang."""
    fc = flang()
    if fc is None:
        return None
    srcs = [os.path.join(FHOST_DIR, f) for f in ("cumc_iface.f90", "model_file.f90", "cumc3d.f90")]
    mod = os.path.join(FHOST_DIR, "_mod")
    os.makedirs(mod, exist_ok=True)
    cmd = [fc, "-O2", "-J", mod, *srcs, "-L", PKG, "-lcumc", "-Wl,-rpath," + PKG, "-Wl,-rpath,$ORIGIN/..",
           "-o", FHOST_EXE + ".tmp"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("flang failed:\n" + r.stdout)
    os.replace(FHOST_EXE + ".tmp", FHOST_EXE)
    return FHOST_EXE
