"""Compile the HIP library (and the Fortran host) in-tree for gfx950."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libcumc.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Common flags.  -packed-fp32-ops off: the compiler must not form v_pk_{add,mul,fma}_f32.
COMMON = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC",
    "-Wno-unused-result", "-Wno-unused-value",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# FP64 drop-in path: no fused multiply-add, every operation rounds as in the reference source.
FP64_FLAGS = ["-ffp-contract=off", "-fno-fast-math"]
# Mixed form: the FP32 reconstruction / Riemann / EMF kernels may contract multiply-adds and use
# the hardware FP32 reciprocal and square root; FP32 subnormals (below 1.2e-38) flush to zero,
# which lets an FP32 divide lower to reciprocal + multiply without exponent rescaling.  FP64
# arithmetic in this unit (the stage tail) keeps IEEE subnormals.
MIXED_FLAGS = ["-ffp-contract=fast", "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero"]

UNITS = [("cumc.hip", FP64_FLAGS), ("cumc_mixed.hip", MIXED_FLAGS), ("cumc_batch.hip", FP64_FLAGS),
         ("cumc_adjoint.hip", FP64_FLAGS)]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


# The diagnostic build: the same sources with the branch-signature hooks compiled in
# (CUMC_BRANCH_SIG; cumc_batch_branch_sig records).  Only the finite-difference diagnostics of the
# reverse mode load it.
LIB_SIG = os.path.join(PKG, "libcumc_sig.so")


def build_library(verbose: bool = False, variant: str = "") -> str:
    """libcumc.so, or with variant="sig" the diagnostic libcumc_sig.so.  The units compile in
    parallel (one hipcc process each)."""
    from concurrent.futures import ThreadPoolExecutor
    inc = ["-I", os.path.join(ROOT, "include")]
    extra = ["-DCUMC_BRANCH_SIG"] if variant == "sig" else []
    target = LIB_SIG if variant == "sig" else LIB
    with tempfile.TemporaryDirectory() as td:
        cmds = [[HIPCC, *COMMON, *flags, *extra, *inc, "-c", "-o", os.path.join(td, src + ".o"),
                 os.path.join(PKG, "csrc", src)] for src, flags in UNITS]
        with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
            outs = list(ex.map(_run, cmds))
        if verbose:
            print("\n".join(outs))
        tmp = target + ".tmp%d" % os.getpid()
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *[c[c.index("-o") + 1] for c in cmds]])
    os.replace(tmp, target)
    return target


def flang() -> str | None:
    for c in ("/opt/rocm/lib/llvm/bin/flang", shutil.which("flang") or "", shutil.which("amdflang") or ""):
        if c and os.path.exists(c):
            return c
    return None


FHOST_DIR = os.path.join(PKG, "fhost")
# the host built with the run-file plugin keeps the plain name; other plugins are cumc3d_<model>
FHOST_EXE = os.path.join(FHOST_DIR, "cumc3d")
FHOST_MODELS = ("file", "riemann")


def fhost_exe(model: str = "file") -> str:
    return FHOST_EXE if model == "file" else FHOST_EXE + "_" + model


H5_LIB = os.path.join(FHOST_DIR, "libcumc_h5.so")


def build_h5profile() -> str:
    """HDF5 profile writer of the host (h5profile.c; opens libhdf5 at run time)."""
    tmp = H5_LIB + ".tmp%d" % os.getpid()
    _run(["gcc", "-O2", "-shared", "-fPIC", "-Wall", "-o", tmp, os.path.join(FHOST_DIR, "h5profile.c"), "-ldl"])
    os.replace(tmp, H5_LIB)
    return H5_LIB


def build_fortran_host(models=FHOST_MODELS) -> str | None:
    """Fortran driver over the C ABI (flang), once per model plugin: the plugin's sources are the
    ones its model/<name>/source.list names, as the reference's build takes model/<m>/src/source.list.
    Returns the run-file host's path, or None without flang."""
    fc = flang()
    if fc is None:
        return None
    build_h5profile()
    for model in models:
        with open(os.path.join(FHOST_DIR, "model", model, "source.list")) as f:
            plugin = [os.path.join(FHOST_DIR, ln.strip()) for ln in f if ln.strip()]
        srcs = [os.path.join(FHOST_DIR, "cumc_iface.f90"), *plugin, os.path.join(FHOST_DIR, "cumc3d.f90")]
        mod = os.path.join(FHOST_DIR, "_mod", model)
        os.makedirs(mod, exist_ok=True)
        exe = fhost_exe(model)
        cmd = [fc, "-O2", "-J", mod, *srcs, "-L", PKG, "-lcumc", "-L", FHOST_DIR, "-lcumc_h5",
               "-Wl,-rpath," + PKG, "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"]
        _run(cmd)
        os.replace(exe + ".tmp", exe)
    return FHOST_EXE
