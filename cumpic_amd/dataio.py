"""Profile files in the reference's HDF5 layout (solver/dataio.f90:7-331, print_hydroprofile).

`write_profile` calls the Fortran host's writer (cumpic_amd/fhost/h5profile.c, built into
fhost/libcumc_h5.so), so Python and the Fortran driver produce the same files:
<dir>/star_weno_data_<n_iter>.hdf5 with datasets xf/yf/zf (faces 0..n, only for axes with more
than one cell), prim (interior, face B in prim(6:8); HDF5 dims (nz, ny, nx, 8)) and eps
(dims (nz, ny, nx)).  `read_profile` reads such a file back through the HDF5 C library
directly (no h5py in this image).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .build import H5_LIB

_HDF5_CANDIDATES = ("libhdf5.so", "/opt/conda/lib/libhdf5.so")


def profile_name(n_iter: int) -> str:
    """File name of profile n_iter (dataio.f90:31-34)."""
    return f"star_weno_data_{int(n_iter)}.hdf5"


def _writer():
    if not os.path.exists(H5_LIB):
        raise RuntimeError(f"{H5_LIB} is not built (cumpic_amd.build.build_h5profile)")
    lib = ctypes.CDLL(H5_LIB)
    f = lib.cumc_h5_profile
    dp = ctypes.POINTER(ctypes.c_double)
    f.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 4 + [dp] * 5
    f.restype = ctypes.c_int
    return f


def write_profile(directory: str, n_iter: int, cfg, xf, yf, zf, prim, eps) -> str:
    """Write one profile from host arrays in the reference's ghost-inclusive layout: faces
    -3..n+3 per axis, prim of shape (nz+6, ny+6, nx+6, 8), eps (nz+6, ny+6, nx+6) (C order of
    the Fortran prim(8,-2:nx+3,-2:ny+3,-2:nz+3)).  Returns the file path."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (xf, yf, zf, prim, eps)]
    want = [(cfg.nx + 7,), (cfg.ny + 7,), (cfg.nz + 7,), cfg.shape + (8,), cfg.shape]
    for a, w in zip(arrs, want):
        if a.shape != w:
            raise ValueError(f"array of shape {a.shape}, expected {w}")
    dp = ctypes.POINTER(ctypes.c_double)
    rc = _writer()(os.fsencode(directory), int(n_iter), cfg.nx, cfg.ny, cfg.nz,
                   *[a.ctypes.data_as(dp) for a in arrs])
    if rc != 0:
        raise RuntimeError("HDF5 profile write failed (libhdf5 not found or directory not writable)")
    return os.path.join(directory, profile_name(n_iter))


def _hdf5():
    env = os.environ.get("CUMC_HDF5_LIB")
    for c in ((env,) if env else ()) + _HDF5_CANDIDATES:
        try:
            lib = ctypes.CDLL(c)
        except OSError:
            continue
        if lib.H5open() >= 0:
            return lib
    raise RuntimeError("the HDF5 C library is not available")


def read_profile(path: str) -> dict:
    """All datasets of a profile file: {name: array in HDF5 (C) dimension order}."""
    h = _hdf5()
    hid = ctypes.c_int64
    for fn, res in (("H5Fopen", hid), ("H5Dopen2", hid), ("H5Dget_space", hid), ("H5Lexists", ctypes.c_int),
                    ("H5Sget_simple_extent_ndims", ctypes.c_int), ("H5Sget_simple_extent_dims", ctypes.c_int),
                    ("H5Dread", ctypes.c_int), ("H5Dclose", ctypes.c_int), ("H5Sclose", ctypes.c_int),
                    ("H5Fclose", ctypes.c_int)):
        getattr(h, fn).restype = res
    h.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid]
    h.H5Lexists.argtypes = [hid, ctypes.c_char_p, hid]
    h.H5Dopen2.argtypes = [hid, ctypes.c_char_p, hid]
    h.H5Dget_space.argtypes = [hid]
    h.H5Sget_simple_extent_ndims.argtypes = [hid]
    h.H5Sget_simple_extent_dims.argtypes = [hid, ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_void_p]
    h.H5Dread.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
    h.H5Dclose.argtypes = [hid]
    h.H5Sclose.argtypes = [hid]
    h.H5Fclose.argtypes = [hid]
    native_double = hid.in_dll(h, "H5T_NATIVE_DOUBLE_g").value
    f = h.H5Fopen(os.fsencode(path), 0, 0)  # H5F_ACC_RDONLY, H5P_DEFAULT
    if f < 0:
        raise RuntimeError(f"cannot open {path}")
    out = {}
    try:
        for name in ("xf", "yf", "zf", "prim", "eps"):
            if h.H5Lexists(f, name.encode(), 0) <= 0:
                continue
            d = h.H5Dopen2(f, name.encode(), 0)
            sp = h.H5Dget_space(d)
            nd = h.H5Sget_simple_extent_ndims(sp)
            dims = (ctypes.c_ulonglong * nd)()
            h.H5Sget_simple_extent_dims(sp, dims, None)
            arr = np.empty(tuple(int(x) for x in dims), dtype=np.float64)
            rc = h.H5Dread(d, native_double, 0, 0, 0, arr.ctypes.data)  # H5S_ALL, H5P_DEFAULT
            h.H5Sclose(sp)
            h.H5Dclose(d)
            if rc < 0:
                raise RuntimeError(f"reading {name} from {path} failed")
            out[name] = arr
    finally:
        h.H5Fclose(f)
    return out
