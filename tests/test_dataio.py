"""Profile files in the reference's HDF5 layout (solver/dataio.f90:7-331): what the writer puts
in a file, read back through the HDF5 C library.  The reference writes prim(:,1:nx,1:ny,1:nz)
and eps(1:nx,1:ny,1:nz) from Fortran arrays, which an HDF5 reader sees with the dimension order
reversed, and the face coordinates x_F(0:n) only for axes with more than one cell
(dataio.f90:52-54, 157-195)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from cumpic_amd import build, ics


@pytest.fixture(scope="module")
def dataio():
    build.build_h5profile()
    from cumpic_amd import dataio as d
    try:
        d._hdf5()
    except RuntimeError:
        pytest.skip("the HDF5 C library is not available")
    return d


def _check(d, prob, tmp_path, n_iter, axes):
    c = prob.cfg
    path = d.write_profile(str(tmp_path), n_iter, c, prob.xf, prob.yf, prob.zf, prob.prim, prob.eps)
    assert os.path.basename(path) == f"star_weno_data_{n_iter}.hdf5"
    got = d.read_profile(path)
    assert set(got) == set(axes) | {"prim", "eps"}
    inner = (slice(3, 3 + c.nz), slice(3, 3 + c.ny), slice(3, 3 + c.nx))
    assert got["prim"].shape == (c.nz, c.ny, c.nx, 8)
    assert got["eps"].shape == (c.nz, c.ny, c.nx)
    assert np.array_equal(got["prim"], prob.prim[inner])
    assert np.array_equal(got["eps"], prob.eps[inner])
    for ax, faces, n in (("xf", prob.xf, c.nx), ("yf", prob.yf, c.ny), ("zf", prob.zf, c.nz)):
        if ax in axes:
            assert np.array_equal(got[ax], faces[3:3 + n + 1])


def test_profile_3d(dataio, tmp_path):
    prob = ics.randomize(ics.small_periodic(9, seed=5100), seed=1, amp=0.2)
    _check(dataio, prob, tmp_path, 0, ("xf", "yf", "zf"))


def test_profile_2d_and_1d_skip_single_cell_axes(dataio, tmp_path):
    _check(dataio, ics.spectral_box(n=12, seed=5200, ndim=2, ny=10), tmp_path, 3, ("xf", "yf"))
    _check(dataio, ics.riemann_1d(n=40), tmp_path, 7, ("xf",))


def test_profile_rejects_wrong_shapes_and_bad_directory(dataio, tmp_path):
    prob = ics.small_periodic(6, seed=5300)
    c = prob.cfg
    with pytest.raises(ValueError):
        dataio.write_profile(str(tmp_path), 0, c, prob.xf, prob.yf, prob.zf, prob.prim[..., :7], prob.eps)
    with pytest.raises(RuntimeError):
        dataio.write_profile(str(tmp_path / "missing" / "dir"), 0, c, prob.xf, prob.yf, prob.zf, prob.prim,
                             prob.eps)
