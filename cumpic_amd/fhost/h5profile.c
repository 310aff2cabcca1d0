/*
 * h5profile.c -- profile output of the Fortran host in the reference's HDF5 layout
 * (solver/dataio.f90:7-331, print_hydroprofile): one file per output,
 * <dir>/star_weno_data_<n_iter>.hdf5, holding
 *   xf(0:nx), yf(0:ny), zf(0:nz)      face coordinates, each only when that axis has more than
 *                                     one cell (dataio.f90:52-54)
 *   prim(1:8, 1:nx, 1:ny, 1:nz)       interior primitives, face B in prim(6:8) (dataio.f90:275-331)
 *   eps(1:nx, 1:ny, 1:nz)             interior specific internal energy (dataio.f90:213-268)
 * as H5T_NATIVE_DOUBLE datasets written from Fortran-ordered arrays, so an HDF5 reader sees the
 * C-order dimensions (nz, ny, nx, 8) and (nz, ny, nx), exactly as the Fortran HDF5 writer
 * leaves them.
 *
 * The HDF5 C library (1.10 or later: 64-bit hid_t) is opened at run time with dlopen -- from
 * CUMC_HDF5_LIB, else the dynamic loader's path, else /opt/conda/lib -- so the host builds
 * and runs without it; a write without the library returns -1.
 */
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;

static struct {
    void *lib;
    herr_t (*open)(void);
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
    herr_t (*sselect_hyperslab)(hid_t, int, const hsize_t *, const hsize_t *, const hsize_t *, const hsize_t *);
    herr_t (*sclose)(hid_t);
    hid_t (*dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    herr_t (*dclose)(hid_t);
    hid_t *native_double;
} H;

enum { H5P_DEFAULT_ = 0, H5S_SELECT_SET_ = 0, H5F_ACC_TRUNC_ = 2 };

static int load(void)
{
    if (H.lib) return 0;
    const char *env = getenv("CUMC_HDF5_LIB");
    const char *cands[] = {env, "libhdf5.so", "/opt/conda/lib/libhdf5.so"};
    for (int i = 0; i < 3 && !H.lib; i++)
        if (cands[i] && cands[i][0]) H.lib = dlopen(cands[i], RTLD_NOW | RTLD_LOCAL);
    if (!H.lib) return -1;
#define SYM(field, name)                                   \
    *(void **)(&H.field) = dlsym(H.lib, name);             \
    if (!H.field) return -1;
    SYM(open, "H5open")
    SYM(fcreate, "H5Fcreate")
    SYM(fclose, "H5Fclose")
    SYM(screate_simple, "H5Screate_simple")
    SYM(sselect_hyperslab, "H5Sselect_hyperslab")
    SYM(sclose, "H5Sclose")
    SYM(dcreate2, "H5Dcreate2")
    SYM(dwrite, "H5Dwrite")
    SYM(dclose, "H5Dclose")
#undef SYM
    if (H.open() < 0) return -1;
    H.native_double = (hid_t *)dlsym(H.lib, "H5T_NATIVE_DOUBLE_g");
    return H.native_double ? 0 : -1;
}

/* dataset `name` of file dims fdims (rank r) from a memory array of dims mdims, taking the block
 * at offset off (count = fdims) */
static int put(hid_t file, const char *name, int r, const hsize_t *fdims, const hsize_t *mdims,
               const hsize_t *off, const double *buf)
{
    int rc = -1;
    hid_t fs = H.screate_simple(r, fdims, NULL);
    hid_t ms = H.screate_simple(r, mdims, NULL);
    hid_t ds = -1;
    if (fs < 0 || ms < 0) goto out;
    if (H.sselect_hyperslab(ms, H5S_SELECT_SET_, off, NULL, fdims, NULL) < 0) goto out;
    ds = H.dcreate2(file, name, *H.native_double, fs, H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_);
    if (ds < 0) goto out;
    if (H.dwrite(ds, *H.native_double, ms, fs, H5P_DEFAULT_, buf) < 0) goto out;
    rc = 0;
out:
    if (ds >= 0) H.dclose(ds);
    if (ms >= 0) H.sclose(ms);
    if (fs >= 0) H.sclose(fs);
    return rc;
}

/* Write one profile.  xf/yf/zf: faces -3..n+3 (n+7 values); prim: prim(8, -2:nx+3, -2:ny+3,
 * -2:nz+3); eps: eps(-2:nx+3, -2:ny+3, -2:nz+3), Fortran order.  Returns 0, or -1 on any
 * failure (library missing, file not creatable, write error). */
int cumc_h5_profile(const char *dir, int n_iter, int nx, int ny, int nz, const double *xf, const double *yf,
                    const double *zf, const double *prim, const double *eps)
{
    if (!dir || nx < 1 || ny < 1 || nz < 1 || load() != 0) return -1;
    char path[4096];
    if (snprintf(path, sizeof path, "%s/star_weno_data_%d.hdf5", dir, n_iter) >= (int)sizeof path) return -1;
    hid_t f = H.fcreate(path, H5F_ACC_TRUNC_, H5P_DEFAULT_, H5P_DEFAULT_);
    if (f < 0) return -1;
    int rc = 0;
    const int n[3] = {nx, ny, nz};
    const double *faces[3] = {xf, yf, zf};
    const char *names[3] = {"xf", "yf", "zf"};
    for (int a = 0; a < 3 && rc == 0; a++) {
        if (n[a] <= 1) continue;
        const hsize_t fd[1] = {(hsize_t)n[a] + 1}, md[1] = {(hsize_t)n[a] + 7}, off[1] = {3};
        rc = put(f, names[a], 1, fd, md, off, faces[a]);
    }
    if (rc == 0) {
        const hsize_t fd[4] = {(hsize_t)nz, (hsize_t)ny, (hsize_t)nx, 8};
        const hsize_t md[4] = {(hsize_t)nz + 6, (hsize_t)ny + 6, (hsize_t)nx + 6, 8};
        const hsize_t off[4] = {3, 3, 3, 0};
        rc = put(f, "prim", 4, fd, md, off, prim);
    }
    if (rc == 0) {
        const hsize_t fd[3] = {(hsize_t)nz, (hsize_t)ny, (hsize_t)nx};
        const hsize_t md[3] = {(hsize_t)nz + 6, (hsize_t)ny + 6, (hsize_t)nx + 6};
        const hsize_t off[3] = {3, 3, 3};
        rc = put(f, "eps", 3, fd, md, off, eps);
    }
    if (H.fclose(f) < 0) rc = -1;
    return rc;
}
