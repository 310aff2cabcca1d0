"""Seeded initial conditions for the benchmark and parity workloads.

Each builder returns a :class:`Problem`: the run configuration, the face
coordinates xF/yF/zF(-3:n+3) and the primitive state in the reference's
ghost-inclusive array layout (``prim`` C-shape (nz+6, ny+6, nx+6, 8), which is
the Fortran ``prim(1:8,-2:nx+3,-2:ny+3,-2:nz+3)``; ``eps`` (nz+6, ny+6, nx+6)).
Face-centred B lives in prim[..., 5:8] on the reference's staggering
(prim(ibx,j,k,l) is B_x on the face x_F(j)), as GET_MODEL fills it
(model/riemann/src/model.f90:4-25, model/torus_cyn/src/model.f90:149-158).

Seeds are ``numpy.random.default_rng(1000 + workload_index)``.
"""
from __future__ import annotations

import dataclasses
import math
import os

import numpy as np

from .config import (CARTESIAN, CYLINDRICAL, IBX, IBY, IBZ, IRHO, ITAU, IVX, IVY, IVZ, MODEL_RIEMANN,
                     MODEL_TORUS_CYN, MODEL_TORUS_SPH, NGHOST, PPMC, SPHERICAL, Config, log_faces,
                     periodic_angle_faces, uniform_faces)

G = NGHOST


def ix(j: int) -> int:
    """Array index of cell (or face) j in a ghost-inclusive axis that starts at cell 1-NGHOST."""
    return j + G - 1


@dataclasses.dataclass
class Problem:
    cfg: Config
    xf: np.ndarray
    yf: np.ndarray
    zf: np.ndarray
    prim: np.ndarray
    eps: np.ndarray
    steps: int = 0
    total_time: float = 1.0e300
    name: str = ""


def _alloc(cfg: Config):
    prim = np.zeros(cfg.shape + (8,), dtype=np.float64)
    eps = np.zeros(cfg.shape, dtype=np.float64)
    return prim, eps


def riemann_1d(n: int = 800, seed: int = 1000, gamma: float = 2.0, cfl: float = 0.8,
               t_end: float = 0.1) -> Problem:
    """1D MHD Riemann problem on [0,1] with seeded left/right states, outflow on all sides.

    rho, p ~ U[0.1,1]; v ~ U[-0.5,0.5]^3; (By, Bz) ~ U[-1,1] with the sign flipped across
    x = 0.5; uniform Bx ~ U[0.5,1].  ny = nz = 1 with unit widths so dx sets the CFL step.
    Draw order: rho_L, rho_R, p_L, p_R, v_L(3), v_R(3), (By, Bz), Bx.
    """
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.1, 1.0, 2)
    p = rng.uniform(0.1, 1.0, 2)
    vl = rng.uniform(-0.5, 0.5, 3)
    vr = rng.uniform(-0.5, 0.5, 3)
    byz = rng.uniform(-1.0, 1.0, 2)
    bx = rng.uniform(0.5, 1.0)
    cfg = Config(nx=n, ny=1, nz=1, coordinate=CARTESIAN, boundary_flag=(1, 1, 1, 1, 1, 1), recon=PPMC,
                 model=MODEL_RIEMANN, ggas=gamma, cfl=cfl)
    xf = uniform_faces(n, 0.0, 1.0)
    yf = uniform_faces(1, 0.0, 1.0)
    zf = uniform_faces(1, 0.0, 1.0)
    prim, eps = _alloc(cfg)
    xc = 0.5 * (xf[1:] + xf[:-1])  # centres of cells j = -2..n+3 (array index j+2)
    left = xc < 0.5
    # GET_MODEL fills [0,n]^3 (riemann/src/model.f90:15-21); ghosts are set by initial_update
    sl = (slice(ix(0), ix(1) + 1), slice(ix(0), ix(1) + 1))
    for side, mask in ((0, left), (1, ~left)):
        cols = np.nonzero(mask)[0]
        cols = cols[(cols >= ix(0)) & (cols <= ix(n))]
        v = vl if side == 0 else vr
        sgn = 1.0 if side == 0 else -1.0
        for c in cols:
            prim[sl + (c, IRHO)] = rho[side]
            prim[sl + (c, IVX)] = v[0]
            prim[sl + (c, IVY)] = v[1]
            prim[sl + (c, IVZ)] = v[2]
            prim[sl + (c, ITAU)] = p[side]
            prim[sl + (c, IBX)] = bx
            prim[sl + (c, IBY)] = sgn * byz[0]
            prim[sl + (c, IBZ)] = sgn * byz[1]
            eps[sl + (c,)] = p[side] / rho[side] / (gamma - 1.0)
    return Problem(cfg, xf, yf, zf, prim, eps, steps=10 ** 9, total_time=t_end, name="riemann1d")


# The riemann plugin's named tests (model/riemann/src/defs.f90:20-22) and its default
# adiabatic index (defs.f90:14).
TORO_1, BRIO_WU, MHD_ROTOR = 1, 2, 3
RIEMANN_GGAS = 5.0 / 3.0


def _cell_centres(f: np.ndarray, n: int) -> np.ndarray:
    """GET_COORD's centre 0.5*(xF(j) + xF(j-1)) of cells j = 0..n (solver/coords.f90:63-65)."""
    return 0.5 * (f[ix(0) + 1:ix(n) + 2] + f[ix(0):ix(n) + 1])


def riemann_test(test: int, n: int = 400, gamma: float = RIEMANN_GGAS, cfl: float = 0.7,
                 recon: int = PPMC, t_end: float | None = None) -> Problem:
    """One of the riemann plugin's three named problems on the unit interval / square, filled
    on [0,n]^3 as GET_MODEL does (model/riemann/src/model.f90:15-21, RIEMANN_PROBLEM :32-57),
    outflow on every side, zero for every primitive a problem leaves unset.

    * TORO_1 (riemann1d.f90:15-47): rho, p = 1, 1 for x <= 0.5, else 0.125, 0.1; no field;
      n x 1 x 1 cells; stop time 0.2.
    * BRIO_WU (riemann1d.f90:54-89): TORO_1's gas with By = +1 / -1 on the two sides and a
      uniform Bx = 0.75; n x 1 x 1 cells; stop time 0.1.
    * MHD_ROTOR (riemann2d.f90:15-63): a disc r <= r0 = 0.1 about (0.5, 0.5) with rho = 10
      rotating rigidly at v = (-(y-0.5), x-0.5)/r0, a linear taper fr = (r1-r)/(r1-r0) out to
      r1 = 0.115 (rho = 1 + 9 fr, v scaled by fr r0/r), rho = 1 at rest beyond; p = 0.5 and
      Bx = 2.5/sqrt(4 pi) everywhere; n x n x 1 cells; stop time 0.15.

    eps = p/rho/(gamma-1) in every case.  Cell coordinates are GET_COORD's centres (ghost
    faces included, as for cell 0).
    """
    if test not in (TORO_1, BRIO_WU, MHD_ROTOR):
        raise ValueError(f"unknown riemann test {test}")
    two_d = test == MHD_ROTOR
    ny = n if two_d else 1
    cfg = Config(nx=n, ny=ny, nz=1, coordinate=CARTESIAN, boundary_flag=(1, 1, 1, 1, 1, 1), recon=recon,
                 model=MODEL_RIEMANN, ggas=gamma, cfl=cfl)
    xf = uniform_faces(n, 0.0, 1.0)
    yf = uniform_faces(ny, 0.0, 1.0)
    zf = uniform_faces(1, 0.0, 1.0)
    prim, eps = _alloc(cfg)
    x = _cell_centres(xf, n)[None, :]
    y = _cell_centres(yf, ny)[:, None]
    shape = (ny + 1, n + 1)
    vx, vy, bx, by = (np.zeros(shape) for _ in range(4))
    if test == MHD_ROTOR:
        r0, r1 = 0.1, 0.115
        r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2)
        fr = (r1 - r) / (r1 - r0)
        inner, outer = r <= r0, r >= r1
        taper = ~(inner | outer)
        rs = np.where(taper, r, 1.0)
        rho = np.where(inner, 10.0, np.where(outer, 1.0, 1.0 + 9.0 * fr))
        vx = np.where(inner, -1.0 * (y - 0.5) / r0, np.where(outer, 0.0, -(fr * 1.0 * (y - 0.5) / rs)))
        vy = np.where(inner, 1.0 * (x - 0.5) / r0, np.where(outer, 0.0, fr * 1.0 * (x - 0.5) / rs))
        p = np.full(shape, 0.5)
        bx[:] = 2.5 / math.sqrt(4.0 * (4.0 * math.atan(1.0)))
        t_def = 0.15
    else:
        left = np.broadcast_to(x <= 0.5, shape)
        rho = np.where(left, 1.0, 0.125)
        p = np.where(left, 1.0, 0.1)
        if test == BRIO_WU:
            by = np.where(left, 1.0, -1.0)
            bx[:] = 0.75
        t_def = 0.2 if test == TORO_1 else 0.1
    sl = (slice(ix(0), ix(1) + 1), slice(ix(0), ix(ny) + 1), slice(ix(0), ix(n) + 1))
    for comp, val in ((IRHO, rho), (IVX, vx), (IVY, vy), (ITAU, p), (IBX, bx), (IBY, by)):
        prim[sl + (comp,)] = val
    eps[sl] = p / rho / (gamma - 1.0)
    name = {TORO_1: "toro1", BRIO_WU: "briowu", MHD_ROTOR: "mhdrotor"}[test]
    return Problem(cfg, xf, yf, zf, prim, eps, steps=10 ** 9,
                   total_time=t_def if t_end is None else t_end, name=name)


def _spectral_field(shape, kmax: float, slope: float, rng, kz_zero: bool = False) -> np.ndarray:
    """Real random field with Gaussian random-phase modes of amplitude |k|^slope on integer
    wavevectors 1 <= |k| <= kmax (zero elsewhere), normalised to unit rms.  shape = (nz, ny, nx);
    kz_zero keeps only k_z = 0 modes."""
    nz, ny, nx = shape
    km = int(math.floor(kmax))
    nkz = nz // 2 + 1
    coeff = np.zeros((ny, nx, nkz), dtype=np.complex128)
    for iz in range(0, min(km, nkz - 1) + 1):
        if kz_zero and iz != 0:
            continue
        for iy in range(-km, km + 1):
            for ixx in range(-km, km + 1):
                k2 = ixx * ixx + iy * iy + iz * iz
                if k2 < 1 or k2 > kmax * kmax:
                    continue
                amp = math.sqrt(k2) ** slope
                coeff[iy % ny, ixx % nx, iz] = amp * (rng.standard_normal() + 1j * rng.standard_normal())
    if ny * nx * nz >= 384 ** 3:
        # large boxes: scipy's multithreaded transform (same coefficients; the sampled field
        # differs from numpy's transform in the last bits, and every consumer is given this array)
        import scipy.fft
        f = scipy.fft.irfftn(coeff, s=(ny, nx, nz), axes=(0, 1, 2), workers=os.cpu_count() or 1)
    else:
        f = np.fft.irfftn(coeff, s=(ny, nx, nz), axes=(0, 1, 2))  # [k, j, l]
    f = np.ascontiguousarray(np.transpose(f, (2, 0, 1)))  # [l, k, j]
    rms = math.sqrt(float(np.mean(f * f)))
    return f / rms if rms > 0 else f


def _curl_faces(ax, ay, az, dx, dy, dz):
    """Face B from edge-centred A on a periodic staggered grid (arrays indexed [l, k, j], entry j
    holding the right face of cell j).

    A_x sits at (xc, yF(k), zF(l)), A_y at (xF(j), yc, zF(l)), A_z at (xF(j), yF(k), zc); B_x at
    the face xF(j), B_y at yF(k), B_z at zF(l).  The area-weighted divergence of the result is
    zero to round-off."""
    def dm(a, axis):
        return a - np.roll(a, 1, axis=axis)
    bx = dm(az, 1) / dy - dm(ay, 0) / dz
    by = dm(ax, 0) / dz - dm(az, 2) / dx
    bz = dm(ay, 2) / dx - dm(ax, 1) / dy
    return bx, by, bz


def _fill_periodic_ghosts(arr, n3):
    """Copy the interior [1,n] periodically into index 0 and the 3-deep ghost shells."""
    for axis, n in enumerate(n3):
        for g in range(1, G + 1):
            lo = [slice(None)] * arr.ndim
            src = [slice(None)] * arr.ndim
            lo[axis], src[axis] = ix(1 - g), ix(n + 1 - g)
            arr[tuple(lo)] = arr[tuple(src)]
            lo[axis], src[axis] = ix(n + g), ix(g)
            arr[tuple(lo)] = arr[tuple(src)]


def spectral_box(n: int = 256, seed: int = 1001, gamma: float = 5.0 / 3.0, cfl: float = 0.4,
                 v_rms: float = 0.1, db_rms: float = 0.1, ndim: int = 3, ny: int | None = None,
                 steps: int = 200) -> Problem:
    """Periodic box on [0,1]^d: rho = 1, mean B_z = 1, p = beta*B^2/2 = 0.5 (beta = 1).

    v and the vector potential A carry Gaussian random-phase modes with |k|^-2 amplitude over
    1 <= |k| <= 8.  v is scaled to rms |v| = v_rms; A is scaled so that rms |dB| = db_rms
    (a literal rms(A) = 0.1 would give rms |dB| of order 2).  B = B_z z-hat + discrete curl A
    on the reference's face staggering, so the area-weighted div B is zero to round-off.
    ndim = 2 restricts the spectrum to k_z = 0 with nz = 1.

    cfl defaults to 0.4: with the reference's dt = cfl*min(dx)/max_i(|v_i| + c_f,i) the unsplit
    3D update sums three directional CFL numbers, and cfl = 0.7 (definition.f90:116) grows
    grid-scale noise into NaN within ~40 steps on the well-resolved 256^3 box.
    """
    rng = np.random.default_rng(seed)
    nx = n
    ny = n if ny is None else ny
    nz = n if ndim == 3 else 1
    cfg = Config(nx=nx, ny=ny, nz=nz, coordinate=CARTESIAN, boundary_flag=(0, 0, 0, 0, 0, 0), recon=PPMC,
                 model=MODEL_RIEMANN, ggas=gamma, cfl=cfl)
    xf = uniform_faces(nx, 0.0, 1.0)
    yf = uniform_faces(ny, 0.0, 1.0)
    zf = uniform_faces(nz, 0.0, 1.0)
    shp = (nz, ny, nx)
    kz0 = ndim == 2
    v = [_spectral_field(shp, 8.0, -2.0, rng, kz_zero=kz0) for _ in range(3)]
    vr = math.sqrt(float(np.mean(v[0] ** 2 + v[1] ** 2 + v[2] ** 2)))
    v = [c * (v_rms / vr) for c in v]
    a = [_spectral_field(shp, 8.0, -2.0, rng, kz_zero=kz0) for _ in range(3)]
    dx, dy, dz = 1.0 / nx, 1.0 / ny, 1.0 / nz
    bx, by, bz = _curl_faces(a[0], a[1], a[2], dx, dy, dz)
    br = math.sqrt(float(np.mean(bx ** 2 + by ** 2 + bz ** 2)))
    s = db_rms / br
    a = [c * s for c in a]
    bx, by, bz = _curl_faces(a[0], a[1], a[2], dx, dy, dz)
    bz = bz + 1.0
    del a
    prim, eps = _alloc(cfg)
    p0 = 0.5
    inner = (slice(ix(1), ix(nz) + 1), slice(ix(1), ix(ny) + 1), slice(ix(1), ix(nx) + 1))
    prim[inner + (IRHO,)] = 1.0
    prim[inner + (IVX,)] = v[0]
    prim[inner + (IVY,)] = v[1]
    prim[inner + (IVZ,)] = v[2]
    prim[inner + (ITAU,)] = p0
    prim[inner + (IBX,)] = bx
    prim[inner + (IBY,)] = by
    prim[inner + (IBZ,)] = bz
    eps[inner] = p0 / 1.0 / (gamma - 1.0)
    del v, bx, by, bz
    _fill_periodic_ghosts(prim, shp)
    _fill_periodic_ghosts(eps, shp)
    return Problem(cfg, xf, yf, zf, prim, eps, steps=steps, name=f"spectral{ndim}d_{n}")


def torus_cyn(nr: int = 256, nphi: int = 128, nzz: int = 256, seed: int = 1003, gamma: float = 5.0 / 3.0,
              cfl: float = 0.4, beta: float = 100.0, rho_cut: float = 0.2, dp_frac: float = 0.01,
              r0: float = 1.0, r1: float = 40.0, zext: float = 20.0, steps: int = 50) -> Problem:
    """Constant-angular-momentum hydrostatic torus in a Paczynski-Wiita potential on a
    cylindrical (R, phi, z) grid with log-spaced R faces xF(j) = r0 (r1/r0)^(j/nr).

    Torus constants follow the torus_cyn plugin (model/torus_cyn/src/defs.f90:14-46: q = 2,
    r_sh = 1, s_in = 3, s_max = 4.7, rho_max = 1, floor rho = 1e-4 rho_max) and its density
    solution (model/torus_cyn/src/model.f90:71-121); eps_floor follows model.f90:246-247.
    A_phi = max(rho - rho_cut*rho_max, 0) at cell centres is averaged to (R,z) corners and
    curled onto faces (model/torus_cyn/src/model.f90:135-158); B is then scaled so that
    max(p)/max(B_cell^2/2) = beta.  A seeded U[-1,1] relative pressure perturbation of size
    dp_frac is applied where rho exceeds the floor.  Boundary flags (1,1,0,0,1,1).
    """
    rng = np.random.default_rng(seed)
    q, r_sh, s_in, s_max, rho_max, rho_fac = 2.0, 1.0, 3.0, 4.7, 1.0, 1.0e-4
    xf = log_faces(nr, r0, r1)
    yf = periodic_angle_faces(nphi)
    zf = uniform_faces(nzz, -zext, zext)

    def phi_pw(r, z):
        return -1.0 / (np.sqrt(r * r + z * z) - r_sh)

    lkep2 = s_max ** (2.0 * q - 1.0) / (s_max - r_sh) ** 2
    c_const = phi_pw(s_in, 0.0) - lkep2 * s_in ** (2.0 - 2.0 * q) / (2.0 - 2.0 * q)
    h_max = -phi_pw(s_max, 0.0) + lkep2 * s_max ** (2.0 - 2.0 * q) / (2.0 - 2.0 * q) + c_const
    k_poly = h_max * (gamma - 1.0) / gamma / rho_max ** (gamma - 1.0)
    rho_a = rho_fac * rho_max
    rho_floor = rho_a
    eps_floor = k_poly * rho_floor ** (gamma - 1.0) / (gamma - 1.0)

    cfg = Config(nx=nr, ny=nphi, nz=nzz, coordinate=CYLINDRICAL, boundary_flag=(1, 1, 0, 0, 1, 1), recon=PPMC,
                 model=MODEL_TORUS_CYN, ggas=gamma, cfl=cfl, rho_floor=rho_floor, eps_floor=eps_floor, r_sh=r_sh)
    prim, eps = _alloc(cfg)
    xc = 0.5 * (xf[1:] + xf[:-1])  # cell j at index ix(j)
    zc = 0.5 * (zf[1:] + zf[:-1])
    dxa = np.diff(xf)
    dza = np.diff(zf)
    R, Z = np.meshgrid(xc, zc, indexing="xy")  # [l, j] over all cells
    h = -phi_pw(R, Z) + lkep2 * R ** (2.0 - 2.0 * q) / (2.0 - 2.0 * q) + c_const
    h = np.maximum(h, 0.0)
    rho_loc = (h * (gamma - 1.0) / gamma / k_poly) ** (1.0 / (gamma - 1.0))
    inside = (rho_loc > rho_a) & (R >= s_in)
    rho2 = np.where(inside, rho_loc, rho_a)
    vphi2 = np.where(inside, np.sqrt(lkep2) * R ** (-q) * R, 0.0)
    p2 = k_poly * rho2 ** gamma

    # magnetic field: corners A(j,l) at (xF(j), zF(l)) from the four surrounding cell centres
    aphi = np.maximum(rho2 - rho_cut * rho_max, 0.0) / rho_max
    ac = 0.25 * (aphi[:-1, :-1] + aphi[:-1, 1:] + aphi[1:, :-1] + aphi[1:, 1:])  # ac[ix(l), ix(j)]
    J = np.arange(0, nr + 1)
    L = np.arange(0, nzz + 1)
    LL, JJ = np.meshgrid(L, J, indexing="ij")
    A = lambda j, l: ac[ix(l), ix(j)]  # noqa: E731
    xF = lambda j: xf[j + G]  # noqa: E731
    bx_face = np.zeros((nzz + 2 * G, nr + 2 * G))
    bz_face = np.zeros((nzz + 2 * G, nr + 2 * G))
    bx_face[ix(LL), ix(JJ)] = -(A(JJ, LL) - A(JJ, LL - 1)) / dza[ix(LL)]
    bz_face[ix(LL), ix(JJ)] = (xF(JJ) * A(JJ, LL) - xF(JJ - 1) * A(JJ - 1, LL)) / (xc[ix(JJ)] * dxa[ix(JJ)])
    il = slice(ix(1), ix(nzz) + 1)
    ij = slice(ix(1), ix(nr) + 1)
    ijm = slice(ix(0), ix(nr - 1) + 1)
    ilm = slice(ix(0), ix(nzz - 1) + 1)
    bcx = 0.5 * (bx_face[il, ij] + bx_face[il, ijm])
    bcz = 0.5 * (bz_face[il, ij] + bz_face[ilm, ij])
    pmag_max = float(np.max(0.5 * (bcx ** 2 + bcz ** 2)))
    pgas_max = float(np.max(p2[il, ij]))
    norm = math.sqrt(pgas_max / pmag_max / beta) if pmag_max > 0 else 0.0

    pert = rng.uniform(-1.0, 1.0, size=(nzz, nphi, nr))
    rho3 = np.broadcast_to(rho2[il, None, ij], (nzz, nphi, nr))
    p3 = p2[il, None, ij] * (1.0 + dp_frac * pert * (rho3 > rho_floor))
    inner = (il, slice(ix(1), ix(nphi) + 1), ij)
    prim[inner + (IRHO,)] = rho3
    prim[inner + (IVY,)] = np.broadcast_to(vphi2[il, None, ij], (nzz, nphi, nr))
    prim[inner + (ITAU,)] = p3
    eps[inner] = p3 / rho3 / (gamma - 1.0)
    prim[:, :, :, IBX] = (bx_face * norm)[:, None, :]
    prim[:, :, :, IBZ] = (bz_face * norm)[:, None, :]
    return Problem(cfg, xf, yf, zf, prim, eps, steps=steps, name="torus_cyn")


def spherical_areas(xf, yf, zf):
    """Face areas of coords.f90:213-219 (GEOM_AREA, spherical) on every face of the
    ghost-inclusive grid, as arrays [l, k, j] (entry j = face xF(j) for ax, the upper face of
    cell j for ay / az; the first entry along an axis without a lower face is left at 1)."""
    def d(a):
        out = np.ones(len(a) - 1)
        out[1:] = a[1:-1] - a[:-2]
        return out
    xF, xFm = xf[1:], xf[:-1]       # cell j: faces j (upper) and j-1
    yF, yFm = yf[1:], yf[:-1]
    dz = zf[1:] - zf[:-1]
    dy = yf[1:] - yf[:-1]
    x2d = xF * xF - xFm * xFm
    cd = np.cos(yFm) - np.cos(yF)
    ax = (xF * xF)[None, None, :] * cd[None, :, None] * dz[:, None, None]
    ay = (0.5 * x2d)[None, None, :] * np.sin(yF)[None, :, None] * dz[:, None, None]
    az = (0.5 * x2d)[None, None, :] * dy[None, :, None] * np.ones(len(dz))[:, None, None]
    return ax, ay, az


def spherical_shell(nr: int = 16, nth: int = 12, nph: int = 8, seed: int = 1006, r=(1.0, 2.0),
                    theta=(0.25 * math.pi, 0.75 * math.pi), phi=(0.0, 0.5 * math.pi),
                    boundary_flag=(1, 1, 1, 1, 0, 0), recon: int = PPMC, gamma: float = 5.0 / 3.0,
                    cfl: float = 0.4, noise: float = 0.1, v_amp: float = 0.1, b_amp: float = 0.3,
                    static: bool = False, smooth: bool = False, steps: int = 4) -> Problem:
    """Spherical (r, theta, phi) shell with uniform faces on r x theta x phi, riemann hooks.

    rho = 1, p = 0.5 with seeded relative noise, seeded velocities of size v_amp, and face B
    = (discrete curl of seeded edge values) / (face area of coords.f90:213-219), scaled to rms
    b_amp, so the area-weighted divergence of find_divb (mhd.f90:15-66) is zero to round-off.
    static: rho = 1, p = 0.5, v = 0, B = 0 exactly (the geometric-source balance state).
    smooth: the noise, velocities and edge values are smooth random fields (modes 1 <= |k| <= 4
    of the index space, periodic) instead of cell-to-cell noise, for long runs."""
    rng = np.random.default_rng(seed)
    xf = uniform_faces(nr, *r)
    yf = uniform_faces(nth, *theta)
    zf = uniform_faces(nph, *phi)
    cfg = Config(nx=nr, ny=nth, nz=nph, coordinate=SPHERICAL, boundary_flag=tuple(boundary_flag), recon=recon,
                 model=MODEL_RIEMANN, ggas=gamma, cfl=cfl)
    prim, eps = _alloc(cfg)
    inner = (slice(ix(1), ix(nph) + 1), slice(ix(1), ix(nth) + 1), slice(ix(1), ix(nr) + 1))
    shp = (nph, nth, nr)
    rho = np.ones(shp)
    p = 0.5 * np.ones(shp)
    if not static:
        def field(shape):
            if smooth:
                return _spectral_field(shape, 4.0, -2.0, rng)
            return rng.uniform(-1, 1, shape)
        rho *= 1.0 + noise * field(shp)
        p *= 1.0 + noise * field(shp)
        for v in (IVX, IVY, IVZ):
            prim[inner + (v,)] = v_amp * field(shp)
        full = cfg.shape
        psi = [field(full) if smooth else rng.standard_normal(full) for _ in range(3)]  # edge values along x, y, z
        if boundary_flag[4] == 0 and boundary_flag[5] == 0:
            # periodic in phi: the same edge values one period apart, so the face fields of the
            # two periodic boundary faces agree
            wrap = (np.arange(full[0]) - G) % nph + G
            psi = [a[wrap] for a in psi]

        def dm(a, axis):
            out = np.zeros_like(a)
            sl, sm = [slice(None)] * 3, [slice(None)] * 3
            sl[axis], sm[axis] = slice(1, None), slice(None, -1)
            out[tuple(sl)] = a[tuple(sl)] - a[tuple(sm)]
            return out
        # axes of the [l, k, j] arrays: 0 = z, 1 = y, 2 = x
        fx = dm(psi[2], 1) - dm(psi[1], 0)
        fy = dm(psi[0], 0) - dm(psi[2], 2)
        fz = dm(psi[1], 2) - dm(psi[0], 1)
        ax, ay, az = spherical_areas(xf, yf, zf)
        b = [fx / ax, fy / ay, fz / az]
        rms = math.sqrt(float(np.mean(b[0] ** 2 + b[1] ** 2 + b[2] ** 2)))
        for c, arr in enumerate(b):
            prim[:, :, :, IBX + c] = arr * (b_amp / rms)
    prim[inner + (IRHO,)] = rho
    prim[inner + (ITAU,)] = p
    eps[inner] = p / rho / (gamma - 1.0)
    return Problem(cfg, xf, yf, zf, prim, eps, steps=steps, name="spherical_shell")


def torus_sph(nr: int = 256, nth: int = 128, nph: int = 128, seed: int = 1007, gamma: float = 5.0 / 3.0,
              cfl: float = 0.4, beta: float = 100.0, rho_cut: float = 0.2, dp_frac: float = 0.01,
              r0: float = 10.0, r1: float = 110.0, theta0: float = 0.2, steps: int = 50) -> Problem:
    """Constant-angular-momentum torus in a Paczynski-Wiita potential on a spherical (r, theta,
    phi) grid with log-spaced r faces, theta in [theta0, pi - theta0], phi periodic.

    Constants of the torus_sph plugin (model/torus_sph/src/defs.f90: q = 2, r_sh = 2, s_in = 21,
    s_max = 40, rho_max = 1, floor rho = 1e-4 rho_max) and its density solution with the
    equatorial radius s = r sin(theta) (model/torus_sph/src/model.f90:71-131); v_phi = Omega s
    inside the torus.  The seeded field follows the plugin's vector potential
    A_phi = (rho/rho_max) (r sin(theta)/3)^3 exp(-r/400) - rho_cut where rho >= rho_cut, averaged
    to the (r, theta) corners and curled onto the faces with the spherical factors, then scaled
    to max(p)/max(B^2/2) = beta.  A seeded U[-1,1] relative pressure perturbation of size
    dp_frac is applied where rho exceeds the floor.  Boundary flags (1,1,1,1,0,0)."""
    rng = np.random.default_rng(seed)
    q, r_sh, s_in, s_max, rho_max, rho_fac = 2.0, 2.0, 21.0, 40.0, 1.0, 1.0e-4
    xf = log_faces(nr, r0, r1)
    yf = uniform_faces(nth, theta0, math.pi - theta0)
    zf = periodic_angle_faces(nph)

    def phi_pw(r):
        return -1.0 / (r - r_sh)

    lkep2 = s_max ** (2.0 * q - 1.0) / (s_max - r_sh) ** 2
    c_const = phi_pw(s_in) - lkep2 * s_in ** (2.0 - 2.0 * q) / (2.0 - 2.0 * q)
    h_max = -phi_pw(s_max) + lkep2 * s_max ** (2.0 - 2.0 * q) / (2.0 - 2.0 * q) + c_const
    k_poly = h_max * (gamma - 1.0) / gamma / rho_max ** (gamma - 1.0)
    rho_a = rho_fac * rho_max
    eps_floor = k_poly * rho_a ** (gamma - 1.0) / (gamma - 1.0)
    cfg = Config(nx=nr, ny=nth, nz=nph, coordinate=SPHERICAL, boundary_flag=(1, 1, 1, 1, 0, 0), recon=PPMC,
                 model=MODEL_TORUS_SPH, ggas=gamma, cfl=cfl, rho_floor=rho_a, eps_floor=eps_floor, r_sh=r_sh)
    prim, eps = _alloc(cfg)
    xc = 0.5 * (xf[1:] + xf[:-1])
    yc = 0.5 * (yf[1:] + yf[:-1])
    TH, R = np.meshgrid(yc, xc, indexing="ij")  # [k, j] over all cells
    s_eq = R * np.sin(TH)
    h = -phi_pw(R) + lkep2 * s_eq ** (2.0 - 2.0 * q) / (2.0 - 2.0 * q) + c_const
    h = np.maximum(h, 0.0)
    rho_loc = (h * (gamma - 1.0) / gamma / k_poly) ** (1.0 / (gamma - 1.0))
    inside = (rho_loc > rho_a) & (s_eq >= s_in)
    rho2 = np.where(inside, rho_loc, rho_a)
    vphi2 = np.where(inside, np.sqrt(lkep2) * s_eq ** (-q) * s_eq, 0.0)
    p2 = k_poly * rho2 ** gamma
    a_c = np.where(rho2 >= rho_cut, rho2 / rho_max * (R / 3.0 * np.sin(TH)) ** 3 * np.exp(-R / 400.0) - rho_cut,
                   0.0)
    # corners (xF(j), yF(k)) from the four surrounding cells; face fields by the discrete curl
    ac = np.zeros_like(a_c)
    ac[:-1, :-1] = 0.25 * (a_c[:-1, :-1] + a_c[:-1, 1:] + a_c[1:, :-1] + a_c[1:, 1:])  # ac[ix(k), ix(j)]
    xF, xFm = xf[1:], xf[:-1]
    yF, yFm = yf[1:], yf[:-1]
    ac_km = np.zeros_like(ac)
    ac_km[1:, :] = ac[:-1, :]
    ac_jm = np.zeros_like(ac)
    ac_jm[:, 1:] = ac[:, :-1]
    bx_face = (np.sin(yF)[:, None] * ac - np.sin(yFm)[:, None] * ac_km) / (xF[None, :] * (np.cos(yFm) - np.cos(yF))[:, None])
    by_face = -(xF[None, :] * ac - xFm[None, :] * ac_jm) / (xc * (xF - xFm))[None, :]
    ik = slice(ix(1), ix(nth) + 1)
    ij = slice(ix(1), ix(nr) + 1)
    bcx = 0.5 * (bx_face[ik, ij] + bx_face[ik, ix(0):ix(nr - 1) + 1])
    bcy = 0.5 * (by_face[ik, ij] + by_face[ix(0):ix(nth - 1) + 1, ij])
    pmag_max = float(np.max(0.5 * (bcx ** 2 + bcy ** 2)))
    pgas_max = float(np.max(p2[ik, ij]))
    norm = math.sqrt(pgas_max / pmag_max / beta) if pmag_max > 0 else 0.0
    pert = rng.uniform(-1.0, 1.0, size=(nph, nth, nr))
    rho3 = np.broadcast_to(rho2[None, ik, ij], (nph, nth, nr))
    p3 = p2[None, ik, ij] * (1.0 + dp_frac * pert * (rho3 > rho_a))
    inner = (slice(ix(1), ix(nph) + 1), ik, ij)
    prim[inner + (IRHO,)] = rho3
    prim[inner + (IVZ,)] = np.broadcast_to(vphi2[None, ik, ij], (nph, nth, nr))
    prim[inner + (ITAU,)] = p3
    eps[inner] = p3 / rho3 / (gamma - 1.0)
    prim[:, :, :, IBX] = (bx_face * norm)[None, :, :]
    prim[:, :, :, IBY] = (by_face * norm)[None, :, :]
    return Problem(cfg, xf, yf, zf, prim, eps, steps=steps, name="torus_sph")


def small_periodic(n: int = 16, seed: int = 1005, ndim: int = 3) -> Problem:
    """Small smooth random periodic state for parity tests (same construction as spectral_box)."""
    return spectral_box(n=n, seed=seed, ndim=ndim, v_rms=0.3, db_rms=0.3, steps=4)


def randomize(prob: Problem, seed: int, amp: float = 0.2) -> Problem:
    """Add seeded cell-to-cell noise to rho, p and v on [0,n]^3 so that limiter and Riemann
    branches are exercised; eps follows from the EOS."""
    rng = np.random.default_rng(seed)
    p = prob.prim
    c = prob.cfg
    inner = (slice(ix(0), ix(c.nz) + 1), slice(ix(0), ix(c.ny) + 1), slice(ix(0), ix(c.nx) + 1))
    shp = p[inner + (0,)].shape
    p[inner + (IRHO,)] *= 1.0 + amp * rng.uniform(-1, 1, shp)
    p[inner + (ITAU,)] *= 1.0 + amp * rng.uniform(-1, 1, shp)
    for v in (IVX, IVY, IVZ):
        p[inner + (v,)] += amp * rng.uniform(-1, 1, shp)
    prob.eps[inner] = p[inner + (ITAU,)] / p[inner + (IRHO,)] / (c.ggas - 1.0)
    return prob


BATCH_SIZES = (64, 96, 128, 192, 256)


def batch_boxes(nbox: int = 48, seed: int = 1004, sizes=BATCH_SIZES, steps: int = 40) -> list[Problem]:
    """Independent 2D periodic boxes of uneven sizes for the batched form.

    Box sizes n x n (nz = 1) are drawn from ``sizes`` by ``default_rng(seed)``; box i carries the
    spectral_box initial condition restricted to k_z = 0 with its own seed ``100 * seed + i``.
    """
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.asarray(sizes), size=nbox)
    out = []
    for i, n in enumerate(ns):
        p = spectral_box(n=int(n), seed=100 * seed + i, ndim=2, steps=steps)
        p.name = f"batch{i}_{int(n)}sq"
        out.append(p)
    return out
