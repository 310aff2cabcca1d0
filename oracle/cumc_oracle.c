/*
 * cumc_oracle.c -- CPU restatement of the CUMC3D ideal-MHD time step
 * (PPM/TVD + HLL + upwind CT + SSP-RK3), written from the reference's Fortran.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library and the
 * timed CPU baseline.  The product never links or calls this file.
 *
 * Every routine below names the reference routine (file:line) it restates.
 * Operation order inside each expression follows the Fortran source left to
 * right; the file is compiled with -ffp-contract=off so no multiply-add is
 * fused.  Arrays keep the reference's ghost-inclusive shapes and loop ranges,
 * including its quirks (see the notes marked Qn).
 */
#include "cumc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { IRHO = 0, IVX = 1, IVY = 2, IVZ = 3, ITAU = 4, IBX = 5, IBY = 6, IBZ = 7, NV = 8 };
enum { IEX = 0, IEY = 1, IEZ = 2 };
#define NGHOST 3

struct orc {
    orc_config c;
    int nx, ny, nz;
    long sx, sy, sz;        /* n + 6 per axis */
    long ncell;
    double *xf_mem, *yf_mem, *zf_mem;
    double *xf, *yf, *zf;   /* xf[j] valid for j in [-3, n+3] */
    double *prim, *cons, *primL, *primR, *flux, *l_rk, *u_old, *sc;
    double *eps;
    double *bcell, *ecell, *ecorn;
    int bfac_xin[NV], bfac_xout[NV], bfac_yin[NV], bfac_yout[NV], bfac_zin[NV], bfac_zout[NV];
    /* cells CUSTOM_CHECKRHO changed since creation (test instrumentation, no reference
       counterpart): density raised to the floor, eps raised to its floor (or to 0) */
    long nfl_rho, nfl_eps;
};

#define ID(j, k, l) ((((long)(l) + 2) * o->sy + ((long)(k) + 2)) * o->sx + ((long)(j) + 2))
#define PR(v, j, k, l) o->prim[ID(j, k, l) * NV + (v)]
#define CO(v, j, k, l) o->cons[ID(j, k, l) * NV + (v)]
#define PL(v, j, k, l) o->primL[ID(j, k, l) * NV + (v)]
#define PRR(v, j, k, l) o->primR[ID(j, k, l) * NV + (v)]
#define FL(v, j, k, l) o->flux[ID(j, k, l) * NV + (v)]
#define LRK(v, j, k, l) o->l_rk[ID(j, k, l) * NV + (v)]
#define UO(v, j, k, l) o->u_old[ID(j, k, l) * NV + (v)]
#define SC(v, j, k, l) o->sc[ID(j, k, l) * NV + (v)]
#define EPS(j, k, l) o->eps[ID(j, k, l)]
#define BC(c, j, k, l) o->bcell[ID(j, k, l) * 3 + (c)]
#define EC(c, j, k, l) o->ecell[ID(j, k, l) * 3 + (c)]
#define ER(c, j, k, l) o->ecorn[ID(j, k, l) * 3 + (c)]

/* Fortran MIN/MAX on non-NaN arguments; the same definition is used on the GPU. */
static inline double fmin2(double a, double b) { return (b < a) ? b : a; }
static inline double fmax2(double a, double b) { return (b > a) ? b : a; }
static inline double fmin3(double a, double b, double c) { return fmin2(fmin2(a, b), c); }
static inline double fmax3(double a, double b, double c) { return fmax2(fmax2(a, b), c); }
/* Fortran SIGN(1.0d0, x) */
static inline double fsign1(double x) { return copysign(1.0, x); }

/* definition.f90:362-367 */
static inline double dotp(double a1, double a2, double a3, double b1, double b2, double b3)
{
    return a1 * b1 + a2 * b2 + a3 * b3;
}

/* definition.f90:380-385 */
static inline double compute_fluxhll(double yl, double yr, double xl, double xr, double sl, double sr)
{
    return (sr * yl - sl * yr + sl * sr * (xr - xl)) / (sr - sl);
}

/* definition.f90:387-392 */
static inline double compute_roe(double xl, double xr, double rhol, double rhor)
{
    return (sqrt(rhol) * xl + sqrt(rhor) * xr) / (sqrt(rhol) + sqrt(rhor));
}

/* definition.f90:394-410 (Q8: discriminant clamped at zero here, not in finddt) */
double orc_signalspeed(double cs, double bn, double bt1, double bt2, double rho)
{
    double a2 = cs * cs;
    double a4 = a2 * a2;
    double b2n = (bn * bn / rho);
    double b2t1 = (bt1 * bt1 / rho);
    double b2t2 = (bt2 * bt2 / rho);
    double b2 = b2n + b2t1 + b2t2;
    double b4 = b2 * b2;
    return sqrt(0.5 * (a2 + b2 + sqrt(fmax2((a4 + b4 + 2.0 * b2 * a2) - 4.0 * a2 * b2n, 0.0))));
}

/* model/{riemann,torus_cyn}/src/eos.f90:7-81 (identical in both plugins) */
static inline double eos_pressure(double g, double rho, double eps) { return rho * eps * (g - 1.0); }
static inline double eos_soundspeed(double g, double p, double rho) { return sqrt(g * p / rho); }
static inline double eos_epsilon(double g, double p, double rho) { return p / rho / (g - 1.0); }

/* ---------------------------------------------------------------------------------------------- */
/* Reconstruction                                                                                 */

/* reconstruct.f90:137-311, PPM (Colella & Woodward 1984) on a non-uniform grid.
 * d = (dm2, dm1, dc, dp1, dp2) cell widths, v = (vm2, vm1, vc, vp1, vp2). */
void orc_ppm(const double d[5], const double v[5], double *vm_out, double *vp_out)
{
    const double dm2 = d[0], dm1 = d[1], dc = d[2], dp1 = d[3], dp2 = d[4];
    const double vm2 = v[0], vm1 = v[1], vc = v[2], vp1 = v[3], vp2 = v[4];
    /* :211-231 */
    double a1R = dc / (dc + dp1);
    double a1L = dm1 / (dm1 + dc);
    double aR = 2.0 * dp1 * dc / (dp1 + dc);
    double aL = 2.0 * dc * dm1 / (dc + dm1);
    double deltaXR = dm1 + dc + dp1 + dp2;
    double deltaXL = dm2 + dm1 + dc + dp1;
    double z1R = (dm1 + dc) / (2.0 * dc + dp1);
    double z1L = (dm2 + dm1) / (2.0 * dm1 + dc);
    double zR = (dp2 + dp1) / (2.0 * dp1 + dc);
    double zL = (dp1 + dc) / (2.0 * dc + dm1);
    double cc1 = dc / (dm1 + dc + dp1);
    double cp1 = dp1 / (dc + dp1 + dp2);
    double cm1 = dm1 / (dm2 + dm1 + dc);
    double cc2 = (2.0 * dm1 + dc) / (dp1 + dc);
    double cp2 = (2.0 * dc + dp1) / (dp2 + dp1);
    double cm2 = (2.0 * dm2 + dm1) / (dc + dm1);
    double cc3 = (dc + 2.0 * dp1) / (dm1 + dc);
    double cp3 = (dp1 + 2.0 * dp2) / (dc + dp1);
    double cm3 = (dm1 + 2.0 * dc) / (dm2 + dm1);
    /* :236-249 */
    double a0m = a1L;
    double a1m = aL * (z1L - zL) / deltaXL;
    double a2m = (-dm1 * z1L) / deltaXL;
    double a3m = dc * zL / deltaXL;
    double a0p = a1R;
    double a1p = aR * (z1R - zR) / deltaXR;
    double a2p = (-dc * z1R) / deltaXR;
    double a3p = dp1 * zR / deltaXR;
    double b0m = cm1 * cm2, b1m = cm1 * cm3;
    double b0c = cc1 * cc2, b1c = cc1 * cc3;
    double b0p = cp1 * cp2, b1p = cp1 * cp3;
    /* :254-256 */
    double deltap1 = b0p * (vp2 - vp1) + b1p * (vp1 - vc);
    double deltac = b0c * (vp1 - vc) + b1c * (vc - vm1);
    double deltam1 = b0m * (vc - vm1) + b1m * (vm1 - vm2);
    /* van Leer limiter :261-280 */
    double dmc, dmp1, dmm1;
    if ((vp1 - vc) * (vc - vm1) > 0.0)
        dmc = fmin3(fabs(deltac), 2.0 * fabs(vc - vm1), 2.0 * fabs(vc - vp1)) * fsign1(deltac);
    else
        dmc = 0.0;
    if ((vp2 - vp1) * (vp1 - vc) > 0.0)
        dmp1 = fmin3(fabs(deltap1), 2.0 * fabs(vp1 - vc), 2.0 * fabs(vp1 - vp2)) * fsign1(deltap1);
    else
        dmp1 = 0.0;
    if ((vc - vm1) * (vm1 - vm2) > 0.0)
        dmm1 = fmin3(fabs(deltam1), 2.0 * fabs(vm1 - vm2), 2.0 * fabs(vm1 - vc)) * fsign1(deltam1);
    else
        dmm1 = 0.0;
    /* :285-286 */
    double vp = vc + a0p * (vp1 - vc) + a1p * (vp1 - vc) + a2p * dmp1 + a3p * dmc;
    double vm = vm1 + a0m * (vc - vm1) + a1m * (vc - vm1) + a2m * dmc + a3m * dmm1;
    double vl = vm, vr = vp;
    /* :295-307 (Q10: second test uses the possibly-modified vm with the original vl, vr) */
    if ((vr - vc) * (vc - vl) <= 0.0) {
        vp = vc;
        vm = vc;
    } else {
        if ((vr - vl) * (vl - 3.0 * vc + 2.0 * vr) < 0.0) vm = 3.0 * vc - 2.0 * vr;
        if ((vp - vm) * (3.0 * vc - 2.0 * vl - vr) < 0.0) vp = 3.0 * vc - 2.0 * vl;
    }
    *vm_out = vm;
    *vp_out = vp;
}

/* reconstruct.f90:58-128, TVD minmod on a non-uniform grid.  d = (dh_m1, dh_c, dh_p1). */
void orc_tvdmm(const double d[3], const double v[3], double *vm_out, double *vp_out)
{
    const double dh_m1 = d[0], dh_c = d[1], dh_p1 = d[2];
    const double vm1 = v[0], vc = v[1], vp1 = v[2];
    double cf = (dh_c + dh_p1) / dh_c;
    double cb = (dh_c + dh_m1) / dh_c;
    double dqf = 2.0 * (vp1 - vc) / cf;
    double dqb = 2.0 * (vc - vm1) / cb;
    double slope = 0.5 * (fsign1(dqf) + fsign1(dqb)) * fmin2(fabs(dqf), fabs(dqb));
    *vp_out = vc + 0.5 * slope;
    *vm_out = vc - 0.5 * slope;
}

/* TVD with the MC (lim 2) or van Leer (lim 3) limiter: the reconstruction options that
 * definition.f90:160-164 declares (TVDMC, TVDVL) and reconstruct.f90 leaves unimplemented.  An
 * extension, not a restatement: the limiter acts on the same width-normalised one-sided
 * differences dqf, dqb as TVD_MM (reconstruct.f90:115-124), with the grid factors cf, cb in
 * place of MC's 2 (so that the face values stay between the neighbouring averages on any grid):
 *   MC: minmod(cf dq+, cb dq-, (dq+ + dq-)/2)      (van Leer 1977, monotonised central)
 *   VL: 2 dq+ dq- / (dq+ + dq-) when dq+ dq- > 0, else 0   (van Leer 1974, harmonic), its size
 *       capped at min(cf |dq+|, cb |dq-|) for the same bound
 * On a uniform grid cf = cb = 2, the cap never binds, and these are the textbook limiters; lim 1
 * is orc_tvdmm. */
void orc_tvd(int lim, const double d[3], const double v[3], double *vm_out, double *vp_out)
{
    const double dh_m1 = d[0], dh_c = d[1], dh_p1 = d[2];
    const double vm1 = v[0], vc = v[1], vp1 = v[2];
    double cf = (dh_c + dh_p1) / dh_c;
    double cb = (dh_c + dh_m1) / dh_c;
    double dqf = 2.0 * (vp1 - vc) / cf;
    double dqb = 2.0 * (vc - vm1) / cb;
    double slope;
    if (lim == 2) {
        double m = fmin2(fmin2(cf * fabs(dqf), cb * fabs(dqb)), 0.5 * fabs(dqf + dqb));
        slope = 0.5 * (fsign1(dqf) + fsign1(dqb)) * m;
    } else if (lim == 3) {
        if (dqf * dqb > 0.0) {
            double h = 2.0 * dqf * dqb / (dqf + dqb);
            slope = fsign1(h) * fmin2(fmin2(fabs(h), cf * fabs(dqf)), cb * fabs(dqb));
        } else {
            slope = 0.0;
        }
    } else {
        slope = 0.5 * (fsign1(dqf) + fsign1(dqb)) * fmin2(fabs(dqf), fabs(dqb));
    }
    *vp_out = vc + 0.5 * slope;
    *vm_out = vc - 0.5 * slope;
}

/* WENO5 face value at the right face of cell c from v = (c-2 .. c+2): Jiang & Shu (1996), the
 * "WENO (Shu 1997)" option definition.f90:169-170 declares and reconstruct.f90 leaves
 * unimplemented (an extension).  Candidate quadratics p0..p2, smoothness indicators b0..b2,
 * linear weights (1, 6, 3)/10 and eps = 1e-6, with the uniform-grid (index-space) coefficients
 * on every grid.  (The sweep falls back to the piecewise-constant profile for a cell whose WENO
 * face density or pressure is not positive: WENO is not positivity-preserving.) */
static double weno5_face(double vm2, double vm1, double vc, double vp1, double vp2)
{
    /* the candidate values times 6; the 1/6 goes into the final normalisation */
    const double p0 = 2.0 * vm2 - 7.0 * vm1 + 11.0 * vc;
    const double p1 = -vm1 + 5.0 * vc + 2.0 * vp1;
    const double p2 = 2.0 * vc + 5.0 * vp1 - vp2;
    const double c13 = 13.0 / 12.0;
    const double t0 = vm2 - 2.0 * vm1 + vc, s0 = vm2 - 4.0 * vm1 + 3.0 * vc;
    const double t1 = vm1 - 2.0 * vc + vp1, s1 = vm1 - vp1;
    const double t2 = vc - 2.0 * vp1 + vp2, s2 = 3.0 * vc - 4.0 * vp1 + vp2;
    const double b0 = c13 * t0 * t0 + 0.25 * s0 * s0;
    const double b1 = c13 * t1 * t1 + 0.25 * s1 * s1;
    const double b2 = c13 * t2 * t2 + 0.25 * s2 * s2;
    const double e0 = 1.0e-6 + b0, e1 = 1.0e-6 + b1, e2 = 1.0e-6 + b2;
    const double a0 = 0.1 / (e0 * e0), a1 = 0.6 / (e1 * e1), a2 = 0.3 / (e2 * e2);
    return (a0 * p0 + a1 * p1 + a2 * p2) / (6.0 * (a0 + a1 + a2));
}

/* WENO5 edge values of the middle cell of v[5]: vp at its right face, vm at its left face (the
 * same formula on the mirrored stencil) */
void orc_weno5(const double v[5], double *vm_out, double *vp_out)
{
    *vp_out = weno5_face(v[0], v[1], v[2], v[3], v[4]);
    *vm_out = weno5_face(v[4], v[3], v[2], v[1], v[0]);
}

/* coords.f90:77-97 */
static inline double cdx(const orc *o, int j) { return o->xf[j] - o->xf[j - 1]; }
static inline double cdy(const orc *o, int k) { return o->yf[k] - o->yf[k - 1]; }
static inline double cdz(const orc *o, int l) { return o->zf[l] - o->zf[l - 1]; }
/* coords.f90:49-69 */
static inline double cxc(const orc *o, int j) { return 0.5 * (o->xf[j] + o->xf[j - 1]); }
static inline double czc(const orc *o, int l) { return 0.5 * (o->zf[l] + o->zf[l - 1]); }
static inline double cyc(const orc *o, int k) { return 0.5 * (o->yf[k] + o->yf[k - 1]); }
/* Fortran x**2 and x**3 with integer exponents: x*x and (x*x)*x (coords.f90 spherical branches) */
static inline double pw2(double x) { return x * x; }
static inline double pw3(double x) { return x * x * x; }
/* spherical radial factor 2 (r_j^3 - r_{j-1}^3) / 3 / (r_j^2 - r_{j-1}^2) (coords.f90:290, 295) */
static inline double sph_a(const orc *o, int j)
{
    return 2.0 * (pw3(o->xf[j]) - pw3(o->xf[j - 1])) / 3.0 / (pw2(o->xf[j]) - pw2(o->xf[j - 1]));
}
/* cos(theta_{k-1}) - cos(theta_k) (coords.f90:215-216, 290, 389) */
static inline double sph_cd(const orc *o, int k) { return cos(o->yf[k - 1]) - cos(o->yf[k]); }

/* reconstruct.f90:18-51 INTERPOLATE: widths along `dir` around index `c` */
static void interpolate(const orc *o, int dir, int c, const double v[5], double *vm, double *vp)
{
    double (*w)(const orc *, int) = (dir == 1) ? cdx : (dir == 2) ? cdy : cdz;
    if (o->c.recon == 1) {
        double d[3] = {w(o, c - 1), w(o, c), w(o, c + 1)};
        orc_tvdmm(d, v + 1, vm, vp);
    } else if (o->c.recon == 2 || o->c.recon == 3) {
        double d[3] = {w(o, c - 1), w(o, c), w(o, c + 1)};
        orc_tvd(o->c.recon, d, v + 1, vm, vp);
    } else if (o->c.recon == 5) {
        orc_weno5(v, vm, vp);
    } else {
        double d[5] = {w(o, c - 2), w(o, c - 1), w(o, c), w(o, c + 1), w(o, c + 2)};
        orc_ppm(d, v, vm, vp);
    }
}

/* ---------------------------------------------------------------------------------------------- */
/* Point conversions                                                                              */

/* ptou_face.f90:99-153 (one side) */
void orc_ptou_face(const double *p, double eps, double *u)
{
    double vsquare = dotp(p[IVX], p[IVY], p[IVZ], p[IVX], p[IVY], p[IVZ]);
    double bsquare = dotp(p[IBX], p[IBY], p[IBZ], p[IBX], p[IBY], p[IBZ]);
    u[IRHO] = p[IRHO];
    u[IVX] = p[IVX] * p[IRHO];
    u[IVY] = p[IVY] * p[IRHO];
    u[IVZ] = p[IVZ] * p[IRHO];
    u[ITAU] = p[IRHO] * (eps + 0.5 * vsquare) + 0.5 * bsquare;
    u[IBX] = p[IBX];
    u[IBY] = p[IBY];
    u[IBZ] = p[IBZ];
}

/* ptou.f90:33-66 (cell: B^2 from cell-centred field) */
void orc_ptou_cell(const double *p, double eps, const double *bc, double *u)
{
    double vsquare = dotp(p[IVX], p[IVY], p[IVZ], p[IVX], p[IVY], p[IVZ]);
    double bsquare = dotp(bc[0], bc[1], bc[2], bc[0], bc[1], bc[2]);
    u[IRHO] = p[IRHO];
    u[IVX] = p[IVX] * p[IRHO];
    u[IVY] = p[IVY] * p[IRHO];
    u[IVZ] = p[IVZ] * p[IRHO];
    u[ITAU] = p[IRHO] * (eps + 0.5 * vsquare) + 0.5 * bsquare;
    u[IBX] = p[IBX];
    u[IBY] = p[IBY];
    u[IBZ] = p[IBZ];
}

/* utop.f90:185-220 (no negativity check, :215) ; p[ITAU] is left untouched */
void orc_utop_cell(const double *u, const double *bc, double *p, double *eps)
{
    p[IRHO] = u[IRHO];
    p[IVX] = u[IVX] / u[IRHO];
    p[IVY] = u[IVY] / u[IRHO];
    p[IVZ] = u[IVZ] / u[IRHO];
    p[IBX] = u[IBX];
    p[IBY] = u[IBY];
    p[IBZ] = u[IBZ];
    double bsquare = dotp(bc[0], bc[1], bc[2], bc[0], bc[1], bc[2]);
    double vsquare = dotp(p[IVX], p[IVY], p[IVZ], p[IVX], p[IVY], p[IVZ]);
    *eps = (u[ITAU] - 0.5 * bsquare) / u[IRHO] - 0.5 * vsquare;
}

static void dir_indices(int dir, int *ivn, int *ivt1, int *ivt2, int *ibn, int *ibt1, int *ibt2)
{
    /* flux.f90:182-206 ; rsolver.f90:73-91 */
    if (dir == 1) { *ivn = IVX; *ivt1 = IVY; *ivt2 = IVZ; *ibn = IBX; *ibt1 = IBY; *ibt2 = IBZ; }
    else if (dir == 2) { *ivn = IVY; *ivt1 = IVZ; *ivt2 = IVX; *ibn = IBY; *ibt1 = IBZ; *ibt2 = IBX; }
    else { *ivn = IVZ; *ivt1 = IVX; *ivt2 = IVY; *ibn = IBZ; *ibt1 = IBX; *ibt2 = IBY; }
}

/* flux.f90:163-262, one side */
void orc_ptof(int dir, const double *p, const double *u, double *f)
{
    int ivn, ivt1, ivt2, ibn, ibt1, ibt2;
    dir_indices(dir, &ivn, &ivt1, &ivt2, &ibn, &ibt1, &ibt2);
    double bsquare = dotp(p[IBX], p[IBY], p[IBZ], p[IBX], p[IBY], p[IBZ]);
    double vdotb = dotp(p[IBX], p[IBY], p[IBZ], p[IVX], p[IVY], p[IVZ]);
    for (int i = IRHO; i <= ITAU; i++) f[i] = u[i] * p[ivn];
    f[ivn] = f[ivn] + p[ITAU] + 0.5 * bsquare - p[ibn] * p[ibn];
    f[ivt1] = f[ivt1] - p[ibn] * p[ibt1];
    f[ivt2] = f[ivt2] - p[ibn] * p[ibt2];
    f[ITAU] = f[ITAU] + (p[ITAU] + 0.5 * bsquare) * p[ivn] - p[ibn] * vdotb;
    f[ibn] = 0.0;
    f[ibt1] = (p[ivn] * p[ibt1] - p[ivt1] * p[ibn]);
    f[ibt2] = (p[ivn] * p[ibt2] - p[ivt2] * p[ibn]);
}

/* rsolver.f90:43-130 HLL_SOLVER for one face; returns the branch taken */
int orc_hll(int dir, double g, const double *pl, const double *pr, double *flux)
{
    int ivn, ivt1, ivt2, ibn, ibt1, ibt2;
    dir_indices(dir, &ivn, &ivt1, &ivt2, &ibn, &ibt1, &ibt2);
    double epsL = eos_epsilon(g, pl[ITAU], pl[IRHO]);
    double epsR = eos_epsilon(g, pr[ITAU], pr[IRHO]);
    double csL = eos_soundspeed(g, pl[ITAU], pl[IRHO]);
    double csR = eos_soundspeed(g, pr[ITAU], pr[IRHO]);
    double uL[NV], uR[NV], fL[NV], fR[NV];
    orc_ptou_face(pl, epsL, uL);
    orc_ptou_face(pr, epsR, uR);
    orc_ptof(dir, pl, uL, fL);
    orc_ptof(dir, pr, uR, fR);
    double cfsL = orc_signalspeed(csL, pl[ibn], pl[ibt1], pl[ibt2], pl[IRHO]);
    double cfsR = orc_signalspeed(csR, pr[ibn], pr[ibt1], pr[ibt2], pr[IRHO]);
    double ubar = compute_roe(pl[ivn], pr[ivn], pl[IRHO], pr[IRHO]);
    double cbar = compute_roe(cfsL, cfsR, pl[IRHO], pr[IRHO]);
    double sL = fmin2(pl[ivn] - cfsL, ubar - cbar);
    double sR = fmax2(pr[ivn] + cfsR, ubar + cbar);
    if (sL >= 0.0) {
        for (int i = 0; i < NV; i++) flux[i] = fL[i];
        return 0;
    } else if (sL <= 0.0 && sR >= 0.0) {
        for (int i = 0; i < NV; i++) flux[i] = compute_fluxhll(fL[i], fR[i], uL[i], uR[i], sL, sR);
        return 1;
    } else if (sR <= 0.0) {
        for (int i = 0; i < NV; i++) flux[i] = fR[i];
        return 2;
    }
    /* NaN speeds: the reference leaves flux() untouched; this restatement writes NaN */
    for (int i = 0; i < NV; i++) flux[i] = NAN;
    return 3;
}

/* Rotated face state of one side (n = normal, t1/t2 = transverse as in flux.f90:182-206) with its
 * conserved vector U and physical flux F in the order rho, m_n, m_t1, m_t2, E, B_t1, B_t2. */
typedef struct {
    double rho, vn, vt1, vt2, p, bt1, bt2, pt, cf;
    double u[7], f[7];
} side7;

static void side_of(int dir, double g, const double *pw, side7 *s)
{
    int ivn, ivt1, ivt2, ibn, ibt1, ibt2;
    dir_indices(dir, &ivn, &ivt1, &ivt2, &ibn, &ibt1, &ibt2);
    double eps = eos_epsilon(g, pw[ITAU], pw[IRHO]);
    double cs = eos_soundspeed(g, pw[ITAU], pw[IRHO]);
    double u8[NV], f8[NV];
    orc_ptou_face(pw, eps, u8);
    orc_ptof(dir, pw, u8, f8);
    double bsquare = dotp(pw[IBX], pw[IBY], pw[IBZ], pw[IBX], pw[IBY], pw[IBZ]);
    s->rho = pw[IRHO];
    s->vn = pw[ivn];
    s->vt1 = pw[ivt1];
    s->vt2 = pw[ivt2];
    s->p = pw[ITAU];
    s->bt1 = pw[ibt1];
    s->bt2 = pw[ibt2];
    s->pt = pw[ITAU] + 0.5 * bsquare;
    s->cf = orc_signalspeed(cs, pw[ibn], pw[ibt1], pw[ibt2], pw[IRHO]);
    const int iu[7] = {IRHO, ivn, ivt1, ivt2, ITAU, ibt1, ibt2};
    for (int i = 0; i < 7; i++) {
        s->u[i] = u8[iu[i]];
        s->f[i] = f8[iu[i]];
    }
}

static void unrotate7(int dir, const double f7[7], double *flux)
{
    int ivn, ivt1, ivt2, ibn, ibt1, ibt2;
    dir_indices(dir, &ivn, &ivt1, &ivt2, &ibn, &ibt1, &ibt2);
    flux[IRHO] = f7[0];
    flux[ivn] = f7[1];
    flux[ivt1] = f7[2];
    flux[ivt2] = f7[3];
    flux[ITAU] = f7[4];
    flux[ibn] = 0.0;
    flux[ibt1] = f7[5];
    flux[ibt2] = f7[6];
}

/* Local Lax-Friedrichs (Rusanov) flux, the LF option definition.f90:139-140 declares and
 * rsolver.f90 leaves unimplemented (an extension):
 *   F = (F_L + F_R)/2 - s (U_R - U_L)/2,  s = max(|v_nL| + c_fL, |v_nR| + c_fR)
 * with the fast speeds of compute_signalspeed (definition.f90:394-410). */
int orc_lf(int dir, double g, const double *pl, const double *pr, double *flux)
{
    side7 L, R;
    side_of(dir, g, pl, &L);
    side_of(dir, g, pr, &R);
    double smax = fmax2(fabs(L.vn) + L.cf, fabs(R.vn) + R.cf);
    double f[7];
    for (int i = 0; i < 7; i++) f[i] = 0.5 * (L.f[i] + R.f[i]) - 0.5 * smax * (R.u[i] - L.u[i]);
    unrotate7(dir, f, flux);
    return 1;
}

/* HLLD flux of Miyoshi & Kusano (2005, J. Comput. Phys. 208, 315), the HLLD option
 * definition.f90:148-149 declares and rsolver.f90 leaves unimplemented (an extension), written
 * from the paper: outer speeds S_L,R = min/max(v_nL, v_nR) -/+ max(c_fL, c_fR) (eq. 67), contact
 * speed S_M and total pressure p_T* (eqs. 38, 41), single-star states (eqs. 43-48; the
 * degenerate case rho (S - v_n)(S - S_M) - B_n^2 -> 0 keeps the transverse v and B, as the paper
 * prescribes), Alfven speeds S*_L,R = S_M -/+ |B_n| / sqrt(rho*) (eq. 51) and double-star states
 * (eqs. 59-63).  Returns the region: 0 F_L, 1 F*_L, 2 F**_L, 3 F**_R, 4 F*_R, 5 F_R, 6 NaN. */
int orc_hlld(int dir, double g, const double *pl, const double *pr, double *flux)
{
    int ivn, ivt1, ivt2, ibn, ibt1, ibt2;
    dir_indices(dir, &ivn, &ivt1, &ivt2, &ibn, &ibt1, &ibt2);
    side7 L, R;
    side_of(dir, g, pl, &L);
    side_of(dir, g, pr, &R);
    const double bn = pl[ibn];
    double f[7];
    const double cmax = fmax2(L.cf, R.cf);
    const double sL = fmin2(L.vn, R.vn) - cmax;
    const double sR = fmax2(L.vn, R.vn) + cmax;
    int region;
    if (sL >= 0.0) {
        for (int i = 0; i < 7; i++) f[i] = L.f[i];
        region = 0;
    } else if (sR <= 0.0) {
        for (int i = 0; i < 7; i++) f[i] = R.f[i];
        region = 5;
    } else if (sL < 0.0 && sR > 0.0) {
        const double sdl = sL - L.vn, sdr = sR - R.vn;
        const double iden = 1.0 / (sdr * R.rho - sdl * L.rho);
        const double sm = (sdr * R.rho * R.vn - sdl * L.rho * L.vn - R.pt + L.pt) * iden;
        const double pts = (sdr * R.rho * L.pt - sdl * L.rho * R.pt + L.rho * R.rho * sdr * sdl * (R.vn - L.vn)) * iden;
        /* single-star states of both sides: us[a][0..6], with v_t*, B_t* */
        double us[2][7], vts[2][2], bts[2][2], sq[2];
        const side7 *S[2] = {&L, &R};
        const double sa[2] = {sL, sR}, sda[2] = {sdl, sdr};
        for (int a = 0; a < 2; a++) {
            const side7 *q = S[a];
            const double sdm = sa[a] - sm;
            const double isdm = 1.0 / sdm;
            const double rhos = q->rho * sda[a] * isdm;
            const double tmp = q->rho * sda[a] * sdm - bn * bn;
            if (fabs(tmp) < 1.0e-8 * fabs(pts)) {
                vts[a][0] = q->vt1;
                vts[a][1] = q->vt2;
                bts[a][0] = q->bt1;
                bts[a][1] = q->bt2;
            } else {
                const double fac = (q->rho * sda[a] * sda[a] - bn * bn) / tmp;
                const double fv = bn * (sm - q->vn) / tmp;
                vts[a][0] = q->vt1 - q->bt1 * fv;
                vts[a][1] = q->vt2 - q->bt2 * fv;
                bts[a][0] = q->bt1 * fac;
                bts[a][1] = q->bt2 * fac;
            }
            const double vdb = q->vn * bn + q->vt1 * q->bt1 + q->vt2 * q->bt2;
            const double vdbs = sm * bn + vts[a][0] * bts[a][0] + vts[a][1] * bts[a][1];
            const double es = (sda[a] * q->u[4] - q->pt * q->vn + pts * sm + bn * (vdb - vdbs)) * isdm;
            us[a][0] = rhos;
            us[a][1] = rhos * sm;
            us[a][2] = rhos * vts[a][0];
            us[a][3] = rhos * vts[a][1];
            us[a][4] = es;
            us[a][5] = bts[a][0];
            us[a][6] = bts[a][1];
            sq[a] = sqrt(rhos);
        }
        const double ssL = sm - fabs(bn) / sq[0];
        const double ssR = sm + fabs(bn) / sq[1];
        if (ssL >= 0.0) {
            for (int i = 0; i < 7; i++) f[i] = L.f[i] + sL * (us[0][i] - L.u[i]);
            region = 1;
        } else if (ssR <= 0.0) {
            for (int i = 0; i < 7; i++) f[i] = R.f[i] + sR * (us[1][i] - R.u[i]);
            region = 4;
        } else {
            /* double-star states (eqs. 59-63) */
            const double sgn = fsign1(bn);
            const double isq = 1.0 / (sq[0] + sq[1]);
            const double vss1 = (sq[0] * vts[0][0] + sq[1] * vts[1][0] + (bts[1][0] - bts[0][0]) * sgn) * isq;
            const double vss2 = (sq[0] * vts[0][1] + sq[1] * vts[1][1] + (bts[1][1] - bts[0][1]) * sgn) * isq;
            const double bss1 = (sq[0] * bts[1][0] + sq[1] * bts[0][0] + sq[0] * sq[1] * (vts[1][0] - vts[0][0]) * sgn) * isq;
            const double bss2 = (sq[0] * bts[1][1] + sq[1] * bts[0][1] + sq[0] * sq[1] * (vts[1][1] - vts[0][1]) * sgn) * isq;
            const double vdbss = sm * bn + vss1 * bss1 + vss2 * bss2;
            const int a = sm >= 0.0 ? 0 : 1;
            const side7 *q = S[a];
            const double vdbs = sm * bn + vts[a][0] * bts[a][0] + vts[a][1] * bts[a][1];
            const double ess = a == 0 ? us[0][4] - sq[0] * (vdbs - vdbss) * sgn : us[1][4] + sq[1] * (vdbs - vdbss) * sgn;
            const double uss[7] = {us[a][0], us[a][1], us[a][0] * vss1, us[a][0] * vss2, ess, bss1, bss2};
            const double ss = a == 0 ? ssL : ssR;
            for (int i = 0; i < 7; i++) {
                const double fs = q->f[i] + sa[a] * (us[a][i] - q->u[i]);
                f[i] = fs + ss * (uss[i] - us[a][i]);
            }
            region = a == 0 ? 2 : 3;
        }
    } else {
        for (int i = 0; i < 7; i++) f[i] = NAN;
        region = 6;
    }
    unrotate7(dir, f, flux);
    return region;
}

/* HLLC flux for MHD in the form of Li (2005, J. Comput. Phys. 203, 344), the HLLC option
 * definition.f90:145-146 declares and rsolver.f90 leaves unimplemented (an extension): outer
 * speeds as HLLD's (Miyoshi & Kusano eq. 67); the HLL state U_hll = (S_R U_R - S_L U_L - F_R +
 * F_L) / (S_R - S_L) gives the star-region field B* = B_hll and velocity v* = m_hll / rho_hll
 * with normal component S_M = m_n,hll / rho_hll; total pressure
 * p_T* = rho_L (S_L - v_nL)(S_M - v_nL) + p_TL (B_n continuous); on each side
 * rho* = rho (S - v_n) / (S - S_M), m_t* = (rho v_t (S - v_n) - (B_n B_t* - B_n B_t)) / (S - S_M),
 * E* = (E (S - v_n) - p_T v_n + p_T* S_M + B_n (v.B - v*.B*)) / (S - S_M), and
 * F* = F + S (U* - U).  Returns the region: 0 F_L, 1 F*_L, 2 F*_R, 3 F_R, 4 NaN. */
int orc_hllc(int dir, double g, const double *pl, const double *pr, double *flux)
{
    int ivn, ivt1, ivt2, ibn, ibt1, ibt2;
    dir_indices(dir, &ivn, &ivt1, &ivt2, &ibn, &ibt1, &ibt2);
    side7 L, R;
    side_of(dir, g, pl, &L);
    side_of(dir, g, pr, &R);
    const double bn = pl[ibn];
    double f[7];
    const double cmax = fmax2(L.cf, R.cf);
    const double sL = fmin2(L.vn, R.vn) - cmax;
    const double sR = fmax2(L.vn, R.vn) + cmax;
    int region;
    if (sL >= 0.0) {
        for (int i = 0; i < 7; i++) f[i] = L.f[i];
        region = 0;
    } else if (sR <= 0.0) {
        for (int i = 0; i < 7; i++) f[i] = R.f[i];
        region = 3;
    } else if (sL < 0.0 && sR > 0.0) {
        const double ids = 1.0 / (sR - sL);
        double uh[7];
        for (int i = 0; i < 7; i++) uh[i] = (sR * R.u[i] - sL * L.u[i] - R.f[i] + L.f[i]) * ids;
        const double irh = 1.0 / uh[0];
        const double sm = uh[1] * irh, vt1 = uh[2] * irh, vt2 = uh[3] * irh;
        const double bt1 = uh[5], bt2 = uh[6];
        const double pts = L.rho * (sL - L.vn) * (sm - L.vn) + L.pt;
        const double vdbs = sm * bn + vt1 * bt1 + vt2 * bt2;
        const int a = sm >= 0.0 ? 0 : 1;
        const side7 *q = a == 0 ? &L : &R;
        const double s = a == 0 ? sL : sR;
        const double sd = s - q->vn, isdm = 1.0 / (s - sm);
        const double vdb = q->vn * bn + q->vt1 * q->bt1 + q->vt2 * q->bt2;
        double us[7];
        us[0] = q->rho * sd * isdm;
        us[1] = us[0] * sm;
        us[2] = (q->rho * q->vt1 * sd - (bn * bt1 - bn * q->bt1)) * isdm;
        us[3] = (q->rho * q->vt2 * sd - (bn * bt2 - bn * q->bt2)) * isdm;
        us[4] = (q->u[4] * sd - q->pt * q->vn + pts * sm + bn * (vdb - vdbs)) * isdm;
        us[5] = bt1;
        us[6] = bt2;
        for (int i = 0; i < 7; i++) f[i] = q->f[i] + s * (us[i] - q->u[i]);
        region = a == 0 ? 1 : 2;
    } else {
        for (int i = 0; i < 7; i++) f[i] = NAN;
        region = 4;
    }
    unrotate7(dir, f, flux);
    return region;
}

/* the face solver of the configuration: 1 LF, 2 (or 0) HLL, 3 HLLC, 4 HLLD */
static int riemann(int solver, int dir, double g, const double *pl, const double *pr, double *flux)
{
    if (solver == 1) return orc_lf(dir, g, pl, pr, flux);
    if (solver == 3) return orc_hllc(dir, g, pl, pr, flux);
    if (solver == 4) return orc_hlld(dir, g, pl, pr, flux);
    return orc_hll(dir, g, pl, pr, flux);
}

/* ---------------------------------------------------------------------------------------------- */
/* Context                                                                                        */

static void setup_bfac(orc *o)
{
    /* setup.f90:121-230 */
    const int *bf = o->c.boundary_flag;
    for (int i = 0; i < NV; i++) {
        o->bfac_xin[i] = o->bfac_xout[i] = o->bfac_yin[i] = 1;
        o->bfac_yout[i] = o->bfac_zin[i] = o->bfac_zout[i] = 1;
    }
    if (bf[0] == 2) { o->bfac_xin[IVX] = -1; o->bfac_xin[IBX] = -1; }
    else if (bf[0] == 3) { o->bfac_xin[IVX] = -1; o->bfac_xin[IBX] = -1; o->bfac_xin[IVZ] = -1; o->bfac_xin[IBZ] = -1; }
    if (bf[1] == 2) { o->bfac_xout[IVX] = -1; o->bfac_xout[IBX] = -1; }
    else if (bf[1] == 3) { o->bfac_xout[IVX] = -1; o->bfac_xout[IBX] = -1; o->bfac_xout[IVZ] = -1; o->bfac_xout[IBZ] = -1; }
    if (bf[2] == 2) { o->bfac_yin[IVY] = -1; o->bfac_yin[IBY] = -1; }
    else if (bf[2] == 3) { o->bfac_yin[IVY] = -1; o->bfac_yin[IBY] = -1; o->bfac_yin[IVZ] = -1; o->bfac_yin[IBZ] = -1; }
    else if (bf[2] == 4) { o->bfac_yin[IVY] = -1; o->bfac_yin[IBX] = -1; o->bfac_yin[IBZ] = -1; }
    if (bf[3] == 2) { o->bfac_yout[IVY] = -1; o->bfac_yout[IBY] = -1; }
    else if (bf[3] == 3) { o->bfac_yout[IVY] = -1; o->bfac_yout[IBY] = -1; o->bfac_yout[IVZ] = -1; o->bfac_yout[IBZ] = -1; }
    else if (bf[3] == 4) { o->bfac_yout[IVY] = -1; o->bfac_yout[IBX] = -1; o->bfac_yout[IBZ] = -1; }
    if (bf[4] == 2) { o->bfac_zin[IVZ] = -1; o->bfac_zin[IBZ] = -1; }
    else if (bf[4] == 4) { o->bfac_zin[IVZ] = -1; o->bfac_zin[IBX] = -1; o->bfac_zin[IBY] = -1; }
    if (bf[5] == 2) { o->bfac_zout[IVZ] = -1; o->bfac_zout[IBZ] = -1; }
    else if (bf[5] == 4) { o->bfac_zout[IVZ] = -1; o->bfac_zout[IBX] = -1; o->bfac_zout[IBY] = -1; }
}

orc *orc_create(const orc_config *cfg, const double *xf, const double *yf, const double *zf)
{
    orc *o = (orc *)calloc(1, sizeof(orc));
    if (!o) return NULL;
    o->c = *cfg;
    o->nx = cfg->nx; o->ny = cfg->ny; o->nz = cfg->nz;
    o->sx = o->nx + 2 * NGHOST; o->sy = o->ny + 2 * NGHOST; o->sz = o->nz + 2 * NGHOST;
    o->ncell = o->sx * o->sy * o->sz;
    o->xf_mem = (double *)malloc(sizeof(double) * (o->nx + 7));
    o->yf_mem = (double *)malloc(sizeof(double) * (o->ny + 7));
    o->zf_mem = (double *)malloc(sizeof(double) * (o->nz + 7));
    memcpy(o->xf_mem, xf, sizeof(double) * (o->nx + 7));
    memcpy(o->yf_mem, yf, sizeof(double) * (o->ny + 7));
    memcpy(o->zf_mem, zf, sizeof(double) * (o->nz + 7));
    o->xf = o->xf_mem + 3; o->yf = o->yf_mem + 3; o->zf = o->zf_mem + 3;
    size_t n8 = (size_t)o->ncell * NV, n3 = (size_t)o->ncell * 3;
    /* zero-initialised: the reference's ALLOCATE leaves them undefined (Q4, Q6) */
    o->prim = (double *)calloc(n8, sizeof(double));
    o->cons = (double *)calloc(n8, sizeof(double));
    o->primL = (double *)calloc(n8, sizeof(double));
    o->primR = (double *)calloc(n8, sizeof(double));
    o->flux = (double *)calloc(n8, sizeof(double));
    o->l_rk = (double *)calloc(n8, sizeof(double));
    o->u_old = (double *)calloc(n8, sizeof(double));
    o->sc = (double *)calloc(n8, sizeof(double));
    o->eps = (double *)calloc(o->ncell, sizeof(double));
    o->bcell = (double *)calloc(n3, sizeof(double));
    o->ecell = (double *)calloc(n3, sizeof(double));
    o->ecorn = (double *)calloc(n3, sizeof(double));
    if (!o->prim || !o->cons || !o->primL || !o->primR || !o->flux || !o->l_rk || !o->u_old || !o->sc ||
        !o->eps || !o->bcell || !o->ecell || !o->ecorn) {
        orc_destroy(o);
        return NULL;
    }
    setup_bfac(o);
    return o;
}

void orc_destroy(orc *o)
{
    if (!o) return;
    free(o->xf_mem); free(o->yf_mem); free(o->zf_mem);
    free(o->prim); free(o->cons); free(o->primL); free(o->primR); free(o->flux);
    free(o->l_rk); free(o->u_old); free(o->sc); free(o->eps);
    free(o->bcell); free(o->ecell); free(o->ecorn);
    free(o);
}

long orc_ncell(const orc *o) { return o->ncell; }

int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

void orc_set_state(orc *o, const double *prim, const double *eps)
{
    memcpy(o->prim, prim, sizeof(double) * o->ncell * NV);
    memcpy(o->eps, eps, sizeof(double) * o->ncell);
}

void orc_get_state(const orc *o, double *prim, double *eps, double *bcell, double *cons)
{
    if (prim) memcpy(prim, o->prim, sizeof(double) * o->ncell * NV);
    if (eps) memcpy(eps, o->eps, sizeof(double) * o->ncell);
    if (bcell) memcpy(bcell, o->bcell, sizeof(double) * o->ncell * 3);
    if (cons) memcpy(cons, o->cons, sizeof(double) * o->ncell * NV);
}

void orc_get_lrk(const orc *o, double *l_rk) { memcpy(l_rk, o->l_rk, sizeof(double) * o->ncell * NV); }

/* ---------------------------------------------------------------------------------------------- */
/* Model plugin hooks                                                                             */

/* riemann/src/custom.f90:54-76 ; torus_cyn/src/custom.f90:86-116 ; torus_sph/src/custom.f90:86-116 */
static void custom_checkrho(orc *o)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    long nr = 0, ne = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : nr, ne)
    for (int l = 1; l <= nz; l++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++) {
                if (o->c.model != 0) {
                    if (PR(IRHO, j, k, l) < o->c.rho_floor) { PR(IRHO, j, k, l) = o->c.rho_floor; nr++; }
                    if (EPS(j, k, l) < o->c.eps_floor) { EPS(j, k, l) = o->c.eps_floor; ne++; }
                } else {
                    ne += EPS(j, k, l) < 0.0;
                    EPS(j, k, l) = fmax2(EPS(j, k, l), 0.0);
                }
            }
    o->nfl_rho += nr;
    o->nfl_eps += ne;
}

void orc_floor_counts(const orc *o, long *rho, long *eps)
{
    *rho = o->nfl_rho;
    *eps = o->nfl_eps;
}

/* torus_cyn/src/custom.f90:121-166 and torus_sph/src/custom.f90:121-158 (Paczynski-Wiita gravity:
 * cylindrical R and z components; spherical radial component) */
static void custom_source(orc *o)
{
    if (o->c.model == 0) return;
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    const double r_sh = o->c.r_sh, rho_floor = o->c.rho_floor;
    if (o->c.model == 2) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int l = 1; l <= nz; l++)
            for (int k = 1; k <= ny; k++)
                for (int j = 1; j <= nx; j++) {
                    double x_loc = cxc(o, j);
                    double diff = PR(IRHO, j, k, l) - rho_floor;
                    double factor = fmax2(fsign1(diff), 0.0);
                    double dphidr = 1.0 / ((x_loc - r_sh) * (x_loc - r_sh));
                    SC(IVX, j, k, l) = SC(IVX, j, k, l) + (-factor * PR(IRHO, j, k, l) * dphidr);
                    SC(ITAU, j, k, l) = SC(ITAU, j, k, l) + (-factor * PR(IRHO, j, k, l) * PR(IVX, j, k, l) * dphidr);
                }
        return;
    }
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 1; l <= nz; l++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++) {
                double x_loc = cxc(o, j), z_loc = czc(o, l);
                double diff = PR(IRHO, j, k, l) - rho_floor;
                double factor = fmax2(fsign1(diff), 0.0);
                double radius = sqrt(x_loc * x_loc + z_loc * z_loc);
                double dphidr = x_loc / radius / ((radius - r_sh) * (radius - r_sh));
                double dphidz = z_loc / radius / ((radius - r_sh) * (radius - r_sh));
                SC(IVX, j, k, l) = SC(IVX, j, k, l) + (-factor * PR(IRHO, j, k, l) * dphidr);
                SC(IVZ, j, k, l) = SC(IVZ, j, k, l) + (-factor * PR(IRHO, j, k, l) * dphidz);
                SC(ITAU, j, k, l) = SC(ITAU, j, k, l) +
                    (-factor * PR(IRHO, j, k, l) * (PR(IVX, j, k, l) * dphidr + PR(IVZ, j, k, l) * dphidz));
            }
}

/* torus_cyn/src/custom.f90:47-81, torus_sph/src/custom.f90:47-70 (ghost inflow clamp: radial
 * ghosts, and for torus_cyn the z ghosts) */
static void custom_boundary(orc *o)
{
    if (o->c.model == 0) return;
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    for (int l = -2; l <= nz + 3; l++)
        for (int k = -2; k <= ny + 3; k++)
            for (int j = 1; j <= 3; j++) {
                PR(IVX, 1 - j, k, l) = fmin2(PR(IVX, 1 - j, k, l), 0.0);
                PR(IVX, nx + j, k, l) = fmax2(PR(IVX, nx + j, k, l), 0.0);
            }
    if (o->c.model != 1) return;
    for (int l = 1; l <= 3; l++)
        for (int k = -2; k <= ny + 3; k++)
            for (int j = -2; j <= nx + 3; j++) {
                PR(IVZ, j, k, 1 - l) = fmin2(PR(IVZ, j, k, 1 - l), 0.0);
                PR(IVZ, j, k, nz + l) = fmax2(PR(IVZ, j, k, nz + l), 0.0);
            }
}

/* ---------------------------------------------------------------------------------------------- */
/* Ghost fill                                                                                     */

/* Periodic ghost source: boundary.f90 copies index n+1-g into 1-g and g into n+g inside loops
 * that run g = 1..3 in order; for n >= 3 that is the index below, and for n < 3 the sequential
 * order makes each ghost a copy of its periodic image, which this wrap also returns. */
static inline int pwrap(int i, int n) { return ((i - 1) % n + n) % n + 1; }

/* One ghost copy of the BOUNDARYP variable set: prim(1:5), bcell(1:3) and the two face-B
 * components transverse to the pass axis (Q3).  `tb1`, `tb2` are those two components. */
static inline void copy_p(orc *o, long dst, long src, const int *fac, int use_fac, int tb1, int tb2)
{
    for (int i = IRHO; i <= ITAU; i++)
        o->prim[dst * NV + i] = use_fac ? fac[i] * o->prim[src * NV + i] : o->prim[src * NV + i];
    for (int c = 0; c < 3; c++)
        o->bcell[dst * 3 + c] = use_fac ? fac[IBX + c] * o->bcell[src * 3 + c] : o->bcell[src * 3 + c];
    o->prim[dst * NV + tb1] = use_fac ? fac[tb1] * o->prim[src * NV + tb1] : o->prim[src * NV + tb1];
    o->prim[dst * NV + tb2] = use_fac ? fac[tb2] * o->prim[src * NV + tb2] : o->prim[src * NV + tb2];
}

/* boundary.f90:349-672 BOUNDARYP (x -> y -> z passes; Q18: pass ranges are part of the semantics) */
static void boundaryp(orc *o)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    const int *bf = o->c.boundary_flag;
    /* x passes :364-452, k,l in [0,n] */
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l <= nz; l++)
        for (int k = 0; k <= ny; k++)
            for (int j = 1; j <= NGHOST; j++) {
                if (bf[0] == 0) copy_p(o, ID(1 - j, k, l), ID(pwrap(1 - j, nx), k, l), NULL, 0, IBY, IBZ);
                else if (bf[0] == 1) copy_p(o, ID(1 - j, k, l), ID(1, k, l), NULL, 0, IBY, IBZ);
                else copy_p(o, ID(1 - j, k, l), ID(j, k, l), o->bfac_xin, 1, IBY, IBZ);
                if (bf[1] == 0) copy_p(o, ID(nx + j, k, l), ID(pwrap(nx + j, nx), k, l), NULL, 0, IBY, IBZ);
                else if (bf[1] == 1) copy_p(o, ID(nx + j, k, l), ID(nx, k, l), NULL, 0, IBY, IBZ);
                else copy_p(o, ID(nx + j, k, l), ID(nx + 1 - j, k, l), o->bfac_xout, 1, IBY, IBZ);
            }
    /* y passes :464-552, j in [-2, n+3], l in [0,n] */
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l <= nz; l++)
        for (int k = 1; k <= NGHOST; k++)
            for (int j = 1 - NGHOST; j <= nx + NGHOST; j++) {
                if (bf[2] == 0) copy_p(o, ID(j, 1 - k, l), ID(j, pwrap(1 - k, ny), l), NULL, 0, IBX, IBZ);
                else if (bf[2] == 1) copy_p(o, ID(j, 1 - k, l), ID(j, 1, l), NULL, 0, IBX, IBZ);
                else copy_p(o, ID(j, 1 - k, l), ID(j, k, l), o->bfac_yin, 1, IBX, IBZ);
                if (bf[3] == 0) copy_p(o, ID(j, ny + k, l), ID(j, pwrap(ny + k, ny), l), NULL, 0, IBX, IBZ);
                else if (bf[3] == 1) copy_p(o, ID(j, ny + k, l), ID(j, ny, l), NULL, 0, IBX, IBZ);
                else copy_p(o, ID(j, ny + k, l), ID(j, ny + 1 - k, l), o->bfac_yout, 1, IBX, IBZ);
            }
    /* z passes :564-652, j,k full */
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 1; l <= NGHOST; l++)
        for (int k = 1 - NGHOST; k <= ny + NGHOST; k++)
            for (int j = 1 - NGHOST; j <= nx + NGHOST; j++) {
                if (bf[4] == 0) copy_p(o, ID(j, k, 1 - l), ID(j, k, pwrap(1 - l, nz)), NULL, 0, IBX, IBY);
                else if (bf[4] == 1) copy_p(o, ID(j, k, 1 - l), ID(j, k, 1), NULL, 0, IBX, IBY);
                else copy_p(o, ID(j, k, 1 - l), ID(j, k, l), o->bfac_zin, 1, IBX, IBY);
                if (bf[5] == 0) copy_p(o, ID(j, k, nz + l), ID(j, k, pwrap(nz + l, nz)), NULL, 0, IBX, IBY);
                else if (bf[5] == 1) copy_p(o, ID(j, k, nz + l), ID(j, k, nz), NULL, 0, IBX, IBY);
                else copy_p(o, ID(j, k, nz + l), ID(j, k, nz + 1 - l), o->bfac_zout, 1, IBX, IBY);
            }
    custom_boundary(o);
}

/* boundary.f90:29-341 BOUNDARY1D applied to eps with all-even parity (fac = 1).
 * Q2: the y-outer periodic test reads boundary_flag(3) (:211). */
static void boundary1d_eps(orc *o)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    const int *bf = o->c.boundary_flag;
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 1; l <= nz; l++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= NGHOST; j++) {
                if (bf[0] == 0) EPS(1 - j, k, l) = EPS(pwrap(1 - j, nx), k, l);
                else if (bf[0] == 1) EPS(1 - j, k, l) = EPS(1, k, l);
                else EPS(1 - j, k, l) = EPS(j, k, l);
                if (bf[1] == 0) EPS(nx + j, k, l) = EPS(pwrap(nx + j, nx), k, l);
                else if (bf[1] == 1) EPS(nx + j, k, l) = EPS(nx, k, l);
                else EPS(nx + j, k, l) = EPS(nx + 1 - j, k, l);
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 1; l <= nz; l++)
        for (int k = 1; k <= NGHOST; k++)
            for (int j = 1 - NGHOST; j <= nx + NGHOST; j++) {
                if (bf[2] == 0) EPS(j, 1 - k, l) = EPS(j, pwrap(1 - k, ny), l);
                else if (bf[2] == 1) EPS(j, 1 - k, l) = EPS(j, 1, l);
                else EPS(j, 1 - k, l) = EPS(j, k, l);
                if (bf[2] == 0) EPS(j, ny + k, l) = EPS(j, pwrap(ny + k, ny), l);
                else if (bf[3] == 1) EPS(j, ny + k, l) = EPS(j, ny, l);
                else if (bf[3] >= 2) EPS(j, ny + k, l) = EPS(j, ny + 1 - k, l);
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 1; l <= NGHOST; l++)
        for (int k = 1 - NGHOST; k <= ny + NGHOST; k++)
            for (int j = 1 - NGHOST; j <= nx + NGHOST; j++) {
                if (bf[4] == 0) EPS(j, k, 1 - l) = EPS(j, k, pwrap(1 - l, nz));
                else if (bf[4] == 1) EPS(j, k, 1 - l) = EPS(j, k, 1);
                else EPS(j, k, 1 - l) = EPS(j, k, l);
                if (bf[5] == 0) EPS(j, k, nz + l) = EPS(j, k, pwrap(nz + l, nz));
                else if (bf[5] == 1) EPS(j, k, nz + l) = EPS(j, k, nz);
                else EPS(j, k, nz + l) = EPS(j, k, nz + 1 - l);
            }
}

/* boundary.f90:5-13 */
void orc_boundary(orc *o)
{
    boundaryp(o);
    boundary1d_eps(o);
}

/* ---------------------------------------------------------------------------------------------- */
/* Conversion sweeps                                                                              */

/* fromrvetou.f90:21-70 FROMRVETOU */
static void fromrvetou(orc *o)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 1; l <= nz; l++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++) {
                BC(0, j, k, l) = 0.5 * (PR(IBX, j, k, l) + PR(IBX, j - 1, k, l));
                BC(1, j, k, l) = 0.5 * (PR(IBY, j, k, l) + PR(IBY, j, k - 1, l));
                BC(2, j, k, l) = 0.5 * (PR(IBZ, j, k, l) + PR(IBZ, j, k, l - 1));
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l <= nz; l++)
        for (int k = 0; k <= ny; k++)
            for (int j = 0; j <= nx; j++) {
                long c = ID(j, k, l);
                orc_ptou_cell(&o->prim[c * NV], o->eps[c], &o->bcell[c * 3], &o->cons[c * NV]);
            }
}

/* fromrvetou.f90:75-127 FROMUTORVE */
static void fromutorve(orc *o)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 1; l <= nz; l++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++) {
                BC(0, j, k, l) = 0.5 * (CO(IBX, j, k, l) + CO(IBX, j - 1, k, l));
                BC(1, j, k, l) = 0.5 * (CO(IBY, j, k, l) + CO(IBY, j, k - 1, l));
                BC(2, j, k, l) = 0.5 * (CO(IBZ, j, k, l) + CO(IBZ, j, k, l - 1));
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l <= nz; l++)
        for (int k = 0; k <= ny; k++)
            for (int j = 0; j <= nx; j++) {
                long c = ID(j, k, l);
                orc_utop_cell(&o->cons[c * NV], &o->bcell[c * 3], &o->prim[c * NV], &o->eps[c]);
            }
}

/* update.f90:12-45 UPDATE: EOS pressure on every cell including ghosts (Q12) */
static void update(orc *o)
{
    const double g = o->c.ggas;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < o->ncell; c++) o->prim[c * NV + ITAU] = eos_pressure(g, o->prim[c * NV + IRHO], o->eps[c]);
}

/* initial.f90:83-140 initial_update */
void orc_initial_update(orc *o)
{
    custom_checkrho(o);
    fromrvetou(o);
    orc_boundary(o);
    update(o);
}

/* ---------------------------------------------------------------------------------------------- */
/* Spatial operator                                                                               */

/* coords.f90:105-160 GEOM_SOURCE, cylindrical and spherical branches */
static inline void geom_source(orc *o, int j, int k, int l)
{
    if (o->c.coordinate == 3) {
        double x_loc = cxc(o, j), y_loc = cyc(o, k);
        double bsquare = dotp(BC(0, j, k, l), BC(1, j, k, l), BC(2, j, k, l), BC(0, j, k, l), BC(1, j, k, l),
                              BC(2, j, k, l));
        SC(IVX, j, k, l) = SC(IVX, j, k, l) +
            (2.0 * PR(ITAU, j, k, l) + BC(0, j, k, l) * BC(0, j, k, l) +
             PR(IRHO, j, k, l) * (PR(IVY, j, k, l) * PR(IVY, j, k, l) + PR(IVZ, j, k, l) * PR(IVZ, j, k, l))) /
                (x_loc);
        SC(IVY, j, k, l) = SC(IVY, j, k, l) +
            (PR(ITAU, j, k, l) + PR(IRHO, j, k, l) * PR(IVZ, j, k, l) * PR(IVZ, j, k, l) + 0.5 * bsquare -
             BC(2, j, k, l) * BC(2, j, k, l)) / (x_loc * tan(y_loc)) -
            (PR(IRHO, j, k, l) * PR(IVX, j, k, l) * PR(IVY, j, k, l) - BC(0, j, k, l) * BC(1, j, k, l)) / (x_loc);
        SC(IVZ, j, k, l) = SC(IVZ, j, k, l) -
            (PR(IRHO, j, k, l) * PR(IVX, j, k, l) * PR(IVZ, j, k, l) - BC(0, j, k, l) * BC(2, j, k, l)) / (x_loc) -
            (PR(IRHO, j, k, l) * PR(IVY, j, k, l) * PR(IVZ, j, k, l) - BC(1, j, k, l) * BC(2, j, k, l)) /
                (x_loc * tan(y_loc));
        return;
    }
    if (o->c.coordinate != 2) return;
    double x_loc = cxc(o, j);
    double bsquare = dotp(BC(0, j, k, l), BC(1, j, k, l), BC(2, j, k, l), BC(0, j, k, l), BC(1, j, k, l), BC(2, j, k, l));
    SC(IVX, j, k, l) = SC(IVX, j, k, l) +
        (PR(ITAU, j, k, l) + PR(IRHO, j, k, l) * PR(IVY, j, k, l) * PR(IVY, j, k, l) + 0.5 * bsquare -
         BC(1, j, k, l) * BC(1, j, k, l)) / (x_loc);
    SC(IVY, j, k, l) = SC(IVY, j, k, l) -
        (PR(IRHO, j, k, l) * PR(IVX, j, k, l) * PR(IVY, j, k, l) - BC(0, j, k, l) * BC(1, j, k, l)) / (x_loc);
}

/* spatial.f90:75-102 GET_SOURCE (Q1: GEOM_SOURCE runs once per equation index i) */
static void get_source(orc *o)
{
    if (o->c.coordinate != 2 && o->c.coordinate != 3) return;
    const int nx = o->nx, ny = o->ny, nz = o->nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 1; l <= nz; l++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++)
                for (int i = 0; i < o->c.geom_source_repeat; i++) geom_source(o, j, k, l);
}

/* coords.f90:231-302 GEOM_FLUX (cartesian / cylindrical / spherical) */
static inline void geom_flux(const orc *o, int dir, int j, int k, int l, double *gp, double *gc, double *gm)
{
    (void)l;
    if (o->c.coordinate == 3) {
        if (dir == 1) {
            *gp = pw2(o->xf[j]);
            *gm = pw2(o->xf[j - 1]);
            *gc = (pw3(o->xf[j]) - pw3(o->xf[j - 1])) / 3.0;
        } else if (dir == 2) {
            *gp = sin(o->yf[k]);
            *gm = sin(o->yf[k - 1]);
            *gc = sph_a(o, j) * sph_cd(o, k);
        } else {
            *gp = 1.0;
            *gm = 1.0;
            *gc = sph_a(o, j) * sph_cd(o, k) * cdz(o, l) / cdy(o, k);
        }
        return;
    }
    if (o->c.coordinate == 2) {
        double x_loc = cxc(o, j);
        if (dir == 1) { *gp = o->xf[j]; *gm = o->xf[j - 1]; *gc = x_loc * cdx(o, j); }
        else if (dir == 2) { *gp = 1.0; *gm = 1.0; *gc = x_loc * cdy(o, k); }
        else { *gp = 1.0; *gm = 1.0; *gc = cdz(o, l); }
    } else {
        *gp = 1.0; *gm = 1.0;
        *gc = (dir == 1) ? cdx(o, j) : (dir == 2) ? cdy(o, k) : cdz(o, l);
    }
}

/* spatial.f90:113-380 GET_FLUXES: one directional sweep */
static void sweep(orc *o, int dir)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    const double g = o->c.ggas;
    /* transverse cell-centred B components reconstructed per direction (:409-410, :487-488, :565-566)
     * and the normal face component copied (:413-414, :491-492, :569-570) */
    const int bt_a = (dir == 1) ? 1 : 0;          /* bcell index */
    const int bt_b = (dir == 3) ? 1 : 2;
    const int ibn = (dir == 1) ? IBX : (dir == 2) ? IBY : IBZ;
    const int dj = (dir == 1), dk = (dir == 2), dl = (dir == 3);

    /* (i) reconstruction on [0,n+1]^3 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l <= nz + 1; l++)
        for (int k = 0; k <= ny + 1; k++)
            for (int j = 0; j <= nx + 1; j++) {
                int c = (dir == 1) ? j : (dir == 2) ? k : l;
                double v[5], vm, vp;
                for (int i = IRHO; i <= ITAU; i++) {
                    for (int s = -2; s <= 2; s++) v[s + 2] = PR(i, j + s * dj, k + s * dk, l + s * dl);
                    interpolate(o, dir, c, v, &vm, &vp);
                    PRR(i, j - dj, k - dk, l - dl) = vm;
                    PL(i, j, k, l) = vp;
                }
                for (int t = 0; t < 2; t++) {
                    int bc = t ? bt_b : bt_a;
                    for (int s = -2; s <= 2; s++) v[s + 2] = BC(bc, j + s * dj, k + s * dk, l + s * dl);
                    interpolate(o, dir, c, v, &vm, &vp);
                    PRR(IBX + bc, j - dj, k - dk, l - dl) = vm;
                    PL(IBX + bc, j, k, l) = vp;
                }
                if (o->c.recon == 5 && (PRR(IRHO, j - dj, k - dk, l - dl) <= 0.0 || PL(IRHO, j, k, l) <= 0.0 ||
                                        PRR(ITAU, j - dj, k - dk, l - dl) <= 0.0 || PL(ITAU, j, k, l) <= 0.0)) {
                    /* WENO positivity fallback: a cell whose face density or pressure is not
                     * positive gets the piecewise-constant profile in every variable */
                    for (int i = IRHO; i <= ITAU; i++) {
                        PRR(i, j - dj, k - dk, l - dl) = PR(i, j, k, l);
                        PL(i, j, k, l) = PR(i, j, k, l);
                    }
                    for (int t = 0; t < 2; t++) {
                        int bc = t ? bt_b : bt_a;
                        PRR(IBX + bc, j - dj, k - dk, l - dl) = BC(bc, j, k, l);
                        PL(IBX + bc, j, k, l) = BC(bc, j, k, l);
                    }
                }
                PRR(ibn, j - dj, k - dk, l - dl) = PR(ibn, j - dj, k - dk, l - dl);
                PL(ibn, j, k, l) = PR(ibn, j, k, l);
            }
    /* (ii) Riemann problem on faces: [0,n] along dir, [0,n+1] across */
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l <= nz + 1 - dl; l++)
        for (int k = 0; k <= ny + 1 - dk; k++)
            for (int j = 0; j <= nx + 1 - dj; j++) {
                long c = ID(j, k, l);
                riemann(o->c.solver, dir, g, &o->primL[c * NV], &o->primR[c * NV], &o->flux[c * NV]);
            }
    /* (iii) flux difference on [1,n]^3 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 1; l <= nz; l++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++)
                for (int i = IRHO; i <= ITAU; i++) {
                    double gp, gc, gm;
                    geom_flux(o, dir, j, k, l, &gp, &gc, &gm);
                    LRK(i, j, k, l) = LRK(i, j, k, l) -
                        (FL(i, j, k, l) * gp - FL(i, j - dj, k - dk, l - dl) * gm) / gc;
                }
    /* (iv) EMF accumulation on [0,n]^3 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l <= nz; l++)
        for (int k = 0; k <= ny; k++)
            for (int j = 0; j <= nx; j++) {
                if (dir == 1) {
                    ER(IEY, j, k, l) = ER(IEY, j, k, l) + 0.25 * (FL(IBZ, j, k, l) + FL(IBZ, j, k, l + 1));
                    ER(IEZ, j, k, l) = ER(IEZ, j, k, l) + 0.25 * (-FL(IBY, j, k, l) - FL(IBY, j, k + 1, l));
                } else if (dir == 2) {
                    ER(IEX, j, k, l) = ER(IEX, j, k, l) + 0.25 * (-FL(IBZ, j, k, l) - FL(IBZ, j, k, l + 1));
                    ER(IEZ, j, k, l) = ER(IEZ, j, k, l) + 0.25 * (FL(IBX, j, k, l) + FL(IBX, j + 1, k, l));
                } else {
                    ER(IEX, j, k, l) = ER(IEX, j, k, l) + 0.25 * (FL(IBY, j, k, l) + FL(IBY, j, k + 1, l));
                    ER(IEY, j, k, l) = ER(IEY, j, k, l) + 0.25 * (-FL(IBX, j, k, l) - FL(IBX, j + 1, k, l));
                }
            }
}

/* coords.f90:310-415 GEOM_CT (cartesian / cylindrical / spherical) */
typedef struct { double bx_ez[3], bx_ey[3], by_ex[3], by_ez[3], bz_ex[3], bz_ey[3]; } ctgeom; /* m, c, p */
static inline void geom_ct(const orc *o, int j, int k, int l, ctgeom *g)
{
    double dx_loc = cdx(o, j), dy_loc = cdy(o, k), dz_loc = cdz(o, l);
    for (int i = 0; i < 3; i++) {
        g->bx_ez[i] = g->bx_ey[i] = g->by_ex[i] = g->by_ez[i] = g->bz_ex[i] = g->bz_ey[i] = 1.0;
    }
    if (o->c.coordinate == 3) {
        double x_loc = cxc(o, j);
        g->bx_ez[0] = sin(o->yf[k - 1]);
        g->bx_ez[1] = o->xf[j] * sph_cd(o, k);
        g->bx_ez[2] = sin(o->yf[k]);
        g->bx_ey[1] = o->xf[j] * sph_cd(o, k) * dz_loc / dy_loc;
        g->by_ex[1] = x_loc * sin(o->yf[k]) * dz_loc;
        g->by_ez[0] = o->xf[j - 1];
        g->by_ez[1] = x_loc * dx_loc;
        g->by_ez[2] = o->xf[j];
        g->bz_ey[0] = o->xf[j - 1];
        g->bz_ey[1] = x_loc * dx_loc;
        g->bz_ey[2] = o->xf[j];
        g->bz_ex[1] = x_loc * dy_loc;
    } else if (o->c.coordinate == 2) {
        double x_loc = cxc(o, j);
        g->bx_ez[1] = dy_loc * o->xf[j];
        g->bx_ey[1] = dz_loc;
        g->by_ex[1] = dz_loc;
        g->by_ez[1] = dx_loc;
        g->bz_ey[0] = o->xf[j - 1];
        g->bz_ey[1] = x_loc * dx_loc;
        g->bz_ey[2] = o->xf[j];
        g->bz_ex[1] = x_loc * dy_loc;
    } else {
        g->bx_ez[1] = dy_loc;
        g->bx_ey[1] = dz_loc;
        g->by_ex[1] = dz_loc;
        g->by_ez[1] = dx_loc;
        g->bz_ey[1] = dx_loc;
        g->bz_ex[1] = dy_loc;
    }
}

/* mhd.f90:73-157 flux_ct (Gardiner & Stone upwind-CT corner EMF, face-B right-hand side) */
static void flux_ct(orc *o)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l <= nz + 1; l++)
        for (int k = 0; k <= ny + 1; k++)
            for (int j = 0; j <= nx + 1; j++) {
                EC(IEX, j, k, l) = PR(IVZ, j, k, l) * BC(1, j, k, l) - PR(IVY, j, k, l) * BC(2, j, k, l);
                EC(IEY, j, k, l) = PR(IVX, j, k, l) * BC(2, j, k, l) - PR(IVZ, j, k, l) * BC(0, j, k, l);
                EC(IEZ, j, k, l) = PR(IVY, j, k, l) * BC(0, j, k, l) - PR(IVX, j, k, l) * BC(1, j, k, l);
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l <= nz; l++)
        for (int k = 0; k <= ny; k++)
            for (int j = 0; j <= nx; j++) {
                ER(IEX, j, k, l) = ER(IEX, j, k, l) * 2.0 -
                    0.25 * (EC(IEX, j, k, l) + EC(IEX, j, k, l + 1) + EC(IEX, j, k + 1, l) + EC(IEX, j, k + 1, l + 1));
                ER(IEY, j, k, l) = ER(IEY, j, k, l) * 2.0 -
                    0.25 * (EC(IEY, j, k, l) + EC(IEY, j, k, l + 1) + EC(IEY, j + 1, k, l) + EC(IEY, j + 1, k, l + 1));
                ER(IEZ, j, k, l) = ER(IEZ, j, k, l) * 2.0 -
                    0.25 * (EC(IEZ, j, k, l) + EC(IEZ, j, k + 1, l) + EC(IEZ, j + 1, k, l) + EC(IEZ, j + 1, k + 1, l));
            }
    /* Q4: ecorn at index -1 is read for ghost faces only; it stays at its initial zero here */
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l <= nz; l++)
        for (int k = 0; k <= ny; k++)
            for (int j = 0; j <= nx; j++) {
                ctgeom g;
                geom_ct(o, j, k, l, &g);
                LRK(IBX, j, k, l) = -((ER(IEZ, j, k, l) * g.bx_ez[2] - ER(IEZ, j, k - 1, l) * g.bx_ez[0]) / g.bx_ez[1] -
                                      (ER(IEY, j, k, l) * g.bx_ey[2] - ER(IEY, j, k, l - 1) * g.bx_ey[0]) / g.bx_ey[1]);
                LRK(IBY, j, k, l) = -((ER(IEX, j, k, l) * g.by_ex[2] - ER(IEX, j, k, l - 1) * g.by_ex[0]) / g.by_ex[1] -
                                      (ER(IEZ, j, k, l) * g.by_ez[2] - ER(IEZ, j - 1, k, l) * g.by_ez[0]) / g.by_ez[1]);
                LRK(IBZ, j, k, l) = -((ER(IEY, j, k, l) * g.bz_ey[2] - ER(IEY, j - 1, k, l) * g.bz_ey[0]) / g.bz_ey[1] -
                                      (ER(IEX, j, k, l) * g.bz_ex[2] - ER(IEX, j, k - 1, l) * g.bz_ex[0]) / g.bz_ex[1]);
            }
}

/* Face fluxes of one sweep (spatial.f90:396-433 reconstruction + Riemann) on the current state,
 * written as 7 components per cell: rho, mx, my, mz, E, then the two face-transverse B fluxes in
 * x, y, z order.  Used to check the HIP flux kernels in isolation. */
void orc_sweep_flux(orc *o, int dir, double *out)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    const int dj = (dir == 1), dk = (dir == 2), dl = (dir == 3);
    const int ta = (dir == 1) ? IBY : IBX, tb = (dir == 3) ? IBY : IBZ;
    memset(out, 0, sizeof(double) * 7 * o->ncell);
    /* reuse the sweep's stage (i)+(ii) without touching l_rk / ecorn: run it on copies */
    double *lrk = o->l_rk, *er = o->ecorn;
    o->l_rk = (double *)calloc((size_t)o->ncell * NV, sizeof(double));
    o->ecorn = (double *)calloc((size_t)o->ncell * 3, sizeof(double));
    sweep(o, dir);
    free(o->l_rk);
    free(o->ecorn);
    o->l_rk = lrk;
    o->ecorn = er;
    for (int l = 0; l <= nz + 1 - dl; l++)
        for (int k = 0; k <= ny + 1 - dk; k++)
            for (int j = 0; j <= nx + 1 - dj; j++) {
                long c = ID(j, k, l);
                for (int i = 0; i < 5; i++) out[c * 7 + i] = o->flux[c * NV + i];
                out[c * 7 + 5] = o->flux[c * NV + ta];
                out[c * 7 + 6] = o->flux[c * NV + tb];
            }
}

/* spatial.f90:11-68 SPATIAL */
void orc_spatial(orc *o)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l <= nz; l++)
        for (int k = 0; k <= ny; k++)
            for (int j = 0; j <= nx; j++) {
                for (int i = 0; i < NV; i++) { SC(i, j, k, l) = 0.0; LRK(i, j, k, l) = 0.0; }
                for (int c = 0; c < 3; c++) ER(c, j, k, l) = 0.0;
            }
    get_source(o);
    custom_source(o);
    sweep(o, 1);
    sweep(o, 2);
    sweep(o, 3);
    /* :365-375 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 1; l <= nz; l++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++)
                for (int i = IRHO; i <= ITAU; i++) LRK(i, j, k, l) = LRK(i, j, k, l) + SC(i, j, k, l);
    flux_ct(o);
}

/* ---------------------------------------------------------------------------------------------- */
/* Time integration                                                                               */

static const double rk20 = 3.0 / 4.0, rk21 = 1.0 / 4.0, rk22 = 1.0 / 4.0; /* definition.f90:180-185 */
static const double rk30 = 1.0 / 3.0, rk31 = 2.0 / 3.0, rk32 = 2.0 / 3.0;

static void rk_combine(orc *o, int stage, double dt)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l <= nz; l++)
        for (int k = 0; k <= ny; k++)
            for (int j = 0; j <= nx; j++)
                for (int i = 0; i < NV; i++) {
                    if (stage == 1)
                        CO(i, j, k, l) = UO(i, j, k, l) + dt * LRK(i, j, k, l);
                    else if (stage == 2)
                        CO(i, j, k, l) = rk20 * UO(i, j, k, l) + rk21 * CO(i, j, k, l) + rk22 * dt * LRK(i, j, k, l);
                    else
                        CO(i, j, k, l) = rk30 * UO(i, j, k, l) + rk31 * CO(i, j, k, l) + rk32 * dt * LRK(i, j, k, l);
                }
}

/* rungekutta.f90:12-148 RUNGEKUTTA */
void orc_rungekutta(orc *o, double dt)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l <= nz; l++)
        for (int k = 0; k <= ny; k++)
            for (int j = 0; j <= nx; j++)
                for (int i = 0; i < NV; i++) UO(i, j, k, l) = CO(i, j, k, l);
    for (int stage = 1; stage <= 3; stage++) {
        orc_spatial(o);
        rk_combine(o, stage, dt);
        fromutorve(o);
        custom_checkrho(o);
        if (stage == 3) custom_checkrho(o); /* :130-135, OPERATOR_SPLIT is empty */
        fromrvetou(o);
        orc_boundary(o);
        update(o);
    }
}

/* rungekutta.f90:161-265 finddt (Q8: no discriminant clamp; Q9: min width / max speed) */
double orc_finddt(orc *o)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    const double g = o->c.ggas, cfl = o->c.cfl;
    double dt_temp = 1.0e10;
#pragma omp parallel for collapse(2) schedule(static) reduction(min : dt_temp)
    for (int l = 1; l <= nz; l++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++) {
                double rho = PR(IRHO, j, k, l);
                double cs_loc = eos_soundspeed(g, PR(ITAU, j, k, l), rho);
                double a2 = cs_loc * cs_loc;
                double a4 = a2 * a2;
                double b2x = (BC(0, j, k, l) * BC(0, j, k, l) / rho);
                double b2y = (BC(1, j, k, l) * BC(1, j, k, l) / rho);
                double b2z = (BC(2, j, k, l) * BC(2, j, k, l) / rho);
                double b2 = b2x + b2y + b2z;
                double b4 = b2 * b2;
                double cfx = sqrt(0.5 * (a2 + b2 + sqrt((a4 + 2.0 * a2 * b2 + b4) - 4.0 * a2 * b2x)));
                double cfy = sqrt(0.5 * (a2 + b2 + sqrt((a4 + 2.0 * a2 * b2 + b4) - 4.0 * a2 * b2y)));
                double cfz = sqrt(0.5 * (a2 + b2 + sqrt((a4 + 2.0 * a2 * b2 + b4) - 4.0 * a2 * b2z)));
                double lambda = fmax3(fabs(PR(IVX, j, k, l)) + cfx, fabs(PR(IVY, j, k, l)) + cfy,
                                      fabs(PR(IVZ, j, k, l)) + cfz);
                double dx_loc = cdx(o, j), dy_loc = cdy(o, k), dz_loc = cdz(o, l);
                if (o->c.coordinate == 2) dy_loc = cxc(o, j) * dy_loc;
                if (o->c.coordinate == 3) {
                    dy_loc = cxc(o, j) * dy_loc;
                    dz_loc = cxc(o, j) * sin(cyc(o, k)) * dz_loc;
                }
                double dt_out = fmin3(dx_loc, dy_loc, dz_loc) * cfl / lambda;
                /* a NaN cell (NaN/Inf state) gives 0, so the minimum reports it whatever the
                   reduction order; Fortran's MIN leaves NaN arguments processor-dependent */
                if (dt_out != dt_out) dt_out = 0.0;
                dt_temp = fmin2(dt_out, dt_temp);
            }
    return dt_temp;
}

/* mhd.f90:15-66 find_divb (Q11: signed max) plus max |div B|; coords.f90:173-223 GEOM_AREA */
void orc_find_divb(const orc *o, double *max_signed, double *max_abs)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    double maxdb = 0.0, maxabs = 0.0;
#pragma omp parallel for collapse(2) schedule(static) reduction(max : maxdb, maxabs)
    for (int l = 1; l <= nz; l++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++) {
                double dx_loc = cdx(o, j), dy_loc = cdy(o, k), dz_loc = cdz(o, l);
                double axp, axm, ayp, aym, azp, azm;
                if (o->c.coordinate == 3) {
                    axp = o->xf[j] * o->xf[j] * sph_cd(o, k) * dz_loc;
                    axm = o->xf[j - 1] * o->xf[j - 1] * sph_cd(o, k) * dz_loc;
                    ayp = 0.5 * (pw2(o->xf[j]) - pw2(o->xf[j - 1])) * sin(o->yf[k]) * dz_loc;
                    aym = 0.5 * (pw2(o->xf[j]) - pw2(o->xf[j - 1])) * sin(o->yf[k - 1]) * dz_loc;
                    azp = 0.5 * (pw2(o->xf[j]) - pw2(o->xf[j - 1])) * dy_loc;
                    azm = azp;
                } else if (o->c.coordinate == 2) {
                    axp = o->xf[j] * dy_loc * dz_loc;
                    axm = o->xf[j - 1] * dy_loc * dz_loc;
                    ayp = dx_loc * dz_loc;
                    aym = dx_loc * dz_loc;
                    azp = 0.5 * (o->xf[j] * o->xf[j] - o->xf[j - 1] * o->xf[j - 1]) * dy_loc;
                    azm = azp;
                } else {
                    axp = dy_loc * dz_loc; axm = axp;
                    ayp = dx_loc * dz_loc; aym = ayp;
                    azp = dx_loc * dy_loc; azm = azp;
                }
                double div_b = (axp * PR(IBX, j, k, l) - axm * PR(IBX, j - 1, k, l)) +
                               (ayp * PR(IBY, j, k, l) - aym * PR(IBY, j, k - 1, l)) +
                               (azp * PR(IBZ, j, k, l) - azm * PR(IBZ, j, k, l - 1));
                maxdb = fmax2(maxdb, div_b);
                maxabs = fmax2(maxabs, fabs(div_b));
            }
    *max_signed = maxdb;
    *max_abs = maxabs;
}

/* main.f90:127-163 (Q20: plus a step cap) */
int orc_run(orc *o, int max_steps, double total_time, double *dts, double *t_out)
{
    double global_time = t_out ? *t_out : 0.0;
    int n = 0;
    while (global_time < total_time && n < max_steps) {
        double dt = orc_finddt(o);
        if (global_time + dt > total_time) dt = total_time - global_time;
        orc_rungekutta(o, dt);
        if (dts) dts[n] = dt;
        n++;
        global_time = global_time + dt;
    }
    if (t_out) *t_out = global_time;
    return n;
}
