/*
 * cumc_oracle.h -- CPU restatement of the CUMC3D ideal-MHD time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path (cumpic_amd/csrc) and the timed CPU baseline of bench.py.  Nothing in the
 * product links or calls it; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.
 *
 * Layout follows the reference's Fortran arrays one to one:
 *   prim(1:8, -2:n+3, -2:n+3, -2:n+3)  -> prim[(cell)*8 + v], v fastest
 *   eps(-2:n+3, ...)                   -> eps[cell]
 *   bcell(ibx:ibz, ...)                -> bcell[(cell)*3 + c]
 *   xf(-3:n+3)                         -> xf[j + 3]
 * with cell = ((l+2)*(ny+6) + (k+2))*(nx+6) + (j+2).
 * Variable order: 0 rho, 1 vx, 2 vy, 3 vz, 4 p (itau), 5 bx, 6 by, 7 bz
 * (solver/setup.f90:49-116).
 */
#ifndef CUMC_ORACLE_H
#define CUMC_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_config {
    int nx, ny, nz;            /* interior cells per axis (definition.f90:100-112) */
    int coordinate;            /* 1 cartesian, 2 cylindrical, 3 spherical (definition.f90:65-72) */
    int boundary_flag[6];      /* 0 periodic, 1 outflow, >=2 parity (definition.f90:87) */
    int recon;                 /* 1 TVDMM, 4 PPMC (definition.f90:157-173); 2 TVDMC, 3 TVDVL, 5 WENO (extensions) */
    int model;                 /* 0 riemann plugin, 1 torus_cyn plugin */
    int geom_source_repeat;    /* 8 reproduces spatial.f90:86-97 (source added per i) */
    double ggas;               /* adiabatic index (model defs.f90:14) */
    double cfl;                /* definition.f90:116 */
    double rho_floor, eps_floor, r_sh; /* torus_cyn/src/defs.f90, model.f90:246-247 */
    int solver;                /* 2 (or 0) HLL (definition.f90:143); 1 LF, 3 HLLC, 4 HLLD (extensions) */
} orc_config;

typedef struct orc orc;

orc *orc_create(const orc_config *cfg, const double *xf, const double *yf, const double *zf);
void orc_destroy(orc *o);
long orc_ncell(const orc *o);
/* OpenMP threads of the loops (n <= 0 leaves them); returns the count now in force. */
int orc_set_threads(int n);
/* full ghost-inclusive arrays in the layout above */
void orc_set_state(orc *o, const double *prim, const double *eps);
void orc_get_state(const orc *o, double *prim, double *eps, double *bcell, double *cons);
void orc_get_lrk(const orc *o, double *l_rk);
void orc_initial_update(orc *o);
double orc_finddt(orc *o);
void orc_rungekutta(orc *o, double dt);
void orc_spatial(orc *o);
void orc_boundary(orc *o);
void orc_sweep_flux(orc *o, int dir, double *out);
void orc_find_divb(const orc *o, double *max_signed, double *max_abs);
/* test instrumentation: cells whose density / eps CUSTOM_CHECKRHO raised to a floor so far */
void orc_floor_counts(const orc *o, long *rho, long *eps);
/* main.f90:127-163 with a step cap: returns steps taken, fills dts (may be NULL) */
int orc_run(orc *o, int max_steps, double total_time, double *dts, double *t_out);

/* cell-level entry points used by the closed-form tests */
void orc_ppm(const double d[5], const double v[5], double *vm_out, double *vp_out);
void orc_tvdmm(const double d[3], const double v[3], double *vm_out, double *vp_out);
/* WENO5 (Jiang-Shu) edge values of the middle cell of v[5] (extension) */
void orc_weno5(const double v[5], double *vm_out, double *vp_out);
/* lim 1 minmod (= orc_tvdmm), 2 MC, 3 van Leer */
void orc_tvd(int lim, const double d[3], const double v[3], double *vm_out, double *vp_out);
double orc_signalspeed(double cs, double bn, double bt1, double bt2, double rho);
/* dir 1,2,3; pl/pr: 8 face primitives; flux: 8 outputs; returns branch 0 L, 1 HLL, 2 R */
int orc_hll(int dir, double ggas, const double *pl, const double *pr, double *flux);
/* local Lax-Friedrichs; HLLD (returns the wave region 0..5, 6 for NaN speeds) */
int orc_lf(int dir, double ggas, const double *pl, const double *pr, double *flux);
int orc_hlld(int dir, double ggas, const double *pl, const double *pr, double *flux);
/* HLLC for MHD (Li 2005; returns the region 0 F_L, 1 F*_L, 2 F*_R, 3 F_R, 4 NaN) */
int orc_hllc(int dir, double ggas, const double *pl, const double *pr, double *flux);
void orc_ptof(int dir, const double *p, const double *u, double *f);
void orc_ptou_face(const double *p, double eps, double *u);
void orc_ptou_cell(const double *p, double eps, const double *bc, double *u);
void orc_utop_cell(const double *u, const double *bc, double *p, double *eps);

#ifdef __cplusplus
}
#endif
#endif
