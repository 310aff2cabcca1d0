/*
 * cumc.h -- C ABI of the MI355X CUMC3D MHD time step (libcumc.so).
 *
 * The reference drives its step through argument-free Fortran subroutines that
 * operate on MODULE DEFINITION globals.  Each entry point below replaces one of
 * them; the comment names the reference routine (file:line).  All arrays crossing
 * this boundary are caller-owned host arrays in the reference's Fortran layout:
 *
 *   prim (1:8, -2:nx+3, -2:ny+3, -2:nz+3)   rho, vx, vy, vz, p, Bx_face, By_face, Bz_face
 *   eps  (-2:nx+3, -2:ny+3, -2:nz+3)
 *   bcell(1:3, -2:nx+3, -2:ny+3, -2:nz+3)
 *   xf(-3:nx+3), yf(-3:ny+3), zf(-3:nz+3)   face coordinates
 *
 * (first index fastest).  Device storage is private.  Every call returns 0 or a
 * CUMC_E* code and is synchronous on return; one HIP stream per context.
 */
#ifndef CUMC_H
#define CUMC_H
#ifdef __cplusplus
extern "C" {
#endif

enum {
    CUMC_OK = 0,
    CUMC_EINVAL = 1,      /* bad configuration or argument */
    CUMC_EHIP = 2,        /* HIP runtime error */
    CUMC_ENONFINITE = 3,  /* non-finite or non-positive time step */
    CUMC_ENOMEM = 4,      /* device allocation failed */
    CUMC_ESTATE = 5       /* call out of order (e.g. no grid set) */
};

enum { CUMC_CARTESIAN = 1, CUMC_CYLINDRICAL = 2, CUMC_SPHERICAL = 3 };  /* solver/definition.f90:65-67 */
enum { CUMC_TVDMM = 1, CUMC_PPMC = 4 };                  /* solver/definition.f90:158,167 */
/* declared by solver/definition.f90:161,164,170 but not implemented by the reference; extensions */
enum { CUMC_TVDMC = 2, CUMC_TVDVL = 3, CUMC_WENO = 5 };
/* Riemann solvers: HLL is the reference's (rsolver.f90:43-130); LF, HLLC and HLLD are declared by
 * definition.f90:140,146,149 without an implementation there, and are extensions here */
enum { CUMC_LF = 1, CUMC_HLL = 2, CUMC_HLLC = 3, CUMC_HLLD = 4 };
enum { CUMC_MODEL_RIEMANN = 0, CUMC_MODEL_TORUS_CYN = 1, CUMC_MODEL_TORUS_SPH = 2 };
enum { CUMC_FP64 = 0, CUMC_MIXED = 1 };

/* Runtime form of the compile-time PARAMETERs of solver/definition.f90:29-185 and the
 * model plugin's defs.f90 (ggas, rho_floor, eps_floor, r_sh). */
typedef struct cumc_config {
    int nx, ny, nz;               /* interior cells of this subdomain (definition.f90:110-112) */
    int coordinate;               /* CUMC_CARTESIAN | CUMC_CYLINDRICAL | CUMC_SPHERICAL (definition.f90:72) */
    int boundary_flag[6];         /* x-in, x-out, y-in, y-out, z-in, z-out (definition.f90:87) */
    int recon;                    /* CUMC_PPMC | CUMC_TVDMM (definition.f90:173) | CUMC_TVDMC | CUMC_TVDVL | CUMC_WENO */
    int model;                    /* which model plugin's in-loop hooks to run: CUMC_MODEL_* */
    int geom_source_repeat;       /* 8 reproduces spatial.f90:86-97; 1 adds the source once */
    double ggas;                  /* adiabatic index */
    double cfl;                   /* definition.f90:116 */
    double rho_floor, eps_floor;  /* torus_cyn floors (model.f90:246-247) */
    double r_sh;                  /* torus_cyn Schwarzschild radius (defs.f90:23) */
    int solver;                   /* CUMC_HLL (or 0) | CUMC_LF | CUMC_HLLC | CUMC_HLLD (definition.f90:152) */
    int precision;                /* CUMC_FP64 | CUMC_MIXED */
} cumc_config;

typedef struct cumc_ctx cumc_ctx;

/* BUILD_HYDRO (solver/setup.f90:249-312) + SETUP_EQNS bfac tables (setup.f90:121-230) +
 * device residency of POPULATE_DEVICE (solver/openacc.f90:28-59).  The configuration is checked
 * on the host before any HIP call: CUMC_EINVAL (and *out = NULL) for sizes below 1 or a plane of
 * 2^29 cells or more with ghosts, an unknown option, gamma <= 1, cfl <= 0, or a boundary flag
 * outside 0..4 or one the reference stops on. */
int cumc_create(const cumc_config *cfg, cumc_ctx **out);
int cumc_destroy(cumc_ctx *ctx);

/* GETGRID / CUSTOM_GRID result (solver/coords.f90:11-41): face coordinates incl. ghosts. */
int cumc_set_grid(cumc_ctx *ctx, const double *xf, const double *yf, const double *zf);

/* POPULATE_DEVICE copy-in of prim and eps (solver/openacc.f90:35-50). */
int cumc_upload(cumc_ctx *ctx, const double *prim, const double *eps);

/* !$ACC UPDATE HOST(prim, eps) of print_hydroprofile (solver/dataio.f90:24); bcell may be NULL. */
int cumc_download(cumc_ctx *ctx, double *prim, double *eps, double *bcell);

/* cumc_download without any device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  The conversion to the host layout is staged through a fixed 32 MiB device
 * buffer that the library holds per GPU for the life of the process, slab by slab of whole planes;
 * one plane of each requested array must fit ((nx+6)*(ny+6)*8 doubles <= 32 MiB with prim, else
 * CUMC_EINVAL).  Call it with the context's device current.  Calls on one GPU share the buffer:
 * calls that overlap in time on the same GPU (from several host threads) race. */
int cumc_download_noalloc(cumc_ctx *ctx, double *prim, double *eps, double *bcell);

/* initial_update (solver/initial.f90:83-140): CUSTOM_CHECKRHO, FROMRVETOU, BOUNDARY, UPDATE(0). */
int cumc_initial_update(cumc_ctx *ctx);

/* finddt (solver/rungekutta.f90:161-265): CFL time step of the current state. */
int cumc_finddt(cumc_ctx *ctx, double *dt);

/* RUNGEKUTTA (solver/rungekutta.f90:12-148): one SSP-RK3 step of size dt. */
int cumc_rungekutta(cumc_ctx *ctx, double dt);

/* The time loop of solver/main.f90:127-163 (finddt, clamp to total_time, RUNGEKUTTA) with a
 * step cap; *time is read and advanced; dts (may be NULL) receives max_steps entries. */
int cumc_run(cumc_ctx *ctx, int max_steps, double total_time, double *time, int *steps_done, double *dts);

/* find_divb (solver/mhd.f90:15-66): reference signed maximum and the maximum |div B|. */
int cumc_find_divb(cumc_ctx *ctx, double *max_signed, double *max_abs);

/* BOUNDARY (solver/boundary.f90:5-13) followed by UPDATE's EOS on the ghost shell. */
int cumc_boundary(cumc_ctx *ctx);

/* Face fluxes of one sweep for the current state (GET_FLUXES reconstruction + RIEMANN,
 * solver/spatial.f90:396-433): out has 7 ghost-inclusive cell arrays (rho, mx, my, mz, E and the
 * two transverse B fluxes in x,y,z order), entry (j,k,l) holding the face at the high side of
 * cell j (dir 1), k (dir 2) or l (dir 3). */
int cumc_sweep_flux(cumc_ctx *ctx, int dir, double *out);

/* Wall time of the last cumc_rungekutta / cumc_run, measured with HIP events on the context
 * stream. */
int cumc_last_step_ms(cumc_ctx *ctx, double *ms);

/* Per-kernel-class timing with HIP events recorded on the context stream around every launch
 * group (no reference counterpart; the reference has no timers).  cumc_profile(ctx, 1) clears
 * and enables it; cumc_kernel_times fills up to n entries of total milliseconds, launch counts
 * and class names, and returns the number of classes. */
int cumc_profile(cumc_ctx *ctx, int on);
int cumc_kernel_times(cumc_ctx *ctx, int n, double *total_ms, long *launches, const char **names);

/* Diagnostic, no reference counterpart: a plain device-to-device copy of `bytes` bytes run `reps`
 * times; reports the copy bandwidth (read + write bytes per second, in GB/s).  It is the
 * known-size kernel against which HBM traffic counters are calibrated. */
int cumc_copy_probe(long bytes, int reps, double *gbps);

/* Diagnostic, no reference counterpart: the stage tail's memory access pattern on this context's
 * layout and tuned column tiles, with no arithmetic -- `nread` FP64 planes read and `nwrite`
 * written per cell of [0,n]^3, run `reps` times; reports (nread + nwrite) * 8 bytes per cell per
 * second in GB/s, the bandwidth ceiling of a kernel with that access pattern.  Supported
 * (nread, nwrite): (26, 20) and (34, 20) (the FP64 tail of stage 1 and of stages 2-3), (34, 0)
 * and (0, 20). */
int cumc_pattern_probe(cumc_ctx *ctx, int nread, int nwrite, int reps, double *gbps);

const char *cumc_error_string(int code);
const char *cumc_last_error(cumc_ctx *ctx);

/* ------------------------------------------------------------------------------------------
 * Launch-shape tuning.  No reference counterpart (OpenACC picks gangs/vectors itself); the
 * shapes change no result bit.
 * ------------------------------------------------------------------------------------------ */
enum { CUMC_PH_FLUX_X = 0, CUMC_PH_FLUX_Y, CUMC_PH_FLUX_Z, CUMC_PH_ECORN, CUMC_PH_TAIL, CUMC_PH_GHOST,
       CUMC_PH_FINDDT, CUMC_PH_N };
typedef struct cumc_tuning {
    int seg;                 /* target faces per pencil segment of the y and z sweeps */
    int zchunk;              /* planes per column tile of the EMF and stage-tail kernels */
    int blocks[CUMC_PH_N];   /* batch only: persistent workgroups per phase, 0 = one per tile */
    int tile_w;              /* column tile of the EMF / stage-tail kernels: 0 = flattened rows of 256
                                cells, or 32 / 64 / 128 = tile_w x (256 / tile_w) columns */
} cumc_tuning;

/* Time candidate shapes on the context's current state (which is saved and restored) and keep
 * the fastest; *out (may be NULL) receives the choice. */
int cumc_autotune(cumc_ctx *ctx, cumc_tuning *out);
int cumc_get_tuning(cumc_ctx *ctx, cumc_tuning *out);
int cumc_set_tuning(cumc_ctx *ctx, const cumc_tuning *t);

/* z-collapsed stepping of a 2D subdomain (nz = 1, periodic or outflow z).  No reference
 * counterpart: the reference fills and computes the three z ghost layers on each side of the
 * single cell (boundary.f90:564-652, spatial.f90:552-589), which for periodic or outflow z are
 * copies of it.  An FP64 context with the riemann hooks and equal z cell widths steps its one
 * interior plane instead,
 * with every z neighbour read from the plane itself, whenever its z ghost planes are the copies a
 * ghost fill wrote and its two periodic B_z faces agree bit for bit; results are bit-identical to
 * the full layout (the z ghost planes are rewritten before any call that reads them).  On by
 * default; cumc_set_zcollapse(ctx, 0) keeps the full layout.  cumc_zcollapse_state reports
 * whether the context may collapse and whether its state is collapsed now. */
int cumc_set_zcollapse(cumc_ctx *ctx, int on);
int cumc_zcollapse_state(cumc_ctx *ctx, int *allowed, int *active);

/* ------------------------------------------------------------------------------------------
 * Batched form: many independent subdomains of uneven sizes advanced together.  No reference
 * counterpart: the reference runs one problem per process (solver/main.f90:127-163).  Every
 * phase of the step is one persistent launch over a work list of (box, tile) items covering all
 * boxes; each box keeps its own CFL time step on the device, and a run can be captured in a
 * hipGraph and replayed.  Boxes must share coordinate, recon and FP64 precision and use the
 * riemann model hooks; each box is an ordinary context (set_grid / upload / initial_update /
 * download / find_divb stay per box) and must be idle while the batch runs.
 * ------------------------------------------------------------------------------------------ */
typedef struct cumc_batch cumc_batch;

int cumc_batch_create(cumc_ctx *const *boxes, int nbox, cumc_batch **out);
int cumc_batch_destroy(cumc_batch *b);

/* nsteps RUNGEKUTTA steps of every box.  dt_in (nsteps*nbox, step-major) freezes the time
 * steps; NULL means each box's own finddt before every step (solver/main.f90:130, no
 * total_time clamp).  dt_out (may be NULL) receives the steps used.  graph_steps > 0 captures
 * that many steps in one hipGraph and replays it (the remainder runs as direct launches);
 * 0 launches every kernel directly. */
int cumc_batch_run(cumc_batch *b, int nsteps, int graph_steps, const double *dt_in, double *dt_out);

/* Time candidate persistent grid sizes and segment lengths on the boxes' current states (saved
 * and restored) and keep the fastest; *out (may be NULL) receives the choice. */
int cumc_batch_autotune(cumc_batch *b, cumc_tuning *out);
int cumc_batch_get_tuning(cumc_batch *b, cumc_tuning *out);
int cumc_batch_set_tuning(cumc_batch *b, const cumc_tuning *t);
int cumc_batch_last_ms(cumc_batch *b, double *ms);
const char *cumc_batch_last_error(cumc_batch *b);

/* Branch signature of every box (nbox values into sig) accumulated by the batched forward
 * kernels since the previous call, which it then clears.  No reference counterpart: a
 * diagnostic of the finite-difference checks of the reverse mode.  Each data-dependent decision
 * of the reference path (the PPM / TVD-minmod limiter switches of reconstruct.f90:58-128 and
 * :254-307, the HLL speed selections and region of rsolver.f90:110-126, the discriminant clamp
 * of definition.f90:394-410, the floors of custom.f90:66-68) enters a hash, so two runs of a box
 * with equal signatures took the same branches.  Recorded only by the diagnostic build of the
 * library (libcumc_sig.so, compiled with CUMC_BRANCH_SIG); libcumc.so returns CUMC_EINVAL. */
int cumc_batch_branch_sig(cumc_batch *b, unsigned long long *sig);

/* ------------------------------------------------------------------------------------------
 * Reverse mode.  No reference counterpart: the reference has no derivatives.  The taped map is
 * initial_update (solver/initial.f90:83-140) followed by nsteps RUNGEKUTTA steps
 * (solver/rungekutta.f90:12-148) with the time steps frozen; limiter, Riemann-region and floor
 * branches are replayed from the forward values.  FP64, Cartesian, riemann model hooks.
 * ------------------------------------------------------------------------------------------ */

/* From the state as uploaded (not yet initial_update'd): run initial_update and nsteps steps,
 * storing the state every `every` steps on the device, and every stage's state in the last
 * `every`-step segment.  Device memory: about 20 * (nsteps/every + 3*every + 8) state planes.
 * dt_in (nsteps) freezes the steps; NULL takes each step's finddt.  dt_out (may be NULL)
 * receives the steps.  Leaves the context at the final state. */
int cumc_tape_run(cumc_ctx *ctx, int nsteps, int every, const double *dt_in, double *dt_out);

/* Adjoint of the taped map.  Seeds: adjoints of the final prim / eps / bcell (same host layouts
 * as cumc_download; NULL = zero).  Outputs: gradients with respect to the uploaded prim / eps
 * (cumc_upload layouts; NULL = not wanted).  Recomputes each `every`-step segment but the last
 * from its checkpoint (every >= nsteps: nothing is recomputed); leaves the context at the taped
 * final state. */
int cumc_tape_backward(cumc_ctx *ctx, const double *seed_prim, const double *seed_eps, const double *seed_bcell,
                       double *grad_prim, double *grad_eps);
int cumc_tape_free(cumc_ctx *ctx);

/* The same for every box of a batch at once (boxes as uploaded, not yet initial_update'd):
 * dt_in / dt_out are nsteps x nbox, step-major.  Each reverse phase is one launch over all boxes;
 * a box's gradient equals its cumc_tape_backward gradient bit for bit.  seed_bcell and grad_prim
 * are arrays of nbox host pointers (cumc_download / cumc_upload layouts; an array or an entry may
 * be NULL).  cumc_batch_last_ms afterwards gives the device time of the backward pass.  The
 * batched backward re-runs each checkpoint segment but the last two (whose stage states the
 * taping forward keeps in the two sets of segment buffers that the overlapped re-runs use). */
int cumc_batch_tape_run(cumc_batch *b, int nsteps, int every, const double *dt_in, double *dt_out);
int cumc_batch_tape_backward(cumc_batch *b, const double *const *seed_bcell, double *const *grad_prim);
int cumc_batch_tape_free(cumc_batch *b);
/* Diagnostic, no reference counterpart: device time of the last backward pass spent re-running
 * each checkpoint segment forward (storing every stage's state) before reversing it; the rest
 * of cumc_batch_last_ms is the reverse sweep proper. */
int cumc_batch_tape_recompute_ms(cumc_batch *b, double *ms);

#ifdef __cplusplus
}
#endif
#endif
