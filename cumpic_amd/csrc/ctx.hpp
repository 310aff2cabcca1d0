// ctx.hpp -- context of one subdomain and the stage launch sequence, shared by the FP64
// (cumc.hip) and mixed-precision (cumc_mixed.hip) translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/cumc.h"
#include "kernels.hpp"

using namespace cumc;

struct cumc_ctx {
    cumc_config cfg;
    int device = 0;  // HIP device current at cumc_create (the stream's device)
    Lay L;
    Model M;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double *pool = nullptr;  // all FP64 fields
    size_t pool_bytes = 0;
    void *fluxpool = nullptr;
    Fields<double> F64;
    Fields<float> F32;
    double *tab[3] = {nullptr, nullptr, nullptr};
    float *tabf[3] = {nullptr, nullptr, nullptr};
    double *dt_dev = nullptr;           // [0] dt of the current step, [1] dt found by stage 3
    unsigned long long *red = nullptr;  // reduction slots
    double *Ub = nullptr;               // third conserved buffer (stage 2 output)
    bool dt_valid = false;              // dt_dev[1] holds finddt of the current state
    bool p64_stale = false;             // mixed form: P / EPS / BC lag the FP32 planes and U0
    bool grid_set = false;
    // z-collapsed stepping of a periodic nz = 1 context (make_view): zc_ok when the configuration
    // allows it (FP64, riemann hooks, periodic z, z tables equal in every cell); zc_on the
    // caller's switch (cumc_set_zcollapse); zfresh when the z ghost planes hold the copies of
    // plane l = 1 the last full-layout ghost fill wrote; zmode while the state lives in the
    // collapsed view (plane 1 current, z ghost planes not yet rewritten: zexpand restores them)
    bool zc_ok = false, zc_on = true, zfresh = false, zmode = false;
    bool zbad = false;   // the entry check found unequal periodic B_z faces (until the next upload)
    bool ztape = false;  // a reverse-mode record holds full-layout states: no collapsed steps
    struct cumc_tape *tape = nullptr;  // reverse-mode record (cumc_adjoint.hip)
    int seg = 24;    // target faces per pencil segment of the y / z sweeps
    int zchunk = 4;  // planes per column tile of the EMF and stage-tail kernels
    int tile_w = 0;   // column tile width of those kernels (0: flattened rows; see col_of)
    double last_ms = 0.0;
    std::string err;
    // per-kernel-class HIP event timing (cumc_profile / cumc_kernel_times)
    bool prof = false;
    std::vector<hipEvent_t> evpool;
    std::vector<int> evcls;
    size_t evn = 0;
    double kc_ms[16] = {0};
    long kc_count[16] = {0};
};

enum { KC_FLUX_X, KC_FLUX_Y, KC_FLUX_Z, KC_ECORN, KC_TAIL, KC_GHOST, KC_FINDDT, KC_N };
static const char *kc_names[KC_N] = {"flux_x", "flux_y", "flux_z", "ecorn", "tail", "ghost", "finddt"};

// record the start (cls >= 0) or end (cls < 0) of a kernel class on the context stream
static inline void mark(cumc_ctx *ctx, int cls)
{
    if (!ctx->prof) return;
    if (ctx->evn == ctx->evpool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        ctx->evpool.push_back(e);
        ctx->evcls.push_back(0);
    }
    ctx->evcls[ctx->evn] = cls;
    hipEventRecord(ctx->evpool[ctx->evn], ctx->stream);
    ctx->evn++;
}

// fold recorded event pairs into the per-class totals (stream must be idle)
static inline void collect(cumc_ctx *ctx)
{
    for (size_t i = 0; i + 1 < ctx->evn; i += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->evpool[i], ctx->evpool[i + 1]) == hipSuccess) {
            const int c = ctx->evcls[i];
            ctx->kc_ms[c] += ms;
            ctx->kc_count[c] += 1;
        }
    }
    ctx->evn = 0;
}

static const char *g_err_names[] = {"ok", "invalid argument", "HIP error", "non-finite time step",
                                    "out of device memory", "call out of order"};

#define HIPCHK(ctx, call)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                    \
            return CUMC_EHIP;                                                                   \
        }                                                                                       \
    } while (0)

static inline int fail(cumc_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

static inline Tabs tabs_of(const cumc_ctx *ctx)
{
    Tabs tb;
    tb.t[0] = ctx->tab[0]; tb.t[1] = ctx->tab[1]; tb.t[2] = ctx->tab[2];
    tb.tf[0] = ctx->tabf[0]; tb.tf[1] = ctx->tabf[1]; tb.tf[2] = ctx->tabf[2];
    tb.len[0] = ctx->L.nx + 7; tb.len[1] = ctx->L.ny + 7; tb.len[2] = ctx->L.nz + 7;
    return tb;
}

static inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }
// 1D launch size for a gx x gy tile grid remapped by xcd_tile (a multiple of the XCD count)
static inline unsigned xcd_blocks(int gx, int gy) { return (unsigned)(((long)gx * gy + NXCD - 1) / NXCD * NXCD); }

// the step result (stage 3 output) becomes the next step's starting state
static inline void swap_u(cumc_ctx *ctx)
{
    double *t = ctx->F64.U0;
    ctx->F64.U0 = ctx->F64.U1;
    ctx->F64.U1 = t;
    ctx->F32.U0 = ctx->F64.U0;
    ctx->F32.U1 = ctx->F64.U1;
}

// frees the reverse-mode record (defined in cumc_adjoint.hip)
void tape_release(cumc_ctx *ctx);

// batch internals used by the batched reverse mode (defined in cumc_batch.hip)
struct batch_tape;
void batch_tape_release(cumc_batch *b);  // defined in cumc_adjoint.hip
batch_tape *&batch_adj(cumc_batch *b);
double &batch_last_ms(cumc_batch *b);
int batch_nbox(const cumc_batch *b);
cumc_ctx *batch_box(const cumc_batch *b, int i);
hipStream_t batch_stream(const cumc_batch *b);
hipStream_t batch_swap_stream(cumc_batch *b, hipStream_t s);
std::string &batch_err(cumc_batch *b);
int batch_refresh(cumc_batch *b);
void batch_freeze_views(cumc_batch *b, bool on);
int batch_reserve_hist(cumc_batch *b, int nsteps);
double *batch_hist(cumc_batch *b);
void batch_set_counter(cumc_batch *b, int v);
void batch_finddt(cumc_batch *b);
void batch_step(cumc_batch *b, int par, int frozen);
void batch_stage(cumc_batch *b, int st, int par);
void batch_set_dt(cumc_batch *b, int m);
void batch_dt_begin(cumc_batch *b, int frozen);
void batch_after_run(cumc_batch *b, int nsteps);

// ghost passes + torus clamp on the FP64 primitive planes (defined in cumc.hip) and on the
// mixed form's FP32 planes (cumc_mixed.hip)
void launch_ghost(cumc_ctx *ctx);
void launch_ghost32(cumc_ctx *ctx);

template <typename R, int RECON, int SOLVER>
void launch_fluxes_t(cumc_ctx *ctx, const Lay &L, const Fields<R> &F, int only_dir)
{
    hipStream_t s = ctx->stream;
    Tabs tb = tabs_of(ctx);
    const Model &M = ctx->M;
    const int npl = flux_planes(L);
    if (!only_dir || only_dir == 1) {
        mark(ctx, KC_FLUX_X);
        const long waves = ((long)(L.ny + 2) * L.px + XS_FACES - 1) / XS_FACES;
        dim3 grid(nblk(waves * 64, 256), npl);
        hipLaunchKernelGGL((k_flux_x<R, RECON, SOLVER>), grid, dim3(256), 0, s, L, F, tb, M);
        mark(ctx, -1);
    }
    const int seg = ctx->seg;
    if (!only_dir || only_dir == 2) {
        mark(ctx, KC_FLUX_Y);
        const int nseg = (L.ny + 1 + seg - 1) / seg;
        if ((long)L.px * npl * nseg < YLANE_WALKERS && L.ny >= 16) {
            // a 2D box whose pencil walks would leave most SIMDs idle: wave-parallel y sweep
            const int gx = ylane_groups(L) * ylane_segs(L);
            hipLaunchKernelGGL((k_flux_ylane<R, RECON, SOLVER>), dim3(xcd_blocks(gx, npl)), dim3(256), 0, s, L, F, tb, M,
                               gx, npl);
        } else {
            const int sl = (L.ny + 1 + nseg - 1) / nseg;
            dim3 grid(nblk((long)L.px * npl, 128), nseg);
            hipLaunchKernelGGL((k_flux_walk<R, 2, RECON, SOLVER>), grid, dim3(128), 0, s, L, F, tb, M, sl);
        }
        mark(ctx, -1);
    }
    if (!only_dir || only_dir == 3) {
        mark(ctx, KC_FLUX_Z);
        // a z-collapsed view has one z face per column: faces 0 and 1 of the single cell are the
        // same face of its periodic images
        const int nseg = L.zc1 ? 1 : (L.nz + 1 + seg - 1) / seg;
        const int sl = L.zc1 ? 1 : (L.nz + 1 + nseg - 1) / nseg;
        dim3 grid(nblk((long)L.px * (L.ny + 2), 128), nseg);
        hipLaunchKernelGGL((k_flux_walk<R, 3, RECON, SOLVER>), grid, dim3(128), 0, s, L, F, tb, M, sl);
        mark(ctx, -1);
    }
}

template <typename R, int SOLVER>
void launch_fluxes_s(cumc_ctx *ctx, const Lay &L, const Fields<R> &F, int only_dir)
{
    switch (ctx->M.recon) {
    case CUMC_TVDMM: launch_fluxes_t<R, 1, SOLVER>(ctx, L, F, only_dir); break;
    case CUMC_TVDMC: launch_fluxes_t<R, 2, SOLVER>(ctx, L, F, only_dir); break;
    case CUMC_TVDVL: launch_fluxes_t<R, 3, SOLVER>(ctx, L, F, only_dir); break;
    case CUMC_WENO: launch_fluxes_t<R, 5, SOLVER>(ctx, L, F, only_dir); break;
    default: launch_fluxes_t<R, 4, SOLVER>(ctx, L, F, only_dir); break;
    }
}

// the three sweeps (or one, only_dir = 1..3) with the configured reconstruction and solver
template <typename R>
void launch_fluxes(cumc_ctx *ctx, const Lay &L, const Fields<R> &F, int only_dir = 0)
{
    switch (ctx->M.solver) {
    case CUMC_LF: launch_fluxes_s<R, 1>(ctx, L, F, only_dir); break;
    case CUMC_HLLC: launch_fluxes_s<R, 3>(ctx, L, F, only_dir); break;
    case CUMC_HLLD: launch_fluxes_s<R, 4>(ctx, L, F, only_dir); break;
    default: launch_fluxes_s<R, 2>(ctx, L, F, only_dir); break;
    }
}
template <typename R>
void launch_fluxes(cumc_ctx *ctx, const Fields<R> &F, int only_dir = 0)
{
    launch_fluxes<R>(ctx, ctx->L, F, only_dir);
}

// comb: stage 1 writes, and stage 2 reads, the combined stage-2 input of a plain forward step
// (rk_w2, comb_out in kernels.hpp); the mixed form combines only without sources and model hooks
template <typename R, int STAGE, int COORD>
void launch_tail_c(cumc_ctx *ctx, const Lay &L, const Fields<R> &F, const double *Uin, double *Uout, bool comb)
{
    const int zc = ctx->zchunk, tw = ctx->tile_w;
    const int gx = col_tiles(L, tw), gy = zchunks(L, zc);
    dim3 grid(xcd_blocks(gx, gy));
    Tabs tb = tabs_of(ctx);
    unsigned long long *dtn = (unsigned long long *)(ctx->dt_dev + 1);
    if constexpr (STAGE < 3 && (sizeof(R) == 8 || COORD == 1))
        if (comb) {
            hipLaunchKernelGGL((k_tail<R, STAGE, COORD, true>), grid, dim3(256), 0, ctx->stream, L, F, tb, ctx->M,
                               ctx->dt_dev, Uin, Uout, dtn, gx, gy, zc, tw);
            return;
        }
    hipLaunchKernelGGL((k_tail<R, STAGE, COORD, false>), grid, dim3(256), 0, ctx->stream, L, F, tb, ctx->M, ctx->dt_dev,
                       Uin, Uout, dtn, gx, gy, zc, tw);
}
template <typename R, int STAGE>
void launch_tail(cumc_ctx *ctx, const Lay &L, const Fields<R> &F, const double *Uin, double *Uout, bool comb)
{
    if (ctx->M.coordinate == CUMC_SPHERICAL) launch_tail_c<R, STAGE, 3>(ctx, L, F, Uin, Uout, comb);
    else if (ctx->M.coordinate == CUMC_CYLINDRICAL) launch_tail_c<R, STAGE, 2>(ctx, L, F, Uin, Uout, comb);
    else launch_tail_c<R, STAGE, 1>(ctx, L, F, Uin, Uout, comb);
}

// The FP64 fields as the kernels address them: the full layout, or (col) the z-collapsed view of
// a periodic nz = 1 context, whose field bases sit at plane l = 1 and whose z stride is 0.
struct View {
    Lay L;
    Fields<double> F;
    double *Ub;
};
static inline View make_view(const cumc_ctx *ctx, bool col)
{
    View v{ctx->L, ctx->F64, ctx->Ub};
    if (col) {
        const long off = 3 * ctx->L.py;  // cell (j, k, 1) = (1 + 2) * py + ...
        v.L.py = 0;
        v.L.zc1 = 1;
        v.F.P += off;
        v.F.EPS += off;
        v.F.BC += off;
        v.F.U0 += off;
        v.F.U1 += off;
        for (int d = 0; d < 3; d++) v.F.F[d] += off;
        v.F.E += off;
        v.Ub += off;
    }
    return v;
}

// ghost passes on the FP64 primitive planes of a view (defined in cumc.hip)
void launch_ghost_l(cumc_ctx *ctx, const Lay &L, const Fields<double> &F);
// whether the next step runs in the z-collapsed view (enters it when allowed and the z ghost
// planes check out), and the return of a collapsed state to the full layout (cumc.hip)
bool zcollapse_begin(cumc_ctx *ctx);
void zexpand(cumc_ctx *ctx);

// one RK stage: fluxes of the three sweeps, corner EMFs, stage tail, ghost fill, on the fields F
// addressed by L (Ub: the stage-2 output buffer in the same addressing).
// Conserved buffers: stage 1 U0 -> U1, stage 2 U1 -> Ub, stage 3 Ub -> U1; with comb (a plain
// forward step) U1 holds the combined stage-2 input between stages 1 and 2.
template <typename R>
void stage_l(cumc_ctx *ctx, const Lay &L, const Fields<R> &F, double *Ub, int st, bool comb = false)
{
    launch_fluxes<R>(ctx, L, F);
    mark(ctx, KC_ECORN);
    const int zc = ctx->zchunk, tw = ctx->tile_w;
    const int gx = col_tiles(L, tw), gy = zchunks(L, zc);
    hipLaunchKernelGGL((k_ecorn<R>), dim3(xcd_blocks(gx, gy)), dim3(256), 0, ctx->stream, L, F, gx, gy, zc, tw);
    mark(ctx, -1);
    mark(ctx, KC_TAIL);
    if (st == 1) {
        launch_tail<R, 1>(ctx, L, F, F.U0, F.U1, comb);
    } else if (st == 2) {
        launch_tail<R, 2>(ctx, L, F, F.U1, Ub, comb);
    } else {
        const double init = 1.0e10;
        unsigned long long bits;
        std::memcpy(&bits, &init, 8);
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, (unsigned long long *)(ctx->dt_dev + 1), bits, 1);
        launch_tail<R, 3>(ctx, L, F, Ub, F.U1, comb);
    }
    mark(ctx, -1);
    mark(ctx, KC_GHOST);
    if constexpr (sizeof(R) == 4) {
        launch_ghost32(ctx);
        ctx->p64_stale = true;
    } else {
        launch_ghost_l(ctx, L, F);
    }
    mark(ctx, -1);
}

template <typename R>
void stage(cumc_ctx *ctx, const Fields<R> &F, int st, bool comb = false)
{
    stage_l<R>(ctx, ctx->L, F, ctx->Ub, st, comb);
}
