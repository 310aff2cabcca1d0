// cumc_batch.hip -- batched form: many independent subdomains of uneven sizes advanced by one
// persistent launch per phase, optionally captured in a hipGraph.
//
// Each phase kernel walks a work list of (box, tile) items that covers every box: item i
// belongs to the box b with pre[b] <= i < pre[b+1] (prefix sums of the per-box tile counts),
// and a grid of G workgroups strides over the items (G is a tuning choice; G = number of
// items gives one workgroup per tile).  The tile bodies are the ones the single-box kernels run
// (kernels.hpp), compiled with the same FP64 flags, so a box advanced in a batch is bit-identical
// to the same box advanced alone.
//
// Per-box time step: every box keeps [0] current and [1] next dt in its own device slots.  A
// one-workgroup kernel at the start of each step moves next -> current (or takes the frozen
// value) and appends it to a device history, so no host round trip sits inside a step and a
// multi-step run can be replayed from one graph.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.hpp"

namespace {

struct BoxDesc {
    Lay L;
    Fields<double> F;  // U0 / U1 as of the start of the run (steps alternate them by parity)
    Tabs tb;
    Model M;
    double *dt;  // [0] current step, [1] next step (stage-3 reduction target)
    double *Ub;
    int fx_gx;             // x sweep: tiles per plane
    int fy_gx, fy_sl;      // y sweep: tiles per segment row, faces per segment
    int fz_gx, fz_sl;      // z sweep
    int c_gx, zc, tw;      // EMF / stage tail column tiles
    unsigned long long *sig;  // branch signature of the box (diagnostic build only)
};

enum { P_FX, P_FY, P_FZ, P_EC, P_TAIL, P_GX, P_GY, P_GZ, P_DT, P_N };

// largest b in [0, nbox) with pre[b] <= it  (pre[0] = 0 <= it < pre[nbox])
__device__ inline int box_of(const int *pre, int nbox, int it)
{
    int lo = 0, hi = nbox;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= it) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ inline Fields<double> fields_at(const BoxDesc &d, int par)
{
    Fields<double> F = d.F;
    if (par) {
        F.U0 = d.F.U1;
        F.U1 = d.F.U0;
    }
    return F;
}

#ifdef CUMC_BRANCH_SIG
__device__ inline unsigned long long br_mix(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// Adds this thread's decision hash, salted with its work item and thread index, to the box's
// signature (one atomic per wavefront) and restarts the hash.  Every lane of the wavefront calls it.
__device__ inline void br_flush(unsigned long long *sig, int it)
{
    unsigned long long h =
        br_mix(br_slot[threadIdx.x] + 0x9E3779B97F4A7C15ull * ((unsigned long long)it * 256u + threadIdx.x + 1u));
    br_slot[threadIdx.x] = BR_SEED;
    for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(sig, h);
}
#define BR_START() (br_slot[threadIdx.x] = BR_SEED)
#define BR_FLUSH(d, it) br_flush((d).sig, it)
#else
#define BR_START() ((void)0)
#define BR_FLUSH(d, it) ((void)0)
#endif

template <int RECON>
__global__ __launch_bounds__(256) void kb_flux_x(const BoxDesc *__restrict__ D, const int *__restrict__ pre, int nbox)
{
    const int total = pre[nbox];
    BR_START();
    for (int it = blockIdx.x; it < total; it += gridDim.x) {
        const int b = box_of(pre, nbox, it);
        const int loc = it - pre[b];
        const BoxDesc &d = D[b];
        flux_x_tile<double, RECON>(d.L, d.F, d.tb, d.M, loc % d.fx_gx, loc / d.fx_gx);
        BR_FLUSH(d, it);
    }
}

template <int DIR, int RECON>
__global__ __launch_bounds__(128) void kb_flux_walk(const BoxDesc *__restrict__ D, const int *__restrict__ pre,
                                                    int nbox)
{
    const int total = pre[nbox];
    BR_START();
    for (int it = blockIdx.x; it < total; it += gridDim.x) {
        const int b = box_of(pre, nbox, it);
        const int loc = it - pre[b];
        const BoxDesc &d = D[b];
        const int gx = DIR == 2 ? d.fy_gx : d.fz_gx;
        const int sl = DIR == 2 ? d.fy_sl : d.fz_sl;
        flux_walk_tile<double, DIR, RECON>(d.L, d.F, d.tb, d.M, sl, loc % gx, loc / gx);
        BR_FLUSH(d, it);
    }
}

__global__ __launch_bounds__(256) void kb_ecorn(const BoxDesc *__restrict__ D, const int *__restrict__ pre, int nbox)
{
    const int total = pre[nbox];
    for (int it = blockIdx.x; it < total; it += gridDim.x) {
        const int b = box_of(pre, nbox, it);
        const int loc = it - pre[b];
        const BoxDesc &d = D[b];
        ecorn_tile<double>(d.L, d.F, loc % d.c_gx, loc / d.c_gx, d.zc, d.tw);
    }
}

// trip counts are block-uniform, so the stage-3 block reduction inside tail_tile is safe
template <int STAGE, int COORD, bool COMB>
__global__ __launch_bounds__(256) void kb_tail(const BoxDesc *__restrict__ D, const int *__restrict__ pre, int nbox,
                                               int par)
{
    const int total = pre[nbox];
    BR_START();
    for (int it = blockIdx.x; it < total; it += gridDim.x) {
        const int b = box_of(pre, nbox, it);
        const int loc = it - pre[b];
        const BoxDesc &d = D[b];
        const Fields<double> F = fields_at(d, par);
        const double *Uin = STAGE == 1 ? F.U0 : (STAGE == 2 ? F.U1 : d.Ub);
        double *Uout = STAGE == 2 ? d.Ub : F.U1;
        tail_tile<double, STAGE, COORD, COMB>(d.L, F, d.tb, d.M, d.dt, Uin, Uout, (unsigned long long *)(d.dt + 1),
                                        loc % d.c_gx, loc / d.c_gx, d.zc, d.tw);
        BR_FLUSH(d, it);
    }
}

template <int AX>
__global__ __launch_bounds__(256) void kb_ghost(const BoxDesc *__restrict__ D, const int *__restrict__ pre, int nbox)
{
    const int total = pre[nbox];
    for (int it = blockIdx.x; it < total; it += gridDim.x) {
        const int b = box_of(pre, nbox, it);
        const long loc = it - pre[b];
        const BoxDesc &d = D[b];
        ghost_thread<AX>(d.L, d.F, d.M, loc * blockDim.x + threadIdx.x);
    }
}

template <int COORD>
__global__ __launch_bounds__(256) void kb_finddt(const BoxDesc *__restrict__ D, const int *__restrict__ pre, int nbox)
{
    const int total = pre[nbox];
    for (int it = blockIdx.x; it < total; it += gridDim.x) {
        const int b = box_of(pre, nbox, it);
        const BoxDesc &d = D[b];
        finddt_tile<COORD>(d.L, d.F, d.tb, d.M, (unsigned long long *)(d.dt + 1), it - pre[b]);
    }
}

// next-dt slots to the reduction identity 1e10 (rungekutta.f90:205)
static __global__ void kb_dt_reset(const BoxDesc *__restrict__ D, int nbox)
{
    for (int b = threadIdx.x; b < nbox; b += blockDim.x) D[b].dt[1] = 1.0e10;
}

// step start: current dt = next dt (or the frozen value), appended to the history at the
// device step counter, which then advances.  One workgroup.
static __global__ void kb_dt_begin(const BoxDesc *__restrict__ D, int nbox, double *hist, int *ctr, int frozen)
{
    const long s = *ctr;
    for (int b = threadIdx.x; b < nbox; b += blockDim.x) {
        double *dt = D[b].dt;
        const double v = frozen ? hist[s * nbox + b] : dt[1];
        dt[0] = v;
        hist[s * nbox + b] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) *ctr = (int)s + 1;
}

static __global__ void kb_set_int(int *p, int v)
{
    if (threadIdx.x == 0) *p = v;
}

}  // namespace

struct cumc_batch {
    std::vector<cumc_ctx *> box;
    int nbox = 0;
    int coord = 1, recon = 4;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    BoxDesc *d_desc = nullptr;
    int *d_pre = nullptr;  // P_N arrays of nbox+1
    int *d_ctr = nullptr;
    unsigned long long *d_sig = nullptr;  // per-box branch signatures (diagnostic build)
    double *d_hist = nullptr;
    int hist_cap = 0;      // steps
    std::vector<BoxDesc> h_desc;
    std::vector<int> h_pre;
    int total[P_N] = {0};
    cumc_tuning tune{};
    // A captured graph bakes in the kernel arguments (descriptor, work-list and history pointers,
    // parity) and every phase's grid size; the descriptors themselves are read on the device at
    // replay.  A graph is reused only while its grid sizes and history pointer are current.
    struct Graph {
        int steps, frozen, par;
        unsigned grid[P_N];
        double *hist;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    double last_ms = 0.0;
    std::string err;
    batch_tape *adj = nullptr;  // batched reverse-mode record (cumc_adjoint.hip)
    // descriptors keep each box's current view (zmode) instead of entering the z-collapsed one
    // where allowed: set while a backward pass runs in the views its tape was recorded in
    bool zfreeze = false;
    // per-phase event timing used by the tuner
    bool prof = false;
    std::vector<hipEvent_t> evp;
    std::vector<int> evc;
    size_t evn = 0;
    double ph_ms[P_N] = {0};
};

#define BCHK(b, call)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            (b)->err = std::string(#call) + ": " + hipGetErrorString(e_);                      \
            return CUMC_EHIP;                                                                   \
        }                                                                                       \
    } while (0)

static int bfail(cumc_batch *b, int code, const std::string &m)
{
    b->err = m;
    return code;
}

static void bmark(cumc_batch *b, int ph)
{
    if (!b->prof) return;
    if (b->evn == b->evp.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        b->evp.push_back(e);
        b->evc.push_back(0);
    }
    b->evc[b->evn] = ph;
    hipEventRecord(b->evp[b->evn], b->stream);
    b->evn++;
}

static void bcollect(cumc_batch *b)
{
    for (size_t i = 0; i + 1 < b->evn; i += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, b->evp[i], b->evp[i + 1]) == hipSuccess) b->ph_ms[b->evc[i]] += ms;
    }
    b->evn = 0;
}

// Destroys the cached graph execs once the batch stream is idle (a replay still in flight keeps
// using its exec until it completes).
static void clear_graphs(cumc_batch *b)
{
    if (b->graphs.empty()) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.exec);
    b->graphs.clear();
}

// tile counts of one box (in the addressing of its descriptor, d.L) under the current tuning
static void box_tiles(const cumc_batch *b, BoxDesc &d, int cnt[P_N])
{
    const Lay &L = d.L;
    const int seg = b->tune.seg > 0 ? b->tune.seg : 24;
    const int zc = b->tune.zchunk > 0 ? b->tune.zchunk : 4;
    const long waves = ((long)(L.ny + 2) * L.px + XS_FACES - 1) / XS_FACES;
    d.fx_gx = (int)nblk(waves * 64, 256);
    cnt[P_FX] = d.fx_gx * flux_planes(L);
    const int nsy = (L.ny + 1 + seg - 1) / seg;
    d.fy_sl = (L.ny + 1 + nsy - 1) / nsy;
    d.fy_gx = (int)nblk((long)L.px * flux_planes(L), 128);
    cnt[P_FY] = d.fy_gx * nsy;
    const int nsz = L.zc1 ? 1 : (L.nz + 1 + seg - 1) / seg;  // one z face per column when collapsed
    d.fz_sl = L.zc1 ? 1 : (L.nz + 1 + nsz - 1) / nsz;
    d.fz_gx = (int)nblk((long)L.px * (L.ny + 2), 128);
    cnt[P_FZ] = d.fz_gx * nsz;
    d.zc = zc;
    d.tw = b->tune.tile_w;
    d.c_gx = col_tiles(L, d.tw);
    cnt[P_EC] = cnt[P_TAIL] = d.c_gx * zchunks(L, zc);
    cnt[P_GX] = (int)nblk(ghost_threads<0>(L), 256);
    cnt[P_GY] = (int)nblk(ghost_threads<1>(L), 256);
    cnt[P_GZ] = (int)nblk(ghost_threads<2>(L), 256);
    cnt[P_DT] = (int)nblk((long)L.px * L.ny, 256);
}

// descriptors and work lists from the boxes' current buffers and the tuning
static int refresh(cumc_batch *b)
{
    const int nb = b->nbox;
    b->h_desc.assign(nb, BoxDesc{});
    b->h_pre.assign((size_t)P_N * (nb + 1), 0);
    for (int i = 0; i < nb; i++) {
        cumc_ctx *c = b->box[i];
        BoxDesc &d = b->h_desc[i];
        // an nz = 1 box steps in its z-collapsed view when it may (zcollapse_begin)
        const View v = make_view(c, b->zfreeze ? c->zmode : zcollapse_begin(c));
        d.L = v.L;
        d.F = v.F;
        d.tb = tabs_of(c);
        d.M = c->M;
        d.dt = c->dt_dev;
        d.Ub = v.Ub;
        d.sig = b->d_sig + i;
        int cnt[P_N];
        box_tiles(b, d, cnt);
        for (int p = 0; p < P_N; p++) b->h_pre[(size_t)p * (nb + 1) + i + 1] = b->h_pre[(size_t)p * (nb + 1) + i] + cnt[p];
    }
    for (int p = 0; p < P_N; p++) {
        const long t = b->h_pre[(size_t)p * (nb + 1) + nb];
        if (t > 0x7fffffffL) return bfail(b, CUMC_EINVAL, "batch work list exceeds 2^31 tiles");
        b->total[p] = (int)t;
    }
    BCHK(b, hipMemcpyAsync(b->d_desc, b->h_desc.data(), sizeof(BoxDesc) * nb, hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipMemcpyAsync(b->d_pre, b->h_pre.data(), sizeof(int) * b->h_pre.size(), hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    return CUMC_OK;
}

static inline const int *pre_of(const cumc_batch *b, int p) { return b->d_pre + (size_t)p * (b->nbox + 1); }

// persistent grid of a phase: the tuned workgroup count, never more than there are items
static inline unsigned grid_of(const cumc_batch *b, int p)
{
    static const int tune_slot[P_N] = {CUMC_PH_FLUX_X, CUMC_PH_FLUX_Y, CUMC_PH_FLUX_Z, CUMC_PH_ECORN, CUMC_PH_TAIL,
                                       CUMC_PH_GHOST,  CUMC_PH_GHOST,  CUMC_PH_GHOST,  CUMC_PH_FINDDT};
    const int want = b->tune.blocks[tune_slot[p]];
    const int t = b->total[p];
    const int g = (want <= 0 || want > t) ? t : want;
    return (unsigned)(g > 0 ? g : 1);
}

static void enqueue_finddt_b(cumc_batch *b)
{
    hipStream_t s = b->stream;
    hipLaunchKernelGGL(kb_dt_reset, dim3(1), dim3(256), 0, s, b->d_desc, b->nbox);
    bmark(b, P_DT);
    if (b->coord == CUMC_SPHERICAL)
        hipLaunchKernelGGL((kb_finddt<3>), dim3(grid_of(b, P_DT)), dim3(256), 0, s, b->d_desc, pre_of(b, P_DT), b->nbox);
    else if (b->coord == CUMC_CYLINDRICAL)
        hipLaunchKernelGGL((kb_finddt<2>), dim3(grid_of(b, P_DT)), dim3(256), 0, s, b->d_desc, pre_of(b, P_DT), b->nbox);
    else
        hipLaunchKernelGGL((kb_finddt<1>), dim3(grid_of(b, P_DT)), dim3(256), 0, s, b->d_desc, pre_of(b, P_DT), b->nbox);
    bmark(b, -1);
}

// comb: stage 1 writes, and stage 2 reads, the combined stage-2 input (a plain step; kernels.hpp)
template <int STAGE, bool COMB>
static void launch_tail_b(cumc_batch *b, int par)
{
    const unsigned g = grid_of(b, P_TAIL);
    if (b->coord == CUMC_SPHERICAL)
        hipLaunchKernelGGL((kb_tail<STAGE, 3, COMB>), dim3(g), dim3(256), 0, b->stream, b->d_desc, pre_of(b, P_TAIL), b->nbox, par);
    else if (b->coord == CUMC_CYLINDRICAL)
        hipLaunchKernelGGL((kb_tail<STAGE, 2, COMB>), dim3(g), dim3(256), 0, b->stream, b->d_desc, pre_of(b, P_TAIL), b->nbox, par);
    else
        hipLaunchKernelGGL((kb_tail<STAGE, 1, COMB>), dim3(g), dim3(256), 0, b->stream, b->d_desc, pre_of(b, P_TAIL), b->nbox, par);
}

// face fluxes of the three sweeps and the corner EMFs of every box (the first half of a stage)
static void enqueue_fluxes_ecorn_b(cumc_batch *b)
{
    hipStream_t s = b->stream;
    const BoxDesc *D = b->d_desc;
    const int nb = b->nbox;
    const bool ppm = b->recon == CUMC_PPMC;
    bmark(b, P_FX);
    if (ppm) hipLaunchKernelGGL((kb_flux_x<4>), dim3(grid_of(b, P_FX)), dim3(256), 0, s, D, pre_of(b, P_FX), nb);
    else hipLaunchKernelGGL((kb_flux_x<1>), dim3(grid_of(b, P_FX)), dim3(256), 0, s, D, pre_of(b, P_FX), nb);
    bmark(b, -1);
    bmark(b, P_FY);
    if (ppm) hipLaunchKernelGGL((kb_flux_walk<2, 4>), dim3(grid_of(b, P_FY)), dim3(128), 0, s, D, pre_of(b, P_FY), nb);
    else hipLaunchKernelGGL((kb_flux_walk<2, 1>), dim3(grid_of(b, P_FY)), dim3(128), 0, s, D, pre_of(b, P_FY), nb);
    bmark(b, -1);
    bmark(b, P_FZ);
    if (ppm) hipLaunchKernelGGL((kb_flux_walk<3, 4>), dim3(grid_of(b, P_FZ)), dim3(128), 0, s, D, pre_of(b, P_FZ), nb);
    else hipLaunchKernelGGL((kb_flux_walk<3, 1>), dim3(grid_of(b, P_FZ)), dim3(128), 0, s, D, pre_of(b, P_FZ), nb);
    bmark(b, -1);
    bmark(b, P_EC);
    hipLaunchKernelGGL(kb_ecorn, dim3(grid_of(b, P_EC)), dim3(256), 0, s, D, pre_of(b, P_EC), nb);
    bmark(b, -1);
}

// one RK stage of every box; par = parity of the step within the run
static void enqueue_stage_b(cumc_batch *b, int st, int par, bool comb)
{
    hipStream_t s = b->stream;
    const BoxDesc *D = b->d_desc;
    const int nb = b->nbox;
    enqueue_fluxes_ecorn_b(b);
    if (st == 3) hipLaunchKernelGGL(kb_dt_reset, dim3(1), dim3(256), 0, s, D, nb);
    bmark(b, P_TAIL);
    if (comb && st < 3) {
        if (st == 1) launch_tail_b<1, true>(b, par);
        else launch_tail_b<2, true>(b, par);
    } else {
        if (st == 1) launch_tail_b<1, false>(b, par);
        else if (st == 2) launch_tail_b<2, false>(b, par);
        else launch_tail_b<3, false>(b, par);
    }
    bmark(b, -1);
    bmark(b, P_GX);
    hipLaunchKernelGGL(kb_ghost<0>, dim3(grid_of(b, P_GX)), dim3(256), 0, s, D, pre_of(b, P_GX), nb);
    hipLaunchKernelGGL(kb_ghost<1>, dim3(grid_of(b, P_GY)), dim3(256), 0, s, D, pre_of(b, P_GY), nb);
    if (b->total[P_GZ] > 0)  // (no z pass when every box is z-collapsed)
        hipLaunchKernelGGL(kb_ghost<2>, dim3(grid_of(b, P_GZ)), dim3(256), 0, s, D, pre_of(b, P_GZ), nb);
    bmark(b, -1);
}

// one RUNGEKUTTA step of every box; par = parity of the step within the run
static void enqueue_step_b(cumc_batch *b, int par, int frozen)
{
    hipLaunchKernelGGL(kb_dt_begin, dim3(1), dim3(256), 0, b->stream, b->d_desc, b->nbox, b->d_hist, b->d_ctr, frozen);
    for (int st = 1; st <= 3; st++) enqueue_stage_b(b, st, par, true);
}

// history capacity for nsteps steps (graphs hold the history pointer)
static int reserve_hist(cumc_batch *b, int nsteps)
{
    if (nsteps <= b->hist_cap) return CUMC_OK;
    clear_graphs(b);
    hipFree(b->d_hist);
    b->d_hist = nullptr;
    b->hist_cap = 0;
    if (hipMalloc(&b->d_hist, sizeof(double) * (size_t)nsteps * b->nbox) != hipSuccess)
        return bfail(b, CUMC_ENOMEM, "dt history allocation failed");
    b->hist_cap = nsteps;
    return CUMC_OK;
}

// ---- helpers for the batched reverse mode (cumc_adjoint.hip) ----
int batch_nbox(const cumc_batch *b) { return b->nbox; }
batch_tape *&batch_adj(cumc_batch *b) { return b->adj; }
double &batch_last_ms(cumc_batch *b) { return b->last_ms; }
cumc_ctx *batch_box(const cumc_batch *b, int i) { return b->box[i]; }
hipStream_t batch_stream(const cumc_batch *b) { return b->stream; }
// the stream the batch's enqueue functions launch on, replaced by s; returns the previous one
hipStream_t batch_swap_stream(cumc_batch *b, hipStream_t s)
{
    hipStream_t o = b->stream;
    b->stream = s;
    return o;
}
std::string &batch_err(cumc_batch *b) { return b->err; }
int batch_refresh(cumc_batch *b) { return refresh(b); }
void batch_freeze_views(cumc_batch *b, bool on) { b->zfreeze = on; }
int batch_reserve_hist(cumc_batch *b, int nsteps) { return reserve_hist(b, nsteps); }
double *batch_hist(cumc_batch *b) { return b->d_hist; }
void batch_set_counter(cumc_batch *b, int v)
{
    hipLaunchKernelGGL(kb_set_int, dim3(1), dim3(64), 0, b->stream, b->d_ctr, v);
}
void batch_finddt(cumc_batch *b) { enqueue_finddt_b(b); }
void batch_step(cumc_batch *b, int par, int frozen) { enqueue_step_b(b, par, frozen); }
// one stage of a taped step: the plain stage states, which the tape stores
void batch_stage(cumc_batch *b, int st, int par) { enqueue_stage_b(b, st, par, false); }
// a step's dt bookkeeping (enqueue_step_b without its stages): record the found dt in the
// history, or take the history entry when frozen
void batch_dt_begin(cumc_batch *b, int frozen)
{
    hipLaunchKernelGGL(kb_dt_begin, dim3(1), dim3(256), 0, b->stream, b->d_desc, b->nbox, b->d_hist, b->d_ctr, frozen);
}
// current dt of every box := history entry of step m (the counter is left at m + 1)
void batch_set_dt(cumc_batch *b, int m)
{
    batch_set_counter(b, m);
    hipLaunchKernelGGL(kb_dt_begin, dim3(1), dim3(256), 0, b->stream, b->d_desc, b->nbox, b->d_hist, b->d_ctr, 1);
}
// host-side view after nsteps steps of a run that started from the boxes' buffers
void batch_after_run(cumc_batch *b, int nsteps)
{
    if (nsteps & 1)
        for (cumc_ctx *c : b->box) swap_u(c);
    for (cumc_ctx *c : b->box) c->dt_valid = true;
}

static int get_graph(cumc_batch *b, int steps, int frozen, int par, hipGraphExec_t *out)
{
    unsigned grid[P_N];
    for (int p = 0; p < P_N; p++) grid[p] = grid_of(b, p);
    for (size_t i = 0; i < b->graphs.size(); i++) {
        auto &g = b->graphs[i];
        if (g.steps != steps || g.frozen != frozen || g.par != par) continue;
        if (g.hist == b->d_hist && std::memcmp(g.grid, grid, sizeof(grid)) == 0) {
            *out = g.exec;
            return CUMC_OK;
        }
        // captured under other grid sizes or another history buffer: drop it once idle
        BCHK(b, hipStreamSynchronize(b->stream));
        BCHK(b, hipGraphExecDestroy(g.exec));
        b->graphs.erase(b->graphs.begin() + (long)i);
        break;
    }
    hipGraph_t graph = nullptr;
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < steps; i++) enqueue_step_b(b, (par + i) & 1, frozen);
    BCHK(b, hipStreamEndCapture(b->stream, &graph));
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    BCHK(b, e);
    cumc_batch::Graph ng{steps, frozen, par, {}, b->d_hist, exec};
    std::memcpy(ng.grid, grid, sizeof(grid));
    b->graphs.push_back(ng);
    *out = exec;
    return CUMC_OK;
}

extern "C" int cumc_batch_create(cumc_ctx *const *boxes, int nbox, cumc_batch **out)
{
    if (!boxes || !out || nbox < 1 || nbox > 65536) return CUMC_EINVAL;
    *out = nullptr;
    for (int i = 0; i < nbox; i++) {
        const cumc_ctx *c = boxes[i];
        if (!c || !c->grid_set) return CUMC_ESTATE;
        if (c->cfg.precision != CUMC_FP64 || c->cfg.model != CUMC_MODEL_RIEMANN) return CUMC_EINVAL;
        if (c->cfg.coordinate != boxes[0]->cfg.coordinate || c->cfg.recon != boxes[0]->cfg.recon) return CUMC_EINVAL;
        // the batched sweeps carry the reference's options only: HLL with PPM or TVD minmod
        if (c->M.solver != CUMC_HLL || (c->cfg.recon != CUMC_PPMC && c->cfg.recon != CUMC_TVDMM)) return CUMC_EINVAL;
        for (int j = 0; j < i; j++)
            if (boxes[j] == c) return CUMC_EINVAL;
    }
    cumc_batch *b = new cumc_batch();
    b->box.assign(boxes, boxes + nbox);
    b->nbox = nbox;
    b->coord = boxes[0]->cfg.coordinate;
    b->recon = boxes[0]->cfg.recon;
    b->tune.seg = 24;
    b->tune.zchunk = 4;
    for (int p = 0; p < CUMC_PH_N; p++) b->tune.blocks[p] = 0;
    auto bad = [&](int code) {
        cumc_batch_destroy(b);
        return code;
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bad(CUMC_EHIP);
    if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) return bad(CUMC_EHIP);
    if (hipMalloc(&b->d_desc, sizeof(BoxDesc) * nbox) != hipSuccess) return bad(CUMC_ENOMEM);
    if (hipMalloc(&b->d_pre, sizeof(int) * P_N * (nbox + 1)) != hipSuccess) return bad(CUMC_ENOMEM);
    if (hipMalloc(&b->d_ctr, 64) != hipSuccess) return bad(CUMC_ENOMEM);
    if (hipMalloc(&b->d_sig, sizeof(unsigned long long) * nbox) != hipSuccess) return bad(CUMC_ENOMEM);
    if (hipMemset(b->d_sig, 0, sizeof(unsigned long long) * nbox) != hipSuccess) return bad(CUMC_EHIP);
    if (refresh(b)) return bad(CUMC_EHIP);
    *out = b;
    return CUMC_OK;
}

extern "C" int cumc_batch_destroy(cumc_batch *b)
{
    if (!b) return CUMC_EINVAL;
    if (b->stream) hipStreamSynchronize(b->stream);
    batch_tape_release(b);
    clear_graphs(b);
    hipFree(b->d_desc);
    hipFree(b->d_pre);
    hipFree(b->d_ctr);
    hipFree(b->d_sig);
    hipFree(b->d_hist);
    for (hipEvent_t e : b->evp) hipEventDestroy(e);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return CUMC_OK;
}

static int run_steps(cumc_batch *b, int nsteps, int graph_steps, int frozen)
{
    hipStream_t s = b->stream;
    hipLaunchKernelGGL(kb_set_int, dim3(1), dim3(64), 0, s, b->d_ctr, 0);
    if (!frozen) enqueue_finddt_b(b);
    int done = 0;
    if (graph_steps > 0) {
        while (nsteps - done >= graph_steps) {
            hipGraphExec_t g;
            int r = get_graph(b, graph_steps, frozen, done & 1, &g);
            if (r) return r;
            BCHK(b, hipGraphLaunch(g, s));
            done += graph_steps;
        }
    }
    for (; done < nsteps; done++) enqueue_step_b(b, done & 1, frozen);
    BCHK(b, hipGetLastError());
    return CUMC_OK;
}

extern "C" int cumc_batch_run(cumc_batch *b, int nsteps, int graph_steps, const double *dt_in, double *dt_out)
{
    if (!b || nsteps < 0 || graph_steps < 0) return CUMC_EINVAL;
    if (nsteps == 0) return CUMC_OK;
    const int nb = b->nbox;
    if (dt_in)
        for (long i = 0; i < (long)nsteps * nb; i++)
            if (!std::isfinite(dt_in[i]) || !(dt_in[i] > 0.0)) return bfail(b, CUMC_ENONFINITE, "frozen dt must be finite and positive");
    int r = reserve_hist(b, nsteps);
    if (r) return r;
    for (cumc_ctx *c : b->box) BCHK(b, hipStreamSynchronize(c->stream));
    r = refresh(b);
    if (r) return r;
    if (dt_in)
        BCHK(b, hipMemcpyAsync(b->d_hist, dt_in, sizeof(double) * (size_t)nsteps * nb, hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipEventRecord(b->ev0, b->stream));
    r = run_steps(b, nsteps, graph_steps, dt_in != nullptr);
    if (r) return r;
    BCHK(b, hipEventRecord(b->ev1, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    float ms = 0.f;
    hipEventElapsedTime(&ms, b->ev0, b->ev1);
    b->last_ms = ms;
    batch_after_run(b, nsteps);
    std::vector<double> h((size_t)nsteps * nb);
    BCHK(b, hipMemcpy(h.data(), b->d_hist, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    if (dt_out) std::memcpy(dt_out, h.data(), sizeof(double) * h.size());
    if (!dt_in)
        for (size_t i = 0; i < h.size(); i++)
            if (!std::isfinite(h[i]) || !(h[i] > 0.0) || h[i] >= 1.0e10)
                return bfail(b, CUMC_ENONFINITE, "box " + std::to_string(i % nb) + " step " + std::to_string(i / nb) +
                                                     ": finddt produced a non-finite step");
    return CUMC_OK;
}

extern "C" int cumc_batch_get_tuning(cumc_batch *b, cumc_tuning *out)
{
    if (!b || !out) return CUMC_EINVAL;
    *out = b->tune;
    return CUMC_OK;
}

extern "C" int cumc_batch_set_tuning(cumc_batch *b, const cumc_tuning *t)
{
    if (!b || !t) return CUMC_EINVAL;
    if (t->seg < 1 || t->zchunk < 1 || (t->tile_w != 0 && t->tile_w != 32 && t->tile_w != 64 && t->tile_w != 128))
        return CUMC_EINVAL;
    for (int p = 0; p < CUMC_PH_N; p++)
        if (t->blocks[p] < 0) return CUMC_EINVAL;
    clear_graphs(b);
    b->tune = *t;
    return refresh(b);
}

extern "C" int cumc_batch_last_ms(cumc_batch *b, double *ms)
{
    if (!b || !ms) return CUMC_EINVAL;
    *ms = b->last_ms;
    return CUMC_OK;
}

extern "C" const char *cumc_batch_last_error(cumc_batch *b) { return b ? b->err.c_str() : "no batch"; }

extern "C" int cumc_batch_branch_sig(cumc_batch *b, unsigned long long *sig)
{
    if (!b || !sig) return CUMC_EINVAL;
#ifdef CUMC_BRANCH_SIG
    const size_t bytes = sizeof(unsigned long long) * b->nbox;
    BCHK(b, hipStreamSynchronize(b->stream));
    BCHK(b, hipMemcpy(sig, b->d_sig, bytes, hipMemcpyDeviceToHost));
    BCHK(b, hipMemset(b->d_sig, 0, bytes));
    return CUMC_OK;
#else
    return bfail(b, CUMC_EINVAL, "branch signatures are recorded only by the diagnostic build libcumc_sig.so");
#endif
}

// ---------------------------------------------------------------------------------------------
// Tuner: saves every box's FP64 fields and dt slots, times candidate shapes phase by phase with
// HIP events on direct launches (two steps per candidate after two warm-up steps), keeps the
// fastest per phase and restores the saved state.
// ---------------------------------------------------------------------------------------------
static int time_candidate(cumc_batch *b, const cumc_tuning &t, double ph[P_N])
{
    b->tune = t;
    int r = refresh(b);
    if (r) return r;
    hipLaunchKernelGGL(kb_set_int, dim3(1), dim3(64), 0, b->stream, b->d_ctr, 0);
    enqueue_finddt_b(b);
    // an even number of steps in parity order 0, 1, 0, 1: the conserved buffers end in the
    // roles they started in, so every candidate times a consistent state
    enqueue_step_b(b, 0, 0);  // warm-up
    enqueue_step_b(b, 1, 0);
    b->prof = true;
    b->evn = 0;
    for (int p = 0; p < P_N; p++) b->ph_ms[p] = 0.0;
    enqueue_step_b(b, 0, 0);
    enqueue_step_b(b, 1, 0);
    b->prof = false;
    BCHK(b, hipGetLastError());
    BCHK(b, hipStreamSynchronize(b->stream));
    bcollect(b);
    for (int p = 0; p < P_N; p++) ph[p] = b->ph_ms[p];
    return CUMC_OK;
}

extern "C" int cumc_batch_autotune(cumc_batch *b, cumc_tuning *out)
{
    if (!b) return CUMC_EINVAL;
    const int nb = b->nbox;
    for (cumc_ctx *c : b->box) BCHK(b, hipStreamSynchronize(c->stream));
    if (int r = reserve_hist(b, 4)) return r;
    // save
    std::vector<void *> save(nb, nullptr);
    std::vector<double> dts(2 * (size_t)nb);
    int rc = CUMC_OK;
    for (int i = 0; i < nb && rc == CUMC_OK; i++) {
        cumc_ctx *c = b->box[i];
        if (hipMalloc(&save[i], c->pool_bytes) != hipSuccess) rc = bfail(b, CUMC_ENOMEM, "tuner state copy");
        else if (hipMemcpyAsync(save[i], c->pool, c->pool_bytes, hipMemcpyDeviceToDevice, b->stream) != hipSuccess ||
                 hipMemcpyAsync(&dts[2 * i], c->dt_dev, 2 * sizeof(double), hipMemcpyDeviceToHost, b->stream) != hipSuccess)
            rc = bfail(b, CUMC_EHIP, "tuner state copy");
    }
    const cumc_tuning orig = b->tune;
    std::vector<double *> u0(nb), u1(nb);
    for (int i = 0; i < nb; i++) {
        u0[i] = b->box[i]->F64.U0;
        u1[i] = b->box[i]->F64.U1;
    }
    cumc_tuning best = orig;
    if (rc == CUMC_OK && hipStreamSynchronize(b->stream) != hipSuccess) rc = bfail(b, CUMC_EHIP, "tuner sync");
    // 1) segment length of the pencil sweeps (one workgroup per tile)
    if (rc == CUMC_OK) {
        const int segs[] = {8, 12, 16, 24, 32, 64, 128, 256};
        double bestt = 1e300;
        for (int sg : segs) {
            cumc_tuning t = orig;
            t.seg = sg;
            for (int p = 0; p < CUMC_PH_N; p++) t.blocks[p] = 0;
            double ph[P_N];
            if ((rc = time_candidate(b, t, ph))) break;
            const double v = ph[P_FY] + ph[P_FZ];
            if (v < bestt) {
                bestt = v;
                best.seg = sg;
            }
        }
    }
    // 2) plane chunk and column tile shape of the EMF / tail kernels
    if (rc == CUMC_OK) {
        const int zcs[] = {1, 2, 4, 8, 32};
        double bestt = 1e300;
        for (int z : zcs) {
            cumc_tuning t = best;
            t.zchunk = z;
            for (int p = 0; p < CUMC_PH_N; p++) t.blocks[p] = 0;
            double ph[P_N];
            if ((rc = time_candidate(b, t, ph))) break;
            const double v = ph[P_EC] + ph[P_TAIL];
            if (v < bestt) {
                bestt = v;
                best.zchunk = z;
            }
        }
    }
    if (rc == CUMC_OK) {
        const int tws[] = {0, 32, 64};
        double bestt = 1e300;
        for (int w : tws) {
            cumc_tuning t = best;
            t.tile_w = w;
            for (int p = 0; p < CUMC_PH_N; p++) t.blocks[p] = 0;
            double ph[P_N];
            if ((rc = time_candidate(b, t, ph))) break;
            const double v = ph[P_EC] + ph[P_TAIL];
            if (v < bestt) {
                bestt = v;
                best.tile_w = w;
            }
        }
    }
    // 3) persistent workgroups per phase
    if (rc == CUMC_OK) {
        const int cands[] = {0, 256, 512, 1024, 2048, 4096};
        double bt[P_N];
        for (int p = 0; p < P_N; p++) bt[p] = 1e300;
        int pick[P_N] = {0};
        for (int g : cands) {
            cumc_tuning t = best;
            for (int p = 0; p < CUMC_PH_N; p++) t.blocks[p] = g;
            double ph[P_N];
            if ((rc = time_candidate(b, t, ph))) break;
            for (int p = 0; p < P_N; p++)
                if (ph[p] < bt[p]) {
                    bt[p] = ph[p];
                    pick[p] = g;
                }
        }
        best.blocks[CUMC_PH_FLUX_X] = pick[P_FX];
        best.blocks[CUMC_PH_FLUX_Y] = pick[P_FY];
        best.blocks[CUMC_PH_FLUX_Z] = pick[P_FZ];
        best.blocks[CUMC_PH_ECORN] = pick[P_EC];
        best.blocks[CUMC_PH_TAIL] = pick[P_TAIL];
        best.blocks[CUMC_PH_GHOST] = pick[P_GX];
        best.blocks[CUMC_PH_FINDDT] = pick[P_DT];
    }
    // restore every FP64 plane (so which of U0/U1 holds the state is as saved) and the dt slots
    for (int i = 0; i < nb; i++) {
        cumc_ctx *c = b->box[i];
        if (save[i]) {
            hipMemcpyAsync(c->pool, save[i], c->pool_bytes, hipMemcpyDeviceToDevice, b->stream);
            hipMemcpyAsync(c->dt_dev, &dts[2 * i], 2 * sizeof(double), hipMemcpyHostToDevice, b->stream);
        }
        c->F64.U0 = u0[i];
        c->F64.U1 = u1[i];
        c->F32.U0 = u0[i];
        c->F32.U1 = u1[i];
    }
    hipError_t e = hipStreamSynchronize(b->stream);
    for (void *p : save) hipFree(p);
    if (rc == CUMC_OK && e != hipSuccess) rc = bfail(b, CUMC_EHIP, std::string("tuner restore: ") + hipGetErrorString(e));
    clear_graphs(b);
    b->tune = rc == CUMC_OK ? best : orig;
    int r2 = refresh(b);
    if (rc == CUMC_OK) rc = r2;
    if (rc == CUMC_OK && out) *out = b->tune;
    return rc;
}
