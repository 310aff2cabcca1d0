// cumc.hip -- C ABI (include/cumc.h) over the HIP kernels of kernels.hpp.
//
// One context = one subdomain on one GPU, one HIP stream.  The context owns every device
// buffer; host arrays crossing the ABI are caller-owned and in the reference's Fortran layout.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ctx.hpp"

// float instantiations live in cumc_mixed.hip (built with FP32 contraction and fast division)
void stage_mixed(cumc_ctx *ctx, int st, bool comb);
static void prim_to_f32(cumc_ctx *ctx);
static void sync_prim64(cumc_ctx *ctx);
void fluxes_mixed(cumc_ctx *ctx, int dir, double *aos_tmp);


// setup.f90:121-230 parity factors
static void setup_bfac(Model &M)
{
    for (int s = 0; s < 6; s++)
        for (int i = 0; i < 8; i++) M.bfac[s][i] = 1;
    const int *bf = M.bflag;
    auto set = [&](int s, std::initializer_list<int> idx) {
        for (int i : idx) M.bfac[s][i] = -1;
    };
    if (bf[0] == 2) set(0, {IVX, IBX});
    else if (bf[0] == 3) set(0, {IVX, IBX, IVZ, IBZ});
    if (bf[1] == 2) set(1, {IVX, IBX});
    else if (bf[1] == 3) set(1, {IVX, IBX, IVZ, IBZ});
    if (bf[2] == 2) set(2, {IVY, IBY});
    else if (bf[2] == 3) set(2, {IVY, IBY, IVZ, IBZ});
    else if (bf[2] == 4) set(2, {IVY, IBX, IBZ});
    if (bf[3] == 2) set(3, {IVY, IBY});
    else if (bf[3] == 3) set(3, {IVY, IBY, IVZ, IBZ});
    else if (bf[3] == 4) set(3, {IVY, IBX, IBZ});
    if (bf[4] == 2) set(4, {IVZ, IBZ});
    else if (bf[4] == 4) set(4, {IVZ, IBX, IBY});
    if (bf[5] == 2) set(5, {IVZ, IBZ});
    else if (bf[5] == 4) set(5, {IVZ, IBX, IBY});
}

extern "C" const char *cumc_error_string(int code)
{
    if (code < 0 || code > 5) return "unknown error";
    return g_err_names[code];
}

extern "C" const char *cumc_last_error(cumc_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

extern "C" int cumc_create(const cumc_config *cfg, cumc_ctx **out)
{
    if (!cfg || !out) return CUMC_EINVAL;
    *out = nullptr;
    const cumc_config &c = *cfg;
    if (c.nx < 1 || c.ny < 1 || c.nz < 1) return CUMC_EINVAL;
    if (c.coordinate != CUMC_CARTESIAN && c.coordinate != CUMC_CYLINDRICAL && c.coordinate != CUMC_SPHERICAL)
        return CUMC_EINVAL;
    if (c.recon != CUMC_PPMC && c.recon != CUMC_TVDMM && c.recon != CUMC_TVDMC && c.recon != CUMC_TVDVL &&
        c.recon != CUMC_WENO)
        return CUMC_EINVAL;
    if (c.solver != 0 && c.solver != CUMC_HLL && c.solver != CUMC_LF && c.solver != CUMC_HLLC && c.solver != CUMC_HLLD)
        return CUMC_EINVAL;
    if (c.model != CUMC_MODEL_RIEMANN && c.model != CUMC_MODEL_TORUS_CYN && c.model != CUMC_MODEL_TORUS_SPH)
        return CUMC_EINVAL;
    if (c.precision != CUMC_FP64 && c.precision != CUMC_MIXED) return CUMC_EINVAL;
    if (!(c.ggas > 1.0) || !(c.cfl > 0.0)) return CUMC_EINVAL;
    const int n3[3] = {c.nx, c.ny, c.nz};
    for (int s = 0; s < 6; s++) {
        const int f = c.boundary_flag[s];
        if (f < 0 || f > 4) return CUMC_EINVAL;
        if (f >= 2 && n3[s / 2] < NG) return CUMC_EINVAL;  // parity copies need NGHOST interior cells
    }
    // setup.f90:147-151 and :206-210: equatorial symmetry on x / axial symmetry on z stop the run
    if (c.boundary_flag[0] == 4 || c.boundary_flag[1] == 4) return CUMC_EINVAL;
    if (c.boundary_flag[4] == 3 || c.boundary_flag[5] == 3) return CUMC_EINVAL;
    // kernels address a plane with 32-bit byte offsets: fewer than 2^29 cells per plane
    // (about 800^3 interior cells; a larger grid is sharded into several subdomains).  Checked
    // in 64-bit arithmetic before any HIP call, so a rejected size touches no device.
    if (c.nx >= (1 << 29) || c.ny >= (1 << 29) || c.nz >= (1 << 29)) return CUMC_EINVAL;
    const long row = ((long)c.nx + 2 * NG + 7) / 8 * 8 * (c.ny + 2 * NG);  // below 2^59
    if (row >= (1L << 29) || (row * (c.nz + 2 * NG) + 63) / 64 * 64 >= (1L << 29)) return CUMC_EINVAL;

    cumc_ctx *ctx = new cumc_ctx();
    ctx->cfg = c;
    if (hipGetDevice(&ctx->device) != hipSuccess) {
        delete ctx;
        return CUMC_EHIP;
    }
    Lay &L = ctx->L;
    L.nx = c.nx;
    L.ny = c.ny;
    L.nz = c.nz;
    L.px = ((c.nx + 2 * NG) + 7) / 8 * 8;
    L.py = L.px * (c.ny + 2 * NG);
    const long cells = L.py * (c.nz + 2 * NG);
    L.ps = (cells + 63) / 64 * 64;
    Model &M = ctx->M;
    M.coordinate = c.coordinate;
    M.recon = c.recon;
    M.solver = c.solver == 0 ? CUMC_HLL : c.solver;
    M.model = c.model;
    M.geom_repeat = c.geom_source_repeat;
    for (int s = 0; s < 6; s++) M.bflag[s] = c.boundary_flag[s];
    setup_bfac(M);
    M.ggas = c.ggas;
    M.cfl = c.cfl;
    M.rho_floor = c.rho_floor;
    M.eps_floor = c.eps_floor;
    M.r_sh = c.r_sh;

    // FP64 state: P 8 + EPS 1 + BC 3 + U0 8 + U1 8 + Ub 8 = 36 planes ; fluxes 21 + EMF 3 planes in R,
    // and in the mixed form 12 more FP32 planes (P32 8, EPS32 1, BC32 3)
    const size_t nplanes = 36;
    ctx->pool_bytes = nplanes * (size_t)L.ps * sizeof(double);
    const bool mixed = c.precision == CUMC_MIXED;
    const size_t rsz = mixed ? sizeof(float) : sizeof(double);
    const size_t fbytes = (mixed ? 36 : 24) * (size_t)L.ps * rsz;
    // every failure from here on goes through cumc_destroy, which frees whatever was created
    if (hipMalloc(&ctx->pool, ctx->pool_bytes) != hipSuccess || hipMalloc(&ctx->fluxpool, fbytes) != hipSuccess ||
        hipMalloc(&ctx->dt_dev, 64) != hipSuccess || hipMalloc(&ctx->red, 64) != hipSuccess) {
        cumc_destroy(ctx);
        return CUMC_ENOMEM;
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        cumc_destroy(ctx);
        return CUMC_EHIP;
    }
    // zero everything once: ghost EMFs at index -1 (mhd.f90:142-148) and initial bcell stay zero
    hipMemset(ctx->pool, 0, ctx->pool_bytes);
    hipMemset(ctx->fluxpool, 0, fbytes);
    hipMemset(ctx->dt_dev, 0, 64);
    double *p = ctx->pool;
    auto take = [&](int n) { double *r = p; p += (size_t)n * L.ps; return r; };
    Fields<double> &F = ctx->F64;
    F.P = take(8);
    F.EPS = take(1);
    F.BC = take(3);
    F.U0 = take(8);
    F.U1 = take(8);
    ctx->Ub = take(8);
    Fields<float> &G = ctx->F32;
    G.P = F.P; G.EPS = F.EPS; G.BC = F.BC; G.U0 = F.U0; G.U1 = F.U1;
    for (int d = 0; d < 3; d++) {
        F.F[d] = (double *)ctx->fluxpool + (size_t)d * 7 * L.ps;
        G.F[d] = (float *)ctx->fluxpool + (size_t)d * 7 * L.ps;
    }
    F.E = (double *)ctx->fluxpool + (size_t)21 * L.ps;
    G.E = (float *)ctx->fluxpool + (size_t)21 * L.ps;
    if (mixed) {
        G.P32 = (float *)ctx->fluxpool + (size_t)24 * L.ps;
        G.EPS32 = (float *)ctx->fluxpool + (size_t)32 * L.ps;
        G.BC32 = (float *)ctx->fluxpool + (size_t)33 * L.ps;
    }
    if (hipDeviceSynchronize() != hipSuccess) { cumc_destroy(ctx); return CUMC_EHIP; }
    *out = ctx;
    return CUMC_OK;
}

extern "C" int cumc_destroy(cumc_ctx *ctx)
{
    if (!ctx) return CUMC_EINVAL;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    tape_release(ctx);
    hipFree(ctx->pool);
    hipFree(ctx->fluxpool);
    hipFree(ctx->dt_dev);
    hipFree(ctx->red);
    for (int a = 0; a < 3; a++) {
        hipFree(ctx->tab[a]);
        hipFree(ctx->tabf[a]);
    }
    for (hipEvent_t e : ctx->evpool) hipEventDestroy(e);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return CUMC_OK;
}

// Host-side geometry tables of one axis (coords.f90:49-97 widths/centres, reconstruct.f90
// coefficients), FP64 with the reference's expressions so that hoisting changes no bit.
static std::vector<double> build_tables(const double *f_in, int n)
{
    const int len = n + 7;
    std::vector<double> t((size_t)NTAB * len, 0.0);
    auto F = [&](int j) { return f_in[j + 3]; };
    auto D = [&](int j) { return F(j) - F(j - 1); };
    auto T = [&](int row, int j) -> double & { return t[(size_t)row * len + j + 3]; };
    for (int j = -3; j <= n + 3; j++) T(TB_F, j) = F(j);
    for (int j = -2; j <= n + 3; j++) {
        T(TB_D, j) = D(j);
        T(TB_C, j) = 0.5 * (F(j) + F(j - 1));
    }
    // spherical factors, each with the expression coords.f90 evaluates per call (x**2 = x*x,
    // x**3 = x*x*x), so the kernels combine bit-identical values
    auto p2 = [](double x) { return x * x; };
    auto p3 = [](double x) { return x * x * x; };
    for (int j = -3; j <= n + 3; j++) {
        T(TB_F2, j) = p2(F(j));
        T(TB_SIN, j) = std::sin(F(j));
    }
    for (int j = -2; j <= n + 3; j++) {
        T(TB_V3, j) = (p3(F(j)) - p3(F(j - 1))) / 3.0;
        T(TB_RA, j) = 2.0 * (p3(F(j)) - p3(F(j - 1))) / 3.0 / (p2(F(j)) - p2(F(j - 1)));
        T(TB_D2, j) = p2(F(j)) - p2(F(j - 1));
        T(TB_CD, j) = std::cos(F(j - 1)) - std::cos(F(j));
        T(TB_TANC, j) = std::tan(0.5 * (F(j) + F(j - 1)));
        T(TB_SINC, j) = std::sin(0.5 * (F(j) + F(j - 1)));
    }
    for (int j = -1; j <= n + 2; j++) {
        double b0, b1;
        ppm_slope_coef(D(j - 1), D(j), D(j + 1), &b0, &b1);
        T(TB_B0C, j) = b0;
        T(TB_B1C, j) = b1;
        const double dc = D(j);
        T(TB_CF, j) = (dc + D(j + 1)) / dc;
        T(TB_CB, j) = (dc + D(j - 1)) / dc;
    }
    for (int j = 0; j <= n + 1; j++) {
        PPMCoef k = ppm_coef(D(j - 2), D(j - 1), D(j), D(j + 1), D(j + 2));
        T(TB_A0M, j) = k.a0m; T(TB_A1M, j) = k.a1m; T(TB_A2M, j) = k.a2m; T(TB_A3M, j) = k.a3m;
        T(TB_A0P, j) = k.a0p; T(TB_A1P, j) = k.a1p; T(TB_A2P, j) = k.a2p; T(TB_A3P, j) = k.a3p;
    }
    return t;
}

// An nz = 1 context with periodic or outflow z sides (both make every z ghost a copy of the one
// cell: boundary.f90:564-652) may step in the z-collapsed view when every z face the reference
// computes equals the others: the z sweep's reconstruction coefficients (PPM a/b rows, TVD cf/cb)
// must be the same in every cell of the z table (true for equal cell widths), since faces 0 and
// 1 of the single cell become one face.  FP64 form with the riemann hooks only (the tori clamp
// their ghost zones; the mixed form keeps FP32 copies of the planes).
static bool zcollapse_allowed(const cumc_ctx *ctx, const std::vector<double> &tz)
{
    const cumc_config &c = ctx->cfg;
    const int *bf = c.boundary_flag;
    if (c.nz != 1 || (bf[4] != 0 && bf[4] != 1) || (bf[5] != 0 && bf[5] != 1)) return false;
    if (c.precision != CUMC_FP64 || c.model != CUMC_MODEL_RIEMANN) return false;
    const int len = c.nz + 7;
    auto row_flat = [&](int row, int j0, int j1) {
        const double *r = tz.data() + (size_t)row * len;
        for (int j = j0; j <= j1; j++)
            if (std::memcmp(&r[j + 3], &r[j0 + 3], sizeof(double)) != 0) return false;
        return true;
    };
    for (int row = TB_A0M; row <= TB_A3P; row++)
        if (!row_flat(row, 0, c.nz + 1)) return false;
    for (int row : {(int)TB_B0C, (int)TB_B1C, (int)TB_CF, (int)TB_CB})
        if (!row_flat(row, -1, c.nz + 2)) return false;
    return true;
}

extern "C" int cumc_set_grid(cumc_ctx *ctx, const double *xf, const double *yf, const double *zf)
{
    if (!ctx || !xf || !yf || !zf) return CUMC_EINVAL;
    const double *fs[3] = {xf, yf, zf};
    const int ns[3] = {ctx->L.nx, ctx->L.ny, ctx->L.nz};
    // every check precedes the first upload, so a rejected grid leaves the previous one in place
    for (int a = 0; a < 3; a++)
        for (int j = 1; j < ns[a] + 7; j++)
            if (!(fs[a][j] > fs[a][j - 1])) return fail(ctx, CUMC_EINVAL, "face coordinates must increase");
    if (ctx->cfg.coordinate == CUMC_CYLINDRICAL && !(xf[3] > 0.0))
        return fail(ctx, CUMC_EINVAL, "cylindrical grid needs R > 0 on the inner face");
    if (ctx->cfg.coordinate == CUMC_SPHERICAL && !(xf[3] > 0.0 && yf[3] > 0.0 && yf[ctx->L.ny + 3] < 3.14159265358979323846))
        return fail(ctx, CUMC_EINVAL, "spherical grid needs r > 0 and 0 < theta < pi on the boundary faces");
    zexpand(ctx);
    ctx->zc_ok = false;
    for (int a = 0; a < 3; a++) {
        std::vector<double> t = build_tables(fs[a], ns[a]);
        if (a == 2) ctx->zc_ok = zcollapse_allowed(ctx, t);
        if (!ctx->tab[a]) HIPCHK(ctx, hipMalloc(&ctx->tab[a], t.size() * sizeof(double)));
        HIPCHK(ctx, hipMemcpy(ctx->tab[a], t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
        std::vector<float> tf(t.begin(), t.end());
        if (!ctx->tabf[a]) HIPCHK(ctx, hipMalloc(&ctx->tabf[a], tf.size() * sizeof(float)));
        HIPCHK(ctx, hipMemcpy(ctx->tabf[a], tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    ctx->grid_set = true;
    return CUMC_OK;
}


static int conv_in(cumc_ctx *ctx, const double *host, int nv, double *dst)
{
    const Lay &L = ctx->L;
    const long ncell = (long)(L.nx + 6) * (L.ny + 6) * (L.nz + 6);
    double *tmp = nullptr;
    HIPCHK(ctx, hipMalloc(&tmp, (size_t)ncell * nv * sizeof(double)));
    hipError_t e = hipMemcpyAsync(tmp, host, (size_t)ncell * nv * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_aos_to_soa, dim3(nblk(ncell, 256)), dim3(256), 0, ctx->stream, L, tmp, nv, dst);
        e = hipStreamSynchronize(ctx->stream);
    }
    hipFree(tmp);
    HIPCHK(ctx, e);
    return CUMC_OK;
}

static int conv_out(cumc_ctx *ctx, const double *src, int nv, double *host)
{
    const Lay &L = ctx->L;
    const long ncell = (long)(L.nx + 6) * (L.ny + 6) * (L.nz + 6);
    double *tmp = nullptr;
    HIPCHK(ctx, hipMalloc(&tmp, (size_t)ncell * nv * sizeof(double)));
    hipLaunchKernelGGL(k_soa_to_aos, dim3(nblk(ncell, 256)), dim3(256), 0, ctx->stream, L, src, nv, tmp);
    hipError_t e = hipMemcpyAsync(host, tmp, (size_t)ncell * nv * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(tmp);
    HIPCHK(ctx, e);
    return CUMC_OK;
}

extern "C" int cumc_upload(cumc_ctx *ctx, const double *prim, const double *eps)
{
    if (!ctx || !prim || !eps) return CUMC_EINVAL;
    ctx->dt_valid = false;
    zexpand(ctx);
    ctx->zfresh = false;  // the uploaded z ghost planes are the caller's
    ctx->zbad = false;
    // bcell outside [1,n]^3 is read before the first BOUNDARY (as the reference's zero-initialised
    // array): a re-upload starts from the same zeros as a fresh context
    HIPCHK(ctx, hipMemsetAsync(ctx->F64.BC, 0, 3 * (size_t)ctx->L.ps * sizeof(double), ctx->stream));
    int r = conv_in(ctx, prim, 8, ctx->F64.P);
    if (r) return r;
    if ((r = conv_in(ctx, eps, 1, ctx->F64.EPS))) return r;
    ctx->p64_stale = false;
    prim_to_f32(ctx);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return CUMC_OK;
}

extern "C" int cumc_download(cumc_ctx *ctx, double *prim, double *eps, double *bcell)
{
    if (!ctx) return CUMC_EINVAL;
    int r;
    zexpand(ctx);
    sync_prim64(ctx);
    if (prim && (r = conv_out(ctx, ctx->F64.P, 8, prim))) return r;
    if (eps && (r = conv_out(ctx, ctx->F64.EPS, 1, eps))) return r;
    if (bcell && (r = conv_out(ctx, ctx->F64.BC, 3, bcell))) return r;
    return CUMC_OK;
}

// ---------------------------------------------------------------------------------------------
// Download without device allocation (cumc_download_noalloc).  The AoS staging goes through a
// fixed 32 MiB file-scope device array, slab by slab of whole planes, instead of a temporary the
// size of the state.  Each GPU the library runs on holds its own copy of the array until the
// process ends; host code reaches it through the address hipGetSymbolAddress returns for the
// current device (looked up once per device), never through the array's name.
// ---------------------------------------------------------------------------------------------
constexpr size_t NOALLOC_STAGING_BYTES = size_t(32) << 20;
static __device__ __align__(256) double scratch[NOALLOC_STAGING_BYTES / sizeof(double)];
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[NOALLOC_MAX_DEVICES];

static int scratch_for(cumc_ctx *ctx, double **out)
{
    int dev = -1;
    HIPCHK(ctx, hipGetDevice(&dev));
    if (dev != ctx->device)
        return fail(ctx, CUMC_EINVAL, "cumc_download_noalloc: make the context's device current first");
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(ctx, CUMC_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ctx, CUMC_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = (double *)p;
    }
    *out = g_scratch_dev[dev];
    return CUMC_OK;
}

// SoA planes -> caller's AoS host array through the staging buffer, nl whole planes at a time
static int stage_out(cumc_ctx *ctx, double *stg, const double *src, int nv, double *host)
{
    const Lay &L = ctx->L;
    const long sx = L.nx + 6, sy = L.ny + 6, sz = L.nz + 6;
    const long per_plane = sx * sy * nv;
    const long cap = (long)(NOALLOC_STAGING_BYTES / sizeof(double)) / per_plane;
    for (long l0 = 0; l0 < sz; l0 += cap) {
        const long nl = (sz - l0 < cap) ? sz - l0 : cap;
        Lay S = L;
        S.nz = (int)nl - 6;  // k_soa_to_aos covers nz + 6 planes
        hipLaunchKernelGGL(k_soa_to_aos, dim3(nblk(sx * sy * nl, 256)), dim3(256), 0, ctx->stream, S,
                           src + l0 * L.py, nv, stg);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipMemcpyAsync(host + l0 * per_plane, stg, (size_t)(nl * per_plane) * sizeof(double),
                                   hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));  // the next slab reuses the buffer
    }
    return CUMC_OK;
}

extern "C" int cumc_download_noalloc(cumc_ctx *ctx, double *prim, double *eps, double *bcell)
{
    if (!ctx) return CUMC_EINVAL;
    const Lay &L = ctx->L;
    const long plane = (long)(L.nx + 6) * (L.ny + 6);
    const int nv_max = prim ? 8 : (bcell ? 3 : 1);
    if (plane * nv_max * (long)sizeof(double) > (long)NOALLOC_STAGING_BYTES)
        return fail(ctx, CUMC_EINVAL, "cumc_download_noalloc: one plane of the requested arrays exceeds the "
                                      "32 MiB staging buffer ((nx+6)(ny+6) <= 524288 with prim)");
    double *stg = nullptr;
    int r = scratch_for(ctx, &stg);
    if (r) return r;
    zexpand(ctx);
    sync_prim64(ctx);
    if (prim && (r = stage_out(ctx, stg, ctx->F64.P, 8, prim))) return r;
    if (eps && (r = stage_out(ctx, stg, ctx->F64.EPS, 1, eps))) return r;
    if (bcell && (r = stage_out(ctx, stg, ctx->F64.BC, 3, bcell))) return r;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return CUMC_OK;
}

// ---------------------------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------------------------
void launch_ghost_l(cumc_ctx *ctx, const Lay &L, const Fields<double> &F)
{
    hipStream_t s = ctx->stream;
    const long nx6 = L.nx + 6, ny6 = L.ny + 6;
    hipLaunchKernelGGL(k_ghost<0>, dim3(nblk(ghost_threads<0>(L), 256)), dim3(256), 0, s, L, F, ctx->M);
    hipLaunchKernelGGL(k_ghost<1>, dim3(nblk(ghost_threads<1>(L), 256)), dim3(256), 0, s, L, F, ctx->M);
    // a z-collapsed view has no z pass: its z ghosts are the plane itself
    if (!L.zc1) hipLaunchKernelGGL(k_ghost<2>, dim3(nblk(ghost_threads<2>(L), 256)), dim3(256), 0, s, L, F, ctx->M);
    if (ctx->M.model == CUMC_MODEL_TORUS_CYN || ctx->M.model == CUMC_MODEL_TORUS_SPH) {
        // inflow clamp on the R / r ghosts; torus_cyn also clamps the z ghosts
        const long nyz = (long)(L.ny + 6) * (L.nz + 6) * 3, nxy = nx6 * ny6 * 3;
        hipLaunchKernelGGL(k_custom_boundary_x<double>, dim3(nblk(nyz, 256)), dim3(256), 0, s, L, F.P);
        if (ctx->M.model == CUMC_MODEL_TORUS_CYN)
            hipLaunchKernelGGL(k_custom_boundary_z<double>, dim3(nblk(nxy, 256)), dim3(256), 0, s, L, F.P);
    }
    if (!L.zc1) ctx->zfresh = true;
}

void launch_ghost(cumc_ctx *ctx) { launch_ghost_l(ctx, ctx->L, ctx->F64); }

// Enters the z-collapsed view when the context allows it and its state is a full-layout state
// whose z ghost planes the last ghost fill wrote (zfresh) and whose two periodic B_z faces agree.
bool zcollapse_begin(cumc_ctx *ctx)
{
    if (!ctx->zc_ok || !ctx->zc_on || ctx->ztape) return false;
    if (ctx->zmode) return true;
    // (unequal periodic B_z faces stay unequal: both evolve from equal right-hand sides)
    if (!ctx->zfresh || ctx->zbad) return false;
    const Lay &L = ctx->L;
    unsigned long long cnt = 1;
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, ctx->red + 2, 0ull, 1);
    hipLaunchKernelGGL(k_zcheck, dim3(nblk((long)(L.nx + 6) * (L.ny + 6), 256)), dim3(256), 0, ctx->stream, L,
                       (const double *)ctx->F64.P, (const double *)ctx->F64.U0, ctx->red + 2);
    if (hipMemcpyAsync(&cnt, ctx->red + 2, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return false;
    ctx->zmode = cnt == 0;
    ctx->zbad = !ctx->zmode;
    return ctx->zmode;
}

// A collapsed state back to the full layout (before anything reads or writes the z ghost planes)
void zexpand(cumc_ctx *ctx)
{
    if (!ctx->zmode) return;
    const Lay &L = ctx->L;
    hipLaunchKernelGGL(k_zexpand, dim3(nblk((long)(L.nx + 6) * (L.ny + 6), 256)), dim3(256), 0, ctx->stream, L,
                       ctx->F64);
    ctx->zmode = false;
    ctx->zfresh = true;
}

// mixed form: the FP32 primitive planes from the FP64 ones (after the host or an FP64 pass set
// P / EPS / BC)
static void prim_to_f32(cumc_ctx *ctx)
{
    if (ctx->cfg.precision != CUMC_MIXED) return;
    const Fields<double> &F = ctx->F64;
    const Fields<float> &G = ctx->F32;
    hipLaunchKernelGGL(k_prim_to_f32, dim3(nblk(ctx->L.ps, 256)), dim3(256), 0, ctx->stream, ctx->L.ps, F.P, F.EPS,
                       F.BC, G.P32, G.EPS32, G.BC32);
}

// mixed form: bring P / EPS / BC up to date with the conserved state U0 after steps (the stage
// tail only writes the FP32 planes), ghost cells included
static void sync_prim64(cumc_ctx *ctx)
{
    if (!ctx->p64_stale) return;
    const Lay &L = ctx->L;
    dim3 grid(nblk((long)L.px * (L.ny + 1), 256), L.nz + 1);
    hipLaunchKernelGGL(k_prim_from_u, grid, dim3(256), 0, ctx->stream, L, ctx->F64, ctx->M, (const double *)ctx->F64.U0);
    launch_ghost(ctx);
    ctx->p64_stale = false;
}


// the three stages of a RUNGEKUTTA step on the stream, dt already in dt_dev[0], in the view zc
// (zcollapse_begin's answer); leaves finddt of the new state in dt_dev[1].  Kernel launches only.
static void enqueue_stages(cumc_ctx *ctx, bool zc)
{
    // stage 1 passes the combined stage-2 input (a plain step: nothing reads the intermediate
    // state); the mixed form takes its sources' density from the stage-input state, so it
    // combines only without sources and model hooks
    if (ctx->cfg.precision == CUMC_MIXED) {
        const bool comb = ctx->M.coordinate == CUMC_CARTESIAN && ctx->M.model == CUMC_MODEL_RIEMANN;
        for (int st = 1; st <= 3; st++) stage_mixed(ctx, st, comb);
    } else {
        const View v = make_view(ctx, zc);
        for (int st = 1; st <= 3; st++) stage_l<double>(ctx, v.L, v.F, v.Ub, st, true);
    }
}

// RUNGEKUTTA body on the stream, dt already in dt_dev[0]; leaves finddt of the new state in
// dt_dev[1]
static void enqueue_step(cumc_ctx *ctx)
{
    enqueue_stages(ctx, ctx->cfg.precision != CUMC_MIXED && zcollapse_begin(ctx));
    swap_u(ctx);
    ctx->dt_valid = true;
}

// cumc_run's steps as hipGraphs: one graph per conserved-buffer parity and view, captured on
// first use within the call and destroyed at its end, so no graph outlives the buffers, view and
// launch shapes it was captured with.  On small grids the ~24 launches of a step cost more than
// their kernels.  The view is settled (zcollapse_begin, which may synchronise the stream) before
// any capture; a capture that fails leaves the step to direct launches.
struct StepGraphs {
    hipStream_t s;
    hipGraphExec_t ex[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    explicit StepGraphs(hipStream_t st) : s(st) {}
    ~StepGraphs()
    {
        // (an error return may leave launches of these graphs in flight)
        bool any = false;
        for (auto &r : ex)
            for (hipGraphExec_t e : r) any = any || e;
        if (!any) return;
        (void)hipStreamSynchronize(s);
        for (auto &r : ex)
            for (hipGraphExec_t e : r)
                if (e) hipGraphExecDestroy(e);
    }
};
static bool capture_stages(cumc_ctx *ctx, bool zc, hipGraphExec_t *out)
{
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
    enqueue_stages(ctx, zc);
    const bool ended = hipStreamEndCapture(ctx->stream, &g) == hipSuccess && g;
    const bool ok = ended && hipGraphInstantiate(out, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) hipGraphDestroy(g);
    if (!ok) *out = nullptr;
    (void)hipGetLastError();  // a refused capture leaves nothing enqueued; the caller launches directly
    return ok;
}

extern "C" int cumc_initial_update(cumc_ctx *ctx)
{
    if (!ctx) return CUMC_EINVAL;
    if (!ctx->grid_set) return fail(ctx, CUMC_ESTATE, "cumc_set_grid must precede cumc_initial_update");
    zexpand(ctx);
    const Lay &L = ctx->L;
    dim3 grid(nblk((long)L.px * (L.ny + 1), 256), L.nz + 1);
    hipLaunchKernelGGL(k_init, grid, dim3(256), 0, ctx->stream, L, ctx->F64, ctx->M, ctx->F64.U0);
    launch_ghost(ctx);
    ctx->p64_stale = false;
    prim_to_f32(ctx);
    ctx->dt_valid = false;
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return CUMC_OK;
}

extern "C" int cumc_boundary(cumc_ctx *ctx)
{
    if (!ctx) return CUMC_EINVAL;
    zexpand(ctx);
    sync_prim64(ctx);
    launch_ghost(ctx);
    prim_to_f32(ctx);
    ctx->dt_valid = false;
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return CUMC_OK;
}

static void enqueue_finddt(cumc_ctx *ctx)
{
    const Lay &L = ctx->L;
    const double init = 1.0e10;
    unsigned long long bits;
    std::memcpy(&bits, &init, 8);
    unsigned long long *out = (unsigned long long *)(ctx->dt_dev + 1);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, out, bits, 1);
    dim3 grid(nblk((long)L.px * L.ny, 256));
    Tabs tb = tabs_of(ctx);
    if (ctx->M.coordinate == CUMC_SPHERICAL)
        hipLaunchKernelGGL((k_finddt<3>), grid, dim3(256), 0, ctx->stream, L, ctx->F64, tb, ctx->M, out);
    else if (ctx->M.coordinate == CUMC_CYLINDRICAL)
        hipLaunchKernelGGL((k_finddt<2>), grid, dim3(256), 0, ctx->stream, L, ctx->F64, tb, ctx->M, out);
    else
        hipLaunchKernelGGL((k_finddt<1>), grid, dim3(256), 0, ctx->stream, L, ctx->F64, tb, ctx->M, out);
    ctx->dt_valid = true;
}

// dt of the current state into dt_dev[0] (and *dt): the value the last stage 3 left in
// dt_dev[1] when the state has not changed since, else a finddt launch
static int finddt_to_current(cumc_ctx *ctx, double *dt)
{
    if (!ctx->dt_valid) {
        sync_prim64(ctx);
        mark(ctx, KC_FINDDT);
        enqueue_finddt(ctx);
        mark(ctx, -1);
    }
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpyAsync(ctx->dt_dev, ctx->dt_dev + 1, sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(dt, ctx->dt_dev + 1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    collect(ctx);
    if (!std::isfinite(*dt) || !(*dt > 0.0) || *dt >= 1.0e10)
        return fail(ctx, CUMC_ENONFINITE, "finddt produced a non-finite step (state has NaN/Inf)");
    return CUMC_OK;
}

extern "C" int cumc_finddt(cumc_ctx *ctx, double *dt)
{
    if (!ctx || !dt) return CUMC_EINVAL;
    if (!ctx->grid_set) return fail(ctx, CUMC_ESTATE, "no grid");
    return finddt_to_current(ctx, dt);
}

extern "C" int cumc_rungekutta(cumc_ctx *ctx, double dt)
{
    if (!ctx) return CUMC_EINVAL;
    if (!ctx->grid_set) return fail(ctx, CUMC_ESTATE, "no grid");
    if (!std::isfinite(dt) || !(dt > 0.0)) return fail(ctx, CUMC_ENONFINITE, "time step must be finite and positive");
    HIPCHK(ctx, hipMemcpyAsync(ctx->dt_dev, &dt, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    enqueue_step(ctx);
    HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    ctx->last_ms = ms;
    return CUMC_OK;
}

extern "C" int cumc_run(cumc_ctx *ctx, int max_steps, double total_time, double *time, int *steps_done, double *dts)
{
    if (!ctx || !time) return CUMC_EINVAL;
    if (!ctx->grid_set) return fail(ctx, CUMC_ESTATE, "no grid");
    double global_time = *time;
    int n = 0;
    // step graphs for longer runs (profiling marks events per kernel class: direct launches)
    bool graphs = max_steps >= 8 && !ctx->prof;
    StepGraphs sg(ctx->stream);
    int par = 0;  // conserved-buffer parity of the step within this call
    HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    while (global_time < total_time && n < max_steps) {
        double dt;
        int r = finddt_to_current(ctx, &dt);
        if (r) { *time = global_time; if (steps_done) *steps_done = n; return r; }
        if (global_time + dt > total_time) {
            dt = total_time - global_time;
            HIPCHK(ctx, hipMemcpyAsync(ctx->dt_dev, &dt, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        }
        if (graphs) {
            const bool zc = ctx->cfg.precision != CUMC_MIXED && zcollapse_begin(ctx);
            hipGraphExec_t &ex = sg.ex[par][zc ? 1 : 0];
            if (!ex && !capture_stages(ctx, zc, &ex)) graphs = false;
            if (graphs) {
                HIPCHK(ctx, hipGraphLaunch(ex, ctx->stream));
                swap_u(ctx);
                ctx->dt_valid = true;
            } else {
                enqueue_step(ctx);
            }
        } else {
            enqueue_step(ctx);
        }
        par ^= 1;
        HIPCHK(ctx, hipGetLastError());
        if (dts) dts[n] = dt;
        n++;
        global_time = global_time + dt;
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    ctx->last_ms = ms;
    *time = global_time;
    if (steps_done) *steps_done = n;
    return CUMC_OK;
}

extern "C" int cumc_find_divb(cumc_ctx *ctx, double *max_signed, double *max_abs)
{
    if (!ctx || !max_signed || !max_abs) return CUMC_EINVAL;
    if (!ctx->grid_set) return fail(ctx, CUMC_ESTATE, "no grid");
    const Lay &L = ctx->L;
    zexpand(ctx);
    sync_prim64(ctx);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, ctx->red, 0ull, 2);
    dim3 grid(nblk((long)L.px * L.ny, 256));
    Tabs tb = tabs_of(ctx);
    if (ctx->M.coordinate == CUMC_SPHERICAL)
        hipLaunchKernelGGL((k_divb<3>), grid, dim3(256), 0, ctx->stream, L, ctx->F64, tb, ctx->red);
    else if (ctx->M.coordinate == CUMC_CYLINDRICAL)
        hipLaunchKernelGGL((k_divb<2>), grid, dim3(256), 0, ctx->stream, L, ctx->F64, tb, ctx->red);
    else
        hipLaunchKernelGGL((k_divb<1>), grid, dim3(256), 0, ctx->stream, L, ctx->F64, tb, ctx->red);
    double r[2];
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpyAsync(r, ctx->red, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    *max_signed = r[0];
    *max_abs = r[1];
    return CUMC_OK;
}

extern "C" int cumc_sweep_flux(cumc_ctx *ctx, int dir, double *out)
{
    if (!ctx || !out || dir < 1 || dir > 3) return CUMC_EINVAL;
    if (!ctx->grid_set) return fail(ctx, CUMC_ESTATE, "no grid");
    const Lay &L = ctx->L;
    const long ncell = (long)(L.nx + 6) * (L.ny + 6) * (L.nz + 6);
    zexpand(ctx);
    double *tmp = nullptr;
    HIPCHK(ctx, hipMalloc(&tmp, (size_t)ncell * 7 * sizeof(double)));
    if (ctx->cfg.precision == CUMC_MIXED) {
        fluxes_mixed(ctx, dir, tmp);
    } else {
        launch_fluxes<double>(ctx, ctx->F64, dir);
        hipLaunchKernelGGL((k_flux_to_aos<double>), dim3(nblk(ncell, 256)), dim3(256), 0, ctx->stream, L,
                           ctx->F64.F[dir - 1], tmp);
    }
    hipError_t e = hipMemcpyAsync(out, tmp, (size_t)ncell * 7 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(tmp);
    HIPCHK(ctx, e);
    return CUMC_OK;
}

extern "C" int cumc_get_tuning(cumc_ctx *ctx, cumc_tuning *out)
{
    if (!ctx || !out) return CUMC_EINVAL;
    std::memset(out, 0, sizeof(*out));
    out->seg = ctx->seg;
    out->zchunk = ctx->zchunk;
    out->tile_w = ctx->tile_w;
    return CUMC_OK;
}

extern "C" int cumc_set_tuning(cumc_ctx *ctx, const cumc_tuning *t)
{
    if (!ctx || !t || t->seg < 1 || t->zchunk < 1) return CUMC_EINVAL;
    if (t->tile_w != 0 && t->tile_w != 32 && t->tile_w != 64 && t->tile_w != 128) return CUMC_EINVAL;
    ctx->tile_w = t->tile_w;
    ctx->seg = t->seg;
    ctx->zchunk = t->zchunk;
    return CUMC_OK;
}

// Tuner: one warm-up step, then one timed step per candidate with per-kernel-class HIP events:
// the segment length is scored by the y + z sweeps, the plane chunk and column tiles by the EMF +
// stage-tail kernels.  The FP64 fields (and the mixed form's FP32 primitive planes), the dt slots and the
// U0/U1 assignment are saved and restored, so the state is as before; the per-class timing
// totals are cleared.
extern "C" int cumc_autotune(cumc_ctx *ctx, cumc_tuning *out)
{
    if (!ctx) return CUMC_EINVAL;
    if (!ctx->grid_set) return fail(ctx, CUMC_ESTATE, "no grid");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    void *save = nullptr;
    const size_t p32_bytes = ctx->F32.P32 ? 12 * (size_t)ctx->L.ps * sizeof(float) : 0;
    if (hipMalloc(&save, ctx->pool_bytes + p32_bytes) != hipSuccess) return fail(ctx, CUMC_ENOMEM, "tuner state copy");
    double dts[2];
    double *u0 = ctx->F64.U0, *u1 = ctx->F64.U1;
    const bool valid = ctx->dt_valid, prof = ctx->prof, stale = ctx->p64_stale;
    const bool zmode0 = ctx->zmode, zfresh0 = ctx->zfresh;
    const int seg0 = ctx->seg, zc0 = ctx->zchunk, tw0 = ctx->tile_w;
    hipMemcpyAsync(save, ctx->pool, ctx->pool_bytes, hipMemcpyDeviceToDevice, ctx->stream);
    if (p32_bytes)
        hipMemcpyAsync((char *)save + ctx->pool_bytes, ctx->F32.P32, p32_bytes, hipMemcpyDeviceToDevice, ctx->stream);
    hipMemcpyAsync(dts, ctx->dt_dev, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    int rc = CUMC_OK;
    double dt;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(ctx, CUMC_EHIP, "tuner state copy");
    if (rc == CUMC_OK) rc = finddt_to_current(ctx, &dt);
    auto timed_step = [&](double *cls) -> int {
        ctx->prof = true;
        ctx->evn = 0;
        for (int c = 0; c < KC_N; c++) ctx->kc_ms[c] = 0.0;
        hipMemcpyAsync(ctx->dt_dev, &dt, sizeof(double), hipMemcpyHostToDevice, ctx->stream);
        enqueue_step(ctx);
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(ctx, CUMC_EHIP, "tuner step");
        collect(ctx);
        for (int c = 0; c < KC_N; c++) cls[c] = ctx->kc_ms[c];
        return CUMC_OK;
    };
    double cls[KC_N];
    int best_seg = seg0, best_zc = zc0, best_tw = tw0;
    if (rc == CUMC_OK) rc = timed_step(cls);  // warm-up
    if (rc == CUMC_OK) {
        const int segs[] = {8, 12, 16, 24, 32, 48, 64, 96};
        double bt = 1e300;
        for (int sg : segs) {
            ctx->seg = sg;
            if ((rc = timed_step(cls))) break;
            const double v = cls[KC_FLUX_Y] + cls[KC_FLUX_Z];
            if (v < bt) { bt = v; best_seg = sg; }
        }
        ctx->seg = best_seg;
    }
    if (rc == CUMC_OK) {
        const int zcs[] = {1, 2, 4, 8, 16, 32};
        double bt = 1e300;
        for (int z : zcs) {
            ctx->zchunk = z;
            if ((rc = timed_step(cls))) break;
            const double v = cls[KC_ECORN] + cls[KC_TAIL];
            if (v < bt) { bt = v; best_zc = z; }
        }
        ctx->zchunk = best_zc;
    }
    if (rc == CUMC_OK) {
        const int tws[] = {0, 32, 64, 128};
        double bt = 1e300;
        for (int w : tws) {
            ctx->tile_w = w;
            if ((rc = timed_step(cls))) break;
            const double v = cls[KC_ECORN] + cls[KC_TAIL];
            if (v < bt) { bt = v; best_tw = w; }
        }
    }
    hipMemcpyAsync(ctx->pool, save, ctx->pool_bytes, hipMemcpyDeviceToDevice, ctx->stream);
    if (p32_bytes)
        hipMemcpyAsync(ctx->F32.P32, (char *)save + ctx->pool_bytes, p32_bytes, hipMemcpyDeviceToDevice, ctx->stream);
    hipMemcpyAsync(ctx->dt_dev, dts, 2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    hipFree(save);
    ctx->F64.U0 = u0;
    ctx->F64.U1 = u1;
    ctx->F32.U0 = u0;
    ctx->F32.U1 = u1;
    ctx->dt_valid = valid;
    ctx->p64_stale = stale;
    ctx->zmode = zmode0;
    ctx->zfresh = zfresh0;
    ctx->prof = prof;
    ctx->evn = 0;
    for (int c = 0; c < KC_N; c++) {
        ctx->kc_ms[c] = 0.0;
        ctx->kc_count[c] = 0;
    }
    if (rc == CUMC_OK && e != hipSuccess) rc = fail(ctx, CUMC_EHIP, std::string("tuner restore: ") + hipGetErrorString(e));
    ctx->seg = rc == CUMC_OK ? best_seg : seg0;
    ctx->zchunk = rc == CUMC_OK ? best_zc : zc0;
    ctx->tile_w = rc == CUMC_OK ? best_tw : tw0;
    if (rc == CUMC_OK && out) cumc_get_tuning(ctx, out);
    return rc;
}

extern "C" int cumc_set_zcollapse(cumc_ctx *ctx, int on)
{
    if (!ctx) return CUMC_EINVAL;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (!on) zexpand(ctx);
    ctx->zc_on = on != 0;
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return CUMC_OK;
}

extern "C" int cumc_zcollapse_state(cumc_ctx *ctx, int *allowed, int *active)
{
    if (!ctx) return CUMC_EINVAL;
    if (allowed) *allowed = ctx->zc_ok && ctx->zc_on && !ctx->ztape;
    if (active) *active = ctx->zmode;
    return CUMC_OK;
}

extern "C" int cumc_last_step_ms(cumc_ctx *ctx, double *ms)
{
    if (!ctx || !ms) return CUMC_EINVAL;
    *ms = ctx->last_ms;
    return CUMC_OK;
}

extern "C" int cumc_profile(cumc_ctx *ctx, int on)
{
    if (!ctx) return CUMC_EINVAL;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->evn = 0;
    for (int c = 0; c < KC_N; c++) {
        ctx->kc_ms[c] = 0.0;
        ctx->kc_count[c] = 0;
    }
    ctx->prof = on != 0;
    return CUMC_OK;
}

extern "C" int cumc_kernel_times(cumc_ctx *ctx, int n, double *total_ms, long *launches, const char **names)
{
    if (!ctx || n < 0) return CUMC_EINVAL;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    collect(ctx);
    for (int c = 0; c < n && c < KC_N; c++) {
        if (total_ms) total_ms[c] = ctx->kc_ms[c];
        if (launches) launches[c] = ctx->kc_count[c];
        if (names) names[c] = kc_names[c];
    }
    return KC_N;
}

// Plain device copy of `bytes` bytes, run `reps` times on a private stream: the known-size
// reference against which HBM traffic counters (FETCH_SIZE / WRITE_SIZE) are calibrated.
// one 16-byte element per thread over n/256 workgroups: on MI355X this reaches about 6.1 TB/s,
// against about 4.6 TB/s for a grid-stride loop over 4096 workgroups
__global__ __launch_bounds__(256) void k_copy_probe(const double2 *__restrict__ a, double2 *__restrict__ b, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// The stage tail's memory access pattern with no arithmetic: NR FP64 planes read and NW written
// at each cell, one (j,k) column per thread over a chunk of planes, with the tail's column tiles
// and XCD tile order.  Its bandwidth is the ceiling the tail itself can reach on this layout.
template <int NR, int NW>
__global__ __launch_bounds__(256) void k_pattern_probe(Lay L, const double *__restrict__ in, double *__restrict__ out,
                                                       int gx, int gy, int zc, int tw)
{
    int bx, by;
    if (!xcd_tile(gx, gy, bx, by)) return;
    int j, k;
    col_of(L, bx, tw, j, k);
    const int l0 = by * zc;
    if (k > L.ny || j > L.nx || l0 > L.nz) return;
    const int l1 = (l0 + zc - 1 < L.nz) ? l0 + zc - 1 : L.nz;
    const long ps = L.ps;
    unsigned g = (unsigned)L.id(j, k, l0);
    for (int l = l0; l <= l1; l++, g += (unsigned)L.py) {
        double a = 0.0;
#pragma unroll
        for (int c = 0; c < NR; c++) a += ldi(in + c * ps, g);
#pragma unroll
        for (int c = 0; c < NW; c++) sti(out + c * ps, g, a + (double)c);
        if (NW == 0 && a == -1.0) sti(out, g, a);  // keeps the loads of a read-only pattern
    }
}

extern "C" int cumc_pattern_probe(cumc_ctx *ctx, int nread, int nwrite, int reps, double *gbps)
{
    if (!ctx || reps < 1 || !gbps) return CUMC_EINVAL;
    const Lay &L = ctx->L;
    const int zc = ctx->zchunk, tw = ctx->tile_w;
    const int gx = col_tiles(L, tw), gy = (L.nz + zc) / zc;
    const dim3 grid(xcd_blocks(gx, gy));
    void (*kern)(Lay, const double *, double *, int, int, int, int) = nullptr;
    if (nread == 26 && nwrite == 20) kern = k_pattern_probe<26, 20>;
    else if (nread == 34 && nwrite == 20) kern = k_pattern_probe<34, 20>;
    else if (nread == 34 && nwrite == 0) kern = k_pattern_probe<34, 0>;
    else if (nread == 0 && nwrite == 20) kern = k_pattern_probe<0, 20>;
    else return fail(ctx, CUMC_EINVAL, "pattern probe: supported (reads, writes) are (26,20) (34,20) (34,0) (0,20)");
    double *in = nullptr, *out = nullptr;
    const size_t pb = (size_t)L.ps * sizeof(double);
    HIPCHK(ctx, hipMalloc(&in, (nread > 0 ? nread : 1) * pb));
    if (hipMalloc(&out, (nwrite > 0 ? nwrite : 1) * pb) != hipSuccess) {
        hipFree(in);
        return fail(ctx, CUMC_ENOMEM, "pattern probe buffers");
    }
    hipMemsetAsync(in, 0, (nread > 0 ? nread : 1) * pb, ctx->stream);
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, ctx->stream, L, in, out, gx, gy, zc, tw);
    hipEventRecord(ctx->ev0, ctx->stream);
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(kern, grid, dim3(256), 0, ctx->stream, L, in, out, gx, gy, zc, tw);
    hipEventRecord(ctx->ev1, ctx->stream);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    float ms = 0.f;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    hipFree(in);
    hipFree(out);
    HIPCHK(ctx, e);
    const double cells = (double)(L.nx + 1) * (L.ny + 1) * (L.nz + 1);
    *gbps = (nread + nwrite) * 8.0 * cells * reps / (ms * 1e-3) / 1e9;
    return CUMC_OK;
}

extern "C" int cumc_copy_probe(long bytes, int reps, double *gbps)
{
    if (bytes < 1024 || reps < 1 || bytes / 16 / 256 >= 0x7fffffffL) return CUMC_EINVAL;
    void *a = nullptr, *b = nullptr;
    hipStream_t s;
    hipEvent_t e0, e1;
    if (hipMalloc(&a, bytes) != hipSuccess) return CUMC_ENOMEM;
    if (hipMalloc(&b, bytes) != hipSuccess) { hipFree(a); return CUMC_ENOMEM; }
    hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipMemsetAsync(a, 0, bytes, s);
    const long n = bytes / 16;
    const dim3 grid((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(k_copy_probe, grid, dim3(256), 0, s, (const double2 *)a, (double2 *)b, n);
    hipEventRecord(e0, s);
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(k_copy_probe, grid, dim3(256), 0, s, (const double2 *)a, (double2 *)b, n);
    hipEventRecord(e1, s);
    hipError_t e = hipStreamSynchronize(s);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    if (gbps) *gbps = 2.0 * (double)n * 16.0 * reps / (ms * 1e-3) / 1e9;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamDestroy(s);
    hipFree(a);
    hipFree(b);
    return e == hipSuccess ? CUMC_OK : CUMC_EHIP;
}
