// kernels.hpp -- HIP kernels of the CUMC3D step for gfx950 (CDNA4, wave64).
//
// Device layout: structure-of-arrays planes, x fastest, with the reference's 3-deep ghost
// shell on every side.  Cell (j,k,l), j in [-2, nx+3], lives at
//     (l+2)*py + (k+2)*px + (j+2)
// of every plane; px >= nx+6 is the padded row pitch, py = px*(ny+6) the plane pitch and ps
// the stride between the planes of one field.  Face-centred quantities use the index of the
// cell on their low side (face j sits between cells j and j+1, as in prim(ibx,j,k,l)).
#pragma once
#include <hip/hip_runtime.h>

#include "physics.hpp"

namespace cumc {

constexpr int NG = 3;
constexpr int NTAB = 23;  // per-axis table rows
enum { TB_A0M = 0, TB_A1M, TB_A2M, TB_A3M, TB_A0P, TB_A1P, TB_A2P, TB_A3P, TB_B0C, TB_B1C, TB_CF, TB_CB,
       TB_F, TB_D, TB_C,
       // spherical factors of coords.f90:139-160, 213-219, 284-298, 388-411 (Fortran x**2 = x*x,
       // x**3 = x*x*x): F^2; (F_j^3 - F_{j-1}^3)/3; 2 (F_j^3 - F_{j-1}^3)/3/(F_j^2 - F_{j-1}^2);
       // F_j^2 - F_{j-1}^2; sin F; cos F_{j-1} - cos F_j; tan and sin of the cell centre
       TB_F2, TB_V3, TB_RA, TB_D2, TB_SIN, TB_CD, TB_TANC, TB_SINC };

struct Lay {
    int nx, ny, nz;
    long px, py, ps;
    // 1: the z-collapsed view of a periodic nz = 1 subdomain.  Its z ghost planes hold copies of
    // the single interior plane l = 1, so the view addresses only that plane (py = 0, field bases
    // moved to l = 1): every z offset reads the plane itself, and the kernels compute the plane
    // once instead of each ghost copy of it.
    int zc1 = 0;
    __host__ __device__ long id(int j, int k, int l) const { return (long)(l + 2) * py + (long)(k + 2) * px + (j + 2); }
};

// planes of face fluxes the x and y sweeps compute: l in [0, nz+1], or the one plane of a
// z-collapsed view
__host__ __device__ inline int flux_planes(const Lay &L) { return L.zc1 ? 1 : L.nz + 2; }
// column-tile chunks along z of the EMF and stage-tail kernels: planes [0, nz], or the interior
// plane l = 1 alone in a z-collapsed view (the reference's l = 0 plane is a ghost copy there)
__host__ __device__ inline int zchunks(const Lay &L, int zc) { return L.zc1 ? 1 : (L.nz + zc) / zc; }

// Element i of a plane whose base is uniform across the wavefront, addressed as base + 32-bit
// lane byte offset: the compiler can then use the scalar-base form of the global load/store
// (one 32-bit offset register per lane shared by every plane) instead of a 64-bit address per
// lane and per plane.  Needs i * sizeof(T) < 2^32: planes of fewer than 2^29 cells (checked
// at context creation).
template <typename T>
__device__ inline T ldi(const T *base, unsigned i)
{
    return *(const T *)((const char *)base + i * (unsigned)sizeof(T));
}
template <typename T>
__device__ inline void sti(T *base, unsigned i, T v)
{
    *(T *)((char *)base + i * (unsigned)sizeof(T)) = v;
}

// Per-axis geometry tables: tab[row * (n + 7) + j + 3] for j in [-3, n+3].
struct Tabs {
    const double *t[3];
    const float *tf[3];  // the same tables rounded to FP32 (read by the FP32 sweeps)
    int len[3];          // n + 7 per axis
    __device__ double at(int ax, int row, int j) const { return ldi(t[ax] + (long)row * len[ax], (unsigned)(j + 3)); }
    // entry in the sweep type: FP32 reads the rounded copy (equal to R(at(...)))
    template <typename R>
    __device__ R atr(int ax, int row, int j) const
    {
        if constexpr (sizeof(R) == 4) return ldi(tf[ax] + (long)row * len[ax], (unsigned)(j + 3));
        else return at(ax, row, j);
    }
};

struct Model {
    int coordinate, recon, model, geom_repeat;
    int solver;  // CUMC_HLL, CUMC_LF, CUMC_HLLC or CUMC_HLLD
    int bflag[6];
    int bfac[6][8];  // x-in, x-out, y-in, y-out, z-in, z-out
    double ggas, cfl, rho_floor, eps_floor, r_sh;
};

// All device fields of one context (FP64 state; flux/EMF planes in the stage type R).
template <typename R>
struct Fields {
    double *P;    // 8 planes: rho vx vy vz p Bx By Bz (face B)
    double *EPS;  // 1 plane
    double *BC;   // 3 planes: cell-centred B
    double *U0;   // 8 planes: conserved state at the start of the step (u_old)
    double *U1;   // 8 planes: conserved state of the current stage (cons)
    R *F[3];      // 7 planes per sweep direction
    R *E;         // 3 planes: corner EMFs
    // mixed form only: the primitive state as FP32 planes, written by the stage tail and the
    // FP32 ghost fill and read by the sweeps and the EMF kernel (which compute in FP32 anyway);
    // P / EPS / BC are then brought up to date from the conserved state when the host asks
    float *P32 = nullptr;    // 8 planes, as P
    float *EPS32 = nullptr;  // 1 plane
    float *BC32 = nullptr;   // 3 planes
};

// primitive / cell-B planes the sweeps and the EMF kernel read: the FP64 state in the FP64
// form, its FP32 copy in the mixed form
__host__ __device__ inline const double *sweep_p(const Fields<double> &F) { return F.P; }
__host__ __device__ inline const float *sweep_p(const Fields<float> &F) { return F.P32; }
__host__ __device__ inline const double *sweep_bc(const Fields<double> &F) { return F.BC; }
__host__ __device__ inline const float *sweep_bc(const Fields<float> &F) { return F.BC32; }


// ------------------------------------------------------------------------------------------
// x sweep: lanes walk x inside one wavefront (consecutive cells), neighbours via shuffles.
// A wave loads 64 consecutive cells of the flattened (j,k) plane and emits the 59 faces whose
// 6-cell stencil lies inside the wave.  GET_FLUXES x-direction (solver/spatial.f90:396-433).
// ------------------------------------------------------------------------------------------
constexpr int XS_FACES = 59;

// Lane-neighbour moves of the wave-parallel sweeps: lane i takes the value of lane i - 1
// (lane_up) or i + 1 (lane_down), the end lane keeping its own, as __shfl_up / __shfl_down with
// delta 1 over 64 lanes.  FP64 values move by DPP whole-wave shifts (wave_shr:1 / wave_shl:1,
// two VALU moves) instead of ds_bpermute through the LDS crossbar: the FP64 sweeps and the
// reverse sweeps leave VALU cycles idle while they wait on the LDS.  FP32 values keep
// ds_bpermute: the FP32 x sweep is VALU-bound and measured 10 % slower with the DPP moves.
// Every lane of the wavefront must be active.
__device__ inline int dpp_up1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xF, 0xF, false); }
__device__ inline int dpp_dn1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xF, 0xF, false); }
__device__ inline float lane_up(float v) { return __shfl_up(v, 1, 64); }
__device__ inline float lane_down(float v) { return __shfl_down(v, 1, 64); }
__device__ inline double lane_up(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)dpp_up1((int)(unsigned)b), hi = (unsigned)dpp_up1((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline double lane_down(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)dpp_dn1((int)(unsigned)b), hi = (unsigned)dpp_dn1((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <typename R, int RECON>
__device__ inline void recon_lane(const Tabs &tb, int ax, int jc, R c, R &vm, R &vp)
{
    R cm1 = lane_up(c);
    R cp1 = lane_down(c);
    if (RECON == 5) {
        weno5_edges<R>(__shfl_up(c, 2, 64), cm1, c, cp1, __shfl_down(c, 2, 64), vm, vp);
    } else if (RECON == 4) {
        R dm = ppm_slope<R>(tb.atr<R>(ax, TB_B0C, jc), tb.atr<R>(ax, TB_B1C, jc), cm1, c, cp1);
        R dmm1 = lane_up(dm);
        R dmp1 = lane_down(dm);
        ppm_edges<R>(tb.atr<R>(ax, TB_A0M, jc), tb.atr<R>(ax, TB_A1M, jc), tb.atr<R>(ax, TB_A2M, jc),
                     tb.atr<R>(ax, TB_A3M, jc), tb.atr<R>(ax, TB_A0P, jc), tb.atr<R>(ax, TB_A1P, jc),
                     tb.atr<R>(ax, TB_A2P, jc), tb.atr<R>(ax, TB_A3P, jc), cm1, c, cp1, dmm1, dm, dmp1, vm, vp);
    } else {
        tvd_any<R, RECON>(tb.atr<R>(ax, TB_CF, jc), tb.atr<R>(ax, TB_CB, jc), cm1, c, cp1, vm, vp);
    }
}

// One 256-thread tile of the x sweep: tile bx of plane l.  Tiles per plane:
// ceil(ceil((ny+2)*px / XS_FACES) * 64 / 256), planes l in [0, nz+1].
template <typename R, int RECON, int SOLVER = 2>
__device__ inline void flux_x_tile(const Lay &L, const Fields<R> &F, const Tabs &tb, const Model &M, int bx, int l)
{
    const int lane = threadIdx.x & 63;
    const long wave = ((long)bx * blockDim.x + threadIdx.x) >> 6;
    const long f = wave * XS_FACES + lane - 2;  // flattened (j+2) + k*px over rows k in [0, ny+1]
    const long nload = (long)(L.ny + 4) * L.px;
    const bool inr = (f >= 0) && (f < nload);
    const long g = (long)(l + 2) * L.py + 2 * L.px + (inr ? f : 0);
    const int k = inr ? (int)(f / L.px) : 0;
    const int jj = inr ? (int)(f - (long)k * L.px) : 0;
    const int j = jj - 2;
    const bool active = lane >= 2 && lane <= 60 && inr && k <= L.ny + 1 && j >= 0 && j <= L.nx;
    const int jc = j < -2 ? -2 : (j > L.nx + 3 ? L.nx + 3 : j);

    const unsigned gu = (unsigned)g;
    const long ps = L.ps;
    const R *SP = sweep_p(F), *SB = sweep_bc(F);
    R q[7];
    q[0] = ldi(SP, gu);
    q[1] = ldi(SP + ps, gu);
    q[2] = ldi(SP + 2 * ps, gu);
    q[3] = ldi(SP + 3 * ps, gu);
    q[4] = ldi(SP + 4 * ps, gu);
    q[5] = ldi(SB + ps, gu);       // By cell
    q[6] = ldi(SB + 2 * ps, gu);   // Bz cell
    const R bn = ldi(SP + 5 * ps, gu);  // Bx on face j

    R vl[7], vr[7];
    if (RECON == 5) {
        R vm[7];
#pragma unroll
        for (int v = 0; v < 7; v++) recon_lane<R, RECON>(tb, 0, jc, q[v], vm[v], vl[v]);
        weno_positivity<R>(q, vm, vl);
#pragma unroll
        for (int v = 0; v < 7; v++) vr[v] = lane_down(vm[v]);
    } else {
#pragma unroll
        for (int v = 0; v < 7; v++) {
            R vm, vp;
            recon_lane<R, RECON>(tb, 0, jc, q[v], vm, vp);
            vl[v] = vp;
            vr[v] = lane_down(vm);
        }
    }
    FaceState<R> sl{vl[0], vl[1], vl[2], vl[3], vl[4], bn, vl[5], vl[6]};
    FaceState<R> sr{vr[0], vr[1], vr[2], vr[3], vr[4], bn, vr[5], vr[6]};
    R fl[7];
    riemann_face<1, R, SOLVER>(R(M.ggas), sl, sr, fl);
    if (active) {
        R *o = F.F[0];
#pragma unroll
        for (int c = 0; c < 7; c++) sti(o + c * ps, gu, fl[c]);
    }
}

template <typename R, int RECON, int SOLVER>
__global__ __launch_bounds__(256) void k_flux_x(Lay L, Fields<R> F, Tabs tb, Model M)
{
    flux_x_tile<R, RECON, SOLVER>(L, F, tb, M, blockIdx.x, blockIdx.y);
}

// ------------------------------------------------------------------------------------------
// y sweep of a box with few y pencils (a 2D box, or a small 3D one, whose pencil walks would
// leave most of the GPU idle), wave-parallel like the x sweep: a wavefront holds 64 consecutive
// cells of one y pencil, neighbours by lane moves, and emits its 59 faces; the four wavefronts
// of a workgroup take neighbouring pencils j, so the rows they load (stride px) meet in the same
// 128-byte lines.  The same reconstruction and Riemann solve as the pencil walk, bit for bit.
// Tile bx of plane l: 4-pencil group bx % njt (j in [0, nx+1]) of segment bx / njt.
// A single context takes this form when its walks would run fewer than YLANE_WALKERS threads
// (one wavefront per SIMD of the 256 CUs); with more walkers the walks keep up (the batched
// 48-box workload, 1.2 M cells in 104 K walkers, measured 51.6 ms for 40 steps with this form in
// XCD order against 51.7 ms with the walks, and 56.5 ms with this form in launch order).
// ------------------------------------------------------------------------------------------
constexpr long YLANE_WALKERS = 1024L * 64;
__host__ __device__ inline int ylane_groups(const Lay &L) { return (L.nx + 2 + 3) / 4; }
__host__ __device__ inline int ylane_segs(const Lay &L) { return (L.ny + 1 + 2 + XS_FACES - 1) / XS_FACES; }

template <typename R, int RECON, int SOLVER = 2>
__device__ inline void flux_ylane_tile(const Lay &L, const Fields<R> &F, const Tabs &tb, const Model &M, int bx, int l)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int njt = ylane_groups(L);
    const int seg = bx / njt, j = (bx % njt) * 4 + w;
    const int k = seg * XS_FACES + lane - 4;  // this lane's cell; lanes 2..60 own face k + 1/2
    // loads stay on the stored cells k in [-2, ny+3] (lanes outside feed no owned face); every
    // lane runs the lane moves
    const int kl = k < -2 ? -2 : (k > L.ny + 3 ? L.ny + 3 : k);
    const int jl = j > L.nx + 1 ? L.nx + 1 : j;
    const int kc = k < -2 ? -2 : (k > L.ny + 3 ? L.ny + 3 : k);
    const bool active = lane >= 2 && lane <= 60 && k >= 0 && k <= L.ny && j <= L.nx + 1;
    const unsigned gu = (unsigned)L.id(jl, kl, l);
    const long ps = L.ps;
    const R *SP = sweep_p(F), *SB = sweep_bc(F);
    R q[7];
    q[0] = ldi(SP, gu);
    q[1] = ldi(SP + ps, gu);
    q[2] = ldi(SP + 2 * ps, gu);
    q[3] = ldi(SP + 3 * ps, gu);
    q[4] = ldi(SP + 4 * ps, gu);
    q[5] = ldi(SB, gu);           // Bx cell
    q[6] = ldi(SB + 2 * ps, gu);  // Bz cell
    const R bn = ldi(SP + 6 * ps, gu);  // By on face k

    R vl[7], vr[7];
    if (RECON == 5) {
        R vm[7];
#pragma unroll
        for (int v = 0; v < 7; v++) recon_lane<R, RECON>(tb, 1, kc, q[v], vm[v], vl[v]);
        weno_positivity<R>(q, vm, vl);
#pragma unroll
        for (int v = 0; v < 7; v++) vr[v] = lane_down(vm[v]);
    } else {
#pragma unroll
        for (int v = 0; v < 7; v++) {
            R vm, vp;
            recon_lane<R, RECON>(tb, 1, kc, q[v], vm, vp);
            vl[v] = vp;
            vr[v] = lane_down(vm);
        }
    }
    // rotated frame n = y, t1 = z, t2 = x; reconstructed cell B order (x, z)
    FaceState<R> sl{vl[0], vl[2], vl[3], vl[1], vl[4], bn, vl[6], vl[5]};
    FaceState<R> sr{vr[0], vr[2], vr[3], vr[1], vr[4], bn, vr[6], vr[5]};
    R fl[7];
    riemann_face<2, R, SOLVER>(R(M.ggas), sl, sr, fl);
    if (active) {
        R *o = F.F[1];
        sti(o, gu, fl[0]);
        sti(o + 1 * ps, gu, fl[3]);
        sti(o + 2 * ps, gu, fl[1]);
        sti(o + 3 * ps, gu, fl[2]);
        sti(o + 4 * ps, gu, fl[4]);
        sti(o + 5 * ps, gu, fl[6]);  // F(Bx)
        sti(o + 6 * ps, gu, fl[5]);  // F(Bz)
    }
}



// ------------------------------------------------------------------------------------------
// y and z sweeps: each thread walks one pencil along the sweep axis, lanes along x, keeping the
// 4-cell value window and 3 limited slopes of every reconstructed variable in registers so that
// each slope, each PPM call and each Riemann solve is computed exactly once.
// (solver/spatial.f90:474-511 and :552-589)
// ------------------------------------------------------------------------------------------
// One 128-thread tile: bx covers px*(nt+2) pencils, segment by of seglen faces along DIR.
template <typename R, int DIR, int RECON, int SOLVER = 2>
__device__ inline void flux_walk_tile(const Lay &L, const Fields<R> &F, const Tabs &tb, const Model &M, int seglen,
                                      int bx, int by)
{
    const long t = (long)bx * blockDim.x + threadIdx.x;
    const int nd = DIR == 2 ? L.ny : L.nz;   // cells along the sweep
    const int nt = DIR == 2 ? L.nz : L.ny;   // second transverse extent
    const int jj = (int)(t % L.px);
    const long tq = t / L.px;                // transverse index 0 .. nt+1
    const int j = jj - 2;
    if (tq > nt + 1 || j < 0 || j > L.nx + 1) return;
    const int s0 = by * seglen;
    if (s0 > nd) return;
    const int s1 = (s0 + seglen - 1 < nd) ? s0 + seglen - 1 : nd;  // last face of the segment
    const unsigned stride = (unsigned)(DIR == 2 ? L.px : L.py);
    const unsigned base = (unsigned)(DIR == 2 ? (long)(tq + 2) * L.py + 2 + j : (long)(tq + 2) * L.px + 2 + j);
    auto at = [&](int c) { return base + (unsigned)(c + 2) * stride; };
    const long ps = L.ps;
    const int ax = DIR - 1;
    const int bA = DIR == 2 ? 0 : 0;        // first transverse cell-B (x order): y-sweep Bx, z-sweep Bx
    const int bB = DIR == 2 ? 2 : 1;        //                                     y-sweep Bz, z-sweep By
    const int ibn = DIR == 2 ? 6 : 7;       // face normal B plane in P

    auto load7 = [&](int c, R *o) {
        const unsigned g = at(c);
        const R *SP = sweep_p(F), *SB = sweep_bc(F);
        o[0] = ldi(SP, g);
        o[1] = ldi(SP + ps, g);
        o[2] = ldi(SP + 2 * ps, g);
        o[3] = ldi(SP + 3 * ps, g);
        o[4] = ldi(SP + 4 * ps, g);
        o[5] = ldi(SB + bA * ps, g);
        o[6] = ldi(SB + bB * ps, g);
    };

    if constexpr (DIR == 3 && RECON != 5 && sizeof(R) == 8) {
        // z-collapsed box (z stride 0): the one z face of the column sees the cell on both
        // sides and a stencil of copies of it, so its edge values take the closed form of
        // const_edges whenever the seven values are finite
        if (L.zc1) {
            R w[7];
            load7(0, w);
            bool fin = true;
#pragma unroll
            for (int v = 0; v < 7; v++) fin = fin && isfinite((double)w[v]);
            if (fin) {
                R vm[7], vp[7], fl[7];
#pragma unroll
                for (int v = 0; v < 7; v++) const_edges<R, RECON>(w[v], vm[v], vp[v]);
                const unsigned gf = at(0);
                const R bn = ldi(sweep_p(F) + ibn * ps, gf);
                FaceState<R> sl{vp[0], vp[3], vp[1], vp[2], vp[4], bn, vp[5], vp[6]};
                FaceState<R> sr{vm[0], vm[3], vm[1], vm[2], vm[4], bn, vm[5], vm[6]};
                riemann_face<3, R, SOLVER>(R(M.ggas), sl, sr, fl);
                R *o = F.F[2];
                sti(o, gf, fl[0]);
                sti(o + 1 * ps, gf, fl[2]);
                sti(o + 2 * ps, gf, fl[3]);
                sti(o + 3 * ps, gf, fl[1]);
                sti(o + 4 * ps, gf, fl[4]);
                sti(o + 5 * ps, gf, fl[5]);  // F(Bx)
                sti(o + 6 * ps, gf, fl[6]);  // F(By)
                return;
            }
        }
    }
    // window: w0 = v(c-1), w1 = v(c), w2 = v(c+1), w3 = v(c+2) ; slopes d0 = dm(c-1), d1 = dm(c), d2 = dm(c+1)
    R w0[7], w1[7], w2[7], w3[7], d0[7], d1[7], vpp[7];
    R wm[RECON == 5 ? 7 : 1];  // WENO5: the value two cells back
    {
        R a[7];
        load7(s0 - 2, a);
        load7(s0 - 1, w0);
        load7(s0, w1);
        load7(s0 + 1, w2);
        if (RECON == 5) {
#pragma unroll
            for (int v = 0; v < 7; v++) wm[v] = a[v];
        }
        if (RECON == 4) {
            R b0 = tb.atr<R>(ax, TB_B0C, s0 - 1), b1 = tb.atr<R>(ax, TB_B1C, s0 - 1);
            R c0 = tb.atr<R>(ax, TB_B0C, s0), c1 = tb.atr<R>(ax, TB_B1C, s0);
#pragma unroll
            for (int v = 0; v < 7; v++) {
                d0[v] = ppm_slope<R>(b0, b1, a[v], w0[v], w1[v]);
                d1[v] = ppm_slope<R>(c0, c1, w0[v], w1[v], w2[v]);
            }
        }
    }
    for (int c = s0; c <= s1 + 1; c++) {
        load7(c + 2, w3);
        R vm[7], vp[7];
        if (RECON == 5) {
#pragma unroll
            for (int v = 0; v < 7; v++) weno5_edges<R>(wm[v], w0[v], w1[v], w2[v], w3[v], vm[v], vp[v]);
            weno_positivity<R>(w1, vm, vp);
        } else if (RECON == 4) {
            R b0 = tb.atr<R>(ax, TB_B0C, c + 1), b1 = tb.atr<R>(ax, TB_B1C, c + 1);
            R a0m = tb.atr<R>(ax, TB_A0M, c), a1m = tb.atr<R>(ax, TB_A1M, c), a2m = tb.atr<R>(ax, TB_A2M, c);
            R a3m = tb.atr<R>(ax, TB_A3M, c), a0p = tb.atr<R>(ax, TB_A0P, c), a1p = tb.atr<R>(ax, TB_A1P, c);
            R a2p = tb.atr<R>(ax, TB_A2P, c), a3p = tb.atr<R>(ax, TB_A3P, c);
#pragma unroll
            for (int v = 0; v < 7; v++) {
                R d2 = ppm_slope<R>(b0, b1, w1[v], w2[v], w3[v]);
                ppm_edges<R>(a0m, a1m, a2m, a3m, a0p, a1p, a2p, a3p, w0[v], w1[v], w2[v], d0[v], d1[v], d2,
                             vm[v], vp[v]);
                d0[v] = d1[v];
                d1[v] = d2;
            }
        } else {
            R cf = tb.atr<R>(ax, TB_CF, c), cb = tb.atr<R>(ax, TB_CB, c);
#pragma unroll
            for (int v = 0; v < 7; v++) tvd_any<R, RECON>(cf, cb, w0[v], w1[v], w2[v], vm[v], vp[v]);
        }
        if (c > s0) {
            // face c-1: left state = vp(c-1), right state = vm(c)
            const unsigned gf = at(c - 1);
            const R bn = ldi(sweep_p(F) + ibn * ps, gf);
            R fl[7];
            if (DIR == 2) {
                // rotated frame n=y, t1=z, t2=x ; reconstructed cell B order (x, z)
                FaceState<R> sl{vpp[0], vpp[2], vpp[3], vpp[1], vpp[4], bn, vpp[6], vpp[5]};
                FaceState<R> sr{vm[0], vm[2], vm[3], vm[1], vm[4], bn, vm[6], vm[5]};
                riemann_face<2, R, SOLVER>(R(M.ggas), sl, sr, fl);
                R *o = F.F[1];
                sti(o, gf, fl[0]);
                sti(o + 1 * ps, gf, fl[3]);
                sti(o + 2 * ps, gf, fl[1]);
                sti(o + 3 * ps, gf, fl[2]);
                sti(o + 4 * ps, gf, fl[4]);
                sti(o + 5 * ps, gf, fl[6]);  // F(Bx)
                sti(o + 6 * ps, gf, fl[5]);  // F(Bz)
            } else {
                // rotated frame n=z, t1=x, t2=y ; reconstructed cell B order (x, y)
                FaceState<R> sl{vpp[0], vpp[3], vpp[1], vpp[2], vpp[4], bn, vpp[5], vpp[6]};
                FaceState<R> sr{vm[0], vm[3], vm[1], vm[2], vm[4], bn, vm[5], vm[6]};
                riemann_face<3, R, SOLVER>(R(M.ggas), sl, sr, fl);
                R *o = F.F[2];
                sti(o, gf, fl[0]);
                sti(o + 1 * ps, gf, fl[2]);
                sti(o + 2 * ps, gf, fl[3]);
                sti(o + 3 * ps, gf, fl[1]);
                sti(o + 4 * ps, gf, fl[4]);
                sti(o + 5 * ps, gf, fl[5]);  // F(Bx)
                sti(o + 6 * ps, gf, fl[6]);  // F(By)
            }
        }
#pragma unroll
        for (int v = 0; v < 7; v++) {
            if (RECON == 5) wm[v] = w0[v];
            vpp[v] = vp[v];
            w0[v] = w1[v];
            w1[v] = w2[v];
            w2[v] = w3[v];
        }
    }
}

template <typename R, int DIR, int RECON, int SOLVER>
__global__ __launch_bounds__(128) void k_flux_walk(Lay L, Fields<R> F, Tabs tb, Model M, int seglen)
{
    flux_walk_tile<R, DIR, RECON, SOLVER>(L, F, tb, M, seglen, blockIdx.x, blockIdx.y);
}

// ------------------------------------------------------------------------------------------
// Corner EMFs on [0,n]^3: face-flux averages accumulated in sweep order (spatial.f90:455-467,
// :533-546, :611-624) then the upwind-CT combination with cell EMFs (mhd.f90:95-125).
// 2.5D: a thread owns one (j,k) column and marches l through a chunk of zc planes (a tuning
// parameter); the plane-(l+1) values it loads are carried into the next iteration, so HBM sees
// each plane once.
// ------------------------------------------------------------------------------------------
constexpr int NXCD = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin over them

// Map the 1D launch index onto a logical (bx, by) tile so that logically consecutive tiles
// (adjacent rows k, which share their k-1 / k+1 halos) run on the same XCD and meet in its
// L2.  The launch has ceil(gx*gy/8)*8 blocks; returns false for the padding blocks.
__device__ inline bool xcd_tile(int gx, int gy, int &bx, int &by)
{
    const unsigned per = gridDim.x / NXCD;
    const unsigned logical = (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
    if (logical >= (unsigned)(gx * gy)) return false;
    bx = (int)(logical % (unsigned)gx);
    by = (int)(logical / (unsigned)gx);
    return true;
}

// Column (j,k) of this thread in column tile bx of the [0,nx] x [0,ny] column set.  tw = 0 takes
// 256 consecutive cells of the (j,k) plane of pitch px (about one row per tile); tw > 0 takes a
// tw x (256/tw) block of columns, so most of the k-1 / k+1 rows a column reads are loaded by the
// same workgroup.  Column kernels run 256-thread workgroups.
__device__ inline void col_of(const Lay &L, int bx, int tw, int &j, int &k)
{
    if (tw == 0) {
        const long t = (long)bx * 256 + threadIdx.x;
        k = (int)(t / L.px);
        j = (int)(t - (long)k * L.px);
    } else {
        const int ntx = (L.nx + tw) / tw;
        const int tx = bx % ntx, ty = bx / ntx;
        j = tx * tw + (int)(threadIdx.x % tw);
        k = ty * (256 / tw) + (int)(threadIdx.x / tw);
    }
}
// the y sweep's wave-parallel form (flux_ylane_tile): XCD-aware tile order, so that the
// 4-pencil groups sharing 128-byte rows run on one XCD and meet in its L2
template <typename R, int RECON, int SOLVER>
__global__ __launch_bounds__(256) void k_flux_ylane(Lay L, Fields<R> F, Tabs tb, Model M, int gx, int gy)
{
    int bx, by;
    if (!xcd_tile(gx, gy, bx, by)) return;
    flux_ylane_tile<R, RECON, SOLVER>(L, F, tb, M, bx, by);
}

__host__ __device__ inline int col_tiles(const Lay &L, int tw)
{
    if (tw == 0) return (int)((L.px * (L.ny + 1) + 255) / 256);
    return ((L.nx + tw) / tw) * ((L.ny + 256 / tw) / (256 / tw));
}

template <typename R>
__device__ inline void ecells(const Lay &L, const Fields<R> &F, long g, R &ex, R &ey, R &ez)
{
    const long ps = L.ps;
    const R *SP = sweep_p(F), *SB = sweep_bc(F);
    const R vx = SP[g + ps], vy = SP[g + 2 * ps], vz = SP[g + 3 * ps];
    const R bx = SB[g], by = SB[g + ps], bz = SB[g + 2 * ps];
    ex = vz * by - vy * bz;
    ey = vx * bz - vz * bx;
    ez = vy * bx - vx * by;
}

// Column tile (bx, by): 256 (j,k) columns of the flattened [0,nx]x[0,ny] plane (pitch px),
// planes [by*zc, by*zc+zc-1] of [0,nz].
template <typename R>
__device__ inline void ecorn_tile(const Lay &L, const Fields<R> &F, int bx, int by, int zc, int tw)
{
    int j, k;
    col_of(L, bx, tw, j, k);
    const int l0 = L.zc1 ? 1 : by * zc;
    if (k > L.ny || j > L.nx || l0 > L.nz) return;
    const int l1 = L.zc1 ? 1 : ((l0 + zc - 1 < L.nz) ? l0 + zc - 1 : L.nz);
    const long sj = 1, sk = L.px, sl = L.py, ps = L.ps;
    const R *Fx = F.F[0], *Fy = F.F[1], *Fz = F.F[2];
    const R q = R(0.25);
    long g = L.id(j, k, l0);
    // cell EMFs of plane l at (j,k), (j,k+1), (j+1,k), (j+1,k+1) and the plane-l B fluxes
    // that the next plane needs again
    R x00, y00, z00, x01, y01, z01, x10, y10, z10, x11, y11, z11;
    ecells(L, F, g, x00, y00, z00);
    ecells(L, F, g + sk, x01, y01, z01);
    ecells(L, F, g + sj, x10, y10, z10);
    ecells(L, F, g + sj + sk, x11, y11, z11);
    R fx6 = Fx[g + 6 * ps], fy6 = Fy[g + 6 * ps];
    for (int l = l0; l <= l1; l++, g += sl) {
        const long gn = g + sl;
        R X00, Y00, Z00, X01, Y01, Z01, X10, Y10, Z10, X11, Y11, Z11;
        ecells(L, F, gn, X00, Y00, Z00);
        ecells(L, F, gn + sk, X01, Y01, Z01);
        ecells(L, F, gn + sj, X10, Y10, Z10);
        ecells(L, F, gn + sj + sk, X11, Y11, Z11);
        const R fx6n = Fx[gn + 6 * ps], fy6n = Fy[gn + 6 * ps];
        // x sweep: ey += (F_x(Bz)(l) + F_x(Bz)(l+1))/4, ez += (-F_x(By)(k) - F_x(By)(k+1))/4
        R ey = R(0) + q * (fx6 + fx6n);
        R ez = R(0) + q * (-Fx[g + 5 * ps] - Fx[g + sk + 5 * ps]);
        // y sweep: ex += (-F_y(Bz)(l) - F_y(Bz)(l+1))/4, ez += (F_y(Bx)(j) + F_y(Bx)(j+1))/4
        R ex = R(0) + q * (-fy6 - fy6n);
        ez = ez + q * (Fy[g + 5 * ps] + Fy[g + sj + 5 * ps]);
        // z sweep: ex += (F_z(By)(k) + F_z(By)(k+1))/4, ey += (-F_z(Bx)(j) - F_z(Bx)(j+1))/4
        ex = ex + q * (Fz[g + 6 * ps] + Fz[g + sk + 6 * ps]);
        ey = ey + q * (-Fz[g + 5 * ps] - Fz[g + sj + 5 * ps]);
        // flux_ct
        ex = ex * R(2.0) - q * (x00 + X00 + x01 + X01);
        ey = ey * R(2.0) - q * (y00 + Y00 + y10 + Y10);
        ez = ez * R(2.0) - q * (z00 + z01 + z10 + z11);
        F.E[g] = ex;
        F.E[g + ps] = ey;
        F.E[g + 2 * ps] = ez;
        x00 = X00; y00 = Y00; z00 = Z00; x01 = X01; y01 = Y01; z01 = Z01;
        x10 = X10; y10 = Y10; z10 = Z10; x11 = X11; y11 = Y11; z11 = Z11;
        fx6 = fx6n;
        fy6 = fy6n;
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_ecorn(Lay L, Fields<R> F, int gx, int gy, int zc, int tw)
{
    int bx, by;
    if (!xcd_tile(gx, gy, bx, by)) return;
    ecorn_tile<R>(L, F, bx, by, zc, tw);
}

// ------------------------------------------------------------------------------------------
// Stage tail on [0,n]^3, one pass: flux divergence + sources (spatial.f90:436-452, :365-375,
// coords.f90:105-137, torus_cyn custom.f90:121-166), face-B right-hand side (mhd.f90:130-153),
// the SSP-RK3 combination (rungekutta.f90:41-51, :75-85, :109-119), FROMUTORVE
// (fromrvetou.f90:75-127, utop.f90:185-220), CUSTOM_CHECKRHO, FROMRVETOU (ptou.f90:33-66),
// UPDATE's EOS (update.f90:12-45) and, in stage 3, the next finddt (rungekutta.f90:205-252).
// A thread marches one (j,k) column over a chunk of planes.  Cell-centred B needs the new
// face B of the lower neighbours; those are recomputed from the same inputs (bit-identical to
// their owners), which is why the stage reads Uin and writes a different buffer Uout.
// ------------------------------------------------------------------------------------------
constexpr double RK20 = 3.0 / 4.0, RK21 = 1.0 / 4.0, RK22 = 1.0 / 4.0;
constexpr double RK30 = 1.0 / 3.0, RK31 = 2.0 / 3.0, RK32 = 2.0 / 3.0;

template <int STAGE>
__device__ inline double rk(double u0, double u1, double dt, double lrk)
{
    if (STAGE == 1) return u0 + dt * lrk;
    if (STAGE == 2) return RK20 * u0 + RK21 * u1 + RK22 * dt * lrk;
    return RK30 * u0 + RK31 * u1 + RK32 * dt * lrk;
}

// Combined stage input (COMB).  Stage 2 uses its input state u1 only in
// (RK20 u0 + RK21 u1) + RK22 dt L, so in a plain forward step stage 1, which has u0 and u1 at
// hand, stores w2 = RK20 u0 + RK21 u1 in place of u1, and stage 2 adds its right-hand side to w2
// by the same operations rk performs, bit for bit, reading 8 FP64 planes instead of 16; stage 2
// stores the plain u2 and stage 3 is unchanged.  (Forming RK30 u0 + RK31 u2 in stage 2 as well
// made stage 2 load u0 after its arithmetic, which measured slower in the mixed form.)  Taped
// steps, whose stage states are stored for the reverse pass, keep the plain states.
__device__ inline double rk_w2(double w, double dt, double lrk) { return w + RK22 * dt * lrk; }
template <int STAGE, bool COMB>
__device__ inline double comb_out(double u0, double u)
{
    if (COMB && STAGE == 1) return RK20 * u0 + RK21 * u;
    return u;
}

// right-hand sides of the three face fields of cell (j,k,l) (mhd.f90:130-153, GEOM_CT of
// coords.f90:338-386 evaluated at the cell)
template <int COORD>
__device__ inline double rhs_bx(const Tabs &tb, int j, int k, int l, double ez, double ez_km, double ey, double ey_lm)
{
    if (COORD == 3) {
        const double cez = tb.at(0, TB_F, j) * tb.at(1, TB_CD, k);
        const double cey = tb.at(0, TB_F, j) * tb.at(1, TB_CD, k) * tb.at(2, TB_D, l) / tb.at(1, TB_D, k);
        return -((ez * tb.at(1, TB_SIN, k) - ez_km * tb.at(1, TB_SIN, k - 1)) / cez - (ey * 1.0 - ey_lm * 1.0) / cey);
    }
    const double cez = COORD == 2 ? tb.at(1, TB_D, k) * tb.at(0, TB_F, j) : tb.at(1, TB_D, k);
    return -((ez * 1.0 - ez_km * 1.0) / cez - (ey * 1.0 - ey_lm * 1.0) / tb.at(2, TB_D, l));
}
template <int COORD>
__device__ inline double rhs_by(const Tabs &tb, int j, int k, int l, double ex, double ex_lm, double ez, double ez_jm)
{
    if (COORD == 3) {
        const double xl = tb.at(0, TB_C, j);
        return -((ex * 1.0 - ex_lm * 1.0) / (xl * tb.at(1, TB_SIN, k) * tb.at(2, TB_D, l)) -
                 (ez * tb.at(0, TB_F, j) - ez_jm * tb.at(0, TB_F, j - 1)) / (xl * tb.at(0, TB_D, j)));
    }
    return -((ex * 1.0 - ex_lm * 1.0) / tb.at(2, TB_D, l) - (ez * 1.0 - ez_jm * 1.0) / tb.at(0, TB_D, j));
}
template <int COORD>
__device__ inline double rhs_bz(const Tabs &tb, int j, int k, double ey, double ey_jm, double ex, double ex_km)
{
    if (COORD == 2 || COORD == 3) {
        const double xl = tb.at(0, TB_C, j);
        return -((ey * tb.at(0, TB_F, j) - ey_jm * tb.at(0, TB_F, j - 1)) / (xl * tb.at(0, TB_D, j)) -
                 (ex * 1.0 - ex_km * 1.0) / (xl * tb.at(1, TB_D, k)));
    }
    return -((ey * 1.0 - ey_jm * 1.0) / tb.at(0, TB_D, j) - (ex * 1.0 - ex_km * 1.0) / tb.at(1, TB_D, k));
}

// CFL time step of one cell (rungekutta.f90:213-247; Q8: no discriminant clamp)
template <int COORD>
__device__ inline double cell_dt(const Tabs &tb, const Model &M, int j, int k, int l, double rho, double vx, double vy,
                                 double vz, double p, double bx, double by, double bz)
{
    const double cs = sqrt(M.ggas * p / rho);
    const double a2 = cs * cs;
    const double a4 = a2 * a2;
    const double b2x = (bx * bx / rho), b2y = (by * by / rho), b2z = (bz * bz / rho);
    const double b2 = b2x + b2y + b2z;
    const double b4 = b2 * b2;
    const double cfx = sqrt(0.5 * (a2 + b2 + sqrt((a4 + 2.0 * a2 * b2 + b4) - 4.0 * a2 * b2x)));
    const double cfy = sqrt(0.5 * (a2 + b2 + sqrt((a4 + 2.0 * a2 * b2 + b4) - 4.0 * a2 * b2y)));
    const double cfz = sqrt(0.5 * (a2 + b2 + sqrt((a4 + 2.0 * a2 * b2 + b4) - 4.0 * a2 * b2z)));
    const double l1 = fabs(vx) + cfx, l2 = fabs(vy) + cfy, l3 = fabs(vz) + cfz;
    const double lambda = fmax2<double>(fmax2<double>(l1, l2), l3);
    double dxl = tb.at(0, TB_D, j), dyl = tb.at(1, TB_D, k), dzl = tb.at(2, TB_D, l);
    if (COORD == 2) dyl = tb.at(0, TB_C, j) * dyl;
    if (COORD == 3) {
        dyl = tb.at(0, TB_C, j) * dyl;
        dzl = tb.at(0, TB_C, j) * tb.at(1, TB_SINC, k) * dzl;
    }
    const double dt_out = fmin2<double>(fmin2<double>(dxl, dyl), dzl) * M.cfl / lambda;
    // a NaN cell step becomes 0: every min reduction then ends at 0 (a NaN would drop out of
    // the comparisons), and the host reports the non-finite state
    const double d = fmin2<double>(dt_out, 1.0e10);
    return d == d ? d : 0.0;
}

__device__ inline double block_min(double v)
{
    __shared__ double red[16];
    for (int o = 32; o > 0; o >>= 1) {
        double w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = red[w] < m ? red[w] : m;
    return m;
}

// Flux-divergence geometry of a column (coords.f90:255-282)
struct ColGeom {
    double dyl, gxp, gxm, gxc, gyp, gym, gyc, gac, xloc, tanc;
};
template <int COORD>
__device__ inline ColGeom col_geom(const Tabs &tb, int j, int k)
{
    ColGeom c;
    const double dxl = tb.at(0, TB_D, j);
    c.dyl = tb.at(1, TB_D, k);
    c.gxp = 1.0; c.gxm = 1.0; c.gxc = dxl; c.gyp = 1.0; c.gym = 1.0; c.gyc = c.dyl; c.gac = 0.0; c.xloc = 0.0;
    c.tanc = 1.0;
    if (COORD == 2) {
        c.xloc = tb.at(0, TB_C, j);
        c.gxp = tb.at(0, TB_F, j);
        c.gxm = tb.at(0, TB_F, j - 1);
        c.gxc = c.xloc * dxl;
        c.gyc = c.xloc * c.dyl;
    }
    if (COORD == 3) {
        c.xloc = tb.at(0, TB_C, j);
        c.tanc = tb.at(1, TB_TANC, k);
        c.gxp = tb.at(0, TB_F2, j);
        c.gxm = tb.at(0, TB_F2, j - 1);
        c.gxc = tb.at(0, TB_V3, j);
        c.gyp = tb.at(1, TB_SIN, k);
        c.gym = tb.at(1, TB_SIN, k - 1);
        c.gac = tb.at(0, TB_RA, j) * tb.at(1, TB_CD, k);
        c.gyc = c.gac;
    }
    return c;
}

// Corner EMFs around cell (j,k,l): own (j,k), (j,k-1), (j-1,k), (j-1,k-1) in plane l
struct CellE {
    double ex0, ey0, ez0, exk, ezk, eyj, ezj, ezjk;
};
// What the update of cell (j,k,l) takes over from cell (j,k,l-1): the plane-(l-1) EMFs
// ex(j,k), ey(j,k), ey(j-1,k), ex(j,k-1), the new face B_z below the cell and the z-face fluxes
// below it
struct TailCarry {
    double exL, eyL, eyjL, exkL, bzL;
    double fzL[5];
};

// Stage tail of one cell of [0,n]^3: flux divergence + sources, face-B right-hand side, RK3
// combination, U->P, floors, P->U, EOS, and (stage 3) its CFL step.  fz holds the z-face fluxes
// above the cell; cr carries from the cell below and is advanced to this cell.
template <typename R, int STAGE, int COORD, bool COMB>
__device__ inline void tail_cell(const Lay &L, const Fields<R> &F, const Tabs &tb, const Model &M, double dt,
                                 const double *__restrict__ Uin, double *__restrict__ Uout, int j, int k, int l,
                                 unsigned g, const ColGeom &cg, const CellE &E, TailCarry &cr, const double fz[5],
                                 double &dtmin)
{
    const unsigned sj = 1, sk = (unsigned)L.px;
    const long ps = L.ps;
    const double *U0 = F.U0;
    // plane c of each field at cell offset i (uniform plane bases, 32-bit lane offsets)
    auto fx = [&](int c, unsigned i) { return (double)ldi(F.F[0] + c * ps, i); };
    auto fy = [&](int c, unsigned i) { return (double)ldi(F.F[1] + c * ps, i); };
    auto u0 = [&](int c, unsigned i) { return ldi(U0 + c * ps, i); };
    auto ui = [&](int c, unsigned i) { return STAGE == 1 ? 0.0 : ldi(Uin + c * ps, i); };
    // the stage's RK combination of component c at offset i with right-hand side r
    auto rkc = [&](int c, unsigned i, double r) {
        if constexpr (COMB && STAGE == 2) return rk_w2(ui(c, i), dt, r);
        else return rk<STAGE>(u0(c, i), ui(c, i), dt, r);
    };
    // stage-input primitives: the FP64 planes, or the FP32 ones of the mixed form
    constexpr bool f32p = sizeof(R) == 4;
    auto pp = [&](int c, unsigned i) { return f32p ? (double)ldi(F.P32 + c * ps, i) : ldi(F.P + c * ps, i); };
    auto bcl = [&](int c, unsigned i) { return f32p ? (double)ldi(F.BC32 + c * ps, i) : ldi(F.BC + c * ps, i); };
    const double gm1 = M.ggas - 1.0;
    const bool interior = j >= 1 && k >= 1 && l >= 1;
    const double ex0 = E.ex0, ey0 = E.ey0, ez0 = E.ez0, exk = E.exk, ezk = E.ezk, eyj = E.eyj, ezj = E.ezj,
                 ezjk = E.ezjk;
    // new face fields: own three, and the x / y faces of the lower neighbours
    const double bxn = rkc(5, g, rhs_bx<COORD>(tb, j, k, l, ez0, ezk, ey0, cr.eyL));
    const double byn = rkc(6, g, rhs_by<COORD>(tb, j, k, l, ex0, cr.exL, ez0, ezj));
    const double bzn = rkc(7, g, rhs_bz<COORD>(tb, j, k, ey0, eyj, ex0, exk));
    // hydro right-hand side on interior cells
    double lrk[5];
    if (interior) {
        double sc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (COORD == 2 || COORD == 3 || M.model != 0) {
            // (the mixed form takes the density from the stage-input U, which holds the
            // FP64 value, so that the density-floor switch of the model gravity sees
            // what the FP64 form sees; it runs COMB only without sources and model hooks)
            static_assert(!(COMB && sizeof(R) == 4 && COORD != 1), "combined inputs with sources in the mixed form");
            const double rho = f32p ? (STAGE == 1 ? u0(0, g) : ui(0, g)) : pp(0, g);
            const double vx = pp(1, g), vy = pp(2, g), vz = pp(3, g);
            const double p = pp(4, g);
            const double bx = bcl(0, g), by = bcl(1, g), bz = bcl(2, g);
            if (COORD == 2) {
                const double bsq = bx * bx + by * by + bz * bz;
                const double tx = (p + rho * vy * vy + 0.5 * bsq - by * by) / (cg.xloc);
                const double ty = (rho * vx * vy - bx * by) / (cg.xloc);
                for (int r = 0; r < M.geom_repeat; r++) {
                    sc[IVX] = sc[IVX] + tx;
                    sc[IVY] = sc[IVY] - ty;
                }
            }
            if (COORD == 3) {
                // coords.f90:139-156
                const double bsq = bx * bx + by * by + bz * bz;
                const double tx = (2.0 * p + bx * bx + rho * (vy * vy + vz * vz)) / (cg.xloc);
                const double ty1 = (p + rho * vz * vz + 0.5 * bsq - bz * bz) / (cg.xloc * cg.tanc);
                const double ty2 = (rho * vx * vy - bx * by) / (cg.xloc);
                const double tz1 = (rho * vx * vz - bx * bz) / (cg.xloc);
                const double tz2 = (rho * vy * vz - by * bz) / (cg.xloc * cg.tanc);
                for (int r = 0; r < M.geom_repeat; r++) {
                    sc[IVX] = sc[IVX] + tx;
                    sc[IVY] = sc[IVY] + ty1 - ty2;
                    sc[IVZ] = sc[IVZ] - tz1 - tz2;
                }
            }
            if (M.model == 2) {
                // torus_sph/src/custom.f90:121-158: radial Paczynski-Wiita gravity
                const double xl = tb.at(0, TB_C, j);
                const double factor = fmax2<double>(sign1<double>(rho - M.rho_floor), 0.0);
                const double dphidr = 1.0 / ((xl - M.r_sh) * (xl - M.r_sh));
                sc[IVX] = sc[IVX] + (-factor * rho * dphidr);
                sc[ITAU] = sc[ITAU] + (-factor * rho * vx * dphidr);
            }
            if (M.model == 1) {
                const double xl = tb.at(0, TB_C, j), zl = tb.at(2, TB_C, l);
                const double factor = fmax2<double>(sign1<double>(rho - M.rho_floor), 0.0);
                const double radius = sqrt(xl * xl + zl * zl);
                const double dphidr = xl / radius / ((radius - M.r_sh) * (radius - M.r_sh));
                const double dphidz = zl / radius / ((radius - M.r_sh) * (radius - M.r_sh));
                sc[IVX] = sc[IVX] + (-factor * rho * dphidr);
                sc[IVZ] = sc[IVZ] + (-factor * rho * dphidz);
                sc[ITAU] = sc[ITAU] + (-factor * rho * (vx * dphidr + vz * dphidz));
            }
        }
        const double dzl = tb.at(2, TB_D, l);
        const double gzc = COORD == 3 ? cg.gac * dzl / cg.dyl : dzl;
#pragma unroll
        for (int i = 0; i < 5; i++) {
            double a = 0.0;
            a = a - (fx(i, g) * cg.gxp - fx(i, g - sj) * cg.gxm) / cg.gxc;
            a = a - (fy(i, g) * cg.gyp - fy(i, g - sk) * cg.gym) / cg.gyc;
            a = a - (fz[i] * 1.0 - cr.fzL[i] * 1.0) / gzc;
            lrk[i] = a + sc[i];
        }
    } else {
#pragma unroll
        for (int i = 0; i < 5; i++) lrk[i] = 0.0;
    }
    double u[5];
#pragma unroll
    for (int i = 0; i < 5; i++) u[i] = rkc(i, g, lrk[i]);
    // cell-centred field from the new face fields (interior) or the ghost value
    double bcx, bcy, bcz;
    if (interior) {
        const double bxm = rkc(5, g - sj, rhs_bx<COORD>(tb, j - 1, k, l, ezj, ezjk, eyj, cr.eyjL));
        const double bym = rkc(6, g - sk, rhs_by<COORD>(tb, j, k - 1, l, exk, cr.exkL, ezk, ezjk));
        bcx = 0.5 * (bxn + bxm);
        bcy = 0.5 * (byn + bym);
        bcz = 0.5 * (bzn + cr.bzL);
        if (f32p) {
            sti(F.BC32, g, (float)bcx);
            sti(F.BC32 + ps, g, (float)bcy);
            sti(F.BC32 + 2 * ps, g, (float)bcz);
        } else {
            sti(F.BC, g, bcx);
            sti(F.BC + ps, g, bcy);
            sti(F.BC + 2 * ps, g, bcz);
        }
    } else {
        bcx = bcl(0, g);
        bcy = bcl(1, g);
        bcz = bcl(2, g);
    }
    // U_to_P, floors, P_to_U, EOS
    const double rho0 = u[0];
    double vx = u[1] / rho0, vy = u[2] / rho0, vz = u[3] / rho0;
    const double bsq = bcx * bcx + bcy * bcy + bcz * bcz;
    double vsq = vx * vx + vy * vy + vz * vz;
    double eps = (u[4] - 0.5 * bsq) / rho0 - 0.5 * vsq;
    double rho = rho0;
    if (interior) {
        if (M.model != 0) {
            if (br(rho < M.rho_floor)) rho = M.rho_floor;
            if (br(eps < M.eps_floor)) eps = M.eps_floor;
        } else {
            eps = fmax2b<double>(eps, 0.0);
        }
    }
    vsq = vx * vx + vy * vy + vz * vz;
    const double p = rho * eps * gm1;
    auto uo = [&](int c, double v) { sti(Uout + c * ps, g, comb_out<STAGE, COMB>(COMB && STAGE == 1 ? u0(c, g) : 0.0, v)); };
    uo(0, rho);
    uo(1, vx * rho);
    uo(2, vy * rho);
    uo(3, vz * rho);
    uo(4, rho * (eps + 0.5 * vsq) + 0.5 * bsq);
    uo(5, bxn);
    uo(6, byn);
    uo(7, bzn);
    if (f32p) {
        sti(F.P32, g, (float)rho);
        sti(F.P32 + ps, g, (float)vx);
        sti(F.P32 + 2 * ps, g, (float)vy);
        sti(F.P32 + 3 * ps, g, (float)vz);
        sti(F.P32 + 4 * ps, g, (float)p);
        sti(F.P32 + 5 * ps, g, (float)bxn);
        sti(F.P32 + 6 * ps, g, (float)byn);
        sti(F.P32 + 7 * ps, g, (float)bzn);
        sti(F.EPS32, g, (float)eps);
    } else {
        sti(F.P, g, rho);
        sti(F.P + ps, g, vx);
        sti(F.P + 2 * ps, g, vy);
        sti(F.P + 3 * ps, g, vz);
        sti(F.P + 4 * ps, g, p);
        sti(F.P + 5 * ps, g, bxn);
        sti(F.P + 6 * ps, g, byn);
        sti(F.P + 7 * ps, g, bzn);
        sti(F.EPS, g, eps);
    }
    if (STAGE == 3 && interior)
        dtmin = fmin2<double>(cell_dt<COORD>(tb, M, j, k, l, rho, vx, vy, vz, p, bcx, bcy, bcz), dtmin);
    // carries
    cr.exL = ex0;
    cr.eyL = ey0;
    cr.eyjL = eyj;
    cr.exkL = exk;
    cr.bzL = bzn;
#pragma unroll
    for (int i = 0; i < 5; i++) cr.fzL[i] = fz[i];
}

// Column tile (bx, by) as in ecorn_tile.  Every thread of the block reaches the stage-3 block
// reduction at the end, so callers may loop over tiles with block-uniform trip counts.  In the
// mixed form (R = float) the stage-input primitives come from, and the new ones go to, the FP32
// planes P32 / EPS32 / BC32.
template <typename R, int STAGE, int COORD, bool COMB>
__device__ inline void tail_tile(const Lay &L, const Fields<R> &F, const Tabs &tb, const Model &M, const double *dtp,
                                 const double *__restrict__ Uin, double *__restrict__ Uout,
                                 unsigned long long *dtnext, int bx, int by, int zc, int tw)
{
    int j, k;
    col_of(L, bx, tw, j, k);
    const int l0 = L.zc1 ? 1 : by * zc;
    const bool col = k <= L.ny && j <= L.nx && l0 <= L.nz;
    double dtmin = 1.0e10;
    if (col) {
        const int l1 = L.zc1 ? 1 : ((l0 + zc - 1 < L.nz) ? l0 + zc - 1 : L.nz);
        const unsigned sj = 1, sk = (unsigned)L.px, sl = (unsigned)L.py;
        const long ps = L.ps;
        const double dt = *dtp;
        auto fzz = [&](int c, unsigned i) { return (double)ldi(F.F[2] + c * ps, i); };
        auto e = [&](int c, unsigned i) { return (double)ldi(F.E + c * ps, i); };
        auto u0 = [&](int c, unsigned i) { return ldi(F.U0 + c * ps, i); };
        auto ui = [&](int c, unsigned i) { return STAGE == 1 ? 0.0 : ldi(Uin + c * ps, i); };
        const ColGeom cg = col_geom<COORD>(tb, j, k);
        // carries from plane l0-1
        unsigned g = (unsigned)L.id(j, k, l0);
        const unsigned gm = g - sl;
        TailCarry cr;
        cr.exL = e(0, gm);
        cr.eyL = e(1, gm);
        cr.eyjL = e(1, gm - sj);
        cr.exkL = e(0, gm - sk);
        cr.bzL = 0.0;
        if (l0 >= 1) {
            const double r = rhs_bz<COORD>(tb, j, k, cr.eyL, cr.eyjL, cr.exL, cr.exkL);
            if constexpr (COMB && STAGE == 2) cr.bzL = rk_w2(ui(7, gm), dt, r);
            else cr.bzL = rk<STAGE>(u0(7, gm), ui(7, gm), dt, r);
        }
#pragma unroll
        for (int i = 0; i < 5; i++) cr.fzL[i] = fzz(i, gm);
        for (int l = l0; l <= l1; l++, g += sl) {
            // corner EMFs around the cell's faces
            CellE E;
            E.ex0 = e(0, g); E.ey0 = e(1, g); E.ez0 = e(2, g);
            E.exk = e(0, g - sk); E.ezk = e(2, g - sk);
            E.eyj = e(1, g - sj); E.ezj = e(2, g - sj);
            E.ezjk = e(2, g - sj - sk);
            double fz[5];
#pragma unroll
            for (int i = 0; i < 5; i++) fz[i] = fzz(i, g);
            tail_cell<R, STAGE, COORD, COMB>(L, F, tb, M, dt, Uin, Uout, j, k, l, g, cg, E, cr, fz, dtmin);
        }
    }
    if (STAGE == 3) {
        const double m = block_min(dtmin);
        if (threadIdx.x == 0) atomicMin(dtnext, (unsigned long long)__double_as_longlong(m));
    }
}
// The stage tail of the mixed form's stages 2 and 3: the arithmetic of tail_tile with 64-bit
// per-lane addresses, which measured faster there.  As in tail_tile's mixed instantiation the
// new primitive state goes to the FP32 planes P32 / EPS32 / BC32 that the next sweeps read (48
// bytes per cell instead of 96); the FP64 P / EPS / BC are rebuilt from the conserved state only
// when the host needs them.
template <typename R, int STAGE, int COORD>
__device__ inline void tail_tile_wide(const Lay &L, const Fields<R> &F, const Tabs &tb, const Model &M, const double *dtp,
                                 const double *__restrict__ Uin, double *__restrict__ Uout,
                                 unsigned long long *dtnext, int bx, int by, int zc, int tw)
{
    int j, k;
    col_of(L, bx, tw, j, k);
    const int l0 = L.zc1 ? 1 : by * zc;
    const bool col = k <= L.ny && j <= L.nx && l0 <= L.nz;
    double dtmin = 1.0e10;
    if (col) {
        const int l1 = L.zc1 ? 1 : ((l0 + zc - 1 < L.nz) ? l0 + zc - 1 : L.nz);
        const long sj = 1, sk = L.px, sl = L.py, ps = L.ps;
        const double dt = *dtp;
        const R *Fx = F.F[0], *Fy = F.F[1], *Fz = F.F[2];
        const R *E = F.E;
        const double *U0 = F.U0;
        const double gm1 = M.ggas - 1.0;
        // flux-divergence geometry of the column (coords.f90:255-282)
        const double dxl = tb.at(0, TB_D, j), dyl = tb.at(1, TB_D, k);
        double gxp = 1.0, gxm = 1.0, gxc = dxl, gyp = 1.0, gym = 1.0, gyc = dyl, gac = 0.0, xloc = 0.0, tanc = 1.0;
        if (COORD == 2) {
            xloc = tb.at(0, TB_C, j);
            gxp = tb.at(0, TB_F, j);
            gxm = tb.at(0, TB_F, j - 1);
            gxc = xloc * dxl;
            gyc = xloc * dyl;
        }
        if (COORD == 3) {
            xloc = tb.at(0, TB_C, j);
            tanc = tb.at(1, TB_TANC, k);
            gxp = tb.at(0, TB_F2, j);
            gxm = tb.at(0, TB_F2, j - 1);
            gxc = tb.at(0, TB_V3, j);
            gyp = tb.at(1, TB_SIN, k);
            gym = tb.at(1, TB_SIN, k - 1);
            gac = tb.at(0, TB_RA, j) * tb.at(1, TB_CD, k);
            gyc = gac;
        }
        // carries from plane l0-1
        long g = L.id(j, k, l0);
        long gm = g - sl;
        double exL = E[gm], eyL = E[gm + ps], eyjL = E[gm - sj + ps], exkL = E[gm - sk];
        double bzL = 0.0;
        // the stage's RK combination of component c at element e with right-hand side r
        auto rkc = [&](int c, long e, double r) {
            return rk<STAGE>(U0[e + c * ps], STAGE == 1 ? 0.0 : Uin[e + c * ps], dt, r);
        };
        if (l0 >= 1) bzL = rkc(7, gm, rhs_bz<COORD>(tb, j, k, eyL, eyjL, exL, exkL));
        double fzL[5];
#pragma unroll
        for (int i = 0; i < 5; i++) fzL[i] = (double)Fz[gm + i * ps];
        for (int l = l0; l <= l1; l++, g += sl) {
            const bool interior = j >= 1 && k >= 1 && l >= 1;
            // corner EMFs around the cell's faces
            const double ex0 = E[g], ey0 = E[g + ps], ez0 = E[g + 2 * ps];
            const double exk = E[g - sk], ezk = E[g - sk + 2 * ps];
            const double eyj = E[g - sj + ps], ezj = E[g - sj + 2 * ps];
            const double ezjk = E[g - sj - sk + 2 * ps];
            // new face fields: own three, and the x / y faces of the lower neighbours
            const double bxn = rkc(5, g, rhs_bx<COORD>(tb, j, k, l, ez0, ezk, ey0, eyL));
            const double byn = rkc(6, g, rhs_by<COORD>(tb, j, k, l, ex0, exL, ez0, ezj));
            const double bzn = rkc(7, g, rhs_bz<COORD>(tb, j, k, ey0, eyj, ex0, exk));
            // hydro right-hand side on interior cells
            double lrk[5];
            double fz[5];
#pragma unroll
            for (int i = 0; i < 5; i++) fz[i] = (double)Fz[g + i * ps];
            if (interior) {
                double sc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
                if (COORD == 2 || COORD == 3 || M.model != 0) {
                    // the stage-input density exactly (U holds the same FP64 value as the primitive
                    // rho), so the density-floor switch of the model gravity sees what the FP64
                    // form sees; velocities and pressure from the FP32 planes
                    const double rho = STAGE == 1 ? U0[g] : Uin[g];
                    const double vx = F.P32[g + ps], vy = F.P32[g + 2 * ps], vz = F.P32[g + 3 * ps];
                    const double p = F.P32[g + 4 * ps];
                    const double bx = F.BC32[g], by = F.BC32[g + ps], bz = F.BC32[g + 2 * ps];
                    if (COORD == 2) {
                        const double bsq = bx * bx + by * by + bz * bz;
                        const double tx = (p + rho * vy * vy + 0.5 * bsq - by * by) / (xloc);
                        const double ty = (rho * vx * vy - bx * by) / (xloc);
                        for (int r = 0; r < M.geom_repeat; r++) {
                            sc[IVX] = sc[IVX] + tx;
                            sc[IVY] = sc[IVY] - ty;
                        }
                    }
                    if (COORD == 3) {
                        // coords.f90:139-156
                        const double bsq = bx * bx + by * by + bz * bz;
                        const double tx = (2.0 * p + bx * bx + rho * (vy * vy + vz * vz)) / (xloc);
                        const double ty1 = (p + rho * vz * vz + 0.5 * bsq - bz * bz) / (xloc * tanc);
                        const double ty2 = (rho * vx * vy - bx * by) / (xloc);
                        const double tz1 = (rho * vx * vz - bx * bz) / (xloc);
                        const double tz2 = (rho * vy * vz - by * bz) / (xloc * tanc);
                        for (int r = 0; r < M.geom_repeat; r++) {
                            sc[IVX] = sc[IVX] + tx;
                            sc[IVY] = sc[IVY] + ty1 - ty2;
                            sc[IVZ] = sc[IVZ] - tz1 - tz2;
                        }
                    }
                    if (M.model == 2) {
                        // torus_sph/src/custom.f90:121-158: radial Paczynski-Wiita gravity
                        const double xl = tb.at(0, TB_C, j);
                        const double factor = fmax2<double>(sign1<double>(rho - M.rho_floor), 0.0);
                        const double dphidr = 1.0 / ((xl - M.r_sh) * (xl - M.r_sh));
                        sc[IVX] = sc[IVX] + (-factor * rho * dphidr);
                        sc[ITAU] = sc[ITAU] + (-factor * rho * vx * dphidr);
                    }
                    if (M.model == 1) {
                        const double xl = tb.at(0, TB_C, j), zl = tb.at(2, TB_C, l);
                        const double factor = fmax2<double>(sign1<double>(rho - M.rho_floor), 0.0);
                        const double radius = sqrt(xl * xl + zl * zl);
                        const double dphidr = xl / radius / ((radius - M.r_sh) * (radius - M.r_sh));
                        const double dphidz = zl / radius / ((radius - M.r_sh) * (radius - M.r_sh));
                        sc[IVX] = sc[IVX] + (-factor * rho * dphidr);
                        sc[IVZ] = sc[IVZ] + (-factor * rho * dphidz);
                        sc[ITAU] = sc[ITAU] + (-factor * rho * (vx * dphidr + vz * dphidz));
                    }
                }
                const double dzl = tb.at(2, TB_D, l);
                const double gzc = COORD == 3 ? gac * dzl / dyl : dzl;
#pragma unroll
                for (int i = 0; i < 5; i++) {
                    double a = 0.0;
                    a = a - ((double)Fx[g + i * ps] * gxp - (double)Fx[g - sj + i * ps] * gxm) / gxc;
                    a = a - ((double)Fy[g + i * ps] * gyp - (double)Fy[g - sk + i * ps] * gym) / gyc;
                    a = a - (fz[i] * 1.0 - fzL[i] * 1.0) / gzc;
                    lrk[i] = a + sc[i];
                }
            } else {
#pragma unroll
                for (int i = 0; i < 5; i++) lrk[i] = 0.0;
            }
            double u[5];
#pragma unroll
            for (int i = 0; i < 5; i++) u[i] = rkc(i, g, lrk[i]);
            // cell-centred field from the new face fields (interior) or the ghost value
            double bcx, bcy, bcz;
            if (interior) {
                const double bxm = rkc(5, g - sj, rhs_bx<COORD>(tb, j - 1, k, l, ezj, ezjk, eyj, eyjL));
                const double bym = rkc(6, g - sk, rhs_by<COORD>(tb, j, k - 1, l, exk, exkL, ezk, ezjk));
                bcx = 0.5 * (bxn + bxm);
                bcy = 0.5 * (byn + bym);
                bcz = 0.5 * (bzn + bzL);
                F.BC32[g] = (float)bcx;
                F.BC32[g + ps] = (float)bcy;
                F.BC32[g + 2 * ps] = (float)bcz;
            } else {
                bcx = F.BC32[g];
                bcy = F.BC32[g + ps];
                bcz = F.BC32[g + 2 * ps];
            }
            // U_to_P, floors, P_to_U, EOS
            const double rho0 = u[0];
            double vx = u[1] / rho0, vy = u[2] / rho0, vz = u[3] / rho0;
            const double bsq = bcx * bcx + bcy * bcy + bcz * bcz;
            double vsq = vx * vx + vy * vy + vz * vz;
            double eps = (u[4] - 0.5 * bsq) / rho0 - 0.5 * vsq;
            double rho = rho0;
            if (interior) {
                if (M.model != 0) {
                    if (rho < M.rho_floor) rho = M.rho_floor;
                    if (eps < M.eps_floor) eps = M.eps_floor;
                } else {
                    eps = fmax2<double>(eps, 0.0);
                }
            }
            vsq = vx * vx + vy * vy + vz * vz;
            const double p = rho * eps * gm1;
            Uout[g] = rho;
            Uout[g + ps] = vx * rho;
            Uout[g + 2 * ps] = vy * rho;
            Uout[g + 3 * ps] = vz * rho;
            Uout[g + 4 * ps] = rho * (eps + 0.5 * vsq) + 0.5 * bsq;
            Uout[g + 5 * ps] = bxn;
            Uout[g + 6 * ps] = byn;
            Uout[g + 7 * ps] = bzn;
            F.P32[g] = (float)rho;
            F.P32[g + ps] = (float)vx;
            F.P32[g + 2 * ps] = (float)vy;
            F.P32[g + 3 * ps] = (float)vz;
            F.P32[g + 4 * ps] = (float)p;
            F.P32[g + 5 * ps] = (float)bxn;
            F.P32[g + 6 * ps] = (float)byn;
            F.P32[g + 7 * ps] = (float)bzn;
            F.EPS32[g] = (float)eps;
            if (STAGE == 3 && interior)
                dtmin = fmin2<double>(cell_dt<COORD>(tb, M, j, k, l, rho, vx, vy, vz, p, bcx, bcy, bcz), dtmin);
            // carries
            exL = ex0;
            eyL = ey0;
            eyjL = eyj;
            exkL = exk;
            bzL = bzn;
#pragma unroll
            for (int i = 0; i < 5; i++) fzL[i] = fz[i];
        }
    }
    if (STAGE == 3) {
        const double m = block_min(dtmin);
        if (threadIdx.x == 0) atomicMin(dtnext, (unsigned long long)__double_as_longlong(m));
    }
}

template <typename R, int STAGE, int COORD, bool COMB>
__global__ __launch_bounds__(256) void k_tail(Lay L, Fields<R> F, Tabs tb, Model M, const double *dtp,
                                              const double *__restrict__ Uin, double *__restrict__ Uout,
                                              unsigned long long *dtnext, int gx, int gy, int zc, int tw)
{
    int bx, by;
    if (!xcd_tile(gx, gy, bx, by)) return;
    // mixed form, Cartesian stages 2-3 reading 34 planes: the 64-bit-address form (1.30 against
    // 1.34 ms per launch on the 256^3 box); the 26-plane stages (stage 1, stage 2 with a combined
    // input), the cylindrical and spherical tails (mixed tails of the 256x128x256 torus: 1.95
    // against 2.12 ms per step) and the FP64 form: 32-bit offsets (mixed stage 1: 1.11 against
    // 1.73 ms)
    if constexpr (sizeof(R) == 4 && STAGE != 1 && COORD == 1 && !(COMB && STAGE == 2))
        tail_tile_wide<R, STAGE, COORD>(L, F, tb, M, dtp, Uin, Uout, dtnext, bx, by, zc, tw);
    else
        tail_tile<R, STAGE, COORD, COMB>(L, F, tb, M, dtp, Uin, Uout, dtnext, bx, by, zc, tw);
}

// initial_update's FROMRVETOU + CUSTOM_CHECKRHO + EOS on [0,n]^3 (initial.f90:103-126)
static __global__ __launch_bounds__(256) void k_init(Lay L, Fields<double> F, Model M, double *Uout)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    const int k = (int)(t / L.px);
    const int j = (int)(t - (long)k * L.px);
    if (k > L.ny || j > L.nx) return;
    const long g = L.id(j, k, l);
    const long sj = 1, sk = L.px, sl = L.py, ps = L.ps;
    const bool interior = j >= 1 && k >= 1 && l >= 1;
    const double *B = F.P + 5 * ps;
    double bx, by, bz;
    if (interior) {
        bx = 0.5 * (B[g] + B[g - sj]);
        by = 0.5 * (B[g + ps] + B[g - sk + ps]);
        bz = 0.5 * (B[g + 2 * ps] + B[g - sl + 2 * ps]);
        F.BC[g] = bx;
        F.BC[g + ps] = by;
        F.BC[g + 2 * ps] = bz;
    } else {
        bx = F.BC[g];
        by = F.BC[g + ps];
        bz = F.BC[g + 2 * ps];
    }
    double rho = F.P[g];
    const double vx = F.P[g + ps], vy = F.P[g + 2 * ps], vz = F.P[g + 3 * ps];
    double eps = F.EPS[g];
    if (interior) {
        if (M.model != 0) {
            if (rho < M.rho_floor) rho = M.rho_floor;
            if (eps < M.eps_floor) eps = M.eps_floor;
        } else {
            eps = fmax2<double>(eps, 0.0);
        }
    }
    const double vsq = vx * vx + vy * vy + vz * vz;
    const double bsq = bx * bx + by * by + bz * bz;
    Uout[g] = rho;
    Uout[g + ps] = vx * rho;
    Uout[g + 2 * ps] = vy * rho;
    Uout[g + 3 * ps] = vz * rho;
    Uout[g + 4 * ps] = rho * (eps + 0.5 * vsq) + 0.5 * bsq;
    Uout[g + 5 * ps] = F.P[g + 5 * ps];
    Uout[g + 6 * ps] = F.P[g + 6 * ps];
    Uout[g + 7 * ps] = F.P[g + 7 * ps];
    F.P[g] = rho;
    F.P[g + 4 * ps] = rho * eps * (M.ggas - 1.0);
    F.EPS[g] = eps;
}

// Mixed form: the FP32 primitive planes from the FP64 ones, every cell of the padded volume.
static __global__ __launch_bounds__(256) void k_prim_to_f32(long ps, const double *__restrict__ P,
                                                            const double *__restrict__ EPS,
                                                            const double *__restrict__ BC, float *__restrict__ P32,
                                                            float *__restrict__ EPS32, float *__restrict__ BC32)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ps) return;
#pragma unroll
    for (int c = 0; c < 8; c++) P32[g + c * ps] = (float)P[g + c * ps];
#pragma unroll
    for (int c = 0; c < 3; c++) BC32[g + c * ps] = (float)BC[g + c * ps];
    EPS32[g] = (float)EPS[g];
}

// Mixed form: the FP64 primitive state rebuilt from the conserved state U (the last stage
// output) on [0,n]^3 -- face B copied on the whole range, cell-centred B, U_to_P with the
// floors and the EOS on the interior, as the stage tail forms them (the ghost fill then
// completes the rest).  Rounding can differ from the tail's own values in the last bit, since
// U already went through P_to_U.
static __global__ __launch_bounds__(256) void k_prim_from_u(Lay L, Fields<double> F, Model M, const double *U)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    const int k = (int)(t / L.px);
    const int j = (int)(t - (long)k * L.px);
    if (k > L.ny || j > L.nx) return;
    const long g = L.id(j, k, l);
    const long sj = 1, sk = L.px, sl = L.py, ps = L.ps;
    const double bxn = U[g + 5 * ps], byn = U[g + 6 * ps], bzn = U[g + 7 * ps];
    F.P[g + 5 * ps] = bxn;
    F.P[g + 6 * ps] = byn;
    F.P[g + 7 * ps] = bzn;
    if (!(j >= 1 && k >= 1 && l >= 1)) return;
    const double bcx = 0.5 * (bxn + U[g - sj + 5 * ps]);
    const double bcy = 0.5 * (byn + U[g - sk + 6 * ps]);
    const double bcz = 0.5 * (bzn + U[g - sl + 7 * ps]);
    F.BC[g] = bcx;
    F.BC[g + ps] = bcy;
    F.BC[g + 2 * ps] = bcz;
    const double rho0 = U[g];
    const double vx = U[g + ps] / rho0, vy = U[g + 2 * ps] / rho0, vz = U[g + 3 * ps] / rho0;
    const double bsq = bcx * bcx + bcy * bcy + bcz * bcz;
    const double vsq = vx * vx + vy * vy + vz * vz;
    double eps = (U[g + 4 * ps] - 0.5 * bsq) / rho0 - 0.5 * vsq;
    double rho = rho0;
    if (M.model != 0) {
        if (rho < M.rho_floor) rho = M.rho_floor;
        if (eps < M.eps_floor) eps = M.eps_floor;
    } else {
        eps = fmax2<double>(eps, 0.0);
    }
    F.P[g] = rho;
    F.P[g + ps] = vx;
    F.P[g + 2 * ps] = vy;
    F.P[g + 3 * ps] = vz;
    F.P[g + 4 * ps] = rho * eps * (M.ggas - 1.0);
    F.EPS[g] = eps;
}

// ------------------------------------------------------------------------------------------
// Ghost fill, one launch per pass in the reference's x -> y -> z order (boundary.f90:349-672 for
// prim/bcell/transverse face B, boundary.f90:29-341 for eps), followed in each pass by UPDATE's
// EOS on the ghost cells whose density and eps the pass finalises (update.f90:27-35).
// ------------------------------------------------------------------------------------------
// source index along an axis of n interior cells for ghost g (1..3) on side s (0 low, 1 high);
// periodic sources are wrapped so that n < NGHOST reproduces the sequential pass order
__device__ inline int ghost_src(int flag, int side, int g, int n)
{
    if (flag == 0) {
        int i = side == 0 ? 1 - g : n + g;
        return ((i - 1) % n + n) % n + 1;
    }
    if (flag == 1) return side == 0 ? 1 : n;
    return side == 0 ? g : n + 1 - g;
}

// BOUNDARY1D eps rule for the y-outer side (Q2 of the reference: tests boundary_flag(3))
__device__ inline bool eps_y_outer(const int *bf, int *flag_eff)
{
    if (bf[2] == 0) { *flag_eff = 0; return true; }
    if (bf[3] == 1) { *flag_eff = 1; return true; }
    if (bf[3] >= 2) { *flag_eff = 2; return true; }
    return false;
}

// number of ghost-fill threads of pass AX
// (a z-collapsed view runs the x and y passes on its interior plane only and has no z pass: its
// z ghosts are the plane itself)
template <int AX>
__host__ __device__ inline long ghost_threads(const Lay &L)
{
    const long nzb = L.zc1 ? 1 : L.nz + 1;
    if (AX == 0) return (long)(L.ny + 1) * nzb * 6;
    if (AX == 1) return (long)(L.nx + 6) * nzb * 6;
    return L.zc1 ? 0 : (long)(L.nx + 6) * (L.ny + 6) * 6;
}

// One ghost cell of pass AX on the primitive planes P (8), EPS, BC (3) of element type T: the
// FP64 state, or the FP32 copy of the mixed form (whose ghost pressure is formed in FP64 and
// rounded, as the stage tail forms the interior one).
template <int AX, typename T>
__device__ inline void ghost_cells(const Lay &L, T *__restrict__ P, T *__restrict__ EPS, T *__restrict__ BC,
                                   const Model &M, long t)
{
    // thread domain: (a, b) over the two other axes, g in 1..3, side in {0,1}
    int na, nb, a0, b0;  // ranges [a0, a0+na) x [b0, b0+nb)
    if (AX == 0) { a0 = 0; na = L.ny + 1; b0 = 0; nb = L.nz + 1; }           // k, l in [0,n]
    else if (AX == 1) { a0 = -2; na = L.nx + 6; b0 = 0; nb = L.nz + 1; }     // j full, l in [0,n]
    else { a0 = -2; na = L.nx + 6; b0 = -2; nb = L.ny + 6; }                 // j, k full
    if (AX != 2 && L.zc1) { b0 = 1; nb = 1; }                                // the interior plane
    const long per = (long)na * nb;
    if (t >= per * 6) return;
    int gs;  // 0..5 = (g-1)*2 + side
    long r;
    if (AX == 0) {
        // x pass: the six ghost cells of one row are adjacent in memory, so let them be
        // neighbouring threads
        gs = (int)(t % 6);
        r = t / 6;
    } else {
        gs = (int)(t / per);
        r = t - (long)gs * per;
    }
    const int g = gs / 2 + 1, side = gs & 1;
    const int a = (int)(r % na) + a0, b = (int)(r / na) + b0;
    const int n = AX == 0 ? L.nx : (AX == 1 ? L.ny : L.nz);
    const int flag = M.bflag[2 * AX + side];
    const int *fac = M.bfac[2 * AX + side];
    const int dsti = side == 0 ? 1 - g : n + g;
    const int srci = ghost_src(flag, side, g, n);
    const long ps = L.ps;
    long dst, src;
    if (AX == 0) { dst = L.id(dsti, a, b); src = L.id(srci, a, b); }
    else if (AX == 1) { dst = L.id(a, dsti, b); src = L.id(a, srci, b); }
    else { dst = L.id(a, b, dsti); src = L.id(a, b, srci); }
    const bool usef = flag >= 2;
    // prim(1:5), bcell(1:3), transverse face B
#pragma unroll
    for (int i = 0; i < 5; i++) P[dst + i * ps] = usef ? (T)fac[i] * P[src + i * ps] : P[src + i * ps];
#pragma unroll
    for (int c = 0; c < 3; c++)
        BC[dst + c * ps] = usef ? (T)fac[5 + c] * BC[src + c * ps] : BC[src + c * ps];
    const int tb1 = AX == 0 ? 6 : 5, tb2 = AX == 2 ? 6 : 7;
    P[dst + tb1 * ps] = usef ? (T)fac[tb1] * P[src + tb1 * ps] : P[src + tb1 * ps];
    P[dst + tb2 * ps] = usef ? (T)fac[tb2] * P[src + tb2 * ps] : P[src + tb2 * ps];
    // eps (BOUNDARY1D ranges: x pass k,l in [1,n]; y pass l in [1,n]; z pass all)
    bool final_region;
    if (AX == 0) final_region = a >= 1 && a <= L.ny && b >= 1 && b <= L.nz;
    else if (AX == 1) final_region = b >= 1 && b <= L.nz;
    else final_region = true;
    if (final_region) {
        int ef = flag;
        bool doit = true;
        if (AX == 1 && side == 1) doit = eps_y_outer(M.bflag, &ef);
        bool same_src = false;  // eps comes from the cell prim(1:5) came from
        if (doit) {
            const int es = ghost_src(ef, side, g, n);
            same_src = es == srci;
            long esrc;
            if (AX == 0) esrc = L.id(es, a, b);
            else if (AX == 1) esrc = L.id(a, es, b);
            else esrc = L.id(a, b, es);
            EPS[dst] = EPS[esrc];
        }
        // EOS.  With eps from the prim source cell the FP64 product equals the source pressure
        // (formed by the same expression), which the FP32 planes then keep as copied.
        if (sizeof(T) == 8 || !same_src) P[dst + 4 * ps] = (T)((double)P[dst] * (double)EPS[dst] * (M.ggas - 1.0));
    }
}

template <int AX>
__device__ inline void ghost_thread(const Lay &L, const Fields<double> &F, const Model &M, long t)
{
    ghost_cells<AX, double>(L, F.P, F.EPS, F.BC, M, t);
}

template <int AX>
__global__ __launch_bounds__(256) void k_ghost(Lay L, Fields<double> F, Model M)
{
    ghost_thread<AX>(L, F, M, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ghost fill of the mixed form's FP32 primitive planes
template <int AX>
__global__ __launch_bounds__(256) void k_ghost32(Lay L, float *P32, float *EPS32, float *BC32, Model M)
{
    ghost_cells<AX, float>(L, P32, EPS32, BC32, M, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// torus_cyn CUSTOM_BOUNDARY (custom.f90:47-81): no inflow through the R ghost zones ...
template <typename T>
static __global__ __launch_bounds__(256) void k_custom_boundary_x(Lay L, T *P)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nyz = (long)(L.ny + 6) * (L.nz + 6);
    if (t >= nyz * 3) return;
    const int g = (int)(t / nyz) + 1;
    const long r = t % nyz;
    const int k = (int)(r % (L.ny + 6)) - 2, l = (int)(r / (L.ny + 6)) - 2;
    T *vx = P + L.ps;
    const long a = L.id(1 - g, k, l), b = L.id(L.nx + g, k, l);
    vx[a] = fmin2<T>(vx[a], T(0));
    vx[b] = fmax2<T>(vx[b], T(0));
}

// ... and through the z ghost zones
template <typename T>
static __global__ __launch_bounds__(256) void k_custom_boundary_z(Lay L, T *P)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ps = L.ps;
    const long nxy = (long)(L.nx + 6) * (L.ny + 6);
    if (t >= nxy * 3) return;
    const int g = (int)(t / nxy) + 1;
    const long r = t % nxy;
    const int j = (int)(r % (L.nx + 6)) - 2, k = (int)(r / (L.nx + 6)) - 2;
    T *vz = P + 3 * ps;
    long a = L.id(j, k, 1 - g), b = L.id(j, k, L.nz + g);
    vz[a] = fmin2<T>(vz[a], T(0));
    vz[b] = fmax2<T>(vz[b], T(0));
}

// ------------------------------------------------------------------------------------------
// Reductions
// ------------------------------------------------------------------------------------------
__device__ inline double wave_min(double v)
{
    for (int o = 32; o > 0; o >>= 1) {
        double w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}
__device__ inline double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) {
        double w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

static __global__ void k_fill_u64(unsigned long long *p, unsigned long long v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// finddt (rungekutta.f90:205-252) on the current state; a thread marches one (j,k) column so a
// launch issues one atomic per block.  Positive doubles order like their bit patterns.
template <int COORD>
__device__ inline void finddt_tile(const Lay &L, const Fields<double> &F, const Tabs &tb, const Model &M,
                                   unsigned long long *out, int bx)
{
    const long t = (long)bx * blockDim.x + threadIdx.x;
    const int k = (int)(t / L.px) + 1;
    const int j = (int)(t - (long)(k - 1) * L.px) + 1;
    double dt = 1.0e10;
    if (k <= L.ny && j <= L.nx) {
        const long ps = L.ps;
        for (int l = 1; l <= L.nz; l++) {
            const long g = L.id(j, k, l);
            dt = fmin2<double>(cell_dt<COORD>(tb, M, j, k, l, F.P[g], F.P[g + ps], F.P[g + 2 * ps], F.P[g + 3 * ps],
                                              F.P[g + 4 * ps], F.BC[g], F.BC[g + ps], F.BC[g + 2 * ps]),
                               dt);
        }
    }
    const double m = block_min(dt);
    if (threadIdx.x == 0) atomicMin(out, (unsigned long long)__double_as_longlong(m));
}

template <int COORD>
__global__ __launch_bounds__(256) void k_finddt(Lay L, Fields<double> F, Tabs tb, Model M, unsigned long long *out)
{
    finddt_tile<COORD>(L, F, tb, M, out, blockIdx.x);
}

// find_divb (mhd.f90:15-66) with GEOM_AREA (coords.f90:173-223)
template <int COORD>
__global__ __launch_bounds__(256) void k_divb(Lay L, Fields<double> F, Tabs tb, unsigned long long *out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = (int)(t / L.px) + 1;
    const int j = (int)(t - (long)(k - 1) * L.px) + 1;
    double ms = 0.0, ma = 0.0;
    if (k <= L.ny && j <= L.nx) {
        for (int l = 1; l <= L.nz; l++) {
            const long g = L.id(j, k, l), ps = L.ps;
            const double dxl = tb.at(0, TB_D, j), dyl = tb.at(1, TB_D, k), dzl = tb.at(2, TB_D, l);
            double axp, axm, ayp, aym, azp, azm;
            if (COORD == 3) {
                const double xp = tb.at(0, TB_F, j), xm = tb.at(0, TB_F, j - 1), cd = tb.at(1, TB_CD, k);
                const double d2 = tb.at(0, TB_D2, j);
                axp = xp * xp * cd * dzl;
                axm = xm * xm * cd * dzl;
                ayp = 0.5 * d2 * tb.at(1, TB_SIN, k) * dzl;
                aym = 0.5 * d2 * tb.at(1, TB_SIN, k - 1) * dzl;
                azp = 0.5 * d2 * dyl;
                azm = azp;
            } else if (COORD == 2) {
                const double xp = tb.at(0, TB_F, j), xm = tb.at(0, TB_F, j - 1);
                axp = xp * dyl * dzl;
                axm = xm * dyl * dzl;
                ayp = dxl * dzl;
                aym = ayp;
                azp = 0.5 * (xp * xp - xm * xm) * dyl;
                azm = azp;
            } else {
                axp = dyl * dzl; axm = axp;
                ayp = dxl * dzl; aym = ayp;
                azp = dxl * dyl; azm = azp;
            }
            const double *B = F.P + 5 * ps;
            const double d = (axp * B[g] - axm * B[g - 1]) + (ayp * B[g + ps] - aym * B[g - L.px + ps]) +
                             (azp * B[g + 2 * ps] - azm * B[g - L.py + 2 * ps]);
            ms = fmax2<double>(ms, d);
            ma = fmax2<double>(ma, fabs(d));
        }
    }
    __shared__ double red[8];
    ms = wave_max(ms);
    ma = wave_max(ma);
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = ms;
        red[4 + (threadIdx.x >> 6)] = ma;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = red[0], b = red[4];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            a = red[w] > a ? red[w] : a;
            b = red[4 + w] > b ? red[4 + w] : b;
        }
        atomicMax(out, (unsigned long long)__double_as_longlong(a));
        atomicMax(out + 1, (unsigned long long)__double_as_longlong(b));
    }
}

// ------------------------------------------------------------------------------------------
// z-collapsed view of a periodic nz = 1 subdomain: entry check and return to the full layout
// ------------------------------------------------------------------------------------------
// Counts the cells of the whole (j, k) plane where face B_z of plane 0 differs from plane 1, in
// P or in U0 (bitwise).  In the reference both faces are the same periodic face, evolved
// separately from equal right-hand sides; the collapsed view keeps one of them.
static __global__ void k_zcheck(Lay L, const double *__restrict__ P, const double *__restrict__ U0,
                                unsigned long long *cnt)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long sx = L.nx + 6;
    if (t >= sx * (L.ny + 6)) return;
    const int j = (int)(t % sx) - 2, k = (int)(t / sx) - 2;
    const long g0 = L.id(j, k, 0) + 7 * L.ps, g1 = L.id(j, k, 1) + 7 * L.ps;
    const bool same = __double_as_longlong(P[g0]) == __double_as_longlong(P[g1]) &&
                      __double_as_longlong(U0[g0]) == __double_as_longlong(U0[g1]);
    if (!same) atomicAdd(cnt, 1ull);
}

// Rewrites the z ghost planes from plane l = 1 as the reference's z ghost pass leaves them
// (boundary.f90:564-652 with periodic z and nz = 1: every ghost plane is a copy of plane 1 in
// rho, v, p, the transverse face fields B_x, B_y, cell-centred B and eps; face B_z is not copied
// by the pass), plus face B_z and the conserved state of plane 0, the periodic face the
// collapsed view evolved as plane 1's own.
static __global__ void k_zexpand(Lay L, Fields<double> F)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long sx = L.nx + 6;
    if (t >= sx * (L.ny + 6)) return;
    const int j = (int)(t % sx) - 2, k = (int)(t / sx) - 2;
    const long ps = L.ps, src = L.id(j, k, 1);
    const int ls[6] = {-2, -1, 0, 2, 3, 4};
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const long dst = L.id(j, k, ls[q]);
#pragma unroll
        for (int i = 0; i < 7; i++) F.P[dst + i * ps] = F.P[src + i * ps];
#pragma unroll
        for (int c = 0; c < 3; c++) F.BC[dst + c * ps] = F.BC[src + c * ps];
        F.EPS[dst] = F.EPS[src];
        if (ls[q] == 0) {
            F.P[dst + 7 * ps] = F.P[src + 7 * ps];
#pragma unroll
            for (int i = 0; i < 8; i++) F.U0[dst + i * ps] = F.U0[src + i * ps];
        }
    }
}

// ------------------------------------------------------------------------------------------
// Layout conversion between the Fortran AoS arrays and the device planes
// ------------------------------------------------------------------------------------------
// aos[(cell)*nv + v] with cell = ((l+2)*(ny+6) + (k+2))*(nx+6) + (j+2)
static __global__ void k_aos_to_soa(Lay L, const double *aos, int nv, double *soa)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long sx = L.nx + 6, sy = L.ny + 6, sz = L.nz + 6;
    if (t >= sx * sy * sz) return;
    const long jj = t % sx, kk = (t / sx) % sy, ll = t / (sx * sy);
    const long g = ll * L.py + kk * L.px + jj;
    for (int v = 0; v < nv; v++) soa[g + v * L.ps] = aos[t * nv + v];
}

static __global__ void k_soa_to_aos(Lay L, const double *soa, int nv, double *aos)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long sx = L.nx + 6, sy = L.ny + 6, sz = L.nz + 6;
    if (t >= sx * sy * sz) return;
    const long jj = t % sx, kk = (t / sx) % sy, ll = t / (sx * sy);
    const long g = ll * L.py + kk * L.px + jj;
    for (int v = 0; v < nv; v++) aos[t * nv + v] = soa[g + v * L.ps];
}

template <typename R>
__global__ void k_flux_to_aos(Lay L, const R *soa, double *aos)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long sx = L.nx + 6, sy = L.ny + 6, sz = L.nz + 6;
    if (t >= sx * sy * sz) return;
    const long jj = t % sx, kk = (t / sx) % sy, ll = t / (sx * sy);
    const long g = ll * L.py + kk * L.px + jj;
    for (int v = 0; v < 7; v++) aos[t * 7 + v] = (double)soa[g + v * L.ps];
}

}  // namespace cumc
