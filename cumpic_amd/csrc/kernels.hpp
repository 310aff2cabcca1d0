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
constexpr int NTAB = 15;  // per-axis table rows
enum { TB_A0M = 0, TB_A1M, TB_A2M, TB_A3M, TB_A0P, TB_A1P, TB_A2P, TB_A3P, TB_B0C, TB_B1C, TB_CF, TB_CB,
       TB_F, TB_D, TB_C };

struct Lay {
    int nx, ny, nz;
    long px, py, ps;
    __host__ __device__ long id(int j, int k, int l) const { return (long)(l + 2) * py + (long)(k + 2) * px + (j + 2); }
};

// Per-axis geometry tables: tab[row * (n + 7) + j + 3] for j in [-3, n+3].
struct Tabs {
    const double *t[3];
    int len[3];  // n + 7 per axis
    __device__ double at(int ax, int row, int j) const { return t[ax][row * len[ax] + j + 3]; }
};

struct Model {
    int coordinate, recon, model, geom_repeat;
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
};

__device__ inline double ld(const double *p) { return __builtin_nontemporal_load(p); }

// ------------------------------------------------------------------------------------------
// x sweep: lanes walk x inside one wavefront (consecutive cells), neighbours via shuffles.
// A wave loads 64 consecutive cells of the flattened (j,k) plane and emits the 59 faces whose
// 6-cell stencil lies inside the wave.  GET_FLUXES x-direction (solver/spatial.f90:396-433).
// ------------------------------------------------------------------------------------------
constexpr int XS_FACES = 59;

template <typename R, int RECON>
__device__ inline void recon_lane(const Tabs &tb, int ax, int jc, R c, R &vm, R &vp)
{
    R cm1 = __shfl_up(c, 1, 64);
    R cp1 = __shfl_down(c, 1, 64);
    if (RECON == 4) {
        R dm = ppm_slope<R>(R(tb.at(ax, TB_B0C, jc)), R(tb.at(ax, TB_B1C, jc)), cm1, c, cp1);
        R dmm1 = __shfl_up(dm, 1, 64);
        R dmp1 = __shfl_down(dm, 1, 64);
        ppm_edges<R>(R(tb.at(ax, TB_A0M, jc)), R(tb.at(ax, TB_A1M, jc)), R(tb.at(ax, TB_A2M, jc)),
                     R(tb.at(ax, TB_A3M, jc)), R(tb.at(ax, TB_A0P, jc)), R(tb.at(ax, TB_A1P, jc)),
                     R(tb.at(ax, TB_A2P, jc)), R(tb.at(ax, TB_A3P, jc)), cm1, c, cp1, dmm1, dm, dmp1, vm, vp);
    } else {
        tvd_edges<R>(R(tb.at(ax, TB_CF, jc)), R(tb.at(ax, TB_CB, jc)), cm1, c, cp1, vm, vp);
    }
}

template <typename R, int RECON>
__global__ __launch_bounds__(256) void k_flux_x(Lay L, Fields<R> F, Tabs tb, Model M)
{
    const int lane = threadIdx.x & 63;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long f = wave * XS_FACES + lane - 2;  // flattened (j+2) + k*px over rows k in [0, ny+1]
    const int l = blockIdx.y;                   // 0 .. nz+1
    const long nload = (long)(L.ny + 4) * L.px;
    const bool inr = (f >= 0) && (f < nload);
    const long g = (long)(l + 2) * L.py + 2 * L.px + (inr ? f : 0);
    const int k = inr ? (int)(f / L.px) : 0;
    const int jj = inr ? (int)(f - (long)k * L.px) : 0;
    const int j = jj - 2;
    const bool active = lane >= 2 && lane <= 60 && inr && k <= L.ny + 1 && j >= 0 && j <= L.nx;
    const int jc = j < -2 ? -2 : (j > L.nx + 3 ? L.nx + 3 : j);

    R q[7];
    q[0] = R(F.P[g]);
    q[1] = R(F.P[g + L.ps]);
    q[2] = R(F.P[g + 2 * L.ps]);
    q[3] = R(F.P[g + 3 * L.ps]);
    q[4] = R(F.P[g + 4 * L.ps]);
    q[5] = R(F.BC[g + L.ps]);       // By cell
    q[6] = R(F.BC[g + 2 * L.ps]);   // Bz cell
    const R bn = R(F.P[g + 5 * L.ps]);  // Bx on face j

    R vl[7], vr[7];
#pragma unroll
    for (int v = 0; v < 7; v++) {
        R vm, vp;
        recon_lane<R, RECON>(tb, 0, jc, q[v], vm, vp);
        vl[v] = vp;
        vr[v] = __shfl_down(vm, 1, 64);
    }
    FaceState<R> sl{vl[0], vl[1], vl[2], vl[3], vl[4], bn, vl[5], vl[6]};
    FaceState<R> sr{vr[0], vr[1], vr[2], vr[3], vr[4], bn, vr[5], vr[6]};
    R fl[7];
    hll_face<1, R>(R(M.ggas), sl, sr, fl);
    if (active) {
        R *o = F.F[0];
#pragma unroll
        for (int c = 0; c < 7; c++) o[g + c * L.ps] = fl[c];
    }
}

// ------------------------------------------------------------------------------------------
// y and z sweeps: each thread walks one pencil along the sweep axis, lanes along x, keeping the
// 4-cell value window and 3 limited slopes of every reconstructed variable in registers so that
// each slope, each PPM call and each Riemann solve is computed exactly once.
// (solver/spatial.f90:474-511 and :552-589)
// ------------------------------------------------------------------------------------------
template <typename R, int DIR, int RECON>
__global__ __launch_bounds__(128) void k_flux_walk(Lay L, Fields<R> F, Tabs tb, Model M, int seglen)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nd = DIR == 2 ? L.ny : L.nz;   // cells along the sweep
    const int nt = DIR == 2 ? L.nz : L.ny;   // second transverse extent
    const int jj = (int)(t % L.px);
    const long tq = t / L.px;                // transverse index 0 .. nt+1
    const int j = jj - 2;
    if (tq > nt + 1 || j < 0 || j > L.nx + 1) return;
    const int s0 = blockIdx.y * seglen;
    if (s0 > nd) return;
    const int s1 = (s0 + seglen - 1 < nd) ? s0 + seglen - 1 : nd;  // last face of the segment
    const long stride = DIR == 2 ? L.px : L.py;
    const long base = DIR == 2 ? (long)(tq + 2) * L.py + 2 + j : (long)(tq + 2) * L.px + 2 + j;
    auto at = [&](int c) { return base + (long)(c + 2) * stride; };
    const int ax = DIR - 1;
    const int bA = DIR == 2 ? 0 : 0;        // first transverse cell-B (x order): y-sweep Bx, z-sweep Bx
    const int bB = DIR == 2 ? 2 : 1;        //                                     y-sweep Bz, z-sweep By
    const int ibn = DIR == 2 ? 6 : 7;       // face normal B plane in P

    auto load7 = [&](int c, R *o) {
        long g = at(c);
        o[0] = R(F.P[g]);
        o[1] = R(F.P[g + L.ps]);
        o[2] = R(F.P[g + 2 * L.ps]);
        o[3] = R(F.P[g + 3 * L.ps]);
        o[4] = R(F.P[g + 4 * L.ps]);
        o[5] = R(F.BC[g + bA * L.ps]);
        o[6] = R(F.BC[g + bB * L.ps]);
    };

    // window: w0 = v(c-1), w1 = v(c), w2 = v(c+1), w3 = v(c+2) ; slopes d0 = dm(c-1), d1 = dm(c), d2 = dm(c+1)
    R w0[7], w1[7], w2[7], w3[7], d0[7], d1[7], vpp[7];
    {
        R a[7];
        load7(s0 - 2, a);
        load7(s0 - 1, w0);
        load7(s0, w1);
        load7(s0 + 1, w2);
        if (RECON == 4) {
            R b0 = R(tb.at(ax, TB_B0C, s0 - 1)), b1 = R(tb.at(ax, TB_B1C, s0 - 1));
            R c0 = R(tb.at(ax, TB_B0C, s0)), c1 = R(tb.at(ax, TB_B1C, s0));
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
        if (RECON == 4) {
            R b0 = R(tb.at(ax, TB_B0C, c + 1)), b1 = R(tb.at(ax, TB_B1C, c + 1));
            R a0m = R(tb.at(ax, TB_A0M, c)), a1m = R(tb.at(ax, TB_A1M, c)), a2m = R(tb.at(ax, TB_A2M, c));
            R a3m = R(tb.at(ax, TB_A3M, c)), a0p = R(tb.at(ax, TB_A0P, c)), a1p = R(tb.at(ax, TB_A1P, c));
            R a2p = R(tb.at(ax, TB_A2P, c)), a3p = R(tb.at(ax, TB_A3P, c));
#pragma unroll
            for (int v = 0; v < 7; v++) {
                R d2 = ppm_slope<R>(b0, b1, w1[v], w2[v], w3[v]);
                ppm_edges<R>(a0m, a1m, a2m, a3m, a0p, a1p, a2p, a3p, w0[v], w1[v], w2[v], d0[v], d1[v], d2,
                             vm[v], vp[v]);
                d0[v] = d1[v];
                d1[v] = d2;
            }
        } else {
            R cf = R(tb.at(ax, TB_CF, c)), cb = R(tb.at(ax, TB_CB, c));
#pragma unroll
            for (int v = 0; v < 7; v++) tvd_edges<R>(cf, cb, w0[v], w1[v], w2[v], vm[v], vp[v]);
        }
        if (c > s0) {
            // face c-1: left state = vp(c-1), right state = vm(c)
            const long gf = at(c - 1);
            const R bn = R(F.P[gf + ibn * L.ps]);
            R fl[7];
            if (DIR == 2) {
                // rotated frame n=y, t1=z, t2=x ; reconstructed cell B order (x, z)
                FaceState<R> sl{vpp[0], vpp[2], vpp[3], vpp[1], vpp[4], bn, vpp[6], vpp[5]};
                FaceState<R> sr{vm[0], vm[2], vm[3], vm[1], vm[4], bn, vm[6], vm[5]};
                hll_face<2, R>(R(M.ggas), sl, sr, fl);
                R *o = F.F[1];
                o[gf] = fl[0];
                o[gf + 1 * L.ps] = fl[3];
                o[gf + 2 * L.ps] = fl[1];
                o[gf + 3 * L.ps] = fl[2];
                o[gf + 4 * L.ps] = fl[4];
                o[gf + 5 * L.ps] = fl[6];  // F(Bx)
                o[gf + 6 * L.ps] = fl[5];  // F(Bz)
            } else {
                // rotated frame n=z, t1=x, t2=y ; reconstructed cell B order (x, y)
                FaceState<R> sl{vpp[0], vpp[3], vpp[1], vpp[2], vpp[4], bn, vpp[5], vpp[6]};
                FaceState<R> sr{vm[0], vm[3], vm[1], vm[2], vm[4], bn, vm[5], vm[6]};
                hll_face<3, R>(R(M.ggas), sl, sr, fl);
                R *o = F.F[2];
                o[gf] = fl[0];
                o[gf + 1 * L.ps] = fl[2];
                o[gf + 2 * L.ps] = fl[3];
                o[gf + 3 * L.ps] = fl[1];
                o[gf + 4 * L.ps] = fl[4];
                o[gf + 5 * L.ps] = fl[5];  // F(Bx)
                o[gf + 6 * L.ps] = fl[6];  // F(By)
            }
        }
#pragma unroll
        for (int v = 0; v < 7; v++) {
            vpp[v] = vp[v];
            w0[v] = w1[v];
            w1[v] = w2[v];
            w2[v] = w3[v];
        }
    }
}

// ------------------------------------------------------------------------------------------
// Corner EMFs on [0,n]^3: face-flux averages accumulated in sweep order (spatial.f90:455-467,
// :533-546, :611-624) then the upwind-CT combination with cell EMFs (mhd.f90:95-125).
// ------------------------------------------------------------------------------------------
template <typename R>
__device__ inline R ecell_x(const Lay &L, const Fields<R> &F, long g)
{
    return R(F.P[g + 3 * L.ps]) * R(F.BC[g + L.ps]) - R(F.P[g + 2 * L.ps]) * R(F.BC[g + 2 * L.ps]);
}
template <typename R>
__device__ inline R ecell_y(const Lay &L, const Fields<R> &F, long g)
{
    return R(F.P[g + 1 * L.ps]) * R(F.BC[g + 2 * L.ps]) - R(F.P[g + 3 * L.ps]) * R(F.BC[g]);
}
template <typename R>
__device__ inline R ecell_z(const Lay &L, const Fields<R> &F, long g)
{
    return R(F.P[g + 2 * L.ps]) * R(F.BC[g]) - R(F.P[g + 1 * L.ps]) * R(F.BC[g + L.ps]);
}

template <typename R>
__global__ __launch_bounds__(256) void k_ecorn(Lay L, Fields<R> F)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    const int k = (int)(t / L.px);
    const int j = (int)(t - (long)k * L.px);
    if (k > L.ny || j > L.nx) return;
    const long g = L.id(j, k, l);
    const long sj = 1, sk = L.px, sl = L.py, ps = L.ps;
    const R *Fx = F.F[0], *Fy = F.F[1], *Fz = F.F[2];
    const R q = R(0.25);
    // x-sweep contributions (Fx plane 5 = F(By), 6 = F(Bz))
    R ey = R(0) + q * (Fx[g + 6 * ps] + Fx[g + sl + 6 * ps]);
    R ez = R(0) + q * (-Fx[g + 5 * ps] - Fx[g + sk + 5 * ps]);
    // y-sweep (Fy plane 5 = F(Bx), 6 = F(Bz))
    R ex = R(0) + q * (-Fy[g + 6 * ps] - Fy[g + sl + 6 * ps]);
    ez = ez + q * (Fy[g + 5 * ps] + Fy[g + sj + 5 * ps]);
    // z-sweep (Fz plane 5 = F(Bx), 6 = F(By))
    ex = ex + q * (Fz[g + 6 * ps] + Fz[g + sk + 6 * ps]);
    ey = ey + q * (-Fz[g + 5 * ps] - Fz[g + sj + 5 * ps]);
    // flux_ct corner EMF
    ex = ex * R(2.0) - q * (ecell_x(L, F, g) + ecell_x(L, F, g + sl) + ecell_x(L, F, g + sk) + ecell_x(L, F, g + sk + sl));
    ey = ey * R(2.0) - q * (ecell_y(L, F, g) + ecell_y(L, F, g + sl) + ecell_y(L, F, g + sj) + ecell_y(L, F, g + sj + sl));
    ez = ez * R(2.0) - q * (ecell_z(L, F, g) + ecell_z(L, F, g + sk) + ecell_z(L, F, g + sj) + ecell_z(L, F, g + sj + sk));
    F.E[g] = ex;
    F.E[g + ps] = ey;
    F.E[g + 2 * ps] = ez;
}

// ------------------------------------------------------------------------------------------
// RK stage update on [0,n]^3: flux divergence + sources (spatial.f90:436-452, :365-375,
// coords.f90:105-137, torus_cyn custom.f90:121-166), face-B right-hand side
// (mhd.f90:130-153) and the SSP-RK3 combination (rungekutta.f90:41-51, :75-85, :109-119).
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

template <typename R, int STAGE, int COORD>
__global__ __launch_bounds__(256) void k_update(Lay L, Fields<R> F, Tabs tb, Model M, const double *dtp)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    const int k = (int)(t / L.px);
    const int j = (int)(t - (long)k * L.px);
    if (k > L.ny || j > L.nx) return;
    const long g = L.id(j, k, l);
    const long sj = 1, sk = L.px, sl = L.py, ps = L.ps;
    const double dt = *dtp;
    const bool interior = j >= 1 && k >= 1 && l >= 1;

    double lrk[8];
    if (interior) {
        const R *Fx = F.F[0], *Fy = F.F[1], *Fz = F.F[2];
        double sc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        const double dxl = tb.at(0, TB_D, j), dyl = tb.at(1, TB_D, k), dzl = tb.at(2, TB_D, l);
        if (COORD == 2) {
            const double xloc = tb.at(0, TB_C, j);
            const double rho = F.P[g], vx = F.P[g + ps], vy = F.P[g + 2 * ps], p = F.P[g + 4 * ps];
            const double bx = F.BC[g], by = F.BC[g + ps], bz = F.BC[g + 2 * ps];
            const double bsq = bx * bx + by * by + bz * bz;
            const double tx = (p + rho * vy * vy + 0.5 * bsq - by * by) / (xloc);
            const double ty = (rho * vx * vy - bx * by) / (xloc);
            for (int r = 0; r < M.geom_repeat; r++) {
                sc[IVX] = sc[IVX] + tx;
                sc[IVY] = sc[IVY] - ty;
            }
        }
        if (M.model == 1) {
            const double xloc = tb.at(0, TB_C, j), zloc = tb.at(2, TB_C, l);
            const double rho = F.P[g], vx = F.P[g + ps], vz = F.P[g + 3 * ps];
            const double diff = rho - M.rho_floor;
            const double factor = fmax2<double>(sign1<double>(diff), 0.0);
            const double radius = sqrt(xloc * xloc + zloc * zloc);
            const double dphidr = xloc / radius / ((radius - M.r_sh) * (radius - M.r_sh));
            const double dphidz = zloc / radius / ((radius - M.r_sh) * (radius - M.r_sh));
            sc[IVX] = sc[IVX] + (-factor * rho * dphidr);
            sc[IVZ] = sc[IVZ] + (-factor * rho * dphidz);
            sc[ITAU] = sc[ITAU] + (-factor * rho * (vx * dphidr + vz * dphidz));
        }
        double gxp = 1.0, gxm = 1.0, gxc = dxl, gyc = dyl;
        if (COORD == 2) {
            const double xloc = tb.at(0, TB_C, j);
            gxp = tb.at(0, TB_F, j);
            gxm = tb.at(0, TB_F, j - 1);
            gxc = xloc * dxl;
            gyc = xloc * dyl;
        }
#pragma unroll
        for (int i = 0; i < 5; i++) {
            double a = 0.0;
            if (COORD == 2)
                a = a - ((double)Fx[g + i * ps] * gxp - (double)Fx[g - sj + i * ps] * gxm) / gxc;
            else
                a = a - ((double)Fx[g + i * ps] * 1.0 - (double)Fx[g - sj + i * ps] * 1.0) / gxc;
            a = a - ((double)Fy[g + i * ps] * 1.0 - (double)Fy[g - sk + i * ps] * 1.0) / gyc;
            a = a - ((double)Fz[g + i * ps] * 1.0 - (double)Fz[g - sl + i * ps] * 1.0) / dzl;
            lrk[i] = a + sc[i];
        }
    } else {
#pragma unroll
        for (int i = 0; i < 5; i++) lrk[i] = 0.0;
    }
    {
        // GEOM_CT factors at cell (j,k,l) (coords.f90:338-386)
        const double dxl = tb.at(0, TB_D, j), dyl = tb.at(1, TB_D, k), dzl = tb.at(2, TB_D, l);
        double bx_ez_c = dyl, bx_ey_c = dzl, by_ex_c = dzl, by_ez_c = dxl, bz_ey_c = dxl, bz_ex_c = dyl;
        double bz_ey_m = 1.0, bz_ey_p = 1.0;
        if (COORD == 2) {
            const double xloc = tb.at(0, TB_C, j);
            bx_ez_c = dyl * tb.at(0, TB_F, j);
            bz_ey_m = tb.at(0, TB_F, j - 1);
            bz_ey_c = xloc * dxl;
            bz_ey_p = tb.at(0, TB_F, j);
            bz_ex_c = xloc * dyl;
        }
        const R *E = F.E;
        const double ex = E[g], ey = E[g + ps], ez = E[g + 2 * ps];
        const double ez_km = E[g - sk + 2 * ps], ey_lm = E[g - sl + ps];
        const double ex_lm = E[g - sl], ez_jm = E[g - sj + 2 * ps];
        const double ey_jm = E[g - sj + ps], ex_km = E[g - sk];
        lrk[IBX] = -((ez * 1.0 - ez_km * 1.0) / bx_ez_c - (ey * 1.0 - ey_lm * 1.0) / bx_ey_c);
        lrk[IBY] = -((ex * 1.0 - ex_lm * 1.0) / by_ex_c - (ez * 1.0 - ez_jm * 1.0) / by_ez_c);
        lrk[IBZ] = -((ey * bz_ey_p - ey_jm * bz_ey_m) / bz_ey_c - (ex * 1.0 - ex_km * 1.0) / bz_ex_c);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const double u0 = F.U0[g + i * ps];
        const double u1 = STAGE == 1 ? 0.0 : F.U1[g + i * ps];
        F.U1[g + i * ps] = rk<STAGE>(u0, u1, dt, lrk[i]);
    }
}

// ------------------------------------------------------------------------------------------
// Stage tail on [0,n]^3: FROMUTORVE (fromrvetou.f90:75-127, utop.f90:185-220), CUSTOM_CHECKRHO,
// FROMRVETOU (fromrvetou.f90:21-70, ptou.f90:33-66) and UPDATE's EOS (update.f90:12-45).
// With INIT the conserved state is built from the uploaded primitives instead
// (initial.f90:103-126).
// ------------------------------------------------------------------------------------------
template <int INIT>
__global__ __launch_bounds__(256) void k_c2p(Lay L, Fields<double> F, Model M, double *Uout)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    const int k = (int)(t / L.px);
    const int j = (int)(t - (long)k * L.px);
    if (k > L.ny || j > L.nx) return;
    const long g = L.id(j, k, l);
    const long sj = 1, sk = L.px, sl = L.py, ps = L.ps;
    const bool interior = j >= 1 && k >= 1 && l >= 1;
    const double *Bsrc = INIT ? F.P + 5 * ps : Uout + 5 * ps;  // face B: prim (initial) or cons
    double bx, by, bz;
    if (interior) {
        bx = 0.5 * (Bsrc[g] + Bsrc[g - sj]);
        by = 0.5 * (Bsrc[g + ps] + Bsrc[g - sk + ps]);
        bz = 0.5 * (Bsrc[g + 2 * ps] + Bsrc[g - sl + 2 * ps]);
        F.BC[g] = bx;
        F.BC[g + ps] = by;
        F.BC[g + 2 * ps] = bz;
    } else {
        bx = F.BC[g];
        by = F.BC[g + ps];
        bz = F.BC[g + 2 * ps];
    }
    double rho, vx, vy, vz, eps;
    if (INIT) {
        rho = F.P[g];
        vx = F.P[g + ps];
        vy = F.P[g + 2 * ps];
        vz = F.P[g + 3 * ps];
        eps = F.EPS[g];
    } else {
        const double *U = Uout;
        const double ur = U[g];
        rho = ur;
        vx = U[g + ps] / ur;
        vy = U[g + 2 * ps] / ur;
        vz = U[g + 3 * ps] / ur;
        const double bsq = bx * bx + by * by + bz * bz;
        const double vsq = vx * vx + vy * vy + vz * vz;
        eps = (U[g + 4 * ps] - 0.5 * bsq) / ur - 0.5 * vsq;
        F.P[g + 5 * ps] = U[g + 5 * ps];
        F.P[g + 6 * ps] = U[g + 6 * ps];
        F.P[g + 7 * ps] = U[g + 7 * ps];
    }
    if (interior) {
        if (M.model == 1) {
            if (rho < M.rho_floor) rho = M.rho_floor;
            if (eps < M.eps_floor) eps = M.eps_floor;
        } else {
            eps = fmax2<double>(eps, 0.0);
        }
    }
    // P_to_U with the cell-centred field
    const double vsq = vx * vx + vy * vy + vz * vz;
    const double bsq = bx * bx + by * by + bz * bz;
    Uout[g] = rho;
    Uout[g + ps] = vx * rho;
    Uout[g + 2 * ps] = vy * rho;
    Uout[g + 3 * ps] = vz * rho;
    Uout[g + 4 * ps] = rho * (eps + 0.5 * vsq) + 0.5 * bsq;
    if (INIT) {
        Uout[g + 5 * ps] = F.P[g + 5 * ps];
        Uout[g + 6 * ps] = F.P[g + 6 * ps];
        Uout[g + 7 * ps] = F.P[g + 7 * ps];
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

template <int AX>
__global__ __launch_bounds__(256) void k_ghost(Lay L, Fields<double> F, Model M)
{
    // thread domain: (a, b) over the two other axes, g in 1..3, side in {0,1}
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int na, nb, a0, b0;  // ranges [a0, a0+na) x [b0, b0+nb)
    if (AX == 0) { a0 = 0; na = L.ny + 1; b0 = 0; nb = L.nz + 1; }           // k, l in [0,n]
    else if (AX == 1) { a0 = -2; na = L.nx + 6; b0 = 0; nb = L.nz + 1; }     // j full, l in [0,n]
    else { a0 = -2; na = L.nx + 6; b0 = -2; nb = L.ny + 6; }                 // j, k full
    const long per = (long)na * nb;
    if (t >= per * 6) return;
    const int gs = (int)(t / per);  // 0..5 = (g-1)*2 + side
    const long r = t - (long)gs * per;
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
    for (int i = 0; i < 5; i++) F.P[dst + i * ps] = usef ? (double)fac[i] * F.P[src + i * ps] : F.P[src + i * ps];
#pragma unroll
    for (int c = 0; c < 3; c++)
        F.BC[dst + c * ps] = usef ? (double)fac[5 + c] * F.BC[src + c * ps] : F.BC[src + c * ps];
    const int tb1 = AX == 0 ? 6 : 5, tb2 = AX == 2 ? 6 : 7;
    F.P[dst + tb1 * ps] = usef ? (double)fac[tb1] * F.P[src + tb1 * ps] : F.P[src + tb1 * ps];
    F.P[dst + tb2 * ps] = usef ? (double)fac[tb2] * F.P[src + tb2 * ps] : F.P[src + tb2 * ps];
    // eps (BOUNDARY1D ranges: x pass k,l in [1,n]; y pass l in [1,n]; z pass all)
    bool final_region;
    if (AX == 0) final_region = a >= 1 && a <= L.ny && b >= 1 && b <= L.nz;
    else if (AX == 1) final_region = b >= 1 && b <= L.nz;
    else final_region = true;
    if (final_region) {
        int ef = flag;
        bool doit = true;
        if (AX == 1 && side == 1) doit = eps_y_outer(M.bflag, &ef);
        if (doit) {
            const int es = ghost_src(ef, side, g, n);
            long esrc;
            if (AX == 0) esrc = L.id(es, a, b);
            else if (AX == 1) esrc = L.id(a, es, b);
            else esrc = L.id(a, b, es);
            F.EPS[dst] = F.EPS[esrc];
        }
        F.P[dst + 4 * ps] = F.P[dst] * F.EPS[dst] * (M.ggas - 1.0);
    }
}

// torus_cyn CUSTOM_BOUNDARY (custom.f90:47-81): no inflow through the R ghost zones ...
__global__ __launch_bounds__(256) void k_custom_boundary_x(Lay L, Fields<double> F)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nyz = (long)(L.ny + 6) * (L.nz + 6);
    if (t >= nyz * 3) return;
    const int g = (int)(t / nyz) + 1;
    const long r = t % nyz;
    const int k = (int)(r % (L.ny + 6)) - 2, l = (int)(r / (L.ny + 6)) - 2;
    double *vx = F.P + L.ps;
    const long a = L.id(1 - g, k, l), b = L.id(L.nx + g, k, l);
    vx[a] = fmin2<double>(vx[a], 0.0);
    vx[b] = fmax2<double>(vx[b], 0.0);
}

// ... and through the z ghost zones
__global__ __launch_bounds__(256) void k_custom_boundary_z(Lay L, Fields<double> F)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ps = L.ps;
    const long nxy = (long)(L.nx + 6) * (L.ny + 6);
    if (t >= nxy * 3) return;
    const int g = (int)(t / nxy) + 1;
    const long r = t % nxy;
    const int j = (int)(r % (L.nx + 6)) - 2, k = (int)(r / (L.nx + 6)) - 2;
    double *vz = F.P + 3 * ps;
    long a = L.id(j, k, 1 - g), b = L.id(j, k, L.nz + g);
    vz[a] = fmin2<double>(vz[a], 0.0);
    vz[b] = fmax2<double>(vz[b], 0.0);
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

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// finddt (rungekutta.f90:205-252); positive doubles order like their bit patterns
template <int COORD>
__global__ __launch_bounds__(256) void k_finddt(Lay L, Fields<double> F, Tabs tb, Model M, unsigned long long *out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y + 1;
    const int k = (int)(t / L.px) + 1;
    const int j = (int)(t - (long)(k - 1) * L.px) + 1;
    double dt = 1.0e10;
    if (k <= L.ny && j <= L.nx) {
        const long g = L.id(j, k, l), ps = L.ps;
        const double g_ = M.ggas;
        const double rho = F.P[g];
        const double cs = sqrt(g_ * F.P[g + 4 * ps] / rho);
        const double a2 = cs * cs;
        const double a4 = a2 * a2;
        const double bx = F.BC[g], by = F.BC[g + ps], bz = F.BC[g + 2 * ps];
        const double b2x = (bx * bx / rho), b2y = (by * by / rho), b2z = (bz * bz / rho);
        const double b2 = b2x + b2y + b2z;
        const double b4 = b2 * b2;
        const double cfx = sqrt(0.5 * (a2 + b2 + sqrt((a4 + 2.0 * a2 * b2 + b4) - 4.0 * a2 * b2x)));
        const double cfy = sqrt(0.5 * (a2 + b2 + sqrt((a4 + 2.0 * a2 * b2 + b4) - 4.0 * a2 * b2y)));
        const double cfz = sqrt(0.5 * (a2 + b2 + sqrt((a4 + 2.0 * a2 * b2 + b4) - 4.0 * a2 * b2z)));
        const double l1 = fabs(F.P[g + ps]) + cfx, l2 = fabs(F.P[g + 2 * ps]) + cfy, l3 = fabs(F.P[g + 3 * ps]) + cfz;
        const double lambda = fmax2<double>(fmax2<double>(l1, l2), l3);
        double dxl = tb.at(0, TB_D, j), dyl = tb.at(1, TB_D, k), dzl = tb.at(2, TB_D, l);
        if (COORD == 2) dyl = tb.at(0, TB_C, j) * dyl;
        const double dt_out = fmin2<double>(fmin2<double>(dxl, dyl), dzl) * M.cfl / lambda;
        dt = fmin2<double>(dt_out, dt);
    }
    __shared__ double red[4];
    dt = wave_min(dt);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dt;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = red[w] < m ? red[w] : m;
        atomicMin(out, (unsigned long long)__double_as_longlong(m));
    }
}

// find_divb (mhd.f90:15-66) with GEOM_AREA (coords.f90:173-223)
template <int COORD>
__global__ __launch_bounds__(256) void k_divb(Lay L, Fields<double> F, Tabs tb, unsigned long long *out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y + 1;
    const int k = (int)(t / L.px) + 1;
    const int j = (int)(t - (long)(k - 1) * L.px) + 1;
    double ms = 0.0, ma = 0.0;
    if (k <= L.ny && j <= L.nx) {
        const long g = L.id(j, k, l), ps = L.ps;
        const double dxl = tb.at(0, TB_D, j), dyl = tb.at(1, TB_D, k), dzl = tb.at(2, TB_D, l);
        double axp, axm, ayp, aym, azp, azm;
        if (COORD == 2) {
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
        ms = fmax2<double>(0.0, d);
        ma = fabs(d);
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
// Layout conversion between the Fortran AoS arrays and the device planes
// ------------------------------------------------------------------------------------------
// aos[(cell)*nv + v] with cell = ((l+2)*(ny+6) + (k+2))*(nx+6) + (j+2)
__global__ void k_aos_to_soa(Lay L, const double *aos, int nv, double *soa)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long sx = L.nx + 6, sy = L.ny + 6, sz = L.nz + 6;
    if (t >= sx * sy * sz) return;
    const long jj = t % sx, kk = (t / sx) % sy, ll = t / (sx * sy);
    const long g = ll * L.py + kk * L.px + jj;
    for (int v = 0; v < nv; v++) soa[g + v * L.ps] = aos[t * nv + v];
}

__global__ void k_soa_to_aos(Lay L, const double *soa, int nv, double *aos)
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
