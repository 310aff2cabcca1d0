// cumc_adjoint.hip -- reverse mode of initial_update + N RUNGEKUTTA steps (FP64, Cartesian,
// riemann model hooks), with checkpoint-every-k recomputation.
//
// The forward map is the exact FP64 step of the other kernels with the time steps frozen (the
// CFL reduction is a max over cells and is not differentiated).  Each kernel below reverses one
// forward write phase: it reads the adjoints of what the phase wrote, adds their contributions
// to the adjoints of what it read, and zeroes the adjoints of the overwritten values, so the
// adjoint arrays always hold d(objective)/d(current contents) of the forward arrays.  Every
// adjoint value has one writer per kernel (gathers, no atomics), so gradients are reproducible
// bit for bit from run to run.
//
// Reverse of one stage (forward: sweeps -> corner EMFs -> tail -> ghost fill):
//   ghost fill z, y, x (kr_ghost)      BOUNDARYP / BOUNDARY1D + UPDATE's EOS on the ghosts
//   tail (kr_tail_a, kr_tail_b)        RK combination, FROMUTORVE, CHECKRHO, FROMRVETOU, EOS
//   flux / EMF gathers (kr_gather)     flux divergence and the face-B right-hand sides
//   corner EMFs (kr_ecorn)             flux_ct (mhd.f90:73-157)
//   sweeps (kr_flux_hll, kr_flux_recon) RIEMANN (rsolver.f90) then PPM / TVD reconstruction
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "adjoint.hpp"
#include "ctx.hpp"

namespace {

// adjoint arrays of one context, FP64 planes in the state layout
struct AdjF {
    double *P;    // 8
    double *BC;   // 3
    double *EPS;  // 1
    double *F[3]; // 7 per sweep (xyz component order of the forward flux planes)
    double *E;    // 3
    double *T;    // 11 temporaries: 0-4 d/d(rhs hydro), 5-7 face-B, 8-10 half cell-B adjoints
    double *H;    // 14: adjoints of the reconstructed left (0-6) and right (7-13) face states
};
constexpr int NADJ = 8 + 3 + 1 + 21 + 3 + 11 + 14;

template <int STAGE> struct RKC;
template <> struct RKC<1> { static constexpr double a0 = 1.0, a1 = 0.0, a2 = 1.0; };
template <> struct RKC<2> { static constexpr double a0 = RK20, a1 = RK21, a2 = RK22; };
template <> struct RKC<3> { static constexpr double a0 = RK30, a1 = RK31, a2 = RK32; };

// ---------------------------------------------------------------------------------------------
// ghost fill, one pass, reversed: a thread owns one line (the six ghost cells of the pass along
// AX at fixed transverse (a, b)); sources are interior along AX, destinations ghosts, so lines
// never share a cell.  rho_out / eps_out: forward values after the stage (at every destination
// whose p adjoint is still live they are the values the pass wrote).
// ---------------------------------------------------------------------------------------------
// Six threads take one line, each its ghost (side, g) = (q / 3, q % 3 + 1): along x the six
// threads of a line are neighbours (its ghosts are adjacent in memory), along y and z thread t
// takes line t mod lines and ghost t / lines (neighbouring threads, neighbouring lines).  When
// both sides are periodic along AX and n >= 6 the six ghosts' source cells are distinct (n+1-g
// and g), so each thread reverses its own ghost and every adjoint sum keeps its single term;
// otherwise the thread of ghost 0 reverses the whole line in order and the others return.
template <int AX>
__host__ __device__ inline long kr_ghost_lines(const Lay &L)
{
    const long nzb = L.zc1 ? 1 : L.nz + 1;
    if (AX == 0) return (long)(L.ny + 1) * nzb;
    if (AX == 1) return (long)(L.nx + 6) * nzb;
    return (long)(L.nx + 6) * (L.ny + 6);
}

template <int AX>
__device__ inline void kr_ghost_line(const Lay &L, const AdjF &A, const Model &M, const double *__restrict__ rho_out,
                                     const double *__restrict__ eps_out, long t6)
{
    int na, nb, a0, b0;
    if (AX == 0) { a0 = 0; na = L.ny + 1; b0 = 0; nb = L.nz + 1; }
    else if (AX == 1) { a0 = -2; na = L.nx + 6; b0 = 0; nb = L.nz + 1; }
    else { a0 = -2; na = L.nx + 6; b0 = -2; nb = L.ny + 6; }
    if (AX != 2 && L.zc1) { b0 = 1; nb = 1; }  // z-collapsed view: the interior plane (no z pass)
    const long nl = (long)na * nb;
    long t;
    int q;
    if (AX == 0) {
        t = t6 / 6;
        q = (int)(t6 - t * 6);
    } else {
        q = (int)(t6 / nl);
        t = t6 - (long)q * nl;
    }
    if (t >= nl || q >= 6) return;
    const int nax = AX == 0 ? L.nx : (AX == 1 ? L.ny : L.nz);
    const bool split = M.bflag[2 * AX] == 0 && M.bflag[2 * AX + 1] == 0 && nax >= 6;
    if (!split && q != 0) return;
    const int a = (int)(t % na) + a0, b = (int)(t / na) + b0;
    const int n = AX == 0 ? L.nx : (AX == 1 ? L.ny : L.nz);
    const long ps = L.ps;
    const double gm1 = M.ggas - 1.0;
    auto cell = [&](int i) {
        if (AX == 0) return L.id(i, a, b);
        if (AX == 1) return L.id(a, i, b);
        return L.id(a, b, i);
    };
    bool final_region;
    if (AX == 0) final_region = a >= 1 && a <= L.ny && b >= 1 && b <= L.nz;
    else if (AX == 1) final_region = b >= 1 && b <= L.nz;
    else final_region = true;
    const int tb1 = AX == 0 ? 6 : 5, tb2 = AX == 2 ? 6 : 7;
    for (int side = 0; side < 2; side++) {
        if (split && side != q / 3) continue;
        const int flag = M.bflag[2 * AX + side];
        const int *fac = M.bfac[2 * AX + side];
        const bool usef = flag >= 2;
        for (int g = 1; g <= 3; g++) {
            if (split && g != q % 3 + 1) continue;
            const long dst = cell(side == 0 ? 1 - g : n + g);
            const long src = cell(ghost_src(flag, side, g, n));
            if (final_region) {
                // p(dst) = rho(dst) * eps(dst) * (gamma - 1)
                const double lp = A.P[dst + 4 * ps];
                A.P[dst + 4 * ps] = 0.0;
                A.P[dst] += lp * eps_out[dst] * gm1;
                A.EPS[dst] += lp * rho_out[dst] * gm1;
                int ef = flag;
                bool doit = true;
                if (AX == 1 && side == 1) doit = eps_y_outer(M.bflag, &ef);
                if (doit) {
                    const long es = cell(ghost_src(ef, side, g, n));
                    const double le = A.EPS[dst];
                    A.EPS[dst] = 0.0;
                    A.EPS[es] += le;
                }
            }
            auto back = [&](double *X, long off, int fi) {
                const double l = X[dst + off];
                X[dst + off] = 0.0;
                X[src + off] += usef ? (double)fac[fi] * l : l;
            };
            back(A.P, tb1 * ps, tb1);
            back(A.P, tb2 * ps, tb2);
            for (int c = 0; c < 3; c++) back(A.BC, c * ps, 5 + c);
            for (int i = 0; i < 5; i++) back(A.P, i * ps, i);
        }
    }
}

template <int AX>
__global__ __launch_bounds__(256) void kr_ghost(Lay L, AdjF A, Model M, const double *__restrict__ rho_out,
                                                const double *__restrict__ eps_out)
{
    kr_ghost_line<AX>(L, A, M, rho_out, eps_out, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------------------------
// stage tail reversed, part a: per cell of [0,n]^3 the chain
//   Uout, P, EPS, bcell  <-  (rho, v, eps, bcell)  <-  (u, bcell)  <-  (U0, Uin, rhs)
// is reversed.  The forward values it needs come from the stored stage output on interior cells
// (k_tail writes rho, v, eps and bcell there and the ghost fill leaves them; with the riemann
// hooks rho is not floored and eps = max(eps0, 0), so eps0 is recovered where eps > 0, and a
// +0 eps marks the floored branch) and are recomputed on the lower ghost layer (j, k or l = 0),
// where the tail's right-hand side is zero and its bcell is the stage input's.  Hydro adjoints
// go to U0 / Uin directly; the face-B parts wait for part b because a face field feeds the
// cell-centred field of two cells.
// ---------------------------------------------------------------------------------------------
template <int STAGE>
__device__ inline void kr_tail_a_cell(const Lay &L, const Fields<double> &F, const double *__restrict__ Pout,
                                      const double *__restrict__ EPSout, const double *__restrict__ BCout,
                                      const Model &M, double dt, const double *__restrict__ U0,
                                      const double *__restrict__ Uin, const AdjF &A, double *lU0, double *lUin,
                                      double *lUout, long t, int l)
{
    const int k = (int)(t / L.px), j = (int)(t - (long)k * L.px);
    if (k > L.ny || j > L.nx) return;
    const long g = L.id(j, k, l), ps = L.ps;
    const bool interior = j >= 1 && k >= 1 && l >= 1;
    const double gm1 = M.ggas - 1.0;
    double rho, vx, vy, vz, bc[3], eps, ekin;  // ekin = (u4 - B^2/2) / rho = eps0 + v^2/2
    bool floored;
    if (interior) {
        rho = Pout[g];
        vx = Pout[g + ps];
        vy = Pout[g + 2 * ps];
        vz = Pout[g + 3 * ps];
        for (int c = 0; c < 3; c++) bc[c] = BCout[g + c * ps];
        eps = EPSout[g];
        floored = eps == 0.0 && !signbit(eps);
        ekin = eps + 0.5 * (vx * vx + vy * vy + vz * vz);
    } else {
        auto uin = [&](long i) { return STAGE == 1 ? 0.0 : Uin[i]; };
        double u[5];
        for (int i = 0; i < 5; i++) u[i] = rk<STAGE>(U0[g + i * ps], uin(g + i * ps), dt, 0.0);
        for (int c = 0; c < 3; c++) bc[c] = F.BC[g + c * ps];
        rho = u[0];
        vx = u[1] / rho;
        vy = u[2] / rho;
        vz = u[3] / rho;
        const double bsq = bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2];
        const double vsq = vx * vx + vy * vy + vz * vz;
        ekin = (u[4] - 0.5 * bsq) / rho;
        eps = ekin - 0.5 * vsq;
        floored = false;
    }
    const double vsq = vx * vx + vy * vy + vz * vz;
    // adjoints of the outputs
    double lP[8], lUo[8], lBc[3];
    for (int v = 0; v < 8; v++) {
        lP[v] = A.P[g + v * ps];
        lUo[v] = lUout[g + v * ps];
        A.P[g + v * ps] = 0.0;
        lUout[g + v * ps] = 0.0;
    }
    const double lE = A.EPS[g];
    A.EPS[g] = 0.0;
    for (int c = 0; c < 3; c++) {
        lBc[c] = interior ? A.BC[g + c * ps] : 0.0;
        if (interior) A.BC[g + c * ps] = 0.0;
    }
    // Uout = (rho, rho v, rho (eps + v^2/2) + B^2/2), P = (rho, v, rho eps (gamma-1)), EPS = eps
    double lrho = lUo[0] + lP[0] + lUo[1] * vx + lUo[2] * vy + lUo[3] * vz + lUo[4] * (eps + 0.5 * vsq) +
                  lP[4] * eps * gm1;
    double lv[3] = {lUo[1] * rho + lP[1] + lUo[4] * rho * vx, lUo[2] * rho + lP[2] + lUo[4] * rho * vy,
                    lUo[3] * rho + lP[3] + lUo[4] * rho * vz};
    const double leps = lE + lUo[4] * rho + lP[4] * rho * gm1;
    double lbc[3];
    for (int c = 0; c < 3; c++) lbc[c] = lBc[c] + lUo[4] * bc[c];  // d(B^2/2)
    const double leps0 = floored ? 0.0 : leps;
    // eps0 = (u4 - B^2/2)/rho - v^2/2
    double lu[5];
    lu[4] = leps0 / rho;
    for (int c = 0; c < 3; c++) lbc[c] -= leps0 * bc[c] / rho;
    lrho -= leps0 * ekin / rho;
    lv[0] -= leps0 * vx;
    lv[1] -= leps0 * vy;
    lv[2] -= leps0 * vz;
    // v = m / rho
    lu[1] = lv[0] / rho;
    lu[2] = lv[1] / rho;
    lu[3] = lv[2] / rho;
    lrho -= (lv[0] * vx + lv[1] * vy + lv[2] * vz) / rho;
    lu[0] = lrho;
    // u = a0 U0 + a1 Uin + a2 dt rhs
    for (int i = 0; i < 5; i++) {
        lU0[g + i * ps] += RKC<STAGE>::a0 * lu[i];
        if (STAGE > 1) lUin[g + i * ps] += RKC<STAGE>::a1 * lu[i];
        A.T[g + i * ps] = interior ? RKC<STAGE>::a2 * dt * lu[i] : 0.0;
    }
    for (int c = 0; c < 3; c++) {
        const double half = interior ? 0.5 * lbc[c] : 0.0;
        A.T[g + (5 + c) * ps] = lP[5 + c] + lUo[5 + c] + half;
        A.T[g + (8 + c) * ps] = half;
        if (!interior) A.BC[g + c * ps] += lbc[c];  // the ghost bcell the tail read
    }
}

// forward fields of a stored state (P 8, EPS 1, bcell 3 planes at s) as the input of a stage
__host__ __device__ inline Fields<double> fields_of_state(Fields<double> F, const double *s, long ps)
{
    F.P = const_cast<double *>(s);
    F.EPS = const_cast<double *>(s) + 8 * ps;
    F.BC = const_cast<double *>(s) + 9 * ps;
    return F;
}

template <int STAGE>
__global__ __launch_bounds__(256) void kr_tail_a(Lay L, Fields<double> F, const double *out, Model M, double dt,
                                                 const double *__restrict__ U0, const double *__restrict__ Uin,
                                                 AdjF A, double *lU0, double *lUin, double *lUout)
{
    kr_tail_a_cell<STAGE>(L, F, out, out + 8 * L.ps, out + 9 * L.ps, M, dt, U0, Uin, A, lU0, lUin, lUout,
                          (long)blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

// part b: face-B adjoints (own cell's share + the upper neighbour's share of its cell field)
template <int STAGE>
__device__ inline void kr_tail_b_cell(const Lay &L, const AdjF &A, double dt, double *lU0, double *lUin, long t, int l)
{
    const int k = (int)(t / L.px), j = (int)(t - (long)k * L.px);
    if (k > L.ny || j > L.nx) return;
    const long g = L.id(j, k, l), ps = L.ps;
    double lb[3];
    lb[0] = A.T[g + 5 * ps] + ((j + 1 <= L.nx && k >= 1 && l >= 1) ? A.T[g + 1 + 8 * ps] : 0.0);
    lb[1] = A.T[g + 6 * ps] + ((k + 1 <= L.ny && j >= 1 && l >= 1) ? A.T[g + L.px + 9 * ps] : 0.0);
    // (in a z-collapsed view the cell above is the cell itself, which reads this face as its lower
    // one too)
    lb[2] = A.T[g + 7 * ps] + (((L.zc1 || l + 1 <= L.nz) && j >= 1 && k >= 1) ? A.T[g + L.py + 10 * ps] : 0.0);
    for (int c = 0; c < 3; c++) {
        lU0[g + (5 + c) * ps] += RKC<STAGE>::a0 * lb[c];
        if (STAGE > 1) lUin[g + (5 + c) * ps] += RKC<STAGE>::a1 * lb[c];
        A.T[g + (5 + c) * ps] = RKC<STAGE>::a2 * dt * lb[c];  // d/d(rhs of the face field)
    }
}

template <int STAGE>
__global__ __launch_bounds__(256) void kr_tail_b(Lay L, AdjF A, double dt, double *lU0, double *lUin)
{
    kr_tail_b_cell<STAGE>(L, A, dt, lU0, lUin, (long)blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

// flux and corner-EMF adjoints gathered from the per-cell right-hand-side adjoints, on [0,n+1]^3
// (hydro flux components; the B-flux components are set by kr_ecorn)
__device__ inline void kr_gather_cell(const Lay &L, const AdjF &A, const Tabs &tb, long t, int l)
{
    const int k = (int)(t / L.px), j = (int)(t - (long)k * L.px);
    if (k > L.ny + 1 || j > L.nx + 1) return;
    const long g = L.id(j, k, l), ps = L.ps;
    // In a z-collapsed view (l = 1 only) the z neighbour l + 1 is the cell itself: its lower-face
    // and lower-edge reads are reads of this face / edge, with this cell's widths.
    const int lu = L.zc1 ? l : l + 1;
    auto inter = [&](int jj, int kk, int ll) {
        return jj >= 1 && jj <= L.nx && kk >= 1 && kk <= L.ny && ll >= 1 && ll <= L.nz;
    };
    auto inbox = [&](int jj, int kk, int ll) { return jj <= L.nx && kk <= L.ny && ll <= L.nz; };
    // flux divergence: lrk(c) = -(F(face c) - F(face c-1)) / d(c)
    if (j <= L.nx)
        for (int i = 0; i < 5; i++)
            A.F[0][g + i * ps] = (inter(j, k, l) ? -A.T[g + i * ps] / tb.at(0, TB_D, j) : 0.0) +
                                 (inter(j + 1, k, l) ? A.T[g + 1 + i * ps] / tb.at(0, TB_D, j + 1) : 0.0);
    if (k <= L.ny)
        for (int i = 0; i < 5; i++)
            A.F[1][g + i * ps] = (inter(j, k, l) ? -A.T[g + i * ps] / tb.at(1, TB_D, k) : 0.0) +
                                 (inter(j, k + 1, l) ? A.T[g + L.px + i * ps] / tb.at(1, TB_D, k + 1) : 0.0);
    if (l <= L.nz)
        for (int i = 0; i < 5; i++)
            A.F[2][g + i * ps] = (inter(j, k, l) ? -A.T[g + i * ps] / tb.at(2, TB_D, l) : 0.0) +
                                 (inter(j, k, lu) ? A.T[g + L.py + i * ps] / tb.at(2, TB_D, lu) : 0.0);
    // corner EMFs from the face-field right-hand sides (mhd.f90:130-153, Cartesian)
    if (inbox(j, k, l)) {
        auto rb = [&](int c, int jj, int kk, int ll) {
            return inbox(jj, kk, ll) ? A.T[L.id(jj, kk, ll) + (5 + c) * ps] : 0.0;
        };
        const double dx = tb.at(0, TB_D, j), dy = tb.at(1, TB_D, k), dz = tb.at(2, TB_D, l);
        const double dx1 = tb.at(0, TB_D, j + 1), dy1 = tb.at(1, TB_D, k + 1), dz1 = tb.at(2, TB_D, lu);
        const double lez = -rb(0, j, k, l) / dy + rb(0, j, k + 1, l) / dy1 + rb(1, j, k, l) / dx - rb(1, j + 1, k, l) / dx1;
        const double ley = rb(0, j, k, l) / dz - rb(0, j, k, lu) / dz1 - rb(2, j, k, l) / dx + rb(2, j + 1, k, l) / dx1;
        const double lex = -rb(1, j, k, l) / dz + rb(1, j, k, lu) / dz1 + rb(2, j, k, l) / dy - rb(2, j, k + 1, l) / dy1;
        A.E[g] = lex;
        A.E[g + ps] = ley;
        A.E[g + 2 * ps] = lez;
    }
}

__global__ __launch_bounds__(256) void kr_gather(Lay L, AdjF A, Tabs tb)
{
    kr_gather_cell(L, A, tb, (long)blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

// flux_ct reversed (forward in ecorn_tile): on [0,n+1]^3 a thread owns the B-flux components of
// the three faces at its index and the cell EMF of its cell
__device__ inline void kr_ecorn_cell(const Lay &L, const Fields<double> &F, const AdjF &A, long t, int l)
{
    const int k = (int)(t / L.px), j = (int)(t - (long)k * L.px);
    if (k > L.ny + 1 || j > L.nx + 1) return;
    const long g = L.id(j, k, l), ps = L.ps;
    auto le = [&](int c, int jj, int kk, int ll) {
        return (jj >= 0 && kk >= 0 && ll >= 0 && jj <= L.nx && kk <= L.ny && ll <= L.nz)
                   ? A.E[L.id(jj, kk, ll) + c * ps] : 0.0;
    };
    if (j <= L.nx) {
        A.F[0][g + 6 * ps] = 0.5 * (le(1, j, k, l) + le(1, j, k, l - 1));    // F_x(Bz) -> ey
        A.F[0][g + 5 * ps] = -0.5 * (le(2, j, k, l) + le(2, j, k - 1, l));   // F_x(By) -> ez
    }
    if (k <= L.ny) {
        A.F[1][g + 6 * ps] = -0.5 * (le(0, j, k, l) + le(0, j, k, l - 1));   // F_y(Bz) -> ex
        A.F[1][g + 5 * ps] = 0.5 * (le(2, j, k, l) + le(2, j - 1, k, l));    // F_y(Bx) -> ez
    }
    if (l <= L.nz) {
        A.F[2][g + 6 * ps] = 0.5 * (le(0, j, k, l) + le(0, j, k - 1, l));    // F_z(By) -> ex
        A.F[2][g + 5 * ps] = -0.5 * (le(1, j, k, l) + le(1, j - 1, k, l));   // F_z(Bx) -> ey
    }
    // cell EMF -v x B at this cell, used by four edges per component with weight -1/4
    const double lx = -0.25 * (le(0, j, k, l) + le(0, j, k, l - 1) + le(0, j, k - 1, l) + le(0, j, k - 1, l - 1));
    const double ly = -0.25 * (le(1, j, k, l) + le(1, j, k, l - 1) + le(1, j - 1, k, l) + le(1, j - 1, k, l - 1));
    const double lz = -0.25 * (le(2, j, k, l) + le(2, j, k - 1, l) + le(2, j - 1, k, l) + le(2, j - 1, k - 1, l));
    const double vx = F.P[g + ps], vy = F.P[g + 2 * ps], vz = F.P[g + 3 * ps];
    const double bx = F.BC[g], by = F.BC[g + ps], bz = F.BC[g + 2 * ps];
    // ex = vz by - vy bz ; ey = vx bz - vz bx ; ez = vy bx - vx by
    A.P[g + ps] += ly * bz - lz * by;
    A.P[g + 2 * ps] += -lx * bz + lz * bx;
    A.P[g + 3 * ps] += lx * by - ly * bx;
    A.BC[g] += -ly * vz + lz * vy;
    A.BC[g + ps] += lx * vz - lz * vx;
    A.BC[g + 2 * ps] += -lx * vy + ly * vx;
}

__global__ __launch_bounds__(256) void kr_ecorn(Lay L, Fields<double> F, AdjF A)
{
    kr_ecorn_cell(L, F, A, (long)blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------------------------
// sweeps reversed.  Pencils along DIR cover the forward flux domain; cells along the pencil run
// from -2 to n+3, faces from 0 to n (face f between cells f and f+1).
// ---------------------------------------------------------------------------------------------
template <int DIR>
struct Pencil {
    long base, stride;
    int n;
    __device__ bool init(const Lay &L, long t)
    {
        // (a z-collapsed view has one plane of x / y pencils and one z face per column: its z
        // pencils are walked as if n = 0, cells 0 and 1 around face 0, all of them the cell itself)
        if (DIR == 1) {
            const int k = (int)(t % (L.ny + 2)), l = (int)(t / (L.ny + 2));
            if (l >= flux_planes(L)) return false;
            base = L.id(-2, k, l); stride = 1; n = L.nx;
        } else if (DIR == 2) {
            const int j = (int)(t % (L.nx + 2)), l = (int)(t / (L.nx + 2));
            if (l >= flux_planes(L)) return false;
            base = L.id(j, -2, l); stride = L.px; n = L.ny;
        } else {
            const int j = (int)(t % (L.nx + 2)), k = (int)(t / (L.nx + 2));
            if (k > L.ny + 1) return false;
            base = L.id(j, k, -2); stride = L.py; n = L.zc1 ? 0 : L.nz;
        }
        return true;
    }
    __device__ long at(int c) const { return base + (long)(c + 2) * stride; }
};

template <int DIR>
__host__ __device__ inline long npencils(const Lay &L)
{
    return DIR == 1 ? (long)(L.ny + 2) * flux_planes(L)
                    : (DIR == 2 ? (long)(L.nx + 2) * flux_planes(L) : (long)(L.nx + 2) * (L.ny + 2));
}

// reconstructed variable v (0..6) of a sweep: P planes 0..4, then the two transverse cell fields
template <int DIR>
__device__ inline const double *qplane(const Fields<double> &F, long ps, int v)
{
    if (v < 5) return F.P + v * ps;
    const int bA = DIR == 1 ? 1 : 0, bB = DIR == 3 ? 1 : 2;
    return F.BC + (v == 5 ? bA : bB) * ps;
}
template <int DIR>
__device__ inline double *qplane_adj(const AdjF &A, long ps, int v)
{
    if (v < 5) return A.P + v * ps;
    const int bA = DIR == 1 ? 1 : 0, bB = DIR == 3 ? 1 : 2;
    return A.BC + (v == 5 ? bA : bB) * ps;
}

// PPM / TVD edge values of cell c from w(c-2..c+2)
template <int RECON>
__device__ inline void recon_cell(const Tabs &tb, int ax, int c, const double *w, double &vm, double &vp)
{
    if (RECON == 4) {
        const double dmm = ppm_slope<double>(tb.at(ax, TB_B0C, c - 1), tb.at(ax, TB_B1C, c - 1), w[0], w[1], w[2]);
        const double dmc = ppm_slope<double>(tb.at(ax, TB_B0C, c), tb.at(ax, TB_B1C, c), w[1], w[2], w[3]);
        const double dmp = ppm_slope<double>(tb.at(ax, TB_B0C, c + 1), tb.at(ax, TB_B1C, c + 1), w[2], w[3], w[4]);
        ppm_edges<double>(tb.at(ax, TB_A0M, c), tb.at(ax, TB_A1M, c), tb.at(ax, TB_A2M, c), tb.at(ax, TB_A3M, c),
                          tb.at(ax, TB_A0P, c), tb.at(ax, TB_A1P, c), tb.at(ax, TB_A2P, c), tb.at(ax, TB_A3P, c), w[1],
                          w[2], w[3], dmm, dmc, dmp, vm, vp);
    } else {
        tvd_edges<double>(tb.at(ax, TB_CF, c), tb.at(ax, TB_CB, c), w[1], w[2], w[3], vm, vp);
    }
}

// face states in the rotated frame of hll_face<DIR> from the reconstructed q order
template <int DIR>
__device__ inline FaceState<double> rot(const double *q, double bn)
{
    if (DIR == 1) return FaceState<double>{q[0], q[1], q[2], q[3], q[4], bn, q[5], q[6]};
    if (DIR == 2) return FaceState<double>{q[0], q[2], q[3], q[1], q[4], bn, q[6], q[5]};
    return FaceState<double>{q[0], q[3], q[1], q[2], q[4], bn, q[5], q[6]};
}
template <int DIR>
__device__ inline void unrot(const FaceState<double> &s, double *q)
{
    if (DIR == 1) { q[0] = s.rho; q[1] = s.vn; q[2] = s.vt1; q[3] = s.vt2; q[4] = s.p; q[5] = s.bt1; q[6] = s.bt2; }
    else if (DIR == 2) { q[0] = s.rho; q[2] = s.vn; q[3] = s.vt1; q[1] = s.vt2; q[4] = s.p; q[6] = s.bt1; q[5] = s.bt2; }
    else { q[0] = s.rho; q[3] = s.vn; q[1] = s.vt1; q[2] = s.vt2; q[4] = s.p; q[5] = s.bt1; q[6] = s.bt2; }
}
// stored flux planes (xyz order) -> rotated hll_face output order
template <int DIR>
__device__ inline void flux_to_rot(const double *o, double *f)
{
    if (DIR == 1) { for (int i = 0; i < 7; i++) f[i] = o[i]; }
    else if (DIR == 2) { f[0] = o[0]; f[3] = o[1]; f[1] = o[2]; f[2] = o[3]; f[4] = o[4]; f[6] = o[5]; f[5] = o[6]; }
    else { f[0] = o[0]; f[2] = o[1]; f[3] = o[2]; f[1] = o[3]; f[4] = o[4]; f[5] = o[5]; f[6] = o[6]; }
}

// HLL reversed at every face of a pencil: writes the left / right state adjoints (q order) to
// H and adds the normal-field adjoint to the face-B plane
template <int DIR, int RECON>
__device__ inline void kr_flux_hll_pencil(const Lay &L, const Fields<double> &F, const Tabs &tb, const Model &M,
                                          const AdjF &A, long t)
{
    if (t >= npencils<DIR>(L)) return;
    Pencil<DIR> pc;
    if (!pc.init(L, t)) return;
    const long ps = L.ps;
    const int ax = DIR - 1;
    const double *bnp = F.P + (4 + DIR) * ps;
    // sliding window as in the forward walk: each cell's slope and edge values computed once;
    // w0..w3 = v(c-1 .. c+2), d0, d1 = dm(c-1), dm(c), vpp = vp(c-1)
    double w0[7], w1[7], w2[7], w3[7], d0[7], d1[7], vpp[7];
    auto load = [&](int c, double *o) {
        const unsigned g = (unsigned)pc.at(c);
        for (int v = 0; v < 7; v++) o[v] = ldi(qplane<DIR>(F, ps, v), g);
    };
    {
        double a[7];
        load(-2, a);
        load(-1, w0);
        load(0, w1);
        load(1, w2);
        if (RECON == 4) {
            const double b0 = tb.at(ax, TB_B0C, -1), b1 = tb.at(ax, TB_B1C, -1);
            const double c0 = tb.at(ax, TB_B0C, 0), c1 = tb.at(ax, TB_B1C, 0);
            for (int v = 0; v < 7; v++) {
                d0[v] = ppm_slope<double>(b0, b1, a[v], w0[v], w1[v]);
                d1[v] = ppm_slope<double>(c0, c1, w0[v], w1[v], w2[v]);
            }
        }
    }
    for (int c = 0; c <= pc.n + 1; c++) {
        load(c + 2, w3);
        double vm[7], vp[7];
        if (RECON == 4) {
            const double b0 = tb.at(ax, TB_B0C, c + 1), b1 = tb.at(ax, TB_B1C, c + 1);
            const double a0m = tb.at(ax, TB_A0M, c), a1m = tb.at(ax, TB_A1M, c), a2m = tb.at(ax, TB_A2M, c);
            const double a3m = tb.at(ax, TB_A3M, c), a0p = tb.at(ax, TB_A0P, c), a1p = tb.at(ax, TB_A1P, c);
            const double a2p = tb.at(ax, TB_A2P, c), a3p = tb.at(ax, TB_A3P, c);
            for (int v = 0; v < 7; v++) {
                const double d2 = ppm_slope<double>(b0, b1, w1[v], w2[v], w3[v]);
                ppm_edges<double>(a0m, a1m, a2m, a3m, a0p, a1p, a2p, a3p, w0[v], w1[v], w2[v], d0[v], d1[v], d2,
                                  vm[v], vp[v]);
                d0[v] = d1[v];
                d1[v] = d2;
            }
        } else {
            const double cf = tb.at(ax, TB_CF, c), cb = tb.at(ax, TB_CB, c);
            for (int v = 0; v < 7; v++) tvd_edges<double>(cf, cb, w0[v], w1[v], w2[v], vm[v], vp[v]);
        }
        if (c > 0) {
            // face c-1: left state vp(c-1), right state vm(c)
            const unsigned gf = (unsigned)pc.at(c - 1);
            const double bn = ldi(bnp, gf);
            double o[7], lf[7];
            for (int i = 0; i < 7; i++) o[i] = ldi(A.F[ax] + i * ps, gf);
            flux_to_rot<DIR>(o, lf);
            FaceState<double> lL{0, 0, 0, 0, 0, 0, 0, 0}, lR{0, 0, 0, 0, 0, 0, 0, 0};
            hll_face_adj(M.ggas, rot<DIR>(vpp, bn), rot<DIR>(vm, bn), lf, lL, lR);
            double hl[7], hr[7];
            unrot<DIR>(lL, hl);
            unrot<DIR>(lR, hr);
            for (int v = 0; v < 7; v++) {
                sti(A.H + v * ps, gf, hl[v]);
                sti(A.H + (7 + v) * ps, gf, hr[v]);
            }
            double *pb = A.P + (4 + DIR) * ps;
            sti(pb, gf, ldi(pb, gf) + (lL.bn + lR.bn));
        }
        for (int v = 0; v < 7; v++) {
            vpp[v] = vp[v];
            w0[v] = w1[v];
            w1[v] = w2[v];
            w2[v] = w3[v];
        }
    }
}

template <int DIR, int RECON>
__global__ __launch_bounds__(64) void kr_flux_hll(Lay L, Fields<double> F, Tabs tb, Model M, AdjF A)
{
    kr_flux_hll_pencil<DIR, RECON>(L, F, tb, M, A, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// reconstruction reversed along a pencil for one variable (blockIdx.y): a sliding window of
// values, slopes and adjoint accumulators; each cell's adjoint is emitted once complete
template <int DIR, int RECON>
__device__ inline void kr_flux_recon_pencil(const Lay &L, const Fields<double> &F, const Tabs &tb, const AdjF &A,
                                            long t, int v)
{
    if (t >= npencils<DIR>(L)) return;
    Pencil<DIR> pc;
    if (!pc.init(L, t)) return;
    const long ps = L.ps;
    const int ax = DIR - 1, n = pc.n;
    const double *W = qplane<DIR>(F, ps, v);
    double *LW = qplane_adj<DIR>(A, ps, v);
    const double *HL = A.H + v * ps, *HR = A.H + (7 + v) * ps;
    auto lvm_of = [&](int c) { return c >= 1 ? HR[pc.at(c - 1)] : 0.0; };
    auto lvp_of = [&](int c) { return c <= n ? HL[pc.at(c)] : 0.0; };
    if (RECON == 4) {
        double w[5], dm[3], aw[4] = {0, 0, 0, 0}, ad[3] = {0, 0, 0};
        for (int i = 0; i < 5; i++) w[i] = W[pc.at(i - 2)];  // w(c-2 .. c+2), c = 0
        auto slope = [&](int s, double a, double b, double c) {
            return ppm_slope<double>(tb.at(ax, TB_B0C, s), tb.at(ax, TB_B1C, s), a, b, c);
        };
        dm[0] = slope(-1, w[0], w[1], w[2]);
        dm[1] = slope(0, w[1], w[2], w[3]);
        dm[2] = slope(1, w[2], w[3], w[4]);
        for (int c = 0; c <= n + 1; c++) {
            ppm_edges_adj(tb.at(ax, TB_A0M, c), tb.at(ax, TB_A1M, c), tb.at(ax, TB_A2M, c), tb.at(ax, TB_A3M, c),
                          tb.at(ax, TB_A0P, c), tb.at(ax, TB_A1P, c), tb.at(ax, TB_A2P, c), tb.at(ax, TB_A3P, c), w[1],
                          w[2], w[3], dm[0], dm[1], dm[2], lvm_of(c), lvp_of(c), aw[1], aw[2], aw[3], ad[0], ad[1],
                          ad[2]);
            // slope c-1 has all its users
            ppm_slope_adj(tb.at(ax, TB_B0C, c - 1), tb.at(ax, TB_B1C, c - 1), w[0], w[1], w[2], ad[0], aw[0], aw[1],
                          aw[2]);
            LW[pc.at(c - 2)] += aw[0];
            aw[0] = aw[1]; aw[1] = aw[2]; aw[2] = aw[3]; aw[3] = 0.0;
            ad[0] = ad[1]; ad[1] = ad[2]; ad[2] = 0.0;
            if (c <= n) {
                w[0] = w[1]; w[1] = w[2]; w[2] = w[3]; w[3] = w[4];
                w[4] = W[pc.at(c + 3)];
                dm[0] = dm[1]; dm[1] = dm[2];
                dm[2] = slope(c + 2, w[2], w[3], w[4]);
            } else {
                w[0] = w[1]; w[1] = w[2]; w[2] = w[3]; w[3] = w[4];
            }
        }
        // now w[0..3] = w(n .. n+3); slopes n+1 and n+2 are complete
        ppm_slope_adj(tb.at(ax, TB_B0C, n + 1), tb.at(ax, TB_B1C, n + 1), w[0], w[1], w[2], ad[0], aw[0], aw[1], aw[2]);
        ppm_slope_adj(tb.at(ax, TB_B0C, n + 2), tb.at(ax, TB_B1C, n + 2), w[1], w[2], w[3], ad[1], aw[1], aw[2], aw[3]);
        for (int i = 0; i < 4; i++) LW[pc.at(n + i)] += aw[i];
    } else {
        double w[3], aw[3] = {0, 0, 0};
        for (int i = 0; i < 3; i++) w[i] = W[pc.at(i - 1)];  // w(c-1 .. c+1), c = 0
        for (int c = 0; c <= n + 1; c++) {
            tvd_edges_adj(tb.at(ax, TB_CF, c), tb.at(ax, TB_CB, c), w[0], w[1], w[2], lvm_of(c), lvp_of(c), aw[0], aw[1],
                          aw[2]);
            LW[pc.at(c - 1)] += aw[0];
            aw[0] = aw[1]; aw[1] = aw[2]; aw[2] = 0.0;
            w[0] = w[1]; w[1] = w[2];
            if (c <= n) w[2] = W[pc.at(c + 2)];
        }
        LW[pc.at(n + 1)] += aw[0];
        LW[pc.at(n + 2)] += aw[1];
    }
}

template <int DIR, int RECON>
__global__ __launch_bounds__(64) void kr_flux_recon(Lay L, Fields<double> F, Tabs tb, AdjF A)
{
    kr_flux_recon_pencil<DIR, RECON>(L, F, tb, A, (long)blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

// The z face of a z-collapsed box reversed: one face per (j, k) column, both of whose cells are
// the column itself.  The forward reconstructs it from a constant stencil, which PPM (flattened:
// (vr - vc)(vc - vl) = 0) and TVD (zero slope) both return unchanged, so the face states are the
// cell values and the Riemann reverse's state adjoints go straight to them.  The forward values
// are recomputed with the same calls as the forward sweep (cell 0's edges; cell 1's are equal,
// its coefficient rows being the same).  Thread t takes column (j, k), j fastest, over
// [0, nx+1] x [0, ny+1] (the forward z sweep's columns).
template <int RECON>
__device__ inline void kr_flux_z1_cell(const Lay &L, const Fields<double> &F, const Tabs &tb, const Model &M,
                                       const AdjF &A, long t)
{
    const long nj = L.nx + 2;
    if (t >= nj * (L.ny + 2)) return;
    const int j = (int)(t % nj), k = (int)(t / nj);
    const unsigned g = (unsigned)L.id(j, k, 1);
    const long ps = L.ps;
    double vm[7], vp[7], o[7], lf[7];
    bool fin = true;
#pragma unroll
    for (int v = 0; v < 7; v++) {
        vm[v] = ldi(qplane<3>(F, ps, v), g);
        fin = fin && isfinite(vm[v]);
    }
    if (fin) {
#pragma unroll
        for (int v = 0; v < 7; v++) const_edges<double, RECON>(vm[v], vm[v], vp[v]);
    } else {
#pragma unroll
        for (int v = 0; v < 7; v++) {
            const double q = vm[v];
            const double w[5] = {q, q, q, q, q};
            recon_cell<RECON>(tb, 2, 0, w, vm[v], vp[v]);
        }
    }
    const double bn = ldi(F.P + 7 * ps, g);
#pragma unroll
    for (int i = 0; i < 7; i++) o[i] = ldi(A.F[2] + i * ps, g);
    flux_to_rot<3>(o, lf);
    FaceState<double> lL{0, 0, 0, 0, 0, 0, 0, 0}, lR{0, 0, 0, 0, 0, 0, 0, 0};
    hll_face_adj(M.ggas, rot<3>(vp, bn), rot<3>(vm, bn), lf, lL, lR);
    double hl[7], hr[7];
    unrot<3>(lL, hl);
    unrot<3>(lR, hr);
    double *pb = A.P + 7 * ps;
    sti(pb, g, ldi(pb, g) + (lL.bn + lR.bn));
#pragma unroll
    for (int v = 0; v < 7; v++) {
        double *lp = qplane_adj<3>(A, ps, v);
        sti(lp, g, ldi(lp, g) + (hl[v] + hr[v]));
    }
}

// ---------------------------------------------------------------------------------------------
// A sweep reversed, wave-parallel like k_flux_x: the lanes of a wavefront hold 64 consecutive
// cells along the sweep (one outer index: plane l for the x and y sweeps, row k for the z
// sweep); each lane recomputes its cell's reconstruction (neighbours by shuffles), reverses the
// HLL solve of its face, then the edge values and slope of its cell, and gathers the adjoint of
// its cell value from the +-1 lanes.  Lanes 5..58 own their cells (54 per wave); the others only
// feed their neighbours.  Along x the rows k in [0, ny+1] are laid end to end at pitch px (row
// padding included) and loaded straight from HBM; along y and z each wavefront takes one pencil
// segment and the workgroup stages a 4-pencil tile through LDS (kr_flux_lane_t_tile).
// ---------------------------------------------------------------------------------------------
constexpr int XR_OWN = 54;
// pencils shorter than this (n < LANE_MIN) are reversed by the per-pencil walks instead
constexpr int LANE_MIN = 8;

// cells of the end-to-end x rows of one plane
__host__ __device__ inline long lane_cells_x(const Lay &L) { return (long)(L.ny + 2) * L.px; }
template <int DIR>
__host__ __device__ inline int lane_outer(const Lay &L)
{
    return DIR == 3 ? L.ny + 2 : flux_planes(L);
}
// 256-thread tiles per outer index: along x, 4 waves of XR_OWN owned cells of the row run; along
// y and z, 4 pencils x one XR_OWN-cell segment (kr_flux_lane_t_tile)
template <int DIR>
__host__ __device__ inline int lane_tiles(const Lay &L)
{
    if (DIR != 1) return ((L.nx + 2 + 3) / 4) * (((DIR == 2 ? L.ny : L.nz) + 6 + XR_OWN - 1) / XR_OWN);
    const long waves = (lane_cells_x(L) + XR_OWN - 1) / XR_OWN;
    return (int)((waves + 3) / 4);
}
template <int DIR>
__host__ __device__ inline bool lane_sweep(const Lay &L)
{
    return DIR == 1 || (DIR == 2 ? L.ny : L.nz) >= LANE_MIN;
}

// The lane-parallel reverse of one wavefront: inputs per lane are its cell's 7 reconstructed
// values q, its face's 7 flux adjoints lfo (stored xyz order; read on face lanes only) and normal
// field bn; outputs are the adjoint increments lw of its cell values and lbn of its face field
// (meaningful on the lanes that own them).  c is the lane's cell along the sweep (-100 off the
// pencils), n the interior cell count along it.  Every lane of the wavefront must call it.
template <int DIR, int RECON>
__device__ inline void lane_reverse(const Tabs &tb, const Model &M, int n, int lane, int c, bool face,
                                    const double *q, const double *lfo, double bn, double *lw, double &lbn)
{
    const int ax = DIR - 1;
    const int cc = c < -2 ? -2 : (c > n + 3 ? n + 3 : c);
    double dm[7], vp[7], vmr[7];
    // forward: edge values of this cell; the right state of face c is vm of lane+1
#pragma unroll
    for (int v = 0; v < 7; v++) {
        const double w = q[v], wm1 = lane_up(w), wp1 = lane_down(w);
        double vm, vpv;
        if (RECON == 4) {
            dm[v] = ppm_slope<double>(tb.at(ax, TB_B0C, cc), tb.at(ax, TB_B1C, cc), wm1, w, wp1);
            const double dmm1 = lane_up(dm[v]), dmp1 = lane_down(dm[v]);
            ppm_edges<double>(tb.at(ax, TB_A0M, cc), tb.at(ax, TB_A1M, cc), tb.at(ax, TB_A2M, cc),
                              tb.at(ax, TB_A3M, cc), tb.at(ax, TB_A0P, cc), tb.at(ax, TB_A1P, cc),
                              tb.at(ax, TB_A2P, cc), tb.at(ax, TB_A3P, cc), wm1, w, wp1, dmm1, dm[v], dmp1, vm, vpv);
        } else {
            tvd_edges<double>(tb.at(ax, TB_CF, cc), tb.at(ax, TB_CB, cc), wm1, w, wp1, vm, vpv);
        }
        vp[v] = vpv;
        vmr[v] = lane_down(vm);
    }
    // HLL reversed at face c
    double lL[7] = {0, 0, 0, 0, 0, 0, 0}, lR[7] = {0, 0, 0, 0, 0, 0, 0};
    lbn = 0.0;
    if (face) {
        double lf[7];
        flux_to_rot<DIR>(lfo, lf);
        FaceState<double> a{0, 0, 0, 0, 0, 0, 0, 0}, b{0, 0, 0, 0, 0, 0, 0, 0};
        hll_face_adj(M.ggas, rot<DIR>(vp, bn), rot<DIR>(vmr, bn), lf, a, b);
        unrot<DIR>(a, lL);
        unrot<DIR>(b, lR);
        lbn = a.bn + b.bn;
    }
    // reconstruction reversed, one variable at a time
#pragma unroll
    for (int v = 0; v < 7; v++) {
        const double lvp = lL[v];
        const double lvm = lane_up(lR[v]);  // from face c-1
        const double w = q[v], wm1 = lane_up(w), wp1 = lane_down(w);
        double am1 = 0.0, a0 = 0.0, ap1 = 0.0, em1 = 0.0, e0 = 0.0, ep1 = 0.0;
        // only cells 0..n+1 are reconstructed and only slopes -1..n+2 are used; elsewhere the
        // table entries are not defined and the forward values there are never read
        const bool recon_cell = c >= 0 && c <= n + 1;
        if (RECON == 4) {
            const double dmm1 = lane_up(dm[v]), dmp1 = lane_down(dm[v]);
            if (recon_cell)
                ppm_edges_adj(tb.at(ax, TB_A0M, cc), tb.at(ax, TB_A1M, cc), tb.at(ax, TB_A2M, cc),
                              tb.at(ax, TB_A3M, cc), tb.at(ax, TB_A0P, cc), tb.at(ax, TB_A1P, cc),
                              tb.at(ax, TB_A2P, cc), tb.at(ax, TB_A3P, cc), wm1, w, wp1, dmm1, dm[v], dmp1, lvm, lvp,
                              am1, a0, ap1, em1, e0, ep1);
            // slope of this cell: its adjoint from the three cells that used it
            const double ldm = e0 + lane_down(em1) + lane_up(ep1);
            double bm1 = 0.0, b0 = 0.0, bp1 = 0.0;
            if (c >= -1 && c <= n + 2)
                ppm_slope_adj(tb.at(ax, TB_B0C, cc), tb.at(ax, TB_B1C, cc), wm1, w, wp1, ldm, bm1, b0, bp1);
            lw[v] = a0 + lane_down(am1) + lane_up(ap1) + b0 + lane_down(bm1) +
                    lane_up(bp1);
        } else {
            if (recon_cell)
                tvd_edges_adj(tb.at(ax, TB_CF, cc), tb.at(ax, TB_CB, cc), wm1, w, wp1, lvm, lvp, am1, a0, ap1);
            lw[v] = a0 + lane_down(am1) + lane_up(ap1);
        }
    }
}

// x sweep: a wavefront's lanes load 64 consecutive cells of a row (coalesced) straight from HBM
template <int RECON>
__device__ inline void kr_flux_lane_x_tile(const Lay &L, const Fields<double> &F, const Tabs &tb, const Model &M,
                                           const AdjF &A, int bx, int o)
{
    const int lane = threadIdx.x & 63;
    const long wave = ((long)bx * blockDim.x + threadIdx.x) >> 6;
    const long f = wave * XR_OWN + lane - 5;
    const bool inr = (f >= 0) && (f < lane_cells_x(L));
    const long fc = inr ? f : 0;
    const int n = L.nx;
    const int k = (int)(fc / L.px);
    const int c = inr ? (int)(fc - (long)k * L.px) - 2 : -100;
    const unsigned g = (unsigned)((long)(o + 2) * L.py + 2 * L.px + fc);
    const long ps = L.ps;
    const bool own = lane >= 5 && lane <= 58 && inr && c <= n + 3;
    // a face exists for c in [0, n]; lanes 2..60 have the full forward stencil
    const bool face = inr && c >= 0 && c <= n && lane >= 2 && lane <= 60;
    double q[7], lfo[7] = {0, 0, 0, 0, 0, 0, 0}, bn = 0.0;
#pragma unroll
    for (int v = 0; v < 7; v++) q[v] = ldi(qplane<1>(F, ps, v), g);
    if (face) {
        bn = ldi(F.P + 5 * ps, g);
#pragma unroll
        for (int i = 0; i < 7; i++) lfo[i] = ldi(A.F[0] + i * ps, g);
    }
    double lw[7], lbn;
    lane_reverse<1, RECON>(tb, M, n, lane, c, face, q, lfo, bn, lw, lbn);
    if (own) {
        if (face) sti(A.P + 5 * ps, g, ldi(A.P + 5 * ps, g) + lbn);
#pragma unroll
        for (int v = 0; v < 7; v++) {
            double *lp = qplane_adj<1>(A, ps, v);
            sti(lp, g, ldi(lp, g) + lw[v]);
        }
    }
}

// y / z sweeps: a 256-thread workgroup takes 4 neighbouring pencils (columns j0..j0+3 at outer
// index o) over 64 consecutive cells.  The 15 input planes of that 64 x 4 tile are loaded row by
// row (4 consecutive cells of an x row per load) into LDS, each wavefront runs the lane-parallel
// reverse on one pencil from LDS, and the cell adjoints return through LDS to row-wise
// read-add-writes of the owned cells.  Every thread reaches the barriers.
constexpr int LT_PITCH = 5;  // LDS row pitch of a 4-pencil tile (one pad double)
constexpr int LT_PLANES = 15;

template <int DIR>
__host__ __device__ inline int lane_t_segs(const Lay &L)
{
    return ((DIR == 2 ? L.ny : L.nz) + 6 + XR_OWN - 1) / XR_OWN;
}

template <int DIR, int RECON>
__device__ inline void kr_flux_lane_t_tile(const Lay &L, const Fields<double> &F, const Tabs &tb, const Model &M,
                                           const AdjF &A, int bx, int o, double (*buf)[64][LT_PITCH])
{
    const int n = DIR == 2 ? L.ny : L.nz;
    // consecutive tiles are neighbouring 4-pencil groups of one segment: four of them read the
    // same 128-byte rows, which meet in one L2 when the caller keeps them on one XCD
    const int njt = (L.nx + 2 + 3) / 4;
    const int seg = bx / njt, j0 = (bx % njt) * 4;
    const long ps = L.ps;
    auto cell = [&](int j, int c) { return (unsigned)(DIR == 2 ? L.id(j, c, o) : L.id(j, o, c)); };
    // row r of the tile is pencil cell f = seg * XR_OWN + r - 5 (c = f - 2)
    const int tr = threadIdx.x >> 2, tp = threadIdx.x & 3;
    const int tf = seg * XR_OWN + tr - 5, tj = j0 + tp;
    const bool tin = tf >= 0 && tf < n + 6 && tj <= L.nx + 1;
    const unsigned tg = tin ? cell(tj, tf - 2) : 0u;
#pragma unroll
    for (int v = 0; v < 7; v++) buf[v][tr][tp] = tin ? ldi(qplane<DIR>(F, ps, v), tg) : 0.0;
#pragma unroll
    for (int i = 0; i < 7; i++) buf[7 + i][tr][tp] = tin ? ldi(A.F[DIR - 1] + i * ps, tg) : 0.0;
    buf[14][tr][tp] = tin ? ldi(F.P + (4 + DIR) * ps, tg) : 0.0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int f = seg * XR_OWN + lane - 5;
    const bool inr = f >= 0 && f < n + 6 && j0 + w <= L.nx + 1;
    const int c = inr ? f - 2 : -100;
    const bool face = inr && c >= 0 && c <= n && lane >= 2 && lane <= 60;
    double q[7], lfo[7];
#pragma unroll
    for (int v = 0; v < 7; v++) {
        q[v] = buf[v][lane][w];
        lfo[v] = buf[7 + v][lane][w];
    }
    const double bn = buf[14][lane][w];
    double lw[7], lbn;
    lane_reverse<DIR, RECON>(tb, M, n, lane, c, face, q, lfo, bn, lw, lbn);
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 7; v++) buf[v][lane][w] = lw[v];
    buf[7][lane][w] = lbn;
    __syncthreads();
    // owned rows 5..58 of the tile; a face exists at c in [0, n]
    if (tin && tr >= 5 && tr <= 58) {
        const int tc = tf - 2;
        if (tc >= 0 && tc <= n) {
            double *pb = A.P + (4 + DIR) * ps;
            sti(pb, tg, ldi(pb, tg) + buf[7][tr][tp]);
        }
#pragma unroll
        for (int v = 0; v < 7; v++) {
            double *lp = qplane_adj<DIR>(A, ps, v);
            sti(lp, tg, ldi(lp, tg) + buf[v][tr][tp]);
        }
    }
    __syncthreads();  // the next tile of a persistent loop reuses buf
}

template <int DIR, int RECON>
__global__ __launch_bounds__(256) void kr_flux_lane(Lay L, Fields<double> F, Tabs tb, Model M, AdjF A)
{
    if constexpr (DIR == 1) {
        kr_flux_lane_x_tile<RECON>(L, F, tb, M, A, blockIdx.x, blockIdx.y);
    } else {
        // XCD-aware tile order (neighbouring tiles share cache lines); padding blocks leave
        __shared__ double buf[LT_PLANES][64][LT_PITCH];
        int bx, o;
        if (!xcd_tile(lane_tiles<DIR>(L), lane_outer<DIR>(L), bx, o)) return;
        kr_flux_lane_t_tile<DIR, RECON>(L, F, tb, M, A, bx, o, buf);
    }
}

// ---------------------------------------------------------------------------------------------
// initial_update reversed (k_init: CHECKRHO, FROMRVETOU, EOS); the ghost fill that follows it is
// reversed by kr_ghost first.  Forward inputs: the uploaded prim / eps (and the zero ghost bcell).
// ---------------------------------------------------------------------------------------------
__device__ inline void kr_init_a_cell(const Lay &L, const Fields<double> &F, const Model &M, const AdjF &A, double *lU,
                                      long t, int l)
{
    const int k = (int)(t / L.px), j = (int)(t - (long)k * L.px);
    if (k > L.ny || j > L.nx) return;
    const long g = L.id(j, k, l), sj = 1, sk = L.px, sl = L.py, ps = L.ps;
    const bool interior = j >= 1 && k >= 1 && l >= 1;
    const double gm1 = M.ggas - 1.0;
    const double *B = F.P + 5 * ps;
    double bc[3];
    if (interior) {
        bc[0] = 0.5 * (B[g] + B[g - sj]);
        bc[1] = 0.5 * (B[g + ps] + B[g - sk + ps]);
        bc[2] = 0.5 * (B[g + 2 * ps] + B[g - sl + 2 * ps]);
    } else {
        bc[0] = F.BC[g]; bc[1] = F.BC[g + ps]; bc[2] = F.BC[g + 2 * ps];
    }
    const double rho = F.P[g], vx = F.P[g + ps], vy = F.P[g + 2 * ps], vz = F.P[g + 3 * ps];
    const double eps_in = F.EPS[g];
    const bool floored = interior && (0.0 > eps_in);
    const double eps = floored ? 0.0 : eps_in;
    const double vsq = vx * vx + vy * vy + vz * vz;
    double lUv[8];
    for (int v = 0; v < 8; v++) {
        lUv[v] = lU[g + v * ps];
        lU[g + v * ps] = 0.0;
    }
    const double lP0 = A.P[g], lP4 = A.P[g + 4 * ps], lE = A.EPS[g];
    const double lrho = lP0 + lUv[0] + lUv[1] * vx + lUv[2] * vy + lUv[3] * vz + lUv[4] * (eps + 0.5 * vsq) +
                        lP4 * eps * gm1;
    A.P[g] = lrho;
    A.P[g + ps] += lUv[1] * rho + lUv[4] * rho * vx;
    A.P[g + 2 * ps] += lUv[2] * rho + lUv[4] * rho * vy;
    A.P[g + 3 * ps] += lUv[3] * rho + lUv[4] * rho * vz;
    A.P[g + 4 * ps] = 0.0;  // p is recomputed from rho and eps
    const double leps = lE + lUv[4] * rho + lP4 * rho * gm1;
    A.EPS[g] = floored ? 0.0 : leps;
    for (int c = 0; c < 3; c++) {
        const double lbc = (interior ? A.BC[g + c * ps] : 0.0) + lUv[4] * bc[c];
        A.T[g + (8 + c) * ps] = interior ? 0.5 * lbc : 0.0;
        if (interior) A.BC[g + c * ps] = 0.0;
        else A.BC[g + c * ps] += lbc;
        A.T[g + (5 + c) * ps] = lUv[5 + c];
    }
}

__global__ __launch_bounds__(256) void kr_init_a(Lay L, Fields<double> F, Model M, AdjF A, double *lU)
{
    kr_init_a_cell(L, F, M, A, lU, (long)blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

__device__ inline void kr_init_b_cell(const Lay &L, const AdjF &A, long t, int l)
{
    const int k = (int)(t / L.px), j = (int)(t - (long)k * L.px);
    if (k > L.ny || j > L.nx) return;
    const long g = L.id(j, k, l), ps = L.ps;
    const bool in = j >= 1 && k >= 1 && l >= 1;
    A.P[g + 5 * ps] += A.T[g + 5 * ps] + (in ? A.T[g + 8 * ps] : 0.0) +
                       ((j + 1 <= L.nx && k >= 1 && l >= 1) ? A.T[g + 1 + 8 * ps] : 0.0);
    A.P[g + 6 * ps] += A.T[g + 6 * ps] + (in ? A.T[g + 9 * ps] : 0.0) +
                       ((k + 1 <= L.ny && j >= 1 && l >= 1) ? A.T[g + L.px + 9 * ps] : 0.0);
    A.P[g + 7 * ps] += A.T[g + 7 * ps] + (in ? A.T[g + 10 * ps] : 0.0) +
                       ((l + 1 <= L.nz && j >= 1 && k >= 1) ? A.T[g + L.py + 10 * ps] : 0.0);
}

__global__ __launch_bounds__(256) void kr_init_b(Lay L, AdjF A)
{
    kr_init_b_cell(L, A, (long)blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------
struct cumc_tape {
    int nsteps = 0, every = 1;
    std::vector<double> dts;
    double *mem = nullptr;     // one allocation
    double *init = nullptr;    // uploaded prim (8) + eps (1)
    double *ckpt = nullptr;    // nck states of SW planes: P 8, EPS 1, BC 3, U 8
    int nck = 0;
    double *seg = nullptr;     // every + 1 states
    double *sa = nullptr;      // every states after stage 1: P/EPS/BC and the stage's U (U1)
    double *sb = nullptr;      // every states after stage 2: P/EPS/BC and the stage's U (Ub)
    // with two sets of segment buffers (the batched backward re-runs one segment while it
    // reverses the next), set s's slot i of seg / sa / sb is slot i + s * set_stride
    int sets = 1, set_stride = 0;
    bool zcol = false;  // recorded in the z-collapsed view (batched tapes): states are plane-1 slabs
    double *adj = nullptr;     // NADJ planes
    double *lu[3] = {nullptr, nullptr, nullptr};  // 8 planes each: U adjoints by role U0, U1, Ub
    AdjF A;
    // seg / sa / sb hold the last checkpoint segment, stored by tape_run's own forward pass, so
    // the backward does not re-run that segment (with every >= nsteps it re-runs nothing)
    bool last_stored = false;
};
static constexpr int SW = 20;  // planes of a stored state

void tape_release(cumc_ctx *ctx)
{
    if (!ctx || !ctx->tape) return;
    hipFree(ctx->tape->mem);
    delete ctx->tape;
    ctx->tape = nullptr;
    ctx->ztape = false;
}

static inline size_t plane_bytes(const cumc_ctx *ctx) { return (size_t)ctx->L.ps * sizeof(double); }

// P, EPS, BC are the first 12 planes of the pool; U of the current step start is F64.U0
static void save_pool_u(cumc_ctx *ctx, double *dst, const double *U)
{
    const size_t pb = plane_bytes(ctx);
    hipMemcpyAsync(dst, ctx->pool, 12 * pb, hipMemcpyDeviceToDevice, ctx->stream);
    hipMemcpyAsync(dst + 12 * ctx->L.ps, U, 8 * pb, hipMemcpyDeviceToDevice, ctx->stream);
}
static void save_state(cumc_ctx *ctx, double *dst) { save_pool_u(ctx, dst, ctx->F64.U0); }
static void load_state(cumc_ctx *ctx, const double *src, bool with_u = true)
{
    const size_t pb = plane_bytes(ctx);
    hipMemcpyAsync(ctx->pool, src, 12 * pb, hipMemcpyDeviceToDevice, ctx->stream);
    if (with_u) hipMemcpyAsync(ctx->F64.U0, src + 12 * ctx->L.ps, 8 * pb, hipMemcpyDeviceToDevice, ctx->stream);
}

static void set_dt(cumc_ctx *ctx, double dt)
{
    hipMemcpyAsync(ctx->dt_dev, &dt, sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    hipStreamSynchronize(ctx->stream);  // dt lives on the host stack
}

static void forward_step(cumc_ctx *ctx, double dt)
{
    set_dt(ctx, dt);
    for (int st = 1; st <= 3; st++) stage<double>(ctx, ctx->F64, st);
    swap_u(ctx);
}

template <int DIR>
static void reverse_sweep(cumc_ctx *ctx, const Fields<double> &F, const AdjF &A)
{
    const Lay &L = ctx->L;
    Tabs tb = tabs_of(ctx);
    hipStream_t s = ctx->stream;
    if (lane_sweep<DIR>(L)) {
        const dim3 grid = DIR == 1 ? dim3(lane_tiles<DIR>(L), lane_outer<DIR>(L))
                                   : dim3(xcd_blocks(lane_tiles<DIR>(L), lane_outer<DIR>(L)));
        if (ctx->M.recon == CUMC_PPMC) hipLaunchKernelGGL((kr_flux_lane<DIR, 4>), grid, dim3(256), 0, s, L, F, tb, ctx->M, A);
        else hipLaunchKernelGGL((kr_flux_lane<DIR, 1>), grid, dim3(256), 0, s, L, F, tb, ctx->M, A);
        return;
    }
    if constexpr (DIR != 1) {
        const long np = npencils<DIR>(L);
        if (ctx->M.recon == CUMC_PPMC) {
            hipLaunchKernelGGL((kr_flux_hll<DIR, 4>), dim3(nblk(np, 64)), dim3(64), 0, s, L, F, tb, ctx->M, A);
            hipLaunchKernelGGL((kr_flux_recon<DIR, 4>), dim3(nblk(np, 64), 7), dim3(64), 0, s, L, F, tb, A);
        } else {
            hipLaunchKernelGGL((kr_flux_hll<DIR, 1>), dim3(nblk(np, 64)), dim3(64), 0, s, L, F, tb, ctx->M, A);
            hipLaunchKernelGGL((kr_flux_recon<DIR, 1>), dim3(nblk(np, 64), 7), dim3(64), 0, s, L, F, tb, A);
        }
    }
}

static void reverse_ghost(cumc_ctx *ctx, const AdjF &A, const double *rho_out, const double *eps_out)
{
    const Lay &L = ctx->L;
    const long n0 = 6 * kr_ghost_lines<0>(L), n1 = 6 * kr_ghost_lines<1>(L), n2 = 6 * kr_ghost_lines<2>(L);
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(kr_ghost<2>, dim3(nblk(n2, 256)), dim3(256), 0, s, L, A, ctx->M, rho_out, eps_out);
    hipLaunchKernelGGL(kr_ghost<1>, dim3(nblk(n1, 256)), dim3(256), 0, s, L, A, ctx->M, rho_out, eps_out);
    hipLaunchKernelGGL(kr_ghost<0>, dim3(nblk(n0, 256)), dim3(256), 0, s, L, A, ctx->M, rho_out, eps_out);
}

// reverse one stage: forward input state `in` (P, EPS, bcell of the stored state before the
// stage), forward output state `out` (after the stage and its ghost fill), U0 / Uin its conserved
// inputs.  No forward kernel runs: the sweeps recompute their reconstruction and Riemann solves
// from `in`, the tail reads `out`.
template <int STAGE>
static void reverse_stage(cumc_ctx *ctx, const double *in, const double *out, double dt, const double *U0,
                          const double *Uin, double *lU0, double *lUin, double *lUout)
{
    cumc_tape *tp = ctx->tape;
    const AdjF &A = tp->A;
    const Lay &L = ctx->L;
    hipStream_t s = ctx->stream;
    Tabs tb = tabs_of(ctx);
    const Fields<double> F = fields_of_state(ctx->F64, in, L.ps);
    reverse_ghost(ctx, A, out, out + 8 * L.ps);
    dim3 g0(nblk((long)L.px * (L.ny + 1), 256), L.nz + 1);
    hipLaunchKernelGGL((kr_tail_a<STAGE>), g0, dim3(256), 0, s, L, F, out, ctx->M, dt, U0, Uin, A, lU0, lUin, lUout);
    hipLaunchKernelGGL((kr_tail_b<STAGE>), g0, dim3(256), 0, s, L, A, dt, lU0, lUin);
    dim3 g1(nblk((long)L.px * (L.ny + 2), 256), L.nz + 2);
    hipLaunchKernelGGL(kr_gather, g1, dim3(256), 0, s, L, A, tb);
    hipLaunchKernelGGL(kr_ecorn, g1, dim3(256), 0, s, L, F, A);
    reverse_sweep<3>(ctx, F, A);
    reverse_sweep<2>(ctx, F, A);
    reverse_sweep<1>(ctx, F, A);
}

// reverse step m: S_m in seg slot i, its stage-1 / stage-2 states in sa / sb slot i, S_{m+1} in
// seg slot i+1; the adjoint of S_{m+1} in A / lu[U1]
static void reverse_step(cumc_ctx *ctx, int i, double dt)
{
    cumc_tape *tp = ctx->tape;
    const size_t st = (size_t)SW * ctx->L.ps;
    const double *Sm = tp->seg + i * st, *Sn = Sm + st, *Sa = tp->sa + i * st, *Sb = tp->sb + i * st;
    const double *U0 = Sm + 12 * ctx->L.ps;
    double *&l0 = tp->lu[0], *&l1 = tp->lu[1], *&lb = tp->lu[2];
    reverse_stage<3>(ctx, Sb, Sn, dt, U0, Sb + 12 * ctx->L.ps, l0, lb, l1);
    reverse_stage<2>(ctx, Sa, Sb, dt, U0, Sa + 12 * ctx->L.ps, l0, l1, lb);
    reverse_stage<1>(ctx, Sm, Sa, dt, U0, nullptr, l0, nullptr, l1);
    // the adjoint of S_m's U becomes the output-role adjoint of step m-1
    double *tmp = l0;
    l0 = l1;
    l1 = tmp;
}

// one allocation per context: uploaded state, checkpoints, segment states, stage saves,
// adjoint arrays and the three conserved-state adjoints
static int tape_alloc(cumc_ctx *ctx, int nsteps, int every, bool zcollapse_ok = false, int sets = 1)
{
    tape_release(ctx);
    // a single context's taped steps and their reversal run in the full layout (the stored states
    // keep their z ghost planes, which its reverse kernels read); the batched tape reverses
    // z-collapsed boxes in their view
    zexpand(ctx);
    cumc_tape *tp = new cumc_tape();
    ctx->tape = tp;
    ctx->ztape = !zcollapse_ok;
    tp->nsteps = nsteps;
    tp->every = every;
    tp->nck = nsteps / every + 2;  // step 0, every, 2 every, ... and the final state
    tp->sets = sets;
    tp->set_stride = sets > 1 ? every + 1 : 0;
    const size_t ps = ctx->L.ps;
    const size_t seg_states = sets > 1 ? (size_t)3 * (every + 1) * sets : (size_t)3 * every + 1;
    const size_t planes = 9 + (size_t)SW * tp->nck + (size_t)SW * seg_states + NADJ + 24;
    if (hipMalloc(&tp->mem, planes * ps * sizeof(double)) != hipSuccess) {
        tape_release(ctx);
        return fail(ctx, CUMC_ENOMEM, "tape allocation");
    }
    double *p = tp->mem;
    auto take = [&](size_t n) { double *r = p; p += n * ps; return r; };
    tp->init = take(9);
    tp->ckpt = take((size_t)SW * tp->nck);
    const size_t per = sets > 1 ? (size_t)(every + 1) * sets : 0;
    tp->seg = take((size_t)SW * (sets > 1 ? per : every + 1));
    tp->sa = take((size_t)SW * (sets > 1 ? per : every));
    tp->sb = take((size_t)SW * (sets > 1 ? per : every));
    tp->adj = take(NADJ);
    for (int r = 0; r < 3; r++) tp->lu[r] = take(8);
    AdjF &A = tp->A;
    double *q = tp->adj;
    auto sub = [&](int n) { double *r = q; q += (size_t)n * ps; return r; };
    A.P = sub(8); A.BC = sub(3); A.EPS = sub(1);
    for (int d = 0; d < 3; d++) A.F[d] = sub(7);
    A.E = sub(3); A.T = sub(11); A.H = sub(14);
    return CUMC_OK;
}

static bool adjoint_supported(const cumc_ctx *ctx, std::string &why)
{
    const cumc_config &c = ctx->cfg;
    if (c.precision != CUMC_FP64) { why = "reverse mode needs the FP64 form"; return false; }
    if (c.coordinate != CUMC_CARTESIAN) { why = "reverse mode covers Cartesian grids"; return false; }
    if (c.model != CUMC_MODEL_RIEMANN) { why = "reverse mode covers the riemann model hooks"; return false; }
    if (ctx->M.solver != CUMC_HLL || (c.recon != CUMC_PPMC && c.recon != CUMC_TVDMM)) {
        why = "reverse mode covers the reference's HLL solver with PPM or TVD minmod";
        return false;
    }
    if (c.boundary_flag[2] != 0 && c.boundary_flag[3] == 0) {
        why = "y-outer periodic with y-inner non-periodic leaves ghost eps stale (BOUNDARY1D); not differentiated";
        return false;
    }
    return true;
}

extern "C" int cumc_tape_run(cumc_ctx *ctx, int nsteps, int every, const double *dt_in, double *dt_out)
{
    if (!ctx || nsteps < 0 || every < 1) return CUMC_EINVAL;
    if (!ctx->grid_set) return fail(ctx, CUMC_ESTATE, "no grid");
    std::string why;
    if (!adjoint_supported(ctx, why)) return fail(ctx, CUMC_EINVAL, why);
    int r = tape_alloc(ctx, nsteps, every);
    if (r) return r;
    cumc_tape *tp = ctx->tape;
    const size_t ps = ctx->L.ps;
    const size_t pb = plane_bytes(ctx);
    // the uploaded state, then initial_update and the steps
    HIPCHK(ctx, hipMemcpyAsync(tp->init, ctx->F64.P, 8 * pb, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(tp->init + 8 * ps, ctx->F64.EPS, pb, hipMemcpyDeviceToDevice, ctx->stream));
    if ((r = cumc_initial_update(ctx))) return r;
    tp->dts.assign(nsteps, 0.0);
    save_state(ctx, tp->ckpt);
    // the last segment's steps store their stage states as the backward's re-run would
    const int s_last = nsteps > 0 ? ((nsteps - 1) / every) * every : 0;
    for (int m = 0; m < nsteps; m++) {
        double dt;
        if (dt_in) {
            dt = dt_in[m];
            if (!std::isfinite(dt) || !(dt > 0.0)) return fail(ctx, CUMC_ENONFINITE, "frozen dt must be finite and positive");
        } else if ((r = cumc_finddt(ctx, &dt))) {
            return r;
        }
        tp->dts[m] = dt;
        if (m < s_last) {
            forward_step(ctx, dt);
        } else {
            const size_t at = (size_t)SW * ps * (m - s_last);
            if (m == s_last) save_state(ctx, tp->seg);
            set_dt(ctx, dt);
            stage<double>(ctx, ctx->F64, 1);
            save_pool_u(ctx, tp->sa + at, ctx->F64.U1);
            stage<double>(ctx, ctx->F64, 2);
            save_pool_u(ctx, tp->sb + at, ctx->Ub);
            stage<double>(ctx, ctx->F64, 3);
            swap_u(ctx);
            save_state(ctx, tp->seg + at + (size_t)SW * ps);
        }
        if ((m + 1) % every == 0) save_state(ctx, tp->ckpt + (size_t)SW * ps * ((m + 1) / every));
    }
    tp->last_stored = nsteps > 0;
    save_state(ctx, tp->ckpt + (size_t)SW * ps * (tp->nck - 1));
    ctx->dt_valid = false;
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (dt_out) std::memcpy(dt_out, tp->dts.data(), sizeof(double) * nsteps);
    return CUMC_OK;
}

static int aos_into(cumc_ctx *ctx, const double *host, int nv, double *dst)
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

static int soa_out(cumc_ctx *ctx, const double *src, int nv, double *host)
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

extern "C" int cumc_tape_backward(cumc_ctx *ctx, const double *seed_prim, const double *seed_eps,
                                  const double *seed_bcell, double *grad_prim, double *grad_eps)
{
    if (!ctx) return CUMC_EINVAL;
    cumc_tape *tp = ctx->tape;
    if (!tp) return fail(ctx, CUMC_ESTATE, "cumc_tape_run must precede cumc_tape_backward");
    const Lay &L = ctx->L;
    const size_t ps = L.ps, pb = plane_bytes(ctx);
    AdjF &A = tp->A;
    HIPCHK(ctx, hipMemsetAsync(A.P, 0, (size_t)NADJ * pb, ctx->stream));
    for (int r = 0; r < 3; r++) HIPCHK(ctx, hipMemsetAsync(tp->lu[r], 0, 8 * pb, ctx->stream));
    int r;
    if (seed_prim && (r = aos_into(ctx, seed_prim, 8, A.P))) return r;
    if (seed_eps && (r = aos_into(ctx, seed_eps, 1, A.EPS))) return r;
    if (seed_bcell && (r = aos_into(ctx, seed_bcell, 3, A.BC))) return r;
    const int N = tp->nsteps, k = tp->every;
    for (int s0 = ((N - 1) / k) * k; N > 0 && s0 >= 0; s0 -= k) {
        const int s1 = s0 + k < N ? s0 + k : N;
        const bool stored = s0 == ((N - 1) / k) * k && tp->last_stored;
        if (!stored) {
            load_state(ctx, tp->ckpt + (size_t)SW * ps * (s0 / k));
            save_state(ctx, tp->seg);
        }
        for (int m = s0; m < s1 && !stored; m++) {
            const size_t at = (size_t)SW * ps * (m - s0);
            set_dt(ctx, tp->dts[m]);
            stage<double>(ctx, ctx->F64, 1);
            save_pool_u(ctx, tp->sa + at, ctx->F64.U1);
            stage<double>(ctx, ctx->F64, 2);
            save_pool_u(ctx, tp->sb + at, ctx->Ub);
            stage<double>(ctx, ctx->F64, 3);
            swap_u(ctx);
            save_state(ctx, tp->seg + at + (size_t)SW * ps);
        }
        for (int m = s1 - 1; m >= s0; m--) reverse_step(ctx, m - s0, tp->dts[m]);
        HIPCHK(ctx, hipGetLastError());
        // re-running an earlier segment overwrote the segment buffers: a later backward pass on
        // this tape must re-run the last segment as well
        if (!stored) tp->last_stored = false;
    }
    // initial_update: its ghost fill, then k_init on the uploaded state (zero ghost bcell)
    const double *S0 = tp->ckpt;
    reverse_ghost(ctx, A, S0, S0 + 8 * ps);
    HIPCHK(ctx, hipMemcpyAsync(ctx->F64.P, tp->init, 8 * pb, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->F64.EPS, tp->init + 8 * ps, pb, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(ctx->F64.BC, 0, 3 * pb, ctx->stream));
    dim3 g0(nblk((long)L.px * (L.ny + 1), 256), L.nz + 1);
    hipLaunchKernelGGL(kr_init_a, g0, dim3(256), 0, ctx->stream, L, ctx->F64, ctx->M, A, tp->lu[1]);
    hipLaunchKernelGGL(kr_init_b, g0, dim3(256), 0, ctx->stream, L, A);
    HIPCHK(ctx, hipGetLastError());
    if (grad_prim && (r = soa_out(ctx, A.P, 8, grad_prim))) return r;
    if (grad_eps && (r = soa_out(ctx, A.EPS, 1, grad_eps))) return r;
    // leave the context at the taped forward result
    load_state(ctx, tp->ckpt + (size_t)SW * ps * (tp->nck - 1));
    ctx->dt_valid = false;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return CUMC_OK;
}

extern "C" int cumc_tape_free(cumc_ctx *ctx)
{
    if (!ctx) return CUMC_EINVAL;
    tape_release(ctx);
    return CUMC_OK;
}

// ---------------------------------------------------------------------------------------------
// Batched reverse mode: the boxes of a cumc_batch taped and reversed together.  Every reverse
// phase is one launch over a work list of (box, tile) items of all boxes, with the same bodies
// as the single-box kernels (so a batched gradient equals the box's own gradient bit for bit);
// the forward recomputation uses the batch's stage launches with the recorded per-box steps.
// ---------------------------------------------------------------------------------------------
namespace {

struct AdjBox {
    Lay L;
    Fields<double> F;  // forward fields; U0 / U1 as of the last descriptor upload
    Tabs tb;
    Model M;
    double *pool, *Ub, *dt;
    AdjF A;
    double *lu[3];
    double *seg, *sa, *sb, *ckpt, *init;
    int gx_c, gx_e;  // 256-cell tiles per plane row set: [0,n]^2 and [0,n+1]^2 column domains
    int gpen[3];     // 64-pencil tiles per sweep
    int glane[3];    // kr_flux_lane tiles per outer index, per sweep
    // z-collapsed view (L.zc1): the column kernels run plane l0 = 1 only, and the fields of a
    // stored state are the slab of zslab cells at offset zoff of each full-layout plane (0, 0
    // and ps in the full layout)
    int l0;
    long zoff, zslab;
};

// work lists; Q_CP: state copies (a z-collapsed box's plane-1 slabs), Q_CW: whole-state copies
// Q_Z1: the z faces of z-collapsed boxes (kbr_flux_z1)
enum { Q_TAIL, Q_GE, Q_GX, Q_GY, Q_GZ, Q_H1, Q_H2, Q_H3, Q_R1, Q_R2, Q_R3, Q_CP, Q_L1, Q_L2, Q_L3, Q_CW, Q_Z1, Q_N };
constexpr int CP_ELEMS = 1024;  // doubles per copy tile (256 threads x 4)
// copy tiles per plane of n elements (a tile never spans two planes)
__host__ __device__ inline long cp_chunks(long n) { return (n + CP_ELEMS - 1) / CP_ELEMS; }

__device__ inline int adj_box_of(const int *pre, int nbox, int it)
{
    int lo = 0, hi = nbox;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= it) lo = mid;
        else hi = mid;
    }
    return lo;
}

// the stored states around stage `stage` (1..3) of segment step i: forward input (its P, EPS,
// bcell planes first), forward output, U0 of the step and the stage's Uin; stage 0 stands for
// initial_update, whose output is checkpoint 0
struct StageStates {
    const double *in, *out, *U0, *Uin;
};
__device__ inline StageStates stage_states(const AdjBox &d, int stage, int i)
{
    const size_t ps = d.L.ps, st = (size_t)SW * ps;
    const double *Sm = d.seg + i * st, *Sa = d.sa + i * st, *Sb = d.sb + i * st;
    StageStates v;
    const long z = d.zoff;
    v.U0 = Sm + 12 * ps + z;
    if (stage == 3) { v.in = Sb + z; v.out = Sm + st + z; v.Uin = Sb + 12 * ps + z; }
    else if (stage == 2) { v.in = Sa + z; v.out = Sb + z; v.Uin = Sa + 12 * ps + z; }
    else if (stage == 1) { v.in = Sm + z; v.out = Sa + z; v.Uin = nullptr; }
    else { v.in = d.ckpt + z; v.out = d.ckpt + z; v.Uin = nullptr; }
    return v;
}

#define FOR_ITEMS(Q)                                                                            \
    const int *pq = pre + (size_t)(Q) * (nbox + 1);                                            \
    const int total = pq[nbox];                                                                 \
    for (int it = blockIdx.x; it < total; it += gridDim.x) {                                    \
        const int b = adj_box_of(pq, nbox, it);                                                 \
        const int loc = it - pq[b];                                                             \
        const AdjBox &d = D[b];

template <int AX>
__global__ __launch_bounds__(256) void kbr_ghost(const AdjBox *__restrict__ D, const int *__restrict__ pre, int nbox,
                                                 int stage, int i)
{
    FOR_ITEMS(Q_GX + AX)
        const double *out = stage_states(d, stage, i).out;
        kr_ghost_line<AX>(d.L, d.A, d.M, out, out + 8 * d.L.ps, (long)loc * 256 + threadIdx.x);
    }
}

// (the batched reverse takes each box's step from the dt history, hist[m * nbox + box]: the
// boxes' own dt slots belong to the forward re-run of another segment running beside it)
template <int STAGE>
__global__ __launch_bounds__(256) void kbr_tail_a(const AdjBox *__restrict__ D, const int *__restrict__ pre, int nbox,
                                                  int q, int i, const double *__restrict__ hist, int m)
{
    FOR_ITEMS(Q_TAIL)
        const StageStates v = stage_states(d, STAGE, i);
        const double dt = hist[(size_t)m * nbox + b];
        const long ps = d.L.ps;
        double *l0 = d.lu[q], *l1 = d.lu[1 - q], *lb = d.lu[2];
        double *lUin = STAGE == 3 ? lb : (STAGE == 2 ? l1 : nullptr);
        double *lUout = STAGE == 2 ? lb : l1;
        kr_tail_a_cell<STAGE>(d.L, fields_of_state(d.F, v.in, ps), v.out, v.out + 8 * ps, v.out + 9 * ps, d.M, dt,
                              v.U0, v.Uin, d.A, l0, lUin, lUout, (long)(loc % d.gx_c) * 256 + threadIdx.x,
                              loc / d.gx_c + d.l0);
    }
}

template <int STAGE>
__global__ __launch_bounds__(256) void kbr_tail_b(const AdjBox *__restrict__ D, const int *__restrict__ pre, int nbox,
                                                  int q, const double *__restrict__ hist, int m)
{
    FOR_ITEMS(Q_TAIL)
        const double dt = hist[(size_t)m * nbox + b];
        double *l0 = d.lu[q], *l1 = d.lu[1 - q], *lb = d.lu[2];
        double *lUin = STAGE == 3 ? lb : (STAGE == 2 ? l1 : nullptr);
        kr_tail_b_cell<STAGE>(d.L, d.A, dt, l0, lUin, (long)(loc % d.gx_c) * 256 + threadIdx.x, loc / d.gx_c + d.l0);
    }
}

__global__ __launch_bounds__(256) void kbr_gather(const AdjBox *__restrict__ D, const int *__restrict__ pre, int nbox)
{
    FOR_ITEMS(Q_GE)
        kr_gather_cell(d.L, d.A, d.tb, (long)(loc % d.gx_e) * 256 + threadIdx.x, loc / d.gx_e + d.l0);
    }
}

__global__ __launch_bounds__(256) void kbr_ecorn(const AdjBox *__restrict__ D, const int *__restrict__ pre, int nbox,
                                                 int stage, int i)
{
    FOR_ITEMS(Q_GE)
        const Fields<double> F = fields_of_state(d.F, stage_states(d, stage, i).in, d.L.ps);
        kr_ecorn_cell(d.L, F, d.A, (long)(loc % d.gx_e) * 256 + threadIdx.x, loc / d.gx_e + d.l0);
    }
}

template <int DIR, int RECON>
__global__ __launch_bounds__(64) void kbr_flux_hll(const AdjBox *__restrict__ D, const int *__restrict__ pre, int nbox,
                                                   int stage, int i)
{
    FOR_ITEMS(Q_H1 + DIR - 1)
        const Fields<double> F = fields_of_state(d.F, stage_states(d, stage, i).in, d.L.ps);
        kr_flux_hll_pencil<DIR, RECON>(d.L, F, d.tb, d.M, d.A, (long)loc * 64 + threadIdx.x);
    }
}

template <int DIR, int RECON>
__global__ __launch_bounds__(64) void kbr_flux_recon(const AdjBox *__restrict__ D, const int *__restrict__ pre,
                                                     int nbox, int stage, int i)
{
    FOR_ITEMS(Q_R1 + DIR - 1)
        const int gp = d.gpen[DIR - 1];
        const Fields<double> F = fields_of_state(d.F, stage_states(d, stage, i).in, d.L.ps);
        kr_flux_recon_pencil<DIR, RECON>(d.L, F, d.tb, d.A, (long)(loc % gp) * 64 + threadIdx.x, loc / gp);
    }
}

template <int RECON>
__global__ __launch_bounds__(256) void kbr_flux_z1(const AdjBox *__restrict__ D, const int *__restrict__ pre, int nbox,
                                                   int stage, int i)
{
    FOR_ITEMS(Q_Z1)
        const Fields<double> F = fields_of_state(d.F, stage_states(d, stage, i).in, d.L.ps);
        kr_flux_z1_cell<RECON>(d.L, F, d.tb, d.M, d.A, (long)loc * 256 + threadIdx.x);
    }
}

__global__ __launch_bounds__(256) void kbr_init_a(const AdjBox *__restrict__ D, const int *__restrict__ pre, int nbox,
                                                  int q)
{
    FOR_ITEMS(Q_TAIL)
        kr_init_a_cell(d.L, d.F, d.M, d.A, d.lu[1 - q], (long)(loc % d.gx_c) * 256 + threadIdx.x, loc / d.gx_c);
    }
}

__global__ __launch_bounds__(256) void kbr_init_b(const AdjBox *__restrict__ D, const int *__restrict__ pre, int nbox)
{
    FOR_ITEMS(Q_TAIL)
        kr_init_b_cell(d.L, d.A, (long)(loc % d.gx_c) * 256 + threadIdx.x, loc / d.gx_c);
    }
}

template <int DIR, int RECON>
__global__ __launch_bounds__(256) void kbr_flux_lane(const AdjBox *__restrict__ D, const int *__restrict__ pre,
                                                     int nbox, int stage, int i)
{
    if constexpr (DIR == 1) {
        FOR_ITEMS(Q_L1)
            const int gl = d.glane[0];
            const Fields<double> F = fields_of_state(d.F, stage_states(d, stage, i).in, d.L.ps);
            kr_flux_lane_x_tile<RECON>(d.L, F, d.tb, d.M, d.A, loc % gl, loc / gl);
        }
    } else {
        // one workgroup per item in an XCD-aware order: the workgroups of one XCD (blockIdx
        // modulo NXCD) take consecutive items, i.e. neighbouring tiles that share cache lines
        __shared__ double buf[LT_PLANES][64][LT_PITCH];
        const int *pq = pre + (size_t)(Q_L1 + DIR - 1) * (nbox + 1);
        const int total = pq[nbox];
        const int per = (int)(gridDim.x / NXCD);
        const int it = (int)(blockIdx.x % NXCD) * per + (int)(blockIdx.x / NXCD);
        if (it < total) {
            const int b = adj_box_of(pq, nbox, it);
            const int loc = it - pq[b];
            const AdjBox &d = D[b];
            const int gl = d.glane[DIR - 1];
            const Fields<double> F = fields_of_state(d.F, stage_states(d, stage, i).in, d.L.ps);
            kr_flux_lane_t_tile<DIR, RECON>(d.L, F, d.tb, d.M, d.A, loc % gl, loc / gl, buf);
        }
    }
}

// state copies between the context fields and the per-box tape (element e of 20 planes: the
// first 12 are P, EPS, bcell of the pool, the last 8 the conserved state: U0 / U1 by par, or Ub)
enum { CP_SAVE_SEG, CP_SAVE_SA, CP_SAVE_SB, CP_LOAD_CK, CP_SAVE_CK, CP_LOAD_INIT, CP_SAVE_INIT };
__host__ __device__ inline bool copy_whole(int mode, int slot)
{
    return mode == CP_SAVE_INIT || mode == CP_LOAD_INIT || (slot == 0 && (mode == CP_SAVE_CK || mode == CP_LOAD_CK));
}

__global__ __launch_bounds__(256) void kbr_copy(const AdjBox *__restrict__ D, const int *__restrict__ pre, int nbox,
                                                int mode, int slot, int par)
{
    // the uploaded state and checkpoint 0 (read by the full-layout reverse of initial_update) are
    // copied whole; a z-collapsed box's other states are the plane-1 slabs of the fields
    const bool whole = copy_whole(mode, slot);
    FOR_ITEMS(whole ? Q_CW : Q_CP)
        const long ps = d.L.ps;
        const long pe = whole ? ps : d.zslab, zo = whole ? 0 : d.zoff;
        // tile loc: plane f of the 20, chunk ch of that plane's pe elements
        const int nch = (int)cp_chunks(pe);
        const int f = loc / nch, ch = loc - f * nch;
        // conserved buffers as full-layout bases (the descriptor's are the view's)
        double *U = (mode == CP_SAVE_SB ? d.Ub : (par ? d.F.U1 : d.F.U0)) - d.zoff;
        double *pool = d.pool;
        const bool lo = f < 12;
        const size_t at = (size_t)slot * SW * ps;
        const long e0 = (long)f * ps + zo;  // first element of the plane's range in the 20 full-layout planes
        double *ctx0 = lo ? pool + e0 : U + (e0 - 12 * ps);
        for (int r = 0; r < CP_ELEMS / 256; r++) {
            const long w = (long)ch * CP_ELEMS + r * 256 + threadIdx.x;
            if (w >= pe) break;
            const long e = e0 + w;
            double *ctxp = ctx0 + w;
            switch (mode) {
            case CP_SAVE_SEG: d.seg[at + e] = *ctxp; break;
            case CP_SAVE_SA: d.sa[at + e] = *ctxp; break;
            case CP_SAVE_SB: d.sb[at + e] = *ctxp; break;
            case CP_LOAD_CK: *ctxp = d.ckpt[at + e]; break;
            case CP_SAVE_CK: d.ckpt[at + e] = *ctxp; break;
            case CP_LOAD_INIT:
                if (e < 9 * ps) pool[e] = d.init[e];
                else if (lo) pool[e] = 0.0;
                break;
            case CP_SAVE_INIT: if (e < 9 * ps) d.init[e] = pool[e]; break;
            }
        }
    }
}

#undef FOR_ITEMS

// The reverse of the last ghost fill's z pass on the cell-field seeds of a z-collapsed box
// (full layout): the seed on each z ghost plane is added to plane 1, which that pass copied there,
// in the order of kr_ghost_line<2> (side 0, then 1; ghost layers 1..3), and cleared.
__global__ __launch_bounds__(256) void kbr_zfold_bc(Lay L, double *BC)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long sx = L.nx + 6;
    if (t >= sx * (L.ny + 6)) return;
    const int j = (int)(t % sx) - 2, k = (int)(t / sx) - 2;
    const long ps = L.ps, src = L.id(j, k, 1);
    for (int c = 0; c < 3; c++) {
        double acc = BC[src + c * ps];
        for (int side = 0; side < 2; side++)
            for (int g = 1; g <= 3; g++) {
                const long dst = L.id(j, k, side == 0 ? 1 - g : 1 + g) + c * ps;
                acc += BC[dst];
                BC[dst] = 0.0;
            }
        BC[src + c * ps] = acc;
    }
}

}  // namespace

struct batch_tape {
    int nsteps = 0, every = 1, nck = 0, recon = 4;
    std::vector<AdjBox> h;
    AdjBox *d = nullptr;
    int *pre = nullptr;
    int total[Q_N] = {0};
    // the same descriptors in the full layout for every box (the reverse of initial_update)
    AdjBox *d_full = nullptr;
    int *pre_full = nullptr;
    int total_full[Q_N] = {0};
    std::vector<double> dts;  // nsteps x nbox
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // per checkpoint segment of the backward: start and end of its forward re-run (on s2) and end
    // of its reverse (on the batch stream)
    std::vector<hipEvent_t> seg_ev;
    double recompute_ms = 0.0;       // the re-runs' total in the last backward pass
    hipStream_t s2 = nullptr;        // the re-runs' stream, default priority like the batch stream
    // the checkpoint segment (g, counted from the last in time) whose stage states each set of
    // segment buffers holds, -1 for none: tape_run's own forward pass stores the last segment in
    // set 0 and, with two sets, the one before it in set 1, so the backward re-runs every segment
    // but those; with every >= nsteps nothing is re-run
    int held[2] = {-1, -1};
};

void batch_tape_release(cumc_batch *b)
{
    batch_tape *&bt = batch_adj(b);
    if (!bt) return;
    hipFree(bt->d);
    hipFree(bt->pre);
    hipFree(bt->d_full);
    hipFree(bt->pre_full);
    if (bt->e0) hipEventDestroy(bt->e0);
    if (bt->e1) hipEventDestroy(bt->e1);
    for (hipEvent_t e : bt->seg_ev) hipEventDestroy(e);
    if (bt->s2) {
        hipStreamSynchronize(bt->s2);
        hipStreamDestroy(bt->s2);
    }
    delete bt;
    bt = nullptr;
}

// adjoint arrays addressed like make_view's fields (bases moved by off cells)
static AdjF adj_view(AdjF A, long off)
{
    A.P += off; A.BC += off; A.EPS += off;
    for (int d = 0; d < 3; d++) A.F[d] += off;
    A.E += off; A.T += off; A.H += off;
    return A;
}

// descriptors from the boxes' current buffers: a box whose state is z-collapsed (zmode) is
// reversed in that view unless `full` (the reverse of initial_update runs in the full layout)
static int bt_upload(cumc_batch *b, batch_tape *bt, bool full = false)
{
    const int nb = batch_nbox(b);
    hipStream_t s = batch_stream(b);
    std::vector<int> pre((size_t)Q_N * (nb + 1), 0);
    for (int i = 0; i < nb; i++) {
        cumc_ctx *c = batch_box(b, i);
        cumc_tape *tp = c->tape;
        AdjBox &a = bt->h[i];
        const bool col = c->zmode && !full;
        const View v = make_view(c, col);
        const long off = col ? 3 * c->L.py : 0;
        const Lay &L = v.L;
        a.L = L;
        a.F = v.F;
        a.tb = tabs_of(c);
        a.M = c->M;
        a.pool = c->pool;
        a.Ub = v.Ub;
        a.dt = c->dt_dev;
        a.A = adj_view(tp->A, off);
        for (int r = 0; r < 3; r++) a.lu[r] = tp->lu[r] + off;
        a.seg = tp->seg; a.sa = tp->sa; a.sb = tp->sb; a.ckpt = tp->ckpt; a.init = tp->init;
        a.l0 = col ? 1 : 0;
        a.zoff = off;
        a.zslab = col ? c->L.py : L.ps;
        a.gx_c = (int)nblk(L.px * (L.ny + 1), 256);
        a.gx_e = (int)nblk(L.px * (L.ny + 2), 256);
        a.gpen[0] = (int)nblk(npencils<1>(L), 64);
        a.gpen[1] = (int)nblk(npencils<2>(L), 64);
        a.gpen[2] = (int)nblk(npencils<3>(L), 64);
        long cnt[Q_N];
        const long nzc = col ? 1 : L.nz + 1;  // planes of the column kernels
        cnt[Q_TAIL] = (long)a.gx_c * nzc;
        cnt[Q_GE] = (long)a.gx_e * (col ? 1 : L.nz + 2);
        cnt[Q_GX] = nblk(6 * kr_ghost_lines<0>(L), 256);
        cnt[Q_GY] = nblk(6 * kr_ghost_lines<1>(L), 256);
        cnt[Q_GZ] = col ? 0 : nblk(6 * kr_ghost_lines<2>(L), 256);
        const bool lane[3] = {lane_sweep<1>(L), lane_sweep<2>(L), lane_sweep<3>(L)};
        a.glane[0] = lane_tiles<1>(L);
        a.glane[1] = lane_tiles<2>(L);
        a.glane[2] = lane_tiles<3>(L);
        const int outer[3] = {lane_outer<1>(L), lane_outer<2>(L), lane_outer<3>(L)};
        for (int dd = 0; dd < 3; dd++) {
            cnt[Q_H1 + dd] = lane[dd] ? 0 : a.gpen[dd];
            cnt[Q_R1 + dd] = lane[dd] ? 0 : 7L * a.gpen[dd];
            cnt[Q_L1 + dd] = lane[dd] ? (long)a.glane[dd] * outer[dd] : 0;
        }
        // a collapsed box's z faces: one column kernel instead of the pencil walks
        cnt[Q_Z1] = col ? nblk((long)(L.nx + 2) * (L.ny + 2), 256) : 0;
        if (col) cnt[Q_H3] = cnt[Q_R3] = 0;
        cnt[Q_CP] = SW * cp_chunks(a.zslab);
        cnt[Q_CW] = SW * cp_chunks(L.ps);
        for (int q = 0; q < Q_N; q++) {
            const long v = (long)pre[(size_t)q * (nb + 1) + i] + cnt[q];
            if (v > 0x7fffffffL) return fail(c, CUMC_EINVAL, "batched reverse work list exceeds 2^31 tiles");
            pre[(size_t)q * (nb + 1) + i + 1] = (int)v;
        }
    }
    int *total = full ? bt->total_full : bt->total;
    for (int q = 0; q < Q_N; q++) total[q] = pre[(size_t)q * (nb + 1) + nb];
    if (hipMemcpyAsync(full ? bt->d_full : bt->d, bt->h.data(), sizeof(AdjBox) * nb, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(full ? bt->pre_full : bt->pre, pre.data(), sizeof(int) * pre.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        batch_err(b) = "batched reverse descriptors upload failed";
        return CUMC_EHIP;
    }
    return CUMC_OK;
}

static inline dim3 bt_grid(const batch_tape *bt, int q, bool full = false)
{
    const int t = full ? bt->total_full[q] : bt->total[q];
    return dim3(t > 0 ? t : 1);
}

static void bt_copy(cumc_batch *b, batch_tape *bt, int mode, int slot, int par, bool full = false)
{
    const int nb = batch_nbox(b);
    hipLaunchKernelGGL(kbr_copy, bt_grid(bt, copy_whole(mode, slot) ? Q_CW : Q_CP, full), dim3(256), 0, batch_stream(b),
                       full ? bt->d_full : bt->d, full ? bt->pre_full : bt->pre, nb, mode, slot, par);
}

template <int DIR>
static void bt_sweep(cumc_batch *b, batch_tape *bt, int stage, int i)
{
    const int nb = batch_nbox(b);
    hipStream_t s = batch_stream(b);
    // boxes whose pencils along DIR are long enough take the wave-parallel kernel, the others
    // the per-pencil walks (each work list is empty when no box takes it)
    if (bt->total[Q_L1 + DIR - 1] > 0) {
        // DIR 2 / 3: one workgroup per item, a multiple of NXCD of them (the kernel's XCD order)
        const dim3 g = DIR == 1 ? bt_grid(bt, Q_L1) : dim3(xcd_blocks(bt->total[Q_L1 + DIR - 1], 1));
        if (bt->recon == CUMC_PPMC)
            hipLaunchKernelGGL((kbr_flux_lane<DIR, 4>), g, dim3(256), 0, s, bt->d, bt->pre, nb, stage, i);
        else
            hipLaunchKernelGGL((kbr_flux_lane<DIR, 1>), g, dim3(256), 0, s, bt->d, bt->pre, nb, stage, i);
    }
    if constexpr (DIR == 3) {
        if (bt->total[Q_Z1] > 0) {
            if (bt->recon == CUMC_PPMC)
                hipLaunchKernelGGL((kbr_flux_z1<4>), bt_grid(bt, Q_Z1), dim3(256), 0, s, bt->d, bt->pre, nb, stage, i);
            else
                hipLaunchKernelGGL((kbr_flux_z1<1>), bt_grid(bt, Q_Z1), dim3(256), 0, s, bt->d, bt->pre, nb, stage, i);
        }
    }
    if constexpr (DIR != 1) {
        if (bt->total[Q_H1 + DIR - 1] > 0) {
            const dim3 gh = bt_grid(bt, Q_H1 + DIR - 1), gr = bt_grid(bt, Q_R1 + DIR - 1);
            if (bt->recon == CUMC_PPMC) {
                hipLaunchKernelGGL((kbr_flux_hll<DIR, 4>), gh, dim3(64), 0, s, bt->d, bt->pre, nb, stage, i);
                hipLaunchKernelGGL((kbr_flux_recon<DIR, 4>), gr, dim3(64), 0, s, bt->d, bt->pre, nb, stage, i);
            } else {
                hipLaunchKernelGGL((kbr_flux_hll<DIR, 1>), gh, dim3(64), 0, s, bt->d, bt->pre, nb, stage, i);
                hipLaunchKernelGGL((kbr_flux_recon<DIR, 1>), gr, dim3(64), 0, s, bt->d, bt->pre, nb, stage, i);
            }
        }
    }
}

static void bt_ghost(cumc_batch *b, batch_tape *bt, int stage, int i, bool full = false)
{
    const int nb = batch_nbox(b);
    hipStream_t s = batch_stream(b);
    const AdjBox *D = full ? bt->d_full : bt->d;
    const int *pre = full ? bt->pre_full : bt->pre;
    // (no z pass when every box is z-collapsed)
    if ((full ? bt->total_full : bt->total)[Q_GZ] > 0)
        hipLaunchKernelGGL(kbr_ghost<2>, bt_grid(bt, Q_GZ, full), dim3(256), 0, s, D, pre, nb, stage, i);
    hipLaunchKernelGGL(kbr_ghost<1>, bt_grid(bt, Q_GY, full), dim3(256), 0, s, D, pre, nb, stage, i);
    hipLaunchKernelGGL(kbr_ghost<0>, bt_grid(bt, Q_GX, full), dim3(256), 0, s, D, pre, nb, stage, i);
}

// stage STAGE of segment slot i (set offset included) of step m reversed
template <int STAGE>
static void bt_reverse_stage(cumc_batch *b, batch_tape *bt, int i, int q, int m)
{
    const int nb = batch_nbox(b);
    hipStream_t s = batch_stream(b);
    const double *hist = batch_hist(b);
    bt_ghost(b, bt, STAGE, i);
    hipLaunchKernelGGL((kbr_tail_a<STAGE>), bt_grid(bt, Q_TAIL), dim3(256), 0, s, bt->d, bt->pre, nb, q, i, hist, m);
    hipLaunchKernelGGL((kbr_tail_b<STAGE>), bt_grid(bt, Q_TAIL), dim3(256), 0, s, bt->d, bt->pre, nb, q, hist, m);
    hipLaunchKernelGGL(kbr_gather, bt_grid(bt, Q_GE), dim3(256), 0, s, bt->d, bt->pre, nb);
    hipLaunchKernelGGL(kbr_ecorn, bt_grid(bt, Q_GE), dim3(256), 0, s, bt->d, bt->pre, nb, STAGE, i);
    bt_sweep<3>(b, bt, STAGE, i);
    bt_sweep<2>(b, bt, STAGE, i);
    bt_sweep<1>(b, bt, STAGE, i);
}

#define BTCHK(b, call)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            batch_err(b) = std::string(#call) + ": " + hipGetErrorString(e_);                  \
            return CUMC_EHIP;                                                                   \
        }                                                                                       \
    } while (0)

extern "C" int cumc_batch_tape_run(cumc_batch *b, int nsteps, int every, const double *dt_in, double *dt_out)
{
    if (!b || nsteps < 0 || every < 1) return CUMC_EINVAL;
    const int nb = batch_nbox(b);
    for (int i = 0; i < nb; i++) {
        std::string why;
        if (!adjoint_supported(batch_box(b, i), why)) {
            batch_err(b) = "box " + std::to_string(i) + ": " + why;
            return CUMC_EINVAL;
        }
    }
    if (dt_in)
        for (long i = 0; i < (long)nsteps * nb; i++)
            if (!std::isfinite(dt_in[i]) || !(dt_in[i] > 0.0)) {
                batch_err(b) = "frozen dt must be finite and positive";
                return CUMC_ENONFINITE;
            }
    for (int i = 0; i < nb; i++) BTCHK(b, hipStreamSynchronize(batch_box(b, i)->stream));
    batch_tape_release(b);
    // two sets of segment buffers when the backward re-runs segments (it re-runs one segment
    // while it reverses the next)
    const int sets = nsteps > every ? 2 : 1;
    for (int i = 0; i < nb; i++) {
        int r = tape_alloc(batch_box(b, i), nsteps, every, true, sets);
        if (r) {
            batch_err(b) = "box " + std::to_string(i) + ": " + batch_box(b, i)->err;
            return r;
        }
    }
    batch_tape *bt = new batch_tape();
    batch_adj(b) = bt;
    bt->nsteps = nsteps;
    bt->every = every;
    bt->nck = nsteps / every + 2;
    bt->recon = batch_box(b, 0)->cfg.recon;
    bt->h.assign(nb, AdjBox{});
    BTCHK(b, hipMalloc(&bt->d, sizeof(AdjBox) * nb));
    BTCHK(b, hipMalloc(&bt->pre, sizeof(int) * Q_N * (nb + 1)));
    BTCHK(b, hipMalloc(&bt->d_full, sizeof(AdjBox) * nb));
    BTCHK(b, hipMalloc(&bt->pre_full, sizeof(int) * Q_N * (nb + 1)));
    BTCHK(b, hipEventCreate(&bt->e0));
    BTCHK(b, hipEventCreate(&bt->e1));
    BTCHK(b, hipStreamCreateWithFlags(&bt->s2, hipStreamNonBlocking));
    int r = batch_refresh(b);
    if (r) return r;
    if ((r = bt_upload(b, bt))) return r;
    hipStream_t s = batch_stream(b);
    bt_copy(b, bt, CP_SAVE_INIT, 0, 0);
    BTCHK(b, hipStreamSynchronize(s));
    for (int i = 0; i < nb; i++)
        if ((r = cumc_initial_update(batch_box(b, i)))) {
            batch_err(b) = "box " + std::to_string(i) + ": " + batch_box(b, i)->err;
            return r;
        }
    bt_copy(b, bt, CP_SAVE_CK, 0, 0);
    // the steps run in each box's z-collapsed view where it may (initial_update left the state
    // in the full layout with fresh z ghost planes)
    BTCHK(b, hipStreamSynchronize(s));
    if ((r = batch_refresh(b)) || (r = bt_upload(b, bt))) return r;
    for (int i = 0; i < nb; i++) batch_box(b, i)->tape->zcol = batch_box(b, i)->zmode;
    if ((r = batch_reserve_hist(b, nsteps > 0 ? nsteps : 1))) return r;
    if (dt_in) BTCHK(b, hipMemcpyAsync(batch_hist(b), dt_in, sizeof(double) * (size_t)nsteps * nb, hipMemcpyHostToDevice, s));
    batch_set_counter(b, 0);
    if (!dt_in) batch_finddt(b);
    // steps of the last segment (and, with two sets, of the one before it) store their stage
    // states as the backward's re-run would (bit-identical: the same kernels on the same state
    // and dt), into the set the backward reads them from
    const int s_last = nsteps > 0 ? ((nsteps - 1) / every) * every : 0;
    const int s_prev = sets > 1 && s_last >= every ? s_last - every : s_last;
    const int stride = batch_box(b, 0)->tape->set_stride;
    for (int m = 0; m < nsteps; m++) {
        if (m < s_prev) {
            batch_step(b, m & 1, dt_in != nullptr);
        } else {
            const bool last = m >= s_last;
            const int par = m & 1, i = m - (last ? s_last : s_prev), off = last ? 0 : stride;
            if (i == 0) bt_copy(b, bt, CP_SAVE_SEG, off, par);
            batch_dt_begin(b, dt_in != nullptr);
            batch_stage(b, 1, par);
            bt_copy(b, bt, CP_SAVE_SA, off + i, 1 - par);
            batch_stage(b, 2, par);
            bt_copy(b, bt, CP_SAVE_SB, off + i, 0);
            batch_stage(b, 3, par);
            bt_copy(b, bt, CP_SAVE_SEG, off + i + 1, 1 - par);
        }
        if ((m + 1) % every == 0) bt_copy(b, bt, CP_SAVE_CK, (m + 1) / every, (m + 1) & 1);
    }
    bt->held[0] = nsteps > 0 ? 0 : -1;
    bt->held[1] = s_prev < s_last ? 1 : -1;
    bt_copy(b, bt, CP_SAVE_CK, bt->nck - 1, nsteps & 1);
    BTCHK(b, hipGetLastError());
    BTCHK(b, hipStreamSynchronize(s));
    bt->dts.assign((size_t)nsteps * nb, 0.0);
    if (nsteps > 0)
        BTCHK(b, hipMemcpy(bt->dts.data(), batch_hist(b), sizeof(double) * bt->dts.size(), hipMemcpyDeviceToHost));
    batch_after_run(b, nsteps);
    if ((r = batch_refresh(b)) || (r = bt_upload(b, bt))) return r;
    if (dt_out && nsteps > 0) std::memcpy(dt_out, bt->dts.data(), sizeof(double) * bt->dts.size());
    if (!dt_in)
        for (size_t i = 0; i < bt->dts.size(); i++)
            if (!std::isfinite(bt->dts[i]) || !(bt->dts[i] > 0.0) || bt->dts[i] >= 1.0e10) {
                batch_err(b) = "box " + std::to_string(i % nb) + ": finddt produced a non-finite step";
                return CUMC_ENONFINITE;
            }
    return CUMC_OK;
}

extern "C" int cumc_batch_tape_backward(cumc_batch *b, const double *const *seed_bcell, double *const *grad_prim)
{
    if (!b) return CUMC_EINVAL;
    batch_tape *bt = batch_adj(b);
    if (!bt) {
        batch_err(b) = "cumc_batch_tape_run must precede cumc_batch_tape_backward";
        return CUMC_ESTATE;
    }
    const int nb = batch_nbox(b);
    hipStream_t s = batch_stream(b);
    // each box is reversed (and its segments re-run) in the view its tape was recorded in,
    // whatever happened to its live state since (uploads, downloads, the collapse switch)
    for (int i = 0; i < nb; i++) {
        cumc_ctx *c = batch_box(b, i);
        if (c->tape->zcol) c->zmode = true;  // the live state is overwritten before it is read
        else zexpand(c);
    }
    batch_freeze_views(b, true);
    int r = batch_refresh(b);
    batch_freeze_views(b, false);
    if (r || (r = bt_upload(b, bt, true)) || (r = bt_upload(b, bt))) return r;
    for (int i = 0; i < nb; i++) {
        cumc_ctx *c = batch_box(b, i);
        cumc_tape *tp = c->tape;
        const size_t pb = plane_bytes(c);
        BTCHK(b, hipMemsetAsync(tp->adj, 0, (size_t)NADJ * pb, c->stream));
        for (int k = 0; k < 3; k++) BTCHK(b, hipMemsetAsync(tp->lu[k], 0, 8 * pb, c->stream));
        if (seed_bcell && seed_bcell[i] && (r = aos_into(c, seed_bcell[i], 3, tp->A.BC))) {
            batch_err(b) = "box " + std::to_string(i) + ": " + c->err;
            return r;
        }
        // a z-collapsed box reverses its steps on plane 1: the seeds on the z ghost planes (the
        // copies the last z ghost pass made) go to plane 1 first
        if (c->zmode)
            hipLaunchKernelGGL(kbr_zfold_bc, dim3(nblk((long)(c->L.nx + 6) * (c->L.ny + 6), 256)), dim3(256), 0,
                               c->stream, c->L, tp->A.BC);
        BTCHK(b, hipGetLastError());
        BTCHK(b, hipStreamSynchronize(c->stream));
    }
    const int N = bt->nsteps, k = bt->every;
    if ((r = batch_reserve_hist(b, N > 0 ? N : 1))) return r;
    if (N > 0)
        BTCHK(b, hipMemcpyAsync(batch_hist(b), bt->dts.data(), sizeof(double) * bt->dts.size(), hipMemcpyHostToDevice, s));
    // Checkpoint segments g = 0, 1, ... from the last in time back.  The forward re-run of a
    // segment stores every stage's output state, so its reverse runs no forward kernel; the last
    // two segments' states are still in sets 0 and 1 from tape_run (held[]), unless an earlier
    // backward re-ran other segments into them.  Segment g uses buffer set g mod sets:
    // the re-run of segment g + 1 (on s2) runs while segment g is reversed (on the batch stream)
    // and waits for the reverse of segment g - 1, the last reader of its set; the reverse of a
    // re-run segment waits for its re-run.  The re-runs own the boxes' fields and dt slots; the
    // reverse reads only stored states, the adjoint arrays and the dt history.
    const int nseg = N > 0 ? (N - 1) / k + 1 : 0;
    const int sets = batch_box(b, 0)->tape->sets, stride = batch_box(b, 0)->tape->set_stride;
    while ((int)bt->seg_ev.size() < 3 * nseg) {
        hipEvent_t e;
        BTCHK(b, hipEventCreate(&e));
        bt->seg_ev.push_back(e);
    }
    std::vector<char> reran(nseg, 0);
    hipStream_t s2 = bt->s2;
    auto seg_start = [&](int g) { return ((N - 1) / k - g) * k; };
    auto rerun = [&](int g) {
        const int s0 = seg_start(g), s1 = s0 + k < N ? s0 + k : N, off = (g % sets) * stride;
        hipStream_t s_b = batch_swap_stream(b, s2);
        hipEventRecord(bt->seg_ev[3 * g], s2);
        bt_copy(b, bt, CP_LOAD_CK, s0 / k, 0);
        bt_copy(b, bt, CP_SAVE_SEG, off, 0);
        for (int m = s0; m < s1; m++) {
            const int i = m - s0, par = i & 1;
            batch_set_dt(b, m);
            batch_stage(b, 1, par);
            bt_copy(b, bt, CP_SAVE_SA, off + i, 1 - par);  // stage-1 output U: U1 for even steps, U0 for odd
            batch_stage(b, 2, par);
            bt_copy(b, bt, CP_SAVE_SB, off + i, 0);
            batch_stage(b, 3, par);
            bt_copy(b, bt, CP_SAVE_SEG, off + i + 1, 1 - par);
        }
        hipEventRecord(bt->seg_ev[3 * g + 1], s2);
        batch_swap_stream(b, s_b);
        reran[g] = 1;
        bt->held[g % sets] = g;
    };
    auto stored = [&](int g) { return bt->held[g % sets] == g; };
    BTCHK(b, hipEventRecord(bt->e0, s));
    BTCHK(b, hipStreamWaitEvent(s2, bt->e0, 0));
    if (nseg > 0 && !stored(0)) rerun(0);
    int q = 0;
    for (int g = 0; g < nseg; g++) {
        if (g + 1 < nseg && !stored(g + 1)) {
            if (g >= 1) BTCHK(b, hipStreamWaitEvent(s2, bt->seg_ev[3 * (g - 1) + 2], 0));
            rerun(g + 1);
        }
        if (reran[g]) BTCHK(b, hipStreamWaitEvent(s, bt->seg_ev[3 * g + 1], 0));
        const int s0 = seg_start(g), s1 = s0 + k < N ? s0 + k : N, off = (g % sets) * stride;
        for (int m = s1 - 1; m >= s0; m--) {
            const int i = off + m - s0;
            bt_reverse_stage<3>(b, bt, i, q, m);
            bt_reverse_stage<2>(b, bt, i, q, m);
            bt_reverse_stage<1>(b, bt, i, q, m);
            q ^= 1;
        }
        BTCHK(b, hipEventRecord(bt->seg_ev[3 * g + 2], s));
        BTCHK(b, hipGetLastError());
    }
    // initial_update, in the full layout (z-collapsed boxes' adjoints are on plane 1, their z
    // ghost planes' zero): its ghost fill, then k_init on the uploaded state
    bt_ghost(b, bt, 0, 0, true);
    bt_copy(b, bt, CP_LOAD_INIT, 0, 0, true);
    hipLaunchKernelGGL(kbr_init_a, bt_grid(bt, Q_TAIL, true), dim3(256), 0, s, bt->d_full, bt->pre_full, nb, q);
    hipLaunchKernelGGL(kbr_init_b, bt_grid(bt, Q_TAIL, true), dim3(256), 0, s, bt->d_full, bt->pre_full, nb);
    BTCHK(b, hipEventRecord(bt->e1, s));
    BTCHK(b, hipGetLastError());
    BTCHK(b, hipStreamSynchronize(s));
    float ms = 0.f;
    hipEventElapsedTime(&ms, bt->e0, bt->e1);
    batch_last_ms(b) = ms;
    bt->recompute_ms = 0.0;
    for (int g = 0; g < nseg; g++) {
        float m = 0.f;
        if (reran[g] && hipEventElapsedTime(&m, bt->seg_ev[3 * g], bt->seg_ev[3 * g + 1]) == hipSuccess)
            bt->recompute_ms += m;
    }
    for (int i = 0; i < nb; i++)
        if (grad_prim && grad_prim[i] && (r = soa_out(batch_box(b, i), batch_box(b, i)->tape->A.P, 8, grad_prim[i]))) {
            batch_err(b) = "box " + std::to_string(i) + ": " + batch_box(b, i)->err;
            return r;
        }
    // leave every box at the taped final state (a collapsed box's in its view; one whose
    // collapse was switched off since returns to the full layout)
    bt_copy(b, bt, CP_LOAD_CK, bt->nck - 1, 0);
    BTCHK(b, hipStreamSynchronize(s));
    for (int i = 0; i < nb; i++) {
        cumc_ctx *c = batch_box(b, i);
        c->dt_valid = false;
        if (!c->zc_on) zexpand(c);
        BTCHK(b, hipStreamSynchronize(c->stream));
    }
    return CUMC_OK;
}

extern "C" int cumc_batch_tape_recompute_ms(cumc_batch *b, double *ms)
{
    if (!b || !ms) return CUMC_EINVAL;
    batch_tape *bt = batch_adj(b);
    if (!bt) return CUMC_ESTATE;
    *ms = bt->recompute_ms;
    return CUMC_OK;
}

extern "C" int cumc_batch_tape_free(cumc_batch *b)
{
    if (!b) return CUMC_EINVAL;
    batch_tape_release(b);
    for (int i = 0; i < batch_nbox(b); i++) tape_release(batch_box(b, i));
    return CUMC_OK;
}
