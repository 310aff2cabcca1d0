// physics.hpp -- point-wise MHD numerics of the CUMC3D step, shared by every kernel.
//
// Each function restates one reference routine (file:line given) with the same
// operation order, so that with -ffp-contract=off and IEEE double division and
// square root the FP64 results are bit-identical to the reference arithmetic.
// Templated on the arithmetic type R (double for the drop-in path, float for the
// mixed-precision reconstruction / Riemann / EMF stages).
#pragma once
#include <hip/hip_runtime.h>

namespace cumc {

enum { IRHO = 0, IVX = 1, IVY = 2, IVZ = 3, ITAU = 4, IBX = 5, IBY = 6, IBZ = 7, NV = 8 };

template <typename R> __host__ __device__ inline R rsqrt_(R x);
template <> __host__ __device__ inline double rsqrt_(double x) { return sqrt(x); }
template <> __host__ __device__ inline float rsqrt_(float x) { return sqrtf(x); }

// Fortran MIN/MAX/SIGN on non-NaN operands (identical definition in the CPU restatement)
template <typename R> __host__ __device__ inline R fmin2(R a, R b) { return (b < a) ? b : a; }
template <typename R> __host__ __device__ inline R fmax2(R a, R b) { return (b > a) ? b : a; }

// Branch signature, compiled only into the diagnostic build (-DCUMC_BRANCH_SIG): each
// data-dependent decision of the reference path (the PPM / TVD-minmod limiter switches, the HLL
// speed selections and region, the signal-speed discriminant clamp, the floors) is folded into a
// per-thread hash kept in LDS, and the batched forward kernels add every thread's hash into a
// per-box sum after each tile.  Two runs with equal sums took the same branch everywhere (up to
// 64-bit hash collisions), so a central difference between them crosses no kink of the map.
// In every other build br() returns its argument and the LDS slot does not exist.
#ifdef CUMC_BRANCH_SIG
constexpr unsigned long long BR_SEED = 0x243F6A8885A308D3ull;
static __shared__ unsigned long long br_slot[256];
#endif
__host__ __device__ inline bool br(bool c)
{
#if defined(CUMC_BRANCH_SIG) && defined(__HIP_DEVICE_COMPILE__)
    unsigned long long &h = br_slot[threadIdx.x];
    h = (h ^ (c ? 0x9E3779B97F4A7C15ull : 0x632BE59BD9B4E019ull)) * 0x100000001B3ull;
#endif
    return c;
}
// MIN / MAX whose selection is a recorded decision.  Operands equal to within a few hundred
// units in the last place count as a tie, recorded as such: a selection between them can follow
// the rounding of the run (as between the two sides of a face whose states are copies), and
// changes the result by less than the band, so it is no kink.
__host__ __device__ inline void br_sel(double a, double b)
{
#if defined(CUMC_BRANCH_SIG) && defined(__HIP_DEVICE_COMPILE__)
    const double w = 4.0e-14 * (fabs(a) + fabs(b));
    if (br(fabs(a - b) > w)) br(b < a);
#endif
}
template <typename R> __host__ __device__ inline R fmin2b(R a, R b)
{
    br_sel((double)a, (double)b);
    return (b < a) ? b : a;
}
template <typename R> __host__ __device__ inline R fmax2b(R a, R b)
{
    br_sel((double)a, (double)b);
    return (b > a) ? b : a;
}
template <typename R> __host__ __device__ inline R fabs_(R a);
template <> __host__ __device__ inline double fabs_(double a) { return fabs(a); }
template <> __host__ __device__ inline float fabs_(float a) { return fabsf(a); }
template <typename R> __host__ __device__ inline R sign1(R x);
template <> __host__ __device__ inline double sign1(double x) { return copysign(1.0, x); }
template <> __host__ __device__ inline float sign1(float x) { return copysignf(1.0f, x); }

// definition.f90:380-385
template <typename R>
__host__ __device__ inline R fluxhll(R yl, R yr, R xl, R xr, R sl, R sr)
{
    return (sr * yl - sl * yr + sl * sr * (xr - xl)) / (sr - sl);
}

// definition.f90:394-410 ; discriminant clamped at zero
template <typename R>
__host__ __device__ inline R signalspeed(R cs, R bn, R bt1, R bt2, R rho)
{
    R a2 = cs * cs;
    R a4 = a2 * a2;
    R b2n = (bn * bn / rho);
    R b2t1 = (bt1 * bt1 / rho);
    R b2t2 = (bt2 * bt2 / rho);
    R b2 = b2n + b2t1 + b2t2;
    R b4 = b2 * b2;
    return rsqrt_<R>(R(0.5) * (a2 + b2 + rsqrt_<R>(fmax2b<R>((a4 + b4 + R(2.0) * b2 * a2) - R(4.0) * a2 * b2n, R(0)))));
}

// PPM interpolation coefficients of one cell (reconstruct.f90:211-249), hoisted out of the
// per-cell call: they depend only on the five widths along the sweep axis.
struct PPMCoef {
    double a0m, a1m, a2m, a3m, a0p, a1p, a2p, a3p, b0c, b1c;
};

inline PPMCoef ppm_coef(double dm2, double dm1, double dc, double dp1, double dp2)
{
    PPMCoef k;
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
    double cc2 = (2.0 * dm1 + dc) / (dp1 + dc);
    double cc3 = (dc + 2.0 * dp1) / (dm1 + dc);
    k.a0m = a1L;
    k.a1m = aL * (z1L - zL) / deltaXL;
    k.a2m = (-dm1 * z1L) / deltaXL;
    k.a3m = dc * zL / deltaXL;
    k.a0p = a1R;
    k.a1p = aR * (z1R - zR) / deltaXR;
    k.a2p = (-dc * z1R) / deltaXR;
    k.a3p = dp1 * zR / deltaXR;
    k.b0c = cc1 * cc2;
    k.b1c = cc1 * cc3;
    return k;
}

// Slope-only coefficients of a cell (needs the three widths around it): the b0c/b1c of cell j
// equal the b0p/b1p of cell j-1 and the b0m/b1m of cell j+1 bit for bit, so one limited slope
// per cell serves the three PPM calls that use it.
inline void ppm_slope_coef(double dm1, double dc, double dp1, double *b0c, double *b1c)
{
    double cc1 = dc / (dm1 + dc + dp1);
    double cc2 = (2.0 * dm1 + dc) / (dp1 + dc);
    double cc3 = (dc + 2.0 * dp1) / (dm1 + dc);
    *b0c = cc1 * cc2;
    *b1c = cc1 * cc3;
}

// reconstruct.f90:254-280: van Leer limited slope of cell c from its neighbours.
template <typename R>
__host__ __device__ inline R ppm_slope(R b0, R b1, R vm1, R vc, R vp1)
{
    R delta = b0 * (vp1 - vc) + b1 * (vc - vm1);
    if (br((vp1 - vc) * (vc - vm1) > R(0))) {
        R a = fabs_<R>(delta), b = R(2.0) * fabs_<R>(vc - vm1), c = R(2.0) * fabs_<R>(vc - vp1);
        br(delta < R(0));
        return fmin2b<R>(fmin2b<R>(a, b), c) * sign1<R>(delta);
    }
    return R(0);
}

// reconstruct.f90:285-307: PPM edge values of cell c from the limited slopes dmm1, dmc, dmp1.
template <typename R>
__host__ __device__ inline void ppm_edges(R a0m, R a1m, R a2m, R a3m, R a0p, R a1p, R a2p, R a3p, R vm1, R vc,
                                          R vp1, R dmm1, R dmc, R dmp1, R &vm_out, R &vp_out)
{
    R vp = vc + a0p * (vp1 - vc) + a1p * (vp1 - vc) + a2p * dmp1 + a3p * dmc;
    R vm = vm1 + a0m * (vc - vm1) + a1m * (vc - vm1) + a2m * dmc + a3m * dmm1;
    R vl = vm, vr = vp;
    if (br((vr - vc) * (vc - vl) <= R(0))) {
        vp = vc;
        vm = vc;
    } else {
        if (br((vr - vl) * (vl - R(3.0) * vc + R(2.0) * vr) < R(0))) vm = R(3.0) * vc - R(2.0) * vr;
        if (br((vp - vm) * (R(3.0) * vc - R(2.0) * vl - vr) < R(0))) vp = R(3.0) * vc - R(2.0) * vl;
    }
    vm_out = vm;
    vp_out = vp;
}

// reconstruct.f90:110-124 TVD minmod with precomputed cf, cb of the cell.
template <typename R>
__host__ __device__ inline void tvd_edges(R cf, R cb, R vm1, R vc, R vp1, R &vm_out, R &vp_out)
{
    R dqf = R(2.0) * (vp1 - vc) / cf;
    R dqb = R(2.0) * (vc - vm1) / cb;
    br(dqf < R(0));
    br(dqb < R(0));
    R slope = R(0.5) * (sign1<R>(dqf) + sign1<R>(dqb)) * fmin2b<R>(fabs_<R>(dqf), fabs_<R>(dqb));
    vp_out = vc + R(0.5) * slope;
    vm_out = vc - R(0.5) * slope;
}

// TVD with the MC (LIM 2: minmod(cf dq+, cb dq-, (dq+ + dq-)/2)) or van Leer (LIM 3: harmonic
// mean, capped at min(cf |dq+|, cb |dq-|)) limiter on the width-normalised differences of TVD_MM
// (reconstruct.f90:115-124): the
// TVDMC / TVDVL options that definition.f90:160-164 declares and the reference leaves
// unimplemented (extensions; restated as orc_tvd).
template <typename R, int LIM>
__host__ __device__ inline void tvd_lim_edges(R cf, R cb, R vm1, R vc, R vp1, R &vm_out, R &vp_out)
{
    R dqf = R(2.0) * (vp1 - vc) / cf;
    R dqb = R(2.0) * (vc - vm1) / cb;
    R slope;
    if (LIM == 2) {
        R m = fmin2<R>(fmin2<R>(cf * fabs_<R>(dqf), cb * fabs_<R>(dqb)), R(0.5) * fabs_<R>(dqf + dqb));
        slope = R(0.5) * (sign1<R>(dqf) + sign1<R>(dqb)) * m;
    } else {
        if (dqf * dqb > R(0)) {
            R h = R(2.0) * dqf * dqb / (dqf + dqb);
            slope = sign1<R>(h) * fmin2<R>(fmin2<R>(fabs_<R>(h), cf * fabs_<R>(dqf)), cb * fabs_<R>(dqb));
        } else {
            slope = R(0);
        }
    }
    vp_out = vc + R(0.5) * slope;
    vm_out = vc - R(0.5) * slope;
}

// WENO5 face value at the right face of the middle cell (Jiang & Shu 1996): the WENO option
// that definition.f90:169-170 declares and the reference leaves unimplemented (an extension;
// restated as orc_weno5 with the same operation order).  Uniform-grid coefficients on every
// grid, linear weights (1, 6, 3)/10, eps = 1e-6.
template <typename R>
__host__ __device__ inline R weno5_face(R vm2, R vm1, R vc, R vp1, R vp2)
{
    // the candidate values times 6; the 1/6 goes into the final normalisation
    const R p0 = R(2.0) * vm2 - R(7.0) * vm1 + R(11.0) * vc;
    const R p1 = -vm1 + R(5.0) * vc + R(2.0) * vp1;
    const R p2 = R(2.0) * vc + R(5.0) * vp1 - vp2;
    const R c13 = R(13.0 / 12.0);
    const R t0 = vm2 - R(2.0) * vm1 + vc, s0 = vm2 - R(4.0) * vm1 + R(3.0) * vc;
    const R t1 = vm1 - R(2.0) * vc + vp1, s1 = vm1 - vp1;
    const R t2 = vc - R(2.0) * vp1 + vp2, s2 = R(3.0) * vc - R(4.0) * vp1 + vp2;
    const R b0 = c13 * t0 * t0 + R(0.25) * s0 * s0;
    const R b1 = c13 * t1 * t1 + R(0.25) * s1 * s1;
    const R b2 = c13 * t2 * t2 + R(0.25) * s2 * s2;
    const R e0 = R(1.0e-6) + b0, e1 = R(1.0e-6) + b1, e2 = R(1.0e-6) + b2;
    const R a0 = R(0.1) / (e0 * e0), a1 = R(0.6) / (e1 * e1), a2 = R(0.3) / (e2 * e2);
    return (a0 * p0 + a1 * p1 + a2 * p2) / (R(6.0) * (a0 + a1 + a2));
}

// WENO5 edge values of cell c from its 5-cell stencil (the left one on the mirrored stencil)
template <typename R>
__host__ __device__ inline void weno5_edges(R vm2, R vm1, R vc, R vp1, R vp2, R &vm_out, R &vp_out)
{
    vp_out = weno5_face<R>(vm2, vm1, vc, vp1, vp2);
    vm_out = weno5_face<R>(vp2, vp1, vc, vm1, vm2);
}

// WENO5 is not positivity-preserving: a cell whose face density (variable 0) or pressure
// (variable 4) is not positive gets the piecewise-constant profile in every variable (the same
// fallback as the restatement's sweep)
template <typename R>
__host__ __device__ inline void weno_positivity(const R *c, R *vm, R *vp)
{
    if (vm[0] <= R(0) || vp[0] <= R(0) || vm[4] <= R(0) || vp[4] <= R(0)) {
        for (int v = 0; v < 7; v++) {
            vm[v] = c[v];
            vp[v] = c[v];
        }
    }
}

// edge values of cell c by a TVD option: 1 minmod (the reference's TVD_MM), 2 MC, 3 van Leer
template <typename R, int RECON>
__host__ __device__ inline void tvd_any(R cf, R cb, R vm1, R vc, R vp1, R &vm_out, R &vp_out)
{
    if (RECON == 1) tvd_edges<R>(cf, cb, vm1, vc, vp1, vm_out, vp_out);
    else tvd_lim_edges<R, RECON>(cf, cb, vm1, vc, vp1, vm_out, vp_out);
}

// Edge values of a cell whose whole stencil holds its own finite value v (the z sweep of a box
// with one periodic or outflow z cell), equal bit for bit to the general reconstruction with
// finite, positive table rows: every difference is +0, so the PPM slopes are +0 and its
// monotonicity branch sets both edges to v; the TVD limiters' slope is +0, so vm = v - 0 = v and
// vp = v + 0, which is v except that -0 becomes +0.  WENO5 has no such form (its weights do not
// sum to one exactly) and is not taken here.
template <typename R, int RECON>
__host__ __device__ inline void const_edges(R v, R &vm_out, R &vp_out)
{
    static_assert(RECON >= 1 && RECON <= 4, "PPM and the TVD options only");
    vm_out = v;
    vp_out = RECON == 4 ? v : v + R(0);
}

// One side of a face in the rotated frame of the sweep: n = normal, t1/t2 = transverse in the
// cyclic order of flux.f90:182-206 (x: n=x,t1=y,t2=z ; y: n=y,t1=z,t2=x ; z: n=z,t1=x,t2=y).
template <typename R>
struct FaceState {
    R rho, vn, vt1, vt2, p, bn, bt1, bt2;
};

// Bit-exact dot products need the reference's (x, y, z) summation order, so the rotated state
// carries which slot is x, y and z.
template <int DIR, typename R>
__host__ __device__ inline R dot_xyz(R an, R at1, R at2, R bn, R bt1, R bt2)
{
    if (DIR == 1) return an * bn + at1 * bt1 + at2 * bt2;        // (x,y,z) = (n,t1,t2)
    else if (DIR == 2) return at2 * bt2 + an * bn + at1 * bt1;   // (x,y,z) = (t2,n,t1)
    else return at1 * bt1 + at2 * bt2 + an * bn;                 // (x,y,z) = (t1,t2,n)
}

// rsolver.f90:43-130 HLL_SOLVER with ptou_face.f90:99-153 and flux.f90:163-262 inlined.
// Outputs, rotated: f[0] rho, f[1] normal momentum, f[2] t1 momentum, f[3] t2 momentum,
// f[4] energy, f[5] F(B_t1), f[6] F(B_t2)  (F(B_n) is identically zero and not stored).
template <int DIR, typename R>
__host__ __device__ inline void hll_face(R g, const FaceState<R> &L, const FaceState<R> &Rr, R f[7])
{
    const R gm1 = g - R(1.0);
    // EOS_EPSILON / EOS_SOUNDSPEED (model eos.f90:34-81)
    R epsL = L.p / L.rho / gm1;
    R epsR = Rr.p / Rr.rho / gm1;
    R csL = rsqrt_<R>(g * L.p / L.rho);
    R csR = rsqrt_<R>(g * Rr.p / Rr.rho);
    // P_to_U_face
    R vsqL = dot_xyz<DIR, R>(L.vn, L.vt1, L.vt2, L.vn, L.vt1, L.vt2);
    R bsqL = dot_xyz<DIR, R>(L.bn, L.bt1, L.bt2, L.bn, L.bt1, L.bt2);
    R vsqR = dot_xyz<DIR, R>(Rr.vn, Rr.vt1, Rr.vt2, Rr.vn, Rr.vt1, Rr.vt2);
    R bsqR = dot_xyz<DIR, R>(Rr.bn, Rr.bt1, Rr.bt2, Rr.bn, Rr.bt1, Rr.bt2);
    R uL[5], uR[5];
    uL[0] = L.rho; uL[1] = L.vn * L.rho; uL[2] = L.vt1 * L.rho; uL[3] = L.vt2 * L.rho;
    uL[4] = L.rho * (epsL + R(0.5) * vsqL) + R(0.5) * bsqL;
    uR[0] = Rr.rho; uR[1] = Rr.vn * Rr.rho; uR[2] = Rr.vt1 * Rr.rho; uR[3] = Rr.vt2 * Rr.rho;
    uR[4] = Rr.rho * (epsR + R(0.5) * vsqR) + R(0.5) * bsqR;
    // P_to_flux
    R vdbL = dot_xyz<DIR, R>(L.bn, L.bt1, L.bt2, L.vn, L.vt1, L.vt2);
    R vdbR = dot_xyz<DIR, R>(Rr.bn, Rr.bt1, Rr.bt2, Rr.vn, Rr.vt1, Rr.vt2);
    R fL[7], fR[7];
    fL[0] = uL[0] * L.vn;
    fL[1] = uL[1] * L.vn + L.p + R(0.5) * bsqL - L.bn * L.bn;
    fL[2] = uL[2] * L.vn - L.bn * L.bt1;
    fL[3] = uL[3] * L.vn - L.bn * L.bt2;
    fL[4] = uL[4] * L.vn + (L.p + R(0.5) * bsqL) * L.vn - L.bn * vdbL;
    fL[5] = (L.vn * L.bt1 - L.vt1 * L.bn);
    fL[6] = (L.vn * L.bt2 - L.vt2 * L.bn);
    fR[0] = uR[0] * Rr.vn;
    fR[1] = uR[1] * Rr.vn + Rr.p + R(0.5) * bsqR - Rr.bn * Rr.bn;
    fR[2] = uR[2] * Rr.vn - Rr.bn * Rr.bt1;
    fR[3] = uR[3] * Rr.vn - Rr.bn * Rr.bt2;
    fR[4] = uR[4] * Rr.vn + (Rr.p + R(0.5) * bsqR) * Rr.vn - Rr.bn * vdbR;
    fR[5] = (Rr.vn * Rr.bt1 - Rr.vt1 * Rr.bn);
    fR[6] = (Rr.vn * Rr.bt2 - Rr.vt2 * Rr.bn);
    // signal speeds (rsolver.f90:110-115)
    R cfsL = signalspeed<R>(csL, L.bn, L.bt1, L.bt2, L.rho);
    R cfsR = signalspeed<R>(csR, Rr.bn, Rr.bt1, Rr.bt2, Rr.rho);
    R sqL = rsqrt_<R>(L.rho), sqR = rsqrt_<R>(Rr.rho);
    R ubar = (sqL * L.vn + sqR * Rr.vn) / (sqL + sqR);
    R cbar = (sqL * cfsL + sqR * cfsR) / (sqL + sqR);
    R sL = fmin2b<R>(L.vn - cfsL, ubar - cbar);
    R sR = fmax2b<R>(Rr.vn + cfsR, ubar + cbar);
    if (br(sL >= R(0))) {
        for (int i = 0; i < 7; i++) f[i] = fL[i];
    } else if (br(sL <= R(0) && sR >= R(0))) {
        for (int i = 0; i < 5; i++) f[i] = fluxhll<R>(fL[i], fR[i], uL[i], uR[i], sL, sR);
        f[5] = fluxhll<R>(fL[5], fR[5], L.bt1, Rr.bt1, sL, sR);
        f[6] = fluxhll<R>(fL[6], fR[6], L.bt2, Rr.bt2, sL, sR);
    } else if (br(sR <= R(0))) {
        for (int i = 0; i < 7; i++) f[i] = fR[i];
    } else {
        for (int i = 0; i < 7; i++) f[i] = R(NAN);
    }
}

// One side of a face with its conserved vector and physical flux in the rotated order rho, m_n,
// m_t1, m_t2, E, B_t1, B_t2 (the expressions of hll_face), total pressure and fast speed: the
// common front of the LF and HLLD solvers.
template <typename R>
struct Side7 {
    R rho, vn, vt1, vt2, p, bt1, bt2, pt, cf;
    R u[7], f[7];
};

template <int DIR, typename R>
__host__ __device__ inline void face_side(R g, const FaceState<R> &S, Side7<R> &o)
{
    const R gm1 = g - R(1.0);
    R eps = S.p / S.rho / gm1;
    R cs = rsqrt_<R>(g * S.p / S.rho);
    R vsq = dot_xyz<DIR, R>(S.vn, S.vt1, S.vt2, S.vn, S.vt1, S.vt2);
    R bsq = dot_xyz<DIR, R>(S.bn, S.bt1, S.bt2, S.bn, S.bt1, S.bt2);
    R vdb = dot_xyz<DIR, R>(S.bn, S.bt1, S.bt2, S.vn, S.vt1, S.vt2);
    o.rho = S.rho; o.vn = S.vn; o.vt1 = S.vt1; o.vt2 = S.vt2; o.p = S.p; o.bt1 = S.bt1; o.bt2 = S.bt2;
    o.u[0] = S.rho;
    o.u[1] = S.vn * S.rho;
    o.u[2] = S.vt1 * S.rho;
    o.u[3] = S.vt2 * S.rho;
    o.u[4] = S.rho * (eps + R(0.5) * vsq) + R(0.5) * bsq;
    o.u[5] = S.bt1;
    o.u[6] = S.bt2;
    o.f[0] = o.u[0] * S.vn;
    o.f[1] = o.u[1] * S.vn + S.p + R(0.5) * bsq - S.bn * S.bn;
    o.f[2] = o.u[2] * S.vn - S.bn * S.bt1;
    o.f[3] = o.u[3] * S.vn - S.bn * S.bt2;
    o.f[4] = o.u[4] * S.vn + (S.p + R(0.5) * bsq) * S.vn - S.bn * vdb;
    o.f[5] = (S.vn * S.bt1 - S.vt1 * S.bn);
    o.f[6] = (S.vn * S.bt2 - S.vt2 * S.bn);
    o.pt = S.p + R(0.5) * bsq;
    o.cf = signalspeed<R>(cs, S.bn, S.bt1, S.bt2, S.rho);
}

// Local Lax-Friedrichs (Rusanov) flux, the LF option of definition.f90:139-140 (an extension;
// restated as orc_lf): F = (F_L + F_R)/2 - s (U_R - U_L)/2, s = max(|v_nL| + c_fL, |v_nR| + c_fR).
template <int DIR, typename R>
__host__ __device__ inline void lf_face(R g, const FaceState<R> &L, const FaceState<R> &Rr, R f[7])
{
    Side7<R> a, b;
    face_side<DIR, R>(g, L, a);
    face_side<DIR, R>(g, Rr, b);
    const R smax = fmax2<R>(fabs_<R>(a.vn) + a.cf, fabs_<R>(b.vn) + b.cf);
    for (int i = 0; i < 7; i++) f[i] = R(0.5) * (a.f[i] + b.f[i]) - R(0.5) * smax * (b.u[i] - a.u[i]);
}

// HLLD flux of Miyoshi & Kusano (2005, J. Comput. Phys. 208, 315), the HLLD option of
// definition.f90:148-149 (an extension; restated as orc_hlld with the same operation order):
// outer speeds min/max(v_nL, v_nR) -/+ max(c_fL, c_fR) (eq. 67); contact speed S_M and p_T*
// (eqs. 38, 41); single-star states (eqs. 43-48, degenerate denominators keep the transverse v
// and B); Alfven speeds S_M -/+ |B_n|/sqrt(rho*) (eq. 51); double-star states (eqs. 59-63).
template <int DIR, typename R>
__host__ __device__ inline void hlld_face(R g, const FaceState<R> &L, const FaceState<R> &Rr, R f[7])
{
    Side7<R> s[2];
    face_side<DIR, R>(g, L, s[0]);
    face_side<DIR, R>(g, Rr, s[1]);
    const R bn = L.bn;
    const R cmax = fmax2<R>(s[0].cf, s[1].cf);
    const R sL = fmin2<R>(s[0].vn, s[1].vn) - cmax;
    const R sR = fmax2<R>(s[0].vn, s[1].vn) + cmax;
    if (sL >= R(0)) {
        for (int i = 0; i < 7; i++) f[i] = s[0].f[i];
        return;
    }
    if (sR <= R(0)) {
        for (int i = 0; i < 7; i++) f[i] = s[1].f[i];
        return;
    }
    if (!(sL < R(0) && sR > R(0))) {
        for (int i = 0; i < 7; i++) f[i] = R(NAN);
        return;
    }
    const R sdl = sL - s[0].vn, sdr = sR - s[1].vn;
    const R iden = R(1.0) / (sdr * s[1].rho - sdl * s[0].rho);
    const R sm = (sdr * s[1].rho * s[1].vn - sdl * s[0].rho * s[0].vn - s[1].pt + s[0].pt) * iden;
    const R pts = (sdr * s[1].rho * s[0].pt - sdl * s[0].rho * s[1].pt + s[0].rho * s[1].rho * sdr * sdl * (s[1].vn - s[0].vn)) * iden;
    // (every array below is indexed by compile-time constants only, so it stays in registers)
    R us[2][7], vts[2][2], bts[2][2], sq[2];
    const R sa[2] = {sL, sR}, sda[2] = {sdl, sdr};
#pragma unroll
    for (int a = 0; a < 2; a++) {
        const Side7<R> &q = s[a];
        const R sdm = sa[a] - sm;
        const R isdm = R(1.0) / sdm;
        const R rhos = q.rho * sda[a] * isdm;
        const R tmp = q.rho * sda[a] * sdm - bn * bn;
        if (fabs_<R>(tmp) < R(1.0e-8) * fabs_<R>(pts)) {
            vts[a][0] = q.vt1;
            vts[a][1] = q.vt2;
            bts[a][0] = q.bt1;
            bts[a][1] = q.bt2;
        } else {
            const R fac = (q.rho * sda[a] * sda[a] - bn * bn) / tmp;
            const R fv = bn * (sm - q.vn) / tmp;
            vts[a][0] = q.vt1 - q.bt1 * fv;
            vts[a][1] = q.vt2 - q.bt2 * fv;
            bts[a][0] = q.bt1 * fac;
            bts[a][1] = q.bt2 * fac;
        }
        const R vdb = q.vn * bn + q.vt1 * q.bt1 + q.vt2 * q.bt2;
        const R vdbs = sm * bn + vts[a][0] * bts[a][0] + vts[a][1] * bts[a][1];
        const R es = (sda[a] * q.u[4] - q.pt * q.vn + pts * sm + bn * (vdb - vdbs)) * isdm;
        us[a][0] = rhos;
        us[a][1] = rhos * sm;
        us[a][2] = rhos * vts[a][0];
        us[a][3] = rhos * vts[a][1];
        us[a][4] = es;
        us[a][5] = bts[a][0];
        us[a][6] = bts[a][1];
        sq[a] = rsqrt_<R>(rhos);
    }
    const R ssL = sm - fabs_<R>(bn) / sq[0];
    const R ssR = sm + fabs_<R>(bn) / sq[1];
    if (ssL >= R(0)) {
        for (int i = 0; i < 7; i++) f[i] = s[0].f[i] + sL * (us[0][i] - s[0].u[i]);
        return;
    }
    if (ssR <= R(0)) {
        for (int i = 0; i < 7; i++) f[i] = s[1].f[i] + sR * (us[1][i] - s[1].u[i]);
        return;
    }
    const R sgn = sign1<R>(bn);
    const R isq = R(1.0) / (sq[0] + sq[1]);
    const R vss1 = (sq[0] * vts[0][0] + sq[1] * vts[1][0] + (bts[1][0] - bts[0][0]) * sgn) * isq;
    const R vss2 = (sq[0] * vts[0][1] + sq[1] * vts[1][1] + (bts[1][1] - bts[0][1]) * sgn) * isq;
    const R bss1 = (sq[0] * bts[1][0] + sq[1] * bts[0][0] + sq[0] * sq[1] * (vts[1][0] - vts[0][0]) * sgn) * isq;
    const R bss2 = (sq[0] * bts[1][1] + sq[1] * bts[0][1] + sq[0] * sq[1] * (vts[1][1] - vts[0][1]) * sgn) * isq;
    const R vdbss = sm * bn + vss1 * bss1 + vss2 * bss2;
    const bool left = sm >= R(0);
    const R vdbsL = sm * bn + vts[0][0] * bts[0][0] + vts[0][1] * bts[0][1];
    const R vdbsR = sm * bn + vts[1][0] * bts[1][0] + vts[1][1] * bts[1][1];
    const R ess = left ? us[0][4] - sq[0] * (vdbsL - vdbss) * sgn : us[1][4] + sq[1] * (vdbsR - vdbss) * sgn;
    const R rs = left ? us[0][0] : us[1][0];
    const R uss[7] = {rs, left ? us[0][1] : us[1][1], rs * vss1, rs * vss2, ess, bss1, bss2};
    const R sa_ = left ? sL : sR, ss = left ? ssL : ssR;
#pragma unroll
    for (int i = 0; i < 7; i++) {
        const R usi = left ? us[0][i] : us[1][i];
        const R fs = (left ? s[0].f[i] : s[1].f[i]) + sa_ * (usi - (left ? s[0].u[i] : s[1].u[i]));
        f[i] = fs + ss * (uss[i] - usi);
    }
}

// HLLC flux for MHD in the form of Li (2005, J. Comput. Phys. 203, 344), the HLLC option of
// definition.f90:145-146 (an extension; restated as orc_hllc with the same operation order): HLL
// state for the star-region field B* and velocity (normal component S_M), p_T* from the left
// state, single-star states on the side of S_M.
template <int DIR, typename R>
__host__ __device__ inline void hllc_face(R g, const FaceState<R> &L, const FaceState<R> &Rr, R f[7])
{
    Side7<R> a, b;
    face_side<DIR, R>(g, L, a);
    face_side<DIR, R>(g, Rr, b);
    const R bn = L.bn;
    const R cmax = fmax2<R>(a.cf, b.cf);
    const R sL = fmin2<R>(a.vn, b.vn) - cmax;
    const R sR = fmax2<R>(a.vn, b.vn) + cmax;
    if (sL >= R(0)) {
        for (int i = 0; i < 7; i++) f[i] = a.f[i];
        return;
    }
    if (sR <= R(0)) {
        for (int i = 0; i < 7; i++) f[i] = b.f[i];
        return;
    }
    if (!(sL < R(0) && sR > R(0))) {
        for (int i = 0; i < 7; i++) f[i] = R(NAN);
        return;
    }
    const R ids = R(1.0) / (sR - sL);
    R uh[7];
#pragma unroll
    for (int i = 0; i < 7; i++) uh[i] = (sR * b.u[i] - sL * a.u[i] - b.f[i] + a.f[i]) * ids;
    const R irh = R(1.0) / uh[0];
    const R sm = uh[1] * irh, vt1 = uh[2] * irh, vt2 = uh[3] * irh;
    const R bt1 = uh[5], bt2 = uh[6];
    const R pts = a.rho * (sL - a.vn) * (sm - a.vn) + a.pt;
    const R vdbs = sm * bn + vt1 * bt1 + vt2 * bt2;
    const bool left = sm >= R(0);
    // the side of S_M, selected element by element (no runtime array index)
    const R s = left ? sL : sR;
    const R qrho = left ? a.rho : b.rho, qvn = left ? a.vn : b.vn, qvt1 = left ? a.vt1 : b.vt1;
    const R qvt2 = left ? a.vt2 : b.vt2, qbt1 = left ? a.bt1 : b.bt1, qbt2 = left ? a.bt2 : b.bt2;
    const R qpt = left ? a.pt : b.pt, qe = left ? a.u[4] : b.u[4];
    const R sd = s - qvn, isdm = R(1.0) / (s - sm);
    const R vdb = qvn * bn + qvt1 * qbt1 + qvt2 * qbt2;
    R us[7];
    us[0] = qrho * sd * isdm;
    us[1] = us[0] * sm;
    us[2] = (qrho * qvt1 * sd - (bn * bt1 - bn * qbt1)) * isdm;
    us[3] = (qrho * qvt2 * sd - (bn * bt2 - bn * qbt2)) * isdm;
    us[4] = (qe * sd - qpt * qvn + pts * sm + bn * (vdb - vdbs)) * isdm;
    us[5] = bt1;
    us[6] = bt2;
#pragma unroll
    for (int i = 0; i < 7; i++) f[i] = (left ? a.f[i] : b.f[i]) + s * (us[i] - (left ? a.u[i] : b.u[i]));
}

// the face solver of the configuration: CUMC_LF (1), CUMC_HLL (2, the reference's), CUMC_HLLC (3),
// CUMC_HLLD (4)
template <int DIR, typename R, int SOLVER>
__host__ __device__ inline void riemann_face(R g, const FaceState<R> &L, const FaceState<R> &Rr, R f[7])
{
    if (SOLVER == 1) lf_face<DIR, R>(g, L, Rr, f);
    else if (SOLVER == 3) hllc_face<DIR, R>(g, L, Rr, f);
    else if (SOLVER == 4) hlld_face<DIR, R>(g, L, Rr, f);
    else hll_face<DIR, R>(g, L, Rr, f);
}

}  // namespace cumc
