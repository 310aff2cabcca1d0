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
    return rsqrt_<R>(R(0.5) * (a2 + b2 + rsqrt_<R>(fmax2<R>((a4 + b4 + R(2.0) * b2 * a2) - R(4.0) * a2 * b2n, R(0)))));
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
    if ((vp1 - vc) * (vc - vm1) > R(0)) {
        R a = fabs_<R>(delta), b = R(2.0) * fabs_<R>(vc - vm1), c = R(2.0) * fabs_<R>(vc - vp1);
        return fmin2<R>(fmin2<R>(a, b), c) * sign1<R>(delta);
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
    if ((vr - vc) * (vc - vl) <= R(0)) {
        vp = vc;
        vm = vc;
    } else {
        if ((vr - vl) * (vl - R(3.0) * vc + R(2.0) * vr) < R(0)) vm = R(3.0) * vc - R(2.0) * vr;
        if ((vp - vm) * (R(3.0) * vc - R(2.0) * vl - vr) < R(0)) vp = R(3.0) * vc - R(2.0) * vl;
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
    R slope = R(0.5) * (sign1<R>(dqf) + sign1<R>(dqb)) * fmin2<R>(fabs_<R>(dqf), fabs_<R>(dqb));
    vp_out = vc + R(0.5) * slope;
    vm_out = vc - R(0.5) * slope;
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
    R sL = fmin2<R>(L.vn - cfsL, ubar - cbar);
    R sR = fmax2<R>(Rr.vn + cfsR, ubar + cbar);
    if (sL >= R(0)) {
        for (int i = 0; i < 7; i++) f[i] = fL[i];
    } else if (sL <= R(0) && sR >= R(0)) {
        for (int i = 0; i < 5; i++) f[i] = fluxhll<R>(fL[i], fR[i], uL[i], uR[i], sL, sR);
        f[5] = fluxhll<R>(fL[5], fR[5], L.bt1, Rr.bt1, sL, sR);
        f[6] = fluxhll<R>(fL[6], fR[6], L.bt2, Rr.bt2, sL, sR);
    } else if (sR <= R(0)) {
        for (int i = 0; i < 7; i++) f[i] = fR[i];
    } else {
        for (int i = 0; i < 7; i++) f[i] = R(NAN);
    }
}

}  // namespace cumc
