// adjoint.hpp -- reverse-mode derivatives of the point numerics in physics.hpp.
//
// Each *_adj function recomputes the forward intermediates of its primal (same inputs, same
// branch decisions) and accumulates  lambda_in += (d out / d in)^T lambda_out.  The derivative
// is that of the real-valued function the primal evaluates, taken on the branch the primal took:
// limiter, HLL-region, min/max and clamp selections are replayed, not differentiated, so the
// result is the derivative of the piecewise-smooth step at points off the switching surfaces.
#pragma once
#include "physics.hpp"

namespace cumc {

// d/dx sqrt(x) * lam with the value s = sqrt(x) known; zero at s = 0 (one-sided choice)
__host__ __device__ inline double dsqrt(double s, double lam) { return s > 0.0 ? lam * 0.5 / s : 0.0; }

// compute_signalspeed (definition.f90:394-410) reverse: cf(cs, bn, bt1, bt2, rho)
__host__ __device__ inline void signalspeed_adj(double cs, double bn, double bt1, double bt2, double rho, double lcf,
                                                double &lcs, double &lbn, double &lbt1, double &lbt2, double &lrho)
{
    const double a2 = cs * cs;
    const double b2n = bn * bn / rho, b2t1 = bt1 * bt1 / rho, b2t2 = bt2 * bt2 / rho;
    const double b2 = b2n + b2t1 + b2t2;
    const double a4 = a2 * a2, b4 = b2 * b2;
    const double disc = (a4 + b4 + 2.0 * b2 * a2) - 4.0 * a2 * b2n;
    const double dc = fmax2<double>(disc, 0.0);
    const double sd = sqrt(dc);
    const double inner = 0.5 * (a2 + b2 + sd);
    const double cf = sqrt(inner);
    const double linner = dsqrt(cf, lcf);
    double la2 = 0.5 * linner, lb2 = 0.5 * linner;
    const double lsd = 0.5 * linner;
    const double ldc = dsqrt(sd, lsd);
    const double ldisc = (0.0 > disc) ? 0.0 : ldc;  // fmax2(disc, 0) picked 0 when 0 > disc
    la2 += ldisc * (2.0 * (a2 + b2) - 4.0 * b2n);
    lb2 += ldisc * 2.0 * (a2 + b2);
    const double lb2n = lb2 - 4.0 * a2 * ldisc;
    const double lb2t1 = lb2, lb2t2 = lb2;
    lcs += la2 * 2.0 * cs;
    lbn += lb2n * 2.0 * bn / rho;
    lbt1 += lb2t1 * 2.0 * bt1 / rho;
    lbt2 += lb2t2 * 2.0 * bt2 / rho;
    lrho -= (lb2n * b2n + lb2t1 * b2t1 + lb2t2 * b2t2) / rho;
}

// One side of a face: P_to_U_face / P_to_flux (ptou_face.f90:99-153, flux.f90:163-262) in the
// rotated frame, as plain formulas of (rho, vn, vt1, vt2, p, bn, bt1, bt2).
struct SideFwd {
    double eps, cs, vsq, bsq, u[5], f[7], vdb;
};

__host__ __device__ inline SideFwd side_fwd(double g, const FaceState<double> &s)
{
    SideFwd o;
    const double gm1 = g - 1.0;
    o.eps = s.p / s.rho / gm1;
    o.cs = sqrt(g * s.p / s.rho);
    o.vsq = s.vn * s.vn + s.vt1 * s.vt1 + s.vt2 * s.vt2;
    o.bsq = s.bn * s.bn + s.bt1 * s.bt1 + s.bt2 * s.bt2;
    o.u[0] = s.rho;
    o.u[1] = s.vn * s.rho;
    o.u[2] = s.vt1 * s.rho;
    o.u[3] = s.vt2 * s.rho;
    o.u[4] = s.rho * (o.eps + 0.5 * o.vsq) + 0.5 * o.bsq;
    o.vdb = s.bn * s.vn + s.bt1 * s.vt1 + s.bt2 * s.vt2;
    o.f[0] = o.u[0] * s.vn;
    o.f[1] = o.u[1] * s.vn + s.p + 0.5 * o.bsq - s.bn * s.bn;
    o.f[2] = o.u[2] * s.vn - s.bn * s.bt1;
    o.f[3] = o.u[3] * s.vn - s.bn * s.bt2;
    o.f[4] = o.u[4] * s.vn + (s.p + 0.5 * o.bsq) * s.vn - s.bn * o.vdb;
    o.f[5] = s.vn * s.bt1 - s.vt1 * s.bn;
    o.f[6] = s.vn * s.bt2 - s.vt2 * s.bn;
    return o;
}

// reverse of side_fwd: lu[5], lf[7] and lcs given, accumulate into ls (cs enters only the
// signal speed, whose adjoint the caller folds into lcs)
__host__ __device__ inline void side_adj(double g, const FaceState<double> &s, const SideFwd &o, const double lu_in[5],
                                         const double lf[7], double lcs, FaceState<double> &ls)
{
    const double gm1 = g - 1.0;
    double lu[5] = {lu_in[0], lu_in[1], lu_in[2], lu_in[3], lu_in[4]};
    double lvn = 0.0, lvt1 = 0.0, lvt2 = 0.0, lp = 0.0, lbn = 0.0, lbt1 = 0.0, lbt2 = 0.0, lrho = 0.0;
    double lbsq = 0.0, lvsq = 0.0, lvdb = 0.0, leps = 0.0;
    // fluxes
    lu[0] += lf[0] * s.vn;
    lvn += lf[0] * o.u[0];
    lu[1] += lf[1] * s.vn;
    lvn += lf[1] * o.u[1];
    lp += lf[1];
    lbsq += 0.5 * lf[1];
    lbn -= 2.0 * s.bn * lf[1];
    lu[2] += lf[2] * s.vn;
    lvn += lf[2] * o.u[2];
    lbn -= lf[2] * s.bt1;
    lbt1 -= lf[2] * s.bn;
    lu[3] += lf[3] * s.vn;
    lvn += lf[3] * o.u[3];
    lbn -= lf[3] * s.bt2;
    lbt2 -= lf[3] * s.bn;
    lu[4] += lf[4] * s.vn;
    lvn += lf[4] * (o.u[4] + s.p + 0.5 * o.bsq);
    lp += lf[4] * s.vn;
    lbsq += 0.5 * lf[4] * s.vn;
    lbn -= lf[4] * o.vdb;
    lvdb -= lf[4] * s.bn;
    lvn += lf[5] * s.bt1;
    lbt1 += lf[5] * s.vn;
    lvt1 -= lf[5] * s.bn;
    lbn -= lf[5] * s.vt1;
    lvn += lf[6] * s.bt2;
    lbt2 += lf[6] * s.vn;
    lvt2 -= lf[6] * s.bn;
    lbn -= lf[6] * s.vt2;
    // v.B
    lbn += lvdb * s.vn;
    lvn += lvdb * s.bn;
    lbt1 += lvdb * s.vt1;
    lvt1 += lvdb * s.bt1;
    lbt2 += lvdb * s.vt2;
    lvt2 += lvdb * s.bt2;
    // conserved
    lrho += lu[0];
    lvn += lu[1] * s.rho;
    lrho += lu[1] * s.vn;
    lvt1 += lu[2] * s.rho;
    lrho += lu[2] * s.vt1;
    lvt2 += lu[3] * s.rho;
    lrho += lu[3] * s.vt2;
    lrho += lu[4] * (o.eps + 0.5 * o.vsq);
    leps += lu[4] * s.rho;
    lvsq += lu[4] * 0.5 * s.rho;
    lbsq += lu[4] * 0.5;
    // squares
    lvn += 2.0 * s.vn * lvsq;
    lvt1 += 2.0 * s.vt1 * lvsq;
    lvt2 += 2.0 * s.vt2 * lvsq;
    lbn += 2.0 * s.bn * lbsq;
    lbt1 += 2.0 * s.bt1 * lbsq;
    lbt2 += 2.0 * s.bt2 * lbsq;
    // eps = p/rho/gm1, cs = sqrt(g p / rho)
    lp += leps / (s.rho * gm1);
    lrho -= leps * o.eps / s.rho;
    if (o.cs > 0.0) {
        lp += lcs * g / (2.0 * o.cs * s.rho);
        lrho -= lcs * o.cs / (2.0 * s.rho);
    }
    ls.rho += lrho;
    ls.vn += lvn;
    ls.vt1 += lvt1;
    ls.vt2 += lvt2;
    ls.p += lp;
    ls.bn += lbn;
    ls.bt1 += lbt1;
    ls.bt2 += lbt2;
}

// HLL_SOLVER (rsolver.f90:43-130) reverse.  lf: adjoint of the rotated flux f[0..6] of
// hll_face; accumulates into lL, lR (bn appears on both sides: the caller adds lL.bn + lR.bn).
__host__ __device__ inline void hll_face_adj(double g, const FaceState<double> &L, const FaceState<double> &R,
                                             const double lf[7], FaceState<double> &lL, FaceState<double> &lR)
{
    const SideFwd oL = side_fwd(g, L), oR = side_fwd(g, R);
    const double cfsL = signalspeed<double>(oL.cs, L.bn, L.bt1, L.bt2, L.rho);
    const double cfsR = signalspeed<double>(oR.cs, R.bn, R.bt1, R.bt2, R.rho);
    const double sqL = sqrt(L.rho), sqR = sqrt(R.rho);
    const double S = sqL + sqR;
    const double ubar = (sqL * L.vn + sqR * R.vn) / S;
    const double cbar = (sqL * cfsL + sqR * cfsR) / S;
    const double A1 = L.vn - cfsL, A2 = ubar - cbar;
    const double B1 = R.vn + cfsR, B2 = ubar + cbar;
    const double sL = fmin2<double>(A1, A2);
    const double sR = fmax2<double>(B1, B2);
    double lfL[7] = {0, 0, 0, 0, 0, 0, 0}, lfR[7] = {0, 0, 0, 0, 0, 0, 0};
    double luL[5] = {0, 0, 0, 0, 0}, luR[5] = {0, 0, 0, 0, 0};
    double lsL = 0.0, lsR = 0.0, lxbt[4] = {0, 0, 0, 0};  // bt1L, bt1R, bt2L, bt2R through the HLL state
    if (sL >= 0.0) {
        for (int i = 0; i < 7; i++) lfL[i] = lf[i];
    } else if (sL <= 0.0 && sR >= 0.0) {
        const double D = sR - sL;
        for (int i = 0; i < 7; i++) {
            const double xl = i < 5 ? oL.u[i] : (i == 5 ? L.bt1 : L.bt2);
            const double xr = i < 5 ? oR.u[i] : (i == 5 ? R.bt1 : R.bt2);
            const double a = oL.f[i], b = oR.f[i];
            const double f = (sR * a - sL * b + sL * sR * (xr - xl)) / D;
            const double l = lf[i];
            lfL[i] += l * sR / D;
            lfR[i] -= l * sL / D;
            const double lx = l * sL * sR / D;
            if (i < 5) {
                luR[i] += lx;
                luL[i] -= lx;
            } else if (i == 5) {
                lxbt[1] += lx;
                lxbt[0] -= lx;
            } else {
                lxbt[3] += lx;
                lxbt[2] -= lx;
            }
            lsR += l * (a + sL * (xr - xl) - f) / D;
            lsL += l * (-b + sR * (xr - xl) + f) / D;
        }
    } else if (sR <= 0.0) {
        for (int i = 0; i < 7; i++) lfR[i] = lf[i];
    }
    // sL = min(A1, A2), sR = max(B1, B2) with the primal's tie rule
    double lA1 = 0.0, lA2 = 0.0, lB1 = 0.0, lB2 = 0.0;
    if (A2 < A1) lA2 = lsL;
    else lA1 = lsL;
    if (B2 > B1) lB2 = lsR;
    else lB1 = lsR;
    const double lubar = lA2 + lB2, lcbar = -lA2 + lB2;
    double lvnL = lA1 + lubar * sqL / S, lvnR = lB1 + lubar * sqR / S;
    double lcfsL = -lA1 + lcbar * sqL / S, lcfsR = lB1 + lcbar * sqR / S;
    const double lsqL = lubar * (L.vn - ubar) / S + lcbar * (cfsL - cbar) / S;
    const double lsqR = lubar * (R.vn - ubar) / S + lcbar * (cfsR - cbar) / S;
    double lcsL = 0.0, lcsR = 0.0;
    FaceState<double> tL{0, 0, 0, 0, 0, 0, 0, 0}, tR{0, 0, 0, 0, 0, 0, 0, 0};
    signalspeed_adj(oL.cs, L.bn, L.bt1, L.bt2, L.rho, lcfsL, lcsL, tL.bn, tL.bt1, tL.bt2, tL.rho);
    signalspeed_adj(oR.cs, R.bn, R.bt1, R.bt2, R.rho, lcfsR, lcsR, tR.bn, tR.bt1, tR.bt2, tR.rho);
    tL.rho += dsqrt(sqL, lsqL);
    tR.rho += dsqrt(sqR, lsqR);
    tL.vn += lvnL;
    tR.vn += lvnR;
    tL.bt1 += lxbt[0];
    tR.bt1 += lxbt[1];
    tL.bt2 += lxbt[2];
    tR.bt2 += lxbt[3];
    side_adj(g, L, oL, luL, lfL, lcsL, tL);
    side_adj(g, R, oR, luR, lfR, lcsR, tR);
    lL.rho += tL.rho; lL.vn += tL.vn; lL.vt1 += tL.vt1; lL.vt2 += tL.vt2;
    lL.p += tL.p; lL.bn += tL.bn; lL.bt1 += tL.bt1; lL.bt2 += tL.bt2;
    lR.rho += tR.rho; lR.vn += tR.vn; lR.vt1 += tR.vt1; lR.vt2 += tR.vt2;
    lR.p += tR.p; lR.bn += tR.bn; lR.bt1 += tR.bt1; lR.bt2 += tR.bt2;
}

// ppm_slope reverse (reconstruct.f90:254-280): accumulate into lvm1, lvc, lvp1
__host__ __device__ inline void ppm_slope_adj(double b0, double b1, double vm1, double vc, double vp1, double l,
                                              double &lvm1, double &lvc, double &lvp1)
{
    if (!((vp1 - vc) * (vc - vm1) > 0.0)) return;
    const double delta = b0 * (vp1 - vc) + b1 * (vc - vm1);
    const double a = fabs(delta), b = 2.0 * fabs(vc - vm1), c = 2.0 * fabs(vc - vp1);
    const double sg = sign1<double>(delta);
    // fmin2(fmin2(a, b), c) with the primal's tie rule
    int pick = 0;
    double m = a;
    if (b < m) { m = b; pick = 1; }
    if (c < m) { m = c; pick = 2; }
    if (pick == 0) {
        // |delta| * sign(delta) = delta
        lvp1 += l * b0;
        lvc += l * (b1 - b0);
        lvm1 -= l * b1;
    } else if (pick == 1) {
        const double d = 2.0 * sign1<double>(vc - vm1) * sg * l;
        lvc += d;
        lvm1 -= d;
    } else {
        const double d = 2.0 * sign1<double>(vc - vp1) * sg * l;
        lvc += d;
        lvp1 -= d;
    }
}

// ppm_edges reverse (reconstruct.f90:285-307): given the adjoints of vm_out / vp_out,
// accumulate into the three cell values and the three slopes
__host__ __device__ inline void ppm_edges_adj(double a0m, double a1m, double a2m, double a3m, double a0p, double a1p,
                                              double a2p, double a3p, double vm1, double vc, double vp1, double dmm1,
                                              double dmc, double dmp1, double lvm, double lvp, double &lvm1,
                                              double &lvc, double &lvp1, double &ldmm1, double &ldmc, double &ldmp1)
{
    const double vp0 = vc + a0p * (vp1 - vc) + a1p * (vp1 - vc) + a2p * dmp1 + a3p * dmc;
    const double vm0 = vm1 + a0m * (vc - vm1) + a1m * (vc - vm1) + a2m * dmc + a3m * dmm1;
    const double vl = vm0, vr = vp0;
    double lvp0 = 0.0, lvm0 = 0.0;
    if ((vr - vc) * (vc - vl) <= 0.0) {
        lvc += lvm + lvp;
    } else {
        double vm = vm0;
        const bool first = (vr - vl) * (vl - 3.0 * vc + 2.0 * vr) < 0.0;
        if (first) vm = 3.0 * vc - 2.0 * vr;
        const bool second = (vp0 - vm) * (3.0 * vc - 2.0 * vl - vr) < 0.0;
        if (second) {
            lvc += 3.0 * lvp;
            lvm0 -= 2.0 * lvp;
        } else {
            lvp0 += lvp;
        }
        if (first) {
            lvc += 3.0 * lvm;
            lvp0 -= 2.0 * lvm;
        } else {
            lvm0 += lvm;
        }
    }
    const double cp = a0p + a1p, cm = a0m + a1m;
    lvc += lvp0 * (1.0 - cp);
    lvp1 += lvp0 * cp;
    ldmp1 += lvp0 * a2p;
    ldmc += lvp0 * a3p;
    lvm1 += lvm0 * (1.0 - cm);
    lvc += lvm0 * cm;
    ldmc += lvm0 * a2m;
    ldmm1 += lvm0 * a3m;
}

// tvd_edges reverse (reconstruct.f90:110-124)
__host__ __device__ inline void tvd_edges_adj(double cf, double cb, double vm1, double vc, double vp1, double lvm,
                                              double lvp, double &lvm1, double &lvc, double &lvp1)
{
    const double dqf = 2.0 * (vp1 - vc) / cf;
    const double dqb = 2.0 * (vc - vm1) / cb;
    const double s = 0.5 * (sign1<double>(dqf) + sign1<double>(dqb));
    lvc += lvp + lvm;
    const double lslope = 0.5 * (lvp - lvm);
    if (s == 0.0) return;
    const double af = fabs(dqf), ab = fabs(dqb);
    // fmin2(|dqf|, |dqb|) picks |dqb| when |dqb| < |dqf|
    if (ab < af) {
        const double l = lslope * s * sign1<double>(dqb) * 2.0 / cb;
        lvc += l;
        lvm1 -= l;
    } else {
        const double l = lslope * s * sign1<double>(dqf) * 2.0 / cf;
        lvp1 += l;
        lvc -= l;
    }
}

}  // namespace cumc
