// Host-side harness of the point numerics and their adjoints (physics.hpp, adjoint.hpp):
// extern "C" wrappers that the CPU tests call through ctypes to compare each adjoint with
// central differences of its primal.  Built by tests/test_adjoint_points.py.
#include "../../cumpic_amd/csrc/adjoint.hpp"

using namespace cumc;

static FaceState<double> fs(const double *a) { return FaceState<double>{a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]}; }
static void put(const FaceState<double> &s, double *a)
{
    a[0] = s.rho; a[1] = s.vn; a[2] = s.vt1; a[3] = s.vt2; a[4] = s.p; a[5] = s.bn; a[6] = s.bt1; a[7] = s.bt2;
}

extern "C" {

void pt_hll(double g, const double *L, const double *R, double *f) { hll_face<1, double>(g, fs(L), fs(R), f); }

int pt_hll_sig(double g, const double *La, const double *Ra)
{
    const FaceState<double> L = fs(La), R = fs(Ra);
    const SideFwd oL = side_fwd(g, L), oR = side_fwd(g, R);
    auto disc = [](const SideFwd &o, const FaceState<double> &s) {
        const double a2 = o.cs * o.cs, b2n = s.bn * s.bn / s.rho, b2 = b2n + s.bt1 * s.bt1 / s.rho + s.bt2 * s.bt2 / s.rho;
        return ((a2 * a2 + b2 * b2 + 2.0 * b2 * a2) - 4.0 * a2 * b2n) < 0.0;
    };
    const double cfsL = signalspeed<double>(oL.cs, L.bn, L.bt1, L.bt2, L.rho);
    const double cfsR = signalspeed<double>(oR.cs, R.bn, R.bt1, R.bt2, R.rho);
    const double sqL = sqrt(L.rho), sqR = sqrt(R.rho);
    const double ubar = (sqL * L.vn + sqR * R.vn) / (sqL + sqR);
    const double cbar = (sqL * cfsL + sqR * cfsR) / (sqL + sqR);
    const double A1 = L.vn - cfsL, A2 = ubar - cbar, B1 = R.vn + cfsR, B2 = ubar + cbar;
    const double sL = fmin2<double>(A1, A2), sR = fmax2<double>(B1, B2);
    const int region = sL >= 0.0 ? 0 : (sR >= 0.0 ? 1 : 2);
    return region | ((A2 < A1) << 2) | ((B2 > B1) << 3) | (disc(oL, L) << 4) | (disc(oR, R) << 5);
}

void pt_hll_adj(double g, const double *L, const double *R, const double *lf, double *lL, double *lR)
{
    FaceState<double> a{0, 0, 0, 0, 0, 0, 0, 0}, b{0, 0, 0, 0, 0, 0, 0, 0};
    hll_face_adj(g, fs(L), fs(R), lf, a, b);
    put(a, lL);
    put(b, lR);
}

double pt_signalspeed(const double *x) { return signalspeed<double>(x[0], x[1], x[2], x[3], x[4]); }

void pt_signalspeed_adj(const double *x, double l, double *g)
{
    g[0] = g[1] = g[2] = g[3] = g[4] = 0.0;
    signalspeed_adj(x[0], x[1], x[2], x[3], x[4], l, g[0], g[1], g[2], g[3], g[4]);
}

// c: a0m a1m a2m a3m a0p a1p a2p a3p ; v: vm1 vc vp1 ; d: dmm1 dmc dmp1
void pt_ppm_edges(const double *c, const double *v, const double *d, double *out)
{
    ppm_edges<double>(c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], v[0], v[1], v[2], d[0], d[1], d[2], out[0], out[1]);
}

int pt_ppm_edges_sig(const double *c, const double *v, const double *d)
{
    const double vc = v[1];
    const double vp0 = vc + c[4] * (v[2] - vc) + c[5] * (v[2] - vc) + c[6] * d[2] + c[7] * d[1];
    const double vm0 = v[0] + c[0] * (vc - v[0]) + c[1] * (vc - v[0]) + c[2] * d[1] + c[3] * d[0];
    if ((vp0 - vc) * (vc - vm0) <= 0.0) return 1;
    double vm = vm0;
    const bool first = (vp0 - vm0) * (vm0 - 3.0 * vc + 2.0 * vp0) < 0.0;
    if (first) vm = 3.0 * vc - 2.0 * vp0;
    const bool second = (vp0 - vm) * (3.0 * vc - 2.0 * vm0 - vp0) < 0.0;
    return 2 | (first << 2) | (second << 3);
}

void pt_ppm_edges_adj(const double *c, const double *v, const double *d, const double *l, double *lv, double *ld)
{
    for (int i = 0; i < 3; i++) lv[i] = ld[i] = 0.0;
    ppm_edges_adj(c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], v[0], v[1], v[2], d[0], d[1], d[2], l[0], l[1], lv[0],
                  lv[1], lv[2], ld[0], ld[1], ld[2]);
}

double pt_ppm_slope(double b0, double b1, const double *v) { return ppm_slope<double>(b0, b1, v[0], v[1], v[2]); }

int pt_ppm_slope_sig(double b0, double b1, const double *v)
{
    if (!((v[2] - v[1]) * (v[1] - v[0]) > 0.0)) return 0;
    const double delta = b0 * (v[2] - v[1]) + b1 * (v[1] - v[0]);
    const double a = fabs(delta), b = 2.0 * fabs(v[1] - v[0]), c = 2.0 * fabs(v[1] - v[2]);
    int pick = 0;
    double m = a;
    if (b < m) { m = b; pick = 1; }
    if (c < m) pick = 2;
    return 1 | (pick << 1) | ((delta < 0.0) << 3);
}

void pt_ppm_slope_adj(double b0, double b1, const double *v, double l, double *lv)
{
    lv[0] = lv[1] = lv[2] = 0.0;
    ppm_slope_adj(b0, b1, v[0], v[1], v[2], l, lv[0], lv[1], lv[2]);
}

void pt_tvd(double cf, double cb, const double *v, double *out) { tvd_edges<double>(cf, cb, v[0], v[1], v[2], out[0], out[1]); }

void pt_tvd_adj(double cf, double cb, const double *v, const double *l, double *lv)
{
    lv[0] = lv[1] = lv[2] = 0.0;
    tvd_edges_adj(cf, cb, v[0], v[1], v[2], l[0], l[1], lv[0], lv[1], lv[2]);
}

}  // extern "C"
