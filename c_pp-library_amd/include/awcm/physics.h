// FP64 ideal-gas Euler physics shared by the HIP kernels and the host-side
// header API: equation of state, physical fluxes, spectral radii, the Roe
// average with its left-eigenvector projection, and the WENO-Z5 Roe-
// characteristic Lax-Friedrichs flux.
//
// Every expression keeps the reference's evaluation order, literal by literal
// (IDEAL_GAS_EQUATIONS.h:17-109, EULER_EQUATION.h:16-143, ROESOLVER.h:9-110,
// WENOZ.h:10-320), so that a build without FMA contraction
// (-ffp-contract=off) reproduces the reference's FP64 results bit for bit.
// Notable literals: the EOS uses (gamma-1) (0.3999999999999999 for 1.4) while
// the Roe projection hard-codes 0.4, the 1-D projection multiplies by
// 0.4/(c*c) and the 2-D one by 0.4/c/c; ABS and MAX are the ternaries of
// MACROS.h:3,6.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define AWCM_HD __host__ __device__ __forceinline__
#else
#define AWCM_HD inline
#endif

namespace awcm {

AWCM_HD double t_abs(double x) { return x < 0 ? -x : x; }
AWCM_HD double t_max(double a, double b) { return a > b ? a : b; }

// ---- quotients that share a divisor -------------------------------------------
// Div(b)(a) is a / b, correctly rounded.  The host divides.  On the device a
// divisor shared by several quotients gets one reciprocal y = RN(1/b): the
// hardware reciprocal, two Newton steps, then whichever of the result and its
// neighbour toward 1/b leaves the smaller exact residual 1 - b*y.  Each quotient
// is q = a*y corrected once with the exact FMA residual, q + (a - b*q)*y
// (P. Markstein, IBM J. Res. Develop. 34(1), 1990, the final step of his
// division): with y = RN(1/b) that is the correctly rounded a/b for operands
// and quotients in the normal range (checked over every pair of significands
// of 4- to 12-bit binary formats by tests/native/markstein_exhaustive.c).  The
// quotient takes the sign of a*y, which also gives a zero quotient the sign
// IEEE division gives it.  Outside that range the result is not IEEE's: a
// divisor of 0, inf or NaN gives NaN quotients where IEEE gives inf, 0 or NaN
// (the state is no longer finite in either case), and subnormal divisors or
// quotients are not rounded correctly; no physical state of these problems
// comes near either.
struct Div {
    double b;
#if defined(__HIP_DEVICE_COMPILE__)
    double y;
    __device__ __forceinline__ explicit Div(double b_) : b(b_) {
        double r = __builtin_amdgcn_rcp(b);
        double e = __builtin_fma(-b, r, 1.0);
        r = __builtin_fma(r, e, r);
        e = __builtin_fma(-b, r, 1.0);
        r = __builtin_fma(r, e, r);
        e = __builtin_fma(-b, r, 1.0);  // exact: 1/b = r + e/b
        // the neighbour toward 1/b is one bit pattern up when e > 0, for either sign
        const double n = __longlong_as_double(__double_as_longlong(r) + (e > 0 ? 1 : -1));
        const double en = __builtin_fma(-b, n, 1.0);
        y = __builtin_fabs(en) < __builtin_fabs(e) ? n : r;
    }
    __device__ __forceinline__ double operator()(double a) const {
        const double q = a * y;
        const double q1 = __builtin_fma(__builtin_fma(-b, q, a), y, q);
        return __builtin_copysign(q1, q);
    }
    // a / b for a dividend that is never zero (no sign of zero to carry over)
    __device__ __forceinline__ double nz(double a) const {
        const double q = a * y;
        return __builtin_fma(__builtin_fma(-b, q, a), y, q);
    }
    // division by 2b: RN(1/(2b)) is RN(1/b)/2 exactly
    __device__ __forceinline__ Div twice() const { Div d(*this); d.b = 2 * b; d.y = 0.5 * y; return d; }
#else
    AWCM_HD explicit Div(double b_) : b(b_) {}
    AWCM_HD double operator()(double a) const { return a / b; }
    AWCM_HD double nz(double a) const { return a / b; }
    AWCM_HD Div twice() const { return Div(2 * b); }
#endif
};

// ---- equation of state ------------------------------------------------------
// d0 divides by the density q[0]
AWCM_HD double eos_p1(double g, const double* q, const Div& d0) { return (g - 1) * (q[2] - d0(0.5 * (q[1] * q[1]))); }
AWCM_HD double eos_p1(double g, const double* q) { return eos_p1(g, q, Div(q[0])); }
AWCM_HD double eos_H1(double g, const double* q, const Div& d0) { return d0(q[2] + eos_p1(g, q, d0)); }
AWCM_HD double eos_H1(double g, const double* q) { return eos_H1(g, q, Div(q[0])); }
AWCM_HD double eos_c1(double g, const double* q, const Div& d0) { return sqrt(d0(g * eos_p1(g, q, d0))); }
AWCM_HD double eos_c1(double g, const double* q) { return eos_c1(g, q, Div(q[0])); }
AWCM_HD double eos_E1(double g, double rho, double u, double p) { return Div(g - 1)(p) + 0.5 * rho * (u * u); }

// d0 divides by the density q[0]
AWCM_HD double eos_p2(double g, const double* q, const Div& d0) {
    return (g - 1) * (q[3] - d0(0.5 * ((q[1] * q[1]) + (q[2] * q[2]))));
}
AWCM_HD double eos_p2(double g, const double* q) { return eos_p2(g, q, Div(q[0])); }
AWCM_HD double eos_H2(double g, const double* q, const Div& d0) { return d0(q[3] + eos_p2(g, q, d0)); }
AWCM_HD double eos_H2(double g, const double* q) { return eos_H2(g, q, Div(q[0])); }
AWCM_HD double eos_c2(double g, const double* q, const Div& d0) { return sqrt(d0(g * eos_p2(g, q, d0))); }
AWCM_HD double eos_c2(double g, const double* q) { return eos_c2(g, q, Div(q[0])); }
AWCM_HD double eos_E2(double g, double rho, double u, double v, double p) {
    return Div(g - 1)(p) + 0.5 * rho * (u * u + v * v);
}

// ---- physical fluxes and spectral radii --------------------------------------
AWCM_HD void flux_f1(double g, const double* q, double* f) {
    const Div d0(q[0]);
    const double rho = q[0], u = d0(q[1]), E = q[2], p = eos_p1(g, q, d0);
    f[0] = rho * u;
    f[1] = rho * u * u + p;
    f[2] = (E + p) * u;
}
// dir 0: F (x), dir 1: G (y)
template <int DIR>
AWCM_HD void flux_f2(double g, const double* q, double* f) {
    const Div d0(q[0]);
    const double rho = q[0], u = d0(q[1]), v = d0(q[2]), E = q[3], p = eos_p2(g, q, d0);
    if (DIR == 0) {
        f[0] = rho * u;
        f[1] = rho * u * u + p;
        f[2] = rho * v * u;
        f[3] = (E + p) * u;
    } else {
        f[0] = rho * v;
        f[1] = rho * u * v;
        f[2] = rho * v * v + p;
        f[3] = (E + p) * v;
    }
}
AWCM_HD double spec_rad1(double g, const double* q) {
    const Div d0(q[0]);
    const double u = d0(q[1]);
    const double c = eos_c1(g, q, d0);
    return t_abs(u) + c;
}
AWCM_HD double spec_rad2x(double g, const double* q) {
    const Div d0(q[0]);
    const double u = d0(q[1]);
    const double c = eos_c2(g, q, d0);
    return t_abs(u) + c;
}
AWCM_HD double spec_rad2y(double g, const double* q) {
    const Div d0(q[0]);
    const double v = d0(q[2]);
    const double c = eos_c2(g, q, d0);
    return t_abs(v) + c;
}

// ---- Roe average and left-eigenvector projection ------------------------------
struct Roe1 { double u, H, c; };
AWCM_HD Roe1 roe_avg1(double g, const double* l, const double* r) {
    const double rl = l[0], rr = r[0];
    const Div dl(l[0]), dr(r[0]);
    const double ul = dl(l[1]), ur = dr(r[1]);
    const double Hl = eos_H1(g, l, dl), Hr = eos_H1(g, r, dr);
    const Div den(sqrt(rl) + sqrt(rr));
    Roe1 a;
    a.u = den(sqrt(rl) * ul + sqrt(rr) * ur);
    a.H = den(sqrt(rl) * Hl + sqrt(rr) * Hr);
    a.c = sqrt((g - 1) * (a.H - 0.5 * a.u * a.u));
    return a;
}
AWCM_HD void roe_proj1(const Roe1& a, const double* b, double* x) {
    double x2 = (a.H - a.u * a.u) * b[0] + a.u * b[1] - b[2];
    x2 *= Div(a.c * a.c).nz(0.4);
    double x3 = b[1] + (a.c - a.u) * b[0] - a.c * x2;
    x3 = Div(2 * a.c)(x3);
    x[0] = b[0] - x2 - x3;
    x[1] = x2;
    x[2] = x3;
}

struct Roe2 { double u, v, H, c; };
AWCM_HD Roe2 roe_avg2(double g, const double* l, const double* r) {
    const double rl = l[0], rr = r[0];
    const Div dl(l[0]), dr(r[0]);
    const double Hl = eos_H2(g, l, dl), Hr = eos_H2(g, r, dr);
    const double ul = dl(l[1]), ur = dr(r[1]);
    const double vl = dl(l[2]), vr = dr(r[2]);
    const Div den(sqrt(rl) + sqrt(rr));
    Roe2 a;
    a.u = den(sqrt(rl) * ul + sqrt(rr) * ur);
    a.v = den(sqrt(rl) * vl + sqrt(rr) * vr);
    a.H = den(sqrt(rl) * Hl + sqrt(rr) * Hr);
    a.c = sqrt(0.4 * (a.H - 0.5 * (a.u * a.u + a.v * a.v)));
    return a;
}
template <int DIR>
AWCM_HD void roe_proj2(const Roe2& a, const double* b, double* x) {
    const double u = a.u, v = a.v, H = a.H, c = a.c;
    const Div dc(c), d2c = dc.twice();
    double x3, x4, x2;
    if (DIR == 0) {
        x3 = (H - u * u - v * v) * b[0] + u * b[1] + v * b[2] - b[3];
        x3 *= dc.nz(dc.nz(0.4));
        x2 = b[2] - v * b[0];
        x4 = b[1] + (c - u) * b[0] - c * x3;
    } else {
        x3 = (H - u * u - v * v) * b[0] + v * b[2] + u * b[1] - b[3];
        x3 *= dc.nz(dc.nz(0.4));
        x2 = b[1] - u * b[0];
        x4 = b[2] + (c - v) * b[0] - c * x3;
    }
    x4 = d2c(x4);
    x[0] = b[0] - x3 - x4;
    x[1] = x2;
    x[2] = x3;
    x[3] = x4;
}

// ---- WENO-Z5 (tau5 = |b0-b2|, eps = 1e-40, ideal weights 0.1/0.6/0.3) ---------
// CR: the three quotients tau / (b_k + eps) correctly rounded on the device as
// well (the 1-D path); without it they are the hardware divide (the 2-D face
// kernel, where they cost a quarter of its divide work).
template <bool CR = false>
AWCM_HD double wenoz_left(double ll, double l, double c, double r, double rr) {
    const double t0 = c - 2 * r + rr, s0 = 3 * c - 4 * r + rr;
    const double t1 = l - 2 * c + r, s1 = l - r;
    const double t2 = ll - 2 * l + c, s2 = ll - 4 * l + 3 * c;
    const double b0 = 13.0 / 12.0 * (t0 * t0) + 0.25 * (s0 * s0);
    const double b1 = 13.0 / 12.0 * (t1 * t1) + 0.25 * (s1 * s1);
    const double b2 = 13.0 / 12.0 * (t2 * t2) + 0.25 * (s2 * s2);
    const double tau = t_abs(b0 - b2);
    // tau >= +0: a zero quotient needs no sign from the dividend
    const double a0 = 0.1 * (1 + (CR ? Div(b0 + 1e-40).nz(tau) : tau / (b0 + 1e-40)));
    const double a1 = 0.6 * (1 + (CR ? Div(b1 + 1e-40).nz(tau) : tau / (b1 + 1e-40)));
    const double a2 = 0.3 * (1 + (CR ? Div(b2 + 1e-40).nz(tau) : tau / (b2 + 1e-40)));
    const Div as(a0 + a1 + a2);
    const double w0 = as.nz(a0), w1 = as.nz(a1), w2 = as.nz(a2);  // a_k >= d_k > 0
    double out = w0 * (11.0 / 6.0 * c + -7.0 / 6.0 * r + 1.0 / 3.0 * rr);
    out += w1 * (1.0 / 3.0 * l + 5.0 / 6.0 * c + -1.0 / 6.0 * r);
    out += w2 * (-1.0 / 6.0 * ll + 5.0 / 6.0 * l + 1.0 / 3.0 * c);
    return out;
}
template <bool CR = false>
AWCM_HD double wenoz_right(double ll, double l, double c, double r, double rr) { return wenoz_left<CR>(rr, r, c, l, ll); }

// ---- WENO5-JS (WENO.h:71-99: a_k = d_k / (1e-6 + b_k)^2, ideal weights 0.1/0.6/0.3) --
AWCM_HD double wenojs_left(double ll, double l, double c, double r, double rr) {
    const double t0 = c - 2 * r + rr, s0 = 3 * c - 4 * r + rr;
    const double t1 = l - 2 * c + r, s1 = l - r;
    const double t2 = ll - 2 * l + c, s2 = ll - 4 * l + 3 * c;
    const double b0 = 13.0 / 12.0 * (t0 * t0) + 0.25 * (s0 * s0);
    const double b1 = 13.0 / 12.0 * (t1 * t1) + 0.25 * (s1 * s1);
    const double b2 = 13.0 / 12.0 * (t2 * t2) + 0.25 * (s2 * s2);
    const double e0 = 1e-6 + b0, e1 = 1e-6 + b1, e2 = 1e-6 + b2;
    const double a0 = Div(e0 * e0).nz(0.1), a1 = Div(e1 * e1).nz(0.6), a2 = Div(e2 * e2).nz(0.3);
    const Div as(a0 + a1 + a2);
    const double w0 = as.nz(a0), w1 = as.nz(a1), w2 = as.nz(a2);  // a_k > 0
    double out = w0 * (11.0 / 6.0 * c + -7.0 / 6.0 * r + 1.0 / 3.0 * rr);
    out += w1 * (1.0 / 3.0 * l + 5.0 / 6.0 * c + -1.0 / 6.0 * r);
    out += w2 * (-1.0 / 6.0 * ll + 5.0 / 6.0 * l + 1.0 / 3.0 * c);
    return out;
}
AWCM_HD double wenojs_right(double ll, double l, double c, double r, double rr) { return wenojs_left(rr, r, c, l, ll); }

// reconstruction of the characteristic flux splitting: WENO-Z5 (WENOZ.h, the
// 2-D drivers and the C1 configuration) or WENO5-JS (WENO.h, the shipped 1-D
// driver's ComputedF)
enum { kRecWenoZ = 0, kRecWeno5 = 1 };
template <int REC, bool CR = false>
AWCM_HD double rec_left(double ll, double l, double c, double r, double rr) {
    return REC == kRecWeno5 ? wenojs_left(ll, l, c, r, rr) : wenoz_left<CR>(ll, l, c, r, rr);
}
template <int REC, bool CR = false>
AWCM_HD double rec_right(double ll, double l, double c, double r, double rr) {
    return REC == kRecWeno5 ? wenojs_right(ll, l, c, r, rr) : wenoz_right<CR>(ll, l, c, r, rr);
}

// ---- Roe-characteristic LF-split WENO numerical flux at one face ----------------
// s[p] (p = 0..5) are the states lll, ll, l, r, rr, rrr around the face l|r.
// WENOZ_EULER_SOLVER_1D::ComputeFlux (WENOZ.h:160-205) with REC = kRecWenoZ,
// WENO_EULER_SOLVER_1D::ComputeFlux (WENO.h:129-174) with REC = kRecWeno5.
// Every quotient is correctly rounded on the device (Div), as on the host.
template <int REC = kRecWenoZ>
AWCM_HD void face_flux1(double g, const double (*s)[3], double alpha, double* flux) {
    double fp[6][3], fm[6][3];
    const Roe1 a = roe_avg1(g, s[2], s[3]);
#pragma unroll
    for (int p = 0; p < 6; p++) {
        double fu[3], RQ[3], Rf[3];
        flux_f1(g, s[p], fu);
        roe_proj1(a, s[p], RQ);
        roe_proj1(a, fu, Rf);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            fp[p][k] = 0.5 * (Rf[k] + alpha * RQ[k]);
            fm[p][k] = 0.5 * (Rf[k] - alpha * RQ[k]);
        }
    }
    double sk[3];
#pragma unroll
    for (int k = 0; k < 3; k++)
        sk[k] = rec_right<REC, true>(fp[0][k], fp[1][k], fp[2][k], fp[3][k], fp[4][k]) +
                rec_left<REC, true>(fm[1][k], fm[2][k], fm[3][k], fm[4][k], fm[5][k]);
    const double r1[3] = {1, a.u - a.c, a.H - a.u * a.c};
    const double r2[3] = {1, a.u, 0.5 * (a.u * a.u)};
    const double r3[3] = {1, a.u + a.c, a.H + a.u * a.c};
#pragma unroll
    for (int k = 0; k < 3; k++) flux[k] = sk[0] * r1[k] + sk[1] * r2[k] + sk[2] * r3[k];
}

// WENOZ_EULER_SOLVER_2D::ComputeFluxX/Y (WENOZ.h:236-320) with REC = kRecWenoZ,
// WENO_EULER_SOLVER_2D::ComputeFluxX/Y (WENO.h:216-311) with REC = kRecWeno5.
template <int DIR, int REC = kRecWenoZ>
AWCM_HD void face_flux2(double g, const double (*s)[4], double alpha, double* flux) {
    double fp[6][4], fm[6][4];
    const Roe2 a = roe_avg2(g, s[2], s[3]);
#pragma unroll
    for (int p = 0; p < 6; p++) {
        double fu[4], RQ[4], Rf[4];
        flux_f2<DIR>(g, s[p], fu);
        roe_proj2<DIR>(a, s[p], RQ);
        roe_proj2<DIR>(a, fu, Rf);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            fp[p][k] = 0.5 * (Rf[k] + alpha * RQ[k]);
            fm[p][k] = 0.5 * (Rf[k] - alpha * RQ[k]);
        }
    }
    double sk[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        sk[k] = rec_right<REC>(fp[0][k], fp[1][k], fp[2][k], fp[3][k], fp[4][k]) +
                rec_left<REC>(fm[1][k], fm[2][k], fm[3][k], fm[4][k], fm[5][k]);
    const double u = a.u, v = a.v, H = a.H, c = a.c;
    double r1[4], r2[4], r3[4], r4[4];
    if (DIR == 0) {
        r1[0] = 1;         r2[0] = 0; r3[0] = 1;                   r4[0] = 1;
        r1[1] = u - c;     r2[1] = 0; r3[1] = u;                   r4[1] = u + c;
        r1[2] = v;         r2[2] = 1; r3[2] = v;                   r4[2] = v;
        r1[3] = H - u * c; r2[3] = v; r3[3] = 0.5 * (u * u + v * v); r4[3] = H + u * c;
    } else {
        r1[0] = 1;         r2[0] = 0; r3[0] = 1;                   r4[0] = 1;
        r1[1] = u;         r2[1] = 1; r3[1] = u;                   r4[1] = u;
        r1[2] = v - c;     r2[2] = 0; r3[2] = v;                   r4[2] = v + c;
        r1[3] = H - v * c; r2[3] = u; r3[3] = 0.5 * (u * u + v * v); r4[3] = H + v * c;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) flux[k] = sk[0] * r1[k] + sk[1] * r2[k] + sk[2] * r3[k] + sk[3] * r4[k];
}

// ---- coarse-level central differences (AWCM_1D.h:21,36-39; AWCM_2D.h:24,43-46) --
constexpr double kD1 = 2.0 / 3.0, kD2 = -1.0 / 12.0;
constexpr double kD01 = -1.0 / 3.0, kD11 = -1.0 / 2.0, kD21 = 1.0, kD31 = -1.0 / 6.0;
constexpr double kD02 = 1.0 / 12.0, kD12 = -2.0 / 3.0, kD22 = 0.0, kD32 = 2.0 / 3.0;

// ---- interpolating-wavelet weights (WAVELET_1D.h:26-42) -------------------------
// weight(i, j) = prod_{k != i} (j + 0.5 - k) / (i - k), i = 0..3, j = -1..3,
// stored as w[i][j + 1]; computed with the same loop so the products round alike.
AWCM_HD void wavelet_weights(double (*w)[5]) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 5; j++) w[i][j] = 1;
    for (int j = -1; j < 4; j++)
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 4; k++) {
                if (k == i) continue;
                w[i][j + 1] *= double(j + 0.5 - k) / double(i - k);
            }
}

}  // namespace awcm
