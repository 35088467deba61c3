// Reverse-mode (vector-Jacobian) products of the 1-D Euler physics in
// awcm/physics.h, for the discrete adjoint of the AWCM step with frozen masks,
// time step and Lax-Friedrichs alpha.
//
// Each *_vjp function differentiates the function of the same name in
// physics.h as a mathematical function of its state arguments (alpha, gamma
// and the Roe literal 0.4 are constants), recomputing whatever forward
// intermediates it needs, and ADDS the products into the output arrays.  The
// adjoint is not bit-matched to anything; it is checked against central
// finite differences of the frozen forward map.
//
// |b0 - b2| in WENO-Z is differentiated with the sign of b0 - b2 and with the
// subgradient 0 where b0 == b2.  Where a stencil is uniform (all smoothness
// indicators 0) the WENO-Z weights are homogeneous of degree 0 in the stencil
// differences, so the map is only directionally differentiable there; the
// product below is the derivative along directions that keep the stencil
// uniform (linear weights), and stencils uniform up to rounding are treated
// as uniform.
#pragma once

#include "physics.h"

namespace awcm {

// f = flux_f1(q): qbar += (df/dq)^T fbar
AWCM_HD void flux_f1_vjp(double g, const double* q, const double* fbar, double* qbar) {
    const double rho = q[0], u = q[1] / q[0], E = q[2];
    const double p = eos_p1(g, q);
    // f0 = rho u, f1 = rho u u + p, f2 = (E + p) u
    const double ub = fbar[0] * rho + fbar[1] * 2.0 * rho * u + fbar[2] * (E + p);
    const double rb = fbar[0] * u + fbar[1] * u * u;
    const double Eb = fbar[2] * u;
    const double pb = fbar[1] + fbar[2] * u;
    // p = (g-1) (q2 - 0.5 q1^2 / q0), u = q1 / q0
    qbar[0] += rb + pb * (g - 1) * 0.5 * u * u - ub * u / rho;
    qbar[1] += -pb * (g - 1) * u + ub / rho;
    qbar[2] += Eb + pb * (g - 1);
}

// out = wenoz_left(x0..x4) (x = ll, l, c, r, rr): xbar[k] += d out / d x_k * obar.
// The smoothness terms and the candidate values are formed from the first
// differences of the stencil (algebraically the same function), so that an
// exactly uniform stencil gives exact zeros, and with them the linear-weight
// derivative, instead of rounding residues of 3c - 4r + rr that the 1e-40
// epsilon would turn into weight derivatives of order 1e30.
AWCM_HD void wenoz_left_vjp(const double* x, double obar, double* xbar) {
    double da = x[1] - x[0], db = x[2] - x[1], dc = x[3] - x[2], dd = x[4] - x[3];
    // A stencil uniform up to rounding (all differences within 64 ulp of its
    // scale) is treated as exactly uniform: its smoothness indicators carry
    // only rounding noise, whose weight derivatives are meaningless and, with
    // the 1e-40 epsilon, unbounded.
    const double scale = t_max(t_max(t_max(t_abs(x[0]), t_abs(x[1])), t_max(t_abs(x[2]), t_abs(x[3]))), t_abs(x[4]));
    const double dmax = t_max(t_max(t_abs(da), t_abs(db)), t_max(t_abs(dc), t_abs(dd)));
    if (dmax <= 64 * 2.220446049250313e-16 * scale) da = db = dc = dd = 0.0;
    const double t0 = dd - dc, s0 = dd - 3 * dc;
    const double t1 = dc - db, s1 = -db - dc;
    const double t2 = db - da, s2 = 3 * db - da;
    const double b0 = 13.0 / 12.0 * (t0 * t0) + 0.25 * (s0 * s0);
    const double b1 = 13.0 / 12.0 * (t1 * t1) + 0.25 * (s1 * s1);
    const double b2 = 13.0 / 12.0 * (t2 * t2) + 0.25 * (s2 * s2);
    const double diff = b0 - b2;
    const double tau = t_abs(diff);
    // subgradient 0 of |b0 - b2| on its kink
    const double sg = diff < 0 ? -1.0 : diff > 0 ? 1.0 : 0.0;
    const double e0 = b0 + 1e-40, e1 = b1 + 1e-40, e2 = b2 + 1e-40;
    const double a0 = 0.1 * (1 + tau / e0), a1 = 0.6 * (1 + tau / e1), a2 = 0.3 * (1 + tau / e2);
    const double as = a0 + a1 + a2;
    const double w0 = a0 / as, w1 = a1 / as, w2 = a2 / as;
    // candidate values minus the centre value c
    const double q0 = -5.0 / 6.0 * dc + 1.0 / 3.0 * dd;
    const double q1 = -1.0 / 3.0 * db - 1.0 / 6.0 * dc;
    const double q2 = 1.0 / 6.0 * da - 2.0 / 3.0 * db;
    // out = c + sum w_k q_k (sum w_k = 1)
    const double qb0 = obar * w0, qb1 = obar * w1, qb2 = obar * w2;
    const double sw = w0 * q0 + w1 * q1 + w2 * q2;
    // w_k = a_k / as
    const double ab0 = obar * (q0 - sw) / as, ab1 = obar * (q1 - sw) / as, ab2 = obar * (q2 - sw) / as;
    // a_k = d_k (1 + tau / e_k)
    const double taub = ab0 * 0.1 / e0 + ab1 * 0.6 / e1 + ab2 * 0.3 / e2;
    double bb0 = -ab0 * 0.1 * tau / (e0 * e0);
    const double bb1 = -ab1 * 0.6 * tau / (e1 * e1);
    double bb2 = -ab2 * 0.3 * tau / (e2 * e2);
    bb0 += taub * sg;
    bb2 -= taub * sg;
    // b_k = 13/12 t_k^2 + 1/4 s_k^2
    const double tb0 = bb0 * (13.0 / 6.0) * t0, sb0 = bb0 * 0.5 * s0;
    const double tb1 = bb1 * (13.0 / 6.0) * t1, sb1 = bb1 * 0.5 * s1;
    const double tb2 = bb2 * (13.0 / 6.0) * t2, sb2 = bb2 * 0.5 * s2;
    // back to the differences, then to the stencil values
    const double dab = -tb2 - sb2 + 1.0 / 6.0 * qb2;
    const double dbb = -tb1 - sb1 + tb2 + 3 * sb2 - 1.0 / 3.0 * qb1 - 2.0 / 3.0 * qb2;
    const double dcb = -tb0 - 3 * sb0 + tb1 - sb1 - 5.0 / 6.0 * qb0 - 1.0 / 6.0 * qb1;
    const double ddb = tb0 + sb0 + 1.0 / 3.0 * qb0;
    xbar[0] += -dab;
    xbar[1] += dab - dbb;
    xbar[2] += dbb - dcb + obar;
    xbar[3] += dcb - ddb;
    xbar[4] += ddb;
}

// x = roe_proj1(a, b): bbar += (dx/db)^T xbar, (u,H,c)bar += (dx/d(u,H,c))^T xbar
AWCM_HD void roe_proj1_vjp(const Roe1& a, const double* b, const double* xbar, double* bbar, double* ub, double* Hb,
                           double* cb) {
    const double u = a.u, H = a.H, c = a.c;
    const double kap = 0.4 / (c * c);
    const double M = (H - u * u) * b[0] + u * b[1] - b[2];
    const double x2 = M * kap;
    const double N = b[1] + (c - u) * b[0] - c * x2;
    const double x3 = N / (2 * c);
    // x0 = b0 - x2 - x3, x1 = x2, x2' = x3
    bbar[0] += xbar[0];
    double x2b = xbar[1] - xbar[0];
    const double x3b = xbar[2] - xbar[0];
    // x3 = N / (2c)
    const double Nb = x3b / (2 * c);
    *cb += -x3b * x3 / c;
    bbar[1] += Nb;
    bbar[0] += Nb * (c - u);
    *cb += Nb * (b[0] - x2);
    *ub += -Nb * b[0];
    x2b += -Nb * c;
    // x2 = kap M, kap = 0.4 / c^2
    const double Mb = x2b * kap;
    *cb += -2.0 * x2b * M * kap / c;
    bbar[0] += Mb * (H - u * u);
    bbar[1] += Mb * u;
    bbar[2] -= Mb;
    *Hb += Mb * b[0];
    *ub += Mb * (b[1] - 2.0 * u * b[0]);
}

// (u,H,c) = roe_avg1(l, r): lbar, rbar += products
AWCM_HD void roe_avg1_vjp(double g, const double* l, const double* r, double ub, double Hb, double cb, double* lbar,
                          double* rbar) {
    const double rl = l[0], rr = r[0];
    const double ul = l[1] / l[0], ur = r[1] / r[0];
    const double Hl = eos_H1(g, l), Hr = eos_H1(g, r);
    const double sl = sqrt(rl), sr = sqrt(rr);
    const double den = sl + sr;
    const double u = (sl * ul + sr * ur) / den;
    const double H = (sl * Hl + sr * Hr) / den;
    const double c = sqrt((g - 1) * (H - 0.5 * u * u));
    // c = sqrt((g-1)(H - u^2/2))
    const double zb = cb / (2 * c);
    Hb += zb * (g - 1);
    ub += -zb * (g - 1) * u;
    // u = (sl ul + sr ur) / den, H = (sl Hl + sr Hr) / den
    const double denb = -(ub * u + Hb * H) / den;
    const double slb = (ub * ul + Hb * Hl) / den + denb;
    const double srb = (ub * ur + Hb * Hr) / den + denb;
    const double ulb = ub * sl / den, urb = ub * sr / den;
    const double Hlb = Hb * sl / den, Hrb = Hb * sr / den;
    // side terms: rho via sqrt, u = m / rho, H = (E + p) / rho
    const double* q[2] = {l, r};
    double* qb[2] = {lbar, rbar};
    const double sb[2] = {slb, srb}, sq[2] = {sl, sr}, vb[2] = {ulb, urb}, hb[2] = {Hlb, Hrb};
    const double uu[2] = {ul, ur}, hh[2] = {Hl, Hr};
    for (int s = 0; s < 2; s++) {
        const double rho = q[s][0];
        double r0 = sb[s] / (2 * sq[s]);
        // u = m / rho
        qb[s][1] += vb[s] / rho;
        r0 += -vb[s] * uu[s] / rho;
        // H = (E + p) / rho, p = (g-1)(E - 0.5 m^2 / rho)
        const double pb = hb[s] / rho;
        qb[s][2] += hb[s] / rho + pb * (g - 1);
        r0 += -hb[s] * hh[s] / rho;
        qb[s][1] += -pb * (g - 1) * uu[s];
        r0 += pb * (g - 1) * 0.5 * uu[s] * uu[s];
        qb[s][0] += r0;
    }
}

// flux = face_flux1(s, alpha): sbar[p][k] += (d flux / d s[p][k])^T fbar
AWCM_HD void face_flux1_vjp(double g, const double (*s)[3], double alpha, const double* fbar, double (*sbar)[3]) {
    const Roe1 a = roe_avg1(g, s[2], s[3]);
    double fp[6][3], fm[6][3];
    for (int p = 0; p < 6; p++) {
        double fu[3], RQ[3], Rf[3];
        flux_f1(g, s[p], fu);
        roe_proj1(a, s[p], RQ);
        roe_proj1(a, fu, Rf);
        for (int k = 0; k < 3; k++) {
            fp[p][k] = 0.5 * (Rf[k] + alpha * RQ[k]);
            fm[p][k] = 0.5 * (Rf[k] - alpha * RQ[k]);
        }
    }
    double sk[3];
    for (int k = 0; k < 3; k++)
        sk[k] = wenoz_right<true>(fp[0][k], fp[1][k], fp[2][k], fp[3][k], fp[4][k]) +
                wenoz_left<true>(fm[1][k], fm[2][k], fm[3][k], fm[4][k], fm[5][k]);
    const double u = a.u, H = a.H, c = a.c;
    // flux = sk0 (1, u-c, H-uc) + sk1 (1, u, u^2/2) + sk2 (1, u+c, H+uc)
    double skb[3];
    skb[0] = fbar[0] + fbar[1] * (u - c) + fbar[2] * (H - u * c);
    skb[1] = fbar[0] + fbar[1] * u + fbar[2] * (0.5 * (u * u));
    skb[2] = fbar[0] + fbar[1] * (u + c) + fbar[2] * (H + u * c);
    double ub = sk[0] * (fbar[1] - fbar[2] * c) + sk[1] * (fbar[1] + fbar[2] * u) + sk[2] * (fbar[1] + fbar[2] * c);
    double Hb = (sk[0] + sk[2]) * fbar[2];
    double cb = (sk[2] - sk[0]) * (fbar[1] + fbar[2] * u);
    // WENO-Z: wenoz_right(x0..x4) = wenoz_left(x4..x0)
    double fpb[6][3] = {}, fmb[6][3] = {};
    for (int k = 0; k < 3; k++) {
        const double xr[5] = {fp[4][k], fp[3][k], fp[2][k], fp[1][k], fp[0][k]};
        double gr[5] = {0, 0, 0, 0, 0};
        wenoz_left_vjp(xr, skb[k], gr);
        for (int p = 0; p < 5; p++) fpb[p][k] += gr[4 - p];
        const double xl[5] = {fm[1][k], fm[2][k], fm[3][k], fm[4][k], fm[5][k]};
        double gl[5] = {0, 0, 0, 0, 0};
        wenoz_left_vjp(xl, skb[k], gl);
        for (int p = 0; p < 5; p++) fmb[p + 1][k] += gl[p];
    }
    // fp = (Rf + alpha RQ)/2, fm = (Rf - alpha RQ)/2, RQ = L s, Rf = L F(s)
    for (int p = 0; p < 6; p++) {
        double Rfb[3], RQb[3];
        for (int k = 0; k < 3; k++) {
            Rfb[k] = 0.5 * (fpb[p][k] + fmb[p][k]);
            RQb[k] = 0.5 * alpha * (fpb[p][k] - fmb[p][k]);
        }
        double fu[3], fub[3] = {0, 0, 0};
        flux_f1(g, s[p], fu);
        roe_proj1_vjp(a, s[p], RQb, sbar[p], &ub, &Hb, &cb);
        roe_proj1_vjp(a, fu, Rfb, fub, &ub, &Hb, &cb);
        flux_f1_vjp(g, s[p], fub, sbar[p]);
    }
    roe_avg1_vjp(g, s[2], s[3], ub, Hb, cb, sbar[2], sbar[3]);
}

}  // namespace awcm
