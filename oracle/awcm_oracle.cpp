// CPU ORACLE FOR THE AWCM COMPRESSIBLE-EULER PATH -- TEST INFRASTRUCTURE ONLY.
//
// A plain C++ restatement of the reference solver (doseongju/C_pp-library),
// written to follow the reference's FP64 operation order term by term so that
// a build with -ffp-contract=off reproduces its results bit for bit.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
// library, and only as the checker / the timed CPU baseline.  The product
// (c_pp-library_amd/) never links, loads or calls anything in oracle/.
//
// Pinning (see tests/test_oracle.py):
//   * every exported entry point is compared bit-exactly with oracle/_ref, a
//     build of the reference's own headers driven by oracle/ref_harness.cpp,
//     on 1-D and 2-D trajectories (masks, levels, fields, dt, t);
//   * committed golden vectors (tests/golden/) made from that build;
//   * closed forms and invariants: dyadic wavelet weights, WENO-Z ideal-weight
//     limit (2,-13,47,27,-3)/60, flux consistency F(Q,..,Q)=F(Q) (WENO-Z and
//     WENO5-JS), Roe eigenvector inversion R*L=I, nested masks, band always
//     active, eps=0 reduces to the uniform-grid WENO-Z solver, the
//     interpolations reproduce cubics / bicubics from the coarsest grid and
//     leave mask_from's points alone, the scalar path equals the Euler path
//     on component 0.
// Entry points beyond the time step: PrepareEuler, the scalar Prepare (1-D
// and 2-D), Interpolate / InterpolateEuler (both dimensions, both forms),
// InterpolateAllGridEuler, and the WENO5-JS stage flux of the shipped 1-D
// driver; each is pinned by the same three means.
//
// Layouts (same as the reference API):
//   1-D state  AoS, ghost 3:  q[(i+3)*3 + c],             i = -3..Nx+2
//   2-D state  AoS, ghost 3:  q[((i+3)+(Nx+6)*(j+3))*4+c]
//   masks      uint8, ghost 3 (same linear index without the component)
//   levels     int32, no ghost: lvl[i] (1-D), lvl[i + Nx*j] (2-D)

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

constexpr int GW = 3;  // ghost width of every ghosted field (AWCM_1D.h:23-27)

// MACROS.h:3,6 -- ternary ABS / MAX (NaN and -0.0 semantics matter)
inline double ABSd(double x) { return x < 0 ? -x : x; }
inline double SQRd(double x) { return x * x; }  // SQR (MACROS.h)
inline double MAXd(double a, double b) { return a > b ? a : b; }

// ---------------------------------------------------------------------------
// Integer MRA geometry: GRID_INT_1D / GRID_INT_2D (MRA_GRID_UNIFORM_1D.h:9-45,
// MRA_GRID_UNIFORM_2D.h:10-60).  dx is an integer division, as there.
struct GridI1 { int s, e, res, d; };
GridI1 mk1(int s, int e, int res) { return GridI1{s, e, res, (e - s) / (res - 1)}; }

struct GridI2 { int is, js, ie, je, ires, jres, dx, dy; };
GridI2 mk2(int is, int js, int ie, int je, int ires, int jres) {
    return GridI2{is, js, ie, je, ires, jres, (ie - is) / (ires - 1), (je - js) / (jres - 1)};
}

// WAVELET_1D.h:26-42 / WAVELET_2D.h:17-33: weight(i, j) = prod_{k != i} (j+0.5-k)/(i-k),
// i = 0..3, j = -1..3; stored as W[i][j+1].
struct Weights {
    double W[4][5];
    Weights() {
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 5; j++) W[i][j] = 1;
        for (int j = -1; j < 4; j++)
            for (int i = 0; i < 4; i++)
                for (int k = 0; k < 4; k++) {
                    if (k == i) continue;
                    W[i][j + 1] *= double(j + 0.5 - k) / double(i - k);
                }
    }
    double operator()(int i, int j) const { return W[i][j + 1]; }
};

// ---------------------------------------------------------------------------
// Physics: IDEAL_GAS_EQUATIONS.h:17-57,69-109; EULER_EQUATION.h:16-26,74-99;
// ROESOLVER.h:9-110; WENOZ.h:10-136.

inline double p1(double g, const double* q) { return (g - 1) * (q[2] - 0.5 * (q[1] * q[1]) / q[0]); }
inline double H1(double g, const double* q) { return (q[2] + p1(g, q)) / q[0]; }
inline double c1(double g, const double* q) { return std::sqrt(g * p1(g, q) / q[0]); }

inline double p2(double g, const double* q) {
    return (g - 1) * (q[3] - 0.5 * ((q[1] * q[1]) + (q[2] * q[2])) / q[0]);
}
inline double H2(double g, const double* q) { return (q[3] + p2(g, q)) / q[0]; }
inline double c2(double g, const double* q) { return std::sqrt(g * p2(g, q) / q[0]); }

inline double E1prim(double g, double rho, double u, double p) { return p / (g - 1) + 0.5 * rho * (u * u); }
inline double E2prim(double g, double rho, double u, double v, double p) {
    return p / (g - 1) + 0.5 * rho * (u * u + v * v);
}

inline void Fu1(double g, const double* Q, double* f) {
    double rho = Q[0], u = Q[1] / Q[0], E = Q[2], p = p1(g, Q);
    f[0] = rho * u;
    f[1] = rho * u * u + p;
    f[2] = (E + p) * u;
}
inline void Fu2(double g, const double* Q, double* f) {
    double rho = Q[0], u = Q[1] / Q[0], v = Q[2] / Q[0], E = Q[3], p = p2(g, Q);
    f[0] = rho * u;
    f[1] = rho * u * u + p;
    f[2] = rho * v * u;
    f[3] = (E + p) * u;
}
inline void Gu2(double g, const double* Q, double* f) {
    double rho = Q[0], u = Q[1] / Q[0], v = Q[2] / Q[0], E = Q[3], p = p2(g, Q);
    f[0] = rho * v;
    f[1] = rho * u * v;
    f[2] = rho * v * v + p;
    f[3] = (E + p) * v;
}
inline double SR1(double g, const double* q) { double u = q[1] / q[0]; double c = c1(g, q); return ABSd(u) + c; }
inline double SRx2(double g, const double* q) { double u = q[1] / q[0]; double c = c2(g, q); return ABSd(u) + c; }
inline double SRy2(double g, const double* q) { double v = q[2] / q[0]; double c = c2(g, q); return ABSd(v) + c; }

inline void roe1(double g, const double* l, const double* r, double& u, double& H, double& c) {
    double rl = l[0], rr = r[0];
    double ul = l[1] / l[0], ur = r[1] / r[0];
    double Hl = H1(g, l), Hr = H1(g, r);
    double den = std::sqrt(rl) + std::sqrt(rr);
    u = std::sqrt(rl) * ul + std::sqrt(rr) * ur;
    u /= den;
    H = std::sqrt(rl) * Hl + std::sqrt(rr) * Hr;
    H /= den;
    c = std::sqrt((g - 1) * (H - 0.5 * u * u));
}
inline void inv1(double u, double H, double c, const double* b, double* x) {
    double x1, x2, x3;
    x2 = (H - u * u) * b[0] + u * b[1] - b[2];
    x2 *= 0.4 / (c * c);
    x3 = b[1] + (c - u) * b[0] - c * x2;
    x3 /= 2 * c;
    x1 = b[0] - x2 - x3;
    x[0] = x1; x[1] = x2; x[2] = x3;
}
inline void roe2(double g, const double* l, const double* r, double& u, double& v, double& H, double& c) {
    double rl = l[0], rr = r[0];
    double Hl = H2(g, l), Hr = H2(g, r);
    double ul = l[1] / l[0], ur = r[1] / r[0];
    double vl = l[2] / l[0], vr = r[2] / r[0];
    double den = std::sqrt(rl) + std::sqrt(rr);
    u = std::sqrt(rl) * ul + std::sqrt(rr) * ur;
    u /= den;
    v = std::sqrt(rl) * vl + std::sqrt(rr) * vr;
    v /= den;
    H = std::sqrt(rl) * Hl + std::sqrt(rr) * Hr;
    H /= den;
    c = std::sqrt(0.4 * (H - 0.5 * (u * u + v * v)));
}
inline void invx2(double u, double v, double H, double c, const double* b, double* x) {
    double x1, x2, x3, x4;
    x3 = (H - u * u - v * v) * b[0] + u * b[1] + v * b[2] - b[3];
    x3 *= 0.4 / c / c;
    x2 = b[2] - v * b[0];
    x4 = b[1] + (c - u) * b[0] - c * x3;
    x4 /= 2 * c;
    x1 = b[0] - x3 - x4;
    x[0] = x1; x[1] = x2; x[2] = x3; x[3] = x4;
}
inline void invy2(double u, double v, double H, double c, const double* b, double* x) {
    double x1, x2, x3, x4;
    x3 = (H - u * u - v * v) * b[0] + v * b[2] + u * b[1] - b[3];
    x3 *= 0.4 / c / c;
    x2 = b[1] - u * b[0];
    x4 = b[2] + (c - v) * b[0] - c * x3;
    x4 /= 2 * c;
    x1 = b[0] - x3 - x4;
    x[0] = x1; x[1] = x2; x[2] = x3; x[3] = x4;
}

// WENOZ5::ReconstructLeft (WENOZ.h:75-104); Right(a..e) = Left(e..a) (WENOZ.h:106-108)
const double WZ_EPS = 1e-40;
const double CM1[3] = {11.0 / 6.0, -7.0 / 6.0, 1.0 / 3.0};
const double C0[3] = {1.0 / 3.0, 5.0 / 6.0, -1.0 / 6.0};
const double C1[3] = {-1.0 / 6.0, 5.0 / 6.0, 1.0 / 3.0};

inline double wenoz_left(double ll, double l, double c, double r, double rr) {
    double d0 = 0.1, d1 = 0.6, d2 = 0.3;
    double t0 = c - 2 * r + rr, s0 = 3 * c - 4 * r + rr;
    double t1 = l - 2 * c + r, s1 = l - r;
    double t2 = ll - 2 * l + c, s2 = ll - 4 * l + 3 * c;
    double b0 = 13.0 / 12.0 * (t0 * t0) + 0.25 * (s0 * s0);
    double b1 = 13.0 / 12.0 * (t1 * t1) + 0.25 * (s1 * s1);
    double b2 = 13.0 / 12.0 * (t2 * t2) + 0.25 * (s2 * s2);
    double tau = ABSd(b0 - b2);
    double a0 = d0 * (1 + tau / (b0 + WZ_EPS));
    double a1 = d1 * (1 + tau / (b1 + WZ_EPS));
    double a2 = d2 * (1 + tau / (b2 + WZ_EPS));
    double as = a0 + a1 + a2;
    double w0 = a0 / as, w1 = a1 / as, w2 = a2 / as;
    double left = w0 * (CM1[0] * c + CM1[1] * r + CM1[2] * rr);
    left += w1 * (C0[0] * l + C0[1] * c + C0[2] * r);
    left += w2 * (C1[0] * ll + C1[1] * l + C1[2] * c);
    return left;
}
inline double wenoz_right(double ll, double l, double c, double r, double rr) { return wenoz_left(rr, r, c, l, ll); }

// WENO5::ReconstructLeft (WENO.h:71-99): a_k = d_k / SQR(eps + b_k), eps = 1e-6;
// Right(a..e) = Left(e..a) (WENO.h:101-104)
const double JS_EPS = 1e-6;
inline double wenojs_left(double ll, double l, double c, double r, double rr) {
    double d0 = 0.1, d1 = 0.6, d2 = 0.3;
    double b0 = 13.0 / 12.0 * SQRd(c - 2 * r + rr) + 0.25 * SQRd(3 * c - 4 * r + rr);
    double b1 = 13.0 / 12.0 * SQRd(l - 2 * c + r) + 0.25 * SQRd(l - r);
    double b2 = 13.0 / 12.0 * SQRd(ll - 2 * l + c) + 0.25 * SQRd(ll - 4 * l + 3 * c);
    double a0 = d0 / SQRd(JS_EPS + b0);
    double a1 = d1 / SQRd(JS_EPS + b1);
    double a2 = d2 / SQRd(JS_EPS + b2);
    double as = a0 + a1 + a2;
    double w0 = a0 / as, w1 = a1 / as, w2 = a2 / as;
    double left = w0 * (CM1[0] * c + CM1[1] * r + CM1[2] * rr);
    left += w1 * (C0[0] * l + C0[1] * c + C0[2] * r);
    left += w2 * (C1[0] * ll + C1[1] * l + C1[2] * c);
    return left;
}
inline double wenojs_right(double ll, double l, double c, double r, double rr) { return wenojs_left(rr, r, c, l, ll); }

// WENOZ_EULER_SOLVER_1D::ComputeFlux (WENOZ.h:160-205), or with js
// WENO_EULER_SOLVER_1D::ComputeFlux (WENO.h:129-174); s[0..5] = lll..rrr
void flux1(double g, const double* const* s, double alpha, double* flux, bool js = false) {
    double fu[6][3], RQ[6][3], Rfu[6][3], fp[6][3], fm[6][3];
    for (int p = 0; p < 6; p++) Fu1(g, s[p], fu[p]);
    double u, H, c;
    roe1(g, s[2], s[3], u, H, c);
    double r1[3] = {1, u - c, H - u * c};
    double r2[3] = {1, u, 0.5 * (u * u)};
    double r3[3] = {1, u + c, H + u * c};
    for (int p = 0; p < 6; p++) inv1(u, H, c, s[p], RQ[p]);
    for (int p = 0; p < 6; p++) inv1(u, H, c, fu[p], Rfu[p]);
    for (int p = 0; p < 6; p++)
        for (int k = 0; k < 3; k++) {
            fp[p][k] = 0.5 * (Rfu[p][k] + alpha * RQ[p][k]);
            fm[p][k] = 0.5 * (Rfu[p][k] - alpha * RQ[p][k]);
        }
    double sk[3];
    for (int k = 0; k < 3; k++) {
        double lr = js ? wenojs_right(fp[0][k], fp[1][k], fp[2][k], fp[3][k], fp[4][k])
                       : wenoz_right(fp[0][k], fp[1][k], fp[2][k], fp[3][k], fp[4][k]);
        double rl = js ? wenojs_left(fm[1][k], fm[2][k], fm[3][k], fm[4][k], fm[5][k])
                       : wenoz_left(fm[1][k], fm[2][k], fm[3][k], fm[4][k], fm[5][k]);
        sk[k] = lr + rl;
    }
    for (int k = 0; k < 3; k++) flux[k] = sk[0] * r1[k] + sk[1] * r2[k] + sk[2] * r3[k];
}

// WENOZ_EULER_SOLVER_2D::ComputeFluxX / ComputeFluxY (WENOZ.h:236-320)
void flux2(double g, int dir, const double* const* s, double alpha, double* flux) {
    double fu[6][4], RQ[6][4], Rfu[6][4], fp[6][4], fm[6][4];
    for (int p = 0; p < 6; p++) {
        if (dir == 0) Fu2(g, s[p], fu[p]); else Gu2(g, s[p], fu[p]);
    }
    double u, v, H, c;
    roe2(g, s[2], s[3], u, v, H, c);
    double r1[4], r2[4], r3[4], r4[4];
    if (dir == 0) {
        r1[0] = 1;         r2[0] = 0; r3[0] = 1;                   r4[0] = 1;
        r1[1] = u - c;     r2[1] = 0; r3[1] = u;                   r4[1] = u + c;
        r1[2] = v;         r2[2] = 1; r3[2] = v;                   r4[2] = v;
        r1[3] = H - u * c; r2[3] = v; r3[3] = 0.5 * (u * u + v * v); r4[3] = H + u * c;
        for (int p = 0; p < 6; p++) invx2(u, v, H, c, s[p], RQ[p]);
        for (int p = 0; p < 6; p++) invx2(u, v, H, c, fu[p], Rfu[p]);
    } else {
        r1[0] = 1;         r2[0] = 0; r3[0] = 1;                   r4[0] = 1;
        r1[1] = u;         r2[1] = 1; r3[1] = u;                   r4[1] = u;
        r1[2] = v - c;     r2[2] = 0; r3[2] = v;                   r4[2] = v + c;
        r1[3] = H - v * c; r2[3] = u; r3[3] = 0.5 * (u * u + v * v); r4[3] = H + v * c;
        for (int p = 0; p < 6; p++) invy2(u, v, H, c, s[p], RQ[p]);
        for (int p = 0; p < 6; p++) invy2(u, v, H, c, fu[p], Rfu[p]);
    }
    for (int p = 0; p < 6; p++)
        for (int k = 0; k < 4; k++) {
            fp[p][k] = 0.5 * (Rfu[p][k] + alpha * RQ[p][k]);
            fm[p][k] = 0.5 * (Rfu[p][k] - alpha * RQ[p][k]);
        }
    double sk[4];
    for (int k = 0; k < 4; k++) {
        double lr = wenoz_right(fp[0][k], fp[1][k], fp[2][k], fp[3][k], fp[4][k]);
        double rl = wenoz_left(fm[1][k], fm[2][k], fm[3][k], fm[4][k], fm[5][k]);
        sk[k] = lr + rl;
    }
    for (int k = 0; k < 4; k++) flux[k] = sk[0] * r1[k] + sk[1] * r2[k] + sk[2] * r3[k] + sk[3] * r4[k];
}

// AWCM_1D/AWCM_2D derivative coefficients (AWCM_1D.h:21,36-39; AWCM_2D.h:24,43-46)
const double D1 = 2.0 / 3.0, D2 = -1.0 / 12.0;
const double D01 = -1.0 / 3.0, D11 = -1.0 / 2.0, D21 = 1.0, D31 = -1.0 / 6.0;
const double D02 = 1.0 / 12.0, D12 = -2.0 / 3.0, D22 = 0, D32 = 2.0 / 3.0;

// ===========================================================================
// 1-D solver (AWCM_1D.h, AWCM_MANAGER_1D.h, WAVELET_1D.h, main_1d_euler.cpp)
// ===========================================================================
struct Solver1D {
    int L, Nxc, Nx, maxlev;
    std::vector<int> dxl;          // MRA_GRID_INT_1D::dx
    std::vector<GridI1> grid, grid_x;
    Weights w;
    double gamma, dx;              // physical spacing (GRID_UNIFORM_1D::dx)
    std::vector<double> Q, Qc, Qt; // Q, Q_copy, Q_temp
    std::vector<uint8_t> mt, m[4];
    std::vector<int> lv[4];
    long oob = 0;                  // stencil reads outside the ghost band
    bool flux_js = false;          // WENO5-JS face fluxes (the shipped driver) in place of WENO-Z

    Solver1D(int L_, int Nxc_, double g, double dx_) : L(L_), Nxc(Nxc_), gamma(g), dx(dx_) {
        // MRA_GRID_INT_1D::Initialize (MRA_GRID_UNIFORM_1D.h:70-98)
        Nx = (Nxc - 1) * int(std::pow(2.0, L - 1)) + 1;
        maxlev = L - 1;
        dxl.resize(L); grid.resize(L); grid_x.resize(L - 1);
        for (int lev = L - 1; lev >= 0; lev--) {
            dxl[lev] = int(std::pow(2.0, L - lev - 1));
            if (lev == L - 1) grid[lev] = mk1(0, Nx - 1, Nx);
            else {
                grid[lev] = mk1(0, Nx - 1, (Nxc - 1) * int(std::pow(2.0, lev)) + 1);
                grid_x[lev] = mk1(dxl[lev + 1], (Nx - 1) - dxl[lev + 1], int((Nx - 1) / std::pow(2.0, L - 1 - lev)));
            }
        }
        Q.assign((Nx + 2 * GW) * 3, 0.0);
        Qc = Q; Qt = Q;
        mt.assign(Nx + 2 * GW, 0);
        for (int k = 0; k < 4; k++) { m[k].assign(Nx + 2 * GW, 0); lv[k].assign(Nx, 0); }
        std::fill(m[0].begin(), m[0].end(), 1);  // AWCM_1D.h:34
    }
    double* q(std::vector<double>& a, int i) { return &a[(i + GW) * 3]; }
    uint8_t& mk(std::vector<uint8_t>& a, int i) { return a[i + GW]; }

    // WAVELET_1D::ForwardWaveletTransformEuler (WAVELET_1D.h:282-335)
    void fwt(std::vector<uint8_t>& mask, double tol) {
        const int ie = Nx - 1;
        for (int lev = maxlev - 1; lev >= 0; lev--) {
            const int d = grid_x[lev].d, dh = d / 2;
            const GridI1& gx = grid_x[lev];
            for (int i = gx.s; i <= gx.e; i += gx.d) {
                if (mk(mask, i) == 0) continue;
                int is = i - d - dh, iend = i + d + dh, k = 1;
                if (is < 0) is = 0, k = 0;
                else if (iend > ie) is = ie - 3 * d, k = 2;
                for (int c = 0; c < 4; c++) q(Q, i)[0] -= w(c, k) * q(Q, is + c * d)[0];
                q(Q, i)[0] *= 0.5;
                if (ABSd(q(Q, i)[0]) < tol) mk(mask, i) = 0;
            }
            const GridI1& g = grid[lev];
            for (int i = g.s + g.d; i <= g.e - g.d; i += g.d) {
                if (mk(mask, i) == 0) continue;
                int is = i - d - dh, iend = i + d + dh, k = 1;
                if (is < 0) is = dh, k = 0;
                else if (iend > ie) is = ie - 3 * d - dh, k = 2;
                for (int c = 0; c < 4; c++) q(Q, i)[0] += w(c, k) * q(Q, is + c * d)[0];
            }
        }
    }
    // WAVELET_1D::InverseWaveletTransformEuler (WAVELET_1D.h:337-391)
    void iwt(std::vector<uint8_t>& mask) {
        const int ie = Nx - 1;
        for (int lev = 0; lev < maxlev; lev++) {
            const int d = grid_x[lev].d, dh = d / 2;
            const GridI1& g = grid[lev];
            for (int i = g.s + g.d; i <= g.e - g.d; i += g.d) {
                if (mk(mask, i) == 0) { q(Q, i)[0] = 0; continue; }
                int is = i - d - dh, iend = i + d + dh, k = 1;
                if (is < 0) is = dh, k = 0;
                else if (iend > ie) is = ie - 3 * d - dh, k = 2;
                for (int c = 0; c < 4; c++) q(Q, i)[0] -= w(c, k) * q(Q, is + c * d)[0];
            }
            const GridI1& gx = grid_x[lev];
            for (int i = gx.s; i <= gx.e; i += gx.d) {
                if (mk(mask, i) == 0) { q(Q, i)[0] = 0; continue; }
                int is = i - d - dh, iend = i + d + dh, k = 1;
                if (is < 0) is = 0, k = 0;
                else if (iend > ie) is = ie - 3 * d, k = 2;
                q(Q, i)[0] *= 2;
                for (int c = 0; c < 4; c++) q(Q, i)[0] += w(c, k) * q(Q, is + c * d)[0];
            }
        }
    }
    // AWCM_MANAGER_1D::CheckReconstruction (AWCM_MANAGER_1D.h:45-71)
    void check_recon(std::vector<uint8_t>& mask) {
        const int ie = Nx - 1;
        for (int lev = maxlev - 1; lev >= 0; lev--) {
            const int d = grid_x[lev].d, dh = d / 2;
            const GridI1& gx = grid_x[lev];
            for (int i = gx.s; i <= gx.e; i += gx.d) {
                if (mk(mask, i) == 0) continue;
                int is = i - d - dh, iend = i + d + dh;
                if (is < 0) is = 0;
                else if (iend > ie) is = ie - 3 * d;
                for (int c = 0; c < 4; c++) mk(mask, is + c * d) = 1;
            }
        }
    }
    // WAVELET_1D::InterpolateEuler(q, mra, from, to) (WAVELET_1D.h:393-431)
    // components p0..p1-1: 1..2 for InterpolateEuler (WAVELET_1D.h:393-431), 0 for the
    // scalar Interpolate(q, mra, from, to) (WAVELET_1D.h:198-233)
    void interp(std::vector<uint8_t>& from, std::vector<uint8_t>& to, int p0 = 1, int p1 = 3) {
        const int ie = Nx - 1;
        std::vector<uint8_t> tmp = to;
        for (int i = 0; i <= ie; i++) if (mk(from, i)) mk(tmp, i) = 0;
        check_recon(tmp);
        for (int lev = 0; lev < maxlev; lev++) {
            const int d = grid_x[lev].d, dh = d / 2;
            const GridI1& gx = grid_x[lev];
            for (int i = gx.s; i <= gx.e; i += gx.d) {
                if (mk(tmp, i) == 0 || mk(from, i) == 1) continue;
                int is = i - d - dh, iend = i + d + dh, k = 1;
                if (is < 0) is = 0, k = 0;
                else if (iend > ie) is = ie - 3 * d, k = 2;
                for (int p = p0; p < p1; p++) {
                    q(Q, i)[p] = 0;
                    for (int c = 0; c < 4; c++) q(Q, i)[p] += w(c, k) * q(Q, is + c * d)[p];
                }
            }
        }
    }
    // AWCM_MANAGER_1D::AddAdjacentZone (AWCM_MANAGER_1D.h:14-43)
    void adjacent(std::vector<uint8_t>& mask) {
        std::vector<uint8_t> snap = mask;
        const GridI1& gf = grid_x[maxlev - 1];
        for (int i = gf.s; i <= gf.e; i += gf.d)
            if (mk(snap, i)) { mk(mask, i - 1) = 1; mk(mask, i + 1) = 1; }
        for (int lev = maxlev - 2; lev >= 0; lev--) {
            const int d = grid[lev].d, dq = d / 4;
            const GridI1& gx = grid_x[lev];
            for (int i = gx.s; i <= gx.e; i += gx.d)
                if (mk(snap, i)) { mk(mask, i - dq) = 1; mk(mask, i + dq) = 1; }
        }
    }
    // AWCM_MANAGER_1D::AddPointForDifferentiation (AWCM_MANAGER_1D.h:73-107)
    void addpoint(std::vector<uint8_t>& mask, const std::vector<int>& level) {
        std::vector<uint8_t> snap = mask;
        const int N = Nx - 1;
        for (int i = 0; i <= Nx - 1; i++) {
            if (mk(snap, i) == 0) continue;
            const int lev = level[i];
            const int d = grid[lev].d;
            int is = i - 2 * d, iend = i + 2 * d;
            if (lev == maxlev) {
                for (int c = -3; c < 4; c++) mk(mask, i + c) = 1;
            } else {
                if (is <= 0) { for (int c = 0; c < 6; c++) mk(mask, c * d) = 1; }
                else if (iend >= N) { for (int c = 0; c < 6; c++) mk(mask, N - c * d) = 1; }
                else { for (int c = 0; c < 5; c++) mk(mask, is + c * d) = 1; }
            }
        }
    }
    // AWCM_MANAGER_1D::ComputeLevel (AWCM_MANAGER_1D.h:109-133)
    void level(std::vector<int>& lvl, std::vector<uint8_t>& mask) {
        for (int i = 1; i <= Nx - 2; i++) {
            if (mk(mask, i) == 0) { lvl[i] = 0; continue; }
            int d = 1;
            for (int lev = maxlev; lev >= 0; lev--) {
                if (i - d < -GW || i + d > Nx - 1 + GW) { oob++; break; }
                if (mk(mask, i - d) || mk(mask, i + d)) { lvl[i] = lev; break; }
                d *= 2;
            }
        }
        lvl[0] = maxlev;
        lvl[Nx - 1] = maxlev;
    }
    // AWCM_1D::BadingBoundary (AWCM_1D.h:73-80)
    void band(std::vector<uint8_t>& mask) {
        for (int i = 0; i <= Nx - 1; i++) {
            if (3 <= i && i <= Nx - 1 - 3) continue;
            mk(mask, i) = 1;
        }
    }
    // AWCM_1D::PrepareEuler (AWCM_1D.h:86-120)
    void prepare(double tol) {
        mt = m[0];
        fwt(m[0], tol);
        band(m[0]);
        adjacent(m[0]);
        check_recon(m[0]);
        for (int k = 1; k <= 3; k++) {
            m[k] = m[k - 1];
            level(lv[k - 1], m[k - 1]);
            adjacent(m[k]);
            addpoint(m[k], lv[k - 1]);
            check_recon(m[k]);
        }
        level(lv[3], m[3]);
        iwt(m[3]);
        interp(mt, m[3]);
    }
    // AWCM_1D::Prepare (AWCM_1D.h:45-71) on the scalar field (component 0): no
    // mask_temp, no AddAdjacentZone on masks 1..3, no interpolation
    void prepare_scalar(double tol) {
        fwt(m[0], tol);
        band(m[0]);
        adjacent(m[0]);
        check_recon(m[0]);
        for (int k = 1; k <= 3; k++) {
            m[k] = m[k - 1];
            level(lv[k - 1], m[k - 1]);
            addpoint(m[k], lv[k - 1]);
            check_recon(m[k]);
        }
        level(lv[3], m[3]);
        iwt(m[3]);
    }
    // AWCM_MANAGER_1D::RemoveResidue (AWCM_MANAGER_1D.h:146-157)
    void post() {
        for (int i = 0; i <= Nx - 1; i++)
            if (mk(m[0], i) == 0) { q(Q, i)[0] = 0; q(Q, i)[1] = 0; q(Q, i)[2] = 0; }
    }
    // BOUNDARY_MANAGER::SetBoundaryCondition (PROBLEM_SHOCKTUBE.h:86-107); the
    // Shu-Osher manager (PROBLEM_ENTROPY_INTERACTING.h:81-102) is identical.
    void bc() {
        for (int i = -1; i >= -GW; i--) for (int c = 0; c < 3; c++) q(Q, i)[c] = q(Q, 0)[c];
        for (int i = Nx; i <= Nx - 1 + GW; i++) for (int c = 0; c < 3; c++) q(Q, i)[c] = q(Q, Nx - 1)[c];
    }
    double alpha(int k) {
        double a = 0;
        for (int i = 0; i <= Nx - 1; i++) {
            if (mk(m[k], i) == 0 || lv[k][i] != maxlev) continue;
            double u = q(Q, i)[1] / q(Q, i)[0];
            double c = c1(gamma, q(Q, i));
            a = MAXd(a, ABSd(u) + c);
        }
        return a;
    }
    // dF at point i of one stage (main_1d_euler.cpp:99-134): the WENO-Z flux
    // pair, or with flux_js the shipped WENO_EULER_SOLVER_1D::ComputedF
    // (WENO.h:176-184), the same pair of WENO5-JS fluxes.
    void dF(int i, int lev, double a, double* out) {
        const int d = dxl[lev];
        auto QT = [&](int ii) -> const double* {
            if (ii < -GW || ii > Nx - 1 + GW) oob++;
            return q(Qt, ii);
        };
        if (lev == maxlev) {
            const double* s0[6] = {QT(i - 3), QT(i - 2), QT(i - 1), QT(i), QT(i + 1), QT(i + 2)};
            const double* s1[6] = {QT(i - 2), QT(i - 1), QT(i), QT(i + 1), QT(i + 2), QT(i + 3)};
            double Fl[3], Fr[3];
            flux1(gamma, s0, a, Fl, flux_js);
            flux1(gamma, s1, a, Fr, flux_js);
            for (int k = 0; k < 3; k++) out[k] = Fr[k] - Fl[k];
            return;
        }
        const int ie = Nx - 1;
        double f_ll[3], f_l[3], f_c[3], f_r[3], f_rr[3];
        if (i - 2 * d > 0 && i + 2 * d < ie) {
            Fu1(gamma, QT(i - 2 * d), f_ll); Fu1(gamma, QT(i - d), f_l); Fu1(gamma, QT(i), f_c);
            Fu1(gamma, QT(i + d), f_r); Fu1(gamma, QT(i + 2 * d), f_rr);
            for (int p = 0; p < 3; p++) out[p] = D1 * (f_r[p] - f_l[p]) + D2 * (f_rr[p] - f_ll[p]);
        } else if (i - 2 * d < 0) {
            // DifferentiateEulerAtLeftBoundary(l, c, r, rr) (AWCM_1D.h:158-168)
            Fu1(gamma, QT(i - d), f_l); Fu1(gamma, QT(i), f_c); Fu1(gamma, QT(i + d), f_r); Fu1(gamma, QT(i + 2 * d), f_rr);
            for (int p = 0; p < 3; p++) out[p] = D01 * f_l[p] + D11 * f_c[p] + D21 * f_r[p] + D31 * f_rr[p];
        } else if (i == 2 * d) {
            // DifferentiateEulerAtLeftBoundary(ll, l, c, r, rr) (AWCM_1D.h:174-185)
            Fu1(gamma, QT(i - 2 * d), f_ll); Fu1(gamma, QT(i - d), f_l); Fu1(gamma, QT(i), f_c);
            Fu1(gamma, QT(i + d), f_r); Fu1(gamma, QT(i + 2 * d), f_rr);
            for (int p = 0; p < 3; p++)
                out[p] = D02 * f_ll[p] + D12 * f_l[p] + D22 * f_c[p] + D32 * f_r[p] + D2 * f_rr[p];
        } else if (i > ie - 2 * d) {
            // DifferentiateEulerAtRightBoundary(ll, l, c, r) = AtLeft(r, c, l, ll) (AWCM_1D.h:169-172)
            Fu1(gamma, QT(i - 2 * d), f_ll); Fu1(gamma, QT(i - d), f_l); Fu1(gamma, QT(i), f_c); Fu1(gamma, QT(i + d), f_r);
            for (int p = 0; p < 3; p++) out[p] = D01 * f_r[p] + D11 * f_c[p] + D21 * f_l[p] + D31 * f_ll[p];
        } else {
            // DifferentiateEulerAtRightBoundary(ll, l, c, r, rr) = AtLeft5(rr, r, c, l, ll) (AWCM_1D.h:186-189)
            Fu1(gamma, QT(i - 2 * d), f_ll); Fu1(gamma, QT(i - d), f_l); Fu1(gamma, QT(i), f_c);
            Fu1(gamma, QT(i + d), f_r); Fu1(gamma, QT(i + 2 * d), f_rr);
            for (int p = 0; p < 3; p++)
                out[p] = D02 * f_rr[p] + D12 * f_r[p] + D22 * f_c[p] + D32 * f_l[p] + D2 * f_ll[p];
        }
    }
    // The three SSP-RK3 stages and RemoveResidue of one driver-loop iteration
    // (main_1d_euler.cpp:97-257), with dt and the stage alphas given.
    void stages(double dt, const double* al) {
        bc();
        Qc = Q; Qt = Q;
        for (int i = 0; i <= Nx - 1; i++) {
            if (mk(m[2], i) == 0) continue;
            const int lev = lv[2][i];
            double f[3]; dF(i, lev, al[0], f);
            const double cc = dt / (dx * dxl[lev]);
            for (int k = 0; k < 3; k++) q(Q, i)[k] = q(Q, i)[k] - cc * f[k];
        }
        bc();
        Qt = Q;
        for (int i = 0; i <= Nx - 1; i++) {
            if (mk(m[1], i) == 0) continue;
            const int lev = lv[1][i];
            double f[3]; dF(i, lev, al[1], f);
            const double cc = dt / (dx * dxl[lev]);
            for (int k = 0; k < 3; k++) q(Q, i)[k] = 0.75 * q(Qc, i)[k] + 0.25 * (q(Q, i)[k] - cc * f[k]);
        }
        bc();
        Qt = Q;
        for (int i = 0; i <= Nx - 1; i++) {
            if (mk(m[0], i) == 0) continue;
            const int lev = lv[0][i];
            double f[3]; dF(i, lev, al[2], f);
            const double cc = dt / (dx * dxl[lev]);
            for (int k = 0; k < 3; k++) q(Q, i)[k] = (q(Qc, i)[k] + 2.0 * (q(Q, i)[k] - cc * f[k])) / 3.0;
        }
        post();
    }
    // Frozen-trajectory tape for finite-difference checks of the discrete
    // adjoint: per step the masks, levels, dt and stage alphas it used.
    struct TapeStep {
        std::vector<uint8_t> mt, m[4];
        std::vector<int> lv[3];
        double dt, al[3];
    };
    std::vector<TapeStep> tape;
    bool recording = false;

    // One iteration of the driver loop (main_1d_euler.cpp:61-260).  Returns dt;
    // *cnt is the print counter (starts at 1), *t the time.
    double step(double tol, double cfl, const double* pt, int npt, int* cnt, double* t) {
        std::vector<uint8_t> m_start = m[0];
        prepare(tol);
        double vmax = 0;
        for (int i = 0; i <= Nx - 1; i++) {
            if (mk(m[3], i) == 0) continue;
            double c = c1(gamma, q(Q, i));
            double u = q(Q, i)[1] / q(Q, i)[0];
            vmax = MAXd(vmax, ABSd(u) + c);
        }
        double dt = cfl * dx / vmax;
        bool do_print = false;
        if (*cnt < npt && *t + dt > pt[*cnt]) { dt = pt[*cnt] - *t; do_print = true; }
        // alphas: stage s uses mask_{3-s} of the stage input (main_1d_euler.cpp:86-95)
        bc();
        double al[3];
        al[0] = alpha(3);
        {
            // stage inputs for alpha(2) and alpha(1) are only known stage by stage
            // (they depend on the stage results), so run the stages inline
            Qc = Q; Qt = Q;
            for (int i = 0; i <= Nx - 1; i++) {
                if (mk(m[2], i) == 0) continue;
                const int lev = lv[2][i];
                double f[3]; dF(i, lev, al[0], f);
                const double cc = dt / (dx * dxl[lev]);
                for (int k = 0; k < 3; k++) q(Q, i)[k] = q(Q, i)[k] - cc * f[k];
            }
            bc();
            al[1] = alpha(2);
            Qt = Q;
            for (int i = 0; i <= Nx - 1; i++) {
                if (mk(m[1], i) == 0) continue;
                const int lev = lv[1][i];
                double f[3]; dF(i, lev, al[1], f);
                const double cc = dt / (dx * dxl[lev]);
                for (int k = 0; k < 3; k++) q(Q, i)[k] = 0.75 * q(Qc, i)[k] + 0.25 * (q(Q, i)[k] - cc * f[k]);
            }
            bc();
            al[2] = alpha(1);
            Qt = Q;
            for (int i = 0; i <= Nx - 1; i++) {
                if (mk(m[0], i) == 0) continue;
                const int lev = lv[0][i];
                double f[3]; dF(i, lev, al[2], f);
                const double cc = dt / (dx * dxl[lev]);
                for (int k = 0; k < 3; k++) q(Q, i)[k] = (q(Qc, i)[k] + 2.0 * (q(Q, i)[k] - cc * f[k])) / 3.0;
            }
            post();
        }
        if (recording) {
            TapeStep s;
            s.mt = m_start;
            for (int k = 0; k < 4; k++) s.m[k] = m[k];
            for (int k = 0; k < 3; k++) s.lv[k] = lv[k];
            s.dt = dt;
            for (int k = 0; k < 3; k++) s.al[k] = al[k];
            tape.push_back(std::move(s));
        }
        *t += dt;
        if (do_print) (*cnt)++;
        return dt;
    }
    // Step n of the tape with masks, levels, dt and alphas frozen: the same
    // arithmetic as step(), since the FWT consults mask_temp at every point
    // before thresholding it (WAVELET_1D.h:282-335).
    void replay(int n) {
        const TapeStep& s = tape[n];
        std::vector<uint8_t> mcopy = s.mt;
        fwt(mcopy, -1.0);
        for (int k = 0; k < 4; k++) m[k] = s.m[k];
        for (int k = 0; k < 3; k++) lv[k] = s.lv[k];
        iwt(m[3]);
        interp(mcopy, m[3]);
        stages(s.dt, s.al);
    }
};

// ===========================================================================
// 2-D solver (AWCM_2D.h, AWCM_MANAGER_2D.h, WAVELET_2D.h, main_2d_euler(dmach|quad).cpp)
// ===========================================================================
enum { BC_QUAD = 0, BC_DMACH = 1 };

struct Solver2D {
    int L, Nxc, Nyc, Nx, Ny, maxlev, P;  // P = row pitch Nx+6
    std::vector<int> mra;
    std::vector<GridI2> grid;
    std::vector<GridI2> sub;  // sub[lev*3 + mu-1]
    Weights w;
    double gamma, dx, x0;
    int bc_kind;
    std::vector<double> Q, Q0, Q1, Q2;
    std::vector<uint8_t> mt, m[4];
    std::vector<int> lv[4];  // level0..3
    std::vector<double> ax, ay;
    long oob = 0;

    Solver2D(int L_, int nxc, int nyc, double g, double dx_, double x0_, int bck)
        : L(L_), Nxc(nxc), Nyc(nyc), gamma(g), dx(dx_), x0(x0_), bc_kind(bck) {
        // MRA_GRID_INT_2D::Initialize (MRA_GRID_UNIFORM_2D.h:87-134)
        Nx = (Nxc - 1) * int(std::pow(2.0, L - 1)) + 1;
        Ny = (Nyc - 1) * int(std::pow(2.0, L - 1)) + 1;
        maxlev = L - 1;
        P = Nx + 2 * GW;
        mra.resize(L); grid.resize(L); sub.resize(L * 3);
        for (int lev = L - 1; lev >= 0; lev--) {
            mra[lev] = int(std::pow(2.0, L - lev - 1));
            const int mm = mra[lev];
            if (lev == L - 1)
                grid[lev] = mk2(0, 0, Nx - 1, Ny - 1, int((Nxc - 1) * std::pow(2.0, lev) + 1),
                                int((Nyc - 1) * std::pow(2.0, lev) + 1));
            else
                grid[lev] = mk2(mm, mm, (Nx - 1) - mm, (Ny - 1) - mm, int((Nxc - 1) * std::pow(2.0, lev) + 1) - 2,
                                int((Nyc - 1) * std::pow(2.0, lev) + 1) - 2);
            if (lev == L - 1) continue;
            sub[lev * 3 + 0] = mk2(int(0.5 * mm), mm, int(Nx - 0.5 * mm - 1), Ny - mm - 1, int(Nx / mm), int(Ny / mm - 1));
            sub[lev * 3 + 1] = mk2(mm, int(0.5 * mm), Nx - mm - 1, int(Ny - 0.5 * mm - 1), int(Nx / mm - 1), int(Ny / mm));
            sub[lev * 3 + 2] = mk2(int(0.5 * mm), int(0.5 * mm), int(Nx - 0.5 * mm - 1), int(Ny - 0.5 * mm - 1),
                                   int(Nx / mm), int(Ny / mm));
        }
        const size_t np = size_t(P) * (Ny + 2 * GW);
        Q.assign(np * 4, 0.0); Q0 = Q; Q1 = Q; Q2 = Q;
        mt.assign(np, 0);
        for (int k = 0; k < 4; k++) { m[k].assign(np, 0); lv[k].assign(size_t(Nx) * Ny, maxlev); }
        std::fill(m[0].begin(), m[0].end(), 1);  // AWCM_2D.h:37-41
        ax.assign(Ny, 0); ay.assign(Nx, 0);
    }
    size_t id(int i, int j) const { return size_t(i + GW) + size_t(P) * (j + GW); }
    double* q(std::vector<double>& a, int i, int j) { return &a[id(i, j) * 4]; }
    uint8_t& mk(std::vector<uint8_t>& a, int i, int j) { return a[id(i, j)]; }
    const GridI2& sg(int lev, int mu) const { return sub[lev * 3 + mu - 1]; }

    // WAVELET_2D::ForwardWaveletTransformationEuler (WAVELET_2D.h:478-624)
    void fwt(std::vector<uint8_t>& mask, double tol) {
        const int ie = Nx - 1, je = Ny - 1;
        for (int lev = maxlev - 1; lev >= 0; lev--) {
            const int d = mra[lev], dh = d / 2;
            {   const GridI2& g = sg(lev, 3);
#pragma omp parallel for
                for (int j = g.js; j <= g.je; j += g.dy)
                    for (int i = g.is; i <= g.ie; i += g.dx) {
                        if (mk(mask, i, j) == 0) continue;
                        int is = i - d - dh, js = j - d - dh, iend = i + d + dh, jend = j + d + dh, k1 = 1, k2 = 1;
                        if (is < 0) is = 0, k1 = 0; else if (iend > ie) is = ie - 3 * d, k1 = 2;
                        if (js < 0) js = 0, k2 = 0; else if (jend > je) js = je - 3 * d, k2 = 2;
                        double& v = q(Q, i, j)[0];
                        int mm = is;
                        for (int c = 0; c < 4; c++, mm += d) v -= q(Q, mm, j)[0] * w(c, k1);
                        int n = js;
                        for (int c = 0; c < 4; c++, n += d) v -= q(Q, i, n)[0] * w(c, k2);
                        n = js;
                        for (int c2 = 0; c2 < 4; c2++, n += d) {
                            mm = is;
                            for (int c1 = 0; c1 < 4; c1++, mm += d) v += q(Q, mm, n)[0] * w(c1, k1) * w(c2, k2);
                        }
                        v *= 0.25;
                        if (ABSd(v) < tol) mk(mask, i, j) = 0;
                    }
            }
            {   const GridI2& g = sg(lev, 1);
#pragma omp parallel for
                for (int j = g.js; j <= g.je; j += g.dy)
                    for (int i = g.is; i <= g.ie; i += g.dx) {
                        if (mk(mask, i, j) == 0) continue;
                        int is = i - d - dh, iend = i + d + dh, k = 1;
                        if (is < 0) is = 0, k = 0; else if (iend > ie) is = ie - 3 * d, k = 2;
                        double& v = q(Q, i, j)[0];
                        int mm = is;
                        for (int c = 0; c < 4; c++, mm += d) v -= q(Q, mm, j)[0] * w(c, k);
                        v *= 0.5;
                        if (ABSd(v) < tol) mk(mask, i, j) = 0;
                    }
            }
            {   const GridI2& g = sg(lev, 2);
#pragma omp parallel for
                for (int j = g.js; j <= g.je; j += g.dy)
                    for (int i = g.is; i <= g.ie; i += g.dx) {
                        if (mk(mask, i, j) == 0) continue;
                        int js = j - d - dh, jend = j + d + dh, k = 1;
                        if (js < 0) js = 0, k = 0; else if (jend > je) js = je - 3 * d, k = 2;
                        double& v = q(Q, i, j)[0];
                        int n = js;
                        for (int c = 0; c < 4; c++, n += d) v -= q(Q, i, n)[0] * w(c, k);
                        v *= 0.5;
                        if (ABSd(v) < tol) mk(mask, i, j) = 0;
                    }
            }
            {   const GridI2& g = grid[lev];
#pragma omp parallel for
                for (int j = g.js; j <= g.je; j += g.dy)
                    for (int i = g.is; i <= g.ie; i += g.dx) {
                        if (mk(mask, i, j) == 0) continue;
                        int is = i - d - dh, js = j - d - dh, iend = i + d + dh, jend = j + d + dh, k1 = 1, k2 = 1;
                        if (is < 0) is = dh, k1 = 0; else if (iend > ie) is = ie - 3 * d - dh, k1 = 2;
                        if (js < 0) js = dh, k2 = 0; else if (jend > je) js = je - 3 * d - dh, k2 = 2;
                        double& v = q(Q, i, j)[0];
                        int mm = is;
                        for (int c = 0; c < 4; c++, mm += d) v += q(Q, mm, j)[0] * w(c, k1);
                        int n = js;
                        for (int c = 0; c < 4; c++, n += d) v += q(Q, i, n)[0] * w(c, k2);
                        n = js;
                        for (int c2 = 0; c2 < 4; c2++, n += d) {
                            mm = is;
                            for (int c1 = 0; c1 < 4; c1++, mm += d) v += q(Q, mm, n)[0] * w(c1, k1) * w(c2, k2);
                        }
                    }
            }
        }
    }
    // WAVELET_2D::InverseWaveletTransformationEuler (WAVELET_2D.h:626-782)
    void iwt(std::vector<uint8_t>& mask) {
        const int ie = Nx - 1, je = Ny - 1;
        for (int lev = 0; lev < maxlev; lev++) {
            const int d = mra[lev], dh = d / 2;
            {   const GridI2& g = grid[lev];
#pragma omp parallel for
                for (int j = g.js; j <= g.je; j += g.dy)
                    for (int i = g.is; i <= g.ie; i += g.dx) {
                        if (mk(mask, i, j) == 0) { q(Q, i, j)[0] = 0; continue; }
                        int is = i - d - dh, js = j - d - dh, iend = i + d + dh, jend = j + d + dh, k1 = 1, k2 = 1;
                        if (is < 0) is = dh, k1 = 0; else if (iend > ie) is = ie - 3 * d - dh, k1 = 2;
                        if (js < 0) js = dh, k2 = 0; else if (jend > je) js = je - 3 * d - dh, k2 = 2;
                        double& v = q(Q, i, j)[0];
                        int mm = is;
                        for (int c = 0; c < 4; c++, mm += d) v -= q(Q, mm, j)[0] * w(c, k1);
                        int n = js;
                        for (int c = 0; c < 4; c++, n += d) v -= q(Q, i, n)[0] * w(c, k2);
                        n = js;
                        for (int c2 = 0; c2 < 4; c2++, n += d) {
                            mm = is;
                            for (int c1 = 0; c1 < 4; c1++, mm += d) v -= q(Q, mm, n)[0] * w(c1, k1) * w(c2, k2);
                        }
                    }
            }
            {   const GridI2& g = sg(lev, 2);
#pragma omp parallel for
                for (int j = g.js; j <= g.je; j += g.dy)
                    for (int i = g.is; i <= g.ie; i += g.dx) {
                        if (mk(mask, i, j) == 0) { q(Q, i, j)[0] = 0; continue; }
                        int js = j - d - dh, jend = j + d + dh, k = 1;
                        if (js < 0) js = 0, k = 0; else if (jend > je) js = je - 3 * d, k = 2;
                        double& v = q(Q, i, j)[0];
                        v *= 2.0;
                        int n = js;
                        for (int c = 0; c < 4; c++, n += d) v += q(Q, i, n)[0] * w(c, k);
                    }
            }
            {   const GridI2& g = sg(lev, 1);
#pragma omp parallel for
                for (int j = g.js; j <= g.je; j += g.dy)
                    for (int i = g.is; i <= g.ie; i += g.dx) {
                        if (mk(mask, i, j) == 0) { q(Q, i, j)[0] = 0; continue; }
                        int is = i - d - dh, iend = i + d + dh, k = 1;
                        if (is < 0) is = 0, k = 0; else if (iend > ie) is = ie - 3 * d, k = 2;
                        double& v = q(Q, i, j)[0];
                        v *= 2.0;
                        int mm = is;
                        for (int c = 0; c < 4; c++, mm += d) v += q(Q, mm, j)[0] * w(c, k);
                    }
            }
            {   const GridI2& g = sg(lev, 3);
#pragma omp parallel for
                for (int j = g.js; j <= g.je; j += g.dy)
                    for (int i = g.is; i <= g.ie; i += g.dx) {
                        if (mk(mask, i, j) == 0) { q(Q, i, j)[0] = 0; continue; }
                        int is = i - d - dh, js = j - d - dh, iend = i + d + dh, jend = j + d + dh, k1 = 1, k2 = 1;
                        if (is < 0) is = 0, k1 = 0; else if (iend > ie) is = ie - 3 * d, k1 = 2;
                        if (js < 0) js = 0, k2 = 0; else if (jend > je) js = je - 3 * d, k2 = 2;
                        double& v = q(Q, i, j)[0];
                        v *= 4.0;
                        int mm = is;
                        for (int c = 0; c < 4; c++, mm += d) v += q(Q, mm, j)[0] * w(c, k1);
                        int n = js;
                        for (int c = 0; c < 4; c++, n += d) v += q(Q, i, n)[0] * w(c, k2);
                        n = js;
                        for (int c2 = 0; c2 < 4; c2++, n += d) {
                            mm = is;
                            for (int c1 = 0; c1 < 4; c1++, mm += d) v -= q(Q, mm, n)[0] * w(c1, k1) * w(c2, k2);
                        }
                    }
            }
        }
    }
    // AWCM_MANAGER_2D::CheckReconstruction (AWCM_MANAGER_2D.h:63-158); serial, as there.
    void check_recon(std::vector<uint8_t>& mask) {
        const int ie = Nx - 1, je = Ny - 1;
        for (int lev = maxlev - 1; lev >= 0; lev--) {
            const int d = mra[lev], dh = d / 2;
            {   const GridI2& g = sg(lev, 3);
                for (int j = g.js; j <= g.je; j += g.dy)
                    for (int i = g.is; i <= g.ie; i += g.dx) {
                        if (mk(mask, i, j) == 0) continue;
                        int is = i - d - dh, js = j - d - dh, iend = i + d + dh, jend = j + d + dh;
                        if (is < 0) is = 0; else if (iend > ie) is = ie - 3 * d;
                        if (js < 0) js = 0; else if (jend > je) js = je - 3 * d;
                        for (int c = 0; c < 4; c++) mk(mask, is + c * d, j) = 1;
                        for (int c = 0; c < 4; c++) mk(mask, i, js + c * d) = 1;
                        for (int c2 = 0; c2 < 4; c2++)
                            for (int c1 = 0; c1 < 4; c1++) mk(mask, is + c1 * d, js + c2 * d) = 1;
                    }
            }
            {   const GridI2& g = sg(lev, 1);
                for (int j = g.js; j <= g.je; j += g.dy)
                    for (int i = g.is; i <= g.ie; i += g.dx) {
                        if (mk(mask, i, j) == 0) continue;
                        int is = i - d - dh, iend = i + d + dh;
                        if (is < 0) is = 0; else if (iend > ie) is = ie - 3 * d;
                        for (int c = 0; c < 4; c++) mk(mask, is + c * d, j) = 1;
                    }
            }
            {   const GridI2& g = sg(lev, 2);
                for (int j = g.js; j <= g.je; j += g.dy)
                    for (int i = g.is; i <= g.ie; i += g.dx) {
                        if (mk(mask, i, j) == 0) continue;
                        int js = j - d - dh, jend = j + d + dh;
                        if (js < 0) js = 0; else if (jend > je) js = je - 3 * d;
                        for (int c = 0; c < 4; c++) mk(mask, i, js + c * d) = 1;
                    }
            }
        }
    }
    // AWCM_MANAGER_2D::AddAdjacentZone (AWCM_MANAGER_2D.h:14-61)
    void adjacent(std::vector<uint8_t>& mask) {
        std::vector<uint8_t> snap = mask;
        for (int mu = 1; mu <= 3; mu++) {
            const GridI2& g = sg(maxlev - 1, mu);
            for (int j = g.js; j <= g.je; j += g.dy)
                for (int i = g.is; i <= g.ie; i += g.dx)
                    if (mk(snap, i, j)) {
                        mk(mask, i - 1, j) = 1; mk(mask, i + 1, j) = 1;
                        mk(mask, i, j - 1) = 1; mk(mask, i, j + 1) = 1;
                    }
        }
        for (int lev = maxlev - 2; lev >= 0; lev--) {
            const int d = grid[lev].dx, dyy = grid[lev].dy;
            const int dh = d / 2, dq = d / 4, dyh = dyy / 2, dyq = dyy / 4;
            for (int mu = 1; mu <= 3; mu++) {
                const GridI2& g = sg(lev, mu);
                for (int j = g.js; j <= g.je; j += g.dy)
                    for (int i = g.is; i <= g.ie; i += g.dx)
                        if (mk(snap, i, j)) {
                            mk(mask, i - dq, j) = 1; mk(mask, i + dq, j) = 1;
                            mk(mask, i, j - dyq) = 1; mk(mask, i, j + dyq) = 1;
                            mk(mask, i - dh, j) = 1; mk(mask, i + dh, j) = 1;
                            mk(mask, i, j - dyh) = 1; mk(mask, i, j + dyh) = 1;
                        }
            }
        }
    }
    // AWCM_MANAGER_2D::AddPointForDifferentiation (AWCM_MANAGER_2D.h:160-213)
    void addpoint(std::vector<uint8_t>& mask, const std::vector<int>& level) {
        std::vector<uint8_t> snap = mask;
        const int NX = Nx - 1, NY = Ny - 1;
        for (int j = 1; j <= Ny - 2; j++)
            for (int i = 1; i <= Nx - 2; i++) {
                if (mk(snap, i, j) == 0) continue;
                const int lev = level[i + size_t(Nx) * j];
                const int d = grid[lev].dx;
                const int is = i - 2 * d, iend = i + 2 * d, js = j - 2 * d, jend = j + 2 * d;
                if (lev == maxlev) {
                    for (int c = -3; c <= 3; c++) { mk(mask, i + c, j) = 1; mk(mask, i, j + c) = 1; }
                } else {
                    if (is < 0) { for (int c = -1; c <= 2; c++) mk(mask, i + d * c, j) = 1; }
                    else if (iend > NX) { for (int c = -2; c <= 1; c++) mk(mask, i + d * c, j) = 1; }
                    else { for (int c = -2; c <= 2; c++) mk(mask, i + d * c, j) = 1; }
                    if (js < 0) { for (int c = -1; c <= 2; c++) mk(mask, i, j + d * c) = 1; }
                    else if (jend > NY) { for (int c = -2; c <= 1; c++) mk(mask, i, j + d * c) = 1; }
                    else { for (int c = -2; c <= 2; c++) mk(mask, i, j + d * c) = 1; }
                }
            }
    }
    // AWCM_MANAGER_2D::ComputeLevel (AWCM_MANAGER_2D.h:215-241)
    void level(std::vector<int>& lvl, std::vector<uint8_t>& mask) {
#pragma omp parallel for
        for (int j = 0; j <= Ny - 1; j++)
            for (int i = 0; i <= Nx - 1; i++) {
                if (mk(mask, i, j) == 0) { lvl[i + size_t(Nx) * j] = 0; continue; }
                int d = 1;
                for (int lev = maxlev; lev >= 0; lev--) {
                    if (i - d < -GW || i + d > Nx - 1 + GW || j - d < -GW || j + d > Ny - 1 + GW) {
#pragma omp atomic
                        oob++;
                        break;
                    }
                    if (mk(mask, i - d, j) || mk(mask, i + d, j) || mk(mask, i, j - d) || mk(mask, i, j + d)) {
                        lvl[i + size_t(Nx) * j] = lev;
                        break;
                    }
                    d *= 2;
                }
            }
    }
    // BOUNDARY_MANAGER_SHOCK_VORTEX::BandingMask (PROBLEM_SHOCK_VORTEX_INTERACTION.h:205-213)
    void band(std::vector<uint8_t>& mask) {
        for (int j = 0; j <= Ny - 1; j++)
            for (int i = 0; i <= Nx - 1; i++) {
                if (2 <= i && i <= Nx - 1 - 2 && j >= 2 && j <= Ny - 1 - 2) continue;
                mk(mask, i, j) = 1;
            }
    }
    // WAVELET_2D::InterpolateEuler(q, mra, from, to) (WAVELET_2D.h:784-907)
    void interp(std::vector<uint8_t>& from, std::vector<uint8_t>& to) {
        const int ie = Nx - 1, je = Ny - 1;
        std::vector<uint8_t> tmp = to;
        for (int j = 0; j <= je; j++)
            for (int i = 0; i <= ie; i++)
                if (mk(from, i, j)) mk(tmp, i, j) = 0;
        check_recon(tmp);
        interp_passes(from, tmp, 1);
    }
    // WAVELET_2D::InterpolateAllGridEuler(q, mra, mask_from) (WAVELET_2D.h:912-1044):
    // mask_temp is every point (ghosts included) except the finest grid's points
    // of mask_from, closed by CheckReconstruction; then the passes of
    // InterpolateEuler over all four components.
    void interpolate_all(std::vector<uint8_t>& from) {
        const int ie = Nx - 1, je = Ny - 1;
        std::vector<uint8_t> tmp(from.size(), 1);
        for (int j = 0; j <= je; j++)
            for (int i = 0; i <= ie; i++)
                if (mk(from, i, j)) mk(tmp, i, j) = 0;
        check_recon(tmp);
        interp_passes(from, tmp, 0);
    }
    // WAVELET_2D::Interpolate(q, mra, from, to) (WAVELET_2D.h:360-473): the
    // passes of InterpolateEuler on the scalar field (component 0 here)
    void interp_scalar(std::vector<uint8_t>& from, std::vector<uint8_t>& to) {
        const int ie = Nx - 1, je = Ny - 1;
        std::vector<uint8_t> tmp = to;
        for (int j = 0; j <= je; j++)
            for (int i = 0; i <= ie; i++)
                if (mk(from, i, j)) mk(tmp, i, j) = 0;
        check_recon(tmp);
        interp_passes(from, tmp, 0, 1);
    }
    // coarse-to-fine d1, d2, d3 predictions of components p0..p1-1 at the points
    // of `tmp` not in `from` (WAVELET_2D.h:796-905, :922-1041 and :372-470)
    void interp_passes(std::vector<uint8_t>& from, std::vector<uint8_t>& tmp, int p0, int p1 = 4) {
        const int ie = Nx - 1, je = Ny - 1;
        for (int lev = 0; lev < maxlev; lev++) {
            const int d = mra[lev], dh = d / 2;
            {   const GridI2& g = sg(lev, 1);
#pragma omp parallel for
                for (int j = g.js; j <= g.je; j += g.dy)
                    for (int i = g.is; i <= g.ie; i += g.dx) {
                        if (mk(tmp, i, j) == 0 || mk(from, i, j) == 1) continue;
                        int is = i - d - dh, iend = i + d + dh, k = 1;
                        if (is < 0) is = 0, k = 0; else if (iend > ie) is = ie - 3 * d, k = 2;
                        for (int p = p0; p < p1; p++) {
                            double& v = q(Q, i, j)[p];
                            v = 0.0;
                            int mm = is;
                            for (int c = 0; c < 4; c++, mm += d) v += q(Q, mm, j)[p] * w(c, k);
                        }
                    }
            }
            {   const GridI2& g = sg(lev, 2);
#pragma omp parallel for
                for (int j = g.js; j <= g.je; j += g.dy)
                    for (int i = g.is; i <= g.ie; i += g.dx) {
                        if (mk(tmp, i, j) == 0 || mk(from, i, j) == 1) continue;
                        int js = j - d - dh, jend = j + d + dh, k = 1;
                        if (js < 0) js = 0, k = 0; else if (jend > je) js = je - 3 * d, k = 2;
                        for (int p = p0; p < p1; p++) {
                            double& v = q(Q, i, j)[p];
                            v = 0.0;
                            int n = js;
                            for (int c = 0; c < 4; c++, n += d) v += q(Q, i, n)[p] * w(c, k);
                        }
                    }
            }
            {   const GridI2& g = sg(lev, 3);
#pragma omp parallel for
                for (int j = g.js; j <= g.je; j += g.dy)
                    for (int i = g.is; i <= g.ie; i += g.dx) {
                        if (mk(tmp, i, j) == 0 || mk(from, i, j) == 1) continue;
                        int is = i - d - dh, js = j - d - dh, iend = i + d + dh, jend = j + d + dh, k1 = 1, k2 = 1;
                        if (is < 0) is = 0, k1 = 0; else if (iend > ie) is = ie - 3 * d, k1 = 2;
                        if (js < 0) js = 0, k2 = 0; else if (jend > je) js = je - 3 * d, k2 = 2;
                        for (int p = p0; p < p1; p++) {
                            double& v = q(Q, i, j)[p];
                            v = 0.0;
                            int mm = is, n = js;
                            for (int c = 0; c < 4; c++, mm += d) v += q(Q, mm, j)[p] * w(c, k1);
                            for (int c = 0; c < 4; c++, n += d) v += q(Q, i, n)[p] * w(c, k2);
                            n = js;
                            for (int c2 = 0; c2 < 4; c2++, n += d) {
                                mm = is;
                                for (int c1 = 0; c1 < 4; c1++, mm += d) v -= q(Q, mm, n)[p] * w(c1, k1) * w(c2, k2);
                            }
                        }
                    }
            }
        }
    }
    // AWCM_2D::PrepareEuler (AWCM_2D.h:114-147)
    void prepare(double tol) {
        mt = m[0];
        fwt(m[0], tol);
        band(m[0]);
        adjacent(m[0]);
        check_recon(m[0]);
        for (int k = 1; k <= 3; k++) {
            m[k] = m[k - 1];
            level(lv[k - 1], m[k - 1]);
            addpoint(m[k], lv[k - 1]);
            check_recon(m[k]);
        }
        iwt(m[3]);
        interp(mt, m[3]);
    }
    // AWCM_2D::Prepare (AWCM_2D.h:52-79) on the scalar field (component 0):
    // PrepareEuler without the mask_temp copy and the interpolation
    void prepare_scalar(double tol) {
        fwt(m[0], tol);
        band(m[0]);
        adjacent(m[0]);
        check_recon(m[0]);
        for (int k = 1; k <= 3; k++) {
            m[k] = m[k - 1];
            level(lv[k - 1], m[k - 1]);
            addpoint(m[k], lv[k - 1]);
            check_recon(m[k]);
        }
        iwt(m[3]);
    }
    // AWCM_MANAGER_2D::RemoveResidueEuler (AWCM_MANAGER_2D.h:267-280): interior only
    void post() {
#pragma omp parallel for
        for (int j = 1; j <= Ny - 2; j++)
            for (int i = 1; i <= Nx - 2; i++)
                if (mk(m[0], i, j) == 0) { double* v = q(Q, i, j); v[0] = v[1] = v[2] = v[3] = 0.0; }
    }

    // ---- the scalar driver's stages (Wavelet/main_2d_advect.cpp:55-200) on
    // component 0: box advection with velocity (1, 1).  Q2 / Q1 / Q0 keep
    // their values between iterations, as the driver's fields do.
    // PROBLEM_BOX_ADVECTION.h:78-151: periodic sides (rows 0..Ny-1 for the x
    // ghosts, columns 0..Nx-1 for the y ghosts); the corners are not written.
    void periodic(std::vector<double>& F) {
        for (int j = 0; j <= Ny - 1; j++)
            for (int c = 0; c < GW; c++) {
                q(F, -1 - c, j)[0] = q(F, Nx - 1 - c, j)[0];
                q(F, Nx + c, j)[0] = q(F, c, j)[0];
            }
        for (int i = 0; i <= Nx - 1; i++)
            for (int c = 0; c < GW; c++) {
                q(F, i, -1 - c)[0] = q(F, i, Ny - 1 - c)[0];
                q(F, i, Ny + c)[0] = q(F, i, c)[0];
            }
    }
    // dF + dG at (i, j): WENO5 right reconstructions on the finest level
    // (main_2d_advect.cpp:63-73), AWCM_2D::Differentiate and its one-sided
    // forms at spacing d below it (:74-103, AWCM_2D.h:95-110)
    double adv_rhs(std::vector<double>& F, int i, int j, int lev, long& oobs) {
        auto f = [&](int a, int b) { return q(F, a, b)[0]; };
        if (lev == maxlev) {
            double a = wenojs_right(f(i - 3, j), f(i - 2, j), f(i - 1, j), f(i, j), f(i + 1, j));
            double b = wenojs_right(f(i - 2, j), f(i - 1, j), f(i, j), f(i + 1, j), f(i + 2, j));
            const double dF = b - a;
            a = wenojs_right(f(i, j - 3), f(i, j - 2), f(i, j - 1), f(i, j), f(i, j + 1));
            b = wenojs_right(f(i, j - 2), f(i, j - 1), f(i, j), f(i, j + 1), f(i, j + 2));
            const double dG = b - a;
            return dF + dG;
        }
        const int d = grid[lev].dx;
        // the reference reads outside its arrays here; such a point is counted
        auto in = [&](int a, int b) { return a >= -GW && a <= Nx - 1 + GW && b >= -GW && b <= Ny - 1 + GW; };
        double dF, dG;
        if (i <= d) {
            if (!in(i - d, j) || !in(i + 2 * d, j)) { oobs++; return 0; }
            dF = D1 * (f(i + d, j) - f(i - d, j)) + D2 * (f(i + 2 * d, j) - f(i - d, j));
        } else if (i >= Nx - 1 - d) {
            if (!in(i - 2 * d, j) || !in(i + d, j)) { oobs++; return 0; }
            dF = D1 * (f(i + d, j) - f(i - d, j)) + D2 * (f(i + d, j) - f(i - 2 * d, j));
        } else {
            if (!in(i - 2 * d, j) || !in(i + 2 * d, j)) { oobs++; return 0; }
            dF = D1 * (f(i + d, j) - f(i - d, j)) + D2 * (f(i + 2 * d, j) - f(i - 2 * d, j));
        }
        if (j <= d) {
            if (!in(i, j - d) || !in(i, j + 2 * d)) { oobs++; return 0; }
            dG = D1 * (f(i, j + d) - f(i, j - d)) + D2 * (f(i, j + 2 * d) - f(i, j - d));
        } else if (j >= Ny - 1 - d) {
            if (!in(i, j - 2 * d) || !in(i, j + d)) { oobs++; return 0; }
            dG = D1 * (f(i, j + d) - f(i, j - d)) + D2 * (f(i, j + d) - f(i, j - 2 * d));
        } else {
            if (!in(i, j - 2 * d) || !in(i, j + 2 * d)) { oobs++; return 0; }
            dG = D1 * (f(i, j + d) - f(i, j - d)) + D2 * (f(i, j + 2 * d) - f(i, j - 2 * d));
        }
        return dF + dG;
    }
    // everything of one driver iteration after awcm.Prepare (main_2d_advect.cpp:50-198)
    void advect_stages(double dt) {
        long oobs = 0;
        periodic(Q);
#pragma omp parallel for reduction(+:oobs)
        for (int j = 0; j <= Ny - 1; j++)
            for (int i = 0; i <= Nx - 1; i++) {
                if (mk(m[2], i, j) == 0) continue;
                const int lev = lv[2][i + size_t(Nx) * j];
                const double h = dx * grid[lev].dx;
                const double r = adv_rhs(Q, i, j, lev, oobs);
                q(Q2, i, j)[0] = q(Q, i, j)[0] - dt / h * r;
            }
        periodic(Q2);
#pragma omp parallel for reduction(+:oobs)
        for (int j = 0; j <= Ny - 1; j++)
            for (int i = 0; i <= Nx - 1; i++) {
                if (mk(m[1], i, j) == 0) continue;
                const int lev = lv[1][i + size_t(Nx) * j];
                const double h = dx * grid[lev].dx;
                const double r = adv_rhs(Q2, i, j, lev, oobs);
                q(Q1, i, j)[0] = 0.75 * q(Q, i, j)[0] + 0.25 * (q(Q2, i, j)[0] - dt / h * r);
            }
        periodic(Q1);
#pragma omp parallel for reduction(+:oobs)
        for (int j = 0; j <= Ny - 1; j++)
            for (int i = 0; i <= Nx - 1; i++) {
                if (mk(m[0], i, j) == 0) continue;
                const int lev = lv[0][i + size_t(Nx) * j];
                const double h = dx * grid[lev].dx;
                const double r = adv_rhs(Q1, i, j, lev, oobs);
                q(Q0, i, j)[0] = q(Q1, i, j)[0] - dt / h * r;
            }
        oob += oobs;
        for (int j = 0; j <= Ny - 1; j++)
            for (int i = 0; i <= Nx - 1; i++)
                if (mk(m[0], i, j)) q(Q, i, j)[0] = (1.0 / 3.0 * q(Q, i, j)[0] + 2.0 / 3.0 * q(Q0, i, j)[0]);
        // AWCM_2D::PostProcess -> AWCM_MANAGER_2D::RemoveResidue (AWCM_MANAGER_2D.h:257-266)
        for (int j = 1; j <= Ny - 2; j++)
            for (int i = 1; i <= Nx - 2; i++)
                if (mk(m[0], i, j) == 0) q(Q, i, j)[0] = 0.0;
    }
    // BOUNDARY_MANAGER_QUAD::SetBoundaryCondition (PROBLEM_QUADRANTS.h:123-203) and
    // BOUNDARY_MANAGER_DMACH::SetBoundaryCondition(t) (PROBLEM_DOUBLEMACH.h:106-245)
    void bc(std::vector<double>& F, double t) {
        if (bc_kind == BC_QUAD) {
            for (int j = 0; j <= Ny - 1; j++) {
                for (int i = -1; i >= -GW; i--) std::memcpy(q(F, i, j), q(F, 0, j), 32);
                for (int i = Nx; i <= Nx - 1 + GW; i++) std::memcpy(q(F, i, j), q(F, Nx - 1, j), 32);
            }
            for (int i = 0; i <= Nx - 1; i++) {
                for (int j = -1; j >= -GW; j--) std::memcpy(q(F, i, j), q(F, i, 0), 32);
                for (int j = Ny; j <= Ny - 1 + GW; j++) std::memcpy(q(F, i, j), q(F, i, Ny - 1), 32);
            }
            return;
        }
        const double g14 = 1.4;
        const double sqrt3 = std::sqrt(3.0);
        // left inflow
        {
            double rho = 8, u = 8.25 * std::sqrt(3.0) * 0.5, v = -8.25 * 0.5, p = 116.5;
            for (int j = 0; j <= Ny - 1; j++)
                for (int i = -1; i >= -GW; i--) {
                    double* o = q(F, i, j);
                    o[0] = rho; o[1] = rho * u; o[2] = rho * v; o[3] = E2prim(g14, rho, u, v, p);
                }
        }
        for (int j = 0; j <= Ny - 1; j++)
            for (int i = Nx; i <= Nx - 1 + GW; i++) std::memcpy(q(F, i, j), q(F, Nx - 1, j), 32);
        const double Epost = E2prim(g14, 8, 8.25 * sqrt3 * 0.5, -8.25 * 0.5, 116.5);
        for (int i = 0; i <= Nx - 1; i++)
            for (int j = -1; j >= -GW; j--) {
                const double x = x0 + (i - 0) * dx;
                double* o = q(F, i, j);
                const double* b = q(F, i, 0);
                if (x > 1.0 / 6.0) { o[0] = b[0]; o[1] = b[1]; o[2] = -b[2]; o[3] = b[3]; }
                else { o[0] = 8; o[1] = 8 * 8.25 * sqrt3 * 0.5; o[2] = -8 * 8.25 * 0.5; o[3] = Epost; }
            }
        const double Epre = E2prim(g14, 1.4, 0, 0, 1);
        for (int i = 0; i <= Nx - 1; i++)
            for (int j = Ny; j <= Ny - 1 + GW; j++) {
                const double x = x0 + (i - 0) * dx;
                double* o = q(F, i, j);
                if (x < 1.0 / 6.0 + (1 + 20 * t) / sqrt3) { o[0] = 8; o[1] = 8 * 8.25 * sqrt3 * 0.5; o[2] = -8 * 8.25 * 0.5; o[3] = Epost; }
                else { o[0] = 1.4; o[1] = 1.4 * 0; o[2] = 1.4 * 0; o[3] = Epre; }
            }
    }
    void alphas(std::vector<double>& F, int k) {
        std::fill(ax.begin(), ax.end(), 0.0);
        std::fill(ay.begin(), ay.end(), 0.0);
#pragma omp parallel for
        for (int j = 0; j <= Ny - 1; j++)
            for (int i = 0; i <= Nx - 1; i++) {
                if (mk(m[k], i, j) == 0) continue;
                ax[j] = MAXd(ax[j], SRx2(gamma, q(F, i, j)));
            }
#pragma omp parallel for
        for (int i = 0; i <= Nx - 1; i++)
            for (int j = 0; j <= Ny - 1; j++) {
                if (mk(m[k], i, j) == 0) continue;
                ay[i] = MAXd(ay[i], SRy2(gamma, q(F, i, j)));
            }
    }
    // dF + dG at (i,j) of one stage (main_2d_euler(dmach).cpp:108-161)
    void rhs(std::vector<double>& F, int i, int j, int lev, double* out, long& oobl) {
        const int d = grid[lev].dx;
        auto QF = [&](int ii, int jj) -> const double* {
            if (ii < -GW || ii > Nx - 1 + GW || jj < -GW || jj > Ny - 1 + GW) oobl++;
            return q(F, ii, jj);
        };
        double dFv[4], dGv[4];
        if (lev == maxlev) {
            double a[4], b[4];
            const double* x0s[6] = {QF(i - 3, j), QF(i - 2, j), QF(i - 1, j), QF(i, j), QF(i + 1, j), QF(i + 2, j)};
            const double* x1s[6] = {QF(i - 2, j), QF(i - 1, j), QF(i, j), QF(i + 1, j), QF(i + 2, j), QF(i + 3, j)};
            flux2(1.4, 0, x0s, ax[j], a);
            flux2(1.4, 0, x1s, ax[j], b);
            for (int k = 0; k < 4; k++) dFv[k] = b[k] - a[k];
            const double* y0s[6] = {QF(i, j - 3), QF(i, j - 2), QF(i, j - 1), QF(i, j), QF(i, j + 1), QF(i, j + 2)};
            const double* y1s[6] = {QF(i, j - 2), QF(i, j - 1), QF(i, j), QF(i, j + 1), QF(i, j + 2), QF(i, j + 3)};
            flux2(1.4, 1, y0s, ay[i], a);
            flux2(1.4, 1, y1s, ay[i], b);
            for (int k = 0; k < 4; k++) dGv[k] = b[k] - a[k];
        } else {
            double Fll[4], Fl[4], Fc[4], Fr[4], Frr[4];
            if (i <= d) {
                Fu2(gamma, QF(i - d, j), Fl); Fu2(gamma, QF(i, j), Fc); Fu2(gamma, QF(i + d, j), Fr); Fu2(gamma, QF(i + 2 * d, j), Frr);
                for (int k = 0; k < 4; k++) dFv[k] = D1 * (Fr[k] - Fl[k]) + D2 * (Frr[k] - Fl[k]);
            } else if (i >= Nx - 1 - d) {
                Fu2(gamma, QF(i - 2 * d, j), Fll); Fu2(gamma, QF(i - d, j), Fl); Fu2(gamma, QF(i, j), Fc); Fu2(gamma, QF(i + d, j), Fr);
                for (int k = 0; k < 4; k++) dFv[k] = D1 * (Fr[k] - Fl[k]) + D2 * (Fr[k] - Fll[k]);
            } else {
                Fu2(gamma, QF(i - 2 * d, j), Fll); Fu2(gamma, QF(i - d, j), Fl); Fu2(gamma, QF(i, j), Fc);
                Fu2(gamma, QF(i + d, j), Fr); Fu2(gamma, QF(i + 2 * d, j), Frr);
                for (int k = 0; k < 4; k++) dFv[k] = D1 * (Fr[k] - Fl[k]) + D2 * (Frr[k] - Fll[k]);
            }
            double Gbb[4], Gb[4], Gc[4], Gt[4], Gtt[4];
            if (j <= d) {
                Gu2(gamma, QF(i, j - d), Gb); Gu2(gamma, QF(i, j), Gc); Gu2(gamma, QF(i, j + d), Gt); Gu2(gamma, QF(i, j + 2 * d), Gtt);
                for (int k = 0; k < 4; k++) dGv[k] = D1 * (Gt[k] - Gb[k]) + D2 * (Gtt[k] - Gb[k]);
            } else if (j >= Ny - 1 - d) {
                Gu2(gamma, QF(i, j - 2 * d), Gbb); Gu2(gamma, QF(i, j - d), Gb); Gu2(gamma, QF(i, j), Gc); Gu2(gamma, QF(i, j + d), Gt);
                for (int k = 0; k < 4; k++) dGv[k] = D1 * (Gt[k] - Gb[k]) + D2 * (Gt[k] - Gbb[k]);
            } else {
                Gu2(gamma, QF(i, j - 2 * d), Gbb); Gu2(gamma, QF(i, j - d), Gb); Gu2(gamma, QF(i, j), Gc);
                Gu2(gamma, QF(i, j + d), Gt); Gu2(gamma, QF(i, j + 2 * d), Gtt);
                for (int k = 0; k < 4; k++) dGv[k] = D1 * (Gt[k] - Gb[k]) + D2 * (Gtt[k] - Gbb[k]);
            }
        }
        for (int k = 0; k < 4; k++) out[k] = dFv[k] + dGv[k];
    }
    // One iteration of main_2d_euler(dmach|quad).cpp:56-350.
    double step(double tol, double cfl, const double* pt, int npt, int* cnt, double* t) {
        prepare(tol);
        double vmax = 0;
#pragma omp parallel
        {
            double vl = 0;
#pragma omp for
            for (int j = 0; j <= Ny - 1; j++)
                for (int i = 0; i <= Nx - 1; i++) {
                    if (mk(m[3], i, j) == 0) continue;
                    vl = MAXd(vl, SRx2(gamma, q(Q, i, j)) + SRy2(gamma, q(Q, i, j)));
                }
#pragma omp critical
            vmax = MAXd(vmax, vl);
        }
        double dt = cfl * dx / vmax;
        bool do_print = false;
        if (*cnt < npt && *t + dt > pt[*cnt]) { dt = pt[*cnt] - *t; do_print = true; }
        const double T = *t;
        long oobs = 0;
        // stage 1: mask2 / level2, reads Q -> Q2
        bc(Q, T);
        alphas(Q, 3);
#pragma omp parallel for reduction(+:oobs)
        for (int j = 0; j <= Ny - 1; j++)
            for (int i = 0; i <= Nx - 1; i++) {
                if (mk(m[2], i, j) == 0) continue;
                const int lev = lv[2][i + size_t(Nx) * j];
                double r[4]; rhs(Q, i, j, lev, r, oobs);
                const double h = dx * grid[lev].dx;
                double* o = q(Q2, i, j); const double* s = q(Q, i, j);
                for (int k = 0; k < 4; k++) o[k] = s[k] - dt / h * r[k];
            }
        // stage 2: mask1 / level1, reads Q2 -> Q1
        bc(Q2, T);
        alphas(Q2, 2);
#pragma omp parallel for reduction(+:oobs)
        for (int j = 0; j <= Ny - 1; j++)
            for (int i = 0; i <= Nx - 1; i++) {
                if (mk(m[1], i, j) == 0) continue;
                const int lev = lv[1][i + size_t(Nx) * j];
                double r[4]; rhs(Q2, i, j, lev, r, oobs);
                const double h = dx * grid[lev].dx;
                double* o = q(Q1, i, j); const double* s = q(Q, i, j); const double* s2 = q(Q2, i, j);
                for (int k = 0; k < 4; k++) o[k] = 0.75 * s[k] + 0.25 * (s2[k] - dt / h * r[k]);
            }
        // stage 3: mask0 / level0, reads Q1 -> Q0
        bc(Q1, T);
        alphas(Q1, 1);
#pragma omp parallel for reduction(+:oobs)
        for (int j = 0; j <= Ny - 1; j++)
            for (int i = 0; i <= Nx - 1; i++) {
                if (mk(m[0], i, j) == 0) continue;
                const int lev = lv[0][i + size_t(Nx) * j];
                double r[4]; rhs(Q1, i, j, lev, r, oobs);
                const double h = dx * grid[lev].dx;
                double* o = q(Q0, i, j); const double* s1 = q(Q1, i, j);
                for (int k = 0; k < 4; k++) o[k] = s1[k] - dt / h * r[k];
            }
        oob += oobs;
#pragma omp parallel for
        for (int j = 0; j <= Ny - 1; j++)
            for (int i = 0; i <= Nx - 1; i++)
                if (mk(m[0], i, j)) {
                    double* o = q(Q, i, j); const double* s0 = q(Q0, i, j);
                    for (int k = 0; k < 4; k++) o[k] = (o[k] + 2.0 * s0[k]) / 3.0;
                }
        post();
        *t += dt;
        if (do_print) (*cnt)++;
        return dt;
    }
};

}  // namespace

// ===========================================================================
// C entry points (loaded with ctypes by the tests and bench.py)
// ===========================================================================
extern "C" {

// ---- initial conditions (Problems/1D, Problems/2D) -------------------------
// SHOCKTUBE::Initialize / ENTROPY_INTERACTING::Initialize generalised to a
// domain [xl, xu] and a split point (PROBLEM_SHOCKTUBE.h:14-54): cell-centred
// grid, ghost 3.  kind 0: two-state Riemann problem, left if x <= split
// (states rho,u,p);  kind 1: Shu-Osher, left if x < split, right rho = 1 +
// amp*sin(5x), u = 0, p = 1 (PROBLEM_ENTROPY_INTERACTING.h:14-49).
int orc_ic1(int kind, int res, double xl, double xu, double split, const double* sl, const double* sr,
            double amp, double gamma, double* q, double* dx_out, double* x0_out) {
    double dxx = (xu - xl) / ((res + 1) - 1);
    double xmin = xl + 0.5 * dxx, xmax = xu - 0.5 * dxx;
    double dx = (xmax - xmin) / (res - 1);
    for (int i = -GW; i <= res - 1 + GW; i++) {
        const double x = xmin + dx * (i - 0);
        double rho, u, p;
        if (kind == 0) {
            if (x <= split) { rho = sl[0]; u = sl[1]; p = sl[2]; }
            else { rho = sr[0]; u = sr[1]; p = sr[2]; }
        } else {
            if (x < split) { rho = sl[0]; u = sl[1]; p = sl[2]; }
            else { rho = 1 + amp * std::sin(5 * x); u = 0; p = 1; }
        }
        double* o = &q[(i + GW) * 3];
        o[0] = rho; o[1] = rho * u; o[2] = E1prim(gamma, rho, u, p);
    }
    *dx_out = dx; *x0_out = xmin;
    return 0;
}

// QUADRANTS::Initialize (PROBLEM_QUADRANTS.h:15-81) with split point s and
// states st[4][4] = (rho,u,v,p) for quadrants 1..4; DOUBLEMACH::Initialize
// (PROBLEM_DOUBLEMACH.h:13-63) with the pre-shock density as a parameter;
// kind 2: BOX_ADVECTION's scalar Gaussian on the cell-centre grid [xl, xu] x
// [yl, yu] (component 0; the others zero).
int orc_ic2(int kind, int nx, int ny, double xl, double xu, double yl, double yu, double s,
            const double* st, double gamma, double* q, double* dx_out, double* dy_out) {
    const double dx = (xu - xl) / (nx - 1), dy = (yu - yl) / (ny - 1);
    const int P = nx + 2 * GW;
    for (int j = -GW; j <= ny - 1 + GW; j++)
        for (int i = -GW; i <= nx - 1 + GW; i++) {
            const double x = xl + (i - 0) * dx, y = yl + (j - 0) * dy;
            double rho, u, v, p;
            if (kind == 0) {
                int qd;
                if (x >= s && y >= s) qd = 0;
                else if (x < s && y >= s) qd = 1;
                else if (x < s && y < s) qd = 2;
                else qd = 3;
                rho = st[qd * 4 + 0]; u = st[qd * 4 + 1]; v = st[qd * 4 + 2]; p = st[qd * 4 + 3];
            } else if (kind == 2) {
                // BOX_ADVECTION (PROBLEM_BOX_ADVECTION.h:15-47): component 0 only
                double* o = &q[(size_t(i + GW) + size_t(P) * (j + GW)) * 4];
                o[0] = std::exp(-50 * SQRd(x - 0.5) - 50 * SQRd(y - 0.5));
                o[1] = o[2] = o[3] = 0;
                continue;
            } else {
                if (y > std::sqrt(3.0) * (x - 1.0 / 6.0)) {
                    rho = 8; u = 8.25 * std::sqrt(3.0) * 0.5; v = -8.25 * 0.5; p = 116.5;
                } else { rho = s; u = 0; v = 0; p = 1; }
            }
            double* o = &q[(size_t(i + GW) + size_t(P) * (j + GW)) * 4];
            o[0] = rho; o[1] = rho * u; o[2] = rho * v; o[3] = E2prim(gamma, rho, u, v, p);
        }
    *dx_out = dx; *dy_out = dy;
    return 0;
}

// FNV-1a (64-bit) over n bytes, continuing from h: the hash the class-API
// probes print (tests/native/api_probe*.cpp)
uint64_t orc_fnv1a(const uint8_t* p, long long n, uint64_t h) {
    for (long long k = 0; k < n; k++) {
        h ^= p[k];
        h *= 1099511628211ull;
    }
    return h;
}

// ---- 1-D solver ---------------------------------------------------------------
void* orc1_create(int num_levels, int nx_coarse, double gamma, double dx) {
    return new Solver1D(num_levels, nx_coarse, gamma, dx);
}
void orc1_destroy(void* h) { delete static_cast<Solver1D*>(h); }
int orc1_nx(void* h) { return static_cast<Solver1D*>(h)->Nx; }
void orc1_set_q(void* h, const double* q) { auto* s = static_cast<Solver1D*>(h); std::memcpy(s->Q.data(), q, s->Q.size() * 8); }
void orc1_get_q(void* h, double* q) { auto* s = static_cast<Solver1D*>(h); std::memcpy(q, s->Q.data(), s->Q.size() * 8); }
// masks: 5 x (Nx+6) bytes in the order mask_temp, mask0..3; levels: 4 x Nx int32
void orc1_get_masks(void* h, uint8_t* masks, int32_t* levels) {
    auto* s = static_cast<Solver1D*>(h);
    const size_t n = s->Nx + 2 * GW;
    std::memcpy(masks, s->mt.data(), n);
    for (int k = 0; k < 4; k++) {
        std::memcpy(masks + (k + 1) * n, s->m[k].data(), n);
        for (int i = 0; i < s->Nx; i++) levels[k * s->Nx + i] = s->lv[k][i];
    }
}
void orc1_set_mask0(void* h, const uint8_t* m0) { auto* s = static_cast<Solver1D*>(h); std::memcpy(s->m[0].data(), m0, s->m[0].size()); }
void orc1_prepare(void* h, double tol) { static_cast<Solver1D*>(h)->prepare(tol); }
// face fluxes of the stages: 0 WENO-Z pair, 1 WENO5-JS pair (the shipped driver)
void orc1_set_flux(void* h, int kind) { static_cast<Solver1D*>(h)->flux_js = kind == 1; }
// the scalar path: the field is component 0 ((Nx+6) doubles)
void orc1_set_scalar(void* h, const double* q) {
    auto* s = static_cast<Solver1D*>(h);
    for (int i = 0; i < s->Nx + 6; i++) s->Q[3 * i] = q[i];
}
void orc1_get_scalar(void* h, double* q) {
    auto* s = static_cast<Solver1D*>(h);
    for (int i = 0; i < s->Nx + 6; i++) q[i] = s->Q[3 * i];
}
void orc1_prepare_scalar(void* h, double tol) { static_cast<Solver1D*>(h)->prepare_scalar(tol); }
// WAVELET_1D::InterpolateEuler(q, mra, from, to) (WAVELET_1D.h:393-431: momentum
// and energy); a null `to` is InterpolateEuler(q, mra, from) (:433-473): every
// point, all three components
void orc1_interpolate_euler(void* h, const uint8_t* mask_from, const uint8_t* mask_to) {
    auto* s = static_cast<Solver1D*>(h);
    const size_t n = size_t(s->Nx) + 6;
    std::vector<uint8_t> from(mask_from, mask_from + n), to(n, 1);
    if (mask_to) {
        to.assign(mask_to, mask_to + n);
        s->interp(from, to, 1, 3);
    } else {
        s->interp(from, to, 0, 3);
    }
}
// WAVELET_1D::Interpolate(q, mra, from, to) (WAVELET_1D.h:198-233); a null `to`
// is Interpolate(q, mra, from) (:235-275): every point
void orc1_interpolate_scalar(void* h, const uint8_t* mask_from, const uint8_t* mask_to) {
    auto* s = static_cast<Solver1D*>(h);
    const size_t n = size_t(s->Nx) + 6;
    std::vector<uint8_t> from(mask_from, mask_from + n), to(n, 1);
    if (mask_to) to.assign(mask_to, mask_to + n);
    s->interp(from, to, 0, 1);
}
double orc1_step(void* h, double tol, double cfl, const double* pt, int npt, int* cnt, double* t) {
    return static_cast<Solver1D*>(h)->step(tol, cfl, pt, npt, cnt, t);
}
long orc1_oob(void* h) { return static_cast<Solver1D*>(h)->oob; }
// Frozen-trajectory tape: record the following orc1_step calls, then replay
// them from any state (finite-difference checks of the discrete adjoint).
void orc1_record(void* h, int on) {
    auto* s = static_cast<Solver1D*>(h);
    s->recording = on != 0;
    if (on) s->tape.clear();
}
int orc1_tape_steps(void* h) { return int(static_cast<Solver1D*>(h)->tape.size()); }
void orc1_tape_scalars(void* h, double* out) {
    auto* s = static_cast<Solver1D*>(h);
    for (size_t n = 0; n < s->tape.size(); n++) {
        out[4 * n] = s->tape[n].dt;
        for (int k = 0; k < 3; k++) out[4 * n + 1 + k] = s->tape[n].al[k];
    }
}
void orc1_replay(void* h, int first, int count) {
    auto* s = static_cast<Solver1D*>(h);
    for (int n = first; n < first + count && n < int(s->tape.size()); n++) s->replay(n);
}
// WENOZ_EULER_SOLVER_1D::ComputeFlux on explicit states (6 x 3 doubles)
void orc_flux1(double gamma, const double* states, double alpha, double* flux) {
    const double* s[6] = {states, states + 3, states + 6, states + 9, states + 12, states + 15};
    flux1(gamma, s, alpha, flux);
}
// WENO_EULER_SOLVER_1D::ComputeFlux on 6 x 3 doubles
void orc_flux1_js(double gamma, const double* states, double alpha, double* flux) {
    const double* s[6] = {states, states + 3, states + 6, states + 9, states + 12, states + 15};
    flux1(gamma, s, alpha, flux, true);
}
// WENOZ_EULER_SOLVER_2D::ComputeFluxX (dir 0) / ComputeFluxY (dir 1) on 6 x 4 doubles
void orc_flux2(double gamma, int dir, const double* states, double alpha, double* flux) {
    const double* s[6] = {states, states + 4, states + 8, states + 12, states + 16, states + 20};
    flux2(gamma, dir, s, alpha, flux);
}
double orc_wenoz_left(double a, double b, double c, double d, double e) { return wenoz_left(a, b, c, d, e); }
double orc_wenoz_right(double a, double b, double c, double d, double e) { return wenoz_right(a, b, c, d, e); }
void orc_weights(double* w) { Weights W; for (int i = 0; i < 4; i++) for (int j = 0; j < 5; j++) w[i * 5 + j] = W.W[i][j]; }

// ---- 2-D solver ---------------------------------------------------------------
void* orc2_create(int num_levels, int nx_coarse, int ny_coarse, double gamma, double dx, double x0, int bc_kind) {
    return new Solver2D(num_levels, nx_coarse, ny_coarse, gamma, dx, x0, bc_kind);
}
void orc2_destroy(void* h) { delete static_cast<Solver2D*>(h); }
int orc2_nx(void* h) { return static_cast<Solver2D*>(h)->Nx; }
int orc2_ny(void* h) { return static_cast<Solver2D*>(h)->Ny; }
void orc2_set_q(void* h, const double* q) { auto* s = static_cast<Solver2D*>(h); std::memcpy(s->Q.data(), q, s->Q.size() * 8); }
void orc2_get_q(void* h, double* q) { auto* s = static_cast<Solver2D*>(h); std::memcpy(q, s->Q.data(), s->Q.size() * 8); }
void orc2_get_masks(void* h, uint8_t* masks, int32_t* levels) {
    auto* s = static_cast<Solver2D*>(h);
    const size_t n = s->m[0].size(), nl = size_t(s->Nx) * s->Ny;
    std::memcpy(masks, s->mt.data(), n);
    for (int k = 0; k < 4; k++) {
        std::memcpy(masks + (k + 1) * n, s->m[k].data(), n);
        for (size_t i = 0; i < nl; i++) levels[k * nl + i] = s->lv[k][i];
    }
}
void orc2_set_mask0(void* h, const uint8_t* m0) { auto* s = static_cast<Solver2D*>(h); std::memcpy(s->m[0].data(), m0, s->m[0].size()); }
void orc2_prepare(void* h, double tol) { static_cast<Solver2D*>(h)->prepare(tol); }
// the scalar field of AWCM_2D::Prepare / WAVELET_2D::Interpolate lives in
// component 0 (reference layout, (Nx+6) x (Ny+6) doubles)
void orc2_set_scalar(void* h, const double* q) {
    auto* s = static_cast<Solver2D*>(h);
    for (size_t k = 0; k < s->m[0].size(); k++) s->Q[4 * k] = q[k];
}
void orc2_get_scalar(void* h, double* q) {
    auto* s = static_cast<Solver2D*>(h);
    for (size_t k = 0; k < s->m[0].size(); k++) q[k] = s->Q[4 * k];
}
void orc2_prepare_scalar(void* h, double tol) { static_cast<Solver2D*>(h)->prepare_scalar(tol); }
// n iterations of the scalar driver (main_2d_advect.cpp:46-199): Prepare, then
// the stages, combine and PostProcess; the stages alone with n_prepare = 0
void orc2_advect(void* h, double tol, double dt, int n, int with_prepare) {
    auto* s = static_cast<Solver2D*>(h);
    for (int k = 0; k < n; k++) {
        if (with_prepare) s->prepare_scalar(tol);
        s->advect_stages(dt);
    }
}
void orc2_interpolate_scalar(void* h, const uint8_t* mask_from, const uint8_t* mask_to) {
    auto* s = static_cast<Solver2D*>(h);
    std::vector<uint8_t> from(mask_from, mask_from + s->m[0].size()), to(mask_to, mask_to + s->m[0].size());
    s->interp_scalar(from, to);
}
void orc2_interpolate_euler(void* h, const uint8_t* mask_from, const uint8_t* mask_to) {
    auto* s = static_cast<Solver2D*>(h);
    std::vector<uint8_t> from(mask_from, mask_from + s->m[0].size()), to(mask_to, mask_to + s->m[0].size());
    s->interp(from, to);
}
void orc2_interpolate_all(void* h, const uint8_t* mask_from) {
    auto* s = static_cast<Solver2D*>(h);
    std::vector<uint8_t> from(mask_from, mask_from + s->m[0].size());
    s->interpolate_all(from);
}
double orc2_step(void* h, double tol, double cfl, const double* pt, int npt, int* cnt, double* t) {
    return static_cast<Solver2D*>(h)->step(tol, cfl, pt, npt, cnt, t);
}
long orc2_oob(void* h) { return static_cast<Solver2D*>(h)->oob; }

// AWCM_MANAGER_1D / AWCM_MANAGER_2D statics on caller arrays: op 0
// AddAdjacentZone, 1 CheckReconstruction, 2 AddPointForDifferentiation (level
// read), 3 ComputeLevel (level in/out).  Ghosted masks, unghosted int levels.
void orc1_mask_op(void* h, int op, uint8_t* mask, int32_t* level) {
    auto* s = static_cast<Solver1D*>(h);
    std::vector<uint8_t> m(mask, mask + s->Nx + 2 * GW);
    std::vector<int> lv(size_t(s->Nx), 0);
    if (level) for (int i = 0; i < s->Nx; i++) lv[i] = level[i];
    if (op == 0) s->adjacent(m);
    else if (op == 1) s->check_recon(m);
    else if (op == 2) s->addpoint(m, lv);
    else s->level(lv, m);
    std::memcpy(mask, m.data(), m.size());
    if (op == 3) for (int i = 0; i < s->Nx; i++) level[i] = lv[i];
}
void orc2_mask_op(void* h, int op, uint8_t* mask, int32_t* level) {
    auto* s = static_cast<Solver2D*>(h);
    const size_t np = size_t(s->P) * (s->Ny + 2 * GW), nl = size_t(s->Nx) * s->Ny;
    std::vector<uint8_t> m(mask, mask + np);
    std::vector<int> lv(nl, 0);
    if (level) for (size_t t = 0; t < nl; t++) lv[t] = level[t];
    if (op == 0) s->adjacent(m);
    else if (op == 1) s->check_recon(m);
    else if (op == 2) s->addpoint(m, lv);
    else s->level(lv, m);
    std::memcpy(mask, m.data(), np);
    if (op == 3) for (size_t t = 0; t < nl; t++) level[t] = lv[t];
}
// AWCM_MANAGER_1D::RemoveResidue (AWCM_MANAGER_1D.h:135-157): finest points
void orc1_remove_residue(void* h, double* q, int ncomp, const uint8_t* mask) {
    auto* s = static_cast<Solver1D*>(h);
    for (int i = 0; i <= s->Nx - 1; i++)
        if (!mask[i + GW]) for (int c = 0; c < ncomp; c++) q[size_t(i + GW) * ncomp + c] = 0.0;
}
// AWCM_MANAGER_2D::RemoveResidue / RemoveResidueEuler (AWCM_MANAGER_2D.h:257-280): interior points
void orc2_remove_residue(void* h, double* q, int ncomp, const uint8_t* mask) {
    auto* s = static_cast<Solver2D*>(h);
    for (int j = 1; j <= s->Ny - 2; j++)
        for (int i = 1; i <= s->Nx - 2; i++)
            if (!mask[s->id(i, j)]) for (int c = 0; c < ncomp; c++) q[s->id(i, j) * ncomp + c] = 0.0;
}

}  // extern "C"
