// Probe of the reference's C++ class surface on the AWCM Euler path.
// TEST INFRASTRUCTURE.  One source, built twice:
//   * against include/compat/ (this library)          -> tests/native/_build/awcm_api_probe
//   * against the reference's own headers, in place    -> oracle/_ref/awcm_api_probe_ref
//     (-DAWCM_PROBE_REFERENCE, oracle/Makefile)
// It names only classes, members and methods both sides define, and prints
// one line per check with FNV-1a hashes and hex doubles, so the two builds
// must print identical text.  tests/golden/api_probe_*.txt hold what the
// reference build printed.
//
//   awcm_api_probe host <dir>     host-side classes; writes Print() files in <dir>
//   awcm_api_probe prepare        AWCM_1D/2D::PrepareEuler on five problems
//   awcm_api_probe manager        AWCM_MANAGER_1D/2D static members
//   awcm_api_probe loop <steps>   the drivers' time loop through the class API
#include <cinttypes>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <type_traits>
#include <unistd.h>

#include "FIELD_UNIFORM_1D.h"
#ifdef AWCM_PROBE_REFERENCE
// FIELD_UNIFORM_1D<T>::AssignAllValues forwards to ARRAY_1D::AssignAllValue,
// which the reference does not define (FIELD_UNIFORM_1D.h:65-68); AWCM_1D's
// constructor needs it for bool.
template <>
void FIELD_UNIFORM_1D<bool>::AssignAllValues(const bool& v) {
    for (int i = 0; i < arr.i_res; i++) arr.val[i] = v;
}
#endif
#include "ARRAY.h"
#include "EULER_EQUATION.h"
#include "OMP_MANAGER.h"
#include "Problems/1D/PROBLEM_ENTROPY_INTERACTING.h"
#include "Problems/1D/PROBLEM_SHOCKTUBE.h"
#include "Problems/2D/PROBLEM_DOUBLEMACH.h"
#include "Problems/2D/PROBLEM_QUADRANTS.h"
#include "ROESOLVER.h"
#include "WENO.h"
#include "WENOZ.h"
#include "Wavelet/AWCM_1D.h"
#include "Wavelet/AWCM_2D.h"

namespace {

typedef VECTOR_3D<double> V3;
typedef VECTOR_4D<double> V4;

struct Hash {
    uint64_t h = 1469598103934665603ull;
    void byte(uint8_t b) {
        h ^= b;
        h *= 1099511628211ull;
    }
    void u64(uint64_t v) {
        for (int k = 0; k < 8; k++) byte(uint8_t(v >> (8 * k)));
    }
    void f(double x) {
        uint64_t v;
        std::memcpy(&v, &x, 8);
        u64(v);
    }
    void i(int v) { u64(uint64_t(uint32_t(v))); }
    void v3(const V3& q) { f(q[0]), f(q[1]), f(q[2]); }
    void v4(const V4& q) { f(q[0]), f(q[1]), f(q[2]), f(q[3]); }
};

std::string hx(double x) {
    uint64_t v;
    std::memcpy(&v, &x, 8);
    char b[24];
    std::snprintf(b, sizeof b, "%016" PRIx64, v);
    return b;
}
void line(const char* key, const Hash& h) { std::printf("%s %016" PRIx64 "\n", key, h.h); }

// splitmix64
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u01() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    double in(double a, double b) { return a + (b - a) * u01(); }
};

V3 state1(Rng& r) {
    EULER_EQUATION_1D e(1.4);
    V3 q;
    e.PrimitiveToConserved(r.in(0.1, 10), r.in(-3, 3), r.in(0.1, 100), q);
    return q;
}
V4 state2(Rng& r) {
    EULER_EQUATION_2D e(1.4);
    V4 q;
    e.PrimitiveToConserved(r.in(0.1, 10), r.in(-3, 3), r.in(-3, 3), r.in(0.1, 100), q);
    return q;
}

// A 7-state stencil: random (kinds 0, 1), smooth around one state (2) or
// uniform (3), the last exercising the degenerate weights.
template <class V, class Gen>
void stencil(Rng& r, int kind, V* s, Gen gen) {
    const V base = gen(r);
    for (int p = 0; p < 7; p++) {
        if (kind <= 1) s[p] = gen(r);
        else if (kind == 2) s[p] = base * (1 + 1e-3 * std::sin(0.7 * p + r.u01()));
        else s[p] = base;
    }
}

// ---------------------------------------------------------------- host mode
void probe_vectors() {
    Rng r(11);
    Hash h3, h4;
    for (int n = 0; n < 64; n++) {
        V3 a = state1(r), b = state1(r);
        const double c = r.in(-2, 2);
        h3.v3(a + b), h3.v3(a - b), h3.v3(a * b), h3.v3(a * c), h3.v3(a / c), h3.v3(c * a), h3.v3(a + c), h3.v3(a - c);
        h3.f(DotProduct(a, b)), h3.v3(CrossProduct(a, b));
        h3.f(a.SquareMagnitude());
        a += b, h3.v3(a);
        a -= b, h3.v3(a);
        a *= c, h3.v3(a);
        V4 x = state2(r), y = state2(r);
        h4.v4(x + y), h4.v4(x - y), h4.f(x * y), h4.v4(c * x), h4.v4(x * c), h4.v4(x / c), h4.f(DotProduct(x, y));
        h4.v4((x + 2.0 * y) / 3.0), h4.v4(0.75 * x + 0.25 * (y - c * (x + y)));
        x += y, h4.v4(x);
        x -= y, h4.v4(x);
    }
    line("vector3d", h3);
    line("vector4d", h4);
}

void probe_eos() {
    Rng r(12);
    IDEAL_GAS_1D g1(1.4);
    IDEAL_GAS_2D g2(1.4);
    EULER_EQUATION_1D e1(1.4);
    EULER_EQUATION_2D e2(1.4);
    Hash a, b, c, d;
    for (int n = 0; n < 64; n++) {
        const double rho = r.in(0.1, 10), u = r.in(-3, 3), v = r.in(-3, 3), p = r.in(0.1, 100);
        a.f(g1.E(rho, u, p)), a.f(g1.e(rho, u, p)), a.f(g1.H(rho, u, p)), a.f(g1.c(rho, p));
        V3 q1;
        e1.PrimitiveToConserved(rho, u, p, q1);
        a.v3(q1), a.f(g1.E(q1)), a.f(g1.e(q1)), a.f(g1.p(q1)), a.f(g1.H(q1)), a.f(g1.c(q1));
        b.f(g2.E(rho, u, v, p)), b.f(g2.e(rho, u, v, p)), b.f(g2.H(rho, u, v, p)), b.f(g2.c(rho, p));
        V4 q2;
        e2.PrimitiveToConserved(rho, u, v, p, q2);
        b.v4(q2), b.f(g2.E(q2)), b.f(g2.e(q2)), b.f(g2.p(q2)), b.f(g2.H(q2)), b.f(g2.c(q2));
        V3 f1;
        e1.ComputeFu(q1, f1);
        double rr, uu, vv, pp;
        e1.ConservedToPrimitive(q1, rr, uu, pp);
        c.v3(f1), c.f(rr), c.f(uu), c.f(pp);
        c.f(e1.Spectral_Radius(q1)), c.f(e1.MaxEigenvalue(q1)), c.f(e1.MinEigenvalue(q1));
        V4 f2, g;
        e2.ComputeFu(q2, f2), e2.ComputeGu(q2, g);
        e2.ConservedToPrimitive(q2, rr, uu, vv, pp);
        d.v4(f2), d.v4(g), d.f(rr), d.f(uu), d.f(vv), d.f(pp);
        d.f(e2.Spectral_RadiusX(q2)), d.f(e2.Spectral_RadiusY(q2)), d.f(e2.MaxEigenvalueX(q2));
        d.f(e2.MinEigenvalueX(q2)), d.f(e2.MaxEigenvalueY(q2)), d.f(e2.MinEigenvalueY(q2));
    }
    line("ideal_gas_1d", a);
    line("ideal_gas_2d", b);
    line("euler_equation_1d", c);
    line("euler_equation_2d", d);
}

void probe_roe() {
    Rng r(13);
    Hash a, b;
    for (int n = 0; n < 64; n++) {
        const V3 l = state1(r), rr = state1(r), bb = state1(r);
        double u, H, c;
        ROE_SOLVER_1D::RoeAverage(l, rr, u, H, c, 1.4);
        V3 x;
        ROE_SOLVER_1D::Inversion(u, H, c, bb, x);
        double x1, x2, x3;
        ROE_SOLVER_1D::Inversion(u, H, c, bb[0], bb[1], bb[2], x1, x2, x3);
        a.f(u), a.f(H), a.f(c), a.v3(x), a.f(x1), a.f(x2), a.f(x3);
        const V4 L = state2(r), R = state2(r), B = state2(r);
        double v;
        ROE_SOLVER_2D::RoeAverage(L, R, u, v, H, c, 1.4);
        V4 X, Y;
        ROE_SOLVER_2D::Inversion_X(u, v, H, c, B, X);
        ROE_SOLVER_2D::Inversion_Y(u, v, H, c, B, Y);
        double y1, y2, y3, y4;
        ROE_SOLVER_2D::Inversion_X(u, v, H, c, B[0], B[1], B[2], B[3], y1, y2, y3, y4);
        b.f(u), b.f(v), b.f(H), b.f(c), b.v4(X), b.v4(Y), b.f(y1), b.f(y2), b.f(y3), b.f(y4);
        ROE_SOLVER_2D::Inversion_Y(u, v, H, c, B[0], B[1], B[2], B[3], y1, y2, y3, y4);
        b.f(y1), b.f(y2), b.f(y3), b.f(y4);
    }
    line("roe_1d", a);
    line("roe_2d", b);
}

void probe_reconstruction() {
    Rng r(14);
    WENOZ5 z;
    WENO5 js;
    Hash a, b;
    for (int n = 0; n < 256; n++) {
        double s[5];
        const int kind = n % 4;
        const double base = r.in(-5, 5);
        for (int p = 0; p < 5; p++)
            s[p] = kind <= 1 ? r.in(-5, 5) : (kind == 2 ? base + 1e-4 * std::sin(p + r.u01()) : base);
        double L, R, L2, R2;
        z.Reconstruct(s[0], s[1], s[2], s[3], s[4], L, R);
        z.ReconstructLeft(s[0], s[1], s[2], s[3], s[4], L2);
        z.ReconstructRight(s[0], s[1], s[2], s[3], s[4], R2);
        a.f(L), a.f(R), a.f(L2), a.f(R2);
        js.Reconstruct(s[0], s[1], s[2], s[3], s[4], L, R);
        js.ReconstructLeft(s[0], s[1], s[2], s[3], s[4], L2);
        js.ReconstructRight(s[0], s[1], s[2], s[3], s[4], R2);
        b.f(L), b.f(R), b.f(L2), b.f(R2);
    }
    line("wenoz5_reconstruct", a);
    line("weno5_reconstruct", b);
}

void probe_fluxes() {
    Rng r(15);
    WENOZ_EULER_SOLVER_1D z1(1.4);
    WENO_EULER_SOLVER_1D j1(1.4);
    WENOZ_EULER_SOLVER_2D z2(1.4);
    WENO_EULER_SOLVER_2D j2(1.4);
    EULER_EQUATION_1D e1(1.4);
    EULER_EQUATION_2D e2(1.4);
    Hash a, b, c, d, e;
    std::string first;
    for (int n = 0; n < 256; n++) {
        V3 s[7];
        stencil(r, n % 4, s, state1);
        double al = 0;
        for (int p = 0; p < 7; p++) al = MAX(al, e1.Spectral_Radius(s[p]));
        al *= 1.1;
        V3 F, G, dF;
        z1.ComputeFlux(s[0], s[1], s[2], s[3], s[4], s[5], al, F);
        j1.ComputeFlux(s[0], s[1], s[2], s[3], s[4], s[5], al, G);
        j1.ComputedF(s[0], s[1], s[2], s[3], s[4], s[5], s[6], al, dF);
        a.v3(F), b.v3(G), b.v3(dF);
        if (n == 0) first = hx(F[0]) + " " + hx(G[2]);
        V4 t[7];
        stencil(r, n % 4, t, state2);
        double ax = 0, ay = 0;
        for (int p = 0; p < 7; p++) ax = MAX(ax, e2.Spectral_RadiusX(t[p])), ay = MAX(ay, e2.Spectral_RadiusY(t[p]));
        V4 X, Y, JX, JY;
        z2.ComputeFluxX(t[0], t[1], t[2], t[3], t[4], t[5], ax, X);
        z2.ComputeFluxY(t[0], t[1], t[2], t[3], t[4], t[5], ay, Y);
        j2.ComputeFluxX(t[0], t[1], t[2], t[3], t[4], t[5], ax, JX);
        j2.ComputeFluxY(t[0], t[1], t[2], t[3], t[4], t[5], ay, JY);
        c.v4(X), c.v4(Y), d.v4(JX), d.v4(JY);
        // alpha = 0 (pure upwinding on the characteristic split) as a corner case
        z2.ComputeFluxX(t[0], t[1], t[2], t[3], t[4], t[5], 0.0, X);
        e.v4(X);
    }
    line("wenoz_euler_1d.flux", a);
    line("weno_euler_1d.flux+dF", b);
    line("wenoz_euler_2d.flux_xy", c);
    line("weno_euler_2d.flux_xy", d);
    line("wenoz_euler_2d.alpha0", e);
    std::printf("flux.first %s\n", first.c_str());
}

void hash_mra1(const MRA_GRID_INT_1D& m, Hash& h) {
    h.i(m.num_levels), h.i(m.Nx_coarse), h.i(m.Nx);
    for (int l = 0; l < m.num_levels; l++) {
        const GRID_INT_1D& g = m.grid[l];
        h.i(m.dx[l]), h.i(g.i_start), h.i(g.i_end), h.i(g.i_res), h.i(g.dx), h.i(g.dx2);
    }
    for (int l = 0; l < m.num_levels - 1; l++) {
        const GRID_INT_1D& g = m.grid_x[l];
        h.i(g.i_start), h.i(g.i_end), h.i(g.i_res), h.i(g.dx), h.i(g.dx2);
    }
    h.i(m.grid_finest.i_res);
}
void hash_gi2(const GRID_INT_2D& g, Hash& h) {
    h.i(g.i_start), h.i(g.j_start), h.i(g.i_end), h.i(g.j_end), h.i(g.i_res), h.i(g.j_res);
    h.i(g.dx), h.i(g.dy), h.i(g.dx2), h.i(g.dy2);
}
void probe_mra() {
    const int c1[3][2] = {{11, 16}, {6, 16}, {3, 6}};
    for (const auto& c : c1) {
        MRA_GRID_INT_1D m;
        m.Initialize(c[0], c[1]);
        Hash h;
        hash_mra1(m, h);
        std::printf("mra_1d.%d.%d Nx=%d %016" PRIx64 "\n", c[0], c[1], m.Nx, h.h);
    }
    const int c2[3][3] = {{7, 65, 17}, {6, 15, 15}, {3, 5, 5}};
    for (const auto& c : c2) {
        MRA_GRID_INT_2D m;
        m.Initialize(c[0], c[1], c[2]);
        Hash h;
        h.i(m.Nx), h.i(m.Ny);
        for (int l = 0; l < m.num_levels; l++) h.i(m.mra[l]), hash_gi2(m.grid[l], h);
        for (int l = 0; l < m.num_levels - 1; l++)
            for (int mu = 1; mu <= 3; mu++) hash_gi2(m.subgrid(l, mu), h);
        hash_gi2(m.grid_finest, h);
        std::printf("mra_2d.%d.%d.%d Nx=%d Ny=%d %016" PRIx64 "\n", c[0], c[1], c[2], m.Nx, m.Ny, h.h);
    }
}

template <class F>
void hash_field1(const F& q, Hash& h) {
    for (int i = q.i_start_g; i <= q.i_end_g; i++) h.v3(q(i));
}
template <class F>
void hash_field2(const F& q, Hash& h) {
    for (int j = q.j_start_g; j <= q.j_end_g; j++)
        for (int i = q.i_start_g; i <= q.i_end_g; i++) h.v4(q(i, j));
}
void hash_grid1(const GRID_UNIFORM_1D& g, Hash& h) {
    h.i(g.i_start), h.i(g.i_res), h.f(g.x_min), h.f(g.x_max), h.f(g.dx), h.f(g.one_over_dx), h.f(g(g.i_end));
}
void hash_grid2(const GRID_UNIFORM_2D& g, Hash& h) {
    h.i(g.i_res), h.i(g.j_res), h.f(g.x_min), h.f(g.y_min), h.f(g.x_max), h.f(g.y_max), h.f(g.dx), h.f(g.dy);
    h.f(g.one_over_dx), h.f(g.one_over_dy), h.f(g(3, 2).x), h.f(g(3, 2).y);
}

void probe_problems() {
    {
        GRID_UNIFORM_1D g, gx;
        FIELD_UNIFORM_1D<V3> Q;
        double tf, gamma;
        SHOCKTUBE::Initialize(241, g, gx, Q, tf, gamma);
        Hash h;
        hash_grid1(g, h), hash_grid1(gx, h), hash_field1(Q, h), h.f(tf), h.f(gamma);
        line("shocktube.241", h);
        ENTROPY_INTERACTING::Initialize(241, g, gx, Q, tf, gamma);
        Hash k;
        hash_grid1(g, k), hash_grid1(gx, k), hash_field1(Q, k), k.f(tf), k.f(gamma);
        line("entropy_interacting.241", k);
    }
    {
        GRID_UNIFORM_2D g, gx, gy, gxy;
        FIELD_UNIFORM_2D<V4> Q;
        double tf, gamma;
        QUADRANTS::Initialize(65, g, gx, gy, gxy, Q, tf, gamma);
        Hash h;
        hash_grid2(g, h), hash_grid2(gx, h), hash_grid2(gy, h), hash_grid2(gxy, h), hash_field2(Q, h), h.f(tf), h.f(gamma);
        line("quadrants.65", h);
        DOUBLEMACH::Initialize(33, g, gx, gy, gxy, Q, tf, gamma);
        Hash k;
        hash_grid2(g, k), hash_grid2(gx, k), hash_grid2(gy, k), hash_grid2(gxy, k), hash_field2(Q, k), k.f(tf), k.f(gamma);
        line("doublemach.33", k);
    }
}

void probe_boundaries() {
    Rng r(16);
    {
        FIELD_UNIFORM_1D<V3> Q(0, 40, 0.0, 1.0, 3);
        for (int i = Q.i_start_g; i <= Q.i_end_g; i++) Q(i) = state1(r);
        BOUNDARY_MANAGER bm(Q);
        bm.SetBoundaryCondition();
        Hash h;
        hash_field1(Q, h);
        for (int i = Q.i_start_g; i <= Q.i_end_g; i++) Q(i) = state1(r);
        BOUNDARY_MANAGER_ENTROPY_INTERATING be(Q);
        be.SetBoundaryCondition();
        hash_field1(Q, h);
        line("boundary_1d", h);
    }
    {
        FIELD_UNIFORM_2D<V4> Q(0, 0, 17, 13, 0.0, 0.0, 1.0, 0.75, 3);
        for (int j = Q.j_start_g; j <= Q.j_end_g; j++)
            for (int i = Q.i_start_g; i <= Q.i_end_g; i++) Q(i, j) = state2(r);
        BOUNDARY_MANAGER_QUAD bq(Q);
        bq.SetBoundaryCondition();
        Hash h;
        hash_field2(Q, h);
        bq.SetDiagonalBoundaryCondition();
        hash_field2(Q, h);
        line("boundary_quad", h);
    }
    {
        GRID_UNIFORM_2D g, gx, gy, gxy;
        FIELD_UNIFORM_2D<V4> Q;
        double tf, gamma;
        DOUBLEMACH::Initialize(9, g, gx, gy, gxy, Q, tf, gamma);
        for (int j = Q.j_start_g; j <= Q.j_end_g; j++)
            for (int i = Q.i_start_g; i <= Q.i_end_g; i++) Q(i, j) = state2(r);
        BOUNDARY_MANAGER_DMACH bd(Q);
        Hash h;
        bd.SetBoundaryCondition(0.05);
        hash_field2(Q, h);
        bd.SetBoundaryCondition(0.17);
        hash_field2(Q, h);
        bd.SetDiagonalBoundaryCondition();
        hash_field2(Q, h);
        line("boundary_dmach", h);
    }
}

void probe_awcm_host() {
    Rng r(17);
    MRA_GRID_INT_1D m1;
    m1.Initialize(4, 8);
    AWCM_1D a1(m1);
    Hash h1;
    for (int n = 0; n < 64; n++) {
        double s[5];
        for (int p = 0; p < 5; p++) s[p] = r.in(-3, 3);
        double v;
        a1.Differentiate(s[0], s[1], s[2], s[3], s[4], v), h1.f(v);
        a1.DifferentiateAtLeftBoundary(s[1], s[2], s[3], s[4], v), h1.f(v);
        a1.DifferentiateAtRightBoundary(s[0], s[1], s[2], s[3], v), h1.f(v);
        V3 q[5], d;
        for (int p = 0; p < 5; p++) q[p] = state1(r);
        a1.DifferentiateEuler(q[0], q[1], q[2], q[3], q[4], d), h1.v3(d);
        a1.DifferentiateEulerAtLeftBoundary(q[1], q[2], q[3], q[4], d), h1.v3(d);
        a1.DifferentiateEulerAtRightBoundary(q[0], q[1], q[2], q[3], d), h1.v3(d);
        a1.DifferentiateEulerAtLeftBoundary(q[0], q[1], q[2], q[3], q[4], d), h1.v3(d);
        a1.DifferentiateEulerAtRightBoundary(q[0], q[1], q[2], q[3], q[4], d), h1.v3(d);
    }
    line("awcm_1d.differentiate", h1);
    MRA_GRID_INT_2D m2;
    m2.Initialize(3, 5, 5);
    AWCM_2D a2(m2);
    Hash h2;
    for (int n = 0; n < 64; n++) {
        V4 q[5], d;
        for (int p = 0; p < 5; p++) q[p] = state2(r);
        a2.Differentiate(q[0], q[1], q[2], q[3], q[4], d), h2.v4(d);
        a2.DifferentiateAtLeftBoundary(q[1], q[2], q[3], q[4], d), h2.v4(d);
        a2.DifferentiateAtRightBoundary(q[0], q[1], q[2], q[3], d), h2.v4(d);
    }
    line("awcm_2d.differentiate", h2);

    // constructor state: masks and levels as the constructors leave them
    Hash h3;
    for (int i = a1.mask0.i_start_g; i <= a1.mask0.i_end_g; i++) h3.byte(a1.mask0(i)), h3.byte(a1.mask1(i));
    for (int j = a2.mask0.j_start_g; j <= a2.mask0.j_end_g; j++)
        for (int i = a2.mask0.i_start_g; i <= a2.mask0.i_end_g; i++) h3.byte(a2.mask0(i, j)), h3.byte(a2.mask3(i, j));
    for (int j = 0; j < m2.Ny; j++)
        for (int i = 0; i < m2.Nx; i++) h3.i(a2.level0(i, j)), h3.i(a2.level2(i, j));
    h3.f(a1.d00), h3.f(a1.d13), h3.f(a1.d24), h3.f(a1.d35), h3.f(a2.d1), h3.f(a2.d2), h3.f(a2.d30);
    line("awcm.constructed", h3);

    // PostProcessEuler on random masks
    FIELD_UNIFORM_1D<V3> Q1(0, m1.Nx, 0.0, 1.0, 3);
    for (int i = Q1.i_start_g; i <= Q1.i_end_g; i++) Q1(i) = state1(r), a1.mask0(i) = (r.next() & 3) != 0;
    a1.PostProcessEuler(Q1);
    Hash h4;
    hash_field1(Q1, h4);
    FIELD_UNIFORM_2D<V4> Q2(0, 0, m2.Nx, m2.Ny, 0.0, 0.0, 1.0, 1.0, 3);
    for (int j = Q2.j_start_g; j <= Q2.j_end_g; j++)
        for (int i = Q2.i_start_g; i <= Q2.i_end_g; i++) Q2(i, j) = state2(r), a2.mask0(i, j) = (r.next() & 3) != 0;
    a2.PostProcessEuler(Q2);
    hash_field2(Q2, h4);
    line("awcm.post_process_euler", h4);
}

void probe_omp() {
    OMP_MANAGER om(1);
    om.Input(2.5);
    om.Input(7);
    Hash h;
    h.f(om.Max_double()), h.f(om.Min_double()), h.f(om.Sum_double()), h.i(om.Max_int()), h.i(om.Min_int()), h.i(om.Sum_int());
    line("omp_manager", h);
}

void file_line(const char* name) {
    std::ifstream f(name, std::ios::binary);
    if (!f) {
        std::printf("file %s absent\n", name);
        return;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string s = ss.str();
    Hash h;
    for (unsigned char ch : s) h.byte(ch);
    std::printf("file %s %zu %016" PRIx64 "\n", name, s.size(), h.h);
}

void probe_print(const char* dir) {
    if (chdir(dir) != 0) {
        std::printf("cannot enter %s\n", dir);
        std::exit(2);
    }
    Rng r(18);
    static char n1[] = "nm", n2[] = "Q";
    static char f_p1d[] = "p1d", f_p1dg[] = "p1dg", f_p1v[] = "p1v", f_p2v[] = "p2v", f_p2vg[] = "p2vg";
    static char f_p2i[] = "p2i", f_m2[] = "m2", f_m1d[] = "m1d", f_m1v[] = "m1v", f_gx[] = "gx", f_g2x[] = "g2x";
    static char f_g2y[] = "g2y", f_arr[] = "arr";
    FIELD_UNIFORM_1D<double> f(0, 12, -1.0, 2.0, 3);
    FIELD_UNIFORM_1D<bool> m(0, 12, -1.0, 2.0, 3);
    for (int i = f.i_start_g; i <= f.i_end_g; i++) f(i) = r.in(-1e3, 1e3), m(i) = (r.next() & 1) != 0;
    f.Print(f_p1d);
    f.Print(f_p1dg, true);
    f.TagName(n1);
    f.Print();
    f.Print(true);
    FIELD_UNIFORM_1D<V3> q(0, 12, -1.0, 2.0, 3);
    for (int i = q.i_start_g; i <= q.i_end_g; i++) q(i) = state1(r);
    q.Print(f_p1v);
    AWCM_MANAGER_1D::Print(f, m, f_m1d);
    AWCM_MANAGER_1D::Print(q, m, f_m1v);
    FIELD_UNIFORM_2D<V4> Q(0, 0, 7, 5, 0.0, 0.0, 3.0, 1.0, 3);
    FIELD_UNIFORM_2D<bool> M(0, 0, 7, 5, 0.0, 0.0, 3.0, 1.0, 3);
    FIELD_UNIFORM_2D<int> I(0, 0, 7, 5, 0.0, 0.0, 3.0, 1.0);
    for (int j = Q.j_start_g; j <= Q.j_end_g; j++)
        for (int i = Q.i_start_g; i <= Q.i_end_g; i++) Q(i, j) = state2(r), M(i, j) = (r.next() & 1) != 0;
    for (int j = 0; j < 5; j++)
        for (int i = 0; i < 7; i++) I(i, j) = int(r.next() % 9);
    Q.Print(f_p2v);
    Q.Print(f_p2vg, true);
    Q.TagName(n2);
    Q.Print();
    I.Print(f_p2i);
    AWCM_MANAGER_2D::Print(Q, M, f_m2);
    f.base_grid.PrintGrid(f_gx);
    Q.grid.PrintGridX(f_g2x);
    Q.grid.PrintGridY(f_g2y);
    ARRAY<double> a(5);
    for (int k = 0; k < 5; k++) a[k] = r.in(0, 1);
    a.Print(f_arr);
    const char* names[] = {"p1d", "p1dg", "nm0000", "nm0001", "p1v", "m1d", "m1v", "p2v", "p2vg", "Q_0000",
                           "p2i", "m2",   "gx",     "g2x",    "g2y", "arr"};
    for (const char* nm : names) file_line(nm);
}

// ---------------------------------------------------------------- device modes
void hash_awcm1(const AWCM_1D& a, const FIELD_UNIFORM_1D<V3>& Q, const char* tag) {
    Hash hq, hm, hl;
    hash_field1(Q, hq);
    const FIELD_UNIFORM_1D<bool>* ms[5] = {&a.mask_temp, &a.mask0, &a.mask1, &a.mask2, &a.mask3};
    const FIELD_UNIFORM_1D<int>* ls[4] = {&a.level0, &a.level1, &a.level2, &a.level3};
    int n0 = 0;
    for (const auto* m : ms)
        for (int i = m->i_start_g; i <= m->i_end_g; i++) hm.byte((*m)(i));
    for (int i = 0; i < a.mra.Nx; i++) n0 += a.mask0(i);
    for (const auto* l : ls)
        for (int i = 0; i < a.mra.Nx; i++) hl.i((*l)(i));
    std::printf("%s active=%d Q=%016" PRIx64 " masks=%016" PRIx64 " levels=%016" PRIx64 "\n", tag, n0, hq.h, hm.h, hl.h);
}
void hash_awcm2(const AWCM_2D& a, const FIELD_UNIFORM_2D<V4>& Q, const char* tag) {
    Hash hq, hm, hl;
    hash_field2(Q, hq);
    const FIELD_UNIFORM_2D<bool>* ms[5] = {&a.mask_temp, &a.mask0, &a.mask1, &a.mask2, &a.mask3};
    const FIELD_UNIFORM_2D<int>* ls[4] = {&a.level0, &a.level1, &a.level2, &a.level3};
    int n0 = 0;
    for (const auto* m : ms)
        for (int j = m->j_start_g; j <= m->j_end_g; j++)
            for (int i = m->i_start_g; i <= m->i_end_g; i++) hm.byte((*m)(i, j));
    for (int j = 0; j < a.mra_grid.Ny; j++)
        for (int i = 0; i < a.mra_grid.Nx; i++) n0 += a.mask0(i, j);
    for (const auto* l : ls)
        for (int j = 0; j < a.mra_grid.Ny; j++)
            for (int i = 0; i < a.mra_grid.Nx; i++) hl.i((*l)(i, j));
    std::printf("%s active=%d Q=%016" PRIx64 " masks=%016" PRIx64 " levels=%016" PRIx64 "\n", tag, n0, hq.h, hm.h, hl.h);
}

struct Cfg1 {
    const char* name;
    int L, nxc;
    double tol;
    bool shu_osher;
};
struct Cfg2 {
    const char* name;
    int L, nxc, nyc;
    double tol;
    bool dmr;
};
// the last is main_1d_euler.cpp's own configuration (:14-17): 15361 points
const Cfg1 kCfg1[] = {{"sod", 6, 16, 1e-4, false}, {"shu_osher", 6, 16, 1e-4, true}, {"sod_driver", 11, 16, 1e-6, false}};
const Cfg2 kCfg2[] = {{"quad", 4, 9, 9, 1e-3, false}, {"dmr", 4, 17, 5, 5e-3, true}};

void init1(const Cfg1& c, MRA_GRID_INT_1D& mra, GRID_UNIFORM_1D& g, GRID_UNIFORM_1D& gx, FIELD_UNIFORM_1D<V3>& Q,
           double& tf, double& gamma) {
    mra.Initialize(c.L, c.nxc);
    if (c.shu_osher) ENTROPY_INTERACTING::Initialize(mra.grid_finest.i_res, g, gx, Q, tf, gamma);
    else SHOCKTUBE::Initialize(mra.grid_finest.i_res, g, gx, Q, tf, gamma);
}
void init2(const Cfg2& c, MRA_GRID_INT_2D& mra, GRID_UNIFORM_2D& g, FIELD_UNIFORM_2D<V4>& Q, double& tf,
           double& gamma) {
    mra.Initialize(c.L, c.nxc, c.nyc);
    GRID_UNIFORM_2D gx, gy, gxy;
    if (c.dmr) DOUBLEMACH::Initialize(mra.grid_finest.j_res, g, gx, gy, gxy, Q, tf, gamma);
    else QUADRANTS::Initialize(mra.grid_finest.j_res, g, gx, gy, gxy, Q, tf, gamma);
}

void run_prepare() {
    for (const Cfg1& c : kCfg1) {
        MRA_GRID_INT_1D mra;
        GRID_UNIFORM_1D g, gx;
        FIELD_UNIFORM_1D<V3> Q;
        double tf, gamma;
        init1(c, mra, g, gx, Q, tf, gamma);
        AWCM_1D awcm(mra);
        char tag[64];
        for (int k = 0; k < 2; k++) {
            awcm.PrepareEuler(Q, c.tol);
            std::snprintf(tag, sizeof tag, "prepare.%s.%d", c.name, k);
            hash_awcm1(awcm, Q, tag);
        }
        std::fflush(stdout);
    }
    for (const Cfg2& c : kCfg2) {
        MRA_GRID_INT_2D mra;
        GRID_UNIFORM_2D g;
        FIELD_UNIFORM_2D<V4> Q;
        double tf, gamma;
        init2(c, mra, g, Q, tf, gamma);
        AWCM_2D awcm(mra);
        char tag[64];
        for (int k = 0; k < 2; k++) {
            awcm.PrepareEuler(Q, c.tol);
            std::snprintf(tag, sizeof tag, "prepare.%s.%d", c.name, k);
            hash_awcm2(awcm, Q, tag);
        }
        // the whole grid from mask1's points, then from mask0's
        // (WAVELET_2D::InterpolateAllGridEuler, as main_2d_euler.cpp:334)
        awcm.wavelet.InterpolateAllGridEuler(Q, mra, awcm.mask1);
        std::snprintf(tag, sizeof tag, "prepare.%s.all_grid1", c.name);
        hash_awcm2(awcm, Q, tag);
        awcm.wavelet.InterpolateAllGridEuler(Q, mra, awcm.mask0);
        std::snprintf(tag, sizeof tag, "prepare.%s.all_grid0", c.name);
        hash_awcm2(awcm, Q, tag);
        std::fflush(stdout);
    }
}

// AWCM_MANAGER_1D / AWCM_MANAGER_2D static members on copies of an AWCM
// object's masks after two PrepareEuler calls: AddAdjacentZone and
// CheckReconstruction of mask0, AddPointForDifferentiation of mask1 with
// level1, ComputeLevel of mask2 into a copy of level0, RemoveResidue of the
// density (scalar) and of the whole field outside mask1.
template <class F>
void hash_mask1(const F& m, Hash& h) {
    for (int i = m.i_start_g; i <= m.i_end_g; i++) h.byte(m(i));
}
template <class F>
void hash_mask2(const F& m, Hash& h) {
    for (int j = m.j_start_g; j <= m.j_end_g; j++)
        for (int i = m.i_start_g; i <= m.i_end_g; i++) h.byte(m(i, j));
}
void run_manager() {
    for (const Cfg1& c : kCfg1) {
        if (c.L > 8) continue;
        MRA_GRID_INT_1D mra;
        GRID_UNIFORM_1D g, gx;
        FIELD_UNIFORM_1D<V3> Q;
        double tf, gamma;
        init1(c, mra, g, gx, Q, tf, gamma);
        AWCM_1D awcm(mra);
        awcm.PrepareEuler(Q, c.tol);
        awcm.PrepareEuler(Q, c.tol);
        const int Nx = mra.Nx;
        auto copy_mask = [&](const FIELD_UNIFORM_1D<bool>& src) {
            FIELD_UNIFORM_1D<bool> m(0, Nx, 0, Nx, 3);
            for (int i = -3; i <= Nx + 2; i++) m(i) = src(i);
            return m;
        };
        auto copy_level = [&](const FIELD_UNIFORM_1D<int>& src) {
            FIELD_UNIFORM_1D<int> l(0, Nx, 0, Nx);
            for (int i = 0; i < Nx; i++) l(i) = src(i);
            return l;
        };
        Hash h;
        FIELD_UNIFORM_1D<bool> a = copy_mask(awcm.mask0);
        AWCM_MANAGER_1D::AddAdjacentZone(a, mra);
        hash_mask1(a, h);
        FIELD_UNIFORM_1D<bool> b = copy_mask(awcm.mask0);
        AWCM_MANAGER_1D::CheckReconstruction(b, mra);
        hash_mask1(b, h);
        FIELD_UNIFORM_1D<bool> d = copy_mask(awcm.mask1);
        AWCM_MANAGER_1D::AddPointForDifferentiation(d, mra, awcm.level1);
        hash_mask1(d, h);
        FIELD_UNIFORM_1D<int> l = copy_level(awcm.level0);
        AWCM_MANAGER_1D::ComputeLevel(l, awcm.mask2, mra);
        for (int i = 0; i < Nx; i++) h.i(l(i));
        FIELD_UNIFORM_1D<double> rho(0, Nx, 0.0, 1.0, 3);
        for (int i = -3; i <= Nx + 2; i++) rho(i) = Q(i)[0];
        AWCM_MANAGER_1D::RemoveResidue(rho, awcm.mask1, mra);
        FIELD_UNIFORM_1D<V3> Qc(0, Nx, 0.0, 1.0, 3);
        for (int i = -3; i <= Nx + 2; i++) Qc(i) = Q(i);
        AWCM_MANAGER_1D::RemoveResidue(Qc, awcm.mask1, mra);
        for (int i = -3; i <= Nx + 2; i++) h.f(rho(i)), h.v3(Qc(i));
        char tag[64];
        std::snprintf(tag, sizeof tag, "manager.%s", c.name);
        line(tag, h);
        std::fflush(stdout);
    }
    for (const Cfg2& c : kCfg2) {
        MRA_GRID_INT_2D mra;
        GRID_UNIFORM_2D g;
        FIELD_UNIFORM_2D<V4> Q;
        double tf, gamma;
        init2(c, mra, g, Q, tf, gamma);
        AWCM_2D awcm(mra);
        awcm.PrepareEuler(Q, c.tol);
        awcm.PrepareEuler(Q, c.tol);
        const int Nx = mra.Nx, Ny = mra.Ny;
        auto copy_mask = [&](const FIELD_UNIFORM_2D<bool>& src) {
            FIELD_UNIFORM_2D<bool> m(0, 0, Nx, Ny, 0, 0, Nx, Ny, 3);
            for (int j = -3; j <= Ny + 2; j++)
                for (int i = -3; i <= Nx + 2; i++) m(i, j) = src(i, j);
            return m;
        };
        Hash h;
        FIELD_UNIFORM_2D<bool> a = copy_mask(awcm.mask0);
        AWCM_MANAGER_2D::AddAdjacentZone(a, mra);
        hash_mask2(a, h);
        FIELD_UNIFORM_2D<bool> b = copy_mask(awcm.mask0);
        AWCM_MANAGER_2D::CheckReconstruction(b, mra);
        hash_mask2(b, h);
        FIELD_UNIFORM_2D<bool> d = copy_mask(awcm.mask1);
        AWCM_MANAGER_2D::AddPointForDifferentiation(d, mra, awcm.level1);
        hash_mask2(d, h);
        FIELD_UNIFORM_2D<int> l(0, 0, Nx, Ny, 0, 0, Nx, Ny);
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) l(i, j) = awcm.level0(i, j);
        AWCM_MANAGER_2D::ComputeLevel(l, awcm.mask2, mra);
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) h.i(l(i, j));
        FIELD_UNIFORM_2D<double> rho(0, 0, Nx, Ny, 0, 0, Nx, Ny, 3);
        FIELD_UNIFORM_2D<V4> Qc(0, 0, Nx, Ny, 0, 0, Nx, Ny, 3);
        for (int j = -3; j <= Ny + 2; j++)
            for (int i = -3; i <= Nx + 2; i++) rho(i, j) = Q(i, j)[0], Qc(i, j) = Q(i, j);
        AWCM_MANAGER_2D::RemoveResidue(rho, awcm.mask1, mra);
        AWCM_MANAGER_2D::RemoveResidueEuler(Qc, awcm.mask1, mra);
        for (int j = -3; j <= Ny + 2; j++)
            for (int i = -3; i <= Nx + 2; i++) h.f(rho(i, j)), h.v4(Qc(i, j));
        char tag[64];
        std::snprintf(tag, sizeof tag, "manager.%s", c.name);
        line(tag, h);
        std::fflush(stdout);
    }
}

// The 1-D time loop of Wavelet/main_1d_euler.cpp through the class API.  With
// use_js the finest-level difference is WENO_EULER_SOLVER_1D::ComputedF, as
// the shipped driver has it; otherwise the WENO-Z face-flux pair.
void loop1(const Cfg1& c, int steps, bool use_js) {
    MRA_GRID_INT_1D mra;
    GRID_UNIFORM_1D g, gx;
    FIELD_UNIFORM_1D<V3> Q, Q_temp, Q_copy;
    double tf, gamma;
    init1(c, mra, g, gx, Q, tf, gamma);
    AWCM_1D awcm(mra);
    const int max_level = mra.num_levels - 1, num_print = 41;
    const double cfl = 0.5, dx = g.dx;
    const IDEAL_GAS_1D EOS(gamma);
    WENO_EULER_SOLVER_1D weno(gamma);
    WENOZ_EULER_SOLVER_1D wenoz(gamma);
    BOUNDARY_MANAGER bc(Q);
    ARRAY<double> pt(num_print);
    for (int p = 0; p < num_print; p++) pt[p] = tf / (num_print - 1) * p;
    double t = 0;
    int cnt = 1;
    auto wave = [&](const V3& q) { return ABS(q[1] / q[0]) + EOS.c(q); };
    auto finest_alpha = [&](const FIELD_UNIFORM_1D<bool>& m, const FIELD_UNIFORM_1D<int>& lv) {
        double a = 0;
        for (int i = g.i_start; i <= g.i_end; i++)
            if (m(i) && lv(i) == max_level) a = MAX(a, wave(Q(i)));
        return a;
    };
    // dF at point i of level spacing d from the frozen copy Q_temp
    auto dflux = [&](int i, int lev, double alpha, V3& dF) {
        const int d = mra.dx[lev];
        const FIELD_UNIFORM_1D<V3>& T = Q_temp;
        if (lev == max_level) {
            if (use_js) {
                weno.ComputedF(T(i - 3), T(i - 2), T(i - 1), T(i), T(i + 1), T(i + 2), T(i + 3), alpha, dF);
            } else {
                V3 Fl, Fr;
                wenoz.ComputeFlux(T(i - 3), T(i - 2), T(i - 1), T(i), T(i + 1), T(i + 2), alpha, Fl);
                wenoz.ComputeFlux(T(i - 2), T(i - 1), T(i), T(i + 1), T(i + 2), T(i + 3), alpha, Fr);
                dF = Fr - Fl;
            }
        } else if (i - 2 * d > 0 && i + 2 * d < g.i_end) {
            awcm.DifferentiateEuler(T(i - 2 * d), T(i - d), T(i), T(i + d), T(i + 2 * d), dF);
        } else if (i - 2 * d < 0) {
            awcm.DifferentiateEulerAtLeftBoundary(T(i - d), T(i), T(i + d), T(i + 2 * d), dF);
        } else if (i == 2 * d) {
            awcm.DifferentiateEulerAtLeftBoundary(T(i - 2 * d), T(i - d), T(i), T(i + d), T(i + 2 * d), dF);
        } else if (i > g.i_end - 2 * d) {
            awcm.DifferentiateEulerAtRightBoundary(T(i - 2 * d), T(i - d), T(i), T(i + d), dF);
        } else {
            awcm.DifferentiateEulerAtRightBoundary(T(i - 2 * d), T(i - d), T(i), T(i + d), T(i + 2 * d), dF);
        }
        return d;
    };
    for (int s = 0; s < steps; s++) {
        awcm.PrepareEuler(Q, c.tol);
        double vmax = 0;
        for (int i = g.i_start; i <= g.i_end; i++)
            if (awcm.mask3(i)) vmax = MAX(vmax, wave(Q(i)));
        double dt = cfl * dx / vmax;
        bool do_print = false;
        if (t + dt > pt[cnt]) dt = pt[cnt] - t, do_print = true;
        bc.SetBoundaryCondition();
        double alpha = finest_alpha(awcm.mask3, awcm.level3);
        Q_copy = Q;
        Q_temp = Q;
        V3 dF;
        for (int i = g.i_start; i <= g.i_end; i++) {
            if (!awcm.mask2(i)) continue;
            const int d = dflux(i, awcm.level2(i), alpha, dF);
            Q(i) = Q(i) - dt / (dx * d) * dF;
        }
        bc.SetBoundaryCondition();
        alpha = finest_alpha(awcm.mask2, awcm.level2);
        Q_temp = Q;
        for (int i = g.i_start; i <= g.i_end; i++) {
            if (!awcm.mask1(i)) continue;
            const int d = dflux(i, awcm.level1(i), alpha, dF);
            Q(i) = 0.75 * Q_copy(i) + 0.25 * (Q(i) - dt / (dx * d) * dF);
        }
        bc.SetBoundaryCondition();
        alpha = finest_alpha(awcm.mask1, awcm.level1);
        Q_temp = Q;
        for (int i = g.i_start; i <= g.i_end; i++) {
            if (!awcm.mask0(i)) continue;
            const int d = dflux(i, awcm.level0(i), alpha, dF);
            Q(i) = (Q_copy(i) + 2.0 * (Q(i) - dt / (dx * d) * dF)) / 3.0;
        }
        awcm.PostProcessEuler(Q);
        t += dt;
        if (do_print) cnt++;
        Hash h;
        hash_field1(Q, h);
        std::printf("loop.%s%s.%d t=%s dt=%s Q=%016" PRIx64 "\n", c.name, use_js ? ".js" : "", s, hx(t).c_str(),
                    hx(dt).c_str(), h.h);
    }
    char tag[64];
    std::snprintf(tag, sizeof tag, "loop.%s%s.final", c.name, use_js ? ".js" : "");
    hash_awcm1(awcm, Q, tag);
    std::fflush(stdout);
}

// The 2-D time loop of Wavelet/main_2d_euler(dmach|quad).cpp through the class
// API: WENO-Z face fluxes at the finest level, the AWCM central differences
// (one-sided within d of the walls) below it.
template <class BC>
void loop2(const Cfg2& c, int steps) {
    MRA_GRID_INT_2D mra;
    GRID_UNIFORM_2D g;
    FIELD_UNIFORM_2D<V4> Q;
    double tf, gamma;
    init2(c, mra, g, Q, tf, gamma);
    const int max_lev = mra.num_levels - 1, num_print = c.dmr ? 21 : 9;
    const double cfl = 0.5, dx = g.dx;
    EULER_EQUATION_2D eq(gamma);
    FIELD_UNIFORM_2D<V4> Q_0(g, 3), Q_1(g, 3), Q_2(g, 3);
    ARRAY_1D<double> alpha_x(0, g.j_res), alpha_y(0, g.i_res);
    AWCM_2D awcm(mra);
    BC bc(Q), bc0(Q_0), bc1(Q_1), bc2(Q_2);
    WENOZ_EULER_SOLVER_2D wz;
    ARRAY<double> pt(num_print);
    for (int p = 0; p < num_print; p++) pt[p] = tf / (num_print - 1) * p;
    double t = 0;
    int cnt = 1;
    auto alphas = [&](const FIELD_UNIFORM_2D<V4>& F, const FIELD_UNIFORM_2D<bool>& m) {
        for (int j = 0; j < g.j_res; j++) alpha_x(j) = 0;
        for (int i = 0; i < g.i_res; i++) alpha_y(i) = 0;
        for (int j = g.j_start; j <= g.j_end; j++)
            for (int i = g.i_start; i <= g.i_end; i++)
                if (m(i, j)) {
                    alpha_x(j) = MAX(alpha_x(j), eq.Spectral_RadiusX(F(i, j)));
                    alpha_y(i) = MAX(alpha_y(i), eq.Spectral_RadiusY(F(i, j)));
                }
    };
    // dF + dG at (i, j) of level lev from F; returns the level spacing
    auto rhs = [&](const FIELD_UNIFORM_2D<V4>& F, int i, int j, int lev, V4& out) {
        const int d = mra.grid[lev].dx;
        V4 dF, dG;
        if (lev == max_lev) {
            V4 a, b;
            wz.ComputeFluxX(F(i - 3, j), F(i - 2, j), F(i - 1, j), F(i, j), F(i + 1, j), F(i + 2, j), alpha_x(j), a);
            wz.ComputeFluxX(F(i - 2, j), F(i - 1, j), F(i, j), F(i + 1, j), F(i + 2, j), F(i + 3, j), alpha_x(j), b);
            dF = b - a;
            wz.ComputeFluxY(F(i, j - 3), F(i, j - 2), F(i, j - 1), F(i, j), F(i, j + 1), F(i, j + 2), alpha_y(i), a);
            wz.ComputeFluxY(F(i, j - 2), F(i, j - 1), F(i, j), F(i, j + 1), F(i, j + 2), F(i, j + 3), alpha_y(i), b);
            dG = b - a;
        } else {
            // only the fluxes the chosen branch reads: the one-sided forms
            // skip the far point that would lie beyond the ghost cells
            V4 f[5];
            const bool li = i <= d, ri = !li && i >= g.i_end - d;
            for (int k = -2; k <= 2; k++)
                if (!(k == -2 && li) && !(k == 2 && ri)) eq.ComputeFu(F(i + k * d, j), f[k + 2]);
            if (li) awcm.DifferentiateAtLeftBoundary(f[1], f[2], f[3], f[4], dF);
            else if (ri) awcm.DifferentiateAtRightBoundary(f[0], f[1], f[2], f[3], dF);
            else awcm.Differentiate(f[0], f[1], f[2], f[3], f[4], dF);
            const bool lj = j <= d, rj = !lj && j >= g.j_end - d;
            for (int k = -2; k <= 2; k++)
                if (!(k == -2 && lj) && !(k == 2 && rj)) eq.ComputeGu(F(i, j + k * d), f[k + 2]);
            if (lj) awcm.DifferentiateAtLeftBoundary(f[1], f[2], f[3], f[4], dG);
            else if (rj) awcm.DifferentiateAtRightBoundary(f[0], f[1], f[2], f[3], dG);
            else awcm.Differentiate(f[0], f[1], f[2], f[3], f[4], dG);
        }
        out = dF + dG;
        return d;
    };
    auto set_bc = [&](BC& b, double tt) {
        if constexpr (std::is_same<BC, BOUNDARY_MANAGER_DMACH>::value) b.SetBoundaryCondition(tt);
        else b.SetBoundaryCondition();
    };
    for (int s = 0; s < steps; s++) {
        awcm.PrepareEuler(Q, c.tol);
        double vmax = 0;
        for (int j = g.j_start; j <= g.j_end; j++)
            for (int i = g.i_start; i <= g.i_end; i++)
                if (awcm.mask3(i, j)) vmax = MAX(vmax, eq.Spectral_RadiusX(Q(i, j)) + eq.Spectral_RadiusY(Q(i, j)));
        double dt = cfl * dx / vmax;
        bool do_print = false;
        if (t + dt > pt[cnt]) dt = pt[cnt] - t, do_print = true;
        V4 r;
        set_bc(bc, t);
        alphas(Q, awcm.mask3);
        for (int j = g.j_start; j <= g.j_end; j++)
            for (int i = g.i_start; i <= g.i_end; i++) {
                if (!awcm.mask2(i, j)) continue;
                const int d = rhs(Q, i, j, awcm.level2(i, j), r);
                Q_2(i, j) = Q(i, j) - dt / (dx * d) * r;
            }
        set_bc(bc2, t);
        alphas(Q_2, awcm.mask2);
        for (int j = g.j_start; j <= g.j_end; j++)
            for (int i = g.i_start; i <= g.i_end; i++) {
                if (!awcm.mask1(i, j)) continue;
                const int d = rhs(Q_2, i, j, awcm.level1(i, j), r);
                Q_1(i, j) = 0.75 * Q(i, j) + 0.25 * (Q_2(i, j) - dt / (dx * d) * r);
            }
        set_bc(bc1, t);
        alphas(Q_1, awcm.mask1);
        for (int j = g.j_start; j <= g.j_end; j++)
            for (int i = g.i_start; i <= g.i_end; i++) {
                if (!awcm.mask0(i, j)) continue;
                const int d = rhs(Q_1, i, j, awcm.level0(i, j), r);
                Q_0(i, j) = Q_1(i, j) - dt / (dx * d) * r;
            }
        for (int j = g.j_start; j <= g.j_end; j++)
            for (int i = g.i_start; i <= g.i_end; i++)
                if (awcm.mask0(i, j)) Q(i, j) = (Q(i, j) + 2.0 * Q_0(i, j)) / 3.0;
        awcm.PostProcessEuler(Q);
        t += dt;
        if (do_print) cnt++;
        Hash h;
        hash_field2(Q, h);
        std::printf("loop.%s.%d t=%s dt=%s Q=%016" PRIx64 "\n", c.name, s, hx(t).c_str(), hx(dt).c_str(), h.h);
    }
    char tag[64];
    std::snprintf(tag, sizeof tag, "loop.%s.final", c.name);
    hash_awcm2(awcm, Q, tag);
    // the end of main_2d_euler.cpp (:334): the whole grid from mask0's points
    awcm.wavelet.InterpolateAllGridEuler(Q, mra, awcm.mask0);
    std::snprintf(tag, sizeof tag, "loop.%s.all_grid", c.name);
    hash_awcm2(awcm, Q, tag);
    std::fflush(stdout);
}

}  // namespace

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "host" && argc > 2) {
        probe_vectors();
        probe_eos();
        probe_roe();
        probe_reconstruction();
        probe_fluxes();
        probe_mra();
        probe_problems();
        probe_boundaries();
        probe_awcm_host();
        probe_print(argv[2]);
        probe_omp();
        return 0;
    }
    if (mode == "prepare") {
        run_prepare();
        return 0;
    }
    if (mode == "manager") {
        run_manager();
        return 0;
    }
    if (mode == "loop" && argc > 2) {
        const int steps = std::atoi(argv[2]);
        loop1(kCfg1[0], steps, false);
        loop1(kCfg1[1], steps, false);
        loop1(kCfg1[0], steps, true);
        loop1(kCfg1[2], steps, true);
        loop2<BOUNDARY_MANAGER_QUAD>(kCfg2[0], steps);
        loop2<BOUNDARY_MANAGER_DMACH>(kCfg2[1], steps);
        return 0;
    }
    std::fprintf(stderr, "usage: %s host <dir> | prepare | manager | loop <steps>\n", argv[0]);
    return 2;
}
