// Probe of the reference's scalar AWCM path: AWCM_2D::Prepare / PostProcess
// and WAVELET_2D::Interpolate, driven by the time loop of
// Wavelet/main_2d_advect.cpp (box advection, WENO5 right reconstructions at
// the finest level, AWCM central differences below it, SSP-RK3); then
// AWCM_1D::Prepare / PostProcess and both WAVELET_1D::Interpolate forms.
// TEST INFRASTRUCTURE.  One source, built twice, as api_probe.cpp:
//   * against include/compat/ (this library)          -> tests/native/_build/awcm_api_probe_advect
//   * against the reference's own headers, in place    -> oracle/_ref/awcm_api_probe_advect_ref
// Its own translation unit because PROBLEM_BOX_ADVECTION.h and
// PROBLEM_SHOCKTUBE.h both declare a BOUNDARY_MANAGER.  It prints one line per
// step with FNV-1a hashes and hex doubles, so the two builds must print
// identical text; tests/golden/api_probe_advect.txt holds the reference's.
//
//   awcm_api_probe_advect <steps>
#include <cinttypes>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

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
#include "Problems/2D/PROBLEM_BOX_ADVECTION.h"
#include "WENO.h"
#include "Wavelet/AWCM_1D.h"
#include "Wavelet/AWCM_2D.h"

namespace {

struct Hash {
    uint64_t h = 1469598103934665603ull;
    void byte(uint8_t b) {
        h ^= b;
        h *= 1099511628211ull;
    }
    void f(double x) {
        uint64_t v;
        std::memcpy(&v, &x, 8);
        for (int k = 0; k < 8; k++) byte(uint8_t(v >> (8 * k)));
    }
    void i(int v) {
        for (int k = 0; k < 4; k++) byte(uint8_t(uint32_t(v) >> (8 * k)));
    }
};

// the whole field (ghosts included), the five masks (ghosts included) and
// the three level fields the scalar Prepare computes
void report(const char* tag, AWCM_2D& a, const FIELD_UNIFORM_2D<double>& Q) {
    Hash hq, hm, hl;
    for (int j = Q.j_start_g; j <= Q.j_end_g; j++)
        for (int i = Q.i_start_g; i <= Q.i_end_g; i++) hq.f(Q(i, j));
    const FIELD_UNIFORM_2D<bool>* ms[5] = {&a.mask_temp, &a.mask0, &a.mask1, &a.mask2, &a.mask3};
    for (const auto* m : ms)
        for (int j = m->j_start_g; j <= m->j_end_g; j++)
            for (int i = m->i_start_g; i <= m->i_end_g; i++) hm.byte((*m)(i, j) ? 1 : 0);
    const FIELD_UNIFORM_2D<int>* ls[3] = {&a.level0, &a.level1, &a.level2};
    int n0 = 0;
    for (int j = 0; j < a.mra_grid.Ny; j++)
        for (int i = 0; i < a.mra_grid.Nx; i++) n0 += a.mask0(i, j);
    for (const auto* l : ls)
        for (int j = 0; j < a.mra_grid.Ny; j++)
            for (int i = 0; i < a.mra_grid.Nx; i++) hl.i((*l)(i, j));
    std::printf("%s active=%d Q=%016" PRIx64 " masks=%016" PRIx64 " levels=%016" PRIx64 "\n", tag, n0, hq.h, hm.h,
                hl.h);
}

struct Cfg {
    const char* name;
    int num_lev, nxc;
    double tol;
};
// the last is main_2d_advect.cpp's own configuration (:10-15): 6 levels,
// 15 x 15 coarse, 449 x 449 points, eps 1e-4
const Cfg kCfg[] = {{"small", 4, 9, 1e-4}, {"driver", 6, 15, 1e-4}};

// dF + dG of F at (i, j) of level lev, the loop body of main_2d_advect.cpp
double rhs(AWCM_2D& awcm, WENO5& weno, const FIELD_UNIFORM_2D<double>& F, const GRID_UNIFORM_2D& grid, int i, int j,
           int lev, int max_lev, int d) {
    double dF = 0, dG = 0;
    if (lev == max_lev) {
        double a, b;
        weno.ReconstructRight(F(i - 3, j), F(i - 2, j), F(i - 1, j), F(i, j), F(i + 1, j), a);
        weno.ReconstructRight(F(i - 2, j), F(i - 1, j), F(i, j), F(i + 1, j), F(i + 2, j), b);
        dF = b - a;
        weno.ReconstructRight(F(i, j - 3), F(i, j - 2), F(i, j - 1), F(i, j), F(i, j + 1), a);
        weno.ReconstructRight(F(i, j - 2), F(i, j - 1), F(i, j), F(i, j + 1), F(i, j + 2), b);
        dG = b - a;
    } else {
        if (i <= d) awcm.DifferentiateAtLeftBoundary(F(i - d, j), F(i, j), F(i + d, j), F(i + 2 * d, j), dF);
        else if (i >= grid.i_end - d)
            awcm.DifferentiateAtRightBoundary(F(i - 2 * d, j), F(i - d, j), F(i, j), F(i + d, j), dF);
        else awcm.Differentiate(F(i - 2 * d, j), F(i - d, j), F(i, j), F(i + d, j), F(i + 2 * d, j), dF);
        if (j <= d) awcm.DifferentiateAtLeftBoundary(F(i, j - d), F(i, j), F(i, j + d), F(i, j + 2 * d), dG);
        else if (j >= grid.j_end - d)
            awcm.DifferentiateAtRightBoundary(F(i, j - 2 * d), F(i, j - d), F(i, j), F(i, j + d), dG);
        else awcm.Differentiate(F(i, j - 2 * d), F(i, j - d), F(i, j), F(i, j + d), F(i, j + 2 * d), dG);
    }
    return dF + dG;
}

void run(const Cfg& c, int steps) {
    MRA_GRID_INT_2D mra_grid;
    mra_grid.Initialize(c.num_lev, c.nxc, c.nxc);
    const int max_lev = c.num_lev - 1;
    GRID_UNIFORM_2D grid, grid_x, grid_y, grid_xy;
    FIELD_UNIFORM_2D<double> Q;
    double t_final;
    BOX_ADVECTION::Initialize(mra_grid.grid_finest.i_res, grid, grid_x, grid_y, grid_xy, Q, t_final);
    FIELD_UNIFORM_2D<double> Q_0(grid, 3), Q_1(grid, 3), Q_2(grid, 3);
    AWCM_2D awcm(mra_grid);
    WENO5 weno;
    const double dt = 0.4 * grid.dx;
    BOUNDARY_MANAGER bc(Q), bc1(Q_1), bc2(Q_2);
    char tag[64];
    for (int s = 0; s < steps; s++) {
        awcm.Prepare(Q, c.tol);
        std::snprintf(tag, sizeof tag, "%s.%d.prepare", c.name, s);
        report(tag, awcm, Q);
        bc.SetBoundaryCondition();
        for (int j = grid.j_start; j <= grid.j_end; j++)
            for (int i = grid.i_start; i <= grid.i_end; i++) {
                if (awcm.mask2(i, j) == false) continue;
                const int lev = awcm.level2(i, j), d = mra_grid.grid[lev].dx;
                const double h = grid.dx * d;
                Q_2(i, j) = Q(i, j) - dt / h * rhs(awcm, weno, Q, grid, i, j, lev, max_lev, d);
            }
        bc2.SetBoundaryCondition();
        for (int j = grid.j_start; j <= grid.j_end; j++)
            for (int i = grid.i_start; i <= grid.i_end; i++) {
                if (awcm.mask1(i, j) == false) continue;
                const int lev = awcm.level1(i, j), d = mra_grid.grid[lev].dx;
                const double h = grid.dx * d;
                Q_1(i, j) = 0.75 * Q(i, j) + 0.25 * (Q_2(i, j) - dt / h * rhs(awcm, weno, Q_2, grid, i, j, lev, max_lev, d));
            }
        bc1.SetBoundaryCondition();
        for (int j = grid.j_start; j <= grid.j_end; j++)
            for (int i = grid.i_start; i <= grid.i_end; i++) {
                if (awcm.mask0(i, j) == false) continue;
                const int lev = awcm.level0(i, j), d = mra_grid.grid[lev].dx;
                const double h = grid.dx * d;
                Q_0(i, j) = Q_1(i, j) - dt / h * rhs(awcm, weno, Q_1, grid, i, j, lev, max_lev, d);
            }
        for (int j = grid.j_start; j <= grid.j_end; j++)
            for (int i = grid.i_start; i <= grid.i_end; i++)
                if (awcm.mask0(i, j) == true) Q(i, j) = (1.0 / 3.0 * Q(i, j) + 2.0 / 3.0 * Q_0(i, j));
        awcm.PostProcess(Q);
        std::snprintf(tag, sizeof tag, "%s.%d.step", c.name, s);
        report(tag, awcm, Q);
    }
    // the end of main_2d_advect.cpp (:216-219): the whole grid from mask0; then
    // from mask1 onto mask3 on a fresh copy
    FIELD_UNIFORM_2D<double> Qb(Q);
    FIELD_UNIFORM_2D<bool> mask_init(grid);
    mask_init.AssignAllValues(true);
    awcm.wavelet.Interpolate(Q, mra_grid, awcm.mask0, mask_init);
    std::snprintf(tag, sizeof tag, "%s.interpolate.all", c.name);
    report(tag, awcm, Q);
    awcm.wavelet.Interpolate(Qb, mra_grid, awcm.mask1, awcm.mask3);
    std::snprintf(tag, sizeof tag, "%s.interpolate.m1m3", c.name);
    report(tag, awcm, Qb);
    std::fflush(stdout);
}

// AWCM_1D::Prepare on a Gaussian plus a plateau (three rounds, the field moved
// two points and damped between them as a stand-in for a driver's stages,
// PostProcess after each), then WAVELET_1D::Interpolate from mask0 onto every
// point and from mask1 onto mask3
void run1(int L, int nxc) {
    MRA_GRID_INT_1D mra;
    mra.Initialize(L, nxc);
    const int Nx = mra.Nx;
    FIELD_UNIFORM_1D<double> q(0, Nx, 0.0, 1.0, 3);
    for (int i = -3; i <= Nx + 2; i++) {
        const double x = (i + 0.5) / Nx;
        q(i) = std::exp(-200 * (x - 0.4) * (x - 0.4)) + (std::fabs(x - 0.7) < 0.1 ? 0.5 : 0.0);
    }
    AWCM_1D awcm(mra);
    char tag[64];
    for (int r = 0; r < 3; r++) {
        awcm.Prepare(q, 1e-4);
        Hash hq, hm, hl;
        for (int i = -3; i <= Nx + 2; i++) hq.f(q(i));
        const FIELD_UNIFORM_1D<bool>* ms[5] = {&awcm.mask_temp, &awcm.mask0, &awcm.mask1, &awcm.mask2, &awcm.mask3};
        for (const auto* m : ms)
            for (int i = -3; i <= Nx + 2; i++) hm.byte((*m)(i) ? 1 : 0);
        const FIELD_UNIFORM_1D<int>* ls[4] = {&awcm.level0, &awcm.level1, &awcm.level2, &awcm.level3};
        for (const auto* l : ls)
            for (int i = 0; i < Nx; i++) hl.i((*l)(i));
        std::printf("scalar1d.L%d.%d.prepare Q=%016" PRIx64 " masks=%016" PRIx64 " levels=%016" PRIx64 "\n", L, r, hq.h,
                    hm.h, hl.h);
        awcm.PostProcess(q);
        FIELD_UNIFORM_1D<double> moved(0, Nx, 0.0, 1.0, 3);
        for (int i = -3; i <= Nx + 2; i++) moved(i) = q(i - 2 < -3 ? Nx + 2 - (-3 - (i - 2)) + 1 : i - 2) * 0.999;
        for (int i = -3; i <= Nx + 2; i++) q(i) = moved(i);
    }
    FIELD_UNIFORM_1D<double> qb(0, Nx, 0.0, 1.0, 3);
    for (int i = -3; i <= Nx + 2; i++) qb(i) = q(i);
    awcm.wavelet.Interpolate(q, mra, awcm.mask0);
    awcm.wavelet.Interpolate(qb, mra, awcm.mask1, awcm.mask3);
    Hash ha, hb;
    for (int i = -3; i <= Nx + 2; i++) ha.f(q(i)), hb.f(qb(i));
    std::snprintf(tag, sizeof tag, "scalar1d.L%d", L);
    std::printf("%s.interpolate.all Q=%016" PRIx64 "\n%s.interpolate.m1m3 Q=%016" PRIx64 "\n", tag, ha.h, tag, hb.h);
    std::fflush(stdout);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <steps>\n", argv[0]);
        return 2;
    }
    const int steps = std::atoi(argv[1]);
    run(kCfg[0], steps);
    run(kCfg[1], steps < 10 ? steps : 10);
    run1(7, 17);
    run1(11, 16);
    return 0;
}
