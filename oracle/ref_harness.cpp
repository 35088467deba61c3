// Harness that drives the REFERENCE's own headers (doseongju/C_pp-library,
// compiled in place from /root/reference by oracle/Makefile) behind the same C
// entry points as the oracle, so tests can compare the two bit for bit.
// TEST INFRASTRUCTURE ONLY; the build output lives in oracle/_ref/.
//
// The reference ships no buildable driver (its main_*.cpp use `void main` and
// sprintf_s), so the time-loop bodies below restate
//   Wavelet/main_1d_euler.cpp:61-260   (with the WENO-Z flux pair in place of
//                                        WENO_EULER_SOLVER_1D::ComputedF,
//                                        WENO.h:176-184)
//   Wavelet/main_2d_euler(dmach).cpp:56-350 and (quad).cpp (same loop)
// while every numerical routine they call is the reference's own class.
//
// Two build fixes are applied without touching the reference files:
//   * `-include cmath` on the command line (the headers call sqrt without it);
//   * FIELD_UNIFORM_1D<bool>::AssignAllValues is given an explicit
//     specialisation here, because the ARRAY_1D::AssignAllValue it forwards to
//     does not exist (FIELD_UNIFORM_1D.h:65-68).

#include "FIELD_UNIFORM_1D.h"
template <>
void FIELD_UNIFORM_1D<bool>::AssignAllValues(const bool& v) {
    for (int i = 0; i < arr.i_res; i++) arr.val[i] = v;
}
#include "WENOZ.h"
#include "WENO.h"
#include "Problems/1D/PROBLEM_SHOCKTUBE.h"
#include "Problems/1D/PROBLEM_ENTROPY_INTERACTING.h"
#include "Wavelet/AWCM_1D.h"
#include "Problems/2D/PROBLEM_QUADRANTS.h"
#include "Problems/2D/PROBLEM_DOUBLEMACH.h"
#include "Wavelet/AWCM_2D.h"

#include <cstdint>
#include <vector>
#include <cstring>

namespace {

struct Ref1D {
    MRA_GRID_INT_1D mra;
    AWCM_1D* awcm;
    FIELD_UNIFORM_1D<VECTOR_3D<double>> Q, Q_temp, Q_copy;
    FIELD_UNIFORM_1D<double> Qs;  // the scalar field of AWCM_1D::Prepare
    BOUNDARY_MANAGER* bcm;
    WENOZ_EULER_SOLVER_1D* wenoz;
    WENO_EULER_SOLVER_1D* weno;
    bool js = false;  // the shipped driver's WENO5-JS ComputedF in place of the WENO-Z pair
    IDEAL_GAS_1D* EOS;
    double gamma, dx;
    int max_level;
    Ref1D(int L, int nxc, double g, double dx_) : gamma(g), dx(dx_) {
        mra.Initialize(L, nxc);
        awcm = new AWCM_1D(mra);
        Q.Initialize(0, mra.Nx, 0.0, 1.0, 3);
        Qs.Initialize(0, mra.Nx, 0.0, 1.0, 3);
        bcm = new BOUNDARY_MANAGER(Q);
        wenoz = new WENOZ_EULER_SOLVER_1D(gamma);
        weno = new WENO_EULER_SOLVER_1D(gamma);
        EOS = new IDEAL_GAS_1D(gamma);
        max_level = L - 1;
    }
    ~Ref1D() { delete awcm; delete bcm; delete wenoz; delete weno; delete EOS; }

    void dF_at(int i, int lev, double alpha_x, VECTOR_3D<double>& dF) {
        const int dx_ = mra.dx[lev];
        const GRID_UNIFORM_1D& grid = Q.base_grid;
        if (lev == max_level && js) {
            // main_1d_euler.cpp:109
            weno->ComputedF(Q_temp(i - 3), Q_temp(i - 2), Q_temp(i - 1), Q_temp(i), Q_temp(i + 1), Q_temp(i + 2),
                            Q_temp(i + 3), alpha_x, dF);
        } else if (lev == max_level) {
            VECTOR_3D<double> F_l, F_r;
            wenoz->ComputeFlux(Q_temp(i - 3), Q_temp(i - 2), Q_temp(i - 1), Q_temp(i), Q_temp(i + 1), Q_temp(i + 2), alpha_x, F_l);
            wenoz->ComputeFlux(Q_temp(i - 2), Q_temp(i - 1), Q_temp(i), Q_temp(i + 1), Q_temp(i + 2), Q_temp(i + 3), alpha_x, F_r);
            dF = (F_r - F_l);
        } else {
            if (i - 2 * dx_ > 0 && i + 2 * dx_ < grid.i_end)
                awcm->DifferentiateEuler(Q_temp(i - 2 * dx_), Q_temp(i - dx_), Q_temp(i), Q_temp(i + dx_), Q_temp(i + 2 * dx_), dF);
            else if (i - 2 * dx_ < 0)
                awcm->DifferentiateEulerAtLeftBoundary(Q_temp(i - dx_), Q_temp(i), Q_temp(i + dx_), Q_temp(i + 2 * dx_), dF);
            else if (i == 2 * dx_)
                awcm->DifferentiateEulerAtLeftBoundary(Q_temp(i - 2 * dx_), Q_temp(i - dx_), Q_temp(i), Q_temp(i + dx_), Q_temp(i + 2 * dx_), dF);
            else if (i > grid.i_end - 2 * dx_)
                awcm->DifferentiateEulerAtRightBoundary(Q_temp(i - 2 * dx_), Q_temp(i - dx_), Q_temp(i), Q_temp(i + dx_), dF);
            else
                awcm->DifferentiateEulerAtRightBoundary(Q_temp(i - 2 * dx_), Q_temp(i - dx_), Q_temp(i), Q_temp(i + dx_), Q_temp(i + 2 * dx_), dF);
        }
    }
    double alpha(FIELD_UNIFORM_1D<bool>& m, FIELD_UNIFORM_1D<int>& l) {
        double a = 0;
        ITERATION_GRID_1D(Q.base_grid, 0) {
            if (m(i) == false || l(i) != max_level) continue;
            double u = Q(i)[1] / Q(i)[0];
            double c = EOS->c(Q(i));
            a = MAX(a, ABS(u) + c);
        }
        return a;
    }
    double step(double tol, double cfl, const double* pt, int npt, int* cnt, double* t) {
        const GRID_UNIFORM_1D& grid = Q.base_grid;
        awcm->PrepareEuler(Q, tol);
        double max_velocity = 0;
        ITERATION_GRID_1D(grid, 0) {
            if (awcm->mask3(i) == false) continue;
            const double c = EOS->c(Q(i));
            const double u = Q(i)[1] / Q(i)[0];
            max_velocity = MAX(max_velocity, (ABS(u) + c));
        }
        double dt = cfl * dx / max_velocity;
        bool do_print = false;
        if (*cnt < npt && *t + dt > pt[*cnt]) { dt = pt[*cnt] - *t; do_print = true; }
        bcm->SetBoundaryCondition();
        double alpha_x = alpha(awcm->mask3, awcm->level3);
        Q_copy = Q;
        Q_temp = Q;
        ITERATION_GRID_1D(grid, 0) {
            if (awcm->mask2(i) == false) continue;
            const int lev = awcm->level2(i);
            const int dx_ = mra.dx[lev];
            VECTOR_3D<double> dF;
            dF_at(i, lev, alpha_x, dF);
            Q(i) = Q(i) - dt / (dx * dx_) * dF;
        }
        bcm->SetBoundaryCondition();
        alpha_x = alpha(awcm->mask2, awcm->level2);
        Q_temp = Q;
        ITERATION_GRID_1D(grid, 0) {
            if (awcm->mask1(i) == false) continue;
            const int lev = awcm->level1(i);
            const int dx_ = mra.dx[lev];
            VECTOR_3D<double> dF;
            dF_at(i, lev, alpha_x, dF);
            Q(i) = 0.75 * Q_copy(i) + 0.25 * (Q(i) - dt / (dx * dx_) * dF);
        }
        bcm->SetBoundaryCondition();
        alpha_x = alpha(awcm->mask1, awcm->level1);
        Q_temp = Q;
        ITERATION_GRID_1D(grid, 0) {
            if (awcm->mask0(i) == false) continue;
            const int lev = awcm->level0(i);
            const int dx_ = mra.dx[lev];
            VECTOR_3D<double> dF;
            dF_at(i, lev, alpha_x, dF);
            Q(i) = (Q_copy(i) + 2.0 * (Q(i) - dt / (dx * dx_) * dF)) / 3.0;
        }
        awcm->PostProcessEuler(Q);
        *t += dt;
        if (do_print) (*cnt)++;
        return dt;
    }
};

enum { BC_QUAD = 0, BC_DMACH = 1 };

struct Ref2D {
    MRA_GRID_INT_2D mra_grid;
    AWCM_2D* awcm;
    GRID_UNIFORM_2D grid;
    FIELD_UNIFORM_2D<VECTOR_4D<double>> Q, Q_0, Q_1, Q_2;
    FIELD_UNIFORM_2D<double> Qs;  // the scalar field of AWCM_2D::Prepare
    BOUNDARY_MANAGER_QUAD *bq, *bq0, *bq1, *bq2;
    BOUNDARY_MANAGER_DMACH *bd, *bd0, *bd1, *bd2;
    EULER_EQUATION_2D* euler_eqn;
    ARRAY_1D<double> alpha_x, alpha_y;
    double gamma;
    int bc_kind, max_lev;
    Ref2D(int L, int nxc, int nyc, double g, double xl, double xu, double yl, double yu, int bck) : gamma(g), bc_kind(bck) {
        mra_grid.Initialize(L, nxc, nyc);
        awcm = new AWCM_2D(mra_grid);
        grid.Initialize(0, 0, mra_grid.Nx, mra_grid.Ny, xl, yl, xu, yu);
        Q.Initialize(grid, 3); Q_0.Initialize(grid, 3); Q_1.Initialize(grid, 3); Q_2.Initialize(grid, 3);
        Qs.Initialize(grid, 3);
        bq = new BOUNDARY_MANAGER_QUAD(Q); bq0 = new BOUNDARY_MANAGER_QUAD(Q_0);
        bq1 = new BOUNDARY_MANAGER_QUAD(Q_1); bq2 = new BOUNDARY_MANAGER_QUAD(Q_2);
        bd = new BOUNDARY_MANAGER_DMACH(Q); bd0 = new BOUNDARY_MANAGER_DMACH(Q_0);
        bd1 = new BOUNDARY_MANAGER_DMACH(Q_1); bd2 = new BOUNDARY_MANAGER_DMACH(Q_2);
        euler_eqn = new EULER_EQUATION_2D(gamma);
        alpha_x.Initialize(0, mra_grid.Ny, true);
        alpha_y.Initialize(0, mra_grid.Nx, true);
        max_lev = L - 1;
    }
    ~Ref2D() {
        delete awcm; delete bq; delete bq0; delete bq1; delete bq2;
        delete bd; delete bd0; delete bd1; delete bd2; delete euler_eqn;
    }
    void set_bc(int which, double t) {
        if (bc_kind == BC_QUAD) { (which == 0 ? bq : which == 1 ? bq1 : bq2)->SetBoundaryCondition(); }
        else { (which == 0 ? bd : which == 1 ? bd1 : bd2)->SetBoundaryCondition(t); }
    }
    void alphas(FIELD_UNIFORM_2D<VECTOR_4D<double>>& F, FIELD_UNIFORM_2D<bool>& m) {
        for (int k = 0; k < alpha_x.i_res; k++) alpha_x.val[k] = 0;
        for (int k = 0; k < alpha_y.i_res; k++) alpha_y.val[k] = 0;
        ITERATION_GRID_2D(grid, 0) {
            if (m(i, j) == 0) continue;
            alpha_x(j) = MAX(alpha_x(j), euler_eqn->Spectral_RadiusX(F(i, j)));
        }
        ITERATION_GRID_2D_IJ(grid, 0) {
            if (m(i, j) == 0) continue;
            alpha_y(i) = MAX(alpha_y(i), euler_eqn->Spectral_RadiusY(F(i, j)));
        }
    }
    void rhs(FIELD_UNIFORM_2D<VECTOR_4D<double>>& Qs, int i, int j, int lev, VECTOR_4D<double>& dF, VECTOR_4D<double>& dG,
             WENOZ_EULER_SOLVER_2D& wenoz_solver) {
        VECTOR_4D<double> F_ll, F_l, F_c, F_r, F_rr, G_bb, G_b, G_c, G_t, G_tt;
        const int d = mra_grid.grid[lev].dx;
        if (lev == max_lev) {
            VECTOR_4D<double> a, b;
            wenoz_solver.ComputeFluxX(Qs(i - 3, j), Qs(i - 2, j), Qs(i - 1, j), Qs(i, j), Qs(i + 1, j), Qs(i + 2, j), alpha_x(j), a);
            wenoz_solver.ComputeFluxX(Qs(i - 2, j), Qs(i - 1, j), Qs(i, j), Qs(i + 1, j), Qs(i + 2, j), Qs(i + 3, j), alpha_x(j), b);
            dF = b - a;
            wenoz_solver.ComputeFluxY(Qs(i, j - 3), Qs(i, j - 2), Qs(i, j - 1), Qs(i, j), Qs(i, j + 1), Qs(i, j + 2), alpha_y(i), a);
            wenoz_solver.ComputeFluxY(Qs(i, j - 2), Qs(i, j - 1), Qs(i, j), Qs(i, j + 1), Qs(i, j + 2), Qs(i, j + 3), alpha_y(i), b);
            dG = b - a;
        } else {
            // The reference evaluates F(Q(i-2d)) / F(Q(i+2d)) even in the one-sided
            // branches, where the value is unused; only the used ones are read here.
            if (i <= d) {
                euler_eqn->ComputeFu(Qs(i - d, j), F_l), euler_eqn->ComputeFu(Qs(i, j), F_c);
                euler_eqn->ComputeFu(Qs(i + d, j), F_r), euler_eqn->ComputeFu(Qs(i + 2 * d, j), F_rr);
                awcm->DifferentiateAtLeftBoundary(F_l, F_c, F_r, F_rr, dF);
            } else if (i >= grid.i_end - d) {
                euler_eqn->ComputeFu(Qs(i - 2 * d, j), F_ll), euler_eqn->ComputeFu(Qs(i - d, j), F_l), euler_eqn->ComputeFu(Qs(i, j), F_c);
                euler_eqn->ComputeFu(Qs(i + d, j), F_r);
                awcm->DifferentiateAtRightBoundary(F_ll, F_l, F_c, F_r, dF);
            } else {
                euler_eqn->ComputeFu(Qs(i - 2 * d, j), F_ll), euler_eqn->ComputeFu(Qs(i - d, j), F_l), euler_eqn->ComputeFu(Qs(i, j), F_c);
                euler_eqn->ComputeFu(Qs(i + d, j), F_r), euler_eqn->ComputeFu(Qs(i + 2 * d, j), F_rr);
                awcm->Differentiate(F_ll, F_l, F_c, F_r, F_rr, dF);
            }
            if (j <= d) {
                euler_eqn->ComputeGu(Qs(i, j - d), G_b), euler_eqn->ComputeGu(Qs(i, j), G_c);
                euler_eqn->ComputeGu(Qs(i, j + d), G_t), euler_eqn->ComputeGu(Qs(i, j + 2 * d), G_tt);
                awcm->DifferentiateAtLeftBoundary(G_b, G_c, G_t, G_tt, dG);
            } else if (j >= grid.j_end - d) {
                euler_eqn->ComputeGu(Qs(i, j - 2 * d), G_bb), euler_eqn->ComputeGu(Qs(i, j - d), G_b), euler_eqn->ComputeGu(Qs(i, j), G_c);
                euler_eqn->ComputeGu(Qs(i, j + d), G_t);
                awcm->DifferentiateAtRightBoundary(G_bb, G_b, G_c, G_t, dG);
            } else {
                euler_eqn->ComputeGu(Qs(i, j - 2 * d), G_bb), euler_eqn->ComputeGu(Qs(i, j - d), G_b), euler_eqn->ComputeGu(Qs(i, j), G_c);
                euler_eqn->ComputeGu(Qs(i, j + d), G_t), euler_eqn->ComputeGu(Qs(i, j + 2 * d), G_tt);
                awcm->Differentiate(G_bb, G_b, G_c, G_t, G_tt, dG);
            }
        }
    }
    double step(double tol, double cfl, const double* pt, int npt, int* cnt, double* t) {
        awcm->PrepareEuler(Q, tol);
        double max_velocity = 0;
        ITERATION_GRID_2D(grid, 0) {
            if (awcm->mask3(i, j) == false) continue;
            max_velocity = MAX(max_velocity, euler_eqn->Spectral_RadiusX(Q(i, j)) + euler_eqn->Spectral_RadiusY(Q(i, j)));
        }
        const double dx = grid.dx;
        double dt = cfl * dx / max_velocity;
        bool do_print = false;
        if (*cnt < npt && *t + dt > pt[*cnt]) { dt = pt[*cnt] - *t; do_print = true; }
        WENOZ_EULER_SOLVER_2D wenoz_solver;
        set_bc(0, *t);
        alphas(Q, awcm->mask3);
        ITERATION_GRID_2D(grid, 0) {
            if (awcm->mask2(i, j) == false) continue;
            const int lev = awcm->level2(i, j);
            const int d = mra_grid.grid[lev].dx;
            const double dxl = grid.dx * d;
            VECTOR_4D<double> dF, dG;
            rhs(Q, i, j, lev, dF, dG, wenoz_solver);
            Q_2(i, j) = Q(i, j) - dt / dxl * (dF + dG);
        }
        set_bc(2, *t);
        alphas(Q_2, awcm->mask2);
        ITERATION_GRID_2D(grid, 0) {
            if (awcm->mask1(i, j) == false) continue;
            const int lev = awcm->level1(i, j);
            const int d = mra_grid.grid[lev].dx;
            const double dxl = grid.dx * d;
            VECTOR_4D<double> dF, dG;
            rhs(Q_2, i, j, lev, dF, dG, wenoz_solver);
            Q_1(i, j) = 0.75 * Q(i, j) + 0.25 * (Q_2(i, j) - dt / dxl * (dF + dG));
        }
        set_bc(1, *t);
        alphas(Q_1, awcm->mask1);
        ITERATION_GRID_2D(grid, 0) {
            if (awcm->mask0(i, j) == false) continue;
            const int lev = awcm->level0(i, j);
            const int d = mra_grid.grid[lev].dx;
            const double dxl = grid.dx * d;
            VECTOR_4D<double> dF, dG;
            rhs(Q_1, i, j, lev, dF, dG, wenoz_solver);
            Q_0(i, j) = Q_1(i, j) - dt / dxl * (dF + dG);
        }
        ITERATION_GRID_2D(grid, 0) {
            if (awcm->mask0(i, j) == true) Q(i, j) = (Q(i, j) + 2.0 * Q_0(i, j)) / 3.0;
        }
        awcm->PostProcessEuler(Q);
        *t += dt;
        if (do_print) (*cnt)++;
        return dt;
    }
};

template <class F>
void copy_mask1(uint8_t* out, const F& m) { for (int k = 0; k < m.arr.i_res; k++) out[k] = m.arr.val[k] ? 1 : 0; }
template <class F>
void copy_mask2(uint8_t* out, const F& m) { for (int k = 0; k < m.arr.ij_res; k++) out[k] = m.arr.val[k] ? 1 : 0; }

}  // namespace

extern "C" {

// ---- reference initial conditions, verbatim problem set-ups -------------------
// kind 0: SHOCKTUBE (Sod on [0,10], split 5); kind 1: ENTROPY_INTERACTING (Shu-Osher on [-5,5])
int ref_ic1(int kind, int res, double* q, double* dx_out, double* x0_out) {
    GRID_UNIFORM_1D grid, grid_x;
    FIELD_UNIFORM_1D<VECTOR_3D<double>> Q;
    double t_final, gamma;
    if (kind == 0) SHOCKTUBE::Initialize(res, grid, grid_x, Q, t_final, gamma);
    else ENTROPY_INTERACTING::Initialize(res, grid, grid_x, Q, t_final, gamma);
    for (int i = -3; i <= res + 2; i++) for (int c = 0; c < 3; c++) q[(i + 3) * 3 + c] = Q(i)[c];
    *dx_out = grid.dx; *x0_out = grid.x_min;
    return 0;
}
// kind 0: QUADRANTS (split 0.8) on res x res; kind 1: DOUBLEMACH ((res-1)*4+1) x res
int ref_ic2(int kind, int res, double* q, double* dx_out, double* dy_out) {
    GRID_UNIFORM_2D grid, gx, gy, gxy;
    FIELD_UNIFORM_2D<VECTOR_4D<double>> Q;
    double t_final, gamma;
    if (kind == 0) QUADRANTS::Initialize(res, grid, gx, gy, gxy, Q, t_final, gamma);
    else DOUBLEMACH::Initialize(res, grid, gx, gy, gxy, Q, t_final, gamma);
    for (int k = 0; k < Q.arr.ij_res; k++) for (int c = 0; c < 4; c++) q[k * 4 + c] = Q.arr.val[k][c];
    *dx_out = grid.dx; *dy_out = grid.dy;
    return 0;
}

void* ref1_create(int num_levels, int nx_coarse, double gamma, double dx) { return new Ref1D(num_levels, nx_coarse, gamma, dx); }
void ref1_destroy(void* h) { delete static_cast<Ref1D*>(h); }
int ref1_nx(void* h) { return static_cast<Ref1D*>(h)->mra.Nx; }
void ref1_set_q(void* h, const double* q) {
    auto* s = static_cast<Ref1D*>(h);
    for (int k = 0; k < s->Q.arr.i_res; k++) for (int c = 0; c < 3; c++) s->Q.arr.val[k][c] = q[k * 3 + c];
}
void ref1_get_q(void* h, double* q) {
    auto* s = static_cast<Ref1D*>(h);
    for (int k = 0; k < s->Q.arr.i_res; k++) for (int c = 0; c < 3; c++) q[k * 3 + c] = s->Q.arr.val[k][c];
}
void ref1_get_masks(void* h, uint8_t* masks, int32_t* levels) {
    auto* s = static_cast<Ref1D*>(h);
    const int n = s->mra.Nx + 6, N = s->mra.Nx;
    copy_mask1(masks, s->awcm->mask_temp);
    copy_mask1(masks + n, s->awcm->mask0); copy_mask1(masks + 2 * n, s->awcm->mask1);
    copy_mask1(masks + 3 * n, s->awcm->mask2); copy_mask1(masks + 4 * n, s->awcm->mask3);
    FIELD_UNIFORM_1D<int>* L[4] = {&s->awcm->level0, &s->awcm->level1, &s->awcm->level2, &s->awcm->level3};
    for (int k = 0; k < 4; k++) for (int i = 0; i < N; i++) levels[k * N + i] = (*L[k])(i);
}
void ref1_set_mask0(void* h, const uint8_t* m0) {
    auto* s = static_cast<Ref1D*>(h);
    for (int k = 0; k < s->awcm->mask0.arr.i_res; k++) s->awcm->mask0.arr.val[k] = m0[k] != 0;
}
void ref1_prepare(void* h, double tol) { auto* s = static_cast<Ref1D*>(h); s->awcm->PrepareEuler(s->Q, tol); }
void ref1_set_flux(void* h, int kind) { static_cast<Ref1D*>(h)->js = kind == 1; }
void ref1_set_scalar(void* h, const double* q) {
    auto* s = static_cast<Ref1D*>(h);
    for (int k = 0; k < s->Qs.arr.i_res; k++) s->Qs.arr.val[k] = q[k];
}
void ref1_get_scalar(void* h, double* q) {
    auto* s = static_cast<Ref1D*>(h);
    for (int k = 0; k < s->Qs.arr.i_res; k++) q[k] = s->Qs.arr.val[k];
}
void ref1_prepare_scalar(void* h, double tol) { auto* s = static_cast<Ref1D*>(h); s->awcm->Prepare(s->Qs, tol); }
void ref1_interpolate_euler(void* h, const uint8_t* mask_from, const uint8_t* mask_to) {
    auto* s = static_cast<Ref1D*>(h);
    FIELD_UNIFORM_1D<bool> from, to;
    from.Initialize(0, s->mra.Nx, 0.0, 1.0, 3);
    to.Initialize(0, s->mra.Nx, 0.0, 1.0, 3);
    for (int k = 0; k < from.arr.i_res; k++) from.arr.val[k] = mask_from[k] != 0;
    if (!mask_to) {
        s->awcm->wavelet.InterpolateEuler(s->Q, s->mra, from);
        return;
    }
    for (int k = 0; k < to.arr.i_res; k++) to.arr.val[k] = mask_to[k] != 0;
    s->awcm->wavelet.InterpolateEuler(s->Q, s->mra, from, to);
}
void ref1_interpolate_scalar(void* h, const uint8_t* mask_from, const uint8_t* mask_to) {
    auto* s = static_cast<Ref1D*>(h);
    // ghost width 3, as the masks arrive (FIELD_UNIFORM_1D's copy constructor
    // would drop the ghosts: FIELD_UNIFORM_1D.h:31-39)
    FIELD_UNIFORM_1D<bool> from, to;
    from.Initialize(0, s->mra.Nx, 0.0, 1.0, 3);
    to.Initialize(0, s->mra.Nx, 0.0, 1.0, 3);
    for (int k = 0; k < from.arr.i_res; k++) from.arr.val[k] = mask_from[k] != 0;
    if (!mask_to) {
        s->awcm->wavelet.Interpolate(s->Qs, s->mra, from);
        return;
    }
    for (int k = 0; k < to.arr.i_res; k++) to.arr.val[k] = mask_to[k] != 0;
    s->awcm->wavelet.Interpolate(s->Qs, s->mra, from, to);
}
double ref1_step(void* h, double tol, double cfl, const double* pt, int npt, int* cnt, double* t) {
    return static_cast<Ref1D*>(h)->step(tol, cfl, pt, npt, cnt, t);
}
void ref_flux1(double gamma, const double* st, double alpha, double* flux) {
    WENOZ_EULER_SOLVER_1D s(gamma);
    VECTOR_3D<double> v[6], f;
    for (int p = 0; p < 6; p++) v[p] = VECTOR_3D<double>(st[p * 3], st[p * 3 + 1], st[p * 3 + 2]);
    s.ComputeFlux(v[0], v[1], v[2], v[3], v[4], v[5], alpha, f);
    for (int c = 0; c < 3; c++) flux[c] = f[c];
}
void ref_flux1_js(double gamma, const double* st, double alpha, double* flux) {
    WENO_EULER_SOLVER_1D s(gamma);
    VECTOR_3D<double> v[6], f;
    for (int p = 0; p < 6; p++) v[p] = VECTOR_3D<double>(st[p * 3], st[p * 3 + 1], st[p * 3 + 2]);
    s.ComputeFlux(v[0], v[1], v[2], v[3], v[4], v[5], alpha, f);
    for (int c = 0; c < 3; c++) flux[c] = f[c];
}
void ref_flux2(double gamma, int dir, const double* st, double alpha, double* flux) {
    WENOZ_EULER_SOLVER_2D s(gamma);
    VECTOR_4D<double> v[6], f;
    for (int p = 0; p < 6; p++) for (int c = 0; c < 4; c++) v[p][c] = st[p * 4 + c];
    if (dir == 0) s.ComputeFluxX(v[0], v[1], v[2], v[3], v[4], v[5], alpha, f);
    else s.ComputeFluxY(v[0], v[1], v[2], v[3], v[4], v[5], alpha, f);
    for (int c = 0; c < 4; c++) flux[c] = f[c];
}
double ref_wenoz_left(double a, double b, double c, double d, double e) { WENOZ5 w; double o; w.ReconstructLeft(a, b, c, d, e, o); return o; }
double ref_wenoz_right(double a, double b, double c, double d, double e) { WENOZ5 w; double o; w.ReconstructRight(a, b, c, d, e, o); return o; }
void ref_weights(double* out) {
    WAVELET_1D w;
    for (int i = 0; i < 4; i++) for (int j = -1; j < 4; j++) out[i * 5 + j + 1] = w.weight(i, j);
}

void* ref2_create(int num_levels, int nx_coarse, int ny_coarse, double gamma, double xl, double xu, double yl, double yu, int bc_kind) {
    return new Ref2D(num_levels, nx_coarse, ny_coarse, gamma, xl, xu, yl, yu, bc_kind);
}
void ref2_destroy(void* h) { delete static_cast<Ref2D*>(h); }
int ref2_nx(void* h) { return static_cast<Ref2D*>(h)->mra_grid.Nx; }
int ref2_ny(void* h) { return static_cast<Ref2D*>(h)->mra_grid.Ny; }
void ref2_set_q(void* h, const double* q) {
    auto* s = static_cast<Ref2D*>(h);
    for (int k = 0; k < s->Q.arr.ij_res; k++) for (int c = 0; c < 4; c++) s->Q.arr.val[k][c] = q[k * 4 + c];
}
void ref2_get_q(void* h, double* q) {
    auto* s = static_cast<Ref2D*>(h);
    for (int k = 0; k < s->Q.arr.ij_res; k++) for (int c = 0; c < 4; c++) q[k * 4 + c] = s->Q.arr.val[k][c];
}
void ref2_get_masks(void* h, uint8_t* masks, int32_t* levels) {
    auto* s = static_cast<Ref2D*>(h);
    const size_t n = s->awcm->mask0.arr.ij_res;
    copy_mask2(masks, s->awcm->mask_temp);
    copy_mask2(masks + n, s->awcm->mask0); copy_mask2(masks + 2 * n, s->awcm->mask1);
    copy_mask2(masks + 3 * n, s->awcm->mask2); copy_mask2(masks + 4 * n, s->awcm->mask3);
    FIELD_UNIFORM_2D<int>* L[4] = {&s->awcm->level0, &s->awcm->level1, &s->awcm->level2, &s->awcm->level3};
    const size_t nl = size_t(s->mra_grid.Nx) * s->mra_grid.Ny;
    for (int k = 0; k < 4; k++) for (size_t i = 0; i < nl; i++) levels[k * nl + i] = L[k]->arr.val[i];
}
void ref2_set_mask0(void* h, const uint8_t* m0) {
    auto* s = static_cast<Ref2D*>(h);
    for (int k = 0; k < s->awcm->mask0.arr.ij_res; k++) s->awcm->mask0.arr.val[k] = m0[k] != 0;
}
void ref2_prepare(void* h, double tol) { auto* s = static_cast<Ref2D*>(h); s->awcm->PrepareEuler(s->Q, tol); }
void ref2_set_scalar(void* h, const double* q) {
    auto* s = static_cast<Ref2D*>(h);
    for (int k = 0; k < s->Qs.arr.ij_res; k++) s->Qs.arr.val[k] = q[k];
}
void ref2_get_scalar(void* h, double* q) {
    auto* s = static_cast<Ref2D*>(h);
    for (int k = 0; k < s->Qs.arr.ij_res; k++) q[k] = s->Qs.arr.val[k];
}
void ref2_prepare_scalar(void* h, double tol) { auto* s = static_cast<Ref2D*>(h); s->awcm->Prepare(s->Qs, tol); }
static void bool_field(FIELD_UNIFORM_2D<bool>& f, const uint8_t* m) {
    for (int k = 0; k < f.arr.ij_res; k++) f.arr.val[k] = m[k] != 0;
}
void ref2_interpolate_scalar(void* h, const uint8_t* mask_from, const uint8_t* mask_to) {
    auto* s = static_cast<Ref2D*>(h);
    FIELD_UNIFORM_2D<bool> from(s->awcm->mask0), to(s->awcm->mask0);
    bool_field(from, mask_from);
    bool_field(to, mask_to);
    s->awcm->wavelet.Interpolate(s->Qs, s->mra_grid, from, to);
}
void ref2_interpolate_euler(void* h, const uint8_t* mask_from, const uint8_t* mask_to) {
    auto* s = static_cast<Ref2D*>(h);
    FIELD_UNIFORM_2D<bool> from(s->awcm->mask0), to(s->awcm->mask0);
    bool_field(from, mask_from);
    bool_field(to, mask_to);
    s->awcm->wavelet.InterpolateEuler(s->Q, s->mra_grid, from, to);
}
void ref2_interpolate_all(void* h, const uint8_t* mask_from) {
    auto* s = static_cast<Ref2D*>(h);
    FIELD_UNIFORM_2D<bool> from(s->awcm->mask0);
    for (int k = 0; k < from.arr.ij_res; k++) from.arr.val[k] = mask_from[k] != 0;
    s->awcm->wavelet.InterpolateAllGridEuler(s->Q, s->mra_grid, from);
}
double ref2_step(void* h, double tol, double cfl, const double* pt, int npt, int* cnt, double* t) {
    return static_cast<Ref2D*>(h)->step(tol, cfl, pt, npt, cnt, t);
}

// AWCM_MANAGER_1D / AWCM_MANAGER_2D statics on caller arrays, the reference's
// own functions (op 0 AddAdjacentZone, 1 CheckReconstruction, 2
// AddPointForDifferentiation, 3 ComputeLevel): the AWCM object's mask3 and
// level3 fields (the reference's own extents) carry the arguments and are
// restored afterwards.
void ref1_mask_op(void* h, int op, uint8_t* mask, int32_t* level) {
    auto* s = static_cast<Ref1D*>(h);
    FIELD_UNIFORM_1D<bool>& m = s->awcm->mask3;
    FIELD_UNIFORM_1D<int>& lv = s->awcm->level3;
    const int n = m.arr.i_res, N = s->mra.Nx;
    std::vector<bool> keep_m(n);
    std::vector<int> keep_l(N);
    for (int k = 0; k < n; k++) { keep_m[k] = m.arr.val[k]; m.arr.val[k] = mask[k] != 0; }
    for (int i = 0; i < N; i++) { keep_l[i] = lv(i); lv(i) = level ? level[i] : 0; }
    if (op == 0) AWCM_MANAGER_1D::AddAdjacentZone(m, s->mra);
    else if (op == 1) AWCM_MANAGER_1D::CheckReconstruction(m, s->mra);
    else if (op == 2) AWCM_MANAGER_1D::AddPointForDifferentiation(m, s->mra, lv);
    else AWCM_MANAGER_1D::ComputeLevel(lv, m, s->mra);
    for (int k = 0; k < n; k++) { mask[k] = m.arr.val[k] ? 1 : 0; m.arr.val[k] = keep_m[k]; }
    for (int i = 0; i < N; i++) { if (op == 3) level[i] = lv(i); lv(i) = keep_l[i]; }
}
void ref2_mask_op(void* h, int op, uint8_t* mask, int32_t* level) {
    auto* s = static_cast<Ref2D*>(h);
    FIELD_UNIFORM_2D<bool>& m = s->awcm->mask3;
    FIELD_UNIFORM_2D<int>& lv = s->awcm->level3;
    const size_t n = m.arr.ij_res, nl = size_t(s->mra_grid.Nx) * s->mra_grid.Ny;
    std::vector<bool> keep_m(n);
    std::vector<int> keep_l(nl);
    for (size_t k = 0; k < n; k++) { keep_m[k] = m.arr.val[k]; m.arr.val[k] = mask[k] != 0; }
    for (size_t k = 0; k < nl; k++) { keep_l[k] = lv.arr.val[k]; lv.arr.val[k] = level ? level[k] : 0; }
    if (op == 0) AWCM_MANAGER_2D::AddAdjacentZone(m, s->mra_grid);
    else if (op == 1) AWCM_MANAGER_2D::CheckReconstruction(m, s->mra_grid);
    else if (op == 2) AWCM_MANAGER_2D::AddPointForDifferentiation(m, s->mra_grid, lv);
    else AWCM_MANAGER_2D::ComputeLevel(lv, m, s->mra_grid);
    for (size_t k = 0; k < n; k++) { mask[k] = m.arr.val[k] ? 1 : 0; m.arr.val[k] = keep_m[k]; }
    for (size_t k = 0; k < nl; k++) { if (op == 3) level[k] = lv.arr.val[k]; lv.arr.val[k] = keep_l[k]; }
}
// RemoveResidue (AWCM_MANAGER_1D.h:135-157, both overloads) on the AWCM
// object's fields, restored afterwards
void ref1_remove_residue(void* h, double* q, int ncomp, const uint8_t* mask) {
    auto* s = static_cast<Ref1D*>(h);
    FIELD_UNIFORM_1D<bool>& m = s->awcm->mask3;
    const int n = m.arr.i_res;
    std::vector<bool> keep_m(n);
    for (int k = 0; k < n; k++) { keep_m[k] = m.arr.val[k]; m.arr.val[k] = mask[k] != 0; }
    if (ncomp == 1) {
        FIELD_UNIFORM_1D<double>& f = s->Qs;
        std::vector<double> keep(f.arr.val, f.arr.val + f.arr.i_res);
        for (int k = 0; k < f.arr.i_res; k++) f.arr.val[k] = q[k];
        AWCM_MANAGER_1D::RemoveResidue(f, m, s->mra);
        for (int k = 0; k < f.arr.i_res; k++) { q[k] = f.arr.val[k]; f.arr.val[k] = keep[k]; }
    } else {
        FIELD_UNIFORM_1D<VECTOR_3D<double>>& f = s->Q;
        std::vector<VECTOR_3D<double>> keep(f.arr.val, f.arr.val + f.arr.i_res);
        for (int k = 0; k < f.arr.i_res; k++) for (int c = 0; c < 3; c++) f.arr.val[k][c] = q[k * 3 + c];
        AWCM_MANAGER_1D::RemoveResidue(f, m, s->mra);
        for (int k = 0; k < f.arr.i_res; k++) { for (int c = 0; c < 3; c++) q[k * 3 + c] = f.arr.val[k][c]; f.arr.val[k] = keep[k]; }
    }
    for (int k = 0; k < n; k++) m.arr.val[k] = keep_m[k];
}
// RemoveResidue / RemoveResidueEuler (AWCM_MANAGER_2D.h:257-280)
void ref2_remove_residue(void* h, double* q, int ncomp, const uint8_t* mask) {
    auto* s = static_cast<Ref2D*>(h);
    FIELD_UNIFORM_2D<bool>& m = s->awcm->mask3;
    const size_t n = m.arr.ij_res;
    std::vector<bool> keep_m(n);
    for (size_t k = 0; k < n; k++) { keep_m[k] = m.arr.val[k]; m.arr.val[k] = mask[k] != 0; }
    if (ncomp == 1) {
        FIELD_UNIFORM_2D<double>& f = s->Qs;
        std::vector<double> keep(f.arr.val, f.arr.val + f.arr.ij_res);
        for (int k = 0; k < f.arr.ij_res; k++) f.arr.val[k] = q[k];
        AWCM_MANAGER_2D::RemoveResidue(f, m, s->mra_grid);
        for (int k = 0; k < f.arr.ij_res; k++) { q[k] = f.arr.val[k]; f.arr.val[k] = keep[k]; }
    } else {
        FIELD_UNIFORM_2D<VECTOR_4D<double>>& f = s->Q;
        std::vector<VECTOR_4D<double>> keep(f.arr.val, f.arr.val + f.arr.ij_res);
        for (int k = 0; k < f.arr.ij_res; k++) for (int c = 0; c < 4; c++) f.arr.val[k][c] = q[k * 4 + c];
        AWCM_MANAGER_2D::RemoveResidueEuler(f, m, s->mra_grid);
        for (int k = 0; k < f.arr.ij_res; k++) { for (int c = 0; c < 4; c++) q[k * 4 + c] = f.arr.val[k][c]; f.arr.val[k] = keep[k]; }
    }
    for (size_t k = 0; k < n; k++) m.arr.val[k] = keep_m[k];
}

}  // extern "C"
