// Initial conditions of the configured problems, host side, in the reference
// layout (AoS, ghost 3).  Each mirrors one reference Initialize routine with
// the domain, split point and states as parameters:
//   awcm_ic_riemann_1d     SHOCKTUBE::Initialize            (PROBLEM_SHOCKTUBE.h:14-54)
//   awcm_ic_shu_osher_1d   ENTROPY_INTERACTING::Initialize  (PROBLEM_ENTROPY_INTERACTING.h:14-49)
//   awcm_ic_quadrants_2d   QUADRANTS::Initialize            (PROBLEM_QUADRANTS.h:15-81)
//   awcm_ic_double_mach_2d DOUBLEMACH::Initialize           (PROBLEM_DOUBLEMACH.h:13-63)
//   awcm_ic_box_advection_2d BOX_ADVECTION::Initialize      (PROBLEM_BOX_ADVECTION.h:15-47)
// Coordinates and conserved variables are formed with the same expressions so
// the fields are bit-identical to the reference's for the same parameters.

#include <cmath>
#include <cstddef>

#include "awcm/physics.h"
#include "awcm_capi.h"

namespace {

constexpr int G = 3;

// Cell-centred 1-D grid of the reference problems: GRID_UNIFORM_1D(0, res,
// xl + dx'/2, xu - dx'/2) with dx' the spacing of the res+1 face grid.
void cell_grid(int res, double xl, double xu, double& xmin, double& dx) {
    const double dxx = (xu - xl) / ((res + 1) - 1);
    xmin = xl + 0.5 * dxx;
    const double xmax = xu - 0.5 * dxx;
    dx = (xmax - xmin) / (res - 1);
}

}  // namespace

extern "C" {

int awcm_ic_riemann_1d(int res, double xl, double xu, double split, const double* left, const double* right,
                       double gamma, double* q, double* dx_out, double* x0_out) {
    if (res < 2 || !left || !right || !q || !(xu > xl)) return AWCM_EINVAL;
    double xmin, dx;
    cell_grid(res, xl, xu, xmin, dx);
    for (int i = -G; i <= res - 1 + G; i++) {
        const double x = xmin + dx * (i - 0);
        const double* s = (x <= split) ? left : right;
        const double rho = s[0], u = s[1], p = s[2];
        double* o = q + size_t(i + G) * 3;
        o[0] = rho;
        o[1] = rho * u;
        o[2] = awcm::eos_E1(gamma, rho, u, p);
    }
    if (dx_out) *dx_out = dx;
    if (x0_out) *x0_out = xmin;
    return AWCM_OK;
}

int awcm_ic_shu_osher_1d(int res, double xl, double xu, double split, const double* left, double amp, double gamma,
                         double* q, double* dx_out, double* x0_out) {
    if (res < 2 || !left || !q || !(xu > xl)) return AWCM_EINVAL;
    double xmin, dx;
    cell_grid(res, xl, xu, xmin, dx);
    for (int i = -G; i <= res - 1 + G; i++) {
        const double x = xmin + dx * (i - 0);
        double rho, u, p;
        if (x < split) { rho = left[0]; u = left[1]; p = left[2]; }
        else { rho = 1 + amp * std::sin(5 * x); u = 0; p = 1; }
        double* o = q + size_t(i + G) * 3;
        o[0] = rho;
        o[1] = rho * u;
        o[2] = awcm::eos_E1(gamma, rho, u, p);
    }
    if (dx_out) *dx_out = dx;
    if (x0_out) *x0_out = xmin;
    return AWCM_OK;
}

int awcm_ic_box_advection_2d(int res, double* q, double* dx_out) {
    if (res < 2 || !q) return AWCM_EINVAL;
    // grid_xy: res+1 points on [0, 1]; the cell grid from dx'/2 to 1 - dx'/2
    const double dxy = (1.0 - 0.0) / ((res + 1) - 1);
    const double xmin = 0.0 + 0.5 * dxy, xmax = 1.0 - 0.5 * dxy;
    const double dx = (xmax - xmin) / (res - 1);
    const int P = res + 2 * G;
    for (int j = -G; j <= res - 1 + G; j++)
        for (int i = -G; i <= res - 1 + G; i++) {
            const double x = xmin + (i - 0) * dx, y = xmin + (j - 0) * dx;
            q[size_t(i + G) + size_t(P) * (j + G)] = std::exp(-50 * ((x - 0.5) * (x - 0.5)) - 50 * ((y - 0.5) * (y - 0.5)));
        }
    if (dx_out) *dx_out = dx;
    return AWCM_OK;
}

int awcm_ic_quadrants_2d(int res, double split, const double* states, double gamma, double* q, double* dx_out) {
    if (res < 2 || !states || !q) return AWCM_EINVAL;
    const double dx = (1.0 - 0.0) / (res - 1), dy = (1.0 - 0.0) / (res - 1);
    const int P = res + 2 * G;
    for (int j = -G; j <= res - 1 + G; j++)
        for (int i = -G; i <= res - 1 + G; i++) {
            const double x = 0.0 + (i - 0) * dx, y = 0.0 + (j - 0) * dy;
            int k;
            if (x >= split && y >= split) k = 0;
            else if (x < split && y >= split) k = 1;
            else if (x < split && y < split) k = 2;
            else k = 3;
            const double rho = states[4 * k], u = states[4 * k + 1], v = states[4 * k + 2], p = states[4 * k + 3];
            double* o = q + (size_t(i + G) + size_t(P) * (j + G)) * 4;
            o[0] = rho;
            o[1] = rho * u;
            o[2] = rho * v;
            o[3] = awcm::eos_E2(gamma, rho, u, v, p);
        }
    if (dx_out) *dx_out = dx;
    return AWCM_OK;
}

int awcm_ic_double_mach_2d(int res, double rho_pre, double gamma, double* q, double* dx_out) {
    if (res < 2 || !q) return AWCM_EINVAL;
    const int nx = (res - 1) * 4 + 1, ny = res;
    const double dx = (4.0 - 0.0) / (nx - 1), dy = (1.0 - 0.0) / (ny - 1);
    const int P = nx + 2 * G;
    for (int j = -G; j <= ny - 1 + G; j++)
        for (int i = -G; i <= nx - 1 + G; i++) {
            const double x = 0.0 + (i - 0) * dx, y = 0.0 + (j - 0) * dy;
            double rho, u, v, p;
            if (y > std::sqrt(3.0) * (x - 1.0 / 6.0)) {
                rho = 8; u = 8.25 * std::sqrt(3.0) * 0.5; v = -8.25 * 0.5; p = 116.5;
            } else {
                rho = rho_pre; u = 0; v = 0; p = 1;
            }
            double* o = q + (size_t(i + G) + size_t(P) * (j + G)) * 4;
            o[0] = rho;
            o[1] = rho * u;
            o[2] = rho * v;
            o[3] = awcm::eos_E2(gamma, rho, u, v, p);
        }
    if (dx_out) *dx_out = dx;
    return AWCM_OK;
}

}  // extern "C"
