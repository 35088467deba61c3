// Host build of the product's physics and adjoint headers, so the CPU tests can
// check the reverse-mode products against finite differences of the forward
// functions.  Compiled by tests/test_adjoint_cpu.py with g++ -ffp-contract=off.
#include "awcm/adjoint.h"

using namespace awcm;

extern "C" {

void shim_face_flux1(double g, const double* s, double alpha, double* flux) {
    face_flux1(g, reinterpret_cast<const double(*)[3]>(s), alpha, flux);
}
void shim_face_flux1_vjp(double g, const double* s, double alpha, const double* fbar, double* sbar) {
    face_flux1_vjp(g, reinterpret_cast<const double(*)[3]>(s), alpha, fbar, reinterpret_cast<double(*)[3]>(sbar));
}
void shim_flux_f1(double g, const double* q, double* f) { flux_f1(g, q, f); }
void shim_flux_f1_vjp(double g, const double* q, const double* fbar, double* qbar) { flux_f1_vjp(g, q, fbar, qbar); }
double shim_wenoz_left(const double* x) { return wenoz_left(x[0], x[1], x[2], x[3], x[4]); }
void shim_wenoz_left_vjp(const double* x, double obar, double* xbar) { wenoz_left_vjp(x, obar, xbar); }

}  // extern "C"
