// Driver of the 1-D kernels' emulation with a workgroup of real threads
// (tests/native/emu1d.cpp built against tests/native/emu_mt): a short Sod run
// recorded on the adaptive grid (k1_solve), its frozen replay and the reverse
// sweep (k1_adjoint, two recompute segment lengths); with a third argument 1,
// the recording alone with the WENO5-JS stage fluxes.  Built with
// ThreadSanitizer, any access of two threads to the same memory between
// barriers, without atomics, is reported.  The run also checks that the
// replay reproduces the recording bit for bit and that the two segment
// lengths give the same gradient.  Prints "ok" and exits 0 on success.
// Test infrastructure only.
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

extern "C" {
void* emu_create(int L, int nxc, double gamma, double dx, const double* pt, int npt);
void emu_destroy(void* h);
void emu_set_q(void* h, const double* q);
void emu_get_q(void* h, double* q);
int emu_record(void* h, double tol, double cfl, double t_final, int max_steps, int K);
void emu_replay(void* h);
void emu_adjoint(void* h, const double* gf, int seg_steps, double* g0);
int emu_status(void* h);
void emu_set_flux(int kind);
}

int main(int argc, char** argv) {
    const int L = argc > 1 ? std::atoi(argv[1]) : 4;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 8;
    const int flux = argc > 3 ? std::atoi(argv[3]) : 0;  // 1: WENO5-JS stage fluxes (no adjoint)
    emu_set_flux(flux);
    const int nxc = 17;
    const int Nx = (nxc - 1) * (1 << (L - 1)) + 1, NP = Nx + 6;
    const double dx = 1.0 / (Nx - 1), g = 1.4;
    // Sod: (rho, u, p) = (1, 0, 1) left of x = 0.5, (0.125, 0, 0.1) right; ghosts copy the ends
    std::vector<double> q(3 * NP);
    for (int k = 0; k < NP; k++) {
        const int i = std::min(std::max(k - 3, 0), Nx - 1);
        const bool left = i * dx < 0.5;
        const double rho = left ? 1.0 : 0.125, p = left ? 1.0 : 0.1;
        q[3 * k] = rho;
        q[3 * k + 1] = 0.0;
        q[3 * k + 2] = p / (g - 1.0);
    }
    const double pt[2] = {0.2, 0.2};
    void* e = emu_create(L, nxc, g, dx, pt, 2);
    emu_set_q(e, q.data());
    const int n = emu_record(e, 1e-4, 0.5, 0.2, steps, 4);
    if (flux) {
        // the replay and the adjoint are the WENO-Z pair's: the recording alone
        const int st = emu_status(e);
        emu_destroy(e);
        std::printf("steps %d status %d\n", n, st);
        if (n != steps || st != 0) return 1;
        std::printf("ok\n");
        return 0;
    }
    std::vector<double> rec(3 * NP), rep(3 * NP), g8(3 * NP), g4(3 * NP), gf(3 * NP);
    emu_get_q(e, rec.data());
    emu_set_q(e, q.data());
    emu_replay(e);
    emu_get_q(e, rep.data());
    for (int k = 0; k < 3 * NP; k++) gf[k] = (k % 7) * 0.125 - 0.375;
    emu_adjoint(e, gf.data(), 8, g8.data());
    emu_adjoint(e, gf.data(), 4, g4.data());
    const int st = emu_status(e);
    emu_destroy(e);
    const bool same_replay = std::memcmp(rec.data(), rep.data(), sizeof(double) * rec.size()) == 0;
    const bool same_grad = std::memcmp(g8.data(), g4.data(), sizeof(double) * g8.size()) == 0;
    std::printf("steps %d status %d replay %s gradient %s\n", n, st, same_replay ? "same" : "DIFFERS",
                same_grad ? "same" : "DIFFERS");
    if (n != steps || st != 0 || !same_replay || !same_grad) return 1;
    std::printf("ok\n");
    return 0;
}
