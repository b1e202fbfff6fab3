// Host build of the device contact-potential math (rigid-ipc_amd/csrc/
// rigc_barrier.hpp) so that its closed-form derivatives can be checked on
// the CPU against the oracle's dual-number derivatives. Test
// infrastructure: the library is loaded by tests/test_closed_forms.py only.
#include "../../rigid-ipc_amd/csrc/rigc_barrier.hpp"

using namespace rigc;

extern "C" {

void cf_body_frame(const double* r, double* dR, double* d2R) {
    BodyFrame F;
    body_frame(r, F);
    for (int c = 0; c < 3; c++)
        for (int e = 0; e < 9; e++) dR[9 * c + e] = F.dR[c][e];
    for (int s = 0; s < 6; s++)
        for (int e = 0; e < 9; e++) d2R[9 * s + e] = F.d2R[s][e];
}

// Distance of the constraint's closest feature pair and the form kind.
double cf_distance(int type, const double* X, int* kind) {
    DistForm f;
    constraint_form(type, X, f);
    *kind = f.kind;
    return f.d;
}

double cf_barrier_local(int type, const double* X, double eps_x, double dhat, double dmin, long long mult, double* g,
                        double* H) {
    BarrierLocal L;
    barrier_local_setup(L, type, X, eps_x, mult, dhat, dmin);
    for (int r = 0; r < 12; r++) g[r] = barrier_local_grad(L, r / 3, r % 3);
    for (int r = 0; r < 12; r++)
        for (int q = 0; q < 12; q++) H[12 * r + q] = barrier_local_hess(L, r / 3, r % 3, q / 3, q % 3);
    return barrier_local_energy(L);
}

void cf_friction_frame(int type, const double* X, double* coords, double* basis) {
    double b[3][2];
    friction_frame(type, X, coords, b);
    for (int k = 0; k < 3; k++) {
        basis[2 * k] = b[k][0];
        basis[2 * k + 1] = b[k][1];
    }
}

}  // extern "C"
