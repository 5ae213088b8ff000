/* A C client of the C ABI (compiled and linked by the CPU tests, not run):
 * the reference's main() sequence -- basis, Hamiltonian, three-pass Lanczos,
 * ground-state energy -- plus observables and the Lanczos record. */
#include <stdio.h>
#include <stdlib.h>

#include "blqh.h"

static void die(const char* what) {
    fprintf(stderr, "%s: %s\n", what, blqh_last_error());
    exit(1);
}

int main(void) {
    blqh_sector s = {2, 8, 8, 4, -1, 4, 4};
    blqh_basis* B;
    if (blqh_basis_create(&s, &B)) die("basis");
    int64_t n = blqh_basis_size(B);
    blqh_params p = {4.0, 4.0, 0, 3.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    blqh_hamil* H;
    if (blqh_hamil_build(B, &p, &H)) die("hamiltonian");
    blqh_lanczos_opts o = {200, 1, 1e-10, 1, -1, NULL, 1e-7, 0, 0.0};
    blqh_lanczos_result r;
    double* psi = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    double alpha[200], beta[201];
    if (blqh_lanczos(H, &o, &r, psi, alpha, beta)) die("lanczos");
    printf("E_gs:= %.10g\n", r.E_rq / 8);
    blqh_observables ob;
    if (blqh_compute_observables(B, psi, &ob)) die("observables");
    double eig[200], first[200], vec0[200];
    if (blqh_tridiag_spectrum(r.iterations, alpha, beta + 1, eig, first, vec0)) die("spectrum");
    if (blqh_record_save("lanczos.txt", n, r.iterations, 1, eig, beta, alpha, vec0, first)) die("record");
    free(psi);
    blqh_hamil_destroy(H);
    blqh_basis_destroy(B);
    return 0;
}
