/* o_sancheck.c -- sanitizer driver for the oracle (test infrastructure only):
 * tests/test_sanitizers.py builds it with the oracle sources under
 * -fsanitize=address,undefined and under -fsanitize=thread and runs it on a
 * small config with two worker threads: fixed steps, one RHS evaluation and
 * its transpose, an adaptive step, and a two-step adjoint march. */
#include <stdio.h>
#include <stdlib.h>
#include "oracle.h"

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    char err[512] = "";
    orc_sim* s = orc_create(argv[1], argc - 2, (const char**)(argv + 2), NULL, 2, err, sizeof err);
    if (!s) { printf("create: %s\n", err); return 2; }
    int R, T, NZ;
    orc_dims(s, &R, &T, &NZ);
    double a, b;
    if (orc_fixed_step(s, 2e-6, &a, &b)) { printf("fixed_step: %s\n", orc_error(s)); return 4; }
    double* f = malloc(sizeof(double) * 2 * (size_t)R * T);
    double* g = malloc(sizeof(double) * 2 * (size_t)R * T);
    double* h = malloc(sizeof(double) * 2 * (size_t)R * T);
    double* k = malloc(sizeof(double) * 2 * (size_t)R * T);
    orc_get(s, "field", f);
    if (orc_nl_response(s, f, g)) { printf("nl_response: %s\n", orc_error(s)); return 4; }
    if (orc_nl_response_vjp(s, f, g, h)) { printf("nl_response_vjp: %s\n", orc_error(s)); return 4; }
    if (orc_propagator(s)) { printf("propagator: %s\n", orc_error(s)); return 4; }
    srand(1);
    for (size_t i = 0; i < 2 * (size_t)R * T; ++i) g[i] = rand() / (double)RAND_MAX - 0.5;
    const double rk[2] = {2e-6, 1.5e-6}, lin[2] = {1.5e-6, 2e-6};
    if (orc_adjoint_march(s, 2, rk, lin, g, NULL, h, k)) { printf("adjoint: %s\n", orc_error(s)); return 4; }
    printf("oracle %d x %d ok, z = %g, maxA2 = %g\n", R, T, orc_scalar(s, "CURRENT_Z"), a);
    free(f); free(g); free(h); free(k);
    orc_destroy(s);
    return 0;
}
