/* gsolver_oracle <config> [NAME VALUE]... -- runs the oracle march like the
 * reference binary (solver_main.cpp:151-263), dumps into the current
 * directory.  Test infrastructure only. */
#include <stdio.h>
#include "oracle.h"

int main(int argc, char** argv)
{
    if (argc < 2) { printf("\n!!! Error: Too few arguements  !!!\n"); return 2; }
    char err[512] = "";
    orc_sim* s = orc_create(argv[1], argc - 2, (const char**)(argv + 2), ".", 1, err, sizeof err);
    if (!s) { printf("\n!!!Error in process 0 occured!!! , %s\n", err); return 2; }
    int rc = orc_run(s, ".", 0);
    if (rc < 0) { printf("\n!!!Error in process 0 occured!!! , %s\n", orc_error(s)); orc_destroy(s); return 2; }
    orc_destroy(s);
    return 0;
}
