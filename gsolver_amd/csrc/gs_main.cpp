// gsolver_amd <config> [NAME VALUE]... [--precision fp64|fp32|mixed]
// Command-line entry point with the reference's contract
// (`gsolver_solid_qdht <config> [NAME VALUE]...`, process_input.cpp:35-43,300):
// reads the config, marches to the last dump plane writing the .udf files into
// the current directory, exits 2 with the reference's message on an error
// (solver_main.cpp:234-247).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gsolver_amd.h"

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::printf("\n!!! Error: Too few arguements  !!!\n");
        return 2;
    }
    int prec = GS_MIXED;
    std::vector<const char*> over;
    for (int i = 2; i < argc; ++i) {
        if (std::strcmp(argv[i], "--precision") == 0 && i + 1 < argc) {
            const std::string v = argv[++i];
            prec = v == "fp64" ? GS_FP64 : (v == "fp32" ? GS_FP32 : GS_MIXED);
            continue;
        }
        over.push_back(argv[i]);
    }
    std::printf("*** UPPE solver (MI355X HIP version) ***\n");
    gs_plan* p = nullptr;
    if (gs_plan_create(argv[1], (int)over.size(), over.data(), prec, ".", &p) != 0) {
        std::printf("\n!!!Error in process 0 occured!!! , %s\n", gs_last_error());
        return 2;
    }
    const int rc = gs_march(p, ".", 0);
    if (rc < 0) {
        std::printf("\n!!!Error in process 0 occured!!! , %s\n", gs_last_error());
        gs_plan_destroy(p);
        return 2;
    }
    double steps = 0, rej = 0;
    gs_get_scalar(p, "STEP_N", &steps);
    gs_get_scalar(p, "n_reject", &rej);
    std::printf("\nDone: %.0f steps (%.0f rejected)\n", steps, rej);
    gs_plan_destroy(p);
    return 0;
}
