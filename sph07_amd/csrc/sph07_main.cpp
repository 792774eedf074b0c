// sph07_main.cpp -- command-line driver: the counterpart of the reference's
// main.cpp (main.cpp:9-59) on the device path, written against the
// reference's class names (include/sph07_amd.hpp).
//
//   sph07_run <deck.mcm> [fp64|fp32c|phases] [max_steps]
//
// Reads the deck, initialises the scene on the GPU (MSimulationInit::
// Initialize), runs MSolution::Solution until the stop test fires (or
// max_steps), printing the reference's status line every status_interval
// steps (MSolution.cpp:299-304) and the wall time at the end
// (MSimulationData.cpp:110-122).  "phases" runs the same FP64 loop one
// MSolution phase call at a time (MSolution.cpp:250-319) instead of the fused
// device-resident steps.  The last line prints the final step, time and dt
// to full precision and the grid of the last neighbour pass.
#include "../../include/sph07_amd.hpp"

#include <chrono>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

using sph07_amd::MSimulationInit;
using sph07_amd::MSolution;

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s <deck.mcm> [fp64|fp32c|phases] [max_steps]\n", argv[0]);
        return 1;
    }
    const char *mode = argc > 2 ? argv[2] : "fp64";
    const bool phases = strcmp(mode, "phases") == 0;
    const int arm = strcmp(mode, "fp32c") == 0 ? SPH_ARM_FP32C : SPH_ARM_FP64;
    const long long max_steps = argc > 3 ? atoll(argv[3]) : -1;
    char err[512];
    void *deck = nullptr;
    if (sph_deck_read(argv[1], &deck, err, sizeof err) != SPH_OK) {
        fprintf(stderr, "ERROR reading %s: %s\n", argv[1], err);
        return 1;
    }
    printf("%s\n", sph_deck_title(deck));
    const auto t0 = std::chrono::steady_clock::now();
    MSimulationInit init;
    int rc = init.Initialize(sph_deck_desc(deck), arm, err, sizeof err);
    if (rc != SPH_OK) {
        fprintf(stderr, "ERROR in initialisation: %s\n", err);
        sph_deck_free(deck);
        return 1;
    }
    void *ctx = init.Context();
    MSolution solution(ctx, sph_deck_desc(deck)->ndim);
    printf("\nProblem Initialisation Complete\n");
    const int interval = sph_deck_status_interval(deck);
    sph_globals g;
    long long total = 0;
    for (;;) {
        int chunk = interval;
        if (max_steps >= 0 && total + chunk > max_steps) chunk = (int)(max_steps - total);
        if (chunk <= 0) break;
        int done = 0;
        rc = phases ? solution.SolutionByPhases(chunk, &done) : solution.Solution(chunk, &done);
        total += done;
        if (rc != SPH_OK) {
            fprintf(stderr, "%s\n", sph_last_error(ctx));
            break;
        }
        sph_get_globals(ctx, &g);
        printf("\tProblem status for time-step %6d\t Time: %10.4e\t dt: %10.4e\n", g.timestep - 1,
               g.ptime, g.dt);
        if (g.done) break;
    }
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    sph_get_globals(ctx, &g);
    sph07_amd::MNeighbourVars &nv = solution.NeighbourVars();
    nv.Refresh();
    printf("\nRun completed: %lld steps, %d particles, %.3f s (%.4e particle-updates/s)\n", total, g.np,
           secs, secs > 0 ? (double)g.np * (double)total / secs : 0.0);
    printf("final: timestep %d ptime %.17g dt %.17g grid %d %d %d\n", g.timestep, g.ptime, g.dt,
           nv.sph_gridlim[0], nv.sph_gridlim[1], nv.sph_gridlim[2]);
    sph_deck_free(deck);
    return rc == SPH_OK ? 0 : 1;
}
