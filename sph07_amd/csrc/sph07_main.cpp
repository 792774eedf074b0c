// sph07_main.cpp -- command-line driver: the counterpart of the reference's
// main.cpp (main.cpp:9-59) on the device path.
//
//   sph07_run <deck.mcm> [fp64|fp32c] [max_steps]
//
// Reads the deck, initialises the scene on the GPU (MSimulationInit::
// Initialize), runs MSolution::Solution until the stop test fires (or
// max_steps), printing the reference's status line every status_interval
// steps (MSolution.cpp:299-304) and the wall time at the end
// (MSimulationData.cpp:110-122).
#include "../../include/sph07_amd.h"

#include <chrono>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s <deck.mcm> [fp64|fp32c] [max_steps]\n", argv[0]);
        return 1;
    }
    const int arm = (argc > 2 && strcmp(argv[2], "fp32c") == 0) ? SPH_ARM_FP32C : SPH_ARM_FP64;
    const long long max_steps = argc > 3 ? atoll(argv[3]) : -1;
    char err[512];
    void *deck = nullptr;
    if (sph_deck_read(argv[1], &deck, err, sizeof err) != SPH_OK) {
        fprintf(stderr, "ERROR reading %s: %s\n", argv[1], err);
        return 1;
    }
    printf("%s\n", sph_deck_title(deck));
    void *ctx = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    int rc = sph_create(sph_deck_desc(deck), arm, &ctx, err, sizeof err);
    if (rc != SPH_OK) {
        fprintf(stderr, "ERROR in initialisation: %s\n", err);
        sph_deck_free(deck);
        return 1;
    }
    printf("\nProblem Initialisation Complete\n");
    const int interval = sph_deck_status_interval(deck);
    sph_globals g;
    long long total = 0;
    for (;;) {
        int chunk = interval;
        if (max_steps >= 0 && total + chunk > max_steps) chunk = (int)(max_steps - total);
        if (chunk <= 0) break;
        int done = 0;
        rc = sph_solution(ctx, chunk, &done);
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
    printf("\nRun completed: %lld steps, %d particles, %.3f s (%.4e particle-updates/s)\n", total, g.np,
           secs, secs > 0 ? (double)g.np * (double)total / secs : 0.0);
    sph_destroy(ctx);
    sph_deck_free(deck);
    return rc == SPH_OK ? 0 : 1;
}
