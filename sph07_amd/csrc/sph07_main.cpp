// sph07_main.cpp -- command-line driver: the counterpart of the reference's
// main.cpp (main.cpp:9-59) on the device path, written against the
// reference's class names (include/sph07_amd.hpp).
//
//   sph07_run <deck.mcm> [fp64|fp32c|phases] [max_steps] [--output] [--dump=<path>]
//
// Reads the deck, initialises the scene on the GPU (MSimulationInit::
// Initialize), runs MSolution::Solution until the stop test fires (or
// max_steps), printing the reference's status line every status_interval
// steps (MSolution.cpp:299-304) and the wall time at the end
// (MSimulationData.cpp:110-122).  "phases" runs the same FP64 loop one
// MSolution phase call at a time (MSolution.cpp:250-319) instead of the fused
// device-resident steps.  The last line prints the final step, time and dt
// to full precision and the grid of the last neighbour pass.
//
// --output writes the reference's state output when the deck asks for the
// EnSight CASE format (control card 3, format 2): the state at t = 0 and at
// the end of the first step whose time passes each state-plot time or the
// end time (MSolution.cpp:284-293), under the deck path as base name, with the
// cylinder plot <deck>.drp at each state and, for a rigid-body deck, one line
// of <deck>.rbt per state.  --dump writes the final state as hexadecimal floats.
#include "../../include/sph07_amd.hpp"

#include <chrono>
#include <string>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

using sph07_amd::MSimulationInit;
using sph07_amd::MSolution;

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s <deck.mcm> [fp64|fp32c|phases] [max_steps] [--output] [--dump=<path>]\n",
                argv[0]);
        return 1;
    }
    const char *pos[3] = {argv[1], "fp64", nullptr};
    int npos = 1;
    bool want_output = false;
    const char *dump = nullptr;
    for (int k = 2; k < argc; k++) {
        if (strcmp(argv[k], "--output") == 0) want_output = true;
        else if (strncmp(argv[k], "--dump=", 7) == 0) dump = argv[k] + 7;
        else if (npos < 3) pos[npos++] = argv[k];
    }
    const char *mode = pos[1];
    const bool phases = strcmp(mode, "phases") == 0;
    const int arm = strcmp(mode, "fp32c") == 0 ? SPH_ARM_FP32C : SPH_ARM_FP64;
    const long long max_steps = pos[2] ? atoll(pos[2]) : -1;
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
    const sph_scene_desc *desc = sph_deck_desc(deck);
    const int interval = sph_deck_status_interval(deck);
    const double stpltime = sph_deck_state_time(deck);
    const std::string base = argv[1];
    void *out = nullptr;
    double nextst = 0.0;
    if (want_output && sph_deck_state_opt(deck) == 2) {
        // MOutput::InitStateOutput + StateOutput at initialisation (MSimulationInit.cpp:88-89)
        if (sph_output_open(ctx, base.c_str(), desc->ndim, desc->nummat, desc->mat_id, &out) != SPH_OK ||
            sph_output_state(out) != SPH_OK) {
            fprintf(stderr, "ERROR in state output: %s\n", out ? sph_output_last_error(out) : "open failed");
            sph_deck_free(deck);
            return 1;
        }
        nextst = stpltime;
    } else if (want_output) {
        printf("State output: the deck's format is not EnSight CASE (2); none written.\n");
    }
    sph_globals g;
    sph_get_globals(ctx, &g);
    long long total = 0;
    int since_status = 0;
    for (;;) {
        int chunk = interval - since_status;
        if (max_steps >= 0 && total + chunk > max_steps) chunk = (int)(max_steps - total);
        if (chunk <= 0) break;
        if (out) {
            // stop at the step that passes the next state time: at most half
            // the remaining time at the current step size per call
            const double left = nextst - g.ptime;
            long long k = left <= 0.0 || g.dt <= 0.0 ? 1 : (long long)(0.5 * left / g.dt);
            if (k < 1) k = 1;
            if (k < chunk) chunk = (int)k;
        }
        int done = 0;
        rc = phases ? solution.SolutionByPhases(chunk, &done) : solution.Solution(chunk, &done);
        total += done;
        since_status += done;
        if (rc != SPH_OK) {
            fprintf(stderr, "%s\n", sph_last_error(ctx));
            break;
        }
        sph_get_globals(ctx, &g);
        if (out && (g.ptime > nextst || g.ptime > desc->endtime)) {
            if (sph_output_state(out) != SPH_OK ||
                sph_output_cylinder_plot(out, (base + ".drp").c_str()) != SPH_OK ||
                (desc->rigid_body && sph_output_rigid(out, (base + ".rbt").c_str()) != SPH_OK)) {
                fprintf(stderr, "ERROR in state output: %s\n", sph_output_last_error(out));
                rc = SPH_ERROR;
                break;
            }
            printf("\tState plot written at time %10.4e\n", g.ptime);
            nextst += stpltime;
        }
        if (since_status >= interval || g.done) {
            printf("\tProblem status for time-step %6d\t Time: %10.4e\t dt: %10.4e\n", g.timestep - 1,
                   g.ptime, g.dt);
            since_status = 0;
        }
        if (g.done) break;
    }
    if (out) sph_output_close(out);
    if (dump && rc == SPH_OK && sph_write_state_dump(ctx, dump) != SPH_OK) {
        fprintf(stderr, "ERROR writing the state dump %s\n", dump);
        rc = SPH_ERROR;
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
