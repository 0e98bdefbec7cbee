/*
 * cpu_baseline.c - times the oracle's step loop on the host (test infrastructure).
 *
 * Used only by bench.py's cpu_baseline leg. Runs InitialConditions and then the loop of
 * src/Simulation.cpp:138-199 on one core in the oracle's REF mode (the reference's own
 * algorithm: one MWC stream, swap-with-last removal, fixed SCMax) for a bounded number of
 * steps, and prints one JSON line with particle-steps/s measured around the step loop only.
 *
 * usage: cpu_baseline nx ny npc_free wall(0 diffuse|1 specular|2 cll) aoa_deg steps
 *                     [warm_steps] [first_sampling_step]
 */
#define _POSIX_C_SOURCE 199309L
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "dsmc_oracle.h"

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char** argv)
{
    if (argc < 7) {
        fprintf(stderr, "usage: %s nx ny npc_free wall aoa steps [warm_steps] [first_sampling]\n",
                argv[0]);
        return 2;
    }
    dsmc_params p;
    or_params_default(&p);
    p.nx = atoi(argv[1]);
    p.ny = atoi(argv[2]);
    p.npc_free = atof(argv[3]);
    p.wall_model = atoi(argv[4]);
    p.attack_angle_deg = atof(argv[5]);
    int steps = atoi(argv[6]);
    int warm = argc > 7 ? atoi(argv[7]) : 0;
    p.first_sampling_step = argc > 8 ? atoi(argv[8]) : 0;
    p.max_per_cell = 1u << 20;
    int st = 0;
    or_ctx* c = or_create(&p, OR_MODE_REF, &st);
    if (!c) { fprintf(stderr, "create failed %d\n", st); return 1; }
    or_init_freestream(c);
    if (warm > 0 && or_step(c, warm)) { fprintf(stderr, "warm-up failed\n"); return 1; }
    double particle_steps = 0.0;
    double t0 = now();
    for (int s = 0; s < steps; ++s) {
        int rc = or_step(c, 1);
        if (rc) { fprintf(stderr, "step failed %d\n", rc); return 1; }
        particle_steps += (double)or_num_particles(c);
    }
    double t1 = now();
    printf("{\"particle_steps_per_s\": %.6e, \"steps\": %d, \"seconds\": %.6f, "
           "\"np_final\": %zu, \"cores\": 1}\n",
           particle_steps / (t1 - t0), steps, t1 - t0, or_num_particles(c));
    or_destroy(c);
    return 0;
}
