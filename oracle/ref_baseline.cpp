// ref_baseline.cpp - times the reference's own step functions on the host (test
// infrastructure; built by oracle/Makefile `ref` next to the reference objects).
//
// Configures the reference's globals for one rank through the harness, runs
// InitialConditions and then `steps` iterations of the loop of src/Simulation.cpp:138-161
// (MoveParticles, BoundaryConditions, Indexing, CollideParticles, Sampling), and prints one
// JSON line with particle-steps/s measured around the step loop only.
//
// usage: ref_baseline nx ny npc_free wall(0 diffuse|1 specular|2 cll) aoa_deg steps
//                     first_sampling_step
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "../include/dsmc_c.h"

extern "C" {
int ref_configure(const dsmc_params* p);
void ref_set_seed(unsigned u, unsigned v);
void ref_initial_conditions();
void ref_clear_cells();
void ref_set_step(int step);
int ref_step();
int ref_np();
}

int main(int argc, char** argv)
{
    if (argc < 8) {
        std::fprintf(stderr, "usage: %s nx ny npc wall aoa steps first_sampling\n", argv[0]);
        return 2;
    }
    dsmc_params p;
    dsmc_params_default(&p);
    p.nx = std::atoi(argv[1]);
    p.ny = std::atoi(argv[2]);
    p.npc_free = std::atof(argv[3]);
    p.wall_model = std::atoi(argv[4]);
    p.attack_angle_deg = std::atof(argv[5]);
    const int steps = std::atoi(argv[6]);
    p.first_sampling_step = std::atoi(argv[7]);
    if (ref_configure(&p) != 0) return 1;
    ref_set_seed(521288629u, 362436069u);
    ref_clear_cells();
    ref_initial_conditions();
    ref_set_step(0);
    double particle_steps = 0.0;
    const auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < steps; ++s) {
        ref_step();
        particle_steps += ref_np();
    }
    const auto t1 = std::chrono::steady_clock::now();
    const double sec = std::chrono::duration<double>(t1 - t0).count();
    std::printf("{\"particle_steps_per_s\": %.6e, \"steps\": %d, \"seconds\": %.6f, "
                "\"np_final\": %d, \"cores\": 1}\n",
                particle_steps / sec, steps, sec, ref_np());
    return 0;
}

// dsmc_params_default without the product library: the reference's defaults.
extern "C" void dsmc_params_default(dsmc_params* p)
{
    *p = dsmc_params{};
    p->x1 = -20.0; p->x2 = 20.0; p->y1 = -20.0; p->y2 = 20.0; p->dz = 0.1;
    p->nx = 100; p->ny = 100; p->dl = 0.1;
    p->box_width = 4.0; p->box_height = 4.0;
    p->dt = 1.0e-5; p->n_step = 100000; p->first_sampling_step = 10000; p->sampling_period = 1;
    p->npc_free = 10; p->molar_mass = 0.040; p->mach = 4; p->n_free = 1e17; p->t_free = 200;
    p->t_wall = 300.0; p->d_ref = 4.17e-10; p->viscosity_index = 0.81; p->t_ref = 273.0;
    p->wall_model = DSMC_WALL_CLL; p->acc_n = 0.8; p->acc_t = 0.8; p->seed = 1;
}
