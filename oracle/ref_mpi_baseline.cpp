// ref_mpi_baseline.cpp - runs the reference's own RunSimulation (src/Simulation.cpp:64-247),
// MPI domain decomposition included, with its compile-time globals set for one workload
// (test infrastructure; built by oracle/Makefile `ref` next to the reference objects).
//
// usage: mpiexec -n R ref_mpi_baseline nx ny npc_free wall(0 diffuse|1 specular|2 cll) aoa_deg
//                                      steps first_sampling_step
// Rank 0 prints one JSON line with the wall-clock of the whole call (setup, initial fill, the
// step loop, the field gather and the output files). Two runs with different step counts
// difference away everything but the steps; the reference's own progress lines (global
// particle count every 100 steps) give the particles of those steps.
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "Boundary.h"
#include "Simulation.h"

extern int NX, NY, NStep, FirstSamplingStep, MPIRank, MPISize;
extern double NPCFree, AttackAngle;

int main(int argc, char** argv)
{
    if (argc < 8) {
        std::fprintf(stderr, "usage: %s nx ny npc wall aoa steps first_sampling\n", argv[0]);
        return 2;
    }
    NX = std::atoi(argv[1]);
    NY = std::atoi(argv[2]);
    NPCFree = std::atof(argv[3]);
    const int wall = std::atoi(argv[4]);
    WallInteraction = wall == 0 ? WallModel::Diffuse : wall == 1 ? WallModel::Specular : WallModel::CLL;
    AttackAngle = std::atof(argv[5]);
    NStep = std::atoi(argv[6]);
    FirstSamplingStep = std::atoi(argv[7]);
    char* args[] = {argv[0], nullptr};
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = RunSimulation(1, args);
    const auto t1 = std::chrono::steady_clock::now();
    if (MPIRank == 0)
        std::printf("{\"seconds\": %.6f, \"steps\": %d, \"ranks\": %d, \"rc\": %d}\n",
                    std::chrono::duration<double>(t1 - t0).count(), NStep, MPISize, rc);
    return rc;
}
