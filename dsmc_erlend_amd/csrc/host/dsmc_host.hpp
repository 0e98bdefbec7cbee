// dsmc_host.hpp - C++ host side of the gfx950 DSMC pipeline, over the C-ABI of dsmc_c.h.
//
// Mirrors the reference's step API (include/Simulation.h, Boundary.h, Geometry.h, Collision.h,
// Sampling.h) on an explicit object instead of global state: the same stage names, the same
// error behaviour (invalid CLL coefficients throw std::invalid_argument as
// src/Boundary.cpp:247-250 does; other failures throw std::runtime_error carrying the
// reference's abort code), and RunSimulation with the reference's command line
// (src/Simulation.cpp:21-62) plus options for what the reference sets at compile time.
#pragma once

#include <stdexcept>
#include <string>

#include "../../../include/dsmc_c.h"

namespace dsmc {
namespace host {

// A failure of the C-ABI, carrying its status code (20, 22, 40, ... as the reference aborts).
class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

const char* WallModelName(int wall_model);

class Simulation {
public:
    // Setup + SetupGeometry (src/Initialization.cpp:210-288).
    explicit Simulation(const dsmc_params& params, int device = 0);
    ~Simulation();
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    void InitialConditions();     // src/Initialization.cpp:290-313
    void MoveParticles();         // src/Boundary.cpp:284-328
    void BoundaryConditions();    // src/Boundary.cpp:94-203
    void Indexing();              // src/Geometry.cpp:37-170
    int CollideParticles();       // src/Collision.cpp:45-56
    void Sampling();              // src/Sampling.cpp:14-31
    void Step(int nsteps);        // the loop body of src/Simulation.cpp:138-161, fused
    // Printing (src/Sampling.cpp:36-268): <results>/<wall model>/DSMC2D_Field.{plt,json}.
    std::string Printing(const std::string& results_directory);

    dsmc_state State() const;
    dsmc_constants Constants() const;
    const dsmc_params& Params() const { return params_; }
    dsmc_ctx* Handle() const { return ctx_; }

private:
    void check(int rc) const;
    dsmc_params params_;
    dsmc_ctx* ctx_ = nullptr;
};

// The reference's RunSimulation (src/Simulation.cpp:65-247) on one GPU: parses the command
// line, runs Setup, InitialConditions and the time loop with progress every 100 steps, and
// writes the field. Returns 0 on success or --help, 2 on a command-line error.
int RunSimulation(int argc, char** argv);

}  // namespace host
}  // namespace dsmc
