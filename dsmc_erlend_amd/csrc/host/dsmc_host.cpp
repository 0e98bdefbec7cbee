// dsmc_host.cpp - C++ host side of the gfx950 DSMC pipeline (see dsmc_host.hpp).
#include "dsmc_host.hpp"

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <string>

namespace dsmc {
namespace host {

const char* WallModelName(int wall_model)
{
    switch (wall_model) {
    case DSMC_WALL_SPECULAR: return "specular";
    case DSMC_WALL_DIFFUSE: return "diffuse";
    case DSMC_WALL_CLL: return "cll";
    }
    throw std::invalid_argument("Unknown wall model");
}

Simulation::Simulation(const dsmc_params& params, int device) : params_(params)
{
    const int rc = dsmc_create(&params_, device, &ctx_);
    if (rc == DSMC_E_CLL_COEFF) {
        dsmc_destroy(ctx_);
        ctx_ = nullptr;
        throw std::invalid_argument("CLL accommodation coefficients must be in [0, 1]");
    }
    if (rc) {
        const std::string msg = ctx_ ? dsmc_last_error(ctx_) : "dsmc_create failed";
        dsmc_destroy(ctx_);
        ctx_ = nullptr;
        throw Error(rc, msg);
    }
}

Simulation::~Simulation() { dsmc_destroy(ctx_); }

void Simulation::check(int rc) const
{
    if (rc) throw Error(rc, dsmc_last_error(ctx_));
}

void Simulation::InitialConditions() { check(dsmc_init_freestream(ctx_)); }
void Simulation::MoveParticles() { check(dsmc_move(ctx_)); }
void Simulation::BoundaryConditions() { check(dsmc_boundary(ctx_)); }
void Simulation::Indexing() { check(dsmc_index(ctx_)); }

int Simulation::CollideParticles()
{
    int64_t nc = 0;
    check(dsmc_collide(ctx_, &nc));
    return (int)nc;
}

void Simulation::Sampling() { check(dsmc_sample(ctx_)); }
void Simulation::Step(int nsteps) { check(dsmc_step(ctx_, nsteps)); }

std::string Simulation::Printing(const std::string& results_directory)
{
    const std::filesystem::path dir =
        std::filesystem::path(results_directory) / WallModelName(params_.wall_model);
    check(dsmc_write_field(ctx_, 0, dir.string().c_str()));
    return (dir / "DSMC2D_Field.plt").string();
}

dsmc_state Simulation::State() const
{
    dsmc_state s;
    check(dsmc_get_state(ctx_, 0, &s));
    return s;
}

dsmc_constants Simulation::Constants() const
{
    dsmc_constants k;
    check(dsmc_get_constants(ctx_, 0, &k));
    return k;
}

namespace {

enum class Parse { Run, Help, Error };

bool parse_number(const char* text, double& out)
{
    char* end = nullptr;
    errno = 0;
    out = std::strtod(text, &end);
    return end != text && *end == '\0' && errno != ERANGE && std::isfinite(out);
}

// A decimal integer in [lo, hi]: digits only (every integer option is non-negative, so a sign,
// a fraction, an exponent or trailing characters are errors).
bool parse_integer(const char* text, unsigned long long lo, unsigned long long hi,
                   unsigned long long& out)
{
    if (!text || !std::isdigit((unsigned char)text[0])) return false;
    char* end = nullptr;
    errno = 0;
    out = std::strtoull(text, &end, 10);
    return *end == '\0' && errno != ERANGE && out >= lo && out <= hi;
}

void usage()
{
    std::cout << "Usage: dsmc_gpu [--angle-of-attack DEGREES] [options]\n"
              << "  --angle-of-attack, --aoa  Flow angle in [-180, 180] degrees\n"
              << "  Positive angles tilt the flow toward negative y.\n"
              << "  The box remains aligned with the x/y axes.\n"
              << "  --wall specular|diffuse|cll   wall interaction (default cll)\n"
              << "  --acc-n X --acc-t X           CLL accommodation (SPARTA acc_n, acc_t)\n"
              << "  --nx N --ny N                 grid (default 100 x 100)\n"
              << "  --npc X                       freestream particles per cell (default 10)\n"
              << "  --steps N --first-sampling N  NStep, FirstSamplingStep\n"
              << "  --mach X --density X --temperature X --wall-temperature X\n"
              << "  --seed N --device N --results DIR\n";
}

// The reference's ParseArguments (src/Simulation.cpp:21-62), extended with the settings the
// reference fixes at compile time.
Parse parse(int argc, char** argv, dsmc_params& p, std::string& results, int& device)
{
    for (int i = 1; i < argc; ++i) {
        const std::string opt = argv[i];
        if (opt == "--help" || opt == "-h") {
            usage();
            return Parse::Help;
        }
        const bool known = opt == "--angle-of-attack" || opt == "--aoa" || opt == "--wall"
            || opt == "--acc-n" || opt == "--acc-t" || opt == "--nx" || opt == "--ny"
            || opt == "--npc" || opt == "--steps" || opt == "--first-sampling"
            || opt == "--mach" || opt == "--density" || opt == "--temperature"
            || opt == "--wall-temperature" || opt == "--seed" || opt == "--device"
            || opt == "--results";
        if (!known) {
            std::cerr << "Unknown option: " << opt << ". Use --help.\n";
            return Parse::Error;
        }
        if (++i == argc) {
            std::cerr << opt << " requires a value.\n";
            return Parse::Error;
        }
        const char* val = argv[i];
        if (opt == "--wall") {
            const std::string w = val;
            if (w == "specular") p.wall_model = DSMC_WALL_SPECULAR;
            else if (w == "diffuse") p.wall_model = DSMC_WALL_DIFFUSE;
            else if (w == "cll") p.wall_model = DSMC_WALL_CLL;
            else {
                std::cerr << "Invalid wall model: " << w << '\n';
                return Parse::Error;
            }
            continue;
        }
        if (opt == "--results") {
            results = val;
            continue;
        }
        // integer options: NX, NY, NStep, FirstSamplingStep, the seed and the device
        struct IntOpt {
            const char* name;
            unsigned long long lo, hi;
        };
        static const IntOpt ints[] = {
            {"--nx", 1, 65535}, {"--ny", 1, 65535}, {"--steps", 0, 2147483647ull},
            {"--first-sampling", 0, 2147483647ull}, {"--seed", 0, ~0ull}, {"--device", 0, 1023}};
        bool handled = false;
        for (const IntOpt& io : ints) {
            if (opt != io.name) continue;
            unsigned long long v = 0;
            if (!parse_integer(val, io.lo, io.hi, v)) {
                std::cerr << "Invalid value for " << opt << ": " << val << ". Expected an integer in ["
                          << io.lo << ", " << io.hi << "].\n";
                return Parse::Error;
            }
            if (opt == "--nx") p.nx = (int32_t)v;
            else if (opt == "--ny") p.ny = (int32_t)v;
            else if (opt == "--steps") p.n_step = (int32_t)v;
            else if (opt == "--first-sampling") p.first_sampling_step = (int32_t)v;
            else if (opt == "--seed") p.seed = (uint64_t)v;
            else device = (int)v;
            handled = true;
        }
        if (handled) continue;
        double x;
        if (!parse_number(val, x)) {
            std::cerr << "Invalid value for " << opt << ": " << val << '\n';
            return Parse::Error;
        }
        if (opt == "--angle-of-attack" || opt == "--aoa") {
            if (x < -180.0 || x > 180.0) {
                std::cerr << "Invalid angle: " << val
                          << ". Expected a finite number in [-180, 180].\n";
                return Parse::Error;
            }
            p.attack_angle_deg = x;
        } else if (opt == "--acc-n") p.acc_n = x;
        else if (opt == "--acc-t") p.acc_t = x;
        else if (opt == "--npc") p.npc_free = x;
        else if (opt == "--mach") p.mach = x;
        else if (opt == "--density") p.n_free = x;
        else if (opt == "--temperature") p.t_free = x;
        else if (opt == "--wall-temperature") p.t_wall = x;
    }
    return Parse::Run;
}

}  // namespace

int RunSimulation(int argc, char** argv)
{
    dsmc_params p;
    dsmc_params_default(&p);
    std::string results = "results";
    int device = 0;
    const Parse a = parse(argc, argv, p, results, device);
    if (a != Parse::Run) return a == Parse::Help ? 0 : 2;
    try {
        Simulation sim(p, device);
        const dsmc_constants k = sim.Constants();
        std::cout << "\nFreestream properties\n-----------------------------\n"
                  << "Box size    = " << p.box_width << " x " << p.box_height << " m\n"
                  << "Wall model  = " << WallModelName(p.wall_model) << '\n';
        if (p.wall_model == DSMC_WALL_CLL)
            std::cout << "CLL normal  = " << p.acc_n << '\n' << "CLL tangent = " << p.acc_t << '\n';
        std::cout << "Angle       = " << p.attack_angle_deg << " degrees\n"
                  << "P_inf       = " << k.p_free << " Pa\n"
                  << "U_inf       = " << k.u_free << " m/s\n"
                  << "MFT_inf     = " << k.mft_free << " s\n"
                  << "MFP_inf     = " << k.mfp_free << " m\n"
                  << "Kn_inf      = " << k.kn_free << "\n"
                  << "DX / lambda = " << k.dx / k.mfp_free << "\n"
                  << "DY / lambda = " << k.dy / k.mfp_free << "\n"
                  << "Dt / MFT    = " << p.dt / k.mft_free << "\n";
        sim.InitialConditions();
        std::cout << "Total particles = " << sim.State().np << "\nInitialConditions\n";
        int done = 0;
        while (done < p.n_step) {
            const int n = std::min(100, p.n_step - done);
            sim.Step(n);
            done += n;
            if (done % 100 == 0) {
                const dsmc_state s = sim.State();
                std::cout << "Time step " << done << ", Number of particles " << s.np
                          << ", Number of collisions " << s.last_collisions << '\n';
            }
        }
        std::cout << "Global field written to " << sim.Printing(results) << '\n';
        // the drag report the reference keeps commented out (src/Simulation.cpp:210-237)
        const dsmc_state s = sim.State();
        std::cout << "\n====================================\n"
                  << "Box collisions = " << s.wall_hits << "\n"
                  << "Sampling time = " << s.drag_sampling_time << " s\n"
                  << "Mean drag force = " << s.mean_drag << " N\n"
                  << "Drag coefficient Cd = " << s.drag_coefficient << "\n"
                  << "====================================\n";
    } catch (const std::invalid_argument& e) {
        std::cerr << "ERROR: " << e.what() << '\n';
        return 40;
    } catch (const Error& e) {
        std::cerr << "ERROR: " << e.what() << '\n';
        return e.code();
    }
    return 0;
}

}  // namespace host
}  // namespace dsmc
