// ref_harness.cpp - C entry points over the reference's own step functions (test infrastructure).
//
// Built by oracle/Makefile together with the reference sources where they lie
// (src/Boundary.cpp, src/Geometry.cpp, src/Collision.cpp, src/Sampling.cpp,
// src/Initialization.cpp, rng/*.cxx) into oracle/_ref/libdsmc_ref.so. Nothing here is part of
// the product. The harness assigns the reference's globals for a single rank (what
// SetupMPIDomain gives with one process, src/MPIManager.cpp:79-230) and calls the reference's
// functions unchanged; tests compare the oracle's MWC mode against it bit for bit.
#include "Boundary.h"
#include "Collision.h"
#include "Geometry.h"
#include "Initialization.h"
#include "RNG.h"
#include "Sampling.h"
#include "SimulationState.h"

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "../include/dsmc_c.h"

extern "C" {

int ref_configure(const dsmc_params* p)
{
    X1 = p->x1; X2 = p->x2; Y1 = p->y1; Y2 = p->y2; DZ = p->dz;
    NX = p->nx; NY = p->ny; DL = p->dl;
    BoxX = p->box_x; BoxY = p->box_y; BoxWidth = p->box_width; BoxHeight = p->box_height;
    Dt = p->dt; NStep = p->n_step; FirstSamplingStep = p->first_sampling_step;
    SamplingPeriod = p->sampling_period; NPCFree = p->npc_free; MolarMass = p->molar_mass;
    MaFree = p->mach; NFree = p->n_free; TFree = p->t_free; AttackAngle = p->attack_angle_deg;
    Tw = p->t_wall; DRef = p->d_ref; ViscosityIndex = p->viscosity_index; Tref = p->t_ref;
    WallInteraction = p->wall_model == DSMC_WALL_DIFFUSE ? WallModel::Diffuse
                    : p->wall_model == DSMC_WALL_SPECULAR ? WallModel::Specular : WallModel::CLL;
    CLLNormalAccommodation = p->acc_n;
    CLLTangentialAccommodation = p->acc_t;
    if (NX > MAX_CELL_X || NY > MAX_CELL_Y) return 2;
    MPIRank = 0; MPISize = 1;
    ProcX = ProcY = 1; RankX = RankY = 0;
    LeftRank = RightRank = BottomRank = TopRank = MPI_PROC_NULL;
    LocalNX = NX; LocalNY = NY;
    GlobalStartCellX = 0; GlobalEndCellX = NX; GlobalStartCellY = 0; GlobalEndCellY = NY;
    LocalX1 = X1; LocalX2 = X2; LocalY1 = Y1; LocalY2 = Y2;
    Setup();
    DragImpulseTotal = 0.0; DragCollisionCount = 0; DragSamplingTime = 0.0;
    return 0;
}

void ref_constants(double* out)
{
    out[0] = MoleculeMass; out[1] = PFree; out[2] = UFree; out[3] = UxFree; out[4] = UyFree;
    out[5] = DX; out[6] = DY; out[7] = DV; out[8] = Weight; out[9] = SigmaRef; out[10] = Omega;
    out[11] = CrRef; out[12] = MFTFree; out[13] = MFPFree; out[14] = KnFree;
}

void ref_set_seed(unsigned u, unsigned v) { SetSeed(u, v); }
double ref_brng() { return brng(); }
int ref_irand_poisson(double e) { return irand_Poisson(e); }
void ref_isotropic(double* r) { v3rand_isotropic(r); }

void ref_initial_conditions() { InitialConditions(); }
int ref_np() { return NP; }

void ref_get_particles(double* x, double* y, double* vx, double* vy, double* vz)
{
    for (int i = 0; i < NP; ++i) {
        x[i] = P[i].X[0]; y[i] = P[i].X[1]; vx[i] = P[i].V[0]; vy[i] = P[i].V[1]; vz[i] = P[i].V[2];
    }
}

int ref_set_particles(int n, const double* x, const double* y, const double* vx,
                      const double* vy, const double* vz)
{
    if (n > MAX_PCL) return 1;
    NP = n;
    for (int i = 0; i < n; ++i) {
        P[i].X[0] = x[i]; P[i].X[1] = y[i]; P[i].V[0] = vx[i]; P[i].V[1] = vy[i]; P[i].V[2] = vz[i];
    }
    return 0;
}

void ref_set_step(int step)
{
    Step = step;
    SamplingMode = (Step > FirstSamplingStep && Step % SamplingPeriod == 0) ? 1 : 0;
}

int ref_move()
{
    try { MoveParticles(); } catch (const std::invalid_argument&) { return 40; }
    return 0;
}

void ref_boundary() { BoundaryConditions(); }

// Indexing aborts through MPI on bad input; callers give it in-domain particles only.
void ref_indexing(uint32_t* counts, uint32_t* perm)
{
    Indexing();
    int w = 0;
    for (int k = 0; k < NX; ++k)
        for (int l = 0; l < NY; ++l) {
            counts[k * NY + l] = (uint32_t)NPC[k][l];
            for (int j = 0; j < NPC[k][l]; ++j) perm[w++] = (uint32_t)IPC[k][l][j];
        }
}

int ref_collide() { return CollideParticles(); }
void ref_sampling() { Sampling(); }
void ref_clear_cells()
{
    std::memset(C, 0, sizeof C);
    SampleStep = 0; SampleTime = 0.0;
}
int ref_sample_step() { return SampleStep; }

void ref_get_cells(double* out)
{
    for (int k = 0; k < NX; ++k)
        for (int l = 0; l < NY; ++l) {
            const CELL& c = C[k][l];
            double* o = out + (size_t)(k * NY + l) * 17;
            o[0] = c.CountNP;
            for (int m = 0; m < 3; ++m) o[1 + m] = c.CountV[m];
            for (int m = 0; m < 3; ++m)
                for (int q = 0; q < 3; ++q) o[4 + 3 * m + q] = c.CountVV[m][q];
            for (int m = 0; m < 3; ++m) o[13 + m] = c.CountV2V[m];
            o[16] = c.CountNC;
        }
}

// Printing (src/Sampling.cpp:36-268) of a field of CELL sums given as 17 doubles per cell in
// k*NY+l order, with SampleStep = sample_step; the reference writes DSMC2D_Field.plt and .json
// into its compiled-in results directory / wall model name (ref_results_directory()).
void ref_printing(const double* cell17, int sample_step)
{
    std::vector<CELL> field((size_t)NX * NY);
    for (size_t i = 0; i < field.size(); ++i) {
        const double* o = cell17 + i * 17;
        CELL& c = field[i];
        c.CountNP = o[0];
        for (int m = 0; m < 3; ++m) c.CountV[m] = o[1 + m];
        for (int m = 0; m < 3; ++m)
            for (int q = 0; q < 3; ++q) c.CountVV[m][q] = o[4 + 3 * m + q];
        for (int m = 0; m < 3; ++m) c.CountV2V[m] = o[13 + m];
        c.CountNC = o[16];
    }
    SampleStep = sample_step;
    Printing(field);
}

const char* ref_results_directory() { return DSMC_RESULTS_DIRECTORY; }

void ref_drag(double* impulse, long long* hits)
{
    *impulse = DragImpulseTotal;
    *hits = DragCollisionCount;
}

// One iteration of the loop of src/Simulation.cpp:138-161 at one rank (no exchange).
int ref_step()
{
    SamplingMode = (Step > FirstSamplingStep && Step % SamplingPeriod == 0) ? 1 : 0;
    if (ref_move()) return 40;
    if (Step > FirstSamplingStep) DragSamplingTime += Dt;
    BoundaryConditions();
    Indexing();
    int nc = CollideParticles();
    if (SamplingMode == 1) Sampling();
    Step++;
    Time += Dt;
    return nc;
}

int ref_intersect_box(double bx, double by, double bw, double bh, double x, double y, double dx,
                      double dy, double* t, double* nx, double* ny)
{
    BoxX = bx; BoxY = by; BoxWidth = bw; BoxHeight = bh;
    return IntersectBox(x, y, dx, dy, *t, *nx, *ny) ? 1 : 0;
}

int ref_inside_box(double bx, double by, double bw, double bh, double x, double y)
{
    BoxX = bx; BoxY = by; BoxWidth = bw; BoxHeight = bh;
    return InsideBox(x, y) ? 1 : 0;
}

void ref_diffuse(double* v, double m, double tw, double nx, double ny)
{
    DiffuseScattering(v, m, tw, nx, ny);
}

int ref_cll(double* v, double m, double tw, double nx, double ny, double an, double at)
{
    try { CLLScattering(v, m, tw, nx, ny, an, at); } catch (const std::invalid_argument&) { return 40; }
    return 0;
}

// GenerateNewParticles on a rectangle with the freestream state; returns the number added.
int ref_generate(double x1, double y1, double x2, double y2, int move)
{
    int before = NP;
    GenerateNewParticles(x1, y1, x2, y2, NFree, UxFree, UyFree, TFree, move);
    return NP - before;
}

}  // extern "C"
