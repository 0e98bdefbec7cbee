/*
 * dsmc_c.h - C-ABI of the MI355X (gfx950) DSMC per-step particle pipeline.
 *
 * The reference solver exposes its step as free C++ functions over global state:
 *   include/Boundary.h:11-21   GenerateNewParticles, BoundaryConditions, *Scattering, MoveParticles
 *   include/Geometry.h:14-22   SetupGeometry, Indexing, InsideBox, IntersectBox
 *   include/Collision.h:4-6    ElasticCollision, BinaryCollisionsInCell, CollideParticles
 *   include/Sampling.h:7-8     Sampling, Printing
 *   include/Simulation.h:4     RunSimulation
 * This header keeps those step names one-to-one, but every call takes an explicit context
 * that owns the device buffers of one or more independent cases (a "batch").
 *
 * Conventions
 *   - Ownership: the context owns every device buffer; the caller owns host arrays.
 *   - Errors: 0 on success, otherwise a DSMC_E_* code. The numeric codes follow the
 *     reference's abort codes (src/Geometry.cpp:82,157; src/Initialization.cpp:218;
 *     src/Boundary.cpp:73-80). dsmc_last_error() returns a message for the last failure.
 *   - Threading: one host thread per context. Calls are stream-ordered on the context's
 *     HIP stream; calls that return host data synchronise.
 *   - Cells are addressed k*NY+l (k along x), as the reference's field gather
 *     (src/MPIManager.cpp:487-489) and writer (src/Sampling.cpp:114-222) order them.
 *   - No PyTorch or C++ types cross this boundary: plain pointers, sizes and PODs.
 */
#ifndef DSMC_C_H
#define DSMC_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Wall models, numbered as the reference's enum class WallModel { Diffuse, Specular, CLL }
 * (include/Boundary.h:3). */
enum {
    DSMC_WALL_DIFFUSE = 0,
    DSMC_WALL_SPECULAR = 1,
    DSMC_WALL_CLL = 2
};

/* Status codes. */
enum {
    DSMC_OK = 0,
    DSMC_E_CAPACITY = 1,        /* particle pool or candidate capacity exceeded (Boundary.cpp:73-80) */
    DSMC_E_ARGUMENT = 2,        /* invalid argument or call order */
    DSMC_E_HIP = 3,             /* HIP runtime error */
    DSMC_E_OUTSIDE_DOMAIN = 20, /* particle outside the domain at indexing (Geometry.cpp:66-86) */
    DSMC_E_CELL_CAPACITY = 22,  /* more than max_per_cell particles in a cell (Geometry.cpp:142-161) */
    DSMC_E_CLL_COEFF = 40,      /* CLL accommodation outside [0,1] or NaN (Initialization.cpp:211-220) */
    DSMC_E_IO = 50              /* field writer could not write its files */
};

/* Parameters of one case. Defaults (dsmc_params_default) are the reference's compile-time
 * globals: src/Geometry.cpp:11-24, src/Initialization.cpp:11-31, src/Boundary.cpp:10-13. */
typedef struct dsmc_params {
    /* domain and grid */
    double x1, x2, y1, y2;     /* domain bounds [m] */
    double dz;                 /* cell thickness [m] */
    int32_t nx, ny;            /* global grid */
    double dl;                 /* injection strip width [m] */
    /* axis-aligned box obstacle */
    double box_x, box_y, box_width, box_height;
    /* time and sampling */
    double dt;                 /* time step [s] */
    int32_t n_step;            /* NStep, used by dsmc_run */
    int32_t first_sampling_step;
    int32_t sampling_period;
    /* gas and freestream */
    double npc_free;           /* target freestream particles per cell */
    double molar_mass;         /* [kg/mol] */
    double mach;
    double n_free;             /* [m^-3] */
    double t_free;             /* [K] */
    double attack_angle_deg;   /* positive tilts the flow toward -y */
    double t_wall;             /* [K] */
    double d_ref;              /* VHS reference diameter [m] */
    double viscosity_index;
    double t_ref;              /* [K] */
    /* wall interaction */
    int32_t wall_model;        /* DSMC_WALL_* */
    double acc_n, acc_t;       /* CLL accommodation, SPARTA acc_n / acc_t convention */
    /* random numbers: Philox4x32-10 key = (seed, case_id) */
    uint64_t seed;
    uint32_t case_id;
    /* capacities */
    uint32_t max_per_cell;     /* 0 selects 1000, as MAX_PCL_IN_CELL (include/Constants.h:3) */
    double capacity_factor;    /* particle pool capacity / initial freestream count; 0 selects 2.0 */
} dsmc_params;

/* Derived constants, as Setup() computes them (src/Initialization.cpp:221-266). */
typedef struct dsmc_constants {
    double molecule_mass, p_free, u_free, ux_free, uy_free;
    double dx, dy, dv, weight;
    double sigma_ref, omega, cr_ref, sc_max;
    double mft_free, mfp_free, kn_free;
} dsmc_constants;

/* Host-visible counters of one case. */
typedef struct dsmc_state {
    int64_t step;               /* Step */
    int64_t np;                 /* particles in the pool */
    int64_t sample_step;        /* SampleStep */
    int64_t last_collisions;    /* collisions of the last collide call (NC) */
    int64_t total_collisions;
    int64_t last_injected;      /* injected particles that survived the last boundary call */
    int64_t last_removed;       /* particles removed by the last boundary call */
    int64_t wall_hits;          /* DragCollisionCount (src/Boundary.cpp:319) */
    double drag_impulse_total;  /* DragImpulseTotal (src/Boundary.cpp:318) */
    double drag_sampling_time;  /* DragSamplingTime (src/Simulation.cpp:144-146) */
    double time;                /* Time */
    double sample_time;         /* SampleTime */
    int64_t particle_steps;     /* sum over completed steps of the particles indexed that step */
    /* The drag report of src/Simulation.cpp:210-237 (commented out in the reference):
     * MeanDrag = DragImpulseTotal / DragSamplingTime and
     * Cd = MeanDrag / (1/2 NFree m UFree^2 BoxHeight DZ); both 0 before any sampling time. */
    double mean_drag;           /* [N] */
    double drag_coefficient;    /* Cd */
} dsmc_state;

/* Stages of one fused step, in launch order (for dsmc_profile_steps). */
enum {
    DSMC_STAGE_BEGIN = 0,   /* step opening: closes the previous fused step, inflow counts */
    DSMC_STAGE_MOVE = 1,    /* free flight + outflow + cell key + histogram */
    DSMC_STAGE_EXTRA = 2,   /* near-box particles (exact box test, wall model) + inflow */
    DSMC_STAGE_SCAN = 3,    /* cell offsets */
    DSMC_STAGE_SCATTER = 4, /* arrival-order scatter into cell segments */
    DSMC_STAGE_GATHER = 5,  /* stable in-segment order, particle gather, NTC draw, sampling */
    DSMC_STAGE_COLLIDE = 6, /* large schedule: cells with collision candidates: collide, sample */
    DSMC_NUM_STAGES = 7
};

/* Launch schedules of the fused step. Both give bit-identical results. */
enum {
    DSMC_SCHEDULE_LARGE = 0, /* open, move, extra, scan, scatter, gather, collision list */
    DSMC_SCHEDULE_SMALL = 1  /* open, move, scan+scatter, gather with collisions; every block
                                holds the whole cell scan in LDS, so at most 12288 cells in
                                all cases together */
};

/* Launch parameters that change speed, never results. */
typedef struct dsmc_tuning {
    int32_t schedule; /* DSMC_SCHEDULE_* */
    int32_t chunk;    /* particles per gather work item (a workgroup's cells: the cells that
                         start in one chunk, at most 4 * chunk + 3): a power of two in
                         [128, 2048] */
    int32_t particles_per_thread; /* of the pool-streaming kernels of the large schedule (the
                         move, keys, fill): 1, 2 or 4 (0 means 1) */
} dsmc_tuning;

typedef struct dsmc_ctx dsmc_ctx;

void dsmc_params_default(dsmc_params* p);

/* Setup + SetupGeometry (src/Initialization.cpp:210-288, src/Geometry.cpp:30-35): validates the
 * parameters (CLL coefficients -> DSMC_E_CLL_COEFF), derives constants, allocates device state. */
int dsmc_create(const dsmc_params* p, int device, dsmc_ctx** out);
/* The same for n independent cases that share one particle pool (case-offset table). */
int dsmc_create_batch(const dsmc_params* cases, int ncases, int device, dsmc_ctx** out);
void dsmc_destroy(dsmc_ctx* ctx);
const char* dsmc_last_error(const dsmc_ctx* ctx);
int dsmc_num_cases(const dsmc_ctx* ctx);

int dsmc_get_params(dsmc_ctx* ctx, int case_index, dsmc_params* out);
int dsmc_get_constants(dsmc_ctx* ctx, int case_index, dsmc_constants* out);
int dsmc_get_state(dsmc_ctx* ctx, int case_index, dsmc_state* out);

/* InitialConditions (src/Initialization.cpp:290-313): freestream fill of every case, then
 * RemoveParticlesInsideBox (src/Geometry.cpp:231-247). Resets step and sampling counters. */
int dsmc_init_freestream(dsmc_ctx* ctx);
/* Warm start: seed each case from per-cell local Maxwellians given as 4 planes of w*h floats
 * (n [m^-3], ux, uy [m/s], T [K]) per case, resampled onto the case's grid. */
int dsmc_init_from_fields(dsmc_ctx* ctx, const float* n_ux_uy_t, int w, int h);
/* The same from fields already on each case grid: 4 doubles (n, ux, uy, T) per cell, k*NY+l
 * order within a case, cases concatenated. */
int dsmc_init_from_cell_fields(dsmc_ctx* ctx, const double* n_ux_uy_t);
/* The same with a temperature per velocity component: 6 doubles (n, ux, uy, Tx, Ty, Tz) per
 * cell, each component drawn from its own local Maxwellian. With Tx = Ty = Tz = T it seeds
 * exactly what dsmc_init_from_cell_fields seeds. No counterpart in the reference (which always
 * starts from the freestream fill): a warm start that keeps the local temperature anisotropy
 * of a sampled field. */
int dsmc_init_from_cell_moments(dsmc_ctx* ctx, const double* n_ux_uy_tx_ty_tz);

/* One stage of the step, in the order of src/Simulation.cpp:143-159. */
int dsmc_move(dsmc_ctx* ctx);                       /* MoveParticles (src/Boundary.cpp:284-328) */
int dsmc_boundary(dsmc_ctx* ctx);                   /* BoundaryConditions (src/Boundary.cpp:94-203) */
int dsmc_index(dsmc_ctx* ctx);                      /* Indexing (src/Geometry.cpp:37-170) */
int dsmc_collide(dsmc_ctx* ctx, int64_t* ncoll);    /* CollideParticles (src/Collision.cpp:45-56) */
int dsmc_sample(dsmc_ctx* ctx);                     /* Sampling (src/Sampling.cpp:14-31) */
/* nsteps of the loop of src/Simulation.cpp:138-199 (SamplingMode, move, boundary, index,
 * collide, sample, Step++), fused and graph-captured. */
int dsmc_step(dsmc_ctx* ctx, int nsteps);
/* Retires one case of a batch: the next step still moves its particles by free flight (no
 * wall interaction) and then drops all of them in that step's outflow test; it gets no inflow
 * and nothing more is added to its CELL sums. Its sums, sampled step count (sample_step), time,
 * sample_time, wall hits and drag keep their values. dsmc_state.step is the batch's shared step
 * counter and keeps advancing for a retired case. The other cases' results do not change. A
 * retired case stays retired until the next dsmc_init_*; the single-stage calls reject it. The
 * reference runs one case per process and has no counterpart: this is for batched sweeps that
 * stop each case once it is steady, so the pool shrinks to the cases still running. */
int dsmc_retire_case(dsmc_ctx* ctx, int case_index);
/* Fed random draws, for parity runs against the reference on identical draws (its one
 * sequential stream: rng/BRNG.cxx:30-34, consumed by src/Boundary.cpp:49-57 and the transforms
 * of rng/PoissonRNG.cxx, RayleighRNG.cxx, DiscreteUniformRNG.cxx and IsotropicVectorRNG.cxx;
 * src/Collision.cpp:27-38). While set, the single-stage dsmc_move, dsmc_boundary
 * and dsmc_collide of a one-case context take their uniforms from u[0..n), each in [0, 1), in
 * the reference's order (particles in pool order; strips left, right, bottom, top with each
 * strip's count draw first; cells in k*NY+l order), on one GPU thread, instead of the
 * counter-based streams. The position carries over from call to call; a call that runs past
 * the end fails with DSMC_E_CAPACITY. n = 0 returns to the counter-based streams. The fused
 * dsmc_step always uses the counter-based streams. */
int dsmc_set_draws(dsmc_ctx* ctx, const double* u, size_t n);
/* The number of fed draws consumed since dsmc_set_draws (0 without fed draws). */
int dsmc_draws_used(dsmc_ctx* ctx, size_t* used);
/* Sets Step (and the derived SamplingMode rule); used to start mid-run in tests. */
int dsmc_set_step(dsmc_ctx* ctx, int64_t step);
/* Runs nsteps fused steps eagerly with HIP events around every stage (the reference has no
 * timers); stage_ms[DSMC_NUM_STAGES] receives the mean milliseconds per step of each stage. */
int dsmc_profile_steps(dsmc_ctx* ctx, int nsteps, double* stage_ms);
/* Launch tuning (the reference runs on CPU ranks and has none): the current choice, a new one
 * (DSMC_E_ARGUMENT if out of range), and an autotuner that times every candidate schedule and
 * chunk (at one particle per thread), then the particles per thread of the best large-schedule
 * candidate, over `nsteps` fused steps each, keeps the fastest and returns it in *best. The
 * autotuner advances the simulation by the steps it times; results do not depend on tuning. */
int dsmc_get_tuning(dsmc_ctx* ctx, dsmc_tuning* out);
int dsmc_set_tuning(dsmc_ctx* ctx, const dsmc_tuning* t);
int dsmc_autotune(dsmc_ctx* ctx, int nsteps, dsmc_tuning* best);

/* Parity and output. The particle state P[] / NP (include/Particle.h:6-9,
 * src/Initialization.cpp:81-82) as five arrays in pool order: dsmc_get_particles copies case
 * case_index's particles out (any array may be NULL; cap is their length, *n the count) and
 * dsmc_set_particles replaces a one-case pool. The pool keeps (x, y) and (vx, vy) as pairs on
 * the device; both calls split or pair them there, bit for bit. */
int dsmc_get_num_particles(dsmc_ctx* ctx, int case_index, size_t* n);
int dsmc_get_particles(dsmc_ctx* ctx, int case_index, double* x, double* y,
                       float* vx, float* vy, float* vz, size_t cap, size_t* n);
int dsmc_set_particles(dsmc_ctx* ctx, int case_index, const double* x, const double* y,
                       const float* vx, const float* vy, const float* vz, size_t n);
/* Per-cell counts (NPC) and the stable permutation (IPC lists concatenated in cell order)
 * of the last dsmc_index; perm[i] is the pre-index pool position of the particle at sorted
 * position i. */
int dsmc_get_index(dsmc_ctx* ctx, int case_index, uint32_t* counts, size_t ncells,
                   uint32_t* perm, size_t n);
/* CELL sums (include/Cell.h:4-10), 17 doubles per cell: CountNP, CountV[3], CountVV[3][3],
 * CountV2V[3], CountNC. */
int dsmc_get_cells(dsmc_ctx* ctx, int case_index, double* cell17, size_t ncells);
/* The CELL sums of case `case_index` added over bx x by blocks of cells (the reference has no
 * counterpart; it serves field comparisons across grids and sampling windows): block (i, j) holds
 * the cells (k, l) with floor((k + 0.5) bx / NX) = i and floor((l + 0.5) by / NY) = j. cell17
 * receives bx*by blocks of 17 doubles in i*by+j order. Differences of two calls are the sums of
 * the steps sampled in between. */
int dsmc_get_binned_cells(dsmc_ctx* ctx, int case_index, int bx, int by, double* cell17);
/* dsmc_get_binned_cells without any allocation: no counterpart in the reference. It is for
 * callers that cannot allocate (code under HIP graph capture, tight serving or sweep loops that
 * poll a running batch, harnesses that forbid hipMalloc). The block sums are staged in a fixed
 * device array of 16384 blocks x 17 doubles (2.2 MB) that each GPU holds until the process
 * ends; bx * by above 16384 is rejected with DSMC_E_ARGUMENT. All calls on one GPU share that
 * array: calls that overlap on one GPU (from other threads or contexts) race. The results equal
 * dsmc_get_binned_cells' bit for bit. */
int dsmc_get_binned_cells_noalloc(dsmc_ctx* ctx, int case_index, int bx, int by, double* cell17);
/* Printing's moment-to-field formulas (src/Sampling.cpp:152-196), 6 doubles per cell:
 * N, Ux, Uy, Tx, Ty, T. */
int dsmc_get_fields(dsmc_ctx* ctx, int case_index, double* out6, size_t ncells);
/* Printing (src/Sampling.cpp:36-268): DSMC2D_Field.plt and DSMC2D_Field.json in directory. */
int dsmc_write_field(dsmc_ctx* ctx, int case_index, const char* directory);

#ifdef __cplusplus
}
#endif

#endif /* DSMC_C_H */
