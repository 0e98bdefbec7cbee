/*
 * dsmc_oracle.h - CPU oracle for the DSMC per-step particle pipeline.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load this library, and only as the checker (or the timed CPU baseline). The product
 * (dsmc_erlend_amd, include/dsmc_c.h) never links or calls it.
 *
 * A plain-C restatement of the reference's step functions (cited per function in
 * dsmc_oracle.c). Every random draw goes through an injectable draw source:
 *   - OR_DS_MWC    the reference's sequential multiply-with-carry stream (rng/BRNG.cxx:15-34);
 *   - OR_DS_PHILOX a Philox4x32-10 stream keyed exactly as the GPU kernels key theirs;
 *   - OR_DS_ARRAY  a caller-supplied array of uniforms ("fed draws").
 * A context runs in one of two modes:
 *   - OR_MODE_REF  the reference's semantics: one global MWC stream, swap-with-last removal
 *                  (src/Boundary.cpp:181-202), fixed SCMax (src/Collision.cpp:23), FP64 state;
 *   - OR_MODE_GPU  the GPU pipeline's semantics: Philox streams keyed by (slot|cell|candidate,
 *                  step, purpose), stable compaction on removal, stable physical sort by cell at
 *                  indexing, per-cell (sigma*g)max raised on demand, velocities rounded to FP32
 *                  at every store, FP64 positions.
 */
#ifndef DSMC_ORACLE_H
#define DSMC_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/dsmc_c.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_MODE_REF = 0, OR_MODE_GPU = 1 };
enum { OR_DS_MWC = 0, OR_DS_PHILOX = 1, OR_DS_ARRAY = 2 };

/* Philox counter word 2 ("purpose") of each stream the GPU pipeline draws from. */
enum {
    OR_P_MOVE = 1,        /* c0 = pool slot,            c1 = step */
    OR_P_INJ_COUNT = 2,   /* c0 = strip (0..3),          c1 = step */
    OR_P_INJECT = 3,      /* c0 = candidate in strip,    c1 = step, c2 |= strip << 8 */
    OR_P_COLLIDE = 4,     /* c0 = cell (k*NY+l),         c1 = step */
    OR_P_INIT_COUNT = 5,  /* c0 = 0,                     c1 = 0 */
    OR_P_INIT = 6,        /* c0 = candidate,             c1 = 0 */
    OR_P_WARM = 7         /* c0 = cell,                  c1 = 0 (warm start seeding) */
};

typedef struct or_draws or_draws;
typedef struct or_ctx or_ctx;

/* ---- draw sources ---- */
or_draws* or_draws_new_mwc(uint32_t u, uint32_t v);
or_draws* or_draws_new_philox(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2);
or_draws* or_draws_new_array(const double* values, size_t n);
void or_draws_free(or_draws* d);
double or_draws_next(or_draws* d);
size_t or_draws_used(const or_draws* d);
/* Order in which frand_Gaussian consumes its two draws under the MWC source: 0 = the log()
 * draw first, 1 = the cos() draw first. The reference leaves it unspecified
 * (src/Boundary.cpp:50); the build recipe pins it against the compiled reference. */
void or_set_mwc_gaussian_order(int cos_first);

/* ---- RNG primitives ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint32_t or_mwc_next(uint32_t* u, uint32_t* v);
int or_irand_uniform(int n, or_draws* d);
int or_irand_poisson(double mean, or_draws* d);
double or_frand_rayleigh(double s, or_draws* d);
double or_frand_gaussian(double mean, double variance, or_draws* d);
void or_v3rand_isotropic(double* r, or_draws* d);

/* ---- geometry and wall models (box given explicitly) ---- */
int or_inside_box(double bx, double by, double bw, double bh, double x, double y);
int or_intersect_box(double bx, double by, double bw, double bh, double x, double y,
                     double dx, double dy, double* t_hit, double* nx, double* ny);
void or_specular(double* v, double nx, double ny);
void or_diffuse(double* v, double m, double tw, double nx, double ny, or_draws* d);
int or_cll(double* v, double m, double tw, double nx, double ny, double an, double at,
           or_draws* d);

/* ---- setup ---- */
void or_params_default(dsmc_params* p);
int or_constants_of(const dsmc_params* p, dsmc_constants* out);

/* ---- one case ---- */
or_ctx* or_create(const dsmc_params* p, int mode, int* status);
void or_destroy(or_ctx* c);
void or_set_mwc_seed(or_ctx* c, uint32_t u, uint32_t v);
void or_get_mwc_state(const or_ctx* c, uint32_t* u, uint32_t* v);
int or_init_freestream(or_ctx* c);
int or_set_particles(or_ctx* c, size_t n, const double* x, const double* y,
                     const double* vx, const double* vy, const double* vz);
size_t or_num_particles(const or_ctx* c);
void or_get_particles(const or_ctx* c, double* x, double* y, double* vx, double* vy, double* vz);
void or_set_step(or_ctx* c, int64_t step);
int or_move(or_ctx* c);
int or_boundary(or_ctx* c);
void or_last_strip_counts(const or_ctx* c, int64_t counts[4]);
int or_index(or_ctx* c);
void or_get_index(const or_ctx* c, uint32_t* counts, uint32_t* perm);
int or_collide(or_ctx* c, int64_t* ncoll);
int or_sample(or_ctx* c);
int or_step(or_ctx* c, int nsteps);
void or_get_cells(const or_ctx* c, double* cell17);
void or_get_fields(const or_ctx* c, double* out6);
void or_get_state(const or_ctx* c, dsmc_state* s);
/* Seed particles from per-cell local Maxwellians on the case grid (n, ux, uy, T per cell,
 * k*NY+l order), GPU mode only: the restatement of the warm-start seeding kernel. */
int or_init_from_cell_fields(or_ctx* c, const double* n_ux_uy_t);

#ifdef __cplusplus
}
#endif

#endif
