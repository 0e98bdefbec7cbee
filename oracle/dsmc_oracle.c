/*
 * dsmc_oracle.c - CPU oracle for the DSMC per-step particle pipeline (test infrastructure).
 *
 * Plain C restatement of the reference's step functions; each function cites the reference
 * file:line it follows. Nothing here is linked into the product. Compile with
 * -ffp-contract=off so every a*b+c is a rounded multiply then a rounded add, as the
 * reference's x86-64 build (no FMA) evaluates it.
 */
#include "dsmc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Old CODATA values the reference keeps (include/Constants.h:9-10). */
#define OR_KB 1.380662E-23
#define OR_NA 6.022045E+23
/* The reference's MWC scale to [0,1) (rng/BRNG.cxx:33). */
#define OR_MWC_SCALE 2.328306435996595e-10
#define OR_TWO_POW_M32 2.3283064365386962890625e-10

/* ======================================================================================
 * Draw sources
 * ====================================================================================== */
struct or_draws {
    int kind;
    uint32_t mu, mv;                 /* MWC state */
    uint32_t key[2], ctr[4], buf[4]; /* Philox stream */
    int pos;
    const double* arr;               /* fed draws */
    size_t n;
    size_t used;
};

static int g_mwc_cos_first = 0;
void or_set_mwc_gaussian_order(int cos_first) { g_mwc_cos_first = cos_first ? 1 : 0; }

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of two 32x32->64 multiplies
 * with the published multipliers and Weyl key increments. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Marsaglia MWC pair (rng/BRNG.cxx:23-28). */
uint32_t or_mwc_next(uint32_t* u, uint32_t* v)
{
    *v = 36969u * (*v & 65535u) + (*v >> 16);
    *u = 18000u * (*u & 65535u) + (*u >> 16);
    return (*v << 16) + *u;
}

or_draws* or_draws_new_mwc(uint32_t u, uint32_t v)
{
    or_draws* d = (or_draws*)calloc(1, sizeof *d);
    d->kind = OR_DS_MWC; d->mu = u; d->mv = v;
    return d;
}

static void philox_init(or_draws* d, uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1,
                        uint32_t c2)
{
    memset(d, 0, sizeof *d);
    d->kind = OR_DS_PHILOX;
    d->key[0] = k0; d->key[1] = k1;
    d->ctr[0] = c0; d->ctr[1] = c1; d->ctr[2] = c2; d->ctr[3] = 0;
    d->pos = 4;
}

or_draws* or_draws_new_philox(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2)
{
    or_draws* d = (or_draws*)malloc(sizeof *d);
    philox_init(d, k0, k1, c0, c1, c2);
    return d;
}

or_draws* or_draws_new_array(const double* values, size_t n)
{
    or_draws* d = (or_draws*)calloc(1, sizeof *d);
    d->kind = OR_DS_ARRAY; d->arr = values; d->n = n;
    return d;
}

void or_draws_free(or_draws* d) { free(d); }
size_t or_draws_used(const or_draws* d) { return d->used; }

/* One uniform. MWC: brng() of rng/BRNG.cxx:30-34, range [0,1). Philox: (z+0.5)*2^-32, range
 * (0,1), exact in FP64; words are consumed in order, the block counter c3 advancing every 4.
 * Array: the next fed value (0.5 once the array is exhausted). */
double or_draws_next(or_draws* d)
{
    d->used++;
    switch (d->kind) {
    case OR_DS_MWC: {
        uint32_t z = or_mwc_next(&d->mu, &d->mv);
        return z * OR_MWC_SCALE;
    }
    case OR_DS_PHILOX: {
        if (d->pos == 4) {
            or_philox4x32_10(d->ctr, d->key, d->buf);
            d->ctr[3]++;
            d->pos = 0;
        }
        uint32_t z = d->buf[d->pos++];
        return ((double)z + 0.5) * OR_TWO_POW_M32;
    }
    default:
        if (d->used - 1 < d->n) return d->arr[d->used - 1];
        return 0.5;
    }
}

/* ======================================================================================
 * Distribution transforms
 * ====================================================================================== */

/* irand_uniform (rng/DiscreteUniformRNG.cxx:11-14): int(N * u). */
int or_irand_uniform(int n, or_draws* d) { return (int)(n * or_draws_next(d)); }

/* irand_Poisson (rng/PoissonRNG.cxx:11-22): Knuth's product of uniforms. */
int or_irand_poisson(double mean, or_draws* d)
{
    double lim = exp(-mean);
    double prod = 1.0;
    int i = 0;
    do {
        i++;
        prod *= or_draws_next(d);
    } while (prod >= lim);
    return i - 1;
}

/* frand_Rayleigh (rng/RayleighRNG.cxx:11-15). */
double or_frand_rayleigh(double s, or_draws* d)
{
    double u = or_draws_next(d);
    return s * sqrt(-2.0 * log(u));
}

/* frand_Gaussian (src/Boundary.cpp:49-50): E + sqrt(-2 V log u1) cos(2 pi u2). The two draws
 * are unsequenced in the reference; the MWC order is selectable, the Philox order is log
 * first. */
double or_frand_gaussian(double mean, double variance, or_draws* d)
{
    double ul, uc;
    if (d->kind == OR_DS_MWC && g_mwc_cos_first) {
        uc = or_draws_next(d);
        ul = or_draws_next(d);
    } else {
        ul = or_draws_next(d);
        uc = or_draws_next(d);
    }
    return mean + sqrt(-2.0 * variance * log(ul)) * cos(2.0 * M_PI * uc);
}

/* v3rand_isotropic (rng/IsotropicVectorRNG.cxx:11-19). */
void or_v3rand_isotropic(double* r, or_draws* d)
{
    double cos_t = 1.0 - 2.0 * or_draws_next(d);
    double sin_t = sqrt(1.0 - cos_t * cos_t);
    double e = 2.0 * M_PI * or_draws_next(d);
    r[0] = cos_t;
    r[1] = sin_t * cos(e);
    r[2] = sin_t * sin(e);
}

/* SampleGaussianPair (src/Boundary.cpp:27-41): no draws at zero variance; the first uniform
 * is redrawn until it lies strictly inside (0,1). */
static void gaussian_pair(double variance, double* first, double* second, or_draws* d)
{
    if (variance == 0.0) { *first = *second = 0.0; return; }
    double u;
    do { u = or_draws_next(d); } while (u <= 0.0 || u >= 1.0);
    double radius = sqrt(-2.0 * variance * log(u));
    double angle = 2.0 * M_PI * or_draws_next(d);
    *first = radius * cos(angle);
    *second = radius * sin(angle);
}

/* ======================================================================================
 * Geometry and wall models
 * ====================================================================================== */

/* InsideBox (src/Geometry.cpp:173-177): strict interior. */
int or_inside_box(double bx, double by, double bw, double bh, double x, double y)
{
    return fabs(x - bx) < bw / 2.0 && fabs(y - by) < bh / 2.0;
}

/* IntersectBox (src/Geometry.cpp:179-229): slab test, first entry in [0,1], exact corner
 * entries resolved to the x face, touching an edge or corner is a miss. */
int or_intersect_box(double bx, double by, double bw, double bh, double x, double y,
                     double dx, double dy, double* t_hit, double* nx, double* ny)
{
    double pos[2] = {x, y};
    double disp[2] = {dx, dy};
    double lo[2] = {bx - bw / 2.0, by - bh / 2.0};
    double hi[2] = {bx + bw / 2.0, by + bh / 2.0};
    double t_in = -INFINITY, t_out = INFINITY;
    int axis_hit = -1;
    double n_hit = 0.0;
    for (int a = 0; a < 2; ++a) {
        if (disp[a] == 0.0) {
            if (pos[a] <= lo[a] || pos[a] >= hi[a]) return 0;
            continue;
        }
        double t0 = (lo[a] - pos[a]) / disp[a];
        double t1 = (hi[a] - pos[a]) / disp[a];
        double n = -1.0;
        if (t0 > t1) { double s = t0; t0 = t1; t1 = s; n = 1.0; }
        if (t0 > t_in) { t_in = t0; axis_hit = a; n_hit = n; }
        if (t1 < t_out) t_out = t1;
    }
    if (axis_hit < 0 || t_in < 0.0 || t_in > 1.0 || t_in >= t_out) return 0;
    *t_hit = t_in;
    *nx = axis_hit == 0 ? n_hit : 0.0;
    *ny = axis_hit == 1 ? n_hit : 0.0;
    return 1;
}

/* SpecularScattering (src/Boundary.cpp:235-241). */
void or_specular(double* v, double nx, double ny)
{
    double vn = v[0] * nx + v[1] * ny;
    v[0] -= 2.0 * vn * nx;
    v[1] -= 2.0 * vn * ny;
}

/* DiffuseScattering (src/Boundary.cpp:205-233): Rayleigh normal, Gaussian tangent and z. */
void or_diffuse(double* v, double m, double tw, double nx, double ny, or_draws* d)
{
    double rtw = OR_KB * tw / m;
    double vn = or_frand_rayleigh(sqrt(rtw), d);
    double vt = or_frand_gaussian(0.0, rtw, d);
    double vz = or_frand_gaussian(0.0, rtw, d);
    double tx = -ny, ty = nx;
    v[0] = vn * nx + vt * tx;
    v[1] = vn * ny + vt * ty;
    v[2] = vz;
}

static int cll_coeff_ok(double an, double at)
{
    return isfinite(an) && an >= 0.0 && an <= 1.0 && isfinite(at) && at >= 0.0 && at <= 1.0;
}

/* CLLScattering (src/Boundary.cpp:243-282), SPARTA acc_n/acc_t convention. Returns 40 on
 * invalid coefficients (the reference throws std::invalid_argument). */
int or_cll(double* v, double m, double tw, double nx, double ny, double an, double at,
           or_draws* d)
{
    if (!cll_coeff_ok(an, at)) return DSMC_E_CLL_COEFF;
    if (an == 0.0 && at == 0.0) { or_specular(v, nx, ny); return 0; }
    double var = OR_KB * tw / m;
    double vn_in = v[0] * nx + v[1] * ny;
    double tx = -ny, ty = nx;
    double vt_in = v[0] * tx + v[1] * ty;
    double vz_in = v[2];
    double g1, g2;
    gaussian_pair(an * var, &g1, &g2, d);
    double keep_n = sqrt(1.0 - an) * fabs(vn_in);
    double vn = hypot(keep_n + g1, g2);
    double gt, gz;
    gaussian_pair(at * var, &gt, &gz, d);
    double keep_t = sqrt(1.0 - at);
    double vt = keep_t * vt_in + gt;
    v[0] = vn * nx + vt * tx;
    v[1] = vn * ny + vt * ty;
    v[2] = keep_t * vz_in + gz;
    return 0;
}

/* ======================================================================================
 * Setup (src/Initialization.cpp:210-288, src/Geometry.cpp:30-35)
 * ====================================================================================== */
void or_params_default(dsmc_params* p)
{
    memset(p, 0, sizeof *p);
    p->x1 = -20.0; p->x2 = 20.0; p->y1 = -20.0; p->y2 = 20.0; p->dz = 0.1;
    p->nx = 100; p->ny = 100; p->dl = 0.1;
    p->box_x = 0.0; p->box_y = 0.0; p->box_width = 4.0; p->box_height = 4.0;
    p->dt = 1.0e-5; p->n_step = 100000; p->first_sampling_step = 10000; p->sampling_period = 1;
    p->npc_free = 10; p->molar_mass = 0.040; p->mach = 4; p->n_free = 1e17; p->t_free = 200;
    p->attack_angle_deg = 0; p->t_wall = 300.0;
    p->d_ref = 4.17e-10; p->viscosity_index = 0.81; p->t_ref = 273.0;
    p->wall_model = DSMC_WALL_CLL; p->acc_n = 0.8; p->acc_t = 0.8;
    p->seed = 1; p->case_id = 0; p->max_per_cell = 0; p->capacity_factor = 0.0;
}

int or_constants_of(const dsmc_params* p, dsmc_constants* k)
{
    if (p->wall_model == DSMC_WALL_CLL && !cll_coeff_ok(p->acc_n, p->acc_t))
        return DSMC_E_CLL_COEFF;
    memset(k, 0, sizeof *k);
    k->molecule_mass = p->molar_mass / OR_NA;
    k->p_free = p->n_free * (OR_KB * p->t_free);
    k->u_free = p->mach * sqrt((5.0 / 3.0) * OR_KB * p->t_free / k->molecule_mass);
    k->ux_free = k->u_free * cos(M_PI * p->attack_angle_deg / 180.0);
    k->uy_free = -k->u_free * sin(M_PI * p->attack_angle_deg / 180.0);
    k->dx = (p->x2 - p->x1) / p->nx;
    k->dy = (p->y2 - p->y1) / p->ny;
    k->dv = k->dx * k->dy * p->dz;
    k->weight = k->dv * p->n_free / p->npc_free;
    k->sigma_ref = M_PI * p->d_ref * p->d_ref;
    k->omega = 2.0 * p->viscosity_index - 1.0;
    /* Gamma from the C library (the reference uses a rational approximation of it). */
    k->cr_ref = sqrt(4.0 * OR_KB * p->t_ref / k->molecule_mass)
                / pow(tgamma(2.5 - p->viscosity_index), 1.0 / k->omega);
    k->sc_max = 9.0 * k->sigma_ref * sqrt(OR_KB * p->t_free / k->molecule_mass);
    double q = 1.0 - k->omega;
    double mrm = pow(4.0 * OR_KB * p->t_free / k->molecule_mass, q / 2.0)
                 * tgamma((q + 3.0) / 2.0) / tgamma(1.5);
    double rate = k->sigma_ref * pow(k->cr_ref, k->omega) * mrm;
    k->mft_free = 1.0 / (p->n_free * rate);
    double cbar = sqrt(8.0 * OR_KB * p->t_free / (M_PI * k->molecule_mass));
    k->mfp_free = cbar * k->mft_free;
    k->kn_free = k->mfp_free / p->box_height;
    return 0;
}

/* ======================================================================================
 * Context
 * ====================================================================================== */
struct or_ctx {
    dsmc_params p;
    dsmc_constants k;
    int mode;
    uint32_t key0, key1;
    uint32_t max_per_cell;
    int ncell;
    /* particles (SoA, FP64; GPU mode stores FP32-representable velocities) */
    size_t np, cap;
    double *x, *y, *vx, *vy, *vz;
    /* index: counts, offsets (CSR), perm (IPC lists concatenated) */
    uint32_t* cnt;
    uint32_t* off;
    uint32_t* perm;
    size_t perm_cap;
    int indexed;
    /* per-cell sums, 17 doubles per cell (include/Cell.h:4-10) */
    double* cells;
    double* sgmax;
    /* counters */
    int64_t step, sample_step;
    int sampling_mode;
    double time, sample_time;
    int64_t last_nc, total_nc, last_inj, last_rem;
    int64_t strip_counts[4];
    int64_t wall_hits;
    int64_t particle_steps;
    double drag_impulse, drag_time;
    or_draws mwc; /* the reference's global stream (REF mode) */
};

static double vstore(const or_ctx* c, double v) { return c->mode == OR_MODE_GPU ? (double)(float)v : v; }

static void reserve(or_ctx* c, size_t n)
{
    if (n <= c->cap) return;
    size_t cap = c->cap ? c->cap : 1024;
    while (cap < n) cap *= 2;
    c->x = (double*)realloc(c->x, cap * sizeof(double));
    c->y = (double*)realloc(c->y, cap * sizeof(double));
    c->vx = (double*)realloc(c->vx, cap * sizeof(double));
    c->vy = (double*)realloc(c->vy, cap * sizeof(double));
    c->vz = (double*)realloc(c->vz, cap * sizeof(double));
    c->cap = cap;
}

or_ctx* or_create(const dsmc_params* p, int mode, int* status)
{
    dsmc_constants k;
    int rc = or_constants_of(p, &k);
    if (status) *status = rc;
    if (rc) return NULL;
    or_ctx* c = (or_ctx*)calloc(1, sizeof *c);
    c->p = *p; c->k = k; c->mode = mode;
    c->key0 = (uint32_t)p->seed;
    c->key1 = (uint32_t)(p->seed >> 32) ^ (p->case_id * 0x9E3779B9u);
    c->max_per_cell = p->max_per_cell ? p->max_per_cell : 1000u;
    c->ncell = p->nx * p->ny;
    c->cnt = (uint32_t*)calloc((size_t)c->ncell, sizeof(uint32_t));
    c->off = (uint32_t*)calloc((size_t)c->ncell + 1, sizeof(uint32_t));
    c->cells = (double*)calloc((size_t)c->ncell * 17, sizeof(double));
    c->sgmax = (double*)malloc((size_t)c->ncell * sizeof(double));
    for (int i = 0; i < c->ncell; ++i) c->sgmax[i] = k.sc_max;
    /* Default MWC state of rng/BRNG.cxx:15. */
    c->mwc.kind = OR_DS_MWC; c->mwc.mu = 521288629u; c->mwc.mv = 362436069u;
    reserve(c, 1024);
    return c;
}

void or_destroy(or_ctx* c)
{
    if (!c) return;
    free(c->x); free(c->y); free(c->vx); free(c->vy); free(c->vz);
    free(c->cnt); free(c->off); free(c->perm); free(c->cells); free(c->sgmax);
    free(c);
}

void or_set_mwc_seed(or_ctx* c, uint32_t u, uint32_t v) { c->mwc.mu = u; c->mwc.mv = v; }
void or_get_mwc_state(const or_ctx* c, uint32_t* u, uint32_t* v) { *u = c->mwc.mu; *v = c->mwc.mv; }

static void stream(const or_ctx* c, or_draws* d, uint32_t c0, uint32_t c1, uint32_t c2)
{
    philox_init(d, c->key0, c->key1, c0, c1, c2);
}

/* The step's SamplingMode rule (src/Simulation.cpp:139-142). */
static int sampling_rule(const or_ctx* c, int64_t step)
{
    return step > c->p.first_sampling_step && step % c->p.sampling_period == 0;
}

void or_set_step(or_ctx* c, int64_t step)
{
    c->step = step;
    c->sampling_mode = sampling_rule(c, step);
}

int or_set_particles(or_ctx* c, size_t n, const double* x, const double* y,
                     const double* vx, const double* vy, const double* vz)
{
    reserve(c, n);
    for (size_t i = 0; i < n; ++i) {
        c->x[i] = x[i]; c->y[i] = y[i];
        c->vx[i] = vstore(c, vx[i]); c->vy[i] = vstore(c, vy[i]); c->vz[i] = vstore(c, vz[i]);
    }
    c->np = n;
    c->indexed = 0;
    return 0;
}

size_t or_num_particles(const or_ctx* c) { return c->np; }

void or_get_particles(const or_ctx* c, double* x, double* y, double* vx, double* vy, double* vz)
{
    size_t n = c->np;
    memcpy(x, c->x, n * sizeof(double)); memcpy(y, c->y, n * sizeof(double));
    memcpy(vx, c->vx, n * sizeof(double)); memcpy(vy, c->vy, n * sizeof(double));
    memcpy(vz, c->vz, n * sizeof(double));
}

static void copy_particle(or_ctx* c, size_t dst, size_t src)
{
    c->x[dst] = c->x[src]; c->y[dst] = c->y[src];
    c->vx[dst] = c->vx[src]; c->vy[dst] = c->vy[src]; c->vz[dst] = c->vz[src];
}

/* vrand_MB (src/Boundary.cpp:52-57) with the bulk velocity (ux, uy, 0). */
static void maxwellian(double* v, double m, double ux, double uy, double t, or_draws* d)
{
    double rt = OR_KB * t / m;
    v[0] = or_frand_gaussian(ux, rt, d);
    v[1] = or_frand_gaussian(uy, rt, d);
    v[2] = or_frand_gaussian(0.0, rt, d);
}

/* Particle count of a generation rectangle (src/Boundary.cpp:62-71). */
static int64_t new_count(const or_ctx* c, double x1, double y1, double x2, double y2, or_draws* d)
{
    double vol = (x2 - x1) * (y2 - y1) * c->p.dz;
    double avg = c->p.n_free * vol / c->k.weight;
    if (avg > 20.0) {
        int64_t n = (int64_t)avg;
        if (or_draws_next(d) < avg - (double)n) n++;
        return n;
    }
    return or_irand_poisson(avg, d);
}

/* One new particle (src/Boundary.cpp:83-90): uniform position in the rectangle
 * (v2rand_uniform_rect, :44-47), Maxwellian velocity, one free flight of Dt if move. */
static void new_particle(or_ctx* c, size_t i, double x1, double y1, double x2, double y2,
                         double ux, double uy, double t, int move, or_draws* d)
{
    double v[3];
    c->x[i] = x1 + (x2 - x1) * or_draws_next(d);
    c->y[i] = y1 + (y2 - y1) * or_draws_next(d);
    maxwellian(v, c->k.molecule_mass, ux, uy, t, d);
    if (move) {
        c->x[i] += c->p.dt * v[0];
        c->y[i] += c->p.dt * v[1];
    }
    c->vx[i] = vstore(c, v[0]); c->vy[i] = vstore(c, v[1]); c->vz[i] = vstore(c, v[2]);
}

/* GenerateNewParticles (src/Boundary.cpp:60-92), REF mode: one global stream. */
static int64_t generate_ref(or_ctx* c, double x1, double y1, double x2, double y2, int move)
{
    int64_t n = new_count(c, x1, y1, x2, y2, &c->mwc);
    reserve(c, c->np + (size_t)n);
    for (int64_t j = 0; j < n; ++j)
        new_particle(c, c->np + (size_t)j, x1, y1, x2, y2, c->k.ux_free, c->k.uy_free,
                     c->p.t_free, move, &c->mwc);
    c->np += (size_t)n;
    return n;
}

/* InitialConditions (src/Initialization.cpp:290-313). */
int or_init_freestream(or_ctx* c)
{
    c->step = 0; c->time = 0.0; c->np = 0; c->sample_step = 0; c->sample_time = 0.0;
    c->sampling_mode = sampling_rule(c, 0);
    memset(c->cells, 0, (size_t)c->ncell * 17 * sizeof(double));
    c->wall_hits = 0; c->drag_impulse = 0.0; c->drag_time = 0.0;
    c->total_nc = 0;
    c->indexed = 0;
    double bx = c->p.box_x, by = c->p.box_y, bw = c->p.box_width, bh = c->p.box_height;
    if (c->mode == OR_MODE_REF) {
        generate_ref(c, c->p.x1, c->p.y1, c->p.x2, c->p.y2, 0);
        /* RemoveParticlesInsideBox (src/Geometry.cpp:231-247): swap with last. */
        for (size_t i = 0; i < c->np;) {
            if (or_inside_box(bx, by, bw, bh, c->x[i], c->y[i])) {
                if (i < c->np - 1) copy_particle(c, i, c->np - 1);
                c->np--;
            } else {
                i++;
            }
        }
        return 0;
    }
    or_draws d;
    stream(c, &d, 0, 0, OR_P_INIT_COUNT);
    int64_t n = new_count(c, c->p.x1, c->p.y1, c->p.x2, c->p.y2, &d);
    reserve(c, (size_t)n);
    size_t w = 0;
    for (int64_t j = 0; j < n; ++j) {
        stream(c, &d, (uint32_t)j, 0, OR_P_INIT);
        new_particle(c, w, c->p.x1, c->p.y1, c->p.x2, c->p.y2, c->k.ux_free, c->k.uy_free,
                     c->p.t_free, 0, &d);
        if (!or_inside_box(bx, by, bw, bh, c->x[w], c->y[w])) w++;
    }
    c->np = w;
    /* The GPU pipeline leaves its freshly filled pool sorted by cell. */
    return or_index(c);
}

/* MoveParticles (src/Boundary.cpp:284-328) for one particle. */
static int move_one(or_ctx* c, size_t i, or_draws* d)
{
    const dsmc_params* p = &c->p;
    double x0 = c->x[i], y0 = c->y[i];
    double dx = p->dt * c->vx[i];
    double dy = p->dt * c->vy[i];
    double t, nx, ny;
    if (!or_intersect_box(p->box_x, p->box_y, p->box_width, p->box_height, x0, y0, dx, dy,
                          &t, &nx, &ny)) {
        c->x[i] = x0 + dx;
        c->y[i] = y0 + dy;
        return 0;
    }
    double xh = x0 + t * dx;
    double yh = y0 + t * dy;
    double v[3] = {c->vx[i], c->vy[i], c->vz[i]};
    double vx_in = v[0];
    int rc = 0;
    switch (p->wall_model) {
    case DSMC_WALL_SPECULAR: or_specular(v, nx, ny); break;
    case DSMC_WALL_DIFFUSE: or_diffuse(v, c->k.molecule_mass, p->t_wall, nx, ny, d); break;
    default: rc = or_cll(v, c->k.molecule_mass, p->t_wall, nx, ny, p->acc_n, p->acc_t, d); break;
    }
    if (rc) return rc;
    c->drag_impulse += c->k.weight * c->k.molecule_mass * (vx_in - v[0]);
    c->wall_hits++;
    double rem = p->dt * (1.0 - t);
    const double eps = 1e-12;
    c->x[i] = xh + rem * v[0] + eps * nx;
    c->y[i] = yh + rem * v[1] + eps * ny;
    c->vx[i] = vstore(c, v[0]); c->vy[i] = vstore(c, v[1]); c->vz[i] = vstore(c, v[2]);
    return 0;
}

int or_move(or_ctx* c)
{
    c->indexed = 0;
    for (size_t i = 0; i < c->np; ++i) {
        int rc;
        if (c->mode == OR_MODE_REF) {
            rc = move_one(c, i, &c->mwc);
        } else {
            or_draws d;
            stream(c, &d, (uint32_t)i, (uint32_t)c->step, OR_P_MOVE);
            rc = move_one(c, i, &d);
        }
        if (rc) return rc;
    }
    return 0;
}

/* The four injection strips of BoundaryConditions (src/Boundary.cpp:100-173) at one rank:
 * left, right, bottom, top. The DLxDL corners are not covered. */
static void strip_rect(const or_ctx* c, int s, double r[4])
{
    const dsmc_params* p = &c->p;
    switch (s) {
    case 0: r[0] = p->x1 - p->dl; r[1] = p->y1; r[2] = p->x1; r[3] = p->y2; break;
    case 1: r[0] = p->x2; r[1] = p->y1; r[2] = p->x2 + p->dl; r[3] = p->y2; break;
    case 2: r[0] = p->x1; r[1] = p->y1 - p->dl; r[2] = p->x2; r[3] = p->y1; break;
    default: r[0] = p->x1; r[1] = p->y2; r[2] = p->x2; r[3] = p->y2 + p->dl; break;
    }
}

/* Closed outflow test of src/Boundary.cpp:183-188. */
static int outside_domain(const or_ctx* c, double x, double y)
{
    return x <= c->p.x1 || x >= c->p.x2 || y <= c->p.y1 || y >= c->p.y2;
}

int or_boundary(or_ctx* c)
{
    c->indexed = 0;
    size_t np0 = c->np;
    if (c->mode == OR_MODE_REF) {
        for (int s = 0; s < 4; ++s) {
            double r[4];
            strip_rect(c, s, r);
            c->strip_counts[s] = generate_ref(c, r[0], r[1], r[2], r[3], 1);
        }
        size_t before = c->np;
        for (size_t i = 0; i < c->np;) {
            if (outside_domain(c, c->x[i], c->y[i])) {
                if (i < c->np - 1) copy_particle(c, i, c->np - 1);
                c->np--;
            } else {
                i++;
            }
        }
        int64_t injected = (int64_t)(before - np0);
        c->last_rem = (int64_t)(before - c->np);
        c->last_inj = injected; /* REF mode cannot tell which removals were new particles */
        return 0;
    }
    /* GPU mode: keep existing survivors in order, then injected survivors in candidate order. */
    size_t w = 0;
    for (size_t i = 0; i < np0; ++i)
        if (!outside_domain(c, c->x[i], c->y[i])) { if (w != i) copy_particle(c, w, i); w++; }
    int64_t removed_old = (int64_t)(np0 - w);
    int64_t inj = 0;
    for (int s = 0; s < 4; ++s) {
        double r[4];
        strip_rect(c, s, r);
        or_draws d;
        stream(c, &d, (uint32_t)s, (uint32_t)c->step, OR_P_INJ_COUNT);
        int64_t n = new_count(c, r[0], r[1], r[2], r[3], &d);
        c->strip_counts[s] = n;
        reserve(c, w + 1);
        for (int64_t j = 0; j < n; ++j) {
            stream(c, &d, (uint32_t)j, (uint32_t)c->step, OR_P_INJECT | ((uint32_t)s << 8));
            reserve(c, w + 1);
            new_particle(c, w, r[0], r[1], r[2], r[3], c->k.ux_free, c->k.uy_free, c->p.t_free,
                         1, &d);
            if (!outside_domain(c, c->x[w], c->y[w])) { w++; inj++; }
        }
    }
    c->np = w;
    c->last_inj = inj;
    c->last_rem = removed_old;
    return 0;
}

void or_last_strip_counts(const or_ctx* c, int64_t counts[4])
{
    for (int s = 0; s < 4; ++s) counts[s] = c->strip_counts[s];
}

/* Indexing (src/Geometry.cpp:37-170) at one rank: k = floor((x-X1)/DX), l = floor((y-Y1)/DY),
 * outside the grid -> 20, a cell already holding max_per_cell -> 22, stable append. GPU mode
 * then reorders the particles physically into cell order. */
int or_index(or_ctx* c)
{
    const dsmc_params* p = &c->p;
    size_t n = c->np;
    if (c->perm_cap < n + 1) {
        c->perm_cap = n + 1024;
        c->perm = (uint32_t*)realloc(c->perm, c->perm_cap * sizeof(uint32_t));
    }
    uint32_t* key = (uint32_t*)malloc((n + 1) * sizeof(uint32_t));
    memset(c->cnt, 0, (size_t)c->ncell * sizeof(uint32_t));
    for (size_t i = 0; i < n; ++i) {
        int k = (int)floor((c->x[i] - p->x1) / c->k.dx);
        int l = (int)floor((c->y[i] - p->y1) / c->k.dy);
        if (k < 0 || k >= p->nx || l < 0 || l >= p->ny) { free(key); return DSMC_E_OUTSIDE_DOMAIN; }
        uint32_t cell = (uint32_t)k * (uint32_t)p->ny + (uint32_t)l;
        if (c->cnt[cell] >= c->max_per_cell) { free(key); return DSMC_E_CELL_CAPACITY; }
        c->cnt[cell]++;
        key[i] = cell;
    }
    c->off[0] = 0;
    for (int cell = 0; cell < c->ncell; ++cell) c->off[cell + 1] = c->off[cell] + c->cnt[cell];
    uint32_t* cur = (uint32_t*)malloc(((size_t)c->ncell + 1) * sizeof(uint32_t));
    memcpy(cur, c->off, (size_t)c->ncell * sizeof(uint32_t));
    for (size_t i = 0; i < n; ++i) c->perm[cur[key[i]]++] = (uint32_t)i;
    free(cur);
    free(key);
    if (c->mode == OR_MODE_GPU) {
        double* tmp = (double*)malloc((n + 1) * sizeof(double));
        double* arrs[5] = {c->x, c->y, c->vx, c->vy, c->vz};
        for (int a = 0; a < 5; ++a) {
            for (size_t i = 0; i < n; ++i) tmp[i] = arrs[a][c->perm[i]];
            memcpy(arrs[a], tmp, n * sizeof(double));
        }
        free(tmp);
    }
    c->indexed = 1;
    return 0;
}

void or_get_index(const or_ctx* c, uint32_t* counts, uint32_t* perm)
{
    if (counts) memcpy(counts, c->cnt, (size_t)c->ncell * sizeof(uint32_t));
    if (perm) memcpy(perm, c->perm, c->np * sizeof(uint32_t));
}

/* Particle at position j of cell list `cell`: an IPC entry (REF) or the sorted slot (GPU). */
static size_t member(const or_ctx* c, int cell, uint32_t j)
{
    size_t pos = (size_t)c->off[cell] + j;
    return c->mode == OR_MODE_GPU ? pos : (size_t)c->perm[pos];
}

/* ElasticCollision (src/Collision.cpp:7-19): isotropic scattering about the centre of mass. */
static void elastic(or_ctx* c, size_t a, size_t b, double cr, or_draws* d)
{
    double nvec[3];
    or_v3rand_isotropic(nvec, d);
    cr *= 0.5;
    double va[3] = {c->vx[a], c->vy[a], c->vz[a]};
    double vb[3] = {c->vx[b], c->vy[b], c->vz[b]};
    double vc[3] = {0.5 * (vb[0] + va[0]), 0.5 * (vb[1] + va[1]), 0.5 * (vb[2] + va[2])};
    double vr[3] = {cr * nvec[0], cr * nvec[1], cr * nvec[2]};
    c->vx[a] = vstore(c, vc[0] + vr[0]); c->vy[a] = vstore(c, vc[1] + vr[1]);
    c->vz[a] = vstore(c, vc[2] + vr[2]);
    c->vx[b] = vstore(c, vc[0] - vr[0]); c->vy[b] = vstore(c, vc[1] - vr[1]);
    c->vz[b] = vstore(c, vc[2] - vr[2]);
}

/* BinaryCollisionsInCell (src/Collision.cpp:21-44): NTC with the fixed SCMax (REF) or the
 * cell's (sigma*g)max raised on demand (GPU). */
static int collide_cell(or_ctx* c, int cell, or_draws* d)
{
    int npc = (int)c->cnt[cell];
    if (npc < 2) return 0;
    double scmax = c->mode == OR_MODE_GPU ? c->sgmax[cell] : c->k.sc_max;
    double npair = 0.5 * npc * (npc - 1) * c->k.weight * scmax * c->p.dt / c->k.dv;
    int nn = (int)npair;
    double frac = npair - nn;
    nn = (or_draws_next(d) < frac) ? nn + 1 : nn;
    int ncoll = 0;
    for (int q = 0; q < nn; ++q) {
        int i = or_irand_uniform(npc, d);
        int j;
        do { j = or_irand_uniform(npc, d); } while (j == i);
        size_t a = member(c, cell, (uint32_t)i), b = member(c, cell, (uint32_t)j);
        double g0 = c->vx[b] - c->vx[a], g1 = c->vy[b] - c->vy[a], g2 = c->vz[b] - c->vz[a];
        double cr = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
        double sigma = c->k.sigma_ref * pow(c->k.cr_ref / cr, c->k.omega);
        double sg = sigma * cr;
        if (c->mode == OR_MODE_GPU && sg > scmax) { scmax = sg; c->sgmax[cell] = sg; }
        if (or_draws_next(d) < sg / scmax) {
            elastic(c, a, b, cr, d);
            ncoll++;
        }
    }
    return ncoll;
}

/* CollideParticles (src/Collision.cpp:45-56). */
int or_collide(or_ctx* c, int64_t* ncoll)
{
    if (!c->indexed) return DSMC_E_ARGUMENT;
    int64_t total = 0;
    for (int cell = 0; cell < c->ncell; ++cell) {
        int nc;
        if (c->mode == OR_MODE_REF) {
            nc = collide_cell(c, cell, &c->mwc);
        } else {
            or_draws d;
            stream(c, &d, (uint32_t)cell, (uint32_t)c->step, OR_P_COLLIDE);
            nc = collide_cell(c, cell, &d);
        }
        total += nc;
        if (c->sampling_mode && nc) c->cells[(size_t)cell * 17 + 16] += nc;
    }
    c->last_nc = total;
    c->total_nc += total;
    if (ncoll) *ncoll = total;
    return 0;
}

/* Sampling (src/Sampling.cpp:14-31): FP64 running sums per cell, particles in list order. */
int or_sample(or_ctx* c)
{
    if (!c->indexed) return DSMC_E_ARGUMENT;
    c->sample_step++;
    c->sample_time += c->p.dt;
    for (int cell = 0; cell < c->ncell; ++cell) {
        double* s = c->cells + (size_t)cell * 17;
        for (uint32_t j = 0; j < c->cnt[cell]; ++j) {
            size_t i = member(c, cell, j);
            double v[3] = {c->vx[i], c->vy[i], c->vz[i]};
            s[0] += 1.0;
            double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            for (int m = 0; m < 3; ++m) {
                s[1 + m] += v[m];
                s[13 + m] += v2 * v[m];
                for (int q = 0; q < 3; ++q) s[4 + 3 * m + q] += v[m] * v[q];
            }
        }
    }
    return 0;
}

/* The time loop of src/Simulation.cpp:138-199 at one rank (the exchange is an identity). */
int or_step(or_ctx* c, int nsteps)
{
    for (int s = 0; s < nsteps; ++s) {
        c->sampling_mode = sampling_rule(c, c->step);
        int rc = or_move(c);
        if (rc) return rc;
        if (c->step > c->p.first_sampling_step) c->drag_time += c->p.dt;
        if ((rc = or_boundary(c))) return rc;
        if ((rc = or_index(c))) return rc;
        int64_t nc;
        if ((rc = or_collide(c, &nc))) return rc;
        if (c->sampling_mode && (rc = or_sample(c))) return rc;
        c->particle_steps += (int64_t)c->np;
        c->step++;
        c->time += c->p.dt;
    }
    c->sampling_mode = sampling_rule(c, c->step);
    return 0;
}

void or_get_cells(const or_ctx* c, double* cell17)
{
    memcpy(cell17, c->cells, (size_t)c->ncell * 17 * sizeof(double));
}

/* Printing's conversion (src/Sampling.cpp:152-196): N, Ux, Uy, Tx, Ty, T per cell. */
void or_get_fields(const or_ctx* c, double* out6)
{
    for (int cell = 0; cell < c->ncell; ++cell) {
        const double* s = c->cells + (size_t)cell * 17;
        double n = 0.0, u[3] = {0, 0, 0}, t[3] = {0, 0, 0};
        if (c->sample_step > 0) n = c->k.weight * s[0] / (double)c->sample_step / c->k.dv;
        if (s[0] > 0.0) {
            for (int m = 0; m < 3; ++m) {
                u[m] = s[1 + m] / s[0];
                t[m] = c->k.molecule_mass * (s[4 + 4 * m] / s[0] - u[m] * u[m]) / OR_KB;
            }
        }
        double* o = out6 + (size_t)cell * 6;
        o[0] = n; o[1] = u[0]; o[2] = u[1]; o[3] = t[0]; o[4] = t[1];
        o[5] = (t[0] + t[1] + t[2]) / 3.0;
    }
}

void or_get_state(const or_ctx* c, dsmc_state* s)
{
    memset(s, 0, sizeof *s);
    s->step = c->step; s->np = (int64_t)c->np; s->sample_step = c->sample_step;
    s->last_collisions = c->last_nc; s->total_collisions = c->total_nc;
    s->last_injected = c->last_inj; s->last_removed = c->last_rem;
    s->wall_hits = c->wall_hits; s->drag_impulse_total = c->drag_impulse;
    s->drag_sampling_time = c->drag_time; s->time = c->time; s->sample_time = c->sample_time;
    s->particle_steps = c->particle_steps;
    /* the drag report of src/Simulation.cpp:210-237 (commented out in the reference):
     * MeanDrag = DragImpulseTotal / DragSamplingTime, Cd = MeanDrag / (RhoFree UFree^2 / 2
     * AreaRef) with RhoFree = NFree MoleculeMass and AreaRef = BoxHeight DZ */
    if (c->drag_time > 0.0) {
        const double mean_drag = c->drag_impulse / c->drag_time;
        const double rho_free = c->p.n_free * c->k.molecule_mass;
        const double area_ref = c->p.box_height * c->p.dz;
        s->mean_drag = mean_drag;
        s->drag_coefficient = mean_drag / (0.5 * rho_free * c->k.u_free * c->k.u_free * area_ref);
    }
}

/* Warm-start seeding (GPU mode): each cell outside the box interior receives a count of
 * n*DV/W (integer part plus a Bernoulli draw, the rule of src/Boundary.cpp:62-71 with Poisson
 * below 20) particles, uniform in the cell, from the Maxwellian of the cell's (ux, uy, T);
 * particles inside the box are dropped. Cells are filled in k*NY+l order, so the pool comes
 * out sorted by cell. */
int or_init_from_cell_fields(or_ctx* c, const double* f)
{
    if (c->mode != OR_MODE_GPU) return DSMC_E_ARGUMENT;
    c->step = 0; c->time = 0.0; c->np = 0; c->sample_step = 0; c->sample_time = 0.0;
    c->sampling_mode = sampling_rule(c, 0);
    memset(c->cells, 0, (size_t)c->ncell * 17 * sizeof(double));
    c->wall_hits = 0; c->drag_impulse = 0.0; c->drag_time = 0.0; c->total_nc = 0;
    c->indexed = 0;
    const dsmc_params* p = &c->p;
    size_t w = 0;
    for (int cell = 0; cell < c->ncell; ++cell) {
        int k = cell / p->ny, l = cell % p->ny;
        double x1 = p->x1 + k * c->k.dx, y1 = p->y1 + l * c->k.dy;
        double x2 = x1 + c->k.dx, y2 = y1 + c->k.dy;
        double n = f[(size_t)cell * 4 + 0], ux = f[(size_t)cell * 4 + 1];
        double uy = f[(size_t)cell * 4 + 2], t = f[(size_t)cell * 4 + 3];
        if (!(n > 0.0) || !(t > 0.0)) continue;
        or_draws d;
        stream(c, &d, (uint32_t)cell, 0, OR_P_WARM);
        double avg = n * c->k.dv / c->k.weight;
        int64_t cnt;
        if (avg > 20.0) {
            cnt = (int64_t)avg;
            if (or_draws_next(&d) < avg - (double)cnt) cnt++;
        } else {
            cnt = or_irand_poisson(avg, &d);
        }
        for (int64_t j = 0; j < cnt; ++j) {
            reserve(c, w + 1);
            new_particle(c, w, x1, y1, x2, y2, ux, uy, t, 0, &d);
            if (!or_inside_box(p->box_x, p->box_y, p->box_width, p->box_height, c->x[w], c->y[w]))
                w++;
        }
    }
    c->np = w;
    return or_index(c);
}
