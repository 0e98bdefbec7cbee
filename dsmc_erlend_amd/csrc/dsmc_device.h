// dsmc_device.h - device-side physics and random numbers of the gfx950 DSMC pipeline.
//
// Every function here is a restatement of one reference function (cited) written for the GPU:
// FP64 arithmetic in the reference's operation order (the library is built with
// -ffp-contract=off, so a*b+c stays a rounded multiply then a rounded add), IEEE division and
// square root, and draws from a counter-based Philox4x32-10 stream instead of the reference's
// single global multiply-with-carry generator.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dsmc {

// Old CODATA constants the reference keeps (include/Constants.h:9-10).
constexpr double kBoltzmann = 1.380662E-23;
constexpr double kAvogadro = 6.022045E+23;
// 2.0*M_PI as the reference's expressions fold it (src/Boundary.cpp:38,50).
constexpr double kTwoPi = 2.0 * 3.14159265358979323846;
constexpr double kTwoPowM32 = 2.3283064365386962890625e-10;
constexpr uint32_t kSentinel = 0xFFFFFFFFu;
// In pools that defer free flight to the gather (Pool::defer): the key, and the scatter's
// pool position, of an entry whose step is already applied in place (a near-box particle
// after its exact move, a new particle) carry this bit; the cell or position is the rest.
constexpr uint32_t kMovedBit = 0x80000000u;
constexpr uint32_t kIndexMask = 0x7FFFFFFFu;

// Philox counter word 2 of each stream.
enum : uint32_t {
    kPMove = 1,       // c0 = pool slot, c1 = step
    kPInjCount = 2,   // c0 = strip, c1 = step
    kPInject = 3,     // c0 = candidate within strip, c1 = step, c2 |= strip << 8
    kPCollide = 4,    // c0 = cell within case, c1 = step
    kPInitCount = 5,  // c0 = 0, c1 = 0
    kPInit = 6,       // c0 = candidate, c1 = 0
    kPWarm = 7        // c0 = cell within case, c1 = 0
};

enum : int32_t { kWallDiffuse = 0, kWallSpecular = 1, kWallCLL = 2 };

// Everything a kernel needs to know about one case, resident in HBM (one entry per case).
struct CaseDev {
    double x1, x2, y1, y2;    // domain
    double dx, dy, dz, dv;    // cell size and volume
    double inv_dx, inv_dy;    // 1/dx, 1/dy: the fast path of cell_of
    double dl;                // strip width
    double box_x, box_y, box_w, box_h;
    double dt, m, weight;
    double sc_max, sigma_ref, cr_ref, omega;
    double ux_free, uy_free, t_free, n_free, t_wall;
    double acc_n, acc_t;
    double strip[4][4];       // injection rectangles x1,y1,x2,y2 (left, right, bottom, top)
    double strip_avg[4];      // mean new particles per strip and step
    double strip_reach[4];    // bound on a candidate's flight into the domain in one step
    int32_t nx, ny;
    int32_t wall;
    int32_t first_sampling_step, sampling_period;
    uint32_t cell_base, ncells;
    uint32_t key0, key1;      // Philox key
    uint32_t max_per_cell;
    int32_t retired;          // nonzero: the next step removes the case's particles unmoved
                              // and gives it no inflow
};

// ---------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and a sequential draw stream over it.
// ---------------------------------------------------------------------------------------
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    uint32_t c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    c[0] = c0; c[1] = c1; c[2] = c2; c[3] = c3;
}

// Uniforms (z + 0.5) * 2^-32 in (0,1), exact in FP64, four per Philox block in word order; the
// block counter (word 3) advances after every block.
struct Stream {
    uint32_t k0, k1, c0, c1, c2, blk;
    uint32_t w0, w1, w2, w3;
    int pos;

    __host__ __device__ Stream(uint32_t key0, uint32_t key1, uint32_t a, uint32_t b, uint32_t c)
        : k0(key0), k1(key1), c0(a), c1(b), c2(c), blk(0), w0(0), w1(0), w2(0), w3(0), pos(4) {}

    __host__ __device__ double next()
    {
        if (pos == 4) {
            uint32_t ctr[4] = {c0, c1, c2, blk};
            philox4x32_10(ctr, k0, k1);
            w0 = ctr[0]; w1 = ctr[1]; w2 = ctr[2]; w3 = ctr[3];
            ++blk;
            pos = 0;
        }
        uint32_t z = pos == 0 ? w0 : pos == 1 ? w1 : pos == 2 ? w2 : w3;
        ++pos;
        return ((double)z + 0.5) * kTwoPowM32;
    }
};

// ---------------------------------------------------------------------------------------
// Distribution transforms
// ---------------------------------------------------------------------------------------
// The transforms below take any source R of uniforms with next(): a Stream, or FedStream for
// runs on fed draws.

// irand_uniform (rng/DiscreteUniformRNG.cxx:11-14)
template <class R>
__device__ inline int irand_uniform(int n, R& s) { return (int)(n * s.next()); }

// irand_Poisson (rng/PoissonRNG.cxx:11-22), Knuth's method.
template <class R>
__host__ __device__ inline int irand_poisson(double mean, R& s)
{
    double lim = exp(-mean);
    double prod = 1.0;
    int i = 0;
    do {
        ++i;
        prod *= s.next();
    } while (prod >= lim);
    return i - 1;
}

// Count of new particles in a rectangle holding `avg` on average (src/Boundary.cpp:62-71).
template <class R>
__host__ __device__ inline uint32_t new_count(double avg, R& s)
{
    if (avg > 20.0) {
        uint32_t n = (uint32_t)avg;
        if (s.next() < avg - (double)n) ++n;
        return n;
    }
    return (uint32_t)irand_poisson(avg, s);
}

// frand_Gaussian (src/Boundary.cpp:49-50); the log() draw is taken first.
template <class R>
__device__ inline double frand_gaussian(double mean, double variance, R& s)
{
    double ul = s.next();
    double uc = s.next();
    return mean + sqrt(-2.0 * variance * log(ul)) * cos(kTwoPi * uc);
}

// frand_Rayleigh (rng/RayleighRNG.cxx:11-15)
template <class R>
__device__ inline double frand_rayleigh(double sigma, R& s)
{
    return sigma * sqrt(-2.0 * log(s.next()));
}

// SampleGaussianPair (src/Boundary.cpp:27-41). Stream uniforms never reach 0 or 1, so the
// open-interval redraw only ever fires on fed draws.
template <class R>
__device__ inline void gaussian_pair(double variance, double& first, double& second, R& s)
{
    if (variance == 0.0) { first = second = 0.0; return; }
    double u;
    do { u = s.next(); } while (u <= 0.0 || u >= 1.0);
    double radius = sqrt(-2.0 * variance * log(u));
    double angle = kTwoPi * s.next();
    first = radius * cos(angle);
    second = radius * sin(angle);
}

// vrand_MB (src/Boundary.cpp:52-57) with bulk velocity (ux, uy, 0).
template <class R>
__device__ inline void maxwellian(double v[3], double m, double ux, double uy, double t, R& s)
{
    double rt = kBoltzmann * t / m;
    v[0] = frand_gaussian(ux, rt, s);
    v[1] = frand_gaussian(uy, rt, s);
    v[2] = frand_gaussian(0.0, rt, s);
}

// The same with a temperature per component (a warm start from anisotropic local fields); with
// tx = ty = tz = t it makes the draws of maxwellian exactly.
template <class R>
__device__ inline void maxwellian3(double v[3], double m, double ux, double uy, double tx,
                                   double ty, double tz, R& s)
{
    v[0] = frand_gaussian(ux, kBoltzmann * tx / m, s);
    v[1] = frand_gaussian(uy, kBoltzmann * ty / m, s);
    v[2] = frand_gaussian(0.0, kBoltzmann * tz / m, s);
}

// ---------------------------------------------------------------------------------------
// Box geometry
// ---------------------------------------------------------------------------------------
// InsideBox (src/Geometry.cpp:173-177)
__host__ __device__ inline bool inside_box(const CaseDev& c, double x, double y)
{
    return fabs(x - c.box_x) < c.box_w / 2.0 && fabs(y - c.box_y) < c.box_h / 2.0;
}

// IntersectBox (src/Geometry.cpp:179-229): first entry into the box along the segment.
__host__ __device__ inline bool intersect_box(const CaseDev& c, double x, double y, double dx,
                                              double dy, double& t_hit, double& nx, double& ny)
{
    const double pos[2] = {x, y};
    const double disp[2] = {dx, dy};
    const double lo[2] = {c.box_x - c.box_w / 2.0, c.box_y - c.box_h / 2.0};
    const double hi[2] = {c.box_x + c.box_w / 2.0, c.box_y + c.box_h / 2.0};
    double t_in = -INFINITY, t_out = INFINITY;
    int axis_hit = -1;
    double n_hit = 0.0;
    for (int a = 0; a < 2; ++a) {
        if (disp[a] == 0.0) {
            if (pos[a] <= lo[a] || pos[a] >= hi[a]) return false;
            continue;
        }
        double t0 = (lo[a] - pos[a]) / disp[a];
        double t1 = (hi[a] - pos[a]) / disp[a];
        double n = -1.0;
        if (t0 > t1) { double s = t0; t0 = t1; t1 = s; n = 1.0; }
        if (t0 > t_in) { t_in = t0; axis_hit = a; n_hit = n; }
        if (t1 < t_out) t_out = t1;
    }
    if (axis_hit < 0 || t_in < 0.0 || t_in > 1.0 || t_in >= t_out) return false;
    t_hit = t_in;
    nx = axis_hit == 0 ? n_hit : 0.0;
    ny = axis_hit == 1 ? n_hit : 0.0;
    return true;
}

// A segment whose bounding interval stays 1e-9 m clear of the box slab on either axis cannot
// enter the box; IntersectBox's divides are then skipped. For |d| < 1 m the slab parameters
// of such a segment lie at least 1e-9 outside [0,1], far beyond the divides' 2-ulp error, so
// the shortcut returns exactly what IntersectBox would. The tests combine with & and | (no
// early exit), so a move kernel issues its position and velocity loads together.
__device__ inline bool clearly_misses_box(const CaseDev& c, double x, double y, double dx,
                                          double dy)
{
    const double adx = fabs(dx), ady = fabs(dy);
    const bool short_step = (adx < 1.0) & (ady < 1.0);
    const bool clear_x = fabs(x - c.box_x) > c.box_w / 2.0 + adx + 1e-9;
    const bool clear_y = fabs(y - c.box_y) > c.box_h / 2.0 + ady + 1e-9;
    return short_step & (clear_x | clear_y);
}

// ---------------------------------------------------------------------------------------
// Wall models (velocity in/out, FP64)
// ---------------------------------------------------------------------------------------
// SpecularScattering (src/Boundary.cpp:235-241)
__device__ inline void specular(double v[3], double nx, double ny)
{
    double vn = v[0] * nx + v[1] * ny;
    v[0] -= 2.0 * vn * nx;
    v[1] -= 2.0 * vn * ny;
}

// DiffuseScattering (src/Boundary.cpp:205-233)
template <class R>
__device__ inline void diffuse(double v[3], double m, double tw, double nx, double ny, R& s)
{
    double rtw = kBoltzmann * tw / m;
    double vn = frand_rayleigh(sqrt(rtw), s);
    double vt = frand_gaussian(0.0, rtw, s);
    double vz = frand_gaussian(0.0, rtw, s);
    double tx = -ny, ty = nx;
    v[0] = vn * nx + vt * tx;
    v[1] = vn * ny + vt * ty;
    v[2] = vz;
}

// CLLScattering (src/Boundary.cpp:243-282), SPARTA acc_n / acc_t. Coefficients are validated
// when the case is created (the reference's Setup check, src/Initialization.cpp:211-220).
template <class R>
__device__ inline void cll(double v[3], double m, double tw, double nx, double ny, double an,
                           double at, R& s)
{
    if (an == 0.0 && at == 0.0) { specular(v, nx, ny); return; }
    double var = kBoltzmann * tw / m;
    double vn_in = v[0] * nx + v[1] * ny;
    double tx = -ny, ty = nx;
    double vt_in = v[0] * tx + v[1] * ty;
    double vz_in = v[2];
    double g1, g2;
    gaussian_pair(an * var, g1, g2, s);
    double keep_n = sqrt(1.0 - an) * fabs(vn_in);
    double vn = hypot(keep_n + g1, g2);
    double gt, gz;
    gaussian_pair(at * var, gt, gz, s);
    double keep_t = sqrt(1.0 - at);
    double vt = keep_t * vt_in + gt;
    v[0] = vn * nx + vt * tx;
    v[1] = vn * ny + vt * ty;
    v[2] = keep_t * vz_in + gz;
}

// The box impact branch of MoveParticles (src/Boundary.cpp:301-326) for a particle whose
// segment hits the box: wall model at the hit point, then the rest of the flight plus a 1e-12
// push along the outward normal. Returns the x-momentum change per unit mass (vx_in - vx_out).
template <class R>
__device__ inline double wall_hit_with(const CaseDev& c, double& x, double& y, double dx,
                                       double dy, double v[3], double t, double nx, double ny,
                                       R& s)
{
    double xh = x + t * dx;
    double yh = y + t * dy;
    double vx_in = v[0];
    if (c.wall == kWallSpecular) specular(v, nx, ny);
    else if (c.wall == kWallDiffuse) diffuse(v, c.m, c.t_wall, nx, ny, s);
    else cll(v, c.m, c.t_wall, nx, ny, c.acc_n, c.acc_t, s);
    double rem = c.dt * (1.0 - t);
    const double eps = 1e-12;
    x = xh + rem * v[0] + eps * nx;
    y = yh + rem * v[1] + eps * ny;
    return vx_in - v[0];
}

// The same with the particle's own stream (slot within the case, step); kept out of line so the
// FP64 transcendental code does not raise the callers' register use.
__device__ __noinline__ double wall_hit(const CaseDev& c, double& x, double& y, double dx,
                                        double dy, double v[3], double t, double nx, double ny,
                                        uint32_t slot, uint32_t step)
{
    Stream s(c.key0, c.key1, slot, step, kPMove);
    return wall_hit_with(c, x, y, dx, dy, v, t, nx, ny, s);
}

// Uniforms read in order from a caller's array (the reference's one sequential stream,
// rng/BRNG.cxx:30-34): parity runs on fed draws, one GPU thread consuming them in the
// reference's order. Past the end it returns 0.5 and records the overrun.
struct FedStream {
    const double* u;
    uint64_t n;
    uint64_t pos;
    bool overrun;

    __device__ double next()
    {
        if (pos < n) return u[pos++];
        overrun = true;
        return 0.5;
    }
};

// The closed outflow test of BoundaryConditions (src/Boundary.cpp:183-188). This and cell_of
// combine their tests with | and & (no early exit), so the case fields load together.
__device__ inline bool outside_domain(const CaseDev& c, double x, double y)
{
    return (x <= c.x1) | (x >= c.x2) | (y <= c.y1) | (y >= c.y2);
}

// Indexing's cell coordinates (src/Geometry.cpp:47-59); false (cell unchanged) when outside
// the grid.
__device__ inline bool cell_of(const CaseDev& c, double x, double y, uint32_t& cell)
{
    // floor of the quotient by a multiply; within 1e-9 of a cell edge (where the product and the
    // rounded quotient could fall on different sides) the division decides. Both lie within
    // 2 * 2^-53 * quotient of the exact quotient, under 1e-9 for grids up to 1e6 cells a side
    // (checked at create), so the floor is the one the division gives.
    const double ax = x - c.x1, ay = y - c.y1;
    double qx = ax * c.inv_dx, qy = ay * c.inv_dy;
    const double fx = qx - floor(qx), fy = qy - floor(qy);
    if ((fx < 1e-9) | (fx > 1.0 - 1e-9)) qx = ax / c.dx;
    if ((fy < 1e-9) | (fy > 1.0 - 1e-9)) qy = ay / c.dy;
    const int k = (int)floor(qx);
    const int l = (int)floor(qy);
    const bool on_grid = (k >= 0) & (k < c.nx) & (l >= 0) & (l < c.ny);
    const uint32_t id = c.cell_base + (uint32_t)k * (uint32_t)c.ny + (uint32_t)l;
    cell = on_grid ? id : cell;
    return on_grid;
}

}  // namespace dsmc
