// dsmc_capi.hip - the C-ABI of include/dsmc_c.h over the gfx950 kernels.
//
// A context owns one particle pool shared by one or more independent cases (a case-offset
// table: each case owns a contiguous cell range, and after every sort its particles occupy a
// contiguous pool range), the device counters, and one HIP stream. The step API keeps the
// reference's stage names (src/Simulation.cpp:143-159); dsmc_step runs the same stages fused
// into three (small grids) or six launches per step, without host synchronisation, captured
// in a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dsmc_c.h"
#include "dsmc_kernels.h"

using namespace dsmc;

// Staging of dsmc_get_binned_cells_noalloc: the block sums of one call in the CELL layout, at
// most kNoAllocMaxBins blocks of 17 doubles (2.2 MB). Each GPU the library runs on holds its own
// copy until the process ends. In host code `scratch` is only a stand-in for the symbol: its
// device address on each GPU comes from hipGetSymbolAddress (scratch_address below).
constexpr int kNoAllocMaxBins = 16384;
static __device__ __align__(256) double scratch[(size_t)kNoAllocMaxBins * 17];

struct dsmc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int ncases = 0;
    std::vector<dsmc_params> params;
    std::vector<dsmc_constants> k;
    std::vector<CaseDev> hcases;
    std::vector<uint32_t> cell_base;  // [ncases+1]
    uint32_t mtot = 0;
    Pool P{};
    Particles buf[2]{};
    int cur = 0;
    uint32_t* perm_sorted = nullptr;
    uint32_t* bcnt = nullptr;
    uint32_t* init_base = nullptr;    // [ncases+1]
    uint32_t* warm_cnt = nullptr;     // [mtot+1]
    uint32_t* warm_off = nullptr;     // [mtot+1]
    uint32_t* warm_starts = nullptr;  // [ncases+1]
    double* warm_fields = nullptr;    // [6*mtot]
    double* bins = nullptr;           // [bins_cap * kSumRows] block sums of dsmc_get_binned_cells
    size_t bins_cap = 0;
    size_t gscratch_words = 0;        // allocated words of P.gscratch
    std::vector<void*> allocs;
    // host-side loop counters (the device keeps Step for its random streams)
    int64_t step = 0;
    std::vector<int64_t> sample_step;
    std::vector<double> drag_time, time, sample_time;
    bool indexed = false;
    bool perm_valid = false;  // perm_sorted holds the last dsmc_index permutation
    bool initialised = false;
    bool primed = false;      // the device holds the current step's inflow counts
    dsmc_tuning tuning{};     // launch schedule and gather chunk
    int ncu = 0;
    double nwork = 0.0;       // particles the grids are sized for
    // graphs of two fused steps, one per starting buffer (A->B->A and B->A->B), so a run of an
    // odd number of steps does not capture again on the next call
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    // events that bound the graph replays one dsmc_step call keeps queued (a ring of two)
    hipEvent_t inflight[2] = {nullptr, nullptr};
    std::string err;
    double* scratch = nullptr;  // this GPU's copy of `scratch` (device address)
    // fed draws (dsmc_set_draws): the caller's uniforms on the device and their read position
    double* fed_u = nullptr;
    size_t fed_cap = 0;
    FedStream* fed = nullptr;
    bool fed_on = false;
};

namespace {

int fail(dsmc_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

// Error path of the entry points that must not allocate: the message is copied into the
// context's error string, whose capacity is reserved at creation.
constexpr size_t kErrCapacity = 256;
int fail_noalloc(dsmc_ctx* c, int code, const char* msg)
{
    if (c) c->err.assign(msg, std::min(std::strlen(msg), kErrCapacity - 1));
    return code;
}

// Device address of `scratch` on `device`, looked up once per GPU (with that GPU current) and
// kept for the life of the process.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch[kMaxDevices] = {};

int scratch_address(dsmc_ctx* c, int device)
{
    if (device < 0 || device >= kMaxDevices)
        return fail(c, DSMC_E_ARGUMENT, "device index outside [0, 63]");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(c, DSMC_E_HIP, std::string("hipGetSymbolAddress(scratch): ")
                                           + hipGetErrorString(e));
        g_scratch[device] = static_cast<double*>(ptr);
    }
    c->scratch = g_scratch[device];
    return 0;
}

#define HIP_TRY(ctx, expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail((ctx), DSMC_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int dalloc(dsmc_ctx* c, T** p, size_t n)
{
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess)
        return fail(c, DSMC_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return 0;
}

template <class T>
int dzero(dsmc_ctx* c, T* p, size_t n)
{
    HIP_TRY(c, hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
    return 0;
}

bool cll_ok(double an, double at)
{
    return std::isfinite(an) && an >= 0.0 && an <= 1.0 && std::isfinite(at) && at >= 0.0
        && at <= 1.0;
}

// Setup (src/Initialization.cpp:221-266) and SetupGeometry (src/Geometry.cpp:30-35).
int derive(const dsmc_params& p, dsmc_constants& k)
{
    if (p.wall_model == DSMC_WALL_CLL && !cll_ok(p.acc_n, p.acc_t)) return DSMC_E_CLL_COEFF;
    if (p.wall_model < 0 || p.wall_model > 2) return DSMC_E_ARGUMENT;
    // at most 1e6 cells a side: the cell index's multiply-by-reciprocal path is exact there
    if (p.nx <= 0 || p.ny <= 0 || p.nx > 1000000 || p.ny > 1000000 || !(p.x2 > p.x1)
        || !(p.y2 > p.y1) || !(p.dt > 0.0) || !(p.npc_free > 0.0) || p.sampling_period <= 0)
        return DSMC_E_ARGUMENT;
    std::memset(&k, 0, sizeof k);
    k.molecule_mass = p.molar_mass / kAvogadro;
    k.p_free = p.n_free * (kBoltzmann * p.t_free);
    k.u_free = p.mach * std::sqrt((5.0 / 3.0) * kBoltzmann * p.t_free / k.molecule_mass);
    k.ux_free = k.u_free * std::cos(M_PI * p.attack_angle_deg / 180.0);
    k.uy_free = -k.u_free * std::sin(M_PI * p.attack_angle_deg / 180.0);
    k.dx = (p.x2 - p.x1) / p.nx;
    k.dy = (p.y2 - p.y1) / p.ny;
    k.dv = k.dx * k.dy * p.dz;
    k.weight = k.dv * p.n_free / p.npc_free;
    k.sigma_ref = M_PI * p.d_ref * p.d_ref;
    k.omega = 2.0 * p.viscosity_index - 1.0;
    k.cr_ref = std::sqrt(4.0 * kBoltzmann * p.t_ref / k.molecule_mass)
               / std::pow(std::tgamma(2.5 - p.viscosity_index), 1.0 / k.omega);
    k.sc_max = 9.0 * k.sigma_ref * std::sqrt(kBoltzmann * p.t_free / k.molecule_mass);
    const double q = 1.0 - k.omega;
    const double mrm = std::pow(4.0 * kBoltzmann * p.t_free / k.molecule_mass, q / 2.0)
                       * std::tgamma((q + 3.0) / 2.0) / std::tgamma(1.5);
    const double rate = k.sigma_ref * std::pow(k.cr_ref, k.omega) * mrm;
    k.mft_free = 1.0 / (p.n_free * rate);
    const double cbar = std::sqrt(8.0 * kBoltzmann * p.t_free / (M_PI * k.molecule_mass));
    k.mfp_free = cbar * k.mft_free;
    k.kn_free = k.mfp_free / p.box_height;
    return 0;
}

// The drag report the reference leaves commented out (src/Simulation.cpp:210-237): the mean
// drag force over the drag sampling time and its coefficient on the box's frontal height.
void drag_report(const dsmc_params& p, const dsmc_constants& k, dsmc_state* s)
{
    s->mean_drag = 0.0;
    s->drag_coefficient = 0.0;
    if (!(s->drag_sampling_time > 0.0)) return;
    s->mean_drag = s->drag_impulse_total / s->drag_sampling_time;
    const double rho_free = p.n_free * k.molecule_mass;
    const double area_ref = p.box_height * p.dz;
    s->drag_coefficient = s->mean_drag / (0.5 * rho_free * k.u_free * k.u_free * area_ref);
}

// The four strips of BoundaryConditions at one rank (src/Boundary.cpp:100-173).
void strips(const dsmc_params& p, double r[4][4])
{
    const double s[4][4] = {{p.x1 - p.dl, p.y1, p.x1, p.y2},
                            {p.x2, p.y1, p.x2 + p.dl, p.y2},
                            {p.x1, p.y1 - p.dl, p.x2, p.y1},
                            {p.x1, p.y2, p.x2, p.y2 + p.dl}};
    std::memcpy(r, s, sizeof s);
}

// Mean new particles of a rectangle (src/Boundary.cpp:62-63).
double rect_avg(const dsmc_params& p, const dsmc_constants& k, const double r[4])
{
    const double vol = (r[2] - r[0]) * (r[3] - r[1]) * p.dz;
    return p.n_free * vol / k.weight;
}

// How far a new particle of each strip can fly into the domain in one step, at most: the bulk
// velocity's inward component plus the largest deviation frand_gaussian can give from stream
// uniforms, which are at least 2^-33 (sqrt(-2 var log 2^-33) = 6.76 standard deviations), with
// a margin far above the rounding of the position update. A candidate further out than this
// cannot enter, whatever its velocity, and the inflow kernels skip its velocity draws (its own
// stream, so no other draw moves).
void strip_reach(const CaseDev& h, double reach[4])
{
    const double var = kBoltzmann * h.t_free / h.m;
    const double dev = std::sqrt(-2.0 * var * std::log(0.5 * kTwoPowM32)) * (1.0 + 1e-6);
    const double inward[4] = {h.ux_free, -h.ux_free, h.uy_free, -h.uy_free};
    const double extent = std::max(std::max(std::fabs(h.x1), std::fabs(h.x2)),
                                   std::max(std::fabs(h.y1), std::fabs(h.y2))) + h.dl;
    for (int s = 0; s < 4; ++s) {
        const double r = h.dt * (inward[s] + dev);
        reach[s] = r + 1e-6 * std::fabs(r) + 1e-9 * (1.0 + extent);
    }
}

// Upper bound of new_count(avg): integer part + 1, or a Poisson tail far beyond any draw.
uint32_t count_bound(double avg)
{
    if (avg > 20.0) return (uint32_t)avg + 1u;
    return (uint32_t)std::ceil(avg + 12.0 * std::sqrt(avg) + 30.0);
}

Particles cur_buf(dsmc_ctx* c) { return c->buf[c->cur]; }
Particles other_buf(dsmc_ctx* c) { return c->buf[1 - c->cur]; }

void invalidate_graph(dsmc_ctx* c)
{
    for (hipGraphExec_t& g : c->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// After a call on fed draws: an array that ran out is an error (the results are not the
// reference's on those draws).
int fed_check(dsmc_ctx* c)
{
    FedStream fs;
    HIP_TRY(c, hipMemcpy(&fs, c->fed, sizeof fs, hipMemcpyDeviceToHost));
    if (fs.overrun)
        return fail(c, DSMC_E_CAPACITY, "the fed draws ran out (dsmc_set_draws with more)");
    return 0;
}

// Synchronises and turns the device error word into a status code.
int sync_check(dsmc_ctx* c)
{
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    DevState st;
    HIP_TRY(c, hipMemcpy(&st, c->P.st, sizeof st, hipMemcpyDeviceToHost));
    if (st.err_code) {
        uint32_t zero[2] = {0, 0};
        (void)hipMemcpy(&c->P.st->err_code, zero, sizeof zero, hipMemcpyHostToDevice);
        char msg[160];
        const char* what = st.err_code == 20   ? "particle outside the global domain"
                           : st.err_code == 22 ? "too many particles in a cell"
                           : st.err_code == 1  ? "particle pool capacity exceeded"
                                               : "device error";
        std::snprintf(msg, sizeof msg, "%s (code %u, index %u)", what, st.err_code, st.err_index);
        c->indexed = false;
        c->primed = false;
        return fail(c, (int)st.err_code, msg);
    }
    return 0;
}

bool sampling_rule(const dsmc_params& p, int64_t step)
{
    return step > p.first_sampling_step && step % p.sampling_period == 0;
}

// Host bookkeeping of one completed step (src/Simulation.cpp:139-161); a retired case's clocks
// stop with its sums.
void account_step(dsmc_ctx* c)
{
    for (int i = 0; i < c->ncases; ++i) {
        if (c->hcases[i].retired) continue;
        const dsmc_params& p = c->params[i];
        if (c->step > p.first_sampling_step) c->drag_time[i] += p.dt;
        if (sampling_rule(p, c->step)) {
            c->sample_step[i] += 1;
            c->sample_time[i] += p.dt;
        }
        c->time[i] += p.dt;
    }
    c->step += 1;
}

// Index pipeline over keys already in P.key with counts in P.cnt -> sorted pool.
void sort_pool(dsmc_ctx* c, uint32_t* perm_sorted)
{
    launch_scan_cells(c->P, c->stream);
    launch_scatter(c->P, c->stream);
    launch_gather(c->P, cur_buf(c), other_buf(c), perm_sorted, false, c->stream);
    c->cur = 1 - c->cur;
}

// ceil(work) blocks within [lo, hi], a multiple of `mult` (hi is rounded down to one first).
int grid_for(double work, int lo, int hi, int mult)
{
    hi = std::max(lo, hi / mult * mult);
    const double w = std::ceil(work / mult) * mult;
    return (int)std::max<double>(lo, std::min<double>(hi, w));
}

bool small_schedule(const dsmc_ctx* c) { return c->tuning.schedule == DSMC_SCHEDULE_SMALL; }

// One fused step: the step opening (which first closes the previous fused step), then
//   large: move (+keys), near-box moves and inflow (+keys), scan, scatter, the gather that
//          samples the cells without collision candidates and lists the others, and the
//          collision list;
//   small: move with near-box moves and inflow (+keys), scan + scatter, and the gather that
//          also collides and samples every cell.
// The step stays pending until the next opening or close_fused_steps().
void enqueue_fused_step(dsmc_ctx* c, LaunchTimer* timer = nullptr)
{
    auto stage = [&](int s) { if (timer) timer->stage = s; };
    const bool small = small_schedule(c);
    stage(DSMC_STAGE_BEGIN);
    launch_step_open(c->P, true, c->stream);
    if (small) {
        stage(DSMC_STAGE_MOVE);
        launch_move_small(c->P, cur_buf(c), c->stream);
        stage(DSMC_STAGE_SCATTER);
        launch_scan_scatter_small(c->P, c->stream);
        stage(DSMC_STAGE_GATHER);
        launch_gather_small(c->P, cur_buf(c), other_buf(c), c->stream);
        c->cur = 1 - c->cur;
        return;
    }
    stage(DSMC_STAGE_MOVE);
    launch_move(c->P, cur_buf(c), true, c->stream);
    stage(DSMC_STAGE_EXTRA);
    launch_extra(c->P, cur_buf(c), true, true, c->stream);
    stage(DSMC_STAGE_SCAN);
    launch_scan_cells(c->P, c->stream);
    stage(DSMC_STAGE_SCATTER);
    launch_scatter(c->P, c->stream);
    stage(DSMC_STAGE_GATHER);
    launch_gather(c->P, cur_buf(c), other_buf(c), nullptr, true, c->stream);
    c->cur = 1 - c->cur;
    stage(DSMC_STAGE_COLLIDE);
    launch_collide_list(c->P, cur_buf(c), c->stream);
}

// Finishes the last fused step, so the device state is complete when an entry point returns.
void close_fused_steps(dsmc_ctx* c)
{
    launch_step_open(c->P, false, c->stream);
}

// A fused step starts from an empty histogram (a step cut short by an error may have left
// counts); the small schedule's scatter also starts its cursors from zero (its gather clears
// both for the next step, the large schedule's scan does so itself).
int prime(dsmc_ctx* c)
{
    if (!c->primed) {
        HIP_TRY(c, hipMemsetAsync(c->P.cnt, 0, sizeof(uint32_t) * c->mtot, c->stream));
        if (small_schedule(c))
            HIP_TRY(c, hipMemsetAsync(c->P.cursor, 0, sizeof(uint32_t) * c->mtot, c->stream));
        c->primed = true;
    }
    return 0;
}

// The gather stages chunk + max_per_cell entries in LDS.
bool gather_fits(const dsmc_ctx* c, int chunk)
{
    uint32_t max_pc = 0;
    for (const CaseDev& h : c->hcases) max_pc = std::max(max_pc, h.max_per_cell);
    return (size_t)chunk + max_pc <= (size_t)kGatherLdsMaxEntries;
}

// Grids of the kernels that stream the pool, sized for c->nwork particles. A multiple of 8
// lets the scatter deal contiguous slices to the 8 XCDs. The kernels that stream the whole pool
// get a thread per particle (their grid-stride loops only run again when the pool outgrows the
// estimate): waves that retire after one load-compute-store make room for fresh ones, which
// keeps more loads in flight than a capped grid whose waves loop (measured on the move
// pattern: 5.2 against 4.2 TB/s). The gather's grid follows the tuning.
void size_pool_grids(dsmc_ctx* c)
{
    Pool& P = c->P;
    const int ppt = std::max(1, (int)c->tuning.particles_per_thread);
    P.grid_p = grid_for(c->nwork / 256.0, 8, c->ncu * 8, 8);
    P.grid_n = grid_for(c->nwork / (256.0 * ppt), 8, 1 << 22, 8);
    P.grid_n4 = grid_for(c->nwork / 1024.0, 8, 1 << 22, 8);
    P.grid_s = grid_for(c->nwork / 1024.0, 1, c->ncu, 1);
}

// Applies a tuning: chunk size, the gather's LDS and grid, the launch schedule.
int apply_tuning(dsmc_ctx* c, const dsmc_tuning& t)
{
    Pool& P = c->P;
    if (t.chunk < 128 || t.chunk > 2048 || (t.chunk & (t.chunk - 1)))
        return fail(c, DSMC_E_ARGUMENT, "chunk must be a power of two in [128, 2048]");
    if (t.schedule != DSMC_SCHEDULE_LARGE && t.schedule != DSMC_SCHEDULE_SMALL)
        return fail(c, DSMC_E_ARGUMENT, "unknown schedule");
    if (t.particles_per_thread != 0 && t.particles_per_thread != 1 && t.particles_per_thread != 2
        && t.particles_per_thread != 4)
        return fail(c, DSMC_E_ARGUMENT, "particles_per_thread must be 1, 2 or 4");
    if (t.schedule == DSMC_SCHEDULE_SMALL && c->mtot > kSmallCells)
        return fail(c, DSMC_E_ARGUMENT, "the small schedule needs at most 12288 cells");
    if (!gather_fits(c, t.chunk))
        return fail(c, DSMC_E_ARGUMENT, "chunk + max_per_cell exceed the gather's 64 KiB of LDS");
    uint32_t max_pc = 0;
    for (const CaseDev& h : c->hcases) max_pc = std::max(max_pc, h.max_per_cell);
    P.chunk = t.chunk;
    P.lds_cap = P.chunk + (int)std::min<uint32_t>(max_pc, (uint32_t)kGatherSpare);
    P.gcap = P.chunk + (int)max_pc;
    const int per_cu = gather_blocks_per_cu(kGatherLdsPerEntry * P.lds_cap);
    P.grid_g = grid_for(c->nwork / t.chunk, 8, c->ncu * per_cu, 8);
    // global staging for chunks that overflow the LDS: one region of 5 * gcap words per block
    const size_t words = (size_t)P.grid_g * 5u * (size_t)P.gcap;
    if (words > c->gscratch_words) {
        if (P.gscratch) {
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            for (auto& a : c->allocs)
                if (a == P.gscratch) a = nullptr;
            (void)hipFree(P.gscratch);
            P.gscratch = nullptr;
        }
        if (int rc = dalloc(c, &P.gscratch, words)) return rc;
        c->gscratch_words = words;
    }
    c->tuning = t;
    c->tuning.particles_per_thread = std::max(1, (int)t.particles_per_thread);
    size_pool_grids(c);
    invalidate_graph(c);
    c->primed = false;
    return 0;
}

int fill_sgmax(dsmc_ctx* c)
{
    std::vector<double> sg(c->mtot);
    for (int i = 0; i < c->ncases; ++i)
        std::fill(sg.begin() + c->cell_base[i], sg.begin() + c->cell_base[i + 1], c->k[i].sc_max);
    HIP_TRY(c, hipMemcpy(c->P.sgmax, sg.data(), sizeof(double) * c->mtot, hipMemcpyHostToDevice));
    return 0;
}

int create_impl(const dsmc_params* cases, int ncases, int device, dsmc_ctx** out)
{
    if (!out) return DSMC_E_ARGUMENT;
    *out = nullptr;
    if (!cases || ncases <= 0 || ncases > 1024) return DSMC_E_ARGUMENT;
    dsmc_ctx* c = new dsmc_ctx();
    *out = c;
    c->device = device;
    c->ncases = ncases;
    c->params.assign(cases, cases + ncases);
    c->k.resize(ncases);
    c->hcases.resize(ncases);
    c->cell_base.resize(ncases + 1);
    uint32_t max_pc = 0, min_pc = 0xFFFFFFFFu;
    double n0 = 0.0, nfree = 0.0;
    uint64_t ncand_max = 0;
    uint64_t mtot = 0;
    for (int i = 0; i < ncases; ++i) {
        const dsmc_params& p = c->params[i];
        int rc = derive(p, c->k[i]);
        if (rc)
            return fail(c, rc, rc == DSMC_E_CLL_COEFF
                                   ? "CLL accommodation coefficients must be in [0, 1]"
                                   : "invalid case parameters");
        const dsmc_constants& k = c->k[i];
        CaseDev& h = c->hcases[i];
        std::memset(&h, 0, sizeof h);
        h.x1 = p.x1; h.x2 = p.x2; h.y1 = p.y1; h.y2 = p.y2;
        h.dx = k.dx; h.dy = k.dy; h.dz = p.dz; h.dv = k.dv; h.dl = p.dl;
        h.inv_dx = 1.0 / k.dx; h.inv_dy = 1.0 / k.dy;
        h.box_x = p.box_x; h.box_y = p.box_y; h.box_w = p.box_width; h.box_h = p.box_height;
        h.dt = p.dt; h.m = k.molecule_mass; h.weight = k.weight;
        h.sc_max = k.sc_max; h.sigma_ref = k.sigma_ref; h.cr_ref = k.cr_ref; h.omega = k.omega;
        h.ux_free = k.ux_free; h.uy_free = k.uy_free; h.t_free = p.t_free; h.n_free = p.n_free;
        h.t_wall = p.t_wall; h.acc_n = p.acc_n; h.acc_t = p.acc_t;
        strips(p, h.strip);
        strip_reach(h, h.strip_reach);
        for (int s = 0; s < 4; ++s) {
            h.strip_avg[s] = rect_avg(p, k, h.strip[s]);
            ncand_max += count_bound(h.strip_avg[s]);
        }
        h.nx = p.nx; h.ny = p.ny; h.wall = p.wall_model;
        h.first_sampling_step = p.first_sampling_step; h.sampling_period = p.sampling_period;
        h.cell_base = (uint32_t)mtot;
        h.ncells = (uint32_t)p.nx * (uint32_t)p.ny;
        h.key0 = (uint32_t)p.seed;
        h.key1 = (uint32_t)(p.seed >> 32) ^ (p.case_id * 0x9E3779B9u);
        h.max_per_cell = p.max_per_cell ? p.max_per_cell : 1000u;
        // the gather stages chunk + max_per_cell entries of 20 B in LDS: keep it within 64 KiB
        if (h.max_per_cell > 2048u) return fail(c, DSMC_E_ARGUMENT, "max_per_cell above 2048");
        max_pc = std::max(max_pc, h.max_per_cell);
        min_pc = std::min(min_pc, h.max_per_cell);
        c->cell_base[i] = (uint32_t)mtot;
        mtot += h.ncells;
        const double factor = p.capacity_factor > 0.0 ? p.capacity_factor : 2.0;
        n0 += factor * p.npc_free * (double)h.ncells;
        nfree += p.npc_free * (double)h.ncells;
    }
    if (mtot >= 0xFFFFFFF0ull) return fail(c, DSMC_E_ARGUMENT, "too many cells");
    c->mtot = (uint32_t)mtot;
    c->cell_base[ncases] = c->mtot;
    const double capd = n0 + (double)ncand_max + 4096.0;
    if (capd > 4.0e9) return fail(c, DSMC_E_ARGUMENT, "particle capacity above 4e9");
    const uint32_t cap = (uint32_t)capd;

    HIP_TRY(c, hipSetDevice(device));
    c->err.reserve(kErrCapacity);
    if (int rc = scratch_address(c, device)) return rc;
    HIP_TRY(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    int ncu = 0;
    HIP_TRY(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));

    Pool& P = c->P;
    P.ncases = ncases;
    P.mtot = c->mtot;
    P.cap = cap;
    P.ncand_max = (uint32_t)ncand_max;
    // cells and pool positions below 2^31 leave the top bit of keys and scatter entries free
    // for kMovedBit, which the deferred flight of the fused step needs
    P.defer = (c->mtot <= kIndexMask && cap <= kIndexMask) ? 1 : 0;
    P.min_max_per_cell = min_pc;
    c->ncu = ncu;
    // Grids are persistent (grid-stride loops), sized for the expected pool with room for
    // growth: a block with no work still costs its dispatch.
    c->nwork = 1.25 * nfree;
    // default tuning: enough gather chunks to give every CU a few (at most 1024 particles
    // each), and the small schedule when the cell scan fits one block's LDS
    int chunk = 1024;
    while (chunk > 128 && nfree / chunk < 4.0 * ncu) chunk /= 2;
    dsmc_tuning t0{};
    t0.chunk = chunk;
    t0.schedule = c->mtot <= kSmallCells ? DSMC_SCHEDULE_SMALL : DSMC_SCHEDULE_LARGE;
    (void)max_pc;
    size_pool_grids(c);
    P.grid_c = grid_for(c->mtot / 256.0, 1, ncu * 8, 1);
    P.grid_l = grid_for(c->mtot / 256.0, 1, ncu * 2, 1);
    int rc = 0;
    CaseDev* dcases = nullptr;
    uint32_t* dcb = nullptr;
    CaseCounters& q = P.cc;
    if ((rc = dalloc(c, &dcases, ncases)) || (rc = dalloc(c, &dcb, ncases + 1))
        || (rc = dalloc(c, &P.case_pstart, ncases + 1)) || (rc = dalloc(c, &P.cand_base, 4 * ncases + 1))
        || (rc = dalloc(c, &P.st, 1)) || (rc = dalloc(c, &P.key, cap))
        || (rc = dalloc(c, &P.hitlist, cap)) || (rc = dalloc(c, &P.perm, cap))
        || (rc = dalloc(c, &P.cnt, c->mtot)) || (rc = dalloc(c, &P.off, (size_t)c->mtot + 1))
        || (rc = dalloc(c, &P.cursor, c->mtot))
        || (rc = dalloc(c, &P.first, ((size_t)cap + c->mtot / 4) / 128 + 4))
        || (rc = dalloc(c, &P.coll_list, c->mtot))
        || (rc = dalloc(c, &P.partial, 2048))
        || (rc = dalloc(c, &P.sums, (size_t)kSumRows * c->mtot))
        || (rc = dalloc(c, &P.sgmax, c->mtot))
        || (rc = dalloc(c, &q.ncoll, ncases)) || (rc = dalloc(c, &q.injected, ncases))
        || (rc = dalloc(c, &q.removed, ncases)) || (rc = dalloc(c, &q.ncoll_last, ncases))
        || (rc = dalloc(c, &q.injected_last, ncases)) || (rc = dalloc(c, &q.removed_last, ncases))
        || (rc = dalloc(c, &q.ncoll_total, ncases)) || (rc = dalloc(c, &q.hits, ncases))
        || (rc = dalloc(c, &q.psteps, ncases)) || (rc = dalloc(c, &q.drag, ncases))
        || (rc = dalloc(c, &c->perm_sorted, cap))
        || (rc = dalloc(c, &c->bcnt, 2 * ((size_t)cap / 4096 + 4)))
        || (rc = dalloc(c, &c->init_base, ncases + 1)))
        return rc;
    for (int b = 0; b < 2; ++b) {
        Particles& pb = c->buf[b];
        if ((rc = dalloc(c, &pb.xy, cap)) || (rc = dalloc(c, &pb.vxy, cap))
            || (rc = dalloc(c, &pb.vz, cap)))
            return rc;
    }
    P.cases = dcases;
    P.cell_base = dcb;
    HIP_TRY(c, hipMemcpy(dcases, c->hcases.data(), sizeof(CaseDev) * ncases, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(dcb, c->cell_base.data(), sizeof(uint32_t) * (ncases + 1),
                         hipMemcpyHostToDevice));
    DevState st0{};
    HIP_TRY(c, hipMemcpy(P.st, &st0, sizeof st0, hipMemcpyHostToDevice));
    if ((rc = dzero(c, P.cnt, c->mtot)) || (rc = dzero(c, P.off, (size_t)c->mtot + 1))
        || (rc = dzero(c, P.case_pstart, ncases + 1))
        || (rc = dzero(c, P.sums, (size_t)kSumRows * c->mtot))
        || (rc = dzero(c, q.ncoll, ncases)) || (rc = dzero(c, q.injected, ncases))
        || (rc = dzero(c, q.removed, ncases)) || (rc = dzero(c, q.ncoll_last, ncases))
        || (rc = dzero(c, q.injected_last, ncases)) || (rc = dzero(c, q.removed_last, ncases))
        || (rc = dzero(c, q.ncoll_total, ncases)) || (rc = dzero(c, q.hits, ncases))
        || (rc = dzero(c, q.psteps, ncases)) || (rc = dzero(c, q.drag, ncases)))
        return rc;
    if ((rc = fill_sgmax(c))) return rc;
    if ((rc = apply_tuning(c, t0))) return rc;
    c->sample_step.assign(ncases, 0);
    c->drag_time.assign(ncases, 0.0);
    c->time.assign(ncases, 0.0);
    c->sample_time.assign(ncases, 0.0);
    return 0;
}

int check_ctx(dsmc_ctx* c, bool need_init)
{
    if (!c || !c->stream) return DSMC_E_ARGUMENT;
    if (need_init && !c->initialised)
        return fail(c, DSMC_E_ARGUMENT, "no particles: call dsmc_init_freestream first");
    HIP_TRY(c, hipSetDevice(c->device));
    return 0;
}

// Grids for the cases still running (1.25 x their freestream particles): the thread-per-
// particle kernels of a batch whose big cases have retired would otherwise dispatch millions of
// blocks with nothing to do every step. The next dsmc_step captures its graph again.
int resize_for_running_cases(dsmc_ctx* c)
{
    double nfree = 0.0;
    for (int i = 0; i < c->ncases; ++i)
        if (!c->hcases[i].retired) nfree += c->params[i].npc_free * (double)c->hcases[i].ncells;
    c->nwork = 1.25 * nfree;
    size_pool_grids(c);
    return apply_tuning(c, c->tuning);
}

int check_single(dsmc_ctx* c)
{
    if (c->ncases != 1)
        return fail(c, DSMC_E_ARGUMENT, "the single-stage API needs a one-case context");
    if (c->hcases[0].retired)
        return fail(c, DSMC_E_ARGUMENT, "the case is retired: initialise the context again");
    return 0;
}

// Writes case i's retired flag into the device case table (between steps: the stream is
// idle, and the captured step graph reads the table when it runs).
int upload_retired(dsmc_ctx* c, int i)
{
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    int32_t* dst = const_cast<int32_t*>(&c->P.cases[i].retired);  // the table the host owns
    HIP_TRY(c, hipMemcpy(dst, &c->hcases[i].retired, sizeof(int32_t), hipMemcpyHostToDevice));
    return 0;
}

int check_case(dsmc_ctx* c, int i)
{
    if (!c || i < 0 || i >= c->ncases) return fail(c, DSMC_E_ARGUMENT, "case index out of range");
    return 0;
}

void reset_counters(dsmc_ctx* c)
{
    c->step = 0;
    std::fill(c->sample_step.begin(), c->sample_step.end(), 0);
    std::fill(c->drag_time.begin(), c->drag_time.end(), 0.0);
    std::fill(c->time.begin(), c->time.end(), 0.0);
    std::fill(c->sample_time.begin(), c->sample_time.end(), 0.0);
}

int reset_device_sums(dsmc_ctx* c)
{
    Pool& P = c->P;
    const size_t n = c->ncases;
    HIP_TRY(c, hipMemsetAsync(P.sums, 0, sizeof(double) * kSumRows * c->mtot, c->stream));
    unsigned long long* ctrs[] = {P.cc.ncoll, P.cc.injected, P.cc.removed, P.cc.ncoll_last,
                                  P.cc.injected_last, P.cc.removed_last, P.cc.ncoll_total,
                                  P.cc.hits, P.cc.psteps};
    for (auto* q : ctrs) HIP_TRY(c, hipMemsetAsync(q, 0, sizeof(unsigned long long) * n, c->stream));
    HIP_TRY(c, hipMemsetAsync(P.cc.drag, 0, sizeof(double) * n, c->stream));
    HIP_TRY(c, hipMemsetAsync(&P.st->step, 0, sizeof(uint32_t), c->stream));
    return 0;
}

// The per-step counters of a single-stage call become the "last step" values.
int publish_stage_counters(dsmc_ctx* c, bool collisions, bool boundary)
{
    const Pool& P = c->P;
    const size_t b = sizeof(unsigned long long) * c->ncases;
    if (collisions) HIP_TRY(c, hipMemcpyAsync(P.cc.ncoll_last, P.cc.ncoll, b, hipMemcpyDeviceToDevice, c->stream));
    if (boundary) {
        HIP_TRY(c, hipMemcpyAsync(P.cc.injected_last, P.cc.injected, b, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(P.cc.removed_last, P.cc.removed, b, hipMemcpyDeviceToDevice, c->stream));
    }
    return 0;
}

// Index pass after a fill in the other buffer, sigma*g max reset, counters reset.
int finish_fill(dsmc_ctx* c, KeyMode mode, const uint32_t* starts, const uint32_t* n_ptr)
{
    // a fresh fill starts every case again, retired ones included
    bool any = false;
    for (int i = 0; i < c->ncases; ++i)
        if (c->hcases[i].retired) {
            c->hcases[i].retired = 0;
            any = true;
            if (int rc = upload_retired(c, i)) return rc;
        }
    if (any)
        if (int rc = resize_for_running_cases(c)) return rc;
    c->cur = 1 - c->cur;  // the generated fill is now the pool being keyed
    launch_keys(c->P, cur_buf(c), mode, starts, n_ptr, c->stream);
    sort_pool(c, c->perm_sorted);
    if (int rc = fill_sgmax(c)) return rc;
    reset_counters(c);
    if (int rc = sync_check(c)) return rc;
    c->initialised = true;
    c->indexed = true;
    c->perm_valid = true;
    c->primed = false;
    return 0;
}

// Warm-start fill from nf values per cell (cases concatenated, k*NY+l within a case): n, ux,
// uy, T (nf = 4) or n, ux, uy, Tx, Ty, Tz (nf = 6).
int init_from_cells(dsmc_ctx* c, const double* cf, int nf)
{
    if (int rc = check_ctx(c, false)) return rc;
    if (!cf) return fail(c, DSMC_E_ARGUMENT, "bad field array");
    invalidate_graph(c);
    Pool& P = c->P;
    int rc = 0;
    if (!c->warm_cnt) {
        if ((rc = dalloc(c, &c->warm_cnt, (size_t)c->mtot + 1)) || (rc = dalloc(c, &c->warm_off, (size_t)c->mtot + 1))
            || (rc = dalloc(c, &c->warm_starts, c->ncases + 1)) || (rc = dalloc(c, &c->warm_fields, 6 * (size_t)c->mtot)))
            return rc;
    }
    for (int i = 0; i < c->ncases; ++i) {
        const double per_particle = c->k[i].dv / c->k[i].weight;
        for (uint32_t q = c->cell_base[i]; q < c->cell_base[i + 1]; ++q) {
            const double* v = cf + (size_t)nf * q;
            bool ok = v[0] >= 0.0 && v[0] * per_particle <= 65536.0;
            for (int a = 0; a < nf; ++a) ok = ok && std::isfinite(v[a]);
            for (int a = 3; a < nf; ++a) ok = ok && v[a] >= 0.0;
            if (!ok)
                return fail(c, DSMC_E_ARGUMENT, "warm-start fields must be finite, n and "
                                                "temperatures >= 0, and at most 65536 particles "
                                                "per cell");
        }
    }
    HIP_TRY(c, hipMemcpyAsync(c->warm_fields, cf, sizeof(double) * nf * c->mtot,
                              hipMemcpyHostToDevice, c->stream));
    launch_warm_count(P, c->warm_fields, nf, c->warm_cnt, c->stream);
    launch_scan(P, c->warm_cnt, c->mtot, c->warm_off, c->warm_off + c->mtot, c->stream);
    uint32_t total = 0;
    HIP_TRY(c, hipMemcpyAsync(&total, c->warm_off + c->mtot, sizeof total, hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    // a fill that does not fit leaves the context as it was: its particles, sums and counters
    if (total > P.cap) return fail(c, DSMC_E_CAPACITY, "warm start exceeds the pool capacity");
    if ((rc = reset_device_sums(c))) return rc;
    launch_warm_generate(P, other_buf(c), c->warm_fields, nf, c->warm_off, c->stream);
    launch_case_starts(P, c->warm_off, c->warm_starts, c->stream);
    return finish_fill(c, kKeyDropInBox, c->warm_starts, c->warm_off + c->mtot);
}

}  // namespace

extern "C" {

void dsmc_params_default(dsmc_params* p)
{
    std::memset(p, 0, sizeof *p);
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

int dsmc_create(const dsmc_params* p, int device, dsmc_ctx** out)
{
    return dsmc_create_batch(p, 1, device, out);
}

int dsmc_create_batch(const dsmc_params* cases, int ncases, int device, dsmc_ctx** out)
{
    return create_impl(cases, ncases, device, out);
}

void dsmc_destroy(dsmc_ctx* c)
{
    if (!c) return;
    if (c->stream) (void)hipSetDevice(c->device);
    invalidate_graph(c);
    for (hipEvent_t e : c->inflight)
        if (e) (void)hipEventDestroy(e);
    for (void* q : c->allocs) (void)hipFree(q);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* dsmc_last_error(const dsmc_ctx* c) { return c ? c->err.c_str() : "null context"; }
int dsmc_num_cases(const dsmc_ctx* c) { return c ? c->ncases : 0; }

int dsmc_get_params(dsmc_ctx* c, int i, dsmc_params* out)
{
    if (int rc = check_case(c, i)) return rc;
    *out = c->params[i];
    return 0;
}

int dsmc_get_constants(dsmc_ctx* c, int i, dsmc_constants* out)
{
    if (int rc = check_case(c, i)) return rc;
    *out = c->k[i];
    return 0;
}

int dsmc_get_state(dsmc_ctx* c, int i, dsmc_state* s)
{
    if (int rc = check_case(c, i)) return rc;
    if (int rc = check_ctx(c, false)) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const CaseCounters& q = c->P.cc;
    uint32_t ps[2] = {0, 0};
    unsigned long long nc = 0, nct = 0, hits = 0, inj = 0, rem = 0, pst = 0;
    double drag = 0.0;
    HIP_TRY(c, hipMemcpy(ps, c->P.case_pstart + i, sizeof ps, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(&nc, q.ncoll_last + i, sizeof nc, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(&nct, q.ncoll_total + i, sizeof nct, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(&hits, q.hits + i, sizeof hits, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(&inj, q.injected_last + i, sizeof inj, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(&rem, q.removed_last + i, sizeof rem, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(&drag, q.drag + i, sizeof drag, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(&pst, q.psteps + i, sizeof pst, hipMemcpyDeviceToHost));
    std::memset(s, 0, sizeof *s);
    s->step = c->step;
    s->np = (int64_t)ps[1] - (int64_t)ps[0];
    s->sample_step = c->sample_step[i];
    s->last_collisions = (int64_t)nc;
    s->total_collisions = (int64_t)nct;
    s->last_injected = (int64_t)inj;
    s->last_removed = (int64_t)rem;
    s->wall_hits = (int64_t)hits;
    s->drag_impulse_total = drag;
    s->drag_sampling_time = c->drag_time[i];
    s->time = c->time[i];
    s->sample_time = c->sample_time[i];
    s->particle_steps = (int64_t)pst;
    drag_report(c->params[i], c->k[i], s);
    return 0;
}

// InitialConditions (src/Initialization.cpp:290-313) for every case: uniform fill of the
// domain (count by the rule of src/Boundary.cpp:62-71), removal of the box interior, then an
// index pass that leaves the pool sorted by cell.
int dsmc_init_freestream(dsmc_ctx* c)
{
    if (int rc = check_ctx(c, false)) return rc;
    invalidate_graph(c);
    std::vector<uint32_t> base(c->ncases + 1, 0);
    uint64_t total = 0;
    for (int i = 0; i < c->ncases; ++i) {
        const CaseDev& h = c->hcases[i];
        const double r[4] = {h.x1, h.y1, h.x2, h.y2};
        const double avg = rect_avg(c->params[i], c->k[i], r);
        Stream rs(h.key0, h.key1, 0u, 0u, kPInitCount);
        base[i] = (uint32_t)total;
        total += new_count(avg, rs);
    }
    base[c->ncases] = (uint32_t)total;
    if (total > c->P.cap) return fail(c, DSMC_E_CAPACITY, "initial fill exceeds the pool capacity");
    HIP_TRY(c, hipMemcpyAsync(c->init_base, base.data(), sizeof(uint32_t) * (c->ncases + 1),
                              hipMemcpyHostToDevice, c->stream));
    if (int rc = reset_device_sums(c)) return rc;
    launch_init_generate(c->P, other_buf(c), c->init_base, (uint32_t)total, c->stream);
    return finish_fill(c, kKeyDropInBox, c->init_base, c->init_base + c->ncases);
}

// Warm start: per-case fields of w*h cells (4 planes: n, ux, uy, T) resampled bilinearly onto
// each case grid, then every cell seeded from its local Maxwellian.
int dsmc_init_from_fields(dsmc_ctx* c, const float* f, int w, int h)
{
    if (int rc = check_ctx(c, false)) return rc;
    if (!f || w <= 0 || h <= 0) return fail(c, DSMC_E_ARGUMENT, "bad field array");
    std::vector<double> cf(4 * (size_t)c->mtot);
    const size_t plane = (size_t)w * h;
    for (int i = 0; i < c->ncases; ++i) {
        const dsmc_params& p = c->params[i];
        const float* src = f + (size_t)i * 4 * plane;
        for (int kx = 0; kx < p.nx; ++kx)
            for (int ly = 0; ly < p.ny; ++ly) {
                // cell centre in field pixel coordinates (pixel (a, b) is x index a, y index b)
                const double fx = ((kx + 0.5) / p.nx) * w - 0.5;
                const double fy = ((ly + 0.5) / p.ny) * h - 0.5;
                int a0 = (int)std::floor(fx), b0 = (int)std::floor(fy);
                double ta = fx - a0, tb = fy - b0;
                int a1 = std::min(a0 + 1, w - 1), b1 = std::min(b0 + 1, h - 1);
                if (a0 < 0) { a0 = 0; ta = 0.0; }
                if (b0 < 0) { b0 = 0; tb = 0.0; }
                const size_t cell = c->cell_base[i] + (size_t)kx * p.ny + ly;
                for (int qn = 0; qn < 4; ++qn) {
                    const float* pl = src + qn * plane;
                    const double v00 = pl[(size_t)a0 * h + b0], v01 = pl[(size_t)a0 * h + b1];
                    const double v10 = pl[(size_t)a1 * h + b0], v11 = pl[(size_t)a1 * h + b1];
                    cf[4 * cell + qn] = (1 - ta) * ((1 - tb) * v00 + tb * v01)
                                        + ta * ((1 - tb) * v10 + tb * v11);
                }
            }
    }
    return dsmc_init_from_cell_fields(c, cf.data());
}

// Seeds every cell from a local Maxwellian given per cell (n, ux, uy, T; k*NY+l order per case,
// cases concatenated): count by the rule of src/Boundary.cpp:62-71, positions uniform in the
// cell, particles inside the box dropped; then an index pass.
int dsmc_init_from_cell_fields(dsmc_ctx* c, const double* cf)
{
    return init_from_cells(c, cf, 4);
}

int dsmc_init_from_cell_moments(dsmc_ctx* c, const double* cm)
{
    return init_from_cells(c, cm, 6);
}

int dsmc_set_step(dsmc_ctx* c, int64_t step)
{
    if (int rc = check_ctx(c, false)) return rc;
    if (step < 0 || step > 0xFFFFFFFFll) return fail(c, DSMC_E_ARGUMENT, "step out of range");
    uint32_t s = (uint32_t)step;
    HIP_TRY(c, hipMemcpyAsync(&c->P.st->step, &s, sizeof s, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->step = step;
    c->primed = false;
    return 0;
}

// MoveParticles (src/Boundary.cpp:284-328).
int dsmc_move(dsmc_ctx* c)
{
    if (int rc = check_ctx(c, true)) return rc;
    if (int rc = check_single(c)) return rc;
    if (c->fed_on) {
        launch_move_fed(c->P, cur_buf(c), c->fed, c->stream);
        c->indexed = false;
        c->perm_valid = false;
        if (int rc = sync_check(c)) return rc;
        return fed_check(c);
    }
    launch_clear_hits(c->P, c->stream);
    launch_move(c->P, cur_buf(c), false, c->stream);
    launch_extra(c->P, cur_buf(c), false, false, c->stream);
    launch_clear_hits(c->P, c->stream);
    c->indexed = false;
    c->perm_valid = false;
    return sync_check(c);
}

// BoundaryConditions (src/Boundary.cpp:94-203): inject the four strips, then remove what left
// the domain. The pool keeps its order: surviving old particles, then surviving new ones.
int dsmc_boundary(dsmc_ctx* c)
{
    if (int rc = check_ctx(c, true)) return rc;
    if (int rc = check_single(c)) return rc;
    if (c->fed_on) {
        launch_inject_fed(c->P, cur_buf(c), c->fed, c->stream);
        if (int rc = sync_check(c)) return rc;
        if (int rc = fed_check(c)) return rc;
    } else {
        launch_step_begin(c->P, c->stream);
        launch_extra(c->P, cur_buf(c), false, true, c->stream);
    }
    DevState st;
    HIP_TRY(c, hipMemcpyAsync(&st, c->P.st, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (st.err_code) return sync_check(c);
    const uint32_t n = st.np + st.ncand;
    const uint32_t np_before = st.np;
    launch_compact(c->P, n, cur_buf(c), other_buf(c), c->bcnt, c->stream);
    c->cur = 1 - c->cur;
    c->indexed = false;
    c->perm_valid = false;
    c->primed = false;
    uint32_t np = 0;
    HIP_TRY(c, hipMemcpyAsync(&np, &c->P.st->np, sizeof np, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const uint32_t ps[2] = {0, np};
    HIP_TRY(c, hipMemcpy(c->P.case_pstart, ps, sizeof ps, hipMemcpyHostToDevice));
    // removed = old + injected - kept
    unsigned long long inj = 0;
    HIP_TRY(c, hipMemcpy(&inj, c->P.cc.injected, sizeof inj, hipMemcpyDeviceToHost));
    const unsigned long long rem = (unsigned long long)np_before + inj - np;
    HIP_TRY(c, hipMemcpy(c->P.cc.removed, &rem, sizeof rem, hipMemcpyHostToDevice));
    if (int rc = publish_stage_counters(c, false, true)) return rc;
    return sync_check(c);
}

// Indexing (src/Geometry.cpp:37-170): cell keys, histogram, scan, and the stable sort of the
// pool into k*NY+l order.
int dsmc_index(dsmc_ctx* c)
{
    if (int rc = check_ctx(c, true)) return rc;
    if (int rc = check_single(c)) return rc;
    launch_keys(c->P, cur_buf(c), kKeyStrict, c->P.case_pstart, &c->P.st->np, c->stream);
    sort_pool(c, c->perm_sorted);
    c->primed = false;
    if (int rc = sync_check(c)) return rc;
    c->indexed = true;
    c->perm_valid = true;
    return 0;
}

// CollideParticles (src/Collision.cpp:45-56).
int dsmc_collide(dsmc_ctx* c, int64_t* ncoll)
{
    if (int rc = check_ctx(c, true)) return rc;
    if (!c->indexed) return fail(c, DSMC_E_ARGUMENT, "dsmc_collide needs a preceding dsmc_index");
    const size_t b = sizeof(unsigned long long) * c->ncases;
    HIP_TRY(c, hipMemsetAsync(c->P.cc.ncoll, 0, b, c->stream));
    if (c->fed_on) {
        if (int rc = check_single(c)) return rc;
        launch_collide_fed(c->P, cur_buf(c), c->fed, c->stream);
    } else {
        launch_collide(c->P, cur_buf(c), c->stream);
    }
    if (int rc = publish_stage_counters(c, true, false)) return rc;
    if (int rc = sync_check(c)) return rc;
    if (c->fed_on)
        if (int rc = fed_check(c)) return rc;
    std::vector<unsigned long long> v(c->ncases), t(c->ncases);
    HIP_TRY(c, hipMemcpy(v.data(), c->P.cc.ncoll, b, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(t.data(), c->P.cc.ncoll_total, b, hipMemcpyDeviceToHost));
    unsigned long long total = 0;
    for (int i = 0; i < c->ncases; ++i) {
        total += v[i];
        t[i] += v[i];
    }
    HIP_TRY(c, hipMemcpy(c->P.cc.ncoll_total, t.data(), b, hipMemcpyHostToDevice));
    c->primed = false;
    if (ncoll) *ncoll = (int64_t)total;
    return 0;
}

// Fed draws: the single-stage move, boundary and collide calls of a one-case context take their
// uniforms from u[0..n) in the reference's order instead of the counter-based streams.
int dsmc_set_draws(dsmc_ctx* c, const double* u, size_t n)
{
    if (int rc = check_ctx(c, false)) return rc;
    if (int rc = check_single(c)) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (n == 0) {
        c->fed_on = false;
        return 0;
    }
    if (!u) return fail(c, DSMC_E_ARGUMENT, "null draws");
    for (size_t i = 0; i < n; ++i)
        if (!(u[i] >= 0.0 && u[i] < 1.0))
            return fail(c, DSMC_E_ARGUMENT, "fed draws must lie in [0, 1)");
    if (!c->fed)
        if (int rc = dalloc(c, &c->fed, 1)) return rc;
    if (n > c->fed_cap) {
        if (c->fed_u) {
            for (auto& a : c->allocs)
                if (a == c->fed_u) a = nullptr;
            (void)hipFree(c->fed_u);
            c->fed_u = nullptr;
            c->fed_cap = 0;
        }
        if (int rc = dalloc(c, &c->fed_u, n)) return rc;
        c->fed_cap = n;
    }
    HIP_TRY(c, hipMemcpy(c->fed_u, u, n * sizeof(double), hipMemcpyHostToDevice));
    const FedStream fs{c->fed_u, (uint64_t)n, 0, false};
    HIP_TRY(c, hipMemcpy(c->fed, &fs, sizeof fs, hipMemcpyHostToDevice));
    c->fed_on = true;
    return 0;
}

int dsmc_draws_used(dsmc_ctx* c, size_t* used)
{
    if (int rc = check_ctx(c, false)) return rc;
    if (!used) return fail(c, DSMC_E_ARGUMENT, "null output");
    *used = 0;
    if (!c->fed_on) return 0;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    FedStream fs;
    HIP_TRY(c, hipMemcpy(&fs, c->fed, sizeof fs, hipMemcpyDeviceToHost));
    *used = (size_t)fs.pos;
    return 0;
}

// Sampling (src/Sampling.cpp:14-31).
int dsmc_sample(dsmc_ctx* c)
{
    if (int rc = check_ctx(c, true)) return rc;
    if (!c->indexed) return fail(c, DSMC_E_ARGUMENT, "dsmc_sample needs a preceding dsmc_index");
    launch_sample(c->P, cur_buf(c), true, c->stream);
    for (int i = 0; i < c->ncases; ++i) {
        if (c->hcases[i].retired) continue;  // no particles: nothing was added
        c->sample_step[i] += 1;
        c->sample_time[i] += c->params[i].dt;
    }
    return sync_check(c);
}

// nsteps iterations of the loop of src/Simulation.cpp:138-161.
int dsmc_step(dsmc_ctx* c, int nsteps)
{
    if (int rc = check_ctx(c, true)) return rc;
    if (nsteps < 0) return fail(c, DSMC_E_ARGUMENT, "negative step count");
    if (int rc = prime(c)) return rc;
    int s = 0;
    // Pairs of steps replay a captured graph (A->B->A, or B->A->B from the other buffer); a
    // leftover step runs eagerly.
    if (nsteps >= 2) {
        hipGraphExec_t& graph = c->graph[c->cur];
        if (!graph) {
            hipGraph_t g = nullptr;
            HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            const int cur0 = c->cur;
            enqueue_fused_step(c);
            enqueue_fused_step(c);
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            c->cur = cur0;
            if (e != hipSuccess) return fail(c, DSMC_E_HIP, "graph capture failed");
            e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                graph = nullptr;
                return fail(c, DSMC_E_HIP, "graph instantiate failed");
            }
        }
        if (!c->inflight[0]) {
            HIP_TRY(c, hipEventCreateWithFlags(&c->inflight[0], hipEventDisableTiming));
            HIP_TRY(c, hipEventCreateWithFlags(&c->inflight[1], hipEventDisableTiming));
        }
        // At most 2 * kPairsInFlight replays stay queued: every kPairsInFlight replays an event
        // is recorded, and the host first waits for the one recorded two marks earlier. The GPU
        // never runs dry, and the queued dispatches (each carries a completion record under a
        // kernel-trace profiler) stay bounded however long the run.
        constexpr int kPairsInFlight = 128;
        int mark = 0;
        for (int pair = 0; s + 2 <= nsteps; s += 2, ++pair) {
            if (pair > 0 && pair % kPairsInFlight == 0) {
                hipEvent_t ev = c->inflight[mark & 1];
                if (mark >= 2) HIP_TRY(c, hipEventSynchronize(ev));
                HIP_TRY(c, hipEventRecord(ev, c->stream));
                ++mark;
            }
            HIP_TRY(c, hipGraphLaunch(graph, c->stream));
            account_step(c);
            account_step(c);
        }
    }
    for (; s < nsteps; ++s) {
        enqueue_fused_step(c);
        account_step(c);
    }
    if (nsteps > 0) close_fused_steps(c);
    c->indexed = true;
    c->perm_valid = false;
    return sync_check(c);
}

// Stops case i of a batch from the next step on (its sums and clocks keep their values). The
// other cases' random streams are keyed by case-local indices, so their results do not change.
int dsmc_retire_case(dsmc_ctx* c, int i)
{
    if (int rc = check_ctx(c, true)) return rc;
    if (int rc = check_case(c, i)) return rc;
    if (c->hcases[i].retired) return 0;
    c->hcases[i].retired = 1;
    if (int rc = upload_retired(c, i)) return rc;
    return resize_for_running_cases(c);
}

// Eager fused steps; each kernel carries a start/stop event pair from its own dispatch, and
// each stage's time is the sum of its kernels' times (no launch gaps).
int dsmc_profile_steps(dsmc_ctx* c, int nsteps, double* stage_ms)
{
    if (int rc = check_ctx(c, true)) return rc;
    if (nsteps <= 0 || !stage_ms) return fail(c, DSMC_E_ARGUMENT, "bad profile arguments");
    double acc[DSMC_NUM_STAGES] = {0};
    int rc = 0;
    {
        LaunchTimer timer;
        set_launch_timer(&timer);
        rc = prime(c);
        for (int s = 0; s < nsteps && !rc; ++s) {
            hipStream_t q = c->stream;
            enqueue_fused_step(c, &timer);
            if (hipStreamSynchronize(q) != hipSuccess) { rc = DSMC_E_HIP; break; }
            timer.harvest(acc);
            account_step(c);
        }
        set_launch_timer(nullptr);
    }
    close_fused_steps(c);
    if (rc) return rc == DSMC_E_HIP ? fail(c, rc, "stream synchronisation failed") : rc;
    for (int i = 0; i < DSMC_NUM_STAGES; ++i) stage_ms[i] = acc[i] / nsteps;
    c->indexed = true;
    c->perm_valid = false;
    return sync_check(c);
}

int dsmc_get_tuning(dsmc_ctx* c, dsmc_tuning* out)
{
    if (int rc = check_ctx(c, false)) return rc;
    if (!out) return fail(c, DSMC_E_ARGUMENT, "null tuning");
    *out = c->tuning;
    return 0;
}

int dsmc_set_tuning(dsmc_ctx* c, const dsmc_tuning* t)
{
    if (int rc = check_ctx(c, false)) return rc;
    if (!t) return fail(c, DSMC_E_ARGUMENT, "null tuning");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return apply_tuning(c, *t);
}

// Every admissible (schedule, chunk) pair at one particle per thread, then 2 and 4 particles
// per thread for the fastest large-schedule pair, runs one graph-warming pair of steps, then
// nsteps timed steps between two events on the stream; the fastest stays applied.
int dsmc_autotune(dsmc_ctx* c, int nsteps, dsmc_tuning* best)
{
    if (int rc = check_ctx(c, true)) return rc;
    if (nsteps < 2) return fail(c, DSMC_E_ARGUMENT, "autotune needs at least 2 steps per candidate");
    std::vector<dsmc_tuning> cand;
    for (int sched = DSMC_SCHEDULE_LARGE; sched <= DSMC_SCHEDULE_SMALL; ++sched) {
        if (sched == DSMC_SCHEDULE_SMALL && c->mtot > kSmallCells) continue;
        for (int chunk = 128; chunk <= 2048; chunk *= 2)
            if (gather_fits(c, chunk)) cand.push_back(dsmc_tuning{sched, chunk, 1});
    }
    const size_t n_first = cand.size();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(c, hipEventCreate(&e0));
    HIP_TRY(c, hipEventCreate(&e1));
    dsmc_tuning pick = c->tuning;
    dsmc_tuning best_large{-1, 0, 1};
    float best_ms = -1.f, best_large_ms = -1.f;
    int rc = 0;
    for (size_t i = 0; i < cand.size(); ++i) {
        const dsmc_tuning t = cand[i];
        if ((rc = dsmc_set_tuning(c, &t))) break;
        if ((rc = dsmc_step(c, 2))) break;
        if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = DSMC_E_HIP; break; }
        if ((rc = dsmc_step(c, nsteps))) break;
        if (hipEventRecord(e1, c->stream) != hipSuccess
            || hipEventSynchronize(e1) != hipSuccess) { rc = DSMC_E_HIP; break; }
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = DSMC_E_HIP; break; }
        if (best_ms < 0.f || ms < best_ms) { best_ms = ms; pick = t; }
        if (t.schedule == DSMC_SCHEDULE_LARGE && (best_large_ms < 0.f || ms < best_large_ms)) {
            best_large_ms = ms;
            best_large = t;
        }
        // second round: the particles per thread of the best large-schedule pair
        if (i + 1 == n_first && best_large.schedule == DSMC_SCHEDULE_LARGE) {
            for (int ppt : {2, 4})
                cand.push_back(dsmc_tuning{best_large.schedule, best_large.chunk, ppt});
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc == DSMC_E_HIP ? fail(c, rc, "autotune timing failed") : rc;
    if ((rc = dsmc_set_tuning(c, &pick))) return rc;
    if (best) *best = pick;
    return 0;
}

int dsmc_get_num_particles(dsmc_ctx* c, int i, size_t* n)
{
    if (int rc = check_case(c, i)) return rc;
    if (int rc = check_ctx(c, false)) return rc;
    uint32_t ps[2] = {0, 0};
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(ps, c->P.case_pstart + i, sizeof ps, hipMemcpyDeviceToHost));
    *n = ps[1] - ps[0];
    return 0;
}

int dsmc_get_particles(dsmc_ctx* c, int i, double* x, double* y, float* vx, float* vy, float* vz,
                       size_t cap, size_t* n)
{
    if (int rc = check_case(c, i)) return rc;
    if (int rc = check_ctx(c, false)) return rc;
    uint32_t ps[2] = {0, 0};
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(ps, c->P.case_pstart + i, sizeof ps, hipMemcpyDeviceToHost));
    const size_t cnt = ps[1] - ps[0];
    if (n) *n = cnt;
    if (cap < cnt) return fail(c, DSMC_E_ARGUMENT, "output arrays too small");
    const Particles b = cur_buf(c);
    const size_t o = ps[0];
    // the pairs are split into separate arrays in the other buffer, which holds nothing
    // between calls (every stage writes it, then makes it current)
    const Particles t = other_buf(c);
    double* tx = reinterpret_cast<double*>(t.xy);
    float* tv = reinterpret_cast<float*>(t.vxy);
    if (x || y || vx || vy) {
        launch_unpack(b, (uint32_t)o, (uint32_t)cnt, tx, tx + cnt, tv, tv + cnt, c->stream);
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    if (x) HIP_TRY(c, hipMemcpy(x, tx, cnt * sizeof(double), hipMemcpyDeviceToHost));
    if (y) HIP_TRY(c, hipMemcpy(y, tx + cnt, cnt * sizeof(double), hipMemcpyDeviceToHost));
    if (vx) HIP_TRY(c, hipMemcpy(vx, tv, cnt * sizeof(float), hipMemcpyDeviceToHost));
    if (vy) HIP_TRY(c, hipMemcpy(vy, tv + cnt, cnt * sizeof(float), hipMemcpyDeviceToHost));
    if (vz) HIP_TRY(c, hipMemcpy(vz, b.vz + o, cnt * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int dsmc_set_particles(dsmc_ctx* c, int i, const double* x, const double* y, const float* vx,
                       const float* vy, const float* vz, size_t n)
{
    if (int rc = check_case(c, i)) return rc;
    if (int rc = check_ctx(c, false)) return rc;
    if (int rc = check_single(c)) return rc;
    if (n + c->P.ncand_max > c->P.cap) return fail(c, DSMC_E_CAPACITY, "too many particles");
    invalidate_graph(c);
    const Particles b = cur_buf(c);
    if (n) {
        // staged as separate arrays in the other buffer, then paired into the current one
        const Particles t = other_buf(c);
        double* tx = reinterpret_cast<double*>(t.xy);
        float* tv = reinterpret_cast<float*>(t.vxy);
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        HIP_TRY(c, hipMemcpy(tx, x, n * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(tx + n, y, n * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(tv, vx, n * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(tv + n, vy, n * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(b.vz, vz, n * sizeof(float), hipMemcpyHostToDevice));
        launch_pack(b, (uint32_t)n, tx, tx + n, tv, tv + n, c->stream);
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    uint32_t nn = (uint32_t)n;
    uint32_t ps[2] = {0, nn};
    HIP_TRY(c, hipMemcpy(&c->P.st->np, &nn, sizeof nn, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->P.case_pstart, ps, sizeof ps, hipMemcpyHostToDevice));
    c->initialised = true;
    c->indexed = false;
    c->perm_valid = false;
    c->primed = false;
    return 0;
}

int dsmc_get_index(dsmc_ctx* c, int i, uint32_t* counts, size_t ncells, uint32_t* perm, size_t n)
{
    if (int rc = check_case(c, i)) return rc;
    if (int rc = check_ctx(c, false)) return rc;
    if (!c->indexed) return fail(c, DSMC_E_ARGUMENT, "no index: call dsmc_index first");
    const uint32_t m = (uint32_t)c->params[i].nx * (uint32_t)c->params[i].ny;
    if (ncells < m) return fail(c, DSMC_E_ARGUMENT, "counts array too small");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::vector<uint32_t> off(m + 1);
    HIP_TRY(c, hipMemcpy(off.data(), c->P.off + c->cell_base[i], sizeof(uint32_t) * (m + 1),
                         hipMemcpyDeviceToHost));
    if (counts)
        for (uint32_t q = 0; q < m; ++q) counts[q] = off[q + 1] - off[q];
    const size_t cnt = off[m] - off[0];
    if (perm) {
        if (!c->perm_valid)
            return fail(c, DSMC_E_ARGUMENT, "the permutation is kept only by dsmc_index");
        if (n < cnt) return fail(c, DSMC_E_ARGUMENT, "perm array too small");
        HIP_TRY(c, hipMemcpy(perm, c->perm_sorted + off[0], cnt * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (size_t q = 0; q < cnt; ++q) perm[q] -= off[0];
    }
    return 0;
}

int dsmc_get_cells(dsmc_ctx* c, int i, double* cell17, size_t ncells)
{
    if (int rc = check_case(c, i)) return rc;
    if (int rc = check_ctx(c, false)) return rc;
    const uint32_t m = (uint32_t)c->params[i].nx * (uint32_t)c->params[i].ny;
    if (ncells < m) return fail(c, DSMC_E_ARGUMENT, "cell array too small");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::vector<double> rows((size_t)kSumRows * m);
    for (int r = 0; r < kSumRows; ++r)
        HIP_TRY(c, hipMemcpy(rows.data() + (size_t)r * m, c->P.sums + (size_t)r * c->mtot + c->cell_base[i],
                             sizeof(double) * m, hipMemcpyDeviceToHost));
    // device rows: NP, V0..2, VV00 VV01 VV02 VV11 VV12 VV22, V2V0..2, NC
    static const int vv[9] = {4, 5, 6, 5, 7, 8, 6, 8, 9};
    for (uint32_t q = 0; q < m; ++q) {
        double* o = cell17 + (size_t)q * 17;
        o[0] = rows[q];
        for (int a = 0; a < 3; ++a) o[1 + a] = rows[(size_t)(1 + a) * m + q];
        for (int a = 0; a < 9; ++a) o[4 + a] = rows[(size_t)vv[a] * m + q];
        for (int a = 0; a < 3; ++a) o[13 + a] = rows[(size_t)(10 + a) * m + q];
        o[16] = rows[(size_t)13 * m + q];
    }
    return 0;
}

// CELL sums of one case added over bx x by blocks of cells on the device; the block of cell
// (k, l) is (floor((k + 0.5) bx / NX), floor((l + 0.5) by / NY)).
int dsmc_get_binned_cells(dsmc_ctx* c, int i, int bx, int by, double* cell17)
{
    if (int rc = check_case(c, i)) return rc;
    if (int rc = check_ctx(c, false)) return rc;
    const dsmc_params& p = c->params[i];
    if (bx < 1 || by < 1 || bx > p.nx || by > p.ny || !cell17)
        return fail(c, DSMC_E_ARGUMENT, "bins must lie in [1, NX] x [1, NY]");
    const size_t nb = (size_t)bx * by;
    if (nb > c->bins_cap) {
        if (c->bins) {
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            for (auto& a : c->allocs)
                if (a == c->bins) a = nullptr;
            (void)hipFree(c->bins);
            c->bins = nullptr;
        }
        if (int rc = dalloc(c, &c->bins, nb * kSumRows)) return rc;
        c->bins_cap = nb;
    }
    launch_bin_sums(c->P, i, bx, by, c->bins, false, c->stream);
    std::vector<double> rows(nb * kSumRows);
    HIP_TRY(c, hipMemcpyAsync(rows.data(), c->bins, sizeof(double) * rows.size(),
                              hipMemcpyDeviceToHost, c->stream));
    if (int rc = sync_check(c)) return rc;
    static const int vv[9] = {4, 5, 6, 5, 7, 8, 6, 8, 9};
    for (size_t q = 0; q < nb; ++q) {
        const double* r = rows.data() + q * kSumRows;
        double* o = cell17 + q * 17;
        o[0] = r[0];
        for (int a = 0; a < 3; ++a) o[1 + a] = r[1 + a];
        for (int a = 0; a < 9; ++a) o[4 + a] = r[vv[a]];
        for (int a = 0; a < 3; ++a) o[13 + a] = r[10 + a];
        o[16] = r[13];
    }
    return 0;
}

// The block sums of dsmc_get_binned_cells without allocating: the kernel writes the CELL layout
// into this GPU's `scratch`, which is copied to the caller's array. Calls that overlap on one
// GPU (other threads, other contexts) share the array and race; each call writes every element
// it copies out, so nothing needs clearing between calls.
int dsmc_get_binned_cells_noalloc(dsmc_ctx* c, int i, int bx, int by, double* cell17)
{
    if (!c || !c->stream) return DSMC_E_ARGUMENT;
    if (i < 0 || i >= c->ncases) return fail_noalloc(c, DSMC_E_ARGUMENT, "case index out of range");
    const dsmc_params& p = c->params[i];
    if (bx < 1 || by < 1 || bx > p.nx || by > p.ny || !cell17)
        return fail_noalloc(c, DSMC_E_ARGUMENT, "bins must lie in [1, NX] x [1, NY]");
    if ((int64_t)bx * by > kNoAllocMaxBins)
        return fail_noalloc(c, DSMC_E_ARGUMENT,
                            "at most 16384 blocks (bx * by) without allocation; "
                            "dsmc_get_binned_cells takes more");
    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) {
        launch_bin_sums(c->P, i, bx, by, c->scratch, true, c->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(cell17, c->scratch, sizeof(double) * 17 * (size_t)bx * by,
                           hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    DevState st;
    if (e == hipSuccess) e = hipMemcpy(&st, c->P.st, sizeof st, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_noalloc(c, DSMC_E_HIP, hipGetErrorString(e));
    if (st.err_code) {
        // an error word left by an earlier step: report it as every synchronising call does
        uint32_t zero[2] = {0, 0};
        (void)hipMemcpy(&c->P.st->err_code, zero, sizeof zero, hipMemcpyHostToDevice);
        char msg[kErrCapacity];
        std::snprintf(msg, sizeof msg, "device error (code %u, index %u)", st.err_code,
                      st.err_index);
        c->indexed = false;
        c->primed = false;
        return fail_noalloc(c, (int)st.err_code, msg);
    }
    return 0;
}

// Printing's conversion (src/Sampling.cpp:152-196).
int dsmc_get_fields(dsmc_ctx* c, int i, double* out6, size_t ncells)
{
    if (int rc = check_case(c, i)) return rc;
    const uint32_t m = (uint32_t)c->params[i].nx * (uint32_t)c->params[i].ny;
    if (ncells < m) return fail(c, DSMC_E_ARGUMENT, "field array too small");
    std::vector<double> cells((size_t)17 * m);
    if (int rc = dsmc_get_cells(c, i, cells.data(), m)) return rc;
    const dsmc_constants& k = c->k[i];
    const int64_t ss = c->sample_step[i];
    for (uint32_t q = 0; q < m; ++q) {
        const double* s = cells.data() + (size_t)q * 17;
        double n = 0.0, u[3] = {0, 0, 0}, t[3] = {0, 0, 0};
        if (ss > 0) n = k.weight * s[0] / (double)ss / k.dv;
        if (s[0] > 0.0)
            for (int a = 0; a < 3; ++a) {
                u[a] = s[1 + a] / s[0];
                t[a] = k.molecule_mass * (s[4 + 4 * a] / s[0] - u[a] * u[a]) / kBoltzmann;
            }
        double* o = out6 + (size_t)q * 6;
        o[0] = n; o[1] = u[0]; o[2] = u[1]; o[3] = t[0]; o[4] = t[1];
        o[5] = (t[0] + t[1] + t[2]) / 3.0;
    }
    return 0;
}

}  // extern "C"
