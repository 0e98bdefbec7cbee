// dsmc_kernels.hip - gfx950 kernels of the DSMC per-step particle pipeline.
//
// One fused step (src/Simulation.cpp:138-199) of the large-pool schedule is these launches
// over one particle pool shared by every case:
//   k_step_open   closes the previous fused step (Step++, counters, pool size) and computes
//                 this step's inflow counts
//   k_move        MoveParticles' free flight (src/Boundary.cpp:284-299) for every particle
//                 whose segment cannot reach the box; the rest go to a hit list. With keys:
//                 the outflow removal (src/Boundary.cpp:181-202), the cell key
//                 (src/Geometry.cpp:47-59) and a wave-aggregated histogram.
//   k_extra       the hit list through IntersectBox and the wall model (src/Boundary.cpp:
//                 301-326, src/Geometry.cpp:179-229), and GenerateNewParticles for the four
//                 inflow strips (src/Boundary.cpp:60-173)
//   k_scan_cells  scan of the cell counts: the CSR offsets of Indexing's IPC lists, scatter
//                 cursors and the gather chunk table (three launches on large grids)
//   k_scatter     arrival-order scatter of pool positions into their cell segments
//   k_gather      per chunk of cells: the stable order of Indexing's IPC append
//                 (src/Geometry.cpp:168, a rank by pool position inside each segment, in LDS),
//                 the particle gather, then per cell the NTC candidate draw of
//                 CollideParticles (src/Collision.cpp:45-56) and Sampling
//                 (src/Sampling.cpp:14-31) of the cells without candidates
//   k_collide_list  the cells with candidates: collisions (src/Collision.cpp:7-44), then
//                 their sampling
// The small-pool schedule folds these into k_move_small, k_scan_scatter_small and
// k_gather_small (which collides in LDS). Kernels loop over device-side counts: a step needs no
// host synchronisation and is captured in a hipGraph.
#include "dsmc_kernels.h"

#include <hip/hip_ext.h>

#include <type_traits>

namespace dsmc {

namespace {

LaunchTimer* g_timer = nullptr;

// Every launch of this file goes through here: with a timer attached, the kernel is launched
// with a start/stop event pair taken from the dispatch itself.
template <typename F, typename... Args>
inline void launch_k(F kernel, dim3 grid, dim3 block, size_t shm, hipStream_t s, Args... args)
{
    if (g_timer) {
        hipEvent_t* ev = g_timer->acquire();
        hipExtLaunchKernelGGL(kernel, grid, block, (uint32_t)shm, s, ev[0], ev[1], 0u, args...);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shm, s, args...);
    }
}

constexpr int kNT = 256;        // threads of the particle and cell kernels
constexpr int kScanNT = 1024;   // threads of the scan kernels
constexpr int kScanItems = 4;   // items per scan thread
constexpr int kScanTile = kScanNT * kScanItems;
static_assert(kScanTile == kScanTileCells, "scan tile");

__device__ inline void raise_error(DevState* st, uint32_t code, uint32_t index)
{
    if (atomicCAS(&st->err_code, 0u, code) == 0u) st->err_index = index;
}

// Largest c in [0, n) with starts[c] <= v (starts ascending, starts[0] == 0).
__device__ inline int find_case(const uint32_t* starts, int n, uint32_t v)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// find_case for the lanes of a wave that ask about ascending, nearby values (consecutive pool
// positions or cells): one binary search for the wave's first active value, on wave-uniform
// (scalar) operands, then a short forward walk per lane that only runs where the wave
// straddles a case boundary. Every lane of the wave must call it; `active` lanes get a result.
__device__ inline int find_case_wave(const uint32_t* starts, int n, uint32_t v, bool active)
{
    const uint64_t m = __ballot(active);
    if (!m) return 0;
    const int first = __ffsll((unsigned long long)m) - 1;
    const uint32_t v0 = (uint32_t)__builtin_amdgcn_readlane((int)v, first);
    int c = find_case(starts, n, v0);
    if (active)
        while (c + 1 < n && starts[c + 1] <= v) ++c;
    return c;
}

// The case starts of a pool of at most 65 cases held one per lane for find_case_full_wave:
// lane i holds starts[i + 1], all ones past the last case. Loaded once per kernel, before its
// loop, by all 64 lanes of the wave; unused (all ones) in larger pools.
__device__ inline uint32_t lane_case_starts(const uint32_t* starts, int n)
{
    const int lane = threadIdx.x & 63;
    return (n <= 65 && lane + 1 < n) ? starts[lane + 1] : 0xFFFFFFFFu;
}

// find_case_wave for a wave whose 64 lanes all execute the call (a loop with a wave-uniform
// bound) on values at or above the first active lane's: with at most 65 cases, whose starts
// the lanes hold (ls, from lane_case_starts), one ballot counts the starts at or below the
// wave's first value, and a wave that straddles case starts takes them by readlane, with no
// memory access. Larger pools search the table in memory.
__device__ inline int find_case_full_wave(const uint32_t* starts, int n, uint32_t ls, uint32_t v,
                                          bool active)
{
    if (n > 65) return find_case_wave(starts, n, v, active);
    const uint64_t m = __ballot(active);
    if (!m) return 0;
    const int first = __ffsll((unsigned long long)m) - 1;
    const uint32_t v0 = (uint32_t)__builtin_amdgcn_readlane((int)v, first);
    int c = __popcll(__ballot(ls <= v0));
    // cases past the wave's first: starts[k] is held by lane k - 1
    for (int k = c + 1; k < n; ++k) {
        const uint32_t sk = (uint32_t)__builtin_amdgcn_readlane((int)ls, k - 1);
        const bool past = active && v >= sk;
        if (__ballot(past) == 0) break;
        if (past) c = k;
    }
    return c;
}

// Calls f(c, uniform) once for the wave. When every active lane holds the same case index,
// c is that index read from the first lane: an SGPR value, so the case table (a __restrict__
// kernel argument) is read with scalar loads ahead of the lanes' own loads, and uniform is
// std::true_type. Otherwise c is each lane's own index and uniform is std::false_type. The
// test is wave-uniform, so f runs on the whole wave either way; every lane must call it.
template <class F>
__device__ __forceinline__ void with_case(int c, bool active, F&& f)
{
    const int cu = __builtin_amdgcn_readfirstlane(c);
    if (__ballot(active && c != cu) == 0) f(cu, std::true_type{});
    else f(c, std::false_type{});
}

__device__ inline bool sampling_step(const CaseDev& c, uint32_t step)
{
    return (int64_t)step > c.first_sampling_step && (int64_t)step % c.sampling_period == 0;
}

__device__ inline uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan over a block of NT threads; `sm` holds NT/64 + 1 words.
template <int NT>
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* sm, uint32_t& total)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) sm[w] = inc;
    __syncthreads();
    if (w == 0) {
        uint32_t s = lane < NW ? sm[lane] : 0u;
        uint32_t si = wave_incl_scan(s);
        if (lane < NW) sm[lane] = si - s;
        if (lane == NW - 1) sm[NW] = si;
    }
    __syncthreads();
    uint32_t res = inc - v + sm[w];
    total = sm[NW];
    __syncthreads();
    return res;
}

// Groups the valid lanes of a wave by key with ballots only (no memory traffic): each lane
// learns its group's leader (lowest lane), its rank in the group, and the leader learns the
// group's size. One pass per distinct key: after a move a wave's keys interleave (particles
// that stayed in a cell between those that crossed to a neighbour), so grouping by key, not
// by runs of equal neighbours, is what keeps the atomics to one per distinct key.
struct WaveGroup {
    int leader;
    uint32_t rank;
    uint32_t size;
};

// Each pass takes the lowest pending lane's key and keeps, on the lanes holding it, the
// group's lane mask (one readlane, one compare and a 64-bit select); leader, rank and size
// follow from that mask once, after the loop. A lane holding the pass's key is always still
// pending: every lane of a key joins its group in the pass that first meets the key.
__device__ inline WaveGroup wave_group(uint32_t key, bool valid)
{
    const int lane = threadIdx.x & 63;
    const uint64_t vmask = __ballot(valid);
    uint64_t pending = vmask;
    uint64_t mine = 0ull;
    while (pending) {
        const int l = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lkey = (uint32_t)__builtin_amdgcn_readlane((int)key, l);
        // the group is the wave compare itself, masked with the valid lanes: written this way
        // the compare's lane mask is the ballot (six vector instructions per pass instead of
        // eight). An invalid lane holding the key may take the mask; it never uses it.
        const uint64_t grp = __ballot(key == lkey) & vmask;
        mine = key == lkey ? grp : mine;
        pending &= ~grp;
    }
    WaveGroup g{lane, 0u, 0u};
    if (valid) {
        g.leader = __ffsll((unsigned long long)mine) - 1;
        g.rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mine >> 32),
                                           __builtin_amdgcn_mbcnt_lo((uint32_t)mine, 0u));
        g.size = (uint32_t)__popcll(mine);
    }
    return g;
}

// Wave-aggregated histogram add: one atomic per distinct key, all issued by one instruction.
__device__ inline void wave_hist_add(uint32_t* cnt, uint32_t key, bool valid)
{
    const WaveGroup g = wave_group(key, valid);
    if (valid && g.leader == (int)(threadIdx.x & 63)) atomicAdd(&cnt[key], g.size);
}

// Wave-aggregated claim of consecutive cursor positions per key, one atomic round trip per
// wave: the group leaders' atomics go out together and each group's base is shuffled back.
__device__ inline uint32_t wave_claim(uint32_t* cursor, uint32_t key, bool valid)
{
    const WaveGroup g = wave_group(key, valid);
    uint32_t base = 0;
    if (valid && g.leader == (int)(threadIdx.x & 63)) base = atomicAdd(&cursor[key], g.size);
    base = (uint32_t)__shfl((int)base, g.leader, 64);
    return base + g.rank;
}

// Wave-aggregated append to a list guarded by a counter.
__device__ inline void wave_append(uint32_t* counter, uint32_t* list, uint32_t value, bool want)
{
    const int lane = threadIdx.x & 63;
    const uint64_t m = __ballot(want);
    if (!m) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
    if (want) list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = value;
}

// Per-case counter add: one atomic per distinct case among the wave's lanes with `pred` (the
// lanes of a wave nearly always share their case; the per-case counters of a batch share a few
// cache lines, so one atomic per particle would queue on one L2 channel). Only lanes in the
// ballot are read, so it is safe in divergent code.
template <bool UNIFORM = false>
__device__ inline void count_add(const Pool& P, unsigned long long* ctr, int c, bool pred)
{
    const int lane = threadIdx.x & 63;
    if (UNIFORM || P.ncases == 1) {  // c is the same on every lane
        const uint64_t m = __ballot(pred);
        if (m && lane == __ffsll((unsigned long long)m) - 1)
            atomicAdd(&ctr[UNIFORM ? c : 0], (unsigned long long)__popcll(m));
        return;
    }
    uint64_t pending = __ballot(pred);
    while (pending) {
        const int l = __ffsll((unsigned long long)pending) - 1;
        const int lc = __builtin_amdgcn_readlane(c, l);
        const uint64_t grp = __ballot(pred && c == lc) & pending;
        if (lane == l) atomicAdd(&ctr[lc], (unsigned long long)__popcll(grp));
        pending &= ~grp;
    }
}

// A wall impact's drag impulse and hit count (DragImpulseTotal, DragCollisionCount) added to
// its case, called by the impacting lanes only: per distinct case among them one atomic each,
// the impulses summed in lane order.
__device__ inline void impact_add(const Pool& P, int c, double impulse)
{
    const int lane = threadIdx.x & 63;
    uint64_t pending = __ballot(true);
    while (pending) {
        const int l = __ffsll((unsigned long long)pending) - 1;
        const int lc = __builtin_amdgcn_readlane(c, l);
        const uint64_t grp = __ballot(c == lc) & pending;
        double sum = 0.0;
        for (uint64_t m = grp; m; m &= m - 1ull)
            sum += __shfl(impulse, __ffsll((unsigned long long)m) - 1, 64);
        if (lane == l) {
            atomicAdd(&P.cc.drag[lc], sum);
            atomicAdd(&P.cc.hits[lc], (unsigned long long)__popcll(grp));
        }
        pending &= ~grp;
    }
}

// Key of a particle at its post-move position: removed (sentinel) if it left the domain
// through the closed outflow test or its case is retired, code 20 if inside the domain but
// off the grid.
__device__ inline uint32_t outflow_key(const Pool& P, const CaseDev& cs, double x, double y,
                                       uint32_t p, bool& removed)
{
    uint32_t k = kSentinel;
    removed = outside_domain(cs, x, y) | (cs.retired != 0);
    const bool on_grid = cell_of(cs, x, y, k);  // k stays the sentinel when off the grid
    if (!removed && !on_grid) raise_error(P.st, 20u, p);
    return removed ? kSentinel : k;
}

// ---------------------------------------------------------------------------------------
// Step begin: strip counts (src/Boundary.cpp:62-71) of every case for the current step,
// candidate offsets, per-step counters cleared. One block, a thread per (case, strip): the
// strips' Poisson draws run side by side, and the candidate offsets are the exclusive scan of
// the counts in (case, strip) order.
template <int NT>
__device__ inline void begin_step_block(const Pool& P, uint32_t step, uint32_t np)
{
    __shared__ uint32_t sm[NT / 64 + 1];
    DevState* st = P.st;
    const int ng = 4 * P.ncases;
    uint32_t carry = 0;
    for (int g0 = 0; g0 < ng; g0 += NT) {
        const int g = g0 + (int)threadIdx.x;
        uint32_t n = 0;
        if (g < ng) {
            const int c = g >> 2, s = g & 3;
            const CaseDev& cs = P.cases[c];
            Stream rs(cs.key0, cs.key1, (uint32_t)s, step, kPInjCount);
            n = cs.retired ? 0u : new_count(cs.strip_avg[s], rs);
            if (s == 0) {
                P.cc.ncoll[c] = 0ull;
                P.cc.injected[c] = 0ull;
                P.cc.removed[c] = 0ull;
            }
        }
        uint32_t tot;
        const uint32_t ex = block_excl_scan<NT>(n, sm, tot);
        if (g < ng) P.cand_base[g] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        uint32_t ncand = carry;
        if (ncand > P.ncand_max || (uint64_t)np + ncand > (uint64_t)P.cap) {
            raise_error(st, 1u, ncand);
            ncand = 0;
        }
        P.cand_base[4 * P.ncases] = ncand;
        st->ncand = ncand;
        st->n_in = np + ncand;
        st->nhits = 0;
        st->ncl = 0;
    }
}

__global__ void __launch_bounds__(kNT) k_step_begin(Pool P)
{
    begin_step_block<kNT>(P, P.st->step, P.st->np);
}

// ---------------------------------------------------------------------------------------
// The step kernels take the case table as a separate __restrict__ argument: no kernel writes
// it, and saying so lets the compiler read a wave-uniform case with scalar loads hoisted
// ahead of the particle loads. Read through P.cases, each field is a vector load issued
// behind its own wait (the kernels' stores might alias it), several round trips per wave.
template <bool KEYS>
__global__ void __launch_bounds__(kNT) k_move(Pool P, Particles b, const CaseDev* __restrict__ cases)
{
    const uint32_t n = P.st->np;
    // with keys in a deferring pool only the key of the new position is written: the gather
    // that follows in the fused step adds the same flight as it copies the particle
    const bool write_pos = !(KEYS && P.defer);
    const uint32_t ls = P.ncases > 1 ? lane_case_starts(P.case_pstart, P.ncases) : 0u;
    for (uint32_t base = blockIdx.x * kNT; base < n; base += gridDim.x * kNT) {
        const uint32_t p = base + threadIdx.x;
        const bool act = p < n;
        // the particle's loads go out before the case lookup, which they do not depend on
        double x0 = 0.0, y0 = 0.0;
        float fvx = 0.0f, fvy = 0.0f;
        if (act) {
            const double2 q2 = b.xy[p];
            const float2 w2 = b.vxy[p];
            x0 = q2.x;
            y0 = q2.y;
            fvx = w2.x;
            fvy = w2.y;
        }
        const int c0 = P.ncases > 1 ? find_case_full_wave(P.case_pstart, P.ncases, ls, p, act) : 0;
        with_case(c0, act, [&](int c, auto uniform) {
            bool deferred = false, removed = false;
            uint32_t k = kSentinel;
            if (act) {
                const CaseDev& cs = cases[c];
                const double dx = cs.dt * (double)fvx;
                const double dy = cs.dt * (double)fvy;
                // a retired case's particles take the plain path: no wall hit, then removal
                deferred = !clearly_misses_box(cs, x0, y0, dx, dy) & (cs.retired == 0);
                if (!deferred) {
                    const double x1 = x0 + dx, y1 = y0 + dy;
                    if (write_pos) {
                        b.xy[p] = make_double2(x1, y1);
                    }
                    if (KEYS) k = outflow_key(P, cs, x1, y1, p, removed);
                }
            }
            wave_append(&P.st->nhits, P.hitlist, p, deferred);
            if (KEYS) {
                count_add<decltype(uniform)::value>(P, P.cc.removed, c, removed);
                wave_hist_add(P.cnt, k, k != kSentinel);
                if (act) P.key[p] = k;  // a deferred particle's key is written by k_extra
            }
        });
    }
}

// The exact move of a near-box particle (src/Boundary.cpp:286-327); returns true on impact.
__device__ inline bool exact_move(const Pool& P, const CaseDev& cs, int c, Particles& b,
                                  uint32_t p, uint32_t slot, uint32_t step, double& x, double& y)
{
    x = b.xy[p].x;
    y = b.xy[p].y;
    double v[3] = {(double)b.vxy[p].x, (double)b.vxy[p].y, (double)b.vz[p]};
    const double dx = cs.dt * v[0];
    const double dy = cs.dt * v[1];
    double t, nx, ny;
    if (!intersect_box(cs, x, y, dx, dy, t, nx, ny)) {
        x = x + dx;
        y = y + dy;
        b.xy[p] = make_double2(x, y);
        return false;
    }
    const double dvx = wall_hit(cs, x, y, dx, dy, v, t, nx, ny, slot, step);
    impact_add(P, c, cs.weight * cs.m * dvx);
    b.xy[p] = make_double2(x, y);
    b.vxy[p] = make_float2((float)v[0], (float)v[1]);
    b.vz[p] = (float)v[2];
    return true;
}

// Near-box particle p: exact move; with keys, its outflow key.
template <bool KEYS>
__device__ inline void near_box_one(const Pool& P, const CaseDev* cases, Particles& b, uint32_t p,
                                    uint32_t step, int& c, uint32_t& k, bool& removed)
{
    if (P.ncases > 1) c = find_case(P.case_pstart, P.ncases, p);
    const CaseDev& cs = cases[c];
    const uint32_t slot = P.ncases > 1 ? p - P.case_pstart[c] : p;
    double x, y;
    exact_move(P, cs, c, b, p, slot, step, x, y);
    if (KEYS) k = outflow_key(P, cs, x, y, p, removed);
}

// Distance of a candidate of strip s outside the domain, across the strip's edge.
__device__ inline double strip_outside(const CaseDev& cs, int s, double px, double py)
{
    return s == 0 ? cs.x1 - px : s == 1 ? px - cs.x2 : s == 2 ? cs.y1 - py : py - cs.y2;
}

// Inflow candidate j of the step (src/Boundary.cpp:83-90, MoveFlag 1), written at pool entry
// np + j (with keys only if it survives the outflow test); with keys, its cell.
template <bool KEYS>
__device__ inline void inject_one(const Pool& P, const CaseDev* cases, Particles& b, uint32_t j,
                                  uint32_t np, uint32_t step, int& c, uint32_t& k, bool& survive)
{
    const int g = find_case(P.cand_base, 4 * P.ncases, j);
    c = g >> 2;
    const int s = g & 3;
    const CaseDev& cs = cases[c];
    Stream rs(cs.key0, cs.key1, j - P.cand_base[g], step, kPInject | ((uint32_t)s << 8));
    const double* r = cs.strip[s];
    double px = r[0] + (r[2] - r[0]) * rs.next();
    double py = r[1] + (r[3] - r[1]) * rs.next();
    if (KEYS && strip_outside(cs, s, px, py) > cs.strip_reach[s]) {
        // beyond the reach no velocity brings the candidate in: dropped without its velocity
        // draws (the fused step's inflow has already keyed such candidates)
        survive = false;
        return;
    }
    double v[3];
    maxwellian(v, cs.m, cs.ux_free, cs.uy_free, cs.t_free, rs);
    px += cs.dt * v[0];
    py += cs.dt * v[1];
    survive = !outside_domain(cs, px, py);
    const uint32_t q = np + j;
    if (!KEYS || survive) {
        b.xy[q] = make_double2(px, py);
        b.vxy[q] = make_float2((float)v[0], (float)v[1]);
        b.vz[q] = (float)v[2];
    }
    if (KEYS && survive) {
        bool gone;
        k = outflow_key(P, cs, px, py, q, gone);
    }
}


// Candidates per block and pass of the fused step's inflow (kInflowPer per thread).
constexpr int kInflowPer = 4;
constexpr int kInflowRange = kInflowPer * kNT;

// The fused step's inflow candidates, a range of kInflowRange per block and pass. First the
// position draws (the first two uniforms of each candidate's own stream) and the reach test:
// a candidate further out than its strip's reach cannot enter and gets the sentinel key at
// once (at C3 about 85% of them). The others are listed in LDS, and the block's threads then
// run the velocity draws on that dense list (inject_one, from the same streams), instead of a
// few lanes of nearly every wave carrying the long FP64 draws while the rest wait.
__device__ inline void inflow_ranges(const Pool& P, const CaseDev* cases, Particles& b)
{
    __shared__ uint32_t list[kInflowRange];
    __shared__ uint32_t nlist;
    const uint32_t ncand = P.st->ncand;
    const uint32_t np = P.st->np;
    const uint32_t step = P.st->step;
    for (uint32_t r0 = blockIdx.x * kInflowRange; r0 < ncand; r0 += gridDim.x * kInflowRange) {
        if (threadIdx.x == 0) nlist = 0u;
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kInflowPer; ++u) {
            const uint32_t j = r0 + u * kNT + threadIdx.x;
            bool within = false;
            if (j < ncand) {
                const int g = find_case(P.cand_base, 4 * P.ncases, j);
                const int s = g & 3;
                const CaseDev& cs = cases[g >> 2];
                Stream rs(cs.key0, cs.key1, j - P.cand_base[g], step,
                          kPInject | ((uint32_t)s << 8));
                const double* r = cs.strip[s];
                const double px = r[0] + (r[2] - r[0]) * rs.next();
                const double py = r[1] + (r[3] - r[1]) * rs.next();
                within = !(strip_outside(cs, s, px, py) > cs.strip_reach[s]);
                if (!within) P.key[np + j] = kSentinel;
            }
            const uint64_t m = __ballot(within);
            uint32_t at = 0;
            if ((threadIdx.x & 63) == 0 && m) at = atomicAdd(&nlist, (uint32_t)__popcll(m));
            at = (uint32_t)__shfl((int)at, 0, 64);
            if (within)
                list[at + (uint32_t)__popcll(m & ((1ull << (threadIdx.x & 63)) - 1ull))] = j;
        }
        __syncthreads();
        const uint32_t n = nlist;
        for (uint32_t t0 = 0; t0 < n; t0 += kNT) {
            const uint32_t t = t0 + threadIdx.x;
            int c = 0;
            uint32_t k = kSentinel;
            bool survive = false;
            if (t < n) {
                const uint32_t j = list[t];
                inject_one<true>(P, cases, b, j, np, step, c, k, survive);
                P.key[np + j] = (P.defer && k != kSentinel) ? (k | kMovedBit) : k;
            }
            count_add(P, P.cc.injected, c, survive);
            wave_hist_add(P.cnt, k, k != kSentinel);
        }
        __syncthreads();
    }
}

template <bool KEYS>
__global__ void __launch_bounds__(kNT) k_extra(Pool P, Particles b, bool inject,
                                               const CaseDev* __restrict__ cases)
{
    const uint32_t nh = P.st->nhits;
    // with keys (the fused step) the candidates go through inflow_ranges after the hit list
    const uint32_t ncand = (inject && !KEYS) ? P.st->ncand : 0u;
    const uint32_t np = P.st->np;
    const uint32_t step = P.st->step;
    const uint32_t total = nh + ncand;
    for (uint32_t base = blockIdx.x * kNT; base < total; base += gridDim.x * kNT) {
        const uint32_t i = base + threadIdx.x;
        int c = 0;
        uint32_t k = kSentinel, q = 0;
        bool removed = false, survive = false, keyed = false;
        if (i < nh) {
            q = P.hitlist[i];
            near_box_one<KEYS>(P, cases, b, q, step, c, k, removed);
            keyed = true;
        } else if (i < total) {
            q = np + (i - nh);
            inject_one<KEYS>(P, cases, b, i - nh, np, step, c, k, survive);
            keyed = true;
        }
        count_add(P, P.cc.injected, c, survive);
        if (KEYS) {
            count_add(P, P.cc.removed, c, removed);
            wave_hist_add(P.cnt, k, k != kSentinel);
            // in a deferring pool these entries are already at their new place: marked, so
            // the gather copies them as they stand
            if (keyed) P.key[q] = (P.defer && k != kSentinel) ? (k | kMovedBit) : k;
        }
    }
    if (KEYS && inject) inflow_ranges(P, cases, b);
}

// Small pools: the move with the exact near-box path inline, and the inflow candidates as
// entries [np, np + ncand) of the same index space; keys and histogram for all.
__global__ void __launch_bounds__(kNT) k_move_small(Pool P, Particles b,
                                                    const CaseDev* __restrict__ cases)
{
    // the first pass's particle is read before the pool size is known (every slot below the
    // capacity is readable; the values are used only for live particles), so the two reads
    // overlap
    const uint32_t i0 = blockIdx.x * kNT + threadIdx.x;
    double fx = 0.0, fy = 0.0;
    float fvx = 0.0f, fvy = 0.0f;
    if (i0 < P.cap) {
        const double2 q2 = b.xy[i0];
        const float2 w2 = b.vxy[i0];
        fx = q2.x;
        fy = q2.y;
        fvx = w2.x;
        fvy = w2.y;
    }
    const uint32_t np = P.st->np;
    const uint32_t total = np + P.st->ncand;
    const uint32_t step = P.st->step;
    const uint32_t ls = P.ncases > 1 ? lane_case_starts(P.case_pstart, P.ncases) : 0u;
    for (uint32_t base = blockIdx.x * kNT; base < total; base += gridDim.x * kNT) {
        const uint32_t i = base + threadIdx.x;
        const bool first = i == i0;
        const int c0 = P.ncases > 1 ? find_case_full_wave(P.case_pstart, P.ncases, ls, i, i < np) : 0;
        with_case(c0, i < np, [&](const int cw, auto) {
            int c = cw;  // the near-box and inflow paths set their particle's own case
            uint32_t k = kSentinel;
            bool removed = false, survive = false;
            if (i < np) {
                const CaseDev& cs = cases[cw];
                const double x0 = first ? fx : b.xy[i].x, y0 = first ? fy : b.xy[i].y;
                const double dx = cs.dt * (double)(first ? fvx : b.vxy[i].x);
                const double dy = cs.dt * (double)(first ? fvy : b.vxy[i].y);
                if (clearly_misses_box(cs, x0, y0, dx, dy) | (cs.retired != 0)) {
                    const double x1 = x0 + dx, y1 = y0 + dy;
                    b.xy[i] = make_double2(x1, y1);
                    k = outflow_key(P, cs, x1, y1, i, removed);
                } else {
                    near_box_one<true>(P, cases, b, i, step, c, k, removed);
                }
            } else if (i < total) {
                inject_one<true>(P, cases, b, i - np, np, step, c, k, survive);
            }
            count_add(P, P.cc.injected, c, survive);
            count_add(P, P.cc.removed, c, removed);
            wave_hist_add(P.cnt, k, k != kSentinel);
            if (i < total) P.key[i] = k;
        });
    }
}

__global__ void k_clear_hits(Pool P) { P.st->nhits = 0; }

// Keys of entries [0, *n_ptr) of a buffer; `starts` gives each case's first entry.
//   kKeyStrict:     Indexing on the pool as it stands, every entry must lie on the grid (20).
//   kKeyDropInBox:  a freshly generated fill; entries inside the box are removed
//                   (RemoveParticlesInsideBox, src/Geometry.cpp:231-247).
__global__ void __launch_bounds__(kNT) k_keys(Pool P, Particles b, int mode,
                                              const uint32_t* starts, const uint32_t* n_ptr)
{
    const uint32_t n = *n_ptr;
    for (uint32_t base = blockIdx.x * kNT; base < n; base += gridDim.x * kNT) {
        const uint32_t p = base + threadIdx.x;
        uint32_t k = kSentinel;
        if (p < n) {
            const int c = P.ncases > 1 ? find_case(starts, P.ncases, p) : 0;
            const CaseDev& cs = P.cases[c];
            const double x = b.xy[p].x, y = b.xy[p].y;
            if (!(mode == kKeyDropInBox && inside_box(cs, x, y)) && !cell_of(cs, x, y, k)) {
                raise_error(P.st, 20u, p);
                k = kSentinel;
            }
            P.key[p] = k;
        }
        wave_hist_add(P.cnt, k, k != kSentinel);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        P.st->n_in = n;
        P.st->ncand = 0;
    }
}

// first-entry table of each case from cell offsets: out[c] = off[cell_base[c]], c <= ncases.
__global__ void k_case_starts(Pool P, const uint32_t* off, uint32_t* out)
{
    for (int c = threadIdx.x; c <= P.ncases; c += blockDim.x) out[c] = off[P.cell_base[c]];
}

// ---------------------------------------------------------------------------------------
// The gather's chunk table. Cell i starts at s(i) = off[i] + i/4, a scale on which a particle
// counts 1 and four cells count 1; chunk w holds the cells that start in [w*chunk, (w+1)*chunk)
// and first[w] is its first cell. A chunk then holds fewer than chunk particles before its last
// cell and at most 4*chunk + 3 cells, so a long run of empty cells (a retired case, a vacuum)
// is spread over many chunks instead of making one block walk all of it. Cell i (offset o,
// count v) marks every chunk boundary in (s(i), s(i+1)]; chunk size is a power of two.
__device__ inline void chunk_marks(const Pool& P, uint32_t i, uint32_t o, uint32_t v)
{
    const uint32_t sh = (uint32_t)__builtin_ctz((uint32_t)P.chunk);
    const uint32_t a = (o + (i >> 2)) >> sh, b = (o + v + ((i + 1) >> 2)) >> sh;
    for (uint32_t w = a + 1; w <= b; ++w) P.first[w] = i + 1;
}

// ---------------------------------------------------------------------------------------
// Scan of the cell histogram into Indexing's CSR offsets. Per tile of 4096 cells the epilogue
// writes off (exclusive) and the scatter cursors, clears the counts for the next step, checks
// the cell capacity (22) and fills the gather chunk table.
__device__ inline void cells_tile(const Pool& P, uint32_t i0, uint32_t hi, const uint32_t v[],
                                  uint32_t o)
{
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        const uint32_t i = i0 + q;
        if (i < hi) {
            P.off[i] = o;
            P.cursor[i] = o;
            P.cnt[i] = 0u;
            if (v[q] > P.min_max_per_cell) {
                const int c = P.ncases > 1 ? find_case(P.cell_base, P.ncases, i) : 0;
                if (v[q] > P.cases[c].max_per_cell) raise_error(P.st, 22u, i);
            }
            chunk_marks(P, i, o, v[q]);
        }
        o += v[q];
    }
}

// Every boundary up to s(mtot) = total + mtot/4 is marked, so chunks [0, s(mtot)/chunk] all
// have their first cell (the last may hold none).
__device__ inline void cells_total(const Pool& P, uint32_t total)
{
    DevState* st = P.st;
    P.off[P.mtot] = total;
    P.first[0] = 0;
    st->np_next = total;
    st->nchunks = ((total + (P.mtot >> 2)) >> __builtin_ctz((uint32_t)P.chunk)) + 1u;
    st->nhits = 0;
}

// One block walks every tile in order (small grids: one launch, no inter-block traffic).
__global__ void __launch_bounds__(kScanNT) k_scan_cells_small(Pool P)
{
    __shared__ uint32_t sm[kScanNT / 64 + 1];
    const uint32_t M = P.mtot;
    uint32_t carry = 0;
    for (uint32_t t0 = 0; t0 < M; t0 += kScanTile) {
        const uint32_t i0 = t0 + threadIdx.x * kScanItems;
        uint32_t v[kScanItems];
        uint32_t mine = 0;
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            v[q] = (i0 + q < M) ? P.cnt[i0 + q] : 0u;
            mine += v[q];
        }
        uint32_t tot;
        const uint32_t ex = block_excl_scan<kScanNT>(mine, sm, tot);
        cells_tile(P, i0, M, v, carry + ex);
        carry += tot;
    }
    if (threadIdx.x == 0) cells_total(P, carry);
}

// Small grids (M <= kSmallCells): every block scans the whole histogram into LDS and
// scatters its pool entries with it; block 0 publishes the offsets and the chunk table. The
// cursors start at zero (the small gather clears them, with the counts, for the next step).
constexpr int kSmallItems = kSmallCells / kScanNT;

__global__ void __launch_bounds__(kScanNT) k_scan_scatter_small(Pool P)
{
    __shared__ uint32_t loff[kSmallCells];
    __shared__ uint32_t sm[kScanNT / 64 + 1];
    // the first pass's key (written by the previous kernel) is read before the scan, which it
    // then overlaps; every slot below the capacity is readable
    const uint32_t p0 = blockIdx.x * kScanNT + threadIdx.x;
    const uint32_t k0 = p0 < P.cap ? P.key[p0] : kSentinel;
    const uint32_t M = P.mtot;
    const uint32_t i0 = threadIdx.x * kSmallItems;
    uint32_t v[kSmallItems];
    uint32_t mine = 0;
#pragma unroll
    for (int q = 0; q < kSmallItems; ++q) {
        v[q] = (i0 + q < M) ? P.cnt[i0 + q] : 0u;
        mine += v[q];
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan<kScanNT>(mine, sm, tot);
    uint32_t o = ex;
#pragma unroll
    for (int q = 0; q < kSmallItems; ++q) {
        const uint32_t i = i0 + q;
        if (i < M) {
            loff[i] = o;
            if (blockIdx.x == 0) {
                P.off[i] = o;
                if (v[q] > P.min_max_per_cell) {
                    const int c = P.ncases > 1 ? find_case(P.cell_base, P.ncases, i) : 0;
                    if (v[q] > P.cases[c].max_per_cell) raise_error(P.st, 22u, i);
                }
                chunk_marks(P, i, o, v[q]);
            }
        }
        o += v[q];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) cells_total(P, tot);
    __syncthreads();
    const uint32_t n = P.st->n_in;
    for (uint32_t base = blockIdx.x * kScanNT; base < n; base += gridDim.x * kScanNT) {
        const uint32_t p = base + threadIdx.x;
        const uint32_t k = p < n ? (p == p0 ? k0 : P.key[p]) : kSentinel;
        const bool valid = k != kSentinel;
        const uint32_t slot = wave_claim(P.cursor, k, valid);
        if (valid) {
            const uint32_t d = loff[k] + slot;
            if (d < P.cap) P.perm[d] = p;
        }
    }
}

// Large grids: per-block sums (k_scan_reduce), their scan, then this rescan per block range.
__global__ void __launch_bounds__(kScanNT) k_scan_cells_top(Pool P, uint32_t nb)
{
    __shared__ uint32_t sm[kScanNT / 64 + 1];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += kScanNT) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nb ? P.partial[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan<kScanNT>(v, sm, tot);
        if (i < nb) P.partial[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) cells_total(P, carry);
}

__global__ void __launch_bounds__(kScanNT) k_scan_cells_down(Pool P, uint32_t per_block)
{
    __shared__ uint32_t sm[kScanNT / 64 + 1];
    const uint32_t M = P.mtot;
    const uint32_t lo = blockIdx.x * per_block;
    const uint32_t hi = min(M, lo + per_block);
    uint32_t carry = P.partial[blockIdx.x];
    for (uint32_t t0 = lo; t0 < hi; t0 += kScanTile) {
        const uint32_t i0 = t0 + threadIdx.x * kScanItems;
        uint32_t v[kScanItems];
        uint32_t mine = 0;
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            v[q] = (i0 + q < hi) ? P.cnt[i0 + q] : 0u;
            mine += v[q];
        }
        uint32_t tot;
        const uint32_t ex = block_excl_scan<kScanNT>(mine, sm, tot);
        cells_tile(P, i0, hi, v, carry + ex);
        carry += tot;
    }
}

// ---------------------------------------------------------------------------------------
// Three-phase exclusive scan for general arrays (compaction, warm start).
__global__ void __launch_bounds__(kScanNT) k_scan_reduce(const uint32_t* in, uint32_t n,
                                                         uint32_t per_block, uint32_t* partial)
{
    __shared__ uint32_t sm[kScanNT / 64 + 1];
    const uint32_t lo = blockIdx.x * per_block;
    const uint32_t hi = min(n, lo + per_block);
    uint32_t acc = 0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kScanNT) acc += in[i];
    uint32_t tot;
    block_excl_scan<kScanNT>(acc, sm, tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanNT) k_scan_top(uint32_t* partial, uint32_t nb,
                                                      uint32_t* total_dst)
{
    __shared__ uint32_t sm[kScanNT / 64 + 1];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += kScanNT) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nb ? partial[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan<kScanNT>(v, sm, tot);
        if (i < nb) partial[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        partial[nb] = carry;
        *total_dst = carry;
    }
}

__global__ void __launch_bounds__(kScanNT) k_scan_down(const uint32_t* in, uint32_t n,
                                                       uint32_t per_block,
                                                       const uint32_t* partial, uint32_t* out)
{
    __shared__ uint32_t sm[kScanNT / 64 + 1];
    const uint32_t lo = blockIdx.x * per_block;
    const uint32_t hi = min(n, lo + per_block);
    uint32_t carry = partial[blockIdx.x];
    for (uint32_t t0 = lo; t0 < hi; t0 += kScanTile) {
        const uint32_t i0 = t0 + threadIdx.x * kScanItems;
        uint32_t v[kScanItems];
        uint32_t mine = 0;
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            v[q] = (i0 + q < hi) ? in[i0 + q] : 0u;
            mine += v[q];
        }
        uint32_t tot;
        uint32_t o = carry + block_excl_scan<kScanNT>(mine, sm, tot);
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            if (i0 + q < hi) out[i0 + q] = o;
            o += v[q];
        }
        carry += tot;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = partial[gridDim.x];
}

// ---------------------------------------------------------------------------------------
// Tiles [0, ntiles) dealt to the 8 XCDs in contiguous slices (blocks are dispatched to XCDs
// round-robin), so one XCD's L2 sees all writes that land in neighbouring cell segments.
struct XcdTiles {
    uint32_t first, end, stride;
};

__device__ inline XcdTiles xcd_tiles(uint32_t ntiles)
{
    const uint32_t nx = (gridDim.x % 8u == 0u) ? 8u : 1u;
    const uint32_t x = blockIdx.x % nx, b = blockIdx.x / nx;
    const uint32_t slice = (ntiles + nx - 1) / nx;
    return {x * slice + b, min(ntiles, (x + 1) * slice), gridDim.x / nx};
}

// kScatterBatch tiles per thread and pass: their keys are grouped first, then every group
// leader's returning atomic of the batch is issued before the first result is used, so one
// round trip to the L2 serves the whole batch.
constexpr int kScatterBatch = 4;

__global__ void __launch_bounds__(kNT) k_scatter(Pool P)
{
    const uint32_t n = P.st->n_in;
    const XcdTiles t = xcd_tiles((n + kNT - 1) / kNT);
    const int lane = threadIdx.x & 63;
    // a deferring pool's keys may carry kMovedBit: grouped and counted by cell, the bit goes
    // on with the pool position
    const uint32_t cell_mask = P.defer ? kIndexMask : 0xFFFFFFFFu;
    for (uint32_t tile = t.first; tile < t.end; tile += kScatterBatch * t.stride) {
        uint32_t p[kScatterBatch], k[kScatterBatch], base[kScatterBatch];
        WaveGroup g[kScatterBatch];
#pragma unroll
        for (int s = 0; s < kScatterBatch; ++s) {
            const uint32_t ts = tile + s * t.stride;
            p[s] = ts * kNT + threadIdx.x;
            k[s] = (ts < t.end && p[s] < n) ? P.key[p[s]] : kSentinel;
        }
#pragma unroll
        for (int s = 0; s < kScatterBatch; ++s) {
            const uint32_t cell = k[s] & cell_mask;
            g[s] = wave_group(cell, k[s] != kSentinel);
            base[s] = 0u;
            if (k[s] != kSentinel && g[s].leader == lane) base[s] = atomicAdd(&P.cursor[cell], g[s].size);
        }
#pragma unroll
        for (int s = 0; s < kScatterBatch; ++s) {
            const uint32_t d = (uint32_t)__shfl((int)base[s], g[s].leader, 64) + g[s].rank;
            if (k[s] != kSentinel && d < P.cap) P.perm[d] = p[s] | (k[s] & ~cell_mask);
        }
    }
}

// ---------------------------------------------------------------------------------------
// Velocities of one cell segment in the gather's LDS copy (or its global staging): three
// arrays.
struct SegVel {
    float* vx;
    float* vy;
    float* vz;
    __device__ void ld(uint32_t i, float& fx, float& fy, float& fz) const
    {
        fx = vx[i];
        fy = vy[i];
        fz = vz[i];
    }
    __device__ void get(uint32_t i, double v[3]) const
    {
        v[0] = (double)vx[i];
        v[1] = (double)vy[i];
        v[2] = (double)vz[i];
    }
    __device__ void set(uint32_t i, const double v[3]) const
    {
        vx[i] = (float)v[0];
        vy[i] = (float)v[1];
        vz[i] = (float)v[2];
    }
};

// Velocities of one cell segment of the pool: (vx, vy) pairs and vz.
struct PoolVel {
    float2* vxy;
    float* vz;
    __device__ void ld(uint32_t i, float& fx, float& fy, float& fz) const
    {
        const float2 w = vxy[i];
        fx = w.x;
        fy = w.y;
        fz = vz[i];
    }
    __device__ void get(uint32_t i, double v[3]) const
    {
        const float2 w = vxy[i];
        v[0] = (double)w.x;
        v[1] = (double)w.y;
        v[2] = (double)vz[i];
    }
    __device__ void set(uint32_t i, const double v[3]) const
    {
        vxy[i] = make_float2((float)v[0], (float)v[1]);
        vz[i] = (float)v[2];
    }
};

// The candidate loop of BinaryCollisionsInCell (src/Collision.cpp:28-43) with
// ElasticCollision (:7-19). Few cells have a candidate in a step, so kernels that run it
// from global memory call it out of line (collide_candidates); the small-pool gather inlines
// it under its VGPR cap.
template <class V, class R>
__device__ __forceinline__ int candidates_body(const CaseDev& cs, V vel, int npc, int nn,
                                               R& rs, double& scmax, bool& raised)
{
    int ncoll = 0;
    for (int k = 0; k < nn; ++k) {
        const int i = irand_uniform(npc, rs);
        int j;
        do { j = irand_uniform(npc, rs); } while (j == i);
        double va[3], vb[3];
        vel.get((uint32_t)i, va);
        vel.get((uint32_t)j, vb);
        const double g0 = vb[0] - va[0], g1 = vb[1] - va[1], g2 = vb[2] - va[2];
        const double cr = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
        const double sigma = cs.sigma_ref * pow(cs.cr_ref / cr, cs.omega);
        const double sg = sigma * cr;
        if (sg > scmax) { scmax = sg; raised = true; }
        if (rs.next() < sg / scmax) {
            const double cos_t = 1.0 - 2.0 * rs.next();
            const double sin_t = sqrt(1.0 - cos_t * cos_t);
            const double e = kTwoPi * rs.next();
            const double nv[3] = {cos_t, sin_t * cos(e), sin_t * sin(e)};
            const double h = cr * 0.5;
            const double vc[3] = {0.5 * (vb[0] + va[0]), 0.5 * (vb[1] + va[1]),
                                  0.5 * (vb[2] + va[2])};
            const double vr[3] = {h * nv[0], h * nv[1], h * nv[2]};
            const double a1[3] = {vc[0] + vr[0], vc[1] + vr[1], vc[2] + vr[2]};
            const double b1[3] = {vc[0] - vr[0], vc[1] - vr[1], vc[2] - vr[2]};
            vel.set((uint32_t)i, a1);
            vel.set((uint32_t)j, b1);
            ++ncoll;
        }
    }
    return ncoll;
}

__device__ __noinline__ int collide_candidates(const CaseDev& cs, PoolVel vel, int npc, int nn,
                                               Stream& rs, double& scmax, bool& raised)
{
    return candidates_body(cs, vel, npc, nn, rs, scmax, raised);
}

// The NTC candidate count of a cell (src/Collision.cpp:22-27) with the cell's (sigma*g)max:
// the integer part of 1/2 N (N-1) W (sigma g)max Dt / V plus a Bernoulli draw for the rest.
// The cell's stream is (cell within case, step, kPCollide); this is its first draw.
template <class R>
__device__ inline int draw_candidates(const Pool& P, const CaseDev& cs, uint32_t cell, int npc,
                                      R& rs)
{
    const double scmax = P.sgmax[cell];
    const double npair = 0.5 * npc * (npc - 1) * cs.weight * scmax * cs.dt / cs.dv;
    const int nn = (int)npair;
    const double frac = npair - nn;
    return (rs.next() < frac) ? nn + 1 : nn;
}

// BinaryCollisionsInCell (src/Collision.cpp:21-44): NTC with the cell's (sigma*g)max, raised
// when a candidate exceeds it.
template <bool INL, class V>
__device__ inline int collide_cell(const Pool& P, const CaseDev& cs, uint32_t cell, V vel,
                                   int npc, uint32_t step)
{
    if (npc < 2) return 0;
    Stream rs(cs.key0, cs.key1, cell - cs.cell_base, step, kPCollide);
    const int nn = draw_candidates(P, cs, cell, npc, rs);
    if (nn == 0) return 0;
    double scmax = P.sgmax[cell];
    bool raised = false;
    int ncoll;
    if constexpr (INL) ncoll = candidates_body(cs, vel, npc, nn, rs, scmax, raised);
    else ncoll = collide_candidates(cs, vel, npc, nn, rs, scmax, raised);
    if (raised) P.sgmax[cell] = scmax;
    return ncoll;
}

// Sampling (src/Sampling.cpp:14-31) of one cell: running FP64 sums, particles added one by
// one in list order. Velocities are loaded B at a time (kSampleBatch from global memory and in
// the small-pool gather, kFusedSampleBatch in the large-pool gather, whose register budget
// keeps 6 blocks per CU), so one latency covers a batch while the sums keep their order.
constexpr int kSampleBatch = 8;
constexpr int kFusedSampleBatch = 4;
// Entries per thread whose particle loads the gather issues at staging (5 x 256 covers a
// chunk of 1024 plus the LDS spare of 256), and entries per thread and pass of the gather's
// load/store batches beyond those (chunks above 1024).
constexpr int kGatherRegs = 5;
constexpr int kGatherBatch = 2;

__device__ inline void sample_add(double s[13], float fx, float fy, float fz)
{
    const double v0 = (double)fx, v1 = (double)fy, v2 = (double)fz;
    const double w2 = v0 * v0 + v1 * v1 + v2 * v2;
    s[0] += 1.0;
    s[1] += v0; s[2] += v1; s[3] += v2;
    s[4] += v0 * v0; s[5] += v0 * v1; s[6] += v0 * v2;
    s[7] += v1 * v1; s[8] += v1 * v2; s[9] += v2 * v2;
    s[10] += w2 * v0; s[11] += w2 * v1; s[12] += w2 * v2;
}

template <int B, class V>
__device__ inline void sample_cell(const Pool& P, uint32_t cell, V vel, int npc)
{
    const size_t M = P.mtot;
    double s[13];
#pragma unroll
    for (int r = 0; r < 13; ++r) s[r] = P.sums[r * M + cell];
    int i = 0;
    for (; B > 1 && i + B <= npc; i += B) {
        float bx[B], by[B], bz[B];
#pragma unroll
        for (int q = 0; q < B; ++q) vel.ld(i + q, bx[q], by[q], bz[q]);
#pragma unroll
        for (int q = 0; q < B; ++q) sample_add(s, bx[q], by[q], bz[q]);
    }
    for (; i < npc; ++i) {
        float fx, fy, fz;
        vel.ld(i, fx, fy, fz);
        sample_add(s, fx, fy, fz);
    }
#pragma unroll
    for (int r = 0; r < 13; ++r) P.sums[r * M + cell] = s[r];
}

// One of four parts of a cell's 13 moment sums, added over the cell's particles in list order
// with sample_add's own expressions, so every sum ends with the same bits as sample_cell's:
// part 0 the count and first moments (rows 0-3), 1 rows 4-6, 2 rows 7-9, 3 the third moments
// (rows 10-12). The gathers give each of a block's four waves one part (a wave-uniform branch),
// which cuts the serial per-particle work of a cell about threefold. `s` holds the part's rows
// on entry (loaded by sums_part_load) and is stored back.
template <int PART>
struct SumPart {
    static constexpr int r0 = PART == 0 ? 0 : PART == 1 ? 4 : PART == 2 ? 7 : 10;
    static constexpr int nr = PART == 0 ? 4 : 3;
};

template <int PART>
__device__ inline void sums_part_load(const Pool& P, uint32_t cell, double s[4])
{
    const size_t M = P.mtot;
#pragma unroll
    for (int r = 0; r < SumPart<PART>::nr; ++r) s[r] = P.sums[(SumPart<PART>::r0 + r) * M + cell];
}

template <int PART>
__device__ inline void sums_part_add(double s[4], float fx, float fy, float fz)
{
    const double v0 = (double)fx, v1 = (double)fy, v2 = (double)fz;
    if (PART == 0) {
        s[0] += 1.0;
        s[1] += v0; s[2] += v1; s[3] += v2;
    } else if (PART == 1) {
        s[0] += v0 * v0; s[1] += v0 * v1; s[2] += v0 * v2;
    } else if (PART == 2) {
        s[0] += v1 * v1; s[1] += v1 * v2; s[2] += v2 * v2;
    } else {
        const double w2 = v0 * v0 + v1 * v1 + v2 * v2;
        s[0] += w2 * v0; s[1] += w2 * v1; s[2] += w2 * v2;
    }
}

// The particles' velocities are read in batches of B before they are added (in order), so a
// batch's reads are in flight together instead of one read's latency per particle.
template <int PART, int B, class V>
__device__ inline void sums_part_add_store(const Pool& P, uint32_t cell, double s[4], V vel,
                                           int npc)
{
    int i = 0;
    for (; i + B <= npc; i += B) {
        float bx[B], by[B], bz[B];
#pragma unroll
        for (int q = 0; q < B; ++q) vel.ld(i + q, bx[q], by[q], bz[q]);
#pragma unroll
        for (int q = 0; q < B; ++q) sums_part_add<PART>(s, bx[q], by[q], bz[q]);
    }
    for (; i < npc; ++i) {
        float fx, fy, fz;
        vel.ld(i, fx, fy, fz);
        sums_part_add<PART>(s, fx, fy, fz);
    }
    const size_t M = P.mtot;
#pragma unroll
    for (int r = 0; r < SumPart<PART>::nr; ++r) P.sums[(SumPart<PART>::r0 + r) * M + cell] = s[r];
}

// Calls f(std::integral_constant<int, part>) with this wave's part (wave index, wave-uniform).
template <class F>
__device__ __forceinline__ void with_wave_part(F&& f)
{
    static_assert(kNT == 256, "one sum part per wave of a 256-thread block");
    switch (threadIdx.x >> 6) {
        case 0: f(std::integral_constant<int, 0>{}); break;
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        default: f(std::integral_constant<int, 3>{}); break;
    }
}

// Collide and sample one cell of the sorted pool; shared by the fused gather and the
// single-stage kernels so both paths compute the same numbers.
template <bool LDS, class V>
__device__ inline void collide_sample_cell(const Pool& P, const CaseDev* cases, uint32_t cell,
                                           V vel, int npc, uint32_t step, bool collide,
                                           bool sample_force, bool sample_rule)
{
    const int c0 = P.ncases > 1 ? find_case(P.cell_base, P.ncases, cell) : 0;
    auto body = [&](const int c, auto) {
        const CaseDev& cs = cases[c];
        const bool sampling = sampling_step(cs, step);
        if (collide) {
            const int ncoll = collide_cell<LDS>(P, cs, cell, vel, npc, step);
            if (ncoll) {
                atomicAdd(&P.cc.ncoll[c], (unsigned long long)ncoll);
                if (sampling) P.sums[(size_t)13 * P.mtot + cell] += (double)ncoll;
            }
        }
        if (sample_force || (sample_rule && sampling))
            sample_cell<LDS ? 1 : kSampleBatch>(P, cell, vel, npc);
    };
    // the collision kernels split on a wave-uniform case (scalar case reads, one ncoll atomic
    // per wave); the small-pool gather, which runs this inline, keeps its register budget
    if constexpr (LDS) body(c0, std::false_type{});
    else with_case(c0, true, body);
}

// Closes a completed fused step (one block): per-case counters, Step++, the new pool size.
// Every load is issued before the first store (a store might alias a later load, so the
// compiler would otherwise wait for each read-modify-write in turn). The new step counter and
// pool size come back through `step` and `np` on every thread.
__device__ inline void close_step_block(const Pool& P, uint32_t& step, uint32_t& np)
{
    __shared__ uint32_t sh[2];
    DevState* st = P.st;
    for (int c = threadIdx.x; c < P.ncases; c += blockDim.x) {
        const uint32_t b0 = P.cell_base[c], b1 = P.cell_base[c + 1];
        const unsigned long long nc =
            __hip_atomic_load(&P.cc.ncoll[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long ps = P.cc.psteps[c], nt = P.cc.ncoll_total[c];
        const unsigned long long inj = P.cc.injected[c], rem = P.cc.removed[c];
        const unsigned long long n = (unsigned long long)(P.off[b1] - P.off[b0]);
        P.cc.psteps[c] = ps + n;
        P.cc.ncoll_last[c] = nc;
        P.cc.ncoll_total[c] = nt + nc;
        P.cc.injected_last[c] = inj;
        P.cc.removed_last[c] = rem;
    }
    if (threadIdx.x == 0) {
        const uint32_t s1 = st->step + 1u, n1 = st->np_next;
        st->step = s1;
        st->np = n1;
        st->pending = 0u;
        sh[0] = s1;
        sh[1] = n1;
    }
    __syncthreads();
    step = sh[0];
    np = sh[1];
}

// Collide, then sample, listed cell i of the sorted pool (CollideParticles then Sampling,
// src/Simulation.cpp:155-159).
__device__ inline void collide_listed(const Pool& P, const CaseDev* cases, const Particles& b,
                                      uint32_t i, uint32_t step)
{
    const uint32_t cell = P.coll_list[i];
    const uint32_t o = P.off[cell];
    const int npc = (int)(P.off[cell + 1] - o);
    collide_sample_cell<false>(P, cases, cell, PoolVel{b.vxy + o, b.vz + o}, npc, step,
                               true, false, true);
}

// Opens a fused step (one block): if the previous fused step is still pending (its last
// kernel leaves the flag), closes it; with `begin`, the inflow counts of the step now current
// follow. One block closes the step, so no block of a wide kernel has to fence its writes (an
// agent-scope fence writes back and invalidates the L2 and costs far more than this launch).
__global__ void __launch_bounds__(kNT) k_step_open(Pool P, int begin)
{
    uint32_t step, np;
    if (P.st->pending) {
        close_step_block(P, step, np);
    } else {
        step = P.st->step;
        np = P.st->np;
    }
    if (begin) begin_step_block<kNT>(P, step, np);
}

// One chunk per iteration: stage its pool positions in LDS, mark every position with the
// bounds of its cell segment, rank each entry within its segment by pool position (the order
// in which Indexing appends to IPC), and gather the particle into that sorted place. Fused:
// then draw each cell's NTC candidate count; cells with candidates go to the collision list
// (collided and sampled by the next kernel), the others are sampled here on the LDS
// velocities, which are then written out. Chunks are dealt to the 8 XCDs in contiguous
// slices, so each XCD's L2 sees neighbouring columns. SMALL (small-pool schedule): cells with
// candidates are collided and sampled right here on the LDS copy, and the cell counts and
// scatter cursors are cleared for the next step. The two schedules' kernels (below) differ in
// their register budgets.
// One chunk: cells [c0, c1) whose particles go to sorted positions [base, base + n), staged in
// sp/seg/lvx/lvy/lvz (the block's LDS, or its region of global scratch for a chunk that
// overflows the LDS).
// Entries of sp[lo, hi) below v: eight reads issued before their compares, so one LDS
// latency covers eight entries.
__device__ __forceinline__ uint32_t rank_in_segment(const uint32_t* sp, uint32_t lo, uint32_t hi,
                                                    uint32_t v)
{
    uint32_t rank = 0, r = lo;
    for (; r + 8 <= hi; r += 8) {
        uint32_t w[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) w[t] = sp[r + t];
#pragma unroll
        for (int t = 0; t < 8; ++t) rank += w[t] < v ? 1u : 0u;
    }
    for (; r < hi; ++r) rank += sp[r] < v ? 1u : 0u;
    return rank;
}

// s_waitcnt vmcnt(0) (gfx9 encoding: expcnt and lgkmcnt left at their maxima, no wait)
constexpr int kWaitVmcnt0 = 0x0F70;
// Grid columns side by side in the gather's band order (3-4 measured best at C3: 1.4% on the
// gather; 8 columns were slower than none).
constexpr uint32_t kBandColumns = 4;

template <bool FUSED, bool SMALL, bool DEFER>
__device__ __forceinline__ void gather_chunk(const Pool& P, const CaseDev* cases, const Particles& in,
                                             const Particles& out, uint32_t* perm_sorted,
                                             uint32_t c0, uint32_t c1, int chunk_case,
                                             uint32_t base, uint32_t n,
                                             uint32_t step, bool any_sampling, uint32_t* sp,
                                             uint32_t* seg, float* lvx, float* lvy, float* lvz,
                                             uint32_t* unsorted)
{
    // a chunk of empty cells only (a retired case, a vacuum) has nothing to gather or sample,
    // and its cells' counts and cursors were not touched this step (they stay zero)
    if (n == 0) return;
    DevState* st = P.st;
    // Deferred flight (the fused step of a deferring pool): an entry without kMovedBit is
    // copied with the step's free flight added, x + dt * vx as the move computes it. The
    // chunk's cells all belong to case chunk_case (the caller cuts chunks at case starts).
    const double dt = DEFER ? cases[chunk_case].dt : 0.0;
    constexpr uint32_t pos_mask = DEFER ? kIndexMask : 0xFFFFFFFFu;
    // The large schedule stores velocities straight to their sorted place (its collisions run
    // afterwards, on the output); the LDS copy is only read by the small schedule's collisions
    // and by the sampling of fused steps.
    const bool keep_lds = SMALL || (FUSED && any_sampling);
    auto flight = [&](uint32_t entry, double& x, double& y, float vx, float vy) {
        if (!DEFER) return;
        const double dx = dt * (double)vx;
        const double dy = dt * (double)vy;
        const bool moved = (entry & kMovedBit) != 0u;
        x = moved ? x : x + dx;
        y = moved ? y : y + dy;
    };
    // Entries q = threadIdx.x + u * kNT, u < kGatherRegs, stay with this thread from staging
    // to store: their pool positions are loaded together, and their particles' loads are
    // issued right away, so those loads are in flight while the segments are marked and the
    // ranks computed. Entries beyond kGatherRegs * kNT (chunks above 1024) take the loop below.
    uint32_t my[kGatherRegs];
    double gx[kGatherRegs], gy[kGatherRegs];
    float gvx[kGatherRegs], gvy[kGatherRegs], gvz[kGatherRegs];
#pragma unroll
    for (int u = 0; u < kGatherRegs; ++u) {
        const uint32_t q = threadIdx.x + u * kNT;
        my[u] = q < n ? P.perm[base + q] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kGatherRegs; ++u) {
        // one entry's loads in flight per thread at a time: the block's requests then cover
        // one 256-entry slice of the chunk at a time, which measured faster than issuing all
        // five entries' loads at once
        if (u > 0) __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        if (threadIdx.x + u * kNT < n) {
            const uint32_t p = my[u] & pos_mask;
            const double2 q2 = in.xy[p];
            const float2 w2 = in.vxy[p];
            gx[u] = q2.x;
            gy[u] = q2.y;
            gvx[u] = w2.x;
            gvy[u] = w2.y;
            gvz[u] = in.vz[p];
        }
    }
#pragma unroll
    for (int u = 0; u < kGatherRegs; ++u) {
        const uint32_t q = threadIdx.x + u * kNT;
        if (q < n) sp[q] = my[u] & pos_mask;
    }
    for (uint32_t q = threadIdx.x + kGatherRegs * kNT; q < n; q += kNT)
        sp[q] = P.perm[base + q] & pos_mask;
    for (uint32_t c = c0 + threadIdx.x; c < c1; c += kNT) {
        const uint32_t s = P.off[c] - base, e = P.off[c + 1] - base;
        for (uint32_t q = s; q < e; ++q) seg[q] = s | (e << 16);
    }
    if (unsorted)
        for (uint32_t w = threadIdx.x; w < (n + 31) / 32; w += kNT) unsorted[w] = 0u;
    __syncthreads();
    // A segment whose entries already stand in pool order (no descent) needs no ranking: its
    // entry q has rank q - lo. Most do when few particles changed cells (the scatter keeps a
    // wave's entries of one cell in lane order); the others are ranked by a scan.
    if (unsorted) {
        for (uint32_t q = threadIdx.x; q < n; q += kNT) {
            const uint32_t lo = seg[q] & 0xFFFFu;
            if (q > lo && sp[q - 1] > sp[q]) atomicOr(&unsorted[lo >> 5], 1u << (lo & 31));
        }
        __syncthreads();
    }
    // the register-held entries: rank in the segment, then store to the sorted place
#pragma unroll
    for (int u = 0; u < kGatherRegs; ++u) {
        const uint32_t q = threadIdx.x + u * kNT;
        if (q < n) {
            const uint32_t sg = seg[q];
            const uint32_t lo = sg & 0xFFFFu, hi = sg >> 16;
            const uint32_t p = my[u] & pos_mask;
            uint32_t rank = q - lo;
            if (!unsorted || ((unsorted[lo >> 5] >> (lo & 31)) & 1u))
                rank = rank_in_segment(sp, lo, hi, p);
            const uint32_t d = lo + rank;
            flight(my[u], gx[u], gy[u], gvx[u], gvy[u]);
            out.xy[base + d] = make_double2(gx[u], gy[u]);
            if (!SMALL) {
                out.vxy[base + d] = make_float2(gvx[u], gvy[u]);
                out.vz[base + d] = gvz[u];
            }
            // the entry's stores complete before the next entry's rank (measured faster than
            // letting the five entries' stores pile up)
            __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
            if (keep_lds) {
                lvx[d] = gvx[u];
                lvy[d] = gvy[u];
                lvz[d] = gvz[u];
            }
            if (perm_sorted) perm_sorted[base + d] = p;
        }
    }
    // entries beyond the register-held batch (chunks above 1024), kGatherBatch per thread:
    // every load of the batch is issued before any store
    for (uint32_t q0 = threadIdx.x + kGatherRegs * kNT; q0 < n; q0 += kGatherBatch * kNT) {
        uint32_t mq[kGatherBatch], d[kGatherBatch];
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u) {
            const uint32_t q = q0 + u * kNT;
            mq[u] = 0u;
            d[u] = kSentinel;
            if (q < n) {
                // the staged position has lost kMovedBit: a deferring pool rereads the entry
                mq[u] = DEFER ? P.perm[base + q] : sp[q];
                const uint32_t sg = seg[q];
                const uint32_t l0 = sg & 0xFFFFu, h0 = sg >> 16;
                uint32_t r = q - l0;
                if (!unsorted || ((unsorted[l0 >> 5] >> (l0 & 31)) & 1u))
                    r = rank_in_segment(sp, l0, h0, mq[u] & pos_mask);
                d[u] = l0 + r;
            }
        }
        double bx[kGatherBatch], by[kGatherBatch];
        float bvx[kGatherBatch], bvy[kGatherBatch], bvz[kGatherBatch];
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u) {
            if (d[u] != kSentinel) {
                const uint32_t p = mq[u] & pos_mask;
                const double2 q2 = in.xy[p];
                const float2 w2 = in.vxy[p];
                bx[u] = q2.x;
                by[u] = q2.y;
                bvx[u] = w2.x;
                bvy[u] = w2.y;
                bvz[u] = in.vz[p];
            }
        }
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u) {
            if (d[u] != kSentinel) {
                flight(mq[u], bx[u], by[u], bvx[u], bvy[u]);
                out.xy[base + d[u]] = make_double2(bx[u], by[u]);
                if (!SMALL) {
                    out.vxy[base + d[u]] = make_float2(bvx[u], bvy[u]);
                    out.vz[base + d[u]] = bvz[u];
                }
                if (keep_lds) {
                    lvx[d[u]] = bvx[u];
                    lvy[d[u]] = bvy[u];
                    lvz[d[u]] = bvz[u];
                }
                if (perm_sorted) perm_sorted[base + d[u]] = mq[u] & pos_mask;
            }
        }
    }
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    const bool wave0 = threadIdx.x < 64;
    if (SMALL) {
        // the cells' collisions (they change the LDS velocities), one lane per cell
        for (uint32_t c = c0 + threadIdx.x; c < c1; c += kNT) {
            P.cnt[c] = 0u;
            P.cursor[c] = 0u;
            const uint32_t s = P.off[c] - base;
            const int npc = (int)(P.off[c + 1] - base - s);
            if (npc > 0)
                collide_sample_cell<true>(P, cases, c, SegVel{lvx + s, lvy + s, lvz + s}, npc,
                                          step, true, false, false);
        }
        __syncthreads();
        // then their sampling: every wave walks the cells and adds its part of the sums
        if (any_sampling) {
            for (uint32_t c = c0 + lane; c < c1; c += 64) {
                const uint32_t s = P.off[c] - base;
                const int npc = (int)(P.off[c + 1] - base - s);
                int ci = chunk_case;
                while (ci + 1 < P.ncases && P.cell_base[ci + 1] <= c) ++ci;
                if (npc > 0 && sampling_step(cases[ci], step)) {
                    with_wave_part([&](auto part) {
                        constexpr int K = decltype(part)::value;
                        double sp4[4];
                        sums_part_load<K>(P, c, sp4);
                        sums_part_add_store<K, kSampleBatch>(P, c, sp4,
                                                             SegVel{lvx + s, lvy + s, lvz + s}, npc);
                    });
                }
            }
        }
        __syncthreads();
    } else if (FUSED) {
        // every wave walks the chunk's cells (one lane per cell) on sampling steps, wave 0
        // alone otherwise; each draws the cell's candidate count (the same deterministic draw),
        // wave 0 lists the cells with candidates and every wave adds its part of the sums of the
        // cells without
        if (wave0 || any_sampling) {
            for (uint32_t c = c0 + lane; c < c1; c += 64) {
                const uint32_t s = P.off[c] - base;
                const int npc = (int)(P.off[c + 1] - base - s);
                // per-lane case index: a with_case split here duplicates the Philox and
                // sampling code and measured slower than the vector loads it saves
                int ci = chunk_case;
                while (ci + 1 < P.ncases && P.cell_base[ci + 1] <= c) ++ci;
                bool listed = false;
                if (npc > 0) {
                    const CaseDev& cs = cases[ci];
                    const bool samp = sampling_step(cs, step);
                    with_wave_part([&](auto part) {
                        constexpr int K = decltype(part)::value;
                        double sp4[4];
                        if (samp) sums_part_load<K>(P, c, sp4);  // in flight during the draw
                        if (npc >= 2) {
                            Stream rs(cs.key0, cs.key1, c - cs.cell_base, step, kPCollide);
                            listed = draw_candidates(P, cs, c, npc, rs) > 0;
                        }
                        if (!listed && samp)
                            sums_part_add_store<K, kFusedSampleBatch>(
                                P, c, sp4, SegVel{lvx + s, lvy + s, lvz + s}, npc);
                    });
                }
                if (wave0) wave_append(&st->ncl, P.coll_list, c, listed);
            }
        }
    }
    if (SMALL) {
        for (uint32_t q = threadIdx.x; q < n; q += kNT) {
            out.vxy[base + q] = make_float2(lvx[q], lvy[q]);
            out.vz[base + q] = lvz[q];
        }
    }
    __syncthreads();
}

template <bool FUSED, bool SMALL, bool DEFER>
__device__ __forceinline__ void gather_body(const Pool& P, const CaseDev* cases, Particles in,
                                            Particles out, uint32_t* perm_sorted)
{
    extern __shared__ uint32_t lds[];
    // one bit per segment start of an LDS-staged chunk: set when the segment is out of order
    __shared__ uint32_t unsorted[(kGatherLdsMaxEntries + 31) / 32];
    const int L = P.lds_cap;
    DevState* st = P.st;
    const uint32_t nch = st->nchunks;
    const uint32_t step = st->step;
    // whether any case samples this step (only then do waves 1-3 join the per-cell work)
    __shared__ int s_any;
    if (threadIdx.x == 0) s_any = 0;
    __syncthreads();
    for (int c = threadIdx.x; c < P.ncases; c += kNT)
        if (sampling_step(cases[c], step)) s_any = 1;
    __syncthreads();
    const bool any_sampling = s_any != 0;
    if (blockIdx.x == 0) {
        for (int c = threadIdx.x; c <= P.ncases; c += kNT) P.case_pstart[c] = P.off[P.cell_base[c]];
        if (threadIdx.x == 0) {
            if (FUSED) st->pending = 1u;  // closed by the next k_step_open
            else st->np = st->np_next;
        }
    }
    // a chunk holds up to chunk - 1 + (its last cell) particles: the LDS keeps a spare of
    // kGatherSpare entries beyond the chunk; a chunk whose last cell overflows it (at most
    // max_per_cell, checked by the scan) is staged in the block's region of global scratch
    const uint32_t G = (uint32_t)P.gcap;
    uint32_t* gs = P.gscratch + (size_t)blockIdx.x * 5u * G;
    const uint32_t nxcd = (gridDim.x % 8u == 0u) ? 8u : 1u;
    const uint32_t xcd = blockIdx.x % nxcd, lane_blk = blockIdx.x / nxcd;
    const uint32_t per_xcd = gridDim.x / nxcd;
    const uint32_t slice = (nch + nxcd - 1) / nxcd;
    const uint32_t w_end = min(nch, (xcd + 1) * slice);
    // Band order (single-case pools): within the slice, blocks of kBandColumns * cstride chunks
    // are taken column by column at the same height, cstride being the chunks per grid column.
    // A chunk's particles come partly from the neighbouring column of the previous pool order;
    // read at about the same time by the neighbour's own chunk, those lines are then still in
    // the XCD's L2. Each chunk is still taken exactly once (a transpose within each block; the
    // slice's tail keeps its order), so the results do not change.
    const uint32_t s0 = xcd * slice;
    const uint32_t ncol = P.ncases == 1 ? (uint32_t)cases[0].nx : 0u;
    const uint32_t cstride = ncol ? max(1u, (nch + ncol / 2) / ncol) : 0u;
    const uint32_t band = kBandColumns * cstride;
    for (uint32_t t = s0 + lane_blk; t < w_end; t += per_xcd) {
        uint32_t w = t;
        if (band) {
            const uint32_t r = t - s0, blk = r / band;
            if (s0 + (blk + 1) * band <= w_end) {
                const uint32_t q = r - blk * band;
                w = s0 + blk * band + (q % kBandColumns) * cstride + q / kBandColumns;
            }
        }
        const uint32_t ca = P.first[w];
        const uint32_t cb = (w + 1 < nch) ? P.first[w + 1] : P.mtot;
        // the chunk's cells go to gather_chunk one case at a time, so each part has a single
        // case (block-uniform case constants); a part without particles returns at once
        int ci = P.ncases > 1 ? find_case(P.cell_base, P.ncases, ca) : 0;
        for (uint32_t c0 = ca; c0 < cb; ++ci) {
            const uint32_t c1 = ci + 1 < P.ncases ? min(cb, P.cell_base[ci + 1]) : cb;
            const uint32_t base = P.off[c0];
            const uint32_t n = P.off[c1] - base;
            if (n <= (uint32_t)L) {
                float* lv = reinterpret_cast<float*>(lds + 2 * L);
                // the in-order check pays off where few particles change cells (small pools)
                gather_chunk<FUSED, SMALL, DEFER>(
                    P, cases, in, out, perm_sorted, c0, c1, ci, base, n, step, any_sampling, lds,
                    lds + L, lv, lv + L, lv + 2 * L,
                    SMALL && L <= kGatherLdsMaxEntries ? unsorted : nullptr);
            } else if (n <= G) {
                float* gv = reinterpret_cast<float*>(gs + 2 * G);
                gather_chunk<FUSED, SMALL, DEFER>(P, cases, in, out, perm_sorted, c0, c1, ci,
                                                  base, n, step, any_sampling, gs, gs + G, gv,
                                                  gv + G, gv + 2 * G, nullptr);
            } else if (threadIdx.x == 0) {
                raise_error(st, 22u, c0);
            }
            c0 = c1;
        }
    }
}

// The large schedule's gather streams the pool at 6 waves per SIMD (its LDS, chunk 1024 plus
// the spare, allows 6 blocks per CU); held to 80 VGPRs for that.
// DEFER: the fused step of a deferring pool, whose move left the free flight to this copy.
template <bool FUSED, bool DEFER>
__global__ void __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(6)))
k_gather(Pool P, Particles in, Particles out, uint32_t* perm_sorted,
         const CaseDev* __restrict__ cases)
{
    gather_body<FUSED, false, DEFER>(P, cases, in, out, perm_sorted);
}

// The small schedule's gather also runs the collision code inline. Its pools are small enough
// (at most 12288 cells) that its grid gives about 2 waves per SIMD in any case, so it gets the
// register budget of 2 waves per SIMD and runs the collision code without spilling.
__global__ void __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(1, 2)))
k_gather_small(Pool P, Particles in, Particles out, const CaseDev* __restrict__ cases)
{
    gather_body<true, true, false>(P, cases, in, out, nullptr);
}

// The cells of the collision list: CollideParticles on the sorted pool, then Sampling of the
// cell (src/Simulation.cpp:155-159); the next k_step_open closes the step.
__global__ void __launch_bounds__(kNT) k_collide_list(Pool P, Particles b,
                                                      const CaseDev* __restrict__ cases)
{
    const uint32_t ncl = P.st->ncl;
    const uint32_t step = P.st->step;
    for (uint32_t i = blockIdx.x * kNT + threadIdx.x; i < ncl; i += gridDim.x * kNT)
        collide_listed(P, cases, b, i, step);
}

// Single-stage CollideParticles and Sampling on the sorted pool, one lane per cell.
__global__ void __launch_bounds__(kNT) k_collide(Pool P, Particles b,
                                                 const CaseDev* __restrict__ cases)
{
    const uint32_t step = P.st->step;
    for (uint32_t cell = blockIdx.x * kNT + threadIdx.x; cell < P.mtot; cell += gridDim.x * kNT) {
        const uint32_t o = P.off[cell];
        const int npc = (int)(P.off[cell + 1] - o);
        if (npc < 2) continue;
        PoolVel vel{b.vxy + o, b.vz + o};
        collide_sample_cell<false>(P, cases, cell, vel, npc, step, true, false, false);
    }
}

__global__ void __launch_bounds__(kNT) k_sample(Pool P, Particles b, bool force,
                                                const CaseDev* __restrict__ cases)
{
    const uint32_t step = P.st->step;
    for (uint32_t cell = blockIdx.x * kNT + threadIdx.x; cell < P.mtot; cell += gridDim.x * kNT) {
        const uint32_t o = P.off[cell];
        const int npc = (int)(P.off[cell + 1] - o);
        if (npc == 0) continue;
        PoolVel vel{b.vxy + o, b.vz + o};
        collide_sample_cell<false>(P, cases, cell, vel, npc, step, false, force, !force);
    }
}

// ---------------------------------------------------------------------------------------
// Fed draws: the single-stage calls of a one-case pool on a caller's uniforms, consumed by one
// thread in the reference's order. MoveParticles (src/Boundary.cpp:284-328): particles in pool
// order, the wall model's draws at each impact. BoundaryConditions (src/Boundary.cpp:60-173):
// strips left, right, bottom, top, each its count draw(s), then per new particle two position
// and six velocity draws. CollideParticles (src/Collision.cpp:45-56): cells in k*NY+l order.
// `fs` carries the position in the array from call to call.
__global__ void k_move_fed(Pool P, Particles b, const CaseDev* __restrict__ cases, FedStream* fs)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const CaseDev& cs = cases[0];
    FedStream r = *fs;
    const uint32_t n = P.st->np;
    for (uint32_t p = 0; p < n; ++p) {
        double x = b.xy[p].x, y = b.xy[p].y;
        double v[3] = {(double)b.vxy[p].x, (double)b.vxy[p].y, (double)b.vz[p]};
        const double dx = cs.dt * v[0], dy = cs.dt * v[1];
        double t, nx, ny;
        if (!intersect_box(cs, x, y, dx, dy, t, nx, ny)) {
            b.xy[p] = make_double2(x + dx, y + dy);
            continue;
        }
        const double dvx = wall_hit_with(cs, x, y, dx, dy, v, t, nx, ny, r);
        P.cc.drag[0] += cs.weight * cs.m * dvx;
        P.cc.hits[0] += 1ull;
        b.xy[p] = make_double2(x, y);
        b.vxy[p] = make_float2((float)v[0], (float)v[1]);
        b.vz[p] = (float)v[2];
    }
    *fs = r;
}

// New particles at pool entries np.. (all of them, as the reference appends them before its
// removal pass); st->ncand their number, the case's injected count the survivors.
__global__ void k_inject_fed(Pool P, Particles b, const CaseDev* __restrict__ cases, FedStream* fs)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const CaseDev& cs = cases[0];
    FedStream r = *fs;
    const uint32_t np = P.st->np;
    uint64_t w = np, kept = 0;
    for (int s = 0; s < 4; ++s) {
        const uint32_t cnt = new_count(cs.strip_avg[s], r);
        const double* q = cs.strip[s];
        for (uint32_t j = 0; j < cnt; ++j, ++w) {
            double px = q[0] + (q[2] - q[0]) * r.next();
            double py = q[1] + (q[3] - q[1]) * r.next();
            double v[3];
            maxwellian(v, cs.m, cs.ux_free, cs.uy_free, cs.t_free, r);
            px += cs.dt * v[0];
            py += cs.dt * v[1];
            kept += outside_domain(cs, px, py) ? 0u : 1u;
            if (w < P.cap) {
                b.xy[w] = make_double2(px, py);
                b.vxy[w] = make_float2((float)v[0], (float)v[1]);
                b.vz[w] = (float)v[2];
            }
        }
    }
    if (w > P.cap) {
        raise_error(P.st, 1u, (uint32_t)min(w, (uint64_t)0xFFFFFFFFu));
        w = np;
    }
    P.st->ncand = (uint32_t)(w - np);
    P.cc.injected[0] = kept;
    *fs = r;
}

__global__ void k_collide_fed(Pool P, Particles b, const CaseDev* __restrict__ cases,
                              FedStream* fs)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const CaseDev& cs = cases[0];
    FedStream r = *fs;
    const bool sampling = sampling_step(cs, P.st->step);
    unsigned long long total = 0;
    for (uint32_t cell = 0; cell < P.mtot; ++cell) {
        const uint32_t o = P.off[cell];
        const int npc = (int)(P.off[cell + 1] - o);
        if (npc < 2) continue;
        const int nn = draw_candidates(P, cs, cell, npc, r);
        if (nn == 0) continue;
        double scmax = P.sgmax[cell];
        bool raised = false;
        const int nc = candidates_body(cs, PoolVel{b.vxy + o, b.vz + o}, npc, nn, r,
                                       scmax, raised);
        if (raised) P.sgmax[cell] = scmax;
        if (nc && sampling) P.sums[(size_t)13 * P.mtot + cell] += (double)nc;
        total += (unsigned long long)nc;
    }
    P.cc.ncoll[0] += total;
    *fs = r;
}

// ---------------------------------------------------------------------------------------
// InitialConditions (src/Initialization.cpp:290-313): uniform positions over the case's
// domain, freestream Maxwellian velocities; candidate j of case c draws from (j, 0, kPInit).
__global__ void __launch_bounds__(kNT) k_init_generate(Pool P, Particles b,
                                                       const uint32_t* init_base, uint32_t total)
{
    for (uint32_t j = blockIdx.x * kNT + threadIdx.x; j < total; j += gridDim.x * kNT) {
        const int c = P.ncases > 1 ? find_case(init_base, P.ncases, j) : 0;
        const CaseDev& cs = P.cases[c];
        Stream rs(cs.key0, cs.key1, j - init_base[c], 0u, kPInit);
        const double px = cs.x1 + (cs.x2 - cs.x1) * rs.next();
        const double py = cs.y1 + (cs.y2 - cs.y1) * rs.next();
        double v[3];
        maxwellian(v, cs.m, cs.ux_free, cs.uy_free, cs.t_free, rs);
        b.xy[j] = make_double2(px, py);
        b.vxy[j] = make_float2((float)v[0], (float)v[1]);
        b.vz[j] = (float)v[2];
    }
}

// Warm start: the particles of one cell's local Maxwellian that land outside the box (the
// cell's stream: count draw(s), then 8 draws per particle). With `b`, also writes them. A cell
// gives nf values: n, ux, uy, T (nf = 4) or n, ux, uy, Tx, Ty, Tz (nf = 6).
__device__ inline uint32_t warm_cell(const CaseDev& cs, uint32_t local, const double* f, int nf,
                                     const Particles* b, uint32_t dst)
{
    const double n = f[0], ux = f[1], uy = f[2], tx = f[3];
    const double ty = nf == 6 ? f[4] : tx, tz = nf == 6 ? f[5] : tx;
    if (!(n > 0.0) || !(tx > 0.0) || !(ty > 0.0) || !(tz > 0.0)) return 0u;
    const int k = (int)(local / (uint32_t)cs.ny), l = (int)(local % (uint32_t)cs.ny);
    const double x1 = cs.x1 + k * cs.dx, y1 = cs.y1 + l * cs.dy;
    const double x2 = x1 + cs.dx, y2 = y1 + cs.dy;
    Stream rs(cs.key0, cs.key1, local, 0u, kPWarm);
    const uint32_t cnt = new_count(n * cs.dv / cs.weight, rs);
    uint32_t kept = 0;
    for (uint32_t j = 0; j < cnt; ++j) {
        const double px = x1 + (x2 - x1) * rs.next();
        const double py = y1 + (y2 - y1) * rs.next();
        double v[3];
        maxwellian3(v, cs.m, ux, uy, tx, ty, tz, rs);
        if (inside_box(cs, px, py)) continue;
        if (b) {
            const uint32_t q = dst + kept;
            b->xy[q] = make_double2(px, py);
            b->vxy[q] = make_float2((float)v[0], (float)v[1]);
            b->vz[q] = (float)v[2];
        }
        ++kept;
    }
    return kept;
}

__global__ void __launch_bounds__(kNT) k_warm_count(Pool P, const double* fields, int nf,
                                                    uint32_t* cell_cnt)
{
    for (uint32_t cell = blockIdx.x * kNT + threadIdx.x; cell < P.mtot; cell += gridDim.x * kNT) {
        const int c = P.ncases > 1 ? find_case(P.cell_base, P.ncases, cell) : 0;
        const CaseDev& cs = P.cases[c];
        cell_cnt[cell] = warm_cell(cs, cell - cs.cell_base, fields + (size_t)nf * cell, nf,
                                   nullptr, 0);
    }
}

__global__ void __launch_bounds__(kNT) k_warm_generate(Pool P, Particles b, const double* fields,
                                                       int nf, const uint32_t* cell_off)
{
    for (uint32_t cell = blockIdx.x * kNT + threadIdx.x; cell < P.mtot; cell += gridDim.x * kNT) {
        const int c = P.ncases > 1 ? find_case(P.cell_base, P.ncases, cell) : 0;
        const CaseDev& cs = P.cases[c];
        warm_cell(cs, cell - cs.cell_base, fields + (size_t)nf * cell, nf, &b, cell_off[cell]);
    }
}

// ---------------------------------------------------------------------------------------
// Moment sums of one case accumulated over bx x by blocks of cells (a cell belongs to the block
// holding its centre): one thread per (block, row), cells added k-major, l-minor, so the sums
// are reproducible. out [bx*by][kSumRows].
__device__ inline int bin_first(int b, int n, int nb)
{
    // smallest k >= 0 with floor((k + 0.5) * nb / n) >= b
    const long long num = 2ll * b * n - nb;
    return num <= 0 ? 0 : (int)((num + 2ll * nb - 1) / (2ll * nb));
}

// CELL17: out holds the blocks in the 17-double CELL layout (include/Cell.h:4-10) instead of the
// kSumRows device rows; a row of the symmetric CountVV goes to both of its places.
template <bool CELL17>
__global__ void __launch_bounds__(kNT) k_bin_sums(Pool P, int c, int bx, int by, double* out)
{
    const CaseDev& cs = P.cases[c];
    const int total = bx * by * kSumRows;
    for (int t = blockIdx.x * kNT + threadIdx.x; t < total; t += gridDim.x * kNT) {
        const int r = t % kSumRows, b = t / kSumRows;
        const int ib = b / by, jb = b % by;
        const int k0 = bin_first(ib, cs.nx, bx), k1 = bin_first(ib + 1, cs.nx, bx);
        const int l0 = bin_first(jb, cs.ny, by), l1 = bin_first(jb + 1, cs.ny, by);
        const double* row = P.sums + (size_t)r * P.mtot + cs.cell_base;
        double acc = 0.0;
        for (int k = k0; k < min(k1, cs.nx); ++k)
            for (int l = l0; l < min(l1, cs.ny); ++l) acc += row[(size_t)k * cs.ny + l];
        if (!CELL17) {
            out[t] = acc;
            continue;
        }
        // rows NP, V0..2, VV00 VV01 VV02 VV11 VV12 VV22, V2V0..2, NC -> CELL places
        constexpr int first[kSumRows] = {0, 1, 2, 3, 4, 5, 6, 8, 9, 12, 13, 14, 15, 16};
        constexpr int second[kSumRows] = {-1, -1, -1, -1, -1, 7, 10, -1, 11, -1, -1, -1, -1, -1};
        double* o = out + (size_t)b * 17;
        o[first[r]] = acc;
        if (second[r] >= 0) o[second[r]] = acc;
    }
}

// ---------------------------------------------------------------------------------------
// Stable compaction in blocks of kScanTile entries (kScanItems consecutive per thread): the
// removal of BoundaryConditions (src/Boundary.cpp:181-202) for a one-case pool, keeping order.
__device__ inline bool keep_entry(const Pool& P, const Particles& in, uint32_t i)
{
    return !outside_domain(P.cases[0], in.xy[i].x, in.xy[i].y);
}

__global__ void __launch_bounds__(kScanNT) k_compact_count(Pool P, uint32_t n, Particles in,
                                                           uint32_t* bcnt)
{
    __shared__ uint32_t sm[kScanNT / 64 + 1];
    const uint32_t i0 = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t mine = 0;
    for (int q = 0; q < kScanItems; ++q)
        if (i0 + q < n && keep_entry(P, in, i0 + q)) ++mine;
    uint32_t tot;
    block_excl_scan<kScanNT>(mine, sm, tot);
    if (threadIdx.x == 0) bcnt[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanNT) k_compact_scatter(Pool P, uint32_t n, Particles in,
                                                             Particles out, const uint32_t* boff)
{
    __shared__ uint32_t sm[kScanNT / 64 + 1];
    const uint32_t i0 = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    bool k[kScanItems];
    uint32_t mine = 0;
    for (int q = 0; q < kScanItems; ++q) {
        k[q] = i0 + q < n && keep_entry(P, in, i0 + q);
        mine += k[q] ? 1u : 0u;
    }
    uint32_t tot;
    uint32_t o = boff[blockIdx.x] + block_excl_scan<kScanNT>(mine, sm, tot);
    for (int q = 0; q < kScanItems; ++q) {
        if (!k[q]) continue;
        const uint32_t i = i0 + q;
        out.xy[o] = in.xy[i];
        out.vxy[o] = in.vxy[i];
        out.vz[o] = in.vz[i];
        ++o;
    }
}

__global__ void __launch_bounds__(kNT) k_unpack(Particles b, uint32_t o, uint32_t n, double* x,
                                                double* y, float* vx, float* vy)
{
    for (uint32_t i = blockIdx.x * kNT + threadIdx.x; i < n; i += gridDim.x * kNT) {
        const double2 q = b.xy[o + i];
        const float2 w = b.vxy[o + i];
        x[i] = q.x;
        y[i] = q.y;
        vx[i] = w.x;
        vy[i] = w.y;
    }
}

__global__ void __launch_bounds__(kNT) k_pack(Particles b, uint32_t n, const double* x,
                                              const double* y, const float* vx, const float* vy)
{
    for (uint32_t i = blockIdx.x * kNT + threadIdx.x; i < n; i += gridDim.x * kNT) {
        b.xy[i] = make_double2(x[i], y[i]);
        b.vxy[i] = make_float2(vx[i], vy[i]);
    }
}

// Block count and per-block range of a three-phase scan over n entries.
inline void scan_shape(uint32_t n, uint32_t& nb, uint32_t& per_block)
{
    uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles == 0) tiles = 1;
    nb = tiles < 1024u ? tiles : 1024u;
    const uint32_t tiles_per_block = (tiles + nb - 1) / nb;
    per_block = tiles_per_block * kScanTile;
    nb = (n + per_block - 1) / per_block;
    if (nb == 0) nb = 1;
}

}  // namespace

// =======================================================================================
void launch_step_begin(const Pool& P, hipStream_t s)
{
    launch_k(k_step_begin, dim3(1), dim3(kNT), 0, s, P);
}

void launch_move(const Pool& P, const Particles& cur, bool keys, hipStream_t s)
{
    if (keys) launch_k(k_move<true>, dim3(P.grid_n), dim3(kNT), 0, s, P, cur, P.cases);
    else launch_k(k_move<false>, dim3(P.grid_n), dim3(kNT), 0, s, P, cur, P.cases);
}

void launch_extra(const Pool& P, const Particles& cur, bool keys, bool inject, hipStream_t s)
{
    if (keys) launch_k(k_extra<true>, dim3(P.grid_p), dim3(kNT), 0, s, P, cur, inject, P.cases);
    else launch_k(k_extra<false>, dim3(P.grid_p), dim3(kNT), 0, s, P, cur, inject, P.cases);
}

void launch_clear_hits(const Pool& P, hipStream_t s)
{
    launch_k(k_clear_hits, dim3(1), dim3(1), 0, s, P);
}

void launch_keys(const Pool& P, const Particles& buf, KeyMode mode, const uint32_t* starts,
                 const uint32_t* n_ptr, hipStream_t s)
{
    launch_k(k_keys, dim3(P.grid_n), dim3(kNT), 0, s, P, buf, (int)mode, starts, n_ptr);
}

void launch_case_starts(const Pool& P, const uint32_t* off, uint32_t* out, hipStream_t s)
{
    launch_k(k_case_starts, dim3(1), dim3(256), 0, s, P, off, out);
}

void launch_scan(const Pool& P, uint32_t* in, uint32_t n, uint32_t* out, uint32_t* total_dst,
                 hipStream_t s)
{
    uint32_t nb, per;
    scan_shape(n, nb, per);
    launch_k(k_scan_reduce, dim3(nb), dim3(kScanNT), 0, s, in, n, per, P.partial);
    launch_k(k_scan_top, dim3(1), dim3(kScanNT), 0, s, P.partial, nb, total_dst);
    launch_k(k_scan_down, dim3(nb), dim3(kScanNT), 0, s, in, n, per, P.partial, out);
}

void launch_scan_cells(const Pool& P, hipStream_t s)
{
    if (P.mtot <= 16u * kScanTile) {
        launch_k(k_scan_cells_small, dim3(1), dim3(kScanNT), 0, s, P);
        return;
    }
    uint32_t nb, per;
    scan_shape(P.mtot, nb, per);
    launch_k(k_scan_reduce, dim3(nb), dim3(kScanNT), 0, s, P.cnt, P.mtot, per, P.partial);
    launch_k(k_scan_cells_top, dim3(1), dim3(kScanNT), 0, s, P, nb);
    launch_k(k_scan_cells_down, dim3(nb), dim3(kScanNT), 0, s, P, per);
}

void launch_scatter(const Pool& P, hipStream_t s)
{
    launch_k(k_scatter, dim3(P.grid_n4), dim3(kNT), 0, s, P);
}

void launch_gather(const Pool& P, const Particles& in, const Particles& out,
                   uint32_t* perm_sorted, bool fused, hipStream_t s)
{
    const size_t lds = (size_t)kGatherLdsPerEntry * P.lds_cap;
    if (fused && P.defer)
        launch_k(k_gather<true, true>, dim3(P.grid_g), dim3(kNT), lds, s, P, in, out,
                 perm_sorted, P.cases);
    else if (fused)
        launch_k(k_gather<true, false>, dim3(P.grid_g), dim3(kNT), lds, s, P, in, out,
                 perm_sorted, P.cases);
    else
        launch_k(k_gather<false, false>, dim3(P.grid_g), dim3(kNT), lds, s, P, in, out,
                 perm_sorted, P.cases);
}

void launch_move_small(const Pool& P, const Particles& cur, hipStream_t s)
{
    launch_k(k_move_small, dim3(P.grid_p), dim3(kNT), 0, s, P, cur, P.cases);
}

void launch_scan_scatter_small(const Pool& P, hipStream_t s)
{
    launch_k(k_scan_scatter_small, dim3(P.grid_s), dim3(kScanNT), 0, s, P);
}

void launch_gather_small(const Pool& P, const Particles& in, const Particles& out, hipStream_t s)
{
    const size_t lds = (size_t)kGatherLdsPerEntry * P.lds_cap;
    launch_k(k_gather_small, dim3(P.grid_g), dim3(kNT), lds, s, P, in, out, P.cases);
}

void launch_unpack(const Particles& b, uint32_t o, uint32_t n, double* x, double* y, float* vx,
                   float* vy, hipStream_t s)
{
    if (n == 0) return;
    const uint32_t nb = std::min<uint32_t>((n + kNT - 1) / kNT, 65536u);
    launch_k(k_unpack, dim3(nb), dim3(kNT), 0, s, b, o, n, x, y, vx, vy);
}

void launch_pack(const Particles& b, uint32_t n, const double* x, const double* y,
                 const float* vx, const float* vy, hipStream_t s)
{
    if (n == 0) return;
    const uint32_t nb = std::min<uint32_t>((n + kNT - 1) / kNT, 65536u);
    launch_k(k_pack, dim3(nb), dim3(kNT), 0, s, b, n, x, y, vx, vy);
}

void launch_step_open(const Pool& P, bool begin, hipStream_t s)
{
    launch_k(k_step_open, dim3(1), dim3(kNT), 0, s, P, (int)begin);
}

void launch_collide_list(const Pool& P, const Particles& cur, hipStream_t s)
{
    launch_k(k_collide_list, dim3(P.grid_l), dim3(kNT), 0, s, P, cur, P.cases);
}

void launch_collide(const Pool& P, const Particles& cur, hipStream_t s)
{
    launch_k(k_collide, dim3(P.grid_c), dim3(kNT), 0, s, P, cur, P.cases);
}

void launch_sample(const Pool& P, const Particles& cur, bool force, hipStream_t s)
{
    launch_k(k_sample, dim3(P.grid_c), dim3(kNT), 0, s, P, cur, force, P.cases);
}

void launch_move_fed(const Pool& P, const Particles& cur, FedStream* fs, hipStream_t s)
{
    launch_k(k_move_fed, dim3(1), dim3(1), 0, s, P, cur, P.cases, fs);
}

void launch_inject_fed(const Pool& P, const Particles& cur, FedStream* fs, hipStream_t s)
{
    launch_k(k_inject_fed, dim3(1), dim3(1), 0, s, P, cur, P.cases, fs);
}

void launch_collide_fed(const Pool& P, const Particles& cur, FedStream* fs, hipStream_t s)
{
    launch_k(k_collide_fed, dim3(1), dim3(1), 0, s, P, cur, P.cases, fs);
}

void launch_init_generate(const Pool& P, const Particles& buf, const uint32_t* init_base,
                          uint32_t total, hipStream_t s)
{
    if (!total) return;
    launch_k(k_init_generate, dim3(P.grid_n), dim3(kNT), 0, s, P, buf, init_base, total);
}

void launch_warm_count(const Pool& P, const double* fields, int nf, uint32_t* cell_cnt,
                       hipStream_t s)
{
    launch_k(k_warm_count, dim3(P.grid_c), dim3(kNT), 0, s, P, fields, nf, cell_cnt);
}

void launch_warm_generate(const Pool& P, const Particles& buf, const double* fields, int nf,
                          const uint32_t* cell_off, hipStream_t s)
{
    launch_k(k_warm_generate, dim3(P.grid_c), dim3(kNT), 0, s, P, buf, fields, nf, cell_off);
}

void launch_bin_sums(const Pool& P, int c, int bx, int by, double* out, bool cell17, hipStream_t s)
{
    const int total = bx * by * kSumRows;
    const dim3 grid((total + kNT - 1) / kNT);
    if (cell17) launch_k(k_bin_sums<true>, grid, dim3(kNT), 0, s, P, c, bx, by, out);
    else launch_k(k_bin_sums<false>, grid, dim3(kNT), 0, s, P, c, bx, by, out);
}

void launch_compact(const Pool& P, uint32_t n, const Particles& in, const Particles& out,
                    uint32_t* bcnt, hipStream_t s)
{
    const uint32_t nblk = n ? (n + kScanTile - 1) / kScanTile : 1u;
    launch_k(k_compact_count, dim3(nblk), dim3(kScanNT), 0, s, P, n, in, bcnt);
    // bcnt holds nblk counts; their exclusive offsets go to bcnt + nblk + 1.
    uint32_t* boff = bcnt + nblk + 1;
    launch_scan(P, bcnt, nblk, boff, &P.st->np, s);
    if (n) launch_k(k_compact_scatter, dim3(nblk), dim3(kScanNT), 0, s, P, n, in, out, boff);
}

void set_launch_timer(LaunchTimer* t) { g_timer = t; }

hipEvent_t* LaunchTimer::acquire()
{
    if (used == pairs.size()) {
        std::array<hipEvent_t, 2> e{};
        (void)hipEventCreate(&e[0]);
        (void)hipEventCreate(&e[1]);
        pairs.push_back(e);
    }
    stage_of.resize(pairs.size());
    stage_of[used] = stage;
    return pairs[used++].data();
}

void LaunchTimer::harvest(double* stage_ms)
{
    for (size_t i = 0; i < used; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pairs[i][0], pairs[i][1]) == hipSuccess)
            stage_ms[stage_of[i]] += ms;
    }
    used = 0;
}

LaunchTimer::~LaunchTimer()
{
    for (auto& e : pairs) {
        (void)hipEventDestroy(e[0]);
        (void)hipEventDestroy(e[1]);
    }
}

int gather_blocks_per_cu(int lds_bytes)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_gather<true, true>, kNT, lds_bytes) != hipSuccess
        || nb < 1)
        nb = 1;
    return nb;
}

}  // namespace dsmc
