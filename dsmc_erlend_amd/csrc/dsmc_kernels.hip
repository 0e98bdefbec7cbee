// dsmc_kernels.hip - gfx950 kernels of the DSMC per-step particle pipeline.
//
// Step (src/Simulation.cpp:138-199) as kernels over one particle pool shared by every case:
//   k_step_begin   strip counts of the inflow (src/Boundary.cpp:62-71), candidate offsets
//   k_move         MoveParticles + IntersectBox + wall model (src/Boundary.cpp:284-328); with
//                  keys: the outflow test (src/Boundary.cpp:181-202) and the cell key
//                  (src/Geometry.cpp:47-59) with a wave-aggregated histogram
//   k_inject       GenerateNewParticles for the four strips (src/Boundary.cpp:60-173)
//   k_scan_*       exclusive scan of the per-cell counts (CSR offsets of the IPC lists)
//   k_scatter      arrival-order scatter of pool positions into their cell segments
//   k_gather       per-chunk LDS sort of each cell segment by pool position (the stable order
//                  of Indexing's IPC append, src/Geometry.cpp:168) and gather of the SoA
//   k_collide      CollideParticles / BinaryCollisionsInCell / ElasticCollision
//                  (src/Collision.cpp:7-56), one lane per cell
//   k_sample       Sampling (src/Sampling.cpp:14-31), one lane per cell, FP64 sums
//   k_step_end     Step++ and per-case counters
// Particle kernels are persistent grid-stride loops over device-side counts, so a whole step
// is launched without host synchronisation (and can be captured in a hipGraph).
#include "dsmc_kernels.h"

namespace dsmc {

namespace {

constexpr int kNT = 256;        // threads of the particle and cell kernels
constexpr int kScanNT = 1024;   // threads of the scan kernels
constexpr int kScanItems = 4;   // items per scan thread
constexpr int kScanTile = kScanNT * kScanItems;

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

// Wave-aggregated histogram: lanes with equal keys are grouped by ballot, one atomic per
// distinct key per wave; each lane gets its arrival slot within the cell.
__device__ inline uint32_t wave_hist_slot(uint32_t* cnt, uint32_t key, bool valid)
{
    const int lane = threadIdx.x & 63;
    uint64_t pending = __ballot(valid);
    uint32_t my = 0;
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lkey = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);
        const uint64_t grp = __ballot(valid && key == lkey) & pending;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&cnt[lkey], (uint32_t)__popcll(grp));
        base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
        if ((grp >> lane) & 1ull) my = base + (uint32_t)__popcll(grp & ((1ull << lane) - 1ull));
        pending &= ~grp;
    }
    return my;
}

// Key and slot of one entry; out-of-domain entries are removed (sentinel), entries inside the
// domain whose cell falls off the grid raise code 20 as Indexing does.
__device__ inline void key_and_slot(const Pool& P, const CaseDev& cs, uint32_t p, bool act,
                                    double x, double y, bool removable, uint32_t* key,
                                    uint32_t* slot)
{
    uint32_t k = kSentinel;
    if (act) {
        if (removable && outside_domain(cs, x, y)) {
            k = kSentinel;
        } else if (!cell_of(cs, x, y, k)) {
            raise_error(P.st, 20u, p);
            k = kSentinel;
        }
    }
    const bool valid = act && k != kSentinel;
    uint32_t s = wave_hist_slot(P.cnt, k, valid);
    if (valid && s >= cs.max_per_cell) raise_error(P.st, 22u, p);
    if (act) { key[p] = k; slot[p] = s; }
}

// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_step_begin(Pool P)
{
    __shared__ uint32_t sm[1024 / 64 + 1];
    DevState* st = P.st;
    const uint32_t step = st->step;
    uint32_t total_all = 0;
    uint32_t carry = 0;
    for (int c0 = 0; c0 < P.ncases; c0 += 1024) {
        const int c = c0 + (int)threadIdx.x;
        uint32_t n4[4] = {0, 0, 0, 0};
        uint32_t mine = 0;
        if (c < P.ncases) {
            const CaseDev& cs = P.cases[c];
            for (int s = 0; s < 4; ++s) {
                Stream rs(cs.key0, cs.key1, (uint32_t)s, step, kPInjCount);
                n4[s] = new_count(cs.strip_avg[s], rs);
                mine += n4[s];
            }
            P.cc.ncoll_step[c] = 0ull;
            P.cc.injected[c] = 0ull;
            P.cc.removed[c] = 0ull;
        }
        uint32_t tot;
        uint32_t ex = block_excl_scan<1024>(mine, sm, tot);
        if (c < P.ncases) {
            uint32_t b = carry + ex;
            for (int s = 0; s < 4; ++s) { P.cand_base[4 * c + s] = b; b += n4[s]; }
        }
        carry += tot;
        total_all = carry;
    }
    if (threadIdx.x == 0) {
        uint32_t ncand = total_all;
        if (ncand > P.ncand_max || (uint64_t)st->np + ncand > (uint64_t)P.cap) {
            raise_error(st, 1u, ncand);
            ncand = 0;
        }
        P.cand_base[4 * P.ncases] = ncand;
        st->ncand = ncand;
        st->n_in = st->np + ncand;
    }
}

// ---------------------------------------------------------------------------------------
template <bool KEYS>
__global__ void __launch_bounds__(kNT) k_move(Pool P, Particles b)
{
    const uint32_t n = P.st->np;
    const uint32_t step = P.st->step;
    for (uint32_t base = blockIdx.x * kNT; base < n; base += gridDim.x * kNT) {
        const uint32_t p = base + threadIdx.x;
        const bool act = p < n;
        int c = 0;
        double px = 0.0, py = 0.0;
        if (act) {
            if (P.ncases > 1) c = find_case(P.case_pstart, P.ncases, p);
            const CaseDev& cs = P.cases[c];
            px = b.x[p];
            py = b.y[p];
            double v[3] = {(double)b.vx[p], (double)b.vy[p], (double)b.vz[p]};
            const double dx = cs.dt * v[0];
            const double dy = cs.dt * v[1];
            double t, nx, ny;
            if (!clearly_misses_box(cs, px, py, dx, dy)
                && intersect_box(cs, px, py, dx, dy, t, nx, ny)) {
                // the stream is keyed by the slot within the case, so a case draws the same
                // numbers alone and inside a batch
                const uint32_t slot = P.ncases > 1 ? p - P.case_pstart[c] : p;
                double dvx = wall_hit(cs, px, py, dx, dy, v, t, nx, ny, slot, step);
                atomicAdd(&P.cc.drag[c], cs.weight * cs.m * dvx);
                atomicAdd(&P.cc.hits[c], 1ull);
                b.vx[p] = (float)v[0];
                b.vy[p] = (float)v[1];
                b.vz[p] = (float)v[2];
            } else {
                px = px + dx;
                py = py + dy;
            }
            b.x[p] = px;
            b.y[p] = py;
            if (KEYS && outside_domain(cs, px, py)) atomicAdd(&P.cc.removed[c], 1ull);
        }
        if (KEYS) key_and_slot(P, P.cases[c], p, act, px, py, true, P.key, P.slot);
    }
}

// ---------------------------------------------------------------------------------------
template <bool KEYS>
__global__ void __launch_bounds__(kNT) k_inject(Pool P, Particles b)
{
    const uint32_t ncand = P.st->ncand;
    const uint32_t np = P.st->np;
    const uint32_t step = P.st->step;
    for (uint32_t base = blockIdx.x * kNT; base < ncand; base += gridDim.x * kNT) {
        const uint32_t j = base + threadIdx.x;
        const bool act = j < ncand;
        int c = 0;
        double px = 0.0, py = 0.0;
        const uint32_t q = np + j;
        if (act) {
            const int g = find_case(P.cand_base, 4 * P.ncases, j);
            c = g >> 2;
            const int s = g & 3;
            const CaseDev& cs = P.cases[c];
            Stream rs(cs.key0, cs.key1, j - P.cand_base[g], step, kPInject | ((uint32_t)s << 8));
            const double* r = cs.strip[s];
            px = r[0] + (r[2] - r[0]) * rs.next();
            py = r[1] + (r[3] - r[1]) * rs.next();
            double v[3];
            maxwellian(v, cs.m, cs.ux_free, cs.uy_free, cs.t_free, rs);
            px += cs.dt * v[0];
            py += cs.dt * v[1];
            const bool survive = !outside_domain(cs, px, py);
            if (survive) atomicAdd(&P.cc.injected[c], 1ull);
            if (!KEYS || survive) {
                b.x[q] = px;
                b.y[q] = py;
                b.vx[q] = (float)v[0];
                b.vy[q] = (float)v[1];
                b.vz[q] = (float)v[2];
            }
        }
        if (KEYS) key_and_slot(P, P.cases[c], q, act, px, py, true, P.key, P.slot);
    }
}

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
        const bool act = p < n;
        int c = 0;
        double px = 0.0, py = 0.0;
        bool drop = false;
        if (act) {
            if (P.ncases > 1) c = find_case(starts, P.ncases, p);
            px = b.x[p];
            py = b.y[p];
            drop = mode == kKeyDropInBox && inside_box(P.cases[c], px, py);
        }
        key_and_slot(P, P.cases[c], p, act && !drop, px, py, false, P.key, P.slot);
        if (act && drop) P.key[p] = kSentinel;
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
// Three-phase exclusive scan: per-block reduce, a one-block scan of the block sums, and a
// rescan that writes the offsets. For the cell histogram it also zeroes the counts for the
// next step and writes the gather chunk table: first[w] is the first cell whose offset is at
// least w*chunk, so chunk w holds the cells that start inside [w*chunk, (w+1)*chunk).
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
                                                      uint32_t* total_dst, uint32_t chunk,
                                                      uint32_t* nchunks_dst, uint32_t* first,
                                                      uint32_t nbins)
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
        if (first) {
            const uint32_t nch = (carry + chunk - 1) / chunk;
            *nchunks_dst = nch;
            first[0] = 0;
            first[nch] = nbins;
        }
    }
}

__global__ void __launch_bounds__(kScanNT) k_scan_down(uint32_t* in, uint32_t n,
                                                       uint32_t per_block,
                                                       const uint32_t* partial, uint32_t* out,
                                                       bool zero_in, uint32_t chunk,
                                                       uint32_t* first, const uint32_t* nchunks)
{
    __shared__ uint32_t sm[kScanNT / 64 + 1];
    const uint32_t lo = blockIdx.x * per_block;
    const uint32_t hi = min(n, lo + per_block);
    uint32_t carry = partial[blockIdx.x];
    const uint32_t nch = first ? *nchunks : 0u;
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
            const uint32_t i = i0 + q;
            if (i < hi) {
                out[i] = o;
                if (zero_in) in[i] = 0u;
                if (first && v[q]) {
                    const uint32_t e = o + v[q];
                    for (uint32_t w = o / chunk + 1; w <= e / chunk && w < nch; ++w) first[w] = i + 1;
                }
            }
            o += v[q];
        }
        carry += tot;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = partial[gridDim.x];
}

// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kNT) k_scatter(Pool P)
{
    const uint32_t n = P.st->n_in;
    for (uint32_t p = blockIdx.x * kNT + threadIdx.x; p < n; p += gridDim.x * kNT) {
        const uint32_t k = P.key[p];
        if (k != kSentinel) {
            const uint32_t d = P.off[k] + P.slot[p];
            if (d < P.cap) P.perm[d] = p;
        }
    }
}

// One chunk per iteration: stage its pool positions in LDS, mark every position with the
// bounds of its cell segment, rank each entry within its segment by pool position (the order
// in which Indexing appends to IPC), and gather the particle into that sorted place.
__global__ void __launch_bounds__(kNT) k_gather(Pool P, Particles in, Particles out,
                                                uint32_t* perm_sorted)
{
    extern __shared__ uint32_t lds[];
    uint32_t* sp = lds;                 // [lds_cap] pool positions in arrival order
    uint32_t* seg = lds + P.lds_cap;    // [lds_cap] segment bounds lo | hi << 16
    DevState* st = P.st;
    const uint32_t nch = st->nchunks;
    if (blockIdx.x == 0) {
        for (int c = threadIdx.x; c <= P.ncases; c += kNT) P.case_pstart[c] = P.off[P.cell_base[c]];
        if (threadIdx.x == 0) st->np = st->np_next;
    }
    for (uint32_t w = blockIdx.x; w < nch; w += gridDim.x) {
        const uint32_t c0 = P.first[w], c1 = P.first[w + 1];
        const uint32_t base = P.off[c0];
        const uint32_t n = P.off[c1] - base;
        if (n > (uint32_t)P.lds_cap) {
            if (threadIdx.x == 0) raise_error(st, 22u, c0);
            continue;
        }
        for (uint32_t q = threadIdx.x; q < n; q += kNT) sp[q] = P.perm[base + q];
        for (uint32_t c = c0 + threadIdx.x; c < c1; c += kNT) {
            const uint32_t s = P.off[c] - base, e = P.off[c + 1] - base;
            for (uint32_t q = s; q < e; ++q) seg[q] = s | (e << 16);
        }
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < n; q += kNT) {
            const uint32_t my = sp[q];
            const uint32_t sg = seg[q];
            const uint32_t lo = sg & 0xFFFFu, hi = sg >> 16;
            uint32_t rank = 0;
            for (uint32_t r = lo; r < hi; ++r) rank += sp[r] < my ? 1u : 0u;
            const uint32_t d = base + lo + rank;
            out.x[d] = in.x[my];
            out.y[d] = in.y[my];
            out.vx[d] = in.vx[my];
            out.vy[d] = in.vy[my];
            out.vz[d] = in.vz[my];
            if (perm_sorted) perm_sorted[d] = my;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// BinaryCollisionsInCell (src/Collision.cpp:21-44) on one cell's sorted segment: NTC with the
// cell's (sigma*g)max, raised when a candidate exceeds it; ElasticCollision (:7-19).
__global__ void __launch_bounds__(kNT) k_collide(Pool P, Particles b)
{
    const uint32_t step = P.st->step;
    for (uint32_t cell = blockIdx.x * kNT + threadIdx.x; cell < P.mtot; cell += gridDim.x * kNT) {
        const uint32_t o = P.off[cell];
        const int npc = (int)(P.off[cell + 1] - o);
        if (npc < 2) continue;
        const int c = P.ncases > 1 ? find_case(P.cell_base, P.ncases, cell) : 0;
        const CaseDev& cs = P.cases[c];
        Stream rs(cs.key0, cs.key1, cell - cs.cell_base, step, kPCollide);
        double scmax = P.sgmax[cell];
        const double npair = 0.5 * npc * (npc - 1) * cs.weight * scmax * cs.dt / cs.dv;
        int nn = (int)npair;
        const double frac = npair - nn;
        nn = (rs.next() < frac) ? nn + 1 : nn;
        int ncoll = 0;
        bool raised = false;
        for (int k = 0; k < nn; ++k) {
            const int i = irand_uniform(npc, rs);
            int j;
            do { j = irand_uniform(npc, rs); } while (j == i);
            const uint32_t a = o + (uint32_t)i, bb = o + (uint32_t)j;
            const double va[3] = {(double)b.vx[a], (double)b.vy[a], (double)b.vz[a]};
            const double vb[3] = {(double)b.vx[bb], (double)b.vy[bb], (double)b.vz[bb]};
            const double g0 = vb[0] - va[0], g1 = vb[1] - va[1], g2 = vb[2] - va[2];
            const double cr = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
            const double sigma = cs.sigma_ref * pow(cs.cr_ref / cr, cs.omega);
            const double sg = sigma * cr;
            if (sg > scmax) { scmax = sg; raised = true; }
            if (rs.next() < sg / scmax) {
                // ElasticCollision: isotropic unit vector, centre-of-mass frame.
                const double cos_t = 1.0 - 2.0 * rs.next();
                const double sin_t = sqrt(1.0 - cos_t * cos_t);
                const double e = kTwoPi * rs.next();
                const double nv[3] = {cos_t, sin_t * cos(e), sin_t * sin(e)};
                const double h = cr * 0.5;
                const double vc[3] = {0.5 * (vb[0] + va[0]), 0.5 * (vb[1] + va[1]),
                                      0.5 * (vb[2] + va[2])};
                const double vr[3] = {h * nv[0], h * nv[1], h * nv[2]};
                b.vx[a] = (float)(vc[0] + vr[0]);
                b.vy[a] = (float)(vc[1] + vr[1]);
                b.vz[a] = (float)(vc[2] + vr[2]);
                b.vx[bb] = (float)(vc[0] - vr[0]);
                b.vy[bb] = (float)(vc[1] - vr[1]);
                b.vz[bb] = (float)(vc[2] - vr[2]);
                ++ncoll;
            }
        }
        if (raised) P.sgmax[cell] = scmax;
        if (ncoll) {
            atomicAdd(&P.cc.ncoll_step[c], (unsigned long long)ncoll);
            if (sampling_step(cs, step)) P.sums[(size_t)13 * P.mtot + cell] += (double)ncoll;
        }
    }
}

// Sampling (src/Sampling.cpp:14-31): running FP64 sums, particles added one by one in the
// cell's list order.
__global__ void __launch_bounds__(kNT) k_sample(Pool P, Particles b, bool force)
{
    const uint32_t step = P.st->step;
    const size_t M = P.mtot;
    for (uint32_t cell = blockIdx.x * kNT + threadIdx.x; cell < P.mtot; cell += gridDim.x * kNT) {
        const uint32_t o = P.off[cell], e = P.off[cell + 1];
        if (o == e) continue;
        if (!force) {
            const int c = P.ncases > 1 ? find_case(P.cell_base, P.ncases, cell) : 0;
            if (!sampling_step(P.cases[c], step)) continue;
        }
        double s[13];
#pragma unroll
        for (int r = 0; r < 13; ++r) s[r] = P.sums[r * M + cell];
        for (uint32_t i = o; i < e; ++i) {
            const double v0 = (double)b.vx[i], v1 = (double)b.vy[i], v2 = (double)b.vz[i];
            const double w2 = v0 * v0 + v1 * v1 + v2 * v2;
            s[0] += 1.0;
            s[1] += v0; s[2] += v1; s[3] += v2;
            s[4] += v0 * v0; s[5] += v0 * v1; s[6] += v0 * v2;
            s[7] += v1 * v1; s[8] += v1 * v2; s[9] += v2 * v2;
            s[10] += w2 * v0; s[11] += w2 * v1; s[12] += w2 * v2;
        }
#pragma unroll
        for (int r = 0; r < 13; ++r) P.sums[r * M + cell] = s[r];
    }
}

__global__ void k_step_end(Pool P)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < P.ncases) {
        P.cc.ncoll_total[c] += P.cc.ncoll_step[c];
        P.cc.psteps[c] += (unsigned long long)(P.case_pstart[c + 1] - P.case_pstart[c]);
    }
    if (c == 0) P.st->step += 1u;
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
        b.x[j] = px;
        b.y[j] = py;
        b.vx[j] = (float)v[0];
        b.vy[j] = (float)v[1];
        b.vz[j] = (float)v[2];
    }
}

// Warm start, pass 1: per cell, the count of the local Maxwellian's particles that land
// outside the box (the cell's stream: count draw(s), then 8 draws per particle).
__device__ inline uint32_t warm_cell(const CaseDev& cs, uint32_t local, const double* f,
                                     const Particles* b, uint32_t dst)
{
    const double n = f[0], ux = f[1], uy = f[2], t = f[3];
    if (!(n > 0.0) || !(t > 0.0)) return 0u;
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
        maxwellian(v, cs.m, ux, uy, t, rs);
        if (inside_box(cs, px, py)) continue;
        if (b) {
            const uint32_t q = dst + kept;
            b->x[q] = px;
            b->y[q] = py;
            b->vx[q] = (float)v[0];
            b->vy[q] = (float)v[1];
            b->vz[q] = (float)v[2];
        }
        ++kept;
    }
    return kept;
}

__global__ void __launch_bounds__(kNT) k_warm_count(Pool P, const double* fields,
                                                    uint32_t* cell_cnt)
{
    for (uint32_t cell = blockIdx.x * kNT + threadIdx.x; cell < P.mtot; cell += gridDim.x * kNT) {
        const int c = P.ncases > 1 ? find_case(P.cell_base, P.ncases, cell) : 0;
        const CaseDev& cs = P.cases[c];
        cell_cnt[cell] = warm_cell(cs, cell - cs.cell_base, fields + 4 * (size_t)cell, nullptr, 0);
    }
}

__global__ void __launch_bounds__(kNT) k_warm_generate(Pool P, Particles b, const double* fields,
                                                       const uint32_t* cell_off)
{
    for (uint32_t cell = blockIdx.x * kNT + threadIdx.x; cell < P.mtot; cell += gridDim.x * kNT) {
        const int c = P.ncases > 1 ? find_case(P.cell_base, P.ncases, cell) : 0;
        const CaseDev& cs = P.cases[c];
        warm_cell(cs, cell - cs.cell_base, fields + 4 * (size_t)cell, &b, cell_off[cell]);
    }
}

// ---------------------------------------------------------------------------------------
// Stable compaction in blocks of kScanTile entries (kScanItems consecutive per thread): the
// removal of BoundaryConditions (src/Boundary.cpp:181-202) for a one-case pool, keeping order.
__device__ inline bool keep_entry(const Pool& P, const Particles& in, uint32_t i)
{
    return !outside_domain(P.cases[0], in.x[i], in.y[i]);
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
        out.x[o] = in.x[i];
        out.y[o] = in.y[i];
        out.vx[o] = in.vx[i];
        out.vy[o] = in.vy[i];
        out.vz[o] = in.vz[i];
        ++o;
    }
}

// Block count and per-block range of a scan over n entries.
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
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1024), 0, s, P);
}

void launch_move(const Pool& P, const Particles& cur, bool keys, hipStream_t s)
{
    if (keys) hipLaunchKernelGGL(k_move<true>, dim3(P.grid_p), dim3(kNT), 0, s, P, cur);
    else hipLaunchKernelGGL(k_move<false>, dim3(P.grid_p), dim3(kNT), 0, s, P, cur);
}

void launch_inject(const Pool& P, const Particles& cur, bool keys, hipStream_t s)
{
    int g = (int)((P.ncand_max + kNT - 1) / kNT);
    if (g > P.grid_p) g = P.grid_p;
    if (g < 1) g = 1;
    if (keys) hipLaunchKernelGGL(k_inject<true>, dim3(g), dim3(kNT), 0, s, P, cur);
    else hipLaunchKernelGGL(k_inject<false>, dim3(g), dim3(kNT), 0, s, P, cur);
}

void launch_keys(const Pool& P, const Particles& buf, KeyMode mode, const uint32_t* starts,
                 const uint32_t* n_ptr, hipStream_t s)
{
    hipLaunchKernelGGL(k_keys, dim3(P.grid_p), dim3(kNT), 0, s, P, buf, (int)mode, starts, n_ptr);
}

void launch_case_starts(const Pool& P, const uint32_t* off, uint32_t* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_case_starts, dim3(1), dim3(256), 0, s, P, off, out);
}

void launch_scan(const Pool& P, uint32_t* in, uint32_t n, uint32_t* out, bool zero_in,
                 uint32_t* total_dst, bool chunks, hipStream_t s)
{
    uint32_t nb, per;
    scan_shape(n, nb, per);
    uint32_t* first = chunks ? P.first : nullptr;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanNT), 0, s, in, n, per, P.partial);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanNT), 0, s, P.partial, nb, total_dst,
                       (uint32_t)P.chunk, &P.st->nchunks, first, n);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kScanNT), 0, s, in, n, per, P.partial, out,
                       zero_in, (uint32_t)P.chunk, first, (const uint32_t*)&P.st->nchunks);
}

void launch_scan_cells(const Pool& P, hipStream_t s)
{
    launch_scan(P, P.cnt, P.mtot, P.off, true, &P.st->np_next, true, s);
}

void launch_scatter(const Pool& P, hipStream_t s)
{
    hipLaunchKernelGGL(k_scatter, dim3(P.grid_p), dim3(kNT), 0, s, P);
}

void launch_gather(const Pool& P, const Particles& in, const Particles& out,
                   uint32_t* perm_sorted, hipStream_t s)
{
    size_t lds = (size_t)2 * P.lds_cap * sizeof(uint32_t);
    hipLaunchKernelGGL(k_gather, dim3(P.grid_p), dim3(kNT), lds, s, P, in, out, perm_sorted);
}

void launch_collide(const Pool& P, const Particles& cur, hipStream_t s)
{
    hipLaunchKernelGGL(k_collide, dim3(P.grid_c), dim3(kNT), 0, s, P, cur);
}

void launch_sample(const Pool& P, const Particles& cur, bool force, hipStream_t s)
{
    hipLaunchKernelGGL(k_sample, dim3(P.grid_c), dim3(kNT), 0, s, P, cur, force);
}

void launch_step_end(const Pool& P, hipStream_t s)
{
    int nb = (P.ncases + 255) / 256;
    hipLaunchKernelGGL(k_step_end, dim3(nb), dim3(256), 0, s, P);
}

void launch_init_generate(const Pool& P, const Particles& buf, const uint32_t* init_base,
                          uint32_t total, hipStream_t s)
{
    if (!total) return;
    hipLaunchKernelGGL(k_init_generate, dim3(P.grid_p), dim3(kNT), 0, s, P, buf, init_base, total);
}

void launch_warm_count(const Pool& P, const double* fields, uint32_t* cell_cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_warm_count, dim3(P.grid_c), dim3(kNT), 0, s, P, fields, cell_cnt);
}

void launch_warm_generate(const Pool& P, const Particles& buf, const double* fields,
                          const uint32_t* cell_off, hipStream_t s)
{
    hipLaunchKernelGGL(k_warm_generate, dim3(P.grid_c), dim3(kNT), 0, s, P, buf, fields, cell_off);
}

void launch_compact(const Pool& P, uint32_t n, const Particles& in, const Particles& out,
                    uint32_t* bcnt, hipStream_t s)
{
    const uint32_t nblk = n ? (n + kScanTile - 1) / kScanTile : 1u;
    hipLaunchKernelGGL(k_compact_count, dim3(nblk), dim3(kScanNT), 0, s, P, n, in, bcnt);
    // bcnt holds nblk counts; their exclusive offsets go to bcnt + nblk + 1.
    uint32_t* boff = bcnt + nblk + 1;
    uint32_t nb, per;
    scan_shape(nblk, nb, per);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanNT), 0, s, bcnt, nblk, per, P.partial);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanNT), 0, s, P.partial, nb, &P.st->np, 1u,
                       (uint32_t*)nullptr, (uint32_t*)nullptr, 0u);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kScanNT), 0, s, bcnt, nblk, per, P.partial,
                       boff, false, 1u, (uint32_t*)nullptr, (const uint32_t*)nullptr);
    if (n) hipLaunchKernelGGL(k_compact_scatter, dim3(nblk), dim3(kScanNT), 0, s, P, n, in, out, boff);
}

}  // namespace dsmc
