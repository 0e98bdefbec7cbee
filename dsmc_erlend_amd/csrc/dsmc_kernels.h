// dsmc_kernels.h - launchers of the gfx950 DSMC kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>
#include <vector>

#include "dsmc_device.h"

namespace dsmc {

// Per-pool counters kept in device memory, so a whole step runs without a host round trip.
struct DevState {
    uint32_t np;          // particles in the pool
    uint32_t np_next;     // particles after the pending sort
    uint32_t n_in;        // keyed entries of the pending sort (pool + injection candidates)
    uint32_t ncand;       // injection candidates of this step (all cases)
    uint32_t nchunks;     // gather chunks of the pending sort
    uint32_t step;        // Step
    uint32_t err_code;    // first error (DSMC_E_* code), 0 if none
    uint32_t err_index;   // particle or cell that raised it
    uint32_t nhits;       // entries of the near-box list of the current move
    uint32_t pending;     // a fused step awaits closing by k_step_open
    uint32_t ncl;         // cells on the collision list of the current step
};

// One particle buffer in three streams: FP64 positions as (x, y) pairs, FP32 velocities as
// (vx, vy) pairs and vz. A particle's 28 bytes take three loads (16, 8 and 4 bytes), and the
// move, which needs x, y, vx and vy, reads two of the streams.
struct Particles {
    double2* xy;
    float2* vxy;
    float* vz;
};

// Per-case counters in device memory.
struct CaseCounters {
    unsigned long long* ncoll;        // collisions of the step in progress
    unsigned long long* injected;     // surviving new particles of the step in progress
    unsigned long long* removed;      // particles that left the domain in the step in progress
    unsigned long long* ncoll_last;   // the same three for the last completed step
    unsigned long long* injected_last;
    unsigned long long* removed_last;
    unsigned long long* ncoll_total;
    unsigned long long* hits;         // wall impacts (DragCollisionCount)
    unsigned long long* psteps;       // sum over steps of the particles indexed that step
    double* drag;                     // DragImpulseTotal
};

// Everything the kernels of one pool see.
struct Pool {
    const CaseDev* cases;      // [ncases]
    int ncases;
    // Nonzero when cells and capacity stay below 2^31: the fused large-pool step's move then
    // writes keys only, and the gather applies the free flight as it copies (kMovedBit marks
    // the entries the near-box and inflow pass moved in place).
    int defer;
    uint32_t mtot;             // total cells of all cases
    const uint32_t* cell_base; // [ncases+1], cell_base[ncases] = mtot
    uint32_t* case_pstart;     // [ncases+1] pool position of each case's first particle
    uint32_t* cand_base;       // [4*ncases+1] first candidate of each (case, strip)
    DevState* st;
    uint32_t cap;              // particle capacity of each buffer
    uint32_t ncand_max;        // bound on injection candidates per step
    uint32_t min_max_per_cell; // smallest max_per_cell over the cases
    uint32_t* key;             // [cap] cell (or kSentinel) of each keyed entry
    uint32_t* hitlist;         // [cap] pool positions near the box, for the exact move
    uint32_t* perm;            // [cap] scatter target
    uint32_t* cnt;             // [mtot] histogram
    uint32_t* off;             // [mtot+1] exclusive cell offsets
    uint32_t* cursor;          // [mtot] scatter cursors
    uint32_t* first;           // [chunks+2] first cell of each gather chunk
    uint32_t* coll_list;       // [mtot] cells with collision candidates this step
    uint32_t* partial;         // partials of the three-phase scan
    double* sums;              // [kSumRows * mtot] moment sums, SoA
    double* sgmax;             // [mtot] (sigma*g)max per cell
    CaseCounters cc;
    int chunk;                 // particles per gather chunk
    int lds_cap;               // chunk + min(max_per_cell, kGatherSpare): entries in LDS
    int gcap;                  // chunk + max_per_cell: entries of a block's global scratch
    uint32_t* gscratch;        // [grid_g][5][gcap] staging of chunks that overflow the LDS
    int grid_p;                // grid of the particle kernels with little work (near box, inflow)
    int grid_n;                // grid of the pool-streaming kernels: one particle per thread
    int grid_n4;               // grid of the scatter: one batch of 4 tiles per thread
    int grid_g;                // grid of the gather (a multiple of 8: one slice per XCD)
    int grid_c;                // grid of the cell kernels
    int grid_l;                // grid of the collision-list kernel
    int grid_s;                // grid of the small-pool scan + scatter (1024-thread blocks)
};

// Rows of the device moment sums: CountNP, CountV[3], the 6 distinct CountVV entries
// (00,01,02,11,12,22), CountV2V[3], CountNC.
constexpr int kSumRows = 14;
// Bytes of LDS per staged particle in the gather (position, segment, three velocities).
constexpr int kGatherLdsPerEntry = 20;
// Most entries a chunk may stage in LDS (the tuning keeps chunk + max_per_cell within 64 KiB).
constexpr int kGatherLdsMaxEntries = 65536 / kGatherLdsPerEntry;
constexpr int kScanTileCells = 4096;
// LDS entries the gather keeps beyond its chunk for the chunk's last cell; a chunk whose last
// cell holds more goes through global scratch instead.
constexpr int kGatherSpare = 256;
// Largest grid the three-launch small-pool schedule handles (its scan lives in LDS).
constexpr uint32_t kSmallCells = 12288;

// --- launchers (all asynchronous on `s`) ---
// Strip counts of the current step, candidate offsets, per-step counters cleared.
void launch_step_begin(const Pool& P, hipStream_t s);
// Free flight of every particle not near the box; near-box particles go to the hit list.
// With keys: outflow removal, cell key and histogram; in a pool with `defer` set, the new
// positions are not written (the fused step's gather applies the flight as it copies).
void launch_move(const Pool& P, const Particles& cur, bool keys, hipStream_t s);
// The hit list (exact IntersectBox and wall model) and, if `inject`, the step's new
// particles; with keys: key and histogram of both.
void launch_extra(const Pool& P, const Particles& cur, bool keys, bool inject, hipStream_t s);
void launch_clear_hits(const Pool& P, hipStream_t s);
enum KeyMode { kKeyStrict = 0, kKeyDropInBox = 1 };
// Keys of entries [0, *n_ptr) of `buf`; `starts` [ncases+1] gives each case's first entry.
void launch_keys(const Pool& P, const Particles& buf, KeyMode mode, const uint32_t* starts,
                 const uint32_t* n_ptr, hipStream_t s);
// out[c] = off[cell_base[c]] for c <= ncases.
void launch_case_starts(const Pool& P, const uint32_t* off, uint32_t* out, hipStream_t s);
// Three-phase exclusive scan of in[0..n) into out[0..n] (out[n] = total, also to *total_dst).
void launch_scan(const Pool& P, uint32_t* in, uint32_t n, uint32_t* out, uint32_t* total_dst,
                 hipStream_t s);
// Scan of the cell histogram: offsets, cursors, chunk table, counts cleared.
void launch_scan_cells(const Pool& P, hipStream_t s);
void launch_scatter(const Pool& P, hipStream_t s);
// Stable in-segment order and particle gather. `fused`: also the NTC candidate draw of every cell;
// cells without candidates are sampled, the others listed for launch_collide_list.
void launch_gather(const Pool& P, const Particles& in, const Particles& out,
                   uint32_t* perm_sorted, bool fused, hipStream_t s);
// Collide, then sample, the listed cells.
void launch_collide_list(const Pool& P, const Particles& cur, hipStream_t s);
// First launch of a fused step (one block): closes the previous fused step if one is pending
// (Step++, counters, pool size), then, with `begin`, the current step's inflow counts.
void launch_step_open(const Pool& P, bool begin, hipStream_t s);
// The small-pool schedule (mtot <= kSmallCells): with k_step_open, four launches per step.
void launch_move_small(const Pool& P, const Particles& cur, hipStream_t s);
void launch_scan_scatter_small(const Pool& P, hipStream_t s);
void launch_gather_small(const Pool& P, const Particles& in, const Particles& out, hipStream_t s);
void launch_collide(const Pool& P, const Particles& cur, hipStream_t s);
void launch_sample(const Pool& P, const Particles& cur, bool force, hipStream_t s);

// Fed draws (one-case pools, one thread, the reference's draw order): move, the four strips'
// new particles (all written from entry np on, st->ncand of them), collisions.
void launch_move_fed(const Pool& P, const Particles& cur, FedStream* fs, hipStream_t s);
void launch_inject_fed(const Pool& P, const Particles& cur, FedStream* fs, hipStream_t s);
void launch_collide_fed(const Pool& P, const Particles& cur, FedStream* fs, hipStream_t s);

void launch_init_generate(const Pool& P, const Particles& buf, const uint32_t* init_base,
                          uint32_t total, hipStream_t s);
// Warm-start fill from nf values per cell: n, ux, uy, T (nf = 4) or n, ux, uy, Tx, Ty, Tz (6).
void launch_warm_count(const Pool& P, const double* fields, int nf, uint32_t* cell_cnt,
                       hipStream_t s);
void launch_warm_generate(const Pool& P, const Particles& buf, const double* fields, int nf,
                          const uint32_t* cell_off, hipStream_t s);
// Stable compaction of `n` entries of `in` into `out` keeping those inside the (single) case's
// domain; the kept count goes to st->np. `bcnt` holds 2*(n/4096 + 2) words.
// Moment sums of case c over bx x by blocks of cells: out [bx*by][kSumRows], or with cell17
// [bx*by][17] in the CELL layout.
void launch_bin_sums(const Pool& P, int c, int bx, int by, double* out, bool cell17,
                     hipStream_t s);
void launch_compact(const Pool& P, uint32_t n, const Particles& in, const Particles& out,
                    uint32_t* bcnt, hipStream_t s);
// Pool entries [o, o + n) of `b` to positions and (vx, vy) as separate arrays in device
// memory, and back (entries [0, n)); vz is one array in both forms.
void launch_unpack(const Particles& b, uint32_t o, uint32_t n, double* x, double* y, float* vx,
                   float* vy, hipStream_t s);
void launch_pack(const Particles& b, uint32_t n, const double* x, const double* y,
                 const float* vx, const float* vy, hipStream_t s);

// Occupancy-derived grids for the given LDS use of the gather.
int gather_blocks_per_cu(int lds_bytes);

// Per-kernel timing: while attached, every launch records a start/stop event pair from the
// dispatch itself; harvest() adds each pair's time to the stage it was launched under.
struct LaunchTimer {
    int stage = 0;
    std::vector<std::array<hipEvent_t, 2>> pairs;
    std::vector<int> stage_of;
    size_t used = 0;
    hipEvent_t* acquire();
    void harvest(double* stage_ms);
    ~LaunchTimer();
};
void set_launch_timer(LaunchTimer* t);

}  // namespace dsmc
