// dsmc_kernels.h - launchers of the gfx950 DSMC kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dsmc_device.h"

namespace dsmc {

// Per-pool counters kept in device memory, so a whole step runs without a host round trip.
struct DevState {
    uint32_t np;         // particles in the pool
    uint32_t np_next;    // particles after the pending sort
    uint32_t n_in;       // keyed entries of the pending sort (pool + injection candidates)
    uint32_t ncand;      // injection candidates of this step (all cases)
    uint32_t nchunks;    // gather chunks of the pending sort
    uint32_t step;       // Step
    uint32_t err_code;   // first error (DSMC_E_* code), 0 if none
    uint32_t err_index;  // particle or cell that raised it
};

// One particle buffer (SoA): FP64 positions, FP32 velocities.
struct Particles {
    double* x;
    double* y;
    float* vx;
    float* vy;
    float* vz;
};

// Per-case counters in device memory.
struct CaseCounters {
    unsigned long long* ncoll_step;   // collisions of the current step
    unsigned long long* ncoll_total;
    unsigned long long* hits;         // wall impacts (DragCollisionCount)
    unsigned long long* injected;     // surviving new particles of the current step
    unsigned long long* removed;      // particles that left the domain this step
    unsigned long long* psteps;       // sum over steps of the particles indexed that step
    double* drag;                     // DragImpulseTotal
};

// Everything the kernels of one pool see.
struct Pool {
    const CaseDev* cases;      // [ncases]
    int ncases;
    uint32_t mtot;             // total cells of all cases
    const uint32_t* cell_base; // [ncases+1], cell_base[ncases] = mtot
    uint32_t* case_pstart;     // [ncases+1] pool position of each case's first particle
    uint32_t* cand_base;       // [4*ncases+1] first candidate of each (case, strip)
    DevState* st;
    uint32_t cap;              // particle capacity of each buffer
    uint32_t ncand_max;        // bound on injection candidates per step
    uint32_t* key;             // [cap] cell (or kSentinel) of each keyed entry
    uint32_t* slot;            // [cap] arrival rank within the cell
    uint32_t* perm;            // [cap] scatter target
    uint32_t* cnt;             // [mtot] histogram
    uint32_t* off;             // [mtot+1] exclusive cell offsets
    uint32_t* first;           // [chunks+1] first cell of each gather chunk
    uint32_t* partial;         // scan partials
    double* sums;              // [kSumRows * mtot] moment sums, SoA
    double* sgmax;             // [mtot] (sigma*g)max per cell
    CaseCounters cc;
    int chunk;                 // particles per gather chunk (T)
    int lds_cap;               // chunk + max_per_cell
    int grid_p;                // persistent grid for particle kernels
    int grid_c;                // persistent grid for cell kernels
};

// Rows of the device moment sums: CountNP, CountV[3], the 6 distinct CountVV entries
// (00,01,02,11,12,22), CountV2V[3], CountNC.
constexpr int kSumRows = 14;

// --- launchers (all asynchronous on `s`) ---
void launch_step_begin(const Pool& P, hipStream_t s);
void launch_move(const Pool& P, const Particles& cur, bool keys, hipStream_t s);
void launch_inject(const Pool& P, const Particles& cur, bool keys, hipStream_t s);
enum KeyMode { kKeyStrict = 0, kKeyDropInBox = 1 };
// Keys of entries [0, *n_ptr) of `buf`; `starts` [ncases+1] gives each case's first entry.
void launch_keys(const Pool& P, const Particles& buf, KeyMode mode, const uint32_t* starts,
                 const uint32_t* n_ptr, hipStream_t s);
// out[c] = off[cell_base[c]] for c <= ncases.
void launch_case_starts(const Pool& P, const uint32_t* off, uint32_t* out, hipStream_t s);
// Exclusive scan of in[0..n) into out[0..n] (out[n] = total, also stored to *total_dst);
// `zero_in` clears the input; `chunks` also writes the gather chunk table.
void launch_scan(const Pool& P, uint32_t* in, uint32_t n, uint32_t* out, bool zero_in,
                 uint32_t* total_dst, bool chunks, hipStream_t s);
void launch_scan_cells(const Pool& P, hipStream_t s);
void launch_scatter(const Pool& P, hipStream_t s);
void launch_gather(const Pool& P, const Particles& in, const Particles& out,
                   uint32_t* perm_sorted, hipStream_t s);
void launch_collide(const Pool& P, const Particles& cur, hipStream_t s);
void launch_sample(const Pool& P, const Particles& cur, bool force, hipStream_t s);
void launch_step_end(const Pool& P, hipStream_t s);

// Initial fill: `init_base` [ncases+1] gives each case's first candidate, `total` candidates.
void launch_init_generate(const Pool& P, const Particles& buf, const uint32_t* init_base,
                          uint32_t total, hipStream_t s);
// Warm start: per-cell counts (`cell_cnt`, zeroed) then particles from per-cell Maxwellians
// given as 4 doubles per cell (n, ux, uy, T) in `fields` [4*mtot].
void launch_warm_count(const Pool& P, const double* fields, uint32_t* cell_cnt, hipStream_t s);
void launch_warm_generate(const Pool& P, const Particles& buf, const double* fields,
                          const uint32_t* cell_off, hipStream_t s);

// Stable compaction of `n` entries of `in` into `out` keeping those inside the (single) case's
// domain; the kept count goes to st->np. `bcnt` holds 2*(n/4096 + 2) words.
void launch_compact(const Pool& P, uint32_t n, const Particles& in, const Particles& out,
                    uint32_t* bcnt, hipStream_t s);

}  // namespace dsmc
