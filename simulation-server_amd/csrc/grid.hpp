// grid.hpp -- device uniform grid (replaces the UniformGrid buffers of
// src/grids/uniform_grid.cuh:30-35) and the shared scan.
#pragma once

#include "bcs_internal.hpp"
#include "device_math.hpp"

#include <hip/hip_runtime.h>

#include <vector>

namespace bcs {

// Optional per-kernel HIP-event timing (profiling mode only).
struct KernelTimer {
    std::vector<hipEvent_t>* ev_begin = nullptr;
    std::vector<hipEvent_t>* ev_end = nullptr;
    int offset = 0;  // kernel id offset of this group
    void begin(int id, hipStream_t st) const { (void)hipEventRecord((*ev_begin)[offset + id], st); }
    void end(int id, hipStream_t st) const { (void)hipEventRecord((*ev_end)[offset + id], st); }
};

struct DeviceGrid {
    GridGeom g{};
    int32_t cap = 0;            // object capacity
    int32_t n = 0;              // objects in the last build
    int shift = 0;              // bucket = key >> shift
    uint64_t nb = 0;            // buckets
    int32_t ntiles = 0;         // scan tiles over nb+1 entries
    uint32_t big_cap = 0;
    uint32_t* keys = nullptr;         // per object, input order (last build; clears its cells next build)
    uint2* items = nullptr;           // bucketed (key, index)
    uint32_t* keys_sorted = nullptr;  // gridCellIds after the sort
    int32_t* perm = nullptr;          // particleIds after the sort
    int32_t* slot_of = nullptr;       // inverse of perm (optional)
    uint2* table = nullptr;           // dense {start, count} per cell id (current where occupied)
    bool clear_table = true;          // zero the entries of cells that emptied (else: consult bm)
    BrickMap bm{};                    // occupancy map (optional; words == nullptr when absent)
    uint32_t* gpos = nullptr;         // per sorted slot: map word written by the last build, or kNoWord
    float4* brec = nullptr;           // per bucketed item q (optional): brec[2q] = (x, y, z, object
                                      // index as int bits), brec[2q+1] = (vx, vy, vz, 0)
    float4* pvs = nullptr;            // the same records per sorted slot u: positions at pvs[u],
                                      // velocities at pvs[cap + u]
    uint32_t* arrival = nullptr;      // per object: its slot within its bucket (histogram order)
    uint32_t* bcount = nullptr;       // bucket histogram [nb+1], zero between builds (the scan clears it)
    uint32_t* bstart = nullptr;       // bucket offsets [nb+2]
    uint64_t* status = nullptr;       // scan tile status
    uint32_t* ctrl = nullptr;         // [0] scan ticket, [1] big-bucket count, [2] run count
    uint32_t* runs = nullptr;         // run start of every occupied cell (unordered; optional)
    uint32_t* biglist = nullptr;
    uint64_t* rk[2] = {nullptr, nullptr};  // radix scratch of buckets beyond the LDS sort (cap each)
    uint32_t* rq[2] = {nullptr, nullptr};
    int ibits = 1;                    // bits of an object index below cap
};

// A block of a launch as one of `count` blocks working for one simulation:
// the hardware block in a launch for one simulation, or one virtual block of
// one simulation in a batched launch over many (for_virtual_blocks).
struct VBlock {
    uint32_t id, count;
};
__device__ __forceinline__ VBlock hw_block() { return VBlock{blockIdx.x, gridDim.x}; }

// Batched launch: simulation k owns the virtual blocks [start[k], start[k+1])
// of start[n] in all; the hardware blocks (a persistent grid, usually fewer)
// take them in turn, in runs of `run` consecutive virtual blocks, and call
// body(k, its virtual block) for each.  The body must keep the whole block
// on the same path (it is a kernel body); a block barrier separates two
// virtual blocks, whose LDS the next one reuses.
template <class Body>
__device__ __forceinline__ void for_virtual_blocks(const uint32_t* __restrict__ start, int32_t n, uint32_t run,
                                                   Body body)
{
    const uint32_t total = start[n];
    const uint32_t stride = gridDim.x * run;
    for (uint32_t r0 = blockIdx.x * run; r0 < total; r0 += stride) {
        const uint32_t r1 = min(r0 + run, total);
        for (uint32_t vb = r0; vb < r1; ++vb) {
            int32_t lo = 0, hi = n - 1;  // the last k with start[k] <= vb
            while (lo < hi) {
                const int32_t mid = (lo + hi + 1) >> 1;
                if (start[mid] <= vb) lo = mid;
                else hi = mid - 1;
            }
            body(lo, VBlock{vb - start[lo], start[lo + 1] - start[lo]});
            __syncthreads();
        }
    }
}

// Arguments of the grid-build kernels (one struct per kernel).
struct KeyArgs {
    GridGeom g;
    const float* pos;
    int64_t stride;
    int32_t n, n_prev;
    int shift;
    int clear_table;
    uint32_t* keys;
    uint2* table;
    BrickMap bm;
    const uint32_t* gpos;
    uint32_t* bcount;
    uint64_t* status;
    int32_t ntiles;
    uint32_t* ctrl;
    uint32_t* flags;
    uint32_t* arrival;
    const int2* order;
    int32_t norder;
    int32_t rounds;  // lane rounds per block (1..4)
};
struct ScanArgs {
    uint32_t* count;
    uint32_t* out;
    int64_t m;
    uint64_t* status;
    int32_t ntiles;
    uint32_t* ctrl;
    uint32_t* biglist;
    uint32_t big_thr, big_cap;
    uint32_t* flags;
    int clear_count;
};
struct ScatterArgs {
    const uint32_t* keys;
    int32_t n;
    int shift;
    const uint32_t* bstart;
    const uint32_t* arrival;
    uint2* items;
    uint32_t* flags;
    const float* xyz;
    const float* vel;
    int64_t stride;
    float4* brec;
};
struct RankOut {
    uint32_t* keys_sorted;
    int32_t* perm;
    int32_t* slot_of;      // nullable
    uint2* table;          // {start, count} per cell
    uint32_t* runs;        // run starts of occupied cells, in no particular order (nullable)
    uint32_t* nruns;
    const float4* brec;    // per-item records (nullable) ...
    float4* pvs;           // ... copied to their sorted slots (position records, then velocity records
    int32_t pvs_stride;    // pvs_stride entries further on)
    BrickMap bm;           // occupancy map (bm.words nullable) ...
    uint32_t* gpos;        // ... and the word each sorted slot wrote (kNoWord: none)
    int shift;             // bucket = key >> shift
    int ibits;             // bits of an object index (radix composite)
    uint64_t* rk[2];       // radix scratch: composite keys, ping-pong (capacity entries each) ...
    uint32_t* rq[2];       // ... and the item position within the bucket
};
struct RankArgs {
    const uint2* items;
    int32_t n;
    int shift;
    const uint32_t* bstart;
    const uint32_t* biglist;
    const uint32_t* ctrl;
    uint32_t big_cap;
    RankOut o;
};
struct OccArgs {
    const uint32_t* keys_sorted;
    int32_t n, cap;
    const uint2* table;
    BrickMap bm;
    uint32_t* gpos;
};
// One grid build: the arguments and block counts of its kernels (occ_blocks
// is 0 when the sort writes the occupancy map itself or there is none).
struct GridLaunch {
    KeyArgs key;
    ScanArgs scan;
    ScatterArgs scatter;
    RankArgs rank;
    OccArgs occ;
    uint32_t key_blocks, scan_blocks, scatter_blocks, rank_blocks, occ_blocks;
};

// Initialisation is enqueued on `st` (the owner's stream, which does not
// synchronise with the legacy NULL stream).
int grid_alloc(DeviceGrid& G, const GridGeom& geom, int32_t cap, bool with_slot, bool with_runs, bool with_map,
               bool with_copies, hipStream_t st);
void grid_free(DeviceGrid& G);
// pos: SoA base pointer (x, y, z at +0, +stride, +2*stride).  vel (same
// layout, nullable) with a grid allocated with copies also writes the
// sorted-slot records pvs.  order (nullable): norder lanes naming every
// object in [0, n) exactly once (.x, -1 for an empty lane), in spatial order;
// the key kernel visits the objects in that order, so that a block's lanes
// share buckets (one histogram add per bucket and block).  The result does
// not depend on it.
int grid_build(DeviceGrid& G, const float* pos, int64_t stride, int32_t n, const float* vel, uint32_t* flags,
               hipStream_t st, const KernelTimer* timer = nullptr, const int2* order = nullptr, int32_t norder = 0);
// The launches grid_build enqueues (sets G.n = n).
int grid_launch_args(DeviceGrid& G, const float* pos, int64_t stride, int32_t n, const float* vel, uint32_t* flags,
                     const int2* order, int32_t norder, GridLaunch* out);
// Batched grid builds of many simulations: device arrays of the per-kernel
// arguments of each simulation, the virtual-block offsets start[j][0..nsim]
// of kernel j (key, scan, scatter, rank, occ; device arrays) and the hardware
// blocks grid[j] of its launch (0: not launched).
struct GridBatch {
    const KeyArgs* key;
    const ScanArgs* scan;
    const ScatterArgs* scatter;
    const RankArgs* rank;
    const OccArgs* occ;
    const uint32_t* start[5];
    uint32_t grid[5];
    int32_t nsim;
    uint32_t run;  // consecutive virtual blocks per hardware block turn
};
void grid_build_batch(const GridBatch& b, hipStream_t st);
// Batched exclusive scans (k_scan over many ScanArgs).
void scan_batch(const ScanArgs* a, const uint32_t* start, int32_t nsim, uint32_t grid, uint32_t run, hipStream_t st);
constexpr int kGridKernels = 5;
// The occupancy map is written by k_rank when buckets span at least
// 2^kRankMapShift cells (at most one 8-cell group in 16 straddles two
// buckets), else by k_occ after it.
constexpr int kRankMapShift = 7;
inline bool grid_map_in_rank(const DeviceGrid& G) { return G.gpos && G.shift >= kRankMapShift; }

// Bucket size from a host sample of the objects (SoA, stride apart), and the
// switch to another bucket size (waits for the stream; buckets start empty).
int grid_shift_for(const GridGeom& g, const float* pos, int64_t stride, int32_t n);
int grid_set_shift(DeviceGrid& G, int shift, hipStream_t st);

int scan_exclusive(const uint32_t* count, uint32_t* out, int64_t m, uint64_t* status, int32_t ntiles, uint32_t* ctrl,
                   uint32_t* flags, hipStream_t st);
int scan_tiles(int64_t m);

}  // namespace bcs
