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
    float4* pvs = nullptr;            // the same records per sorted slot
    uint32_t* arrival = nullptr;      // per object: its slot within its bucket (histogram order)
    uint32_t* bcount = nullptr;       // bucket histogram [nb+1], zero between builds (the scan clears it)
    uint32_t* bstart = nullptr;       // bucket offsets [nb+2]
    uint64_t* status = nullptr;       // scan tile status
    uint32_t* ctrl = nullptr;         // [0] scan ticket, [1] big-bucket count, [2] run count
    uint32_t* runs = nullptr;         // run start of every occupied cell (unordered; optional)
    uint32_t* biglist = nullptr;
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
// the key kernel visits the objects in that order, so that its bucket
// counters are hit in address order.  The result does not depend on it.
int grid_build(DeviceGrid& G, const float* pos, int64_t stride, int32_t n, const float* vel, uint32_t* flags,
               hipStream_t st, const KernelTimer* timer = nullptr, const int2* order = nullptr, int32_t norder = 0);
constexpr int kGridKernels = 5;
// The occupancy map is written by k_rank when buckets span at least
// 2^kRankMapShift cells (few 8-cell groups straddle two buckets), else by
// k_occ after it.
constexpr int kRankMapShift = 8;
inline bool grid_map_in_rank(const DeviceGrid& G) { return G.gpos && G.shift >= kRankMapShift; }

// Bucket size from a host sample of the objects (SoA, stride apart), and the
// switch to another bucket size (waits for the stream; buckets start empty).
int grid_shift_for(const GridGeom& g, const float* pos, int64_t stride, int32_t n);
int grid_set_shift(DeviceGrid& G, int shift, hipStream_t st);

int scan_exclusive(const uint32_t* count, uint32_t* out, int64_t m, uint64_t* status, int32_t ntiles, uint32_t* ctrl,
                   uint32_t* flags, hipStream_t st);
int scan_tiles(int64_t m);

}  // namespace bcs
