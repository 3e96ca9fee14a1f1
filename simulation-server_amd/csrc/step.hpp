// step.hpp -- kernel-argument blocks and launchers of the physics step.
#pragma once

#include "grid.hpp"

namespace bcs {

constexpr int kMaxEndings = 16;
constexpr int kBlock = 256;  // threads per cell-chunk block
constexpr int kBsLds = 1024;       // model-vertex radii held in LDS by the contact search (more: read globally)
constexpr int kContactBatch = 4;   // sorted-slot records loaded together by the contact search
constexpr int kContactWindow = 16; // stream positions tested before their contacts are applied (<= 32)
constexpr int kSurrStatSlots = 256, kSurrStatStride = 16;  // surrogate tallies: 256 slots, 64 B apart
constexpr int kCellBatch = 6;      // neighbour-cell table entries loaded together by the contact search
constexpr uint32_t kXcdRun = 16;   // consecutive chunks per XCD run in the block -> chunk mapping
constexpr uint32_t kXcdRunTiles = 256;  // the same for the sorted-slot tiles of the contact search
constexpr int kSpringLds = 1024;   // springs of one cell type held in LDS by k_cells (more: read globally)
constexpr int kOrderBits = 15;      // Morton bits of the cell processing order
constexpr int kReorderPeriod = 16;  // frames between updates of the order

// Folded type tables (src/meta_factory/blood_cell_factory.hpp:172-282).
struct TypeTable {
    int32_t nt;
    int32_t np, nc, nm;
    int32_t ppc[kMaxTypes], cnt[kMaxTypes], pstart[kMaxTypes], cstart[kMaxTypes], mstart[kMaxTypes];
    int32_t end_spheres[kMaxTypes];  // vein-end variant with ending spheres (block-sync variant)
    Div32 ppc_div[kMaxTypes];        // division by ppc[t] (model index of a partner)
};

// Physics constants in the form the kernels use them.
struct PhysConst {
    float dt, k, d, vk, vd, vcd, vcfi, vdamp, cs, cd, csh, brake, bint;
    float initv[3];
    float impact2, minimal2;
    float G[3];
    float cyl_r, spawn_y;
    float thr_upper, thr_lower, thr_right, thr_left, thr_front, thr_back;
    int32_t use_blood_flow, reaction, brake_on, nend;
    float aabb_margin;      // absolute inflation of the triangle-cell boxes (rounding slack)
    float rmax;             // largest contact radius of the model table (+inf: unusable as a bound)
    int32_t rk4;            // particle springs and integration of the reference's Runge-Kutta build
    int32_t r_uniform;      // every radius a partner's model index can select is rmax, and none clamps
    uint64_t seed;
    float endc[3 * kMaxEndings];
    float endr[kMaxEndings];
};

// The learned spring surrogate inside the cell kernel (k_cells<.., true>):
// for the cells of type `type` the spring force S of each particle is
// replaced by the network's prediction when the cell's largest error
// |d_hat - (dt^2/4) S| is at most tol (else the plain S stays: fallback).
struct SurrArgs {
    const float* D;        // network output per cell of `type`: [cell - cstart][3a + axis] (ld 128),
                           // normalised units (times out_scale: spring displacement); null: no prediction
    float* S_out;          // sampling: (dt^2/4) S per particle of `type`, same layout; null: none
    uint8_t* touched;      // per particle: 1 if it had a contact in this frame's search (null: not kept)
    uint32_t* stats;       // kSurrStatSlots slots of kSurrStatStride words (spreading the atomics): [0]
                           // accepted cells, [1] fallback cells, [2] max error of accepted cells (float
                           // bits), [3] max error of all predicted cells (float bits); summed on the host
    int32_t type;          // the served type (ppc 42); -1: none
    float scale;           // dt^2 / 4
    float out_scale;       // network units -> displacement
    float tol;             // accepted when the cell's error <= tol (tol <= 0: every cell falls back)
};

struct StepArgs {
    TypeTable tt;
    PhysConst pc;
    SurrArgs surr;
    // particles (SoA, stride np)
    float* pos;
    float* vel;
    float* frc;
    float* centers;  // SoA, stride nc
    // model tables
    const float* bs;
    const float* r0;
    const float* model;   // 3*nm interleaved
    const int32_t* sp_off;  // CSR per model vertex, ascending neighbour
    const int32_t* sp_b;
    const float* sp_L;
    // work chunks (whole cells of one type), taken in the order of corder
    const int32_t* chunk_type;
    const int32_t* chunk_first;
    const int32_t* chunk_n;
    const int32_t* corder;  // processing order of the cells (a permutation within each type)
    const int2* chunk_pi;   // per chunk c, lane l at [c * kBlock + l]: {particle, blood cell}, from corder
    int32_t nchunks;
    // particle grid
    GridGeom pg;
    const uint32_t* p_keys;
    const int32_t* p_perm;
    const int32_t* p_slot;
    const uint2* p_table;   // {start, count} per cell (current where occupied in p_map)
    BrickMap p_map;         // 2-bit brick occupancy map of the particle grid
    const uint32_t* p_key_of;  // per particle: its cell key in the last build (gridCellIds by particle)
    const float4* pvs;      // per sorted slot u, as of the last grid build: pvs[u] = (x, y, z,
                            // particle perm[u] as int bits), pvs[np + u] = (vx, vy, vz, 0)
    float4* fslot;          // split frames: per sorted slot, F after the springs, then after the contacts
    int32_t ntiles;         // split frames: slot tiles of kBlock lanes (k_contacts blocks)
    // triangle grid + vein
    GridGeom tg;
    const uint2* t_table;
    const float4* tdata;    // per sorted triangle: (v0 | id), e1, e2
    int32_t ntri;
    int32_t nv;
    const float* X;  // SoA stride nv
    const uint32_t* tidx;
    // vein hits
    float4* hits;     // 3 float4 per hit: (c0, particle), (c1, triangle), (c2, -)
    int32_t* vhead;   // per vein vertex: first hit entry (3h + corner) of its list, -1 = none
    int32_t* vnext;   // per hit entry: next entry of the same vertex, -1 = end
    uint32_t* hctrl;  // [0] hit count, [4] hits of the last frame
    const float4* t_aabb;     // per occupied triangle cell key k: [2k] = (lo, run start), [2k+1] = (hi, run count)
    const uint32_t* t_nmask;  // per triangle cell: occupied cells of its neighbourhood
    uint32_t* best;           // per particle: min (cell rank << 27 | slot) of its ray hits
    uint4* work;              // ray work items: (particle, first slot, cell rank, count), one
                              // region of region_cap items per k_raycand block
    uint32_t* wcount;         // items per region (last frame)
    uint32_t region_cap;
    uint32_t* flags;
    uint64_t* step;
};

// Vein-side arguments.
struct VeinArgs {
    int32_t nv;
    float* X;
    float* V;
    float* Fv;
    const int32_t* nbr;  // 9*nv slot-major
    const float* nbl;
    float dt, vk, vd;
    const float4* hits;
    int32_t* vhead;
    const int32_t* vnext;
    uint32_t* hctrl;
    const uint32_t* tidx;
    uint32_t* flags;
    uint64_t* step;
};

// Triangle data and cell boxes of one frame (the triangle role of k_vein_prep).
struct TriBoxArgs {
    const uint4* trun;  // per occupied cell r: {run start, cell key, count, 1}; w = 0 ends the list
    const uint4* trec;  // per sorted slot u: {i0, i1, i2, triangle}
    int32_t ntri;
    const float* X;
    int32_t nv;
    float4* tdata;
    float4* aabb;
};
// k_vein_prep: the first nsb blocks compute vein springs, the rest the triangle data.
struct VeinPrepArgs {
    VeinArgs V;
    TriBoxArgs B;
    uint32_t nsb;
};
// Processing order of the cells (k_order_key, scan, k_order_scatter).
struct OrderKeyArgs {
    TypeTable tt;
    GridGeom g;
    const float* centers;
    uint32_t* okey;
    uint32_t* ocount;
    uint64_t* ostatus;
    int32_t otiles;
    uint32_t* octrl;
};
struct OrderScatterArgs {
    int32_t nc;
    const uint32_t* okey;
    uint32_t* ocount;
    const uint32_t* ostart;
    int32_t* corder;
    uint32_t* flags;
};

}  // namespace bcs
