// step.hpp -- kernel-argument blocks and launchers of the physics step.
#pragma once

#include "grid.hpp"

namespace bcs {

constexpr int kMaxEndings = 16;
constexpr int kBlock = 256;  // threads per cell-chunk block

// Folded type tables (src/meta_factory/blood_cell_factory.hpp:172-282).
struct TypeTable {
    int32_t nt;
    int32_t np, nc, nm;
    int32_t ppc[kMaxTypes], cnt[kMaxTypes], pstart[kMaxTypes], cstart[kMaxTypes], mstart[kMaxTypes];
    int32_t end_spheres[kMaxTypes];  // vein-end variant with ending spheres (block-sync variant)
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
    uint64_t seed;
    float endc[3 * kMaxEndings];
    float endr[kMaxEndings];
};

struct StepArgs {
    TypeTable tt;
    PhysConst pc;
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
    // work chunks (whole cells of one type)
    const int32_t* chunk_type;
    const int32_t* chunk_first;
    const int32_t* chunk_n;
    // particle grid
    GridGeom pg;
    const uint32_t* p_keys;
    const int32_t* p_perm;
    const int32_t* p_slot;
    const uint32_t* p_start;
    const float* xs;
    const float* vs;
    int64_t s_stride;
    // triangle grid + vein
    GridGeom tg;
    const uint32_t* t_keys;
    const int32_t* t_perm;
    const uint32_t* t_start;
    int32_t ntri;
    int32_t nv;
    const float* X;  // SoA stride nv
    const uint32_t* tidx;
    // vein hits
    uint32_t* vcount;
    float4* hits;     // 3 float4 per hit
    uint32_t* hctrl;  // [0] hit count, [1] vein-scan ticket
    uint64_t* vstatus;
    int32_t vntiles;
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
    uint32_t* vcount;
    uint32_t* vstart;   // nv+1
    const float4* hits;
    float4* seg;        // ordered contributions: (c.xyz, key)
    uint32_t* hctrl;
    const uint32_t* tidx;
    uint32_t* flags;
    uint64_t* step;
};

}  // namespace bcs
