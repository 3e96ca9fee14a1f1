// step.hip -- physics kernels of one frame for gfx950.
//
// The particle part of a frame is three kernels.  k_raycand and k_cells give
// one block a chunk of whole blood cells of one type (<= 256 particles).
//   k_raycand  vein ray candidates: the triangle cells of the 27-cell
//              neighbourhood whose box the ray can reach, as work items of
//              <= 16 sorted triangles (needs only the frame's starting x, v
//              and the triangle data: runs beside the particle grid build)
//   k_ray      eight lanes per work item, two triangles each: first
//              Moller-Trumbore hit of the item,
//              combined per particle with atomicMin over (cell-visit rank,
//              sorted slot) = the reference's first hit in x->y->z order
//              (src/simulation/vein_collisions.cuh:60-93)
//   k_cells    centres        calculateBloodCellsCenters   src/objects/blood_cells.cu:44-61
//              springs        gatherForcesKernel (Heun)     src/objects/blood_cells.cu:66-120
//              contacts       calculateParticleCollisions   src/simulation/particle_collisions.cuh:104-269
//              hit response   detectVeinCollisions          src/simulation/vein_collisions.cu:234-275
//              integrate      propagateParticleForcesKernel src/objects/blood_cells.cu:155-179 (fused)
//              vein end       handleVeinEnds*               src/simulation/vein_end.cu:57-138 (fused)
// Reads that the reference shares between threads come from snapshots: the
// cell's F from LDS (loaded before any write), neighbour x/v from the sorted
// copies made by the grid build; x and v are only written by k_cells, which
// never reads another block's particles from x or v.
// Vein-vertex contributions of ray hits are recorded and summed per vertex
// in particle-index order by k_vein (deterministic; the reference adds them
// without atomics).
#include "step.hpp"
#include "surrogate.hpp"

namespace bcs {

__device__ __forceinline__ int32_t model_index(const TypeTable& tt, int t, int32_t id, uint32_t& fl)
{
    // bloodCellmodelStart + (id - particlesStart) % particlesInBloodCell, C remainder
    int32_t m = tt.mstart[t] + (id - tt.pstart[t]) % tt.ppc[t];
    if (m < 0) { fl |= BCS_FLAG_RADIUS_CLAMPED; m = 0; }
    if (m >= tt.nm) { fl |= BCS_FLAG_RADIUS_CLAMPED; m = tt.nm - 1; }
    return m;
}

// model_index for one launch type t held in registers (the contact search
// evaluates it for every partner of a particle): same result, the C remainder
// by multiply-shift division.
struct ModelIx {
    int32_t pstart, mstart, ppc, nm;
    Div32 div;
};
__device__ __forceinline__ ModelIx model_ix(const TypeTable& tt, int t)
{
    return ModelIx{tt.pstart[t], tt.mstart[t], tt.ppc[t], tt.nm, tt.ppc_div[t]};
}
__device__ __forceinline__ int32_t model_index(const ModelIx& mi, int32_t id, uint32_t& fl)
{
    const int32_t d = id - mi.pstart;
    const uint32_t ad = d < 0 ? (uint32_t)(-(int64_t)d) : (uint32_t)d;
    const uint32_t r = ad - udiv(ad, mi.div) * (uint32_t)mi.ppc;
    int32_t m = mi.mstart + (d < 0 ? -(int32_t)r : (int32_t)r);
    if (m < 0) { fl |= BCS_FLAG_RADIUS_CLAMPED; m = 0; }
    if (m >= mi.nm) { fl |= BCS_FLAG_RADIUS_CLAMPED; m = mi.nm - 1; }
    return m;
}

__device__ __forceinline__ int type_of_index(const TypeTable& tt, int32_t s)
{
    for (int t = 0; t < tt.nt; ++t)
        if (s >= tt.pstart[t] && s < tt.pstart[t] + tt.cnt[t] * tt.ppc[t]) return t;
    return tt.nt - 1;
}

// Chunk of this block.  Workgroups are dispatched round-robin over the 8
// XCDs, so block b runs on XCD b % 8.  The blocks of the 8 * ceil(nchunks / 8)
// grid are cut into super-groups of 8 * RUN; in each, XCD x takes a run of
// RUN consecutive chunks (the last, partial super-group splits the same way
// with shorter runs), a bijection onto [0, grid).  With the chunks in a
// spatial (Morton) order a run is one compact region, whose grid-table lines
// stay in that XCD's L2, while a crowded region of the vessel is shared out
// over all XCDs.  Blocks past the last chunk return -1.
template <uint32_t RUN = kXcdRun>
__device__ __forceinline__ int chunk_of_block(int32_t nchunks, uint32_t block)
{
    constexpr uint32_t S = 8u * RUN;
    const uint32_t nb = 8u * (((uint32_t)nchunks + 7u) >> 3);
    const uint32_t sg = block / S, w = block % S;
    const uint32_t x = w & 7u, q = w >> 3;
    const uint32_t run = (sg + 1u) * S <= nb ? RUN : (nb - sg * S) >> 3;
    const uint32_t c = sg * S + x * run + q;
    return c < (uint32_t)nchunks ? (int)c : -1;
}

// Particle of vertex a of the cl-th cell of chunk (T, first).
__device__ __forceinline__ int32_t chunk_particle(const StepArgs& A, int T, int first, int cl, int a, int p,
                                                  int32_t& cell)
{
    cell = A.corder[A.tt.cstart[T] + first + cl];
    return A.tt.pstart[T] + (cell - A.tt.cstart[T]) * p + a;
}

// springMassForceWithDampingForParticle, src/simulation/physics.cuh:24-27 (the
// Runge-Kutta build evaluates it at three perturbed states per spring).
__device__ __forceinline__ float spring_rk(const PhysConst& pc, f3 dP, f3 dv, float L, uint32_t& fl)
{
    return (len(dP) - L) * pc.k + dot(normalize(dP, fl), dv) * pc.d;
}

// Position increment of a particle, v0 -> v (propagateParticleForcesKernel,
// src/objects/blood_cells.cu:155-179): Heun 0.5 dt (v + v0), or in the
// Runge-Kutta build (k1 + 2 k2 + 2 k3 + k4) / 6 with k1 = dt v,
// k2 = dt (v + k1 / 2), k3 = dt (v + k2 / 2), k4 = dt (v + k3).
__device__ __forceinline__ f3 position_step(const PhysConst& pc, f3 v0, f3 v)
{
    if (pc.rk4) {
        const f3 k1 = smul(pc.dt, v);
        const f3 k2 = smul(pc.dt, add(v, sdiv(k1, 2.0f)));
        const f3 k3 = smul(pc.dt, add(v, sdiv(k2, 2.0f)));
        const f3 k4 = smul(pc.dt, add(v, k3));
        return sdiv(add(add(add(k1, smul(2.0f, k2)), smul(2.0f, k3)), k4), 6.0f);
    }
    return smul(0.5f * pc.dt, add(v, v0));
}

// addResilientForceOnCollision, src/simulation/physics.cuh:133-145 (increment).
__device__ __forceinline__ f3 resilient(const PhysConst& pc, f3 relPos, f3 relVel, float d2, float radius, float coef,
                                        uint32_t& fl)
{
    f3 dir = normalize(relPos, fl);
    f3 tang = sub(relVel, smul(dot(relVel, dir), dir));
    f3 spring = smul(-pc.cs * (radius * 2.0f - sqrtf(d2)), dir);
    f3 damp = smul(pc.cd, relVel);
    f3 shear = smul(pc.csh, tang);
    return smul(coef, add(add(spring, damp), shear));
}

// The correctly rounded square root of d2 and quotients v / s in the range
// where they need none of their range handling: d2 in [2^-60, 2^60] and
// every |v| component >= 2^-96, so that no operand or quotient is zero, tiny,
// huge or denormal (and a normalize() of v cannot produce a NaN).  There the
// compiler's sequences for sqrtf() and '/' reduce to what is written here --
// the hardware square-root estimate moved to the correctly rounded root by
// the signs of its two neighbours' residuals, and for n / s the refined
// reciprocal of s, then two residual corrections -- so every value equals
// the one sqrtf() and '/' give.  The reciprocal is refined once for the three
// quotients.
__device__ __forceinline__ bool inrange_ok(f3 v, float d2)
{
    return d2 >= 0x1p-60f && d2 <= 0x1p60f && fminf(fminf(fabsf(v.x), fabsf(v.y)), fabsf(v.z)) >= 0x1p-96f;
}
__device__ __forceinline__ float root_inrange(float d2)
{
    const float s0 = __builtin_amdgcn_sqrtf(d2);
    const float sdn = __uint_as_float(__float_as_uint(s0) - 1u), sup = __uint_as_float(__float_as_uint(s0) + 1u);
    const float edn = fmaf(-sdn, s0, d2), eup = fmaf(-sup, s0, d2);
    const float s = edn <= 0.0f ? sdn : s0;
    return eup > 0.0f ? sup : s;
}
__device__ __forceinline__ f3 quot_inrange(f3 v, float s)
{
    float r = __builtin_amdgcn_rcpf(s);
    r = fmaf(fmaf(-s, r, 1.0f), r, r);
    auto quo = [&](float n) {
        const float q0 = n * r;
        const float q1 = fmaf(fmaf(-s, q0, n), r, q0);
        return fmaf(fmaf(-s, q1, n), r, q1);
    };
    return mk(quo(v.x), quo(v.y), quo(v.z));
}

// resilient() for a contact within that range (inrange_ok(relPos, d2)): the
// same values.
__device__ __forceinline__ f3 resilient_inrange(const PhysConst& pc, f3 relPos, f3 relVel, float d2, float radius,
                                                float coef)
{
    const float s = root_inrange(d2);
    const f3 dir = quot_inrange(relPos, s);
    f3 tang = sub(relVel, smul(dot(relVel, dir), dir));
    f3 spring = smul(-pc.cs * (radius * 2.0f - s), dir);
    f3 damp = smul(pc.cd, relVel);
    f3 shear = smul(pc.csh, tang);
    return smul(coef, add(add(spring, damp), shear));
}

// realCollisionDetection, src/simulation/vein_collisions.cu:11-45, with
// edge1 = v1 - v0 and edge2 = v2 - v0 precomputed (the identical float
// subtractions, done once per frame by k_vein_prep).  Before the correctly
// rounded 1/a, two exact rejections of u = RN(RN(1/a) * d), d = dot(s, h):
//   sign(d) != sign(a), d != 0 and |d| >= |a| 2^-100  =>  u < 0 (no underflow)
//   same signs and |d| > 1.001 |a|                   =>  u > 1 (rounding < 2^-22)
// so exactly the candidates the reference rejects at this test are skipped.
__device__ __forceinline__ bool ray_tri_t(f3 v0, f3 edge1, f3 edge2, f3 orig, f3 dir, float& t_out)
{
    const float EPS = 0.000001f;
    f3 h = cross(dir, edge2);
    float a = dot(edge1, h);
    if (a > -EPS && a < EPS) return false;
    f3 s = sub(orig, v0);
    const float d = dot(s, h);
    const float aa = fabsf(a), ad = fabsf(d);
    const bool opposite = (a < 0.0f) != (d < 0.0f);
    if (d != 0.0f && opposite && ad >= aa * 0x1p-100f) return false;
    if (!opposite && ad > aa * 1.001f) return false;
    float f = 1.0f / a;
    float u = f * d;
    if (u < 0.0f || u > 1.0f) return false;
    f3 q = cross(s, edge1);
    float v = f * dot(dir, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = f * dot(edge2, q);
    t_out = t;
    return t > EPS;
}

// The hit of ray_tri_t with the reference's normal and reflected direction.
__device__ __forceinline__ bool ray_tri_pre(f3 v0, f3 edge1, f3 edge2, f3 orig, f3 dir, float& t_out, f3& normal,
                                            f3& refl, uint32_t& fl)
{
    float t;
    if (!ray_tri_t(v0, edge1, edge2, orig, dir, t)) return false;
    t_out = t;
    normal = normalize(cross(edge2, edge1), fl);
    refl = sub(dir, smul(2.0f * dot(dir, normal), normal));
    return true;
}

// calculateBaricentric (edges v1-v0, v2-v1), src/simulation/vein_collisions.cu:47-61.
__device__ __forceinline__ f3 bary(f3 point, f3 v0, f3 v1, f3 v2)
{
    f3 e0 = sub(v1, v0), e1 = sub(v2, v1), e2 = sub(point, v0);
    float d00 = dot(e0, e0), d01 = dot(e0, e1), d11 = dot(e1, e1), d20 = dot(e2, e0), d21 = dot(e2, e1);
    float denom = d00 * d11 - d01 * d01;
    f3 r;
    r.x = (d11 * d20 - d01 * d21) / denom;
    r.y = (d00 * d21 - d01 * d20) / denom;
    r.z = 1.0f - r.x - r.y;
    return r;
}

// Conservative ray / box test (slab method) on a box inflated by `m`: false
// only if the ray (t >= 0) certainly misses every point of the box.  Used to
// skip triangle cells that cannot hold a hit; the Moller-Trumbore test itself
// is unchanged, so the first hit in visit order is unchanged.
__device__ __forceinline__ bool ray_box(f3 o, f3 inv, f3 dir, float4 lo, float4 hi, float m)
{
    float tmin = 0.0f, tmax = 3.0e38f;
    const float os[3] = {o.x, o.y, o.z}, is[3] = {inv.x, inv.y, inv.z}, ds[3] = {dir.x, dir.y, dir.z};
    const float ls[3] = {lo.x - m, lo.y - m, lo.z - m}, hs[3] = {hi.x + m, hi.y + m, hi.z + m};
    bool out = false;  // branch-free: selects instead of a branch per axis
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const bool flat = ds[d] == 0.0f;  // ray parallel to the slab: inside it or a miss
        out = out || (flat && (os[d] < ls[d] || os[d] > hs[d]));
        const float t1 = (ls[d] - os[d]) * is[d], t2 = (hs[d] - os[d]) * is[d];
        tmin = flat ? tmin : fmaxf(tmin, fminf(t1, t2));
        tmax = flat ? tmax : fminf(tmax, fmaxf(t1, t2));
    }
    return !out && tmax >= tmin;
}

// Vein-end condition of one particle (src/simulation/vein_end.cu:57-90,
// thresholds :12-18).
__device__ __forceinline__ bool vein_end_flag(const PhysConst& pc, bool spheres, f3 p)
{
    bool tele = false;
    if (spheres) {
        for (int e = 0; e < pc.nend; ++e) {
            f3 dd = mk(p.x - pc.endc[3 * e], p.y - pc.endc[3 * e + 1], p.z - pc.endc[3 * e + 2]);
            tele = tele || dot(dd, dd) <= pc.endr[e] * pc.endr[e];
        }
    }
    tele = tele || p.y <= pc.thr_lower || p.y >= pc.thr_upper || p.x <= pc.thr_left || p.x >= pc.thr_right ||
           p.z <= pc.thr_back || p.z >= pc.thr_front;
    return tele;
}

// Teleported position of model vertex a of a cell (vein_end.cu:103-106):
// ((U - 0.5) * 1.2) * R + m_a - m_0 in x and z, minSpawnY + m_a.y - m_0.y in y.
__device__ __forceinline__ f3 teleport_pos(const PhysConst& pc, const float* model, int m0, int a, float u1, float u2)
{
    const float* ma = model + 3 * (m0 + a);
    const float* mz = model + 3 * m0;
    f3 r;
    r.x = (u1 - 0.5f) * 1.2f * pc.cyl_r + ma[0] - mz[0];
    r.y = pc.spawn_y + ma[1] - mz[1];
    r.z = (u2 - 0.5f) * 1.2f * pc.cyl_r + ma[2] - mz[2];
    return r;
}

constexpr uint32_t kNoHit = 0xFFFFFFFFu;
constexpr uint32_t kScanSerial = 0xFFFFFFFEu;  // work list full: k_cells scans serially
constexpr uint32_t kSlotBits = 27;             // best = rank << 27 | sorted triangle slot
constexpr uint32_t kItemTris = 16;
constexpr uint32_t kRayLanes = 8;  // k_ray lanes per work item
// k_ray blocks per work region: about kRayBlocks blocks in all, 1 .. 8 per
// region (few regions: more blocks share one; many: one block each)
constexpr uint32_t kRayBlocks = 8192, kRaySplitMax = 8;
constexpr uint32_t kBoxBatch = 8;  // triangle-cell boxes k_raycand loads at once
constexpr uint32_t kRayLdsItems = 2048;  // work items a k_raycand_ray block holds (8 bytes each)
static_assert(kBlock <= 256 && kItemTris < 32, "an LDS ray item packs lane (8 bits), rank (5), count (5)");

// Triangle-grid neighbourhood of a particle (face clipped, unsigned ids as
// in src/simulation/vein_collisions.cu:80-232).
struct RayCells {
    uint32_t cellId;
    uint32_t xId, yId, zId;  // unclamped (saturated) cell coordinates
    int xl, xh, yl, yh, zl, zh;
    int64_t cX, cXY;
};

__device__ __forceinline__ RayCells ray_cells(const GridGeom& g, f3 x1)
{
    RayCells rc;
    uint32_t dummy = 0;
    rc.cellId = cell_key(g, x1.x, x1.y, x1.z, dummy);
    rc.xId = sat_u32((x1.x - g.minv[0]) / g.cell[0]);
    rc.yId = sat_u32((x1.y - g.minv[1]) / g.cell[1]);
    rc.zId = sat_u32((x1.z - g.minv[2]) / g.cell[2]);
    axis_range_u(rc.xId, g.cnt[0], rc.xl, rc.xh);
    axis_range_u(rc.yId, g.cnt[1], rc.yl, rc.yh);
    axis_range_u(rc.zId, g.cnt[2], rc.zl, rc.zh);
    rc.cX = g.cnt[0];
    rc.cXY = (int64_t)g.cnt[0] * g.cnt[1];
    return rc;
}

// Occupied triangle cells of the neighbourhood as a mask over the visit rank
// r = (dx+1)*9 + (dy+1)*3 + (dz+1).  When the particle's cell coordinates are
// those of its (clamped) key and every axis has at least two cells, the
// neighbourhood is the cell's own and the mask precomputed per cell
// (k_tri_nmask) is returned; otherwise it is formed key by key.
__device__ __forceinline__ uint32_t tri_occupied(const StepArgs& A, const RayCells& rc)
{
    const GridGeom& g = A.tg;
    uint32_t xc, yc, zc;
    grid_coords(g, rc.cellId, xc, yc, zc);
    if (xc == rc.xId && yc == rc.yId && zc == rc.zId && g.cnt[0] >= 2 && g.cnt[1] >= 2 && g.cnt[2] >= 2)
        return A.t_nmask[rc.cellId];
    uint32_t m = 0;
    for (int dx = rc.xl; dx <= rc.xh; ++dx)
        for (int dy = rc.yl; dy <= rc.yh; ++dy)
            for (int dz = rc.zl; dz <= rc.zh; ++dz) {
                const int64_t nk = (int64_t)rc.cellId + dz * rc.cXY + dy * rc.cX + dx;
                if (nk < 0 || (uint64_t)nk >= g.cell_count) continue;
                if (A.t_table[nk].y != 0u) m |= 1u << ((dx + 1) * 9 + (dy + 1) * 3 + (dz + 1));
            }
    return m;
}

// Table entries of the 9 cells of x-slab dx (independent loads) and the mask
// of occupied cells whose box the ray can reach.
__device__ __forceinline__ uint32_t slab_cells(const StepArgs& A, const RayCells& rc, int dx, f3 x1, f3 inv, f3 dir,
                                               float bm, uint2 (&ent)[9])
{
    const GridGeom& g = A.tg;
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        const int dy = c / 3 - 1, dz = c % 3 - 1;
        const int64_t nk = (int64_t)rc.cellId + dz * rc.cXY + dy * rc.cX + dx;
        const bool in = dx >= rc.xl && dx <= rc.xh && dy >= rc.yl && dy <= rc.yh && dz >= rc.zl && dz <= rc.zh &&
                        nk >= 0 && (uint64_t)nk < g.cell_count;
        ent[c] = in ? A.t_table[nk] : make_uint2(0u, 0u);
    }
    uint32_t mask = 0;
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        if (ent[c].y != 0u && ent[c].x < (uint32_t)A.ntri) {
            const int64_t nk = (int64_t)rc.cellId + (c % 3 - 1) * rc.cXY + (c / 3 - 1) * rc.cX + dx;
            const float4 lo = A.t_aabb[2 * nk], hi = A.t_aabb[2 * nk + 1];
            if (ray_box(x1, inv, dir, lo, hi, bm)) mask |= 1u << c;
        }
    }
    return mask;
}

// Neighbour-cell ranges (clipped to the grid) of a particle.
struct NbrRange {
    int xl, xh, yl, yh, zl, zh;
};

// Occupied cells of the 27-neighbourhood of cell `cellId` as a mask over the
// visit rank r = (dx+1)*9 + (dy+1)*3 + (dz+1) (the reference's x, y, z loop
// order), restricted to the clipped ranges and to keys in [0, cellCount).  The
// reference forms neighbour keys as cellId + dz*cX*cY + dy*cX + dx; when every
// allowed offset stays inside the grid on each axis (the usual case) that key
// is the cell (x+dx, y+dy, z+dz) and the brick map is read by coordinates,
// else every key is decoded on its own.  The own cell is dropped when it holds
// exactly one object and that object is the particle itself (own_is_i).
__device__ __forceinline__ uint32_t neighbour_mask(const BrickMap& bm, const GridGeom& g, uint32_t cellId,
                                                   const NbrRange& nr, int64_t cX, int64_t cXY, bool own_is_i)
{
    uint32_t x, y, z;
    key_coords(bm, cellId, x, y, z);
    uint32_t mask = 0;
    const bool inside = (int64_t)x + nr.xl >= 0 && (int64_t)x + nr.xh < bm.cnt[0] && (int64_t)y + nr.yl >= 0 &&
                        (int64_t)y + nr.yh < bm.cnt[1] && (int64_t)z + nr.zl >= 0 && (int64_t)z + nr.zh < bm.cnt[2];
    if (inside) {
        const uint32_t xa = x + nr.xl, xb = x + nr.xh;
        // the words of the low and high x ends of the 9 (dy, dz) rows, all
        // loaded before any is used (one round trip); a row clipped away
        // loads the particle's own row instead and is skipped below
        uint32_t wa_[9], wb_[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            const int dy = c / 3 - 1, dz = c % 3 - 1;
            const bool row = dy >= nr.yl && dy <= nr.yh && dz >= nr.zl && dz <= nr.zh;
            const uint32_t yy = row ? y + dy : y, zz = row ? z + dz : z;
            wa_[c] = bm.words[brick_word(bm, xa, yy, zz)];
            wb_[c] = bm.words[brick_word(bm, xb, yy, zz)];
        }
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            const int dy = c / 3 - 1, dz = c % 3 - 1;
            if (dy < nr.yl || dy > nr.yh || dz < nr.zl || dz > nr.zh) continue;
            const uint32_t wa = wa_[c], wb = wb_[c];  // the same word when xa and xb share a brick
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                if (dx < nr.xl || dx > nr.xh) continue;
                const uint32_t xx = x + dx;
                const uint32_t w = ((xx ^ xa) >> 3) ? wb : wa;
                const uint32_t code = (w >> (2u * (xx & 7u))) & 3u;
                const bool skip_one = own_is_i && dx == 0 && dy == 0 && dz == 0;
                if (code > (skip_one ? 1u : 0u)) mask |= 1u << ((dx + 1) * 9 + c);
            }
        }
    } else {
        for (int dx = nr.xl; dx <= nr.xh; ++dx)
            for (int dy = nr.yl; dy <= nr.yh; ++dy)
                for (int dz = nr.zl; dz <= nr.zh; ++dz) {
                    const int64_t nk = (int64_t)cellId + dz * cXY + dy * cX + dx;
                    if (nk < 0 || (uint64_t)nk >= g.cell_count) continue;
                    const uint32_t code = cell_code(bm, (uint32_t)nk);
                    const bool skip_one = own_is_i && dx == 0 && dy == 0 && dz == 0;
                    if (code > (skip_one ? 1u : 0u)) mask |= 1u << ((dx + 1) * 9 + (dy + 1) * 3 + (dz + 1));
                }
    }
    return mask;
}

__device__ __forceinline__ f3 inv_dir(f3 dir)
{
    return mk(dir.x != 0.0f ? 1.0f / dir.x : 0.0f, dir.y != 0.0f ? 1.0f / dir.y : 0.0f,
              dir.z != 0.0f ? 1.0f / dir.z : 0.0f);
}

__device__ __forceinline__ float box_margin(const StepArgs& A) { return A.pc.aabb_margin; }

// Serial first-hit scan of one ray (used only when the work list is full).
__device__ bool ray_first_hit_serial(const StepArgs& A, f3 x1, f3 dir, float& rt, f3& normal, f3& refl, int32_t& obj,
                                     uint32_t& fl)
{
    if (!(dir.x != 0.0f || dir.y != 0.0f || dir.z != 0.0f)) return false;
    const RayCells rc = ray_cells(A.tg, x1);
    const f3 inv = inv_dir(dir);
    const float bm = box_margin(A);
    for (int dx = -1; dx <= 1; ++dx) {
        uint2 ent[9];
        const uint32_t mask = slab_cells(A, rc, dx, x1, inv, dir, bm, ent);
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            if (!((mask >> c) & 1u)) continue;
            const uint32_t u1 = min(ent[c].x + ent[c].y, (uint32_t)A.ntri);
            for (uint32_t u = ent[c].x; u < u1; ++u) {
                const float4 a = A.tdata[3 * (size_t)u], e1 = A.tdata[3 * (size_t)u + 1], e2 = A.tdata[3 * (size_t)u + 2];
                if (ray_tri_pre(mk(a.x, a.y, a.z), mk(e1.x, e1.y, e1.z), mk(e2.x, e2.y, e2.z), x1, dir, rt, normal, refl,
                                fl)) {
                    obj = __float_as_int(a.w);
                    return true;
                }
            }
        }
    }
    return false;
}


// Lane g of the kRayLanes lanes of work item `it` tests the item's triangles
// it.y + g*K .. + K-1 (K = kItemTris / kRayLanes), their loads issued
// together; returns the first hitting k, or K.
__device__ __forceinline__ uint32_t item_lane_first_hit(const StepArgs& A, uint4 it, uint32_t g, f3 x1, f3 dir)
{
    constexpr uint32_t K = kItemTris / kRayLanes;
    const uint32_t u0 = it.y + g * K;
    const uint32_t u1 = min(it.y + it.w, (uint32_t)A.ntri);
    float4 ta[K], te1[K], te2[K];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        if (u0 + k < u1) {
            ta[k] = A.tdata[3 * (size_t)(u0 + k)];
            te1[k] = A.tdata[3 * (size_t)(u0 + k) + 1];
            te2[k] = A.tdata[3 * (size_t)(u0 + k) + 2];
        }
    }
    uint32_t first = K;
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        float rt;
        if (first == K && u0 + k < u1 &&
            ray_tri_t(mk(ta[k].x, ta[k].y, ta[k].z), mk(te1[k].x, te1[k].y, te1[k].z), mk(te2[k].x, te2[k].y, te2[k].z),
                      x1, dir, rt))
            first = k;
    }
    return first;
}

// The item's first hit in sorted order is the first hit of its lowest hitting
// lane (the reference's scan stops there).  Every lane of the wave calls this;
// returns whether this lane holds the item's first hit.
__device__ __forceinline__ bool item_first_lane(bool hit, uint32_t g)
{
    constexpr uint32_t G = kRayLanes;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t m = __ballot(hit);
    const uint64_t gm = (m >> (lane & ~(G - 1u))) & (G == 64 ? ~0ull : ((1ull << G) - 1ull));
    return hit && g == (uint32_t)__builtin_ctzll(gm);
}

// Vein ray candidates (part of detectVeinCollisions, src/simulation/vein_collisions.cu:181-233):
// per particle, the occupied triangle cells of its 27-neighbourhood whose box
// the ray from x along v/|v| can reach, split into work items of <= 16 sorted
// triangles written to this block's own region of the work list.  Needs only
// the frame's starting x, v and the triangle data, so it runs beside the
// particle grid build.
template <bool FUSED>
__device__ __forceinline__ void raycand_block(const StepArgs& A, VBlock vb)
{
    __shared__ uint32_t s_wsum[kBlock / 64];
    __shared__ uint32_t s_end;
    const TypeTable& tt = A.tt;
    const int l = threadIdx.x;
    const int chunk = chunk_of_block(A.nchunks, vb.id);
    if (chunk < 0) {
        if (l == 0) A.wcount[vb.id] = 0u;
        return;
    }
    // the lane's particle (-1: no slot), independent of the chunk's tables
    const int2 pcl = A.chunk_pi[(size_t)chunk * kBlock + l];
    const bool act = pcl.x >= 0;
    const int64_t np = tt.np;
    const int32_t i = act ? pcl.x : 0;
    uint32_t fl = 0;
    uint32_t nitems = 0, pmask = 0;
    f3 dir = mk(0.f, 0.f, 0.f), x1 = dir;
    RayCells rc{};
    if (act) {
        x1 = load3(A.pos, np, i);
        const f3 v1 = load3(A.vel, np, i);
        // ---- vein ray candidates: occupied triangle cells whose box the ray
        // can reach (pmask), split into work items of <= 16 sorted triangles
        dir = normalize(v1, fl);
        if (dir.x != 0.0f || dir.y != 0.0f || dir.z != 0.0f) {
            rc = ray_cells(A.tg, x1);
            const uint32_t om = tri_occupied(A, rc);
            if (om) {
                const f3 inv = inv_dir(dir);
                const float bm = box_margin(A);
                // boxes of up to kBoxBatch occupied cells loaded together,
                // then tested; a passing cell's triangle count is in hi.w
                uint32_t m = om;
                while (m) {
                    float4 lo[kBoxBatch], hi[kBoxBatch];
                    int rr[kBoxBatch];
#pragma unroll
                    for (int k = 0; k < (int)kBoxBatch; ++k) {
                        rr[k] = -1;
                        if (m) {
                            const int r = __builtin_ctz(m);
                            m &= m - 1u;
                            rr[k] = r;
                            const int dx = r / 9 - 1, c = r % 9;
                            const int64_t nk = (int64_t)rc.cellId + (c % 3 - 1) * rc.cXY + (c / 3 - 1) * rc.cX + dx;
                            lo[k] = A.t_aabb[2 * nk];
                            hi[k] = A.t_aabb[2 * nk + 1];
                        }
                    }
#pragma unroll
                    for (int k = 0; k < (int)kBoxBatch; ++k) {
                        const bool pass = rr[k] >= 0 && ray_box(x1, inv, dir, lo[k], hi[k], bm);
                        pmask |= pass ? 1u << (rr[k] & 31) : 0u;
                        nitems += pass ? (min(__float_as_uint(hi[k].w), (uint32_t)A.ntri) + kItemTris - 1) / kItemTris : 0u;
                    }
                }
            }
        }
    }
    // ---- work-list space: every block owns region vb.id of region_cap
    // items; offsets from a block-wide exclusive scan (no global atomics)
    const int lane = l & 63, wv = l >> 6;
    uint32_t incl = nitems;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63) s_wsum[wv] = incl;
    if (l == 0) s_end = 0u;
    __syncthreads();
    uint32_t wave_off = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w)
        if (w < wv) wave_off += s_wsum[w];
    const uint32_t off = wave_off + incl - nitems;  // within the region
    // particles whose items fit form a prefix of the scan order; the region's
    // count ends with the last of them (no unwritten slots below it)
    const uint32_t cap = FUSED ? min(A.region_cap, kRayLdsItems) : A.region_cap;
    const bool fits = (uint64_t)off + nitems <= (uint64_t)cap;
    if (act && nitems > 0 && fits) atomicMax(&s_end, off + nitems);
    if (!FUSED) {
        __syncthreads();
        if (l == 0) A.wcount[vb.id] = s_end;
        if (act) {
            A.best[i] = (nitems > 0 && !fits) ? kScanSerial : kNoHit;
            if (nitems > 0 && !fits) fl |= BCS_FLAG_RAY_SERIAL;
            if (nitems > 0 && fits) {
                uint4* out = A.work + (size_t)vb.id * A.region_cap;
                uint32_t w = off;
                uint32_t m = pmask;
                while (m) {
                    const int r = __builtin_ctz(m);
                    m &= m - 1u;
                    const int dx = r / 9 - 1, c = r % 9;
                    const int64_t nk = (int64_t)rc.cellId + (c % 3 - 1) * rc.cXY + (c / 3 - 1) * rc.cX + dx;
                    const uint2 ent = A.t_table[nk];
                    for (uint32_t u = ent.x; u < ent.x + ent.y && w < off + nitems; u += kItemTris)
                        out[w++] = make_uint4((uint32_t)i, u, (uint32_t)r, min(kItemTris, ent.x + ent.y - u));
                }
            }
        }
    } else {
        // the block's items, rays and per-particle minima stay in LDS; the
        // block tests its own items (kRayLanes lanes per item)
        // item: {lane | rank << 8 | count << 13, first sorted slot}
        __shared__ uint2 s_items[kRayLdsItems];
        __shared__ float s_ray[6][kBlock];
        __shared__ uint32_t s_best[kBlock];
        s_best[l] = (act && nitems > 0 && !fits) ? kScanSerial : kNoHit;
        if (act && nitems > 0 && !fits) fl |= BCS_FLAG_RAY_SERIAL;
        s_ray[0][l] = x1.x; s_ray[1][l] = x1.y; s_ray[2][l] = x1.z;
        s_ray[3][l] = dir.x; s_ray[4][l] = dir.y; s_ray[5][l] = dir.z;
        if (act && nitems > 0 && fits) {
            uint32_t w = off;
            uint32_t m = pmask;
            while (m) {
                const int r = __builtin_ctz(m);
                m &= m - 1u;
                const int dx = r / 9 - 1, c = r % 9;
                const int64_t nk = (int64_t)rc.cellId + (c % 3 - 1) * rc.cXY + (c / 3 - 1) * rc.cX + dx;
                const uint2 ent = A.t_table[nk];
                for (uint32_t u = ent.x; u < ent.x + ent.y && w < off + nitems; u += kItemTris)
                    s_items[w++] = make_uint2((uint32_t)l | (uint32_t)r << 8 | min(kItemTris, ent.x + ent.y - u) << 13, u);
            }
        }
        __syncthreads();
        const uint32_t total = s_end;
        if (l == 0) A.wcount[vb.id] = total;
        constexpr uint32_t G = kRayLanes, K = kItemTris / kRayLanes;
        const uint32_t g = l % G;
        // base is block-uniform, so every lane reaches the ballot
        for (uint32_t base = 0; base < total; base += kBlock / G) {
            const uint32_t w = base + l / G;
            uint32_t first = K;
            uint4 it = make_uint4(0u, 0u, 0u, 0u);
            if (w < total) {
                const uint2 q = s_items[w];
                it = make_uint4(q.x & 0xFFu, q.y, (q.x >> 8) & 31u, q.x >> 13);
                const uint32_t pl = it.x;
                first = item_lane_first_hit(A, it, g, mk(s_ray[0][pl], s_ray[1][pl], s_ray[2][pl]),
                                            mk(s_ray[3][pl], s_ray[4][pl], s_ray[5][pl]));
            }
            if (item_first_lane(first < K, g)) atomicMin(&s_best[it.x], (it.z << kSlotBits) | (it.y + g * K + first));
        }
        __syncthreads();
        if (act) A.best[i] = s_best[l];
    }
    if (fl) atomicOr(A.flags, fl);
}

__global__ void __launch_bounds__(kBlock) k_raycand(StepArgs A) { raycand_block<false>(A, hw_block()); }

// k_raycand and k_ray in one launch: each block keeps its work items in LDS
// (at most kRayLdsItems; rays that do not fit are scanned serially by k_cells).
__global__ void __launch_bounds__(kBlock) k_raycand_ray(StepArgs A) { raycand_block<true>(A, hw_block()); }

template <bool FUSED>
__device__ __forceinline__ void respond_integrate(const StepArgs& A, bool act, int32_t i, int T, int a, int cl,
                                                  int32_t cell, f3 x1, f3 v1, f3 F, uint32_t fl, int* s_tele);

static_assert(sizeof(StepArgs) <= 4096, "kernel arguments exceed 4 KiB");

// Contact forces on particle i from the occupied neighbour cells in `mask`
// (rank r = (dx+1)*9 + (dy+1)*3 + (dz+1), from the occupancy map), visited in
// the reference order (x, then y, then z; calculateParticleCollisions,
// src/simulation/particle_collisions.cuh:104-269): each contribution is added
// to F in visit order, slot by slot.  Tl is the launch type of i: the radius
// of every partner is looked up through Tl's model.  Returns the number of
// contacts.
// How the contact search drops neighbour cells out of reach: not at all
// (stale grids), per lane, or per grid cell (the union over the lanes of a
// wave that share the particle's cell, which keeps their candidate streams,
// and so their loads and trip counts, identical; lanes in sorted-slot order).
constexpr int kPruneNone = 0, kPruneLane = 1, kPruneCell = 2;

template <int PRUNE, class Radius>
__device__ __forceinline__ uint32_t contact_sum(const StepArgs& A, int32_t i, int Tl, uint32_t key_of, uint32_t mask,
                                                f3 x1, f3 v1, f3& F, uint32_t& fl, Radius radius, uint32_t* s_hit)
{
    const TypeTable& tt = A.tt;
    const PhysConst& pc = A.pc;
    const int64_t np = tt.np;
    const GridGeom& g = A.pg;
    const int64_t cX = g.cnt[0], cXY = (int64_t)g.cnt[0] * g.cnt[1];
    // table entries {first slot, end slot} of the next kCellBatch occupied
    // cells in visit order (empty: {0, 0}), taken off mask: the raw
    // {start, count} entries are all loaded first (one round trip)
    auto load_cells = [&](uint2 (&ent)[kCellBatch]) {
#pragma unroll
        for (int k = 0; k < kCellBatch; ++k) {
            ent[k] = make_uint2(0u, 0u);
            if (mask) {
                const int r = __builtin_ctz(mask);
                mask &= mask - 1u;
                const int dx = r / 9 - 1, c = r % 9;
                ent[k] = A.p_table[(int64_t)key_of + (c % 3 - 1) * cXY + (c / 3 - 1) * cX + dx];
            }
        }
#pragma unroll
        for (int k = 0; k < kCellBatch; ++k) ent[k].y = min(ent[k].x + ent[k].y, (uint32_t)np);
    };
    const ModelIx mi = model_ix(tt, Tl);
    const float r1 = radius(model_index(mi, i, fl));
    // A partner farther than r1 + rmax (rmax: the largest radius of the
    // table) cannot be in contact: d2 > rb^2 implies d2 > (r1 + r2)^2, every
    // rounding being monotonic on these non-negative values, so its radius is
    // not looked up (nor a model index of it outside the table flagged, as
    // the oracle does).
    const float rb = r1 + pc.rmax;
    const float bound2 = rb * rb;
    // With a grid built from the current positions (every fused frame),
    // neighbour cells whose box lies farther than rb from x1 hold no such
    // partner and can be dropped from the mask.  The gaps to the cell faces
    // are shrunk by m (which covers the rounding of every position's cell
    // index and of the gaps themselves) and the squared distance gets a
    // relative slack, so a cell is only dropped when it certainly holds no
    // partner.  On a stale grid (the phase API after propagateAll) partners
    // may have left their cells, and every cell is searched.
    if (PRUNE != kPruneNone) {
        const int32_t id[3] = {sat_i32((x1.x - g.minv[0]) / g.cell[0]), sat_i32((x1.y - g.minv[1]) / g.cell[1]),
                               sat_i32((x1.z - g.minv[2]) / g.cell[2])};
        const bool own = id[0] >= 0 && id[0] < g.cnt[0] && id[1] >= 0 && id[1] < g.cnt[1] && id[2] >= 0 &&
                         id[2] < g.cnt[2] && (int64_t)id[0] + id[1] * cX + id[2] * cXY == (int64_t)key_of;
        uint32_t keep = 0x7FFFFFFu;  // all 27
        if (own && bound2 < 3.0e38f) {
            const float xs[3] = {x1.x, x1.y, x1.z};
            float gl[3], gh[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float m = 1e-6f * (fabsf(g.minv[a]) + (float)g.cnt[a] * g.cell[a]) + 1e-6f;
                const float f = xs[a] - (g.minv[a] + (float)id[a] * g.cell[a]);
                gl[a] = fmaxf(f - m, 0.0f);
                gh[a] = fmaxf(g.cell[a] - f - m, 0.0f);
            }
            const float lim = bound2 * 1.00001f;
            keep = 0u;
#pragma unroll
            for (int r = 0; r < 27; ++r) {
                const int dx = r / 9 - 1, dy = (r / 3) % 3 - 1, dz = r % 3 - 1;
                const float gx = dx < 0 ? gl[0] : (dx > 0 ? gh[0] : 0.0f);
                const float gy = dy < 0 ? gl[1] : (dy > 0 ? gh[1] : 0.0f);
                const float gz = dz < 0 ? gl[2] : (dz > 0 ? gh[2] : 0.0f);
                if (gx * gx + gy * gy + gz * gz <= lim) keep |= 1u << r;
            }
        }
        if (PRUNE == kPruneCell) {
            // union over the lanes of the same cell: they are adjacent
            // (sorted slots, equal keys), so a segmented OR scan up and down
            // the wave (lanes past the end hold key 0xFFFFFFFF)
            const int lane = threadIdx.x & 63;
            uint32_t up = keep, dn = keep;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t ku = __shfl_up(key_of, o, 64), vu = __shfl_up(up, o, 64);
                const uint32_t kd = __shfl_down(key_of, o, 64), vd = __shfl_down(dn, o, 64);
                if (lane >= o && ku == key_of) up |= vu;
                if (lane + o < 64 && kd == key_of) dn |= vd;
            }
            keep = up | dn;
        }
        mask &= keep;
    }
    uint32_t ncontact = 0;
    bool more = mask != 0u;
    // The occupied cells in visit order, kCellBatch at a time: their
    // table entries are loaded together, then each cell's run of
    // sorted-slot records is read 4 slots per load batch.
    // Contributions are added in visit order, slot by slot.
    while (more) {
        uint2 ent[kCellBatch];
        load_cells(ent);
        // One stream of candidate slots over these cells in visit order
        // (stream position p lies in cell k when cum[k] <= p < cum[k+1]);
        // each load batch takes the next kContactBatch of them, across
        // cell boundaries (a cell of one or two objects costs no round
        // trip of its own).
        uint32_t cum[kCellBatch + 1];
        cum[0] = 0u;
#pragma unroll
        for (int k = 0; k < kCellBatch; ++k) cum[k + 1] = cum[k] + (ent[k].y > ent[k].x ? ent[k].y - ent[k].x : 0u);
        const uint32_t total = cum[kCellBatch];
        // Windows of kContactWindow stream positions.  Phase 1 tests them,
        // kContactBatch records at a time, and lists the slots of the
        // partners in contact (in LDS, s_hit[k * kBlock + lane]).  Phase 2
        // then takes each lane's list in stream order, adding the contacts
        // to F in visit order: the wave runs the force code once per contact
        // of its busiest lane, instead of once per stream position at which
        // any lane has a contact.
        // Stream cursor: the current run holds positions up to rend[0]
        // (exclusive), at slot = position + rd[0]; the runs after it follow
        // in rd[1..], rend[1..].  Positions only grow, so a position past
        // the current run's end shifts the next run in (an empty run is
        // passed over the same way).
        uint32_t rd[kCellBatch], rend[kCellBatch];
#pragma unroll
        for (int k = 0; k < kCellBatch; ++k) {
            rd[k] = ent[k].x - cum[k];
            rend[k] = cum[k + 1];
        }
        auto cursor_slot = [&](uint32_t pp) {
#pragma unroll
            for (int s = 0; s < kCellBatch - 1; ++s) {  // at most kCellBatch - 1 shifts (pp < total)
                if (pp < rend[0]) break;
#pragma unroll
                for (int k = 0; k < kCellBatch - 1; ++k) {
                    rd[k] = rd[k + 1];
                    rend[k] = rend[k + 1];
                }
            }
            return pp + rd[0];
        };
        uint32_t* hit_list = s_hit + threadIdx.x;
        for (uint32_t p0 = 0; p0 < total; p0 += kContactWindow) {
            uint32_t nh = 0u;
#pragma unroll
            for (int g = 0; g < kContactWindow / kContactBatch; ++g) {
                const uint32_t b0 = p0 + (uint32_t)(g * kContactBatch);
                if (b0 >= total) break;
                const int nval = (int)min((uint32_t)kContactBatch, total - b0);
                uint32_t slot[kContactBatch];
#pragma unroll
                for (int t = 0; t < kContactBatch; ++t) slot[t] = t < nval ? cursor_slot(b0 + (uint32_t)t) : 0u;
                float4 q[kContactBatch];
#pragma unroll
                for (int t = 0; t < kContactBatch; ++t)
                    if (t < nval) q[t] = A.pvs[slot[t]];
#pragma unroll
                for (int t = 0; t < kContactBatch; ++t) {
                    if (t >= nval) continue;
                    const int32_t j = __float_as_int(q[t].w);
                    // (with the grid of this frame x1 is the particle's own
                    // record, whose d2 = 0 never makes a contact below)
                    if (PRUNE != kPruneCell && j == i) continue;
                    const f3 rel = sub(x1, mk(q[t].x, q[t].y, q[t].z));
                    const float d2 = dot(rel, rel);
                    if (d2 <= bound2) {
                        // with one radius throughout, minD = r1 + rmax and
                        // minD^2 is bound2 itself
                        bool touch = d2 >= 0.0001f;
                        if (!pc.r_uniform) {
                            const float r2 = radius(model_index(mi, j, fl));
                            const float minD = r1 + r2;
                            touch = touch && d2 <= minD * minD;
                        }
                        if (touch) {
                            hit_list[nh * kBlock] = slot[t];
                            ++nh;
                        }
                    }
                }
            }
            // the next contact's record is loaded before this one's force
            // is computed (the loads of contact h + 1 overlap the math of h)
            float4 qn = make_float4(0.f, 0.f, 0.f, 0.f), wn = qn;
            if (nh) {
                const uint32_t sl = hit_list[0];
                qn = A.pvs[sl];
                wn = A.pvs[np + sl];
            }
            for (uint32_t h = 0; h < nh; ++h) {
                const float4 q = qn, w = wn;
                if (h + 1u < nh) {
                    const uint32_t sl = hit_list[(h + 1u) * kBlock];
                    qn = A.pvs[sl];
                    wn = A.pvs[np + sl];
                }
                const f3 rel = sub(x1, mk(q.x, q.y, q.z));
                const float d2 = dot(rel, rel);
                const f3 rv = sub(v1, mk(w.x, w.y, w.z));
                F = add(F, inrange_ok(rel, d2) ? resilient_inrange(pc, rel, rv, d2, r1, 0.5f)
                                               : resilient(pc, rel, rv, d2, r1, 0.5f, fl));
            }
            ncontact += nh;
        }
        more = mask != 0u;
    }
    return ncontact;
}

// What one launch of the cell kernel does: the whole particle part of a
// frame, or (split frames) only the centres, springs and environment, the
// contact search and the response then running as k_contacts and k_respond.
constexpr int kCellsWhole = 0, kCellsSprings = 1;

template <bool FUSED, bool SURR = false, int MODE = kCellsWhole>
__device__ __forceinline__ void cells_body(const StepArgs& A, VBlock vb)
{
    __shared__ uint32_t s_err[SURR ? kBlock : 1];  // per cell of the chunk: the surrogate's largest error
    __shared__ uint32_t s_stat[4];                 // the chunk's accepted / fallback cells, largest errors
    __shared__ float sx[kBlock], sy[kBlock], sz[kBlock];
    __shared__ float svx[kBlock], svy[kBlock], svz[kBlock];
    __shared__ float sfx[kBlock], sfy[kBlock], sfz[kBlock];
    __shared__ int s_tele[kBlock];
    __shared__ float s_bs[kBsLds];
    __shared__ uint32_t s_hit[MODE == kCellsWhole ? kContactWindow * kBlock : 1];  // contact lists
    __shared__ int32_t s_spoff[kBlock + 1];  // the block's cell type: spring list offsets per vertex,
    __shared__ int32_t s_spb[kSpringLds];    // partner vertex and rest length per spring
    __shared__ float s_spL[kSpringLds];

    const TypeTable& tt = A.tt;
    const PhysConst& pc = A.pc;
    const int l = threadIdx.x;
    const int chunk = chunk_of_block(A.nchunks, vb.id);
    if (chunk < 0) return;
    // contact radii per model vertex (cellModelsBoundingSpheres), from LDS
    // when the table fits
    const bool bs_lds = tt.nm <= kBsLds;
    if (bs_lds)
        for (int t = l; t < tt.nm; t += kBlock) s_bs[t] = A.bs[t];
    auto radius = [&](int32_t m) -> float { return bs_lds ? s_bs[m] : A.bs[m]; };
    const int T = A.chunk_type[chunk];
    const int ncell = A.chunk_n[chunk];
    const int p = tt.ppc[T];
    // the lane's {particle, blood cell} (particle -1: no slot, l >= ncell * p),
    // a load independent of the chunk's tables
    const int2 pcl = A.chunk_pi[(size_t)chunk * kBlock + l];
    const bool act = pcl.x >= 0;
    const int a = act ? l % p : 0;
    const int cl = act ? l / p : 0;
    const int64_t np = tt.np;
    const int32_t i = act ? pcl.x : 0, cell = act ? pcl.y : 0;
    uint32_t key_of = 0;
    int32_t slot_i = 0;
    // occupied neighbour cells of the contact search (rank r = (dx+1)*9 +
    // (dy+1)*3 + (dz+1)), from the occupancy map: its words are loaded before
    // the barrier, so that they arrive during the centre and the springs
    const GridGeom& g = A.pg;
    const int64_t cX = g.cnt[0], cXY = (int64_t)g.cnt[0] * g.cnt[1];
    uint32_t mask = 0u;
    if (act) {
        const float px = A.pos[i], py = A.pos[np + i], pz = A.pos[2 * np + i];
        sx[l] = px; sy[l] = py; sz[l] = pz;
        svx[l] = A.vel[i]; svy[l] = A.vel[np + i]; svz[l] = A.vel[2 * np + i];
        sfx[l] = A.frc[i]; sfy[l] = A.frc[np + i]; sfz[l] = A.frc[2 * np + i];
        if (MODE == kCellsWhole) {
            key_of = A.p_key_of[i];
            if (tt.nt > 1) slot_i = A.p_slot[i];
            // the cell the grid build gave i (gridCellIds of its slot); the
            // face clipping uses the current position, as the reference does
            const int32_t xId = sat_i32((px - g.minv[0]) / g.cell[0]);
            const int32_t yId = sat_i32((py - g.minv[1]) / g.cell[1]);
            const int32_t zId = sat_i32((pz - g.minv[2]) / g.cell[2]);
            NbrRange nr;
            axis_range_i(xId, g.cnt[0], nr.xl, nr.xh);
            axis_range_i(yId, g.cnt[1], nr.yl, nr.yh);
            axis_range_i(zId, g.cnt[2], nr.zl, nr.zh);
            mask = neighbour_mask(A.p_map, g, key_of, nr, cX, cXY, true);
        } else {
            slot_i = A.p_slot[i];  // where F after the springs goes
        }
    }
    // spring lists of this type's model vertices (one type per block), from
    // LDS when they fit
    const int m0 = tt.mstart[T];
    const int sp0 = A.sp_off[m0];
    const bool sp_lds = A.sp_off[m0 + p] - sp0 <= kSpringLds;
    if (sp_lds) {
        const int nsp = A.sp_off[m0 + p] - sp0;
        for (int t = l; t <= p; t += kBlock) s_spoff[t] = A.sp_off[m0 + t] - sp0;
        for (int t = l; t < nsp; t += kBlock) {
            s_spb[t] = A.sp_b[sp0 + t];
            s_spL[t] = A.sp_L[sp0 + t];
        }
    }
    const float r0a = act ? A.r0[m0 + a] : 1.0f;
    if (l < ncell) s_tele[l] = 0;
    if (SURR && l < ncell) s_err[l] = 0u;
    if (SURR && l < 4) s_stat[l] = 0u;
    __syncthreads();

    uint32_t fl = 0;
    f3 x1 = mk(0.f, 0.f, 0.f), v1 = x1, F = x1, F0 = x1, cen = x1, newF = x1;
    const int base = cl * p;
    if (act) {
        x1 = mk(sx[l], sy[l], sz[l]);
        v1 = mk(svx[l], svy[l], svz[l]);
        F0 = mk(sfx[l], sfy[l], sfz[l]);
        // ---- cell centre (sequential sum in vertex order)
        for (int b = 0; b < p; ++b) cen = add(cen, mk(sx[base + b], sy[base + b], sz[base + b]));
        cen = sdiv(cen, (float)p);
        if (a == 0) {
            A.centers[cell] = cen.x;
            A.centers[tt.nc + cell] = cen.y;
            A.centers[2 * (int64_t)tt.nc + cell] = cen.z;
        }
        // ---- springs + environment (F of neighbours from the LDS snapshot)
        const int ma = m0 + a;
        auto spring_sum = [&](auto spb, auto spL, int e0, int e1) {
        for (int e = e0; e < e1; ++e) {
            const int nb = base + spb(e);
            const float L = spL(e);
            f3 dP = sub(x1, mk(sx[nb], sy[nb], sz[nb]));
            f3 dv = sub(v1, mk(svx[nb], svy[nb], svz[nb]));
            f3 f2 = mk(sfx[nb], sfy[nb], sfz[nb]);
            // normalize(-dP) == -normalize(dP) and len(dP) is the same
            // sqrtf(dot(dP, dP)) bit for bit (negation commutes with every
            // rounding here), so one square root and three divisions serve all
            // three; a NaN quotient is replaced by (0,0,0) for both, as in
            // normalize().
            const float d2s = dot(dP, dP);
            float sl;
            f3 nrm, shift;
            if (inrange_ok(dP, d2s)) {
                sl = root_inrange(d2s);
                nrm = quot_inrange(dP, sl);
                shift = smul(-1.0f, nrm);
            } else {
                sl = sqrtf(d2s);
                nrm = sdiv(dP, sl);
                shift = smul(-1.0f, nrm);
                if (__builtin_isnan(nrm.x) || __builtin_isnan(nrm.y) || __builtin_isnan(nrm.z)) {
                    fl |= BCS_FLAG_NAN_NORMALIZE;
                    nrm = mk(0.f, 0.f, 0.f);
                    shift = nrm;
                }
            }
            if (pc.rk4) {
                const f3 k1x = smul(pc.dt, dv), k1v = smul(pc.dt, F0);
                const f3 k2x = smul(pc.dt, add(dv, sdiv(k1x, 2.0f)));
                const f3 k2v =
                    smul(pc.dt * spring_rk(pc, add(dP, sdiv(k1x, 2.0f)), add(dv, sdiv(k1v, 2.0f)), L, fl), shift);
                const f3 k3x = smul(pc.dt, add(dv, sdiv(k2x, 2.0f)));
                const f3 k3v =
                    smul(pc.dt * spring_rk(pc, add(dP, sdiv(k2x, 2.0f)), add(dv, sdiv(k2v, 2.0f)), L, fl), shift);
                const f3 k4v = smul(pc.dt * spring_rk(pc, add(dP, k3x), add(dv, k3v), L, fl), shift);
                newF = add(newF, add(add(smul(2.0f, k2v), smul(2.0f, k3v)), k4v));
            } else {
                f3 dvv = add(dv, smul(pc.dt, sub(F0, f2)));
                float mag = (sl - L) * pc.k + dot(nrm, dvv) * pc.d;
                newF = add(newF, smul(mag, shift));
            }
        }
        };
        if (sp_lds)
            spring_sum([&](int e) { return s_spb[e]; }, [&](int e) { return s_spL[e]; }, s_spoff[a], s_spoff[a + 1]);
        else
            spring_sum([&](int e) { return A.sp_b[e]; }, [&](int e) { return A.sp_L[e]; }, A.sp_off[ma],
                       A.sp_off[ma + 1]);
    }
    if constexpr (SURR) {
        // ---- learned springs: the served type's cells take the network's
        // spring displacement when the whole cell is within tolerance of the
        // plain solve (T is one per block, so the barrier is uniform)
        const SurrArgs& sa = A.surr;
        if (T == sa.type) {
            const int64_t row = (int64_t)(cell - tt.cstart[T]) * kSurrOutPad + 3 * a;
            const f3 d = smul(sa.scale, newF);  // the plain spring displacement (dt^2/4) S
            if (act && sa.S_out) {
                sa.S_out[row] = d.x;
                sa.S_out[row + 1] = d.y;
                sa.S_out[row + 2] = d.z;
            }
            if (sa.D) {
                f3 dh = mk(0.f, 0.f, 0.f);
                if (act) {
                    dh = smul(sa.out_scale, mk(sa.D[row], sa.D[row + 1], sa.D[row + 2]));
                    atomicMax(&s_err[cl], __float_as_uint(len(sub(dh, d))));
                }
                __syncthreads();
                if (act) {
                    const float ce = __uint_as_float(s_err[cl]);
                    const bool ok = sa.tol > 0.0f && ce <= sa.tol;
                    if (ok) newF = sdiv(dh, sa.scale);
                    if (a == 0) {  // the block's tallies (errors are >= 0: float order = bit order)
                        atomicAdd(&s_stat[ok ? 0 : 1], 1u);
                        atomicMax(&s_stat[3], __float_as_uint(ce));
                        if (ok) atomicMax(&s_stat[2], __float_as_uint(ce));
                    }
                }
                __syncthreads();
                if (l < 4 && s_stat[l] != 0u) {  // one global update per tally and block, in the block's slot
                    uint32_t* slot = sa.stats + (size_t)(vb.id % kSurrStatSlots) * kSurrStatStride;
                    if (l < 2) atomicAdd(&slot[l], s_stat[l]);
                    else atomicMax(&slot[l], s_stat[l]);
                }
            }
        }
    }
    uint32_t ncontact = 0;
    if (act) {
        const float ratio = len(sub(x1, cen)) / r0a;
        f3 env;
        const f3 G = mk(pc.G[0], pc.G[1], pc.G[2]);
        if (pc.brake_on && ratio > pc.brake) env = sub(G, smul(pc.vdamp * ratio * pc.bint, v1));
        else env = sub(G, smul(pc.vdamp, v1));
        newF = add(newF, env);
        F = sdiv(add(F0, newF), pc.rk4 ? 6.0f : 2.0f);

        if constexpr (MODE == kCellsWhole) {
            // ---- particle-particle contacts over the 27 neighbour cells
            int Tl = T;
            if (tt.nt > 1) Tl = type_of_index(tt, slot_i);
            ncontact = contact_sum<FUSED ? kPruneLane : kPruneNone>(A, i, Tl, key_of, mask, x1, v1, F, fl, radius,
                                                                   s_hit);
        } else {
            A.fslot[slot_i] = make_float4(F.x, F.y, F.z, 0.0f);
        }
        if (MODE == kCellsWhole && SURR && A.surr.touched) A.surr.touched[i] = ncontact > 0u ? 1 : 0;
    }
    if constexpr (MODE == kCellsWhole) respond_integrate<FUSED>(A, act, i, T, a, cl, cell, x1, v1, F, fl, s_tele);
    else if (fl) atomicOr(A.flags, fl);
}
template <bool FUSED, bool SURR = false>
__global__ void __launch_bounds__(kBlock) k_cells(StepArgs A)
{
    cells_body<FUSED, SURR>(A, hw_block());
}
template <bool SURR>
__global__ void __launch_bounds__(kBlock) k_springs(StepArgs A)
{
    cells_body<true, SURR, kCellsSprings>(A, hw_block());
}

// Contact search of a split frame, one lane per sorted slot s of the particle
// grid.  The lanes of a wave hold particles of the same or adjacent grid
// cells, so they walk nearly the same neighbour cells and read the same
// records at the same time (in the cell kernel, whose lanes follow the
// particles of whole blood cells, every lane reads records of its own).  F
// of the particle after springs and environment comes from fslot[s], and F
// after the contacts goes back there.
template <bool SURR>
__device__ __forceinline__ void contacts_body(const StepArgs& A, VBlock vb)
{
    // per lane: the slots of its contacts in a window (the contact radii are
    // read from the global table: the one-radius shortcut rarely needs them,
    // and the kernel then keeps no LDS table and no barrier for it)
    __shared__ uint32_t s_hit[kContactWindow * kBlock];
    const TypeTable& tt = A.tt;
    const int l = threadIdx.x;
    // runs of kXcdRunTiles tiles per XCD: a z-plane of the vessel spans tens
    // of tiles, so a run holds a few neighbouring planes, whose records the
    // z +- 1 neighbour cells read again from the same L2
    const int tile = chunk_of_block<kXcdRunTiles>(A.ntiles, vb.id);
    if (tile < 0) return;
    auto radius = [&](int32_t m) -> float { return A.bs[m]; };
    const int64_t np = tt.np;
    const int64_t u = (int64_t)tile * kBlock + l;
    const bool act = u < np;
    const GridGeom& g = A.pg;
    const int64_t cX = g.cnt[0], cXY = (int64_t)g.cnt[0] * g.cnt[1];
    uint32_t fl = 0, mask = 0u, key_of = 0xFFFFFFFFu;  // lanes past the end: no key, no cells
    int32_t i = 0;
    f3 x1 = mk(0.f, 0.f, 0.f), v1 = x1, F = x1;
    if (act) {
        const float4 q = A.pvs[u], w = A.pvs[np + u], f = A.fslot[u];
        i = __float_as_int(q.w);
        x1 = mk(q.x, q.y, q.z);
        v1 = mk(w.x, w.y, w.z);
        F = mk(f.x, f.y, f.z);
        key_of = A.p_keys[u];
        const int32_t xId = sat_i32((x1.x - g.minv[0]) / g.cell[0]);
        const int32_t yId = sat_i32((x1.y - g.minv[1]) / g.cell[1]);
        const int32_t zId = sat_i32((x1.z - g.minv[2]) / g.cell[2]);
        NbrRange nr;
        axis_range_i(xId, g.cnt[0], nr.xl, nr.xh);
        axis_range_i(yId, g.cnt[1], nr.yl, nr.yh);
        axis_range_i(zId, g.cnt[2], nr.zl, nr.zh);
        mask = neighbour_mask(A.p_map, g, key_of, nr, cX, cXY, true);
    }
    // every lane of the wave takes part (the pruning exchanges masks between
    // lanes); a lane past the end has an empty mask
    const int Tl = act && tt.nt > 1 ? type_of_index(tt, (int32_t)u) : 0;
    const uint32_t nc = contact_sum<kPruneCell>(A, i, Tl, key_of, mask, x1, v1, F, fl, radius, s_hit);
    // w: whether the particle touched another (for the surrogate's next
    // input, written per particle by k_respond)
    if (act) A.fslot[u] = make_float4(F.x, F.y, F.z, __uint_as_float(nc > 0u ? 1u : 0u));
    (void)SURR;
    if (fl) atomicOr(A.flags, fl);
}
template <bool SURR>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7))) k_contacts(StepArgs A)
{
    contacts_body<SURR>(A, hw_block());
}

// Last part of a split frame, per chunk of whole blood cells: the vein hit
// response, integration and vein end, with F after the contacts from fslot.
__device__ __forceinline__ void respond_body(const StepArgs& A, VBlock vb)
{
    __shared__ int s_tele[kBlock];
    const int l = threadIdx.x;
    const int chunk = chunk_of_block(A.nchunks, vb.id);
    if (chunk < 0) return;
    const TypeTable& tt = A.tt;
    const int T = A.chunk_type[chunk];
    const int ncell = A.chunk_n[chunk];
    const int p = tt.ppc[T];
    const int2 pcl = A.chunk_pi[(size_t)chunk * kBlock + l];
    const bool act = pcl.x >= 0;
    const int a = act ? l % p : 0;
    const int cl = act ? l / p : 0;
    const int32_t i = act ? pcl.x : 0, cell = act ? pcl.y : 0;
    const int64_t np = tt.np;
    f3 x1 = mk(0.f, 0.f, 0.f), v1 = x1, F = x1;
    if (act) {
        x1 = load3(A.pos, np, i);
        v1 = load3(A.vel, np, i);
        const float4 f = A.fslot[A.p_slot[i]];
        F = mk(f.x, f.y, f.z);
        if (A.surr.touched) A.surr.touched[i] = __float_as_uint(f.w) != 0u ? 1 : 0;
    }
    if (l < ncell) s_tele[l] = 0;
    respond_integrate<true>(A, act, i, T, a, cl, cell, x1, v1, F, 0u, s_tele);
}
__global__ void __launch_bounds__(kBlock) k_respond(StepArgs A) { respond_body(A, hw_block()); }

// kRayLanes lanes per work item; the item's first hit is folded into the
// particle's minimum (cell rank, sorted slot).  The hit test alone decides;
// the normal of the winning triangle is formed by k_cells.  The items of work
// region r taken by block `sub` of its `split` blocks.
__device__ __forceinline__ void ray_region(const StepArgs& A, uint32_t r, uint32_t sub, uint32_t split)
{
    constexpr uint32_t G = kRayLanes, K = kItemTris / kRayLanes;
    static_assert(G * K == kItemTris && G <= 64 && (G & (G - 1)) == 0, "lanes per item divide the item");
    const uint32_t n = min(A.wcount[r], A.region_cap);
    const uint4* items = A.work + (size_t)r * A.region_cap;
    const int64_t np = A.tt.np;
    const uint32_t g = threadIdx.x % G;
    const uint32_t per_block = blockDim.x / G;
    uint32_t fl = 0;
    // base is block-uniform, so every lane reaches the ballot
    for (uint32_t base = sub * per_block; base < n; base += split * per_block) {
        const uint32_t w = base + threadIdx.x / G;
        uint32_t first = K;
        uint4 it = make_uint4(0u, 0u, 0u, 0u);
        if (w < n) {
            it = items[w];
            const int32_t i = (int32_t)it.x;
            const f3 x1 = load3(A.pos, np, i);
            const f3 dir = normalize(load3(A.vel, np, i), fl);
            first = item_lane_first_hit(A, it, g, x1, dir);
        }
        if (item_first_lane(first < K, g)) atomicMin(&A.best[it.x], (it.z << kSlotBits) | (it.y + g * K + first));
    }
    if (fl) atomicOr(A.flags, fl);
}

__device__ __forceinline__ void ray_body(const StepArgs& A, uint32_t split, VBlock vb)
{
    // `split` blocks share the work region of k_raycand block r; block
    // b = s * nregions + r runs on the XCD (b % 8 == r % 8) whose L2 holds it
    const uint32_t nreg = vb.count / split;
    ray_region(A, vb.id % nreg, vb.id / nreg, split);
}
__global__ void __launch_bounds__(256) k_ray(StepArgs A, uint32_t split) { ray_body(A, split, hw_block()); }

// Vein hit response of particle i (detectVeinCollisions response,
// src/simulation/vein_collisions.cu:234-275), then for FUSED frames the Heun
// integration and the vein end of its cell (src/objects/blood_cells.cu:150-179,
// src/simulation/vein_end.cu:12-173); writes x (FUSED), v and F.  The whole
// block calls it (barriers); s_tele[0 .. ncell) is zero on entry.
template <bool FUSED>
__device__ __forceinline__ void respond_integrate(const StepArgs& A, bool act, int32_t i, int T, int a, int cl,
                                                  int32_t cell, f3 x1, f3 v1, f3 F, uint32_t fl, int* s_tele)
{
    const TypeTable& tt = A.tt;
    const PhysConst& pc = A.pc;
    const int64_t np = tt.np;
    if (act) {
        const uint32_t best = A.best[i];
        const f3 dir = normalize(v1, fl);
        float rt = 1e10f;
        f3 normal = mk(0.f, 0.f, 0.f), refl = mk(0.f, 0.f, 0.f);
        int32_t obj = 0;
        bool hit = false;
        if (best == kScanSerial) {
            hit = ray_first_hit_serial(A, x1, dir, rt, normal, refl, obj, fl);
        } else if (best != kNoHit) {
            const uint32_t u = best & ((1u << kSlotBits) - 1u);
            const float4 ta = A.tdata[3 * (size_t)u], e1 = A.tdata[3 * (size_t)u + 1], e2 = A.tdata[3 * (size_t)u + 2];
            hit = ray_tri_pre(mk(ta.x, ta.y, ta.z), mk(e1.x, e1.y, e1.z), mk(e2.x, e2.y, e2.z), x1, dir, rt, normal, refl,
                              fl);
            obj = __float_as_int(ta.w);
        }
        const f3 relPos = sub(x1, add(x1, smul(rt, dir)));
        const float d2 = dot(relPos, relPos);
        if (hit && d2 <= pc.impact2) {
            if (d2 > pc.minimal2) {
                if (pc.reaction) {
                    f3 resp = sdiv(smul(-1.0f * dot(F, normal), normal), dot(normal, normal));
                    F = add(F, resp);
                } else {
                    const float rad = A.bs[model_index(tt, T, i, fl)];
                    F = add(F, resilient(pc, relPos, v1, d2, rad, 0.5f, fl));
                }
            }
            const int64_t nv = A.nv;
            const float speed = len(v1);
            const f3 dv = smul(1.0f, sub(smul(pc.vcd * speed, refl), v1));
            const f3 ds = smul(pc.vcfi, v1);
            v1 = add(v1, dv);
            const uint32_t i0 = A.tidx[3 * obj], i1 = A.tidx[3 * obj + 1], i2 = A.tidx[3 * obj + 2];
            const f3 v0 = load3(A.X, nv, i0), va = load3(A.X, nv, i1), vb = load3(A.X, nv, i2);
            const f3 bc = bary(add(x1, smul(rt, dir)), v0, va, vb);
            const f3 c0 = smul(bc.x, ds), c1 = smul(bc.y, ds), c2 = smul(bc.z, ds);
            const uint32_t h = atomicAdd(&A.hctrl[0], 1u);
            A.hits[3 * (size_t)h + 0] = make_float4(c0.x, c0.y, c0.z, __int_as_float(i));
            A.hits[3 * (size_t)h + 1] = make_float4(c1.x, c1.y, c1.z, __int_as_float(obj));
            A.hits[3 * (size_t)h + 2] = make_float4(c2.x, c2.y, c2.z, 0.0f);
            // push entry 3h + k (corner k) on the list of its vertex
            const uint32_t vs[3] = {i0, i1, i2};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int32_t e = (int32_t)(3 * h + k);
                A.vnext[e] = atomicExch(&A.vhead[vs[k]], e);
            }
        }
    }
    if (FUSED) {
        // ---- integrate (Heun / trapezoid): v0 = v/gpuCount; v = v0 + dt*F; x += 0.5*dt*(v + v0)
        if (act) {
            const f3 v0 = sdiv(v1, 1.0f);
            v1 = add(v0, smul(pc.dt, F));
            x1 = add(x1, position_step(pc, v0, v1));
        }
        __syncthreads();  // s_tele initialised
        if (act && pc.use_blood_flow && vein_end_flag(pc, tt.end_spheres[T] != 0, x1)) s_tele[cl] = 1;
        __syncthreads();
        if (act && pc.use_blood_flow && s_tele[cl]) {
            const uint64_t stp = *A.step;
            const float u1 = uniform01(pc.seed, stp, (uint64_t)cell, 0u);
            const float u2 = uniform01(pc.seed, stp, (uint64_t)cell, 1u);
            x1 = teleport_pos(pc, A.model, tt.mstart[T], a, u1, u2);
            v1 = mk(pc.initv[0], pc.initv[1], pc.initv[2]);
        }
        if (act) {
            A.pos[i] = x1.x; A.pos[np + i] = x1.y; A.pos[2 * np + i] = x1.z;
        }
    }
    if (act) {
        A.vel[i] = v1.x; A.vel[np + i] = v1.y; A.vel[2 * np + i] = v1.z;
        A.frc[i] = F.x; A.frc[np + i] = F.y; A.frc[2 * np + i] = F.z;
    }
    if (fl) atomicOr(A.flags, fl);
}

// propagateParticleForcesKernel, src/objects/blood_cells.cu:155-179 (API path).
__global__ void __launch_bounds__(256) k_integrate(StepArgs A)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t np = A.tt.np;
    if (i >= np) return;
    const f3 F = load3(A.frc, np, i);
    const f3 v0 = sdiv(load3(A.vel, np, i), 1.0f);
    const f3 v = add(v0, smul(A.pc.dt, F));
    store3(A.vel, np, i, v);
    store3(A.pos, np, i, add(load3(A.pos, np, i), position_step(A.pc, v0, v)));
}

// HandleVeinEnd, src/simulation/vein_end.cu:141-173 (API path; one block per
// chunk of whole cells so the per-cell OR is a block-local reduction).
__global__ void __launch_bounds__(kBlock) k_vein_end(StepArgs A)
{
    __shared__ int s_tele[kBlock];
    const TypeTable& tt = A.tt;
    const int l = threadIdx.x;
    const int chunk = chunk_of_block(A.nchunks, blockIdx.x);
    if (chunk < 0) return;
    const int T = A.chunk_type[chunk], first = A.chunk_first[chunk], ncell = A.chunk_n[chunk];
    const int p = tt.ppc[T];
    const bool act = l < ncell * p;
    const int a = act ? l % p : 0, cl = act ? l / p : 0;
    const int64_t np = tt.np;
    int32_t cell = 0;
    const int32_t i = act ? chunk_particle(A, T, first, cl, a, p, cell) : 0;
    if (l < ncell) s_tele[l] = 0;
    __syncthreads();
    f3 x = mk(0.f, 0.f, 0.f);
    if (act) {
        x = load3(A.pos, np, i);
        if (vein_end_flag(A.pc, tt.end_spheres[T] != 0, x)) s_tele[cl] = 1;
    }
    __syncthreads();
    if (act && s_tele[cl]) {
        const uint64_t stp = *A.step;
        const float u1 = uniform01(A.pc.seed, stp, (uint64_t)cell, 0u);
        const float u2 = uniform01(A.pc.seed, stp, (uint64_t)cell, 1u);
        store3(A.pos, np, i, teleport_pos(A.pc, A.model, tt.mstart[T], a, u1, u2));
        store3(A.vel, np, i, mk(A.pc.initv[0], A.pc.initv[1], A.pc.initv[2]));
    }
}

__global__ void k_bump_step(uint64_t* step) { *step += 1ull; }

// Vein springs: gatherForcesKernel, src/objects/vein_triangles.cu:126-154.
// Forces were zeroed by the previous integrate, so Fv = 0 + spring force.
__device__ __forceinline__ void vein_springs_at(const VeinArgs& A, int32_t id)
{
    if (id >= A.nv) return;
    const int64_t nv = A.nv;
    uint32_t fl = 0;
    const f3 P = load3(A.X, nv, id);
    const f3 Vv = load3(A.V, nv, id);
    f3 force = mk(0.f, 0.f, 0.f);
#pragma unroll
    for (int s = 0; s < kVeinSlots; ++s) {
        const int32_t nb = A.nbr[s * nv + id];
        if (nb != -1) {
            const float L = A.nbl[s * nv + id];
            const f3 NP = load3(A.X, nv, nb);
            const f3 NV = load3(A.V, nv, nb);
            // len(P - NP), normalize(P - NP) and normalize(NP - P) from one
            // square root and three divisions: NP - P = -(P - NP) exactly and
            // negation commutes with every rounding here; a NaN quotient gives
            // (0,0,0) for both, as normalize() does.
            const f3 dP = sub(P, NP);
            const float sl = sqrtf(dot(dP, dP));
            f3 nrm = sdiv(dP, sl);
            f3 back = smul(-1.0f, nrm);
            if (__builtin_isnan(nrm.x) || __builtin_isnan(nrm.y) || __builtin_isnan(nrm.z)) {
                fl |= BCS_FLAG_NAN_NORMALIZE;
                nrm = mk(0.f, 0.f, 0.f);
                back = nrm;
            }
            const float sf = (sl - L) * A.vk + dot(nrm, sub(Vv, NV)) * A.vd;
            force = add(force, smul(sf, back));
        }
    }
    const f3 F0 = load3(A.Fv, nv, id);
    store3(A.Fv, nv, id, add(F0, force));
    if (fl) atomicOr(A.flags, fl);
}

__global__ void __launch_bounds__(256) k_vein_springs(VeinArgs A)
{
    vein_springs_at(A, blockIdx.x * blockDim.x + threadIdx.x);
}

// Per vertex: add the recorded contributions in (particle, corner) order, then
// (mode & 2) integrate: V += dt*Fv; X += dt*V; Fv = 0
// (propagateForcesIntoPositionsKernel, src/objects/vein_triangles.cu:88-117).
template <int MODE>
__device__ __forceinline__ void vein_body(const VeinArgs& A, VBlock vb)
{
    const int32_t id = vb.id * blockDim.x + threadIdx.x;
    if (MODE & 1) {
        if (id == 0) {  // hits consumed (the hit buffer is only read through the lists)
            A.hctrl[4] = A.hctrl[0];
            A.hctrl[0] = 0u;
        }
    }
    if (MODE & 4) {
        if (id == 0) *A.step += 1ull;
    }
    if (id >= A.nv) return;
    const int64_t nv = A.nv;
    f3 F = load3(A.Fv, nv, id);
    if (MODE & 1) {
        const int32_t head = A.vhead[id];
        if (head >= 0) {
            // the vertex's entries in ascending (particle, corner) key order,
            // picked by selection from its list (lists are a few entries long)
            int32_t n = 0;
            for (int32_t e = head; e >= 0 && n < (1 << 20); e = A.vnext[e]) ++n;
            int64_t last = -1;
            for (int32_t r = 0; r < n; ++r) {
                int64_t best = 0x7FFFFFFFFFFFFFFFll;
                int32_t be = -1;
                int32_t c = 0;
                for (int32_t e = head; e >= 0 && c < n; e = A.vnext[e], ++c) {
                    const int64_t key = (int64_t)__float_as_int(A.hits[3 * (size_t)(e / 3)].w) * 4 + (e % 3);
                    if (key > last && key < best) {
                        best = key;
                        be = e;
                    }
                }
                if (be < 0) break;
                const float4 cv = A.hits[be];
                F = add(F, mk(cv.x, cv.y, cv.z));
                last = best;
            }
            A.vhead[id] = -1;
        }
    }
    if (MODE & 2) {
        const f3 V = add(load3(A.V, nv, id), smul(A.dt, F));
        store3(A.V, nv, id, V);
        const f3 Xn = add(load3(A.X, nv, id), smul(A.dt, V));
        store3(A.X, nv, id, Xn);
        store3(A.Fv, nv, id, mk(0.f, 0.f, 0.f));
    } else {
        store3(A.Fv, nv, id, F);
    }
}
template <int MODE>
__global__ void __launch_bounds__(256) k_vein(VeinArgs A)
{
    vein_body<MODE>(A, hw_block());
}

// Triangle centres, calculateCentersKernel, src/objects/vein_triangles.cu:14-27.
__global__ void __launch_bounds__(256) k_tri_centers(const float* __restrict__ X, int32_t nv,
                                                     const uint32_t* __restrict__ idx, int32_t ntri,
                                                     float* __restrict__ c)
{
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    const f3 a = load3(X, nv, idx[3 * t]), b = load3(X, nv, idx[3 * t + 1]), d = load3(X, nv, idx[3 * t + 2]);
    c[t] = (a.x + b.x + d.x) / 3.0f;
    c[ntri + t] = (a.y + b.y + d.y) / 3.0f;
    c[2 * (int64_t)ntri + t] = (a.z + b.z + d.z) / 3.0f;
}

// Per frame, one wavefront per occupied triangle cell: every lane prepares
// one triangle of the cell's run -- (v0 | triangle id), v1 - v0, v2 - v0 from
// the current vein positions, stored at its sorted slot -- and the wave
// reduces the cell's box (exact vertex bounds), stored under the cell key.
// Non-finite coordinates widen the box to everything.
// Blocks [0, nblocks) of 256 threads share the occupied triangle cells, one
// wave per cell at a time: run record, then the sorted triangles' records,
// then their vertices (three dependent loads).
__device__ __forceinline__ void tri_box_blocks(const TriBoxArgs& B, uint32_t block, uint32_t nblocks)
{
    const uint4* __restrict__ trun = B.trun;
    const uint4* __restrict__ trec = B.trec;
    const int32_t ntri = B.ntri;
    const float* __restrict__ X = B.X;
    const int32_t nv = B.nv;
    float4* __restrict__ tdata = B.tdata;
    float4* __restrict__ aabb = B.aabb;
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (block * 256u + threadIdx.x) >> 6;
    const uint32_t nwaves = (nblocks * 256u) >> 6;
    for (uint32_t r = wave; r < (uint32_t)ntri; r += nwaves) {
        const uint4 ru = trun[r];
        if (ru.w == 0u) break;  // wave-uniform: the occupied cells are a prefix
        const uint32_t s0 = ru.x, key = ru.y, cnt = ru.z;
        float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
        for (uint32_t t = lane; t < cnt && s0 + t < (uint32_t)ntri; t += 64) {
            const uint32_t u = s0 + t;
            const uint4 tr = trec[u];
            const int32_t tri = (int32_t)tr.w;
            const f3 v0 = load3(X, nv, tr.x), v1 = load3(X, nv, tr.y), v2 = load3(X, nv, tr.z);
            const f3 e1 = sub(v1, v0), e2 = sub(v2, v0);
            tdata[3 * (size_t)u] = make_float4(v0.x, v0.y, v0.z, __int_as_float(tri));
            tdata[3 * (size_t)u + 1] = make_float4(e1.x, e1.y, e1.z, 0.0f);
            tdata[3 * (size_t)u + 2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
            const f3 vs[3] = {v0, v1, v2};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float q[3] = {vs[c].x, vs[c].y, vs[c].z};
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    if (q[d] == q[d] && fabsf(q[d]) < 3.0e38f) {
                        lo[d] = fminf(lo[d], q[d]);
                        hi[d] = fmaxf(hi[d], q[d]);
                    } else {
                        lo[d] = -3.0e38f;
                        hi[d] = 3.0e38f;
                    }
                }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                lo[d] = fminf(lo[d], __shfl_xor(lo[d], off, 64));
                hi[d] = fmaxf(hi[d], __shfl_xor(hi[d], off, 64));
            }
        }
        if (lane == 0) {  // the cell's run {start, count} rides in the w lanes
            aabb[2 * (size_t)key] = make_float4(lo[0], lo[1], lo[2], __uint_as_float(s0));
            aabb[2 * (size_t)key + 1] = make_float4(hi[0], hi[1], hi[2], __uint_as_float(cnt));
        }
    }
}

// Run and sorted-triangle records of the triangle grid for tri_box_blocks,
// rebuilt with the grid (one thread per sorted slot u < ntri): trun[u] is the
// record of occupied cell u, or zero past the last one.
__global__ void __launch_bounds__(256) k_tri_records(const uint32_t* __restrict__ runs, const uint32_t* __restrict__ ctrl,
                                                     const uint2* __restrict__ table,
                                                     const uint32_t* __restrict__ keys_sorted,
                                                     const int32_t* __restrict__ perm, int32_t ntri,
                                                     const uint32_t* __restrict__ idx, uint4* __restrict__ trun,
                                                     uint4* __restrict__ trec)
{
    const int32_t u = blockIdx.x * 256 + threadIdx.x;
    if (u >= ntri) return;
    const uint32_t nr = ctrl[2];
    uint4 ru = make_uint4(0u, 0u, 0u, 0u);
    if ((uint32_t)u < nr) {
        const uint32_t s0 = runs[u];
        const uint32_t key = keys_sorted[s0];
        ru = make_uint4(s0, key, table[key].y, 1u);
    }
    trun[u] = ru;
    const int32_t tri = perm[u];
    trec[u] = make_uint4(idx[3 * tri], idx[3 * tri + 1], idx[3 * tri + 2], (uint32_t)tri);
}

// Vein springs and triangle data in one launch (both need only the updated
// vein): the first nsb blocks take the vertices, the rest the triangle cells.
__device__ __forceinline__ void vein_prep_body(const VeinPrepArgs& P, VBlock vb)
{
    if (vb.id < P.nsb) vein_springs_at(P.V, vb.id * 256u + threadIdx.x);
    else tri_box_blocks(P.B, vb.id - P.nsb, vb.count - P.nsb);
}
__global__ void __launch_bounds__(256) k_vein_prep(VeinPrepArgs P) { vein_prep_body(P, hw_block()); }

// ---- processing order of the cells: by type, then by the Morton code of the
// cell centre (previous frame) on a 1024^3 lattice over the particle-grid box,
// cut to its top kOrderBits bits.  A counting sort; the order within a bin is
// arbitrary (it changes which block handles a cell, never a result).
__device__ __forceinline__ uint32_t spread3(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t lattice10(float c, float lo, float ext)
{
    float t = (c - lo) / ext * 1024.0f;
    if (!(t >= 0.0f)) t = 0.0f;  // NaN too
    if (t > 1023.0f) t = 1023.0f;
    return (uint32_t)t;
}

__device__ __forceinline__ void order_key_body(const OrderKeyArgs& O, VBlock vb)
{
    const TypeTable& tt = O.tt;
    const GridGeom& g = O.g;
    const float* __restrict__ centers = O.centers;
    const int32_t c = vb.id * blockDim.x + threadIdx.x;
    for (int32_t t = c; t < O.otiles; t += (int32_t)(vb.count * blockDim.x)) O.ostatus[t] = 0ull;
    if (c == 0) O.octrl[0] = 0u;
    if (c >= tt.nc) return;
    int T = 0;
    while (T + 1 < tt.nt && c >= tt.cstart[T + 1]) ++T;
    const int64_t nc = tt.nc;
    const uint32_t qx = lattice10(centers[c], g.minv[0], g.extent[0]);
    const uint32_t qy = lattice10(centers[nc + c], g.minv[1], g.extent[1]);
    const uint32_t qz = lattice10(centers[2 * nc + c], g.minv[2], g.extent[2]);
    const uint32_t m = (spread3(qx) | (spread3(qy) << 1) | (spread3(qz) << 2)) >> (30 - kOrderBits);
    const uint32_t key = ((uint32_t)T << kOrderBits) | m;
    O.okey[c] = key;
    atomicAdd(&O.ocount[key], 1u);
}
__global__ void __launch_bounds__(256) k_order_key(OrderKeyArgs O) { order_key_body(O, hw_block()); }

__device__ __forceinline__ void order_scatter_body(const OrderScatterArgs& O, VBlock vb)
{
    const int32_t c = vb.id * blockDim.x + threadIdx.x;
    if (c >= O.nc) return;
    const uint32_t key = O.okey[c];
    const uint32_t old = atomicSub(&O.ocount[key], 1u);  // counts return to zero
    const uint32_t pos = O.ostart[key] + old - 1u;
    if (old == 0u || pos >= (uint32_t)O.nc) {
        atomicOr(O.flags, BCS_FLAG_SCAN_TIMEOUT);
        return;
    }
    O.corder[pos] = c;
}
__global__ void __launch_bounds__(256) k_order_scatter(OrderScatterArgs O) { order_scatter_body(O, hw_block()); }

// Per triangle-grid cell: mask of the occupied cells of its neighbourhood
// (ranges of axis_range_u on the cell's own coordinates), over the visit rank.
__global__ void __launch_bounds__(256) k_tri_nmask(GridGeom g, const uint2* __restrict__ table,
                                                   uint32_t* __restrict__ nmask)
{
    const uint64_t cell = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= g.cell_count) return;
    const uint32_t k = (uint32_t)cell;
    uint32_t x, y, z;
    grid_coords(g, k, x, y, z);
    int xl, xh, yl, yh, zl, zh;
    axis_range_u(x, g.cnt[0], xl, xh);
    axis_range_u(y, g.cnt[1], yl, yh);
    axis_range_u(z, g.cnt[2], zl, zh);
    const int64_t cX = g.cnt[0], cXY = (int64_t)g.cnt[0] * g.cnt[1];
    uint32_t m = 0;
    for (int dx = xl; dx <= xh; ++dx)
        for (int dy = yl; dy <= yh; ++dy)
            for (int dz = zl; dz <= zh; ++dz) {
                const int64_t nk = (int64_t)k + dz * cXY + dy * cX + dx;
                if (nk < 0 || (uint64_t)nk >= g.cell_count) continue;
                if (table[nk].y != 0u) m |= 1u << ((dx + 1) * 9 + (dy + 1) * 3 + (dz + 1));
            }
    nmask[k] = m;
}

// Contact listing in sorted-slot order (calculateParticleCollisions visit
// order; the force is not applied).  pass 0 counts, pass 1 writes pairs at
// offs[i].
__global__ void __launch_bounds__(256) k_contact_pairs(StepArgs A, int pass, int32_t* counts, const int64_t* offs,
                                                  int32_t* pairs)
{
    const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const TypeTable& tt = A.tt;
    if (s >= tt.np) return;
    uint32_t fl = 0;
    const int64_t np = tt.np;
    const GridGeom& g = A.pg;
    const int Tl = type_of_index(tt, s);
    const int32_t i = A.p_perm[s];
    const f3 x1 = load3(A.pos, np, i);
    const float r1 = A.bs[model_index(tt, Tl, i, fl)];
    const uint32_t cellId = A.p_keys[s];
    const int32_t xId = sat_i32((x1.x - g.minv[0]) / g.cell[0]);
    const int32_t yId = sat_i32((x1.y - g.minv[1]) / g.cell[1]);
    const int32_t zId = sat_i32((x1.z - g.minv[2]) / g.cell[2]);
    int xl, xh, yl, yh, zl, zh;
    axis_range_i(xId, g.cnt[0], xl, xh);
    axis_range_i(yId, g.cnt[1], yl, yh);
    axis_range_i(zId, g.cnt[2], zl, zh);
    const int64_t cX = g.cnt[0], cXY = (int64_t)g.cnt[0] * g.cnt[1];
    int32_t c = 0;
    for (int dx = xl; dx <= xh; ++dx)
        for (int dy = yl; dy <= yh; ++dy)
            for (int dz = zl; dz <= zh; ++dz) {
                const int64_t nk = (int64_t)cellId + dz * cXY + dy * cX + dx;
                if (nk < 0 || (uint64_t)nk >= g.cell_count) continue;
                if (cell_code(A.p_map, (uint32_t)nk) == 0u) continue;
                const uint2 e = A.p_table[nk];
                for (uint32_t u = e.x; u < min(e.x + e.y, (uint32_t)np); ++u) {
                    const int32_t j = A.p_perm[u];
                    if (j == i) continue;
                    const float r2 = A.bs[model_index(tt, Tl, j, fl)];
                    const f3 rel = sub(x1, load3(A.pos, np, j));
                    const float d2 = dot(rel, rel);
                    const float minD = r1 + r2;
                    if (d2 <= minD * minD && d2 >= 0.0001f) {
                        if (pass == 1) pairs[offs[i] + c] = j;
                        ++c;
                    }
                }
            }
    if (pass == 0) counts[i] = c;
}

// Sorted-slot records from the current state (when particles moved after the
// last grid build): slot u keeps the particle perm[u] of that build.
__global__ void __launch_bounds__(256) k_refresh_copies(const float* __restrict__ pos, const float* __restrict__ vel,
                                                        int32_t n, const int32_t* __restrict__ perm,
                                                        float4* __restrict__ pvs)
{
    const int32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const int32_t i = perm[u];
    if (i < 0 || i >= n) return;
    const f3 x = load3(pos, n, i), v = load3(vel, n, i);
    pvs[u] = make_float4(x.x, x.y, x.z, __int_as_float(i));
    pvs[(size_t)n + u] = make_float4(v.x, v.y, v.z, 0.0f);
}

// Frame export: SoA -> xyz interleaved with a given stride (3 or 6 floats).
__global__ void __launch_bounds__(256) k_export(const float* __restrict__ soa, int64_t n, int64_t first, int64_t count,
                                                int stride, float* __restrict__ out)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    out[stride * r] = soa[first + r];
    out[stride * r + 1] = soa[n + first + r];
    out[stride * r + 2] = soa[2 * n + first + r];
}

// Gather of cell-table starts (parity views of the grid).
__global__ void __launch_bounds__(256) k_gather_start(const uint2* __restrict__ table, const uint32_t* __restrict__ idx,
                                                      int32_t n, uint32_t* __restrict__ out)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = table[idx[i]].x;
}

// Dense per-cell-id ranges in the reference's gridCellStarts / gridCellEnds
// layout (src/grids/uniform_grid.cu:51-80), from the sorted keys: every entry
// first set to the empty range (0, -1), then each occupied cell's first and
// last slot written by the slot where its run starts or ends.
__global__ void __launch_bounds__(256) k_fill_ranges(int32_t* __restrict__ starts, int32_t* __restrict__ ends,
                                                     uint64_t m)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
        starts[i] = 0;
        ends[i] = -1;
    }
}
__global__ void __launch_bounds__(256) k_dense_ranges(const uint32_t* __restrict__ keys, int32_t n,
                                                      int32_t* __restrict__ starts, int32_t* __restrict__ ends)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) starts[k] = i;
    if (i == n - 1 || keys[i + 1] != k) ends[k] = i;
}

// Plain copy (HBM counter calibration).
__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ------------------------------------------------------------ launchers --

int launch_order(const TypeTable& tt, const GridGeom& g, const float* centers, uint32_t* okey, uint32_t* ocount,
                 uint32_t* ostart, uint64_t* ostatus, int32_t otiles, uint32_t* octrl, int32_t* corder,
                 uint32_t* flags, hipStream_t st)
{
    if (tt.nc <= 0) return BCS_OK;
    const int nb = (tt.nc + 255) / 256;
    const OrderKeyArgs ka{tt, g, centers, okey, ocount, ostatus, otiles, octrl};
    hipLaunchKernelGGL(k_order_key, dim3(std::max(nb, (otiles + 255) / 256)), dim3(256), 0, st, ka);
    const int rc = scan_exclusive(ocount, ostart, (int64_t)tt.nt << kOrderBits, ostatus, otiles, octrl, flags, st);
    if (rc) return rc;
    const OrderScatterArgs sa{tt.nc, okey, ocount, ostart, corder, flags};
    hipLaunchKernelGGL(k_order_scatter, dim3(nb), dim3(256), 0, st, sa);
    return BCS_OK;
}

void launch_refresh_copies(const float* pos, const float* vel, int32_t n, const int32_t* perm, float4* pvs,
                           hipStream_t st)
{
    if (n > 0 && pvs)
        hipLaunchKernelGGL(k_refresh_copies, dim3((n + 255) / 256), dim3(256), 0, st, pos, vel, n, perm, pvs);
}

void launch_export(const float* soa, int64_t n, int64_t first, int64_t count, int stride, float* out, hipStream_t st)
{
    if (count > 0)
        hipLaunchKernelGGL(k_export, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, soa, n, first, count, stride,
                           out);
}

void launch_tri_nmask(const DeviceGrid& tg, uint32_t* nmask, hipStream_t st)
{
    hipLaunchKernelGGL(k_tri_nmask, dim3((unsigned)((tg.g.cell_count + 255) / 256)), dim3(256), 0, st, tg.g, tg.table,
                       nmask);
}

// {particle, blood cell} of every lane of every chunk, from the processing
// order (run whenever corder changes).
__device__ __forceinline__ void chunk_map_body(const StepArgs& A, int2* __restrict__ out, VBlock vb)
{
    const int chunk = (int)vb.id, l = threadIdx.x;
    const int T = A.chunk_type[chunk];
    const int p = A.tt.ppc[T];
    const bool act = l < A.chunk_n[chunk] * p;
    int32_t cell = 0;
    const int32_t i = act ? chunk_particle(A, T, A.chunk_first[chunk], l / p, l % p, p, cell) : -1;
    out[(size_t)chunk * kBlock + l] = make_int2(i, act ? cell : 0);
}
__global__ void __launch_bounds__(kBlock) k_chunk_map(StepArgs A, int2* __restrict__ out)
{
    chunk_map_body(A, out, hw_block());
}

static dim3 chunk_grid(int nchunks) { return dim3(8 * ((nchunks + 7) / 8)); }

void launch_chunk_map(const StepArgs& A, int2* out, hipStream_t st)
{
    if (A.nchunks > 0) hipLaunchKernelGGL(k_chunk_map, dim3(A.nchunks), dim3(kBlock), 0, st, A, out);
}

void launch_raycand(const StepArgs& A, int nchunks, hipStream_t st)
{
    if (nchunks > 0) hipLaunchKernelGGL(k_raycand, chunk_grid(nchunks), dim3(kBlock), 0, st, A);
}
uint32_t ray_split(int nchunks)
{
    const uint32_t nreg = chunk_grid(nchunks).x;
    return nreg ? std::max(1u, std::min(kRaySplitMax, kRayBlocks / nreg)) : 1u;
}
// The fused launch pays off once k_ray would run one block per region: the
// cell blocks then fill the device by themselves.
bool ray_fuse_auto(int nchunks) { return ray_split(nchunks) == 1u; }
void launch_raycand_ray(const StepArgs& A, int nchunks, hipStream_t st)
{
    if (nchunks > 0) hipLaunchKernelGGL(k_raycand_ray, chunk_grid(nchunks), dim3(kBlock), 0, st, A);
}
void launch_ray(const StepArgs& A, hipStream_t st)
{
    if (A.nchunks <= 0) return;
    const uint32_t split = ray_split(A.nchunks);
    hipLaunchKernelGGL(k_ray, dim3(chunk_grid(A.nchunks).x * split), dim3(256), 0, st, A, split);
}
void launch_cells(const StepArgs& A, int nchunks, bool fused, hipStream_t st)
{
    if (nchunks <= 0) return;
    if (fused) hipLaunchKernelGGL(k_cells<true>, chunk_grid(nchunks), dim3(kBlock), 0, st, A);
    else hipLaunchKernelGGL(k_cells<false>, chunk_grid(nchunks), dim3(kBlock), 0, st, A);
}
void launch_integrate(const StepArgs& A, hipStream_t st)
{
    if (A.tt.np > 0) hipLaunchKernelGGL(k_integrate, dim3((A.tt.np + 255) / 256), dim3(256), 0, st, A);
}
void launch_vein_end(const StepArgs& A, int nchunks, hipStream_t st)
{
    if (nchunks > 0) hipLaunchKernelGGL(k_vein_end, chunk_grid(nchunks), dim3(kBlock), 0, st, A);
}
void launch_bump(uint64_t* step, hipStream_t st) { hipLaunchKernelGGL(k_bump_step, dim3(1), dim3(1), 0, st, step); }
void launch_vein_springs(const VeinArgs& V, hipStream_t st)
{
    if (V.nv > 0) hipLaunchKernelGGL(k_vein_springs, dim3((V.nv + 255) / 256), dim3(256), 0, st, V);
}
void launch_vein(const VeinArgs& V, int mode, hipStream_t st)
{
    const dim3 grid((std::max(V.nv, 1) + 255) / 256);
    switch (mode) {
    case 1: hipLaunchKernelGGL(k_vein<1>, grid, dim3(256), 0, st, V); break;
    case 2: hipLaunchKernelGGL(k_vein<2>, grid, dim3(256), 0, st, V); break;
    case 3: hipLaunchKernelGGL(k_vein<3>, grid, dim3(256), 0, st, V); break;
    case 7: hipLaunchKernelGGL(k_vein<7>, grid, dim3(256), 0, st, V); break;
    default: break;
    }
}
void launch_tri_centers(const float* X, int32_t nv, const uint32_t* idx, int32_t ntri, float* c, hipStream_t st)
{
    if (ntri > 0) hipLaunchKernelGGL(k_tri_centers, dim3((ntri + 255) / 256), dim3(256), 0, st, X, nv, idx, ntri, c);
}
void launch_contacts(const StepArgs& A, int pass, int32_t* counts, const int64_t* offs, int32_t* pairs, hipStream_t st)
{
    if (A.tt.np > 0)
        hipLaunchKernelGGL(k_contact_pairs, dim3((A.tt.np + 255) / 256), dim3(256), 0, st, A, pass, counts, offs, pairs);
}
static TriBoxArgs tri_box_args(int32_t ntri, const uint4* trun, const uint4* trec, const float* X, int32_t nv,
                               float4* tdata, float4* aabb)
{
    return TriBoxArgs{trun, trec, ntri, X, nv, tdata, aabb};
}
constexpr uint32_t kTriBoxBlocks = 1024;
void launch_tri_records(const DeviceGrid& tg, const uint32_t* idx, uint4* trun, uint4* trec, hipStream_t st)
{
    if (tg.n > 0)
        hipLaunchKernelGGL(k_tri_records, dim3((tg.n + 255) / 256), dim3(256), 0, st, tg.runs, tg.ctrl, tg.table,
                           tg.keys_sorted, tg.perm, tg.n, idx, trun, trec);
}
VeinPrepArgs vein_prep_args(const VeinArgs& V, int32_t ntri, const uint4* trun, const uint4* trec, float4* tdata,
                            float4* aabb, uint32_t* blocks)
{
    const uint32_t nsb = (uint32_t)((std::max(V.nv, 0) + 255) / 256);
    const uint32_t ntb = ntri > 0 ? kTriBoxBlocks : 0u;
    *blocks = nsb + ntb;
    return VeinPrepArgs{V, tri_box_args(ntri, trun, trec, V.X, V.nv, tdata, aabb), nsb};
}
void launch_vein_prep(const VeinArgs& V, int32_t ntri, const uint4* trun, const uint4* trec, float4* tdata,
                      float4* aabb, hipStream_t st)
{
    uint32_t blocks = 0;
    const VeinPrepArgs P = vein_prep_args(V, ntri, trun, trec, tdata, aabb, &blocks);
    if (blocks > 0) hipLaunchKernelGGL(k_vein_prep, dim3(blocks), dim3(256), 0, st, P);
}
void launch_gather_start(const uint2* table, const uint32_t* idx, int32_t n, uint32_t* out, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(k_gather_start, dim3((n + 255) / 256), dim3(256), 0, st, table, idx, n, out);
}
void launch_dense_ranges(const uint32_t* keys_sorted, int32_t n, uint64_t cell_count, int32_t* starts,
                         int32_t* ends, hipStream_t st)
{
    if (cell_count > 0) {
        const uint64_t blocks = std::min<uint64_t>((cell_count + 255) / 256, 65536);
        hipLaunchKernelGGL(k_fill_ranges, dim3((uint32_t)blocks), dim3(256), 0, st, starts, ends, cell_count);
    }
    if (n > 0) hipLaunchKernelGGL(k_dense_ranges, dim3((n + 255) / 256), dim3(256), 0, st, keys_sorted, n, starts, ends);
}
void launch_copy(const void* src, void* dst, int64_t n16, hipStream_t st)
{
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, st, (const float4*)src, (float4*)dst, n16);
}

// Network input of every cell of the surrogate's type, one wave per cell
// (lane a = the cell's particle a): 126 centroid-relative positions /
// pos_scale, 126 velocities relative to the cell's mean velocity / vel_scale,
// 42 contact flags of the last frame, zeros to 320 columns (bf16 X); f32
// (optional) receives the first 294 columns unscaled.
__global__ void __launch_bounds__(256) k_surr_features(const float* __restrict__ pos, const float* __restrict__ vel,
                                                       int64_t np, int32_t pstart, int32_t ncells,
                                                       const uint8_t* __restrict__ touched, float pos_scale,
                                                       float vel_scale, uint16_t* __restrict__ X,
                                                       float* __restrict__ f32)
{
    const int lane = threadIdx.x & 63;
    const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= ncells) return;  // whole waves: no barrier below
    const bool act = lane < kSurrPpc;
    const int64_t i = pstart + c * kSurrPpc + lane;
    float v[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float flag = 0.f;
    if (act) {
        v[0] = pos[i];
        v[1] = pos[np + i];
        v[2] = pos[2 * np + i];
        v[3] = vel[i];
        v[4] = vel[np + i];
        v[5] = vel[2 * np + i];
        flag = touched && touched[i] ? 1.0f : 0.0f;
    }
    float mean[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        float t = v[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
        mean[k] = t / (float)kSurrPpc;
    }
    uint16_t* row = X + c * kSurrInPad;
    if (act) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float rp = v[d] - mean[d], rv = v[3 + d] - mean[3 + d];
            row[3 * lane + d] = bf16_bits(rp / pos_scale);
            row[126 + 3 * lane + d] = bf16_bits(rv / vel_scale);
            if (f32) {
                f32[c * kSurrIn + 3 * lane + d] = rp;
                f32[c * kSurrIn + 126 + 3 * lane + d] = rv;
            }
        }
        row[252 + lane] = bf16_bits(flag);
        if (f32) f32[c * kSurrIn + 252 + lane] = flag;
    }
    if (lane < kSurrInPad - kSurrIn) row[kSurrIn + lane] = 0;
}

void launch_surr_features(const float* pos, const float* vel, int64_t np, int32_t pstart, int32_t ncells,
                          const uint8_t* touched, float pos_scale, float vel_scale, uint16_t* X, float* f32,
                          hipStream_t st)
{
    if (ncells > 0)
        hipLaunchKernelGGL(k_surr_features, dim3((unsigned)((ncells + 3) / 4)), dim3(256), 0, st, pos, vel, np, pstart,
                           ncells, touched, pos_scale, vel_scale, X, f32);
}
void launch_cells_surr(const StepArgs& A, int nchunks, hipStream_t st)
{
    if (nchunks > 0) hipLaunchKernelGGL((k_cells<true, true>), chunk_grid(nchunks), dim3(kBlock), 0, st, A);
}
// The particle part of a split frame: k_springs (per chunk), k_contacts (per
// tile of sorted slots), k_respond (per chunk).  part: 0, 1 or 2.
void launch_split_part(const StepArgs& A, int part, bool surr, hipStream_t st)
{
    if (A.nchunks <= 0) return;
    if (part == 0) {
        if (surr) hipLaunchKernelGGL(k_springs<true>, chunk_grid(A.nchunks), dim3(kBlock), 0, st, A);
        else hipLaunchKernelGGL(k_springs<false>, chunk_grid(A.nchunks), dim3(kBlock), 0, st, A);
    } else if (part == 1) {
        if (A.ntiles <= 0) return;
        if (surr) hipLaunchKernelGGL(k_contacts<true>, chunk_grid(A.ntiles), dim3(kBlock), 0, st, A);
        else hipLaunchKernelGGL(k_contacts<false>, chunk_grid(A.ntiles), dim3(kBlock), 0, st, A);
    } else {
        hipLaunchKernelGGL(k_respond, chunk_grid(A.nchunks), dim3(kBlock), 0, st, A);
    }
}

// ---- batched launches: one launch per kernel for many simulations ----------
// Each kernel below runs the body of its single-simulation form on the
// virtual blocks of every simulation k (arguments a[k]); see for_virtual_blocks.

__global__ void __launch_bounds__(kBlock) k_raycand_batch(const StepArgs* __restrict__ a,
                                                          const uint32_t* __restrict__ start, int32_t nsim, uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { raycand_block<false>(a[k], vb); });
}
__global__ void __launch_bounds__(kBlock) k_raycand_ray_batch(const StepArgs* __restrict__ a,
                                                              const uint32_t* __restrict__ start, int32_t nsim,
                                                              uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { raycand_block<true>(a[k], vb); });
}
__global__ void __launch_bounds__(256) k_ray_batch(const StepArgs* __restrict__ a, const uint32_t* __restrict__ split,
                                                   const uint32_t* __restrict__ start, int32_t nsim, uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { ray_body(a[k], split[k], vb); });
}
__global__ void __launch_bounds__(kBlock) k_cells_batch(const StepArgs* __restrict__ a,
                                                        const uint32_t* __restrict__ start, int32_t nsim, uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { cells_body<true>(a[k], vb); });
}
__global__ void __launch_bounds__(kBlock) k_springs_batch(const StepArgs* __restrict__ a,
                                                          const uint32_t* __restrict__ start, int32_t nsim, uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { cells_body<true, false, kCellsSprings>(a[k], vb); });
}
__global__ void __launch_bounds__(kBlock) k_contacts_batch(const StepArgs* __restrict__ a,
                                                           const uint32_t* __restrict__ start, int32_t nsim, uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { contacts_body<false>(a[k], vb); });
}
__global__ void __launch_bounds__(kBlock) k_respond_batch(const StepArgs* __restrict__ a,
                                                          const uint32_t* __restrict__ start, int32_t nsim, uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { respond_body(a[k], vb); });
}
__global__ void __launch_bounds__(256) k_vein_batch(const VeinArgs* __restrict__ a, const uint32_t* __restrict__ start,
                                                    int32_t nsim, uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { vein_body<7>(a[k], vb); });
}
__global__ void __launch_bounds__(256) k_vein_prep_batch(const VeinPrepArgs* __restrict__ a,
                                                         const uint32_t* __restrict__ start, int32_t nsim, uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { vein_prep_body(a[k], vb); });
}
__global__ void __launch_bounds__(256) k_order_key_batch(const OrderKeyArgs* __restrict__ a,
                                                         const uint32_t* __restrict__ start, int32_t nsim, uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { order_key_body(a[k], vb); });
}
__global__ void __launch_bounds__(256) k_order_scatter_batch(const OrderScatterArgs* __restrict__ a,
                                                             const uint32_t* __restrict__ start, int32_t nsim,
                                                             uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { order_scatter_body(a[k], vb); });
}
__global__ void __launch_bounds__(kBlock) k_chunk_map_batch(const StepArgs* __restrict__ a, int2* const* __restrict__ out,
                                                            const uint32_t* __restrict__ start, int32_t nsim,
                                                            uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { chunk_map_body(a[k], out[k], vb); });
}

// Blocks of each single-simulation launch, as the batched launches split
// them (the same counts the launchers above use).
uint32_t cell_blocks(int nchunks) { return nchunks > 0 ? chunk_grid(nchunks).x : 0u; }
uint32_t tile_blocks(int64_t np) { return np > 0 ? chunk_grid((int)((np + kBlock - 1) / kBlock)).x : 0u; }
uint32_t ray_blocks(int nchunks) { return nchunks > 0 ? chunk_grid(nchunks).x * ray_split(nchunks) : 0u; }
uint32_t vein_blocks(int32_t nv) { return (uint32_t)((std::max(nv, 1) + 255) / 256); }
uint32_t order_key_blocks(const TypeTable& tt, int32_t otiles)
{
    return tt.nc > 0 ? (uint32_t)std::max((tt.nc + 255) / 256, (otiles + 255) / 256) : 0u;
}
uint32_t order_scatter_blocks(const TypeTable& tt) { return tt.nc > 0 ? (uint32_t)((tt.nc + 255) / 256) : 0u; }

void batch_raycand(bool fused, const StepArgs* a, const uint32_t* start, int32_t nsim, uint32_t grid, uint32_t run,
                   hipStream_t st)
{
    if (!grid) return;
    if (fused) hipLaunchKernelGGL(k_raycand_ray_batch, dim3(grid), dim3(kBlock), 0, st, a, start, nsim, run);
    else hipLaunchKernelGGL(k_raycand_batch, dim3(grid), dim3(kBlock), 0, st, a, start, nsim, run);
}
void batch_ray(const StepArgs* a, const uint32_t* split, const uint32_t* start, int32_t nsim, uint32_t grid,
               uint32_t run, hipStream_t st)
{
    if (grid) hipLaunchKernelGGL(k_ray_batch, dim3(grid), dim3(256), 0, st, a, split, start, nsim, run);
}
void batch_cells(const StepArgs* a, const uint32_t* start, int32_t nsim, uint32_t grid, uint32_t run, hipStream_t st)
{
    if (grid) hipLaunchKernelGGL(k_cells_batch, dim3(grid), dim3(kBlock), 0, st, a, start, nsim, run);
}
void batch_split_part(const StepArgs* a, int part, const uint32_t* start, int32_t nsim, uint32_t grid, uint32_t run,
                      hipStream_t st)
{
    if (!grid) return;
    if (part == 0) hipLaunchKernelGGL(k_springs_batch, dim3(grid), dim3(kBlock), 0, st, a, start, nsim, run);
    else if (part == 1) hipLaunchKernelGGL(k_contacts_batch, dim3(grid), dim3(kBlock), 0, st, a, start, nsim, run);
    else hipLaunchKernelGGL(k_respond_batch, dim3(grid), dim3(kBlock), 0, st, a, start, nsim, run);
}
void batch_vein(const VeinArgs* a, const uint32_t* start, int32_t nsim, uint32_t grid, uint32_t run, hipStream_t st)
{
    if (grid) hipLaunchKernelGGL(k_vein_batch, dim3(grid), dim3(256), 0, st, a, start, nsim, run);
}
void batch_vein_prep(const VeinPrepArgs* a, const uint32_t* start, int32_t nsim, uint32_t grid, uint32_t run,
                     hipStream_t st)
{
    if (grid) hipLaunchKernelGGL(k_vein_prep_batch, dim3(grid), dim3(256), 0, st, a, start, nsim, run);
}
void batch_order(const OrderKeyArgs* ka, const uint32_t* kstart, uint32_t kgrid, const ScanArgs* sa,
                 const uint32_t* sstart, uint32_t sgrid, const OrderScatterArgs* oa, const uint32_t* ostart,
                 uint32_t ogrid, const StepArgs* a, int2* const* out, const uint32_t* cstart, uint32_t cgrid,
                 int32_t nsim, uint32_t run, hipStream_t st)
{
    if (kgrid) hipLaunchKernelGGL(k_order_key_batch, dim3(kgrid), dim3(256), 0, st, ka, kstart, nsim, run);
    scan_batch(sa, sstart, nsim, sgrid, run, st);
    if (ogrid) hipLaunchKernelGGL(k_order_scatter_batch, dim3(ogrid), dim3(256), 0, st, oa, ostart, nsim, run);
    if (cgrid) hipLaunchKernelGGL(k_chunk_map_batch, dim3(cgrid), dim3(kBlock), 0, st, a, out, cstart, nsim, run);
}

}  // namespace bcs
