// device_math.hpp -- float3 arithmetic with the reference's operand order
// (src/utilities/math.cuh:11-104).  Every source file of libbcs is compiled
// with -ffp-contract=off and IEEE division/sqrt, so each helper below rounds
// exactly like the corresponding reference expression evaluated op by op.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bcs {

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 smul(float a, f3 v) { return f3{a * v.x, a * v.y, a * v.z}; }
__device__ __forceinline__ f3 sdiv(f3 v, float a) { return f3{v.x / a, v.y / a, v.z / a}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 u, f3 v)
{
    return f3{u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}
__device__ __forceinline__ float len(f3 v) { return sqrtf(dot(v, v)); }

// normalize: v / sqrtf(dot(v,v)); NaN -> (0,0,0) with a sticky flag bit.
__device__ __forceinline__ f3 normalize(f3 v, uint32_t& fl)
{
    f3 vn = sdiv(v, sqrtf(dot(v, v)));
    if (__builtin_isnan(vn.x) || __builtin_isnan(vn.y) || __builtin_isnan(vn.z)) {
        fl |= BCS_FLAG_NAN_NORMALIZE;
        return f3{0.f, 0.f, 0.f};
    }
    return vn;
}

// static_cast<int>(float) with saturation and NaN -> 0.
__host__ __device__ __forceinline__ int32_t sat_i32(float q)
{
    if (q != q) return 0;
    if (q >= 2147483647.0f) return 2147483647;
    if (q <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)q;
}
__device__ __forceinline__ uint32_t sat_u32(float q)
{
    if (q != q || q <= 0.0f) return 0u;
    if (q >= 4294967295.0f) return 4294967295u;
    return (uint32_t)q;
}

__device__ __forceinline__ f3 load3(const float* __restrict__ a, size_t n, size_t i)
{
    return f3{a[i], a[n + i], a[2 * n + i]};
}
__device__ __forceinline__ void store3(float* __restrict__ a, size_t n, size_t i, f3 v)
{
    a[i] = v.x;
    a[n + i] = v.y;
    a[2 * n + i] = v.z;
}

// SplitMix64 finalizer (Steele, Lea, Flood 2014) and the per-cell teleport
// draw in (0, 1] with 24 random bits.
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float uniform01(uint64_t seed, uint64_t step, uint64_t cell, uint32_t k)
{
    uint64_t h = mix64(seed ^ mix64(step ^ mix64(cell * 2ull + (uint64_t)k)));
    return (float)((h >> 40) + 1ull) * 0x1p-24f;
}

// Unsigned 32-bit division by a run-time constant d (Granlund & Montgomery,
// PLDI 1994: round-up multiplier with its 33rd bit folded into an add):
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, t = mulhi(m, n),
// n / d = (t + ((n - t) >> 1)) >> (l - 1); s == 255 marks d == 1.
struct Div32 {
    uint32_t m, s;
};
__host__ __device__ inline Div32 make_div32(uint32_t d)
{
    if (d <= 1u) return Div32{0u, 255u};
    uint32_t l = 0;
    while (l < 32u && (1ull << l) < d) ++l;
    const uint64_t m = ((1ull << 32) * ((1ull << l) - d)) / d + 1ull;
    return Div32{(uint32_t)m, l - 1u};
}
__device__ __forceinline__ uint32_t udiv(uint32_t n, Div32 d)
{
    if (d.s == 255u) return n;
    const uint32_t t = __umulhi(d.m, n);
    return (t + ((n - t) >> 1)) >> d.s;
}

// Uniform grid geometry (src/grids/uniform_grid.cuh:15-65, bounds from
// src/meta_factory/vein_factory.hpp:46-86).
struct GridGeom {
    float minv[3];
    float maxv[3];
    float extent[3];
    float cell[3];
    int32_t cnt[3];
    uint32_t pad;
    uint64_t cell_count;
    Div32 div_x, div_y;  // division by cnt[0] and cnt[1] (key -> cell coordinates)
};

// Cell coordinates of a key in [0, cellCount) (key = x + cnt0 * (y + cnt1 * z)).
__device__ __forceinline__ void grid_coords(const GridGeom& g, uint32_t key, uint32_t& x, uint32_t& y, uint32_t& z)
{
    const uint32_t q = udiv(key, g.div_x);
    x = key - q * (uint32_t)g.cnt[0];
    z = udiv(q, g.div_y);
    y = q - z * (uint32_t)g.cnt[1];
}

// calculateIdForCell, src/grids/uniform_grid.cu:24-36 (max/min macros of
// :20-21, upper clamp in world units).  Keys outside [0, cellCount) are
// clamped to cellCount-1 and flagged.
__host__ __device__ __forceinline__ uint32_t cell_key(const GridGeom& g, float x, float y, float z, uint32_t& fl)
{
    if (x < g.minv[0] || x > g.maxv[0] || y < g.minv[1] || y > g.maxv[1] || z < g.minv[2] || z > g.maxv[2])
        fl |= BCS_FLAG_OUT_OF_GRID;
    float p[3] = {x, y, z};
    int64_t id[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float q = (p[d] - g.minv[d]) / g.cell[d];
        float m = (0.0f > q) ? 0.0f : q;
        float c = (g.extent[d] > m) ? m : g.extent[d];
        id[d] = (int64_t)sat_i32(c);
    }
    int64_t key = id[2] * (int64_t)g.cnt[0] * (int64_t)g.cnt[1] + id[1] * (int64_t)g.cnt[0] + id[0];
    if (key < 0 || (uint64_t)key >= g.cell_count) {
        fl |= BCS_FLAG_OUT_OF_GRID;
        key = (int64_t)g.cell_count - 1;
    }
    return (uint32_t)key;
}

// Occupancy map of a grid in bricks of 8 x 8 x 8 cells: 2 bits per cell
// (0 empty, 1 one object, 2 more), one 16-bit word per 8-cell x-row of a
// brick and 64 words -- one 128-byte line -- per brick, so the 3 x 3 x 3
// neighbourhood of a cell lies in one to eight lines (two on average).
struct BrickMap {
    uint16_t* words;
    int32_t cnt[3];  // grid cells per axis
    uint32_t nbx, nby;
    Div32 div_x, div_y;
};
constexpr uint32_t kNoWord = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t brick_word(const BrickMap& b, uint32_t x, uint32_t y, uint32_t z)
{
    const uint32_t br = (x >> 3) + b.nbx * ((y >> 3) + b.nby * (z >> 3));
    return br * 64u + (z & 7u) * 8u + (y & 7u);
}
// Cell coordinates of a key in [0, cellCount).
__device__ __forceinline__ void key_coords(const BrickMap& b, uint32_t key, uint32_t& x, uint32_t& y, uint32_t& z)
{
    const uint32_t q = udiv(key, b.div_x);
    x = key - q * (uint32_t)b.cnt[0];
    z = udiv(q, b.div_y);
    y = q - z * (uint32_t)b.cnt[1];
}
__device__ __forceinline__ uint32_t cell_code(const BrickMap& b, uint32_t key)
{
    uint32_t x, y, z;
    key_coords(b, key, x, y, z);
    return ((uint32_t)b.words[brick_word(b, x, y, z)] >> (2u * (x & 7u))) & 3u;
}

// Face-clipped neighbour offsets of one axis
// (src/simulation/particle_collisions.cuh:270-413, signed ids).
__device__ __forceinline__ void axis_range_i(int32_t id, int32_t cnt, int& lo, int& hi)
{
    if (id < 1) { lo = 0; hi = 1; }
    else if (id > cnt - 2) { lo = -1; hi = 0; }
    else { lo = -1; hi = 1; }
}
// Same with unsigned ids (src/simulation/vein_collisions.cu:82-232).
__device__ __forceinline__ void axis_range_u(uint32_t id, int32_t cnt, int& lo, int& hi)
{
    if (id < 1u) { lo = 0; hi = 1; }
    else if (id > (uint32_t)(cnt - 2)) { lo = -1; hi = 0; }
    else { lo = -1; hi = 1; }
}

}  // namespace bcs
