// grid.hip -- uniform-grid build for gfx950: cell hash, stable counting sort
// and cell-start table.
//
// Replaces UniformGrid::calculateGrid (src/grids/uniform_grid.cu:129-155):
// calculateCellIdKernel (:38-49), thrust::stable_sort_by_key (:144-147) and
// calculateStartAndEndOfCellKernel (:51-80).
//
// Sort: keys are split into buckets of 2^shift consecutive cell ids.
//   k_key      key per object, bucket histogram (per block in LDS, one add per bucket)
//   k_scan     exclusive scan of the histogram (single pass, decoupled look-back)
//   k_scatter  (key, index) pairs into their bucket segments
//   k_rank     in-bucket rank by counting over (key, index): the output order is
//              exactly thrust's stable order (ties by object index); run starts
//              write the dense cell-start table
//              buckets larger than the threshold are then shared out over the
//              blocks: bitonic sort of the (key, index) pairs in LDS (up to
//              kSortLds), else an LSD radix sort of the bucket in one block
//              (8-bit digits of (key - bucket base) << index bits | index,
//              stable passes through two global scratch arrays)
// The optional 2-bit brick occupancy map (one word per 8-cell x-group):
//   buckets of 2^7 cells or more: written by k_rank (the first run of every
//              group writes the group's word; a group straddling two buckets
//              ORs its codes in, rare at these bucket sizes)
//   smaller buckets (many straddling groups): k_occ from the sorted keys
// Cell table: dense {start, count} per cell id (8 bytes, one load per lookup),
// written for occupied cells only.  With clear_table (the triangle grid) k_key
// first zeroes the entries of the previous build's keys, so every entry is
// current or {0, 0}; without it (the particle grid) k_key zeroes the map
// words the previous build wrote and readers consult the map first.
#include "grid.hpp"

#include <algorithm>
#include <cstdio>

namespace bcs {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;  // per thread, loaded and stored as uint4
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr uint32_t kRankThreshold = 64;
constexpr uint32_t kSortLds = 2048;  // bucket size sorted in LDS (larger ones: radix sort through global scratch)
constexpr uint64_t kStatusAgg = 1ull << 62;
constexpr uint64_t kStatusInc = 2ull << 62;
constexpr uint64_t kSpinLimit = 1ull << 22;

// Lanes of the key kernel: up to kKeyRounds rounds of one lane per thread
// per block, so that a block counts the objects of rounds * 256 lanes (with a
// lane order: neighbouring blood cells) into its buckets before it adds to the
// histogram, once per distinct bucket.  Builds of fewer than kKeyRoundsFrom
// lanes take one round (enough blocks to fill the GPU).
constexpr int kKeyRounds = 4;
constexpr int64_t kKeyRoundsFrom = 1 << 20;
constexpr uint32_t kKeyHashBits = 10;
constexpr uint32_t kKeyHash = 1u << kKeyHashBits;  // LDS bucket table (room for every lane's bucket)
constexpr uint32_t kKeyEmpty = 0xFFFFFFFFu;
static_assert(kKeyHash >= 256u * kKeyRounds && (1u << 8) >= 256u, "the bucket table must hold a bucket per lane");

__device__ __forceinline__ void key_body(const KeyArgs& a, VBlock vb)
{
    __shared__ uint32_t s_bkt[kKeyHash];  // bucket of each table entry (kKeyEmpty: free)
    __shared__ uint32_t s_cnt[kKeyHash];  // the block's objects in it, then the block's first slot in it
    const GridGeom& g = a.g;
    const float* __restrict__ pos = a.pos;
    const int64_t stride = a.stride;
    const int32_t n = a.n, n_prev = a.n_prev;
    uint32_t* __restrict__ keys = a.keys;
    uint2* __restrict__ table = a.table;
    const BrickMap& bm = a.bm;
    const uint32_t* __restrict__ gpos = a.gpos;
    const int2* __restrict__ order = a.order;
    const uint32_t tid = threadIdx.x;
    {
        const int64_t t0 = (int64_t)vb.id * blockDim.x + tid, nthreads = (int64_t)vb.count * blockDim.x;
        for (int64_t t = t0; t < a.ntiles; t += nthreads) a.status[t] = 0ull;
    }
    if (vb.id == 0 && tid == 0) {
        a.ctrl[0] = 0u;  // scan ticket
        a.ctrl[1] = 0u;  // big-bucket count
        a.ctrl[2] = 0u;  // occupied cells (run list)
    }
    // the table in use: an entry per lane of the block's rounds
    const uint32_t hbits = a.rounds > 1 ? kKeyHashBits : 8u, hsize = 1u << hbits;
    for (uint32_t e = tid; e < hsize; e += blockDim.x) {
        s_bkt[e] = kKeyEmpty;
        s_cnt[e] = 0u;
    }
    __syncthreads();
    uint32_t fl = 0;
    int32_t obj[kKeyRounds];
    uint32_t ent[kKeyRounds], rank[kKeyRounds];
#pragma unroll
    for (int r = 0; r < kKeyRounds; ++r) {
        obj[r] = -1;
        ent[r] = kKeyHash;
        rank[r] = 0u;
    }
#pragma unroll
    for (int r = 0; r < kKeyRounds; ++r) {
        if (r >= a.rounds) break;  // (uniform)
        const int64_t gid = ((int64_t)vb.id * a.rounds + r) * blockDim.x + tid;
        // every load first (one round trip), then the clears and the new key
        uint32_t old = 0xFFFFFFFFu, w = kNoWord;
        float x = 0.f, y = 0.f, z = 0.f;
        if (gid < n_prev) {  // the previous build's cell of this object
            if (a.clear_table) old = keys[gid];
            if (gpos) w = gpos[gid];  // map word written by the previous build (indexed by its sorted slots)
        }
        // the object of this lane: gid, or with `order` the object of lane
        // gid (objects in spatial order, so that the block's lanes share
        // buckets; -1: no object)
        int64_t i = gid < n ? gid : -1;
        if (order) i = gid < a.norder ? (int64_t)order[gid].x : -1;
        const bool have = i >= 0 && i < n;
        if (have) {
            x = pos[i];
            y = pos[stride + i];
            z = pos[2 * stride + i];
        }
        if ((uint64_t)old < g.cell_count) table[old] = make_uint2(0u, 0u);
        if (w != kNoWord) bm.words[w] = 0;
        obj[r] = have ? (int32_t)i : -1;
        if (have) {
            const uint32_t k = cell_key(g, x, y, z, fl);
            keys[i] = k;
            const uint32_t b = k >> a.shift;
            // the bucket's entry in the block's table (open addressing); its
            // count before this object is the object's place among the
            // block's objects of that bucket
            uint32_t h = (b * 2654435761u) >> (32u - hbits);
            for (uint32_t probe = 0; probe < hsize; ++probe) {
                const uint32_t prev = atomicCAS(&s_bkt[h], kKeyEmpty, b);
                if (prev == kKeyEmpty || prev == b) {
                    ent[r] = h;
                    break;
                }
                h = (h + 1u) & (hsize - 1u);
            }
            // (a full table cannot occur: it has an entry per lane; the
            // histogram is then counted into directly)
            rank[r] = ent[r] < kKeyHash ? atomicAdd(&s_cnt[ent[r]], 1u) : atomicAdd(&a.bcount[b], 1u);
        }
    }
    __syncthreads();
    // one histogram add per bucket of the block; the entry becomes the
    // block's first slot in the bucket (slot order within a bucket is free:
    // the ranking orders by key and index)
    for (uint32_t e = tid; e < hsize; e += blockDim.x)
        if (s_cnt[e]) s_cnt[e] = atomicAdd(&a.bcount[s_bkt[e]], s_cnt[e]);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kKeyRounds; ++r)
        if (obj[r] >= 0) a.arrival[obj[r]] = ent[r] < kKeyHash ? s_cnt[ent[r]] + rank[r] : rank[r];
    if (fl) atomicOr(a.flags, fl);
}
__global__ void __launch_bounds__(256) k_key(KeyArgs a) { key_body(a, hw_block()); }

// Exclusive scan of `count[0..m)` into out[0..m] (out[m] = total), single
// pass with decoupled look-back over tiles taken in ticket order.  Entries
// larger than big_thr are appended to biglist (when biglist != nullptr).
// clear_count: the histogram is zeroed as it is read (ready for the next
// build).
__device__ __forceinline__ void scan_body(const ScanArgs& a)
{
    uint32_t* __restrict__ count = a.count;
    uint32_t* __restrict__ out = a.out;
    const int64_t m = a.m;
    uint64_t* __restrict__ status = a.status;
    const int32_t ntiles = a.ntiles;
    uint32_t* __restrict__ ctrl = a.ctrl;
    uint32_t* __restrict__ biglist = a.biglist;
    const uint32_t big_thr = a.big_thr, big_cap = a.big_cap;
    uint32_t* __restrict__ flags = a.flags;
    const int clear_count = a.clear_count;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave[kScanThreads / 64];
    __shared__ uint32_t s_prefix;
    const int tid = threadIdx.x;
    if (tid == 0) s_tile = atomicAdd(&ctrl[0], 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    if (tile >= ntiles) {  // stale ticket counter: never index past the status array
        if (tid == 0) atomicOr(flags, BCS_FLAG_SCAN_TIMEOUT);
        return;
    }
    const int64_t base = tile * kScanTile + (int64_t)tid * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
    if (base + kScanItems <= m) {  // whole chunk in range: 16-byte loads (base is 64-byte aligned)
        const uint4* c4 = reinterpret_cast<const uint4*>(count + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            const uint4 t = c4[q];
            v[4 * q] = t.x;
            v[4 * q + 1] = t.y;
            v[4 * q + 2] = t.z;
            v[4 * q + 3] = t.w;
        }
        if (clear_count) {  // the histogram starts the next build at zero
            uint4* w4 = reinterpret_cast<uint4*>(count + base);
#pragma unroll
            for (int q = 0; q < kScanItems / 4; ++q) w4[q] = make_uint4(0u, 0u, 0u, 0u);
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) v[k] = (base + k < m) ? count[base + k] : 0u;
        if (clear_count)
            for (int k = 0; k < kScanItems; ++k)
                if (base + k < m) count[base + k] = 0u;
    }
    const int lane = tid & 63, wave = tid >> 6;
    uint32_t nbig = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        nbig += (biglist && base + k < m && v[k] > big_thr) ? 1u : 0u;
        sum += v[k];
    }
    if (biglist) {  // big buckets appended with one counter update per wave
        uint32_t binc = nbig;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(binc, d, 64);
            if (lane >= d) binc += o;
        }
        const uint32_t wtotal = __shfl(binc, 63, 64);
        uint32_t wbase = 0;
        if (lane == 0 && wtotal) wbase = atomicAdd(&ctrl[1], wtotal);
        wbase = __shfl(wbase, 0, 64);
        uint32_t slot = wbase + binc - nbig;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
            if (base + k < m && v[k] > big_thr) {
                if (slot < big_cap) biglist[slot] = (uint32_t)(base + k);
                else atomicOr(flags, BCS_FLAG_SCAN_TIMEOUT);  // more than n / big_thr: inconsistent counts
                ++slot;
            }
    }
    // wave64 inclusive scan of the per-thread sums
    uint32_t inc = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    uint32_t wave_off = 0, agg = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        if (w < wave) wave_off += s_wave[w];
        agg += s_wave[w];
    }
    if (tid == 0) {
        uint32_t prefix = 0;
        if (tile == 0) {
            __hip_atomic_store(&status[0], kStatusInc | (uint64_t)agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&status[tile], kStatusAgg | (uint64_t)agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int64_t j = tile - 1;
            uint64_t spins = 0;
            while (j >= 0) {
                uint64_t w = __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                uint64_t st = w & (3ull << 62);
                if (st == 0) {
                    if (++spins > kSpinLimit) {
                        atomicOr(flags, BCS_FLAG_SCAN_TIMEOUT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                prefix += (uint32_t)(w & 0xFFFFFFFFull);
                if (st == kStatusInc) break;
                --j;
            }
            __hip_atomic_store(&status[tile], kStatusInc | (uint64_t)(prefix + agg), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        s_prefix = prefix;
    }
    __syncthreads();
    uint32_t run = s_prefix + wave_off + (inc - sum);
    uint32_t o[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        o[k] = run;
        run += v[k];
    }
    if (base + kScanItems <= m + 1) {  // out holds m + 1 entries
        uint4* o4 = reinterpret_cast<uint4*>(out + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) o4[q] = make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
            if (base + k <= m) out[base + k] = o[k];
    }
}
// (the tile comes from the ticket counter, not from the block index)
__global__ void __launch_bounds__(kScanThreads) k_scan(ScanArgs a) { scan_body(a); }

// brec (nullable): the object's position and velocity travel with its item
// (read here in object order, copied to the sorted slot by k_rank).
__device__ __forceinline__ void scatter_body(const ScatterArgs& a, VBlock vb)
{
    const uint32_t* __restrict__ keys = a.keys;
    const int32_t n = a.n;
    const uint32_t* __restrict__ bstart = a.bstart;
    const float* __restrict__ xyz = a.xyz;
    const float* __restrict__ vel = a.vel;
    const int64_t stride = a.stride;
    float4* __restrict__ brec = a.brec;
    const int32_t i = vb.id * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 rx, rv;
    if (brec) {
        rx = make_float4(xyz[i], xyz[stride + i], xyz[2 * stride + i], __int_as_float(i));
        rv = make_float4(vel[i], vel[stride + i], vel[2 * stride + i], 0.0f);
    }
    const uint32_t k = keys[i];
    const uint32_t arr = a.arrival[i];
    const uint32_t b = k >> a.shift;
    const uint32_t s0 = bstart[b], s1 = bstart[b + 1];
    const uint32_t pos = s0 + arr;
    if (pos >= s1 || s1 > (uint32_t)n) {  // inconsistent histogram: never write out of bounds
        atomicOr(a.flags, BCS_FLAG_SCAN_TIMEOUT);
        return;
    }
    a.items[pos] = make_uint2(k, (uint32_t)i);
    if (brec) {
        brec[2 * (size_t)pos] = rx;
        brec[2 * (size_t)pos + 1] = rv;
    }
}
__global__ void __launch_bounds__(256) k_scatter(ScatterArgs a) { scatter_body(a, hw_block()); }


// The 8-cell x-group of `key` (cells x&~7 .. x|7 of its row): keys [gs, ge)
// and the map word holding their 2-bit codes.
struct OccGroup {
    uint32_t gs, ge, word, bit;  // bit: position of key's code in the word
};
__device__ __forceinline__ OccGroup occ_group(const BrickMap& bm, uint32_t key)
{
    uint32_t x, y, z;
    key_coords(bm, key, x, y, z);
    OccGroup q;
    q.gs = key - (x & 7u);
    q.ge = q.gs + min(8u, (uint32_t)bm.cnt[0] - (x & ~7u));
    q.word = brick_word(bm, x, y, z);
    q.bit = 2u * (x & 7u);
    return q;
}
// Record key k2 (one object) in the group masks: seen1 = keys present,
// seen2 = keys with two or more objects.
__device__ __forceinline__ void occ_note(uint32_t k2, uint32_t gs, uint32_t ge, uint32_t& seen1, uint32_t& seen2)
{
    if (k2 >= gs && k2 < ge) {
        const uint32_t b = 1u << (k2 - gs);
        seen2 |= seen1 & b;
        seen1 |= b;
    }
}
// The group's word: code 2 for keys of two or more objects, 1 for one, 0 else.
__device__ __forceinline__ uint32_t occ_word(uint32_t seen1, uint32_t seen2)
{
    uint32_t w = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) w |= (((seen2 >> k) & 1u) ? 2u : ((seen1 >> k) & 1u)) << (2 * k);
    return w;
}

// Occupancy map entry of sorted slot pos.  A run start whose group lies in
// one bucket is written whole by the group's first run (owner, with the
// group's masks); a group that straddles two buckets, or one whose masks are
// not known (known = false), gets each run's code by an atomic OR.  Every
// slot records the word it wrote (or kNoWord), so that the next build clears
// exactly these words.
__device__ __forceinline__ void map_write(const RankOut& o, uint32_t key, uint32_t pos, bool first, uint32_t count,
                                          bool known, bool owner, uint32_t seen1, uint32_t seen2)
{
    if (!o.gpos) return;
    uint32_t w = kNoWord;
    if (first) {
        const OccGroup q = occ_group(o.bm, key);
        const bool straddle = (q.gs >> o.shift) != ((q.ge - 1u) >> o.shift);
        if (straddle || !known) {
            uint32_t* cont = reinterpret_cast<uint32_t*>(o.bm.words) + (q.word >> 1);
            atomicOr(cont, (count > 1u ? 2u : 1u) << (q.bit + 16u * (q.word & 1u)));
            w = q.word;
        } else if (owner) {
            o.bm.words[q.word] = (uint16_t)occ_word(seen1, seen2);
            w = q.word;
        }
    }
    o.gpos[pos] = w;
}

// Object idx goes to sorted slot pos; (ra, rb) is its record (when o.pvs).
__device__ __forceinline__ void rank_write(const RankOut& o, uint32_t key, int32_t idx, uint32_t pos, bool first,
                                           uint32_t count, float4 ra, float4 rb)
{
    o.keys_sorted[pos] = key;
    o.perm[pos] = idx;
    if (o.pvs) {
        o.pvs[pos] = ra;
        o.pvs[(size_t)o.pvs_stride + pos] = rb;
    }
    if (o.slot_of) o.slot_of[idx] = (int32_t)pos;
    if (first) {
        o.table[key] = make_uint2(pos, count);
        if (o.runs) o.runs[atomicAdd(o.nruns, 1u)] = pos;
    }
}

// The same with the record read from item q (bucketed position).
__device__ __forceinline__ void rank_write(const RankOut& o, uint32_t key, int32_t idx, uint32_t pos, bool first,
                                           uint32_t count, uint32_t q)
{
    float4 ra = make_float4(0.f, 0.f, 0.f, 0.f), rb = ra;
    if (o.pvs) {
        ra = o.brec[2 * (size_t)q];
        rb = o.brec[2 * (size_t)q + 1];
    }
    rank_write(o, key, idx, pos, first, count, ra, rb);
}

// Exclusive scan of one value per thread over a 256-thread block (wave64
// shuffles, then the four wave totals through tmp[0..3]).
__device__ __forceinline__ uint32_t block_scan256(uint32_t v, uint32_t* tmp)
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63u) tmp[wave] = x;
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t w = 0; w < wave; ++w) add += tmp[w];
    __syncthreads();  // tmp is reused by the next scan
    return x - v + add;
}

// One block sorts the m items of one bucket (base = its first cell key) by
// the composite c = (key - base) << ibits | index, which orders exactly like
// the stable sort by key (indices are distinct): least significant 8-bit
// digit first, each pass stable -- a digit histogram in LDS, then tiles of
// 256 items in order, each item placed after the items of its digit in
// earlier tiles, earlier waves of its tile (LDS counts) and lower lanes of its
// wave (ballot match on the 8 digit bits).  Every pass reads what other
// threads of this block wrote in the previous one; the block barriers order
// those global accesses.  Returns the pass parity holding the result.
__device__ __forceinline__ int radix_bucket(uint32_t* lds, const uint2* __restrict__ items, uint32_t s0, uint32_t m,
                                            uint32_t base, const RankOut& o)
{
    uint32_t* off = lds;         // [256] running digit offsets
    uint32_t* wc = lds + 256;    // [4][256] digit counts of each wave of the current tile
    uint32_t* tmp = lds + 1280;  // [4] scan scratch
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const int ib = o.ibits;
    const int passes = (o.shift + ib + 7) / 8;
    for (uint32_t t = tid; t < m; t += blockDim.x) {
        const uint2 it = items[s0 + t];
        o.rk[0][s0 + t] = ((uint64_t)(it.x - base) << ib) | (uint64_t)it.y;
        o.rq[0][s0 + t] = t;
    }
    for (uint32_t i = tid; i < 1024u; i += blockDim.x) wc[i] = 0u;
    __syncthreads();
    for (int p = 0; p < passes; ++p) {
        const uint64_t* __restrict__ sk = o.rk[p & 1];
        const uint32_t* __restrict__ sq = o.rq[p & 1];
        uint64_t* __restrict__ dk = o.rk[(p + 1) & 1];
        uint32_t* __restrict__ dq = o.rq[(p + 1) & 1];
        const int sh = 8 * p;
        off[tid] = 0u;
        __syncthreads();
        for (uint32_t t = tid; t < m; t += blockDim.x) atomicAdd(&off[(uint32_t)(sk[s0 + t] >> sh) & 255u], 1u);
        __syncthreads();
        const uint32_t start = block_scan256(off[tid], tmp);
        off[tid] = start;
        __syncthreads();
        for (uint32_t tile = 0; tile < m; tile += blockDim.x) {
            const uint32_t t = tile + tid;
            const bool valid = t < m;
            const uint64_t k = valid ? sk[s0 + t] : 0ull;
            const uint32_t q = valid ? sq[s0 + t] : 0u;
            const uint32_t d = (uint32_t)(k >> sh) & 255u;
            uint64_t same = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const uint32_t bit = (d >> b) & 1u;
                const uint64_t bb = __ballot(bit);
                same &= bit ? bb : ~bb;
            }
            const uint32_t below = (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
            if (valid && below == 0u) wc[wave * 256u + d] = (uint32_t)__popcll(same);
            __syncthreads();
            if (valid) {
                uint32_t pos = off[d] + below;
                for (uint32_t w = 0; w < wave; ++w) pos += wc[w * 256u + d];
                dk[s0 + pos] = k;
                dq[s0 + pos] = q;
            }
            __syncthreads();
            off[tid] += wc[tid] + wc[256u + tid] + wc[512u + tid] + wc[768u + tid];
            wc[tid] = 0u;
            wc[256u + tid] = 0u;
            wc[512u + tid] = 0u;
            wc[768u + tid] = 0u;
            __syncthreads();
        }
    }
    return passes & 1;
}

__device__ __forceinline__ void big_buckets(uint64_t* sv, uint16_t* sq, const uint2* __restrict__ items,
                                             int32_t n, const uint32_t* __restrict__ bstart,
                                             const uint32_t* __restrict__ biglist, const uint32_t* __restrict__ ctrl,
                                             uint32_t big_cap, const RankOut& o, VBlock vb)
{
    const uint32_t nbig = min(ctrl[1], big_cap);
    const uint32_t tid = threadIdx.x;
    for (uint32_t e = vb.id; e < nbig; e += vb.count) {
        const uint32_t b = biglist[e];
        const uint32_t s0 = bstart[b], s1 = bstart[b + 1];
        if (s1 < s0 || s1 > (uint32_t)n) continue;
        const uint32_t m = s1 - s0;
        if (m <= kSortLds) {
            // (key << 32 | index) orders exactly like the stable sort by key
            uint32_t P = 1;
            while (P < m) P <<= 1;
            // sq: the item position (within the bucket) of each entry
            for (uint32_t t = tid; t < P; t += blockDim.x) {
                sv[t] = t < m ? ((uint64_t)items[s0 + t].x << 32) | items[s0 + t].y : ~0ull;
                sq[t] = (uint16_t)t;
            }
            __syncthreads();
            for (uint32_t k = 2; k <= P; k <<= 1) {
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint32_t t = tid; t < P; t += blockDim.x) {
                        const uint32_t u = t ^ j;
                        if (u > t) {
                            const uint64_t a = sv[t], c = sv[u];
                            if ((a > c) == ((t & k) == 0)) {
                                sv[t] = c;
                                sv[u] = a;
                                const uint16_t qt = sq[t];
                                sq[t] = sq[u];
                                sq[u] = qt;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
            for (uint32_t t = tid; t < m; t += blockDim.x) {
                const uint64_t v = sv[t];
                const uint32_t key = (uint32_t)(v >> 32);
                const bool first = t == 0 || (uint32_t)(sv[t - 1] >> 32) != key;
                uint32_t count = 0;
                if (first) {
                    count = 1;
                    while (t + count < m && (uint32_t)(sv[t + count] >> 32) == key) ++count;
                }
                rank_write(o, key, (int32_t)(uint32_t)v, s0 + t, first, count, s0 + sq[t]);
                if (first) sq[t] = (uint16_t)count;  // run length at each run start (sq[t] was read above)
            }
            if (o.gpos) {
                // the group's first run builds its word, run by run, from the
                // sorted keys and the run lengths
                __syncthreads();
                for (uint32_t t = tid; t < m; t += blockDim.x) {
                    const uint32_t key = (uint32_t)(sv[t] >> 32);
                    const bool first = t == 0 || (uint32_t)(sv[t - 1] >> 32) != key;
                    bool owner = false;
                    uint32_t seen1 = 0, seen2 = 0, count = 0;
                    if (first) {
                        count = sq[t];
                        const OccGroup q = occ_group(o.bm, key);
                        owner = t == 0 || (uint32_t)(sv[t - 1] >> 32) < q.gs;
                        for (uint32_t u = t; owner && u < m;) {
                            const uint32_t k2 = (uint32_t)(sv[u] >> 32);
                            if (k2 >= q.ge) break;
                            const uint32_t c = sq[u];
                            seen1 |= 1u << (k2 - q.gs);
                            seen2 |= c > 1u ? 1u << (k2 - q.gs) : 0u;
                            u += max(c, 1u);  // c >= 1 at run starts; the loop always ends
                        }
                    }
                    map_write(o, key, s0 + t, first, count, true, owner, seen1, seen2);
                }
            }
            __syncthreads();  // sv is reused by the next bucket
            continue;
        }
        // larger buckets: radix sort through the global scratch, then every
        // slot writes its object; a run start finds its run's end by binary
        // search over the sorted composites
        const uint32_t base = b << o.shift;
        const int par = radix_bucket(reinterpret_cast<uint32_t*>(sv), items, s0, m, base, o);
        const uint64_t* __restrict__ fk = o.rk[par];
        const uint32_t* __restrict__ fq = o.rq[par];
        const int ib = o.ibits;
        for (uint32_t t = tid; t < m; t += blockDim.x) {
            const uint64_t c = fk[s0 + t];
            const uint32_t kl = (uint32_t)(c >> ib);
            const uint32_t key = base + kl;
            const int32_t idx = (int32_t)(uint32_t)(c & ((1ull << ib) - 1ull));
            const bool first = t == 0 || (uint32_t)(fk[s0 + t - 1] >> ib) != kl;
            uint32_t count = 0;
            if (first) {
                uint32_t lo = t + 1, hi = m;  // first position with a larger key
                while (lo < hi) {
                    const uint32_t mid = lo + ((hi - lo) >> 1);
                    if ((uint32_t)(fk[s0 + mid] >> ib) > kl) hi = mid;
                    else lo = mid + 1;
                }
                count = lo - t;
            }
            rank_write(o, key, idx, s0 + t, first, count, s0 + fq[s0 + t]);
            map_write(o, key, s0 + t, first, count, false, false, 0u, 0u);
        }
        __syncthreads();  // the scratch and LDS are reused by the next bucket
    }
}

// Rank of item (key, idx) among the items get(s0 .. s1-1) of its bucket.
// With OCC, the same pass notes the keys of the item's 8-cell group [gs, ge)
// (seen1, seen2) and whether a smaller key of the group is present (earlier).
template <bool OCC, class Get>
__device__ __forceinline__ void count_rank(Get get, uint32_t s0, uint32_t s1, uint32_t key, int32_t idx,
                                           uint32_t& rank, uint32_t& same, bool& first, uint32_t gs = 0u,
                                           uint32_t ge = 0u, uint32_t* seen1 = nullptr, uint32_t* seen2 = nullptr,
                                           bool* earlier = nullptr)
{
    uint32_t before = 0;  // same key, lower index
    uint32_t m1 = 0, m2 = 0;
    bool e = false;
#pragma unroll 4
    for (uint32_t u = s0; u < s1; ++u) {
        const uint2 it = get(u);
        const uint32_t eq = it.x == key ? 1u : 0u;
        rank += it.x < key ? 1u : 0u;
        same += eq;
        before += eq & ((int32_t)it.y < idx ? 1u : 0u);
        if (OCC) {
            e = e || (it.x >= gs && it.x < key);
            occ_note(it.x, gs, ge, m1, m2);
        }
    }
    rank += before;
    first = before == 0u;
    if (OCC) {
        *seen1 = m1;
        *seen2 = m2;
        *earlier = e;
    }
}

// In-bucket ranking: every thread first ranks its item if the bucket holds at
// most kRankThreshold objects (counting over the bucket), then the blocks
// share out the larger buckets (big_buckets).  The small buckets of a block's
// items lie within kRankThreshold items of the block's range; that window is
// read once into LDS.
__device__ __forceinline__ void rank_body(const RankArgs& R, VBlock vb)
{
    const uint2* __restrict__ items = R.items;
    const int32_t n = R.n;
    const int shift = R.shift;
    const uint32_t* __restrict__ bstart = R.bstart;
    const RankOut& o = R.o;
    __shared__ uint64_t sv[kSortLds];  // the item window first, then the big-bucket sorts
    __shared__ uint16_t sq[kSortLds];
    static_assert(256 + 2 * kRankThreshold <= kSortLds, "item window exceeds the sort buffer");
    const int32_t B = vb.id * blockDim.x;
    const int32_t q = B + threadIdx.x;
    const int32_t w0 = max(B - (int32_t)kRankThreshold, 0);
    const int32_t w1 = min(B + (int32_t)blockDim.x + (int32_t)kRankThreshold, n);
    uint2* win = reinterpret_cast<uint2*>(sv);
    {  // at most 2 items per thread (window <= 256 + 2 * kRankThreshold <= 512); both loads in flight together
        static_assert(2 * kRankThreshold <= 256, "window exceeds two items per thread");
        const int32_t t0 = w0 + (int32_t)threadIdx.x, t1 = t0 + 256;
        uint2 a = make_uint2(0u, 0u), c = a;
        if (t0 < w1) a = items[t0];
        if (t1 < w1) c = items[t1];
        if (t0 < w1) win[t0 - w0] = a;
        if (t1 < w1) win[t1 - w0] = c;
    }
    __syncthreads();
    if (q < n) {
        const uint2 me = win[q - w0];
        const uint32_t key = me.x;
        const int32_t idx = (int32_t)me.y;
        const uint32_t b = key >> shift;
        float4 ra = make_float4(0.f, 0.f, 0.f, 0.f), rb = ra;
        if (o.pvs) {  // issued before the count, used after it
            ra = o.brec[2 * (size_t)q];
            rb = o.brec[2 * (size_t)q + 1];
        }
        const uint32_t s0 = bstart[b], s1 = bstart[b + 1];
        if (s1 >= s0 && s1 <= (uint32_t)n && s1 - s0 <= kRankThreshold) {
            uint32_t rank = 0, same = 0;
            bool first = true;
            const bool in_win = s0 >= (uint32_t)w0 && s1 <= (uint32_t)w1;
            auto get = [&](uint32_t u) { return in_win ? win[u - w0] : items[u]; };
            // (inconsistent offsets, flagged by k_scatter, read the items directly)
            if (o.gpos) {
                // the group's first run (no smaller key of its group in the
                // bucket) builds the group's word from the bucket's items,
                // noted in the ranking pass
                const OccGroup g8 = occ_group(o.bm, key);
                uint32_t seen1 = 0, seen2 = 0;
                bool earlier = false;
                count_rank<true>(get, s0, s1, key, idx, rank, same, first, g8.gs, g8.ge, &seen1, &seen2, &earlier);
                rank_write(o, key, idx, s0 + rank, first, same, ra, rb);
                map_write(o, key, s0 + rank, first, same, true, first && !earlier, seen1, seen2);
            } else {
                count_rank<false>(get, s0, s1, key, idx, rank, same, first);
                rank_write(o, key, idx, s0 + rank, first, same, ra, rb);
            }
        }
    }
    __syncthreads();  // the window's LDS is reused by the big-bucket sort
    big_buckets(sv, sq, items, n, bstart, R.biglist, R.ctrl, R.big_cap, o, vb);
}
__global__ void __launch_bounds__(256) k_rank(RankArgs R) { rank_body(R, hw_block()); }

// Brick occupancy map: the first sorted slot of each 8-cell x-group (cells
// x&~7 .. x|7 of one row) builds the group's 2-bit codes from the sorted keys
// and writes its word; every slot records the word it wrote (or kNoWord) so
// that the next build can clear exactly these words.
__device__ __forceinline__ void occ_body(const OccArgs& a, VBlock vb)
{
    const uint32_t* __restrict__ keys_sorted = a.keys_sorted;
    const int32_t n = a.n;
    const uint2* __restrict__ table = a.table;
    const BrickMap& bm = a.bm;
    const int32_t u = vb.id * blockDim.x + threadIdx.x;
    if (u >= a.cap) return;
    uint32_t w = kNoWord;
    if (u < n) {
        const uint32_t key = keys_sorted[u];
        const uint32_t prev = u > 0 ? keys_sorted[u - 1] : 0u;
        uint32_t x, y, z;
        key_coords(bm, key, x, y, z);
        const uint32_t gs = key - (x & 7u);
        if (u == 0 || prev < gs) {
            const uint32_t ge = gs + min(8u, (uint32_t)bm.cnt[0] - (x & ~7u));
            uint32_t code = 0;
            int32_t v = u;
            uint32_t k2 = key;  // keys_sorted[v]
            while (v < n && k2 < ge) {
                // run length from the next sorted key; the table only for runs of 2+
                const uint32_t nxt = v + 1 < n ? keys_sorted[v + 1] : 0xFFFFFFFFu;
                const bool multi = nxt == k2;
                code |= (multi ? 2u : 1u) << (2u * (k2 - gs));
                if (multi) {
                    v += (int32_t)max(table[k2].y, 2u);
                    k2 = v < n ? keys_sorted[v] : 0xFFFFFFFFu;
                } else {
                    v += 1;
                    k2 = nxt;
                }
            }
            w = brick_word(bm, x, y, z);
            bm.words[w] = (uint16_t)code;
        }
    }
    a.gpos[u] = w;
}
__global__ void __launch_bounds__(256) k_occ(OccArgs a) { occ_body(a, hw_block()); }

// ---- batched launches: one launch per kernel for many simulations ----------
// a[k]: the arguments of simulation k; start: its virtual-block offsets.

__global__ void __launch_bounds__(256) k_key_batch(const KeyArgs* __restrict__ a, const uint32_t* __restrict__ start,
                                                   int32_t nsim, uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { key_body(a[k], vb); });
}
__global__ void __launch_bounds__(kScanThreads) k_scan_batch(const ScanArgs* __restrict__ a,
                                                             const uint32_t* __restrict__ start, int32_t nsim,
                                                             uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock) { scan_body(a[k]); });
}
__global__ void __launch_bounds__(256) k_scatter_batch(const ScatterArgs* __restrict__ a,
                                                       const uint32_t* __restrict__ start, int32_t nsim, uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { scatter_body(a[k], vb); });
}
__global__ void __launch_bounds__(256) k_rank_batch(const RankArgs* __restrict__ a, const uint32_t* __restrict__ start,
                                                    int32_t nsim, uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { rank_body(a[k], vb); });
}
__global__ void __launch_bounds__(256) k_occ_batch(const OccArgs* __restrict__ a, const uint32_t* __restrict__ start,
                                                   int32_t nsim, uint32_t run)
{
    for_virtual_blocks(start, nsim, run, [&](int32_t k, VBlock vb) { occ_body(a[k], vb); });
}

// ------------------------------------------------------------------ host --

static int choose_shift(uint64_t cell_count, int64_t cap)
{
    // aim for ~8 objects per bucket if they filled the whole grid uniformly
    const double target = std::max<double>(1.0, (double)std::max<int64_t>(cap, 1) / 8.0);
    int s = 0;
    while (s < 24 && (double)(cell_count >> s) > target) ++s;
    return s;
}

#define GRID_HIP(call)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return fail(BCS_ERR_HIP, std::string("grid: ") + hipGetErrorString(e_)); \
    } while (0)

// Bucket arrays of a given shift: histogram (zeroed), offsets, scan tile
// status and the big-bucket list.
static int alloc_buckets(DeviceGrid& G, int shift, hipStream_t st)
{
    G.shift = shift;
    G.nb = ((G.g.cell_count - 1) >> G.shift) + 1;
    G.ntiles = (int32_t)((G.nb + 1 + kScanTile - 1) / kScanTile);
    const size_t n = (size_t)std::max(G.cap, 1);
    GRID_HIP(hipMalloc(&G.bcount, (G.nb + 1) * 4));
    GRID_HIP(hipMemsetAsync(G.bcount, 0, (G.nb + 1) * 4, st));
    GRID_HIP(hipMalloc(&G.bstart, (G.nb + 2) * 4));
    GRID_HIP(hipMalloc(&G.status, (size_t)std::max(G.ntiles, 1) * 8));
    G.big_cap = (uint32_t)std::min<uint64_t>(G.nb, (uint64_t)n / kRankThreshold + 1);
    GRID_HIP(hipMalloc(&G.biglist, (size_t)(G.big_cap + 1) * 4));
    return BCS_OK;
}

int grid_alloc(DeviceGrid& G, const GridGeom& geom, int32_t cap, bool with_slot, bool with_runs, bool with_map,
               bool with_copies, hipStream_t st)
{
    G = DeviceGrid{};
    G.g = geom;
    G.cap = cap;
    if (geom.cell_count == 0 || geom.cell_count > 0xFFFFFFFFull)
        return fail(BCS_ERR_GRID_TOO_LARGE, "grid: cell count " + std::to_string(geom.cell_count) +
                                                " outside [1, 2^32) (choose a larger cell size)");
    G.shift = choose_shift(geom.cell_count, cap);
    const size_t n = (size_t)std::max(cap, 1);
    GRID_HIP(hipMalloc(&G.keys, n * 4));
    GRID_HIP(hipMalloc(&G.items, n * 8));
    GRID_HIP(hipMalloc(&G.arrival, n * 4));
    GRID_HIP(hipMalloc(&G.keys_sorted, n * 4));
    if (with_runs) GRID_HIP(hipMalloc(&G.runs, n * 4));
    if (with_copies) {
        GRID_HIP(hipMalloc(&G.brec, n * 32));
        GRID_HIP(hipMalloc(&G.pvs, n * 32));
    }
    GRID_HIP(hipMalloc(&G.perm, n * 4));
    if (with_slot) GRID_HIP(hipMalloc(&G.slot_of, n * 4));
    for (int k = 0; k < 2; ++k) {
        GRID_HIP(hipMalloc(&G.rk[k], n * 8));
        GRID_HIP(hipMalloc(&G.rq[k], n * 4));
    }
    G.ibits = 1;
    while (G.ibits < 31 && (1ull << G.ibits) < (uint64_t)n) ++G.ibits;
    GRID_HIP(hipMemsetAsync(G.keys, 0xFF, n * 4, st));
    hipError_t e = hipMalloc(&G.table, geom.cell_count * 8);
    if (e != hipSuccess)
        return fail(BCS_ERR_OUT_OF_MEMORY, "grid: cell table of " + std::to_string(geom.cell_count * 8) + " bytes");
    GRID_HIP(hipMemsetAsync(G.table, 0, geom.cell_count * 8, st));
    if (with_map) {
        const uint64_t nbx = ((uint64_t)geom.cnt[0] + 7) / 8, nby = ((uint64_t)geom.cnt[1] + 7) / 8,
                       nbz = ((uint64_t)geom.cnt[2] + 7) / 8;
        const uint64_t words = nbx * nby * nbz * 64;
        if (words >= kNoWord) return fail(BCS_ERR_GRID_TOO_LARGE, "grid: occupancy map exceeds 2^32 words");
        e = hipMalloc(&G.bm.words, words * 2);
        if (e != hipSuccess)
            return fail(BCS_ERR_OUT_OF_MEMORY, "grid: occupancy map of " + std::to_string(words * 2) + " bytes");
        GRID_HIP(hipMemsetAsync(G.bm.words, 0, words * 2, st));
        for (int d = 0; d < 3; ++d) G.bm.cnt[d] = geom.cnt[d];
        G.bm.div_x = geom.div_x;
        G.bm.div_y = geom.div_y;
        G.bm.nbx = (uint32_t)nbx;
        G.bm.nby = (uint32_t)nby;
        GRID_HIP(hipMalloc(&G.gpos, n * 4));
        GRID_HIP(hipMemsetAsync(G.gpos, 0xFF, n * 4, st));
    }
    GRID_HIP(hipMalloc(&G.ctrl, 16 * 4));
    GRID_HIP(hipMemsetAsync(G.ctrl, 0, 16 * 4, st));
    return alloc_buckets(G, G.shift, st);
}

int grid_set_shift(DeviceGrid& G, int shift, hipStream_t st)
{
    if (shift == G.shift) return BCS_OK;
    GRID_HIP(hipStreamSynchronize(st));  // no build in flight uses the old arrays
    void* old[] = {G.bcount, G.bstart, G.status, G.biglist};
    for (void* p : old)
        if (p) (void)hipFree(p);
    G.bcount = G.bstart = G.biglist = nullptr;
    G.status = nullptr;
    return alloc_buckets(G, shift, st);
}

int grid_shift_for(const GridGeom& g, const float* pos, int64_t stride, int32_t n)
{
    // Sample up to 2^20 objects (every k-th), form their keys on the host,
    // and for each bucket size (2^s cells) estimate the object-weighted mean
    // bucket size w = sum(c^2)/sum(c) -- the bucket an object typically sorts
    // in -- with the sampling bias of sum(c^2) removed (E[c_s^2] = f^2 c^2 +
    // f(1 - f) c for a fraction f).  The bucket size minimising
    // 18 * buckets + n * w is taken among those with w <= 128 and at most 2^27
    // buckets (the finest allowed one if none has w <= 128).  The weights are
    // fitted to MI355X timings of the whole grid build (scan, histogram and
    // ranking kernels) at 2.4e5, 6.6e6 and 4.2e7 objects; with the cap at 128
    // the 6.6e6-object crowded scene takes 2^7-cell buckets (2^5 under a cap
    // of 32 was 4.6% slower per frame).
    if (n <= 0 || !pos) return choose_shift(g.cell_count, n);
    const int64_t step = std::max<int64_t>(1, (int64_t)n >> 20);
    std::vector<uint32_t> keys;
    keys.reserve((size_t)(n / step + 1));
    for (int64_t i = 0; i < n; i += step) {
        int64_t id[3];
        for (int d = 0; d < 3; ++d) {
            const float q = (pos[d * stride + i] - g.minv[d]) / g.cell[d];
            const float m = q > 0.0f ? q : 0.0f;
            const float c = m < g.extent[d] ? m : g.extent[d];
            id[d] = (c == c && c < 2.0e9f) ? (int64_t)c : 0;
        }
        int64_t k = (id[2] * g.cnt[1] + id[1]) * g.cnt[0] + id[0];
        if (k < 0 || (uint64_t)k >= g.cell_count) k = (int64_t)g.cell_count - 1;
        keys.push_back((uint32_t)k);
    }
    std::sort(keys.begin(), keys.end());
    const double f = (double)keys.size() / (double)n;
    int best = -1, finest = -1;
    double best_cost = 0.0;
    for (int sh = 0; sh <= 24; ++sh) {
        const uint64_t nb = ((g.cell_count - 1) >> sh) + 1;
        if (nb > (1ull << 27)) continue;  // too many buckets
        if (finest < 0) finest = sh;
        double s1 = 0, s2 = 0;
        for (size_t a = 0; a < keys.size();) {
            size_t b = a + 1;
            while (b < keys.size() && (keys[b] >> sh) == (keys[a] >> sh)) ++b;
            const double c = (double)(b - a);
            s1 += c;
            s2 += c * c;
            a = b;
        }
        const double weighted = (s2 - (1.0 - f) * s1) / std::max(f * s1, 1e-30);
        if (weighted > 128.0) break;  // coarser buckets only grow
        const double cost = 18.0 * (double)nb + (double)n * weighted;
        if (best < 0 || cost < best_cost) {
            best = sh;
            best_cost = cost;
        }
    }
    return best >= 0 ? best : std::max(finest, 0);
}

void grid_free(DeviceGrid& G)
{
    void* ptrs[] = {G.keys, G.items, G.keys_sorted, G.runs, G.perm, G.slot_of, G.brec, G.pvs, G.arrival,
                    G.table, G.bm.words, G.gpos, G.bcount, G.bstart, G.status, G.ctrl, G.biglist,
                    G.rk[0], G.rk[1], G.rq[0], G.rq[1]};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    G = DeviceGrid{};
}

int grid_launch_args(DeviceGrid& G, const float* pos, int64_t stride, int32_t n, const float* vel, uint32_t* flags,
                     const int2* order, int32_t norder, GridLaunch* L)
{
    if (n < 0 || n > G.cap) return fail(BCS_ERR_INVALID_ARGUMENT, "grid: object count exceeds capacity");
    // With clear_table, lane gid clears the cell of object gid while, with an
    // order, it writes the key of another object: one lane could then read a
    // key that another has already replaced.  The two are never combined.
    if (order && G.clear_table)
        return fail(BCS_ERR_INVALID_ARGUMENT, "grid: a lane order needs a grid without table clearing");
    // Every slot below the capacity clears its previous cell (slots never
    // built hold the sentinel key): nothing here depends on the previous
    // build's size, so a captured graph of this sequence stays valid.
    const int32_t n_prev = G.cap;
    G.n = n;
    const int blocks = (std::max(std::max(n, n_prev), 1) + 255) / 256;
    const int key_lanes = std::max(std::max(std::max(n, n_prev), order ? norder : 0), 1);
    const int key_rounds = key_lanes >= kKeyRoundsFrom ? kKeyRounds : 1;
    const int key_blocks = (key_lanes + 256 * key_rounds - 1) / (256 * key_rounds);
    *L = GridLaunch{};
    L->key = KeyArgs{G.g, pos, stride, n, n_prev, G.shift, G.clear_table ? 1 : 0, G.keys, G.table, G.bm, G.gpos,
                     G.bcount, G.status, G.ntiles, G.ctrl, flags, G.arrival, order, norder, key_rounds};
    L->key_blocks = (uint32_t)key_blocks;  // (the tile status clear strides over every block)
    L->scan = ScanArgs{G.bcount, G.bstart, (int64_t)G.nb, G.status, G.ntiles, G.ctrl, G.biglist, kRankThreshold,
                       G.big_cap, flags, 1};
    L->scan_blocks = (uint32_t)G.ntiles;
    float4* brec = vel ? G.brec : nullptr;
    L->scatter = ScatterArgs{G.keys, n, G.shift, G.bstart, G.arrival, G.items, flags, pos, vel, stride, brec};
    L->scatter_blocks = (uint32_t)blocks;
    const bool rank_map = grid_map_in_rank(G);
    RankOut o{G.keys_sorted, G.perm, G.slot_of, G.table, G.runs, G.ctrl + 2, brec, brec ? G.pvs : nullptr, G.cap,
              G.bm, rank_map ? G.gpos : nullptr, G.shift, G.ibits, {G.rk[0], G.rk[1]}, {G.rq[0], G.rq[1]}};
    L->rank = RankArgs{G.items, n, G.shift, G.bstart, G.biglist, G.ctrl, G.big_cap, o};
    L->rank_blocks = (uint32_t)blocks;
    L->occ = OccArgs{G.keys_sorted, n, G.cap, G.table, G.bm, G.gpos};
    L->occ_blocks = (G.gpos && !rank_map) ? (uint32_t)((std::max(G.cap, 1) + 255) / 256) : 0u;
    return BCS_OK;
}

int grid_build(DeviceGrid& G, const float* pos, int64_t stride, int32_t n, const float* vel, uint32_t* flags,
               hipStream_t st, const KernelTimer* timer, const int2* order, int32_t norder)
{
    GridLaunch L;
    const int rc = grid_launch_args(G, pos, stride, n, vel, flags, order, norder, &L);
    if (rc) return rc;
    auto tb = [&](int id) { if (timer) timer->begin(id, st); };
    auto te = [&](int id) { if (timer) timer->end(id, st); };
    tb(0);
    hipLaunchKernelGGL(k_key, dim3(L.key_blocks), dim3(256), 0, st, L.key);
    te(0);
    tb(1);
    hipLaunchKernelGGL(k_scan, dim3(L.scan_blocks), dim3(kScanThreads), 0, st, L.scan);
    te(1);
    tb(2);
    hipLaunchKernelGGL(k_scatter, dim3(L.scatter_blocks), dim3(256), 0, st, L.scatter);
    te(2);
    tb(3);
    hipLaunchKernelGGL(k_rank, dim3(L.rank_blocks), dim3(256), 0, st, L.rank);
    te(3);
    tb(4);
    if (L.occ_blocks) hipLaunchKernelGGL(k_occ, dim3(L.occ_blocks), dim3(256), 0, st, L.occ);
    te(4);
    GRID_HIP(hipGetLastError());
    return BCS_OK;
}

void grid_build_batch(const GridBatch& b, hipStream_t st)
{
    if (b.grid[0]) hipLaunchKernelGGL(k_key_batch, dim3(b.grid[0]), dim3(256), 0, st, b.key, b.start[0], b.nsim, b.run);
    if (b.grid[1])
        hipLaunchKernelGGL(k_scan_batch, dim3(b.grid[1]), dim3(kScanThreads), 0, st, b.scan, b.start[1], b.nsim, b.run);
    if (b.grid[2])
        hipLaunchKernelGGL(k_scatter_batch, dim3(b.grid[2]), dim3(256), 0, st, b.scatter, b.start[2], b.nsim, b.run);
    if (b.grid[3]) hipLaunchKernelGGL(k_rank_batch, dim3(b.grid[3]), dim3(256), 0, st, b.rank, b.start[3], b.nsim, b.run);
    if (b.grid[4]) hipLaunchKernelGGL(k_occ_batch, dim3(b.grid[4]), dim3(256), 0, st, b.occ, b.start[4], b.nsim, b.run);
}

void scan_batch(const ScanArgs* a, const uint32_t* start, int32_t nsim, uint32_t grid, uint32_t run, hipStream_t st)
{
    if (grid) hipLaunchKernelGGL(k_scan_batch, dim3(grid), dim3(kScanThreads), 0, st, a, start, nsim, run);
}

// Generic exclusive scan used by the vein-hit segments (same kernel).
int scan_exclusive(const uint32_t* count, uint32_t* out, int64_t m, uint64_t* status, int32_t ntiles, uint32_t* ctrl,
                   uint32_t* flags, hipStream_t st)
{
    // count is only read here (clear_count = 0)
    const ScanArgs a{const_cast<uint32_t*>(count), out, m, status, ntiles, ctrl, nullptr, 0u, 0u, flags, 0};
    hipLaunchKernelGGL(k_scan, dim3(ntiles), dim3(kScanThreads), 0, st, a);
    GRID_HIP(hipGetLastError());
    return BCS_OK;
}

int scan_tiles(int64_t m) { return (int)((m + 1 + kScanTile - 1) / kScanTile); }

}  // namespace bcs
