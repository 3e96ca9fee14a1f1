/*
 * bcs_oracle.c -- CPU oracle for the blood-cell physics step.
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py, always as the checker (or the timed CPU
 * baseline), never as the product.  It restates the reference frame
 * (src/main.cu:175-208) sequentially in plain C, in two precisions:
 *   orc_f32_*  float  -- the parity oracle (bit-comparable to the HIP path,
 *                        which is compiled without FP contraction)
 *   orc_f64_*  double -- the same algorithm in double for drift studies
 * Parity pins (tests/): golden tables from the reference's own headers
 * (oracle/_ref), numpy/scipy stable sorts and brute-force contact sets,
 * closed forms for springs, collisions, integrators and the ray test.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/bcs.h"

/* ------------------------------------------------ shared helpers ------ */

/* SplitMix64 finalizer (Steele, Lea, Flood 2014). */
static inline uint64_t orc_mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Teleport draw k (0 or 1) of blood cell `cell` at frame `step`: a float in
 * (0, 1] with 24 random bits, the range of curand_uniform. */
double orc_uniform(uint64_t seed, uint64_t step, uint64_t cell, uint32_t k)
{
    uint64_t h = orc_mix64(seed ^ orc_mix64(step ^ orc_mix64(cell * 2ull + (uint64_t)k)));
    return (double)((float)((h >> 40) + 1ull) * 0x1p-24f);
}

/* Vein AABB +/- margins, src/meta_factory/vein_factory.hpp:21-86. */
void orc_grid_bounds(const bcs_vein_mesh* mesh, const bcs_physics* ph, float bmin[3], float bmax[3])
{
    int d, i;
    for (d = 0; d < 3; ++d) {
        float lo = mesh->positions[d], hi = mesh->positions[d];
        for (i = 1; i < mesh->vertex_count; ++i) {
            float v = mesh->positions[3 * i + d];
            if (v < lo) lo = v;
            if (v > hi) hi = v;
        }
        float margin = d == 1 ? ph->grid_y_margin : ph->grid_xz_margin;
        bmin[d] = lo - margin;
        bmax[d] = hi + margin;
    }
}

static int orc_cmp_u32(const void* a, const void* b)
{
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* calculateSpringLengths, src/meta_factory/vein_factory.hpp:130-174: for every
 * vertex, the neighbour entries of all incident triangles (duplicates kept),
 * sorted ascending, first 9 kept; length = |p_i - p_nb| in float. */
void orc_vein_neighbors(const bcs_vein_mesh* mesh, int32_t* ids, float* lengths)
{
    int32_t nv = mesh->vertex_count, nt = mesh->triangle_count, i, j;
    int64_t* off = (int64_t*)calloc((size_t)nv + 1, sizeof(int64_t));
    for (i = 0; i < 3 * nt; ++i) off[mesh->indices[i] + 1] += 2;
    for (i = 0; i < nv; ++i) off[i + 1] += off[i];
    uint32_t* lst = (uint32_t*)malloc((size_t)(off[nv] > 0 ? off[nv] : 1) * sizeof(uint32_t));
    int64_t* fill = (int64_t*)calloc((size_t)nv, sizeof(int64_t));
    for (i = 0; i < nt; ++i) {
        uint32_t a = mesh->indices[3 * i], b = mesh->indices[3 * i + 1], c = mesh->indices[3 * i + 2];
        lst[off[a] + fill[a]++] = b; lst[off[a] + fill[a]++] = c;
        lst[off[b] + fill[b]++] = a; lst[off[b] + fill[b]++] = c;
        lst[off[c] + fill[c]++] = a; lst[off[c] + fill[c]++] = b;
    }
    for (j = 0; j < 9; ++j)
        for (i = 0; i < nv; ++i) { ids[j * nv + i] = -1; lengths[j * nv + i] = -1.0f; }
    for (i = 0; i < nv; ++i) {
        int64_t n = off[i + 1] - off[i];
        qsort(lst + off[i], (size_t)n, sizeof(uint32_t), orc_cmp_u32);
        int64_t m = n < 9 ? n : 9;
        for (j = 0; j < m; ++j) {
            uint32_t nb = lst[off[i] + j];
            float dx = mesh->positions[3 * i] - mesh->positions[3 * nb];
            float dy = mesh->positions[3 * i + 1] - mesh->positions[3 * nb + 1];
            float dz = mesh->positions[3 * i + 2] - mesh->positions[3 * nb + 2];
            float len = sqrtf(dx * dx + dy * dy + dz * dz);
            ids[j * nv + i] = (int32_t)nb;
            lengths[j * nv + i] = len < 0 ? -len : len;
        }
    }
    free(off); free(lst); free(fill);
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------ instantiations ------ */

#define R float
#define ORC(name) orc_f32_##name
#include "oracle_impl.h"
#undef R
#undef ORC

#define R double
#define ORC(name) orc_f64_##name
#include "oracle_impl.h"
#undef R
#undef ORC
