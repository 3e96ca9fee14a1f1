/*
 * oracle_impl.h -- precision-generic body of the CPU oracle.  Included twice by
 * bcs_oracle.c with R = float (parity oracle) and R = double (drift rerun).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a sequential restatement of one
 * reference frame (src/main.cu:175-208) used by tests/, the smoke check and
 * the CPU-baseline leg of bench.py as the checker.  The product path never
 * links or calls it.
 *
 * Every arithmetic expression keeps the reference's literal operand order so
 * that an IEEE, non-contracted FP32 evaluation is reproducible bit for bit
 * (the build compiles with -ffp-contract=off).  Reference races are replaced
 * by deterministic rules (snapshot reads of F in the cell-spring phase,
 * particle-index-ordered vein force sums, one teleport draw per cell).
 */

#define V3 ORC(v3)
typedef struct { R x, y, z; } V3;

static inline V3 ORC(mk)(R x, R y, R z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
/* src/utilities/math.cuh:11-34 */
static inline V3 ORC(add)(V3 a, V3 b) { return ORC(mk)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 ORC(sub)(V3 a, V3 b) { return ORC(mk)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 ORC(smul)(R a, V3 v) { return ORC(mk)(a * v.x, a * v.y, a * v.z); }
static inline V3 ORC(sdiv)(V3 v, R a) { return ORC(mk)(v.x / a, v.y / a, v.z / a); }
/* src/utilities/math.cuh:52-88 */
static inline R ORC(dot)(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 ORC(cross)(V3 u, V3 v)
{
    return ORC(mk)(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}
static inline R ORC(sqrt_r)(R x) { return sizeof(R) == 4 ? (R)sqrtf((float)x) : (R)sqrt((double)x); }
static inline R ORC(len)(V3 v) { return ORC(sqrt_r)(ORC(dot)(v, v)); }

/* src/utilities/math.cuh:95-104: v / sqrt(dot(v,v)); NaN -> (0,0,0). */
static inline V3 ORC(normalize)(V3 v, uint32_t* flags)
{
    V3 vn = ORC(sdiv)(v, ORC(sqrt_r)(ORC(dot)(v, v)));
    if (isnan(vn.x) || isnan(vn.y) || isnan(vn.z)) {
        *flags |= BCS_FLAG_NAN_NORMALIZE;
        return ORC(mk)(0, 0, 0);
    }
    return vn;
}

typedef struct ORC(grid) {
    R minv[3], maxv[3], extent[3];   /* bounds and width/height/depth */
    R cell[3];
    int32_t cnt[3];                  /* cellCountX/Y/Z */
    uint64_t cell_count;
    int32_t n;                       /* objects in the last build */
    int32_t cap;
    uint32_t* keys;                  /* per object (input order) */
    uint32_t* keys_sorted;
    int32_t* perm;
    int32_t n_occ;
    uint32_t* occ_keys;
    int32_t* occ_start;
    int32_t* occ_end;
    uint32_t* tmp_k;
    int32_t* tmp_v;
} ORC(grid);

typedef struct ORC(sim) {
    bcs_physics ph;
    int32_t nt;                      /* types */
    int32_t ppc[16], cnt[16], pstart[16], cstart[16], mstart[16], gstart[16];
    int32_t np, nc, nm, ngraph;
    R* graph;                        /* dense spring graph, accumulatedGraphSizes layout */
    R* bs;                           /* collision radii per model vertex */
    R* r0;                           /* initial radii per model vertex */
    R* model;                        /* 3*nm xyz interleaved model vertices */
    /* particles (SoA) */
    R *px, *py, *pz, *vx, *vy, *vz, *fx, *fy, *fz;
    R *cx, *cy, *cz;                 /* blood-cell centres */
    /* vein */
    int32_t nv, ntri, nend;
    R *X, *Y, *Z, *VX, *VY, *VZ, *FX, *FY, *FZ;
    uint32_t* idx;
    int32_t* nbr;                    /* 9*nv slot-major */
    R* nbl;
    R *tcx, *tcy, *tcz;              /* triangle centres */
    R* endc; R* endr;
    ORC(grid) pg, tg;
    uint64_t step;
    uint32_t flags;
    /* scratch for vein hits (particle order) */
    int32_t* hit_tri;
    R* hit_w;                        /* 9 per particle: bary_k * ds */
} ORC(sim);

/* ---------------------------------------------------------------- grid --- */

/* (int) conversion with saturation and NaN -> 0, the semantics of the
 * reference's static_cast<int>(float) on its target hardware. */
static inline int32_t ORC(sat_i32)(R q)
{
    if (q != q) return 0;
    if (q >= (R)2147483647.0) return 2147483647;
    if (q <= (R)-2147483648.0) return (-2147483647 - 1);
    return (int32_t)q;
}
static inline uint32_t ORC(sat_u32)(R q)
{
    if (q != q || q <= (R)0) return 0u;
    if (q >= (R)4294967295.0) return 4294967295u;
    return (uint32_t)q;
}

/* calculateIdForCell, src/grids/uniform_grid.cu:24-36, with the max/min
 * macros of :20-21 (upper clamp in world units).  Keys outside
 * [0, cellCount) are clamped to cellCount-1 and flagged (defined behaviour
 * where the reference writes out of bounds). */
static uint32_t ORC(cell_key)(const ORC(grid)* g, R x, R y, R z, uint32_t* flags)
{
    R p[3]; p[0] = x; p[1] = y; p[2] = z;
    int64_t id[3];
    int d;
    if (x < g->minv[0] || x > g->maxv[0] || y < g->minv[1] || y > g->maxv[1] ||
        z < g->minv[2] || z > g->maxv[2])
        *flags |= BCS_FLAG_OUT_OF_GRID;
    for (d = 0; d < 3; ++d) {
        R q = (p[d] - g->minv[d]) / g->cell[d];
        R m = ((R)0 > q) ? (R)0 : q;                 /* max(0, q) */
        R c = (g->extent[d] > m) ? m : g->extent[d];  /* min(ext, m) */
        id[d] = (int64_t)ORC(sat_i32)(c);
    }
    int64_t key = id[2] * (int64_t)g->cnt[0] * (int64_t)g->cnt[1] + id[1] * (int64_t)g->cnt[0] + id[0];
    if (key < 0 || (uint64_t)key >= g->cell_count) {
        *flags |= BCS_FLAG_OUT_OF_GRID;
        key = (int64_t)g->cell_count - 1;
    }
    return (uint32_t)key;
}

static void ORC(grid_init)(ORC(grid)* g, const float bmin[3], const float bmax[3], const float cell[3], int32_t cap)
{
    int d;
    memset(g, 0, sizeof(*g));
    for (d = 0; d < 3; ++d) {
        g->minv[d] = (R)bmin[d];
        g->maxv[d] = (R)bmax[d];
        g->extent[d] = g->maxv[d] - g->minv[d];
        g->cell[d] = (R)cell[d];
        /* src/grids/uniform_grid.cu:86-88: static_cast<int>(width / cellWidth) */
        g->cnt[d] = ORC(sat_i32)(g->extent[d] / g->cell[d]);
    }
    g->cell_count = (uint64_t)g->cnt[0] * (uint64_t)g->cnt[1] * (uint64_t)g->cnt[2];
    g->cap = cap;
    size_t n = (size_t)(cap > 0 ? cap : 1);
    g->keys = (uint32_t*)calloc(n, 4);
    g->keys_sorted = (uint32_t*)calloc(n, 4);
    g->perm = (int32_t*)calloc(n, 4);
    g->occ_keys = (uint32_t*)calloc(n, 4);
    g->occ_start = (int32_t*)calloc(n, 4);
    g->occ_end = (int32_t*)calloc(n, 4);
    g->tmp_k = (uint32_t*)calloc(n, 4);
    g->tmp_v = (int32_t*)calloc(n, 4);
}

static void ORC(grid_free)(ORC(grid)* g)
{
    free(g->keys); free(g->keys_sorted); free(g->perm); free(g->occ_keys);
    free(g->occ_start); free(g->occ_end); free(g->tmp_k); free(g->tmp_v);
}

/* UniformGrid::calculateGrid, src/grids/uniform_grid.cu:129-155: key per
 * object, thrust::stable_sort_by_key (restated as an LSD radix sort, which is
 * stable), then start/end of every occupied cell.  The reference's last-slot
 * quirk (:76-79 writes cellStarts instead of cellEnds) and its never-cleared
 * tables are replaced by exact ranges of occupied cells. */
static void ORC(grid_build)(ORC(grid)* g, int32_t n, const R* x, const R* y, const R* z, uint32_t* flags)
{
    int32_t i;
    g->n = n;
    uint32_t fl = 0;
#pragma omp parallel for reduction(| : fl) schedule(static)
    for (i = 0; i < n; ++i) g->keys[i] = ORC(cell_key)(g, x[i], y[i], z[i], &fl);
    *flags |= fl;
    uint32_t *ka = g->keys_sorted, *kb = g->tmp_k;
    int32_t *va = g->perm, *vb = g->tmp_v;
    for (i = 0; i < n; ++i) { ka[i] = g->keys[i]; va[i] = i; }
    int pass;
    for (pass = 0; pass < 4; ++pass) {
        int shift = pass * 8;
        size_t cntb[257];
        memset(cntb, 0, sizeof(cntb));
        for (i = 0; i < n; ++i) cntb[((ka[i] >> shift) & 255u) + 1]++;
        int b;
        for (b = 0; b < 256; ++b) cntb[b + 1] += cntb[b];
        for (i = 0; i < n; ++i) {
            size_t pos = cntb[(ka[i] >> shift) & 255u]++;
            kb[pos] = ka[i];
            vb[pos] = va[i];
        }
        uint32_t* tk = ka; ka = kb; kb = tk;
        int32_t* tv = va; va = vb; vb = tv;
    }
    /* after 4 passes the data is back in keys_sorted / perm */
    int32_t no = 0;
    for (i = 0; i < n; ++i) {
        if (i == 0 || g->keys_sorted[i] != g->keys_sorted[i - 1]) {
            g->occ_keys[no] = g->keys_sorted[i];
            g->occ_start[no] = i;
            no++;
        }
        g->occ_end[no - 1] = i;
    }
    g->n_occ = no;
}

/* Occupied range of a cell key (binary search over occupied keys).  Returns 0
 * for an empty cell. */
static int ORC(grid_range)(const ORC(grid)* g, int64_t key, int32_t* a, int32_t* b)
{
    if (key < 0 || (uint64_t)key >= g->cell_count) return 0;
    int32_t lo = 0, hi = g->n_occ - 1;
    uint32_t k = (uint32_t)key;
    while (lo <= hi) {
        int32_t mid = lo + ((hi - lo) >> 1);
        uint32_t mk = g->occ_keys[mid];
        if (mk == k) { *a = g->occ_start[mid]; *b = g->occ_end[mid]; return 1; }
        if (mk < k) lo = mid + 1; else hi = mid - 1;
    }
    return 0;
}

/* Face-clipped neighbour range of one axis: src/simulation/particle_collisions.cuh:270-413. */
static inline void ORC(axis_range_i)(int32_t id, int32_t cnt, int* lo, int* hi)
{
    if (id < 1) { *lo = 0; *hi = 1; }
    else if (id > cnt - 2) { *lo = -1; *hi = 0; }
    else { *lo = -1; *hi = 1; }
}
/* Same with unsigned ids: src/simulation/vein_collisions.cu:82-232. */
static inline void ORC(axis_range_u)(uint32_t id, int32_t cnt, int* lo, int* hi)
{
    if (id < 1u) { *lo = 0; *hi = 1; }
    else if (id > (uint32_t)(cnt - 2)) { *lo = -1; *hi = 0; }
    else { *lo = -1; *hi = 1; }
}

/* --------------------------------------------------------------- tables --- */

static int ORC(type_of_particle)(const ORC(sim)* s, int32_t i)
{
    int t;
    for (t = 0; t < s->nt; ++t)
        if (i >= s->pstart[t] && i < s->pstart[t] + s->cnt[t] * s->ppc[t]) return t;
    return s->nt - 1;
}

/* C remainder (truncating), as the reference's `%` on ints. */
static inline int32_t ORC(cmod)(int32_t a, int32_t m) { return a % m; }

/* bloodCellmodelStart + (id - particlesStart) % particlesInBloodCell with the
 * launch type's parameters (src/simulation/particle_collisions.cuh:124,76);
 * indices outside the model table are clamped and flagged. */
static inline int32_t ORC(model_index)(const ORC(sim)* s, int t, int32_t id, uint32_t* flags)
{
    int32_t m = s->mstart[t] + ORC(cmod)(id - s->pstart[t], s->ppc[t]);
    if (m < 0) { *flags |= BCS_FLAG_RADIUS_CLAMPED; m = 0; }
    if (m >= s->nm) { *flags |= BCS_FLAG_RADIUS_CLAMPED; m = s->nm - 1; }
    return m;
}

ORC(sim)* ORC(create)(const bcs_physics* ph, const bcs_cell_type* types, int32_t nt, const bcs_vein_mesh* mesh)
{
    if (nt < 1 || nt > 16 || !ph || !types || !mesh) return NULL;
    ORC(sim)* s = (ORC(sim)*)calloc(1, sizeof(ORC(sim)));
    s->ph = *ph;
    s->nt = nt;
    int t, a, b;
    int32_t np = 0, nc = 0, nm = 0, ng = 0;
    /* src/meta_factory/blood_cell_factory.hpp:172-282 */
    for (t = 0; t < nt; ++t) {
        s->ppc[t] = types[t].particles_in_cell;
        s->cnt[t] = types[t].count;
        s->pstart[t] = np; s->cstart[t] = nc; s->mstart[t] = nm; s->gstart[t] = ng;
        np += types[t].count * types[t].particles_in_cell;
        nc += types[t].count;
        nm += types[t].particles_in_cell;
        ng += types[t].particles_in_cell * types[t].particles_in_cell;
    }
    s->np = np; s->nc = nc; s->nm = nm; s->ngraph = ng;
    s->graph = (R*)calloc((size_t)(ng > 0 ? ng : 1), sizeof(R));
    s->bs = (R*)calloc((size_t)nm, sizeof(R));
    s->r0 = (R*)calloc((size_t)nm, sizeof(R));
    s->model = (R*)calloc((size_t)nm * 3, sizeof(R));
    /* springGraph, src/meta_factory/blood_cell_factory.hpp:292-328 */
    for (t = 0; t < nt; ++t) {
        int32_t p = types[t].particles_in_cell;
        int k;
        for (k = 0; k < types[t].spring_count; ++k) {
            int32_t st = types[t].spring_start[k], en = types[t].spring_end[k];
            float L = types[t].spring_length[k] * 1.0f; /* springLengthCoefficient */
            s->graph[s->gstart[t] + st * p + en] = (R)L;
            s->graph[s->gstart[t] + en * p + st] = (R)L;
        }
    }
    /* generateBoundingSpheres, src/simulation/simulation_controller.cu:93-153.
     * Radii are computed in float (double for the pairwise distances, as the
     * reference's pow/sqrt) and stored as float for both precisions. */
    for (t = 0; t < nt; ++t) {
        int32_t p = types[t].particles_in_cell, m0 = s->mstart[t];
        const float* mv = types[t].vertices;
        float bsf[256];
        for (a = 0; a < p; ++a) bsf[a] = FLT_MAX;
        float cxf = 0.f, cyf = 0.f, czf = 0.f;
        for (a = 0; a < p; ++a) {
            for (b = 0; b < p; ++b) {
                if (a != b) {
                    float dx = mv[3 * a] - mv[3 * b], dy = mv[3 * a + 1] - mv[3 * b + 1], dz = mv[3 * a + 2] - mv[3 * b + 2];
                    float length = (float)(sqrt(pow((double)dx, 2) + pow((double)dy, 2) + pow((double)dz, 2)) /
                                           (double)(2 * ph->bounding_spheres_coeff));
                    if (length < bsf[a]) bsf[a] = length;
                }
            }
            cxf = cxf + mv[3 * a]; cyf = cyf + mv[3 * a + 1]; czf = czf + mv[3 * a + 2];
        }
        cxf = cxf / (float)p; cyf = cyf / (float)p; czf = czf / (float)p;
        for (a = 0; a < p; ++a) {
            float dx = mv[3 * a] - cxf, dy = mv[3 * a + 1] - cyf, dz = mv[3 * a + 2] - czf;
            float r0 = sqrtf(dx * dx + dy * dy + dz * dz);
            s->r0[m0 + a] = (R)r0;
            s->bs[m0 + a] = (R)(types[t].contact_radius > 0.f ? types[t].contact_radius : bsf[a]);
            s->model[3 * (m0 + a)] = (R)mv[3 * a];
            s->model[3 * (m0 + a) + 1] = (R)mv[3 * a + 1];
            s->model[3 * (m0 + a) + 2] = (R)mv[3 * a + 2];
        }
    }
    size_t npa = (size_t)(np > 0 ? np : 1);
    s->px = (R*)calloc(npa, sizeof(R)); s->py = (R*)calloc(npa, sizeof(R)); s->pz = (R*)calloc(npa, sizeof(R));
    s->vx = (R*)calloc(npa, sizeof(R)); s->vy = (R*)calloc(npa, sizeof(R)); s->vz = (R*)calloc(npa, sizeof(R));
    s->fx = (R*)calloc(npa, sizeof(R)); s->fy = (R*)calloc(npa, sizeof(R)); s->fz = (R*)calloc(npa, sizeof(R));
    size_t nca = (size_t)(nc > 0 ? nc : 1);
    s->cx = (R*)calloc(nca, sizeof(R)); s->cy = (R*)calloc(nca, sizeof(R)); s->cz = (R*)calloc(nca, sizeof(R));
    s->hit_tri = (int32_t*)calloc(npa, sizeof(int32_t));
    s->hit_w = (R*)calloc(npa * 9, sizeof(R));

    /* vein: src/objects/vein_triangles.cu:37-69 */
    int32_t nv = mesh->vertex_count, ntri = mesh->triangle_count, i;
    s->nv = nv; s->ntri = ntri;
    size_t nva = (size_t)(nv > 0 ? nv : 1), nta = (size_t)(ntri > 0 ? ntri : 1);
    s->X = (R*)calloc(nva, sizeof(R)); s->Y = (R*)calloc(nva, sizeof(R)); s->Z = (R*)calloc(nva, sizeof(R));
    s->VX = (R*)calloc(nva, sizeof(R)); s->VY = (R*)calloc(nva, sizeof(R)); s->VZ = (R*)calloc(nva, sizeof(R));
    s->FX = (R*)calloc(nva, sizeof(R)); s->FY = (R*)calloc(nva, sizeof(R)); s->FZ = (R*)calloc(nva, sizeof(R));
    for (i = 0; i < nv; ++i) {
        s->X[i] = (R)mesh->positions[3 * i]; s->Y[i] = (R)mesh->positions[3 * i + 1]; s->Z[i] = (R)mesh->positions[3 * i + 2];
    }
    s->idx = (uint32_t*)malloc(nta * 3 * sizeof(uint32_t));
    memcpy(s->idx, mesh->indices, (size_t)ntri * 3 * sizeof(uint32_t));
    s->nend = mesh->ending_count;
    s->endc = (R*)calloc((size_t)(s->nend > 0 ? s->nend : 1) * 3, sizeof(R));
    s->endr = (R*)calloc((size_t)(s->nend > 0 ? s->nend : 1), sizeof(R));
    for (i = 0; i < s->nend; ++i) {
        s->endc[3 * i] = (R)mesh->ending_centers[3 * i];
        s->endc[3 * i + 1] = (R)mesh->ending_centers[3 * i + 1];
        s->endc[3 * i + 2] = (R)mesh->ending_centers[3 * i + 2];
        s->endr[i] = (R)mesh->ending_radii[i];
    }
    /* vein neighbour tables (float, both precisions): calculateSpringLengths,
     * src/meta_factory/vein_factory.hpp:130-174 */
    s->nbr = (int32_t*)malloc(nva * 9 * sizeof(int32_t));
    s->nbl = (R*)malloc(nva * 9 * sizeof(R));
    {
        int32_t* tmpi = (int32_t*)malloc(nva * 9 * sizeof(int32_t));
        float* tmpl = (float*)malloc(nva * 9 * sizeof(float));
        orc_vein_neighbors(mesh, tmpi, tmpl);
        for (i = 0; i < 9 * nv; ++i) { s->nbr[i] = tmpi[i]; s->nbl[i] = (R)tmpl[i]; }
        free(tmpi); free(tmpl);
    }
    /* triangle centres: calculateCentersKernel, src/objects/vein_triangles.cu:14-27 */
    s->tcx = (R*)calloc(nta, sizeof(R)); s->tcy = (R*)calloc(nta, sizeof(R)); s->tcz = (R*)calloc(nta, sizeof(R));
    for (i = 0; i < ntri; ++i) {
        uint32_t a0 = s->idx[3 * i], a1 = s->idx[3 * i + 1], a2 = s->idx[3 * i + 2];
        s->tcx[i] = (s->X[a0] + s->X[a1] + s->X[a2]) / (R)3;
        s->tcy[i] = (s->Y[a0] + s->Y[a1] + s->Y[a2]) / (R)3;
        s->tcz[i] = (s->Z[a0] + s->Z[a1] + s->Z[a2]) / (R)3;
    }
    float bmin[3], bmax[3];
    orc_grid_bounds(mesh, ph, bmin, bmax);
    ORC(grid_init)(&s->pg, bmin, bmax, ph->particle_cell, np);
    ORC(grid_init)(&s->tg, bmin, bmax, ph->triangle_cell, ntri);
    s->step = 0;
    s->flags = 0;
    return s;
}

void ORC(destroy)(ORC(sim)* s)
{
    if (!s) return;
    free(s->graph); free(s->bs); free(s->r0); free(s->model);
    free(s->px); free(s->py); free(s->pz); free(s->vx); free(s->vy); free(s->vz);
    free(s->fx); free(s->fy); free(s->fz); free(s->cx); free(s->cy); free(s->cz);
    free(s->hit_tri); free(s->hit_w);
    free(s->X); free(s->Y); free(s->Z); free(s->VX); free(s->VY); free(s->VZ);
    free(s->FX); free(s->FY); free(s->FZ); free(s->idx); free(s->nbr); free(s->nbl);
    free(s->tcx); free(s->tcy); free(s->tcz); free(s->endc); free(s->endr);
    ORC(grid_free)(&s->pg); ORC(grid_free)(&s->tg);
    free(s);
}

int32_t ORC(particle_count)(const ORC(sim)* s) { return s->np; }

void ORC(tables)(const ORC(sim)* s, R* bs, R* r0, R* graph)
{
    if (bs) memcpy(bs, s->bs, (size_t)s->nm * sizeof(R));
    if (r0) memcpy(r0, s->r0, (size_t)s->nm * sizeof(R));
    if (graph) memcpy(graph, s->graph, (size_t)s->ngraph * sizeof(R));
}

void ORC(set_particles)(ORC(sim)* s, const R* p, const R* v, const R* f)
{
    size_t n = (size_t)s->np;
    if (p) { memcpy(s->px, p, n * sizeof(R)); memcpy(s->py, p + n, n * sizeof(R)); memcpy(s->pz, p + 2 * n, n * sizeof(R)); }
    if (v) { memcpy(s->vx, v, n * sizeof(R)); memcpy(s->vy, v + n, n * sizeof(R)); memcpy(s->vz, v + 2 * n, n * sizeof(R)); }
    if (f) { memcpy(s->fx, f, n * sizeof(R)); memcpy(s->fy, f + n, n * sizeof(R)); memcpy(s->fz, f + 2 * n, n * sizeof(R)); }
}

void ORC(get_particles)(const ORC(sim)* s, R* p, R* v, R* f)
{
    size_t n = (size_t)s->np;
    if (p) { memcpy(p, s->px, n * sizeof(R)); memcpy(p + n, s->py, n * sizeof(R)); memcpy(p + 2 * n, s->pz, n * sizeof(R)); }
    if (v) { memcpy(v, s->vx, n * sizeof(R)); memcpy(v + n, s->vy, n * sizeof(R)); memcpy(v + 2 * n, s->vz, n * sizeof(R)); }
    if (f) { memcpy(f, s->fx, n * sizeof(R)); memcpy(f + n, s->fy, n * sizeof(R)); memcpy(f + 2 * n, s->fz, n * sizeof(R)); }
}

void ORC(set_vein)(ORC(sim)* s, const R* p, const R* v, const R* f)
{
    size_t n = (size_t)s->nv;
    if (p) { memcpy(s->X, p, n * sizeof(R)); memcpy(s->Y, p + n, n * sizeof(R)); memcpy(s->Z, p + 2 * n, n * sizeof(R)); }
    if (v) { memcpy(s->VX, v, n * sizeof(R)); memcpy(s->VY, v + n, n * sizeof(R)); memcpy(s->VZ, v + 2 * n, n * sizeof(R)); }
    if (f) { memcpy(s->FX, f, n * sizeof(R)); memcpy(s->FY, f + n, n * sizeof(R)); memcpy(s->FZ, f + 2 * n, n * sizeof(R)); }
}

void ORC(get_vein)(const ORC(sim)* s, R* p, R* v, R* f)
{
    size_t n = (size_t)s->nv;
    if (p) { memcpy(p, s->X, n * sizeof(R)); memcpy(p + n, s->Y, n * sizeof(R)); memcpy(p + 2 * n, s->Z, n * sizeof(R)); }
    if (v) { memcpy(v, s->VX, n * sizeof(R)); memcpy(v + n, s->VY, n * sizeof(R)); memcpy(v + 2 * n, s->VZ, n * sizeof(R)); }
    if (f) { memcpy(f, s->FX, n * sizeof(R)); memcpy(f + n, s->FY, n * sizeof(R)); memcpy(f + 2 * n, s->FZ, n * sizeof(R)); }
}

void ORC(get_centers)(const ORC(sim)* s, R* c)
{
    size_t n = (size_t)s->nc;
    memcpy(c, s->cx, n * sizeof(R)); memcpy(c + n, s->cy, n * sizeof(R)); memcpy(c + 2 * n, s->cz, n * sizeof(R));
}

void ORC(get_triangle_centers)(const ORC(sim)* s, R* c)
{
    size_t n = (size_t)s->ntri;
    memcpy(c, s->tcx, n * sizeof(R)); memcpy(c + n, s->tcy, n * sizeof(R)); memcpy(c + 2 * n, s->tcz, n * sizeof(R));
}

void ORC(set_step_index)(ORC(sim)* s, uint64_t step) { s->step = step; }
uint64_t ORC(get_step_index)(const ORC(sim)* s) { return s->step; }
uint32_t ORC(flags)(ORC(sim)* s, int clear) { uint32_t f = s->flags; if (clear) s->flags = 0; return f; }

void ORC(calculate_grid)(ORC(sim)* s, int which)
{
    if (which == BCS_GRID_PARTICLES) ORC(grid_build)(&s->pg, s->np, s->px, s->py, s->pz, &s->flags);
    else ORC(grid_build)(&s->tg, s->ntri, s->tcx, s->tcy, s->tcz, &s->flags);
}

/* Recompute triangle centres from the current vein positions (used only when
 * dynamic_triangle_grid is set; the reference does this per frame only in its
 * multi-GPU mode, src/main.cu:167-173). */
void ORC(update_triangle_centers)(ORC(sim)* s)
{
    int32_t i;
    for (i = 0; i < s->ntri; ++i) {
        uint32_t a0 = s->idx[3 * i], a1 = s->idx[3 * i + 1], a2 = s->idx[3 * i + 2];
        s->tcx[i] = (s->X[a0] + s->X[a1] + s->X[a2]) / (R)3;
        s->tcy[i] = (s->Y[a0] + s->Y[a1] + s->Y[a2]) / (R)3;
        s->tcz[i] = (s->Z[a0] + s->Z[a1] + s->Z[a2]) / (R)3;
    }
}

int32_t ORC(get_grid)(const ORC(sim)* s, int which, uint32_t* keys_sorted, int32_t* perm,
                      uint32_t* occ_keys, int32_t* occ_start, int32_t* occ_end)
{
    const ORC(grid)* g = which == BCS_GRID_PARTICLES ? &s->pg : &s->tg;
    size_t n = (size_t)g->n;
    if (keys_sorted) memcpy(keys_sorted, g->keys_sorted, n * 4);
    if (perm) memcpy(perm, g->perm, n * 4);
    if (occ_keys) memcpy(occ_keys, g->occ_keys, (size_t)g->n_occ * 4);
    if (occ_start) memcpy(occ_start, g->occ_start, (size_t)g->n_occ * 4);
    if (occ_end) memcpy(occ_end, g->occ_end, (size_t)g->n_occ * 4);
    return g->n_occ;
}

/* ------------------------------------------------------------ physics --- */

/* Vein springs: gatherForcesKernel, src/objects/vein_triangles.cu:126-154,
 * springMassForceWithDampingForVein, src/simulation/physics.cuh:38-41. */
static void ORC(vein_springs)(ORC(sim)* s)
{
    const R vk = (R)s->ph.vein_k_sniff, vd = (R)s->ph.vein_d_fact;
    int32_t id, nv = s->nv;
    uint32_t fl = 0;
#pragma omp parallel for reduction(| : fl) schedule(static)
    for (id = 0; id < nv; ++id) {
        V3 P = ORC(mk)(s->X[id], s->Y[id], s->Z[id]);
        V3 Vv = ORC(mk)(s->VX[id], s->VY[id], s->VZ[id]);
        V3 force = ORC(mk)(0, 0, 0);
        int sl;
        for (sl = 0; sl < 9; ++sl) {
            int32_t nb = s->nbr[sl * nv + id];
            if (nb != -1) {
                R L = s->nbl[sl * nv + id];
                V3 NP = ORC(mk)(s->X[nb], s->Y[nb], s->Z[nb]);
                V3 NV = ORC(mk)(s->VX[nb], s->VY[nb], s->VZ[nb]);
                R sf = (ORC(len)(ORC(sub)(P, NP)) - L) * vk +
                       ORC(dot)(ORC(normalize)(ORC(sub)(P, NP), &fl), ORC(sub)(Vv, NV)) * vd;
                force = ORC(add)(force, ORC(smul)(sf, ORC(normalize)(ORC(sub)(NP, P), &fl)));
            }
        }
        s->FX[id] = s->FX[id] + force.x;
        s->FY[id] = s->FY[id] + force.y;
        s->FZ[id] = s->FZ[id] + force.z;
    }
    s->flags |= fl;
}

/* calculateBloodCellsCenters, src/objects/blood_cells.cu:44-61. */
static void ORC(centers)(ORC(sim)* s)
{
    int t;
    for (t = 0; t < s->nt; ++t) {
        int32_t c, p = s->ppc[t];
#pragma omp parallel for schedule(static)
        for (c = 0; c < s->cnt[t]; ++c) {
            int32_t first = s->pstart[t] + c * p, a;
            V3 center = ORC(mk)(0, 0, 0);
            for (a = 0; a < p; ++a) center = ORC(add)(center, ORC(mk)(s->px[first + a], s->py[first + a], s->pz[first + a]));
            center = ORC(sdiv)(center, (R)p);
            s->cx[s->cstart[t] + c] = center.x;
            s->cy[s->cstart[t] + c] = center.y;
            s->cz[s->cstart[t] + c] = center.z;
        }
    }
}

/* Cell springs + environment: gatherForcesKernel, src/objects/blood_cells.cu:66-120,
 * calculateParticlesSpringForceComponent (Heun branch), physics.cuh:53-78,
 * accumulateEnvironmentForcesForParticles, physics.cuh:102-120.  F of
 * neighbours is read from the pre-phase snapshot (the reference reads it while
 * neighbours overwrite it). */
/* springMassForceWithDampingForParticle, src/simulation/physics.cuh:24-27. */
static inline R ORC(spring_particle)(const ORC(sim)* s, V3 dP, V3 dv, R L, uint32_t* fl)
{
    return (ORC(len)(dP) - L) * (R)s->ph.particle_k_sniff +
           ORC(dot)(ORC(normalize)(dP, fl), dv) * (R)s->ph.particle_d_fact;
}

static void ORC(cell_springs)(ORC(sim)* s)
{
    const R k = (R)s->ph.particle_k_sniff, d = (R)s->ph.particle_d_fact, dt = (R)s->ph.dt;
    const R vdamp = (R)s->ph.viscous_damping, brake = (R)s->ph.max_cell_size_factor_before_brake,
            bint = (R)s->ph.big_particle_braking_intensity;
    const V3 G = ORC(mk)((R)s->ph.gravity[0], (R)s->ph.gravity[1], (R)s->ph.gravity[2]);
    size_t n = (size_t)s->np;
    R* sfx = (R*)malloc((n ? n : 1) * sizeof(R));
    R* sfy = (R*)malloc((n ? n : 1) * sizeof(R));
    R* sfz = (R*)malloc((n ? n : 1) * sizeof(R));
    memcpy(sfx, s->fx, n * sizeof(R)); memcpy(sfy, s->fy, n * sizeof(R)); memcpy(sfz, s->fz, n * sizeof(R));
    uint32_t fl = 0;
    int t;
    for (t = 0; t < s->nt; ++t) {
        int32_t p = s->ppc[t], total = s->cnt[t] * p, it;
#pragma omp parallel for reduction(| : fl) schedule(static)
        for (it = 0; it < total; ++it) {
            int32_t i = s->pstart[t] + it, a = it % p, cell = s->cstart[t] + it / p, b;
            V3 position = ORC(mk)(s->px[i], s->py[i], s->pz[i]);
            V3 velocity = ORC(mk)(s->vx[i], s->vy[i], s->vz[i]);
            V3 initialForce = ORC(mk)(sfx[i], sfy[i], sfz[i]);
            V3 newForce = ORC(mk)(0, 0, 0);
            V3 radius = ORC(sub)(position, ORC(mk)(s->cx[cell], s->cy[cell], s->cz[cell]));
            R initialRadius = s->r0[s->mstart[t] + a];
            for (b = 0; b < p; ++b) {
                R L = s->graph[s->gstart[t] + b * p + a];
                if (L != (R)0) {
                    int32_t nb = i - a + b;
                    V3 dP = ORC(sub)(position, ORC(mk)(s->px[nb], s->py[nb], s->pz[nb]));
                    V3 dv = ORC(sub)(velocity, ORC(mk)(s->vx[nb], s->vy[nb], s->vz[nb]));
                    V3 f2 = ORC(mk)(sfx[nb], sfy[nb], sfz[nb]);
                    V3 shift = ORC(normalize)(ORC(smul)((R)-1, dP), &fl);
                    if (s->ph.particle_rk4) {
                        /* USE_RUNGE_KUTTA_FOR_PARTICLE branch of
                         * calculateParticlesSpringForceComponent (physics.cuh:57-71):
                         * returns 2 k2_v + 2 k3_v + k4_v (k1_v is not part of it) */
                        V3 k1x = ORC(smul)(dt, dv);
                        V3 k1v = ORC(smul)(dt, initialForce);
                        V3 k2x = ORC(smul)(dt, ORC(add)(dv, ORC(sdiv)(k1x, (R)2)));
                        V3 k2v = ORC(smul)(dt * ORC(spring_particle)(s, ORC(add)(dP, ORC(sdiv)(k1x, (R)2)),
                                                                     ORC(add)(dv, ORC(sdiv)(k1v, (R)2)), L, &fl), shift);
                        V3 k3x = ORC(smul)(dt, ORC(add)(dv, ORC(sdiv)(k2x, (R)2)));
                        V3 k3v = ORC(smul)(dt * ORC(spring_particle)(s, ORC(add)(dP, ORC(sdiv)(k2x, (R)2)),
                                                                     ORC(add)(dv, ORC(sdiv)(k2v, (R)2)), L, &fl), shift);
                        V3 k4v = ORC(smul)(dt * ORC(spring_particle)(s, ORC(add)(dP, k3x), ORC(add)(dv, k3v), L, &fl),
                                           shift);
                        newForce = ORC(add)(newForce, ORC(add)(ORC(add)(ORC(smul)((R)2, k2v), ORC(smul)((R)2, k3v)), k4v));
                        (void)f2;
                    } else {
                        V3 dvv = ORC(add)(dv, ORC(smul)(dt, ORC(sub)(initialForce, f2)));
                        R mag = (ORC(len)(dP) - L) * k + ORC(dot)(ORC(normalize)(dP, &fl), dvv) * d;
                        newForce = ORC(add)(newForce, ORC(smul)(mag, shift));
                    }
                }
            }
            R ratio = ORC(len)(radius) / initialRadius;
            V3 env;
            if (s->ph.enable_big_cells_brake && ratio > brake)
                env = ORC(sub)(G, ORC(smul)(vdamp * ratio * bint, velocity));
            else
                env = ORC(sub)(G, ORC(smul)(vdamp, velocity));
            newForce = ORC(add)(newForce, env);
            /* (F + S + E) / 6 in the Runge-Kutta build (blood_cells.cu:114-119) */
            V3 out = ORC(sdiv)(ORC(add)(initialForce, newForce), s->ph.particle_rk4 ? (R)6 : (R)2);
            s->fx[i] = out.x; s->fy[i] = out.y; s->fz[i] = out.z;
        }
    }
    s->flags |= fl;
    free(sfx); free(sfy); free(sfz);
}

/* addResilientForceOnCollision, src/simulation/physics.cuh:133-145 (returns
 * the increment coef*(spring+damping+shear)). */
static inline V3 ORC(resilient)(const ORC(sim)* s, V3 relPos, V3 relVel, R d2, R radius, R coef, uint32_t* fl)
{
    V3 dir = ORC(normalize)(relPos, fl);
    V3 tang = ORC(sub)(relVel, ORC(smul)(ORC(dot)(relVel, dir), dir));
    V3 spring = ORC(smul)(-(R)s->ph.collision_spring_coeff * (radius * (R)2 - ORC(sqrt_r)(d2)), dir);
    V3 damp = ORC(smul)((R)s->ph.collision_damping_coeff, relVel);
    V3 shear = ORC(smul)((R)s->ph.collision_shear_coeff, tang);
    return ORC(smul)(coef, ORC(add)(ORC(add)(spring, damp), shear));
}

static int ORC(type_of_slot)(const ORC(sim)* s, int32_t slot)
{
    return ORC(type_of_particle)(s, slot);
}

/* Particle collisions over the 27 neighbour cells:
 * calculateParticleCollisions<UniformGrid>, src/simulation/particle_collisions.cuh:104-269,
 * detectCollisionsInNeighborCells :53-83, detectCollision :26-38.
 * If `pairs` is non-NULL the contact (i, j) pairs are recorded instead of
 * applied (counts[i] contacts of particle i in visit order). */
static void ORC(particle_collisions_impl)(ORC(sim)* s, int32_t* counts, int32_t** lists)
{
    const ORC(grid)* g = &s->pg;
    const int64_t cX = g->cnt[0], cXY = (int64_t)g->cnt[0] * g->cnt[1];
    int32_t sl;
    uint32_t fl = 0;
    /* Largest radius of the table (+inf unless every radius is finite and
     * >= 0).  A partner farther than r1 + rmax cannot touch (d2 > (r1 + rmax)^2
     * implies d2 > (r1 + r2)^2, every rounding being monotonic), so its
     * radius is not looked up, and an index of it outside the table is not
     * flagged: BCS_FLAG_RADIUS_CLAMPED reports partners within that reach. */
    R rmax = (R)0;
    {
        int32_t m;
        for (m = 0; m < s->nm; ++m) {
            R r = s->bs[m];
            if (!(r >= (R)0 && r <= (R)3.0e38f)) { rmax = (R)INFINITY; break; }
            if (r > rmax) rmax = r;
        }
    }
#pragma omp parallel for reduction(| : fl) schedule(dynamic, 256)
    for (sl = 0; sl < s->np; ++sl) {
        int t = ORC(type_of_slot)(s, sl);
        int32_t i = g->perm[sl];
        V3 p1 = ORC(mk)(s->px[i], s->py[i], s->pz[i]);
        V3 v1 = ORC(mk)(s->vx[i], s->vy[i], s->vz[i]);
        uint32_t cellId = g->keys_sorted[sl];
        int32_t xId = ORC(sat_i32)((s->px[i] - g->minv[0]) / g->cell[0]);
        int32_t yId = ORC(sat_i32)((s->py[i] - g->minv[1]) / g->cell[1]);
        int32_t zId = ORC(sat_i32)((s->pz[i] - g->minv[2]) / g->cell[2]);
        int xl, xh, yl, yh, zl, zh;
        ORC(axis_range_i)(xId, g->cnt[0], &xl, &xh);
        ORC(axis_range_i)(yId, g->cnt[1], &yl, &yh);
        ORC(axis_range_i)(zId, g->cnt[2], &zl, &zh);
        int32_t bmi = ORC(model_index)(s, t, i, &fl);
        R r1 = s->bs[bmi];
        R rb = r1 + rmax;
        R bound2 = rb * rb;
        V3 F = ORC(mk)(s->fx[i], s->fy[i], s->fz[i]);
        int32_t ncont = 0;
        int dx, dy, dz;
        for (dx = xl; dx <= xh; ++dx)
            for (dy = yl; dy <= yh; ++dy)
                for (dz = zl; dz <= zh; ++dz) {
                    int64_t nk = (int64_t)cellId + dz * cXY + dy * cX + dx;
                    int32_t a, b, u;
                    if (!ORC(grid_range)(g, nk, &a, &b)) continue;
                    for (u = a; u <= b; ++u) {
                        int32_t j = g->perm[u];
                        if (j == i) continue;
                        V3 p2 = ORC(mk)(s->px[j], s->py[j], s->pz[j]);
                        V3 rel = ORC(sub)(p1, p2);
                        R d2 = ORC(dot)(rel, rel);
                        if (!(d2 <= bound2)) continue;
                        int32_t smi = ORC(model_index)(s, t, j, &fl);
                        R r2 = s->bs[smi];
                        R minD = r1 + r2;
                        if (d2 <= minD * minD && d2 >= (R)0.0001f) {
                            if (lists) {
                                if (lists[i]) lists[i][ncont] = j;
                                ncont++;
                            } else {
                                V3 rv = ORC(sub)(v1, ORC(mk)(s->vx[j], s->vy[j], s->vz[j]));
                                F = ORC(add)(F, ORC(resilient)(s, rel, rv, d2, r1, (R)0.5f, &fl));
                            }
                        }
                    }
                }
        if (lists) counts[i] = ncont;
        else { s->fx[i] = F.x; s->fy[i] = F.y; s->fz[i] = F.z; }
    }
    s->flags |= fl;
}

/* Moller-Trumbore as realCollisionDetection, src/simulation/vein_collisions.cu:11-45. */
static inline int ORC(ray_tri)(V3 v0, V3 v1, V3 v2, V3 orig, V3 dir, R* t_out, V3* normal, V3* refl, uint32_t* fl)
{
    const R EPS = (R)0.000001f;
    V3 edge1 = ORC(sub)(v1, v0);
    V3 edge2 = ORC(sub)(v2, v0);
    V3 h = ORC(cross)(dir, edge2);
    R a = ORC(dot)(edge1, h);
    if (a > -EPS && a < EPS) return 0;
    R f = (R)1 / a;
    V3 sv = ORC(sub)(orig, v0);
    R u = f * ORC(dot)(sv, h);
    if (u < (R)0 || u > (R)1) return 0;
    V3 q = ORC(cross)(sv, edge1);
    R v = f * ORC(dot)(dir, q);
    if (v < (R)0 || u + v > (R)1) return 0;
    R t = f * ORC(dot)(edge2, q);
    if (t > EPS) {
        *t_out = t;
        *normal = ORC(normalize)(ORC(cross)(edge2, edge1), fl);
        *refl = ORC(sub)(dir, ORC(smul)((R)2 * ORC(dot)(dir, *normal), *normal));
        return 1;
    }
    return 0;
}

/* calculateBaricentric (non-textbook edges), src/simulation/vein_collisions.cu:47-61. */
static inline V3 ORC(bary)(V3 point, V3 v0, V3 v1, V3 v2)
{
    V3 e0 = ORC(sub)(v1, v0), e1 = ORC(sub)(v2, v1), e2 = ORC(sub)(point, v0);
    R d00 = ORC(dot)(e0, e0), d01 = ORC(dot)(e0, e1), d11 = ORC(dot)(e1, e1);
    R d20 = ORC(dot)(e2, e0), d21 = ORC(dot)(e2, e1);
    R denom = d00 * d11 - d01 * d01;
    V3 r;
    r.x = (d11 * d20 - d01 * d21) / denom;
    r.y = (d00 * d21 - d01 * d20) / denom;
    r.z = (R)1 - r.x - r.y;
    return r;
}

/* Vein collisions: detectVeinCollisions<UniformGrid>,
 * src/simulation/vein_collisions.cu:63-277, calculateSideCollisions,
 * src/simulation/vein_collisions.cuh:60-93 (first hit in x->y->z cell order).
 * Vein-vertex force contributions are summed afterwards in particle-index
 * order (the reference adds them without atomics). */
static void ORC(vein_collisions)(ORC(sim)* s)
{
    const ORC(grid)* g = &s->tg;
    const int64_t cX = g->cnt[0], cXY = (int64_t)g->cnt[0] * g->cnt[1];
    const R impact2 = (R)s->ph.vein_impact_distance * (R)s->ph.vein_impact_distance;
    const R minimal2 = (R)s->ph.vein_impact_minimal_force_distance * (R)s->ph.vein_impact_minimal_force_distance;
    const R vcd = (R)s->ph.velocity_collision_damping, vcfi = (R)s->ph.vein_collision_force_intensity;
    int32_t i;
    uint32_t fl = 0;
#pragma omp parallel for reduction(| : fl) schedule(dynamic, 256)
    for (i = 0; i < s->np; ++i) {
        s->hit_tri[i] = -1;
        V3 velocity = ORC(mk)(s->vx[i], s->vy[i], s->vz[i]);
        V3 pos = ORC(mk)(s->px[i], s->py[i], s->pz[i]);
        V3 dir = ORC(normalize)(velocity, &fl);
        R rt = (R)1e10f;
        V3 normal = ORC(mk)(0, 0, 0), refl = ORC(mk)(0, 0, 0);
        int32_t obj = 0;
        int hit = 0;
        uint32_t cellId = ORC(cell_key)(g, pos.x, pos.y, pos.z, &fl);
        uint32_t xId = ORC(sat_u32)((pos.x - g->minv[0]) / g->cell[0]);
        uint32_t yId = ORC(sat_u32)((pos.y - g->minv[1]) / g->cell[1]);
        uint32_t zId = ORC(sat_u32)((pos.z - g->minv[2]) / g->cell[2]);
        int xl, xh, yl, yh, zl, zh, dx, dy, dz;
        ORC(axis_range_u)(xId, g->cnt[0], &xl, &xh);
        ORC(axis_range_u)(yId, g->cnt[1], &yl, &yh);
        ORC(axis_range_u)(zId, g->cnt[2], &zl, &zh);
        for (dx = xl; dx <= xh && !hit; ++dx)
            for (dy = yl; dy <= yh && !hit; ++dy)
                for (dz = zl; dz <= zh && !hit; ++dz) {
                    int64_t nk = (int64_t)cellId + dz * cXY + dy * cX + dx;
                    int32_t a, b, u;
                    if (!ORC(grid_range)(g, nk, &a, &b)) continue;
                    for (u = a; u <= b; ++u) {
                        int32_t tri = g->perm[u];
                        uint32_t i0 = s->idx[3 * tri], i1 = s->idx[3 * tri + 1], i2 = s->idx[3 * tri + 2];
                        V3 v0 = ORC(mk)(s->X[i0], s->Y[i0], s->Z[i0]);
                        V3 v1 = ORC(mk)(s->X[i1], s->Y[i1], s->Z[i1]);
                        V3 v2 = ORC(mk)(s->X[i2], s->Y[i2], s->Z[i2]);
                        if (!ORC(ray_tri)(v0, v1, v2, pos, dir, &rt, &normal, &refl, &fl)) continue;
                        obj = tri;
                        hit = 1;
                        break;
                    }
                }
        V3 relPos = ORC(sub)(pos, ORC(add)(pos, ORC(smul)(rt, dir)));
        R d2 = ORC(dot)(relPos, relPos);
        if (hit && d2 <= impact2) {
            V3 F = ORC(mk)(s->fx[i], s->fy[i], s->fz[i]);
            if (d2 > minimal2) {
                if (s->ph.enable_reaction_force) {
                    V3 resp = ORC(sdiv)(ORC(smul)((R)-1.0f * ORC(dot)(F, normal), normal), ORC(dot)(normal, normal));
                    F = ORC(add)(F, resp);
                } else {
                    int t = ORC(type_of_particle)(s, i);
                    R rad = s->bs[ORC(model_index)(s, t, i, &fl)];
                    F = ORC(add)(F, ORC(resilient)(s, relPos, velocity, d2, rad, (R)0.5f, &fl));
                }
            }
            s->fx[i] = F.x; s->fy[i] = F.y; s->fz[i] = F.z;
            R speed = ORC(len)(velocity);
            V3 dv = ORC(smul)((R)1, ORC(sub)(ORC(smul)(vcd * speed, refl), velocity));
            s->vx[i] = s->vx[i] + dv.x; s->vy[i] = s->vy[i] + dv.y; s->vz[i] = s->vz[i] + dv.z;
            V3 ds = ORC(smul)(vcfi, velocity);
            uint32_t i0 = s->idx[3 * obj], i1 = s->idx[3 * obj + 1], i2 = s->idx[3 * obj + 2];
            V3 v0 = ORC(mk)(s->X[i0], s->Y[i0], s->Z[i0]);
            V3 v1 = ORC(mk)(s->X[i1], s->Y[i1], s->Z[i1]);
            V3 v2 = ORC(mk)(s->X[i2], s->Y[i2], s->Z[i2]);
            V3 bc = ORC(bary)(ORC(add)(pos, ORC(smul)(rt, dir)), v0, v1, v2);
            V3 c0 = ORC(smul)(bc.x, ds), c1 = ORC(smul)(bc.y, ds), c2 = ORC(smul)(bc.z, ds);
            s->hit_tri[i] = obj;
            R* w = s->hit_w + 9 * (size_t)i;
            w[0] = c0.x; w[1] = c0.y; w[2] = c0.z;
            w[3] = c1.x; w[4] = c1.y; w[5] = c1.z;
            w[6] = c2.x; w[7] = c2.y; w[8] = c2.z;
        }
    }
    s->flags |= fl;
    /* ordered accumulation (particle index, then vertex 0,1,2) */
    for (i = 0; i < s->np; ++i) {
        int32_t tri = s->hit_tri[i];
        if (tri < 0) continue;
        const R* w = s->hit_w + 9 * (size_t)i;
        int kk;
        for (kk = 0; kk < 3; ++kk) {
            uint32_t vi = s->idx[3 * tri + kk];
            s->FX[vi] = s->FX[vi] + w[3 * kk];
            s->FY[vi] = s->FY[vi] + w[3 * kk + 1];
            s->FZ[vi] = s->FZ[vi] + w[3 * kk + 2];
        }
    }
}

/* propagateParticleForcesKernel (Heun), src/objects/blood_cells.cu:155-179. */
static void ORC(integrate_particles)(ORC(sim)* s)
{
    const R dt = (R)s->ph.dt;
    int32_t i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < s->np; ++i) {
        V3 F = ORC(mk)(s->fx[i], s->fy[i], s->fz[i]);
        V3 v0 = ORC(sdiv)(ORC(mk)(s->vx[i], s->vy[i], s->vz[i]), (R)1);
        V3 v = ORC(add)(v0, ORC(smul)(dt, F));
        s->vx[i] = v.x; s->vy[i] = v.y; s->vz[i] = v.z;
        V3 dp;
        if (s->ph.particle_rk4) {
            /* USE_RUNGE_KUTTA_FOR_PARTICLE branch, blood_cells.cu:166-171 */
            V3 k1 = ORC(smul)(dt, v);
            V3 k2 = ORC(smul)(dt, ORC(add)(v, ORC(sdiv)(k1, (R)2)));
            V3 k3 = ORC(smul)(dt, ORC(add)(v, ORC(sdiv)(k2, (R)2)));
            V3 k4 = ORC(smul)(dt, ORC(add)(v, k3));
            dp = ORC(sdiv)(ORC(add)(ORC(add)(ORC(add)(k1, ORC(smul)((R)2, k2)), ORC(smul)((R)2, k3)), k4), (R)6);
        } else {
            dp = ORC(smul)((R)0.5f * dt, ORC(add)(v, v0));
        }
        s->px[i] = s->px[i] + dp.x; s->py[i] = s->py[i] + dp.y; s->pz[i] = s->pz[i] + dp.z;
    }
}

/* propagateForcesIntoPositionsKernel (explicit Euler) + zeroed forces,
 * src/objects/vein_triangles.cu:88-117. */
static void ORC(integrate_vein)(ORC(sim)* s)
{
    const R dt = (R)s->ph.dt;
    int32_t i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < s->nv; ++i) {
        s->VX[i] = s->VX[i] + dt * s->FX[i];
        s->VY[i] = s->VY[i] + dt * s->FY[i];
        s->VZ[i] = s->VZ[i] + dt * s->FZ[i];
        s->X[i] = s->X[i] + dt * s->VX[i];
        s->Y[i] = s->Y[i] + dt * s->VY[i];
        s->Z[i] = s->Z[i] + dt * s->VZ[i];
        s->FX[i] = 0; s->FY[i] = 0; s->FZ[i] = 0;
    }
}

/* Vein-end teleport, src/simulation/vein_end.cu:57-173.  The whole cell
 * teleports if any of its particles meets the condition; one (U1, U2) pair
 * per cell from the counter RNG (the reference shares one curand state
 * between the cell's threads). */
static void ORC(vein_end)(ORC(sim)* s)
{
    const ORC(grid)* g = &s->pg;
    const R my = (R)s->ph.grid_y_margin, mxz = (R)s->ph.grid_xz_margin;
    const R upper = g->maxv[1] - (R)3 * my / (R)4;
    const R lower = g->minv[1] + my / (R)2;
    const R right = g->maxv[0] - mxz / (R)2;
    const R left = g->minv[0] + mxz / (R)2;
    const R front = g->maxv[2] - mxz / (R)2;
    const R back = g->minv[2] + mxz / (R)2;
    const R cr = (R)s->ph.cylinder_radius, spawn = (R)s->ph.min_spawn_y;
    int t;
    for (t = 0; t < s->nt; ++t) {
        int32_t p = s->ppc[t], c;
        /* SelectSynchronizationType, src/simulation/vein_end.cu:23-30: the
         * warp variant (thresholds only) when count*ppc <= 32 or 32 % ppc == 0 */
        int spheres = !((s->cnt[t] * p <= 32) || (32 % p == 0));
#pragma omp parallel for schedule(static)
        for (c = 0; c < s->cnt[t]; ++c) {
            int32_t first = s->pstart[t] + c * p, a;
            int tele = 0;
            for (a = 0; a < p && !tele; ++a) {
                int32_t i = first + a;
                R x = s->px[i], y = s->py[i], z = s->pz[i];
                int e;
                if (spheres) {
                    for (e = 0; e < s->nend; ++e) {
                        V3 dd = ORC(mk)(x - s->endc[3 * e], y - s->endc[3 * e + 1], z - s->endc[3 * e + 2]);
                        if (ORC(dot)(dd, dd) <= s->endr[e] * s->endr[e]) tele = 1;
                    }
                }
                if (y <= lower || y >= upper || x <= left || x >= right || z <= back || z >= front) tele = 1;
            }
            if (!tele) continue;
            int32_t cell = s->cstart[t] + c;
            R u1 = (R)orc_uniform(s->ph.seed, s->step, (uint64_t)cell, 0);
            R u2 = (R)orc_uniform(s->ph.seed, s->step, (uint64_t)cell, 1);
            const R* m = s->model + 3 * (size_t)s->mstart[t];
            for (a = 0; a < p; ++a) {
                int32_t i = first + a;
                s->px[i] = (u1 - (R)0.5f) * (R)1.2f * cr + m[3 * a] - m[0];
                s->py[i] = spawn + m[3 * a + 1] - m[1];
                s->pz[i] = (u2 - (R)0.5f) * (R)1.2f * cr + m[3 * a + 2] - m[2];
                s->vx[i] = (R)s->ph.init_velocity[0];
                s->vy[i] = (R)s->ph.init_velocity[1];
                s->vz[i] = (R)s->ph.init_velocity[2];
            }
        }
    }
}

/* SimulationController::calculateNextFrame, simulation_controller.cu:246-313. */
void ORC(calculate_next_frame)(ORC(sim)* s)
{
    ORC(vein_springs)(s);
    ORC(centers)(s);
    ORC(cell_springs)(s);
    ORC(particle_collisions_impl)(s, NULL, NULL);
    ORC(vein_collisions)(s);
}

/* SimulationController::propagateAll, simulation_controller.cu:315-331. */
void ORC(propagate_all)(ORC(sim)* s)
{
    ORC(integrate_particles)(s);
    ORC(integrate_vein)(s);
    if (s->ph.use_blood_flow) ORC(vein_end)(s);
    s->step++;
}

/* One frame of src/main.cu:175-208.  The triangle grid is rebuilt every
 * frame, as the reference does, on the creation-time centres (frozen in the
 * reference's single-GPU mode) unless dynamic_triangle_grid is set. */
void ORC(step)(ORC(sim)* s, int32_t n)
{
    int32_t k;
    for (k = 0; k < n; ++k) {
        ORC(calculate_grid)(s, BCS_GRID_PARTICLES);
        if (s->ph.dynamic_triangle_grid) ORC(update_triangle_centers)(s);
        ORC(calculate_grid)(s, BCS_GRID_TRIANGLES);
        ORC(calculate_next_frame)(s);
        ORC(propagate_all)(s);
    }
}

/* Contact pairs on the current particle grid (no state change). */
int64_t ORC(contacts)(ORC(sim)* s, int32_t* counts, int32_t* pairs, int64_t capacity)
{
    int32_t n = s->np, i;
    int32_t** lists = (int32_t**)calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t*));
    ORC(particle_collisions_impl)(s, counts, lists);
    int64_t total = 0;
    for (i = 0; i < n; ++i) total += counts[i];
    if (pairs && total <= capacity) {
        for (i = 0; i < n; ++i) lists[i] = (int32_t*)malloc((size_t)(counts[i] > 0 ? counts[i] : 1) * 4);
        ORC(particle_collisions_impl)(s, counts, lists);
        int64_t off = 0;
        for (i = 0; i < n; ++i) {
            memcpy(pairs + off, lists[i], (size_t)counts[i] * 4);
            off += counts[i];
            free(lists[i]);
        }
    }
    free(lists);
    return total;
}

#undef V3
