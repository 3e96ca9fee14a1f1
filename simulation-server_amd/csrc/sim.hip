// sim.hip -- the C-ABI of include/bcs.h: device state, phase sequencing and
// the graph-replayed fused frame.
//
// Replaces, for one GPU: SimulationController (src/simulation/simulation_controller.cu:42-331),
// BloodCells (src/objects/blood_cells.cu:16-42), VeinTriangles
// (src/objects/vein_triangles.cu:37-86), VeinNeighbors
// (src/objects/vein_neighbors.cu:8-54) and the two UniformGrid objects of
// src/main.cu:130-133.  All device memory is owned by the bcs_sim object and
// all work is enqueued on its single stream (default priority).
#include "step.hpp"
#include "surrogate.hpp"

#include <algorithm>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace bcs {
void launch_raycand(const StepArgs& A, int nchunks, hipStream_t st);
void launch_raycand_ray(const StepArgs& A, int nchunks, hipStream_t st);
bool ray_fuse_auto(int nchunks);
void launch_ray(const StepArgs& A, hipStream_t st);
void launch_cells(const StepArgs& A, int nchunks, bool fused, hipStream_t st);
void launch_integrate(const StepArgs& A, hipStream_t st);
void launch_vein_end(const StepArgs& A, int nchunks, hipStream_t st);
void launch_bump(uint64_t* step, hipStream_t st);
void launch_vein_springs(const VeinArgs& V, hipStream_t st);
void launch_vein(const VeinArgs& V, int mode, hipStream_t st);
void launch_tri_centers(const float* X, int32_t nv, const uint32_t* idx, int32_t ntri, float* c, hipStream_t st);
void launch_contacts(const StepArgs& A, int pass, int32_t* counts, const int64_t* offs, int32_t* pairs, hipStream_t st);
void launch_copy(const void* src, void* dst, int64_t n16, hipStream_t st);
void launch_gather_start(const uint2* table, const uint32_t* idx, int32_t n, uint32_t* out, hipStream_t st);
void launch_dense_ranges(const uint32_t* keys_sorted, int32_t n, uint64_t cell_count, int32_t* starts,
                         int32_t* ends, hipStream_t st);
void launch_vein_prep(const VeinArgs& V, int32_t ntri, const uint4* trun, const uint4* trec, float4* tdata,
                      float4* aabb, hipStream_t st);
void launch_tri_records(const DeviceGrid& tg, const uint32_t* idx, uint4* trun, uint4* trec, hipStream_t st);
void launch_tri_nmask(const DeviceGrid& tg, uint32_t* nmask, hipStream_t st);
void launch_export(const float* soa, int64_t n, int64_t first, int64_t count, int stride, float* out, hipStream_t st);
void launch_refresh_copies(const float* pos, const float* vel, int32_t n, const int32_t* perm, float4* pvs,
                           hipStream_t st);
void launch_chunk_map(const StepArgs& A, int2* out, hipStream_t st);
void launch_surr_features(const float* pos, const float* vel, int64_t np, int32_t pstart, int32_t ncells,
                          const uint8_t* touched, float pos_scale, float vel_scale, uint16_t* X, float* f32,
                          hipStream_t st);
void launch_cells_surr(const StepArgs& A, int nchunks, hipStream_t st);
void launch_split_part(const StepArgs& A, int part, bool surr, hipStream_t st);
int launch_order(const TypeTable& tt, const GridGeom& g, const float* centers, uint32_t* okey, uint32_t* ocount,
                 uint32_t* ostart, uint64_t* ostatus, int32_t otiles, uint32_t* octrl, int32_t* corder,
                 uint32_t* flags, hipStream_t st);
// batched launches (ensembles)
VeinPrepArgs vein_prep_args(const VeinArgs& V, int32_t ntri, const uint4* trun, const uint4* trec, float4* tdata,
                            float4* aabb, uint32_t* blocks);
uint32_t ray_split(int nchunks);
uint32_t cell_blocks(int nchunks);
uint32_t tile_blocks(int64_t np);
void batch_split_part(const StepArgs* a, int part, const uint32_t* start, int32_t nsim, uint32_t grid, uint32_t run,
                      hipStream_t st);
uint32_t ray_blocks(int nchunks);
uint32_t vein_blocks(int32_t nv);
uint32_t order_key_blocks(const TypeTable& tt, int32_t otiles);
uint32_t order_scatter_blocks(const TypeTable& tt);
void batch_raycand(bool fused, const StepArgs* a, const uint32_t* start, int32_t nsim, uint32_t grid, uint32_t run,
                   hipStream_t st);
void batch_ray(const StepArgs* a, const uint32_t* split, const uint32_t* start, int32_t nsim, uint32_t grid,
               uint32_t run, hipStream_t st);
void batch_cells(const StepArgs* a, const uint32_t* start, int32_t nsim, uint32_t grid, uint32_t run, hipStream_t st);
void batch_vein(const VeinArgs* a, const uint32_t* start, int32_t nsim, uint32_t grid, uint32_t run, hipStream_t st);
void batch_vein_prep(const VeinPrepArgs* a, const uint32_t* start, int32_t nsim, uint32_t grid, uint32_t run,
                     hipStream_t st);
void batch_order(const OrderKeyArgs* ka, const uint32_t* kstart, uint32_t kgrid, const ScanArgs* sa,
                 const uint32_t* sstart, uint32_t sgrid, const OrderScatterArgs* oa, const uint32_t* ostart,
                 uint32_t ogrid, const StepArgs* a, int2* const* out, const uint32_t* cstart, uint32_t cgrid,
                 int32_t nsim, uint32_t run, hipStream_t st);
}  // namespace bcs

using namespace bcs;

#define HIPCHK(call)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess) return fail(BCS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct bcs_sim {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;        // second branch of the frame (vein springs, triangle data)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipStream_t third = nullptr;       // third branch: the surrogate's network (when one is attached)
    hipEvent_t ev_join3 = nullptr;
    bcs_physics phys{};
    int32_t nt = 0, np = 0, nc = 0, nm = 0, nv = 0, ntri = 0;
    float bmin[3]{}, bmax[3]{};
    TypeTable tt{};
    PhysConst pc{};
    int nchunks = 0;
    std::vector<void*> allocs;
    // particles
    float *pos = nullptr, *vel = nullptr, *frc = nullptr, *centers = nullptr;
    float *bs = nullptr, *r0 = nullptr, *model = nullptr, *sp_L = nullptr;
    int32_t *sp_off = nullptr, *sp_b = nullptr, *chunk_type = nullptr, *chunk_first = nullptr, *chunk_n = nullptr;
    // cell processing order (spatial; see k_order_key)
    int32_t* corder = nullptr;
    int2* chunk_pi = nullptr;  // per chunk and lane: {particle, blood cell} (follows corder)
    uint32_t *okey = nullptr, *ocount = nullptr, *ostart = nullptr, *octrl = nullptr;
    uint64_t* ostatus = nullptr;
    int32_t otiles = 0;
    int32_t frames_since_order = 0;
    // vein
    float *X = nullptr, *V = nullptr, *Fv = nullptr, *nbl = nullptr, *tcent = nullptr;
    int32_t* nbr = nullptr;
    uint32_t* tidx = nullptr;
    uint32_t* hctrl = nullptr;
    float4* hits = nullptr;                     // 3 float4 per hit
    int32_t *vhead = nullptr, *vnext = nullptr;  // per-vertex lists of hit entries
    float4* t_aabb = nullptr;  // triangle-cell boxes by cell key (per frame)
    uint32_t* t_nmask = nullptr;  // per triangle cell: occupied neighbourhood mask
    uint4* trun = nullptr;  // per occupied triangle cell: run record (k_tri_records)
    uint4* trec = nullptr;  // per sorted triangle: vertex indices and triangle
    float4* tdata = nullptr;   // per sorted triangle: (v0 | id), e1, e2 (per frame)
    uint32_t* best = nullptr;  // per particle first-hit key
    uint4* work = nullptr;     // ray work items
    uint32_t* wcount = nullptr;  // per k_raycand block: work items
    uint32_t region_cap = 0;     // work items per block region
    int nblocks = 0;             // k_raycand / k_cells grid (8 * ceil(nchunks / 8))
    // grids
    DeviceGrid pg, tg;
    bool tg_built = false;
    bool copies_current = false;  // packed position/velocity copies match pos/vel
    // misc
    uint32_t* flags = nullptr;
    uint64_t* step = nullptr;
    hipGraphExec_t graph_exec = nullptr;    // frame without the leading vein update ...
    hipGraphExec_t graph_exec_b = nullptr;  // ... and with it (the previous frame's k_vein)
    hipGraphExec_t graph_exec_m = nullptr;  // kGraphFrames frames with it, in one launch
    bool use_graph = true;
    int32_t ray_mode = BCS_RAY_AUTO;
    bool ray_fused() const
    {
        return ray_mode == BCS_RAY_FUSED || (ray_mode == BCS_RAY_AUTO && ray_fuse_auto(nchunks));
    }
    int32_t contact_mode = BCS_CONTACT_AUTO;
    float4* fslot = nullptr;  // split frames: F per sorted slot
    bool contact_split() const
    {
        return contact_mode == BCS_CONTACT_SLOTS || (contact_mode == BCS_CONTACT_AUTO && cell_blocks(nchunks) > 4096u);
    }
    hipGraph_t graph = nullptr, graph_b = nullptr, graph_m = nullptr;
    std::vector<float> host_bs;
    std::vector<float> host_smallest;
    // dense gridCellStarts / gridCellEnds views (bcs_sim_grid_view), per grid
    int32_t* dense_start[2] = {nullptr, nullptr};
    int32_t* dense_end[2] = {nullptr, nullptr};
    // learned spring surrogate (bcs_sim_attach_surrogate, bcs_sim_surrogate_sample)
    bcs_surrogate* surr = nullptr;
    SurrNet surr_net{};
    float surr_tol = 0.0f;
    int32_t surr_type = -1;    // the first type of 42 particles per cell
    int32_t surr_cells = 0;
    bool sampling = false;     // this frame writes the plain spring displacements (sf_S)
    bool recording = false;    // this frame keeps the contact flags (touched)
    uint16_t *sf_X = nullptr, *sf_a0 = nullptr, *sf_a1 = nullptr;
    float *sf_out = nullptr, *sf_S = nullptr, *sf_f32 = nullptr;
    uint8_t* touched = nullptr;
    uint32_t* surr_stats = nullptr;
    bool surr_active() const { return surr_type >= 0 && (surr || sampling || recording); }

    ~bcs_sim()
    {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (graph_exec_b) (void)hipGraphExecDestroy(graph_exec_b);
        if (graph_b) (void)hipGraphDestroy(graph_b);
        if (graph_exec_m) (void)hipGraphExecDestroy(graph_exec_m);
        if (graph_m) (void)hipGraphDestroy(graph_m);
        grid_free(pg);
        grid_free(tg);
        for (void* p : allocs) (void)hipFree(p);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (ev_join3) (void)hipEventDestroy(ev_join3);
        if (third) (void)hipStreamDestroy(third);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <class T>
    int dalloc(T** p, size_t count)
    {
        hipError_t e = hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) return fail(BCS_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back((void*)*p);
        return BCS_OK;
    }

    StepArgs step_args() const
    {
        StepArgs A{};
        A.tt = tt;
        A.pc = pc;
        A.pos = pos; A.vel = vel; A.frc = frc; A.centers = centers;
        A.bs = bs; A.r0 = r0; A.model = model; A.sp_off = sp_off; A.sp_b = sp_b; A.sp_L = sp_L;
        A.chunk_type = chunk_type; A.chunk_first = chunk_first; A.chunk_n = chunk_n;
        A.corder = corder; A.chunk_pi = chunk_pi; A.nchunks = nchunks;
        A.pg = pg.g; A.p_keys = pg.keys_sorted; A.p_perm = pg.perm; A.p_slot = pg.slot_of; A.p_table = pg.table; A.p_map = pg.bm; A.p_key_of = pg.keys;
        A.pvs = pg.pvs;
        A.fslot = fslot;
        A.ntiles = (int32_t)((np + kBlock - 1) / kBlock);
        A.tg = tg.g; A.t_table = tg.table; A.tdata = tdata; A.ntri = ntri;
        A.nv = nv; A.X = X; A.tidx = tidx;
        A.hits = hits; A.vhead = vhead; A.vnext = vnext; A.hctrl = hctrl;
        A.t_aabb = t_aabb; A.t_nmask = t_nmask;
        A.best = best; A.work = work; A.wcount = wcount; A.region_cap = region_cap;
        A.flags = flags; A.step = step;
        A.surr = SurrArgs{};
        A.surr.type = -1;
        if (surr_active()) {
            A.surr.D = (surr && !sampling) ? sf_out : nullptr;
            A.surr.S_out = sampling ? sf_S : nullptr;
            A.surr.touched = touched;
            A.surr.stats = surr_stats;
            A.surr.type = surr_type;
            A.surr.scale = (pc.dt * pc.dt) * 0.25f;
            A.surr.out_scale = surr_net.out_scale;
            A.surr.tol = surr_tol;
        }
        return A;
    }
    VeinArgs vein_args() const
    {
        VeinArgs Va{};
        Va.nv = nv; Va.X = X; Va.V = V; Va.Fv = Fv; Va.nbr = nbr; Va.nbl = nbl;
        Va.dt = phys.dt; Va.vk = phys.vein_k_sniff; Va.vd = phys.vein_d_fact;
        Va.hits = hits; Va.vhead = vhead; Va.vnext = vnext; Va.hctrl = hctrl; Va.tidx = tidx;
        Va.flags = flags; Va.step = step;
        return Va;
    }
};

namespace {

GridGeom make_geom(const float bmin[3], const float bmax[3], const float cell[3])
{
    GridGeom g{};
    for (int d = 0; d < 3; ++d) {
        g.minv[d] = bmin[d];
        g.maxv[d] = bmax[d];
        g.extent[d] = bmax[d] - bmin[d];
        g.cell[d] = cell[d];
        float q = g.extent[d] / g.cell[d];  // static_cast<int>(width / cellWidth), uniform_grid.cu:86-88
        g.cnt[d] = (q >= 2147483647.0f) ? 2147483647 : (q > 0 ? (int32_t)q : 0);
    }
    g.cell_count = (uint64_t)g.cnt[0] * (uint64_t)g.cnt[1] * (uint64_t)g.cnt[2];
    g.div_x = make_div32((uint32_t)std::max(g.cnt[0], 1));
    g.div_y = make_div32((uint32_t)std::max(g.cnt[1], 1));
    return g;
}

int set_dev(bcs_sim* s)
{
    if (!s) return fail(BCS_ERR_INVALID_ARGUMENT, "null simulation");
    HIPCHK(hipSetDevice(s->device));
    return BCS_OK;
}

#define RC(x)                 \
    do {                      \
        int rc_ = (x);        \
        if (rc_) return rc_;  \
    } while (0)

// Triangle grid on the triangle centres: frozen after creation (the
// reference's single-GPU behaviour) unless dynamic_triangle_grid is set.  The
// per-triangle data and cell boxes are rebuilt every frame (k_vein_prep).
int build_triangle_grid(bcs_sim* s)
{
    // host-side sizes before any launch: the grid holds every triangle, the
    // per-triangle arrays and the cell table were allocated for this mesh
    if (s->ntri < 0 || s->ntri > s->tg.cap || !s->tcent || !s->trun || !s->trec || !s->t_nmask ||
        (s->ntri > 0 && (!s->tidx || !s->tg.table)) || s->tg.g.cell_count == 0 ||
        s->tg.g.cell_count > 0xFFFFFFFFull)
        return fail(BCS_ERR_STATE, "triangle grid: buffers do not cover the mesh");
    const bool dyn = s->phys.dynamic_triangle_grid != 0;
    if (dyn || !s->tg_built) launch_tri_centers(s->X, s->nv, s->tidx, s->ntri, s->tcent, s->stream);
    RC(grid_build(s->tg, s->tcent, s->ntri, s->ntri, nullptr, s->flags, s->stream));
    launch_tri_nmask(s->tg, s->t_nmask, s->stream);
    launch_tri_records(s->tg, s->tidx, s->trun, s->trec, s->stream);
    s->tg_built = true;
    return BCS_OK;
}

int enqueue_particle_grid(bcs_sim* s, const KernelTimer* timer = nullptr)
{
    RC(grid_build(s->pg, s->pos, s->np, s->np, s->vel, s->flags, s->stream, timer, s->chunk_pi,
                  s->nchunks * kBlock));
    s->copies_current = true;
    return BCS_OK;
}


// Network input of the served cells from the frame's starting state, then
// the network itself (the spring displacements k_cells checks and uses).
void enqueue_surrogate(bcs_sim* s, hipStream_t st)
{
    launch_surr_features(s->pos, s->vel, s->np, s->tt.pstart[s->surr_type], s->surr_cells, s->touched,
                         s->surr_net.pos_scale, s->surr_net.vel_scale, s->sf_X, nullptr, st);
    surrogate_infer(s->surr_net, s->sf_X, s->surr_cells, s->sf_a0, s->sf_a1, s->sf_out, st);
}

// One fused frame.  Main branch: particle grid, then k_cells (centres,
// springs, contacts, vein hit response, integrate, vein end).  Vein branch,
// beside the grid build: triangle data and boxes, vein ray candidates, ray
// test, vein springs.  Then the vein sum + integrate (k_vein).
// lead_vein: the previous frame's vein update (k_vein) has not run yet; it
// heads the vein branch (it needs nothing from this frame's particle grid).
// tail_vein: this frame's vein update runs at its end (else the caller runs
// it, as the next frame's lead or on its own).
int enqueue_fused_frame(bcs_sim* s, const KernelTimer* timer = nullptr, bool lead_vein = false,
                        bool tail_vein = true)
{
    hipStream_t st = s->stream;
    auto tb = [&](int id) { if (timer) timer->begin(id, st); };
    auto te = [&](int id) { if (timer) timer->end(id, st); };
    if (s->phys.dynamic_triangle_grid) RC(build_triangle_grid(s));
    const StepArgs A = s->step_args();
    const VeinArgs Va = s->vein_args();
    if (!timer) {
        // The vein ray test needs only the frame's starting x, v and the vein
        // state, and the vein springs only the vein state: a second branch
        // runs them while the particle grid is built.
        HIPCHK(hipEventRecord(s->ev_fork, st));
        HIPCHK(hipStreamWaitEvent(s->side, s->ev_fork, 0));
        if (lead_vein) launch_vein(Va, 7, s->side);
        launch_vein_prep(Va, s->ntri, s->trun, s->trec, s->tdata, s->t_aabb, s->side);
        if (s->ray_fused()) {
            launch_raycand_ray(A, s->nchunks, s->side);
        } else {
            launch_raycand(A, s->nchunks, s->side);
            launch_ray(A, s->side);
        }
        // the network's springs need only the frame's starting x, v and the
        // last frame's contact flags: a third branch beside the other two
        const bool net = s->surr && !s->sampling;
        if (net) {
            HIPCHK(hipStreamWaitEvent(s->third, s->ev_fork, 0));
            enqueue_surrogate(s, s->third);
            HIPCHK(hipEventRecord(s->ev_join3, s->third));
        }
        HIPCHK(hipEventRecord(s->ev_join, s->side));
        RC(enqueue_particle_grid(s));
        HIPCHK(hipStreamWaitEvent(st, s->ev_join, 0));
        if (net) HIPCHK(hipStreamWaitEvent(st, s->ev_join3, 0));
    } else {
        if (lead_vein) launch_vein(Va, 7, st);
        RC(enqueue_particle_grid(s, timer));
        tb(kGridKernels + 0);
        launch_vein_prep(Va, s->ntri, s->trun, s->trec, s->tdata, s->t_aabb, st);
        te(kGridKernels + 0);
        // fused: the one launch is timed in the k_raycand slot (named
        // k_raycand_ray), the k_ray slot times no launch
        tb(kGridKernels + 1);
        if (s->ray_fused()) launch_raycand_ray(A, s->nchunks, st);
        else launch_raycand(A, s->nchunks, st);
        te(kGridKernels + 1);
        tb(kGridKernels + 2);
        if (!s->ray_fused()) launch_ray(A, st);
        te(kGridKernels + 2);
        if (s->surr && !s->sampling) enqueue_surrogate(s, st);
    }
    if (s->contact_split()) {
        // springs per blood-cell chunk, contacts per tile of sorted slots,
        // response and integration per chunk again
        for (int part = 0; part < 3; ++part) {
            tb(kGridKernels + 5 + part);
            launch_split_part(A, part, s->surr_active(), st);
            te(kGridKernels + 5 + part);
        }
    } else {
        tb(kGridKernels + 3);
        if (s->surr_active()) launch_cells_surr(A, s->nchunks, st);
        else launch_cells(A, s->nchunks, true, st);
        te(kGridKernels + 3);
    }
    tb(kGridKernels + 4);
    if (tail_vein) launch_vein(Va, 7, st);
    te(kGridKernels + 4);
    HIPCHK(hipGetLastError());
    return BCS_OK;
}

const char* kKernelNames =
    "k_key;k_scan;k_scatter;k_rank;k_occ;k_vein_prep;k_raycand;k_ray;k_cells;k_vein;k_springs;k_contacts;k_respond";
constexpr int kFusedKernels = kGridKernels + 8;
constexpr int32_t kGraphFrames = 16;  // frames per multi-frame graph launch

}  // namespace

extern "C" int bcs_device_count(int32_t* count)
{
    if (!count) return fail(BCS_ERR_INVALID_ARGUMENT, "null count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return BCS_OK;
}

extern "C" int bcs_sim_create(const bcs_physics* phys, const bcs_cell_type* types, int32_t n_types,
                              const bcs_vein_mesh* mesh, int32_t device, bcs_sim** out)
{
    if (!phys || !types || !mesh || !out) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_create: null argument");
    *out = nullptr;
    if (n_types < 1 || n_types > kMaxTypes) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_create: 1..16 cell types");
    if (mesh->vertex_count < 1 || mesh->triangle_count < 1 || !mesh->positions || !mesh->indices)
        return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_create: empty vein mesh");
    if (mesh->ending_count < 0 || mesh->ending_count > kMaxEndings)
        return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_create: at most 16 vein endings");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(BCS_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return fail(BCS_ERR_NO_DEVICE, "device index out of range");
    std::unique_ptr<bcs_sim> s(new bcs_sim());
    s->device = device;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming));
    HIPCHK(hipStreamCreateWithFlags(&s->third, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&s->ev_join3, hipEventDisableTiming));
    s->phys = *phys;
    s->nt = n_types;
    // ---- type tables (blood_cell_factory.hpp:172-282)
    TypeTable& tt = s->tt;
    tt.nt = n_types;
    int64_t np = 0, nc = 0, nm = 0;
    for (int t = 0; t < n_types; ++t) {
        const bcs_cell_type& c = types[t];
        if (c.count < 0 || c.particles_in_cell < 1 || c.particles_in_cell > kBlock || !c.vertices)
            return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_create: bad cell type (1 <= ppc <= 256)");
        for (int k = 0; k < c.spring_count; ++k)
            if (c.spring_start[k] < 0 || c.spring_end[k] < 0 || c.spring_start[k] >= c.particles_in_cell ||
                c.spring_end[k] >= c.particles_in_cell)
                return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_create: ill-formed spring");
        tt.ppc[t] = c.particles_in_cell;
        tt.ppc_div[t] = make_div32((uint32_t)c.particles_in_cell);
        tt.cnt[t] = c.count;
        tt.pstart[t] = (int32_t)np;
        tt.cstart[t] = (int32_t)nc;
        tt.mstart[t] = (int32_t)nm;
        tt.end_spheres[t] = !((int64_t)c.count * c.particles_in_cell <= 32 || (32 % c.particles_in_cell) == 0);
        np += (int64_t)c.count * c.particles_in_cell;
        nc += c.count;
        nm += c.particles_in_cell;
    }
    if (np < 1) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_create: no particles");
    if (np > (1 << 29)) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_create: more than 2^29 particles");
    tt.np = (int32_t)np; tt.nc = (int32_t)nc; tt.nm = (int32_t)nm;
    s->np = tt.np; s->nc = tt.nc; s->nm = tt.nm;
    s->nv = mesh->vertex_count;
    s->ntri = mesh->triangle_count;
    if (s->ntri >= (1 << 27)) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_create: at most 2^27 - 1 triangles");
    for (int64_t k = 0; k < 3 * (int64_t)s->ntri; ++k)
        if (mesh->indices[k] >= (uint32_t)s->nv) return fail(BCS_ERR_INVALID_ARGUMENT, "vein index out of range");
    // ---- host tables
    std::vector<float> hbs(nm), hr0(nm), hmodel(3 * nm);
    RC(bcs_model_tables(phys, types, n_types, hbs.data(), hr0.data()));
    s->host_bs = hbs;
    std::vector<int32_t> hoff(nm + 1, 0), hb;
    std::vector<float> hL;
    for (int t = 0; t < n_types; ++t) {
        const bcs_cell_type& c = types[t];
        const int p = c.particles_in_cell;
        std::memcpy(hmodel.data() + 3 * tt.mstart[t], c.vertices, sizeof(float) * 3 * p);
        // dense springGraph, blood_cell_factory.hpp:292-328 (later springs overwrite)
        std::vector<float> dense((size_t)p * p, 0.f);
        for (int k = 0; k < c.spring_count; ++k) {
            float L = c.spring_length[k] * 1.0f;
            dense[(size_t)c.spring_start[k] * p + c.spring_end[k]] = L;
            dense[(size_t)c.spring_end[k] * p + c.spring_start[k]] = L;
        }
        // CSR of column a in ascending b (the order of gatherForcesKernel's loop)
        for (int a = 0; a < p; ++a) {
            for (int b = 0; b < p; ++b) {
                float L = dense[(size_t)b * p + a];
                if (L != 0.f) {
                    hb.push_back(b);
                    hL.push_back(L);
                }
            }
            hoff[tt.mstart[t] + a + 1] = (int32_t)hb.size();
        }
    }
    float bmin[3], bmax[3];
    RC(bcs_grid_bounds(mesh, phys, bmin, bmax));
    std::memcpy(s->bmin, bmin, sizeof(bmin));
    std::memcpy(s->bmax, bmax, sizeof(bmax));
    // ---- physics constants (physics.hpp, simulation.hpp, vein_end.cu:12-18)
    PhysConst& pc = s->pc;
    pc.dt = phys->dt; pc.k = phys->particle_k_sniff; pc.d = phys->particle_d_fact;
    pc.vk = phys->vein_k_sniff; pc.vd = phys->vein_d_fact;
    pc.vcd = phys->velocity_collision_damping; pc.vcfi = phys->vein_collision_force_intensity;
    pc.vdamp = phys->viscous_damping; pc.cs = phys->collision_spring_coeff; pc.cd = phys->collision_damping_coeff;
    pc.csh = phys->collision_shear_coeff; pc.brake = phys->max_cell_size_factor_before_brake;
    pc.bint = phys->big_particle_braking_intensity;
    for (int d = 0; d < 3; ++d) { pc.initv[d] = phys->init_velocity[d]; pc.G[d] = phys->gravity[d]; }
    pc.impact2 = phys->vein_impact_distance * phys->vein_impact_distance;
    pc.minimal2 = phys->vein_impact_minimal_force_distance * phys->vein_impact_minimal_force_distance;
    pc.cyl_r = phys->cylinder_radius;
    pc.spawn_y = phys->min_spawn_y;
    const float my = phys->grid_y_margin, mxz = phys->grid_xz_margin;
    pc.thr_upper = bmax[1] - 3.0f * my / 4.0f;
    pc.thr_lower = bmin[1] + my / 2.0f;
    pc.thr_right = bmax[0] - mxz / 2.0f;
    pc.thr_left = bmin[0] + mxz / 2.0f;
    pc.thr_front = bmax[2] - mxz / 2.0f;
    pc.thr_back = bmin[2] + mxz / 2.0f;
    pc.use_blood_flow = phys->use_blood_flow;
    pc.reaction = phys->enable_reaction_force;
    pc.brake_on = phys->enable_big_cells_brake;
    pc.rk4 = phys->particle_rk4 ? 1 : 0;
    pc.seed = phys->seed;
    pc.nend = mesh->ending_count;
    {
        float mabs = 0.f;
        for (int d = 0; d < 3; ++d) mabs = std::max(mabs, std::max(std::fabs(bmin[d]), std::fabs(bmax[d])));
        pc.aabb_margin = 0.05f + 1e-5f * mabs;
        // contact-search bound: +inf unless every radius is finite and >= 0
        float rmax = 0.0f;
        for (float r : s->host_bs)
            if (!(r >= 0.0f && r <= 3.0e38f)) rmax = INFINITY;
            else rmax = std::max(rmax, r);
        pc.rmax = rmax;
        // The contact search may skip the partner radius lookup when it
        // cannot change a result: every radius equal (so r2 = rmax) and no
        // partner model index clamped (a partner below the launch type's
        // first particle selects mstart - ((pstart - j) mod ppc), which stays
        // in the table when mstart >= ppc - 1).
        bool uni = rmax < INFINITY;
        for (float r : s->host_bs) uni = uni && r == rmax;
        for (int t = 0; t < n_types; ++t)
            if (tt.pstart[t] > 0 && tt.mstart[t] < tt.ppc[t] - 1) uni = false;
        pc.r_uniform = uni ? 1 : 0;
    }
    for (int e = 0; e < mesh->ending_count; ++e) {
        for (int d = 0; d < 3; ++d) pc.endc[3 * e + d] = mesh->ending_centers[3 * e + d];
        pc.endr[e] = mesh->ending_radii[e];
    }
    // smallestRadiusInType (simulation_controller.cuh:41)
    for (int t = 0; t < n_types; ++t) {
        float m = 3.4e38f;
        for (int a = 0; a < tt.ppc[t]; ++a) m = std::min(m, hbs[tt.mstart[t] + a]);
        s->host_smallest.push_back(m);
    }
    // ---- work chunks (whole cells of one type per block)
    std::vector<int32_t> ct, cf, cn;
    for (int t = 0; t < n_types; ++t) {
        const int cpb = kBlock / tt.ppc[t];
        for (int c0 = 0; c0 < tt.cnt[t]; c0 += cpb) {
            ct.push_back(t);
            cf.push_back(c0);
            cn.push_back(std::min(cpb, tt.cnt[t] - c0));
        }
    }
    s->nchunks = (int)ct.size();
    // ---- device allocations
    const size_t N = (size_t)np, NV = (size_t)s->nv, NT = (size_t)s->ntri;
    RC(s->dalloc(&s->pos, 3 * N)); RC(s->dalloc(&s->vel, 3 * N)); RC(s->dalloc(&s->frc, 3 * N));
    RC(s->dalloc(&s->centers, 3 * (size_t)nc));
    RC(s->dalloc(&s->bs, nm)); RC(s->dalloc(&s->r0, nm)); RC(s->dalloc(&s->model, 3 * nm));
    RC(s->dalloc(&s->sp_off, nm + 1)); RC(s->dalloc(&s->sp_b, hb.size())); RC(s->dalloc(&s->sp_L, hL.size()));
    RC(s->dalloc(&s->chunk_type, ct.size())); RC(s->dalloc(&s->chunk_first, cf.size())); RC(s->dalloc(&s->chunk_n, cn.size()));
    RC(s->dalloc(&s->chunk_pi, ct.size() * kBlock));
    {
        const int64_t nbins = (int64_t)n_types << kOrderBits;
        s->otiles = scan_tiles(nbins);
        RC(s->dalloc(&s->corder, nc)); RC(s->dalloc(&s->okey, nc));
        RC(s->dalloc(&s->ocount, nbins)); RC(s->dalloc(&s->ostart, nbins + 1));
        RC(s->dalloc(&s->ostatus, s->otiles)); RC(s->dalloc(&s->octrl, 4));
    }
    RC(s->dalloc(&s->X, 3 * NV)); RC(s->dalloc(&s->V, 3 * NV)); RC(s->dalloc(&s->Fv, 3 * NV));
    RC(s->dalloc(&s->nbr, kVeinSlots * NV)); RC(s->dalloc(&s->nbl, kVeinSlots * NV));
    RC(s->dalloc(&s->tidx, 3 * NT)); RC(s->dalloc(&s->tcent, 3 * NT));
    RC(s->dalloc(&s->hits, 3 * N)); RC(s->dalloc(&s->vnext, 3 * N)); RC(s->dalloc(&s->vhead, NV));
    RC(s->dalloc(&s->hctrl, 16));
    RC(s->dalloc(&s->flags, 4)); RC(s->dalloc(&s->step, 1));
    RC(s->dalloc(&s->tdata, 3 * NT));
    // ray work list: 8 items per particle, split evenly over the k_raycand blocks
    s->nblocks = 8 * ((s->nchunks + 7) / 8);
    {
        const int32_t items = phys->ray_work_items;
        const uint64_t nb = (uint64_t)std::max(s->nblocks, 1);
        const uint64_t per = items > 0 ? std::max<uint64_t>((uint64_t)items / nb, 1)
                                       : std::max<uint64_t>(std::max<uint64_t>(8 * N, 65536) / nb, 1024);
        s->region_cap = (uint32_t)std::min<uint64_t>(per, 0x7FFFFFFFull);
    }
    RC(s->dalloc(&s->best, N));
    RC(s->dalloc(&s->work, (size_t)s->region_cap * std::max(s->nblocks, 1)));
    RC(s->dalloc(&s->wcount, std::max(s->nblocks, 1)));
    // ---- uploads
    HIPCHK(hipMemsetAsync(s->pos, 0, 12 * std::max<size_t>(N, 1), s->stream));
    HIPCHK(hipMemsetAsync(s->vel, 0, 12 * std::max<size_t>(N, 1), s->stream));
    HIPCHK(hipMemsetAsync(s->frc, 0, 12 * std::max<size_t>(N, 1), s->stream));
    HIPCHK(hipMemsetAsync(s->centers, 0, 12 * std::max<int64_t>(nc, 1), s->stream));
    HIPCHK(hipMemcpyAsync(s->bs, hbs.data(), 4 * nm, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->r0, hr0.data(), 4 * nm, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->model, hmodel.data(), 12 * nm, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->sp_off, hoff.data(), 4 * (nm + 1), hipMemcpyHostToDevice, s->stream));
    if (!hb.empty()) {
        HIPCHK(hipMemcpyAsync(s->sp_b, hb.data(), 4 * hb.size(), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(s->sp_L, hL.data(), 4 * hL.size(), hipMemcpyHostToDevice, s->stream));
    }
    {
        std::vector<int32_t> ident(std::max<int64_t>(nc, 1));
        for (int64_t c = 0; c < nc; ++c) ident[c] = (int32_t)c;
        HIPCHK(hipMemcpyAsync(s->corder, ident.data(), 4 * (size_t)nc, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemsetAsync(s->ocount, 0, 4 * ((size_t)n_types << kOrderBits), s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));  // ident dies here
    }
    if (!ct.empty()) {
        HIPCHK(hipMemcpyAsync(s->chunk_type, ct.data(), 4 * ct.size(), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(s->chunk_first, cf.data(), 4 * cf.size(), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(s->chunk_n, cn.data(), 4 * cn.size(), hipMemcpyHostToDevice, s->stream));
    }
    launch_chunk_map(s->step_args(), s->chunk_pi, s->stream);  // after corder and the chunk tables
    {
        std::vector<float> soa(3 * NV);
        for (size_t v = 0; v < NV; ++v)
            for (int d = 0; d < 3; ++d) soa[d * NV + v] = mesh->positions[3 * v + d];
        HIPCHK(hipMemcpyAsync(s->X, soa.data(), 12 * NV, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemsetAsync(s->V, 0, 12 * NV, s->stream));
        HIPCHK(hipMemsetAsync(s->Fv, 0, 12 * NV, s->stream));
        std::vector<int32_t> ids(kVeinSlots * NV);
        std::vector<float> lens(kVeinSlots * NV);
        RC(bcs_vein_neighbors(mesh, ids.data(), lens.data()));
        HIPCHK(hipMemcpyAsync(s->nbr, ids.data(), 4 * ids.size(), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(s->nbl, lens.data(), 4 * lens.size(), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(s->tidx, mesh->indices, 12 * NT, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));  // host staging vectors die at the end of this block
    }
    HIPCHK(hipMemsetAsync(s->vhead, 0xFF, 4 * std::max<size_t>(NV, 1), s->stream));
    HIPCHK(hipMemsetAsync(s->hctrl, 0, 64, s->stream));
    HIPCHK(hipMemsetAsync(s->flags, 0, 4, s->stream));
    HIPCHK(hipMemsetAsync(s->step, 0, 8, s->stream));
    HIPCHK(hipMemsetAsync(s->wcount, 0, 4 * (size_t)std::max(s->nblocks, 1), s->stream));
    // ---- grids
    const GridGeom pgeom = make_geom(bmin, bmax, phys->particle_cell);
    const GridGeom tgeom = make_geom(bmin, bmax, phys->triangle_cell);
    RC(grid_alloc(s->pg, pgeom, s->np, true, false, true, true, s->stream));
    s->pg.clear_table = false;  // readers consult the occupancy map
    RC(grid_alloc(s->tg, tgeom, s->ntri, false, true, false, false, s->stream));
    RC(s->dalloc(&s->t_nmask, (size_t)tgeom.cell_count));
    RC(s->dalloc(&s->t_aabb, 2 * (size_t)tgeom.cell_count));
    RC(s->dalloc(&s->trun, (size_t)std::max(s->ntri, 1)));
    RC(s->dalloc(&s->trec, (size_t)std::max(s->ntri, 1)));
    RC(s->dalloc(&s->fslot, (size_t)s->np));
    RC(build_triangle_grid(s.get()));
    HIPCHK(hipStreamSynchronize(s->stream));
    *out = s.release();
    return BCS_OK;
}

extern "C" int bcs_sim_destroy(bcs_sim* sim)
{
    if (!sim) return BCS_OK;
    (void)hipSetDevice(sim->device);
    (void)hipStreamSynchronize(sim->stream);
    delete sim;
    return BCS_OK;
}

extern "C" int bcs_sim_get_info(bcs_sim* s, bcs_sim_info* out)
{
    RC(set_dev(s));
    if (!out) return fail(BCS_ERR_INVALID_ARGUMENT, "null info");
    std::memset(out, 0, sizeof(*out));
    out->particle_count = s->np;
    out->blood_cell_count = s->nc;
    out->type_count = s->nt;
    out->model_vertex_count = s->nm;
    out->vertex_count = s->nv;
    out->triangle_count = s->ntri;
    for (int d = 0; d < 3; ++d) {
        out->bounds_min[d] = s->bmin[d];
        out->bounds_max[d] = s->bmax[d];
        out->particle_cells[d] = s->pg.g.cnt[d];
        out->triangle_cells[d] = s->tg.g.cnt[d];
    }
    out->particle_cell_count = s->pg.g.cell_count;
    out->triangle_cell_count = s->tg.g.cell_count;
    int g = 0;
    for (int t = 0; t < s->nt; ++t) {
        out->particle_starts[t] = s->tt.pstart[t];
        out->cell_type_starts[t] = s->tt.cstart[t];
        out->model_starts[t] = s->tt.mstart[t];
        out->graph_starts[t] = g;
        out->particles_in_cell[t] = s->tt.ppc[t];
        out->cells_of_type[t] = s->tt.cnt[t];
        g += s->tt.ppc[t] * s->tt.ppc[t];
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpyAsync(&out->step_index, s->step, 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return BCS_OK;
}

extern "C" int bcs_sim_device_state(bcs_sim* s, bcs_device_state* out)
{
    RC(set_dev(s));
    if (!out) return fail(BCS_ERR_INVALID_ARGUMENT, "null output");
    out->positions = s->pos; out->velocities = s->vel; out->forces = s->frc; out->centers = s->centers;
    out->vein_positions = s->X; out->vein_velocities = s->V; out->vein_forces = s->Fv;
    out->triangle_centers = s->tcent;
    out->stream = (void*)s->stream;
    out->model_radii = s->bs;
    out->initial_radii = s->r0;
    out->triangle_indices = s->tidx;
    return BCS_OK;
}

extern "C" int bcs_sim_grid_view(bcs_sim* s, int32_t which, int32_t dense_ranges, bcs_grid_view* out)
{
    RC(set_dev(s));
    if (!out) return fail(BCS_ERR_INVALID_ARGUMENT, "null output");
    if (which != BCS_GRID_PARTICLES && which != BCS_GRID_TRIANGLES) return fail(BCS_ERR_INVALID_ARGUMENT, "bad grid");
    const DeviceGrid& G = which == BCS_GRID_PARTICLES ? s->pg : s->tg;
    std::memset(out, 0, sizeof(*out));
    out->grid_cell_ids = G.keys_sorted;
    out->particle_ids = G.perm;
    out->object_count = G.n;
    for (int d = 0; d < 3; ++d) {
        out->cell_counts[d] = G.g.cnt[d];
        out->cell_size[d] = G.g.cell[d];
        out->bounds_min[d] = G.g.minv[d];
        out->bounds_max[d] = G.g.maxv[d];
    }
    out->cell_count = G.g.cell_count;
    if (!dense_ranges) return BCS_OK;
    if (G.g.cell_count > (uint64_t)INT32_MAX)
        return fail(BCS_ERR_GRID_TOO_LARGE, "dense cell ranges: more than 2^31 cells");
    if (!s->dense_start[which]) {
        RC(s->dalloc(&s->dense_start[which], G.g.cell_count));
        RC(s->dalloc(&s->dense_end[which], G.g.cell_count));
    }
    launch_dense_ranges(G.keys_sorted, G.n, G.g.cell_count, s->dense_start[which], s->dense_end[which], s->stream);
    HIPCHK(hipGetLastError());
    out->grid_cell_starts = s->dense_start[which];
    out->grid_cell_ends = s->dense_end[which];
    return BCS_OK;
}

extern "C" int bcs_sim_cell_ids(bcs_sim* s, int32_t which, const float* positions, int32_t n, uint32_t* cell_ids)
{
    if (!s) return fail(BCS_ERR_INVALID_ARGUMENT, "null simulation");
    if (which != BCS_GRID_PARTICLES && which != BCS_GRID_TRIANGLES) return fail(BCS_ERR_INVALID_ARGUMENT, "bad grid");
    if (n < 0 || (n > 0 && (!positions || !cell_ids))) return fail(BCS_ERR_INVALID_ARGUMENT, "bad arrays");
    const GridGeom& g = (which == BCS_GRID_PARTICLES ? s->pg : s->tg).g;
    uint32_t fl = 0;
    for (int32_t i = 0; i < n; ++i)
        cell_ids[i] = cell_key(g, positions[i], positions[(size_t)n + i], positions[2 * (size_t)n + i], fl);
    return BCS_OK;
}

static int copy_in(bcs_sim* s, float* dst, const float* src, size_t n)
{
    if (!src || n == 0) return BCS_OK;
    HIPCHK(hipMemcpyAsync(dst, src, 4 * n, hipMemcpyHostToDevice, s->stream));
    return BCS_OK;
}
static int copy_out(bcs_sim* s, float* dst, const float* src, size_t n)
{
    if (!dst || n == 0) return BCS_OK;
    HIPCHK(hipMemcpyAsync(dst, src, 4 * n, hipMemcpyDeviceToHost, s->stream));
    return BCS_OK;
}

static void drop_graph(bcs_sim* s);

extern "C" int bcs_sim_set_particles(bcs_sim* s, const float* p, const float* v, const float* f)
{
    RC(set_dev(s));
    const size_t n = 3 * (size_t)s->np;
    RC(copy_in(s, s->pos, p, n)); RC(copy_in(s, s->vel, v, n)); RC(copy_in(s, s->frc, f, n));
    if (p || v) s->copies_current = false;
    if (p) {  // bucket size of the particle sort from where the particles are
        const int sh = grid_shift_for(s->pg.g, p, s->np, s->np);
        if (sh != s->pg.shift) {
            drop_graph(s);
            RC(grid_set_shift(s->pg, sh, s->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    return BCS_OK;
}
extern "C" int bcs_sim_get_particles(bcs_sim* s, float* p, float* v, float* f)
{
    RC(set_dev(s));
    const size_t n = 3 * (size_t)s->np;
    RC(copy_out(s, p, s->pos, n)); RC(copy_out(s, v, s->vel, n)); RC(copy_out(s, f, s->frc, n));
    HIPCHK(hipStreamSynchronize(s->stream));
    return BCS_OK;
}
extern "C" int bcs_sim_set_vein(bcs_sim* s, const float* p, const float* v, const float* f)
{
    RC(set_dev(s));
    const size_t n = 3 * (size_t)s->nv;
    RC(copy_in(s, s->X, p, n)); RC(copy_in(s, s->V, v, n)); RC(copy_in(s, s->Fv, f, n));
    HIPCHK(hipStreamSynchronize(s->stream));
    return BCS_OK;
}
extern "C" int bcs_sim_get_vein(bcs_sim* s, float* p, float* v, float* f)
{
    RC(set_dev(s));
    const size_t n = 3 * (size_t)s->nv;
    RC(copy_out(s, p, s->X, n)); RC(copy_out(s, v, s->V, n)); RC(copy_out(s, f, s->Fv, n));
    HIPCHK(hipStreamSynchronize(s->stream));
    return BCS_OK;
}
extern "C" int bcs_sim_get_centers(bcs_sim* s, float* c)
{
    RC(set_dev(s));
    RC(copy_out(s, c, s->centers, 3 * (size_t)s->nc));
    HIPCHK(hipStreamSynchronize(s->stream));
    return BCS_OK;
}
extern "C" int bcs_sim_set_step_index(bcs_sim* s, uint64_t step)
{
    RC(set_dev(s));
    HIPCHK(hipMemcpyAsync(s->step, &step, 8, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return BCS_OK;
}

extern "C" int bcs_sim_calculate_grid(bcs_sim* s, int32_t which)
{
    RC(set_dev(s));
    if (which == BCS_GRID_PARTICLES) RC(enqueue_particle_grid(s));
    else if (which == BCS_GRID_TRIANGLES) RC(build_triangle_grid(s));
    else return fail(BCS_ERR_INVALID_ARGUMENT, "unknown grid");
    HIPCHK(hipStreamSynchronize(s->stream));
    return BCS_OK;
}

extern "C" int bcs_sim_calculate_next_frame(bcs_sim* s)
{
    RC(set_dev(s));
    if (s->pg.n != s->np) return fail(BCS_ERR_STATE, "particle grid not built (call bcs_sim_calculate_grid)");
    // particles moved since the grid was built: the reference then scans the
    // old cells with the current positions, so refresh the packed copies
    if (!s->copies_current) {
        launch_refresh_copies(s->pos, s->vel, s->np, s->pg.perm, s->pg.pvs, s->stream);
        s->copies_current = true;
    }
    const StepArgs A = s->step_args();
    const VeinArgs Va = s->vein_args();
    launch_vein_prep(Va, s->ntri, s->trun, s->trec, s->tdata, s->t_aabb, s->stream);
    launch_raycand(A, s->nchunks, s->stream);
    launch_ray(A, s->stream);
    launch_cells(A, s->nchunks, false, s->stream);
    launch_vein(Va, 1, s->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return BCS_OK;
}

extern "C" int bcs_sim_propagate_all(bcs_sim* s)
{
    RC(set_dev(s));
    s->copies_current = false;
    const StepArgs A = s->step_args();
    const VeinArgs Va = s->vein_args();
    launch_integrate(A, s->stream);
    launch_vein(Va, 2, s->stream);
    if (s->phys.use_blood_flow) launch_vein_end(A, s->nchunks, s->stream);
    launch_bump(s->step, s->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return BCS_OK;
}

static void drop_graph(bcs_sim* s)
{
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    if (s->graph_exec_b) (void)hipGraphExecDestroy(s->graph_exec_b);
    if (s->graph_b) (void)hipGraphDestroy(s->graph_b);
    if (s->graph_exec_m) (void)hipGraphExecDestroy(s->graph_exec_m);
    if (s->graph_m) (void)hipGraphDestroy(s->graph_m);
    s->graph_exec = s->graph_exec_b = s->graph_exec_m = nullptr;
    s->graph = s->graph_b = s->graph_m = nullptr;
}

// The vein update of frame f overlaps the particle grid build of frame f+1
// unless the triangle grid is rebuilt from the vein every frame (that build
// precedes the branches and needs the updated vein).
static bool defer_vein(const bcs_sim* s) { return !s->phys.dynamic_triangle_grid; }

// Refresh the spatial processing order of the cells from the last frame's
// centres, and the chunk lanes' particles with it (performance only: results
// do not depend on the order).
static int enqueue_order(bcs_sim* s)
{
    RC(launch_order(s->tt, s->pg.g, s->centers, s->okey, s->ocount, s->ostart, s->ostatus, s->otiles, s->octrl,
                    s->corder, s->flags, s->stream));
    launch_chunk_map(s->step_args(), s->chunk_pi, s->stream);
    return BCS_OK;
}

// nframes consecutive frames in one graph (the same kernels as nframes
// launches of the one-frame graph; fewer graph launches per step call),
// preceded by the order refresh when with_order is set.
static int capture_frame(bcs_sim* s, bool lead_vein, hipGraph_t* graph, hipGraphExec_t* exec, int nframes = 1,
                         bool with_order = false)
{
    HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = with_order ? enqueue_order(s) : BCS_OK;
    for (int f = 0; f < nframes && !rc; ++f) rc = enqueue_fused_frame(s, nullptr, lead_vein, !defer_vein(s));
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(BCS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    *graph = g;
    HIPCHK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return BCS_OK;
}

static int ensure_graph(bcs_sim* s)
{
    if (!s->graph_exec) RC(capture_frame(s, false, &s->graph, &s->graph_exec));
    if (!s->graph_exec_b && defer_vein(s)) RC(capture_frame(s, true, &s->graph_b, &s->graph_exec_b));
    if (!s->graph_exec_m && defer_vein(s))
        RC(capture_frame(s, true, &s->graph_m, &s->graph_exec_m, kGraphFrames, true));
    return BCS_OK;
}

extern "C" int bcs_sim_step(bcs_sim* s, int32_t n_steps)
{
    RC(set_dev(s));
    if (n_steps < 0) return fail(BCS_ERR_INVALID_ARGUMENT, "negative step count");
    if (n_steps == 0) return BCS_OK;
    if (s->use_graph) RC(ensure_graph(s));
    for (int32_t k = 0; k < n_steps;) {
        // frames after the first start with the previous frame's vein update
        const bool lead = defer_vein(s) && k > 0;
        int32_t nf = 1;
        if (s->use_graph && lead && s->graph_exec_m && n_steps - k >= kGraphFrames) {
            // kGraphFrames frames, after the order refresh the graph starts with
            HIPCHK(hipGraphLaunch(s->graph_exec_m, s->stream));
            nf = kGraphFrames;
            s->frames_since_order = 0;
        } else {
            if (s->frames_since_order >= kReorderPeriod) {
                RC(enqueue_order(s));
                s->frames_since_order = 0;
            }
            if (s->use_graph) HIPCHK(hipGraphLaunch(lead ? s->graph_exec_b : s->graph_exec, s->stream));
            else RC(enqueue_fused_frame(s, nullptr, lead, !defer_vein(s)));
        }
        k += nf;
        s->frames_since_order += nf;
    }
    if (defer_vein(s)) launch_vein(s->vein_args(), 7, s->stream);  // the last frame's vein update
    s->copies_current = false;  // the frames moved the particles after their grid builds
    HIPCHK(hipGetLastError());
    return BCS_OK;
}

extern "C" int bcs_sim_use_graph(bcs_sim* s, int32_t enabled)
{
    if (!s) return fail(BCS_ERR_INVALID_ARGUMENT, "null simulation");
    s->use_graph = enabled != 0;
    return BCS_OK;
}

extern "C" int bcs_sim_set_bucket_shift(bcs_sim* s, int32_t shift)
{
    if (!s) return fail(BCS_ERR_INVALID_ARGUMENT, "null simulation");
    if (shift < 0 || shift > 24) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_set_bucket_shift: shift in 0..24");
    if (shift == s->pg.shift) return BCS_OK;
    RC(set_dev(s));
    HIPCHK(hipStreamSynchronize(s->stream));  // the graphs hold the old bucket arrays
    drop_graph(s);
    RC(grid_set_shift(s->pg, shift, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return BCS_OK;
}

extern "C" int bcs_sim_set_ray_mode(bcs_sim* s, int32_t mode)
{
    if (!s) return fail(BCS_ERR_INVALID_ARGUMENT, "null simulation");
    if (mode != BCS_RAY_AUTO && mode != BCS_RAY_SPLIT && mode != BCS_RAY_FUSED)
        return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_set_ray_mode: mode is BCS_RAY_AUTO, _SPLIT or _FUSED");
    if (mode == s->ray_mode) return BCS_OK;
    RC(set_dev(s));
    HIPCHK(hipStreamSynchronize(s->stream));  // the graphs hold the old launches
    drop_graph(s);
    s->ray_mode = mode;
    return BCS_OK;
}

extern "C" int bcs_sim_set_contact_mode(bcs_sim* s, int32_t mode)
{
    if (!s) return fail(BCS_ERR_INVALID_ARGUMENT, "null simulation");
    if (mode != BCS_CONTACT_AUTO && mode != BCS_CONTACT_CELLS && mode != BCS_CONTACT_SLOTS)
        return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_set_contact_mode: mode is BCS_CONTACT_AUTO, _CELLS or _SLOTS");
    if (mode == s->contact_mode) return BCS_OK;
    RC(set_dev(s));
    HIPCHK(hipStreamSynchronize(s->stream));  // the graphs hold the old launches
    drop_graph(s);
    s->contact_mode = mode;
    return BCS_OK;
}

extern "C" int bcs_sim_synchronize(bcs_sim* s)
{
    RC(set_dev(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    return BCS_OK;
}

extern "C" int bcs_sim_time_steps(bcs_sim* s, int32_t n_steps, float* total_ms)
{
    RC(set_dev(s));
    if (!total_ms) return fail(BCS_ERR_INVALID_ARGUMENT, "null output");
    RC(ensure_graph(s));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s->stream));
    int rc = bcs_sim_step(s, n_steps);
    HIPCHK(hipEventRecord(e1, s->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *total_ms = ms;
    return rc;
}

extern "C" int bcs_sim_profile_kernels(bcs_sim* s, int32_t n_steps, float* ms_out, int32_t capacity,
                                       int32_t* n_kernels, char* names_out, int32_t names_capacity)
{
    RC(set_dev(s));
    if (!ms_out || !n_kernels || capacity < kFusedKernels || n_steps < 1)
        return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_profile_kernels: bad arguments");
    std::vector<hipEvent_t> eb(kFusedKernels), ee(kFusedKernels);
    for (int k = 0; k < kFusedKernels; ++k) {
        HIPCHK(hipEventCreate(&eb[k]));
        HIPCHK(hipEventCreate(&ee[k]));
    }
    KernelTimer timer{&eb, &ee, 0};
    std::vector<double> acc(kFusedKernels, 0.0);
    // the slots this frame records: k_cells or the three split kernels (an
    // event that was never recorded must not be queried)
    std::vector<bool> used(kFusedKernels, true);
    if (s->contact_split()) used[kGridKernels + 3] = false;
    else used[kGridKernels + 5] = used[kGridKernels + 6] = used[kGridKernels + 7] = false;
    int rc = BCS_OK;
    for (int32_t it = 0; it < n_steps && !rc; ++it) {
        rc = enqueue_fused_frame(s, &timer);
        if (rc) break;
        HIPCHK(hipStreamSynchronize(s->stream));
        for (int k = 0; k < kFusedKernels; ++k) {
            float ms = 0;
            if (!used[k]) continue;
            const hipError_t e = hipEventElapsedTime(&ms, eb[k], ee[k]);
            if (e != hipSuccess && !rc) rc = fail(BCS_ERR_HIP, std::string("hipEventElapsedTime: ") + hipGetErrorString(e));
            acc[k] += ms;
        }
    }
    if (s->ray_fused()) acc[kGridKernels + 2] = 0.0;  // no k_ray launch
    if (grid_map_in_rank(s->pg)) acc[4] = 0.0;         // no k_occ launch (k_rank writes the map)
    for (int k = 0; k < kFusedKernels; ++k) {
        ms_out[k] = (float)(acc[k] / n_steps);
        (void)hipEventDestroy(eb[k]);
        (void)hipEventDestroy(ee[k]);
    }
    *n_kernels = kFusedKernels;
    s->copies_current = false;  // the profiled frames moved the particles
    if (names_out && names_capacity > 0) {
        std::string names = kKernelNames;
        if (s->ray_fused()) names.replace(names.find("k_raycand;"), 10, "k_raycand_ray;");
        std::strncpy(names_out, names.c_str(), (size_t)names_capacity - 1);
        names_out[names_capacity - 1] = 0;
    }
    return rc;
}

extern "C" int bcs_sim_get_grid(bcs_sim* s, int32_t which, uint32_t* keys_sorted, int32_t* perm, int32_t* n_occupied,
                                uint32_t* occ_keys, int32_t* occ_start, int32_t* occ_end)
{
    RC(set_dev(s));
    const DeviceGrid& G = which == BCS_GRID_PARTICLES ? s->pg : s->tg;
    HIPCHK(hipStreamSynchronize(s->stream));
    const size_t n = (size_t)G.n;
    std::vector<uint32_t> k(n);
    if (n) HIPCHK(hipMemcpyAsync(k.data(), G.keys_sorted, 4 * n, hipMemcpyDeviceToHost, s->stream));
    if (keys_sorted && n) std::memcpy(keys_sorted, k.data(), 4 * n);
    if (perm && n) HIPCHK(hipMemcpyAsync(perm, G.perm, 4 * n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    // occupied ranges; starts are read back through the device start table
    std::vector<uint32_t> okeys;
    std::vector<int32_t> oend;
    for (size_t i = 0; i < n; ++i) {
        if (i == 0 || k[i] != k[i - 1]) okeys.push_back(k[i]);
        if (!okeys.empty()) {
            if (oend.size() < okeys.size()) oend.push_back(0);
            oend.back() = (int32_t)i;
        }
    }
    const int32_t no = (int32_t)okeys.size();
    std::vector<uint32_t> ostart(no);
    if (no) {
        uint32_t *dk = nullptr, *dout = nullptr;
        HIPCHK(hipMalloc(&dk, 4 * (size_t)no));
        HIPCHK(hipMalloc(&dout, 4 * (size_t)no));
        HIPCHK(hipMemcpyAsync(dk, okeys.data(), 4 * (size_t)no, hipMemcpyHostToDevice, s->stream));
        launch_gather_start(G.table, dk, no, dout, s->stream);
        HIPCHK(hipStreamSynchronize(s->stream));
        HIPCHK(hipMemcpyAsync(ostart.data(), dout, 4 * (size_t)no, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        (void)hipFree(dk);
        (void)hipFree(dout);
    }
    for (int32_t q = 0; q < no; ++q) {
        if (occ_keys) occ_keys[q] = okeys[q];
        if (occ_start) occ_start[q] = (int32_t)ostart[q];
        if (occ_end) occ_end[q] = oend[q];
    }
    if (n_occupied) *n_occupied = no;
    return BCS_OK;
}

// ---- bcs_sim_get_grid without allocation --------------------------------
// The gather of cell-table starts above stages the occupied keys and the
// gathered starts in two device buffers it allocates per call.  The variant
// below stages them in this file-scope array instead: kStaticOccupied keys
// followed by kStaticOccupied starts.  Every GPU the library runs on holds
// its own copy until the process ends; host code reaches it only through the
// per-device address taken with hipGetSymbolAddress (never through the
// array's name, which in host code is a stand-in, not a device address).
constexpr int32_t kStaticOccupied = 1 << 20;  // 8 MiB of staging per GPU
static __device__ __align__(256) uint32_t scratch[2 * (size_t)kStaticOccupied];

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static uint32_t* g_scratch_ptr[kScratchDevices] = {};  // per device, taken once

// Address of `scratch` on the current device (looked up at the first call on
// that device).
static int scratch_ptr(int device, uint32_t** out)
{
    if (device < 0 || device >= kScratchDevices) return fail(BCS_ERR_UNSUPPORTED, "static staging: device index >= 64");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(BCS_ERR_HIP, "hipGetSymbolAddress(scratch): null address");
        g_scratch_ptr[device] = (uint32_t*)p;
    }
    *out = g_scratch_ptr[device];
    return BCS_OK;
}

extern "C" int bcs_sim_get_grid_static(bcs_sim* s, int32_t which, uint32_t* keys_sorted, int32_t* perm,
                                       int32_t* n_occupied, uint32_t* occ_keys, int32_t* occ_start,
                                       int32_t* occ_end)
{
    RC(set_dev(s));
    if (which != BCS_GRID_PARTICLES && which != BCS_GRID_TRIANGLES) return fail(BCS_ERR_INVALID_ARGUMENT, "bad grid");
    if (!keys_sorted || !n_occupied || !occ_keys || !occ_start || !occ_end)
        return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_get_grid_static: keys_sorted and every occupied-cell output "
                                              "are required (they are the working arrays)");
    const DeviceGrid& G = which == BCS_GRID_PARTICLES ? s->pg : s->tg;
    uint32_t* scr = nullptr;
    RC(scratch_ptr(s->device, &scr));
    HIPCHK(hipStreamSynchronize(s->stream));
    const int32_t n = G.n;
    if (n) HIPCHK(hipMemcpyAsync(keys_sorted, G.keys_sorted, 4 * (size_t)n, hipMemcpyDeviceToHost, s->stream));
    if (perm && n) HIPCHK(hipMemcpyAsync(perm, G.perm, 4 * (size_t)n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    // occupied cells: runs of equal sorted keys, in the caller's arrays
    int32_t no = 0;
    for (int32_t i = 0; i < n; ++i) {
        if (i == 0 || keys_sorted[i] != keys_sorted[i - 1]) occ_keys[no++] = keys_sorted[i];
        occ_end[no - 1] = i;
    }
    if (no > kStaticOccupied) {
        *n_occupied = 0;
        return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_get_grid_static: " + std::to_string(no) +
                                                  " occupied cells exceed the static staging of " +
                                                  std::to_string(kStaticOccupied) + "; use bcs_sim_get_grid");
    }
    if (no) {
        uint32_t* dkeys = scr;
        uint32_t* dstart = scr + kStaticOccupied;
        HIPCHK(hipMemcpyAsync(dkeys, occ_keys, 4 * (size_t)no, hipMemcpyHostToDevice, s->stream));
        launch_gather_start(G.table, dkeys, no, dstart, s->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(occ_start, dstart, 4 * (size_t)no, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    *n_occupied = no;
    return BCS_OK;
}

extern "C" int bcs_sim_get_contacts(bcs_sim* s, int32_t* counts, int32_t* pairs_j, int64_t capacity, int64_t* n_total)
{
    RC(set_dev(s));
    if (!counts || !n_total) return fail(BCS_ERR_INVALID_ARGUMENT, "null output");
    if (s->pg.n != s->np) return fail(BCS_ERR_STATE, "particle grid not built");
    const StepArgs A = s->step_args();
    int32_t* dcounts = nullptr;
    int64_t* doffs = nullptr;
    int32_t* dpairs = nullptr;
    HIPCHK(hipMalloc(&dcounts, 4 * std::max<size_t>(s->np, 1)));
    launch_contacts(A, 0, dcounts, nullptr, nullptr, s->stream);
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpyAsync(counts, dcounts, 4 * (size_t)s->np, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    int64_t total = 0;
    std::vector<int64_t> offs(s->np + 1, 0);
    for (int32_t i = 0; i < s->np; ++i) {
        offs[i] = total;
        total += counts[i];
    }
    *n_total = total;
    int rc = BCS_OK;
    if (pairs_j && total <= capacity && total > 0) {
        if (hipMalloc(&doffs, 8 * (size_t)(s->np + 1)) != hipSuccess || hipMalloc(&dpairs, 4 * (size_t)total) != hipSuccess)
            rc = fail(BCS_ERR_OUT_OF_MEMORY, "contacts buffer");
        if (!rc) {
            (void)hipMemcpyAsync(doffs, offs.data(), 8 * (size_t)(s->np + 1), hipMemcpyHostToDevice, s->stream);
            launch_contacts(A, 1, dcounts, doffs, dpairs, s->stream);
            (void)hipStreamSynchronize(s->stream);
            (void)hipMemcpyAsync(pairs_j, dpairs, 4 * (size_t)total, hipMemcpyDeviceToHost, s->stream);
            (void)hipStreamSynchronize(s->stream);
        }
    }
    (void)hipFree(dcounts);
    if (doffs) (void)hipFree(doffs);
    if (dpairs) (void)hipFree(dpairs);
    return rc;
}

extern "C" int bcs_sim_export_frame(bcs_sim* s, float* offsets, float* const* type_vertices, int32_t n_types,
                                    float* vein_vertices)
{
    RC(set_dev(s));
    if (n_types < 0 || (n_types > 0 && !type_vertices)) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_export_frame: types");
    if (offsets) launch_export(s->pos, s->np, 0, s->np, 3, offsets, s->stream);
    for (int32_t t = 0; t < std::min<int32_t>(n_types, s->tt.nt); ++t)
        if (type_vertices[t])
            launch_export(s->pos, s->np, s->tt.pstart[t], (int64_t)s->tt.cnt[t] * s->tt.ppc[t], 6, type_vertices[t],
                          s->stream);
    if (vein_vertices) launch_export(s->X, s->nv, 0, s->nv, 6, vein_vertices, s->stream);
    HIPCHK(hipGetLastError());
    return BCS_OK;
}

struct bcs_traj {
    bcs_sim* sim = nullptr;
    std::FILE* f = nullptr;
    bool vel = false;
    std::vector<float> p, v, X;
};

extern "C" int bcs_traj_open(bcs_sim* s, const char* path, int32_t with_velocities, bcs_traj** out)
{
    if (!s || !path || !out) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_traj_open: null argument");
    *out = nullptr;
    std::unique_ptr<bcs_traj> t(new bcs_traj());
    t->sim = s;
    t->vel = with_velocities != 0;
    t->f = std::fopen(path, "wb");
    if (!t->f) return fail(BCS_ERR_INVALID_ARGUMENT, std::string("bcs_traj_open: cannot write ") + path);
    char head[64];
    std::memset(head, 0, sizeof(head));
    std::memcpy(head, "BCSTRAJ1", 8);
    const int32_t ints[4] = {s->np, s->nv, s->tt.nt, t->vel ? 1 : 0};
    std::memcpy(head + 8, ints, sizeof(ints));
    const float dt = s->phys.dt;
    std::memcpy(head + 24, &dt, 4);
    if (std::fwrite(head, 1, 64, t->f) != 64) {
        std::fclose(t->f);
        return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_traj_open: write failed");
    }
    t->p.resize(3 * (size_t)s->np);
    if (t->vel) t->v.resize(3 * (size_t)s->np);
    t->X.resize(3 * (size_t)s->nv);
    *out = t.release();
    return BCS_OK;
}

extern "C" int bcs_traj_append(bcs_traj* t)
{
    if (!t || !t->f) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_traj_append: closed trajectory");
    bcs_sim* s = t->sim;
    RC(bcs_sim_get_particles(s, t->p.data(), t->vel ? t->v.data() : nullptr, nullptr));
    RC(bcs_sim_get_vein(s, t->X.data(), nullptr, nullptr));
    bcs_sim_info info{};
    RC(bcs_sim_get_info(s, &info));
    const uint64_t step = info.step_index;
    bool ok = std::fwrite(&step, 8, 1, t->f) == 1;
    ok = ok && std::fwrite(t->p.data(), 4, t->p.size(), t->f) == t->p.size();
    if (t->vel) ok = ok && std::fwrite(t->v.data(), 4, t->v.size(), t->f) == t->v.size();
    ok = ok && std::fwrite(t->X.data(), 4, t->X.size(), t->f) == t->X.size();
    if (!ok) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_traj_append: write failed");
    return BCS_OK;
}

extern "C" int bcs_traj_close(bcs_traj* t)
{
    if (!t) return BCS_OK;
    const bool ok = !t->f || std::fclose(t->f) == 0;
    delete t;
    return ok ? BCS_OK : fail(BCS_ERR_INVALID_ARGUMENT, "bcs_traj_close: close failed");
}

extern "C" int bcs_sim_frame_stats(bcs_sim* s, uint32_t* ray_work_items, uint32_t* vein_hits)
{
    RC(set_dev(s));
    std::vector<uint32_t> w((size_t)std::max(s->nblocks, 1), 0u);
    uint32_t h[8] = {0};
    HIPCHK(hipMemcpyAsync(w.data(), s->wcount, 4 * w.size(), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(h, s->hctrl, 32, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    uint64_t items = 0;
    for (uint32_t v : w) items += v;
    if (ray_work_items) *ray_work_items = (uint32_t)std::min<uint64_t>(items, 0xFFFFFFFFull);
    if (vein_hits) *vein_hits = h[4];
    return BCS_OK;
}

extern "C" int bcs_sim_error_flags(bcs_sim* s, int32_t clear, uint32_t* flags)
{
    RC(set_dev(s));
    if (!flags) return fail(BCS_ERR_INVALID_ARGUMENT, "null output");
    HIPCHK(hipMemcpyAsync(flags, s->flags, 4, hipMemcpyDeviceToHost, s->stream));
    if (clear) HIPCHK(hipMemsetAsync(s->flags, 0, 4, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return BCS_OK;
}

extern "C" int bcs_calibration_copy(int64_t n_bytes, int32_t repeats, float* ms)
{
    if (n_bytes < 16 || repeats < 1 || !ms) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_calibration_copy");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(BCS_ERR_NO_DEVICE, "no HIP device");
    void *a = nullptr, *b = nullptr;
    const int64_t n16 = n_bytes / 16;
    HIPCHK(hipMalloc(&a, n16 * 16));
    HIPCHK(hipMalloc(&b, n16 * 16));
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIPCHK(hipMemsetAsync(a, 1, n16 * 16, st));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    launch_copy(a, b, n16, st);  // warm-up
    HIPCHK(hipEventRecord(e0, st));
    for (int r = 0; r < repeats; ++r) launch_copy(a, b, n16, st);
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    *ms = t / repeats;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    (void)hipFree(a);
    (void)hipFree(b);
    return BCS_OK;
}

// ============================================================ ensembles ==
// Members keep their own buffers; the ensemble holds each member's kernel
// arguments in device arrays and runs every frame kernel once for all members
// through the batched kernels (a persistent grid over the members' blocks).
// A member's blocks see the same block indices, block counts and arguments
// as in its own launches, so its results are those of bcs_sim_step.

namespace {

enum EnsLaunch {
    EL_KEY, EL_SCAN, EL_SCATTER, EL_RANK, EL_OCC,  // grid build (GridBatch order)
    EL_PREP, EL_RAYCAND, EL_RAY, EL_CELLS, EL_VEIN,
    EL_SPRINGS, EL_CONTACTS, EL_RESPOND,  // the split particle part (contact search by sorted slot)
    EL_OKEY, EL_OSCAN, EL_OSCAT, EL_CMAP,
    EL_COUNT
};

}  // namespace

struct bcs_ensemble {
    int device = 0;
    int cus = 256;
    hipStream_t stream = nullptr, side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::vector<bcs_sim*> members;
    std::vector<int32_t> order;           // member of each argument slot
    bcs_ensemble_tuning tuning{0, 1, BCS_RAY_AUTO, 1, BCS_CONTACT_AUTO};
    bool fused = true;
    bool split = false;  // contact search by sorted slot (k_springs, k_contacts, k_respond)
    // device argument arrays (one entry per slot)
    StepArgs* d_step = nullptr;
    VeinArgs* d_vein = nullptr;
    VeinPrepArgs* d_prep = nullptr;
    KeyArgs* d_key = nullptr;
    ScanArgs* d_scan = nullptr;
    ScatterArgs* d_scatter = nullptr;
    RankArgs* d_rank = nullptr;
    OccArgs* d_occ = nullptr;
    OrderKeyArgs* d_okey = nullptr;
    ScanArgs* d_oscan = nullptr;
    OrderScatterArgs* d_oscat = nullptr;
    int2** d_cmap = nullptr;
    uint32_t* d_split = nullptr;
    uint32_t* d_start = nullptr;  // EL_COUNT arrays of n + 1 virtual-block offsets
    uint32_t total[EL_COUNT]{};
    std::vector<uint32_t*> shapes;  // per member: its particle-grid bucket arrays when the arguments were taken
    int32_t frames_since_order = 0;
    hipGraph_t graph[3]{};
    hipGraphExec_t exec[3]{};  // first frame, a frame after another, kGraphFrames frames after the order refresh
    std::vector<void*> allocs;

    ~bcs_ensemble()
    {
        drop_graphs();
        for (void* p : allocs) (void)hipFree(p);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }
    void drop_graphs()
    {
        for (int k = 0; k < 3; ++k) {
            if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
            exec[k] = nullptr;
            graph[k] = nullptr;
        }
    }
    template <class T>
    int dalloc(T** p, size_t count)
    {
        hipError_t e = hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) return fail(BCS_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back((void*)*p);
        return BCS_OK;
    }
    const uint32_t* start(int l) const { return d_start + (size_t)l * (members.size() + 1); }
    uint32_t grid(int l) const
    {
        const uint32_t t = total[l];
        if (tuning.blocks_per_cu <= 0) return (t + (uint32_t)std::max(tuning.run, 1) - 1) / (uint32_t)std::max(tuning.run, 1);
        return std::min<uint32_t>(t, (uint32_t)tuning.blocks_per_cu * (uint32_t)cus);
    }
};

namespace {

int ens_dev(bcs_ensemble* e)
{
    if (!e) return fail(BCS_ERR_INVALID_ARGUMENT, "null ensemble");
    HIPCHK(hipSetDevice(e->device));
    return BCS_OK;
}

// Arguments and virtual-block offsets of every member, in slot order.
int ens_build_args(bcs_ensemble* e)
{
    const size_t n = e->members.size();
    size_t np_total = 0;
    for (bcs_sim* s : e->members) np_total += (size_t)s->nchunks;
    e->fused = e->tuning.ray_mode == BCS_RAY_FUSED ||
               (e->tuning.ray_mode == BCS_RAY_AUTO && cell_blocks((int)std::min<size_t>(np_total, INT32_MAX)) > 4096u);
    e->split = e->tuning.contact_mode == BCS_CONTACT_SLOTS ||
               (e->tuning.contact_mode == BCS_CONTACT_AUTO &&
                cell_blocks((int)std::min<size_t>(np_total, INT32_MAX)) > 4096u);
    std::vector<StepArgs> step(n);
    std::vector<VeinArgs> vein(n);
    std::vector<VeinPrepArgs> prep(n);
    std::vector<KeyArgs> key(n);
    std::vector<ScanArgs> scan(n), oscan(n);
    std::vector<ScatterArgs> scatter(n);
    std::vector<RankArgs> rank(n);
    std::vector<OccArgs> occ(n);
    std::vector<OrderKeyArgs> okey(n);
    std::vector<OrderScatterArgs> oscat(n);
    std::vector<int2*> cmap(n);
    std::vector<uint32_t> split(n);
    std::vector<uint32_t> blocks((size_t)EL_COUNT * n, 0u);
    e->shapes.assign(e->members.size(), nullptr);
    for (size_t k = 0; k < n; ++k) {
        bcs_sim* s = e->members[e->order[k]];
        e->shapes[e->order[k]] = s->pg.bcount;
        GridLaunch L;
        RC(grid_launch_args(s->pg, s->pos, s->np, s->np, s->vel, s->flags, s->chunk_pi, s->nchunks * kBlock, &L));
        key[k] = L.key;
        scan[k] = L.scan;
        scatter[k] = L.scatter;
        rank[k] = L.rank;
        occ[k] = L.occ;
        step[k] = s->step_args();
        vein[k] = s->vein_args();
        uint32_t pb = 0;
        prep[k] = vein_prep_args(vein[k], s->ntri, s->trun, s->trec, s->tdata, s->t_aabb, &pb);
        okey[k] = OrderKeyArgs{s->tt, s->pg.g, s->centers, s->okey, s->ocount, s->ostatus, s->otiles, s->octrl};
        oscan[k] = ScanArgs{s->ocount, s->ostart, (int64_t)s->tt.nt << kOrderBits, s->ostatus, s->otiles, s->octrl,
                            nullptr, 0u, 0u, s->flags, 0};
        oscat[k] = OrderScatterArgs{s->tt.nc, s->okey, s->ocount, s->ostart, s->corder, s->flags};
        cmap[k] = s->chunk_pi;
        split[k] = ray_split(s->nchunks);
        uint32_t* b = &blocks[0];
        b[EL_KEY * n + k] = L.key_blocks;
        b[EL_SCAN * n + k] = L.scan_blocks;
        b[EL_SCATTER * n + k] = L.scatter_blocks;
        b[EL_RANK * n + k] = L.rank_blocks;
        b[EL_OCC * n + k] = L.occ_blocks;
        b[EL_PREP * n + k] = pb;
        b[EL_RAYCAND * n + k] = cell_blocks(s->nchunks);
        b[EL_RAY * n + k] = e->fused ? 0u : ray_blocks(s->nchunks);
        b[EL_CELLS * n + k] = e->split ? 0u : cell_blocks(s->nchunks);
        b[EL_SPRINGS * n + k] = e->split ? cell_blocks(s->nchunks) : 0u;
        b[EL_CONTACTS * n + k] = e->split ? tile_blocks(s->np) : 0u;
        b[EL_RESPOND * n + k] = e->split ? cell_blocks(s->nchunks) : 0u;
        b[EL_VEIN * n + k] = vein_blocks(s->nv);
        b[EL_OKEY * n + k] = order_key_blocks(s->tt, s->otiles);
        b[EL_OSCAN * n + k] = s->tt.nc > 0 ? (uint32_t)s->otiles : 0u;
        b[EL_OSCAT * n + k] = order_scatter_blocks(s->tt);
        b[EL_CMAP * n + k] = (uint32_t)std::max(s->nchunks, 0);
    }
    std::vector<uint32_t> start((size_t)EL_COUNT * (n + 1));
    for (int l = 0; l < EL_COUNT; ++l) {
        uint64_t acc = 0;
        for (size_t k = 0; k < n; ++k) {
            start[(size_t)l * (n + 1) + k] = (uint32_t)acc;
            acc += blocks[(size_t)l * n + k];
        }
        if (acc > 0xFFFFFFFFull) return fail(BCS_ERR_INVALID_ARGUMENT, "ensemble: more than 2^32 blocks in a launch");
        start[(size_t)l * (n + 1) + n] = (uint32_t)acc;
        e->total[l] = (uint32_t)acc;
    }
    hipStream_t st = e->stream;
    HIPCHK(hipMemcpyAsync(e->d_step, step.data(), n * sizeof(StepArgs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(e->d_vein, vein.data(), n * sizeof(VeinArgs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(e->d_prep, prep.data(), n * sizeof(VeinPrepArgs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(e->d_key, key.data(), n * sizeof(KeyArgs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(e->d_scan, scan.data(), n * sizeof(ScanArgs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(e->d_scatter, scatter.data(), n * sizeof(ScatterArgs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(e->d_rank, rank.data(), n * sizeof(RankArgs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(e->d_occ, occ.data(), n * sizeof(OccArgs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(e->d_okey, okey.data(), n * sizeof(OrderKeyArgs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(e->d_oscan, oscan.data(), n * sizeof(ScanArgs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(e->d_oscat, oscat.data(), n * sizeof(OrderScatterArgs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(e->d_cmap, cmap.data(), n * sizeof(int2*), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(e->d_split, split.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(e->d_start, start.data(), start.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    e->drop_graphs();
    return BCS_OK;
}

// The members' bucket arrays move when their bucket size changes: take the
// arguments again.
int ens_check_shapes(bcs_ensemble* e)
{
    for (size_t m = 0; m < e->members.size(); ++m)
        if (e->members[m]->pg.bcount != e->shapes[m]) return ens_build_args(e);
    return BCS_OK;
}

int ens_enqueue_order(bcs_ensemble* e)
{
    const int32_t n = (int32_t)e->members.size();
    const uint32_t run = (uint32_t)std::max(e->tuning.run, 1);
    batch_order(e->d_okey, e->start(EL_OKEY), e->grid(EL_OKEY), e->d_oscan, e->start(EL_OSCAN), e->grid(EL_OSCAN),
                e->d_oscat, e->start(EL_OSCAT), e->grid(EL_OSCAT), e->d_step, e->d_cmap, e->start(EL_CMAP),
                e->grid(EL_CMAP), n, run, e->stream);
    HIPCHK(hipGetLastError());
    return BCS_OK;
}

// One frame of every member (enqueue_fused_frame, batched).
int ens_enqueue_frame(bcs_ensemble* e, bool lead_vein)
{
    const int32_t n = (int32_t)e->members.size();
    const uint32_t run = (uint32_t)std::max(e->tuning.run, 1);
    hipStream_t st = e->stream;
    HIPCHK(hipEventRecord(e->ev_fork, st));
    HIPCHK(hipStreamWaitEvent(e->side, e->ev_fork, 0));
    if (lead_vein) batch_vein(e->d_vein, e->start(EL_VEIN), n, e->grid(EL_VEIN), run, e->side);
    batch_vein_prep(e->d_prep, e->start(EL_PREP), n, e->grid(EL_PREP), run, e->side);
    batch_raycand(e->fused, e->d_step, e->start(EL_RAYCAND), n, e->grid(EL_RAYCAND), run, e->side);
    if (!e->fused) batch_ray(e->d_step, e->d_split, e->start(EL_RAY), n, e->grid(EL_RAY), run, e->side);
    HIPCHK(hipEventRecord(e->ev_join, e->side));
    GridBatch gb{e->d_key, e->d_scan, e->d_scatter, e->d_rank, e->d_occ,
                 {e->start(EL_KEY), e->start(EL_SCAN), e->start(EL_SCATTER), e->start(EL_RANK), e->start(EL_OCC)},
                 {e->grid(EL_KEY), e->grid(EL_SCAN), e->grid(EL_SCATTER), e->grid(EL_RANK), e->grid(EL_OCC)},
                 n, run};
    grid_build_batch(gb, st);
    HIPCHK(hipStreamWaitEvent(st, e->ev_join, 0));
    if (e->split) {
        batch_split_part(e->d_step, 0, e->start(EL_SPRINGS), n, e->grid(EL_SPRINGS), run, st);
        batch_split_part(e->d_step, 1, e->start(EL_CONTACTS), n, e->grid(EL_CONTACTS), run, st);
        batch_split_part(e->d_step, 2, e->start(EL_RESPOND), n, e->grid(EL_RESPOND), run, st);
    } else {
        batch_cells(e->d_step, e->start(EL_CELLS), n, e->grid(EL_CELLS), run, st);
    }
    HIPCHK(hipGetLastError());
    return BCS_OK;
}

int ens_capture(bcs_ensemble* e, int which)
{
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = BCS_OK;
    if (which == 2) {
        rc = ens_enqueue_order(e);
        for (int f = 0; f < kGraphFrames && !rc; ++f) rc = ens_enqueue_frame(e, true);
    } else {
        rc = ens_enqueue_frame(e, which == 1);
    }
    hipGraph_t g = nullptr;
    hipError_t err = hipStreamEndCapture(e->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (err != hipSuccess) return fail(BCS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(err));
    e->graph[which] = g;
    HIPCHK(hipGraphInstantiate(&e->exec[which], g, nullptr, nullptr, 0));
    return BCS_OK;
}

int ens_step(bcs_ensemble* e, int32_t n_steps)
{
    RC(ens_check_shapes(e));
    for (int k = 0; k < 3; ++k)
        if (!e->exec[k]) RC(ens_capture(e, k));
    for (int32_t k = 0; k < n_steps;) {
        const bool lead = k > 0;
        int32_t nf = 1;
        if (lead && n_steps - k >= kGraphFrames) {
            HIPCHK(hipGraphLaunch(e->exec[2], e->stream));
            nf = kGraphFrames;
            e->frames_since_order = 0;
        } else {
            if (e->frames_since_order >= kReorderPeriod) {
                RC(ens_enqueue_order(e));
                e->frames_since_order = 0;
            }
            HIPCHK(hipGraphLaunch(e->exec[lead ? 1 : 0], e->stream));
        }
        k += nf;
        e->frames_since_order += nf;
    }
    // the last frame's vein update
    batch_vein(e->d_vein, e->start(EL_VEIN), (int32_t)e->members.size(), e->grid(EL_VEIN),
               (uint32_t)std::max(e->tuning.run, 1), e->stream);
    HIPCHK(hipGetLastError());
    for (bcs_sim* s : e->members) {
        s->copies_current = false;
        s->frames_since_order = e->frames_since_order;
    }
    return BCS_OK;
}

int ens_set_tuning(bcs_ensemble* e, const bcs_ensemble_tuning& t)
{
    if (t.blocks_per_cu < 0 || t.blocks_per_cu > 64 || t.run < 1 || t.run > 64)
        return fail(BCS_ERR_INVALID_ARGUMENT, "ensemble tuning: blocks_per_cu in 0..64, run in 1..64");
    if (t.ray_mode != BCS_RAY_AUTO && t.ray_mode != BCS_RAY_SPLIT && t.ray_mode != BCS_RAY_FUSED)
        return fail(BCS_ERR_INVALID_ARGUMENT, "ensemble tuning: ray_mode is BCS_RAY_AUTO, _SPLIT or _FUSED");
    if (t.contact_mode != BCS_CONTACT_AUTO && t.contact_mode != BCS_CONTACT_CELLS && t.contact_mode != BCS_CONTACT_SLOTS)
        return fail(BCS_ERR_INVALID_ARGUMENT, "ensemble tuning: contact_mode is BCS_CONTACT_AUTO, _CELLS or _SLOTS");
    HIPCHK(hipStreamSynchronize(e->stream));
    e->tuning = t;
    e->order.resize(e->members.size());
    for (size_t k = 0; k < e->order.size(); ++k) e->order[k] = (int32_t)k;
    if (t.largest_first)
        std::stable_sort(e->order.begin(), e->order.end(),
                         [&](int32_t a, int32_t b) { return e->members[a]->np > e->members[b]->np; });
    return ens_build_args(e);
}

}  // namespace

extern "C" int bcs_ensemble_create(bcs_sim* const* members, int32_t n_members, bcs_ensemble** out)
{
    if (!members || !out || n_members < 1) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_ensemble_create: no members");
    *out = nullptr;
    const int device = members[0] ? members[0]->device : 0;
    for (int32_t k = 0; k < n_members; ++k) {
        bcs_sim* s = members[k];
        if (!s) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_ensemble_create: null member");
        if (s->device != device) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_ensemble_create: members on two devices");
        if (!defer_vein(s))
            return fail(BCS_ERR_UNSUPPORTED, "bcs_ensemble_create: members need a static triangle grid");
        for (int32_t j = 0; j < k; ++j)
            if (members[j] == s) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_ensemble_create: a member twice");
    }
    std::unique_ptr<bcs_ensemble> e(new bcs_ensemble());
    e->device = device;
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop{};
    HIPCHK(hipGetDeviceProperties(&prop, device));
    e->cus = std::max(prop.multiProcessorCount, 1);
    for (int32_t k = 0; k < n_members; ++k) {
        HIPCHK(hipStreamSynchronize(members[k]->stream));
        e->members.push_back(members[k]);
    }
    HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming));
    const size_t n = (size_t)n_members;
    RC(e->dalloc(&e->d_step, n));
    RC(e->dalloc(&e->d_vein, n));
    RC(e->dalloc(&e->d_prep, n));
    RC(e->dalloc(&e->d_key, n));
    RC(e->dalloc(&e->d_scan, n));
    RC(e->dalloc(&e->d_scatter, n));
    RC(e->dalloc(&e->d_rank, n));
    RC(e->dalloc(&e->d_occ, n));
    RC(e->dalloc(&e->d_okey, n));
    RC(e->dalloc(&e->d_oscan, n));
    RC(e->dalloc(&e->d_oscat, n));
    RC(e->dalloc(&e->d_cmap, n));
    RC(e->dalloc(&e->d_split, n));
    RC(e->dalloc(&e->d_start, (size_t)EL_COUNT * (n + 1)));
    e->frames_since_order = 0;
    for (bcs_sim* s : e->members) e->frames_since_order = std::max(e->frames_since_order, s->frames_since_order);
    RC(ens_set_tuning(e.get(), e->tuning));
    *out = e.release();
    return BCS_OK;
}

extern "C" int bcs_ensemble_destroy(bcs_ensemble* e)
{
    if (!e) return BCS_OK;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    delete e;
    return BCS_OK;
}

extern "C" int bcs_ensemble_set_tuning(bcs_ensemble* e, const bcs_ensemble_tuning* t)
{
    RC(ens_dev(e));
    if (!t) return fail(BCS_ERR_INVALID_ARGUMENT, "null tuning");
    return ens_set_tuning(e, *t);
}

extern "C" int bcs_ensemble_get_tuning(bcs_ensemble* e, bcs_ensemble_tuning* t)
{
    if (!e || !t) return fail(BCS_ERR_INVALID_ARGUMENT, "null argument");
    *t = e->tuning;
    return BCS_OK;
}

extern "C" int bcs_ensemble_step(bcs_ensemble* e, int32_t n_steps)
{
    RC(ens_dev(e));
    if (n_steps < 0) return fail(BCS_ERR_INVALID_ARGUMENT, "negative step count");
    if (n_steps == 0) return BCS_OK;
    return ens_step(e, n_steps);
}

extern "C" int bcs_ensemble_synchronize(bcs_ensemble* e)
{
    RC(ens_dev(e));
    HIPCHK(hipStreamSynchronize(e->stream));
    return BCS_OK;
}

extern "C" int bcs_ensemble_time_steps(bcs_ensemble* e, int32_t n_steps, float* total_ms)
{
    RC(ens_dev(e));
    if (!total_ms || n_steps < 1) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_ensemble_time_steps: bad arguments");
    RC(ens_check_shapes(e));
    for (int k = 0; k < 3; ++k)
        if (!e->exec[k]) RC(ens_capture(e, k));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, e->stream));
    int rc = ens_step(e, n_steps);
    HIPCHK(hipEventRecord(e1, e->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *total_ms = ms;
    return rc;
}

extern "C" int bcs_ensemble_autotune(bcs_ensemble* e, int32_t frames_per_trial, bcs_ensemble_tuning* best,
                                     char* report, int32_t report_capacity)
{
    RC(ens_dev(e));
    if (frames_per_trial < 1) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_ensemble_autotune: frames_per_trial >= 1");
    static const int32_t kBpc[] = {0, 2, 4, 8, 16};
    static const int32_t kRun[] = {1, 4};
    static const int32_t kRay[] = {BCS_RAY_SPLIT, BCS_RAY_FUSED};
    static const int32_t kLargest[] = {0, 1};
    std::string js = "{\"space\": {\"blocks_per_cu\": [0, 2, 4, 8, 16], \"run\": [1, 4], \"ray_mode\": [\"split\", "
                     "\"fused\"], \"largest_first\": [0, 1]}, \"frames_per_trial\": " +
                     std::to_string(frames_per_trial) + ", \"trials\": [";
    bcs_ensemble_tuning chosen = e->tuning;
    float chosen_ms = 3.4e38f;
    bool firstj = true;
    for (int32_t bpc : kBpc)
        for (int32_t run : kRun)
            for (int32_t ray : kRay)
                for (int32_t lf : kLargest) {
                    if (bpc == 0 && run != 1) continue;  // without persistence a run only merges blocks
                    const bcs_ensemble_tuning t{bpc, run, ray, lf, e->tuning.contact_mode};
                    RC(ens_set_tuning(e, t));
                    RC(ens_step(e, 1));  // graphs captured, caches warm
                    float ms = 0.f;
                    RC(bcs_ensemble_time_steps(e, frames_per_trial, &ms));
                    const float per = ms / (float)frames_per_trial;
                    char buf[160];
                    std::snprintf(buf, sizeof(buf),
                                  "%s{\"blocks_per_cu\": %d, \"run\": %d, \"ray_mode\": \"%s\", \"largest_first\": %d, "
                                  "\"ms_per_frame\": %.5f}",
                                  firstj ? "" : ", ", bpc, run, ray == BCS_RAY_SPLIT ? "split" : "fused", lf, per);
                    js += buf;
                    firstj = false;
                    if (per < chosen_ms) {
                        chosen_ms = per;
                        chosen = t;
                    }
                }
    RC(ens_set_tuning(e, chosen));
    char buf[200];
    std::snprintf(buf, sizeof(buf),
                  "], \"chosen\": {\"blocks_per_cu\": %d, \"run\": %d, \"ray_mode\": \"%s\", \"largest_first\": %d, "
                  "\"ms_per_frame\": %.5f}}",
                  chosen.blocks_per_cu, chosen.run, chosen.ray_mode == BCS_RAY_SPLIT ? "split" : "fused",
                  chosen.largest_first, chosen_ms);
    js += buf;
    if (best) *best = chosen;
    if (report && report_capacity > 0) {
        const size_t len = std::min<size_t>(js.size(), (size_t)report_capacity - 1);
        std::memcpy(report, js.data(), len);
        report[len] = '\0';
    }
    return BCS_OK;
}

// ============================================================ surrogate ==

namespace {

// Buffers of the served type (the first with 42 particles per cell).
int surr_alloc(bcs_sim* s)
{
    if (s->surr_type < 0) {
        for (int t = 0; t < s->nt; ++t)
            if (s->tt.ppc[t] == kSurrPpc && s->tt.cnt[t] > 0) {
                s->surr_type = t;
                break;
            }
        if (s->surr_type < 0) return fail(BCS_ERR_UNSUPPORTED, "surrogate: no cell type of 42 particles");
    }
    s->surr_cells = s->tt.cnt[s->surr_type];
    const size_t nc = (size_t)s->surr_cells;
    if (!s->sf_X) {
        RC(s->dalloc(&s->sf_X, nc * kSurrInPad));
        RC(s->dalloc(&s->sf_a0, nc * kSurrHidden));
        RC(s->dalloc(&s->sf_a1, nc * kSurrHidden));
        RC(s->dalloc(&s->sf_out, nc * kSurrOutPad));
        RC(s->dalloc(&s->sf_S, nc * kSurrOutPad));
        RC(s->dalloc(&s->sf_f32, nc * kSurrIn));
        RC(s->dalloc(&s->touched, (size_t)std::max(s->np, 1)));
        RC(s->dalloc(&s->surr_stats, (size_t)kSurrStatSlots * kSurrStatStride));
        HIPCHK(hipMemsetAsync(s->touched, 0, (size_t)std::max(s->np, 1), s->stream));
        HIPCHK(hipMemsetAsync(s->surr_stats, 0, 4 * (size_t)kSurrStatSlots * kSurrStatStride, s->stream));
    }
    return BCS_OK;
}

// One frame enqueued directly (what one frame of bcs_sim_step does).
int enqueue_frame_direct(bcs_sim* s)
{
    if (s->frames_since_order >= kReorderPeriod) {
        RC(enqueue_order(s));
        s->frames_since_order = 0;
    }
    RC(enqueue_fused_frame(s, nullptr, false, true));
    s->frames_since_order += 1;
    s->copies_current = false;
    return BCS_OK;
}

}  // namespace

extern "C" int bcs_sim_attach_surrogate(bcs_sim* s, bcs_surrogate* net, float tolerance)
{
    RC(set_dev(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    drop_graph(s);
    if (!net) {
        s->surr = nullptr;
        return BCS_OK;
    }
    SurrNet n{};
    int dev = -1;
    RC(bcs_surrogate_net(net, &n, &dev));
    if (dev != s->device) return fail(BCS_ERR_INVALID_ARGUMENT, "surrogate: network on another device");
    if (s->pc.rk4) return fail(BCS_ERR_UNSUPPORTED, "surrogate: it learns the Heun springs (particle_rk4 is set)");
    RC(surr_alloc(s));
    s->surr = net;
    s->surr_net = n;
    s->surr_tol = tolerance;
    HIPCHK(hipStreamSynchronize(s->stream));
    return BCS_OK;
}

extern "C" int bcs_sim_surrogate_sample(bcs_sim* s, float* features, float* targets, int32_t capacity, int32_t* n_out)
{
    RC(set_dev(s));
    if (!n_out || capacity < 0 || (capacity > 0 && (!features || !targets)))
        return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_sim_surrogate_sample: bad arguments");
    if (s->pc.rk4) return fail(BCS_ERR_UNSUPPORTED, "surrogate: it learns the Heun springs (particle_rk4 is set)");
    RC(surr_alloc(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    // frame 1 keeps the contact flags; frame 2 samples the network input of
    // its starting state and the plain spring displacements it computes
    s->recording = true;
    int rc = enqueue_frame_direct(s);
    if (!rc) {
        launch_surr_features(s->pos, s->vel, s->np, s->tt.pstart[s->surr_type], s->surr_cells, s->touched, 1.0f, 1.0f,
                             s->sf_X, s->sf_f32, s->stream);
        s->sampling = true;
        rc = enqueue_frame_direct(s);
    }
    s->sampling = false;
    s->recording = false;
    RC(rc);
    HIPCHK(hipStreamSynchronize(s->stream));
    const int32_t n = std::min(capacity, s->surr_cells);
    if (n > 0) {
        HIPCHK(hipMemcpy(features, s->sf_f32, (size_t)n * kSurrIn * 4, hipMemcpyDeviceToHost));
        std::vector<float> S((size_t)n * kSurrOutPad);
        HIPCHK(hipMemcpy(S.data(), s->sf_S, S.size() * 4, hipMemcpyDeviceToHost));
        for (int32_t m = 0; m < n; ++m)
            std::memcpy(targets + (size_t)m * kSurrOut, S.data() + (size_t)m * kSurrOutPad, kSurrOut * 4);
    }
    *n_out = n;
    return BCS_OK;
}

extern "C" int bcs_sim_surrogate_stats(bcs_sim* s, int32_t reset, uint64_t* accepted, uint64_t* fallback,
                                       float* max_error_accepted, float* max_error_all)
{
    RC(set_dev(s));
    if (!s->surr_stats) return fail(BCS_ERR_STATE, "surrogate: none attached");
    HIPCHK(hipStreamSynchronize(s->stream));
    std::vector<uint32_t> h((size_t)kSurrStatSlots * kSurrStatStride);
    HIPCHK(hipMemcpy(h.data(), s->surr_stats, 4 * h.size(), hipMemcpyDeviceToHost));
    uint64_t acc = 0, fb = 0;
    uint32_t e_acc = 0, e_all = 0;  // non-negative floats: bit order is value order
    for (int k = 0; k < kSurrStatSlots; ++k) {
        const uint32_t* w = h.data() + (size_t)k * kSurrStatStride;
        acc += w[0];
        fb += w[1];
        e_acc = std::max(e_acc, w[2]);
        e_all = std::max(e_all, w[3]);
    }
    if (accepted) *accepted = acc;
    if (fallback) *fallback = fb;
    float f2, f3;
    std::memcpy(&f2, &e_acc, 4);
    std::memcpy(&f3, &e_all, 4);
    if (max_error_accepted) *max_error_accepted = f2;
    if (max_error_all) *max_error_all = f3;
    if (reset) {
        HIPCHK(hipMemsetAsync(s->surr_stats, 0, 4 * h.size(), s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    return BCS_OK;
}
