/*
 * bcs.h -- C-ABI of the MI355X (gfx950) blood-cell physics step.
 *
 * The reference exposes this path as C++ classes whose sizes are compile-time
 * constants (src/simulation/simulation_controller.cuh:24-67,
 * src/grids/uniform_grid.cuh:15-65, src/objects/blood_cells.cuh:16-74,
 * src/objects/vein_triangles.cuh:24-124).  This header replaces them with an
 * opaque-handle C interface: plain pointers and sizes, no C++ or torch types,
 * every call returns a status code instead of exiting the process
 * (reference: CUDACHECK -> exit(EXIT_FAILURE), src/utilities/cuda_handle_error.cuh:18-25).
 *
 * Conventions
 *  - Vector quantities are structure-of-arrays: x[n], y[n], z[n] contiguous,
 *    i.e. one float buffer of 3*n values (reference cudaVec3,
 *    src/utilities/cuda_vec3.cuh:11-73).
 *  - Host pointers are caller-owned and copied; device pointers returned by
 *    bcs_sim_device_state() stay owned by the simulation.
 *  - One simulation = one HIP stream = one host thread at a time.
 */
#ifndef BCS_H
#define BCS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define BCS_OK                    0
#define BCS_ERR_INVALID_ARGUMENT  1
#define BCS_ERR_HIP               2
#define BCS_ERR_OUT_OF_MEMORY     3
#define BCS_ERR_GRID_TOO_LARGE    4
#define BCS_ERR_NO_DEVICE         5
#define BCS_ERR_STATE             6
#define BCS_ERR_UNSUPPORTED       7

/* ---- device error word (bit flags, sticky until cleared) ---------------
 * Reference: device printf on out-of-grid positions
 * (src/grids/uniform_grid.cu:26-29) and on NaN in normalize
 * (src/utilities/math.cuh:98-101). */
#define BCS_FLAG_OUT_OF_GRID        0x1u  /* position outside [min,max] or key clamped */
#define BCS_FLAG_NAN_NORMALIZE      0x2u  /* normalize() produced NaN and returned 0 */
#define BCS_FLAG_SCAN_TIMEOUT       0x4u  /* a bounded device wait gave up */
#define BCS_FLAG_RADIUS_CLAMPED     0x8u  /* mixed-type radius index left the model table */
#define BCS_FLAG_RAY_SERIAL         0x10u /* ray work list full: some rays were scanned serially (same result) */

/* ---- grid selector ----------------------------------------------------- */
#define BCS_GRID_PARTICLES 0
#define BCS_GRID_TRIANGLES 1

/* ---- physics / simulation configuration --------------------------------
 * Field-for-field runtime mirror of src/config/simulation.hpp:3-21 and
 * src/config/physics.hpp:4-34 (+ cylinderRadius, src/meta_factory/vein_factory.hpp:88).
 * bcs_default_physics() fills the reference's default values. */
typedef struct bcs_physics {
    /* src/config/simulation.hpp */
    float dt;                          /* dt = 0.008 */
    int32_t use_blood_flow;            /* useBloodFlow = 1 (vein-end teleport) */
    int32_t enable_reaction_force;     /* enableReactionForce = 1 */
    int32_t enable_big_cells_brake;    /* enableBigCellsBrake = 1 */
    float particle_cell[3];            /* cellWidth/Height/Depth = 2 */
    float triangle_cell[3];            /* cellWidthTriangles... = 25 */
    int32_t bounding_spheres_coeff;    /* boundingSpheresCoeff = 3 */
    float grid_y_margin;               /* gridYMargin = 40 */
    float grid_xz_margin;              /* gridXZMargin = 40 */
    float min_spawn_y;                 /* minSpawnY = -20 */
    /* src/config/physics.hpp */
    float velocity_collision_damping;  /* 0.96 */
    float particle_k_sniff;            /* 25 */
    float vein_k_sniff;                /* 0.1 */
    float particle_d_fact;             /* 10 */
    float vein_d_fact;                 /* 0.7 */
    float vein_collision_force_intensity; /* 0.005 */
    float viscous_damping;             /* 0.39 */
    float collision_spring_coeff;      /* 6 */
    float collision_damping_coeff;     /* 6 */
    float collision_shear_coeff;       /* 4 */
    float max_cell_size_factor_before_brake; /* 1.5 */
    float big_particle_braking_intensity;    /* 0.8 */
    float init_velocity[3];            /* (0,-80,0): teleport velocity */
    float vein_impact_distance;        /* 6 */
    float vein_impact_minimal_force_distance; /* 0.001 */
    float gravity[3];                  /* (0,-25,0) */
    /* src/meta_factory/vein_factory.hpp:88 */
    float cylinder_radius;             /* 50: teleport spread */
    /* defined by this implementation (no reference counterpart) */
    uint64_t seed;                     /* counter-based RNG seed for the vein-end teleport */
    int32_t dynamic_triangle_grid;     /* 0: triangle centres frozen after creation
                                          (reference single-GPU behaviour, src/main.cu:164-176);
                                          1: recompute centres + triangle grid every frame */
    int32_t ray_work_items;            /* ray work-list capacity in items: 0 = 8 per particle (at
                                          least 65,536 in all and 1,024 per cell block); n > 0 = n in
                                          all.  Rays that do not fit are scanned serially (same
                                          result, BCS_FLAG_RAY_SERIAL). */
    int32_t particle_rk4;              /* 1: the reference's USE_RUNGE_KUTTA_FOR_PARTICLE build
                                          (src/simulation/physics.cuh:10,56-73,
                                          src/objects/blood_cells.cu:114-119,166-171); 0: Heun (its default) */
} bcs_physics;

/* ---- blood-cell type ---------------------------------------------------
 * Runtime form of BloodCellDef<Count, ParticlesInCell, ..., Springs, Vertices>
 * (src/meta_factory/blood_cells_def_type.hpp:20-46).  Spring i joins vertices
 * spring_start[i] and spring_end[i] with rest length spring_length[i]
 * (0 means "no spring", as in src/objects/blood_cells.cu:14,93). */
typedef struct bcs_cell_type {
    int32_t count;                 /* number of cells of this type */
    int32_t particles_in_cell;     /* ppc */
    int32_t spring_count;
    const int32_t* spring_start;
    const int32_t* spring_end;
    const float* spring_length;
    const float* vertices;         /* 3*ppc floats, xyz interleaved: model vertex positions */
    float contact_radius;          /* > 0: every vertex uses this collision radius;
                                      <= 0: derived like the reference (min neighbour
                                      distance / (2*boundingSpheresCoeff),
                                      src/simulation/simulation_controller.cu:93-153) */
    int32_t reserved;
} bcs_cell_type;

/* ---- vein mesh ---------------------------------------------------------
 * Runtime form of veinPositions / veinIndices / VeinEndingCenters /
 * VeinEndingRadii (src/config/vein_definition.hpp:12,15965,31768-31776). */
typedef struct bcs_vein_mesh {
    int32_t vertex_count;
    const float* positions;        /* 3*vertex_count, xyz interleaved */
    int32_t triangle_count;
    const uint32_t* indices;       /* 3*triangle_count */
    int32_t ending_count;
    const float* ending_centers;   /* 3*ending_count, xyz interleaved */
    const float* ending_radii;     /* ending_count */
} bcs_vein_mesh;

/* Derived, read-only description of a created simulation (the tables the
 * reference folds at compile time in src/meta_factory/blood_cell_factory.hpp:172-328
 * and src/meta_factory/vein_factory.hpp:46-88, plus grid geometry from
 * src/grids/uniform_grid.cu:83-90). */
typedef struct bcs_sim_info {
    int32_t particle_count;          /* particleCount */
    int32_t blood_cell_count;        /* bloodCellCount */
    int32_t type_count;              /* bloodCellTypeCount */
    int32_t model_vertex_count;      /* particleDistinctCellsCount */
    int32_t vertex_count;            /* vein vertices */
    int32_t triangle_count;
    float bounds_min[3];             /* minX, minY, minZ */
    float bounds_max[3];             /* maxX, maxY, maxZ */
    int32_t particle_cells[3];       /* cellCountX/Y/Z of the particle grid */
    int32_t triangle_cells[3];       /* cellCountX/Y/Z of the triangle grid */
    uint64_t particle_cell_count;
    uint64_t triangle_cell_count;
    int32_t particle_starts[16];     /* particleStarts */
    int32_t cell_type_starts[16];    /* bloodCellTypesStarts */
    int32_t model_starts[16];        /* bloodCellModelStarts */
    int32_t graph_starts[16];        /* accumulatedGraphSizes */
    int32_t particles_in_cell[16];
    int32_t cells_of_type[16];
    uint64_t step_index;             /* frames completed (drives the teleport RNG) */
} bcs_sim_info;

/* Raw device pointers (for zero-copy use by a caller that already holds the
 * state on the GPU, e.g. torch tensors built over them).  All SoA float[3*n]. */
typedef struct bcs_device_state {
    float* positions;     /* particles: x[N] y[N] z[N] */
    float* velocities;
    float* forces;
    float* centers;       /* blood-cell centres, 3*bloodCellCount */
    float* vein_positions;
    float* vein_velocities;
    float* vein_forces;
    float* triangle_centers;
    void* stream;         /* hipStream_t the simulation enqueues on */
    /* per model vertex (sum of ppc over the types, in type order); read
     * only: the contact search takes its reach and whether every radius is
     * the same from these radii at creation */
    float* model_radii;   /* collision radii: cellModelsBoundingSpheres
                             (src/simulation/simulation_controller.cuh:42) */
    float* initial_radii; /* distance to the model centroid: BloodCells::initialRadiuses
                             (src/objects/blood_cells.cuh:33) */
    uint32_t* triangle_indices; /* 3 per triangle: VeinTriangles::indices
                                   (src/objects/vein_triangles.cuh:60) */
} bcs_device_state;

/* Device view of one grid: the public members of the reference's
 * UniformGrid (src/grids/uniform_grid.cuh:23-35) after its last build. */
typedef struct bcs_grid_view {
    uint32_t* grid_cell_ids;    /* gridCellIds: cell id of every sorted slot (object_count) */
    int32_t* particle_ids;      /* particleIds: object index of every sorted slot (object_count) */
    int32_t* grid_cell_starts;  /* gridCellStarts: first slot of every cell id (cell_count), or NULL */
    int32_t* grid_cell_ends;    /* gridCellEnds: last slot (inclusive) of every cell id, or NULL */
    int32_t object_count;       /* objectCount */
    int32_t cell_counts[3];     /* cellCountX/Y/Z */
    uint64_t cell_count;        /* cellCount */
    float cell_size[3];         /* cellWidth/Height/Depth */
    float bounds_min[3];        /* minX, minY, minZ */
    float bounds_max[3];        /* maxX, maxY, maxZ */
} bcs_grid_view;

typedef struct bcs_sim bcs_sim;

/* ---- configuration helpers (host only, no GPU needed) ------------------- */

/* Reference defaults of src/config/physics.hpp:4-34 and src/config/simulation.hpp:3-21. */
int bcs_default_physics(bcs_physics* out);

/* Reference type folding: merge user definitions that share (ppc, spring list)
 * and sort them with the reference comparator
 * (src/meta_factory/blood_cell_factory.hpp:21-162, boost::mp11 mp_sort quicksort).
 * order_out[k] = index into `user` of the k-th folded type, counts_out[k] = its
 * merged count; *n_out = number of folded types. */
int bcs_fold_reference_types(const bcs_cell_type* user, int32_t n_user,
                             int32_t* order_out, int32_t* counts_out, int32_t* n_out);

/* Grid bounds = vein AABB +/- margins (src/meta_factory/vein_factory.hpp:21-86). */
int bcs_grid_bounds(const bcs_vein_mesh* mesh, const bcs_physics* phys,
                    float bounds_min[3], float bounds_max[3]);

/* Vein spring-neighbour tables: 9 slots per vertex, slot-major
 * (ids[s*nv + v], lengths[s*nv + v]); unused = -1 / -1.0f
 * (src/meta_factory/vein_factory.hpp:130-174). */
int bcs_vein_neighbors(const bcs_vein_mesh* mesh, int32_t* ids, float* lengths);

/* Per-model-vertex collision radii and initial radii for a type list in final
 * order (src/simulation/simulation_controller.cu:93-153).  Arrays sized
 * sum(ppc). */
int bcs_model_tables(const bcs_physics* phys, const bcs_cell_type* types, int32_t n_types,
                     float* bounding_spheres, float* initial_radii);

/* ---- simulation lifecycle (replaces SimulationController + BloodCells +
 *      VeinTriangles + two UniformGrid objects, src/main.cu:122-138) -------- */

/* Types are taken in final order (no re-sorting; use bcs_fold_reference_types
 * to reproduce the reference order).  At most 16 types. */
int bcs_sim_create(const bcs_physics* phys, const bcs_cell_type* types, int32_t n_types,
                   const bcs_vein_mesh* mesh, int32_t device, bcs_sim** out);
int bcs_sim_destroy(bcs_sim* sim);
int bcs_sim_get_info(bcs_sim* sim, bcs_sim_info* out);
int bcs_sim_device_state(bcs_sim* sim, bcs_device_state* out);

/* Host <-> device state transfer (SoA, 3*n floats each; NULL = skip). */
int bcs_sim_set_particles(bcs_sim* sim, const float* positions, const float* velocities,
                          const float* forces);
int bcs_sim_get_particles(bcs_sim* sim, float* positions, float* velocities, float* forces);
int bcs_sim_set_vein(bcs_sim* sim, const float* positions, const float* velocities,
                     const float* forces);
int bcs_sim_get_vein(bcs_sim* sim, float* positions, float* velocities, float* forces);
int bcs_sim_get_centers(bcs_sim* sim, float* centers);
int bcs_sim_set_step_index(bcs_sim* sim, uint64_t step_index);

/* Reference step API ---------------------------------------------------- */

/* UniformGrid::calculateGrid (src/grids/uniform_grid.cu:129-155) on the
 * particle positions (BCS_GRID_PARTICLES) or the triangle centres
 * (BCS_GRID_TRIANGLES; src/main.cu:175-176). */
int bcs_sim_calculate_grid(bcs_sim* sim, int32_t which);

/* SimulationController::calculateNextFrame
 * (src/simulation/simulation_controller.cu:246-313): vein springs, cell
 * centres, cell springs, particle collisions, vein collisions. */
int bcs_sim_calculate_next_frame(bcs_sim* sim);

/* SimulationController::propagateAll
 * (src/simulation/simulation_controller.cu:315-331): particle integrate, vein
 * integrate (+ zero vein forces), vein-end teleport. */
int bcs_sim_propagate_all(bcs_sim* sim);

/* One whole frame of src/main.cu:175-208 (grids + calculateNextFrame +
 * propagateAll) repeated n times through fused kernels replayed from a HIP
 * graph.  Asynchronous: returns after enqueueing. */
int bcs_sim_step(bcs_sim* sim, int32_t n_steps);

/* Replay frames from the HIP graph (enabled != 0, the default) or enqueue the
 * same fused kernels directly per frame (enabled == 0, e.g. for profilers that
 * do not follow graph launches).  Results are identical either way.  No
 * counterpart in the reference (its frames are direct launches). */
int bcs_sim_use_graph(bcs_sim* sim, int32_t enabled);

/* How a frame runs the vein ray test (detectVeinCollisions' triangle scan,
 * src/simulation/vein_collisions.cu:181-233): BCS_RAY_SPLIT = candidates, then
 * the ray test as a second launch spread over up to 8 blocks per cell block;
 * BCS_RAY_FUSED = one launch in which each cell block tests its own
 * candidates from LDS; BCS_RAY_AUTO (the default) fuses when the split launch
 * would get one block per cell block (more than 4,096 cell blocks).  Results
 * are identical in every mode.  No counterpart in the reference. */
#define BCS_RAY_AUTO 0
#define BCS_RAY_SPLIT 1
#define BCS_RAY_FUSED 2
int bcs_sim_set_ray_mode(bcs_sim* sim, int32_t mode);

/* Where a frame runs the particle-particle contact search
 * (calculateParticleCollisions, src/simulation/particle_collisions.cuh:104-269):
 * BCS_CONTACT_CELLS = inside the cell kernel, whose lanes follow the particles
 * of whole blood cells; BCS_CONTACT_SLOTS = in a kernel of its own whose lanes
 * follow the sorted slots of the particle grid (neighbouring lanes share their
 * neighbour cells), between a springs kernel and a response kernel;
 * BCS_CONTACT_AUTO (the default) = SLOTS above 4,096 cell blocks (large
 * workloads), else CELLS.  Results are identical in every mode.  No
 * counterpart in the reference. */
#define BCS_CONTACT_AUTO 0
#define BCS_CONTACT_CELLS 1
#define BCS_CONTACT_SLOTS 2
int bcs_sim_set_contact_mode(bcs_sim* sim, int32_t mode);

/* Bucket size of the particle sort: 2^shift grid cells per bucket (0..24).
 * bcs_sim_set_particles chooses it from the particles' positions; this call
 * overrides that choice until the next bcs_sim_set_particles.  Buckets of 2^8
 * cells or more have the sort write the occupancy map itself; smaller ones
 * build it in a pass of its own.  Results are identical for every size.  No
 * counterpart in the reference (thrust's sort has no such parameter). */
int bcs_sim_set_bucket_shift(bcs_sim* sim, int32_t shift);

int bcs_sim_synchronize(bcs_sim* sim);

/* Time n fused steps with HIP events on the simulation's stream (ms total). */
int bcs_sim_time_steps(bcs_sim* sim, int32_t n_steps, float* total_ms);

/* Per-kernel average durations (ms) of the fused step over n steps, measured
 * with HIP events around each launch on the simulation's stream.  names_out
 * receives a ';'-separated list of kernel names matching ms_out order.  In
 * fused ray mode the candidate-and-test launch is named k_raycand_ray and
 * the k_ray entry reads 0 (no launch). */
int bcs_sim_profile_kernels(bcs_sim* sim, int32_t n_steps, float* ms_out, int32_t capacity,
                            int32_t* n_kernels, char* names_out, int32_t names_capacity);

/* Parity views ------------------------------------------------------------ */

/* Sorted grid of the last bcs_sim_calculate_grid / step: keys_sorted[n] and
 * perm[n] (gridCellIds / particleIds after thrust::stable_sort_by_key), and
 * the occupied cells with inclusive [start, end] ranges
 * (gridCellStarts / gridCellEnds).  Any output may be NULL.  occ arrays need
 * capacity n. */
int bcs_sim_get_grid(bcs_sim* sim, int32_t which, uint32_t* keys_sorted, int32_t* perm,
                     int32_t* n_occupied, uint32_t* occ_keys, int32_t* occ_start,
                     int32_t* occ_end);

/* bcs_sim_get_grid for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc);
 * no counterpart in the reference.  The same outputs, but it calls no
 * hipMalloc or hipFree: the occupied-cell gather is staged in a static
 * device array of the library (2^20 occupied cells, 8 MiB per GPU, held
 * until the process ends).  keys_sorted, n_occupied and the three occ_*
 * arrays are required (capacity n; they are the host working arrays).  More
 * than 2^20 occupied cells: BCS_ERR_INVALID_ARGUMENT, nothing written.  All
 * calls on one GPU share the static array: calls that overlap on one GPU
 * (from two threads) race. */
int bcs_sim_get_grid_static(bcs_sim* sim, int32_t which, uint32_t* keys_sorted, int32_t* perm,
                            int32_t* n_occupied, uint32_t* occ_keys, int32_t* occ_start, int32_t* occ_end);

/* Contact pairs of the particle-collision phase on the current grid and
 * state: pairs (i, j) with d^2 <= (r_i + r_j)^2 and d^2 >= 1e-4, in visit
 * order per i (src/simulation/particle_collisions.cuh:26-83).  counts[i] per
 * particle; pairs_j flattened by particle index.  *n_total is always set;
 * pairs_j may be NULL to only count. */
int bcs_sim_get_contacts(bcs_sim* sim, int32_t* counts, int32_t* pairs_j, int64_t capacity,
                         int64_t* n_total);

/* Device pointers of a grid's sorted arrays (src/grids/uniform_grid.cuh:32-35).
 * With dense_ranges != 0 the view also holds per-cell-id start and end
 * arrays laid out as the reference's gridCellStarts / gridCellEnds, filled
 * from the current grid at this call (empty cells: start 0, end -1, so an
 * inclusive loop over them runs no iteration; the reference leaves them
 * stale).  They are owned by the simulation, allocated at the first such
 * call (cellCount int32 each; BCS_ERR_GRID_TOO_LARGE above 2^31 cells) and
 * refreshed only by this call, on the simulation's stream.  The frame
 * kernels themselves use a compact table and never read them. */
int bcs_sim_grid_view(bcs_sim* sim, int32_t which, int32_t dense_ranges, bcs_grid_view* out);

/* UniformGrid::calculateCellId (src/grids/uniform_grid.cuh:60,
 * src/grids/uniform_grid.cu:24-36,157-160) for n host positions (SoA: x[n],
 * y[n], z[n]) on the particle or triangle grid; the same keys the grid build
 * computes.  Host computation, no GPU work. */
int bcs_sim_cell_ids(bcs_sim* sim, int32_t which, const float* positions, int32_t n, uint32_t* cell_ids);

/* Counters of the last completed frame: ray work items (triangle runs of
 * <= 16 whose cell box the ray can reach) and particle-vein hits. */
int bcs_sim_frame_stats(bcs_sim* sim, uint32_t* ray_work_items, uint32_t* vein_hits);

/* Frame export in the renderer's buffer layouts
 * (calculatePositionsKernel / calculateTriangleVerticesKernel,
 * src/graphics/glcontroller.cu:23-50, 168-202):
 *   offsets            3 floats per particle, xyz interleaved (all particles);
 *   type_vertices[t]   6 floats per particle of type t, xyz at 0..2 of each
 *                      group of 6 (3..5, the normals, are left untouched);
 *   vein_vertices      6 floats per vein vertex, xyz at 0..2 (3..5 untouched).
 * All are DEVICE pointers (e.g. buffers a renderer mapped for HIP); any may be
 * NULL.  Enqueued on the simulation's stream: asynchronous. */
int bcs_sim_export_frame(bcs_sim* sim, float* offsets, float* const* type_vertices, int32_t n_types,
                         float* vein_vertices);

/* Trajectory files (no reference counterpart: the reference only renders).
 * Layout, little endian: a 64-byte header
 *   char magic[8] = "BCSTRAJ1"; int32 particles, vein_vertices, types, flags
 *   (bit 0: velocities stored); float dt; int32 reserved[9]
 * then per frame: uint64 step index, particle positions (SoA, 3N f32),
 * [particle velocities (3N f32)], vein positions (SoA, 3Nv f32).
 * bcs_traj_append copies the current state from the device (synchronises the
 * stream) and appends one frame. */
typedef struct bcs_traj bcs_traj;
int bcs_traj_open(bcs_sim* sim, const char* path, int32_t with_velocities, bcs_traj** out);
int bcs_traj_append(bcs_traj* traj);
int bcs_traj_close(bcs_traj* traj);

/* ---- batched ensembles ---------------------------------------------------
 * Many independent simulations (one vein each, of any size) stepped together:
 * each frame runs every step kernel ONCE for all members -- a persistent grid
 * whose blocks take the members' blocks in turn -- instead of once per
 * member, and the frames are replayed from HIP graphs.  Every member's
 * results are bit-identical to stepping it alone (bcs_sim_step).  No
 * counterpart in the reference, which runs one vein per process
 * (src/main.cu:102-138, src/simulation/simulation_controller.cu:246-331). */
typedef struct bcs_ensemble bcs_ensemble;

/* Launch shape of the batched kernels (results never depend on it). */
typedef struct bcs_ensemble_tuning {
    int32_t blocks_per_cu;  /* hardware blocks of each batched launch per compute unit
                               (0: one hardware block per member block, no persistence) */
    int32_t run;            /* consecutive member blocks a hardware block takes per turn (>= 1) */
    int32_t ray_mode;       /* BCS_RAY_SPLIT or BCS_RAY_FUSED for every member
                               (BCS_RAY_AUTO: fused when the members hold more than 4,096 cell blocks) */
    int32_t largest_first;  /* 1: the blocks of larger members come first in each launch */
    int32_t contact_mode;   /* BCS_CONTACT_CELLS or BCS_CONTACT_SLOTS for every member
                               (BCS_CONTACT_AUTO: slots when the members hold more than 4,096 cell blocks) */
} bcs_ensemble_tuning;

/* The members stay owned by the caller and must outlive the ensemble; all on
 * one device, with static triangle grids (dynamic_triangle_grid = 0).  While
 * the ensemble exists, step the members only through it (their state is read
 * and written as usual with bcs_sim_get_* / bcs_sim_set_* after
 * bcs_ensemble_synchronize). */
int bcs_ensemble_create(bcs_sim* const* members, int32_t n_members, bcs_ensemble** out);
int bcs_ensemble_destroy(bcs_ensemble* ens);
int bcs_ensemble_set_tuning(bcs_ensemble* ens, const bcs_ensemble_tuning* tuning);
int bcs_ensemble_get_tuning(bcs_ensemble* ens, bcs_ensemble_tuning* tuning);
/* n frames of every member (asynchronous; the same frames bcs_sim_step runs). */
int bcs_ensemble_step(bcs_ensemble* ens, int32_t n_steps);
int bcs_ensemble_synchronize(bcs_ensemble* ens);
/* n frames timed with HIP events on the ensemble's stream (ms in all). */
int bcs_ensemble_time_steps(bcs_ensemble* ens, int32_t n_steps, float* total_ms);
/* Times frames_per_trial frames (after one untimed frame) under every tuning
 * of a fixed search space -- blocks_per_cu in {0, 2, 4, 8, 16}, run in {1, 4},
 * ray_mode in {SPLIT, FUSED}, largest_first in {0, 1} -- and keeps the
 * fastest.  The members advance by all the frames run.  report (optional)
 * receives one JSON object: the search space, every trial's ms per frame and
 * the choice. */
int bcs_ensemble_autotune(bcs_ensemble* ens, int32_t frames_per_trial, bcs_ensemble_tuning* best, char* report,
                          int32_t report_capacity);

/* ---- learned spring surrogate -------------------------------------------
 * A per-cell MLP 294 -> 256 -> 256 -> 256 -> 126 (GELU) in BF16 on the matrix
 * cores with FP32 accumulation, trained on frames of the plain method, that
 * stands in for the spring sub-step (gatherForcesKernel,
 * src/objects/blood_cells.cu:66-120; calculateParticlesSpringForceComponent,
 * src/simulation/physics.cuh:53-78) of 42-particle cells.  No counterpart in
 * the reference.  One row per cell:
 *   input  (294)  centroid-relative positions (126: particle a, axis k at
 *                 3a + k), velocities relative to the cell's mean velocity
 *                 (126), contact flags of the previous frame (42: 1 = the
 *                 particle touched another one)
 *   output (126)  spring displacements d = (dt^2/4) S: the spring force S's
 *                 share of the frame's Heun position update
 * Rows are in physical units; the network works on them divided by the
 * scales (positions, velocities; outputs by out_scale). */
typedef struct bcs_surrogate bcs_surrogate;
#define BCS_SURR_INPUTS 294
#define BCS_SURR_HIDDEN 256
#define BCS_SURR_OUTPUTS 126

/* Weights uniform in +-1/sqrt(fan_in) from seed (SplitMix64). */
int bcs_surrogate_create(int32_t device, uint64_t seed, bcs_surrogate** out);
int bcs_surrogate_destroy(bcs_surrogate* net);
/* weights[l]: [out][in] row-major fp32 (256 x 294, 256 x 256, 256 x 256,
 * 126 x 256); biases[l]: out entries.  get_weights skips null entries. */
int bcs_surrogate_set_weights(bcs_surrogate* net, const float* const* weights, const float* const* biases);
int bcs_surrogate_get_weights(bcs_surrogate* net, float* const* weights, float* const* biases);
int bcs_surrogate_set_scales(bcs_surrogate* net, float pos_scale, float vel_scale, float out_scale);
int bcs_surrogate_get_scales(bcs_surrogate* net, float* pos_scale, float* vel_scale, float* out_scale);
/* n rows of features (n x 294) -> displacements (n x 126), physical units. */
int bcs_surrogate_predict(bcs_surrogate* net, const float* features, int32_t n, float* out);
/* Mean-squared-error training with Adam (betas 0.9 / 0.999, eps 1e-8) on n
 * samples in physical units, epochs passes in a seeded order, batches of
 * `batch` rows (a multiple of 128).  loss_per_epoch (optional, epochs
 * entries): mean squared error in scaled units. */
int bcs_surrogate_train(bcs_surrogate* net, const float* features, const float* targets, int32_t n, int32_t epochs,
                        int32_t batch, float lr, uint64_t seed, float* loss_per_epoch);
/* ms per inference pass over n_cells rows, averaged over reps passes. */
int bcs_surrogate_time_inference(bcs_surrogate* net, int32_t n_cells, int32_t reps, float* ms_per_pass);

/* Frames of bcs_sim_step then give the first 42-particle type's springs
 * from the network: each frame computes the network input from its starting
 * state (and the previous frame's contact flags), runs the network, and in
 * the cell kernel compares every cell's predicted displacements with the
 * plain spring solve of the same state; a cell whose largest particle error
 * is at most `tolerance` (a length) uses the prediction, any other cell the
 * plain springs (with tolerance <= 0 every cell falls back and the frames
 * equal plain frames bit for bit).  NULL detaches.  Attach again after
 * changing the scales. */
int bcs_sim_attach_surrogate(bcs_sim* sim, bcs_surrogate* net, float tolerance);
/* Training rows from the plain method: runs two frames (the first keeps the
 * contact flags) and returns, for up to `capacity` cells of the first
 * 42-particle type, the second frame's network input (features, capacity x
 * 294) and plain spring displacements (targets, capacity x 126). */
int bcs_sim_surrogate_sample(bcs_sim* sim, float* features, float* targets, int32_t capacity, int32_t* n_rows);
/* Cells that used the prediction / fell back since the last reset, and the
 * largest cell error among the accepted / all predicted cells. */
int bcs_sim_surrogate_stats(bcs_sim* sim, int32_t reset, uint64_t* accepted, uint64_t* fallback,
                            float* max_error_accepted, float* max_error_all);

/* Device error word (BCS_FLAG_*); clear != 0 resets it. */
int bcs_sim_error_flags(bcs_sim* sim, int32_t clear, uint32_t* flags);

/* Last error message of the calling thread ("" if none). */
const char* bcs_last_error(void);

/* Number of HIP devices visible (0 on a machine without a GPU). */
int bcs_device_count(int32_t* count);

/* Plain device-to-device copy of n_bytes through this library's own copy
 * kernel, on a fresh allocation (HBM counter calibration).  Returns elapsed ms. */
int bcs_calibration_copy(int64_t n_bytes, int32_t repeats, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* BCS_H */
