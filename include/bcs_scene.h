/*
 * bcs_scene.h -- host-side scene construction for the blood-cell physics step
 * (no GPU needed).
 *
 * The reference defines its scene at compile time: cell presets
 * (src/config/blood_cell_presets.hpp:13-318), the user cell list
 * (src/config/blood_cells_definition.hpp:13-25), a generated vein mesh
 * (src/config/vein_definition.hpp) and a time-seeded random initial placement
 * (src/simulation/simulation_controller.cu:156-243).  This header builds the
 * same inputs at run time, reproducibly from a seed: cell models, vein meshes
 * and initial particle states, and the preset workloads the benchmarks use.
 */
#ifndef BCS_SCENE_H
#define BCS_SCENE_H

#include "bcs.h"

#ifdef __cplusplus
extern "C" {
#endif

#define BCS_VEIN_CYLINDER 0   /* straight tube along -y from y = 0 */
#define BCS_VEIN_Y        1   /* trunk along -y that splits into two branches */

#define BCS_CELL_ICOSAHEDRON 0 /* 12 vertices, 30 edge springs */
#define BCS_CELL_DISC42      1 /* icosphere (42 vertices, 120 edges) flattened + 20 cross springs */

#define BCS_PLACE_RANDOM  0    /* uniform in the lumen, non-overlapping (rejection) */
#define BCS_PLACE_LATTICE 1    /* jittered cubic lattice at a packing fraction, from the inlet */

typedef struct bcs_scene_cell {
    int32_t kind;              /* BCS_CELL_* */
    int32_t count;
    float radius;              /* model radius */
    float thickness;           /* DISC42 flattening factor along y */
    float contact_radius;      /* > 0 overrides the derived collision radius */
} bcs_scene_cell;

typedef struct bcs_scene_params {
    int32_t vein_kind;
    float radius;              /* cylinder / trunk radius */
    float length;              /* cylinder / trunk length */
    int32_t segments;          /* cylinder: vertices per ring */
    int32_t rings;             /* cylinder: bands along the axis */
    float branch_radius;
    float branch_length;
    float branch_angle_deg;    /* angle of each branch to the trunk axis */
    float mesh_spacing;        /* Y: marching-tetrahedra voxel size */
    float scale;               /* uniform geometric scale applied to the vein */
    int32_t n_cells;           /* entries used in cells[] (<= 4) */
    bcs_scene_cell cells[4];
    int32_t placement;         /* BCS_PLACE_* */
    float packing_fraction;    /* lattice: target volume fraction of bounding spheres */
    float init_speed;          /* along the trunk axis (-y) */
    float init_noise_sigma;    /* N(0, sigma) per component, per cell */
    uint64_t seed;
    bcs_physics physics;       /* full physics block (bcs_default_physics + overrides) */
} bcs_scene_params;

typedef struct bcs_scene bcs_scene;

typedef struct bcs_scene_view {
    bcs_physics physics;
    int32_t n_types;
    const bcs_cell_type* types;
    bcs_vein_mesh mesh;
    int32_t particle_count;
    const float* positions;    /* SoA 3*particle_count */
    const float* velocities;
    const float* forces;
    const float* cell_centers; /* SoA 3*cell_count: placement centroids */
    int32_t cell_count;
} bcs_scene_view;

/* Benchmark workloads:
 *   0: 500 icosahedron cells (r 2.0) in a straight cylinder r 40 x 400 (64 x 100 bands), dt 0.01
 *   1: 20,000 icosahedron cells in a Y vein (trunk r 60 x 600, branches r 42 at +/-30 deg), dt 0.005
 *   2: 200,000 cells (70% DISC42, 30% icosahedron) in the Y vein scaled 3x, 35% lattice packing
 *   4: 1,000,000 DISC42 cells in the preset-2 vein
 * cell_fraction scales the cell counts (1.0 = full preset).  Common physics:
 * k = 50, damping 0.5, contact radius 0.5, particle grid cell 1.0. */
int bcs_scene_preset(int32_t preset, double cell_fraction, uint64_t seed, bcs_scene** out);
int bcs_scene_default_params(int32_t preset, bcs_scene_params* out);
int bcs_scene_build(const bcs_scene_params* params, bcs_scene** out);
/* Parameters of vein `index` (0 <= index < n_veins) of a batched ensemble
 * of straight veins: radius uniform in [20, 80], length 10 radius, cell
 * counts log-uniform in [500, 50000] scaled by cell_fraction (70% 42-particle
 * discs, 30% 12-particle icosahedra, 35% lattice packing), radii and counts
 * paired by rank so that every vein holds its cells, dt 0.005; the other
 * physics as preset 0.  Reproducible from seed. */
int bcs_scene_ensemble_params(int32_t n_veins, int32_t index, uint64_t seed, double cell_fraction,
                              bcs_scene_params* out);
int bcs_scene_get_view(bcs_scene* scene, bcs_scene_view* out);
int bcs_scene_destroy(bcs_scene* scene);

/* Signed distance to the lumen wall of a scene's analytic vein (< 0 inside). */
int bcs_scene_lumen_sdf(bcs_scene* scene, const float* points, int32_t n, float* out);

#ifdef __cplusplus
}
#endif

#endif /* BCS_SCENE_H */
