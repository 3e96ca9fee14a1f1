/*
 * bcs_refconfig.h -- run-time loader for the reference's configuration
 * headers (host only, no GPU needed).
 *
 * The reference fixes its scene at compile time in three headers:
 *   definitions  src/config/blood_cells_definition.hpp:13-25
 *                (UserDefinedBloodCellList: BloodCellDef<Count, ParticlesInCell,
 *                 IndicesInCell, Color, Springs, Vertices, Indices, Normals>)
 *   presets      src/config/blood_cell_presets.hpp:13-318
 *                (mp_list<Spring<Start, End, Length[, p]>...> and
 *                 mp_list<mp_float3<X, Y, Z[, p]>...> per preset)
 *   vein         src/config/vein_definition.hpp:12,15965,31768-31776
 *                (veinPositions, veinIndices, VeinEndingCenters, VeinEndingRadii)
 * This loader reads the same files as text -- a user's own presets and
 * meshes in that format included -- and returns the cell types in the
 * definition order (fold them with bcs_fold_reference_types) and the vein mesh.
 * Fixed-point values decode as the reference's templates do
 * (src/meta_factory/mp_helpers.hpp:23-35, blood_cells_def_type.hpp:40-46):
 * float(X) / 10^(p-1) in single precision, p = 7 unless given; mesh positions
 * are static_cast<float> of the double literals.
 */
#ifndef BCS_REFCONFIG_H
#define BCS_REFCONFIG_H

#include "bcs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bcs_ref_config bcs_ref_config;

/* Any path may be NULL (that part is skipped); the presets file is needed
 * when a definitions file is given.  Errors: BCS_ERR_INVALID_ARGUMENT with a
 * message naming the file and the construct that could not be read. */
int bcs_ref_config_load(const char* definitions_path, const char* presets_path, const char* vein_path,
                        bcs_ref_config** out);

/* Views into the loaded data (valid until bcs_ref_config_destroy).  Types are
 * in definition order, contact_radius 0 (derived as in the reference). */
int bcs_ref_config_get(bcs_ref_config* cfg, int32_t* n_types, const bcs_cell_type** types, bcs_vein_mesh* mesh);

int bcs_ref_config_destroy(bcs_ref_config* cfg);

#ifdef __cplusplus
}
#endif

#endif /* BCS_REFCONFIG_H */
