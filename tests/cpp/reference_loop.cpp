// reference_loop.cpp -- the reference's object construction (src/main.cu:122-138)
// and frame loop (src/main.cu:158-208) compiled against bcs_sim.hpp.  The
// lines between the "verbatim" markers are the reference's own; what its
// compile-time headers provide (the cell list, the vein, particleCount,
// cellWidth, ...) comes from a run-time configuration set just before.
//
// usage: reference_loop <cell_fraction> <frames> <out.bin> [fused]
//   builds the config-0 scene (or a fraction of it), runs the frames, writes
//   the final SoA particle positions, velocities and forces as raw float32 to
//   <out.bin>, and prints the grid members and calculateCellId of particle 0.
// exit status: 0 done, 3 no usable HIP device, 1 other error.
#include "bcs_scene.h"
#include "bcs_sim.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>

namespace {

// The scene as the process-wide configuration (what the reference's
// src/config headers hold), plus its initial state.
bcs::Configuration scene_configuration(const bcs_scene_view& v)
{
    bcs::Configuration c;
    for (int t = 0; t < v.n_types; ++t) {
        const bcs_cell_type& ct = v.types[t];
        bcs::CellTypeDef d;
        d.count = ct.count;
        d.particlesInCell = ct.particles_in_cell;
        d.springStart.assign(ct.spring_start, ct.spring_start + ct.spring_count);
        d.springEnd.assign(ct.spring_end, ct.spring_end + ct.spring_count);
        d.springLength.assign(ct.spring_length, ct.spring_length + ct.spring_count);
        d.vertices.assign(ct.vertices, ct.vertices + 3 * ct.particles_in_cell);
        d.contactRadius = ct.contact_radius;
        c.types.push_back(d);
    }
    const bcs_vein_mesh& m = v.mesh;
    c.veinPositions.assign(m.positions, m.positions + 3 * (size_t)m.vertex_count);
    c.veinIndices.assign(m.indices, m.indices + 3 * (size_t)m.triangle_count);
    c.endingCenters.assign(m.ending_centers, m.ending_centers + 3 * (size_t)m.ending_count);
    c.endingRadii.assign(m.ending_radii, m.ending_radii + m.ending_count);
    c.physics = v.physics;
    c.hasPhysics = true;
    return c;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <cell_fraction> <frames> <out.bin> [fused]\n", argv[0]);
        return 1;
    }
    const double fraction = std::atof(argv[1]);
    const int frames = std::atoi(argv[2]);
    const bool fused = argc > 4;
    bcs_scene* scene = nullptr;
    bcs::check(bcs_scene_preset(0, fraction, 1, &scene), "bcs_scene_preset");
    bcs_scene_view v{};
    bcs::check(bcs_scene_get_view(scene, &v), "bcs_scene_get_view");
    bcs::configure(scene_configuration(v));

    // the names the reference's compile-time headers define
    const int particleGridGpu = 0, veinGridGpu = 0;
    const int particleCount = v.particle_count, triangleCount = v.mesh.triangle_count;
    const float cellWidth = v.physics.particle_cell[0], cellHeight = v.physics.particle_cell[1],
                cellDepth = v.physics.particle_cell[2];
    const float cellWidthTriangles = v.physics.triangle_cell[0], cellHeightTriangles = v.physics.triangle_cell[1],
                cellDepthTriangles = v.physics.triangle_cell[2];

    try {
        // ---- verbatim: src/main.cu:121-132 (render meshes and the NoGrid branch left out) ----
        // Create blood cells
        BloodCells bloodCells;

        // Create vein triangles
        VeinTriangles triangles;
        // Create grids
        UniformGrid particleGrid(particleGridGpu, particleCount, cellWidth, cellHeight, cellDepth);
        UniformGrid triangleCentersGrid(veinGridGpu, triangleCount, cellWidthTriangles, cellHeightTriangles, cellDepthTriangles);
        // ---- end verbatim ----
        // the reference draws its initial state inside the controller; here it is the scene's
        const size_t n3 = 3 * (size_t)particleCount;
        bloodCells.setState(std::vector<float>(v.positions, v.positions + n3),
                            std::vector<float>(v.velocities, v.velocities + n3),
                            std::vector<float>(v.forces, v.forces + n3));
        {
            NoGrid noGrid;
            try {
                sim::SimulationController c(bloodCells, triangles, &particleGrid, &noGrid);
                std::fprintf(stderr, "NoGrid was accepted\n");
                return 1;
            } catch (const bcs::Error& e) {
                if (e.status != BCS_ERR_UNSUPPORTED) throw;
                std::printf("NoGrid: unsupported\n");
            }
        }
        // ---- verbatim: src/main.cu:138 ----
        // Create the main simulation controller and inject its dependencies
        sim::SimulationController simulationController(bloodCells, triangles, &particleGrid, &triangleCentersGrid);
        // ---- end verbatim ----
        std::printf("cell id of particle 0 %u\n",
                    particleGrid.calculateCellId(bcs::Vec3f{v.positions[0], v.positions[particleCount],
                                                            v.positions[2 * (size_t)particleCount]}));
        std::printf("particle grid %d %d %d cells %lld; triangle grid %d %d %d\n", particleGrid.cellCountX,
                    particleGrid.cellCountY, particleGrid.cellCountZ, particleGrid.cellCount,
                    triangleCentersGrid.cellCountX, triangleCentersGrid.cellCountY, triangleCentersGrid.cellCountZ);
        std::printf("vertex of triangle 5 corner 2 %u\n", triangles.getIndex(0, 5, vertex2));
        bool shouldBeRunning = frames > 0;
        int frameCount = 0;
        if (fused) {
            simulationController.step(frames);
            simulationController.synchronize();
        } else {
            while (shouldBeRunning) {
                // ---- verbatim: src/main.cu:175-176, 198-199, 207-208 ----
                particleGrid.calculateGrid(bloodCells.particles.positions[particleGridGpu], particleCount);
                triangleCentersGrid.calculateGrid(triangles.centers, triangleCount);
                // Calculate particle positions using CUDA
                simulationController.calculateNextFrame();
                // Propagate forces into velocities and positions
                simulationController.propagateAll();
                // ---- end verbatim ----
                shouldBeRunning = ++frameCount < frames;
            }
        }
        std::printf("device arrays %s\n", particleGrid.gridCellIds[0] && particleGrid.particleIds[0] &&
                                                  bloodCells.initialRadiuses[0] && triangles.indices[0] &&
                                                  simulationController.cellModelsBoundingSpheres[0]
                                              ? "set"
                                              : "missing");
        std::vector<float> p, vel, f;
        simulationController.getParticles(p, vel, f);
        std::ofstream out(argv[3], std::ios::binary);
        out.write(reinterpret_cast<const char*>(p.data()), p.size() * 4);
        out.write(reinterpret_cast<const char*>(vel.data()), vel.size() * 4);
        out.write(reinterpret_cast<const char*>(f.data()), f.size() * 4);
        std::printf("frames %d particles %d flags 0x%x smallest radius %g initial cell 0 %g %g %g\n", frames,
                    particleCount, simulationController.errorFlags(), simulationController.smallestRadiusInType.at(0),
                    simulationController.initialCellPositions.at(0).x, simulationController.initialCellPositions.at(0).y,
                    simulationController.initialCellPositions.at(0).z);
    } catch (const bcs::Error& e) {
        bcs_scene_destroy(scene);
        std::fprintf(stderr, "%s\n", e.what());
        return (e.status == BCS_ERR_NO_DEVICE || e.status == BCS_ERR_HIP) ? 3 : 1;
    }
    bcs_scene_destroy(scene);
    return 0;
}
