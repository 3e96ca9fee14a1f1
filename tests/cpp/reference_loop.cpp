// reference_loop.cpp -- the reference's frame loop (src/main.cu:158-208)
// written against bcs_sim.hpp: build the config-0 scene (or a fraction of it),
// run N frames with the four reference calls, write the final SoA particle
// positions, velocities and forces as raw float32 to argv[3].
//
// usage: reference_loop <cell_fraction> <frames> <out.bin> [fused]
// exit status: 0 done, 3 no usable HIP device, 1 other error.
#include "bcs_scene.h"
#include "bcs_sim.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>

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

    std::vector<bcs::CellTypeDef> defs;
    for (int t = 0; t < v.n_types; ++t) {
        const bcs_cell_type& c = v.types[t];
        bcs::CellTypeDef d;
        d.count = c.count;
        d.particlesInCell = c.particles_in_cell;
        d.springStart.assign(c.spring_start, c.spring_start + c.spring_count);
        d.springEnd.assign(c.spring_end, c.spring_end + c.spring_count);
        d.springLength.assign(c.spring_length, c.spring_length + c.spring_count);
        d.vertices.assign(c.vertices, c.vertices + 3 * c.particles_in_cell);
        d.contactRadius = c.contact_radius;
        defs.push_back(d);
    }
    BloodCells bloodCells(defs, /*fold=*/false);  // scene types are already in final order
    const size_t n3 = 3 * (size_t)v.particle_count;
    bloodCells.setState(std::vector<float>(v.positions, v.positions + n3),
                        std::vector<float>(v.velocities, v.velocities + n3),
                        std::vector<float>(v.forces, v.forces + n3));
    const bcs_vein_mesh& m = v.mesh;
    VeinTriangles triangles(std::vector<float>(m.positions, m.positions + 3 * (size_t)m.vertex_count),
                            std::vector<uint32_t>(m.indices, m.indices + 3 * (size_t)m.triangle_count),
                            std::vector<float>(m.ending_centers, m.ending_centers + 3 * (size_t)m.ending_count),
                            std::vector<float>(m.ending_radii, m.ending_radii + m.ending_count));
    const bcs_physics ph = v.physics;
    const int particleCount = bloodCells.particleCount, triangleCount = triangles.triangleCount;
    UniformGrid particleGrid(0, particleCount, ph.particle_cell[0], ph.particle_cell[1], ph.particle_cell[2]);
    UniformGrid triangleCentersGrid(0, triangleCount, ph.triangle_cell[0], ph.triangle_cell[1], ph.triangle_cell[2]);
    try {
        sim::SimulationController simulationController(bloodCells, triangles, particleGrid, triangleCentersGrid, &ph);
        if (fused) {
            simulationController.step(frames);
            simulationController.synchronize();
        } else {
            for (int f = 0; f < frames; ++f) {
                particleGrid.calculateGrid(bloodCells.particles.positions[0], particleCount);
                triangleCentersGrid.calculateGrid(triangles.centers, triangleCount);
                simulationController.calculateNextFrame();
                simulationController.propagateAll();
            }
        }
        std::vector<float> p, vel, f;
        simulationController.getParticles(p, vel, f);
        std::ofstream out(argv[3], std::ios::binary);
        out.write(reinterpret_cast<const char*>(p.data()), p.size() * 4);
        out.write(reinterpret_cast<const char*>(vel.data()), vel.size() * 4);
        out.write(reinterpret_cast<const char*>(f.data()), f.size() * 4);
        std::printf("frames %d particles %d flags 0x%x smallest radius %g\n", frames, particleCount,
                    simulationController.errorFlags(), simulationController.smallestRadiusInType.at(0));
    } catch (const bcs::Error& e) {
        bcs_scene_destroy(scene);
        std::fprintf(stderr, "%s\n", e.what());
        return (e.status == BCS_ERR_NO_DEVICE || e.status == BCS_ERR_HIP) ? 3 : 1;
    }
    bcs_scene_destroy(scene);
    return 0;
}
