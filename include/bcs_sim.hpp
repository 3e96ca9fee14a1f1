/*
 * bcs_sim.hpp -- C++ drop-in for the reference's step classes, header-only
 * over the C-ABI of bcs.h.
 *
 * A reference frame (src/main.cu:158-208) reads
 *
 *     particleGrid.calculateGrid(bloodCells.particles.positions[0], particleCount);
 *     triangleCentersGrid.calculateGrid(triangles.centers, triangleCount);
 *     simulationController.calculateNextFrame();
 *     simulationController.propagateAll();
 *
 * and the same four lines compile against this header.  The classes keep the
 * reference's names, constructor arguments and call order:
 *   UniformGrid            src/grids/uniform_grid.cuh:15-65
 *   BloodCells             src/objects/blood_cells.cuh:16-74
 *   VeinTriangles          src/objects/vein_triangles.cuh:24-124
 *   sim::SimulationController  src/simulation/simulation_controller.cuh:24-67
 * Differences: cell types and the vein mesh are run-time data (the reference
 * folds them at compile time); device buffers belong to the controller's
 * context and are exposed as SoA views once the controller is built; errors
 * throw bcs::Error instead of exiting the process; there is one GPU.
 */
#ifndef BCS_SIM_HPP
#define BCS_SIM_HPP

#include "bcs.h"
#include "bcs_refconfig.h"

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace bcs {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), status(status) {}
    int status;
};

inline void check(int rc, const char* what)
{
    if (rc != BCS_OK) throw Error(rc, std::string(what) + ": " + bcs_last_error());
}

/* Device SoA view of n float triples (reference cudaVec3,
 * src/utilities/cuda_vec3.cuh:11-73): x[n], y[n], z[n]. */
struct Vec3View {
    float* x = nullptr;
    float* y = nullptr;
    float* z = nullptr;
    int n = 0;
    static Vec3View soa(float* base, int n) { return Vec3View{base, base + n, base + 2 * (size_t)n, n}; }
};

/* One cell type (runtime BloodCellDef, src/meta_factory/blood_cells_def_type.hpp:20-46). */
struct CellTypeDef {
    int count = 0;
    int particlesInCell = 0;
    std::vector<int> springStart, springEnd;
    std::vector<float> springLength;
    std::vector<float> vertices;  // 3 * particlesInCell, xyz interleaved
    float contactRadius = 0.0f;   // <= 0: derived as in the reference
};

}  // namespace bcs

class UniformGrid;
namespace sim {
class SimulationController;
}

/* Particle SoA buffers of one GPU (reference Particles, src/objects/particles.cuh:10-36). */
struct Particles {
    bcs::Vec3View positions[1], velocities[1], forces[1];
};

/* BloodCells: the cell types (in final, folded order) and the particle state. */
class BloodCells {
public:
    /* fold = true merges and orders the definitions like the reference's
     * compile-time folding (src/meta_factory/blood_cell_factory.hpp:21-162). */
    explicit BloodCells(std::vector<bcs::CellTypeDef> userTypes, bool fold = true)
    {
        if (!fold) {
            types = std::move(userTypes);
        } else {
            std::vector<bcs_cell_type> c = rawTypes(userTypes);
            std::vector<int32_t> order(c.size()), counts(c.size());
            int32_t n = 0;
            bcs::check(bcs_fold_reference_types(c.data(), (int32_t)c.size(), order.data(), counts.data(), &n),
                       "bcs_fold_reference_types");
            for (int32_t k = 0; k < n; ++k) {
                bcs::CellTypeDef t = userTypes[order[k]];
                t.count = counts[k];
                types.push_back(t);
            }
        }
        for (const auto& t : types) {
            particleCount += t.count * t.particlesInCell;
            bloodCellCount += t.count;
        }
    }

    /* Initial state, SoA host arrays of 3 * particleCount, copied to the
     * device when the controller is built (empty velocities/forces: zero). */
    void setState(const std::vector<float>& pos, const std::vector<float>& vel, const std::vector<float>& force)
    {
        hostPos = pos;
        hostVel = vel;
        hostForce = force;
    }

    std::vector<bcs_cell_type> rawTypes() const { return rawTypes(types); }

    int particleCount = 0;
    int bloodCellCount = 0;
    std::vector<bcs::CellTypeDef> types;
    Particles particles;             // device views (valid once the controller exists)
    bcs::Vec3View particleCenters;   // device view of the cell centres

private:
    friend class sim::SimulationController;
    static std::vector<bcs_cell_type> rawTypes(const std::vector<bcs::CellTypeDef>& ts)
    {
        std::vector<bcs_cell_type> out;
        for (const auto& t : ts) {
            bcs_cell_type c{};
            c.count = t.count;
            c.particles_in_cell = t.particlesInCell;
            c.spring_count = (int32_t)t.springLength.size();
            c.spring_start = t.springStart.data();
            c.spring_end = t.springEnd.data();
            c.spring_length = t.springLength.data();
            c.vertices = t.vertices.data();
            c.contact_radius = t.contactRadius;
            out.push_back(c);
        }
        return out;
    }
    std::vector<float> hostPos, hostVel, hostForce;
};

/* VeinTriangles: the vein mesh and its vertex state. */
class VeinTriangles {
public:
    VeinTriangles(std::vector<float> positions, std::vector<uint32_t> indices, std::vector<float> endingCenters = {},
                  std::vector<float> endingRadii = {})
        : vertexCount((int)(positions.size() / 3)), triangleCount((int)(indices.size() / 3)),
          hostPositions(std::move(positions)), hostIndices(std::move(indices)), hostEndCenters(std::move(endingCenters)),
          hostEndRadii(std::move(endingRadii))
    {
    }

    bcs_vein_mesh mesh() const
    {
        bcs_vein_mesh m{};
        m.vertex_count = vertexCount;
        m.positions = hostPositions.data();
        m.triangle_count = triangleCount;
        m.indices = hostIndices.data();
        m.ending_count = (int32_t)hostEndRadii.size();
        m.ending_centers = hostEndCenters.data();
        m.ending_radii = hostEndRadii.data();
        return m;
    }
    int getIndex(int triangle, int vertex) const { return (int)hostIndices[3 * (size_t)triangle + vertex]; }

    int vertexCount, triangleCount;
    bcs::Vec3View positions, velocities, forces;  // device views (valid once the controller exists)
    bcs::Vec3View centers;                        // triangle centres

private:
    std::vector<float> hostPositions;
    std::vector<uint32_t> hostIndices;
    std::vector<float> hostEndCenters, hostEndRadii;
};

/* UniformGrid: cell sizes at construction; calculateGrid sorts the given
 * objects (the particle positions or the triangle centres of the controller
 * the grid is attached to). */
class UniformGrid {
public:
    UniformGrid(int gpuId, int objectCount, float cellWidth, float cellHeight, float cellDepth)
        : gpuId(gpuId), objectCount(objectCount), cellWidth(cellWidth), cellHeight(cellHeight), cellDepth(cellDepth)
    {
    }

    void calculateGrid(const bcs::Vec3View& positions, int count)
    {
        if (!ctx) throw bcs::Error(BCS_ERR_STATE, "UniformGrid: not attached to a SimulationController");
        if (positions.x != attachedTo.x || count != objectCount)
            throw bcs::Error(BCS_ERR_INVALID_ARGUMENT,
                             "UniformGrid::calculateGrid: objects must be the attached controller's state");
        bcs::check(bcs_sim_calculate_grid(ctx, which), "calculateGrid");
    }

    /* gridCellIds / particleIds (sorted) and the occupied cells with inclusive
     * [start, end] ranges (gridCellStarts / gridCellEnds at those cells). */
    void download(std::vector<uint32_t>& cellIds, std::vector<int32_t>& particleIds, std::vector<uint32_t>& occupied,
                  std::vector<int32_t>& starts, std::vector<int32_t>& ends) const
    {
        if (!ctx) throw bcs::Error(BCS_ERR_STATE, "UniformGrid: not attached");
        cellIds.resize(objectCount);
        particleIds.resize(objectCount);
        occupied.resize(objectCount);
        starts.resize(objectCount);
        ends.resize(objectCount);
        int32_t n = 0;
        bcs::check(bcs_sim_get_grid(ctx, which, cellIds.data(), particleIds.data(), &n, occupied.data(), starts.data(),
                                    ends.data()),
                   "bcs_sim_get_grid");
        occupied.resize(n);
        starts.resize(n);
        ends.resize(n);
    }

    int gpuId, objectCount;
    float cellWidth, cellHeight, cellDepth;

private:
    friend class sim::SimulationController;
    bcs_sim* ctx = nullptr;
    int32_t which = BCS_GRID_PARTICLES;
    bcs::Vec3View attachedTo;
};

namespace bcs {

/* Cell types (definition order) and vein mesh read from configuration headers
 * in the reference's format (src/config/blood_cells_definition.hpp,
 * blood_cell_presets.hpp, vein_definition.hpp); BloodCells(types) folds them
 * like the reference. */
struct ReferenceConfig {
    std::vector<CellTypeDef> types;
    std::vector<float> veinPositions, endingCenters, endingRadii;
    std::vector<uint32_t> veinIndices;
    VeinTriangles vein() const { return VeinTriangles(veinPositions, veinIndices, endingCenters, endingRadii); }
};

inline ReferenceConfig loadReferenceConfig(const char* definitions, const char* presets, const char* vein)
{
    bcs_ref_config* c = nullptr;
    check(bcs_ref_config_load(definitions, presets, vein, &c), "bcs_ref_config_load");
    ReferenceConfig out;
    int32_t n = 0;
    const bcs_cell_type* t = nullptr;
    bcs_vein_mesh m{};
    const int rc = bcs_ref_config_get(c, &n, &t, &m);
    if (rc == BCS_OK) {
        for (int32_t k = 0; k < n; ++k) {
            CellTypeDef d;
            d.count = t[k].count;
            d.particlesInCell = t[k].particles_in_cell;
            d.springStart.assign(t[k].spring_start, t[k].spring_start + t[k].spring_count);
            d.springEnd.assign(t[k].spring_end, t[k].spring_end + t[k].spring_count);
            d.springLength.assign(t[k].spring_length, t[k].spring_length + t[k].spring_count);
            d.vertices.assign(t[k].vertices, t[k].vertices + 3 * t[k].particles_in_cell);
            out.types.push_back(d);
        }
        if (m.vertex_count > 0) {
            out.veinPositions.assign(m.positions, m.positions + 3 * (size_t)m.vertex_count);
            out.veinIndices.assign(m.indices, m.indices + 3 * (size_t)m.triangle_count);
            if (m.ending_count > 0) {
                out.endingCenters.assign(m.ending_centers, m.ending_centers + 3 * (size_t)m.ending_count);
                out.endingRadii.assign(m.ending_radii, m.ending_radii + m.ending_count);
            }
        }
    }
    bcs_ref_config_destroy(c);
    check(rc, "bcs_ref_config_get");
    return out;
}

}  // namespace bcs

namespace sim {

/* SimulationController: owns the device context (all buffers, one stream). */
class SimulationController {
public:
    SimulationController(BloodCells& bloodCells, VeinTriangles& triangles, UniformGrid& particleGrid,
                         UniformGrid& triangleGrid, const bcs_physics* physics = nullptr)
        : cells(bloodCells), veins(triangles), pgrid(particleGrid), tgrid(triangleGrid)
    {
        bcs_physics ph{};
        if (physics) ph = *physics;
        else bcs::check(bcs_default_physics(&ph), "bcs_default_physics");
        ph.particle_cell[0] = particleGrid.cellWidth;
        ph.particle_cell[1] = particleGrid.cellHeight;
        ph.particle_cell[2] = particleGrid.cellDepth;
        ph.triangle_cell[0] = triangleGrid.cellWidth;
        ph.triangle_cell[1] = triangleGrid.cellHeight;
        ph.triangle_cell[2] = triangleGrid.cellDepth;
        if (particleGrid.objectCount != bloodCells.particleCount || triangleGrid.objectCount != triangles.triangleCount)
            throw bcs::Error(BCS_ERR_INVALID_ARGUMENT, "SimulationController: grid object counts do not match");
        const std::vector<bcs_cell_type> raw = bloodCells.rawTypes();
        const bcs_vein_mesh mesh = triangles.mesh();
        bcs::check(bcs_sim_create(&ph, raw.data(), (int32_t)raw.size(), &mesh, particleGrid.gpuId, &ctx),
                   "bcs_sim_create");
        bcs_device_state ds{};
        bcs::check(bcs_sim_device_state(ctx, &ds), "bcs_sim_device_state");
        const int np = bloodCells.particleCount, nc = bloodCells.bloodCellCount;
        bloodCells.particles.positions[0] = bcs::Vec3View::soa(ds.positions, np);
        bloodCells.particles.velocities[0] = bcs::Vec3View::soa(ds.velocities, np);
        bloodCells.particles.forces[0] = bcs::Vec3View::soa(ds.forces, np);
        bloodCells.particleCenters = bcs::Vec3View::soa(ds.centers, nc);
        triangles.positions = bcs::Vec3View::soa(ds.vein_positions, triangles.vertexCount);
        triangles.velocities = bcs::Vec3View::soa(ds.vein_velocities, triangles.vertexCount);
        triangles.forces = bcs::Vec3View::soa(ds.vein_forces, triangles.vertexCount);
        triangles.centers = bcs::Vec3View::soa(ds.triangle_centers, triangles.triangleCount);
        particleGrid.ctx = ctx;
        particleGrid.which = BCS_GRID_PARTICLES;
        particleGrid.attachedTo = bloodCells.particles.positions[0];
        triangleGrid.ctx = ctx;
        triangleGrid.which = BCS_GRID_TRIANGLES;
        triangleGrid.attachedTo = triangles.centers;
        if (!bloodCells.hostPos.empty())
            bcs::check(bcs_sim_set_particles(ctx, bloodCells.hostPos.data(),
                                             bloodCells.hostVel.empty() ? nullptr : bloodCells.hostVel.data(),
                                             bloodCells.hostForce.empty() ? nullptr : bloodCells.hostForce.data()),
                       "bcs_sim_set_particles");
        // public tables of the reference controller (simulation_controller.cuh:40-42)
        int nm = 0;
        for (const auto& t : raw) nm += t.particles_in_cell;
        cellModelsBoundingSpheres.resize(nm);
        std::vector<float> r0(nm);
        bcs::check(bcs_model_tables(&ph, raw.data(), (int32_t)raw.size(), cellModelsBoundingSpheres.data(), r0.data()),
                   "bcs_model_tables");
        int m = 0;
        for (const auto& t : raw) {
            float s = 3.4e38f;
            for (int a = 0; a < t.particles_in_cell; ++a) s = std::min(s, cellModelsBoundingSpheres[m + a]);
            smallestRadiusInType.push_back(s);
            m += t.particles_in_cell;
        }
    }
    SimulationController(const SimulationController&) = delete;
    SimulationController& operator=(const SimulationController&) = delete;
    ~SimulationController()
    {
        pgrid.ctx = nullptr;
        tgrid.ctx = nullptr;
        if (ctx) bcs_sim_destroy(ctx);
    }

    void calculateNextFrame() { bcs::check(bcs_sim_calculate_next_frame(ctx), "calculateNextFrame"); }
    void propagateAll() { bcs::check(bcs_sim_propagate_all(ctx), "propagateAll"); }

    /* n whole frames (both grids, calculateNextFrame, propagateAll) through the
     * fused, graph-replayed kernels; asynchronous until synchronize(). */
    void step(int n) { bcs::check(bcs_sim_step(ctx, n), "step"); }
    void synchronize() { bcs::check(bcs_sim_synchronize(ctx), "synchronize"); }

    void getParticles(std::vector<float>& pos, std::vector<float>& vel, std::vector<float>& force)
    {
        const size_t n = 3 * (size_t)cells.particleCount;
        pos.resize(n);
        vel.resize(n);
        force.resize(n);
        bcs::check(bcs_sim_get_particles(ctx, pos.data(), vel.data(), force.data()), "bcs_sim_get_particles");
    }
    uint32_t errorFlags(bool clear = false)
    {
        uint32_t f = 0;
        bcs::check(bcs_sim_error_flags(ctx, clear ? 1 : 0, &f), "bcs_sim_error_flags");
        return f;
    }
    bcs_sim* handle() const { return ctx; }

    std::vector<float> smallestRadiusInType;
    std::vector<float> cellModelsBoundingSpheres;

private:
    BloodCells& cells;
    VeinTriangles& veins;
    UniformGrid& pgrid;
    UniformGrid& tgrid;
    bcs_sim* ctx = nullptr;
};

}  // namespace sim

#endif /* BCS_SIM_HPP */
