/*
 * bcs_sim.hpp -- C++ drop-in for the reference's step classes, header-only
 * over the C-ABI of bcs.h.
 *
 * The reference builds its objects (src/main.cu:122-138) and runs a frame
 * (src/main.cu:158-208) as
 *
 *     BloodCells bloodCells;
 *     VeinTriangles triangles;
 *     UniformGrid particleGrid(particleGridGpu, particleCount, cellWidth, cellHeight, cellDepth);
 *     UniformGrid triangleCentersGrid(veinGridGpu, triangleCount, cellWidthTriangles, ...);
 *     sim::SimulationController simulationController(bloodCells, triangles, &particleGrid, &triangleCentersGrid);
 *     ...
 *     particleGrid.calculateGrid(bloodCells.particles.positions[particleGridGpu], particleCount);
 *     triangleCentersGrid.calculateGrid(triangles.centers, triangleCount);
 *     simulationController.calculateNextFrame();
 *     simulationController.propagateAll();
 *
 * and the same lines compile against this header.  The classes keep the
 * reference's names, constructor arguments, public members and call order:
 *   UniformGrid, NoGrid, Grid  src/grids/uniform_grid.cuh:15-65, no_grid.cuh:9-27,
 *                              grid_definition.hpp:6
 *   BloodCells                 src/objects/blood_cells.cuh:16-74
 *   VeinTriangles              src/objects/vein_triangles.cuh:24-124
 *   sim::SimulationController  src/simulation/simulation_controller.cuh:24-67
 * Differences: cell types, vein mesh and physics are run-time data -- the
 * default constructors read them from the process-wide configuration set by
 * bcs::configure() where the reference reads its compile-time headers;
 * device buffers belong to the controller's context and the device views are
 * valid once the controller is built; errors throw bcs::Error instead of
 * exiting the process; there is one GPU (index 0 of every per-GPU array).
 */
#ifndef BCS_SIM_HPP
#define BCS_SIM_HPP

#include "bcs.h"
#include "bcs_refconfig.h"

#include <algorithm>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <variant>
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

/* x, y, z of one point (the reference uses glm::vec3 / float3). */
struct Vec3f {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};

/* Run-time form of what the reference's compile-time headers define: the
 * user cell list (src/config/blood_cells_definition.hpp), the vein mesh and
 * endings (src/config/vein_definition.hpp) and the physics constants
 * (src/config/physics.hpp, simulation.hpp). */
struct Configuration {
    std::vector<CellTypeDef> types;  // definition order; BloodCells() folds them like the reference
    std::vector<float> veinPositions, endingCenters, endingRadii;
    std::vector<uint32_t> veinIndices;
    bcs_physics physics{};
    bool hasPhysics = false;         // false: bcs_default_physics()
};

inline Configuration& configuration()
{
    static Configuration c;
    return c;
}

/* Sets the configuration the default constructors of BloodCells and
 * VeinTriangles (and SimulationController's physics) read. */
inline void configure(const Configuration& c) { configuration() = c; }

}  // namespace bcs

/* Triangle corner (reference src/utilities/vertex_index_enum.hpp). */
enum VertexIndex { vertex0 = 0, vertex1 = 1, vertex2 = 2 };

class UniformGrid;
class NoGrid;
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
    /* The cell types of the process-wide configuration, folded like the
     * reference (the reference's BloodCells() reads its compile-time list). */
    BloodCells() : BloodCells(bcs::configuration().types, true) {}

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
    float* initialRadiuses[1] = {nullptr};  // device, per model vertex: distance to the model centroid

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
    /* The vein of the process-wide configuration (the reference's
     * VeinTriangles() reads its compile-time mesh). */
    VeinTriangles()
        : VeinTriangles(bcs::configuration().veinPositions, bcs::configuration().veinIndices,
                        bcs::configuration().endingCenters, bcs::configuration().endingRadii)
    {
    }

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
    /* Absolute vertex index of a triangle corner (src/objects/vein_triangles.cuh:76-79;
     * host side: the reference's is a device function over indices[gpuId]). */
    unsigned int getIndex(int gpuId, int triangleIndex, VertexIndex vertexIndex) const
    {
        (void)gpuId;
        return hostIndices[3 * (size_t)triangleIndex + (size_t)vertexIndex];
    }

    int vertexCount, triangleCount;
    bcs::Vec3View positions, velocities, forces;  // device views (valid once the controller exists)
    bcs::Vec3View centers;                        // triangle centres
    unsigned int* indices[1] = {nullptr};         // device, 3 per triangle

private:
    std::vector<float> hostPositions;
    std::vector<uint32_t> hostIndices;
    std::vector<float> hostEndCenters, hostEndRadii;
};

/* UniformGrid (src/grids/uniform_grid.cuh:15-65): cell sizes at
 * construction; calculateGrid sorts the objects of the controller the grid
 * is attached to (its particle positions or its triangle centres).  The
 * public members mirror the reference's; the device arrays are valid once a
 * controller has been built with this grid. */
class UniformGrid {
public:
    UniformGrid(int gpuId, int objectCount, float cellWidth, float cellHeight, float cellDepth)
        : cellWidth(cellWidth), cellHeight(cellHeight), cellDepth(cellDepth), objectCount(objectCount), gpuId(gpuId)
    {
    }

    /* src/grids/uniform_grid.cu:129-155 */
    void calculateGrid(const bcs::Vec3View& positions, int count)
    {
        if (!ctx) throw bcs::Error(BCS_ERR_STATE, "UniformGrid: not attached to a SimulationController");
        if (positions.x != attachedTo.x || count != objectCount)
            throw bcs::Error(BCS_ERR_INVALID_ARGUMENT,
                             "UniformGrid::calculateGrid: objects must be the attached controller's state");
        bcs::check(bcs_sim_calculate_grid(ctx, which), "calculateGrid");
        if (denseRanges) refreshView();
    }

    /* calculateIdForCell of one position (src/grids/uniform_grid.cuh:60,
     * src/grids/uniform_grid.cu:24-36): the key the grid build gives it.  Host
     * side; V3 is any type with float members x, y, z (float3, glm::vec3). */
    template <class V3>
    unsigned int calculateCellId(const V3& position) const
    {
        if (!ctx) throw bcs::Error(BCS_ERR_STATE, "UniformGrid: not attached");
        const float p[3] = {(float)position.x, (float)position.y, (float)position.z};
        uint32_t id = 0;
        bcs::check(bcs_sim_cell_ids(ctx, which, p, 1, &id), "calculateCellId");
        return id;
    }

    /* Keep gridCellStarts / gridCellEnds as dense per-cell-id arrays,
     * refreshed after every calculateGrid (cellCount int32 each; empty cells
     * hold start 0, end -1).  Off by default: the frame kernels do not need
     * them. */
    void useDenseCellRanges(bool on = true)
    {
        denseRanges = on;
        if (ctx && on) refreshView();
    }

    /* gridCellIds / particleIds (sorted) and the occupied cells with inclusive
     * [start, end] ranges (gridCellStarts / gridCellEnds at those cells). */
    void download(std::vector<uint32_t>& cellIds, std::vector<int32_t>& particleIdsOut,
                  std::vector<uint32_t>& occupied, std::vector<int32_t>& starts, std::vector<int32_t>& ends) const
    {
        if (!ctx) throw bcs::Error(BCS_ERR_STATE, "UniformGrid: not attached");
        cellIds.resize(objectCount);
        particleIdsOut.resize(objectCount);
        occupied.resize(objectCount);
        starts.resize(objectCount);
        ends.resize(objectCount);
        int32_t n = 0;
        bcs::check(bcs_sim_get_grid(ctx, which, cellIds.data(), particleIdsOut.data(), &n, occupied.data(),
                                    starts.data(), ends.data()),
                   "bcs_sim_get_grid");
        occupied.resize(n);
        starts.resize(n);
        ends.resize(n);
    }

    float cellWidth, cellHeight, cellDepth;
    int cellCountX = 0, cellCountY = 0, cellCountZ = 0;
    long long cellCount = 0;
    int objectCount;
    /* device, per GPU (one here); gridCellStarts / gridCellEnds stay null
     * unless useDenseCellRanges() was called */
    unsigned int* gridCellIds[1] = {nullptr};
    int* particleIds[1] = {nullptr};
    int* gridCellStarts[1] = {nullptr};
    int* gridCellEnds[1] = {nullptr};

private:
    friend class sim::SimulationController;
    void attach(bcs_sim* c, int32_t w, const bcs::Vec3View& objects)
    {
        ctx = c;
        which = w;
        attachedTo = objects;
        refreshView();
    }
    void detach()
    {
        ctx = nullptr;
        gridCellIds[0] = nullptr;
        particleIds[0] = nullptr;
        gridCellStarts[0] = gridCellEnds[0] = nullptr;
    }
    void refreshView()
    {
        bcs_grid_view v{};
        bcs::check(bcs_sim_grid_view(ctx, which, denseRanges ? 1 : 0, &v), "bcs_sim_grid_view");
        cellCountX = v.cell_counts[0];
        cellCountY = v.cell_counts[1];
        cellCountZ = v.cell_counts[2];
        cellCount = (long long)v.cell_count;
        gridCellIds[0] = v.grid_cell_ids;
        particleIds[0] = v.particle_ids;
        gridCellStarts[0] = v.grid_cell_starts;
        gridCellEnds[0] = v.grid_cell_ends;
    }
    int gpuId;
    bcs_sim* ctx = nullptr;
    int32_t which = BCS_GRID_PARTICLES;
    bool denseRanges = false;
    bcs::Vec3View attachedTo;
};

/* NoGrid (src/grids/no_grid.cuh:9-27): the reference's brute-force search
 * without a grid.  Accepted by the Grid variant; a controller built with it
 * reports BCS_ERR_UNSUPPORTED (only the uniform-grid step is built here). */
class NoGrid {
public:
    NoGrid() {}
    void calculateGrid(const bcs::Vec3View&, int) {}
    template <class V3>
    int calculateCellId(const V3&) const
    {
        return 0;
    }
};

/* src/grids/grid_definition.hpp:6 */
using Grid = std::variant<UniformGrid*, NoGrid*>;

namespace bcs {

/* Configuration read from headers in the reference's format
 * (src/config/blood_cells_definition.hpp, blood_cell_presets.hpp,
 * vein_definition.hpp): cell types in definition order and the vein mesh;
 * physics stay the defaults.  configure(loadReferenceConfig(...)) makes the
 * default constructors use them, as the reference's compile-time headers do. */
inline Configuration loadReferenceConfig(const char* definitions, const char* presets, const char* vein)
{
    bcs_ref_config* c = nullptr;
    check(bcs_ref_config_load(definitions, presets, vein, &c), "bcs_ref_config_load");
    Configuration out;
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
    /* src/simulation/simulation_controller.cuh:27.  Both grids must be
     * UniformGrid (the NoGrid search is not built).  Physics: `physics`, else
     * the process-wide configuration's, else bcs_default_physics(); the cell
     * sizes always come from the grids. */
    SimulationController(BloodCells& bloodCells, VeinTriangles& triangles, Grid particleGrid, Grid triangleGrid,
                         const bcs_physics* physics = nullptr)
        : cells(bloodCells), veins(triangles), pgrid(uniform(particleGrid, "particleGrid")),
          tgrid(uniform(triangleGrid, "triangleGrid"))
    {
        bcs_physics ph{};
        if (physics) ph = *physics;
        else if (bcs::configuration().hasPhysics) ph = bcs::configuration().physics;
        else bcs::check(bcs_default_physics(&ph), "bcs_default_physics");
        ph.particle_cell[0] = pgrid.cellWidth;
        ph.particle_cell[1] = pgrid.cellHeight;
        ph.particle_cell[2] = pgrid.cellDepth;
        ph.triangle_cell[0] = tgrid.cellWidth;
        ph.triangle_cell[1] = tgrid.cellHeight;
        ph.triangle_cell[2] = tgrid.cellDepth;
        if (pgrid.objectCount != bloodCells.particleCount || tgrid.objectCount != triangles.triangleCount)
            throw bcs::Error(BCS_ERR_INVALID_ARGUMENT, "SimulationController: grid object counts do not match");
        const std::vector<bcs_cell_type> raw = bloodCells.rawTypes();
        const bcs_vein_mesh mesh = triangles.mesh();
        bcs::check(bcs_sim_create(&ph, raw.data(), (int32_t)raw.size(), &mesh, pgrid.gpuId, &ctx), "bcs_sim_create");
        bcs_device_state ds{};
        bcs::check(bcs_sim_device_state(ctx, &ds), "bcs_sim_device_state");
        const int np = bloodCells.particleCount, nc = bloodCells.bloodCellCount;
        bloodCells.particles.positions[0] = bcs::Vec3View::soa(ds.positions, np);
        bloodCells.particles.velocities[0] = bcs::Vec3View::soa(ds.velocities, np);
        bloodCells.particles.forces[0] = bcs::Vec3View::soa(ds.forces, np);
        bloodCells.particleCenters = bcs::Vec3View::soa(ds.centers, nc);
        bloodCells.initialRadiuses[0] = ds.initial_radii;
        triangles.positions = bcs::Vec3View::soa(ds.vein_positions, triangles.vertexCount);
        triangles.velocities = bcs::Vec3View::soa(ds.vein_velocities, triangles.vertexCount);
        triangles.forces = bcs::Vec3View::soa(ds.vein_forces, triangles.vertexCount);
        triangles.centers = bcs::Vec3View::soa(ds.triangle_centers, triangles.triangleCount);
        triangles.indices[0] = ds.triangle_indices;
        cellModelsBoundingSpheres[0] = ds.model_radii;
        pgrid.attach(ctx, BCS_GRID_PARTICLES, bloodCells.particles.positions[0]);
        tgrid.attach(ctx, BCS_GRID_TRIANGLES, triangles.centers);
        if (!bloodCells.hostPos.empty())
            bcs::check(bcs_sim_set_particles(ctx, bloodCells.hostPos.data(),
                                             bloodCells.hostVel.empty() ? nullptr : bloodCells.hostVel.data(),
                                             bloodCells.hostForce.empty() ? nullptr : bloodCells.hostForce.data()),
                       "bcs_sim_set_particles");
        /* initialCellPositions (simulation_controller.cuh:40): the point each
         * cell's model was placed at, x of its first particle minus the
         * model's first vertex (simulation_controller.cu:224-243) */
        initialCellPositions.assign(nc, bcs::Vec3f{});
        if (!bloodCells.hostPos.empty()) {
            int first = 0, c = 0;
            for (const auto& t : bloodCells.types)
                for (int k = 0; k < t.count; ++k, ++c, first += t.particlesInCell) {
                    initialCellPositions[c].x = bloodCells.hostPos[first] - t.vertices[0];
                    initialCellPositions[c].y = bloodCells.hostPos[(size_t)np + first] - t.vertices[1];
                    initialCellPositions[c].z = bloodCells.hostPos[2 * (size_t)np + first] - t.vertices[2];
                }
        }
        /* smallestRadiusInType (simulation_controller.cu:126-137): the
         * smallest derived radius of each type's model */
        std::vector<bcs_cell_type> derived = raw;
        int nm = 0;
        for (auto& t : derived) {
            t.contact_radius = 0.0f;
            nm += t.particles_in_cell;
        }
        std::vector<float> bs(nm), r0(nm);
        bcs::check(bcs_model_tables(&ph, derived.data(), (int32_t)derived.size(), bs.data(), r0.data()),
                   "bcs_model_tables");
        int m = 0;
        for (const auto& t : derived) {
            float smallest = std::numeric_limits<float>::max();
            for (int a = 0; a < t.particles_in_cell; ++a) smallest = std::min(smallest, bs[m + a]);
            smallestRadiusInType.push_back(smallest);
            m += t.particles_in_cell;
        }
    }
    SimulationController(const SimulationController&) = delete;
    SimulationController& operator=(const SimulationController&) = delete;
    ~SimulationController()
    {
        pgrid.detach();
        tgrid.detach();
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

    /* public members of the reference controller (simulation_controller.cuh:40-42) */
    std::vector<bcs::Vec3f> initialCellPositions;
    std::vector<float> smallestRadiusInType;
    float* cellModelsBoundingSpheres[1] = {nullptr};  // device, per model vertex

private:
    static UniformGrid& uniform(Grid g, const char* which)
    {
        if (auto* u = std::get_if<UniformGrid*>(&g); u && *u) return **u;
        throw bcs::Error(BCS_ERR_UNSUPPORTED,
                         std::string("SimulationController: ") + which + " must be a UniformGrid (NoGrid is not built)");
    }

    BloodCells& cells;
    VeinTriangles& veins;
    UniformGrid& pgrid;
    UniformGrid& tgrid;
    bcs_sim* ctx = nullptr;
};

}  // namespace sim

#endif /* BCS_SIM_HPP */
