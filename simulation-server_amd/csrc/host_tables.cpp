// host_tables.cpp -- run-time restatement of the reference's compile-time
// configuration folding (src/meta_factory/*.hpp) and of the host-side
// initialisation tables (src/simulation/simulation_controller.cu:93-153).
#include "bcs_internal.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

namespace bcs {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

}  // namespace bcs

using bcs::fail;

extern "C" const char* bcs_last_error(void) { return bcs::g_last_error.c_str(); }

// src/config/physics.hpp:4-34, src/config/simulation.hpp:3-21, vein_factory.hpp:88
extern "C" int bcs_default_physics(bcs_physics* out)
{
    if (!out) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_default_physics: null output");
    std::memset(out, 0, sizeof(*out));
    out->dt = 0.008f;
    out->use_blood_flow = 1;
    out->enable_reaction_force = 1;
    out->enable_big_cells_brake = 1;
    for (int d = 0; d < 3; ++d) {
        out->particle_cell[d] = 2.0f;
        out->triangle_cell[d] = 25.0f;
    }
    out->bounding_spheres_coeff = 3;
    out->grid_y_margin = 40.0f;
    out->grid_xz_margin = 40.0f;
    out->min_spawn_y = -20.0f;
    out->velocity_collision_damping = 0.96f;
    out->particle_k_sniff = 25.0f;
    out->vein_k_sniff = 0.1f;
    out->particle_d_fact = 10.0f;
    out->vein_d_fact = 0.7f;
    out->vein_collision_force_intensity = 0.005f;
    out->viscous_damping = 0.39f;
    out->collision_spring_coeff = 6.0f;
    out->collision_damping_coeff = 6.0f;
    out->collision_shear_coeff = 4.0f;
    out->max_cell_size_factor_before_brake = 1.5f;
    out->big_particle_braking_intensity = 0.8f;
    out->init_velocity[0] = 0.0f;
    out->init_velocity[1] = -80.0f;
    out->init_velocity[2] = 0.0f;
    out->vein_impact_distance = 6.0f;
    out->vein_impact_minimal_force_distance = 0.001f;
    out->gravity[0] = 0.0f;
    out->gravity[1] = -25.0f;
    out->gravity[2] = 0.0f;
    out->cylinder_radius = 50.0f;
    out->seed = 0x5eedULL;
    out->dynamic_triangle_grid = 0;
    return BCS_OK;
}

namespace {

// IsDuplicate, src/meta_factory/blood_cell_factory.hpp:53-57: same particle
// count and the same spring list (compared element-wise here; the reference
// compares the Spring<...> types, i.e. the same integers and precision).
bool is_duplicate(const bcs_cell_type& a, const bcs_cell_type& b)
{
    if (a.particles_in_cell != b.particles_in_cell || a.spring_count != b.spring_count) return false;
    for (int i = 0; i < a.spring_count; ++i) {
        if (a.spring_start[i] != b.spring_start[i] || a.spring_end[i] != b.spring_end[i] ||
            a.spring_length[i] != b.spring_length[i])
            return false;
    }
    return true;
}

// isPowerOfTwo + orderBloodCells, blood_cell_factory.hpp:119-146.  The
// "power of two" test is inverted (n & (n-1)) and the parity tests use `& 0`,
// so only the first two rules can fire.
bool ref_is_pow2(int n) { return (n & (n - 1)) != 0; }
bool ref_order(int p1, int p2)
{
    if (ref_is_pow2(p1) && !ref_is_pow2(p2)) return true;
    if (!ref_is_pow2(p1) && ref_is_pow2(p2)) return false;
    if ((p1 & 0) && (p2 & 1)) return true;
    if ((p1 & 1) && (p2 & 0)) return false;
    return true;
}

// boost::mp11 mp_sort (Libraries/include/boost/mp11/algorithm.hpp:581-623):
// quicksort with the first element as pivot; elements U with P<U, pivot>
// go (in order) before it, the rest after.
void mp_sort_like(std::vector<int>& items, const std::vector<int>& ppc)
{
    if (items.size() <= 1) return;
    int pivot = items[0];
    std::vector<int> s1, s2;
    for (size_t i = 1; i < items.size(); ++i) {
        if (ref_order(ppc[items[i]], ppc[pivot])) s1.push_back(items[i]);
        else s2.push_back(items[i]);
    }
    mp_sort_like(s1, ppc);
    mp_sort_like(s2, ppc);
    items.clear();
    items.insert(items.end(), s1.begin(), s1.end());
    items.push_back(pivot);
    items.insert(items.end(), s2.begin(), s2.end());
}

}  // namespace

// FoldedBloodCellList + mp_unique_if + mp_sort, blood_cell_factory.hpp:82-162.
extern "C" int bcs_fold_reference_types(const bcs_cell_type* user, int32_t n_user, int32_t* order_out,
                                        int32_t* counts_out, int32_t* n_out)
{
    if (!user || n_user < 0 || !order_out || !counts_out || !n_out)
        return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_fold_reference_types: null argument");
    // folded count of entry i = sum of counts of all entries duplicate with it
    std::vector<int> folded_count(n_user, 0);
    for (int i = 0; i < n_user; ++i)
        for (int j = 0; j < n_user; ++j)
            if (is_duplicate(user[i], user[j])) folded_count[i] += user[j].count;
    // mp_unique_if keeps the first of every duplicate class
    std::vector<int> uniq;
    for (int i = 0; i < n_user; ++i) {
        bool dup = false;
        for (int k : uniq) dup = dup || is_duplicate(user[k], user[i]);
        if (!dup) uniq.push_back(i);
    }
    std::vector<int> ppc(n_user);
    for (int i = 0; i < n_user; ++i) ppc[i] = user[i].particles_in_cell;
    mp_sort_like(uniq, ppc);
    for (size_t k = 0; k < uniq.size(); ++k) {
        order_out[k] = uniq[k];
        counts_out[k] = folded_count[uniq[k]];
    }
    *n_out = (int32_t)uniq.size();
    return BCS_OK;
}

// minX..maxZ, src/meta_factory/vein_factory.hpp:21-86
extern "C" int bcs_grid_bounds(const bcs_vein_mesh* mesh, const bcs_physics* phys, float bmin[3], float bmax[3])
{
    if (!mesh || !phys || !bmin || !bmax || mesh->vertex_count < 1 || !mesh->positions)
        return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_grid_bounds: empty mesh");
    for (int d = 0; d < 3; ++d) {
        float lo = mesh->positions[d], hi = mesh->positions[d];
        for (int i = 1; i < mesh->vertex_count; ++i) {
            float v = mesh->positions[3 * i + d];
            lo = std::min(lo, v);
            hi = std::max(hi, v);
        }
        float margin = d == 1 ? phys->grid_y_margin : phys->grid_xz_margin;
        bmin[d] = lo - margin;
        bmax[d] = hi + margin;
    }
    return BCS_OK;
}

// calculateSpringLengths, src/meta_factory/vein_factory.hpp:130-174
extern "C" int bcs_vein_neighbors(const bcs_vein_mesh* mesh, int32_t* ids, float* lengths)
{
    if (!mesh || !ids || !lengths) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_vein_neighbors: null argument");
    const int nv = mesh->vertex_count, nt = mesh->triangle_count;
    for (int i = 0; i < 3 * nt; ++i)
        if (mesh->indices[i] >= (uint32_t)nv) return fail(BCS_ERR_INVALID_ARGUMENT, "vein index out of range");
    std::vector<int64_t> off(nv + 1, 0);
    for (int i = 0; i < 3 * nt; ++i) off[mesh->indices[i] + 1] += 2;
    for (int i = 0; i < nv; ++i) off[i + 1] += off[i];
    std::vector<uint32_t> lst(off[nv]);
    std::vector<int64_t> fill(off.begin(), off.end() - 1);
    for (int t = 0; t < nt; ++t) {
        uint32_t a = mesh->indices[3 * t], b = mesh->indices[3 * t + 1], c = mesh->indices[3 * t + 2];
        lst[fill[a]++] = b; lst[fill[a]++] = c;
        lst[fill[b]++] = a; lst[fill[b]++] = c;
        lst[fill[c]++] = a; lst[fill[c]++] = b;
    }
    for (int s = 0; s < bcs::kVeinSlots; ++s)
        for (int i = 0; i < nv; ++i) {
            ids[(size_t)s * nv + i] = -1;
            lengths[(size_t)s * nv + i] = -1.0f;
        }
    const float* P = mesh->positions;
    for (int i = 0; i < nv; ++i) {
        std::sort(lst.begin() + off[i], lst.begin() + off[i + 1]);
        int64_t m = std::min<int64_t>(off[i + 1] - off[i], bcs::kVeinSlots);
        for (int64_t j = 0; j < m; ++j) {
            uint32_t nb = lst[off[i] + j];
            float dx = P[3 * i] - P[3 * nb], dy = P[3 * i + 1] - P[3 * nb + 1], dz = P[3 * i + 2] - P[3 * nb + 2];
            float len = std::sqrt(dx * dx + dy * dy + dz * dz);
            ids[(size_t)j * nv + i] = (int32_t)nb;
            lengths[(size_t)j * nv + i] = len < 0 ? -len : len;
        }
    }
    return BCS_OK;
}

// generateBoundingSpheres, src/simulation/simulation_controller.cu:93-153:
// collision radius = min distance to another model vertex / (2*coeff)
// (pow/sqrt in double, stored as float); initial radius = |m_a - centroid|.
extern "C" int bcs_model_tables(const bcs_physics* phys, const bcs_cell_type* types, int32_t n_types,
                                float* bounding_spheres, float* initial_radii)
{
    if (!phys || !types || n_types < 1) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_model_tables: bad arguments");
    int m0 = 0;
    for (int t = 0; t < n_types; ++t) {
        const int p = types[t].particles_in_cell;
        const float* mv = types[t].vertices;
        if (p < 1 || !mv) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_model_tables: type without vertices");
        float cx = 0.f, cy = 0.f, cz = 0.f;
        for (int a = 0; a < p; ++a) {
            float best = FLT_MAX;
            for (int b = 0; b < p; ++b) {
                if (a == b) continue;
                float dx = mv[3 * a] - mv[3 * b], dy = mv[3 * a + 1] - mv[3 * b + 1], dz = mv[3 * a + 2] - mv[3 * b + 2];
                float len = (float)(std::sqrt(std::pow((double)dx, 2) + std::pow((double)dy, 2) + std::pow((double)dz, 2)) /
                                    (double)(2 * phys->bounding_spheres_coeff));
                if (len < best) best = len;
            }
            if (bounding_spheres)
                bounding_spheres[m0 + a] = types[t].contact_radius > 0.f ? types[t].contact_radius : best;
            cx = cx + mv[3 * a];
            cy = cy + mv[3 * a + 1];
            cz = cz + mv[3 * a + 2];
        }
        cx = cx / (float)p;
        cy = cy / (float)p;
        cz = cz / (float)p;
        for (int a = 0; a < p; ++a) {
            float dx = mv[3 * a] - cx, dy = mv[3 * a + 1] - cy, dz = mv[3 * a + 2] - cz;
            if (initial_radii) initial_radii[m0 + a] = std::sqrt(dx * dx + dy * dy + dz * dz);
        }
        m0 += p;
    }
    return BCS_OK;
}
