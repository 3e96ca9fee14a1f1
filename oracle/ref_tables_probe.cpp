// ref_tables_probe.cpp -- dumps the reference's own compile-time tables.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile (`make ref`) against
// the reference's headers where they lie under /root/reference; it is the one
// part of the reference path that builds here (host-only C++17 template
// folding: src/meta_factory/blood_cell_factory.hpp, src/meta_factory/vein_factory.hpp,
// src/config/*.hpp).  Output: a JSON document on stdout with the folded
// blood-cell tables and grid bounds, plus binary dumps of the reference vein
// mesh and its spring-neighbour tables into the directory given as argv[1].
#include "meta_factory/blood_cell_factory.hpp"
#include "meta_factory/vein_factory.hpp"
#include "config/physics.hpp"

#include <cstdio>
#include <string>
#include <type_traits>

template <class Def>
static void dump_def(const char* indent)
{
    using Springs = typename Def::List;
    using Verts = typename Def::Vertices;
    std::printf("%s{\"count\": %d, \"ppc\": %d, \"springs\": [", indent, Def::count, Def::particlesInCell);
    bool first = true;
    mp_for_each<mp_iota_c<mp_size<Springs>::value>>([&](auto j) {
        using S = mp_at_c<Springs, j>;
        std::printf("%s[%d, %d, %.9g]", first ? "" : ", ", S::start, S::end, (double)S::length);
        first = false;
    });
    std::printf("], \"vertices\": [");
    first = true;
    mp_for_each<mp_iota_c<mp_size<Verts>::value>>([&](auto j) {
        using V = mp_at_c<Verts, j>;
        std::printf("%s%.9g, %.9g, %.9g", first ? "" : ", ", (double)V::x, (double)V::y, (double)V::z);
        first = false;
    });
    std::printf("]}");
}

template <class A, class B>
static constexpr bool same_def()
{
    return A::particlesInCell == B::particlesInCell && std::is_same_v<typename A::List, typename B::List>;
}

template <class T, std::size_t N>
static void dump_int_array(const char* name, const std::array<T, N>& a)
{
    std::printf("\"%s\": [", name);
    for (std::size_t i = 0; i < N; ++i) std::printf("%s%d", i ? ", " : "", (int)a[i]);
    std::printf("],\n");
}

int main(int argc, char** argv)
{
    std::string outdir = argc > 1 ? argv[1] : ".";
    std::printf("{\n\"user_types\": [\n");
    constexpr int nuser = mp_size<UserDefinedBloodCellList>::value;
    mp_for_each<mp_iota_c<nuser>>([&](auto i) {
        dump_def<mp_at_c<UserDefinedBloodCellList, i>>(i ? ",\n  " : "  ");
    });
    std::printf("\n],\n\"folded_types\": [\n");
    mp_for_each<mp_iota_c<bloodCellTypeCount>>([&](auto i) {
        using D = mp_at_c<BloodCellList, i>;
        int match = -1;
        mp_for_each<mp_iota_c<nuser>>([&](auto u) {
            if (match < 0 && same_def<D, mp_at_c<UserDefinedBloodCellList, u>>()) match = (int)u;
        });
        std::printf("%s{\"count\": %d, \"ppc\": %d, \"first_user_index\": %d}", i ? ",\n  " : "  ", D::count,
                    D::particlesInCell, match);
    });
    std::printf("\n],\n");
    std::printf("\"particleCount\": %d,\n\"bloodCellCount\": %d,\n\"particleDistinctCellsCount\": %d,\n"
                "\"totalGraphSize\": %d,\n\"bloodCellTypeCount\": %d,\n",
                particleCount, bloodCellCount, particleDistinctCellsCount, totalGraphSize, bloodCellTypeCount);
    dump_int_array("particleStarts", particleStarts);
    dump_int_array("bloodCellTypesStarts", bloodCellTypesStarts);
    dump_int_array("bloodCellModelStarts", bloodCellModelStarts);
    dump_int_array("accumulatedGraphSizes", accumulatedGraphSizes);
    std::printf("\"springGraph\": [");
    for (int i = 0; i < totalGraphSize; ++i) std::printf("%s%.9g", i ? ", " : "", (double)springGraph[i]);
    std::printf("],\n");
    std::printf("\"bounds_min\": [%.9g, %.9g, %.9g],\n\"bounds_max\": [%.9g, %.9g, %.9g],\n", (double)minX,
                (double)minY, (double)minZ, (double)maxX, (double)maxY, (double)maxZ);
    std::printf("\"extent\": [%.9g, %.9g, %.9g],\n", (double)width, (double)height, (double)depth);
    std::printf("\"particle_cell\": [%d, %d, %d],\n\"triangle_cell\": [%d, %d, %d],\n", cellWidth, cellHeight,
                cellDepth, cellWidthTriangles, cellHeightTriangles, cellDepthTriangles);
    std::printf("\"particle_cell_counts\": [%d, %d, %d],\n", (int)(width / cellWidth), (int)(height / cellHeight),
                (int)(depth / cellDepth));
    std::printf("\"triangle_cell_counts\": [%d, %d, %d],\n", (int)(width / cellWidthTriangles),
                (int)(height / cellHeightTriangles), (int)(depth / cellDepthTriangles));
    std::printf("\"vein_vertex_count\": %d,\n\"vein_triangle_count\": %d,\n", veinPositionCount, triangleCount);
    std::printf("\"vein_endings\": [");
    mp_for_each<mp_iota_c<veinEndingCenterCount>>([&](auto i) {
        using C = mp_at_c<VeinEndingCenters, i>;
        using Rr = mp_at_c<VeinEndingRadii, i>;
        std::printf("%s[%.9g, %.9g, %.9g, %.9g]", i ? ", " : "", (double)C::x, (double)C::y, (double)C::z,
                    (double)Rr::value);
    });
    std::printf("],\n");
    std::printf("\"physics\": {\"dt\": %.9g, \"particle_k_sniff\": %.9g, \"particle_d_fact\": %.9g, "
                "\"vein_k_sniff\": %.9g, \"vein_d_fact\": %.9g, \"viscous_damping\": %.9g, \"cylinderRadius\": %.9g, "
                "\"minSpawnY\": %.9g, \"boundingSpheresCoeff\": %d}\n",
                (double)dt, (double)particle_k_sniff, (double)particle_d_fact, (double)vein_k_sniff,
                (double)vein_d_fact, (double)viscous_damping, (double)cylinderRadius, (double)minSpawnY,
                boundingSpheresCoeff);
    std::printf("}\n");

    // binary dumps: mesh and neighbour tables (slot-major)
    {
        std::FILE* f = std::fopen((outdir + "/vein_positions.f32").c_str(), "wb");
        for (auto v : veinPositions) { float xyz[3] = {v.x, v.y, v.z}; std::fwrite(xyz, 4, 3, f); }
        std::fclose(f);
        f = std::fopen((outdir + "/vein_indices.u32").c_str(), "wb");
        std::fwrite(veinIndices.data(), 4, veinIndices.size(), f);
        std::fclose(f);
        auto tables = calculateSpringLengths();
        auto& ids = std::get<0>(tables);
        auto& lens = std::get<1>(tables);
        f = std::fopen((outdir + "/vein_neighbor_ids.i32").c_str(), "wb");
        for (int s = 0; s < veinVertexMaxNeighbors; ++s) std::fwrite(ids[s].data(), 4, ids[s].size(), f);
        std::fclose(f);
        f = std::fopen((outdir + "/vein_neighbor_lengths.f32").c_str(), "wb");
        for (int s = 0; s < veinVertexMaxNeighbors; ++s) std::fwrite(lens[s].data(), 4, lens[s].size(), f);
        std::fclose(f);
    }
    return 0;
}
