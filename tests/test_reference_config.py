"""Loader for configuration headers in the reference's format
(include/bcs_refconfig.h).  Against the reference's own files, where they are
present, the loaded cell types, folding, spring graph, mesh and grid bounds
equal the tables the reference's compiled probe produced (tests/golden); the
error paths are checked on small synthetic files everywhere."""
import ctypes as C
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
REFCFG = Path("/root/reference/src/config")
F32 = np.float32

needs_reference = pytest.mark.skipif(not (REFCFG / "vein_definition.hpp").exists(),
                                     reason="the reference's configuration headers are not on this machine")


@pytest.fixture(scope="module")
def golden():
    return (json.loads((GOLDEN / "ref_tables.json").read_text()),
            json.loads((GOLDEN / "ref_vein_tables.json").read_text()))


@pytest.fixture(scope="module")
def loaded(bcs):
    if not (REFCFG / "vein_definition.hpp").exists():
        pytest.skip("the reference's configuration headers are not on this machine")
    return bcs.load_reference_config(REFCFG / "blood_cells_definition.hpp", REFCFG / "blood_cell_presets.hpp",
                                     REFCFG / "vein_definition.hpp")


def _fold(bcs, specs):
    lib = bcs.load_library()
    arr = bcs.cell_type_array(specs)
    n = len(specs)
    order = np.zeros(n, np.int32)
    counts = np.zeros(n, np.int32)
    nout = C.c_int32(0)
    bcs.check(lib.bcs_fold_reference_types(arr, n, bcs.iptr(order), bcs.iptr(counts), C.byref(nout)), "fold")
    return order[: nout.value].tolist(), counts[: nout.value].tolist()


@needs_reference
def test_cell_types_equal_the_compiled_reference(loaded, golden):
    types, _ = loaded
    ref = golden[0]["user_types"]
    assert len(types) == len(ref)
    for t, u in zip(types, ref):
        sp = np.array(u["springs"])
        assert (t.count, t.ppc) == (u["count"], u["ppc"])
        np.testing.assert_array_equal(t.s_start, sp[:, 0].astype(np.int32))
        np.testing.assert_array_equal(t.s_end, sp[:, 1].astype(np.int32))
        np.testing.assert_array_equal(t.s_len.view(np.uint32), sp[:, 2].astype(F32).view(np.uint32))
        np.testing.assert_array_equal(t.vertices.view(np.uint32), np.array(u["vertices"], F32).view(np.uint32))


@needs_reference
def test_folding_and_spring_graph_equal_the_compiled_reference(bcs, orc, loaded, golden):
    types, mesh = loaded
    order, counts = _fold(bcs, types)
    ref = golden[0]
    assert order == [t["first_user_index"] for t in ref["folded_types"]]
    assert counts == [t["count"] for t in ref["folded_types"]]
    folded = [bcs.CellTypeSpec(c, types[i].vertices, np.stack([types[i].s_start, types[i].s_end, types[i].s_len], 1))
              for i, c in zip(order, counts)]
    for f, i in zip(folded, order):
        f.s_len = types[i].s_len
    o = orc.Oracle(bcs.default_physics(), folded, mesh)
    _, _, graph = o.tables()
    g = np.array(ref["springGraph"], F32)
    np.testing.assert_array_equal(graph[: g.size], g)
    assert o.np == ref["particleCount"]


@needs_reference
def test_mesh_and_bounds_equal_the_compiled_reference(bcs, loaded, golden):
    _, mesh = loaded
    ref, vein = golden
    assert mesh.positions.size // 3 == ref["vein_vertex_count"] == vein["vertex_count"]
    assert mesh.indices.size // 3 == ref["vein_triangle_count"]
    assert hashlib.sha256(mesh.positions.tobytes()).hexdigest() == vein["mesh_positions_sha256"]
    assert hashlib.sha256(mesh.indices.tobytes()).hexdigest() == vein["mesh_indices_sha256"]
    ends = np.array(ref["vein_endings"], F32)
    np.testing.assert_array_equal(mesh.endc.reshape(-1, 3), ends[:, :3])
    np.testing.assert_array_equal(mesh.endr, ends[:, 3])
    lib = bcs.load_library()
    lo = np.zeros(3, F32)
    hi = np.zeros(3, F32)
    m = mesh.c()
    ph = bcs.default_physics()
    bcs.check(lib.bcs_grid_bounds(C.byref(m), C.byref(ph), bcs.fptr(lo), bcs.fptr(hi)), "bounds")
    np.testing.assert_array_equal(lo, np.array(ref["bounds_min"], F32))
    np.testing.assert_array_equal(hi, np.array(ref["bounds_max"], F32))
    # the 9-slot vein spring table of the loaded mesh = the reference's
    nv = mesh.positions.size // 3
    ids = np.zeros(9 * nv, np.int32)
    lens = np.zeros(9 * nv, F32)
    bcs.check(lib.bcs_vein_neighbors(C.byref(m), bcs.iptr(ids), bcs.fptr(lens)), "neighbours")
    assert hashlib.sha256(ids.tobytes()).hexdigest() == vein["ids_sha256"]
    assert hashlib.sha256(lens.tobytes()).hexdigest() == vein["lengths_sha256"]


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return p


def test_synthetic_files_and_fixed_point_decoding(bcs, tmp_path):
    presets = _write(tmp_path, "p.hpp", """
        // a two-vertex model
        using Rod_Springs = mp_list<
            Spring<0, 1, 15000000>,          /* 1.5 at the default precision 7 */
            Spring<1, 0, 25, 2>
        >;
        using Rod_Vertices = mp_list<
            mp_float3<1000000, 0, -2500000>,
            mp_float3<-5, 7, 123, 3>
        >;""")
    defs = _write(tmp_path, "d.hpp", """
        using UserDefinedBloodCellList = mp_list<
        BloodCellDef<3, 2, 0, 0, preset::Rod_Springs, preset::Rod_Vertices, preset::Rod_Indices, preset::Rod_Normals>
        >;""")
    vein = _write(tmp_path, "v.hpp", """
        inline constexpr std::array<cvec, 3> veinPositions {
        static_cast<float>(1.5), static_cast<float>(-2), static_cast<float>(3e-1),
        static_cast<float>(0), static_cast<float>(1), static_cast<float>(0),
        static_cast<float>(0.1), static_cast<float>(0.2), static_cast<float>(0.3)
        };
        inline constexpr std::array<unsigned int, 3> veinIndices { 0, 1, 2 };
        using VeinEndingCenters = mp_list< mp_float3<10, 20, -30, 2> >;
        using VeinEndingRadii = mp_list< mp_float<5, 1> >;""")
    types, mesh = bcs.load_reference_config(defs, presets, vein)
    (t,) = types
    assert (t.count, t.ppc) == (3, 2)
    assert t.s_start.tolist() == [0, 1] and t.s_end.tolist() == [1, 0]
    assert t.s_len.tolist() == [F32(15000000) / F32(1000000), F32(25) / F32(10)]
    assert t.vertices.tolist() == [F32(1), F32(0), F32(-2.5), F32(-5) / F32(100), F32(7) / F32(100), F32(123) / F32(100)]
    assert mesh.positions.tolist() == [F32(1.5), F32(-2), F32(0.3), 0, 1, 0, F32(0.1), F32(0.2), F32(0.3)]
    assert mesh.indices.tolist() == [0, 1, 2]
    assert mesh.endc.tolist() == [1.0, 2.0, -3.0] and mesh.endr.tolist() == [5.0]


def test_errors_name_the_problem(bcs, tmp_path):
    bad_defs = _write(tmp_path, "d.hpp", "using L = mp_list< BloodCellDef<1, 2, 0, 0, preset::Nope_Springs, "
                                         "preset::Nope_Vertices, preset::I, preset::N> >;")
    presets = _write(tmp_path, "p.hpp", "using X = mp_list< Spring<0, 1, 1> >;")
    with pytest.raises(bcs.BcsError, match="no list Nope_Springs"):
        bcs.load_reference_config(bad_defs, presets)
    with pytest.raises(bcs.BcsError, match="cannot read"):
        bcs.load_reference_config(vein=tmp_path / "missing.hpp")
    bad_vein = _write(tmp_path, "v.hpp", "veinPositions { static_cast<float>(1), static_cast<float>(2) }; "
                                         "veinIndices { 0 };")
    with pytest.raises(bcs.BcsError, match="triples"):
        bcs.load_reference_config(vein=bad_vein)
