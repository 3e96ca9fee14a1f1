"""Host-side tables of libbcs (and of the oracle) against the reference's own
compile-time tables (tests/golden, dumped by the reference probe built from
the reference headers) and against independent numpy/scipy computations."""
import ctypes as C
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest
from scipy.spatial.distance import pdist, squareform

from conftest import ref_dir

GOLDEN = Path(__file__).resolve().parent / "golden"
F32 = np.float32


@pytest.fixture(scope="module")
def ref_tables():
    return json.loads((GOLDEN / "ref_tables.json").read_text())


def _user_specs(bcs, tables):
    out = []
    for u in tables["user_types"]:
        out.append(bcs.CellTypeSpec(u["count"], u["vertices"], u["springs"], 0.0))
    return out


def _fold(bcs, specs):
    lib = bcs.load_library()
    arr = bcs.cell_type_array(specs)
    n = len(specs)
    order = np.zeros(n, np.int32)
    counts = np.zeros(n, np.int32)
    nout = C.c_int32(0)
    bcs.check(lib.bcs_fold_reference_types(arr, n, bcs.iptr(order), bcs.iptr(counts), C.byref(nout)), "fold")
    return order[: nout.value].tolist(), counts[: nout.value].tolist()


def test_reference_type_folding_matches_compiled_reference(bcs, ref_tables):
    order, counts = _fold(bcs, _user_specs(bcs, ref_tables))
    ref = ref_tables["folded_types"]
    assert order == [t["first_user_index"] for t in ref]
    assert counts == [t["count"] for t in ref]
    # derived starts
    ppc = [ref_tables["user_types"][i]["ppc"] for i in order]
    starts = np.concatenate([[0], np.cumsum(np.array(counts) * np.array(ppc))[:-1]]).tolist()
    assert starts == ref_tables["particleStarts"]
    assert np.concatenate([[0], np.cumsum(counts)[:-1]]).tolist() == ref_tables["bloodCellTypesStarts"]
    assert np.concatenate([[0], np.cumsum(ppc)[:-1]]).tolist() == ref_tables["bloodCellModelStarts"]
    assert np.concatenate([[0], np.cumsum(np.square(ppc))[:-1]]).tolist() == ref_tables["accumulatedGraphSizes"]


def _dummy(bcs, count, ppc, springs):
    verts = np.arange(3 * ppc, dtype=F32)
    return bcs.CellTypeSpec(count, verts, springs)


def test_reference_sort_quirks_hand_cases(bcs):
    """mp_sort is a quicksort around the first element; the inverted
    power-of-two test sends powers of two to the END and keeps the relative
    order of same-class types reversed around the pivot."""
    A = _dummy(bcs, 5, 20, [[0, 1, 1.0]])
    B = _dummy(bcs, 7, 16, [[0, 1, 1.0]])
    Cdup = _dummy(bcs, 11, 20, [[0, 1, 1.0]])  # duplicate of A: merged into A
    D = _dummy(bcs, 3, 12, [[0, 2, 1.0]])
    assert _fold(bcs, [A, B]) == ([0, 1], [5, 7])
    assert _fold(bcs, [B, A]) == ([1, 0], [5, 7])
    assert _fold(bcs, [A, B, Cdup]) == ([0, 1], [16, 7])
    # two non-powers of two: the non-pivot goes before the pivot -> reversed
    assert _fold(bcs, [A, D]) == ([1, 0], [3, 5])
    assert _fold(bcs, [D, A, B]) == ([1, 0, 2], [5, 3, 7])


def test_spring_graph_matches_compiled_reference(bcs, orc, ref_tables):
    specs = _user_specs(bcs, ref_tables)
    order, counts = _fold(bcs, specs)
    folded = []
    for i, c in zip(order, counts):
        s = specs[i]
        folded.append(bcs.CellTypeSpec(c, s.vertices, np.stack([s.s_start, s.s_end, s.s_len], 1)))
    mesh = bcs.MeshSpec(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], F32), np.array([0, 1, 2], np.uint32))
    o = orc.Oracle(bcs.default_physics(), folded, mesh)
    _, _, graph = o.tables()
    ref = np.array(ref_tables["springGraph"], F32)
    np.testing.assert_array_equal(graph[: ref.size], ref)
    assert o.np == ref_tables["particleCount"]


def test_model_tables_against_scipy(bcs, orc, ref_tables):
    """Collision radii = min pairwise model distance / (2*3) (double, then
    float); initial radii = |m - centroid| in float; product == oracle."""
    lib = bcs.load_library()
    specs = _user_specs(bcs, ref_tables) + bcs.Scene(2, 0.0001, seed=0).types
    for s in specs:
        s.contact_radius = 0.0
    arr = bcs.cell_type_array(specs)
    nm = sum(s.ppc for s in specs)
    bsp = np.zeros(nm, F32)
    r0p = np.zeros(nm, F32)
    ph = bcs.default_physics()
    bcs.check(lib.bcs_model_tables(C.byref(ph), arr, len(specs), bcs.fptr(bsp), bcs.fptr(r0p)), "model tables")
    m0 = 0
    for s in specs:
        v = s.vertices.reshape(-1, 3)
        d = squareform(pdist(v.astype(np.float64)))
        np.fill_diagonal(d, np.inf)
        np.testing.assert_allclose(bsp[m0:m0 + s.ppc], (d.min(1) / 6.0).astype(F32), rtol=1e-6)
        c = np.zeros(3, F32)
        for a in range(s.ppc):
            c = (c + v[a]).astype(F32)
        c = (c / F32(s.ppc)).astype(F32)
        np.testing.assert_allclose(r0p[m0:m0 + s.ppc], np.linalg.norm(v - c, axis=1), rtol=1e-6)
        m0 += s.ppc
    mesh = bcs.MeshSpec(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], F32), np.array([0, 1, 2], np.uint32))
    o = orc.Oracle(ph, specs, mesh)
    bso, r0o, _ = o.tables()
    np.testing.assert_array_equal(bso, bsp)
    np.testing.assert_array_equal(r0o, r0p)


def _ref_mesh(bcs, tables):
    d = ref_dir()
    if d is None or not (d / "vein_positions.f32").exists():
        pytest.skip("reference probe outputs (oracle/_ref) not built here")
    pos = np.fromfile(d / "vein_positions.f32", F32)
    idx = np.fromfile(d / "vein_indices.u32", np.uint32)
    endings = np.array(tables["vein_endings"], F32)
    return bcs.MeshSpec(pos, idx, endings[:, :3], endings[:, 3])


def test_reference_mesh_dump_matches_golden(bcs, ref_tables):
    golden = json.loads((GOLDEN / "ref_vein_tables.json").read_text())
    mesh = _ref_mesh(bcs, ref_tables)
    assert hashlib.sha256(mesh.positions.tobytes()).hexdigest() == golden["mesh_positions_sha256"]
    assert hashlib.sha256(mesh.indices.tobytes()).hexdigest() == golden["mesh_indices_sha256"]


def test_vein_neighbor_tables_match_reference(bcs, orc, ref_tables):
    golden = json.loads((GOLDEN / "ref_vein_tables.json").read_text())
    mesh = _ref_mesh(bcs, ref_tables)
    nv = mesh.positions.size // 3
    lib = bcs.load_library()
    ids = np.zeros(9 * nv, np.int32)
    lens = np.zeros(9 * nv, F32)
    m = mesh.c()
    bcs.check(lib.bcs_vein_neighbors(C.byref(m), bcs.iptr(ids), bcs.fptr(lens)), "neighbors")
    assert hashlib.sha256(ids.tobytes()).hexdigest() == golden["ids_sha256"]
    assert hashlib.sha256(lens.tobytes()).hexdigest() == golden["lengths_sha256"]
    oi, ol = orc.vein_neighbors(mesh)
    np.testing.assert_array_equal(oi, ids)
    np.testing.assert_array_equal(ol.view(np.uint32), lens.view(np.uint32))
    for v, row_i, row_l in zip(golden["sample_vertices"], golden["sample_ids"], golden["sample_lengths"]):
        assert [int(ids[s * nv + v]) for s in range(9)] == row_i
        np.testing.assert_array_equal([lens[s * nv + v] for s in range(9)], np.array(row_l, F32))


def test_grid_bounds_and_cell_counts_match_reference(bcs, orc, ref_tables):
    mesh = _ref_mesh(bcs, ref_tables)
    lib = bcs.load_library()
    ph = bcs.default_physics()
    lo = np.zeros(3, F32)
    hi = np.zeros(3, F32)
    m = mesh.c()
    bcs.check(lib.bcs_grid_bounds(C.byref(m), C.byref(ph), bcs.fptr(lo), bcs.fptr(hi)), "bounds")
    np.testing.assert_array_equal(lo, np.array(ref_tables["bounds_min"], F32))
    np.testing.assert_array_equal(hi, np.array(ref_tables["bounds_max"], F32))
    olo, ohi = orc.grid_bounds(mesh, ph)
    np.testing.assert_array_equal(olo, lo)
    np.testing.assert_array_equal(ohi, hi)
    ext = (hi - lo).astype(F32)
    np.testing.assert_array_equal(ext, np.array(ref_tables["extent"], F32))
    assert (ext / F32(2)).astype(np.int32).tolist() == ref_tables["particle_cell_counts"]
    assert (ext / F32(25)).astype(np.int32).tolist() == ref_tables["triangle_cell_counts"]


def test_default_physics_matches_reference(bcs, ref_tables):
    ph = bcs.default_physics()
    r = ref_tables["physics"]
    assert ph.dt == F32(r["dt"]) and ph.particle_k_sniff == r["particle_k_sniff"]
    assert ph.particle_d_fact == r["particle_d_fact"] and ph.vein_k_sniff == F32(r["vein_k_sniff"])
    assert ph.vein_d_fact == F32(r["vein_d_fact"]) and ph.viscous_damping == F32(r["viscous_damping"])
    assert ph.cylinder_radius == r["cylinderRadius"] and ph.min_spawn_y == r["minSpawnY"]
    assert ph.bounding_spheres_coeff == r["boundingSpheresCoeff"]
    assert list(ph.particle_cell) == ref_tables["particle_cell"]
    assert list(ph.triangle_cell) == ref_tables["triangle_cell"]
