"""Scene generator (host only): mesh sizes of the benchmark configurations,
mesh topology, reproducibility from the seed, cells inside the lumen and the
initial velocity law."""
import ctypes as C

import numpy as np
import pytest


def edge_use(idx):
    tri = idx.reshape(-1, 3).astype(np.int64)
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    directed = e[:, 0] * (1 << 32) + e[:, 1]
    und = np.sort(e, axis=1)
    key = und[:, 0] * (1 << 32) + und[:, 1]
    _, counts = np.unique(key, return_counts=True)
    return counts, directed


def test_cylinder_sizes_and_topology(bcs):
    sc = bcs.Scene(0, 1.0, seed=0)
    assert sc.mesh.indices.size // 3 == 12800  # 64 segments x 100 bands x 2
    assert sc.mesh.positions.size // 3 == 6464  # 101 rings x 64
    assert sc.particle_count == 6000 and len(sc.types) == 1 and sc.types[0].ppc == 12
    counts, directed = edge_use(sc.mesh.indices)
    # open tube: interior edges shared by two triangles, the two end rings by one
    assert set(np.unique(counts)) == {1, 2}
    assert (counts == 1).sum() == 2 * 64
    # consistent orientation: no directed edge repeats
    assert np.unique(directed).size == directed.size


def test_y_vein_is_a_consistently_oriented_surface(bcs):
    sc = bcs.Scene(1, 0.02, seed=0)
    nt = sc.mesh.indices.size // 3
    assert 180_000 <= nt <= 220_000
    counts, directed = edge_use(sc.mesh.indices)
    assert counts.max() == 2  # a 2-manifold (with boundary at the open ends)
    assert np.unique(directed).size == directed.size
    assert (counts == 1).sum() < 0.01 * counts.size


def test_config1_full_counts(bcs):
    p = bcs.Scene.params(1)
    assert p.cells[0].count == 20000 and p.n_cells == 1
    assert p.physics.dt == np.float32(0.005)


def test_seed_reproducibility(bcs):
    a = bcs.Scene(0, 0.2, seed=7)
    b = bcs.Scene(0, 0.2, seed=7)
    c = bcs.Scene(0, 0.2, seed=8)
    assert np.array_equal(a.positions, b.positions) and np.array_equal(a.velocities, b.velocities)
    assert not np.array_equal(a.positions, c.positions)


@pytest.mark.parametrize("preset,frac", [(0, 0.5), (1, 0.02), (2, 0.002)])
def test_particles_start_inside_the_lumen(bcs, preset, frac):
    lib = bcs.load_library()
    h = C.c_void_p()
    assert lib.bcs_scene_preset(preset, C.c_double(frac), C.c_uint64(1), C.byref(h)) == bcs.BCS_OK
    try:
        v = bcs.SceneView()
        assert lib.bcs_scene_get_view(h, C.byref(v)) == bcs.BCS_OK
        n = v.particle_count
        assert n > 0
        pos = np.ctypeslib.as_array(v.positions, shape=(3 * n,)).reshape(3, n).T.copy()
        sdf = np.zeros(n, np.float32)
        assert lib.bcs_scene_lumen_sdf(h, bcs.fptr(np.ascontiguousarray(pos.reshape(-1))), n, bcs.fptr(sdf)) == 0
        assert (sdf < 0).all(), f"{(sdf >= 0).sum()} particles outside the lumen"
    finally:
        lib.bcs_scene_destroy(h)


def test_cells_are_rigid_copies_of_the_model(bcs):
    sc = bcs.Scene(0, 0.1, seed=2)
    t = sc.types[0]
    n = sc.particle_count
    P = sc.positions.reshape(3, n).T.reshape(-1, t.ppc, 3)
    M = t.vertices.reshape(t.ppc, 3)
    d_model = np.linalg.norm(M[:, None] - M[None], axis=-1)
    d_cells = np.linalg.norm(P[:, :, None] - P[:, None], axis=-1)
    assert np.allclose(d_cells, d_model[None], atol=1e-4)


def test_initial_velocity_law(bcs):
    """Config 0: velocity 0.5 along the axis (-y) plus N(0, 0.05) per cell."""
    sc = bcs.Scene(0, 1.0, seed=0)
    n = sc.particle_count
    V = sc.velocities.reshape(3, n).T.reshape(-1, 12, 3)
    assert np.allclose(V, V[:, :1])  # one velocity per cell
    v = V[:, 0]
    assert abs(v[:, 1].mean() + 0.5) < 0.01
    assert abs(v[:, 0].std() - 0.05) < 0.01 and abs(v[:, 2].std() - 0.05) < 0.01


def test_lattice_packing_fraction(bcs):
    p = bcs.Scene.params(2)
    assert p.placement == bcs.PLACE_LATTICE and abs(p.packing_fraction - 0.35) < 1e-6
    assert p.cells[0].kind == bcs.CELL_DISC42 and p.cells[0].count == 140000
    assert p.cells[1].kind == bcs.CELL_ICOSAHEDRON and p.cells[1].count == 60000
