"""GPU parity: the HIP path (libbcs, through its C-ABI) against the CPU oracle
on the same seeded inputs.  Integer/index outputs must be identical; the
floating-point state is compared bit for bit (the library is compiled without
FP contraction and with IEEE division/sqrt), and the 100-step trajectory is
additionally held to the stated bar of 1e-5 relative."""
import numpy as np
import pytest

from conftest import assert_bits_equal

pytestmark = pytest.mark.gpu


def _scene(bcs, kind, seed=0):
    if kind == "c0":
        return bcs.Scene(0, 1.0, seed=seed)
    if kind == "c1_small":
        return bcs.Scene(1, 0.05, seed=seed)
    if kind == "mixed":
        p = bcs.Scene.params(0)
        p.n_cells = 2
        p.cells[0] = bcs.SceneCell(bcs.CELL_DISC42, 60, 2.0, 0.35, 0.5)
        p.cells[1] = bcs.SceneCell(bcs.CELL_ICOSAHEDRON, 90, 2.0, 1.0, 0.5)
        p.seed = seed
        return bcs.Scene(params=p)
    raise ValueError(kind)


def _pair(bcs, orc, sc):
    sim = bcs.Simulation.from_scene(sc)
    o = orc.Oracle(sc.physics, sc.types, sc.mesh)
    o.set_particles(sc.positions, sc.velocities, sc.forces)
    return sim, o


@pytest.mark.parametrize("kind", ["c0", "mixed", "c1_small"])
def test_particle_grid_matches_oracle(bcs, orc, kind):
    sc = _scene(bcs, kind)
    sim, o = _pair(bcs, orc, sc)
    sim.calculate_grid(bcs.GRID_PARTICLES)
    o.calculate_grid(0)
    g, r = sim.get_grid(0), o.get_grid(0)
    for name, a, b in zip(("keys_sorted", "perm", "occ_keys", "occ_start", "occ_end"), g, r):
        assert_bits_equal(a, b, name)


@pytest.mark.parametrize("shift", [None, 4, 8, 12])
@pytest.mark.parametrize("squeeze", [0.25, 0.03, 0.002])
def test_particle_grid_crowded_matches_oracle(bcs, orc, squeeze, shift):
    """Particles pulled together so that sort buckets hold hundreds to all
    6,000 objects (LDS bitonic path and the tiled fallback), with the
    automatic bucket size and with buckets of 2^4 cells (occupancy map from
    its own pass), 2^8 and 2^12 cells (map written by the sort)."""
    sc = _scene(bcs, "c0")
    p = sc.positions.reshape(3, -1).copy()
    c = p.mean(axis=1, keepdims=True)
    p = (c + (p - c) * np.float32(squeeze)).astype(np.float32)
    sc.positions = p.reshape(-1)
    sim, o = _pair(bcs, orc, sc)
    if shift is not None:
        sim.set_bucket_shift(shift)
    sim.calculate_grid(0)
    o.calculate_grid(0)
    g, r = sim.get_grid(0), o.get_grid(0)
    for name, a, b in zip(("keys_sorted", "perm", "occ_keys", "occ_start", "occ_end"), g, r):
        assert_bits_equal(a, b, f"crowded x{squeeze} {name}")
    o.calculate_grid(1)
    sim.calculate_next_frame()
    o.calculate_next_frame()
    assert_bits_equal(sim.get_particles()[2], o.get_particles()[2], f"crowded x{squeeze} forces")


@pytest.mark.parametrize("kind", ["c0", "c1_small"])
def test_triangle_grid_matches_oracle(bcs, orc, kind):
    sc = _scene(bcs, kind)
    sim, o = _pair(bcs, orc, sc)
    o.calculate_grid(1)
    g, r = sim.get_grid(1), o.get_grid(1)
    for name, a, b in zip(("keys_sorted", "perm", "occ_keys", "occ_start", "occ_end"), g, r):
        assert_bits_equal(a, b, "triangle " + name)


@pytest.mark.parametrize("kind", ["c0", "mixed", "c1_small"])
def test_contact_sets_match_oracle(bcs, orc, kind):
    sc = _scene(bcs, kind)
    # squeeze the cells along the axis so that many particles are in contact
    p = sc.positions.reshape(3, -1).copy()
    p[1] = p[1] * np.float32(0.15) - np.float32(30)
    sc.positions = p.reshape(-1)
    sim, o = _pair(bcs, orc, sc)
    sim.calculate_grid(0)
    o.calculate_grid(0)
    gc, gp = sim.get_contacts()
    oc, opairs = o.contacts()
    assert_bits_equal(gc, oc, "contact counts")
    assert_bits_equal(gp, opairs, "contact pairs (visit order)")
    assert gp.size > 0
    # the forces of the same crowded state (contacts added in visit order)
    o.calculate_grid(1)
    sim.calculate_next_frame()
    o.calculate_next_frame()
    gpos, gv, gf = sim.get_particles()
    opos, ov, of = o.get_particles()
    assert_bits_equal(gf, of, "crowded forces")
    assert_bits_equal(gv, ov, "crowded velocities")


@pytest.mark.parametrize("kind", ["c0", "mixed"])
def test_phases_match_oracle(bcs, orc, kind):
    """calculateNextFrame then propagateAll, state compared after each."""
    sc = _scene(bcs, kind)
    sim, o = _pair(bcs, orc, sc)
    for frame in range(3):
        sim.calculate_grid(0)
        o.calculate_grid(0)
        o.calculate_grid(1)
        sim.calculate_next_frame()
        o.calculate_next_frame()
        gp, gv, gf = sim.get_particles()
        op, ov, of = o.get_particles()
        assert_bits_equal(gf, of, f"frame {frame} forces after calculateNextFrame")
        assert_bits_equal(gv, ov, f"frame {frame} velocities after calculateNextFrame")
        assert_bits_equal(sim.get_centers(), o.get_centers(), f"frame {frame} centres")
        assert_bits_equal(sim.get_vein()[2], o.get_vein()[2], f"frame {frame} vein forces")
        sim.propagate_all()
        o.propagate_all()
        gp, gv, gf = sim.get_particles()
        op, ov, of = o.get_particles()
        assert_bits_equal(gp, op, f"frame {frame} positions after propagateAll")
        assert_bits_equal(gv, ov, f"frame {frame} velocities after propagateAll")
        X, V, Fv = sim.get_vein()
        oX, oV, oF = o.get_vein()
        assert_bits_equal(X, oX, f"frame {frame} vein positions")
        assert_bits_equal(V, oV, f"frame {frame} vein velocities")


def test_stale_grid_uses_current_positions(bcs, orc):
    """calculateNextFrame after propagateAll without a new grid: the old cells
    are scanned with the moved particles, as in the reference."""
    sc = _scene(bcs, "mixed")
    p = sc.positions.reshape(3, -1).copy()
    p[1] = p[1] * np.float32(0.2) - np.float32(30)
    sc.positions = p.reshape(-1)
    sim, o = _pair(bcs, orc, sc)
    sim.calculate_grid(0)
    o.calculate_grid(0)
    o.calculate_grid(1)
    for frame in range(3):
        sim.calculate_next_frame()
        o.calculate_next_frame()
        sim.propagate_all()
        o.propagate_all()
    for x, y, n in zip(sim.get_particles(), o.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, "stale grid " + n)


@pytest.mark.parametrize("history", ["fused", "phases"])
def test_stale_grid_after_frames(bcs, orc, history):
    """Two frames (fused or by phases), a phase frame, then calculateNextFrame
    on the now stale grid: the search is centred on each particle's stored
    cell (gridCellIds) and clipped by its current position, as in the
    reference."""
    sc = _scene(bcs, "mixed")
    sim, o = _pair(bcs, orc, sc)
    if history == "fused":
        sim.step(2)
        o.step(2)
    else:
        for _ in range(2):
            sim.calculate_grid(0)
            o.calculate_grid(0)
            o.calculate_grid(1)
            sim.calculate_next_frame()
            o.calculate_next_frame()
            sim.propagate_all()
            o.propagate_all()
    sim.calculate_grid(0)
    o.calculate_grid(0)
    o.calculate_grid(1)
    sim.calculate_next_frame()
    o.calculate_next_frame()
    sim.propagate_all()
    o.propagate_all()
    sim.calculate_next_frame()
    o.calculate_next_frame()
    for x, y, n in zip(sim.get_particles(), o.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, f"stale grid after {history} frames: {n}")


def test_fused_100_steps_c0(bcs, orc):
    """Config-0 workload, 100 fused graph-replayed frames vs the oracle."""
    sc = _scene(bcs, "c0")
    sim, o = _pair(bcs, orc, sc)
    sim.step(100)
    o.step(100)
    gp, gv, gf = sim.get_particles()
    op, ov, of = o.get_particles()
    rel = np.abs(gp - op) / np.maximum(np.abs(op), 1e-30)
    assert rel.max() <= 1e-5, f"max relative position error {rel.max():.3e}"
    assert_bits_equal(gp, op, "positions after 100 frames")
    assert_bits_equal(gv, ov, "velocities after 100 frames")
    assert_bits_equal(gf, of, "forces after 100 frames")
    assert_bits_equal(sim.get_vein()[0], o.get_vein()[0], "vein positions after 100 frames")
    assert sim.get_info().step_index == 100


@pytest.mark.parametrize("mode", ["auto", "fused"])
def test_fused_1000_steps_c0_bit_exact(bcs, orc, mode):
    """The whole config-0 run (1,000 frames, teleports included) stays bit
    for bit on the oracle's trajectory, with either ray launch."""
    sc = _scene(bcs, "c0")
    sim, o = _pair(bcs, orc, sc)
    if mode == "fused":
        sim.set_ray_mode(bcs.RAY_FUSED)
    for chunk in range(4):
        sim.step(250)
        o.step(250)
        gp, gv, gf = sim.get_particles()
        op, ov, of = o.get_particles()
        assert_bits_equal(gp, op, f"positions after {250 * (chunk + 1)} frames")
        assert_bits_equal(gv, ov, f"velocities after {250 * (chunk + 1)} frames")
    assert_bits_equal(sim.get_vein()[0], o.get_vein()[0], "vein after 1000 frames")
    assert sim.get_info().step_index == 1000


def test_ray_work_overflow_falls_back_to_serial_scan(bcs, orc):
    """A work list of 10 items per cell block overflows near the wall: those
    rays are scanned serially in k_cells_b, with the same result."""
    sc = _scene(bcs, "c0")
    ph = sc.physics
    ph.ray_work_items = 240
    sim = bcs.Simulation.from_scene(sc, physics=ph)
    o = orc.Oracle(ph, sc.types, sc.mesh)
    o.set_particles(sc.positions, sc.velocities, sc.forces)
    sim.step(60)
    o.step(60)
    for x, y, n in zip(sim.get_particles(), o.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, "serial ray fallback " + n)
    assert sim.error_flags() & bcs.FLAG_RAY_SERIAL


@pytest.mark.parametrize("mode", ["split", "fused"])
def test_ray_modes_match_oracle(bcs, orc, mode):
    """The ray test as two launches (candidates, then 8 lanes per item over
    several blocks per cell block) or as one launch (each cell block tests its
    own candidates from LDS): 40 graph frames against the oracle, with the
    serial fallback exercised by a short work list in a second run."""
    sc = _scene(bcs, "c0")
    for items in (0, 240):
        ph = sc.physics
        ph.ray_work_items = items
        sim = bcs.Simulation.from_scene(sc, physics=ph)
        sim.set_ray_mode(bcs.RAY_SPLIT if mode == "split" else bcs.RAY_FUSED)
        o = orc.Oracle(ph, sc.types, sc.mesh)
        o.set_particles(sc.positions, sc.velocities, sc.forces)
        sim.step(40)
        o.step(40)
        for x, y, n in zip(sim.get_particles(), o.get_particles(), ("pos", "vel", "force")):
            assert_bits_equal(x, y, f"{mode} ray test, work list {items}: {n}")
        assert_bits_equal(sim.get_vein()[0], o.get_vein()[0], f"{mode} ray test, work list {items}: vein")
        assert sim.frame_stats()["ray_work_items"] > 0, "the last frame should have ray candidates"
        if items:
            assert sim.error_flags() & bcs.FLAG_RAY_SERIAL


def test_ray_mode_switch_mid_run(bcs, orc):
    """Switching the ray mode between step calls drops the captured frame
    graphs; the run stays bit-identical to the oracle.  An unknown mode is
    rejected and leaves the simulation usable."""
    sc = _scene(bcs, "mixed")
    sim, o = _pair(bcs, orc, sc)
    sim.step(20)
    sim.set_ray_mode(bcs.RAY_FUSED)
    sim.step(20)
    sim.set_ray_mode(bcs.RAY_AUTO)
    sim.step(5)
    with pytest.raises(bcs.BcsError):
        sim.set_ray_mode(7)
    sim.step(3)
    o.step(48)
    for x, y, n in zip(sim.get_particles(), o.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, "ray mode switch " + n)
    assert_bits_equal(sim.get_vein()[0], o.get_vein()[0], "ray mode switch vein")


@pytest.mark.parametrize("mode", ["cells", "slots"])
def test_contact_modes_match_oracle(bcs, orc, mode):
    """The contact search inside the cell kernel (lanes by blood cell) or in a
    kernel of its own (lanes by sorted grid slot, between a springs kernel and
    a response kernel): 40 graph frames of the mixed scene against the oracle,
    then a switch to the other mode for 8 more."""
    sc = _scene(bcs, "mixed")
    sim, o = _pair(bcs, orc, sc)
    first, second = (bcs.CONTACT_CELLS, bcs.CONTACT_SLOTS) if mode == "cells" else (bcs.CONTACT_SLOTS,
                                                                                       bcs.CONTACT_CELLS)
    sim.set_contact_mode(first)
    sim.step(40)
    o.step(40)
    for x, y, n in zip(sim.get_particles(), o.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, f"contacts by {mode}, 40 frames: {n}")
    sim.set_contact_mode(second)
    sim.step(8)
    o.step(8)
    for x, y, n in zip(sim.get_particles(), o.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, f"contacts by {mode}, then switched: {n}")
    assert_bits_equal(sim.get_vein()[0], o.get_vein()[0], f"contacts by {mode}: vein")
    with pytest.raises(bcs.BcsError):
        sim.set_contact_mode(5)


@pytest.mark.parametrize("mode", ["cells", "slots"])
def test_contacts_on_a_lattice_match_oracle(bcs, orc, mode):
    """Positions snapped to a lattice of 1/4: many contact pairs then differ
    in one or two coordinates by exactly 0, which the contact force routes
    through the fully range-handled square root and quotients, while the
    others take the short in-range sequences.  Two frames against the oracle
    (after them the particles have left the lattice), with contacts checked
    to occur at both kinds of pairs."""
    sc = _scene(bcs, "mixed")
    p = np.round(np.asarray(sc.positions, np.float32) * 4.0) / 4.0
    n = p.size // 3
    sim = bcs.Simulation.from_scene(sc)
    sim.set_contact_mode(bcs.CONTACT_CELLS if mode == "cells" else bcs.CONTACT_SLOTS)
    sim.set_particles(p, sc.velocities, sc.forces)
    o = orc.Oracle(sc.physics, sc.types, sc.mesh)
    o.set_particles(p, sc.velocities, sc.forces)
    sim.calculate_grid()
    counts, pairs = sim.get_contacts()
    P = p.reshape(3, n)
    owner = np.repeat(np.arange(n), counts)
    zero_axis = np.any(P[:, owner] == P[:, pairs], axis=0)
    assert zero_axis.any() and (~zero_axis).any(), "lattice scene lacks one kind of contact"
    sim.step(2)
    o.step(2)
    for x, y, name in zip(sim.get_particles(), o.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, f"lattice positions, contacts by {mode}: {name}")


@pytest.mark.parametrize("shift", [4, 8, 11])
def test_bucket_sizes_keep_parity(bcs, orc, shift):
    """40 graph frames with buckets of 2^4 cells (occupancy map from its own
    pass), 2^8 and 2^11 cells (map written by the sort, groups straddling two
    buckets ORed in): the same trajectory as the oracle."""
    sc = _scene(bcs, "mixed")
    sim, o = _pair(bcs, orc, sc)
    sim.set_bucket_shift(shift)
    sim.step(40)
    o.step(40)
    for x, y, n in zip(sim.get_particles(), o.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, f"buckets of 2^{shift} cells: {n}")
    with pytest.raises(bcs.BcsError):
        sim.set_bucket_shift(25)


def test_fused_40_steps_mixed_types(bcs, orc):
    """Two cell types (discs and icosahedra): 40 fused frames, across two
    refreshes of the spatial processing order of the cells."""
    sc = _scene(bcs, "mixed")
    sim, o = _pair(bcs, orc, sc)
    sim.step(40)
    o.step(40)
    for x, y, n in zip(sim.get_particles(), o.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, "mixed 40 frames " + n)
    assert_bits_equal(sim.get_centers(), o.get_centers(), "mixed 40 frames centres")
    assert_bits_equal(sim.get_vein()[0], o.get_vein()[0], "mixed 40 frames vein")


def test_fused_equals_phase_api(bcs):
    sc = _scene(bcs, "mixed")
    a = bcs.Simulation.from_scene(sc)
    b = bcs.Simulation.from_scene(sc)
    a.step(5)
    for _ in range(5):
        b.calculate_grid(0)
        b.calculate_next_frame()
        b.propagate_all()
    for x, y, n in zip(a.get_particles(), b.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, "fused vs phase " + n)


def test_graph_replay_equals_direct_launch(bcs):
    sc = _scene(bcs, "c1_small")
    a = bcs.Simulation.from_scene(sc)
    b = bcs.Simulation.from_scene(sc)
    b.use_graph(False)
    for _ in range(3):  # replays interleaved with host synchronisation
        a.step(4)
        b.step(4)
    for x, y, n in zip(a.get_particles(), b.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, "graph vs direct " + n)
    assert_bits_equal(a.get_vein()[0], b.get_vein()[0], "graph vs direct vein")
    assert a.get_info().step_index == b.get_info().step_index == 12


def test_multi_frame_graph_equals_direct_launch(bcs):
    """step(21) replays the first-frame graph, one 16-frame graph and four
    one-frame graphs; direct launches of the same frames give the same bits."""
    sc = _scene(bcs, "c1_small")
    a = bcs.Simulation.from_scene(sc)
    b = bcs.Simulation.from_scene(sc)
    b.use_graph(False)
    a.step(21)
    b.step(21)
    for x, y, n in zip(a.get_particles(), b.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, "16-frame graph vs direct " + n)
    assert_bits_equal(a.get_vein()[0], b.get_vein()[0], "16-frame graph vs direct vein")
    assert a.get_info().step_index == b.get_info().step_index == 21


def test_determinism_two_runs(bcs):
    sc = _scene(bcs, "c1_small")
    a = bcs.Simulation.from_scene(sc)
    b = bcs.Simulation.from_scene(sc)
    a.step(30)
    b.step(30)
    for x, y, n in zip(a.get_particles(), b.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, "rerun " + n)
    assert_bits_equal(a.get_vein()[0], b.get_vein()[0], "rerun vein")


def test_c1_full_one_frame(bcs, orc):
    """Full config-1 sizes (240k particles, ~200k triangles): one frame."""
    sc = bcs.Scene(1, 1.0, seed=0)
    sim, o = _pair(bcs, orc, sc)
    sim.step(2)
    o.step(2)
    for x, y, n in zip(sim.get_particles(), o.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, "c1 " + n)
    assert_bits_equal(sim.get_vein()[0], o.get_vein()[0], "c1 vein")


def test_c2_full_two_frames(bcs, orc):
    """Full config-2 sizes (6.6M particles of two cell types, ~1.8M triangles,
    ~26k cell blocks): two frames.  At this size the default ray mode runs
    candidates and ray test as one launch, and the particle grid uses small
    buckets with crowded big ones."""
    sc = bcs.Scene(2, 1.0, seed=0)
    sim, o = _pair(bcs, orc, sc)
    sim.step(2)
    o.step(2)
    for x, y, n in zip(sim.get_particles(), o.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, "c2 " + n)
    assert_bits_equal(sim.get_vein()[0], o.get_vein()[0], "c2 vein")
    assert "k_raycand_ray" in sim.profile_kernels(1), "config 2 should run the fused ray launch"


def test_displaced_vein_keeps_parity(bcs, orc):
    """Vein moved away from its creation position (the frozen triangle grid
    keeps its creation-time cells): the per-frame triangle-cell boxes follow
    the moved triangles, so the culled ray scan still finds exactly the
    oracle's hits."""
    sc = _scene(bcs, "c0")
    sim, o = _pair(bcs, orc, sc)
    X, V, F = sim.get_vein()
    X = X.reshape(3, -1).copy()
    X[0] += np.float32(3.0)
    X[2] -= np.float32(1.5)
    X = X.reshape(-1)
    sim.set_vein(X, V, F)
    o.set_vein(X, V, F)
    sim.step(20)
    o.step(20)
    for x, y, n in zip(sim.get_particles(), o.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, "displaced vein " + n)
    assert_bits_equal(sim.get_vein()[0], o.get_vein()[0], "displaced vein X")


@pytest.mark.parametrize("mode", ["auto", "fused"])
def test_dynamic_triangle_grid_parity(bcs, orc, mode):
    """Triangle grid and its records rebuilt from the vein every frame (the
    vein update then ends each frame), with either ray launch."""
    sc = _scene(bcs, "c0")
    ph = sc.physics
    ph.dynamic_triangle_grid = 1
    sim = bcs.Simulation.from_scene(sc, physics=ph)
    if mode == "fused":
        sim.set_ray_mode(bcs.RAY_FUSED)
    o = orc.Oracle(ph, sc.types, sc.mesh)
    o.set_particles(sc.positions, sc.velocities, sc.forces)
    sim.step(20)
    o.step(20)
    for x, y, n in zip(sim.get_particles(), o.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, f"dynamic triangle grid ({mode}) " + n)


def _hip_to_host(ptr, n, dtype):
    """Copies n elements of a device array to a numpy array (hipMemcpy)."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    out = np.empty(n, dtype)
    rc = hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), ctypes.c_size_t(out.nbytes),
                       ctypes.c_int(2))  # hipMemcpyDeviceToHost
    assert rc == 0, f"hipMemcpy failed: {rc}"
    return out


@pytest.mark.parametrize("which", [0, 1])
def test_grid_view_and_cell_ids_match_oracle(bcs, orc, which):
    """The reference UniformGrid's public arrays: gridCellIds / particleIds of
    the sorted grid, the dense gridCellStarts / gridCellEnds per cell id
    (empty cells (0, -1)), and calculateCellId of the objects' positions."""
    sc = _scene(bcs, "mixed")
    sim, o = _pair(bcs, orc, sc)
    sim.step(3)
    o.step(3)
    sim.calculate_grid(0)
    o.calculate_grid(0)
    keys, perm, okeys, ostart, oend = o.get_grid(which)
    v = sim.grid_view(which, dense_ranges=True)
    sim.synchronize()
    n = v.object_count
    assert n == keys.size
    assert_bits_equal(_hip_to_host(v.grid_cell_ids, n, np.uint32), keys, "gridCellIds")
    assert_bits_equal(_hip_to_host(v.particle_ids, n, np.int32), perm, "particleIds")
    m = int(v.cell_count)
    starts = _hip_to_host(v.grid_cell_starts, m, np.int32)
    ends = _hip_to_host(v.grid_cell_ends, m, np.int32)
    exp_s = np.zeros(m, np.int32)
    exp_e = np.full(m, -1, np.int32)
    exp_s[okeys] = ostart
    exp_e[okeys] = oend
    assert_bits_equal(starts, exp_s, "gridCellStarts")
    assert_bits_equal(ends, exp_e, "gridCellEnds")
    objs = o.get_particles()[0] if which == 0 else o.get_triangle_centers()
    per_obj = np.empty(n, np.uint32)
    per_obj[perm] = keys
    if which == 0:
        # positions moved since the build: recompute on the oracle's current grid
        o.calculate_grid(0)
        keys, perm, *_ = o.get_grid(0)
        per_obj[perm] = keys
    assert_bits_equal(sim.cell_ids(objs, which), per_obj, "calculateCellId")
