"""GPU parity on the behaviours that decide results only in some scenes: the
ray search's first hit in cell visit order, the launch type's model table in
the contact search (with radii derived per model vertex), which types test
the ending spheres, the three runtime switches, the reference's own default
cell types and physics, and the full config-1 and config-2 sizes over 100
frames and in config 2's crowded regime.  Every case runs the HIP path
through the C-ABI and the CPU oracle on the same inputs; the state must be
bit-identical, and positions are also held to 1e-5 relative."""
import json
from pathlib import Path

import numpy as np
import pytest

import test_oracle_pins as pins
from conftest import assert_bits_equal

pytestmark = pytest.mark.gpu

F32 = np.float32
GOLDEN = Path(__file__).resolve().parent / "golden"


def _run_pair(bcs, orc, ph, types, mesh, pos, vel=None, force=None, frames=1):
    vel = np.zeros_like(pos) if vel is None else vel
    force = np.zeros_like(pos) if force is None else force
    sim = bcs.Simulation(ph, types, mesh)
    sim.set_particles(pos, vel, force)
    o = orc.Oracle(ph, types, mesh)
    o.set_particles(pos, vel, force)
    if frames:
        sim.step(frames)
        o.step(frames)
    return sim, o


def _compare(sim, o, what, rel_bar=True):
    gp, gv, gf = sim.get_particles()
    op, ov, of = o.get_particles()
    if rel_bar:
        rel = np.abs(gp.astype(np.float64) - op) / np.maximum(np.abs(op), 1e-30)
        assert rel.max() <= 1e-5, f"{what}: max relative position error {rel.max():.3e}"
    assert_bits_equal(gp, op, f"{what}: positions")
    assert_bits_equal(gv, ov, f"{what}: velocities")
    assert_bits_equal(gf, of, f"{what}: forces")
    for a, b, n in zip(sim.get_vein(), o.get_vein(), ("X", "V", "Fv")):
        assert_bits_equal(a, b, f"{what}: vein {n}")


# ------------------------------------------------------ hand-checked scenes --

@pytest.mark.parametrize("order", ["ACB", "AC", "DE"])
def test_first_hit_scenes_match_oracle(bcs, orc, order):
    """The first-hit scenes of the oracle pins (x-outer cell order, the z
    neighbour before the own cell, sorted slots within a cell): one phase
    frame, vein forces compared before they are integrated, then 5 frames."""
    table = {"A": pins.TRI_A, "B": pins.TRI_B, "C": pins.TRI_C, "D": pins.TRI_D, "E": pins.TRI_E}
    ph, types, mesh, pos, vel = pins.first_hit_scene(bcs, [table[c] for c in order])
    sim = bcs.Simulation(ph, types, mesh)
    sim.set_particles(pos, vel, np.zeros_like(pos))
    o = orc.Oracle(ph, types, mesh)
    o.set_particles(pos, vel, np.zeros_like(pos))
    sim.calculate_grid(0)
    o.calculate_grid(0)
    o.calculate_grid(1)
    sim.calculate_next_frame()
    o.calculate_next_frame()
    Fv = sim.get_vein()[2]
    assert np.count_nonzero(Fv) > 0
    assert_bits_equal(Fv, o.get_vein()[2], f"first hit {order}: vein forces")
    assert_bits_equal(sim.get_particles()[1], o.get_particles()[1], f"first hit {order}: velocity")
    sim.propagate_all()
    o.propagate_all()
    sim.step(5)
    o.step(5)
    _compare(sim, o, f"first hit {order}", rel_bar=False)


def test_launch_type_radius_scene_matches_oracle(bcs, orc):
    ph, types, mesh, pos = pins.launch_type_scene(bcs)
    sim = bcs.Simulation(ph, types, mesh)
    sim.set_particles(pos, np.zeros_like(pos), np.zeros_like(pos))
    o = orc.Oracle(ph, types, mesh)
    o.set_particles(pos, np.zeros_like(pos), np.zeros_like(pos))
    sim.calculate_grid(0)
    o.calculate_grid(0)
    gc, gp = sim.get_contacts()
    oc, op = o.contacts()
    assert_bits_equal(gc, oc, "launch type: contact counts")
    assert_bits_equal(gp, op, "launch type: contact pairs")
    np.testing.assert_array_equal(gc, [1, 0, 0, 0, 0])
    o.calculate_grid(1)
    sim.calculate_next_frame()
    o.calculate_next_frame()
    assert_bits_equal(sim.get_particles()[2], o.get_particles()[2], "launch type: forces")
    sim.propagate_all()
    o.propagate_all()
    sim.step(10)
    o.step(10)
    _compare(sim, o, "launch type, 11 frames", rel_bar=False)


def test_teleport_selection_scene_matches_oracle(bcs, orc):
    ph, types, mesh, pos = pins.teleport_scene(bcs)
    sim, o = _run_pair(bcs, orc, ph, types, mesh, pos, frames=1)
    _compare(sim, o, "teleport selection", rel_bar=False)
    X = sim.get_particles()[0].reshape(3, -1)
    # a frame under gravity moves every particle a little; only the teleported
    # cell (type 1, cell 0: particles 80..91) jumps to the spawn plane
    jumped = np.flatnonzero(np.abs(X[1] - pos.reshape(3, -1)[1]) > 1.0)
    np.testing.assert_array_equal(jumped, np.arange(80, 92))


# ------------------------------------------------- derived radii, switches --

def mixed_derived_scene(bcs, seed=0):
    """Discs (42 particles) and icosahedra (12) in the config-0 cylinder with
    collision radii derived per model vertex (contact_radius 0), as the
    reference does: the discs' radii differ from vertex to vertex and from
    the icosahedra's, so the launch type's model table changes contacts."""
    p = bcs.Scene.params(0)
    p.n_cells = 2
    p.cells[0] = bcs.SceneCell(bcs.CELL_DISC42, 60, 2.0, 0.35, 0.0)
    p.cells[1] = bcs.SceneCell(bcs.CELL_ICOSAHEDRON, 90, 2.0, 1.0, 0.0)
    p.seed = seed
    sc = bcs.Scene(params=p)
    # squeeze the cells together along the axis so that many pairs touch
    q = sc.positions.reshape(3, -1).copy()
    q[1] = q[1] * F32(0.3) - F32(60)
    sc.positions = q.reshape(-1)
    return sc


def test_mixed_derived_radii_40_frames(bcs, orc):
    sc = mixed_derived_scene(bcs)
    o = orc.Oracle(sc.physics, sc.types, sc.mesh)
    bs, _, _ = o.tables()
    assert len(np.unique(bs)) > 3, "derived radii should differ per model vertex"
    o.set_particles(sc.positions, sc.velocities, sc.forces)
    o.calculate_grid(0)
    oc, op = o.contacts()
    sim = bcs.Simulation.from_scene(sc)
    sim.calculate_grid(0)
    gc, gp = sim.get_contacts()
    assert_bits_equal(gc, oc, "derived radii: contact counts")
    assert_bits_equal(gp, op, "derived radii: contact pairs")
    assert gp.size > 100
    sim.step(40)
    o.step(40)
    _compare(sim, o, "derived radii, 40 frames")
    assert (sim.error_flags() & bcs.FLAG_RADIUS_CLAMPED) == (o.flags() & bcs.FLAG_RADIUS_CLAMPED)


@pytest.mark.parametrize("mode", ["cells", "slots"])
def test_one_radius_with_clamped_partner_indices(bcs, orc, mode):
    """One contact radius (0.5) for every model vertex, but the icosahedra
    first: a disc's partner among them selects the model index
    mstart - ((pstart - j) mod 42) of the discs' block, which falls below 0
    for many, so the radius lookup must still run and clamp (the one-radius
    shortcut of the contact search applies only when no index can clamp).
    Contact pairs, 40 frames and the clamp flag against the oracle."""
    p = bcs.Scene.params(0)
    p.n_cells = 2
    p.cells[0] = bcs.SceneCell(bcs.CELL_ICOSAHEDRON, 90, 2.0, 1.0, 0.5)
    p.cells[1] = bcs.SceneCell(bcs.CELL_DISC42, 60, 2.0, 0.35, 0.5)
    sc = bcs.Scene(params=p)
    q = sc.positions.reshape(3, -1).copy()
    q[1] = q[1] * F32(0.3) - F32(60)
    sc.positions = q.reshape(-1)
    o = orc.Oracle(sc.physics, sc.types, sc.mesh)
    assert len(np.unique(o.tables()[0])) == 1, "one contact radius expected"
    o.set_particles(sc.positions, sc.velocities, sc.forces)
    o.calculate_grid(0)
    oc, op = o.contacts()
    sim = bcs.Simulation.from_scene(sc)
    sim.set_contact_mode(bcs.CONTACT_CELLS if mode == "cells" else bcs.CONTACT_SLOTS)
    sim.calculate_grid(0)
    gc, gp = sim.get_contacts()
    assert_bits_equal(gc, oc, "one radius, clamped indices: contact counts")
    assert_bits_equal(gp, op, "one radius, clamped indices: contact pairs")
    sim.step(40)
    o.step(40)
    _compare(sim, o, "one radius, clamped indices, 40 frames", rel_bar=False)
    assert o.flags() & bcs.FLAG_RADIUS_CLAMPED, "the scene should clamp partner indices"
    assert (sim.error_flags() & bcs.FLAG_RADIUS_CLAMPED) == (o.flags() & bcs.FLAG_RADIUS_CLAMPED)


@pytest.mark.parametrize("switch", ["enable_reaction_force", "enable_big_cells_brake", "use_blood_flow"])
def test_switch_off_40_frames(bcs, orc, switch):
    sc = mixed_derived_scene(bcs, seed=2)
    ph = sc.physics
    ph.use_blood_flow = 1
    setattr(ph, switch, 0)
    sim, o = _run_pair(bcs, orc, ph, sc.types, sc.mesh, sc.positions, sc.velocities, sc.forces, frames=40)
    _compare(sim, o, f"{switch} = 0, 40 frames")


@pytest.mark.parametrize("mode", ["slots", "cells"])
def test_rk4_build_40_frames(bcs, orc, mode):
    """The reference's USE_RUNGE_KUTTA_FOR_PARTICLE build (particle_rk4 = 1):
    springs from three perturbed states and the Runge-Kutta position update,
    40 frames of the mixed scene against the oracle, in both contact modes.
    The surrogate refuses this build."""
    sc = mixed_derived_scene(bcs, seed=3)
    ph = sc.physics
    ph.particle_rk4 = 1
    sim = bcs.Simulation.from_scene(sc, physics=ph)
    sim.set_contact_mode(bcs.CONTACT_SLOTS if mode == "slots" else bcs.CONTACT_CELLS)
    o = orc.Oracle(ph, sc.types, sc.mesh)
    o.set_particles(sc.positions, sc.velocities, sc.forces)
    sim.step(40)
    o.step(40)
    _compare(sim, o, f"Runge-Kutta build ({mode}), 40 frames")
    with pytest.raises(bcs.BcsError):
        sim.attach_surrogate(bcs.Surrogate(seed=0), 1e-3)


# ------------------------------------------- the reference's default types --

def reference_default_scene(bcs, seed=0):
    """The reference's default cell list folded as its compile-time tables
    fold it (500 20-particle 'blood dust' cells, then 100 20-particle white
    cells, from tests/golden/ref_tables.json), its default physics
    (bcs_default_physics: dt 0.008, particle cell 2, triangle cell 25, k 25,
    d 10, collision radii derived per vertex), and its initial placement
    rule (src/simulation/simulation_controller.cu:199-221: centroid x, z =
    (U - 0.5) * 1.2 * cylinderRadius, y = minSpawnY - 180 U, velocity
    (0, -80, 0)) drawn from a seeded numpy generator.  The vein is a Y
    bifurcation of trunk radius 50 like the reference's mesh (which cannot
    travel): generated, 500 long, branches r 35 x 250 at +/-30 deg."""
    ref = json.loads((GOLDEN / "ref_tables.json").read_text())
    user = ref["user_types"]
    types = []
    for f in ref["folded_types"]:
        u = user[f["first_user_index"]]
        types.append(bcs.CellTypeSpec(f["count"], u["vertices"], u["springs"], contact_radius=0.0))
    p = bcs.Scene.params(1)
    p.radius, p.length = 50.0, 500.0
    p.branch_radius, p.branch_length = 35.0, 250.0
    p.mesh_spacing = 5.0
    p.cells[0] = bcs.SceneCell(bcs.CELL_ICOSAHEDRON, 1, 2.0, 1.0, 0.5)
    p.n_cells = 1
    mesh = bcs.Scene(params=p).mesh
    ph = bcs.default_physics()
    ph.seed = seed
    rng = np.random.default_rng(seed)
    nc = sum(t.count for t in types)
    U = rng.uniform(0, 1, size=(nc, 3)).astype(F32)
    cr = F32(ph.cylinder_radius)
    cen = np.stack([(U[:, 0] - F32(0.5)) * F32(1.2) * cr,
                    F32(ph.min_spawn_y) - F32(180.0) * U[:, 1],
                    (U[:, 2] - F32(0.5)) * F32(1.2) * cr], 1)
    pos = []
    c = 0
    for t in types:
        m = t.vertices.reshape(-1, 3)
        for _ in range(t.count):
            pos.append(cen[c] + m)
            c += 1
    P = np.concatenate(pos).astype(F32)
    n = P.shape[0]
    vel = np.zeros((3, n), F32)
    vel[:] = np.array(ph.init_velocity[:], F32)[:, None]
    return ph, types, mesh, np.ascontiguousarray(P.T).reshape(-1), vel.reshape(-1)


def test_reference_defaults_100_frames(bcs, orc):
    ph, types, mesh, pos, vel = reference_default_scene(bcs)
    assert sum(t.count * t.ppc for t in types) == 12000
    sim, o = _run_pair(bcs, orc, ph, types, mesh, pos, vel, frames=100)
    _compare(sim, o, "reference defaults, 100 frames")
    info = sim.get_info()
    assert list(info.particle_starts[:2]) == [0, 10000]
    assert list(info.model_starts[:2]) == [0, 20]


# -------------------------------------------------- full sizes, 100 frames --

@pytest.mark.timeout(900)
@pytest.mark.parametrize("config", [1, 2])
def test_full_config_100_frames(bcs, orc, config):
    """Full config-1 (240k particles) and config-2 (6.6M particles, two cell
    types) sizes: 100 graph-replayed frames against the oracle (OpenMP on
    the host cores), bit for bit and within 1e-5 relative."""
    sc = bcs.Scene(config, 1.0, seed=0)
    sim, o = _run_pair(bcs, orc, sc.physics, sc.types, sc.mesh, sc.positions, sc.velocities, sc.forces,
                       frames=0)
    for done in range(10, 101, 10):
        sim.step(10)
        o.step(10)
        print(f"config {config}: {done} frames", flush=True)
    _compare(sim, o, f"config {config}, 100 frames")
    assert sim.get_info().step_index == 100


@pytest.mark.timeout(900)
def test_config2_crowded_regime(bcs, orc):
    """Config 2 after 300 GPU frames, when cells crowd at the spawn plane
    (the big-bucket sort paths and long contact lists): the GPU state is
    loaded into a fresh oracle and both run 2 more frames."""
    sc = bcs.Scene(2, 1.0, seed=0)
    sim = bcs.Simulation.from_scene(sc)
    sim.step(300)
    p, v, f = sim.get_particles()
    X, V, Fv = sim.get_vein()
    o = orc.Oracle(sc.physics, sc.types, sc.mesh)
    o.set_particles(p, v, f)
    o.set_vein(X, V, Fv)
    o.set_step_index(300)
    o.calculate_grid(0)
    counts, _ = o.contacts()
    assert counts.max() >= 20, f"expected a crowded state, max contacts {counts.max()}"
    sim.step(2)
    o.step(2)
    _compare(sim, o, "config 2 crowded regime (frames 301-302)")


# ------------------------------------------------------------ FP64 drift --

def test_fp64_drift_config2_subset_1000_frames(bcs, orc):
    """Drift of the FP32 step against an FP64 rerun of the same step (the
    oracle in double precision) over the first 1,000 frames of a named
    config-2 subset: the config-2 vein and cell mix with 2% of its cells
    (4,000 cells, 132,000 particles), so that the FP64 rerun fits a test on
    the host cores.  Printed at frames 10, 50, 100, 250, 500 and 1,000, in
    units of the cell radius r = 2: the median, 99th percentile and largest
    particle distance |x32 - x64|, the distance of the two particle
    centroids, and the relative difference of the mean speed.  The step is
    chaotic (stiff springs, attractive contacts, teleports at the vein end):
    single particles decorrelate within some 100 frames while the
    aggregates stay close.  Bounds: at frame 10 every particle within
    1e-3 r; at frame 1,000 everything finite, centroids within r and mean
    speeds within 5%."""
    sc = bcs.Scene(2, 0.02, seed=0)
    sim = bcs.Simulation.from_scene(sc)
    o = orc.Oracle(sc.physics, sc.types, sc.mesh, "f64")
    o.set_particles(sc.positions, sc.velocities, sc.forces)
    n = sim.np
    done = 0
    rows = {}
    for mark in (10, 50, 100, 250, 500, 1000):
        sim.step(mark - done)
        o.step(mark - done)
        done = mark
        g, gv, _ = (a.reshape(3, n).astype(np.float64) for a in sim.get_particles())
        d, dv, _ = (a.reshape(3, n) for a in o.get_particles())
        assert np.isfinite(g).all() and np.isfinite(gv).all()
        dist = np.sqrt(((g - d) ** 2).sum(axis=0)) / 2.0
        cen = float(np.linalg.norm(g.mean(axis=1) - d.mean(axis=1)) / 2.0)
        s32, s64 = np.sqrt((gv ** 2).sum(axis=0)).mean(), np.sqrt((dv ** 2).sum(axis=0)).mean()
        rows[mark] = (float(np.median(dist)), float(np.percentile(dist, 99)), float(dist.max()), cen,
                      float(abs(s32 - s64) / s64))
        print(f"config-2 subset, frame {mark}: |x32 - x64| / r median {rows[mark][0]:.2e}, p99 {rows[mark][1]:.2e}, "
              f"max {rows[mark][2]:.2e}; centroids {cen:.2e} r; mean speed {rows[mark][4]:.2e} rel", flush=True)
    assert rows[10][2] < 1e-3, rows[10]
    assert rows[1000][3] < 1.0 and rows[1000][4] < 0.05, rows[1000]
