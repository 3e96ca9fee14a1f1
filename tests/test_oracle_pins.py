"""Pins of the CPU oracle that do not come from the oracle's own formulas:
hand-derived cell ids, numpy's stable sort, brute-force contact sets and
closed forms for springs, collisions, the ray test, the barycentric quirk,
the integrators and the vein springs.  A plausible slip in the oracle (a
dropped term, a swapped operand, a wrong sign or index) fails one of these."""
import numpy as np
import pytest

from conftest import soa

F32 = np.float32


def _phys(bcs, **kw):
    ph = bcs.default_physics()
    ph.use_blood_flow = 0
    for k, v in kw.items():
        if isinstance(v, (tuple, list)):
            arr = getattr(ph, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(ph, k, v)
    return ph


def _point_type(bcs, n, r=0.5):
    """n single-particle cells (no springs) with contact radius r."""
    return bcs.CellTypeSpec(n, [0.0, 0.0, 0.0], [], contact_radius=r)


def _tri_mesh(bcs, verts, tris=((0, 1, 2),)):
    return bcs.MeshSpec(np.asarray(verts, F32), np.asarray(tris, np.uint32))


def _grid_keys(orc, o, n):
    o.calculate_grid(0)
    keys, perm, *_ = o.get_grid(0)
    out = np.empty(n, np.uint32)
    out[perm] = keys
    return out


# ---------------------------------------------------------------- cell ids --

def test_cell_ids_hand_cases(bcs, orc):
    # mesh AABB [0,10]^3, margins 40 -> bounds [-40, 50]^3, extent 90, cell 2 -> 45^3 cells
    mesh = _tri_mesh(bcs, [[0, 0, 0], [10, 0, 0], [0, 10, 10]])
    ph = _phys(bcs, particle_cell=(2.0, 2.0, 2.0))
    pts = np.array([
        [-40, -40, -40],      # min corner -> 0
        [-41, -40, -40],      # below min: max(0, q) clamps -> 0 (flagged)
        [-38, -40, -40],      # exactly on the first cell boundary -> 1
        [-38.001, -40, -40],  # just below it -> 0
        [0, 0, 0],            # ix=iy=iz=20 -> 20*2025 + 20*45 + 20
        [49.9, 49.9, 49.9],   # last cell -> 91124
        [200, -40, -40],      # world-unit upper clamp: ix = int(min(90, 120)) = 90 -> key 90
        [-40, -37.5, -40],    # iy = 1 -> 45
        [-40, -40, -36.2],    # iz = 1 -> 2025
    ], F32)
    expect = np.array([0, 0, 1, 0, 41420, 91124, 90, 45, 2025], np.uint32)
    o = orc.Oracle(ph, [_point_type(bcs, len(pts))], mesh)
    o.set_particles(np.ascontiguousarray(pts.T).reshape(-1))
    keys = _grid_keys(orc, o, len(pts))
    np.testing.assert_array_equal(keys, expect)
    assert o.flags() & 0x1  # out-of-grid positions flagged


def test_cell_id_beyond_grid_is_clamped_and_flagged(bcs, orc):
    mesh = _tri_mesh(bcs, [[0, 0, 0], [10, 0, 0], [0, 10, 10]])
    ph = _phys(bcs, particle_cell=(2.0, 2.0, 2.0))
    pts = np.array([[50, 50, 50], [0, 0, 0]], F32)  # q = 45 on every axis -> key 93195 >= 45^3
    o = orc.Oracle(ph, [_point_type(bcs, 2)], mesh)
    o.set_particles(np.ascontiguousarray(pts.T).reshape(-1))
    keys = _grid_keys(orc, o, 2)
    assert keys[0] == 45 ** 3 - 1 and keys[1] == 41420
    assert o.flags() & 0x1


# ------------------------------------------------------------- sort + ranges --

def test_sort_is_numpy_stable_argsort(bcs, orc):
    rng = np.random.default_rng(7)
    n = 5000
    mesh = _tri_mesh(bcs, [[0, 0, 0], [10, 0, 0], [0, 10, 10]])
    ph = _phys(bcs, particle_cell=(7.0, 7.0, 7.0))  # coarse cells -> many ties
    pts = rng.uniform(-39, 49, size=(n, 3)).astype(F32)
    o = orc.Oracle(ph, [_point_type(bcs, n)], mesh)
    o.set_particles(np.ascontiguousarray(pts.T).reshape(-1))
    o.calculate_grid(0)
    keys_sorted, perm, ok, os_, oe = o.get_grid(0)
    per = np.empty(n, np.uint32)
    per[perm] = keys_sorted
    order = np.argsort(per, kind="stable")
    np.testing.assert_array_equal(perm, order)
    np.testing.assert_array_equal(keys_sorted, per[order])
    u, first, counts = np.unique(keys_sorted, return_index=True, return_counts=True)
    np.testing.assert_array_equal(ok, u)
    np.testing.assert_array_equal(os_, first)
    np.testing.assert_array_equal(oe, first + counts - 1)


# ---------------------------------------------------------------- contacts --

def test_contact_set_equals_brute_force(bcs, orc):
    """Grid scan over 27 cells finds exactly the O(N^2) contact set when the
    cell is at least the contact distance (r_i + r_j = 1.0)."""
    sc = bcs.Scene(0, 0.25, seed=3)  # 125 icosahedron cells, contact radius 0.5
    o = orc.Oracle(sc.physics, sc.types, sc.mesh)
    # squeeze the cells together so that there are many contacts
    p = soa(sc.positions, sc.particle_count).copy()
    p[1] = p[1] * F32(0.15) - F32(30)
    o.set_particles(p.reshape(-1), sc.velocities, sc.forces)
    o.calculate_grid(0)
    counts, pairs = o.contacts()
    n = sc.particle_count
    P = p.T.astype(F32)
    rel = (P[:, None, :] - P[None, :, :]).astype(F32)
    d2 = (rel[..., 0] * rel[..., 0] + rel[..., 1] * rel[..., 1]) + rel[..., 2] * rel[..., 2]
    hit = (d2 <= F32(1.0)) & (d2 >= F32(0.0001))
    np.fill_diagonal(hit, False)
    offs = np.concatenate([[0], np.cumsum(counts)])
    assert hit.sum() > 100
    for i in range(n):
        got = np.sort(pairs[offs[i]:offs[i + 1]])
        np.testing.assert_array_equal(got, np.flatnonzero(hit[i]), err_msg=f"particle {i}")


# ----------------------------------------------------------- closed forms --

def _two_particle_spring(bcs, x1, v1, rest=2.0):
    verts = [-1.0, 0.0, 0.0, 1.0, 0.0, 0.0]
    return bcs.CellTypeSpec(1, verts, [[0, 1, rest]], contact_radius=0.5)


def test_spring_and_environment_closed_form(bcs, orc):
    # far-away vein so that no ray hits and no contacts happen
    mesh = _tri_mesh(bcs, [[500, 500, 500], [510, 500, 500], [500, 510, 510]])
    ph = _phys(bcs, dt=0.01, particle_k_sniff=50.0, particle_d_fact=0.5, particle_cell=(5.0, 5.0, 5.0),
               triangle_cell=(25.0, 25.0, 25.0), grid_xz_margin=600.0, grid_y_margin=600.0)
    t = _two_particle_spring(bcs, 3.0, 1.0)
    o = orc.Oracle(ph, [t], mesh)
    pos = np.array([0, 3, 0, 0, 0, 0], F32)  # SoA: x = [0, 3], y = [0, 0], z = [0, 0]
    vel = np.array([0, 1, 0, 0, 0, 0], F32)
    o.set_particles(pos, vel, np.zeros(6, F32))
    o.calculate_grid(0)
    o.calculate_grid(1)
    o.calculate_next_frame()
    _, v, f = o.get_particles()
    F = soa(f, 2)
    # particle 0: stretched spring |dP| = 3 > 2: (3-2)*50 + dot(n(dP), dv)*0.5 = 50.5 toward particle 1
    # environment G - 0.39 v (ratio |x - c| / R0 = 1.5 is not > 1.5); F = (0 + S + E) / 2
    np.testing.assert_allclose(F[:, 0], [50.5 / 2, -12.5, 0.0], rtol=1e-6)
    np.testing.assert_allclose(F[:, 1], [(-50.5 - 0.39) / 2, -12.5, 0.0], rtol=1e-6)


def test_big_cell_brake_branch(bcs, orc):
    mesh = _tri_mesh(bcs, [[500, 500, 500], [510, 500, 500], [500, 510, 510]])
    ph = _phys(bcs, dt=0.01, particle_k_sniff=50.0, particle_d_fact=0.5, particle_cell=(5.0, 5.0, 5.0),
               grid_xz_margin=600.0, grid_y_margin=600.0)
    t = _two_particle_spring(bcs, 4.0, 0.0)
    o = orc.Oracle(ph, [t], mesh)
    pos = np.array([0, 4, 0, 0, 0, 0], F32)
    vel = np.array([2, 2, 0, 0, 0, 0], F32)  # equal velocities: no damping term
    o.set_particles(pos, vel, np.zeros(6, F32))
    o.calculate_grid(0)
    o.calculate_grid(1)
    o.calculate_next_frame()
    F = soa(o.get_particles()[2], 2)
    # |dP| = 4: spring (4-2)*50 = 100; ratio = 2/1 = 2 > 1.5: env = G - 0.39*2*0.8*v
    ex = 100.0 - 0.39 * 2.0 * 0.8 * 2.0
    np.testing.assert_allclose(F[:, 0], [ex / 2, -12.5, 0.0], rtol=1e-6)


def test_collision_force_is_attractive_as_in_reference(bcs, orc):
    """Two contact particles at distance 0.8 (< r_i + r_j = 1.0), at rest:
    F_i += 0.5 * (-6 * (2*0.5 - 0.8) * d_hat) -> pulls i TOWARD j (the
    reference's sign; src/simulation/physics.cuh:139)."""
    mesh = _tri_mesh(bcs, [[500, 500, 500], [510, 500, 500], [500, 510, 510]])
    ph = _phys(bcs, dt=0.01, particle_cell=(1.0, 1.0, 1.0), grid_xz_margin=600.0, grid_y_margin=600.0)
    o = orc.Oracle(ph, [_point_type(bcs, 2)], mesh)
    pos = np.array([0.8, 0.0, 0, 0, 0, 0], F32)
    o.set_particles(pos, np.zeros(6, F32), np.zeros(6, F32))
    o.calculate_grid(0)
    o.calculate_grid(1)
    o.calculate_next_frame()
    F = soa(o.get_particles()[2], 2)
    np.testing.assert_allclose(F[:, 0], [-0.6, -12.5, 0.0], rtol=1e-6)
    np.testing.assert_allclose(F[:, 1], [0.6, -12.5, 0.0], rtol=1e-6)
    counts, pairs = o.contacts()
    np.testing.assert_array_equal(counts, [1, 1])


def test_collision_damping_and_shear(bcs, orc):
    mesh = _tri_mesh(bcs, [[500, 500, 500], [510, 500, 500], [500, 510, 510]])
    ph = _phys(bcs, dt=0.01, particle_cell=(1.0, 1.0, 1.0), grid_xz_margin=600.0, grid_y_margin=600.0,
               viscous_damping=0.0, gravity=(0.0, 0.0, 0.0))
    o = orc.Oracle(ph, [_point_type(bcs, 2)], mesh)
    pos = np.array([0.8, 0.0, 0, 0, 0, 0], F32)
    vel = np.array([1.0, 0.0, 2.0, 0.0, 0, 0], F32)  # v0 = (1, 2, 0), v1 = 0
    o.set_particles(pos, vel, np.zeros(6, F32))
    o.calculate_grid(0)
    o.calculate_grid(1)
    o.calculate_next_frame()
    F = soa(o.get_particles()[2], 2)
    # rv = (1, 2, 0); n = (1,0,0); tangential = (0, 2, 0)
    # 0.5 * (spring (-1.2, 0, 0) + damping 6*(1,2,0) + shear 4*(0,2,0)) = (2.4, 10, 0)
    np.testing.assert_allclose(F[:, 0], [2.4, 10.0, 0.0], rtol=1e-6, atol=1e-6)


def test_ray_hit_reflection_reaction_and_barycentric_quirk(bcs, orc):
    verts = [[-10, -5, -10], [10, -5, -10], [0, -5, 10]]
    mesh = _tri_mesh(bcs, verts)
    ph = _phys(bcs, dt=0.01, particle_cell=(2.0, 2.0, 2.0), triangle_cell=(25.0, 25.0, 25.0))
    o = orc.Oracle(ph, [_point_type(bcs, 1)], mesh)
    o.set_particles(np.zeros(3, F32), np.array([0, -10, 0], F32), np.zeros(3, F32))
    o.calculate_grid(0)
    o.calculate_grid(1)
    o.calculate_next_frame()
    _, v, f = o.get_particles()
    # hit at t = 5 (<= 6): reflection about n = (0,1,0): v = v + (0.96*|v|*refl - v) = (0, 9.6, 0)
    np.testing.assert_allclose(v, [0, 9.6, 0], rtol=1e-6)
    # reaction force removes the normal component of F = (G - 0.39 v)/2 = (0, -10.55, 0)
    np.testing.assert_allclose(f, [0, 0, 0], atol=1e-6)
    # vein: ds = 0.005 * v_old = (0, -0.05, 0); reference barycentric of (0,-5,0)
    # with edges (v1-v0, v2-v1): (0.75, 0.5, -0.25)  [textbook: (0.25, 0.25, 0.5)]
    Fv = soa(o.get_vein()[2], 3)
    np.testing.assert_allclose(Fv[1], [-0.0375, -0.025, 0.0125], rtol=1e-5)
    np.testing.assert_allclose(Fv[0], 0, atol=1e-7)
    np.testing.assert_allclose(Fv[2], 0, atol=1e-7)


def test_ray_misses(bcs, orc):
    verts = [[-10, -5, -10], [10, -5, -10], [0, -5, 10]]
    mesh = _tri_mesh(bcs, verts)
    ph = _phys(bcs, dt=0.01, particle_cell=(2.0, 2.0, 2.0), triangle_cell=(25.0, 25.0, 25.0))
    cases = [
        ([0, 0, 0], [0, 10, 0]),      # moving away (t < 0)
        ([0, 0, 0], [10, 0, 0]),      # parallel to the plane
        ([30, 0, 0], [0, -10, 0]),    # outside the triangle (u > 1)
        ([0, 0, 0], [0, 0, 0]),       # zero velocity: normalize -> 0, no hit
        ([0, 3, 0], [0, -10, 0]),     # hit at t = 8 > 6: no response
    ]
    for p0, v0 in cases:
        o = orc.Oracle(ph, [_point_type(bcs, 1)], mesh)
        o.set_particles(np.array(p0, F32), np.array(v0, F32), np.zeros(3, F32))
        o.calculate_grid(0)
        o.calculate_grid(1)
        o.calculate_next_frame()
        _, v, _ = o.get_particles()
        np.testing.assert_array_equal(v, np.array(v0, F32), err_msg=str((p0, v0)))
        np.testing.assert_array_equal(o.get_vein()[2], 0)


def test_heun_integrator_constant_force(bcs, orc):
    mesh = _tri_mesh(bcs, [[500, 500, 500], [510, 500, 500], [500, 510, 510]])
    ph = _phys(bcs, dt=0.01, grid_xz_margin=600.0, grid_y_margin=600.0)
    o = orc.Oracle(ph, [_point_type(bcs, 1)], mesh)
    x0, v0, F = np.array([1, 2, 3], F32), np.array([4, -5, 6], F32), np.array([10, 20, -30], F32)
    o.set_particles(x0, v0, F)
    o.propagate_all()
    x, v, f = o.get_particles()
    dt = 0.01
    np.testing.assert_allclose(v, v0 + dt * F, rtol=1e-6)
    np.testing.assert_allclose(x, x0 + dt * v0 + 0.5 * dt * dt * F, rtol=1e-6)  # exact for constant F
    np.testing.assert_array_equal(f, F)  # forces persist (not reset)


def test_rk4_spring_closed_form(bcs, orc):
    """The Runge-Kutta build's spring (physics.cuh:57-71) on the pair of
    test_spring_and_environment_closed_form, worked along x by hand.
    Particle 0: dP = -3, dv = -1, F0 = 0, shift = +x; k1x = dt dv = -0.01,
    k1v = dt F0 = 0; spring(dP', dv') = (|dP'| - 2) 50 + sign(dP') dv' 0.5:
    k2v = dt spring(-3.005, -1) = 0.5075; k2x = dt (dv + k1x / 2) = -0.01005;
    k3v = dt spring(-3.005025, -1 + k2v / 2) = 0.50624375;
    k3x = dt (dv + k2x / 2) = -0.01005025;
    k4v = dt spring(-3 + k3x, -1 + k3v) = 0.50749390625;
    S = 2 k2v + 2 k3v + k4v (k1v left out), F = (F0 + S + E) / 6.
    Particle 1 mirrors it, with E = G - 0.39 v."""
    mesh = _tri_mesh(bcs, [[500, 500, 500], [510, 500, 500], [500, 510, 510]])
    ph = _phys(bcs, dt=0.01, particle_k_sniff=50.0, particle_d_fact=0.5, particle_cell=(5.0, 5.0, 5.0),
               triangle_cell=(25.0, 25.0, 25.0), grid_xz_margin=600.0, grid_y_margin=600.0)
    ph.particle_rk4 = 1
    t = _two_particle_spring(bcs, 3.0, 1.0)
    o = orc.Oracle(ph, [t], mesh)
    o.set_particles(np.array([0, 3, 0, 0, 0, 0], F32), np.array([0, 1, 0, 0, 0, 0], F32), np.zeros(6, F32))
    o.calculate_grid(0)
    o.calculate_grid(1)
    o.calculate_next_frame()
    F = soa(o.get_particles()[2], 2)
    S = 2 * 0.5075 + 2 * 0.50624375 + 0.50749390625
    np.testing.assert_allclose(F[:, 0], [S / 6, -25.0 / 6, 0.0], rtol=2e-6)
    np.testing.assert_allclose(F[:, 1], [(-S - 0.39) / 6, -25.0 / 6, 0.0], rtol=2e-6)
    # the Heun build on the same pair divides by 2 and uses dv + dt (F0 - F0') instead
    ph.particle_rk4 = 0
    o2 = orc.Oracle(ph, [t], mesh)
    o2.set_particles(np.array([0, 3, 0, 0, 0, 0], F32), np.array([0, 1, 0, 0, 0, 0], F32), np.zeros(6, F32))
    o2.calculate_grid(0)
    o2.calculate_grid(1)
    o2.calculate_next_frame()
    np.testing.assert_allclose(soa(o2.get_particles()[2], 2)[:, 0], [50.5 / 2, -12.5, 0.0], rtol=1e-6)


def test_rk4_integrator_closed_form(bcs, orc):
    """The Runge-Kutta build's position update (blood_cells.cu:166-171) with
    the updated velocity v = v0 + dt F: k1 = dt v, k2 = dt (v + k1/2),
    k3 = dt (v + k2/2), k4 = dt (v + k3), dx = (k1 + 2k2 + 2k3 + k4)/6
    = dt v (6 + 3 dt + dt^2 + dt^3/4) / 6."""
    mesh = _tri_mesh(bcs, [[500, 500, 500], [510, 500, 500], [500, 510, 510]])
    ph = _phys(bcs, dt=0.01, grid_xz_margin=600.0, grid_y_margin=600.0)
    ph.particle_rk4 = 1
    o = orc.Oracle(ph, [_point_type(bcs, 1)], mesh)
    x0, v0, F = np.array([1, 2, 3], F32), np.array([4, -5, 6], F32), np.array([10, 20, -30], F32)
    o.set_particles(x0, v0, F)
    o.propagate_all()
    x, v, f = o.get_particles()
    dt = 0.01
    vn = v0.astype(np.float64) + dt * F
    np.testing.assert_allclose(v, vn, rtol=1e-6)
    np.testing.assert_allclose(x, x0 + dt * vn * (6 + 3 * dt + dt ** 2 + dt ** 3 / 4) / 6, rtol=1e-6)
    np.testing.assert_array_equal(f, F)


def test_vein_springs_and_euler(bcs, orc):
    # triangle (0,0,0), (4,0,0), (0,3,0); vertex 1 moved to (5,0,0)
    verts = np.array([[0, 0, 0], [4, 0, 0], [0, 3, 0]], F32)
    mesh = _tri_mesh(bcs, verts)
    ph = _phys(bcs, dt=0.5, triangle_cell=(25.0, 25.0, 25.0), grid_xz_margin=600.0, grid_y_margin=600.0)
    o = orc.Oracle(ph, [_point_type(bcs, 1)], mesh)
    o.set_particles(np.array([300, 300, 300], F32), np.zeros(3, F32), np.zeros(3, F32))
    X = np.ascontiguousarray(verts.T).reshape(-1).copy()
    X[1] = 5.0
    V = np.zeros(9, F32)
    V[1] = 1.0  # vertex 1 moving +x at 1
    o.set_vein(X, V, np.zeros(9, F32))
    o.calculate_grid(0)
    o.calculate_grid(1)
    o.calculate_next_frame()
    Fv = soa(o.get_vein()[2], 3)
    # vertex 0: neighbours 1 (twice? no: one triangle lists each neighbour once) and 2
    # spring 0-1: (|X0-X1| - 4)*0.1 + dot(n(X0-X1), V0-V1)*0.7 = 0.1 + 0.7 = 0.8 along n(X1-X0) = +x
    np.testing.assert_allclose(Fv[:, 0], [0.8, 0, 0], atol=1e-6)
    # vertex 2: spring 2-1: |X2-X1| = sqrt(34), L = 5
    d = np.array([5, -3, 0]) / np.sqrt(34)
    s = (np.sqrt(34) - 5) * 0.1 + np.dot(-d, [-1, 0, 0]) * 0.7
    np.testing.assert_allclose(Fv[:, 2], s * d, rtol=1e-5)
    o.propagate_all()
    X2, V2, F2 = o.get_vein()
    np.testing.assert_allclose(soa(V2, 3)[:, 0], [0.4, 0, 0], atol=1e-6)  # V += dt*F
    np.testing.assert_allclose(soa(X2, 3)[:, 0], [0.2, 0, 0], atol=1e-6)  # X += dt*V(new)
    np.testing.assert_array_equal(F2, 0)


def test_rest_shape_cell_force_recurrence(bcs, orc):
    """A cell at its rest shape (centroid at the origin, zero velocity) has
    zero spring force; its first frame is F = (0 + G)/2 exactly and the
    following frames follow F_n = (F_{n-1} + G - mu*v_{n-1})/2, v_n = v_{n-1} + dt*F_n."""
    sc = bcs.Scene(0, 0.002, seed=0)  # one icosahedron cell
    t = sc.types[0]
    mesh = _tri_mesh(bcs, [[500, 500, 500], [510, 500, 500], [500, 510, 510]])
    ph = _phys(bcs, dt=0.01, particle_k_sniff=50.0, particle_d_fact=0.5, particle_cell=(1.0, 1.0, 1.0),
               grid_xz_margin=600.0, grid_y_margin=600.0)
    o = orc.Oracle(ph, [bcs.CellTypeSpec(1, t.vertices, np.stack([t.s_start, t.s_end, t.s_len], 1), 0.5)], mesh)
    pos = np.ascontiguousarray(t.vertices.reshape(-1, 3).T).reshape(-1)
    o.set_particles(pos, np.zeros_like(pos), np.zeros_like(pos))
    o.step(1)
    _, v, f = o.get_particles()
    assert np.all(soa(f, 12)[1] == F32(-12.5)) and np.all(soa(f, 12)[0] == 0) and np.all(soa(f, 12)[2] == 0)
    Fy, vy = -12.5, -0.125
    for _ in range(20):
        o.step(1)
        Fy = (Fy + (-25.0 - 0.39 * vy)) / 2
        vy = vy + 0.01 * Fy
    _, v, f = o.get_particles()
    np.testing.assert_allclose(soa(f, 12)[1], Fy, rtol=1e-4)
    np.testing.assert_allclose(soa(v, 12)[1], vy, rtol=1e-4)


def test_teleport_is_rigid_model_translate(bcs, orc):
    sc = bcs.Scene(0, 0.01, seed=0)  # 5 cells
    ph = sc.physics
    ph.use_blood_flow = 1
    o = orc.Oracle(ph, sc.types, sc.mesh)
    n = sc.particle_count
    p = soa(sc.positions, n).copy()
    p[1, :12] -= F32(600.0)  # cell 0 far below the lower threshold
    o.set_particles(p.reshape(-1), sc.velocities, sc.forces)
    o.set_step_index(17)
    o.propagate_all()
    x, v, _ = o.get_particles()
    X, V = soa(x, n), soa(v, n)
    m = sc.types[0].vertices.reshape(-1, 3)
    u1, u2 = orc.uniform(ph.seed, 17, 0, 0), orc.uniform(ph.seed, 17, 0, 1)
    assert 0 < u1 <= 1 and 0 < u2 <= 1
    for a in range(12):
        ex = [(F32(u1) - F32(0.5)) * F32(1.2) * F32(ph.cylinder_radius) + m[a, 0] - m[0, 0],
              F32(ph.min_spawn_y) + m[a, 1] - m[0, 1],
              (F32(u2) - F32(0.5)) * F32(1.2) * F32(ph.cylinder_radius) + m[a, 2] - m[0, 2]]
        np.testing.assert_array_equal(X[:, a], np.array(ex, F32))
        np.testing.assert_array_equal(V[:, a], np.array(ph.init_velocity[:], F32))
    # the other cells only integrated (no teleport)
    assert np.all(X[1, 12:] > -500)


def test_oracle_is_deterministic(bcs, orc):
    sc = bcs.Scene(0, 0.2, seed=5)
    res = []
    for _ in range(2):
        o = orc.Oracle(sc.physics, sc.types, sc.mesh)
        o.set_particles(sc.positions, sc.velocities, sc.forces)
        o.step(20)
        res.append(o.get_particles())
    for a, b in zip(*res):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_fp64_rerun_drift_is_small_over_100_steps(bcs, orc):
    sc = bcs.Scene(0, 0.2, seed=5)
    a = orc.Oracle(sc.physics, sc.types, sc.mesh, "f32")
    b = orc.Oracle(sc.physics, sc.types, sc.mesh, "f64")
    for o in (a, b):
        o.set_particles(sc.positions, sc.velocities, sc.forces)
        o.step(100)
    pa, pb = a.get_particles()[0], b.get_particles()[0]
    drift = np.abs(pa.astype(np.float64) - pb)
    assert np.isfinite(drift).all()
    assert drift.max() < 1e-2, drift.max()


# ------------------------------------- first hit, launch type, teleport --
#
# Scenes whose outcome depends on the cell visit order of the ray search, on
# the launch type's model table in the contact search and on which cell types
# test the ending spheres.  The expected values are worked out by hand in the
# comments; a nearest-hit search, a z-outer loop, a radius from the
# particle's own type, a floor (not truncating) remainder or ending spheres
# tested for every type fails one of them.

def _anchored_mesh(tris, lo=0.0, hi=100.0):
    """Separate-vertex triangles plus two anchor vertices (in no triangle) at
    (lo, lo, lo) and (hi, hi, hi), so that with zero margins the grid bounds
    are exactly [lo, hi]^3."""
    verts = [v for t in tris for v in t] + [[lo, lo, lo], [hi, hi, hi]]
    idx = [[3 * k, 3 * k + 1, 3 * k + 2] for k in range(len(tris))]
    return np.asarray(verts, F32), np.asarray(idx, np.uint32)


# Three horizontal triangles on the ray x = z = 15, y = 15 + t (particle at
# (15, 15, 15) moving +y; triangle grid cell 10 on [0, 100]^3, so the particle
# is in cell (1, 1, 1)).  Centroids:
#   A  y = 17 (t = 2)  centroid (15, 17, 15)    -> cell (1, 1, 1)  dx=0,  dz=0
#   C  y = 18 (t = 3)  centroid (15, 18, 9.33)  -> cell (1, 1, 0)  dx=0,  dz=-1
#   B  y = 19 (t = 4)  centroid (8, 19, 16.67)  -> cell (0, 1, 1)  dx=-1, dz=0
# The x-outer, y, z-inner visit order reaches B's cell (dx = -1) first, so B
# is the first hit although A is nearer (nearest hit: A; z-outer order: C;
# triangle-index order: A).
TRI_A = [[14, 17, 14], [17, 17, 14], [14, 17, 17]]
TRI_C = [[12, 18, 4], [18, 18, 4], [15, 18, 20]]
TRI_B = [[2, 19, 14], [20, 19, 14], [2, 19, 22]]


def first_hit_scene(bcs, tris):
    verts, idx = _anchored_mesh(tris)
    mesh = bcs.MeshSpec(verts, idx)
    ph = _phys(bcs, dt=0.01, particle_cell=(5.0, 5.0, 5.0), triangle_cell=(10.0, 10.0, 10.0),
               grid_xz_margin=0.0, grid_y_margin=0.0)
    types = [_point_type(bcs, 1)]
    pos = np.array([15, 15, 15], F32)
    vel = np.array([0, 10, 0], F32)
    return ph, types, mesh, pos, vel


def _one_frame(orc, ph, types, mesh, pos, vel, force=None):
    o = orc.Oracle(ph, types, mesh)
    o.set_particles(pos, vel, np.zeros_like(pos) if force is None else force)
    o.calculate_grid(0)
    o.calculate_grid(1)
    o.calculate_next_frame()
    return o


def test_first_hit_follows_x_outer_cell_order(bcs, orc):
    ph, types, mesh, pos, vel = first_hit_scene(bcs, [TRI_A, TRI_C, TRI_B])
    o = _one_frame(orc, ph, types, mesh, pos, vel)
    _, v, f = o.get_particles()
    # every candidate is a horizontal plane: reflection (0, -1, 0),
    # v = v + (0.96 * 10 * (0, -1, 0) - v) = (0, -9.6, 0); reaction force
    # removes all of F = (G - 0.39 v) / 2, which has only a y component
    np.testing.assert_allclose(v, [0, -9.6, 0], rtol=1e-6)
    np.testing.assert_allclose(f, [0, 0, 0], atol=1e-6)
    Fv = soa(o.get_vein()[2], mesh.positions.size // 3)
    # ds = 0.005 * (0, 10, 0) = (0, 0.05, 0) on B's corners (vertices 6, 7, 8)
    # with the reference's barycentric of (15, 19, 15) in B:
    #   e0 = (18, 0, 0), e1 = v2 - v1 = (-18, 0, 8), e2 = (13, 0, 1)
    #   d00 = 324, d01 = -324, d11 = 388, d20 = 234, d21 = -226, den = 20736
    #   x = (388*234 - 324*226) / 20736 = 17568/20736, y = 2592/20736, z = 1 - x - y
    x, y = 17568 / 20736, 2592 / 20736
    np.testing.assert_allclose(Fv[1, 6:9], [0.05 * x, 0.05 * y, 0.05 * (1 - x - y)], rtol=1e-5)
    np.testing.assert_array_equal(Fv[:, 0:6], 0)  # A and C untouched
    np.testing.assert_array_equal(Fv[[0, 2], 6:9], 0)


def test_first_hit_without_the_x_neighbour_is_the_z_neighbour(bcs, orc):
    """With B removed the dx = 0 column is next: C's cell (dz = -1) comes
    before A's (dz = 0), so the farther C still beats the nearer A."""
    ph, types, mesh, pos, vel = first_hit_scene(bcs, [TRI_A, TRI_C])
    o = _one_frame(orc, ph, types, mesh, pos, vel)
    Fv = soa(o.get_vein()[2], mesh.positions.size // 3)
    # barycentric of (15, 18, 15) in C: e0 = (6, 0, 0), e1 = (-3, 0, 16),
    # e2 = (3, 0, 11): d00 = 36, d01 = -18, d11 = 265, d20 = 18, d21 = 167,
    # den = 36*265 - 324 = 9216; x = (265*18 + 18*167)/9216 = 7776/9216,
    # y = (36*167 + 18*18)/9216 = 6336/9216
    x, y = 7776 / 9216, 6336 / 9216
    np.testing.assert_allclose(Fv[1, 3:6], [0.05 * x, 0.05 * y, 0.05 * (1 - x - y)], rtol=1e-5)
    np.testing.assert_array_equal(Fv[:, 0:3], 0)


# Two triangles in the particle's own cell (1, 1, 1): D (index 0, y = 19,
# t = 4) and E (index 1, y = 17, t = 2).  Within a cell the stable sort keeps
# index order, so D is tested first and wins although E is nearer.
TRI_D = [[13, 19, 13], [19, 19, 13], [13, 19, 19]]
TRI_E = [[14, 17, 14], [17, 17, 14], [14, 17, 17]]


def test_first_hit_within_a_cell_follows_sorted_slots(bcs, orc):
    ph, types, mesh, pos, vel = first_hit_scene(bcs, [TRI_D, TRI_E])
    o = _one_frame(orc, ph, types, mesh, pos, vel)
    Fv = soa(o.get_vein()[2], mesh.positions.size // 3)
    # barycentric of (15, 19, 15) in D: e0 = (6, 0, 0), e1 = (-6, 0, 6),
    # e2 = (2, 0, 2): d00 = 36, d01 = -36, d11 = 72, d20 = 12, d21 = 0,
    # den = 1296; x = 864/1296, y = 432/1296, z = 0
    np.testing.assert_allclose(Fv[1, 0:2], [0.05 * 864 / 1296, 0.05 * 432 / 1296], rtol=1e-5)
    assert abs(Fv[1, 2]) < 1e-8
    np.testing.assert_array_equal(Fv[:, 3:6], 0)


# Launch-type radius.  Type 0: three particles in one cell, model vertices at
# x = 0, 1, 3 -> derived radii min distance / (2*3) = [1/6, 1/6, 1/3].
# Type 1: two particles, model vertices at x = 0, 2.4 -> [0.4, 0.4].
# Model table bs = [1/6, 1/6, 1/3, 0.4, 0.4]; particle starts [0, 3].
# Positions (grid cell 1, z cell boundary at z = 0):
#   P3 (0, 0, -0.3)  P1 (10, 0, -0.3)  P2 (20, 0, -0.3)  -> z cell 99
#   P0 (0, 0, 0.35)                                      -> z cell 100
#   P4 (30, 0, 20)                                       -> z cell 120
# Sorted slots: [P3, P1, P2, P0, P4]; slots 0-2 run with type 0's tables,
# slots 3-4 with type 1's (src/simulation/particle_collisions.cuh:124,76).
#   P0 in slot 3 (type 1): own radius bs[3 + (0-3) % 2] = bs[3 - 1] = 1/3
#       (C remainder -3 % 2 = -1), partner P3: bs[3 + 0 % 2] = 0.4;
#       0.65 <= 1/3 + 0.4 -> contact, spring (2/3 - 0.65) * -6 * 0.5 = -0.05 in z
#   P3 in slot 0 (type 0): bs[0 + 3 % 3] = 1/6, partner P0 bs[0] = 1/6;
#       0.65 > 1/3 -> no contact
# Own-type radii (1/6 + 0.4 = 0.567) give no contact at all; a floor
# remainder gives P0 radius 0.4 and a force of -0.45.

def launch_type_scene(bcs):
    mesh = _tri_mesh(bcs, [[500, 500, 500], [510, 500, 500], [500, 510, 510]])
    ph = _phys(bcs, dt=0.01, particle_cell=(1.0, 1.0, 1.0), grid_xz_margin=600.0, grid_y_margin=600.0)
    t0 = bcs.CellTypeSpec(1, [0, 0, 0, 1, 0, 0, 3, 0, 0], [], contact_radius=0.0)
    t1 = bcs.CellTypeSpec(1, [0, 0, 0, 2.4, 0, 0], [], contact_radius=0.0)
    P = np.array([[0, 0, 0.35], [10, 0, -0.3], [20, 0, -0.3], [0, 0, -0.3], [30, 0, 20]], F32)
    return ph, [t0, t1], mesh, np.ascontiguousarray(P.T).reshape(-1)


def test_contact_radius_comes_from_the_launch_type(bcs, orc):
    ph, types, mesh, pos = launch_type_scene(bcs)
    o = orc.Oracle(ph, types, mesh)
    bs, _, _ = o.tables()
    np.testing.assert_allclose(bs, [1 / 6, 1 / 6, 1 / 3, 0.4, 0.4], rtol=1e-6)
    o.set_particles(pos, np.zeros_like(pos), np.zeros_like(pos))
    o.calculate_grid(0)
    keys, perm, *_ = o.get_grid(0)
    np.testing.assert_array_equal(perm, [3, 1, 2, 0, 4])
    counts, pairs = o.contacts()
    np.testing.assert_array_equal(counts, [1, 0, 0, 0, 0])
    np.testing.assert_array_equal(pairs, [3])
    o.calculate_grid(1)
    o.calculate_next_frame()
    F = soa(o.get_particles()[2], 5)
    # cell springs (no springs, v = 0): F = G / 2; then the contact of P0
    np.testing.assert_allclose(F[:, 0], [0, -12.5, -0.05], rtol=1e-3, atol=1e-7)
    np.testing.assert_array_equal(F[:, 3], np.array([0, -12.5, 0], F32))


# Teleport selection (src/simulation/vein_end.cu:23-30): a type whose
# particles fill whole 32-lane groups (32 % ppc == 0) or which has at most 32
# particles checks only the thresholds; the others also check the ending
# spheres.  One cell of each type sits inside the ending sphere at (50, 0, 0),
# well inside the thresholds; only the ppc-12 type with 10 cells teleports.

def teleport_scene(bcs):
    mesh = bcs.MeshSpec(np.array([[-100, -100, -100], [100, -100, -100], [0, 100, 100]], F32),
                        np.array([[0, 1, 2]], np.uint32), [[50.0, 0.0, 0.0]], [5.0])
    ph = _phys(bcs, dt=0.01)
    ph.use_blood_flow = 1

    def ring(ppc):
        a = np.arange(ppc) * (2 * np.pi / ppc)
        return np.stack([np.cos(a), np.sin(a), np.zeros(ppc)], 1).astype(F32).reshape(-1)

    types = [bcs.CellTypeSpec(10, ring(8), [], 0.5),    # 32 % 8 == 0: thresholds only
             bcs.CellTypeSpec(10, ring(12), [], 0.5),   # 120 particles, 32 % 12 != 0: spheres too
             bcs.CellTypeSpec(2, ring(12), [], 0.5)]    # 24 <= 32 particles: thresholds only
    pos = []
    for t in types:
        for c in range(t.count):
            centre = np.array([50.0, 0.0, 0.0]) if c == 0 else np.array([-60.0 + 4 * c, 30.0 * len(pos) / 100, 0.0])
            pos.append(t.vertices.reshape(-1, 3) + centre)
    P = np.concatenate(pos).astype(F32)
    return ph, types, mesh, np.ascontiguousarray(P.T).reshape(-1)


def test_teleport_ending_spheres_only_for_block_synchronised_types(bcs, orc):
    ph, types, mesh, pos = teleport_scene(bcs)
    o = orc.Oracle(ph, types, mesh)
    n = pos.size // 3
    o.set_particles(pos, np.zeros_like(pos), np.zeros_like(pos))
    o.propagate_all()
    X, V = soa(o.get_particles()[0], n), soa(o.get_particles()[1], n)
    P0 = soa(pos, n)
    moved = np.flatnonzero(np.any(X != P0, axis=0))
    # type 1 starts at particle 80; its cell 0 is particles 80..91
    np.testing.assert_array_equal(moved, np.arange(80, 92))
    np.testing.assert_array_equal(X[1, 80:92], np.float32(ph.min_spawn_y) + (P0[1, 80:92] - P0[1, 80]))
    np.testing.assert_array_equal(V[:, 80:92], np.repeat(np.array(ph.init_velocity[:], F32)[:, None], 12, 1))


# Runtime switches (src/config/simulation.hpp:6-8).

def test_reaction_force_off_adds_the_resilient_force(bcs, orc):
    """enableReactionForce = 0 (src/simulation/vein_collisions.cu:245-249):
    the wall pushes with addResilientForceOnCollision on relativePosition =
    pos - hit = (0, 5, 0): dir (0, 1, 0), spring -6 (2*0.5 - 5) = 24,
    damping 6 * (0, -10, 0), no shear: 0.5 * (0, -36, 0) = (0, -18, 0) added
    to F = (G - 0.39 v) / 2 = (0, -10.55, 0)."""
    verts = [[-10, -5, -10], [10, -5, -10], [0, -5, 10]]
    mesh = _tri_mesh(bcs, verts)
    ph = _phys(bcs, dt=0.01, particle_cell=(2.0, 2.0, 2.0), triangle_cell=(25.0, 25.0, 25.0),
               enable_reaction_force=0)
    o = _one_frame(orc, ph, [_point_type(bcs, 1)], mesh, np.zeros(3, F32), np.array([0, -10, 0], F32))
    _, v, f = o.get_particles()
    np.testing.assert_allclose(f, [0, -28.55, 0], rtol=1e-6)
    np.testing.assert_allclose(v, [0, 9.6, 0], rtol=1e-6)  # the reflection is unchanged


def test_big_cells_brake_off(bcs, orc):
    """enableBigCellsBrake = 0 (src/simulation/physics.cuh:102-120): ratio 2
    no longer brakes: env = G - 0.39 v, F_x = (100 - 0.39*2) / 2."""
    mesh = _tri_mesh(bcs, [[500, 500, 500], [510, 500, 500], [500, 510, 510]])
    ph = _phys(bcs, dt=0.01, particle_k_sniff=50.0, particle_d_fact=0.5, particle_cell=(5.0, 5.0, 5.0),
               grid_xz_margin=600.0, grid_y_margin=600.0, enable_big_cells_brake=0)
    t = _two_particle_spring(bcs, 4.0, 0.0)
    o = orc.Oracle(ph, [t], mesh)
    o.set_particles(np.array([0, 4, 0, 0, 0, 0], F32), np.array([2, 2, 0, 0, 0, 0], F32), np.zeros(6, F32))
    o.calculate_grid(0)
    o.calculate_grid(1)
    o.calculate_next_frame()
    F = soa(o.get_particles()[2], 2)
    np.testing.assert_allclose(F[:, 0], [(100.0 - 0.78) / 2, -12.5, 0.0], rtol=1e-6)


def test_blood_flow_off_keeps_cells_beyond_the_thresholds(bcs, orc):
    """useBloodFlow = 0 (src/simulation/simulation_controller.cu:328): no
    teleport; the cell below the lower threshold only integrates."""
    sc = bcs.Scene(0, 0.01, seed=0)
    ph = sc.physics
    ph.use_blood_flow = 0
    o = orc.Oracle(ph, sc.types, sc.mesh)
    n = sc.particle_count
    p = soa(sc.positions, n).copy()
    p[1, :12] -= F32(600.0)
    v = soa(sc.velocities, n)
    o.set_particles(p.reshape(-1), sc.velocities, np.zeros(3 * n, F32))
    o.propagate_all()
    X = soa(o.get_particles()[0], n)
    dt = F32(ph.dt)
    np.testing.assert_array_equal(X, p + F32(0.5) * dt * (v + v))
