"""The learned spring surrogate (bcs_surrogate): the MFMA network against a
numpy reference of the same arithmetic (bf16 inputs, weights and hidden
activations, fp32 accumulation), its training, the training rows sampled
from the plain method against the oracle, the bit-exact fallback, and the
per-cell error bound of surrogate frames."""
import numpy as np
import pytest

from conftest import assert_bits_equal

pytestmark = pytest.mark.gpu

F32 = np.float32


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def gelu_tanh(z):
    """GELU, tanh form: 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))."""
    return 0.5 * z * (1.0 + np.tanh(np.sqrt(2.0 / np.pi) * (z + 0.044715 * z ** 3)))


def test_gelu_tanh_reference_values():
    """The reference activation at known points: GELU(0) = 0, the odd part
    GELU(x) - GELU(-x) = x, the tanh-form values at 1 and -1 (0.841192,
    -0.158808), and within 2e-4 of the erf form x Phi(x) there (0.841345,
    -0.158655)."""
    assert gelu_tanh(0.0) == 0.0
    x = np.linspace(-6, 6, 101)
    np.testing.assert_allclose(gelu_tanh(x) - gelu_tanh(-x), x, atol=1e-12)
    np.testing.assert_allclose(gelu_tanh(np.array([1.0, -1.0])), [0.8411920, -0.1588080], atol=1e-6)
    np.testing.assert_allclose(gelu_tanh(np.array([1.0, -1.0])), [0.8413447, -0.1586553], atol=2e-4)


def reference_forward(ws, bs, x):
    """x: rows already divided by the feature scales."""
    h = bf16(x).astype(np.float64)
    for l in range(3):
        z = h @ bf16(ws[l]).astype(np.float64).T + bs[l]
        h = bf16(gelu_tanh(z).astype(np.float32)).astype(np.float64)
    return h @ bf16(ws[3]).astype(np.float64).T + bs[3]


def _features(rng, n):
    x = np.zeros((n, 294), F32)
    x[:, :126] = rng.normal(0, 1.5, (n, 126))
    x[:, 126:252] = rng.normal(0, 0.3, (n, 126))
    x[:, 252:] = rng.integers(0, 2, (n, 42))
    return x


def test_mlp_matches_numpy_reference(bcs):
    rng = np.random.default_rng(1)
    net = bcs.Surrogate(seed=3)
    net.set_scales(2.0, 0.5, 1e-3)
    ws, bs = net.get_weights()
    assert [w.shape for w in ws] == [(256, 294), (256, 256), (256, 256), (126, 256)]
    x = _features(rng, 1000)  # 8 row tiles, the last one partial
    got = net.predict(x)
    xs = x.copy()
    xs[:, :126] /= F32(2.0)
    xs[:, 126:252] /= F32(0.5)
    ref = reference_forward(ws, bs, xs) * 1e-3
    err = np.abs(got - ref)
    assert err.max() <= 0.02 * np.abs(ref).max(), (err.max(), np.abs(ref).max())
    assert err.mean() <= 0.002 * np.abs(ref).mean(), (err.mean(), np.abs(ref).mean())
    # the weights round-trip, and new weights change the prediction as numpy says
    ws2 = [w * F32(0.5) for w in ws]
    net.set_weights(ws2, bs)
    w3, b3 = net.get_weights()
    for a, b in zip(w3, ws2):
        assert_bits_equal(a, b, "weights round trip")
    ref2 = reference_forward(ws2, bs, xs) * 1e-3
    got2 = net.predict(x)
    assert np.abs(got2 - ref2).max() <= 0.02 * np.abs(ref2).max()


def test_training_fits_a_teacher_network(bcs):
    rng = np.random.default_rng(2)
    teacher = bcs.Surrogate(seed=11)
    x = _features(rng, 16384)
    y = teacher.predict(x)
    student = bcs.Surrogate(seed=12)
    loss = student.train(x, y, epochs=12, batch=1024, lr=2e-3, seed=5)
    assert np.all(np.isfinite(loss))
    assert loss[-1] < 0.2 * loss[0], loss
    assert np.all(np.diff(loss[2:]) < 0.05 * loss[0]), loss  # no divergence after the start


def _disc_scene(bcs, n_discs=40, n_ico=30, seed=0, squeeze=0.3):
    p = bcs.Scene.params(0)
    p.n_cells = 2
    p.cells[0] = bcs.SceneCell(bcs.CELL_DISC42, n_discs, 2.0, 0.35, 0.5)
    p.cells[1] = bcs.SceneCell(bcs.CELL_ICOSAHEDRON, n_ico, 2.0, 1.0, 0.5)
    p.seed = seed
    sc = bcs.Scene(params=p)
    q = sc.positions.reshape(3, -1).copy()
    q[1] = q[1] * F32(squeeze) - F32(60)
    sc.positions = q.reshape(-1)
    return sc


def test_sampled_rows_match_the_oracle(bcs, orc):
    """Isolated, compressed discs far from the wall: the sampled target of a
    particle is (dt^2/4) S with S = 2 F' - F0 - E, where F' is the force after
    the oracle's calculateNextFrame (no contacts, no wall), F0 the force before
    it and E = G - 0.39 v the environment force."""
    p = bcs.Scene.params(0)
    p.n_cells = 1
    p.cells[0] = bcs.SceneCell(bcs.CELL_DISC42, 8, 2.0, 0.35, 0.01)  # contact only below 0.02: none
    p.radius = 60.0
    sc = bcs.Scene(params=p)
    n = sc.particle_count
    q = sc.positions.reshape(3, -1).copy()
    for c in range(8):  # cells 12 apart along x at the axis, each squeezed to 80%
        sl = slice(42 * c, 42 * (c + 1))
        cen = q[:, sl].mean(axis=1, keepdims=True)
        q[:, sl] = (q[:, sl] - cen) * F32(0.8) + np.array([[-42 + 12 * c], [-100], [0]], F32)
    sc.positions = q.reshape(-1).astype(F32)
    sim = bcs.Simulation.from_scene(sc)
    feat, targ = sim.surrogate_sample()
    assert feat.shape == (8, 294) and targ.shape == (8, 126)
    o = orc.Oracle(sc.physics, sc.types, sc.mesh)
    o.set_particles(sc.positions, sc.velocities, sc.forces)
    o.step(1)
    x0, v0, f0 = (a.reshape(3, n).astype(np.float64) for a in o.get_particles())
    o.calculate_grid(0)
    o.calculate_grid(1)
    o.calculate_next_frame()
    f1 = o.get_particles()[2].reshape(3, n).astype(np.float64)
    assert o.contacts()[1].size == 0  # isolated: no contact (and |x| + 2 < 60: no wall hit)
    G = np.array(sc.physics.gravity[:], np.float64)[:, None]
    E = G - np.float64(F32(sc.physics.viscous_damping)) * v0
    S = 2 * f1 - f0 - E
    dt = np.float64(F32(sc.physics.dt))
    d = (dt * dt / 4) * S  # (3, n)
    exp = d.T.reshape(8, 42, 3).reshape(8, 126)
    np.testing.assert_allclose(targ, exp, rtol=1e-4, atol=1e-6 * np.abs(exp).max())
    assert np.abs(exp).max() > 1e-7
    # features: centroid-relative positions and velocities of the frame's start
    rel = (x0.T.reshape(8, 42, 3) - x0.T.reshape(8, 42, 3).mean(axis=1, keepdims=True)).reshape(8, 126)
    np.testing.assert_allclose(feat[:, :126], rel, atol=1e-4)
    relv = (v0.T.reshape(8, 42, 3) - v0.T.reshape(8, 42, 3).mean(axis=1, keepdims=True)).reshape(8, 126)
    np.testing.assert_allclose(feat[:, 126:252], relv, atol=1e-4)
    assert np.all(feat[:, 252:] == 0)  # no contacts


def test_zero_tolerance_falls_back_bit_exactly(bcs, orc):
    """Tolerance 0: every disc cell falls back to the plain springs; the
    frames are the plain frames bit for bit (against the oracle)."""
    sc = _disc_scene(bcs)
    sim = bcs.Simulation.from_scene(sc)
    net = bcs.Surrogate(seed=1)
    net.set_scales(2.0, 1.0, 1e-4)
    sim.attach_surrogate(net, 0.0)
    sim.step(40)
    o = orc.Oracle(sc.physics, sc.types, sc.mesh)
    o.set_particles(sc.positions, sc.velocities, sc.forces)
    o.step(40)
    for x, y, name in zip(sim.get_particles(), o.get_particles(), ("pos", "vel", "force")):
        assert_bits_equal(x, y, f"tolerance 0, 40 frames: {name}")
    assert_bits_equal(sim.get_vein()[0], o.get_vein()[0], "tolerance 0: vein")
    st = sim.surrogate_stats()
    assert st["accepted"] == 0 and st["fallback"] == 40 * 40


def test_surrogate_frames_respect_the_tolerance(bcs, orc):
    """A network trained on sampled rows of the same scene: surrogate frames
    use it only for cells within tolerance of the plain solve; the largest
    accepted error stays within it, and a tighter tolerance accepts fewer
    cells."""
    sc = _disc_scene(bcs, n_discs=120, n_ico=0, seed=3, squeeze=0.5)
    sampler = bcs.Simulation.from_scene(sc)
    feats, targs = [], []
    for _ in range(6):
        f, t = sampler.surrogate_sample()
        feats.append(f)
        targs.append(t)
    x = np.concatenate(feats)
    y = np.concatenate(targs)
    net = bcs.Surrogate(seed=2)
    net.set_scales(float(np.abs(x[:, :126]).std()), float(max(np.abs(x[:, 126:252]).std(), 1e-6)),
                   float(max(y.std(), 1e-9)))
    loss = net.train(x, y, epochs=40, batch=128, lr=1e-3, seed=1)
    assert loss[-1] < loss[0]
    counts = {}
    for tol in (2e-3, 1e-7):
        sim = bcs.Simulation.from_scene(sc)
        sim.attach_surrogate(net, tol)
        sim.step(25)
        st = sim.surrogate_stats()
        assert st["accepted"] + st["fallback"] == 120 * 25
        assert st["max_error_accepted"] <= tol
        counts[tol] = st["accepted"]
        p = sim.get_particles()[0]
        assert np.all(np.isfinite(p))
    assert counts[2e-3] >= counts[1e-7]
    assert counts[2e-3] > 0
