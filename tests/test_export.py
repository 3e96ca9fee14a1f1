"""Frame export in the renderer's layouts and trajectory files."""
import ctypes as C

import numpy as np
import pytest

from conftest import assert_bits_equal


class DeviceBuffer:
    """A float32 device buffer from the HIP runtime (test scaffolding)."""

    hip = None

    def __init__(self, n, fill):
        if DeviceBuffer.hip is None:
            DeviceBuffer.hip = C.CDLL("libamdhip64.so")
        self.n = n
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(4 * n)) == 0
        host = np.full(n, fill, np.float32)
        assert self.hip.hipMemcpy(self.ptr, host.ctypes.data_as(C.c_void_p), C.c_size_t(4 * n), 1) == 0

    def read(self):
        host = np.empty(self.n, np.float32)
        assert self.hip.hipMemcpy(host.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(4 * self.n), 2) == 0
        return host

    def free(self):
        self.hip.hipFree(self.ptr)


def _write_fake(path, n, nv, frames, vel):
    head = bytearray(64)
    head[:8] = b"BCSTRAJ1"
    head[8:24] = np.array([n, nv, 2, 1 if vel else 0], "<i4").tobytes()
    head[24:28] = np.array([0.005], "<f4").tobytes()
    with open(path, "wb") as f:
        f.write(head)
        for k in range(frames):
            f.write(np.array([10 * k], "<u8").tobytes())
            f.write(np.full(3 * n, k, "<f4").tobytes())
            if vel:
                f.write(np.full(3 * n, -k, "<f4").tobytes())
            f.write(np.full(3 * nv, 0.5 * k, "<f4").tobytes())


@pytest.mark.parametrize("vel", [False, True])
def test_trajectory_reader_layout(bcs, tmp_path, vel):
    p = tmp_path / "t.bcstraj"
    _write_fake(p, 5, 4, 3, vel)
    t = bcs.read_trajectory(p)
    assert (t["particles"], t["vein_vertices"], t["types"]) == (5, 4, 2)
    assert t["steps"].tolist() == [0, 10, 20]
    assert t["positions"].shape == (3, 3, 5) and (t["positions"][2] == 2).all()
    assert (t["vein"][1] == 0.5).all()
    assert (t["velocities"] is not None) == vel
    with open(p, "ab") as f:
        f.write(b"xx")
    with pytest.raises(ValueError, match="truncated"):
        bcs.read_trajectory(p)


@pytest.mark.gpu
def test_export_frame_layouts(bcs):
    sc = bcs.Scene(params=_mixed(bcs))
    sim = bcs.Simulation.from_scene(sc)
    sim.step(7)
    info = sim.get_info()
    n, nv = info.particle_count, info.vertex_count
    sentinel = -7.25
    offs = DeviceBuffer(3 * n, sentinel)
    per_type = [DeviceBuffer(6 * info.cells_of_type[t] * info.particles_in_cell[t], sentinel)
                for t in range(info.type_count)]
    vein = DeviceBuffer(6 * nv, sentinel)
    sim.export_frame(offs.ptr.value, [b.ptr.value for b in per_type], vein.ptr.value)
    sim.synchronize()
    p = sim.get_particles()[0].reshape(3, n)
    X = sim.get_vein()[0].reshape(3, nv)
    assert_bits_equal(offs.read().reshape(n, 3), p.T.copy(), "offsets (stride 3)")
    for t in range(info.type_count):
        a = info.particle_starts[t]
        m = info.cells_of_type[t] * info.particles_in_cell[t]
        b = per_type[t].read().reshape(m, 6)
        assert_bits_equal(b[:, :3].copy(), p[:, a: a + m].T.copy(), f"type {t} vertices (stride 6)")
        assert (b[:, 3:] == np.float32(sentinel)).all()
    v = vein.read().reshape(nv, 6)
    assert_bits_equal(v[:, :3].copy(), X.T.copy(), "vein vertices (stride 6)")
    assert (v[:, 3:] == np.float32(sentinel)).all()
    for b in [offs, vein] + per_type:
        b.free()


@pytest.mark.gpu
def test_trajectory_frames_equal_the_oracle(bcs, orc, tmp_path):
    sc = bcs.Scene(params=_mixed(bcs))
    sim = bcs.Simulation.from_scene(sc)
    o = orc.Oracle(sc.physics, sc.types, sc.mesh)
    o.set_particles(sc.positions, sc.velocities, sc.forces)
    path = tmp_path / "run.bcstraj"
    expect = []
    with sim.open_trajectory(path, with_velocities=True) as tr:
        for k in range(4):
            tr.append()
            op, ov, _ = o.get_particles()
            expect.append((op, ov, o.get_vein()[0]))
            sim.step(3)
            o.step(3)
    t = bcs.read_trajectory(path)
    assert t["steps"].tolist() == [0, 3, 6, 9]
    assert t["dt"] == np.float32(sc.physics.dt)
    for k, (op, ov, oX) in enumerate(expect):
        assert_bits_equal(t["positions"][k].reshape(-1), op, f"frame {k} positions")
        assert_bits_equal(t["velocities"][k].reshape(-1), ov, f"frame {k} velocities")
        assert_bits_equal(t["vein"][k].reshape(-1), oX, f"frame {k} vein")


def _mixed(bcs):
    p = bcs.Scene.params(0)
    p.n_cells = 2
    p.cells[0] = bcs.SceneCell(bcs.CELL_DISC42, 40, 2.0, 0.35, 0.5)
    p.cells[1] = bcs.SceneCell(bcs.CELL_ICOSAHEDRON, 60, 2.0, 1.0, 0.5)
    p.seed = 4
    return p
