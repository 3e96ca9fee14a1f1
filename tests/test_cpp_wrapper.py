"""The C++ drop-in header (include/bcs_sim.hpp): the reference's frame loop
(src/main.cu:158-208) compiled against it.  On the CPU the program builds and
reports the missing device through bcs::Error; on the GPU its frames -- the
four reference calls per frame, or the fused step -- equal the oracle's."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import assert_bits_equal

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def loop_exe(tmp_path_factory, bcs):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("cpp") / "reference_loop"
    lib = bcs.LIB_PATH.parent
    cmd = [gxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}",
           str(ROOT / "tests" / "cpp" / "reference_loop.cpp"), f"-L{lib}", "-lbcs", f"-Wl,-rpath,{lib}", "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_reference_loop_builds_and_runs_or_reports_no_device(loop_exe, tmp_path):
    r = subprocess.run([str(loop_exe), "0.05", "2", str(tmp_path / "o.bin")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode in (0, 3), r.stderr
    assert "NoGrid: unsupported" in r.stdout  # refused before any device work
    if r.returncode == 3:
        assert "no HIP device" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["phases", "fused"])
def test_reference_loop_matches_oracle(loop_exe, tmp_path, bcs, orc, mode):
    frames = 12
    out = tmp_path / "o.bin"
    args = [str(loop_exe), "0.2", str(frames), str(out)] + (["fused"] if mode == "fused" else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "NoGrid: unsupported" in r.stdout and "device arrays set" in r.stdout, r.stdout
    data = np.fromfile(out, dtype=np.float32)
    sc = bcs.Scene(0, 0.2, seed=1)
    o = orc.Oracle(sc.physics, sc.types, sc.mesh)
    o.set_particles(sc.positions, sc.velocities, sc.forces)
    o.calculate_grid(0)
    keys, perm, *_ = o.get_grid(0)
    key0 = int(keys[np.flatnonzero(perm == 0)[0]])
    assert f"cell id of particle 0 {key0}\n" in r.stdout, r.stdout
    assert f"vertex of triangle 5 corner 2 {int(sc.mesh.indices[17])}\n" in r.stdout
    o.step(frames)
    op, ov, of = o.get_particles()
    n3 = op.size
    assert data.size == 3 * n3
    assert_bits_equal(data[:n3], op, f"C++ loop ({mode}) positions")
    assert_bits_equal(data[n3:2 * n3], ov, f"C++ loop ({mode}) velocities")
    assert_bits_equal(data[2 * n3:], of, f"C++ loop ({mode}) forces")
