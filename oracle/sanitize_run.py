"""The oracle (test infrastructure) under AddressSanitizer and
UndefinedBehaviorSanitizer: every hand-computed pin of
tests/test_oracle_pins.py, then 30 frames of a config-0 scene, of the mixed
scene with derived radii and the vein end on, 10 frames of the reference's
default types and 10 frames in double precision, all on
build/liboracle_asan.so.  Run from the repository root after
`make -C oracle asan`, with the sanitizer runtime preloaded:

  LD_PRELOAD=$(gcc -print-file-name=libasan.so) ASAN_OPTIONS=detect_leaks=0 \\
      python oracle/sanitize_run.py

A finding aborts the process with the sanitizer's report."""
import importlib
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402
import oracle_py  # noqa: E402

oracle_py.LIB_PATH = ROOT / "oracle" / "build" / "liboracle_asan.so"
bcs = importlib.import_module("simulation-server_amd")
import test_gpu_quirks as quirks  # noqa: E402
import test_oracle_pins as pins  # noqa: E402

for name in sorted(n for n in dir(pins) if n.startswith("test_")):
    fn = getattr(pins, name)
    args = [bcs, oracle_py][: fn.__code__.co_argcount]
    fn(*args)
    print("ok", name, flush=True)
sc = bcs.Scene(0, 0.5, seed=1)
o = oracle_py.Oracle(sc.physics, sc.types, sc.mesh)
o.set_particles(sc.positions, sc.velocities, sc.forces)
o.step(30)
o.contacts()
print("ok config-0 scene, 30 frames", flush=True)
sc = quirks.mixed_derived_scene(bcs)
ph = sc.physics
ph.use_blood_flow = 1
o = oracle_py.Oracle(ph, sc.types, sc.mesh)
o.set_particles(sc.positions, sc.velocities, sc.forces)
o.step(30)
print("ok mixed scene with derived radii, 30 frames", flush=True)
ph2, types, mesh, pos, vel = quirks.reference_default_scene(bcs)
o = oracle_py.Oracle(ph2, types, mesh)
o.set_particles(pos, vel, np.zeros_like(pos))
o.step(10)
print("ok reference default types, 10 frames", flush=True)
o = oracle_py.Oracle(sc.physics, sc.types, sc.mesh, "f64")
o.set_particles(sc.positions, sc.velocities, sc.forces)
o.step(10)
print("ok double precision, 10 frames", flush=True)
