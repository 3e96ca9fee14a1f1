"""Benchmark of the MI355X blood-cell physics step (one JSON line on stdout).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2] [--no-cpu]

Workload (default): config 2 of the plain method -- 200,000 cells (70%
42-particle discs, 30% 12-particle icosahedra; 6.6M particles, FP32) packed
at 35% of the lumen volume in a Y-bifurcating vein scaled 3x (trunk r 180 x
1800, branches r 126 x 1200 at +/-30 deg, ~1.8M triangles), dt = 0.005,
seeded placement.  A step is one full frame (particle grid, vein springs,
cell springs, particle contacts, vein ray collisions, integration, vein-end
teleport), replayed from HIP graphs; all state is resident in HBM during the
timed region.  `value` is particle-steps/s over frames W+1 .. W+K.

Roofline (dominant kernel, HBM-bound): algorithmic bytes per launch / its
average duration measured with HIP events on the library's stream.  The
algorithmic bytes per launch of each kernel are listed in kernel_bytes();
`traffic` comes from the committed rocprofv3 PMC summary of the same
workload when present.

cpu_baseline: the CPU oracle (the C restatement of the reference frame,
OpenMP) on a bounded sample of the same workload, on all of this host's
cores, with a one-thread sample beside it.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s

WORKLOADS = {
    0: "config 0: 500 icosahedron cells (6,000 particles) in a straight cylinder r 40 x 400 (12,800 triangles), dt 0.01",
    1: "config 1: 20,000 icosahedron cells (240,000 particles) in a Y-bifurcating vein (trunk r 60 x 600, branches r 42 x 400 at +/-30 deg), dt 0.005",
    2: "config 2: 200,000 cells (70% 42-particle discs, 30% 12-particle icosahedra) in the Y vein scaled 3x, 35% lattice packing, dt 0.005",
    4: "config 4 geometry: 1,000,000 42-particle discs in the config-2 vein, dt 0.005 (plain method)",
}


def step_bytes(np_, n_occ, nc, nv, nt):
    """Algorithmic HBM bytes of one plain frame: positions, velocities and
    forces read and written once (72 / particle), the sorted key, index and
    grid keys (16 / particle), cell table entries (16 / occupied cell), cell
    centres (24 / blood cell), vein state, springs and triangle data
    (144 / vertex + 16 / triangle):
    88 N_p + 16 N_occ + 24 N_c + 144 N_v + 16 N_t."""
    return 88 * np_ + 16 * n_occ + 24 * nc + 144 * nv + 16 * nt


def kernel_bytes(name, np_, n_occ, nc, nv, nt, n_items=0, map_in_rank=False):
    """Algorithmic HBM bytes of one launch of each frame kernel (every datum
    counted once, as a perfect cache would move it):
      k_key        x read 12, key written 4, previous map-word slot read 4 (the
                   words it names cleared: counted in k_occ)       20 / particle
      k_scatter    key read 4, x, v read 24, (key, index) written 8, the
                   32-byte (x, index | v) record written beside it 68 / particle
      k_rank       (key, index) and record read 40, sorted key, index, inverse
                   written 12, record written to its sorted slot 32
                   + {start, count} per occupied cell 8           84 / particle + 8 / cell
      k_occ        sorted key read 4, map-word slot written 4; count read 8
                   and 2-bit code written (<= 2 bytes) per occupied cell
                                                                   8 / particle + 10 / cell
                   (map_in_rank: with buckets of 2^8 cells or more k_rank
                   writes the map instead -- map-word slot 4 / particle and
                   code 2 / cell, its keys and counts already in registers)
      k_raycand    x, v read 24, first-hit key written 4; work items written
                   16 each (triangle-cell boxes counted in k_vein_prep)
                                                                   28 / particle + 16 / item
      k_ray        work item 16 + x, v 24 per item (triangle data counted in k_vein_prep)
      k_raycand_ray  k_raycand and k_ray in one launch, the work items in LDS:
                   x, v read 24, first-hit key written 4          28 / particle
      k_cells      x, v, F read 36, own key 4, sorted-slot records of the
                   contact candidates read 32, first-hit key read 4, x, v, F
                   written 36; centre written 12 per blood cell; table entry
                   + map code read 10 per occupied cell
                                                                   112 / particle + 12 / blood cell
                                                                   + 10 / cell
      k_vein_prep  vein springs + triangle data: X, V read 24, 9 neighbour ids + lengths 72,
                   Fv read + written 24 per vertex; perm + indices 16, the
                   float4 triangle data written 48 and one box (32) per
                   occupied triangle cell, counted per triangle 12 120 / vertex + 76 / triangle
      k_vein       X, V, Fv read 36, X, V, Fv written 36, offsets 4
                                                                   76 / vertex"""
    table = {
        "k_key": 20 * np_,
        "k_scatter": 68 * np_,
        "k_rank": 84 * np_ + 8 * n_occ + ((4 * np_ + 2 * n_occ) if map_in_rank else 0),
        "k_occ": 8 * np_ + 10 * n_occ,
        "k_raycand": 28 * np_ + 16 * n_items,
        "k_ray": 40 * n_items,
        "k_raycand_ray": 28 * np_,
        "k_cells": 112 * np_ + 12 * nc + 10 * n_occ,
        "k_vein_prep": 120 * nv + 76 * nt,
        "k_vein": 76 * nv,
    }
    return table.get(name)


def host_cores():
    """Cores this process may run on (the affinity mask), and the machine's."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return usable, os.cpu_count() or usable


def _time_oracle(oracle_py, scene, threads, warmup, target_s, max_steps):
    oracle_py.set_threads(threads)
    o = oracle_py.Oracle(scene.physics, scene.types, scene.mesh)
    o.set_particles(scene.positions, scene.velocities, scene.forces)
    if warmup:
        o.step(warmup)
    n = 0
    t0 = time.perf_counter()
    while n < max_steps and (n == 0 or time.perf_counter() - t0 < target_s):
        o.step(1)
        n += 1
    dt = time.perf_counter() - t0
    o.close()
    return n, dt


def cpu_baseline(scene, target_s=12.0, max_steps=400):
    """The oracle on every usable host core (OpenMP), then on one thread, each
    on a bounded sample: whole frames of the same workload from its initial
    state until about target_s seconds have passed (at least one frame)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_py

    usable, machine = host_cores()
    n, dt = _time_oracle(oracle_py, scene, usable, 1, target_s, max_steps)
    n1, dt1 = _time_oracle(oracle_py, scene, 1, 0, target_s, max_steps)
    return {
        "value": scene.particle_count * n / dt,
        "unit": "particle-steps/s",
        "cores": usable,
        "kind": "port",
        "sample": (f"frames 2..{n + 1} of the same workload from its initial state (frame 1 untimed) on {usable} "
                   f"OpenMP threads, one per usable core ({machine} logical CPUs on the host), {dt:.1f} s"),
        "single_thread": {
            "value": scene.particle_count * n1 / dt1,
            "unit": "particle-steps/s",
            "cores": 1,
            "sample": f"frames 1..{n1} of the same workload from its initial state on one thread, {dt1:.1f} s",
        },
    }


def pmc_traffic(kernel, config_id):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE runs scaled by the copy-kernel calibration)."""
    p = ROOT / "profiles" / f"frame_kernels_config{config_id}_pmc.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
        return float(d["kernels"][kernel]["hbm_bytes_per_launch"])
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--config", type=int, default=2, choices=sorted(WORKLOADS))
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--no-graph", action="store_true",
                    help="enqueue the fused frame kernels directly instead of replaying the HIP graph")
    ap.add_argument("--calibrate-bytes", type=int, default=0,
                    help="also run the plain copy kernel (k_copy) over this many bytes, for HBM counter calibration")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on exactly one GPU (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    bcs = importlib.import_module("simulation-server_amd")
    if bcs.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    scene = bcs.Scene(args.config, 1.0, seed=args.seed)
    sim = bcs.Simulation.from_scene(scene, device=0)
    if args.no_graph:
        sim.use_graph(False)
    info = sim.info
    np_, nc, nv, nt = info.particle_count, info.blood_cell_count, info.vertex_count, info.triangle_count

    # warm-up (also builds and instantiates the frame graph)
    sim.step(max(args.warmup, 1))
    # timed region: K graph-replayed frames between two synchronisations
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(args.steps, sync=False)
    sim.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = np_ * args.steps / wall

    flags = sim.error_flags()
    frame_stats = sim.frame_stats()
    _, perm, occ_keys, _, _ = sim.get_grid(0)
    n_occ = len(occ_keys)
    # per-kernel durations (HIP events on the library stream, un-graphed frames)
    kern = sim.profile_kernels(args.profile_steps)
    dominant = max(kern, key=kern.get)
    dom_ms = kern[dominant]
    map_in_rank = kern.get("k_occ", 0.0) == 0.0
    dom_bytes = kernel_bytes(dominant, np_, n_occ, nc, nv, nt, frame_stats.get("ray_work_items", 0), map_in_rank)
    achieved = dom_bytes / (dom_ms * 1e-3) if dom_bytes else None
    roof = {
        "kernel": dominant,
        "bound": "hbm",
        "achieved": achieved / 1e9 if achieved else None,
        "peak": HBM_PEAK / 1e9,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK if achieved else None,
        "traffic": pmc_traffic(dominant, args.config),
        "algorithmic_bytes_per_launch": dom_bytes,
        "avg_launch_ms": dom_ms,
    }
    sb = step_bytes(np_, n_occ, nc, nv, nt)
    step_roof = {
        "bytes_per_step": sb,
        "achieved_GBps": sb / (ms_per_step * 1e-3) / 1e9,
        "frac": sb / (ms_per_step * 1e-3) / HBM_PEAK,
    }
    out = {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "profile_steps": args.profile_steps,
        "map_in_rank": map_in_rank,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator, seed %d)" % args.seed,
        "config": {
            "workload": WORKLOADS[args.config],
            "frames_timed": f"{args.warmup + 1}..{args.warmup + args.steps}",
            "particles": np_,
            "cells": nc,
            "vein_vertices": nv,
            "vein_triangles": nt,
            "occupied_grid_cells": n_occ,
            "particle_grid": list(info.particle_cells),
            "triangle_grid": list(info.triangle_cells),
            "dt": scene.physics.dt,
            "frame_launch": "direct" if args.no_graph else "hip-graph",
        },
        "roofline": roof,
        "step_roofline": step_roof,
        "kernel_ms": kern,
        "dominant_kernel": dominant,
        "error_flags": flags,
        "frame_stats": frame_stats,
    }
    if args.calibrate_bytes:
        cms = bcs.calibration_copy(args.calibrate_bytes, 3)
        out["calibration_copy"] = {"bytes_read": args.calibrate_bytes, "bytes_written": args.calibrate_bytes,
                                   "ms": cms, "GBps": 2 * args.calibrate_bytes / (cms * 1e-3) / 1e9}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
