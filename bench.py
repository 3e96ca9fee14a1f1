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

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s

WORKLOADS = {
    0: "config 0: 500 icosahedron cells (6,000 particles) in a straight cylinder r 40 x 400 (12,800 triangles), dt 0.01",
    1: "config 1: 20,000 icosahedron cells (240,000 particles) in a Y-bifurcating vein (trunk r 60 x 600, branches r 42 x 400 at +/-30 deg), dt 0.005",
    2: "config 2: 200,000 cells (70% 42-particle discs, 30% 12-particle icosahedra) in the Y vein scaled 3x, 35% lattice packing, dt 0.005",
    3: "config 3: batched ensemble of 64 straight veins (radius U(20, 80), length 10 radius, 500-50,000 cells "
       "log-uniform, 70% 42-particle discs / 30% 12-particle icosahedra), dt 0.005, one launch per frame kernel",
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
                   (map_in_rank: with buckets of 2^7 cells or more k_rank
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
      (the default frame runs k_cells as the three kernels below)
      k_springs    x, v, F read 36, sorted slot read 4, F after the springs
                   written to its slot 16; centre written 12 per blood cell
                                                                   56 / particle + 12 / blood cell
      k_contacts   own sorted-slot record read 32 (each record, as the
                   candidate of others, counted once), sorted key 4, F read
                   and written 16 + 16; table entry + map code read 10 per
                   occupied cell                                   68 / particle + 10 / cell
      k_respond    x, v read 24, sorted slot 4, F from its slot 16, first-hit
                   key 4, x, v, F written 36                       84 / particle
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
        "k_springs": 56 * np_ + 12 * nc,
        "k_contacts": 68 * np_ + 10 * n_occ,
        "k_respond": 84 * np_,
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


def bench_ensemble(bcs, args):
    """Config 3: the 64 veins as one batched ensemble, autotuned, then timed;
    beside it the same veins stepped one after another through their own
    launches (bcs_sim_step per vein), the way separate processes of the
    reference would run them on one GPU."""
    n_veins = 64
    sims = []
    for k in range(n_veins):
        sc = bcs.Scene(params=bcs.ensemble_scene_params(n_veins, k, seed=args.seed))
        sims.append(bcs.Simulation.from_scene(sc, device=0))
    np_total = sum(s.np for s in sims)
    ens = bcs.Ensemble(sims)
    tune = ens.autotune(frames_per_trial=args.tune_frames)
    adv = len(tune["trials"]) * (1 + args.tune_frames)  # frames the trials advanced every vein
    # the veins crowd at their inlets like config 2, where the contact search by
    # sorted slot is the faster form: the ensemble and the solo runs both use it
    chosen = {k: v for k, v in ens.get_tuning().items() if k != "contact_mode"}
    ens.set_tuning(**chosen, contact_mode=bcs.CONTACT_SLOTS)
    for s in sims:
        s.set_contact_mode(bcs.CONTACT_SLOTS)
    ens.step(max(args.warmup, 1))
    ens.synchronize()
    t0 = time.perf_counter()
    ens.step(args.steps, sync=False)
    ens.synchronize()
    wall = time.perf_counter() - t0

    # Launch batching against one step call per vein, in the same regime: K
    # ensemble frames, K frames of each vein through its own bcs_sim_step
    # (veins one after another), K ensemble frames; the solo time is compared
    # with the mean of the two ensemble windows around it.
    K = max(1, min(args.steps, 100))

    def ens_window():
        ens.synchronize()
        t = time.perf_counter()
        ens.step(K, sync=False)
        ens.synchronize()
        return (time.perf_counter() - t) / K

    e1 = ens_window()
    for s in sims:
        s.step(1)
    for s in sims:
        s.synchronize()
    t1 = time.perf_counter()
    for s in sims:
        s.step(K, sync=False)
    for s in sims:
        s.synchronize()
    solo = (time.perf_counter() - t1) / K
    for s in sims:
        s.step(1)
    e2 = ens_window()
    ms_per_step = 1e3 * wall / args.steps
    flags = 0
    for s in sims:
        flags |= s.error_flags()
    return {
        "metric": "particle-steps/s",
        "value": np_total * args.steps / wall,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator, seed %d)" % args.seed,
        "config": {
            "workload": WORKLOADS[3],
            "frames_timed": "%d..%d (the %d autotuning trials ran frames 1..%d)" % (
                adv + args.warmup + 1, adv + args.warmup + args.steps, len(tune["trials"]), adv),
            "veins": n_veins,
            "particles": np_total,
            "particles_per_vein_min_max": [min(s.np for s in sims), max(s.np for s in sims)],
            "frame_launch": "hip-graph, batched kernels",
        },
        "autotune": tune,
        "solo_launches": {
            "frames": K,
            "ms_per_frame_all_veins": 1e3 * solo,
            "ensemble_ms_per_frame_before_after": [1e3 * e1, 1e3 * e2],
            "particle_steps_per_s": np_total / solo,
            "note": "the same contact mode; each vein stepped by its own bcs_sim_step call (its own graphs), "
                    "veins one after another, between two ensemble windows of the same length",
        },
        "ensemble_speedup_over_solo_launches": solo / (0.5 * (e1 + e2)),
        "error_flags": flags,
    }


SURR_FLOP_PER_CELL = 2 * (294 * 256 + 256 * 256 + 256 * 256 + 256 * 126)  # 477,184
MFMA_BF16_PEAK = 2.5e15  # MI355X dense BF16, FLOP/s


def bench_surrogate(bcs, args):
    """Config 4: the learned spring surrogate.  Training rows come from the
    plain method on config 2 (the 140,000 disc cells, sampled twice after 20
    frames: 200,000 rows kept); the network is trained on them, then deployed
    on the config-4 geometry (1,000,000 discs, 42M particles) with the
    per-cell check against the plain springs at tolerance 1e-3 of the cell
    radius, for args.surr_frames frames after the plain frames."""
    t_all = time.perf_counter()
    src = bcs.Simulation.from_scene(bcs.Scene(2, 1.0, seed=args.seed), device=0)
    if args.no_graph:
        src.use_graph(False)
    src.step(20)
    feats, targs = [], []
    while sum(len(f) for f in feats) < 200000:
        f, t = src.surrogate_sample(capacity=200000)
        feats.append(f)
        targs.append(t)
    x = np.concatenate(feats)[:200000]
    y = np.concatenate(targs)[:200000]
    src.close()
    net = bcs.Surrogate(seed=args.seed)
    scales = (float(np.abs(x[:, :126]).std()), float(max(np.abs(x[:, 126:252]).std(), 1e-6)),
              float(max(np.abs(y).std(), 1e-12)))
    net.set_scales(*scales)
    t0 = time.perf_counter()
    loss = net.train(x, y, epochs=args.surr_epochs, batch=4096, lr=1e-3, seed=args.seed)
    train_s = time.perf_counter() - t0
    pred = net.predict(x[:20000])
    rel_err = float(np.abs(pred - y[:20000]).max() / max(np.abs(y[:20000]).max(), 1e-30))

    scene = bcs.Scene(4, 1.0, seed=args.seed)
    ncell = scene.types[0].count
    radius = 2.0
    tol = 1e-3 * radius
    out = {"training": {"rows": int(len(x)), "epochs": args.surr_epochs, "seconds": train_s,
                        "loss_first_last": [float(loss[0]), float(loss[-1])], "scales": scales,
                        "max_abs_error_over_max_target_on_20k_rows": rel_err}}
    runs = {}
    for mode in ("plain", "surrogate"):
        sim = bcs.Simulation.from_scene(scene, device=0)
        if args.no_graph:
            sim.use_graph(False)
        if mode == "surrogate":
            sim.attach_surrogate(net, tol)
        sim.step(max(args.warmup, 1))
        if mode == "surrogate":
            sim.surrogate_stats(reset=True)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.step(args.surr_frames, sync=False)
        sim.synchronize()
        wall = time.perf_counter() - t0
        r = {"frames": args.surr_frames, "ms_per_frame": 1e3 * wall / args.surr_frames,
             "cell_steps_per_s": ncell * args.surr_frames / wall,
             "particle_steps_per_s": scene.particle_count * args.surr_frames / wall}
        if mode == "surrogate":
            st = sim.surrogate_stats()
            r.update(st)
            r["tolerance"] = tol
            r["accepted_fraction"] = st["accepted"] / max(st["accepted"] + st["fallback"], 1)
        runs[mode] = r
        sim.close()
    ms = net.time_inference(ncell, reps=5)
    inf = {"cells": ncell, "ms_per_pass": ms, "cell_steps_per_s": ncell / (ms * 1e-3),
           "tflops": SURR_FLOP_PER_CELL * ncell / (ms * 1e-3) / 1e12,
           "mfma_frac": SURR_FLOP_PER_CELL * ncell / (ms * 1e-3) / MFMA_BF16_PEAK}
    out.update({
        "metric": "cell-steps/s (42-particle cells, surrogate frames vs plain frames)",
        "steps": args.surr_frames,
        "warmup": max(args.warmup, 1),
        "value": runs["surrogate"]["cell_steps_per_s"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "dtype": "bf16 network (fp32 accumulate), f32 step",
        "data": "synthetic (seeded scene generator, seed %d)" % args.seed,
        "config": {"workload": WORKLOADS[4] + ", learned springs", "cells": ncell,
                   "particles": scene.particle_count},
        "plain": runs["plain"],
        "surrogate": runs["surrogate"],
        "speedup_vs_plain": runs["surrogate"]["cell_steps_per_s"] / runs["plain"]["cell_steps_per_s"],
        "inference": inf,
        "roofline": {"bound": "mfma", "achieved": inf["tflops"], "peak": MFMA_BF16_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": inf["mfma_frac"], "traffic": None,
                     "flop_per_cell_step": SURR_FLOP_PER_CELL},
        "seconds_total": time.perf_counter() - t_all,
    })
    return out


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
    ap.add_argument("--contact-mode", choices=["auto", "slots", "cells"], default="auto",
                    help="where frames run the contact search (results are identical)")
    ap.add_argument("--direct-timed", action="store_true",
                    help="warm-up frames from the HIP graphs, timed frames launched directly (a counter "
                         "profiler that does not follow graph replays then sees the timed frames only)")
    ap.add_argument("--surrogate", action="store_true",
                    help="config 4 with the learned spring surrogate (trained on config-2 rows first)")
    ap.add_argument("--surr-epochs", type=int, default=8)
    ap.add_argument("--surr-frames", type=int, default=200)
    ap.add_argument("--tune-frames", type=int, default=5,
                    help="config 3: frames timed per autotuning trial")
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
    if args.config == 3:
        print(json.dumps(bench_ensemble(bcs, args)))
        return
    if args.surrogate:
        print(json.dumps(bench_surrogate(bcs, args)))
        return
    scene = bcs.Scene(args.config, 1.0, seed=args.seed)
    sim = bcs.Simulation.from_scene(scene, device=0)
    if args.no_graph:
        sim.use_graph(False)
    sim.set_contact_mode({"auto": bcs.CONTACT_AUTO, "slots": bcs.CONTACT_SLOTS, "cells": bcs.CONTACT_CELLS}[args.contact_mode])
    info = sim.info
    np_, nc, nv, nt = info.particle_count, info.blood_cell_count, info.vertex_count, info.triangle_count

    # warm-up (also builds and instantiates the frame graph)
    sim.step(max(args.warmup, 1))
    if args.direct_timed:
        sim.use_graph(False)
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
            "frame_launch": "direct" if args.no_graph else ("hip-graph warm-up, direct timed frames"
                                                             if args.direct_timed else "hip-graph"),
            "contact_mode": args.contact_mode,
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
