"""Summaries of rocprofv3 runs of bench.py, as the text tables kept under
profiles/.

  python simulation-server_amd/prof_summary.py trace BENCH_JSON TRACE_DIR OUT.md
      per-kernel dispatch durations from a `rocprofv3 --kernel-trace --stats
      -f csv` run, next to the HIP-event means bench.py measured
  python simulation-server_amd/prof_summary.py pmc BENCH_JSON FETCH_DIR WRITE_DIR OUT.md OUT.json
      per-kernel FETCH_SIZE and WRITE_SIZE (KiB per dispatch) from two
      `rocprofv3 --pmc` runs, scaled by what the same counters read for the
      library's plain copy kernel k_copy over a known number of bytes, and
      compared with each kernel's algorithmic bytes per launch

BENCH_JSON is the JSON line the profiled bench.py run printed (its workload
sizes give the algorithmic bytes).  Dispatches of the W warm-up frames are
left out of the means: for a kernel launched on every frame, the first
calls * W / frames dispatches.
"""
from __future__ import annotations

import csv
import json
import sys
from collections import defaultdict
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def _short(name: str) -> str:
    """Kernel name without return type, namespace and argument list
    (template arguments kept): 'void bcs::k_cells<true>(StepArgs)' ->
    'k_cells<true>'."""
    name = name.strip()
    if name.startswith("void "):
        name = name[5:]
    name = name.replace("(anonymous namespace)::", "")
    if name.startswith("bcs::"):
        name = name[5:]
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:i].strip()
    return name


def _rows(directory: Path, suffix: str):
    files = sorted(directory.rglob(f"*{suffix}"))
    if not files:
        raise SystemExit(f"{directory}: no *{suffix}")
    out = []
    for f in files:
        with open(f, newline="") as fh:
            out.extend(csv.DictReader(fh))
    return out


def _frames(bench: dict) -> tuple[int, int]:
    warm = int(bench["warmup"])
    total = warm + int(bench["steps"]) + int(bench.get("profile_steps", 0))
    return warm, total


def _skip(calls: int, warm: int, total: int) -> int:
    return (calls * warm) // total if calls >= total else 0


def _algorithmic(bench: dict) -> dict:
    import bench as B

    c = bench["config"]
    n_items = bench.get("frame_stats", {}).get("ray_work_items", 0)
    names = set(bench.get("kernel_ms", {})) | {"k_key", "k_scatter", "k_rank", "k_occ", "k_raycand", "k_ray",
                                                "k_raycand_ray", "k_cells", "k_vein_prep", "k_vein"}
    out = {}
    for n in names:
        b = B.kernel_bytes(n, c["particles"], c["occupied_grid_cells"], c["cells"], c["vein_vertices"],
                           c["vein_triangles"], n_items, map_in_rank=bench.get("map_in_rank", False))
        if b:
            out[n] = b
    return out


def _base(kname: str) -> str:
    """k_cells<true> -> k_cells, k_vein<7> -> k_vein."""
    return kname.split("<")[0]


def trace(bench_json: str, trace_dir: str, out_md: str, command: str = "") -> None:
    bench = json.loads(Path(bench_json).read_text().strip().splitlines()[-1])
    warm, total = _frames(bench)
    per = defaultdict(list)
    for r in _rows(Path(trace_dir), "kernel_trace.csv"):
        per[_short(r["Kernel_Name"])].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    ev = bench.get("kernel_ms", {})
    lines = [
        f"# Frame kernels: dispatch durations ({bench['config']['workload']})",
        "",
        "rocprofv3 kernel trace of bench.py with directly launched frames (the same kernels the HIP graphs replay):",
        "",
        "```",
        command or "(command not recorded)",
        "```",
        "",
        f"Frames: {warm} warm-up, {bench['steps']} timed, then {bench.get('profile_steps', 0)} event-timed by "
        f"bench.py itself. Means leave out the warm-up frames' dispatches. 'events' is bench.py's HIP-event mean "
        f"per launch on the library stream (it includes the launch gap, a few us).",
        "",
        "| kernel | dispatches | mean us | min us | max us | total ms | events mean us |",
        "|---|---:|---:|---:|---:|---:|---:|",
    ]
    rows = []
    for k, d in per.items():
        d = d[_skip(len(d), warm, total):]
        if not d:
            continue
        rows.append((sum(d), k, len(d), sum(d) / len(d), min(d), max(d)))
    for tot, k, n, mean, lo, hi in sorted(rows, reverse=True):
        e = ev.get(k, ev.get(_base(k)))
        lines.append(f"| {k} | {n} | {mean / 1e3:.1f} | {lo / 1e3:.1f} | {hi / 1e3:.1f} | {tot / 1e6:.2f} | "
                     f"{'' if e is None else f'{e * 1e3:.1f}'} |")
    all_ns = sum(r[0] for r in rows)
    lines += ["", f"Sum of kernel time over the counted dispatches: {all_ns / 1e6:.2f} ms."]
    Path(out_md).write_text("\n".join(lines) + "\n")


def pmc(bench_json: str, fetch_dir: str, write_dir: str, out_md: str, out_json: str, commands: str = "") -> None:
    bench = json.loads(Path(bench_json).read_text().strip().splitlines()[-1])
    warm, total = _frames(bench)
    vals = {}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        per = defaultdict(list)
        for r in _rows(Path(d), "counter_collection.csv"):
            if r["Counter_Name"] != counter:
                continue
            per[_short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        for k, v in per.items():
            v = v[_skip(len(v), warm, total):] or v
            vals.setdefault(k, {})[counter] = sum(v) / len(v)
    cal = bench.get("calibration_copy")
    if not cal or "k_copy" not in vals:
        raise SystemExit("the PMC runs need --calibrate-bytes (k_copy) for the scale factors")
    kib = cal["bytes_read"] / 1024.0
    f_scale = kib / vals["k_copy"]["FETCH_SIZE"]
    w_scale = cal["bytes_written"] / 1024.0 / vals["k_copy"]["WRITE_SIZE"]
    alg = _algorithmic(bench)
    rows = []
    for k, v in vals.items():
        f, w = v.get("FETCH_SIZE", 0.0), v.get("WRITE_SIZE", 0.0)
        hbm = (f * f_scale + w * w_scale) * 1024.0
        a = alg.get(_base(k)) if k != "k_copy" else None
        rows.append((f, k, w, hbm, a))
    rows.sort(reverse=True)
    lines = [
        f"# Frame kernels: HBM traffic from rocprofv3 PMC counters ({bench['config']['workload']})",
        "",
        "Two runs, one counter each:",
        "",
        "```",
        commands or "(commands not recorded)",
        "```",
        "",
        f"Calibration: the plain copy kernel `k_copy` over {cal['bytes_read'] / 2**20:.0f} MiB reads FETCH_SIZE = "
        f"{vals['k_copy']['FETCH_SIZE']:.0f} KiB and WRITE_SIZE = {vals['k_copy']['WRITE_SIZE']:.0f} KiB per "
        f"dispatch for {kib:.0f} KiB moved each way, so FETCH_SIZE is scaled by {f_scale:.3f} and WRITE_SIZE by "
        f"{w_scale:.3f}.",
        "",
        (f"Per-dispatch means over the directly launched frames {bench['config'].get('frames_timed', '')} "
         f"and the {bench.get('profile_steps', 0)} event-timed ones (the {warm} warm-up frames ran from HIP graphs, "
         f"which the counter runs do not see)"
         if "direct timed" in bench["config"].get("frame_launch", "") else
         f"Per-dispatch means ({warm} warm-up frames left out).") +
        " Algorithmic bytes per launch: bench.py kernel_bytes() at this workload's sizes.",
        "",
        "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes / launch (scaled) | algorithmic bytes / launch | traffic / algorithmic |",
        "|---|---:|---:|---:|---:|---:|",
    ]
    js = {"workload": bench["config"]["workload"], "fetch_scale": f_scale, "write_scale": w_scale, "kernels": {}}
    for f, k, w, hbm, a in rows:
        ratio = f"{hbm / a:.2f}" if a else ""
        lines.append(f"| {k} | {f:.0f} | {w:.0f} | {hbm / 1e6:.2f} MB | {'' if not a else f'{a / 1e6:.2f} MB'} | "
                     f"{ratio} |")
        js["kernels"][_base(k)] = {
            "fetch_kib": f, "write_kib": w, "hbm_bytes_per_launch": hbm, "algorithmic_bytes_per_launch": a}
    Path(out_md).write_text("\n".join(lines) + "\n")
    Path(out_json).write_text(json.dumps(js, indent=1) + "\n")


if __name__ == "__main__":
    mode = sys.argv[1] if len(sys.argv) > 1 else ""
    if mode == "trace" and len(sys.argv) >= 5:
        trace(*sys.argv[2:5], command=" ".join(sys.argv[5:]))
    elif mode == "pmc" and len(sys.argv) >= 7:
        pmc(*sys.argv[2:7], commands=" ".join(sys.argv[7:]).replace(" ;; ", "\n"))
    else:
        print(__doc__)
        sys.exit(2)
