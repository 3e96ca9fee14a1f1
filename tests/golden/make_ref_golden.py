"""Write the golden fixtures of the reference's compile-time tables.

Runs the reference table probe (oracle/_ref/ref_tables, built by
`make -C oracle ref` from the reference's own headers in place) and stores:
  ref_tables.json       folded blood-cell tables, spring graph, grid bounds,
                        the user cell list (springs + vertices) and physics
  ref_vein_tables.json  sha256 digests and sample rows of the 9-slot vein
                        spring-neighbour tables of the reference vein mesh
Data only: numbers the reference's code produces, not its source.
"""
import hashlib
import json
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
REF = ROOT / "oracle" / "_ref"


def main():
    probe = REF / "ref_tables"
    if not probe.exists():
        sys.exit("oracle/_ref/ref_tables missing: run `make -C oracle ref` where /root/reference exists")
    out = subprocess.run([str(probe), str(REF)], capture_output=True, text=True, check=True).stdout
    tables = json.loads(out)
    (HERE / "ref_tables.json").write_text(json.dumps(tables, indent=1) + "\n")
    nv = tables["vein_vertex_count"]
    ids = np.fromfile(REF / "vein_neighbor_ids.i32", np.int32)
    lens = np.fromfile(REF / "vein_neighbor_lengths.f32", np.float32)
    assert ids.size == 9 * nv and lens.size == 9 * nv
    sample = [0, 1, 100, 7777, nv - 1]
    vein = {
        "vertex_count": nv,
        "triangle_count": tables["vein_triangle_count"],
        "ids_sha256": hashlib.sha256(ids.tobytes()).hexdigest(),
        "lengths_sha256": hashlib.sha256(lens.tobytes()).hexdigest(),
        "mesh_positions_sha256": hashlib.sha256((REF / "vein_positions.f32").read_bytes()).hexdigest(),
        "mesh_indices_sha256": hashlib.sha256((REF / "vein_indices.u32").read_bytes()).hexdigest(),
        "sample_vertices": sample,
        "sample_ids": [[int(ids[s * nv + v]) for s in range(9)] for v in sample],
        "sample_lengths": [[float(lens[s * nv + v]) for s in range(9)] for v in sample],
        "slots_used_histogram": np.bincount((ids.reshape(9, nv) != -1).sum(0), minlength=10).tolist(),
    }
    (HERE / "ref_vein_tables.json").write_text(json.dumps(vein, indent=1) + "\n")
    print("wrote", HERE / "ref_tables.json", HERE / "ref_vein_tables.json")


if __name__ == "__main__":
    main()
