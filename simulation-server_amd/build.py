"""Build libbcs.so (HIP for gfx950 + host C++) in-tree.

Every translation unit is compiled with IEEE single-precision semantics:
no FP contraction (so device results round like the CPU oracle, op by op),
correctly rounded division and square root, denormals preserved, and no packed
FP32 instructions.  Output: simulation-server_amd/lib/libbcs.so
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
OUT = HERE / "lib"
OBJ = HERE / "build" / "obj"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BCS_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-ffp-contract=off", "-fno-fast-math",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

SOURCES = ["host_tables.cpp", "scenes.cpp", "refconfig.cpp", "grid.hip", "step.hip", "sim.hip", "surrogate.hip"]


def _compile(src: Path, obj: Path) -> tuple[Path, str]:
    cmd = [HIPCC, *COMMON]
    if src.suffix == ".hip":
        cmd += ["-x", "hip", *DEVICE]
    else:
        cmd += ["-x", "c++"]
    cmd += ["-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj, res.stderr


def build(verbose: bool = False) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    OUT.mkdir(parents=True, exist_ok=True)
    srcs = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(len(srcs), max(1, (os.cpu_count() or 2)))) as ex:
        futs = [ex.submit(_compile, s, OBJ / (s.name + ".o")) for s in srcs]
        for f in futs:
            obj, err = f.result()
            objs.append(obj)
            if verbose and err.strip():
                print(err, file=sys.stderr)
    lib = OUT / "libbcs.so"
    tmp = OUT / "libbcs.so.tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, lib)
    return lib


def build_oracle() -> Path:
    """CPU oracle (test infrastructure) and, where the reference tree is
    present, the reference's own table probe (oracle/_ref)."""
    oracle = ROOT / "oracle"
    res = subprocess.run(["make", "-C", str(oracle), "all"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{res.stdout}\n{res.stderr}")
    if Path("/root/reference/src/meta_factory/blood_cell_factory.hpp").exists():
        res = subprocess.run(["make", "-C", str(oracle), "ref"], capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"reference probe build failed\n{res.stdout}\n{res.stderr}")
        probe = oracle / "_ref" / "ref_tables"
        with open(oracle / "_ref" / "ref_tables.json", "w") as fh:
            subprocess.run([str(probe), str(oracle / "_ref")], stdout=fh, check=True)
    return oracle / "build" / "liboracle.so"


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_oracle())
