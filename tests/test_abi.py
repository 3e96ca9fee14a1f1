"""The C-ABI library: it loads without a GPU, exports every function that
include/*.h declares, and its host-side entry points and error paths behave
as documented (no compute calls -- there is no GPU here)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(bcs_[a-z0-9_]+)\s*\(", re.M)


def declared_functions():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names.update(DECL.findall(text))
    return sorted(names)


def test_every_declared_function_is_exported(bcs):
    lib = C.CDLL(str(bcs.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 30, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared in include/*.h but not exported: {missing}"


def test_no_device_is_reported_not_fatal(bcs):
    lib = bcs.load_library()
    n = C.c_int32(-1)
    assert lib.bcs_device_count(C.byref(n)) == bcs.BCS_OK
    if n.value > 0:
        return  # a GPU is present: nothing to check here
    sc = bcs.Scene(0, 0.05, seed=3)
    types = bcs.cell_type_array(sc.types)
    meshc = sc.mesh.c()
    h = C.c_void_p()
    rc = lib.bcs_sim_create(C.byref(sc.physics), types, len(sc.types), C.byref(meshc), 0, C.byref(h))
    assert rc == bcs.BCS_ERR_NO_DEVICE
    assert b"no HIP device" in lib.bcs_last_error()
    assert not h.value


def test_null_and_invalid_arguments(bcs):
    lib = bcs.load_library()
    assert lib.bcs_default_physics(None) == bcs.BCS_ERR_INVALID_ARGUMENT
    assert lib.bcs_sim_step(None, 1) == bcs.BCS_ERR_INVALID_ARGUMENT
    assert lib.bcs_sim_destroy(None) in (bcs.BCS_OK, bcs.BCS_ERR_INVALID_ARGUMENT)
    assert lib.bcs_sim_use_graph(None, 0) == bcs.BCS_ERR_INVALID_ARGUMENT
    assert lib.bcs_sim_set_ray_mode(None, 0) == bcs.BCS_ERR_INVALID_ARGUMENT
    assert lib.bcs_sim_set_bucket_shift(None, 8) == bcs.BCS_ERR_INVALID_ARGUMENT
    ms = C.c_float()
    assert lib.bcs_calibration_copy(C.c_int64(8), 1, C.byref(ms)) == bcs.BCS_ERR_INVALID_ARGUMENT
    assert lib.bcs_scene_preset(99, C.c_double(1.0), C.c_uint64(0), C.byref(C.c_void_p())) == bcs.BCS_ERR_INVALID_ARGUMENT


def test_host_tables_need_no_gpu(bcs):
    sc = bcs.Scene(0, 0.02, seed=0)
    lib = bcs.load_library()
    meshc = sc.mesh.c()
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    assert lib.bcs_grid_bounds(C.byref(meshc), C.byref(sc.physics), bcs.fptr(lo), bcs.fptr(hi)) == bcs.BCS_OK
    assert np.all(hi > lo)
    nv = sc.mesh.positions.size // 3
    ids = np.zeros(9 * nv, np.int32)
    lens = np.zeros(9 * nv, np.float32)
    assert lib.bcs_vein_neighbors(C.byref(meshc), bcs.iptr(ids), bcs.fptr(lens)) == bcs.BCS_OK
    assert (ids >= -1).all() and (ids < nv).all()
