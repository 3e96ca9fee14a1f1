"""ctypes binding of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker / CPU baseline.  The
product path (libbcs) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import importlib
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
LIB_PATH = HERE / "build" / "liboracle.so"

if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
bcs = importlib.import_module("simulation-server_amd")  # ABI struct definitions only

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(str(LIB_PATH))
        P = C.POINTER
        vp = C.c_void_p
        for pre, R in (("orc_f32_", C.c_float), ("orc_f64_", C.c_double)):
            rp = P(R)
            getattr(L, pre + "create").argtypes = [P(bcs.Physics), P(bcs.CellType), C.c_int32, P(bcs.VeinMesh)]
            getattr(L, pre + "create").restype = vp
            getattr(L, pre + "destroy").argtypes = [vp]
            for n in ("set_particles", "get_particles", "set_vein", "get_vein"):
                getattr(L, pre + n).argtypes = [vp, rp, rp, rp]
            getattr(L, pre + "get_centers").argtypes = [vp, rp]
            getattr(L, pre + "get_triangle_centers").argtypes = [vp, rp]
            getattr(L, pre + "tables").argtypes = [vp, rp, rp, rp]
            getattr(L, pre + "calculate_grid").argtypes = [vp, C.c_int]
            getattr(L, pre + "get_grid").argtypes = [vp, C.c_int, P(C.c_uint32), P(C.c_int32), P(C.c_uint32),
                                                     P(C.c_int32), P(C.c_int32)]
            getattr(L, pre + "get_grid").restype = C.c_int32
            getattr(L, pre + "calculate_next_frame").argtypes = [vp]
            getattr(L, pre + "propagate_all").argtypes = [vp]
            getattr(L, pre + "step").argtypes = [vp, C.c_int32]
            getattr(L, pre + "contacts").argtypes = [vp, P(C.c_int32), P(C.c_int32), C.c_int64]
            getattr(L, pre + "contacts").restype = C.c_int64
            getattr(L, pre + "flags").argtypes = [vp, C.c_int]
            getattr(L, pre + "flags").restype = C.c_uint32
            getattr(L, pre + "set_step_index").argtypes = [vp, C.c_uint64]
            getattr(L, pre + "update_triangle_centers").argtypes = [vp]
            getattr(L, pre + "particle_count").argtypes = [vp]
            getattr(L, pre + "particle_count").restype = C.c_int32
        L.orc_vein_neighbors.argtypes = [P(bcs.VeinMesh), P(C.c_int32), P(C.c_float)]
        L.orc_grid_bounds.argtypes = [P(bcs.VeinMesh), P(bcs.Physics), P(C.c_float), P(C.c_float)]
        L.orc_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32]
        L.orc_uniform.restype = C.c_double
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        _LIB = L
    return _LIB


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return lib().orc_get_threads()


def uniform(seed, step, cell, k) -> float:
    return lib().orc_uniform(seed, step, cell, k)


def vein_neighbors(mesh: "bcs.MeshSpec"):
    nv = mesh.positions.size // 3
    ids = np.empty(9 * nv, np.int32)
    lens = np.empty(9 * nv, np.float32)
    m = mesh.c()
    lib().orc_vein_neighbors(C.byref(m), ids.ctypes.data_as(C.POINTER(C.c_int32)),
                             lens.ctypes.data_as(C.POINTER(C.c_float)))
    return ids, lens


def grid_bounds(mesh, physics):
    lo = np.empty(3, np.float32)
    hi = np.empty(3, np.float32)
    m = mesh.c()
    lib().orc_grid_bounds(C.byref(m), C.byref(physics), lo.ctypes.data_as(C.POINTER(C.c_float)),
                          hi.ctypes.data_as(C.POINTER(C.c_float)))
    return lo, hi


class Oracle:
    """One CPU oracle instance (float32 or float64 arithmetic)."""

    def __init__(self, physics, types, mesh, precision: str = "f32"):
        self.L = lib()
        self.pre = "orc_f32_" if precision == "f32" else "orc_f64_"
        self.dtype = np.float32 if precision == "f32" else np.float64
        self.R = C.c_float if precision == "f32" else C.c_double
        self._types = bcs.cell_type_array(types)
        self._specs = list(types)
        self._mesh = mesh
        self._meshc = mesh.c()
        self.h = self._f("create")(C.byref(physics), self._types, len(types), C.byref(self._meshc))
        if not self.h:
            raise RuntimeError("oracle create failed")
        self.np = self._f("particle_count")(self.h)
        self.nv = mesh.positions.size // 3
        self.ntri = mesh.indices.size // 3
        self.nc = sum(t.count for t in types)
        self.nm = sum(t.ppc for t in types)

    def _f(self, name):
        return getattr(self.L, self.pre + name)

    def _p(self, a):
        return a.ctypes.data_as(C.POINTER(self.R))

    def close(self):
        if self.h:
            self._f("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arr(self, a):
        return None if a is None else np.ascontiguousarray(a, self.dtype)

    def set_particles(self, p=None, v=None, f=None):
        a = [self._arr(x) for x in (p, v, f)]
        self._f("set_particles")(self.h, *[None if x is None else self._p(x) for x in a])

    def get_particles(self):
        a = [np.empty(3 * self.np, self.dtype) for _ in range(3)]
        self._f("get_particles")(self.h, *[self._p(x) for x in a])
        return tuple(a)

    def set_vein(self, p=None, v=None, f=None):
        a = [self._arr(x) for x in (p, v, f)]
        self._f("set_vein")(self.h, *[None if x is None else self._p(x) for x in a])

    def get_vein(self):
        a = [np.empty(3 * self.nv, self.dtype) for _ in range(3)]
        self._f("get_vein")(self.h, *[self._p(x) for x in a])
        return tuple(a)

    def get_centers(self):
        c = np.empty(3 * self.nc, self.dtype)
        self._f("get_centers")(self.h, self._p(c))
        return c

    def get_triangle_centers(self):
        c = np.empty(3 * self.ntri, self.dtype)
        self._f("get_triangle_centers")(self.h, self._p(c))
        return c

    def tables(self):
        bs = np.empty(self.nm, self.dtype)
        r0 = np.empty(self.nm, self.dtype)
        g = np.empty(max(sum(t.ppc * t.ppc for t in self._specs), 1), self.dtype)
        self._f("tables")(self.h, self._p(bs), self._p(r0), self._p(g))
        return bs, r0, g

    def calculate_grid(self, which=0):
        self._f("calculate_grid")(self.h, which)

    def get_grid(self, which=0):
        n = self.np if which == 0 else self.ntri
        keys = np.empty(max(n, 1), np.uint32)
        perm = np.empty(max(n, 1), np.int32)
        ok = np.empty(max(n, 1), np.uint32)
        os_ = np.empty(max(n, 1), np.int32)
        oe = np.empty(max(n, 1), np.int32)
        P = C.POINTER
        no = self._f("get_grid")(self.h, which, keys.ctypes.data_as(P(C.c_uint32)), perm.ctypes.data_as(P(C.c_int32)),
                                 ok.ctypes.data_as(P(C.c_uint32)), os_.ctypes.data_as(P(C.c_int32)),
                                 oe.ctypes.data_as(P(C.c_int32)))
        return keys[:n], perm[:n], ok[:no], os_[:no], oe[:no]

    def calculate_next_frame(self):
        self._f("calculate_next_frame")(self.h)

    def propagate_all(self):
        self._f("propagate_all")(self.h)

    def step(self, n=1):
        self._f("step")(self.h, int(n))

    def update_triangle_centers(self):
        self._f("update_triangle_centers")(self.h)

    def contacts(self):
        counts = np.empty(max(self.np, 1), np.int32)
        P = C.POINTER
        total = self._f("contacts")(self.h, counts.ctypes.data_as(P(C.c_int32)), None, 0)
        pairs = np.empty(max(total, 1), np.int32)
        self._f("contacts")(self.h, counts.ctypes.data_as(P(C.c_int32)), pairs.ctypes.data_as(P(C.c_int32)), total)
        return counts[: self.np], pairs[:total]

    def flags(self, clear=False) -> int:
        return int(self._f("flags")(self.h, 1 if clear else 0))

    def set_step_index(self, k: int):
        self._f("set_step_index")(self.h, C.c_uint64(k))
