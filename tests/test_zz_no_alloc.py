"""bcs_sim_get_grid_static: the download of a grid (sorted keys,
permutation, occupied cells with inclusive ranges) that allocates nothing --
it stages through the library's static device array -- against
bcs_sim_get_grid, which stages through buffers it allocates.  Two calls in a
row, the second on a smaller input, on both grids; then a grid with more
occupied cells than the static staging holds is refused with an error."""
import ctypes as C

import numpy as np
import pytest

from conftest import assert_bits_equal

pytestmark = pytest.mark.gpu

STATIC_OCCUPIED = 1 << 20


def _get_grid_static(bcs, sim, which):
    lib = sim.lib
    f = lib.bcs_sim_get_grid_static
    f.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                  C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    n = sim.np if which == 0 else sim.info.triangle_count
    keys = np.empty(n, np.uint32)
    perm = np.empty(n, np.int32)
    okeys = np.empty(n, np.uint32)
    ostart = np.empty(n, np.int32)
    oend = np.empty(n, np.int32)
    no = C.c_int32(-1)
    bcs.check(f(sim.h, which, bcs.uptr(keys), bcs.iptr(perm), C.byref(no), bcs.uptr(okeys), bcs.iptr(ostart),
                bcs.iptr(oend)), "bcs_sim_get_grid_static")
    k = no.value
    return keys, perm, okeys[:k], ostart[:k], oend[:k]


@pytest.mark.parametrize("which", [0, 1])
def test_static_grid_download_equals_allocating_one(bcs, which):
    for fraction in (1.0, 0.2):  # the second call has the smaller input
        sc = bcs.Scene(0, fraction, seed=3)
        sim = bcs.Simulation.from_scene(sc)
        sim.step(4)
        sim.calculate_grid(0)
        a = _get_grid_static(bcs, sim, which)
        b = sim.get_grid(which)
        for name, x, y in zip(("keys", "perm", "occ_keys", "occ_start", "occ_end"), a, b):
            assert_bits_equal(x, y, f"fraction {fraction}, grid {which}: {name}")
        assert a[2].size > 0
        sim.close()


def test_static_grid_download_refuses_more_than_its_staging(bcs):
    sc = bcs.Scene(2, 0.3, seed=0)  # about 1.47M occupied particle cells
    sim = bcs.Simulation.from_scene(sc)
    sim.calculate_grid(0)
    assert sim.get_grid(0)[2].size > STATIC_OCCUPIED
    with pytest.raises(bcs.BcsError, match="static staging"):
        _get_grid_static(bcs, sim, 0)
    # the triangle grid of the same simulation still fits, and the call after
    # the refusal works
    a = _get_grid_static(bcs, sim, 1)
    b = sim.get_grid(1)
    for name, x, y in zip(("keys", "perm", "occ_keys", "occ_start", "occ_end"), a, b):
        assert_bits_equal(x, y, f"after the refusal: {name}")
