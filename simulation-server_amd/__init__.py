"""Python (ctypes) binding of libbcs -- the MI355X blood-cell physics step.

The product is the C-ABI shared library ``lib/libbcs.so`` (include/bcs.h,
include/bcs_scene.h).  This module only marshals numpy arrays across that
ABI; every computation happens in the library (HIP kernels for the step, host
C++ for scene generation).  Importing it never falls back to anything else:
if the library is missing, ``load_library`` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "lib" / "libbcs.so"

BCS_OK = 0
BCS_ERR_INVALID_ARGUMENT = 1
BCS_ERR_HIP = 2
BCS_ERR_OUT_OF_MEMORY = 3
BCS_ERR_GRID_TOO_LARGE = 4
BCS_ERR_NO_DEVICE = 5
BCS_ERR_STATE = 6
BCS_ERR_UNSUPPORTED = 7
GRID_PARTICLES = 0
GRID_TRIANGLES = 1

FLAG_OUT_OF_GRID = 0x1
FLAG_NAN_NORMALIZE = 0x2
FLAG_SCAN_TIMEOUT = 0x4
FLAG_RADIUS_CLAMPED = 0x8
FLAG_RAY_SERIAL = 0x10

RAY_AUTO, RAY_SPLIT, RAY_FUSED = 0, 1, 2
CONTACT_AUTO, CONTACT_CELLS, CONTACT_SLOTS = 0, 1, 2

VEIN_CYLINDER, VEIN_Y = 0, 1
CELL_ICOSAHEDRON, CELL_DISC42 = 0, 1
PLACE_RANDOM, PLACE_LATTICE = 0, 1

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)


class Physics(C.Structure):
    _fields_ = [
        ("dt", C.c_float),
        ("use_blood_flow", C.c_int32),
        ("enable_reaction_force", C.c_int32),
        ("enable_big_cells_brake", C.c_int32),
        ("particle_cell", C.c_float * 3),
        ("triangle_cell", C.c_float * 3),
        ("bounding_spheres_coeff", C.c_int32),
        ("grid_y_margin", C.c_float),
        ("grid_xz_margin", C.c_float),
        ("min_spawn_y", C.c_float),
        ("velocity_collision_damping", C.c_float),
        ("particle_k_sniff", C.c_float),
        ("vein_k_sniff", C.c_float),
        ("particle_d_fact", C.c_float),
        ("vein_d_fact", C.c_float),
        ("vein_collision_force_intensity", C.c_float),
        ("viscous_damping", C.c_float),
        ("collision_spring_coeff", C.c_float),
        ("collision_damping_coeff", C.c_float),
        ("collision_shear_coeff", C.c_float),
        ("max_cell_size_factor_before_brake", C.c_float),
        ("big_particle_braking_intensity", C.c_float),
        ("init_velocity", C.c_float * 3),
        ("vein_impact_distance", C.c_float),
        ("vein_impact_minimal_force_distance", C.c_float),
        ("gravity", C.c_float * 3),
        ("cylinder_radius", C.c_float),
        ("seed", C.c_uint64),
        ("dynamic_triangle_grid", C.c_int32),
        ("ray_work_items", C.c_int32),
        ("particle_rk4", C.c_int32),
    ]


class CellType(C.Structure):
    _fields_ = [
        ("count", C.c_int32),
        ("particles_in_cell", C.c_int32),
        ("spring_count", C.c_int32),
        ("spring_start", i32p),
        ("spring_end", i32p),
        ("spring_length", f32p),
        ("vertices", f32p),
        ("contact_radius", C.c_float),
        ("reserved", C.c_int32),
    ]


class VeinMesh(C.Structure):
    _fields_ = [
        ("vertex_count", C.c_int32),
        ("positions", f32p),
        ("triangle_count", C.c_int32),
        ("indices", u32p),
        ("ending_count", C.c_int32),
        ("ending_centers", f32p),
        ("ending_radii", f32p),
    ]


class SimInfo(C.Structure):
    _fields_ = [
        ("particle_count", C.c_int32),
        ("blood_cell_count", C.c_int32),
        ("type_count", C.c_int32),
        ("model_vertex_count", C.c_int32),
        ("vertex_count", C.c_int32),
        ("triangle_count", C.c_int32),
        ("bounds_min", C.c_float * 3),
        ("bounds_max", C.c_float * 3),
        ("particle_cells", C.c_int32 * 3),
        ("triangle_cells", C.c_int32 * 3),
        ("particle_cell_count", C.c_uint64),
        ("triangle_cell_count", C.c_uint64),
        ("particle_starts", C.c_int32 * 16),
        ("cell_type_starts", C.c_int32 * 16),
        ("model_starts", C.c_int32 * 16),
        ("graph_starts", C.c_int32 * 16),
        ("particles_in_cell", C.c_int32 * 16),
        ("cells_of_type", C.c_int32 * 16),
        ("step_index", C.c_uint64),
    ]


class DeviceState(C.Structure):
    _fields_ = [
        ("positions", C.c_void_p),
        ("velocities", C.c_void_p),
        ("forces", C.c_void_p),
        ("centers", C.c_void_p),
        ("vein_positions", C.c_void_p),
        ("vein_velocities", C.c_void_p),
        ("vein_forces", C.c_void_p),
        ("triangle_centers", C.c_void_p),
        ("stream", C.c_void_p),
        ("model_radii", C.c_void_p),
        ("initial_radii", C.c_void_p),
        ("triangle_indices", C.c_void_p),
    ]


class GridView(C.Structure):
    _fields_ = [
        ("grid_cell_ids", C.c_void_p),
        ("particle_ids", C.c_void_p),
        ("grid_cell_starts", C.c_void_p),
        ("grid_cell_ends", C.c_void_p),
        ("object_count", C.c_int32),
        ("cell_counts", C.c_int32 * 3),
        ("cell_count", C.c_uint64),
        ("cell_size", C.c_float * 3),
        ("bounds_min", C.c_float * 3),
        ("bounds_max", C.c_float * 3),
    ]


class SceneCell(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("count", C.c_int32),
        ("radius", C.c_float),
        ("thickness", C.c_float),
        ("contact_radius", C.c_float),
    ]


class SceneParams(C.Structure):
    _fields_ = [
        ("vein_kind", C.c_int32),
        ("radius", C.c_float),
        ("length", C.c_float),
        ("segments", C.c_int32),
        ("rings", C.c_int32),
        ("branch_radius", C.c_float),
        ("branch_length", C.c_float),
        ("branch_angle_deg", C.c_float),
        ("mesh_spacing", C.c_float),
        ("scale", C.c_float),
        ("n_cells", C.c_int32),
        ("cells", SceneCell * 4),
        ("placement", C.c_int32),
        ("packing_fraction", C.c_float),
        ("init_speed", C.c_float),
        ("init_noise_sigma", C.c_float),
        ("seed", C.c_uint64),
        ("physics", Physics),
    ]


class EnsembleTuning(C.Structure):
    _fields_ = [
        ("blocks_per_cu", C.c_int32),
        ("run", C.c_int32),
        ("ray_mode", C.c_int32),
        ("largest_first", C.c_int32),
        ("contact_mode", C.c_int32),
    ]


class SceneView(C.Structure):
    _fields_ = [
        ("physics", Physics),
        ("n_types", C.c_int32),
        ("types", C.POINTER(CellType)),
        ("mesh", VeinMesh),
        ("particle_count", C.c_int32),
        ("positions", f32p),
        ("velocities", f32p),
        ("forces", f32p),
        ("cell_centers", f32p),
        ("cell_count", C.c_int32),
    ]


_LIB = None


def _sig(lib):
    P = C.POINTER
    vp = C.c_void_p
    lib.bcs_last_error.restype = C.c_char_p
    lib.bcs_default_physics.argtypes = [P(Physics)]
    lib.bcs_fold_reference_types.argtypes = [P(CellType), C.c_int32, i32p, i32p, i32p]
    lib.bcs_grid_bounds.argtypes = [P(VeinMesh), P(Physics), f32p, f32p]
    lib.bcs_vein_neighbors.argtypes = [P(VeinMesh), i32p, f32p]
    lib.bcs_model_tables.argtypes = [P(Physics), P(CellType), C.c_int32, f32p, f32p]
    lib.bcs_sim_create.argtypes = [P(Physics), P(CellType), C.c_int32, P(VeinMesh), C.c_int32, P(vp)]
    lib.bcs_sim_destroy.argtypes = [vp]
    lib.bcs_sim_get_info.argtypes = [vp, P(SimInfo)]
    lib.bcs_sim_device_state.argtypes = [vp, P(DeviceState)]
    for n in ("bcs_sim_set_particles", "bcs_sim_get_particles", "bcs_sim_set_vein", "bcs_sim_get_vein"):
        getattr(lib, n).argtypes = [vp, f32p, f32p, f32p]
    lib.bcs_sim_get_centers.argtypes = [vp, f32p]
    lib.bcs_sim_set_step_index.argtypes = [vp, C.c_uint64]
    lib.bcs_sim_calculate_grid.argtypes = [vp, C.c_int32]
    lib.bcs_sim_calculate_next_frame.argtypes = [vp]
    lib.bcs_sim_propagate_all.argtypes = [vp]
    lib.bcs_sim_step.argtypes = [vp, C.c_int32]
    lib.bcs_sim_use_graph.argtypes = [vp, C.c_int32]
    lib.bcs_sim_set_ray_mode.argtypes = [vp, C.c_int32]
    lib.bcs_sim_set_contact_mode.argtypes = [vp, C.c_int32]
    lib.bcs_sim_set_bucket_shift.argtypes = [vp, C.c_int32]
    lib.bcs_sim_synchronize.argtypes = [vp]
    lib.bcs_sim_time_steps.argtypes = [vp, C.c_int32, f32p]
    lib.bcs_sim_profile_kernels.argtypes = [vp, C.c_int32, f32p, C.c_int32, i32p, C.c_char_p, C.c_int32]
    lib.bcs_sim_get_grid.argtypes = [vp, C.c_int32, u32p, i32p, i32p, u32p, i32p, i32p]
    lib.bcs_sim_get_contacts.argtypes = [vp, i32p, i32p, C.c_int64, P(C.c_int64)]
    lib.bcs_sim_error_flags.argtypes = [vp, C.c_int32, u32p]
    lib.bcs_sim_grid_view.argtypes = [vp, C.c_int32, C.c_int32, P(GridView)]
    lib.bcs_ensemble_create.argtypes = [P(vp), C.c_int32, P(vp)]
    lib.bcs_ensemble_destroy.argtypes = [vp]
    lib.bcs_ensemble_set_tuning.argtypes = [vp, P(EnsembleTuning)]
    lib.bcs_ensemble_get_tuning.argtypes = [vp, P(EnsembleTuning)]
    lib.bcs_ensemble_step.argtypes = [vp, C.c_int32]
    lib.bcs_ensemble_synchronize.argtypes = [vp]
    lib.bcs_ensemble_time_steps.argtypes = [vp, C.c_int32, f32p]
    lib.bcs_ensemble_autotune.argtypes = [vp, C.c_int32, P(EnsembleTuning), C.c_char_p, C.c_int32]
    lib.bcs_scene_ensemble_params.argtypes = [C.c_int32, C.c_int32, C.c_uint64, C.c_double, P(SceneParams)]
    lib.bcs_surrogate_create.argtypes = [C.c_int32, C.c_uint64, P(vp)]
    lib.bcs_surrogate_destroy.argtypes = [vp]
    lib.bcs_surrogate_set_weights.argtypes = [vp, P(f32p), P(f32p)]
    lib.bcs_surrogate_get_weights.argtypes = [vp, P(f32p), P(f32p)]
    lib.bcs_surrogate_set_scales.argtypes = [vp, C.c_float, C.c_float, C.c_float]
    lib.bcs_surrogate_get_scales.argtypes = [vp, f32p, f32p, f32p]
    lib.bcs_surrogate_predict.argtypes = [vp, f32p, C.c_int32, f32p]
    lib.bcs_surrogate_train.argtypes = [vp, f32p, f32p, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_uint64, f32p]
    lib.bcs_surrogate_time_inference.argtypes = [vp, C.c_int32, C.c_int32, f32p]
    lib.bcs_sim_attach_surrogate.argtypes = [vp, vp, C.c_float]
    lib.bcs_sim_surrogate_sample.argtypes = [vp, f32p, f32p, C.c_int32, i32p]
    lib.bcs_sim_surrogate_stats.argtypes = [vp, C.c_int32, P(C.c_uint64), P(C.c_uint64), f32p, f32p]
    lib.bcs_sim_cell_ids.argtypes = [vp, C.c_int32, f32p, C.c_int32, u32p]
    lib.bcs_sim_frame_stats.argtypes = [vp, u32p, u32p]
    lib.bcs_device_count.argtypes = [i32p]
    lib.bcs_calibration_copy.argtypes = [C.c_int64, C.c_int32, f32p]
    lib.bcs_scene_preset.argtypes = [C.c_int32, C.c_double, C.c_uint64, P(vp)]
    lib.bcs_scene_default_params.argtypes = [C.c_int32, P(SceneParams)]
    lib.bcs_scene_build.argtypes = [P(SceneParams), P(vp)]
    lib.bcs_scene_get_view.argtypes = [vp, P(SceneView)]
    lib.bcs_scene_destroy.argtypes = [vp]
    lib.bcs_scene_lumen_sdf.argtypes = [vp, f32p, C.c_int32, f32p]
    lib.bcs_sim_export_frame.argtypes = [vp, vp, P(vp), C.c_int32, vp]
    lib.bcs_traj_open.argtypes = [vp, C.c_char_p, C.c_int32, P(vp)]
    lib.bcs_traj_append.argtypes = [vp]
    lib.bcs_traj_close.argtypes = [vp]
    lib.bcs_ref_config_load.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, P(vp)]
    lib.bcs_ref_config_get.argtypes = [vp, i32p, P(P(CellType)), P(VeinMesh)]
    lib.bcs_ref_config_destroy.argtypes = [vp]
    return lib


def load_library(path: str | os.PathLike | None = None):
    """Load libbcs.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"libbcs not built: {p} missing (run __graft_entry__.build())")
    lib = _sig(C.CDLL(str(p)))
    if path is None:
        _LIB = lib
    return lib


class BcsError(RuntimeError):
    pass


def check(rc: int, what: str = ""):
    if rc != BCS_OK:
        msg = load_library().bcs_last_error().decode(errors="replace")
        raise BcsError(f"{what}: status {rc}: {msg}")


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(f32p)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(i32p)


def uptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(u32p)


def default_physics() -> Physics:
    ph = Physics()
    check(load_library().bcs_default_physics(C.byref(ph)), "bcs_default_physics")
    return ph


def device_count() -> int:
    n = C.c_int32(0)
    check(load_library().bcs_device_count(C.byref(n)), "bcs_device_count")
    return n.value


class CellTypeSpec:
    """Owning numpy form of one bcs_cell_type."""

    def __init__(self, count, vertices, springs, contact_radius=0.0):
        self.count = int(count)
        self.vertices = np.ascontiguousarray(np.asarray(vertices, np.float32).reshape(-1))
        sp = np.asarray(springs, dtype=np.float64).reshape(-1, 3) if len(springs) else np.zeros((0, 3))
        self.s_start = np.ascontiguousarray(sp[:, 0].astype(np.int32))
        self.s_end = np.ascontiguousarray(sp[:, 1].astype(np.int32))
        self.s_len = np.ascontiguousarray(sp[:, 2].astype(np.float32))
        self.ppc = self.vertices.size // 3
        self.contact_radius = float(contact_radius)

    def c(self) -> CellType:
        return CellType(self.count, self.ppc, len(self.s_start), self.s_start.ctypes.data_as(i32p),
                        self.s_end.ctypes.data_as(i32p), self.s_len.ctypes.data_as(f32p),
                        self.vertices.ctypes.data_as(f32p), C.c_float(self.contact_radius), 0)


def cell_type_array(specs):
    arr = (CellType * len(specs))()
    for k, s in enumerate(specs):
        arr[k] = s.c()
    return arr


class MeshSpec:
    """Owning numpy form of one bcs_vein_mesh (positions interleaved xyz)."""

    def __init__(self, positions, indices, ending_centers=(), ending_radii=()):
        self.positions = np.ascontiguousarray(np.asarray(positions, np.float32).reshape(-1))
        self.indices = np.ascontiguousarray(np.asarray(indices, np.uint32).reshape(-1))
        self.endc = np.ascontiguousarray(np.asarray(ending_centers, np.float32).reshape(-1))
        self.endr = np.ascontiguousarray(np.asarray(ending_radii, np.float32).reshape(-1))
        if self.endc.size == 0:
            self.endc = np.zeros(3, np.float32)
        if self.endr.size == 0:
            self.endr = np.zeros(1, np.float32)
        self.n_end = len(np.asarray(ending_radii).reshape(-1))

    def c(self) -> VeinMesh:
        return VeinMesh(self.positions.size // 3, fptr(self.positions), self.indices.size // 3, uptr(self.indices),
                        self.n_end, fptr(self.endc), fptr(self.endr))


class Scene:
    """A generated scene (vein mesh, cell types, initial state), copied into numpy."""

    def __init__(self, preset=None, cell_fraction=1.0, seed=0, params: SceneParams | None = None):
        lib = load_library()
        h = C.c_void_p()
        if params is not None:
            check(lib.bcs_scene_build(C.byref(params), C.byref(h)), "bcs_scene_build")
        else:
            check(lib.bcs_scene_preset(int(preset), C.c_double(cell_fraction), C.c_uint64(seed), C.byref(h)),
                  "bcs_scene_preset")
        try:
            v = SceneView()
            check(lib.bcs_scene_get_view(h, C.byref(v)), "bcs_scene_get_view")
            self.physics = Physics.from_buffer_copy(v.physics)
            self.types = []
            for k in range(v.n_types):
                t = v.types[k]
                p = t.particles_in_cell
                verts = np.ctypeslib.as_array(t.vertices, shape=(3 * p,)).copy()
                ns = t.spring_count
                sa = np.ctypeslib.as_array(t.spring_start, shape=(ns,)).copy() if ns else np.zeros(0, np.int32)
                sb = np.ctypeslib.as_array(t.spring_end, shape=(ns,)).copy() if ns else np.zeros(0, np.int32)
                sl = np.ctypeslib.as_array(t.spring_length, shape=(ns,)).copy() if ns else np.zeros(0, np.float32)
                spec = CellTypeSpec(t.count, verts, np.stack([sa, sb, sl], 1) if ns else [], t.contact_radius)
                spec.s_len = sl.astype(np.float32)  # keep exact float lengths
                self.types.append(spec)
            m = v.mesh
            pos = np.ctypeslib.as_array(m.positions, shape=(3 * m.vertex_count,)).copy()
            idx = np.ctypeslib.as_array(m.indices, shape=(3 * m.triangle_count,)).copy()
            endc = np.ctypeslib.as_array(m.ending_centers, shape=(3 * m.ending_count,)).copy() if m.ending_count else []
            endr = np.ctypeslib.as_array(m.ending_radii, shape=(m.ending_count,)).copy() if m.ending_count else []
            self.mesh = MeshSpec(pos, idx, endc, endr)
            n = v.particle_count
            self.positions = np.ctypeslib.as_array(v.positions, shape=(3 * n,)).copy() if n else np.zeros(0, np.float32)
            self.velocities = np.ctypeslib.as_array(v.velocities, shape=(3 * n,)).copy() if n else np.zeros(0, np.float32)
            self.forces = np.ctypeslib.as_array(v.forces, shape=(3 * n,)).copy() if n else np.zeros(0, np.float32)
            nc = v.cell_count
            self.cell_centers = np.ctypeslib.as_array(v.cell_centers, shape=(3 * nc,)).copy() if nc else np.zeros(0, np.float32)
            self.particle_count = n
        finally:
            lib.bcs_scene_destroy(h)

    @staticmethod
    def params(preset: int) -> SceneParams:
        p = SceneParams()
        check(load_library().bcs_scene_default_params(int(preset), C.byref(p)), "bcs_scene_default_params")
        return p


class Simulation:
    """GPU simulation (bcs_sim) built from cell types, a vein mesh and physics."""

    def __init__(self, physics: Physics, types, mesh: MeshSpec, device: int = 0):
        self.lib = load_library()
        self._types = cell_type_array(types)
        self._specs = list(types)
        self._mesh = mesh
        self._meshc = mesh.c()
        self.h = C.c_void_p()
        check(self.lib.bcs_sim_create(C.byref(physics), self._types, len(types), C.byref(self._meshc), device,
                                      C.byref(self.h)), "bcs_sim_create")
        self.info = self.get_info()
        self.np = self.info.particle_count
        self.nv = self.info.vertex_count
        self.nc = self.info.blood_cell_count

    @classmethod
    def from_scene(cls, scene: Scene, device: int = 0, physics: Physics | None = None):
        sim = cls(physics or scene.physics, scene.types, scene.mesh, device)
        sim.set_particles(scene.positions, scene.velocities, scene.forces)
        return sim

    def close(self):
        if self.h:
            self.lib.bcs_sim_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_info(self) -> SimInfo:
        info = SimInfo()
        check(self.lib.bcs_sim_get_info(self.h, C.byref(info)), "bcs_sim_get_info")
        return info

    def device_state(self) -> DeviceState:
        d = DeviceState()
        check(self.lib.bcs_sim_device_state(self.h, C.byref(d)), "bcs_sim_device_state")
        return d

    @staticmethod
    def _opt(a):
        return None if a is None else fptr(np.ascontiguousarray(a, np.float32))

    def set_particles(self, p=None, v=None, f=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (p, v, f)]
        check(self.lib.bcs_sim_set_particles(self.h, *[None if a is None else fptr(a) for a in arrs]), "set_particles")

    def get_particles(self):
        p, v, f = (np.empty(3 * self.np, np.float32) for _ in range(3))
        check(self.lib.bcs_sim_get_particles(self.h, fptr(p), fptr(v), fptr(f)), "get_particles")
        return p, v, f

    def set_vein(self, p=None, v=None, f=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (p, v, f)]
        check(self.lib.bcs_sim_set_vein(self.h, *[None if a is None else fptr(a) for a in arrs]), "set_vein")

    def get_vein(self):
        p, v, f = (np.empty(3 * self.nv, np.float32) for _ in range(3))
        check(self.lib.bcs_sim_get_vein(self.h, fptr(p), fptr(v), fptr(f)), "get_vein")
        return p, v, f

    def get_centers(self):
        c = np.empty(3 * self.nc, np.float32)
        check(self.lib.bcs_sim_get_centers(self.h, fptr(c)), "get_centers")
        return c

    def set_step_index(self, k: int):
        check(self.lib.bcs_sim_set_step_index(self.h, C.c_uint64(k)), "set_step_index")

    def calculate_grid(self, which=GRID_PARTICLES):
        check(self.lib.bcs_sim_calculate_grid(self.h, which), "calculate_grid")

    def calculate_next_frame(self):
        check(self.lib.bcs_sim_calculate_next_frame(self.h), "calculate_next_frame")

    def propagate_all(self):
        check(self.lib.bcs_sim_propagate_all(self.h), "propagate_all")

    def step(self, n=1, sync=True):
        check(self.lib.bcs_sim_step(self.h, int(n)), "step")
        if sync:
            self.synchronize()

    def use_graph(self, enabled: bool = True):
        check(self.lib.bcs_sim_use_graph(self.h, 1 if enabled else 0), "use_graph")

    def set_ray_mode(self, mode: int):
        """RAY_AUTO, RAY_SPLIT or RAY_FUSED: how frames run the vein ray test
        (identical results)."""
        check(self.lib.bcs_sim_set_ray_mode(self.h, int(mode)), "set_ray_mode")

    def set_contact_mode(self, mode: int):
        """CONTACT_AUTO, CONTACT_CELLS or CONTACT_SLOTS: where frames run the
        particle contact search (identical results)."""
        check(self.lib.bcs_sim_set_contact_mode(self.h, int(mode)), "set_contact_mode")

    def set_bucket_shift(self, shift: int):
        """2^shift grid cells per bucket of the particle sort (identical
        results; the automatic choice returns with the next set_particles)."""
        check(self.lib.bcs_sim_set_bucket_shift(self.h, int(shift)), "set_bucket_shift")

    def synchronize(self):
        check(self.lib.bcs_sim_synchronize(self.h), "synchronize")

    def time_steps(self, n) -> float:
        ms = C.c_float(0)
        check(self.lib.bcs_sim_time_steps(self.h, int(n), C.byref(ms)), "time_steps")
        return ms.value

    def profile_kernels(self, n):
        ms = np.zeros(32, np.float32)
        nk = C.c_int32(0)
        names = C.create_string_buffer(1024)
        check(self.lib.bcs_sim_profile_kernels(self.h, int(n), fptr(ms), 32, C.byref(nk), names, 1024),
              "profile_kernels")
        return dict(zip(names.value.decode().split(";"), ms[: nk.value].tolist()))

    def get_grid(self, which=GRID_PARTICLES):
        n = self.np if which == GRID_PARTICLES else self.info.triangle_count
        keys = np.empty(max(n, 1), np.uint32)
        perm = np.empty(max(n, 1), np.int32)
        occk = np.empty(max(n, 1), np.uint32)
        occs = np.empty(max(n, 1), np.int32)
        occe = np.empty(max(n, 1), np.int32)
        no = C.c_int32(0)
        check(self.lib.bcs_sim_get_grid(self.h, which, uptr(keys), iptr(perm), C.byref(no), uptr(occk), iptr(occs),
                                        iptr(occe)), "get_grid")
        k = no.value
        return keys[:n], perm[:n], occk[:k], occs[:k], occe[:k]

    def get_contacts(self):
        counts = np.empty(max(self.np, 1), np.int32)
        total = C.c_int64(0)
        check(self.lib.bcs_sim_get_contacts(self.h, iptr(counts), None, 0, C.byref(total)), "get_contacts(count)")
        pairs = np.empty(max(total.value, 1), np.int32)
        check(self.lib.bcs_sim_get_contacts(self.h, iptr(counts), iptr(pairs), total.value, C.byref(total)),
              "get_contacts")
        return counts[: self.np], pairs[: total.value]

    def grid_view(self, which=GRID_PARTICLES, dense_ranges=False) -> GridView:
        """Device pointers of a grid's sorted arrays (and, with dense_ranges,
        per-cell-id start/end arrays refreshed from the current grid)."""
        v = GridView()
        check(self.lib.bcs_sim_grid_view(self.h, which, 1 if dense_ranges else 0, C.byref(v)), "grid_view")
        return v

    def cell_ids(self, positions, which=GRID_PARTICLES):
        """Cell ids of SoA host positions (x[n], y[n], z[n]) on a grid."""
        p = np.ascontiguousarray(positions, np.float32)
        n = p.size // 3
        out = np.empty(max(n, 1), np.uint32)
        check(self.lib.bcs_sim_cell_ids(self.h, which, fptr(p), n, uptr(out)), "cell_ids")
        return out[:n]

    def attach_surrogate(self, net, tolerance):
        """Springs of the 42-particle cells from `net` (a Surrogate) where a
        cell's largest error against the plain solve is <= tolerance; None
        detaches."""
        self._surrogate = net
        check(self.lib.bcs_sim_attach_surrogate(self.h, net.h if net is not None else None, C.c_float(tolerance)),
              "attach_surrogate")

    def surrogate_sample(self, capacity=None):
        """Runs two frames; returns (features n x 294, targets n x 126) of the
        second frame's 42-particle cells (network input and plain spring
        displacements)."""
        cap = int(capacity if capacity is not None else self.np)
        feat = np.empty((max(cap, 1), 294), np.float32)
        targ = np.empty((max(cap, 1), 126), np.float32)
        n = C.c_int32(0)
        check(self.lib.bcs_sim_surrogate_sample(self.h, fptr(feat), fptr(targ), cap, C.byref(n)), "surrogate_sample")
        return feat[: n.value], targ[: n.value]

    def surrogate_stats(self, reset=False) -> dict:
        a, f = C.c_uint64(0), C.c_uint64(0)
        ea, e_all = C.c_float(0), C.c_float(0)
        check(self.lib.bcs_sim_surrogate_stats(self.h, 1 if reset else 0, C.byref(a), C.byref(f), C.byref(ea),
                                               C.byref(e_all)), "surrogate_stats")
        return {"accepted": a.value, "fallback": f.value, "max_error_accepted": ea.value, "max_error_all": e_all.value}

    def frame_stats(self):
        w, h = C.c_uint32(0), C.c_uint32(0)
        check(self.lib.bcs_sim_frame_stats(self.h, C.byref(w), C.byref(h)), "frame_stats")
        return {"ray_work_items": w.value, "vein_hits": h.value}

    def export_frame(self, offsets=0, type_vertices=(), vein_vertices=0):
        """Frame export into DEVICE buffers given as integer addresses (e.g.
        torch tensor data_ptr()); asynchronous on the simulation's stream."""
        tv = (C.c_void_p * max(len(type_vertices), 1))(*[C.c_void_p(int(a)) for a in type_vertices])
        check(self.lib.bcs_sim_export_frame(self.h, C.c_void_p(int(offsets) or None), tv, len(type_vertices),
                                            C.c_void_p(int(vein_vertices) or None)), "export_frame")

    def open_trajectory(self, path, with_velocities=False):
        return Trajectory(self, path, with_velocities)

    def error_flags(self, clear=False) -> int:
        f = C.c_uint32(0)
        check(self.lib.bcs_sim_error_flags(self.h, 1 if clear else 0, C.byref(f)), "error_flags")
        return f.value


class Ensemble:
    """Many simulations stepped together (bcs_ensemble): every frame kernel
    runs once for all members.  The members stay usable as Simulation objects
    (read their state after synchronize()); keep them alive meanwhile."""

    def __init__(self, sims):
        self.lib = load_library()
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.h.value for s in self.sims])
        self.h = C.c_void_p()
        check(self.lib.bcs_ensemble_create(arr, len(self.sims), C.byref(self.h)), "bcs_ensemble_create")

    def close(self):
        if self.h:
            self.lib.bcs_ensemble_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1, sync=True):
        check(self.lib.bcs_ensemble_step(self.h, int(n)), "ensemble step")
        if sync:
            self.synchronize()

    def synchronize(self):
        check(self.lib.bcs_ensemble_synchronize(self.h), "ensemble synchronize")

    def time_steps(self, n) -> float:
        ms = C.c_float(0)
        check(self.lib.bcs_ensemble_time_steps(self.h, int(n), C.byref(ms)), "ensemble time_steps")
        return ms.value

    def set_tuning(self, blocks_per_cu=0, run=1, ray_mode=RAY_AUTO, largest_first=1, contact_mode=CONTACT_AUTO):
        t = EnsembleTuning(int(blocks_per_cu), int(run), int(ray_mode), int(largest_first), int(contact_mode))
        check(self.lib.bcs_ensemble_set_tuning(self.h, C.byref(t)), "ensemble set_tuning")

    def get_tuning(self) -> dict:
        t = EnsembleTuning()
        check(self.lib.bcs_ensemble_get_tuning(self.h, C.byref(t)), "ensemble get_tuning")
        return {"blocks_per_cu": t.blocks_per_cu, "run": t.run, "ray_mode": t.ray_mode,
                "largest_first": t.largest_first, "contact_mode": t.contact_mode}

    def autotune(self, frames_per_trial=10) -> dict:
        """Times every tuning of the search space and keeps the fastest;
        returns the report (search space, trials, choice)."""
        import json

        best = EnsembleTuning()
        buf = C.create_string_buffer(16384)
        check(self.lib.bcs_ensemble_autotune(self.h, int(frames_per_trial), C.byref(best), buf, len(buf)),
              "ensemble autotune")
        return json.loads(buf.value.decode())


SURR_SHAPES = [(256, 294), (256, 256), (256, 256), (126, 256)]


class Surrogate:
    """The learned spring surrogate (bcs_surrogate): MLP 294-256-256-256-126
    (GELU), BF16 on the matrix cores with FP32 accumulation."""

    def __init__(self, seed=0, device=0):
        self.lib = load_library()
        self.h = C.c_void_p()
        check(self.lib.bcs_surrogate_create(int(device), C.c_uint64(seed), C.byref(self.h)), "bcs_surrogate_create")

    def close(self):
        if self.h:
            self.lib.bcs_surrogate_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_weights(self):
        ws = [np.empty(sh, np.float32) for sh in SURR_SHAPES]
        bs = [np.empty(sh[0], np.float32) for sh in SURR_SHAPES]
        wp = (f32p * 4)(*[fptr(w) for w in ws])
        bp = (f32p * 4)(*[fptr(b) for b in bs])
        check(self.lib.bcs_surrogate_get_weights(self.h, wp, bp), "surrogate get_weights")
        return ws, bs

    def set_weights(self, ws, bs):
        ws = [np.ascontiguousarray(w, np.float32) for w in ws]
        bs = [np.ascontiguousarray(b, np.float32) for b in bs]
        for w, b, sh in zip(ws, bs, SURR_SHAPES):
            assert w.shape == sh and b.shape == (sh[0],)
        wp = (f32p * 4)(*[fptr(w) for w in ws])
        bp = (f32p * 4)(*[fptr(b) for b in bs])
        check(self.lib.bcs_surrogate_set_weights(self.h, wp, bp), "surrogate set_weights")

    def set_scales(self, pos_scale, vel_scale, out_scale):
        check(self.lib.bcs_surrogate_set_scales(self.h, C.c_float(pos_scale), C.c_float(vel_scale),
                                                C.c_float(out_scale)), "surrogate set_scales")

    def get_scales(self):
        a, b, c = C.c_float(0), C.c_float(0), C.c_float(0)
        check(self.lib.bcs_surrogate_get_scales(self.h, C.byref(a), C.byref(b), C.byref(c)), "surrogate get_scales")
        return a.value, b.value, c.value

    def predict(self, features):
        x = np.ascontiguousarray(features, np.float32).reshape(-1, 294)
        out = np.empty((max(x.shape[0], 1), 126), np.float32)
        check(self.lib.bcs_surrogate_predict(self.h, fptr(x), x.shape[0], fptr(out)), "surrogate predict")
        return out[: x.shape[0]]

    def train(self, features, targets, epochs=10, batch=4096, lr=1e-3, seed=0):
        x = np.ascontiguousarray(features, np.float32).reshape(-1, 294)
        y = np.ascontiguousarray(targets, np.float32).reshape(-1, 126)
        loss = np.zeros(max(epochs, 1), np.float32)
        check(self.lib.bcs_surrogate_train(self.h, fptr(x), fptr(y), x.shape[0], int(epochs), int(batch),
                                           C.c_float(lr), C.c_uint64(seed), fptr(loss)), "surrogate train")
        return loss

    def time_inference(self, n_cells, reps=10) -> float:
        ms = C.c_float(0)
        check(self.lib.bcs_surrogate_time_inference(self.h, int(n_cells), int(reps), C.byref(ms)),
              "surrogate time_inference")
        return ms.value


def ensemble_scene_params(n_veins: int, index: int, seed: int = 0, cell_fraction: float = 1.0) -> SceneParams:
    """Scene parameters of vein `index` of an n_veins batched ensemble
    (bcs_scene_ensemble_params)."""
    p = SceneParams()
    check(load_library().bcs_scene_ensemble_params(int(n_veins), int(index), C.c_uint64(seed),
                                                    C.c_double(cell_fraction), C.byref(p)),
          "bcs_scene_ensemble_params")
    return p


class Trajectory:
    """Appends frames of a simulation to a trajectory file (bcs_traj_*)."""

    def __init__(self, sim, path, with_velocities=False):
        self.lib = sim.lib
        self.h = C.c_void_p()
        check(self.lib.bcs_traj_open(sim.h, str(path).encode(), 1 if with_velocities else 0, C.byref(self.h)),
              "traj_open")

    def append(self):
        check(self.lib.bcs_traj_append(self.h), "traj_append")

    def close(self):
        if self.h:
            check(self.lib.bcs_traj_close(self.h), "traj_close")
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_trajectory(path):
    """Reads a trajectory file: dict with particles, vein_vertices, types, dt,
    steps[F] (uint64) and positions[F, 3, N], velocities[F, 3, N] or None,
    vein[F, 3, Nv] (float32)."""
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size < 64 or raw[:8].tobytes() != b"BCSTRAJ1":
        raise ValueError(f"{path}: not a BCSTRAJ1 trajectory")
    n, nv, nt, flags = np.frombuffer(raw[8:24].tobytes(), dtype="<i4")
    dt = float(np.frombuffer(raw[24:28].tobytes(), dtype="<f4")[0])
    vel = bool(flags & 1)
    frame = 8 + 4 * (3 * n * (2 if vel else 1) + 3 * nv)
    body = raw[64:]
    if body.size % frame:
        raise ValueError(f"{path}: truncated frame")
    nf = body.size // frame
    fr = body.reshape(nf, frame)
    steps = np.frombuffer(fr[:, :8].tobytes(), dtype="<u8")
    f32 = np.frombuffer(fr[:, 8:].tobytes(), dtype="<f4").reshape(nf, -1)
    pos = f32[:, : 3 * n].reshape(nf, 3, n)
    o = 3 * n
    velo = None
    if vel:
        velo = f32[:, o: o + 3 * n].reshape(nf, 3, n)
        o += 3 * n
    vein = f32[:, o: o + 3 * nv].reshape(nf, 3, nv)
    return {"particles": int(n), "vein_vertices": int(nv), "types": int(nt), "dt": dt, "steps": steps,
            "positions": pos, "velocities": velo, "vein": vein}


def load_reference_config(definitions=None, presets=None, vein=None):
    """Cell types (definition order) and vein mesh from configuration headers
    in the reference's format (src/config/blood_cells_definition.hpp,
    blood_cell_presets.hpp, vein_definition.hpp).  Returns (types, mesh);
    either is None when its files were not given."""
    lib = load_library()
    enc = lambda p: None if p is None else str(p).encode()
    h = C.c_void_p()
    check(lib.bcs_ref_config_load(enc(definitions), enc(presets), enc(vein), C.byref(h)), "bcs_ref_config_load")
    try:
        n = C.c_int32(0)
        tp = C.POINTER(CellType)()
        m = VeinMesh()
        check(lib.bcs_ref_config_get(h, C.byref(n), C.byref(tp), C.byref(m)), "bcs_ref_config_get")
        types = None
        if definitions is not None:
            types = []
            for k in range(n.value):
                t = tp[k]
                ppc, ns = t.particles_in_cell, t.spring_count
                verts = np.ctypeslib.as_array(t.vertices, shape=(3 * ppc,)).copy()
                sa = np.ctypeslib.as_array(t.spring_start, shape=(ns,)).copy()
                sb = np.ctypeslib.as_array(t.spring_end, shape=(ns,)).copy()
                sl = np.ctypeslib.as_array(t.spring_length, shape=(ns,)).copy()
                spec = CellTypeSpec(t.count, verts, np.stack([sa, sb, sl], 1), 0.0)
                spec.s_len = sl.astype(np.float32)
                types.append(spec)
        mesh = None
        if vein is not None:
            pos = np.ctypeslib.as_array(m.positions, shape=(3 * m.vertex_count,)).copy()
            idx = np.ctypeslib.as_array(m.indices, shape=(3 * m.triangle_count,)).copy()
            ne = m.ending_count
            endc = np.ctypeslib.as_array(m.ending_centers, shape=(3 * ne,)).copy() if ne else []
            endr = np.ctypeslib.as_array(m.ending_radii, shape=(ne,)).copy() if ne else []
            mesh = MeshSpec(pos, idx, endc, endr)
        return types, mesh
    finally:
        lib.bcs_ref_config_destroy(h)


def calibration_copy(n_bytes: int, repeats: int = 10) -> float:
    ms = C.c_float(0)
    check(load_library().bcs_calibration_copy(C.c_int64(n_bytes), repeats, C.byref(ms)), "calibration_copy")
    return ms.value
