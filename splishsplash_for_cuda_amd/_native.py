"""ctypes binding of libdfsph_hip.so (include/dfsph_hip.h) and libdfsph_host.so
(include/dfsph_host.h).

The libraries are built in-tree by ``splishsplash_for_cuda_amd.build``.  There
is no fallback: if a library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(HERE, "libdfsph_hip.so")
# DFSPH_HOST_LIB: another build of the host library (the sanitizer build of oracle/sanitize.sh)
HOST_LIB = os.environ.get("DFSPH_HOST_LIB") or os.path.join(HERE, "libdfsph_host.so")

DFSPH_OK = 0
DFSPH_ERR_HIP = 1
DFSPH_ERR_ARG = 2
DFSPH_ERR_CAPACITY = 3
DFSPH_ERR_NEIGHBOR_OVERFLOW = 4
DFSPH_ERR_STATE = 5
DFSPH_ERR_NO_MODEL = 6

WARMSTART_REFERENCE = 0
WARMSTART_LEARNED = 1
WARMSTART_NONE = 2

LISTS_AUTO = 0
LISTS_NO_LDS = 1
LISTS_GLOBAL = 2
LEARNED_ALWAYS = 0
LEARNED_ABOVE_FLOORS = 1

FIELD_POS = 0
FIELD_VEL = 1
FIELD_MASS = 2
FIELD_DENSITY = 3
FIELD_FACTOR = 4
FIELD_DENSITY_ADV = 5
FIELD_KAPPA = 6
FIELD_KAPPAV = 7
FIELD_ACC = 8
FIELD_ID = 9
FIELD_NBR_FLUID = 10
FIELD_NBR_BOUNDARY = 11
FIELD_BOUNDARY_POS = 12
FIELD_BOUNDARY_PSI = 13
FIELD_TAG = 14
FIELD_KAPPA_TOTAL = 15
FIELD_KAPPAV_TOTAL = 16
_VECTOR_FIELDS = {FIELD_POS, FIELD_VEL, FIELD_ACC, FIELD_BOUNDARY_POS}
_UINT_FIELDS = {FIELD_ID, FIELD_NBR_FLUID, FIELD_NBR_BOUNDARY, FIELD_TAG}


class DFSPHParams(C.Structure):
    _fields_ = [
        ("particle_radius", C.c_double),
        ("density0", C.c_double),
        ("viscosity", C.c_double),
        ("gravity", C.c_double * 3),
        ("dt", C.c_double),
        ("max_iter_v", C.c_uint32),
        ("max_iter", C.c_uint32),
        ("min_iter_v", C.c_uint32),
        ("min_iter", C.c_uint32),
        ("max_error_v", C.c_double),
        ("max_error", C.c_double),
        ("use_cfl", C.c_int32),
        ("cfl_factor", C.c_double),
        ("dt_min", C.c_double),
        ("dt_max", C.c_double),
        ("enable_divergence", C.c_int32),
        ("check_positions_mode", C.c_int32),
        ("max_neighbors", C.c_uint32),
        ("fluid_capacity", C.c_uint32),
        ("warm_start_mode", C.c_int32),
        ("check_interval", C.c_uint32),
        ("list_mode", C.c_int32),
        ("learned_gate", C.c_int32),
    ]


class DFSPHStepStats(C.Structure):
    _fields_ = [
        ("iters_v", C.c_uint32),
        ("iters", C.c_uint32),
        ("err_v", C.c_double),
        ("err", C.c_double),
        ("dt", C.c_double),
        ("time", C.c_double),
        ("n_fluid", C.c_uint32),
        ("removed", C.c_uint32),
        ("max_neighbors_seen", C.c_uint32),
        ("overflow", C.c_uint32),
        ("neighbors_fluid_total", C.c_uint64),
        ("neighbors_boundary_total", C.c_uint64),
    ]


class DFSPHInfo(C.Structure):
    _fields_ = [
        ("n_fluid", C.c_uint32),
        ("n_boundary", C.c_uint32),
        ("fluid_capacity", C.c_uint32),
        ("max_neighbors", C.c_uint32),
        ("grid_dim", C.c_uint32 * 3),
        ("grid_origin", C.c_double * 3),
        ("cell_size", C.c_double),
        ("h", C.c_double),
        ("h_future", C.c_double),
        ("h_past", C.c_double),
        ("time", C.c_double),
        ("bbox_min", C.c_double * 3),
        ("bbox_max", C.c_double * 3),
        ("max_neighbors_seen", C.c_uint32),
        ("overflow", C.c_uint32),
        ("escaped", C.c_uint32),
        ("unstaged_blocks", C.c_uint32),
        ("neighbors_fluid_total", C.c_uint64),
        ("neighbors_boundary_total", C.c_uint64),
        ("neighbors_fluid_unstaged", C.c_uint64),
        ("sweep_block", C.c_uint32),
    ]


class DFSPHCudaStats(C.Structure):
    _fields_ = [
        ("iters_v", C.c_uint32),
        ("iters", C.c_uint32),
        ("err_v", C.c_double),
        ("err", C.c_double),
        ("h", C.c_double),
        ("time", C.c_double),
        ("n_fluid", C.c_uint32),
        ("removed", C.c_int32),
        ("inflow_added", C.c_uint32),
        ("outflow_removed", C.c_uint32),
        ("max_neighbors_seen", C.c_uint32),
        ("overflow", C.c_uint32),
    ]


# every function declared in include/dfsph_hip.h, with its ctypes signature
BODY_DYNAMIC = 0
BODY_KINEMATIC = 1
BODY_SINE = 2


class DFSPHRigidBody(C.Structure):
    _fields_ = [("type", C.c_int32), ("mass", C.c_double), ("inertia", C.c_double * 3),
                ("position", C.c_double * 3), ("velocity", C.c_double * 3), ("rotation", C.c_double * 4),
                ("angular_velocity", C.c_double * 3), ("axis", C.c_double * 3), ("amplitude", C.c_double),
                ("period", C.c_double)]


def rigid_body(type=BODY_DYNAMIC, mass=0.0, inertia=(0, 0, 0), position=(0, 0, 0), velocity=(0, 0, 0),
               rotation=(0, 0, 0, 1), angular_velocity=(0, 0, 0), axis=(0, 0, 0), amplitude=0.0,
               period=0.0) -> DFSPHRigidBody:
    b = DFSPHRigidBody()
    b.type = type
    b.mass = mass
    b.inertia = (C.c_double * 3)(*inertia)
    b.position = (C.c_double * 3)(*position)
    b.velocity = (C.c_double * 3)(*velocity)
    b.rotation = (C.c_double * 4)(*rotation)
    b.angular_velocity = (C.c_double * 3)(*angular_velocity)
    b.axis = (C.c_double * 3)(*axis)
    b.amplitude = amplitude
    b.period = period
    return b


class DFSPHSurface(C.Structure):
    """dfsph_surface: a BufferFluidSurface (plane, cuboid, cylinder, mesh, sphere, star)."""
    _fields_ = [
        ("type", C.c_int32),
        ("reversed", C.c_int32),
        ("o", C.c_float * 3),
        ("n", C.c_float * 3),
        ("center", C.c_float * 3),
        ("half", C.c_float * 3),
        ("radius", C.c_float),
        ("radius_internal", C.c_float),
        ("points", C.c_uint32),
        ("direction", C.c_float * 3),
        ("mesh", C.c_uint32),
    ]


SURFACE_PLANE, SURFACE_CUBOID, SURFACE_CYLINDER, SURFACE_MESH, SURFACE_SPHERE, SURFACE_STAR = 0, 1, 2, 3, 4, 5
TAG_ACTIVE, TAG_ACTIVE_NEIGHBORS, TAG_UNTAGGED, TAG_REMOVAL = 10000000, 20000000, 40000000, 310000000
TAG_AIR = 50000000


def surface(kind, reversed=False, o=(0, 0, 0), n=(0, 1, 0), center=(0, 0, 0), half=(0, 0, 0), radius=0.0,
            radius_internal=0.0, points=0, direction=(0, 0, 1), mesh=0):
    """A DFSPHSurface; kind is 'plane', 'cuboid', 'cylinder', 'mesh', 'sphere' or 'star'
    (cylinder / star: half height in half[1]; mesh: the id of Solver.mesh_upload, moved by center)."""
    s = DFSPHSurface()
    s.type = {"plane": SURFACE_PLANE, "cuboid": SURFACE_CUBOID, "cylinder": SURFACE_CYLINDER, "mesh": SURFACE_MESH,
              "sphere": SURFACE_SPHERE, "star": SURFACE_STAR}[kind]
    s.mesh = mesh
    s.reversed = 1 if reversed else 0
    s.o = (C.c_float * 3)(*o)
    s.n = (C.c_float * 3)(*n)
    s.center = (C.c_float * 3)(*center)
    s.half = (C.c_float * 3)(*half)
    s.radius = radius
    s.radius_internal = radius_internal
    s.points = points
    s.direction = (C.c_float * 3)(*direction)
    return s


_P = C.c_void_p
_F32P = C.POINTER(C.c_float)
_U32P = C.POINTER(C.c_uint32)
_U8P = C.POINTER(C.c_uint8)
_D = C.c_double
_U32 = C.c_uint32
_I = C.c_int

HIP_SIGNATURES = {
    "dfsph_set_restriction": (_I, [_P, _I, _U32, _U32, C.c_int64]),
    "dfsph_tag_neighbors": (_I, [_P, _U32, _U32, _I, _U32P]),
    "dfsph_remove_density_peaks": (_I, [_P, C.c_float, C.c_float, _U32P, C.POINTER(C.c_double)]),
    "dfsph_get_params": (_I, [_P, C.POINTER(DFSPHParams)]),
    "dfsph_count_tagged": (_I, [_P, _U32, _U32P]),
    "dfsph_order_active_first": (_I, [_P, _U32P, _U32P]),
    "dfsph_remove_tagged": (_I, [_P, _U32, _U32P]),
    "dfsph_surface_inside": (_I, [_P, C.POINTER(DFSPHSurface), _F32P, _U32, _U8P]),
    "dfsph_mesh_upload": (_I, [_P, _F32P, _U32, _U32P, _U32, _U32P]),
    "dfsph_open_boundary_init": (_I, [_P, C.POINTER(DFSPHSurface), C.POINTER(DFSPHSurface),
                                      C.POINTER(DFSPHSurface), _F32P, _U32, _U32P]),
    "dfsph_open_boundary_apply": (_I, [_P, C.POINTER(C.c_double), C.c_float, C.c_float, _I, _I, _U32P, _U32P]),
    "dfsph_add_rigid_body": (_I, [_P, C.POINTER(DFSPHRigidBody), _F32P, _F32P, _U32, _U32P]),
    "dfsph_set_rigid_body": (_I, [_P, _U32, C.POINTER(DFSPHRigidBody)]),
    "dfsph_get_rigid_body": (_I, [_P, _U32, C.POINTER(DFSPHRigidBody), C.POINTER(_D), C.POINTER(_D)]),
    "dfsph_rigid_body_forces": (_I, [_P, _U32, _F32P, _U32]),
    "dfsph_rigid_body_count": (_I, [_P, _U32P]),
    "dfsph_default_params": (None, [C.POINTER(DFSPHParams)]),
    "dfsph_last_error": (C.c_char_p, []),
    "dfsph_create": (_I, [C.POINTER(DFSPHParams), C.POINTER(C.c_void_p)]),
    "dfsph_destroy": (_I, [_P]),
    "dfsph_set_stream": (_I, [_P, _P]),
    "dfsph_get_stream": (_I, [_P, C.POINTER(C.c_void_p)]),
    "dfsph_upload_boundary": (_I, [_P, _F32P, _F32P, _U32]),
    "dfsph_upload_fluid": (_I, [_P, _F32P, _F32P, _F32P, _U32P, _U32]),
    "dfsph_neighbor_search": (_I, [_P, _I]),
    "dfsph_compute_density_factor": (_I, [_P]),
    "dfsph_divergence_solve": (_I, [_P, _U32, _D, _U32P, C.POINTER(_D)]),
    "dfsph_external_forces": (_I, [_P]),
    "dfsph_cfl": (_I, [_P, _D, _D, _D, C.POINTER(_D)]),
    "dfsph_get_time": (_I, [_P, C.POINTER(_D), C.POINTER(_D), C.POINTER(_D)]),
    "dfsph_fluid_count": (_I, [_P, C.POINTER(C.c_uint32)]),
    "dfsph_update_time_step": (_I, [_P, _D]),
    "dfsph_update_vel": (_I, [_P]),
    "dfsph_pressure_solve": (_I, [_P, _U32, _D, _U32P, C.POINTER(_D)]),
    "dfsph_update_pos": (_I, [_P]),
    "dfsph_step_end": (_I, [_P]),
    "dfsph_set_time": (_I, [_P, _D, _D, _D]),
    "dfsph_check_particles_positions": (_I, [_P, _I, _U32P]),
    "dfsph_remove_particles": (_I, [_P, _U8P, _U32P]),
    "dfsph_add_particles": (_I, [_P, _F32P, _F32P, _F32P, _U32P, _U32]),
    "dfsph_step": (_I, [_P, C.POINTER(DFSPHStepStats)]),
    "dfsph_download": (_I, [_P, _I, _P, _U32]),
    "dfsph_upload_field": (_I, [_P, _I, _P, _U32]),
    "dfsph_get_neighbors": (_I, [_P, _U32P, _U32P, _U32P, _U32]),
    "dfsph_get_info": (_I, [_P, C.POINTER(DFSPHInfo)]),
    "dfsph_synchronize": (_I, [_P]),
    "dfsph_change_max_particle_number": (_I, [_P, _U32]),
    "dfsph_load_fluid": (_I, [_P, _F32P, _U32, C.c_float, C.c_float, _I, C.POINTER(C.c_uint32)]),
    "dfsph_neighbor_stats": (_I, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_uint64)]),
    "dfsph_profile_enable": (_I, [_P, _U32]),
    "dfsph_profile_reset": (_I, [_P]),
    "dfsph_profile_read": (_I, [_P, _U32, _U32P, C.POINTER(_D)]),
    "dfsph_mlp_set_params": (_I, [_P, _F32P, _F32P]),
    "dfsph_mlp_get_params": (_I, [_P, _F32P, _F32P]),
    "dfsph_mlp_forward": (_I, [_P, _F32P, _U32, _F32P, _I]),
    "dfsph_mlp_forward_noalloc": (_I, [_P, _F32P, _U32, _F32P, _I]),
    "dfsph_mlp_loss_grad": (_I, [_P, _F32P, _F32P, _U32, _F32P, C.POINTER(_D), _I]),
    "dfsph_mlp_harvest": (_I, [_P, _U32, _U32]),
    "dfsph_mlp_dataset_size": (_I, [_P, _U32P]),
    "dfsph_mlp_dataset_download": (_I, [_P, _F32P, _F32P, _U32]),
    "dfsph_mlp_dataset_upload": (_I, [_P, _F32P, _F32P, _U32]),
    "dfsph_mlp_fit_normalization": (_I, [_P]),
    "dfsph_mlp_train": (_I, [_P, _U32, _U32, C.c_float, _U32, _F32P, _F32P]),
    "dfsph_mlp_step_features": (_I, [_P, _F32P, _F32P, _F32P, _U32]),
}

MLP_NPARAM = 9282
MLP_NNORM = 28
MLP_NOALLOC_MAX = 65536
MLP_NFEAT = 12

KERNEL_NAMES = ["sort", "neighbors", "div_ws_init", "div_init", "div_jacobi", "div_loop_end", "xsph",
                "pressure_ws", "dens_init", "dens_jacobi", "dens_loop_end", "update_pos", "features", "mlp_infer"]

HOST_SIGNATURES = {
    "dfsph_cuda_new": (C.c_void_p, [C.POINTER(DFSPHParams), _F32P, _U32, _F32P, _F32P, _U32, C.c_char_p, _U32]),
    "dfsph_cuda_delete": (None, [_P]),
    "dfsph_cuda_step": (_I, [_P, C.POINTER(DFSPHCudaStats)]),
    "dfsph_cuda_reset": (_I, [_P]),
    "dfsph_cuda_set_open_boundaries": (_I, [_P, C.POINTER(DFSPHSurface), C.POINTER(DFSPHSurface),
                                            C.POINTER(DFSPHSurface), _F32P, _U32, C.c_float, C.c_float, _I, _I]),
    "dfsph_cuda_set_divergence": (None, [_P, _I]),
    "dfsph_cuda_set_dynamic_window": (_I, [_P, C.POINTER(DFSPHSurface), C.POINTER(DFSPHSurface), _F32P, _U32, _I,
                                           _I, C.c_double, C.c_double, C.c_double]),
    "dfsph_cuda_dynamic_window_state": (_I, [_P, C.POINTER(C.c_double), C.POINTER(C.c_uint32),
                                             C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "dfsph_cuda_load_rest_fluid": (_I, [_P, C.POINTER(DFSPHSurface), _F32P, _U32, C.c_float, C.c_float, C.c_float,
                                        _U32, _U32P, _U32P, C.POINTER(C.c_double)]),
    "dfsph_cuda_handle": (C.c_void_p, [_P]),
    "dfsph_cuda_save": (_I, [_P, C.c_char_p]),
    "dfsph_cuda_load": (_I, [_P, C.c_char_p]),
    "dfsph_cuda_last_error": (C.c_char_p, []),
    "dfsph_cuda_save_text": (_I, [_P, C.c_char_p, _I, _I]),
    "dfsph_cuda_load_text": (_I, [_P, C.c_char_p, _I]),
    "dfsph_cuda_new_from_text": (C.c_void_p, [C.POINTER(DFSPHParams), C.c_char_p, _I]),
    "dfsph_scene_read": (C.c_void_p, [C.c_char_p, C.c_char_p]),
    "dfsph_scene_free": (None, [_P]),
    "dfsph_scene_params": (None, [_P, C.POINTER(DFSPHParams)]),
    "dfsph_scene_fluid_count": (_U32, [_P]),
    "dfsph_scene_fluid": (None, [_P, _F32P, _F32P]),
    "dfsph_scene_boundary_count": (_U32, [_P]),
    "dfsph_scene_boundary": (None, [_P, _F32P]),
    "dfsph_cuda_new_from_scene": (C.c_void_p, [_P, C.POINTER(DFSPHParams)]),
    "dfsph_scene_body_count": (_U32, [_P]),
    "dfsph_scene_body": (_U32, [_P, _U32, C.POINTER(DFSPHRigidBody), _F32P]),
    "dfsph_particles_write_text": (_I, [C.c_char_p, _F32P, _F32P, _F32P, _U32, _I, _I]),
    "dfsph_particles_read_text": (C.c_int64, [C.c_char_p, _F32P, _F32P, _F32P, _U32]),
}

_hip = None
_host = None


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def load_hip() -> C.CDLL:
    """Load libdfsph_hip.so (raises if it has not been built)."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise RuntimeError("libdfsph_hip.so is missing: run `python -m splishsplash_for_cuda_amd.build`")
        lib = C.CDLL(HIP_LIB)
        _bind(lib, HIP_SIGNATURES)
        _hip = lib
    return _hip


def load_host() -> C.CDLL:
    global _host
    if _host is None:
        load_hip()
        if not os.path.exists(HOST_LIB):
            raise RuntimeError("libdfsph_host.so is missing: run `python -m splishsplash_for_cuda_amd.build`")
        lib = C.CDLL(HOST_LIB)
        _bind(lib, HOST_SIGNATURES)
        _host = lib
    return _host


class DFSPHError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("DFSPH error %d: %s" % (code, msg))
        self.code = code


def _check(rc: int, what: str = ""):
    if rc != DFSPH_OK:
        msg = load_hip().dfsph_last_error().decode(errors="replace")
        raise DFSPHError(rc, (what + ": " if what else "") + msg)


def _f32(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a


def _ptr(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def default_params(**kw) -> DFSPHParams:
    p = DFSPHParams()
    load_hip().dfsph_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "gravity":
            p.gravity = (C.c_double * 3)(*v)
        else:
            setattr(p, k, v)
    return p


class Solver:
    """One DFSPH simulation on the GPU (a libdfsph_hip.so handle).

    Methods map one to one onto the C ABI, which mirrors the reference's
    cuda_* functions (DFSPH_core_cuda.h:21-83)."""

    def __init__(self, params: DFSPHParams | None = None, **kw):
        self.lib = load_hip()
        self.params = params if params is not None else default_params(**kw)
        h = C.c_void_p()
        _check(self.lib.dfsph_create(C.byref(self.params), C.byref(h)), "dfsph_create")
        self.h = h
        self._owns = True

    def close(self):
        if getattr(self, "h", None) and getattr(self, "_owns", False):
            self.lib.dfsph_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream):
        """dfsph_set_stream: queue the solver's work on a HIP stream (an int handle such as
        torch.cuda.Stream().cuda_stream; 0 or None: the handle's own stream)"""
        _check(self.lib.dfsph_set_stream(self.h, C.c_void_p(int(stream or 0))), "dfsph_set_stream")

    def get_stream(self) -> int:
        p = C.c_void_p()
        _check(self.lib.dfsph_get_stream(self.h, C.byref(p)), "dfsph_get_stream")
        return int(p.value or 0)

    # ---- data in ----
    def upload_boundary(self, xyz, psi=None):
        x = _f32(xyz).reshape(-1)
        p = _f32(psi)
        _check(self.lib.dfsph_upload_boundary(self.h, _ptr(x, C.c_float), _ptr(p, C.c_float), x.size // 3),
               "dfsph_upload_boundary")

    def upload_fluid(self, xyz, vel=None, mass=None, ids=None):
        x = _f32(xyz).reshape(-1)
        v = None if vel is None else _f32(vel).reshape(-1)
        m = _f32(mass)
        i = None if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
        _check(self.lib.dfsph_upload_fluid(self.h, _ptr(x, C.c_float), _ptr(v, C.c_float), _ptr(m, C.c_float),
                                           _ptr(i, C.c_uint32), x.size // 3), "dfsph_upload_fluid")

    def add_particles(self, xyz, vel=None, mass=None, ids=None):
        x = _f32(xyz).reshape(-1)
        v = None if vel is None else _f32(vel).reshape(-1)
        m = _f32(mass)
        i = None if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
        _check(self.lib.dfsph_add_particles(self.h, _ptr(x, C.c_float), _ptr(v, C.c_float), _ptr(m, C.c_float),
                                            _ptr(i, C.c_uint32), x.size // 3), "dfsph_add_particles")

    def upload_field(self, field, values):
        a = np.ascontiguousarray(values, dtype=np.uint32 if field in _UINT_FIELDS else np.float32).reshape(-1)
        _check(self.lib.dfsph_upload_field(self.h, field, a.ctypes.data_as(C.c_void_p), a.size), "dfsph_upload_field")

    # ---- the step's phases ----
    def neighbor_search(self, sort=True):
        _check(self.lib.dfsph_neighbor_search(self.h, 1 if sort else 0), "dfsph_neighbor_search")

    def compute_density_factor(self):
        _check(self.lib.dfsph_compute_density_factor(self.h), "dfsph_compute_density_factor")

    def divergence_solve(self, max_iter=None, max_error=None):
        it = C.c_uint32()
        err = C.c_double()
        _check(self.lib.dfsph_divergence_solve(self.h, self.params.max_iter_v if max_iter is None else max_iter,
                                               self.params.max_error_v if max_error is None else max_error,
                                               C.byref(it), C.byref(err)), "dfsph_divergence_solve")
        return it.value, err.value

    def external_forces(self):
        _check(self.lib.dfsph_external_forces(self.h), "dfsph_external_forces")

    def cfl(self, min_dt=None, cfl_factor=None, max_dt=None, wait=True):
        """CFL step size, applied on the device; returned when wait (a host
        round trip), else left there for get_time()."""
        out = C.c_double()
        p = self.params
        _check(self.lib.dfsph_cfl(self.h, p.dt_min if min_dt is None else min_dt,
                                  p.cfl_factor if cfl_factor is None else cfl_factor,
                                  p.dt_max if max_dt is None else max_dt, C.byref(out) if wait else None),
               "dfsph_cfl")
        return out.value if wait else None

    def get_time(self):
        """(h, h_future, time) as the device holds them."""
        a, b, t = C.c_double(), C.c_double(), C.c_double()
        _check(self.lib.dfsph_get_time(self.h, C.byref(a), C.byref(b), C.byref(t)), "dfsph_get_time")
        return a.value, b.value, t.value

    def fluid_count(self):
        n = C.c_uint32()
        _check(self.lib.dfsph_fluid_count(self.h, C.byref(n)), "dfsph_fluid_count")
        return n.value

    def update_time_step(self, h_future):
        _check(self.lib.dfsph_update_time_step(self.h, h_future), "dfsph_update_time_step")

    def update_vel(self):
        _check(self.lib.dfsph_update_vel(self.h), "dfsph_update_vel")

    def pressure_solve(self, max_iter=None, max_error=None):
        it = C.c_uint32()
        err = C.c_double()
        _check(self.lib.dfsph_pressure_solve(self.h, self.params.max_iter if max_iter is None else max_iter,
                                             self.params.max_error if max_error is None else max_error,
                                             C.byref(it), C.byref(err)), "dfsph_pressure_solve")
        return it.value, err.value

    def update_pos(self):
        _check(self.lib.dfsph_update_pos(self.h), "dfsph_update_pos")

    def step_end(self):
        _check(self.lib.dfsph_step_end(self.h), "dfsph_step_end")

    def set_time(self, h, h_past, time):
        _check(self.lib.dfsph_set_time(self.h, h, h_past, time), "dfsph_set_time")

    def check_particles_positions(self, mode=0):
        c = C.c_uint32()
        _check(self.lib.dfsph_check_particles_positions(self.h, mode, C.byref(c)), "dfsph_check_particles_positions")
        return c.value

    def remove_particles(self, keep):
        k = np.ascontiguousarray(keep, dtype=np.uint8)
        n = C.c_uint32()
        _check(self.lib.dfsph_remove_particles(self.h, _ptr(k, C.c_uint8), C.byref(n)), "dfsph_remove_particles")
        return n.value

    # ---- dynamic window / open boundary ----
    def set_restriction(self, mode, count_active=0, count_active_neighbors=0, true_particle_count=-1):
        _check(self.lib.dfsph_set_restriction(self.h, mode, count_active, count_active_neighbors,
                                              true_particle_count), "dfsph_set_restriction")

    def tag_neighbors(self, tag_in, tag_out, untagged_only=True):
        c = C.c_uint32()
        _check(self.lib.dfsph_tag_neighbors(self.h, tag_in, tag_out, 1 if untagged_only else 0, C.byref(c)),
               "dfsph_tag_neighbors")
        return c.value

    def remove_density_peaks(self, limit, boundary_limit):
        """dfsph_remove_density_peaks: one extraction pass; returns (removed, mean density of the kept)"""
        c, a = C.c_uint32(), C.c_double()
        _check(self.lib.dfsph_remove_density_peaks(self.h, float(limit), float(boundary_limit), C.byref(c),
                                                   C.byref(a)), "dfsph_remove_density_peaks")
        return c.value, a.value

    def get_params(self) -> "DFSPHParams":
        p = DFSPHParams()
        _check(self.lib.dfsph_get_params(self.h, C.byref(p)), "dfsph_get_params")
        return p

    def count_tagged(self, tag):
        c = C.c_uint32()
        _check(self.lib.dfsph_count_tagged(self.h, tag, C.byref(c)), "dfsph_count_tagged")
        return c.value

    def order_active_first(self):
        a, b = C.c_uint32(), C.c_uint32()
        _check(self.lib.dfsph_order_active_first(self.h, C.byref(a), C.byref(b)), "dfsph_order_active_first")
        return a.value, b.value

    def remove_tagged(self, tag):
        c = C.c_uint32()
        _check(self.lib.dfsph_remove_tagged(self.h, tag, C.byref(c)), "dfsph_remove_tagged")
        return c.value

    def mesh_upload(self, vertices, faces) -> int:
        """dfsph_mesh_upload: a closed triangle mesh for 'mesh' surfaces; returns its id"""
        v = _f32(vertices).reshape(-1)
        f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1)
        m = C.c_uint32()
        _check(self.lib.dfsph_mesh_upload(self.h, _ptr(v, C.c_float), v.size // 3, _ptr(f, C.c_uint32), f.size // 3,
                                          C.byref(m)), "dfsph_mesh_upload")
        return m.value

    def surface_inside(self, surf, xyz):
        x = _f32(xyz).reshape(-1)
        n = x.size // 3
        out = np.zeros(n, np.uint8)
        _check(self.lib.dfsph_surface_inside(self.h, C.byref(surf), _ptr(x, C.c_float), n, _ptr(out, C.c_uint8)),
               "dfsph_surface_inside")
        return out.astype(bool)

    def open_boundary_init(self, boundary, interior, fluid_surface, inflow_xyz):
        x = _f32(inflow_xyz).reshape(-1)
        k = C.c_uint32()
        _check(self.lib.dfsph_open_boundary_init(self.h, C.byref(boundary), C.byref(interior), C.byref(fluid_surface),
                                                 _ptr(x, C.c_float), x.size // 3, C.byref(k)),
               "dfsph_open_boundary_init")
        return k.value

    def open_boundary_apply(self, allowed_new_distance, allowed_new_density, use_inflow=True, use_outflow=True,
                            window_displacement=None):
        a, r = C.c_uint32(), C.c_uint32()
        d = None if window_displacement is None else (C.c_double * 3)(*window_displacement)
        _check(self.lib.dfsph_open_boundary_apply(self.h, d, allowed_new_distance, allowed_new_density,
                                                  1 if use_inflow else 0, 1 if use_outflow else 0,
                                                  C.byref(a), C.byref(r)), "dfsph_open_boundary_apply")
        return a.value, r.value

    # ---- rigid bodies ----
    def add_rigid_body(self, body: "DFSPHRigidBody", xyz_body, psi=None) -> int:
        x = _f32(xyz_body).reshape(-1)
        p = None if psi is None else _f32(psi)
        i = C.c_uint32()
        _check(self.lib.dfsph_add_rigid_body(self.h, C.byref(body), _ptr(x, C.c_float), _ptr(p, C.c_float),
                                             x.size // 3, C.byref(i)), "dfsph_add_rigid_body")
        return i.value

    def set_rigid_body(self, i: int, body: "DFSPHRigidBody"):
        _check(self.lib.dfsph_set_rigid_body(self.h, i, C.byref(body)), "dfsph_set_rigid_body")

    def rigid_body(self, i: int):
        """(DFSPHRigidBody state, force[3], torque[3]) after the last step"""
        b = DFSPHRigidBody()
        F = (C.c_double * 3)()
        T = (C.c_double * 3)()
        _check(self.lib.dfsph_get_rigid_body(self.h, i, C.byref(b), F, T), "dfsph_get_rigid_body")
        return b, np.array(F[:]), np.array(T[:])

    def rigid_body_forces(self, i: int, n: int):
        out = np.zeros((n, 3), np.float32)
        _check(self.lib.dfsph_rigid_body_forces(self.h, i, _ptr(out, C.c_float), 3 * n), "dfsph_rigid_body_forces")
        return out

    def rigid_body_count(self) -> int:
        n = C.c_uint32()
        _check(self.lib.dfsph_rigid_body_count(self.h, C.byref(n)), "dfsph_rigid_body_count")
        return n.value

    def step(self, raise_on_overflow=True):
        st = DFSPHStepStats()
        rc = self.lib.dfsph_step(self.h, C.byref(st))
        if rc != DFSPH_OK and (raise_on_overflow or rc != DFSPH_ERR_NEIGHBOR_OVERFLOW):
            _check(rc, "dfsph_step")
        return st

    def synchronize(self):
        _check(self.lib.dfsph_synchronize(self.h), "dfsph_synchronize")

    # ---- learned warm start model ----
    def mlp_set_params(self, params, norm=None):
        w = np.ascontiguousarray(params, np.float32)
        assert w.size == MLP_NPARAM
        nr = None if norm is None else np.ascontiguousarray(norm, np.float32)
        _check(self.lib.dfsph_mlp_set_params(self.h, _ptr(w, C.c_float), _ptr(nr, C.c_float)), "dfsph_mlp_set_params")

    def mlp_get_params(self):
        w = np.zeros(MLP_NPARAM, np.float32)
        nr = np.zeros(MLP_NNORM, np.float32)
        _check(self.lib.dfsph_mlp_get_params(self.h, _ptr(w, C.c_float), _ptr(nr, C.c_float)), "dfsph_mlp_get_params")
        return w, nr

    def mlp_forward(self, x, normalise=True):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_NFEAT)
        y = np.zeros((x.shape[0], 2), np.float32)
        _check(self.lib.dfsph_mlp_forward(self.h, _ptr(x, C.c_float), x.shape[0], _ptr(y, C.c_float),
                                          1 if normalise else 0), "dfsph_mlp_forward")
        return y

    def mlp_forward_noalloc(self, x, normalise=True):
        """dfsph_mlp_forward_noalloc: the same outputs, staged in the library's
        static device array (no device allocation)"""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_NFEAT)
        y = np.zeros((x.shape[0], 2), np.float32)
        _check(self.lib.dfsph_mlp_forward_noalloc(self.h, _ptr(x, C.c_float), x.shape[0], _ptr(y, C.c_float),
                                                  1 if normalise else 0), "dfsph_mlp_forward_noalloc")
        return y

    def mlp_loss_grad(self, x, t, normalise=False):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_NFEAT)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 2)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = C.c_double()
        _check(self.lib.dfsph_mlp_loss_grad(self.h, _ptr(x, C.c_float), _ptr(t, C.c_float), x.shape[0],
                                            _ptr(g, C.c_float), C.byref(loss), 1 if normalise else 0),
               "dfsph_mlp_loss_grad")
        return loss.value, g

    def mlp_harvest(self, capacity, stride=1):
        _check(self.lib.dfsph_mlp_harvest(self.h, capacity, stride), "dfsph_mlp_harvest")

    def mlp_dataset(self):
        n = C.c_uint32()
        _check(self.lib.dfsph_mlp_dataset_size(self.h, C.byref(n)), "dfsph_mlp_dataset_size")
        x = np.zeros((n.value, MLP_NFEAT), np.float32)
        t = np.zeros((n.value, 2), np.float32)
        if n.value:
            _check(self.lib.dfsph_mlp_dataset_download(self.h, _ptr(x, C.c_float), _ptr(t, C.c_float), n.value),
                   "dfsph_mlp_dataset_download")
        return x, t

    def mlp_dataset_upload(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_NFEAT)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 2)
        _check(self.lib.dfsph_mlp_dataset_upload(self.h, _ptr(x, C.c_float), _ptr(t, C.c_float), x.shape[0]),
               "dfsph_mlp_dataset_upload")

    def mlp_fit_normalization(self):
        _check(self.lib.dfsph_mlp_fit_normalization(self.h), "dfsph_mlp_fit_normalization")

    def mlp_train(self, steps, batch=4096, lr=2e-3, seed=0):
        a = C.c_float()
        b = C.c_float()
        _check(self.lib.dfsph_mlp_train(self.h, steps, batch, lr, seed, C.byref(a), C.byref(b)), "dfsph_mlp_train")
        return a.value, b.value

    def mlp_step_features(self):
        n = self.n
        f = np.zeros((n, MLP_NFEAT), np.float32)
        pv = np.zeros(n, np.float32)
        pp = np.zeros(n, np.float32)
        _check(self.lib.dfsph_mlp_step_features(self.h, _ptr(f, C.c_float), _ptr(pv, C.c_float), _ptr(pp, C.c_float),
                                                n), "dfsph_mlp_step_features")
        return f, pv, pp

    # ---- per-kernel event timing ----
    def profile_enable(self, mask=(1 << len(KERNEL_NAMES)) - 1):
        _check(self.lib.dfsph_profile_enable(self.h, mask), "dfsph_profile_enable")

    def profile_reset(self):
        _check(self.lib.dfsph_profile_reset(self.h), "dfsph_profile_reset")

    def profile(self):
        """{kernel name: (launches, total ms)}"""
        out = {}
        for k, name in enumerate(KERNEL_NAMES):
            n = C.c_uint32()
            ms = C.c_double()
            _check(self.lib.dfsph_profile_read(self.h, k, C.byref(n), C.byref(ms)), "dfsph_profile_read")
            out[name] = (n.value, ms.value)
        return out

    # ---- data out ----
    def load_fluid(self, xyz, min_distance, max_density, interpolate_velocity=False):
        """dfsph_load_fluid: add the candidates far enough from the fluid; returns the count added"""
        x = _f32(xyz).reshape(-1)
        a = C.c_uint32()
        _check(self.lib.dfsph_load_fluid(self.h, _ptr(x, C.c_float), x.size // 3, float(min_distance),
                                         float(max_density), 1 if interpolate_velocity else 0, C.byref(a)),
               "dfsph_load_fluid")
        return a.value

    def change_max_particle_number(self, max_particles):
        _check(self.lib.dfsph_change_max_particle_number(self.h, int(max_particles)),
               "dfsph_change_max_particle_number")

    def neighbor_stats(self):
        """(max_seen, overflow, fluid_total, boundary_total) of the last neighbor search"""
        ms, of = C.c_uint32(), C.c_uint32()
        nf, nb = C.c_uint64(), C.c_uint64()
        _check(self.lib.dfsph_neighbor_stats(self.h, C.byref(ms), C.byref(of), C.byref(nf), C.byref(nb)),
               "dfsph_neighbor_stats")
        return ms.value, of.value, nf.value, nb.value

    def info(self) -> DFSPHInfo:
        inf = DFSPHInfo()
        _check(self.lib.dfsph_get_info(self.h, C.byref(inf)), "dfsph_get_info")
        return inf

    @property
    def n(self) -> int:
        return self.info().n_fluid

    def download(self, field):
        inf = self.info()
        n = inf.n_boundary if field in (FIELD_BOUNDARY_POS, FIELD_BOUNDARY_PSI) else inf.n_fluid
        cnt = 3 * n if field in _VECTOR_FIELDS else n
        out = np.empty(cnt, dtype=np.uint32 if field in _UINT_FIELDS else np.float32)
        _check(self.lib.dfsph_download(self.h, field, out.ctypes.data_as(C.c_void_p), cnt), "dfsph_download")
        return out.reshape(-1, 3) if field in _VECTOR_FIELDS else out

    def neighbors(self):
        """(n_fluid, n_boundary, list[n, max]) in device order."""
        inf = self.info()
        n, mx = inf.n_fluid, inf.max_neighbors
        nf = np.zeros(n, np.uint32)
        nb = np.zeros(n, np.uint32)
        L = np.zeros((n, mx), np.uint32)
        _check(self.lib.dfsph_get_neighbors(self.h, _ptr(nf, C.c_uint32), _ptr(nb, C.c_uint32),
                                            _ptr(L, C.c_uint32), mx), "dfsph_get_neighbors")
        return nf, nb, L

    def state_by_id(self):
        """Fluid fields reordered by particle id (dict of numpy arrays)."""
        ids = self.download(FIELD_ID)
        order = np.argsort(ids, kind="stable")
        out = {"id": ids[order]}
        for name, f in (("pos", FIELD_POS), ("vel", FIELD_VEL), ("mass", FIELD_MASS), ("density", FIELD_DENSITY),
                        ("factor", FIELD_FACTOR), ("density_adv", FIELD_DENSITY_ADV), ("kappa", FIELD_KAPPA),
                        ("kappaV", FIELD_KAPPAV), ("acc", FIELD_ACC)):
            out[name] = self.download(f)[order]
        return out


class HostSolver:
    """SPH::DFSPHCUDA (the C++ host class) through libdfsph_host.so."""

    def __init__(self, params: DFSPHParams, boundary_xyz, fluid_xyz, fluid_vel=None):
        self.lib = load_host()
        b = _f32(boundary_xyz).reshape(-1)
        f = _f32(fluid_xyz).reshape(-1)
        v = None if fluid_vel is None else _f32(fluid_vel).reshape(-1)
        err = C.create_string_buffer(512)
        s = self.lib.dfsph_cuda_new(C.byref(params), _ptr(b, C.c_float), b.size // 3, _ptr(f, C.c_float),
                                    _ptr(v, C.c_float), f.size // 3, err, 512)
        if not s:
            raise DFSPHError(DFSPH_ERR_STATE, err.value.decode(errors="replace"))
        self._adopt(s, params)

    def _adopt(self, s, params):
        self.s = s
        self.params = params
        # a Solver view over the handle owned by the C++ object (not owning)
        self.view = Solver.__new__(Solver)
        self.view.lib = load_hip()
        self.view.params = params
        self.view.h = C.c_void_p(self.lib.dfsph_cuda_handle(s))
        self.view._owns = False

    @classmethod
    def _wrap(cls, s, params):
        lib = load_host()
        if not s:
            raise DFSPHError(DFSPH_ERR_STATE, lib.dfsph_cuda_last_error().decode(errors="replace"))
        self = cls.__new__(cls)
        self.lib = lib
        self._adopt(s, params)
        return self

    @classmethod
    def from_scene(cls, scene: "Scene", params: DFSPHParams = None) -> "HostSolver":
        """dfsph_cuda_new_from_scene: the scene's fluid blocks and rigid bodies
        (params None: the scene's Configuration)."""
        p = scene.params() if params is None else params
        return cls._wrap(load_host().dfsph_cuda_new_from_scene(scene.h, C.byref(p)), p)

    @classmethod
    def from_text(cls, params: DFSPHParams, folder: str, load_velocities: bool = True) -> "HostSolver":
        """dfsph_cuda_new_from_text: boundary, fluid and h from the text files of folder."""
        s = load_host().dfsph_cuda_new_from_text(C.byref(params), folder.encode(), 1 if load_velocities else 0)
        return cls._wrap(s, params)

    def save_text(self, folder: str, liquid: bool = True, boundaries: bool = True):
        rc = self.lib.dfsph_cuda_save_text(self.s, folder.encode(), 1 if liquid else 0, 1 if boundaries else 0)
        if rc != DFSPH_OK:
            raise DFSPHError(rc, self.lib.dfsph_cuda_last_error().decode(errors="replace"))

    def load_text(self, folder: str, load_velocities: bool = True):
        rc = self.lib.dfsph_cuda_load_text(self.s, folder.encode(), 1 if load_velocities else 0)
        if rc != DFSPH_OK:
            raise DFSPHError(rc, self.lib.dfsph_cuda_last_error().decode(errors="replace"))

    def step(self, raise_on_overflow=True) -> DFSPHCudaStats:
        st = DFSPHCudaStats()
        rc = self.lib.dfsph_cuda_step(self.s, C.byref(st))
        if rc != DFSPH_OK and (raise_on_overflow or rc != DFSPH_ERR_NEIGHBOR_OVERFLOW):
            raise DFSPHError(rc, self.lib.dfsph_cuda_last_error().decode(errors="replace"))
        return st

    def save(self, path):
        rc = self.lib.dfsph_cuda_save(self.s, path.encode())
        if rc != DFSPH_OK:
            raise DFSPHError(rc, self.lib.dfsph_cuda_last_error().decode(errors="replace"))

    def load(self, path):
        rc = self.lib.dfsph_cuda_load(self.s, path.encode())
        if rc != DFSPH_OK:
            raise DFSPHError(rc, self.lib.dfsph_cuda_last_error().decode(errors="replace"))

    def set_open_boundaries(self, boundary, interior, fluid_surface, inflow_xyz, allowed_new_distance=0.0,
                            allowed_new_density=700.0, use_inflow=True, use_outflow=True):
        c = _f32(inflow_xyz).reshape(-1)
        rc = self.lib.dfsph_cuda_set_open_boundaries(self.s, C.byref(boundary), C.byref(interior),
                                                     C.byref(fluid_surface), _ptr(c, C.c_float), c.size // 3,
                                                     allowed_new_distance, allowed_new_density,
                                                     1 if use_inflow else 0, 1 if use_outflow else 0)
        if rc != DFSPH_OK:
            raise DFSPHError(rc, self.lib.dfsph_cuda_last_error().decode(errors="replace"))

    def set_dynamic_window(self, simulation, fluid, background_xyz, tracked_body=-1, wall_body=-1, trigger=0.0,
                           max_displacement=0.0, band=0.0):
        """dfsph_cuda_set_dynamic_window: the simulated region follows rigid body tracked_body, its walls
        (kinematic body wall_body) with it"""
        b = _f32(background_xyz).reshape(-1)
        rc = self.lib.dfsph_cuda_set_dynamic_window(self.s, C.byref(simulation), C.byref(fluid), _ptr(b, C.c_float),
                                                    b.size // 3, int(tracked_body), int(wall_body), float(trigger),
                                                    float(max_displacement), float(band))
        if rc != DFSPH_OK:
            raise DFSPHError(rc, self.lib.dfsph_cuda_last_error().decode(errors="replace"))

    def window_state(self):
        """(total displacement[3], moves, loaded and removed by the last move)"""
        d = (C.c_double * 3)()
        m, l, r = C.c_uint32(), C.c_uint32(), C.c_uint32()
        rc = self.lib.dfsph_cuda_dynamic_window_state(self.s, d, C.byref(m), C.byref(l), C.byref(r))
        if rc != DFSPH_OK:
            raise DFSPHError(rc, self.lib.dfsph_cuda_last_error().decode(errors="replace"))
        return np.array(d[:]), m.value, l.value, r.value

    def load_rest_fluid(self, fluid, background_xyz, density_start=0.0, density_end=0.0, step_density=0.0,
                        stab_iters=0):
        """dfsph_cuda_load_rest_fluid: the fluid becomes the background inside `fluid`, extracted and
        settled at rest; returns (particles, extraction passes, mean density after the last pass)"""
        b = _f32(background_xyz).reshape(-1)
        n, k, a = C.c_uint32(), C.c_uint32(), C.c_double()
        rc = self.lib.dfsph_cuda_load_rest_fluid(self.s, C.byref(fluid), _ptr(b, C.c_float), b.size // 3,
                                                 float(density_start), float(density_end), float(step_density),
                                                 int(stab_iters), C.byref(n), C.byref(k), C.byref(a))
        if rc != DFSPH_OK:
            raise DFSPHError(rc, self.lib.dfsph_cuda_last_error().decode(errors="replace"))
        return n.value, k.value, a.value

    def set_divergence(self, enable: bool):
        self.lib.dfsph_cuda_set_divergence(self.s, 1 if enable else 0)

    def close(self):
        if getattr(self, "s", None):
            self.view.h = None
            self.lib.dfsph_cuda_delete(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """A SPlisHSPlasH scene file read by the C++ SceneLoader (dfsph_scene_read)."""

    def __init__(self, path: str, base_path: str = None):
        self.lib = load_host()
        self.h = self.lib.dfsph_scene_read(path.encode(), None if base_path is None else base_path.encode())
        if not self.h:
            raise DFSPHError(DFSPH_ERR_ARG, self.lib.dfsph_cuda_last_error().decode(errors="replace"))

    def params(self) -> DFSPHParams:
        p = DFSPHParams()
        self.lib.dfsph_scene_params(self.h, C.byref(p))
        return p

    def fluid(self):
        n = self.lib.dfsph_scene_fluid_count(self.h)
        x = np.zeros((n, 3), np.float32)
        v = np.zeros((n, 3), np.float32)
        self.lib.dfsph_scene_fluid(self.h, _ptr(x, C.c_float), _ptr(v, C.c_float))
        return x, v

    def boundary(self):
        n = self.lib.dfsph_scene_boundary_count(self.h)
        x = np.zeros((n, 3), np.float32)
        self.lib.dfsph_scene_boundary(self.h, _ptr(x, C.c_float))
        return x

    def bodies(self):
        """[(DFSPHRigidBody, particles (n, 3) in the body frame)] of the dynamic rigid bodies"""
        out = []
        for i in range(self.lib.dfsph_scene_body_count(self.h)):
            b = DFSPHRigidBody()
            n = self.lib.dfsph_scene_body(self.h, i, C.byref(b), None)
            x = np.zeros((n, 3), np.float32)
            self.lib.dfsph_scene_body(self.h, i, None, _ptr(x, C.c_float))
            out.append((b, x))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.dfsph_scene_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_particles_text(path: str, xyz, mass, vel=None, has_factor_computation=True,
                         velocity_impacted_by_fluid_solver=True):
    """UnifiedParticleSet::write_to_file for one set (dfsph_particles_write_text)."""
    lib = load_host()
    x = _f32(xyz).reshape(-1)
    m = _f32(mass).reshape(-1)
    v = None if vel is None else _f32(vel).reshape(-1)
    if x.size != 3 * m.size or (v is not None and v.size != x.size):
        raise ValueError("xyz, vel must hold 3 floats per mass")
    rc = lib.dfsph_particles_write_text(path.encode(), _ptr(x, C.c_float), _ptr(v, C.c_float), _ptr(m, C.c_float),
                                        m.size, int(bool(has_factor_computation)),
                                        int(bool(velocity_impacted_by_fluid_solver)))
    if rc != DFSPH_OK:
        raise DFSPHError(rc, lib.dfsph_cuda_last_error().decode(errors="replace"))


def read_particles_text(path: str):
    """UnifiedParticleSet::load_from_file for one set: (xyz, vel, mass)."""
    lib = load_host()
    n = lib.dfsph_particles_read_text(path.encode(), None, None, None, 0)
    if n < 0:
        raise DFSPHError(DFSPH_ERR_ARG, lib.dfsph_cuda_last_error().decode(errors="replace"))
    x = np.zeros((n, 3), np.float32)
    v = np.zeros((n, 3), np.float32)
    m = np.zeros(n, np.float32)
    if n:
        got = lib.dfsph_particles_read_text(path.encode(), _ptr(x, C.c_float), _ptr(v, C.c_float),
                                            _ptr(m, C.c_float), n)
        if got != n:
            raise DFSPHError(DFSPH_ERR_ARG, lib.dfsph_cuda_last_error().decode(errors="replace"))
    return x, v, m
