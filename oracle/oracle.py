"""ctypes loader of the CPU oracle (liboracle.so, built by oracle/Makefile).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Never used by the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB: another build of the same library (the sanitizer build of oracle/sanitize.sh)
LIB = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

MODE_GPU = 0
MODE_CPU = 1
F_POS, F_VEL, F_MASS, F_DENSITY, F_FACTOR, F_DENSITY_ADV, F_KAPPA, F_KAPPAV, F_ACC, F_ID = range(10)
_VEC = {F_POS, F_VEL, F_ACC}

MLP_SIZES = [(64, 12), (64,), (64, 64), (64,), (64, 64), (64,), (2, 64), (2,)]
MLP_NPARAM = sum(int(np.prod(s)) for s in MLP_SIZES)


class OrParams(C.Structure):
    _fields_ = [
        ("particle_radius", C.c_double), ("density0", C.c_double), ("viscosity", C.c_double),
        ("gravity", C.c_double * 3), ("dt", C.c_double), ("mode", C.c_int), ("kernel_table", C.c_int),
        ("fp32_predicate", C.c_int), ("enable_divergence", C.c_int),
        ("max_iter_v", C.c_uint32), ("max_iter", C.c_uint32), ("min_iter_v", C.c_uint32), ("min_iter", C.c_uint32),
        ("max_error_v", C.c_double), ("max_error", C.c_double), ("use_cfl", C.c_int),
        ("cfl_factor", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double), ("cfl_seed", C.c_double),
        ("remove_escaped", C.c_int), ("max_neighbors", C.c_uint32),
    ]


class OrSurface(C.Structure):
    _fields_ = [("type", C.c_int), ("reversed", C.c_int), ("o", C.c_double * 3), ("n", C.c_double * 3),
                ("center", C.c_double * 3), ("half", C.c_double * 3), ("radius", C.c_double),
                ("radius_internal", C.c_double), ("points", C.c_uint32), ("dir", C.c_double * 3),
                ("tri", C.POINTER(C.c_double)), ("ntri", C.c_uint32)]


SURFACE_TYPES = {"plane": 0, "cuboid": 1, "cylinder": 2, "mesh": 3, "sphere": 4, "star": 5}


def surface(kind, reversed=False, o=(0, 0, 0), n=(0, 1, 0), center=(0, 0, 0), half=(0, 0, 0), radius=0.0,
            radius_internal=0.0, points=0, direction=(0, 0, 1), vertices=None, faces=None):
    """an or_surface; 'mesh' takes vertices (nv, 3) and faces (nf, 3), moved by center"""
    s = OrSurface()
    if kind == "mesh":
        v = np.asarray(vertices, np.float64).reshape(-1, 3)
        f = np.asarray(faces, np.int64).reshape(-1, 3)
        tri = np.ascontiguousarray(v[f].reshape(-1))
        s._tri = tri                       # kept alive with the structure
        s.tri = tri.ctypes.data_as(C.POINTER(C.c_double))
        s.ntri = len(f)
    s.type = SURFACE_TYPES[kind]
    s.reversed = 1 if reversed else 0
    s.o = (C.c_double * 3)(*o)
    s.n = (C.c_double * 3)(*n)
    s.center = (C.c_double * 3)(*center)
    s.half = (C.c_double * 3)(*half)
    s.radius = radius
    s.radius_internal = radius_internal
    s.points = points
    s.dir = (C.c_double * 3)(*direction)
    return s


class OrStats(C.Structure):
    _fields_ = [("iters_v", C.c_uint32), ("iters", C.c_uint32), ("err_v", C.c_double), ("err", C.c_double),
                ("dt_used", C.c_double), ("removed", C.c_uint32), ("n_fluid", C.c_uint32),
                ("max_nbr", C.c_uint32), ("overflow", C.c_uint32)]


_lib = None
_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)
_UP = C.POINTER(C.c_uint32)


class OrBodyDesc(C.Structure):
    _fields_ = [("type", C.c_int), ("mass", C.c_double), ("inertia", C.c_double * 3), ("position", C.c_double * 3),
                ("velocity", C.c_double * 3), ("rotation", C.c_double * 4), ("angular_velocity", C.c_double * 3),
                ("axis", C.c_double * 3), ("amplitude", C.c_double), ("period", C.c_double)]


def body_desc(type=0, mass=0.0, inertia=(0, 0, 0), position=(0, 0, 0), velocity=(0, 0, 0), rotation=(0, 0, 0, 1),
              angular_velocity=(0, 0, 0), axis=(0, 0, 0), amplitude=0.0, period=0.0) -> OrBodyDesc:
    d = OrBodyDesc()
    d.type = type
    d.mass = mass
    d.inertia = (C.c_double * 3)(*inertia)
    d.position = (C.c_double * 3)(*position)
    d.velocity = (C.c_double * 3)(*velocity)
    d.rotation = (C.c_double * 4)(*rotation)
    d.angular_velocity = (C.c_double * 3)(*angular_velocity)
    d.axis = (C.c_double * 3)(*axis)
    d.amplitude = amplitude
    d.period = period
    return d


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        sig = {
            "or_default_params": (None, [C.POINTER(OrParams)]),
            "or_create": (C.c_void_p, [C.POINTER(OrParams)]),
            "or_destroy": (None, [C.c_void_p]),
            "or_set_fluid": (C.c_int, [C.c_void_p, C.c_uint32, _DP, _DP, _DP, _UP]),
            "or_set_boundary": (C.c_int, [C.c_void_p, C.c_uint32, _DP, _DP]),
            "or_get_boundary_psi": (C.c_int, [C.c_void_p, _DP]),
            "or_num_fluid": (C.c_uint32, [C.c_void_p]),
            "or_get": (C.c_int, [C.c_void_p, C.c_int, _DP]),
            "or_set": (C.c_int, [C.c_void_p, C.c_int, _DP]),
            "or_get_ids": (C.c_int, [C.c_void_p, _UP]),
            "or_get_time": (None, [C.c_void_p, _DP, _DP, _DP, _DP]),
            "or_set_time_step": (None, [C.c_void_p, C.c_double]),
            "or_neighbor_search": (C.c_int, [C.c_void_p]),
            "or_divergence_solve": (C.c_uint32, [C.c_void_p, _DP]),
            "or_external_forces": (None, [C.c_void_p]),
            "or_cfl": (C.c_double, [C.c_void_p]),
            "or_update_time_step": (None, [C.c_void_p, C.c_double]),
            "or_update_vel": (None, [C.c_void_p]),
            "or_pressure_solve": (C.c_uint32, [C.c_void_p, _DP]),
            "or_update_pos": (None, [C.c_void_p]),
            "or_step_end": (None, [C.c_void_p]),
            "or_check_positions": (C.c_uint32, [C.c_void_p, C.c_int]),
            "or_step": (C.c_int, [C.c_void_p, C.POINTER(OrStats)]),
            "or_get_neighbors": (C.c_int, [C.c_void_p, _UP, _UP, _UP]),
            "or_neighbors_fp32": (C.c_uint32, [C.c_uint32, _FP, C.c_uint32, _FP, C.c_float, C.c_int, C.c_uint32,
                                               _UP, _UP, _UP]),
            "or_kernel_W": (C.c_double, [C.c_double, C.c_double, C.c_int]),
            "or_kernel_gradW": (None, [_DP, C.c_double, C.c_int, _DP]),
            "or_mlp_forward": (None, [C.c_uint32, _DP, _DP, _DP]),
            "or_features": (None, [C.c_void_p, _DP]),
            "or_warm_start_targets": (None, [C.c_void_p, _DP]),
            "or_mlp_loss_grad": (C.c_double, [C.c_uint32, _DP, _DP, _DP, _DP]),
            "or_set_restriction": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_longlong, _UP]),
            "or_surface_inside": (C.c_int, [C.POINTER(OrSurface), _DP, _DP]),
            "or_inflow": (C.c_uint32, [C.c_void_p, C.POINTER(OrSurface), C.c_uint32, _DP, C.c_double, C.c_double,
                                       C.c_double, C.POINTER(C.c_uint8), _DP, _DP, _DP]),
            "or_outflow": (C.c_uint32, [C.c_void_p, C.POINTER(OrSurface), C.POINTER(OrSurface),
                                        C.POINTER(C.c_uint8)]),
            "or_rest_extract": (C.c_uint32, [C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_uint8),
                                             C.POINTER(C.c_uint8), _DP, _DP]),
            "or_add_body": (C.c_int, [C.c_void_p, C.POINTER(OrBodyDesc), C.c_uint32, _DP, _DP]),
            "or_set_body": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(OrBodyDesc)]),
            "or_get_body": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(OrBodyDesc), _DP, _DP]),
            "or_body_forces": (C.c_int, [C.c_void_p, C.c_uint32, _DP]),
            "or_body_psi": (C.c_int, [C.c_void_p, C.c_uint32, _DP]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def params(**kw) -> OrParams:
    p = OrParams()
    lib().or_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "gravity":
            p.gravity = (C.c_double * 3)(*v)
        else:
            setattr(p, k, v)
    return p


class Oracle:
    """FP64 CPU restatement of the DFSPH step (see dfsph_oracle.h)."""

    def __init__(self, p: OrParams | None = None, **kw):
        self.L = lib()
        self.p = p if p is not None else params(**kw)
        self.s = self.L.or_create(C.byref(self.p))

    def __del__(self):
        try:
            if self.s:
                self.L.or_destroy(self.s)
                self.s = None
        except Exception:
            pass

    def set_boundary(self, xyz, psi=None):
        x = _d(xyz).reshape(-1)
        p = None if psi is None else _d(psi)
        self.L.or_set_boundary(self.s, x.size // 3, x.ctypes.data_as(_DP), None if p is None else p.ctypes.data_as(_DP))

    def boundary_psi(self, n):
        out = np.zeros(n)
        self.L.or_get_boundary_psi(self.s, out.ctypes.data_as(_DP))
        return out

    def set_fluid(self, xyz, vel=None, mass=None, ids=None):
        x = _d(xyz).reshape(-1)
        v = None if vel is None else _d(vel).reshape(-1)
        m = None if mass is None else _d(mass)
        i = None if ids is None else np.ascontiguousarray(ids, np.uint32)
        self.L.or_set_fluid(self.s, x.size // 3, x.ctypes.data_as(_DP), None if v is None else v.ctypes.data_as(_DP),
                            None if m is None else m.ctypes.data_as(_DP), None if i is None else i.ctypes.data_as(_UP))

    @property
    def n(self):
        return self.L.or_num_fluid(self.s)

    def get(self, field):
        n = self.n
        out = np.zeros(3 * n if field in _VEC else n)
        self.L.or_get(self.s, field, out.ctypes.data_as(_DP))
        return out.reshape(-1, 3) if field in _VEC else out

    def set(self, field, values):
        a = _d(values).reshape(-1)
        self.L.or_set(self.s, field, a.ctypes.data_as(_DP))

    def ids(self):
        out = np.zeros(self.n, np.uint32)
        self.L.or_get_ids(self.s, out.ctypes.data_as(_UP))
        return out

    def time(self):
        v = [C.c_double() for _ in range(4)]
        self.L.or_get_time(self.s, *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def state(self):
        return {"id": self.ids(), "pos": self.get(F_POS), "vel": self.get(F_VEL), "mass": self.get(F_MASS),
                "density": self.get(F_DENSITY), "factor": self.get(F_FACTOR),
                "density_adv": self.get(F_DENSITY_ADV), "kappa": self.get(F_KAPPA), "kappaV": self.get(F_KAPPAV),
                "acc": self.get(F_ACC)}

    def neighbor_search(self):
        return self.L.or_neighbor_search(self.s)

    def neighbors(self):
        n = self.n
        mx = self.p.max_neighbors
        nf = np.zeros(n, np.uint32)
        nb = np.zeros(n, np.uint32)
        L = np.zeros((n, mx), np.uint32)
        self.L.or_get_neighbors(self.s, nf.ctypes.data_as(_UP), nb.ctypes.data_as(_UP), L.ctypes.data_as(_UP))
        return nf, nb, L

    def divergence_solve(self):
        e = C.c_double()
        it = self.L.or_divergence_solve(self.s, C.byref(e))
        return it, e.value

    def external_forces(self):
        self.L.or_external_forces(self.s)

    def cfl(self):
        return self.L.or_cfl(self.s)

    def update_time_step(self, hf):
        self.L.or_update_time_step(self.s, hf)

    def update_vel(self):
        self.L.or_update_vel(self.s)

    def pressure_solve(self):
        e = C.c_double()
        it = self.L.or_pressure_solve(self.s, C.byref(e))
        return it, e.value

    def update_pos(self):
        self.L.or_update_pos(self.s)

    def step_end(self):
        self.L.or_step_end(self.s)

    def check_positions(self, remove):
        return self.L.or_check_positions(self.s, 1 if remove else 0)

    def step(self):
        st = OrStats()
        self.L.or_step(self.s, C.byref(st))
        return st

    def features(self):
        """learned warm-start features (n, 12) of the current state"""
        out = np.zeros((self.n, 12))
        self.L.or_features(self.s, out.ctypes.data_as(_DP))
        return out

    def warm_start_targets(self):
        """training targets (n, 2) after a step's solves"""
        out = np.zeros((self.n, 2))
        self.L.or_warm_start_targets(self.s, out.ctypes.data_as(_DP))
        return out

    def add_body(self, desc, x_body, psi=None):
        x = _d(x_body).reshape(-1)
        p = None if psi is None else _d(psi)
        r = self.L.or_add_body(self.s, C.byref(desc), x.size // 3, x.ctypes.data_as(_DP),
                               None if p is None else p.ctypes.data_as(_DP))
        if r < 0:
            raise ValueError("invalid body")
        self._bodies = getattr(self, "_bodies", []) + [x.size // 3]
        return r

    def set_body(self, i, desc):
        if self.L.or_set_body(self.s, i, C.byref(desc)) != 0:
            raise ValueError("invalid body")

    def body(self, i):
        """(OrBodyDesc state, force[3], torque[3]) of body i after the last step"""
        d = OrBodyDesc()
        F = np.zeros(3)
        T = np.zeros(3)
        self.L.or_get_body(self.s, i, C.byref(d), F.ctypes.data_as(_DP), T.ctypes.data_as(_DP))
        return d, F, T

    def body_forces(self, i):
        out = np.zeros(3 * self._bodies[i])
        self.L.or_body_forces(self.s, i, out.ctypes.data_as(_DP))
        return out.reshape(-1, 3)

    def body_psi(self, i):
        out = np.zeros(self._bodies[i])
        self.L.or_body_psi(self.s, i, out.ctypes.data_as(_DP))
        return out

    def set_restriction(self, mode, count_active=0, count_active_neighbors=0, true_particle_count=-1, tags=None):
        t = None if tags is None else np.ascontiguousarray(tags, np.uint32)
        self.L.or_set_restriction(self.s, mode, count_active, count_active_neighbors, true_particle_count,
                                  None if t is None else t.ctypes.data_as(_UP))

    def inflow(self, boundary, cand, min_dist, max_density, mass):
        """(accept[nc] bool, vel[nc,3], nearest fluid distance[nc], density[nc])"""
        c = _d(cand).reshape(-1)
        nc = c.size // 3
        acc = np.zeros(nc, np.uint8)
        vel = np.zeros((nc, 3))
        dmin = np.zeros(nc)
        dens = np.zeros(nc)
        self.L.or_inflow(self.s, C.byref(boundary), nc, c.ctypes.data_as(_DP), min_dist, max_density, mass,
                         acc.ctypes.data_as(C.POINTER(C.c_uint8)), vel.ctypes.data_as(_DP), dmin.ctypes.data_as(_DP),
                         dens.ctypes.data_as(_DP))
        return acc.astype(bool), vel, dmin, dens

    def rest_extract(self, limit, boundary_limit, air=None):
        """one extraction pass of the rest-fluid loader on the current lists
        (air: flags of the air particles, or None): (keep, mean density of the
        kept candidates, densities)"""
        n = self.n
        keep = np.zeros(n, np.uint8)
        avg = C.c_double()
        rho = np.zeros(n)
        a = None if air is None else np.ascontiguousarray(air, np.uint8)
        self.L.or_rest_extract(self.s, float(limit), float(boundary_limit),
                               None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint8)),
                               keep.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(avg), rho.ctypes.data_as(_DP))
        return keep.astype(bool), avg.value, rho

    def outflow_keep(self, interior, surf):
        keep = np.zeros(self.n, np.uint8)
        self.L.or_outflow(self.s, C.byref(interior), C.byref(surf), keep.ctypes.data_as(C.POINTER(C.c_uint8)))
        return keep.astype(bool)


def surface_inside(surf, xyz):
    """(inside[n] bool, margin[n]) of or_surface_inside"""
    x = _d(xyz).reshape(-1, 3)
    ins = np.zeros(len(x), bool)
    mg = np.zeros(len(x))
    m = C.c_double()
    for i in range(len(x)):
        row = np.ascontiguousarray(x[i])
        ins[i] = bool(lib().or_surface_inside(C.byref(surf), row.ctypes.data_as(_DP), C.byref(m)))
        mg[i] = m.value
    return ins, mg


def neighbors_fp32(xf, xb, hk, brute=False, max_nbr=90):
    """FP32-predicate neighbor lists (fluid first, then boundary), row-major."""
    xf = np.ascontiguousarray(xf, np.float32).reshape(-1)
    xb = np.ascontiguousarray(xb, np.float32).reshape(-1)
    nf_ = xf.size // 3
    nf = np.zeros(nf_, np.uint32)
    nb = np.zeros(nf_, np.uint32)
    L = np.zeros((nf_, max_nbr), np.uint32)
    mx = lib().or_neighbors_fp32(nf_, xf.ctypes.data_as(_FP), xb.size // 3, xb.ctypes.data_as(_FP), C.c_float(hk),
                                 1 if brute else 0, max_nbr, nf.ctypes.data_as(_UP), nb.ctypes.data_as(_UP),
                                 L.ctypes.data_as(_UP))
    return nf, nb, L, mx


def kernel_W(r, hk, table=False):
    return lib().or_kernel_W(r, hk, 1 if table else 0)


def kernel_gradW(r3, hk, table=False):
    a = _d(r3)
    out = np.zeros(3)
    lib().or_kernel_gradW(a.ctypes.data_as(_DP), hk, 1 if table else 0, out.ctypes.data_as(_DP))
    return out


def mlp_forward(params_flat, x):
    p = _d(params_flat)
    x = _d(x).reshape(-1, 12)
    y = np.zeros((x.shape[0], 2))
    lib().or_mlp_forward(x.shape[0], p.ctypes.data_as(_DP), x.ctypes.data_as(_DP), y.ctypes.data_as(_DP))
    return y


def mlp_loss_grad(params_flat, x, t):
    p = _d(params_flat)
    x = _d(x).reshape(-1, 12)
    t = _d(t).reshape(-1, 2)
    g = np.zeros(MLP_NPARAM)
    loss = lib().or_mlp_loss_grad(x.shape[0], p.ctypes.data_as(_DP), x.ctypes.data_as(_DP), t.ctypes.data_as(_DP),
                                  g.ctypes.data_as(_DP))
    return loss, g
