"""Learned warm start: training-set harvest, GPU training, model files.

The network (12 -> 64 -> 64 -> 64 -> 2, SiLU) predicts a correction to each
particle's reference warm starts of a step, from features of the step's
initial state (see k_features in csrc/dfsph_kernels.hip).  Its targets are
the increments that the plain solver's Jacobi iterations add after the
reference warm start, in dt-invariant units (kappaV h, kappa h_f^2), harvested
from runs with the reference warm start.  The solver adds the correction to
the reference warm start, bounded so that it asks for no more pressure than
the larger of the reference warm start and the particle's converged total
of the last step (half of it for the divergence solver): a network that
outputs zero is the reference warm start, exactly.  The plain solver keeps iterating to the same tolerances when the
learned warm start is used: the network only changes where the Jacobi
iterations start.
"""
from __future__ import annotations

import os

import numpy as np

from . import scene as scenes
from ._native import MLP_NPARAM, WARMSTART_LEARNED, WARMSTART_REFERENCE, Solver, default_params

SIZES = [(64, 12), (64,), (64, 64), (64,), (64, 64), (64,), (2, 64), (2,)]
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "models", "warmstart_mlp.npz")


def init_params(seed=0):
    """Glorot-uniform weights, zero biases (flat FP32 layout of the C ABI)."""
    rng = np.random.default_rng(seed)
    out = []
    for s in SIZES:
        if len(s) == 2:
            lim = np.sqrt(6.0 / (s[0] + s[1]))
            out.append(rng.uniform(-lim, lim, size=s).ravel())
        else:
            out.append(np.zeros(s))
    p = np.concatenate(out).astype(np.float32)
    assert p.size == MLP_NPARAM
    return p


def solver_params(sc, warm_start=WARMSTART_REFERENCE, tol=0.1, min_iter_v=3, min_iter=2, max_neighbors=128, **kw):
    """The dam-break settings of the benchmark (tolerances 0.1 %, CFL 0.4)."""
    return default_params(particle_radius=sc.particle_radius, dt=0.001, use_cfl=1, cfl_factor=1.0, dt_min=1e-4,
                          dt_max=5e-3, max_error=tol, max_error_v=tol, max_neighbors=max_neighbors,
                          warm_start_mode=warm_start, min_iter_v=min_iter_v, min_iter=min_iter, **kw)


def make_solver(sc, params):
    s = Solver(params)
    s.upload_boundary(sc.boundary)
    s.upload_fluid(sc.fluid)
    return s


def harvest_scene(sc, steps, max_samples, **kw):
    """Run the plain solver on one scene and collect (features, targets): the
    targets are the increments of each step's solves after the reference
    warm start, in dt-invariant units."""
    s = make_solver(sc, solver_params(sc, **kw))
    stride = max(1, int(np.ceil(sc.n_fluid * steps / max_samples)))
    s.mlp_harvest(max_samples, stride)
    for _ in range(steps):
        s.step()
    x, t = s.mlp_dataset()
    s.close()
    return x, t


def build_training_set(n_scenes=32, n_range=(50_000, 200_000), total=2_000_000, steps=40, seed=0, log=None):
    """Training pairs from n_scenes seeded dam breaks (seed 1000 + k), with random
    particle counts in n_range and block aspect ratios in [0.5, 2]."""
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    per = total // n_scenes
    for k in range(n_scenes):
        n_target = int(rng.uniform(*n_range))
        sc = scenes.seeded_dam_break(k, n_target)
        x, t = harvest_scene(sc, steps, per)
        xs.append(x)
        ts.append(t)
        if log:
            log("scene %d: %d particles, %d samples" % (k, sc.n_fluid, len(x)))
    return np.concatenate(xs), np.concatenate(ts)


def train(x, t, steps=3000, batch=4096, lr=2e-3, seed=0, init_seed=0):
    """Fit the network on the GPU; returns (params, norm, (first batch loss, last batch loss))."""
    s = Solver()
    s.mlp_set_params(init_params(init_seed))
    s.mlp_dataset_upload(x, t)
    s.mlp_fit_normalization()
    losses = s.mlp_train(steps, batch, lr, seed)
    p, nrm = s.mlp_get_params()
    s.close()
    return p, nrm, losses


def save(path, params, norm, **meta):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=np.asarray(params, np.float32), norm=np.asarray(norm, np.float32),
             **{k: np.asarray(v) for k, v in meta.items()})


def load(path=DEFAULT_MODEL):
    with np.load(path, allow_pickle=False) as z:
        return z["params"], z["norm"]


def attach(solver, path=DEFAULT_MODEL, params=None, norm=None):
    """Load a trained model into a solver (which then needs warm_start_mode = learned)."""
    if params is None:
        params, norm = load(path)
    solver.mlp_set_params(params, norm)


def compare_iterations(sc, params, norm, steps, warmup=0, **kw):
    """Iterations per step of the same scene with the reference, learned and no
    warm start.  Errors are reported relative to rho0 in percent (the
    divergence error as rho_dot h / rho0), as the maximum over the measured
    steps; `rho_diff_pct` is the mean |rho - rho_reference| / rho0 of the final
    density fields, matched by particle id."""
    out = {}
    from ._native import WARMSTART_NONE
    rho0 = 1000.0
    for name, ws in (("reference", WARMSTART_REFERENCE), ("learned", WARMSTART_LEARNED), ("none", WARMSTART_NONE)):
        prm = solver_params(sc, warm_start=ws, **kw)
        rho0 = prm.density0
        s = make_solver(sc, prm)
        if ws == WARMSTART_LEARNED:
            s.mlp_set_params(params, norm)
        iv, ip, ev, ep = [], [], [], []
        for k in range(warmup + steps):
            st = s.step()
            if k >= warmup:
                iv.append(st.iters_v)
                ip.append(st.iters)
                ev.append(st.err_v * st.dt / rho0 * 100.0)
                ep.append(st.err / rho0 * 100.0)
        out[name] = {"iters_v": float(np.mean(iv)), "iters": float(np.mean(ip)), "err_v": float(np.max(ev)),
                     "err": float(np.max(ep)), "state": s.state_by_id()}
        s.close()
    ref = out["reference"]["state"]
    for name, r in out.items():
        st = r["state"]
        common, a, b = np.intersect1d(st["id"], ref["id"], assume_unique=True, return_indices=True)
        r["rho_diff_pct"] = float(np.mean(np.abs(st["density"][a] - ref["density"][b])) / rho0 * 100.0)
    return out
