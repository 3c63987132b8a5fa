"""Benchmark of the DFSPH time step on one MI355X.

Workload (default): the plain DFSPH dam break of 1,048,576 fluid particles
(128 x 128 x 64 block, r = 0.01 m) in a 6 x 3 x 2 m open-top tank with
600,801 wall particles at spacing r; FP32 bulk arithmetic, FP64 convergence
reductions, divergence and density tolerances 0.1 %, CFL 0.4 with dt in
[0.1, 5] ms.  A "step" is one DFSPHCUDA::step() (neighbor search, both
solvers, forces, integration, position check) over the whole particle set.

Prints ONE JSON line: particle-steps/s over the timed steps, the roofline of
the dominant kernel (algorithmic bytes per launch / its average duration from
HIP events recorded on the solver's stream during the timed steps), and the
CPU baseline (the FP64 OpenMP oracle in the reference CPU semantics, timed on
a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import splishsplash_for_cuda_amd as S  # noqa: E402
from splishsplash_for_cuda_amd import scene  # noqa: E402
from splishsplash_for_cuda_amd._native import KERNEL_NAMES  # noqa: E402

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, AMD published

# Compulsory per-particle state of each sweep in bytes (FP32 SoA; uniform
# fluid mass not counted), the index per stored neighbor (2 B in the 16-bit
# lists of LDS-staged blocks, 4 B in 32-bit lists) and the 16 B boundary
# position + psi are added per launch:
#   B = N_f * S_k + 2 * sum n_fluid,staged + 4 * (sum n_fluid,unstaged + sum n_boundary) + 16 * N_b
# Algorithmic bytes per launch, the per-unit figures of the hot-path contract:
# B_k = N_f S_k + 4 (sum n_f + sum n_b) + 16 N_b -- S_k the compulsory
# per-particle FP32 state of the reference kernel (mass uniform, not counted),
# one 32-bit index per stored neighbor, the boundary position + psi once.
# Fused kernels count the reference kernels they replace.
SWEEP_STATE_BYTES = {
    "div_ws_init": 44 + 44,   # fused rho/alpha/div pass (x12 v12 r; rho alpha rho_adv 12 w; kappaV8 rw) + divergence warm start (x12 kappaV4 rho_adv4 v24 rw)
    "div_init": 32,           # x12 v12 r; rho_adv4 kappaV4 w
    "div_jacobi": 52,         # x12 v24 rw, rho_adv4, alpha4, kappaV8 rw
    "div_loop_end": 28,       # x12 v12 r; rho_adv4 w
    "xsph": 40,               # x12 v12 rho4 r; acc12 w (fluid neighbors only, no boundary term)
    "pressure_ws": 44 + 40,   # update_vel (acc12, v24 rw, kappa8 rw) + pressure warm start (x12 kappa4 v24 rw)
    "dens_init": 44,          # x12 v12 rho4 r; alpha8 rw; rho_adv4 kappa4 w
    "dens_jacobi": 44,        # x12 v24 rw, rho_adv4, alpha4
    "dens_loop_end": 44,      # x12 v12 rho4 alpha4 r; rho_adv4 w; kappa8 rw
}
# the reference kernels a fused launch replaces (each reads the neighbor lists once)
SWEEP_PASSES = {"div_ws_init": 2, "pressure_ws": 1}


def kernel_bytes(name, n_f, n_b, sum_nf, sum_nb):
    s = SWEEP_STATE_BYTES[name]
    passes = SWEEP_PASSES.get(name, 1)
    if name == "xsph":
        return n_f * s + 4 * sum_nf
    return n_f * s + passes * (4 * (sum_nf + sum_nb) + 16 * n_b)


def n_f_mean_over(inf0, inf1):
    """fluid particles per step over the timed region (open boundaries change the count)"""
    return (inf0.n_fluid + inf1.n_fluid) / 2.0


def make_params(sc, mode, max_neighbors):
    return S.default_params(
        particle_radius=sc.particle_radius, dt=0.001, use_cfl=1, cfl_factor=1.0, dt_min=1e-4, dt_max=5e-3,
        max_error=0.1, max_error_v=0.1, max_neighbors=max_neighbors,
        warm_start_mode=S.WARMSTART_LEARNED if mode == "learned" else S.WARMSTART_REFERENCE)


def cpu_baseline(sc, budget_s, max_steps, tank=None):
    """FP64 OpenMP oracle in the reference CPU semantics (TimeStepDFSPH) on the
    same scene: whole steps until budget_s of CPU time has been used (the wave
    tank: with its two rigid bodies, without the open boundary)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as om
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = om.Oracle(particle_radius=sc.particle_radius, dt=0.001, mode=om.MODE_CPU, use_cfl=1, cfl_factor=1.0,
                  dt_min=1e-4, dt_max=5e-3, cfl_seed=0.1, max_error=0.1, max_error_v=0.1, min_iter_v=1,
                  max_neighbors=160, remove_escaped=1)
    o.set_boundary(sc.boundary.astype(np.float64))
    o.set_fluid(sc.fluid.astype(np.float64))
    if tank is not None:
        for desc, pts in tank.bodies:
            o.add_body(om.body_desc(**desc), pts.astype(np.float64))
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        o.step()
        steps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    t = time.perf_counter() - t0
    return {"value": sc.n_fluid * steps / t, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "first %d step(s) of the same %d-particle scene, FP64 restatement of the reference CPU "
                      "TimeStepDFSPH with OpenMP (%.1f s)" % (steps, sc.n_fluid, t)}


def learned_leg(sc, p, tank, steps, warmup):
    """The same workload through SPH::DFSPHCUDA::step() with the committed
    learned warm start: ms per step, throughput and iterations per step, run
    after the plain solver is closed (untimed setup, the same warm-up)."""
    from splishsplash_for_cuda_amd import learned
    p.warm_start_mode = S.WARMSTART_LEARNED
    if tank is not None:
        from splishsplash_for_cuda_amd import wave_tank
        hs = wave_tank.make_solver(tank, p, learned=True)
    else:
        hs = S.HostSolver(p, sc.boundary, sc.fluid)
        learned.attach(hs.view)
    for _ in range(warmup + 5):
        hs.step()
    inf0 = hs.view.info()
    iv, ip = [], []
    hs.view.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = hs.step()
        iv.append(st.iters_v)
        ip.append(st.iters)
    hs.view.synchronize()
    elapsed = time.perf_counter() - t0
    inf1 = hs.view.info()
    hs.close()
    return {"ms_per_step": 1e3 * elapsed / steps, "value": n_f_mean_over(inf0, inf1) * steps / elapsed,
            "iters_per_step": {"divergence_mean": float(np.mean(iv)), "density_mean": float(np.mean(ip))}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c4", "c5"],
                    help="c5: the open-boundary wave tank (piston, absorbing end, floating box)")
    ap.add_argument("--mode", default="plain", choices=["plain", "learned"])
    ap.add_argument("--path", default="host", choices=["host", "abi"],
                    help="host: SPH::DFSPHCUDA::step() (C++); abi: dfsph_step")
    ap.add_argument("--min-iters", default="3,2",
                    help="minimum divergence,density Jacobi iterations (the reference GPU solver's are 3,2)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned-leg", action="store_true",
                    help="plain mode: skip the learned run of the same workload reported beside it")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the solver is single-GPU (multi-GPU slab decomposition is "
              "design-only); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    tank = None
    if args.config == "c5":
        from splishsplash_for_cuda_amd import wave_tank
        tank = wave_tank.build()
        sc = scene.Scene("C5_wave_tank", tank.particle_radius, tank.fluid, tank.walls)
    else:
        sc = {"c1": scene.config1, "c2": scene.config2, "c4": scene.config4}[args.config]()
    max_nbr = 90 if args.config == "c1" else 128
    p = make_params(sc, args.mode, max_nbr)
    p.min_iter_v, p.min_iter = (int(v) for v in args.min_iters.split(","))
    if tank is not None:
        if args.path != "host":
            print("c5 runs through SPH::DFSPHCUDA::step() (open boundaries): use --path host", file=sys.stderr)
            sys.exit(2)
        hs = wave_tank.make_solver(tank, p, learned=args.mode == "learned")
        solver = hs.view
        step = hs.step
    elif args.path == "host":
        hs = S.HostSolver(p, sc.boundary, sc.fluid)
        solver = hs.view
        step = hs.step
    else:
        solver = S.Solver(p)
        solver.upload_boundary(sc.boundary)
        solver.upload_fluid(sc.fluid)
        step = solver.step
    if args.mode == "learned" and tank is None:
        from splishsplash_for_cuda_amd import learned
        learned.attach(solver)

    for _ in range(args.warmup):
        step()
    # find the dominant sweep over a few profiled steps (outside the timed region)
    solver.profile_enable((1 << len(KERNEL_NAMES)) - 1)
    solver.profile_reset()
    for _ in range(5):
        step()
    prof = solver.profile()
    solver.profile_enable(0)
    dominant = max((k for k in SWEEP_STATE_BYTES), key=lambda k: prof[k][1])

    # timed region: event timing of the dominant sweep only
    solver.profile_enable(1 << KERNEL_NAMES.index(dominant))
    solver.profile_reset()
    inf0 = solver.info()
    iters_v, iters = [], []
    solver.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = step()
        iters_v.append(st.iters_v)
        iters.append(st.iters)
    solver.synchronize()
    elapsed = time.perf_counter() - t0
    inf1 = solver.info()
    launches, ms = solver.profile()[dominant]
    solver.profile_enable(0)

    n_f = (inf0.n_fluid + inf1.n_fluid) / 2.0
    sum_nf = (inf0.neighbors_fluid_total + inf1.neighbors_fluid_total) / 2.0
    sum_nb = (inf0.neighbors_boundary_total + inf1.neighbors_boundary_total) / 2.0
    sum_nfu = (inf0.neighbors_fluid_unstaged + inf1.neighbors_fluid_unstaged) / 2.0
    nbytes = kernel_bytes(dominant, n_f, inf1.n_boundary, sum_nf, sum_nb)
    avg_s = ms / 1e3 / max(launches, 1)
    achieved = nbytes / avg_s / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (dominant, args.config))
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_launch")

    value = n_f_mean_over(inf0, inf1) * args.steps / elapsed
    line = {
        "metric": "particle-steps/s at ρ-err≤0.1%; pressure iters/step learned vs plain DFSPH",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("%s: %s DFSPH dam break, %d fluid + %d boundary particles, r=%g m, tolerances 0.1%%, "
                         "CFL 0.4 dt in [0.1,5] ms, steps %d-%d of the run"
                         % (args.config.upper(), args.mode, sc.n_fluid, sc.n_boundary, sc.particle_radius,
                            args.warmup + 5, args.warmup + 5 + args.steps)) if tank is None else
                        ("C5: %s DFSPH open-boundary wave tank 10x2x1 m, %d fluid (start) + %d wall particles, "
                         "r=%g m, sinusoidal piston (0.1 m, 1.5 s) and a floating box (500 kg/m^3) as rigid "
                         "bodies, absorbing open boundary with per-step inflow / outflow compaction, tolerances "
                         "0.1%%, CFL dt in [0.1,5] ms, steps %d-%d of the run"
                         % (args.mode, sc.n_fluid, sc.n_boundary, sc.particle_radius, args.warmup + 5,
                            args.warmup + 5 + args.steps)),
            "path": "SPH::DFSPHCUDA::step()" if args.path == "host" else "dfsph_step",
            "iters_per_step": {"divergence_mean": float(np.mean(iters_v)), "density_mean": float(np.mean(iters)),
                               "divergence_p95": float(np.percentile(iters_v, 95)),
                               "density_p95": float(np.percentile(iters, 95))},
            "max_neighbors": max_nbr,
            "min_iters": [int(p.min_iter_v), int(p.min_iter)],
            "unstaged_blocks": int(inf1.unstaged_blocks),
            "fluid_neighbors_unstaged_frac": float(sum_nfu / max(sum_nf, 1.0)),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "bytes_model": "N_f*S_k + 4*(sum n_f + sum n_b) + 16*N_b, S_k = %d B" % SWEEP_STATE_BYTES[dominant],
            "traffic": traffic,
            "algorithmic_bytes_per_launch": nbytes,
            "avg_launch_us": avg_s * 1e6,
            "launches": launches,
        },
        "kernel_share": {k: round(v[1] / max(sum(x[1] for x in prof.values()), 1e-9), 4) for k, v in prof.items()},
    }
    if args.path == "host":
        hs.close()
    else:
        solver.close()
    if args.mode == "plain" and args.path == "host" and not args.no_learned_leg:
        lrn = learned_leg(sc, p, tank, args.steps, args.warmup)
        lrn["speedup_vs_plain"] = line["ms_per_step"] / lrn["ms_per_step"]
        line["learned_vs_plain"] = lrn
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(sc, args.cpu_budget, 1000, tank)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
