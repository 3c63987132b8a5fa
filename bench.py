"""Benchmark of the DFSPH time step on one MI355X.

Workload (default, --config c4): the dam break of 4,194,304 fluid particles
(256 x 128 x 128 block, r = 7.5 mm) with a 0.5 m cube obstacle on the floor
of a closed 6 x 3 x 2 m tank (walls and obstacle sampled at spacing 2r);
FP32 bulk arithmetic, FP64 convergence reductions, divergence and density
tolerances 0.1 %, CFL 0.4 with dt in [0.1, 5] ms, minimum iterations (3, 2)
as the reference GPU solver.  A "step" is one DFSPHCUDA::step() (neighbor
search, both solvers, forces, integration, position check) over the whole
particle set.

Legs, each on its own solver, the same steps:
  plain    the reference warm start: warm-up, then the timed steps (no
           profiling events inside the timed region), then the same number
           of profiled steps for the roofline and the kernel shares: every
           solver sweep carries a HIP event pair in its own dispatch
           (hipExtLaunchKernel), so its duration is the kernel's execution;
           the sort and the search are bracketed by event records;
  learned  the committed learned warm start, timed and profiled the same way;
           its ms per step, iterations per step, error maxima, the time of its
           feature sweep and inference, and the final density field against
           the plain run's (mean |rho_learned - rho_plain| / rho0 over the
           particles matched by id, after the same number of steps);
  cpu      the FP64 OpenMP oracle in the reference CPU semantics, timed on a
           bounded sample of the same scene (context only).

Prints ONE JSON line.
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
RHO0 = 1000.0

# Algorithmic bytes per launch of each neighbor sweep:
#   B_k = N_f * S_k + 4 * (sum n_f + sum n_b) + 16 * N_b
# S_k the compulsory per-particle FP32 state of the reference kernel that the
# sweep replaces (mass uniform, not counted), one 32-bit index per stored
# neighbor, the boundary position + psi read once.  A fused sweep counts the
# reference kernels it replaces (each reads the lists once).
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
SWEEP_PASSES = {"div_ws_init": 2, "pressure_ws": 1}
SCENES = {"c1": scene.config1, "c2": scene.config2, "c4": scene.config4}


def kernel_bytes(name, n_f, n_b, sum_nf, sum_nb):
    s = SWEEP_STATE_BYTES[name]
    passes = SWEEP_PASSES.get(name, 1)
    if name == "xsph":
        return n_f * s + 4 * sum_nf
    return n_f * s + passes * (4 * (sum_nf + sum_nb) + 16 * n_b)


def make_params(sc, mode, max_neighbors, min_iters, floor_only=False, gate="above-floors"):
    p = S.default_params(
        particle_radius=sc.particle_radius, dt=0.001, use_cfl=1, cfl_factor=1.0, dt_min=1e-4, dt_max=5e-3,
        max_error=0.1, max_error_v=0.1, max_neighbors=max_neighbors,
        warm_start_mode=S.WARMSTART_LEARNED if mode == "learned" else S.WARMSTART_REFERENCE,
        learned_gate=S.LEARNED_ABOVE_FLOORS if gate == "above-floors" else S.LEARNED_ALWAYS)
    p.min_iter_v, p.min_iter = min_iters
    if floor_only:
        # cost of a step whose solves stop at their floors: what a perfect warm
        # start would cost (tolerances not enforced)
        p.max_iter_v, p.max_iter = min_iters
    return p


def make_solver(sc, p, tank, mode):
    if tank is not None:
        from splishsplash_for_cuda_amd import wave_tank
        return wave_tank.make_solver(tank, p, learned=mode == "learned")
    hs = S.HostSolver(p, sc.boundary, sc.fluid)
    if mode == "learned":
        from splishsplash_for_cuda_amd import learned
        learned.attach(hs.view)
    return hs


def run_leg(sc, p, tank, mode, steps, warmup, prof_steps):
    """One solver through SPH::DFSPHCUDA::step(): warm-up, timed steps (no
    events), profiled steps (events on every kernel); the final density field
    by particle id."""
    hs = make_solver(sc, p, tank, mode)
    solver = hs.view
    for _ in range(warmup):
        hs.step()
    inf0 = solver.info()
    rec = {k: [] for k in ("iters_v", "iters", "err_v", "err", "h", "max_seen", "overflow")}
    solver.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = hs.step()
        rec["iters_v"].append(st.iters_v)
        rec["iters"].append(st.iters)
        rec["err_v"].append(st.err_v)
        rec["err"].append(st.err)
        rec["h"].append(st.h)
        rec["max_seen"].append(st.max_neighbors_seen)
        rec["overflow"].append(st.overflow)
    solver.synchronize()
    elapsed = time.perf_counter() - t0
    inf1 = solver.info()
    # profiled pass: the same number of steps, events around every kernel
    solver.profile_enable((1 << len(KERNEL_NAMES)) - 1)
    solver.profile_reset()
    for _ in range(prof_steps):
        hs.step()
    solver.synchronize()
    prof = solver.profile()
    inf2 = solver.info()
    solver.profile_enable(0)
    ids = solver.download(S.FIELD_ID)
    rho = solver.download(S.FIELD_DENSITY)
    density = np.full(int(ids.max()) + 1 if ids.size else 0, np.nan)
    density[ids] = rho
    hs.close()
    n_f = (inf0.n_fluid + inf1.n_fluid) / 2.0
    # per-step errors as fractions of their tolerances' units: density
    # err / rho0, divergence err * h / rho0 (the thresholds of the solves,
    # DFSPH_core_cuda.cu:822,1372), in percent
    err_d = 100.0 * np.asarray(rec["err"]) / RHO0
    err_v = 100.0 * np.asarray(rec["err_v"]) * np.asarray(rec["h"]) / RHO0
    return {
        "elapsed": elapsed, "n_f": n_f, "inf0": inf0, "inf1": inf1, "inf2": inf2, "prof": prof,
        "prof_steps": prof_steps, "rec": rec, "density": density, "err_density_pct": err_d,
        "err_divergence_pct": err_v, "time": inf2.time,
    }


def roofline_of(leg, peak=HBM_PEAK_GBS):
    """Dominant sweep and sum over the solver sweeps, from the profiled pass."""
    inf1, inf2, prof = leg["inf1"], leg["inf2"], leg["prof"]
    n_f = (inf1.n_fluid + inf2.n_fluid) / 2.0
    n_b = inf2.n_boundary
    sum_nf = (inf1.neighbors_fluid_total + inf2.neighbors_fluid_total) / 2.0
    sum_nb = (inf1.neighbors_boundary_total + inf2.neighbors_boundary_total) / 2.0
    dominant = max(SWEEP_STATE_BYTES, key=lambda k: prof[k][1])
    launches, ms = prof[dominant]
    nbytes = kernel_bytes(dominant, n_f, n_b, sum_nf, sum_nb)
    avg_s = ms / 1e3 / max(launches, 1)
    tot_b = sum(kernel_bytes(k, n_f, n_b, sum_nf, sum_nb) * prof[k][0] for k in SWEEP_STATE_BYTES)
    tot_s = sum(prof[k][1] for k in SWEEP_STATE_BYTES) / 1e3
    return dominant, {
        "bound": "hbm",
        "kernel": dominant,
        "achieved": nbytes / avg_s / 1e9,
        "peak": peak,
        "unit": "GB/s",
        "frac": nbytes / avg_s / 1e9 / peak,
        "traffic": None,
        "bytes_model": "N_f*S_k + 4*(sum n_f + sum n_b) + 16*N_b, S_k = %d B" % SWEEP_STATE_BYTES[dominant],
        "algorithmic_bytes_per_launch": nbytes,
        "avg_launch_us": avg_s * 1e6,
        "launches": launches,
        "solver_sweeps": {"achieved": tot_b / max(tot_s, 1e-12) / 1e9,
                          "frac": tot_b / max(tot_s, 1e-12) / 1e9 / peak,
                          "bytes": tot_b, "seconds": tot_s},
    }


def iters_summary(rec):
    iv, ip = np.asarray(rec["iters_v"]), np.asarray(rec["iters"])
    return {"divergence_mean": float(iv.mean()), "density_mean": float(ip.mean()),
            "divergence_p95": float(np.percentile(iv, 95)), "density_p95": float(np.percentile(ip, 95))}


def errors_summary(leg):
    return {"density_max_pct": float(leg["err_density_pct"].max()),
            "divergence_max_pct": float(leg["err_divergence_pct"].max()),
            "density_mean_pct": float(leg["err_density_pct"].mean()),
            "divergence_mean_pct": float(leg["err_divergence_pct"].mean())}


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


def workload_text(args, sc, tank, first, last):
    if tank is not None:
        return ("C5: %s DFSPH open-boundary wave tank 10x2x1 m, %d fluid (start) + %d wall particles, r=%g m, "
                "sinusoidal piston (0.1 m, 1.5 s) and a floating box (500 kg/m^3) as rigid bodies, absorbing open "
                "boundary with per-step inflow / outflow compaction, tolerances 0.1%%, CFL dt in [0.1,5] ms, "
                "steps %d-%d of the run" % (args.mode, sc.n_fluid, sc.n_boundary, sc.particle_radius, first, last))
    what = {"c1": "dam break", "c2": "dam break",
            "c4": "dam break with a 0.5 m cube obstacle"}[args.config]
    return ("%s: %s DFSPH %s, %d fluid + %d boundary particles, r=%g m, tolerances 0.1%%, CFL 0.4 dt in "
            "[0.1,5] ms, min iters (%s), steps %d-%d of the run"
            % (args.config.upper(), args.mode, what, sc.n_fluid, sc.n_boundary, sc.particle_radius,
               args.min_iters, first, last))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c4", choices=["c1", "c2", "c4", "c5"],
                    help="c4: the 4M-particle dam break with obstacle; c5: the open-boundary wave tank")
    ap.add_argument("--mode", default="plain", choices=["plain", "learned"])
    ap.add_argument("--min-iters", default="3,2",
                    help="minimum divergence,density Jacobi iterations (the reference GPU solver's are 3,2)")
    ap.add_argument("--prof-steps", type=int, default=0,
                    help="profiled steps after the timed ones (0: as many as timed, at most 50)")
    ap.add_argument("--floor-only", action="store_true",
                    help="stop both solves at their minimum iterations: the cost of a step with a perfect warm "
                         "start (tolerances not enforced; not a valid throughput)")
    ap.add_argument("--learned-gate", default="above-floors", choices=["above-floors", "always"],
                    help="learned warm start every step, or only after a step whose solves ran above their "
                         "minimum iterations (the reference warm start otherwise)")
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
        sc = SCENES[args.config]()
    max_nbr = 90 if args.config == "c1" else 128
    min_iters = tuple(int(v) for v in args.min_iters.split(","))
    prof_steps = args.prof_steps or min(args.steps, 50)
    p = make_params(sc, args.mode, max_nbr, min_iters, args.floor_only, args.learned_gate)

    leg = run_leg(sc, p, tank, args.mode, args.steps, args.warmup, prof_steps)
    dominant, roof = roofline_of(leg)
    tfile = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (dominant, args.config))
    if os.path.exists(tfile):
        roof["traffic"] = json.load(open(tfile)).get("bytes_per_launch")
    prof = leg["prof"]
    total_ms = max(sum(v[1] for v in prof.values()), 1e-9)
    rec = leg["rec"]
    first, last = args.warmup + 1, args.warmup + args.steps
    line = {
        "metric": "particle-steps/s at ρ-err≤0.1%; pressure iters/step learned vs plain DFSPH",
        "value": leg["n_f"] * args.steps / leg["elapsed"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * leg["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": workload_text(args, sc, tank, first, last),
            "path": "SPH::DFSPHCUDA::step()",
            "iters_per_step": iters_summary(rec),
            "errors": errors_summary(leg),
            "tolerances_pct": {"density": 0.1, "divergence": 0.1},
            "max_neighbors": max_nbr,
            "max_neighbors_seen": int(max(rec["max_seen"])),
            "overflow": int(max(rec["overflow"])),
            "min_iters": list(min_iters),
            "unstaged_blocks": int(leg["inf1"].unstaged_blocks),
            "sweep_block": int(leg["inf1"].sweep_block),
            "neighbors_per_particle": {
                "fluid": round(leg["inf1"].neighbors_fluid_total / max(leg["inf1"].n_fluid, 1), 2),
                "boundary": round(leg["inf1"].neighbors_boundary_total / max(leg["inf1"].n_fluid, 1), 2)},
            "timing": "timed steps without profiling events; roofline and kernel shares from %d further "
                      "profiled steps (HIP events in each sweep's dispatch, on the solver's stream)" % prof_steps,
        },
        "roofline": roof,
        "kernel_share": {k: round(v[1] / total_ms, 4) for k, v in prof.items() if v[0]},
        "kernel_ms_per_step": {k: round(v[1] / prof_steps, 4) for k, v in prof.items() if v[0]},
    }
    if args.floor_only:
        line["config"]["floor_only"] = "solves stopped at their minimum iterations: cost ceiling, not a valid throughput"
    if args.mode == "plain" and not args.no_learned_leg and not args.floor_only:
        pl = make_params(sc, "learned", max_nbr, min_iters, gate=args.learned_gate)
        lrn = run_leg(sc, pl, tank, "learned", args.steps, args.warmup, prof_steps)
        ms = 1e3 * lrn["elapsed"] / args.steps
        lp = lrn["prof"]
        nd = min(leg["density"].size, lrn["density"].size)
        both = ~np.isnan(leg["density"][:nd]) & ~np.isnan(lrn["density"][:nd])
        dfield = np.abs(lrn["density"][:nd][both] - leg["density"][:nd][both]) / RHO0
        line["learned_vs_plain"] = {
            "ms_per_step": ms,
            "value": lrn["n_f"] * args.steps / lrn["elapsed"],
            "speedup_vs_plain": line["ms_per_step"] / ms,
            "learned_gate": args.learned_gate,
            "iters_per_step": iters_summary(lrn["rec"]),
            "errors": errors_summary(lrn),
            "max_neighbors_seen": int(max(lrn["rec"]["max_seen"])),
            "overflow": int(max(lrn["rec"]["overflow"])),
            "overhead_ms_per_step": {k: round(lp[k][1] / prof_steps, 4) for k in ("features", "mlp_infer")},
            "density_field_vs_plain": {
                "after_steps": args.warmup + args.steps + prof_steps,
                "mean_abs_rel": float(dfield.mean()) if dfield.size else None,
                "max_abs_rel": float(dfield.max()) if dfield.size else None,
                "matched": int(both.sum()),
                "time_plain_s": leg["time"], "time_learned_s": lrn["time"]},
        }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(sc, args.cpu_budget, 1000, tank)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
