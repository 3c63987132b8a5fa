"""The 4,194,304-particle dam break with a 0.5 m cube obstacle (256 x 128 x 128
block, r = 7.5 mm, closed 6 x 3 x 2 m tank, walls and obstacle at spacing 2r),
the configuration the learned-vs-plain comparison is stated on.

  first steps   3 steps against the FP64 oracle, each started by both sides
                from the oracle's state; the device computes its own boundary
                psi: psi rel 2e-6, density 2e-6 of rho0, factor 1e-4, equal
                iteration counts, velocities 5e-5 of (max|v| + g h) for 99 %
                of the particles, positions one FP32 ulp at 6 m (3.2e-5 r at
                r = 7.5 mm, bound 7e-5 r), kappa / kappaV as in the C2 test
  learned       30 steps after 10 of warm-up, plain and learned warm start at
                the reference floors (3, 2): every step of both within the
                0.1 % tolerances, neighbor lists never truncated, learned
                iterations no more than the plain solver's, the final
                density fields within 1 % of rho0 of each other (mean by id)
"""
import numpy as np
import pytest

import splishsplash_for_cuda_amd as S
from splishsplash_for_cuda_amd import learned, scene
from test_gpu_parity import _assert_kappa, _phase_compare

pytestmark = pytest.mark.gpu

TOL = 0.1


@pytest.fixture(scope="module")
def c4():
    return scene.config4()


@pytest.mark.timeout(900)
def test_config4_first_steps_match_oracle(oracle_mod, c4):
    report = []
    w = _phase_compare(oracle_mod, c4, 0.001, 3, report, resync=True, gpu_psi=True, max_neighbors=128)
    for row in report:
        print(row, flush=True)
    print("C4 3-step worst relative errors:", w, flush=True)
    assert w["psi"] < 2e-6
    assert w["density"] < 2e-6
    assert w["factor"] < 1e-4
    assert w["vel_div"] < 1e-2
    assert w["vel_frac"] < 0.01 and w["vel"] < 1e-2
    assert w["pos"] < 7e-5
    _assert_kappa(w)


def _run(sc, ws, steps, warmup):
    p = learned.solver_params(sc, warm_start=ws, tol=TOL, min_iter_v=3, min_iter=2, max_neighbors=128)
    s = learned.make_solver(sc, p)
    if ws == S.WARMSTART_LEARNED:
        learned.attach(s)
    iv, ip, ev, ep, ms = [], [], [], [], []
    for k in range(warmup + steps):
        st = s.step()
        if k >= warmup:
            iv.append(st.iters_v)
            ip.append(st.iters)
            ev.append(st.err_v * st.dt / 1000.0 * 100.0)
            ep.append(st.err / 1000.0 * 100.0)
            ms.append(st.max_neighbors_seen)
            assert st.overflow == 0
            # within the tolerances unless the iteration cap ended the solve
            assert st.iters_v >= p.max_iter_v or ev[-1] <= TOL * 1.0001, (k, ev[-1])
            assert st.iters >= p.max_iter or ep[-1] <= TOL * 1.0001, (k, ep[-1])
    out = {"iters_v": float(np.mean(iv)), "iters": float(np.mean(ip)), "err_v": max(ev), "err": max(ep),
           "max_seen": max(ms), "state": s.state_by_id()}
    s.close()
    return out


@pytest.mark.timeout(600)
def test_config4_learned_vs_plain(c4):
    plain = _run(c4, S.WARMSTART_REFERENCE, 30, 10)
    lrn = _run(c4, S.WARMSTART_LEARNED, 30, 10)
    a, b = plain["state"], lrn["state"]
    common, ia, ib = np.intersect1d(a["id"], b["id"], assume_unique=True, return_indices=True)
    drho = float(np.mean(np.abs(a["density"][ia] - b["density"][ib])) / 1000.0 * 100.0)
    print("C4 steps 11-40: plain %.2f/%.2f, learned %.2f/%.2f div/dens iterations per step; max errors plain "
          "%.4f/%.4f %%, learned %.4f/%.4f %%; density fields %.3f %% apart; max neighbors %d/%d"
          % (plain["iters_v"], plain["iters"], lrn["iters_v"], lrn["iters"], plain["err_v"], plain["err"],
             lrn["err_v"], lrn["err"], drho, plain["max_seen"], lrn["max_seen"]), flush=True)
    assert len(common) == c4.n_fluid
    assert lrn["iters_v"] <= plain["iters_v"] and lrn["iters"] <= plain["iters"] + 0.05
    assert drho < 1.0
