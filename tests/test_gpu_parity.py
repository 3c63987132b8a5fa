"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Tolerances (FP32 device arithmetic vs the FP64 oracle, same inputs):
  neighbor lists       bit-exact as sorted sets (FP32 predicate in both)
  boundary psi         rel 2e-6
  density              rel 2e-6 of rho0
  factor (alpha)       rel 1e-4 of max |alpha|
  divergence source    1e-4 of max |rho_adv| + 1e-3 rho0/h * 1e-6
  velocities           5e-5 of (max |v| + g h) after a full step
  positions            1e-6 of the particle radius after one step
  iteration counts     equal
"""
import numpy as np
import pytest

import splishsplash_for_cuda_amd as S
from splishsplash_for_cuda_amd import scene
from gpu_util import boundary_index_map, gpu_neighbor_sets, max_rel, oracle_neighbor_sets

pytestmark = pytest.mark.gpu


def _solver(sc, psi=None, **kw):
    p = S.default_params(particle_radius=sc.particle_radius, **kw)
    s = S.Solver(p)
    s.upload_boundary(sc.boundary, psi)
    s.upload_fluid(sc.fluid)
    return s


def _oracle(om, sc, **kw):
    o = om.Oracle(particle_radius=sc.particle_radius, fp32_predicate=1, **kw)
    o.set_boundary(sc.boundary.astype(np.float64))
    o.set_fluid(sc.fluid.astype(np.float64))
    return o


@pytest.mark.parametrize("sort", [True, False])
def test_neighbor_lists_bit_exact_config1(oracle_mod, sort):
    sc = scene.config1()
    s = _solver(sc)
    s.neighbor_search(sort=sort)
    inf = s.info()
    assert inf.overflow == 0
    hk = np.float32(4 * sc.particle_radius)
    nf, nb, L, mx = oracle_mod.neighbors_fp32(sc.fluid, sc.boundary, hk, brute=True)
    want = oracle_neighbor_sets(nf, nb, L)
    got = gpu_neighbor_sets(s, boundary_index_map(s, sc.boundary))
    assert len(got) == len(want) == sc.n_fluid
    diff = [i for i in want if got[i] != want[i]]
    assert not diff, "%d particles differ, first %d" % (len(diff), diff[0])
    assert inf.max_neighbors_seen == mx


@pytest.mark.parametrize("list_mode", [S.LISTS_AUTO, S.LISTS_NO_LDS, S.LISTS_GLOBAL])
def test_neighbor_lists_bit_exact_every_list_form(oracle_mod, list_mode):
    """The same sets from the LDS-staged, the 16-bit unstaged and the 32-bit
    list forms (a dam break after a few steps: dense and sparse blocks)."""
    sc = scene.seeded_dam_break(5, 6000)
    s = _solver(sc, dt=0.002, list_mode=list_mode)
    for _ in range(30):
        s.step()
    s.neighbor_search()
    pos = s.download(S.FIELD_POS).astype(np.float32)
    ids = s.download(S.FIELD_ID)
    inf = s.info()
    blocks = (inf.n_fluid + inf.sweep_block - 1) // inf.sweep_block
    assert inf.sweep_block in (256, 512)
    if list_mode == S.LISTS_AUTO:
        assert inf.unstaged_blocks < blocks
    else:
        assert inf.unstaged_blocks == blocks
    if list_mode == S.LISTS_GLOBAL:
        assert inf.neighbors_fluid_unstaged == inf.neighbors_fluid_total
    else:
        assert inf.neighbors_fluid_unstaged == 0
    hk = np.float32(4 * sc.particle_radius)
    nf, nb, L, _ = oracle_mod.neighbors_fp32(pos, sc.boundary, hk, brute=True)
    want = {int(ids[i]): v for i, v in oracle_neighbor_sets(nf, nb, L).items()}
    want = {k: (frozenset(int(ids[j]) for j in f), b) for k, (f, b) in want.items()}
    got = gpu_neighbor_sets(s, boundary_index_map(s, sc.boundary))
    assert got == want


def test_list_forms_give_the_same_trajectory():
    """LDS-staged, unstaged 16-bit and 32-bit global lists hold the same
    neighbors in the same order (cells in one fixed order, each cell's
    particles by index), so every sum runs in the same order: bit-identical
    trajectories over 20 steps."""
    sc = scene.seeded_dam_break(5, 6000)
    out = {}
    for mode in (S.LISTS_AUTO, S.LISTS_NO_LDS, S.LISTS_GLOBAL):
        s = _solver(sc, dt=0.002, list_mode=mode)
        its = [(st.iters_v, st.iters) for st in (s.step() for _ in range(20))]
        out[mode] = (s.state_by_id(), its)
    a = out[S.LISTS_AUTO]
    for mode in (S.LISTS_NO_LDS, S.LISTS_GLOBAL):
        b = out[mode]
        for f in ("pos", "vel", "density", "kappa", "kappaV"):
            assert np.array_equal(a[0][f], b[0][f]), (mode, f)
        assert a[1] == b[1], mode


def test_neighbor_lists_bit_exact_random_cloud(oracle_mod):
    """Ragged densities, duplicates and pairs at exactly h."""
    rng = np.random.default_rng(21)
    r = 0.02
    hk = np.float32(4 * r)
    # about 85 neighbors in the bulk, about 180 in the crowded corner (cap 255)
    fl = rng.uniform(0.1, 0.5, size=(2500, 3)).astype(np.float32)
    fl[:150] = rng.uniform(0.1, 0.25, size=(150, 3)).astype(np.float32)   # crowded corner
    fl[400] = fl[401]
    fl[500] = fl[501] + np.float32([hk, 0, 0])
    b = scene.box_surface((0, 0, 0), (0.6, 0.6, 0.6), 2 * r).astype(np.float32)
    sc = scene.Scene("cloud", r, fl, b)
    s = _solver(sc, max_neighbors=255)
    s.neighbor_search()
    nf, nb, L, mx = oracle_mod.neighbors_fp32(fl, b, hk, brute=True, max_nbr=255)
    assert mx <= 255
    got = gpu_neighbor_sets(s, boundary_index_map(s, b))
    want = oracle_neighbor_sets(nf, nb, L)
    assert got == want


def test_boundary_psi_matches_oracle(oracle_mod):
    sc = scene.config1()
    s = _solver(sc)
    o = oracle_mod.Oracle(particle_radius=sc.particle_radius)
    o.set_boundary(sc.boundary.astype(np.float64))
    psi_o = o.boundary_psi(sc.n_boundary)
    bmap = boundary_index_map(s, sc.boundary)
    psi_g = s.download(S.FIELD_BOUNDARY_PSI)
    assert max_rel(psi_g, psi_o[bmap]) < 2e-6


def _by_id(state):
    return state


def _resync(s, o, om):
    """Start both sides from the oracle's fluid state rounded to FP32 (the
    device's precision), matched by id."""
    ids = s.download(S.FIELD_ID)
    where = {int(v): k for k, v in enumerate(o.ids())}
    idx = np.array([where[int(v)] for v in ids])
    for fs, fo in ((S.FIELD_POS, om.F_POS), (S.FIELD_VEL, om.F_VEL), (S.FIELD_KAPPA, om.F_KAPPA),
                   (S.FIELD_KAPPAV, om.F_KAPPAV)):
        v32 = o.get(fo).astype(np.float32)
        o.set(fo, v32.astype(np.float64))
        s.upload_field(fs, v32[idx])


def _phase_compare(om, sc, dt, steps, report, resync=False, gpu_psi=False, **kw):
    """Both sides step phase by phase; per step the worst relative error of
    each field.  gpu_psi: the device computes the boundary psi itself (its
    k_boundary_psi, checked against the oracle's first), otherwise both use
    the oracle's."""
    o = _oracle(om, sc, dt=dt, **kw)
    psi = o.boundary_psi(sc.n_boundary)
    s = _solver(sc, psi=None if gpu_psi else psi.astype(np.float32), dt=dt, **kw)
    rho0, g = 1000.0, 9.81
    worst = {}
    if gpu_psi:
        bmap = boundary_index_map(s, sc.boundary)
        worst["psi"] = max_rel(s.download(S.FIELD_BOUNDARY_PSI), psi[bmap])
    for k in range(steps):
        if resync and k:
            _resync(s, o, om)
        s.neighbor_search()
        o.neighbor_search()
        itv_g, _ = s.divergence_solve()
        itv_o, _ = o.divergence_solve()
        gs, os_ = s.state_by_id(), o.state()
        assert itv_g == itv_o, (k, itv_g, itv_o)
        vscale = np.abs(os_["vel"]).max() + g * dt
        e = {
            "density": max_rel(gs["density"], os_["density"], rho0),
            "factor": max_rel(gs["factor"], os_["factor"]),
            "vel_div": max_rel(gs["vel"], os_["vel"], vscale),
        }
        s.external_forces()
        o.external_forces()
        e["acc"] = max_rel(s.state_by_id()["acc"], o.get(om.F_ACC), g)
        s.update_time_step(dt)
        o.update_time_step(dt)
        s.update_vel()
        o.update_vel()
        it_g, _ = s.pressure_solve()
        it_o, _ = o.pressure_solve()
        assert it_g == it_o, (k, it_g, it_o)
        s.update_pos()
        o.update_pos()
        s.step_end()
        o.step_end()
        gs, os_ = s.state_by_id(), o.state()
        vscale = np.abs(os_["vel"]).max() + g * dt
        e["vel"] = max_rel(gs["vel"], os_["vel"], vscale)
        e["pos"] = max_rel(gs["pos"], os_["pos"], sc.particle_radius)
        # share of particles beyond the one-step bounds (a discrete decision
        # taken differently in FP32 and FP64 moves a particle and its shells)
        ev = np.abs(gs["vel"] - os_["vel"]).max(1) / vscale
        ep = np.abs(gs["pos"] - os_["pos"]).max(1) / sc.particle_radius
        e["vel_frac"] = float((ev > 5e-5).mean())
        e["pos_frac"] = float((ep > 1e-6).mean())
        # warm-start accumulators carried to the next step, relative to their
        # largest magnitude; and the share of particles beyond 1e-3 of it
        for f in ("kappa", "kappaV"):
            ks = max(np.abs(os_[f]).max(), 1e-30)
            e[f] = max_rel(gs[f], os_[f])
            e[f + "_frac"] = float((np.abs(gs[f] - os_[f]) > 1e-3 * ks).mean())
        for kk, v in e.items():
            worst[kk] = max(worst.get(kk, 0.0), v)
        report.append((k, itv_g, it_g, e))
    return worst


def test_config2_first_steps_match_oracle(oracle_mod):
    """The 1M-particle dam break (1,048,576 fluid + 600,801 boundary
    particles), first 20 steps against the FP64 oracle, each step started by
    both sides from the oracle's state: the one-step tolerances of the C1
    trajectory hold at full size."""
    sc = scene.config2()
    report = []
    w = _phase_compare(oracle_mod, sc, 0.001, 20, report, resync=True, gpu_psi=True, max_neighbors=128)
    for row in report[::5]:
        print(row)
    print("C2 20-step worst relative errors:", w)
    assert w["psi"] < 2e-6
    assert w["density"] < 2e-6
    assert w["factor"] < 1e-4
    assert w["vel_frac"] < 0.01 and w["vel"] < 1e-2
    # positions: one FP32 ulp of x += h v at |x| up to 6 m is 4.8e-5 r (r = 1 cm)
    assert w["pos"] < 5e-5
    _assert_kappa(w)


def test_one_step_phases_match_oracle(oracle_mod):
    sc = scene.small_dam_break()
    report = []
    w = _phase_compare(oracle_mod, sc, 0.002, 1, report)
    print("one-step worst relative errors:", w)
    assert w["density"] < 2e-6
    assert w["factor"] < 1e-4
    assert w["vel_div"] < 5e-5
    assert w["acc"] < 1e-5
    assert w["vel"] < 5e-5
    assert w["pos"] < 1e-6


def test_short_trajectory_matches_oracle(oracle_mod):
    """Config-1 dam break, first 10 steps, each started by both sides from the
    oracle's state: the one-step tolerances hold along the trajectory (a
    discrete decision taken differently in FP32 and FP64, e.g. the
    warm-start gate of a particle whose source is ~0, does not carry over)."""
    sc = scene.config1()
    report = []
    w = _phase_compare(oracle_mod, sc, 0.002, 10, report, resync=True)
    for row in report:
        print(row)
    print("10-step worst relative errors:", w)
    assert w["density"] < 2e-6
    assert w["factor"] < 1e-4
    # velocities: every particle within the one-step bound except those moved
    # by a flipped discrete decision (at most 1 %), whose error stays bounded;
    # positions: the FP32 rounding of x += h v at |x| up to 4 m (~0.4 r 1e-4)
    assert w["vel_frac"] < 0.01 and w["vel"] < 1e-2
    assert w["pos"] < 2e-5
    _assert_kappa(w)


def _assert_kappa(w):
    """kappa / kappaV after a step (the next step's warm starts): within 2e-3
    of the largest magnitude for 99.9 % of the particles, 2e-2 at worst (the
    FP32 sums of a Jacobi loop's residuals; a particle whose warm-start gate
    or source clamp flips between FP32 and FP64 takes a different increment)."""
    for f in ("kappa", "kappaV"):
        assert w[f + "_frac"] < 1e-3, (f, w[f + "_frac"])
        assert w[f] < 2e-2, (f, w[f])


def test_free_running_trajectory_aggregates_match_oracle(oracle_mod):
    """Config-1 dam break, 10 free-running steps: mean density, kinetic energy
    and the fluid's centroid agree (per-particle fields diverge once a
    discrete decision differs, so the aggregates are what stays comparable)."""
    om = oracle_mod
    sc = scene.config1()
    dt = 0.002
    o = _oracle(om, sc, dt=dt)
    psi = o.boundary_psi(sc.n_boundary)
    s = _solver(sc, psi=psi.astype(np.float32), dt=dt)
    for _ in range(10):
        s.step()
        o.step()
    gs, os_ = s.state_by_id(), o.state()
    assert gs["density"].mean() == pytest.approx(os_["density"].mean(), rel=1e-5)
    ke_g, ke_o = (gs["vel"] ** 2).sum(), (os_["vel"] ** 2).sum()
    assert ke_g == pytest.approx(ke_o, rel=1e-3)
    assert np.abs(gs["pos"].mean(0) - os_["pos"].mean(0)).max() < 1e-4 * sc.particle_radius


def test_full_step_entry_point_matches_phase_sequence():
    """dfsph_step (fused path) gives the same state as the phase-by-phase ABI calls."""
    sc = scene.small_dam_break()
    a = _solver(sc, dt=0.002)
    b = _solver(sc, dt=0.002)
    for _ in range(5):
        a.step()
        b.neighbor_search()
        b.divergence_solve()
        b.external_forces()
        b.update_time_step(0.002)
        b.update_vel()
        b.pressure_solve()
        b.update_pos()
        b.step_end()
        b.check_particles_positions(2)
    sa, sb = a.state_by_id(), b.state_by_id()
    for k in ("pos", "vel", "kappa", "kappaV", "density"):
        assert np.array_equal(sa[k], sb[k]), k


@pytest.mark.parametrize("use_cfl", [0, 1])
def test_host_class_step_matches_abi_step(use_cfl):
    """SPH::DFSPHCUDA::step() (C++ host over the C ABI) == dfsph_step, and
    the step statistics read from the device's loop records (no stream
    synchronisation) equal the synchronised device state."""
    sc = scene.small_dam_break()
    p = S.default_params(particle_radius=sc.particle_radius, dt=0.002, use_cfl=use_cfl)
    host = S.HostSolver(p, sc.boundary, sc.fluid)
    abi = S.Solver(p)
    abi.upload_boundary(sc.boundary)
    abi.upload_fluid(sc.fluid)
    for _ in range(5):
        st = host.step()
        sa = abi.step()
        assert (st.iters_v, st.iters) == (sa.iters_v, sa.iters)
        assert (st.h, st.time) == (sa.dt, sa.time)
        assert st.n_fluid == sa.n_fluid == sc.n_fluid
        inf = abi.info()
        assert (sa.dt, sa.time) == (inf.h_past, inf.time)
        assert sa.max_neighbors_seen == inf.max_neighbors_seen
        assert sa.neighbors_fluid_total == inf.neighbors_fluid_total
    x1, x2 = host.view.state_by_id(), abi.state_by_id()
    assert np.array_equal(x1["pos"], x2["pos"]) and np.array_equal(x1["vel"], x2["vel"])
    host.close()


def test_particle_leaving_the_box_in_a_step_is_removed():
    """A particle that update_pos carries out of the bounding box is removed
    at the end of the step (the escaped count of the step end's record), in
    the host class and in dfsph_step alike."""
    sc = scene.small_dam_break()
    # no pressure solves: nothing but XSPH damps the planted velocity
    p = S.default_params(particle_radius=sc.particle_radius, dt=0.002, use_cfl=0, enable_divergence=0, max_iter=0)
    host = S.HostSolver(p, sc.boundary, sc.fluid)
    abi = S.Solver(p)
    abi.upload_boundary(sc.boundary)
    abi.upload_fluid(sc.fluid)
    for s in (host.view, abi):
        s.step() if s is abi else host.step()
        ids = s.download(S.FIELD_ID)
        vel = s.download(S.FIELD_VEL).copy()
        vel[int(np.nonzero(ids == 7)[0][0])] = (0.0, 5000.0, 0.0)
        s.upload_field(S.FIELD_VEL, vel)
    st, sa = host.step(), abi.step()
    assert st.removed == sa.removed == 1
    assert st.n_fluid == sa.n_fluid == sc.n_fluid - 1
    for s in (host.view, abi):
        assert 7 not in set(s.download(S.FIELD_ID).tolist())
    st, sa = host.step(), abi.step()
    assert st.removed == sa.removed == 0
    host.close()


def test_runs_are_bit_reproducible():
    sc = scene.config1()
    out = []
    for _ in range(2):
        s = _solver(sc, dt=0.002)
        for _ in range(5):
            s.step()
        out.append(s.state_by_id())
    for k in ("id", "pos", "vel", "kappa", "kappaV"):
        assert np.array_equal(out[0][k], out[1][k]), k


def test_profiled_run_is_the_same_run_and_times_every_phase():
    """dfsph_profile_*: the sweeps launched with their timing events in the
    dispatch (hipExtLaunchKernel) compute what the plain launches do, bit
    for bit, and every phase of a step is timed once per launch."""
    from splishsplash_for_cuda_amd._native import KERNEL_NAMES
    sc = scene.config1()
    ref = _solver(sc, dt=0.002)
    for _ in range(5):
        ref.step()
    s = _solver(sc, dt=0.002)
    s.profile_enable()
    s.profile_reset()
    its = [s.step() for _ in range(5)]
    s.synchronize()
    prof = s.profile()
    a, b = ref.state_by_id(), s.state_by_id()
    for k in ("id", "pos", "vel", "kappa", "kappaV"):
        assert np.array_equal(a[k], b[k]), k
    once = ("sort", "neighbors", "div_ws_init", "div_init", "xsph", "pressure_ws", "dens_init", "update_pos")
    for name in once:
        n, ms = prof[name]
        assert n == 5, (name, n)
        assert 0.0 < ms < 1e3, (name, ms)
    # one velocity sweep and one loop-end sweep per Jacobi iteration (the
    # velocity sweep queued ahead that exits at once is not counted)
    nv = sum(st.iters_v for st in its)
    nd = sum(st.iters for st in its)
    assert prof["div_loop_end"][0] == nv and prof["dens_loop_end"][0] == nd
    assert prof["div_jacobi"][0] == nv and prof["dens_jacobi"][0] == nd
    assert all(prof[k][1] > 0.0 for k in ("div_loop_end", "dens_loop_end", "div_jacobi", "dens_jacobi"))
    assert set(prof) == set(KERNEL_NAMES)
    s.profile_enable(0)


def test_caller_stream_gives_the_same_run():
    """dfsph_set_stream: a run whose steps move to a caller's stream after
    two steps (and back to the handle's own for the last) ends bit for bit
    like a run on the handle's own stream."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    sc = scene.config1()
    ref = _solver(sc, dt=0.002)
    for _ in range(6):
        ref.step()
    s = _solver(sc, dt=0.002)
    own = s.get_stream()
    st = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0           # default priority
    try:
        s.step()
        s.step()
        s.set_stream(st.value)
        assert s.get_stream() == st.value
        for _ in range(3):
            s.step()
        s.set_stream(0)
        assert s.get_stream() == own
        s.step()
        a, b = ref.state_by_id(), s.state_by_id()
    finally:
        s.close()
        hip.hipStreamDestroy(st)
    for k in ("id", "pos", "vel", "kappa", "kappaV"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("case", ["reference", "learned", "body"])
def test_checkpoint_resume_is_bit_exact(tmp_path, case):
    """handleSimulationSave / handleSimulationLoad: a run resumed from a
    snapshot continues bit for bit like the uninterrupted one, with the
    reference or the learned warm start (whose bound reads the last step's
    totals, which the snapshot keeps) and with a dynamic rigid body (its
    state is in the snapshot)."""
    sc = scene.small_dam_break()
    ws = S.WARMSTART_LEARNED if case == "learned" else S.WARMSTART_REFERENCE
    p = S.default_params(particle_radius=sc.particle_radius, dt=0.002, warm_start_mode=ws)

    def make():
        hs = S.HostSolver(p, sc.boundary, sc.fluid)
        if case == "learned":
            from splishsplash_for_cuda_amd import learned
            learned.attach(hs.view)
        if case == "body":
            b = scene.box_surface((-0.06, -0.06, -0.06), (0.06, 0.06, 0.06), 2 * sc.particle_radius)
            hs.view.add_rigid_body(S.rigid_body(mass=0.8, inertia=(2e-3, 2e-3, 2e-3), position=(0.7, 0.6, 0.3),
                                                velocity=(0.0, -0.2, 0.0)), b.astype(np.float32))
        return hs

    a = make()
    for _ in range(3):
        a.step()
    path = str(tmp_path / "snap.bin")
    a.save(path)
    for _ in range(4):
        a.step()
    b = make()
    b.load(path)
    for _ in range(4):
        b.step()
    sa, sb = a.view.state_by_id(), b.view.state_by_id()
    for k in ("pos", "vel", "kappa", "kappaV"):
        assert np.array_equal(sa[k], sb[k]), k
    if case == "body":
        ra, rb = a.view.rigid_body(0)[0], b.view.rigid_body(0)[0]
        for f in ("position", "velocity", "rotation", "angular_velocity"):
            assert list(getattr(ra, f)) == list(getattr(rb, f)), f
    a.close()
    b.close()


def test_escaped_particle_is_removed_order_kept():
    sc = scene.small_dam_break()
    s = _solver(sc, dt=0.002)
    s.step()
    st = s.state_by_id()
    x = st["pos"].copy()
    # push particle id 7 out of the tank through the ceiling (device order upload)
    ids = s.download(S.FIELD_ID)
    pos_dev = s.download(S.FIELD_POS).copy()
    slot = int(np.nonzero(ids == 7)[0][0])
    pos_dev[slot, 1] = 50.0
    s.upload_field(S.FIELD_POS, pos_dev)
    before = s.download(S.FIELD_ID)
    removed = s.check_particles_positions(2)
    after = s.download(S.FIELD_ID)
    assert removed == 1
    assert after.tolist() == [i for i in before.tolist() if i != 7]
    assert s.n == sc.n_fluid - 1
    st = s.step()
    assert st.n_fluid == sc.n_fluid - 1
    del x


def test_remove_and_add_particles():
    sc = scene.small_dam_break()
    s = _solver(sc, dt=0.002)
    n = sc.n_fluid
    keep = np.ones(n, np.uint8)
    keep[::5] = 0
    ids = s.download(S.FIELD_ID)
    assert s.remove_particles(keep) == n - (n + 4) // 5
    assert s.download(S.FIELD_ID).tolist() == ids[keep.astype(bool)].tolist()
    s.add_particles(np.array([[0.3, 0.9, 0.3]], np.float32), ids=np.array([99999], np.uint32))
    assert s.download(S.FIELD_ID)[-1] == 99999
    s.step()


def test_neighbor_overflow_is_reported():
    r = 0.025
    rng = np.random.default_rng(2)
    fl = (np.float32(0.5) + rng.uniform(-0.01, 0.01, size=(200, 3))).astype(np.float32)
    b = scene.box_surface((0, 0, 0), (1, 1, 1), 2 * r).astype(np.float32)
    s = _solver(scene.Scene("crowd", r, fl, b), max_neighbors=90)
    with pytest.raises(S.DFSPHError) as e:
        s.step()
    assert e.value.code == 4
    assert s.info().overflow == 200


def test_cfl_matches_oracle(oracle_mod):
    sc = scene.small_dam_break()
    o = _oracle(oracle_mod, sc, dt=0.002, dt_min=1e-5, dt_max=0.05)
    psi = o.boundary_psi(sc.n_boundary)
    s = _solver(sc, psi=psi.astype(np.float32), dt=0.002)
    vel = np.tile(np.float32([2.0, -1.0, 0.5]), (sc.n_fluid, 1))
    o.set_fluid(sc.fluid.astype(np.float64), vel.astype(np.float64))
    s.upload_fluid(sc.fluid, vel)
    for x in (s, o):
        x.neighbor_search()
        x.divergence_solve()
        x.external_forces()
    hg = s.cfl(1e-5, 1.0, 0.05)
    ho = o.cfl()
    assert hg == pytest.approx(ho, rel=1e-5)


@pytest.mark.parametrize("ws", ["reference", "learned"])
def test_water_at_rest_stays_at_rest(ws):
    """Still water in a closed tank (1 x 0.8 x 1 m, 0.3 m deep, r = 12.5 mm)
    for 1 s: the largest speed stays below 1 m/s while the lattice start
    settles and below 0.4 m/s over the last half second with the reference
    warm start (measured 0.50 and 0.36), below 1 and 0.6 m/s with the
    committed learned model (measured 0.58 and 0.38), and every step meets the
    tolerances.  The water is outside the learned model's training set (dam
    breaks)."""
    r = 0.0125
    d = 2 * r
    walls = scene.box_surface((0, 0, 0), (1.0, 0.8, 1.0), d).astype(np.float32)
    fl = scene.fluid_block((d, d, d), (39, 12, 39), d).astype(np.float32)
    lw = ws == "learned"
    p = S.default_params(particle_radius=r, dt=0.001, use_cfl=1, dt_max=0.002, max_error=0.05, max_error_v=0.1,
                         warm_start_mode=S.WARMSTART_LEARNED if lw else S.WARMSTART_REFERENCE)
    s = S.Solver(p)
    s.upload_boundary(walls)
    s.upload_fluid(fl)
    if lw:
        from splishsplash_for_cuda_amd import learned
        learned.attach(s)
    t, vmax = 0.0, []
    while t < 1.0:
        st = s.step()
        t = st.time
        assert st.err / 1000.0 * 100 <= 0.05 * 1.0001 and st.err_v * st.dt / 1000.0 * 100 <= 0.1 * 1.0001
        vmax.append(float(np.abs(s.download(S.FIELD_VEL)).max()))
    late = vmax[len(vmax) // 2:]
    print("max speed over 1 s: %.3f m/s, over the last 0.5 s %.3f (last %.3f)" % (max(vmax), max(late), vmax[-1]))
    bound, bound_late = (1.0, 0.6) if lw else (1.0, 0.4)
    assert max(vmax) < bound and max(late) < bound_late and st.n_fluid == len(fl)
    s.close()


def test_empty_and_single_particle_fluids():
    """Edge cases of the step: a boundary with no fluid steps without error
    (time advances, nothing appears); one particle far from the walls falls
    freely, v = g t and y = y0 + g t^2 / 2 summed over the steps (no
    neighbors: no pressure, no viscosity), every solve at zero error."""
    sc = scene.small_dam_break()
    s = S.Solver(S.default_params(particle_radius=sc.particle_radius, dt=0.002))
    s.upload_boundary(sc.boundary)
    s.upload_fluid(np.zeros((0, 3), np.float32))
    for _ in range(3):
        st = s.step()
        assert st.n_fluid == 0
    assert s.info().n_fluid == 0 and s.get_time()[2] == pytest.approx(0.006, rel=1e-6)
    s.close()
    lo, hi = sc.boundary.min(0), sc.boundary.max(0)
    x0 = 0.5 * (lo + hi)
    s = S.Solver(S.default_params(particle_radius=sc.particle_radius, dt=0.002))
    s.upload_boundary(sc.boundary)
    s.upload_fluid(x0[None, :].astype(np.float32))
    y, v = float(x0[1]), 0.0
    for _ in range(20):
        st = s.step()
        assert st.err == 0.0 and st.err_v == 0.0
        v -= 9.81 * 0.002
        y += 0.002 * v                       # symplectic Euler, as the step integrates
    x = s.download(S.FIELD_POS)[0]
    vel = s.download(S.FIELD_VEL)[0]
    assert vel[1] == pytest.approx(v, rel=1e-5) and abs(vel[0]) < 1e-7 and abs(vel[2]) < 1e-7
    assert x[1] == pytest.approx(y, abs=1e-5) and x[0] == pytest.approx(x0[0], abs=1e-6)
    s.close()
