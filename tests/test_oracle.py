"""Pins of the CPU oracle (oracle/dfsph_oracle.c) before it is trusted as the
checker of the HIP path.  Each check is independent of the oracle's own
formulas: closed forms, quadrature, finite differences, brute force,
symmetry invariants and conservation laws."""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------- kernel (A3)

def test_kernel_golden_values(oracle_mod):
    """Cubic spline at 11 q points against tests/golden/kernel_values.json
    (exact rational evaluation of the closed form)."""
    g = json.load(open(os.path.join(HERE, "golden", "kernel_values.json")))
    hk = g["hk"]
    for q, w in zip(g["q"], g["W"]):
        assert oracle_mod.kernel_W(q * hk, hk) == pytest.approx(w, rel=1e-13, abs=1e-12)


def test_kernel_normalised(oracle_mod):
    hk = 0.1
    r = np.linspace(0.0, hk, 200001)
    w = np.array([oracle_mod.kernel_W(x, hk) for x in r[::100]])
    rr = r[::100]
    integral = np.trapezoid(4.0 * np.pi * rr * rr * w, rr)
    assert integral == pytest.approx(1.0, rel=2e-5)
    wt = np.array([oracle_mod.kernel_W(x, hk, table=True) for x in rr])
    assert np.trapezoid(4.0 * np.pi * rr * rr * wt, rr) == pytest.approx(1.0, rel=2e-3)


def test_gradient_is_derivative_of_W(oracle_mod):
    hk = 0.1
    rng = np.random.default_rng(3)
    for _ in range(40):
        d = rng.normal(size=3)
        d *= rng.uniform(0.05, 0.97) * hk / np.linalg.norm(d)
        g = oracle_mod.kernel_gradW(d, hk)
        fd = np.zeros(3)
        eps = 1e-7
        for a in range(3):
            e = np.zeros(3)
            e[a] = eps
            fd[a] = (oracle_mod.kernel_W(np.linalg.norm(d + e), hk) - oracle_mod.kernel_W(np.linalg.norm(d - e), hk)) / (2 * eps)
        assert np.allclose(g, fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())
        assert np.allclose(oracle_mod.kernel_gradW(-d, hk), -g, rtol=0, atol=0)


def test_gradient_continuous_at_half_and_zero_outside(oracle_mod):
    hk = 0.1
    a = oracle_mod.kernel_gradW([0.5 * hk * (1 - 1e-9), 0, 0], hk)[0]
    b = oracle_mod.kernel_gradW([0.5 * hk * (1 + 1e-9), 0, 0], hk)[0]
    assert a == pytest.approx(b, rel=1e-6)
    assert oracle_mod.kernel_gradW([1.01 * hk, 0, 0], hk).tolist() == [0, 0, 0]
    assert oracle_mod.kernel_gradW([1e-8, 0, 0], hk).tolist() == [0, 0, 0]


def test_table_kernel_close_to_analytic(oracle_mod):
    hk = 0.1
    for q in np.linspace(0.0, 0.999, 37):
        assert oracle_mod.kernel_W(q * hk, hk, table=True) == pytest.approx(oracle_mod.kernel_W(q * hk, hk), rel=2e-3, abs=1e-3 * oracle_mod.kernel_W(0, hk))


# ---------------------------------------------------------------- neighbors (A8)

def _sets(nf, nb, L):
    return [(frozenset(L[i, :nf[i]].tolist()), frozenset(L[i, nf[i]:nf[i] + nb[i]].tolist())) for i in range(len(nf))]


def test_neighbors_grid_equals_brute_force(oracle_mod):
    rng = np.random.default_rng(11)
    hk = np.float32(0.1)
    xf = rng.uniform(0.0, 0.6, size=(3000, 3)).astype(np.float32)
    xf[100] = xf[200]                      # duplicate position
    xf[300] = xf[301] + np.float32([hk, 0, 0])   # exactly h apart
    xb = rng.uniform(-0.05, 0.65, size=(800, 3)).astype(np.float32)
    a = oracle_mod.neighbors_fp32(xf, xb, hk, brute=True, max_nbr=255)
    b = oracle_mod.neighbors_fp32(xf, xb, hk, brute=False, max_nbr=255)
    assert a[3] <= 255
    assert _sets(*a[:3]) == _sets(*b[:3])
    # self excluded, duplicate included, strict < at exactly h
    s = _sets(*a[:3])
    assert 100 not in s[100][0] and 200 in s[100][0]
    assert 301 not in s[300][0]


# ---------------------------------------------------------------- density / factor (A9)

def _far_boundary():
    return np.array([[50.0, 50.0, 50.0]])


def test_lattice_density_and_factor(oracle_mod):
    """Interior particle of a perfect cubic lattice: density and factor from an
    independent enumeration of lattice offsets."""
    r = 0.025
    d = 2 * r
    hk = 4 * r
    n = 11
    g = np.stack(np.meshgrid(*(np.arange(n),) * 3, indexing="ij"), -1).reshape(-1, 3) * d
    o = oracle_mod.Oracle(particle_radius=r, max_iter_v=0, min_iter_v=0)
    o.set_boundary(_far_boundary())
    o.set_fluid(g)
    o.neighbor_search()
    o.divergence_solve()
    c = (n // 2) * (n * n + n + 1)
    m = 0.8 * d ** 3 * 1000.0
    offs = [np.array(v) * d for v in np.ndindex(5, 5, 5)]
    offs = [v - 2 * d for v in offs]
    offs = [v for v in offs if 0 < np.dot(v, v) < hk * hk]
    rho = m * oracle_mod.kernel_W(0.0, hk) + sum(m * oracle_mod.kernel_W(np.linalg.norm(v), hk) for v in offs)
    sg2 = sum(float(np.dot(m * oracle_mod.kernel_gradW(-v, hk), m * oracle_mod.kernel_gradW(-v, hk))) for v in offs)
    alpha = -(1.0 / o.p.dt) / sg2
    assert o.get(oracle_mod.F_DENSITY)[c] == pytest.approx(rho, rel=1e-12)
    assert o.get(oracle_mod.F_FACTOR)[c] == pytest.approx(alpha, rel=1e-12)
    # |j|^2 in {1, 2, 3} lattice offsets (6 + 12 + 8); the 6 at exactly 2 spacings
    # (|r| = h) are excluded by the strict test
    assert len(offs) == 26
    # the 6 ties at |r| = h fall either way with the rounding of the lattice
    # coordinates; they contribute W = gradW = 0
    nf, nb, _ = o.neighbors()
    assert 26 <= nf[c] <= 32 and nb[c] == 0


@pytest.mark.parametrize("motion", ["translation", "rotation"])
def test_divergence_source_vanishes_for_rigid_motion(oracle_mod, motion):
    r = 0.025
    rng = np.random.default_rng(5)
    x = np.stack(np.meshgrid(*(np.arange(9),) * 3, indexing="ij"), -1).reshape(-1, 3) * 2 * r
    x = x + rng.uniform(-0.2 * r, 0.2 * r, size=x.shape)
    if motion == "translation":
        v = np.tile([0.3, -1.2, 0.7], (len(x), 1))
    else:
        w = np.array([0.4, -1.1, 2.0])
        v = np.cross(w, x - x.mean(0)) + np.array([0.1, 0.2, 0.3])
    o = oracle_mod.Oracle(particle_radius=r, max_iter_v=0, min_iter_v=0)
    o.set_boundary(_far_boundary())
    o.set_fluid(x, v)
    o.neighbor_search()
    o.divergence_solve()
    radv = o.get(oracle_mod.F_DENSITY_ADV)
    rho = o.get(oracle_mod.F_DENSITY)
    assert np.abs(radv).max() < 1e-9 * rho.max() / o.p.dt


def test_divergence_source_positive_under_compression(oracle_mod):
    r = 0.025
    x = np.stack(np.meshgrid(*(np.arange(9),) * 3, indexing="ij"), -1).reshape(-1, 3) * 2 * r
    v = -(x - x.mean(0))               # uniform compression: density must grow
    o = oracle_mod.Oracle(particle_radius=r, max_iter_v=0, min_iter_v=0)
    o.set_boundary(_far_boundary())
    o.set_fluid(x, v)
    o.neighbor_search()
    o.divergence_solve()
    radv = o.get(oracle_mod.F_DENSITY_ADV)
    nf, nb, _ = o.neighbors()
    interior = (nf + nb) >= 20
    assert interior.sum() > 50
    assert (radv[interior] > 0).all()
    # analytic value for the lattice: rho_adv = -rho * div v = 3 rho for v = -x (SPH estimate within 20%)
    rho = o.get(oracle_mod.F_DENSITY)
    c = 4 * (81 + 9 + 1)
    assert radv[c] == pytest.approx(3 * rho[c], rel=0.2)


def test_jacobi_update_conserves_fluid_momentum(oracle_mod):
    r = 0.025
    rng = np.random.default_rng(9)
    x = np.stack(np.meshgrid(*(np.arange(8),) * 3, indexing="ij"), -1).reshape(-1, 3) * 2 * r
    x = x + rng.uniform(-0.1 * r, 0.1 * r, size=x.shape)
    v = -(x - x.mean(0)) + rng.normal(scale=0.05, size=x.shape)
    o = oracle_mod.Oracle(particle_radius=r, max_iter_v=1, min_iter_v=1)
    o.set_boundary(_far_boundary())
    o.set_fluid(x, v)
    o.neighbor_search()
    o.divergence_solve()
    dv = o.get(oracle_mod.F_VEL) - v
    m = o.get(oracle_mod.F_MASS)[:, None]
    assert np.abs(dv).max() > 1e-3
    assert np.abs((m * dv).sum(0)).max() < 1e-12 * np.abs(m * dv).sum()


def test_xsph_zero_for_uniform_velocity(oracle_mod):
    r = 0.025
    x = np.stack(np.meshgrid(*(np.arange(6),) * 3, indexing="ij"), -1).reshape(-1, 3) * 2 * r
    v = np.tile([0.5, 0.1, -0.2], (len(x), 1))
    o = oracle_mod.Oracle(particle_radius=r, max_iter_v=0, min_iter_v=0, gravity=(0.0, -9.81, 0.0))
    o.set_boundary(_far_boundary())
    o.set_fluid(x, v)
    o.neighbor_search()
    o.divergence_solve()
    o.external_forces()
    acc = o.get(oracle_mod.F_ACC)
    assert np.abs(acc - np.array([0.0, -9.81, 0.0])).max() < 1e-12


def test_cfl_formula(oracle_mod):
    r = 0.025
    x = np.stack(np.meshgrid(*(np.arange(4),) * 3, indexing="ij"), -1).reshape(-1, 3) * 2 * r
    v = np.tile([3.0, 0.0, 0.0], (len(x), 1))
    o = oracle_mod.Oracle(particle_radius=r, max_iter_v=0, min_iter_v=0, gravity=(0.0, 0.0, 0.0), viscosity=0.0,
                          dt=0.001, dt_min=1e-5, dt_max=0.05)
    o.set_boundary(_far_boundary())
    o.set_fluid(x, v)
    o.neighbor_search()
    o.divergence_solve()
    o.external_forces()
    assert o.cfl() == pytest.approx(0.4 * 2 * r / 3.0, rel=1e-12)


def test_escaped_particles_removed_in_order(oracle_mod):
    r = 0.025
    box = np.array([[0, 0, 0], [1, 1, 1.0]])
    x = np.array([[0.5, 0.5, 0.5], [0.2, 0.2, 0.2], [0.3, 1.5, 0.3], [0.7, 0.7, 0.7], [-0.1, 0.5, 0.5]])
    o = oracle_mod.Oracle(particle_radius=r)
    o.set_boundary(box)
    o.set_fluid(x, ids=np.array([10, 11, 12, 13, 14]))
    assert o.check_positions(True) == 2
    assert o.ids().tolist() == [10, 11, 13]


# ---------------------------------------------------------------- solvers (A10-A16)

def _column(r=0.025, nx=6, ny=14, nz=6):
    d = 2 * r
    x = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3) * d + d
    tank = (nx * d + d, ny * d * 1.6 + d, nz * d + d)
    from splishsplash_for_cuda_amd.scene import box_surface
    return x, box_surface((0, 0, 0), tank, d)


def test_hydrostatic_column_settles(oracle_mod):
    r = 0.025
    x, b = _column(r)
    o = oracle_mod.Oracle(particle_radius=r, dt=0.002, max_error=0.1, max_error_v=0.1)
    o.set_boundary(b)
    o.set_fluid(x)
    eta = 0.1 * 0.01 * 1000.0
    for _ in range(100):
        st = o.step()
        assert st.err <= eta or st.iters == o.p.max_iter
        assert st.err_v <= eta / 0.002 or st.iters_v == o.p.max_iter_v
        assert st.iters >= 2 and st.iters_v >= 3
        assert st.removed == 0
    v = o.get(oracle_mod.F_VEL)
    x = o.get(oracle_mod.F_POS)
    rho = o.get(oracle_mod.F_DENSITY)
    nf, nb, _ = o.neighbors()
    # far below the free-fall speed of the column (sqrt(2 g 0.7) ~ 3.7 m/s)
    assert np.abs(v).max() < 1.0
    assert x[:, 1].min() > 0.02            # no particle crosses the floor layer
    assert np.median(rho[(nf + nb) > 30]) == pytest.approx(1000.0, rel=0.01)


def test_cpu_and_gpu_modes_agree_on_a_first_step(oracle_mod):
    """With zero warm starts and equal minimum iterations the two semantic modes
    give the same first step; only the stored kappa differs by the last
    residual term (the GPU path accumulates it in its last loop end)."""
    r = 0.025
    x, b = _column(r)
    res = {}
    for mode in (oracle_mod.MODE_GPU, oracle_mod.MODE_CPU):
        o = oracle_mod.Oracle(particle_radius=r, dt=0.002, mode=mode)
        o.set_boundary(b)
        o.set_fluid(x)
        st = o.step()
        res[mode] = (o.state(), st)
    a, b_ = res[0][0], res[1][0]
    assert np.array_equal(a["pos"], b_["pos"]) and np.array_equal(a["vel"], b_["vel"])
    h = 0.002
    extra = a["density_adv"] * a["factor"]
    assert np.allclose(a["kappa"] - extra, b_["kappa"] / (h * h), rtol=1e-9, atol=1e-12)


# ---------------------------------------------------------------- MLP (A22)

def test_mlp_gradient_matches_finite_differences(oracle_mod):
    rng = np.random.default_rng(1)
    p = rng.normal(scale=0.3, size=oracle_mod.MLP_NPARAM)
    x = rng.normal(size=(7, 12))
    t = rng.normal(size=(7, 2))
    loss, g = oracle_mod.mlp_loss_grad(p, x, t)
    y = oracle_mod.mlp_forward(p, x)
    assert loss == pytest.approx(0.5 * ((y - t) ** 2).sum() / 7, rel=1e-12)
    idx = rng.choice(oracle_mod.MLP_NPARAM, size=60, replace=False)
    for k in idx:
        e = np.zeros_like(p)
        e[k] = 1e-6
        lp, _ = oracle_mod.mlp_loss_grad(p + e, x, t)
        lm, _ = oracle_mod.mlp_loss_grad(p - e, x, t)
        assert g[k] == pytest.approx((lp - lm) / 2e-6, rel=1e-4, abs=1e-8)


def test_mlp_silu_forward_small_case(oracle_mod):
    """A 1-active-path network computed by hand: all weights zero except one
    chain, so y0 = w4 silu(w3 silu(w2 silu(w1 x0 + b1) + b2) + b3) + b4."""
    p = np.zeros(oracle_mod.MLP_NPARAM)
    sizes = [int(np.prod(s)) for s in oracle_mod.MLP_SIZES]
    off = np.cumsum([0] + sizes)
    p[off[0] + 0] = 1.5      # W1[0,0]
    p[off[1] + 0] = -0.2     # b1[0]
    p[off[2] + 0] = 0.7      # W2[0,0]
    p[off[4] + 0] = -1.3     # W3[0,0]
    p[off[5] + 0] = 0.1      # b3[0]
    p[off[6] + 0] = 2.0      # W4[0,0]
    p[off[7] + 1] = 0.25     # b4[1]
    silu = lambda z: z / (1 + math.exp(-z))
    x = np.zeros((1, 12))
    x[0, 0] = 0.8
    a1 = silu(1.5 * 0.8 - 0.2)
    a2 = silu(0.7 * a1)
    a3 = silu(-1.3 * a2 + 0.1)
    y = oracle_mod.mlp_forward(p, x)
    assert y[0, 0] == pytest.approx(2.0 * a3 + 2.0 * 0.0, rel=1e-14)
    # the other hidden units have zero pre-activation -> silu(0) = 0, except
    # W2/W3 rows with zero inputs: b terms only
    assert y[0, 1] == pytest.approx(0.25, rel=1e-14)


# ---------------------------------------------------------------- open boundary / dynamic window (A19, A20)

def test_surfaces_closed_forms(oracle_mod):
    om = oracle_mod
    S = om.surface
    cases = [
        (S("sphere", center=(1, 2, 3), radius=0.5), [(1, 2, 3.4), (1.2, 2.2, 3)], [(1, 2, 3.6), (1.4, 2.4, 3)]),
        (S("plane", o=(0, 1, 0), n=(0, 1, 0)), [(5, 1.0, 0), (0, 1.01, 9)], [(0, 0.99, 0)]),
        (S("cuboid", center=(0, 0, 0), half=(1, 2, 3)), [(0.9, -1.9, 2.9)], [(1.1, 0, 0), (0, 2.1, 0), (0, 0, -3.1)]),
        (S("cylinder", center=(0, 0, 0), half=(0, 1, 0), radius=0.5), [(0.4, 0.9, 0), (0, -0.99, 0.45)],
         [(0.4, 1.1, 0), (0.6, 0, 0), (0.3, 0, 0.45)]),
    ]
    for surf, inside, outside in cases:
        for pts, want in ((inside, True), (outside, False)):
            got, _ = om.surface_inside(surf, pts)
            assert got.tolist() == [want] * len(pts)
            surf.reversed = 1
            got, _ = om.surface_inside(surf, pts)
            assert got.tolist() == [not want] * len(pts)
            surf.reversed = 0


def test_star_surface_vertices_and_area(oracle_mod):
    """The linear-sided star: its border passes through the outer vertices
    (angle 0 from the first point's direction) and the inner ones (half a
    period), and the enclosed area is that of the star polygon,
    P r_e r_i sin(pi / P) (Monte Carlo, 200k points)."""
    om = oracle_mod
    P, re, ri = 5, 2.0, 1.0
    st = om.surface("star", center=(0, 0, 0), half=(0, 1, 0), radius=re, radius_internal=ri, points=P,
                    direction=(0, 0, 1))
    a0 = 2 * math.pi / P

    def at(angle, rad):        # angle from +z towards +x
        return (rad * math.sin(angle), 0.0, rad * math.cos(angle))

    ins, _ = om.surface_inside(st, [at(0, 0.99 * re), at(a0 / 2, 0.99 * ri), at(a0, 0.99 * re), at(0.3, 0.5)])
    assert ins.all()
    out, _ = om.surface_inside(st, [at(0, 1.01 * re), at(a0 / 2, 1.01 * ri), at(a0 / 2, 1.5), at(a0 / 2, 1.9)])
    assert not out.any()
    rng = np.random.default_rng(3)
    p = np.zeros((200_000, 3))
    p[:, 0] = rng.uniform(-re, re, len(p))
    p[:, 2] = rng.uniform(-re, re, len(p))
    ins, _ = om.surface_inside(st, p)
    area = ins.mean() * (2 * re) ** 2
    assert area == pytest.approx(P * re * ri * math.sin(a0 / 2), rel=0.01)


def test_inflow_rules_by_hand(oracle_mod):
    """inflow_with_predefined_positions: rejected near existing fluid; with one
    fluid and one boundary particle in range the velocity is the fluid's
    times m W_f / (m W_f + psi W_b); rejected when the velocity points
    against the boundary normal (towards the sphere's outside)."""
    om = oracle_mod
    r = 0.025
    hk = 4 * r
    o = om.Oracle(particle_radius=r)
    o.set_boundary(np.array([[0.0, 0.0, 1.0 + 0.06]]), psi=np.array([0.3]))
    o.set_fluid(np.array([[0.0, 0.0, 1.0 - 0.05]]), vel=np.array([[0.0, 0.0, -2.0]]))
    m = o.get(om.F_MASS)[0]
    ball = om.surface("sphere", center=(0, 0, 0), radius=1.2)       # normal: towards the centre
    cand = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 1.0 - 0.05 + 0.02], [5.0, 5.0, 5.0]])
    acc, vel, dmin, dens = o.inflow(ball, cand, 1.75 * r, 700.0, m)
    assert dmin[0] == pytest.approx(0.05) and dmin[1] == pytest.approx(0.02) and dmin[2] == 100.0
    wf = om.kernel_W(0.05, hk) * m
    wb = om.kernel_W(0.06, hk) * 0.3
    assert vel[0] == pytest.approx([0, 0, -2.0 * wf / (wf + wb)])
    assert dens[0] == pytest.approx(om.kernel_W(0.0, hk) * m + wf + wb)
    assert acc.tolist() == [True, False, True]          # 0: inward velocity, 1: too close, 2: empty space
    o.set(om.F_VEL, np.array([[0.0, 0.0, 2.0]]))       # outward: against the normal
    acc, _, _, _ = o.inflow(ball, cand[:1], 1.75 * r, 700.0, m)
    assert not acc[0]
    acc, _, _, _ = o.inflow(ball, cand[2:], 1.75 * r, 0.5 * om.kernel_W(0.0, hk) * m, m)
    assert not acc[0]                                   # density cap below the self contribution


def test_outflow_removes_outside_both_shapes(oracle_mod):
    om = oracle_mod
    o = om.Oracle(particle_radius=0.025)
    o.set_boundary(np.array([[9.0, 9.0, 9.0]]))
    x = np.array([[0.0, 0.5, 0.0], [0.0, 2.0, 0.0], [1.4, 0.5, 0.0], [1.4, 2.0, 0.0]])
    o.set_fluid(x)
    interior = om.surface("cylinder", center=(0, 0, 0), half=(0, 10, 0), radius=1.0)
    surf = om.surface("plane", o=(0, 1.0, 0), n=(0, -1, 0))          # inside below y = 1
    assert o.outflow_keep(interior, surf).tolist() == [True, True, True, False]


def _restricted(om, x, b, mode, **kw):
    o = om.Oracle(particle_radius=0.025, dt=0.002)
    o.set_boundary(b)
    o.set_fluid(x)
    o.set_restriction(mode, **kw)
    return o


def test_restriction_full_set_equals_unrestricted(oracle_mod):
    om = oracle_mod
    x, b = _column()
    n = len(x)
    ref = _restricted(om, x, b, 0)
    a = _restricted(om, x, b, 2, count_active=n)
    t = _restricted(om, x, b, 1, tags=np.full(n, 10000000, np.uint32))
    for _ in range(3):
        ref.step(); a.step(); t.step()
    for o in (a, t):
        for f in (om.F_POS, om.F_VEL, om.F_DENSITY, om.F_KAPPA, om.F_KAPPAV):
            assert np.array_equal(o.get(f), ref.get(f))


def test_restriction_freezes_the_inactive(oracle_mod):
    """Mode 2 with the lower half active: the particles beyond the source
    prefix keep position, velocity and density; the active-neighbor band
    gets densities but no velocity; errors average over true_particle_count."""
    om = oracle_mod
    x, b = _column()
    order = np.argsort(x[:, 1], kind="stable")
    x = x[order]
    n = len(x)
    ca, can = n // 2, 72
    o = _restricted(om, x, b, 2, count_active=ca, count_active_neighbors=can, true_particle_count=ca + can)
    o.step()
    x0, v0 = x.copy(), np.zeros_like(x)
    xs, vs, rho = o.get(om.F_POS), o.get(om.F_VEL), o.get(om.F_DENSITY)
    assert np.array_equal(xs[ca:], x0[ca:]) and np.array_equal(vs[ca:], v0[ca:])
    assert np.all(rho[ca + can:] == 0.0)            # never computed
    assert np.all(rho[ca:ca + can] > 0.0)           # source band: densities
    assert np.abs(vs[:ca]).max() > 0.0              # the active part moved
