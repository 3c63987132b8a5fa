"""Pins of the oracle's rigid bodies (oracle/dfsph_oracle.c, or_add_body ...)
before it checks the HIP path.  Each check is independent of the oracle's own
formulas:
  * free fall and torque-free spin: closed forms of the semi-implicit Euler
    step (v_k = v_0 + k h g, c_k = c_0 + k h v_0 + h^2 g k (k + 1) / 2;
    spin about a principal axis keeps w and turns the body by
    2 atan(h |w| / 2) per step);
  * the sinusoidal motion: c(t) = c0 + A sin(2 pi t / T) axis at every step;
  * Newton's third law: in a first step (no warm start) with no viscosity,
    the fluid's momentum change is h (M g - F) and its angular momentum
    change about the body center h (sum (x - c) x m g - T), F and T being
    the force and torque the fluid put on a container body;
  * psi of a body particle equals the wall psi of the same points
    (FluidModel::computeBoundaryPsi) when the body is far from everything.
"""
import math

import numpy as np
import pytest

R = 0.025
D = 2 * R


def box_surface(lo, hi, spacing):
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    cnt = np.maximum(np.rint((hi - lo) / spacing).astype(int), 1)
    ax = [lo[a] + (hi[a] - lo[a]) * np.arange(cnt[a] + 1) / cnt[a] for a in range(3)]
    g = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)
    on = np.zeros(len(g), bool)
    for a in range(3):
        on |= np.isclose(g[:, a], lo[a]) | np.isclose(g[:, a], hi[a])
    return g[on]


def far_walls():
    return box_surface((-1.0, -1.0, -1.0), (4.0, 4.0, 3.0), D)


def _oracle(om, **kw):
    p = dict(particle_radius=R, viscosity=0.0, dt=0.002, use_cfl=0)
    p.update(kw)
    return om.Oracle(**p)


def test_free_fall_closed_form(oracle_mod):
    om = oracle_mod
    o = _oracle(om)
    o.set_boundary(far_walls())
    o.set_fluid(np.array([[0.0, 3.5, 0.0]]))
    box = box_surface((-0.1, -0.1, -0.1), (0.1, 0.1, 0.1), D)
    v0 = (0.3, 0.5, -0.2)
    c0 = (1.0, 2.0, 1.5)
    b = om.body_desc(type=0, mass=4.0, inertia=(0.03, 0.03, 0.03), position=c0, velocity=v0)
    o.add_body(b, box)
    h, g = 0.002, np.array([0.0, -9.81, 0.0])
    for k in range(1, 26):
        o.step()
        d, F, T = o.body(0)
        vk = np.array(v0) + k * h * g
        ck = np.array(c0) + k * h * np.array(v0) + h * h * g * k * (k + 1) / 2
        assert np.allclose(d.velocity[:], vk, rtol=0, atol=1e-12)
        assert np.allclose(d.position[:], ck, rtol=0, atol=1e-12)
        assert np.abs(F).max() == 0.0 and np.abs(T).max() == 0.0
        assert np.allclose(d.rotation[:], [0, 0, 0, 1], atol=0)


def test_spin_about_principal_axis(oracle_mod):
    om = oracle_mod
    o = _oracle(om, gravity=(0.0, 0.0, 0.0))
    o.set_boundary(far_walls())
    o.set_fluid(np.array([[0.0, 3.5, 0.0]]))
    box = box_surface((-0.1, -0.05, -0.15), (0.1, 0.05, 0.15), D)
    w = 3.0
    b = om.body_desc(type=0, mass=2.0, inertia=(0.02, 0.05, 0.03), position=(1, 1, 1), angular_velocity=(0, w, 0))
    o.add_body(b, box)
    h = 0.002
    for k in range(1, 31):
        o.step()
        d, _, _ = o.body(0)
        th = 2 * k * math.atan(h * w / 2)
        assert np.allclose(d.angular_velocity[:], [0, w, 0], atol=1e-13)
        assert np.allclose(d.rotation[:], [0, math.sin(th / 2), 0, math.cos(th / 2)], atol=1e-13)
        assert np.allclose(d.position[:], [1, 1, 1], atol=0)


def test_sine_motion(oracle_mod):
    om = oracle_mod
    o = _oracle(om)
    o.set_boundary(far_walls())
    o.set_fluid(np.array([[0.0, 3.5, 0.0]]))
    plate = box_surface((-0.01, -0.2, -0.2), (0.01, 0.2, 0.2), D)
    A, T = 0.1, 1.5
    b = om.body_desc(type=2, position=(1.0, 1.0, 1.0), axis=(2.0, 0.0, 0.0), amplitude=A, period=T)
    o.add_body(b, plate)
    d, _, _ = o.body(0)
    assert np.allclose(d.position[:], [1, 1, 1]) and d.velocity[0] == pytest.approx(A * 2 * math.pi / T)
    t = 0.0
    for k in range(40):
        o.step()
        t += 0.002
        d, _, _ = o.body(0)
        assert d.position[0] == pytest.approx(1.0 + A * math.sin(2 * math.pi * t / T), abs=1e-12)
        assert d.velocity[0] == pytest.approx(A * 2 * math.pi / T * math.cos(2 * math.pi * t / T), abs=1e-12)
        assert d.position[1] == 1.0 and d.position[2] == 1.0


def _container_case(om, jitter_seed=4):
    o = _oracle(om)
    o.set_boundary(far_walls())
    tank = box_surface((0.0, 0.0, 0.0), (0.6, 0.6, 0.6), D)
    c = np.array([0.3, 0.3, 0.3])
    rng = np.random.default_rng(jitter_seed)
    g = np.stack(np.meshgrid(*(np.arange(10),) * 3, indexing="ij"), -1).reshape(-1, 3) * D + D
    x = g + rng.uniform(-0.05 * R, 0.05 * R, size=g.shape)
    o.set_fluid(x)
    o.add_body(om.body_desc(type=1, position=tuple(c)), tank - c)
    return o, x, c


def test_fluid_momentum_change_balances_the_container_force(oracle_mod):
    om = oracle_mod
    o, x, c = _container_case(om)
    h = 0.002
    o.neighbor_search()
    o.divergence_solve()
    o.external_forces()
    o.update_time_step(h)
    o.update_vel()
    o.pressure_solve()
    o.update_pos()
    _, F, T = o.body(0)
    m = o.get(om.F_MASS)[:, None]
    v = o.get(om.F_VEL)
    g = np.array([0.0, -9.81, 0.0])
    P = (m * v).sum(0)
    M = m.sum()
    assert F[1] < -0.01 * M * 9.81                     # the fluid presses on the floor
    assert np.allclose(P, h * (M * g - F), rtol=0, atol=1e-10 * h * M * 9.81)
    L = (np.cross(x - c, m * v)).sum(0)
    Lg = h * np.cross(x - c, m * g).sum(0)
    assert np.allclose(L, Lg - h * T, rtol=0, atol=1e-6 * np.abs(Lg).max())


def test_body_psi_equals_wall_psi(oracle_mod):
    om = oracle_mod
    box = box_surface((-0.1, -0.1, -0.1), (0.1, 0.1, 0.1), D)
    a = _oracle(om)
    a.set_boundary(box + 1.0)
    psi_wall = a.boundary_psi(len(box))
    b = _oracle(om)
    b.set_boundary(far_walls())
    b.add_body(om.body_desc(type=1, position=(1.0, 1.0, 1.0)), box)
    assert np.allclose(b.body_psi(0), psi_wall, rtol=1e-12)
    assert psi_wall.min() > 0
