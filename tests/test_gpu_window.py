"""Parity of the open boundary and the dynamic-window restriction (HIP path
through the C ABI) with the CPU oracle.

Tolerances:
  surface inside tests      exact, except points within 1e-5 of a border
  inflow acceptance         exact set, except candidates within 1e-4 of the
                            distance or density threshold
  inflow velocity           1e-5 of max |v|
  outflow removal           exact set by id
  tags / active-first order exact
  restricted steps          density 2e-6 rho0, velocity 5e-5 (max |v| + g h),
                            position 1e-5 r (FP32 positions), iteration
                            counts equal, frozen particles bit-identical
"""
import numpy as np
import pytest

import splishsplash_for_cuda_amd as S
from splishsplash_for_cuda_amd import scene
from gpu_util import max_rel

pytestmark = pytest.mark.gpu

TAG_ACTIVE, TAG_NBR, TAG_UNTAGGED = S.TAG_ACTIVE, S.TAG_ACTIVE_NEIGHBORS, S.TAG_UNTAGGED


def _solver(sc, psi=None, **kw):
    p = S.default_params(particle_radius=sc.particle_radius, **kw)
    s = S.Solver(p)
    s.upload_boundary(sc.boundary, psi)
    s.upload_fluid(sc.fluid)
    return s


def _pair(om, kind, **kw):
    return S.surface(kind, **kw), om.surface(kind, **kw)


def test_surface_inside_matches_oracle(oracle_mod):
    om = oracle_mod
    rng = np.random.default_rng(5)
    pts = rng.uniform(-2.5, 2.5, size=(20000, 3)).astype(np.float32)
    shapes = [
        ("plane", dict(o=(0.1, -0.2, 0.3), n=(0.6, 0.8, 0.0))),
        ("cuboid", dict(center=(0.2, 0.1, -0.3), half=(1.0, 0.7, 1.5))),
        ("cylinder", dict(center=(0.0, 0.5, 0.0), half=(0.0, 1.2, 0.0), radius=1.5)),
        ("sphere", dict(center=(0.3, -0.2, 0.1), radius=1.7)),
        ("star", dict(center=(0.0, 0.0, 0.0), half=(0.0, 2.0, 0.0), radius=1.75, radius_internal=1.0, points=5,
                      direction=(0.0, 0.0, 1.0))),
        ("star", dict(center=(0.1, 0.0, -0.2), half=(0.0, 1.0, 0.0), radius=2.2, radius_internal=0.9, points=7,
                      direction=(0.6, 0.0, 0.8))),
    ]
    s = _solver(scene.small_dam_break())
    for kind, kw in shapes:
        for rev in (False, True):
            gs, os_ = _pair(om, kind, reversed=rev, **kw)
            got = s.surface_inside(gs, pts)
            want, margin = om.surface_inside(os_, pts.astype(np.float64))
            ok = margin > 1e-5
            assert ok.mean() > 0.99
            assert np.array_equal(got[ok], want[ok]), (kind, rev)
            assert 0.05 < want.mean() < 0.95


def test_mesh_surface_inside_matches_oracle(oracle_mod):
    """DFSPH_SURFACE_MESH (MeshPerso::isInside, BufferFluidSurface.cu:224-270)
    on the device against the oracle's ray parity, for a box, an octahedron
    and a non-convex star prism, moved by their centers and reversed: the
    same answer away from the faces and from rays grazing an edge (oracle
    margin > 1e-5)."""
    import meshes
    om = oracle_mod
    rng = np.random.default_rng(11)
    c = (0.25, -0.1, 0.4)
    pts = (np.array(c) + rng.uniform(-0.7, 0.7, size=(20000, 3))).astype(np.float32)
    s = _solver(scene.small_dam_break())
    shapes = [meshes.cube_mesh((-0.4, -0.25, -0.3), (0.4, 0.25, 0.3)), meshes.octahedron(0.5),
              meshes.star_prism(7, 0.55, 0.25, 0.3, direction=(0.6, 0.0, 0.8))]
    for v, f in shapes:
        mid = s.mesh_upload(v, f)
        for rev in (False, True):
            gs = S.surface("mesh", mesh=mid, center=c, reversed=rev)
            os_ = om.surface("mesh", vertices=v, faces=f, center=c, reversed=rev)
            got = s.surface_inside(gs, pts)
            want, margin = om.surface_inside(os_, pts.astype(np.float64))
            ok = margin > 1e-5
            assert ok.mean() > 0.99
            assert np.array_equal(got[ok], want[ok]), (len(f), rev)
            assert 0.02 < want.mean() < 0.98
    with pytest.raises(S.DFSPHError):
        s.surface_inside(S.surface("mesh", mesh=99), pts[:4])
    v, f = meshes.cube_mesh((0, 0, 0), (1, 1, 1))
    with pytest.raises(S.DFSPHError):
        s.mesh_upload(v, f[:3])                        # not closed: fewer than 4 faces
    bad = f.copy()
    bad[0, 0] = 8                                      # a vertex index past the 8 vertices
    with pytest.raises(S.DFSPHError):
        s.mesh_upload(v, bad)


def test_open_boundary_with_a_mesh_interior_equals_the_cuboid(oracle_mod):
    """OpenBoundariesSimple with its interior given as a closed box mesh
    instead of the cuboid (half 0.45 m across the 0.6 m basin, so part of
    the inflow ring lies outside it): the same inflow and outflow over a few
    steps (identical states, particles added)."""
    import meshes
    states = []
    for kind in ("cuboid", "mesh"):
        sc, cand, R = _basin()
        s = _solver(sc, dt=0.002)
        if kind == "cuboid":
            interior = S.surface("cuboid", center=(0.0, 0.0, 0.0), half=(0.45, 10.0, 0.45))
        else:
            v, f = meshes.cube_mesh((-0.45, -10.0, -0.45), (0.45, 10.0, 0.45))
            interior = S.surface("mesh", mesh=s.mesh_upload(v, f))
        boundary = S.surface("cylinder", center=(0, 0, 0), half=(0, 10.0, 0), radius=R)
        fl = S.surface("plane", o=(0, 0.5, 0), n=(0, -1, 0))
        s.open_boundary_init(boundary, interior, fl, cand)
        for _ in range(4):
            s.open_boundary_apply(1.75 * sc.particle_radius, 700.0)
            s.step()
        states.append(s.state_by_id())
        assert len(states[-1]["id"]) > sc.n_fluid
    for k in ("id", "pos", "vel"):
        assert np.array_equal(states[0][k], states[1][k]), k


def _basin():
    """A cylindrical basin (r = 1.5 m boundary), fluid up to y = 0.5: the open
    boundary of OpenBoundariesSimple's cylinder configuration at a small size."""
    r = 0.04
    d = 2 * r
    R = 0.6
    # boundary: floor disk and the cylinder wall
    g = np.arange(-R, R + 1e-9, d)
    fx, fz = np.meshgrid(g, g, indexing="ij")
    floor = np.stack([fx.ravel(), np.zeros(fx.size), fz.ravel()], 1)
    floor = floor[np.hypot(floor[:, 0], floor[:, 2]) <= R + d]
    th = np.arange(0, 2 * np.pi, d / (R + d))
    hs = np.arange(d, 1.0, d)
    T, H = np.meshgrid(th, hs, indexing="ij")
    wall = np.stack([(R + d) * np.cos(T).ravel(), H.ravel(), (R + d) * np.sin(T).ravel()], 1)
    b = np.concatenate([floor, wall]).astype(np.float32)
    # fluid: inside radius R - 3r, up to y = 0.5, jittered lattice
    rng = np.random.default_rng(2)
    fl = np.stack(np.meshgrid(g, np.arange(d, 0.5, d), g, indexing="ij"), -1).reshape(-1, 3)
    fl = fl[np.hypot(fl[:, 0], fl[:, 2]) < R - 3 * r]
    fl = fl + rng.uniform(-0.01 * r, 0.01 * r, fl.shape)
    # inflow candidates: a ring of positions at radii R - 2r .. R - 0.5 r
    rr = np.array([R - 2 * r, R - r])
    tt = np.arange(0, 2 * np.pi, d / R)
    yy = np.arange(d, 0.7, d)
    RR, TT, YY = np.meshgrid(rr, tt, yy, indexing="ij")
    cand = np.stack([(RR * np.cos(TT)).ravel(), YY.ravel(), (RR * np.sin(TT)).ravel()], 1).astype(np.float32)
    return scene.Scene("basin", r, fl.astype(np.float32), b), cand, R


def test_open_boundary_inflow_and_outflow_match_oracle(oracle_mod):
    om = oracle_mod
    sc, cand, R = _basin()
    r = sc.particle_radius
    s = _solver(sc, dt=0.002, fluid_capacity=sc.n_fluid + len(cand) + 64)
    for _ in range(5):
        s.step()
    kw_b = dict(center=(0, 0, 0), half=(0, 10, 0), radius=R)
    kw_i = dict(center=(0, 0, 0), half=(0, 10, 0), radius=R - 3 * r)
    kw_s = dict(o=(0, 0.5, 0), n=(0, -1, 0))
    gb, ob = _pair(om, "cylinder", **kw_b)
    gi, oi = _pair(om, "cylinder", **kw_i)
    gsu, osu = _pair(om, "plane", **kw_s)
    kept = s.open_boundary_init(gb, gi, gsu, cand)
    # candidates kept: outside the interior and below the fluid surface
    c64 = cand.astype(np.float64)
    keep_o = ~om.surface_inside(oi, c64)[0] & om.surface_inside(osu, c64)[0]
    assert kept == int(keep_o.sum()) and kept > 0
    kc = c64[keep_o]

    # the oracle sees the device state before the call
    st0 = s.state_by_id()
    o = om.Oracle(particle_radius=r)
    bpos = s.download(S.FIELD_BOUNDARY_POS).astype(np.float64)
    o.set_boundary(bpos, s.download(S.FIELD_BOUNDARY_PSI).astype(np.float64))
    o.set_fluid(st0["pos"].astype(np.float64), vel=st0["vel"].astype(np.float64),
                mass=st0["mass"].astype(np.float64), ids=st0["id"])
    m0 = float(s.download(S.FIELD_MASS)[0])
    acc, vel, dmin, dens = o.inflow(ob, kc, 1.75 * r, 700.0, m0)
    n0 = s.info().n_fluid
    added, removed = s.open_boundary_apply(1.75 * r, 700.0, use_inflow=True, use_outflow=False)
    sharp = (np.abs(dmin - 1.75 * r) > 1e-4 * r) & (np.abs(dens - 700.0) > 1e-4 * 700.0)
    ids = s.download(S.FIELD_ID)
    pos = s.download(S.FIELD_POS)
    v = s.download(S.FIELD_VEL)
    assert s.info().n_fluid == n0 + added and removed == 0
    new = np.argsort(ids)[n0:]                         # fresh ids after every old one
    got_pos = {tuple(p) for p in pos[new].astype(np.float32)}
    want_idx = np.nonzero(acc & sharp)[0]
    want_pos = {tuple(p) for p in kc[want_idx].astype(np.float32)}
    assert want_pos <= got_pos
    assert len(got_pos - want_pos) <= int((~sharp).sum())
    assert added > 0
    # velocities of the accepted candidates
    key = {tuple(p): i for i, p in zip(new, pos[new].astype(np.float32))}
    vs = np.abs(vel).max() + 1e-6
    for c in want_idx:
        i = key[tuple(kc[c].astype(np.float32))]
        assert np.abs(v[i] - vel[c]).max() < 1e-5 * vs
    # outflow on the new state, with a narrower interior and a lower surface
    gi2, oi2 = _pair(om, "cylinder", center=(0, 0, 0), half=(0, 10, 0), radius=R - 6 * r)
    gs2, os2 = _pair(om, "plane", o=(0, 0.3, 0), n=(0, -1, 0))
    s.open_boundary_init(gb, gi2, gs2, cand)
    st1 = s.state_by_id()
    o2 = om.Oracle(particle_radius=r)
    o2.set_boundary(bpos)
    o2.set_fluid(st1["pos"].astype(np.float64), ids=st1["id"])
    keep = o2.outflow_keep(oi2, os2)
    _, removed = s.open_boundary_apply(1.75 * r, 700.0, use_inflow=False, use_outflow=True)
    assert removed == int((~keep).sum()) and removed > 0
    assert set(s.download(S.FIELD_ID).tolist()) == set(st1["id"][keep].tolist())


def test_inflow_grows_the_fluid_capacity(oracle_mod):
    """An inflow beyond the capacity grows it (change_max_particle_number,
    OpenBoundariesSimple.cu:570-573) instead of failing: a solver made with
    room for only 10 more particles takes the same inflow and then runs the
    same steps, bit for bit, as one made with room for all of them; an
    explicit dfsph_change_max_particle_number mid-run changes nothing either."""
    om = oracle_mod
    sc, cand, R = _basin()
    r = sc.particle_radius
    kw = dict(center=(0, 0, 0), half=(0, 10, 0))
    out = []
    for cap in (sc.n_fluid + 10, 3 * sc.n_fluid):
        s = _solver(sc, dt=0.002, fluid_capacity=cap)
        for _ in range(5):
            s.step()
        gb, _ = _pair(om, "cylinder", radius=R, **kw)
        gi, _ = _pair(om, "cylinder", radius=R - 3 * r, **kw)
        gsu, _ = _pair(om, "plane", o=(0, 0.5, 0), n=(0, -1, 0))
        s.open_boundary_init(gb, gi, gsu, cand)
        added, _ = s.open_boundary_apply(1.75 * r, 700.0, use_inflow=True, use_outflow=False)
        inf = s.info()
        assert added > 10 and inf.n_fluid == sc.n_fluid + added <= inf.fluid_capacity
        if cap < sc.n_fluid + added:
            assert inf.fluid_capacity >= 2 * cap
        for k in range(6):
            if k == 3 and cap > sc.n_fluid + 10:
                s.change_max_particle_number(4 * sc.n_fluid)
                assert s.info().fluid_capacity == 4 * sc.n_fluid
            s.step()
        out.append((added, s.state_by_id()))
        s.close()
    assert out[0][0] == out[1][0]
    for f in ("id", "pos", "vel", "kappa", "kappaV", "density"):
        assert np.array_equal(out[0][1][f], out[1][1][f]), f


def test_window_displacement_moves_surfaces_and_inflow(oracle_mod):
    """Moving the window by d is the same as building everything shifted by d."""
    sc, cand, R = _basin()
    r = sc.particle_radius
    res = []
    for shift in (np.zeros(3), np.array([0.0, 0.125, 0.0])):     # 0.125: exact in binary
        s = _solver(sc, dt=0.002, fluid_capacity=sc.n_fluid + len(cand) + 64)
        for _ in range(3):
            s.step()
        c0 = tuple(-shift)
        gb = S.surface("cylinder", center=c0, half=(0, 10, 0), radius=R)
        gi = S.surface("cylinder", center=c0, half=(0, 10, 0), radius=R - 6 * r)
        gsu = S.surface("plane", o=tuple(np.array([0, 0.3, 0]) - shift), n=(0, -1, 0))
        s.open_boundary_init(gb, gi, gsu, cand - shift.astype(np.float32))
        a, rm = s.open_boundary_apply(1.75 * r, 700.0, window_displacement=tuple(shift))
        res.append((a, rm, np.sort(s.download(S.FIELD_POS)[:, 1])))
    assert res[0][0] == res[1][0] and res[0][1] == res[1][1]
    assert res[0][1] > 0
    assert np.abs(res[0][2] - res[1][2]).max() < 1e-5


def _column(n=(8, 16, 8), r=0.025):
    d = 2 * r
    x = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1).reshape(-1, 3) * d + d
    x = x + np.random.default_rng(4).uniform(-0.01 * r, 0.01 * r, x.shape)
    tank = (n[0] * d + d, n[1] * d * 1.6 + d, n[2] * d + d)
    return scene.Scene("column", r, x.astype(np.float32), scene.box_surface((0, 0, 0), tank, d).astype(np.float32))


def test_tag_neighbors_and_active_first_order(oracle_mod):
    om = oracle_mod
    sc = _column()
    s = _solver(sc, dt=0.002)
    s.neighbor_search()
    pos = s.download(S.FIELD_POS)
    n = len(pos)
    tags = np.full(n, TAG_UNTAGGED, np.uint32)
    act = pos[:, 1] < np.median(pos[:, 1]) - 0.1
    tags[act] = TAG_ACTIVE
    s.upload_field(S.FIELD_TAG, tags)
    changed = s.tag_neighbors(TAG_ACTIVE, TAG_NBR, untagged_only=True)
    got = s.download(S.FIELD_TAG)
    # expected: untagged fluid neighbors of the active set (FP32 predicate)
    nf, nb, L, _ = om.neighbors_fp32(pos, sc.boundary, np.float32(4 * sc.particle_radius), max_nbr=90)
    want = tags.copy()
    for i in np.nonzero(act)[0]:
        for j in L[i, :nf[i]]:
            if tags[j] == TAG_UNTAGGED:
                want[j] = TAG_NBR
    assert np.array_equal(got, want)
    assert changed == int((want != tags).sum()) > 0
    assert s.count_tagged(TAG_ACTIVE) == int(act.sum())
    ids0 = s.download(S.FIELD_ID)
    ca, can = s.order_active_first()
    assert (ca, can) == (int((want == TAG_ACTIVE).sum()), int((want == TAG_NBR).sum()))
    t1, ids1 = s.download(S.FIELD_TAG), s.download(S.FIELD_ID)
    assert np.all(t1[:ca] == TAG_ACTIVE) and np.all(t1[ca:ca + can] == TAG_NBR)
    for cls in (TAG_ACTIVE, TAG_NBR, TAG_UNTAGGED):       # stable inside every class
        assert np.array_equal(ids1[t1 == cls], ids0[want == cls])
    assert s.remove_tagged(TAG_NBR) == can
    assert s.info().n_fluid == n - can


@pytest.mark.parametrize("mode", [1, 2])
def test_restricted_steps_match_oracle(oracle_mod, mode):
    om = oracle_mod
    sc = _column()
    r, dt = sc.particle_radius, 0.002
    o0 = om.Oracle(particle_radius=r)
    o0.set_boundary(sc.boundary.astype(np.float64))
    psi = o0.boundary_psi(sc.n_boundary)
    s = _solver(sc, psi=psi.astype(np.float32), dt=dt)
    s.neighbor_search()
    pos = s.download(S.FIELD_POS)
    n = len(pos)
    tags = np.full(n, TAG_UNTAGGED, np.uint32)
    tags[pos[:, 1] < np.median(pos[:, 1])] = TAG_ACTIVE
    s.upload_field(S.FIELD_TAG, tags)
    s.tag_neighbors(TAG_ACTIVE, TAG_NBR)
    if mode == 2:
        ca, can = s.order_active_first()
        s.set_restriction(2, ca, can, ca + can)
    else:
        s.set_restriction(1, 0, 0, -1)
    st = s.state_by_id()
    # the oracle in the device's current order (mode 2 depends on it)
    ids = s.download(S.FIELD_ID)
    order = np.argsort(np.argsort(ids))
    o = om.Oracle(particle_radius=r, dt=dt, fp32_predicate=1)
    o.set_boundary(sc.boundary.astype(np.float64), psi)
    o.set_fluid(st["pos"][order].astype(np.float64), ids=ids)
    t_dev = s.download(S.FIELD_TAG)
    if mode == 2:
        o.set_restriction(2, ca, can, ca + can)
    else:
        o.set_restriction(1, tags=t_dev)
    frozen = st["pos"].copy()
    for k in range(3):
        sg = s.step()
        so = o.step()
        assert (sg.iters_v, sg.iters) == (so.iters_v, so.iters), k
    gs = s.state_by_id()
    os_ = {f: o.get(fld) for f, fld in (("pos", om.F_POS), ("vel", om.F_VEL), ("density", om.F_DENSITY))}
    oid = o.ids()
    perm = np.argsort(oid)
    for f in os_:
        os_[f] = os_[f][perm]
    vscale = np.abs(os_["vel"]).max() + 9.81 * dt
    assert max_rel(gs["vel"], os_["vel"], vscale) < 5e-5
    assert max_rel(gs["pos"], os_["pos"], r) < 1e-5      # FP32 rounding of x += h v over 3 steps
    src = np.isin(gs["id"], ids[np.isin(t_dev, (TAG_ACTIVE, TAG_NBR))])
    assert max_rel(gs["density"][src], os_["density"][src], 1000.0) < 2e-6
    # everything outside the active set is frozen
    inactive = ~np.isin(gs["id"], ids[t_dev == TAG_ACTIVE])
    assert np.array_equal(gs["pos"][inactive], frozen[inactive])


def test_load_fluid_matches_oracle(oracle_mod):
    """dfsph_load_fluid (the dynamic window's loading) against the oracle's
    inflow rule without the direction test (a plane has no normal): the same
    accepted candidates except within 1e-4 of a threshold, velocities zero,
    the new particles appended in candidate order with fresh ids."""
    om = oracle_mod
    sc, cand, R = _basin()
    r = sc.particle_radius
    s = _solver(sc, dt=0.002)
    for _ in range(5):
        s.step()
    st0 = s.state_by_id()
    o = om.Oracle(particle_radius=r)
    o.set_boundary(s.download(S.FIELD_BOUNDARY_POS).astype(np.float64),
                   s.download(S.FIELD_BOUNDARY_PSI).astype(np.float64))
    o.set_fluid(st0["pos"].astype(np.float64), vel=st0["vel"].astype(np.float64),
                mass=st0["mass"].astype(np.float64), ids=st0["id"])
    m0 = float(s.download(S.FIELD_MASS)[0])
    any_dir = om.surface("plane", o=(0, 0, 0), n=(0, 0, 0))
    acc, _, dmin, dens = o.inflow(any_dir, cand.astype(np.float64), 1.75 * r, 1000.0, m0)
    n0 = s.info().n_fluid
    added = s.load_fluid(cand, 1.75 * r, 1000.0)
    sharp = (np.abs(dmin - 1.75 * r) > 1e-4 * r) & (np.abs(dens - 1000.0) > 1e-4 * 1000.0)
    ids = s.download(S.FIELD_ID)
    pos = s.download(S.FIELD_POS)
    v = s.download(S.FIELD_VEL)
    assert s.info().n_fluid == n0 + added and added > 0
    new = np.argsort(ids)[n0:]
    got = {tuple(p) for p in pos[new].astype(np.float32)}
    want = {tuple(p) for p in cand[acc & sharp].astype(np.float32)}
    assert want <= got and len(got - want) <= int((~sharp).sum())
    assert not v[new].any()
    s.step()


def _window_scene():
    """A 2.4 x 0.8 x 0.6 m basin (static walls) with a background buffer of
    0.3 m of water at r = 15 mm on the lattice of spacing 2 r, the simulated
    region x in [0.31, 1.25] (a cuboid about x = 0.78), its x walls (two
    plates at 0.33 and 1.23, one lattice spacing outside the first fluid
    layer) a kinematic body, and a 0.16 m box half immersed at the region's
    center, the fluid inside and next to it taken out."""
    r = 0.015
    d = 2 * r
    L, H, W = 2.4, 0.8, 0.6
    walls = scene.box_surface((0, 0, 0), (L, H, W), d).astype(np.float32)
    counts = (int(round((L - 2 * d) / d)) + 1, 10, int(round((W - 2 * d) / d)) + 1)
    bg = scene.fluid_block((d, d, d), counts, d)
    level = float(bg[:, 1].max()) + r
    cx, plate, half = 0.78, 0.45, 0.47
    ys = np.linspace(-H / 2, H / 2, int(round(H / d)) + 1)
    zs = np.linspace(-W / 2, W / 2, int(round(W / d)) + 1)
    Y, Z = np.meshgrid(ys, zs, indexing="ij")
    plates = np.concatenate([np.stack([np.full(Y.size, sx * plate), Y.ravel(), Z.ravel()], 1) for sx in (-1, 1)])
    side = 0.16
    box_c = np.array([cx, level, W / 2])
    fluid = bg[(np.abs(bg[:, 0] - cx) < plate - 0.5 * d) & (np.abs(bg - box_c).max(1) >= 0.5 * side + 1.5 * r)]
    return dict(r=r, walls=walls, bg=bg.astype(np.float32), fluid=fluid.astype(np.float32), level=level, cx=cx,
                half=half, plates=plates.astype(np.float32), wall_center=(cx, H / 2, W / 2), side=side, z=W / 2)


def _window_solver(p, sc):
    hs = S.HostSolver(p, sc["walls"], sc["fluid"])
    h = 0.5 * sc["side"]
    box = scene.box_surface((-h, -h, -h), (h, h, h), 2 * sc["r"]).astype(np.float32)
    hs.view.add_rigid_body(S.rigid_body(type=S.BODY_KINEMATIC, position=(sc["cx"], sc["level"], sc["z"])), box)
    hs.view.add_rigid_body(S.rigid_body(type=S.BODY_KINEMATIC, position=sc["wall_center"]), sc["plates"])
    window = S.surface("cuboid", center=(sc["cx"], 0.4, sc["z"]), half=(sc["half"], 1.0, 1.0))
    fluid_surf = S.surface("plane", o=(0, sc["level"], 0), n=(0, -1, 0))
    hs.set_dynamic_window(window, fluid_surf, sc["bg"], tracked_body=0, wall_body=1)
    return hs


def test_dynamic_window_follows_a_moving_body():
    """DFSPHCUDA::applyDynamicWindow over a calm basin: the simulated region
    (x in [0.31, 1.25] of a 2.4 m basin, 0.3 m of water at r = 15 mm, walled
    in x by two plates that move with it) follows a kinematic box pushed
    along x at 0.3 m/s for 1.5 s.  The region moves when the box is more
    than 2.5 h ahead, by at most 3.5 h; each move carries the plates along,
    removes the fluid left behind and along the new edge, loads the
    background buffer along the edges of the moved region and settles it.
    Checked: at least two moves, the region within 2.5 h of the box at the
    end, particles loaded, every step within the tolerances, the fluid
    count within 15 % of the start (measured: +7.7 % after two moves, the
    band loads a little more than the settled fluid it replaces), every
    particle inside the moved region, the plates moved with it, speeds below
    3 m/s (measured 0.88)."""
    sc = _window_scene()
    r, cx, half = sc["r"], sc["cx"], sc["half"]
    hk = 4 * r
    p = S.default_params(particle_radius=r, dt=0.002, use_cfl=1, dt_max=0.004, max_error=0.1, max_error_v=0.1,
                         max_neighbors=128)
    hs = _window_solver(p, sc)
    n0 = len(sc["fluid"])
    t, vmax, loaded, seen = 0.0, 0.0, 0, 0
    v = 0.3
    while t < 1.5:
        hs.view.set_rigid_body(0, S.rigid_body(type=S.BODY_KINEMATIC, position=(cx + v * t, sc["level"], sc["z"]),
                                               velocity=(v, 0.0, 0.0)))
        st = hs.step()
        t = st.time
        assert st.err / p.density0 * 100 <= p.max_error * 1.0001
        assert st.err_v * st.h / p.density0 * 100 <= p.max_error_v * 1.0001
        disp, moves, l, _ = hs.window_state()
        if moves != seen:
            loaded += l
            seen = moves
        vmax = max(vmax, float(np.abs(hs.view.download(S.FIELD_VEL)).max()))
    disp, moves, _, _ = hs.window_state()
    x = hs.view.download(S.FIELD_POS)
    n1 = len(x)
    ins = np.abs(x[:, 0] - (cx + disp[0])) <= half
    wall, _, _ = hs.view.rigid_body(1)
    print("moves %d, window at %+.3f m, box at %+.3f m, fluid %d -> %d, loaded %d, max speed %.2f m/s, "
          "inside %.4f" % (moves, disp[0], v * t, n0, n1, loaded, vmax, ins.mean()))
    assert moves >= 2 and loaded > 0
    assert abs(v * t - disp[0]) <= 2.5 * hk + 1e-6 and disp[1] == 0.0 and abs(disp[2]) < 1e-6
    assert wall.position[0] == pytest.approx(cx + disp[0], abs=1e-5)
    assert abs(n1 - n0) < 0.15 * n0
    assert ins.all()
    assert vmax < 3.0
    hs.close()
