"""The rest-fluid initializer (RestFLuidLoaderInterface,
RestFluidLoader.{h,cu}): its extraction pass against the oracle, and the
loader fitting a settled fluid to a new container."""
import numpy as np
import pytest

import splishsplash_for_cuda_amd as S
from splishsplash_for_cuda_amd import scene

pytestmark = pytest.mark.gpu


def _box(r=0.015, side=0.6):
    """a closed box of wall particles at spacing 2 r"""
    d = 2 * r
    walls = scene.box_surface((0, 0, 0), (side, side, side), d).astype(np.float32)
    return r, d, walls


def test_remove_density_peaks_matches_oracle(oracle_mod):
    """dfsph_remove_density_peaks against the oracle's pass on a jittered,
    crowded block against two walls (some points within r of a wall
    particle), at two limits: the same particles go, except those whose
    density or boundary density lies within 1e-4 of a threshold or within
    1e-4 of a listed neighbor's (FP32 against FP64 densities)."""
    om = oracle_mod
    r, d, walls = _box()
    rng = np.random.default_rng(7)
    g = np.arange(10) * 0.8 * d + 0.5 * r
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + rng.normal(0, 0.15 * r, (1000, 3))
    x = x.astype(np.float32)
    for limit in (1900.0, 1300.0):
        s = S.Solver(particle_radius=r, dt=0.002, max_neighbors=128)
        s.upload_boundary(walls)
        psi = s.download(S.FIELD_BOUNDARY_PSI)
        bpos = s.download(S.FIELD_BOUNDARY_POS)
        s.upload_fluid(x)
        s.neighbor_search(True)
        ids_before = s.download(S.FIELD_ID)
        removed, avg = s.remove_density_peaks(limit, 700.0)
        kept_ids = set(s.download(S.FIELD_ID).tolist())
        o = om.Oracle(particle_radius=r)
        o.set_boundary(bpos.astype(np.float64), psi.astype(np.float64))
        o.set_fluid(x.astype(np.float64))
        o.neighbor_search()
        keep, oavg, rho = o.rest_extract(limit, 700.0)
        nf, nb, L = o.neighbors()
        # boundary densities for the sharpness test
        xb = bpos.astype(np.float64)
        rb = np.zeros(len(x))
        for i in range(len(x)):
            for k in range(nb[i]):
                j = L[i, nf[i] + k]
                rb[i] += psi[j] * om.kernel_W(float(np.linalg.norm(x[i] - xb[j])), 4 * r)
        sharp = (np.abs(rho - limit) > 1e-4 * limit) & (np.abs(rb - 700.0) > 1e-4 * 700.0)
        for i in range(len(x)):
            for k in range(nf[i]):
                j = L[i, k]
                if abs(rho[i] - rho[j]) <= 1e-4 * rho[i]:
                    sharp[i] = sharp[j] = False
        gpu_keep = np.array([i in kept_ids for i in range(len(x))])     # ids are the upload order
        assert np.array_equal(np.sort(ids_before), np.arange(len(x)))
        assert (~keep).sum() > 10, limit
        assert np.array_equal(gpu_keep[sharp], keep[sharp]), limit
        assert abs(removed - (~keep).sum()) <= (~sharp).sum()
        if sharp.all():
            assert avg == pytest.approx(oavg, rel=1e-5)
        s.close()


def _run(hs, steps):
    vmax, err = [], []
    for _ in range(steps):
        st = hs.step()
        vmax.append(float(np.linalg.norm(hs.view.download(S.FIELD_VEL), axis=1).max()))
        err.append(st.err / 1000.0 * 100)
    return np.array(vmax), np.array(err)


def test_rest_fluid_loader_fits_a_settled_fluid_to_a_new_box():
    """The use RestFLuidLoaderInterface is made for: a fluid that has settled
    in one container (a 1 x 0.6 x 1 m tank, 0.24 m deep after 4 s at
    r = 15 mm) is the background buffer for another (a 0.6 m box whose walls
    cut through it), filled to 0.18 m.  Taken as it is, the cut fluid
    overlaps the new walls and blows up (peak speed above 10 m/s within 150
    steps).  Through the loader: the points overlapping or crowding the walls
    are extracted, the rest settled and left at rest (velocities zero), no
    point within r of a wall particle, and the first 150 steps stay below
    1 m/s, within 1.5 x of a plain lattice fill's peak."""
    r = 0.015
    d = 2 * r
    H = 0.6
    p = S.default_params(particle_radius=r, dt=0.002, use_cfl=1, dt_max=0.004, max_error=0.1, max_error_v=0.1,
                         max_neighbors=128)
    LA = 1.0
    walls_a = scene.box_surface((0, 0, 0), (LA, H, LA), d).astype(np.float32)
    counts = (int(round((LA - 2 * d) / d)) + 1, 10, int(round((LA - 2 * d) / d)) + 1)
    hs = S.HostSolver(p, walls_a, scene.fluid_block((d, d, d), counts, d).astype(np.float32))
    for _ in range(1000):
        hs.step()
    settled = hs.view.download(S.FIELD_POS)
    hs.close()
    side, off, level = 0.6, 0.2, 0.18
    walls = scene.box_surface((0, 0, 0), (side, H, side), d).astype(np.float32)
    bg = settled - np.array([off, 0.0, off], np.float32)
    bg = bg[(bg[:, 0] > 0) & (bg[:, 0] < side) & (bg[:, 2] > 0) & (bg[:, 2] < side)]
    counts = (int(round((side - 2 * d) / d)) + 1, int(round((level - d) / d)), int(round((side - 2 * d) / d)) + 1)
    lat = scene.fluid_block((d, d, d), counts, d).astype(np.float32)
    hs0 = S.HostSolver(p, walls, lat)
    v_lat, _ = _run(hs0, 150)
    hs0.close()
    hs_cut = S.HostSolver(p, walls, bg[bg[:, 1] < level])
    v_cut, _ = _run(hs_cut, 150)
    hs_cut.close()
    hs = S.HostSolver(p, walls, lat[:1])
    fluid = S.surface("plane", o=(0, level, 0), n=(0, -1, 0))
    n, passes, avg = hs.load_rest_fluid(fluid, bg)
    x = hs.view.download(S.FIELD_POS)
    v = hs.view.download(S.FIELD_VEL)
    dmin = min(float(np.linalg.norm(walls - xi, axis=1).min()) for xi in x)
    v1, e1 = _run(hs, 150)
    print("lattice fill: peak %.3f m/s; cut as it is: peak %.1f m/s; loader: %d of %d points in %d passes (mean "
          "density %.1f), nearest wall %.2f r, peak %.3f m/s, max density error %.3f %%" % (
              v_lat.max(), v_cut.max(), n, int((bg[:, 1] < level).sum()), passes, avg, dmin / r, v1.max(), e1.max()))
    assert v_cut.max() > 10.0
    assert not v.any()
    assert dmin > r
    assert v1.max() < 1.0 and v1.max() < 1.5 * v_lat.max()
    hs.close()


def test_rest_fluid_loader_edge_cases():
    """An empty background leaves an empty fluid (no pass fails on zero
    particles) and the solver's time as it was; a background entirely above
    the fluid surface is all air and leaves nothing either; a bad step size
    is refused."""
    r, d, walls = _box()
    p = S.default_params(particle_radius=r, dt=0.002, max_neighbors=128)
    hs = S.HostSolver(p, walls, np.array([[0.3, 0.3, 0.3]], np.float32))
    for _ in range(3):
        hs.step()
    t0 = hs.view.info().time
    below = S.surface("plane", o=(0, 0.2, 0), n=(0, -1, 0))
    n, passes, avg = hs.load_rest_fluid(below, np.zeros((0, 3), np.float32))
    assert n == 0 and hs.view.info().n_fluid == 0
    assert hs.view.info().time == t0 > 0.0
    g = np.arange(0.25, 0.4, d)
    above = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    n, passes, avg = hs.load_rest_fluid(below, above)
    assert n == 0 and hs.view.info().n_fluid == 0
    with pytest.raises(S.DFSPHError):
        hs.load_rest_fluid(below, above, density_start=900.0, density_end=1000.0)
    hs.close()
