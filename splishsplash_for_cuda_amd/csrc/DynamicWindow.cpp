// DynamicWindow.cpp -- see DynamicWindow.h.
#include "DynamicWindow.h"

#include <algorithm>
#include <cmath>
#include <string>

#include "DFSPH_CUDA.h"

namespace SPH {

namespace {

void wcheck(int rc, const char* what)
{
    if (rc != DFSPH_OK) throw DFSPHError(rc, std::string("dynamic window: ") + what + ": " + dfsph_last_error());
}

void shift(dfsph_surface& s, const double d[3])
{
    for (int a = 0; a < 3; ++a) {
        s.o[a] += (float)d[a];
        s.center[a] += (float)d[a];
    }
}

}  // namespace

DynamicWindow::DynamicWindow(void* handle, const DynamicWindowParams& p) : m_h(handle), m_p(p)
{
    const int t = p.simulation.type;
    if (t != DFSPH_SURFACE_CUBOID && t != DFSPH_SURFACE_CYLINDER && t != DFSPH_SURFACE_SPHERE)
        throw DFSPHError(DFSPH_ERR_ARG, "dynamic window: the region is a cuboid, a cylinder or a sphere");
    if (p.simulation.reversed) throw DFSPHError(DFSPH_ERR_ARG, "dynamic window: the region cannot be reversed");
    if (p.background.size() % 3) throw DFSPHError(DFSPH_ERR_ARG, "dynamic window: background is xyz triples");
    dfsph_params sp;
    wcheck(dfsph_get_params(m_h, &sp), "dfsph_get_params");
    m_r = sp.particle_radius;
    m_hk = 4.0 * m_r;
    m_rho0 = sp.density0;
    if (m_p.trigger <= 0.0) m_p.trigger = 2.5 * m_hk;                   // DFSPH_CUDA.cpp:4017
    if (m_p.max_displacement <= 0.0) m_p.max_displacement = 3.5 * m_hk; // :4021
    if (m_p.band <= 0.0) m_p.band = m_p.max_displacement + 1.5 * m_hk;  // S_bufferInterior, DynamicWindow.cu:305-307
    if (m_p.min_distance <= 0.0f) m_p.min_distance = (float)(1.75 * m_r);
    if (m_p.max_density <= 0.0f) m_p.max_density = (float)(0.7 * m_rho0);   // the inflow's 700 (DFSPH_CUDA.cpp:3903-3906)
}

dfsph_surface DynamicWindow::simulationSurface() const
{
    dfsph_surface s = m_p.simulation;
    shift(s, m_state.displacement);
    return s;
}

dfsph_surface DynamicWindow::shrunk(const dfsph_surface& in, double by)
{
    dfsph_surface s = in;
    const float b = (float)by;
    if (s.type == DFSPH_SURFACE_CUBOID) {
        // a cuboid region shrinks along the faces it moves through (x and z);
        // its y extent (floor to above the surface) stays
        s.half[0] = std::max(0.0f, s.half[0] - b);
        s.half[2] = std::max(0.0f, s.half[2] - b);
    } else {
        s.radius = std::max(0.0f, s.radius - b);
    }
    return s;
}

dfsph_surface DynamicWindow::fluidInteriorSurface() const { return shrunk(simulationSurface(), 3.0 * m_r); }

dfsph_surface DynamicWindow::innerSurface() const { return shrunk(fluidInteriorSurface(), m_p.band); }

dfsph_surface DynamicWindow::fluidSurface() const
{
    dfsph_surface s = m_p.fluid;
    shift(s, m_state.displacement);
    return s;
}

std::vector<uint8_t> DynamicWindow::inside(const dfsph_surface& s, const std::vector<float>& xyz) const
{
    std::vector<uint8_t> in(xyz.size() / 3);
    if (!in.empty()) wcheck(dfsph_surface_inside(m_h, &s, xyz.data(), (uint32_t)in.size(), in.data()), "inside");
    return in;
}

void DynamicWindow::proposedDisplacement(double out[3]) const
{
    out[0] = out[1] = out[2] = 0.0;
    if (m_p.tracked_body < 0) return;
    dfsph_rigid_body b;
    wcheck(dfsph_get_rigid_body(m_h, (uint32_t)m_p.tracked_body, &b, nullptr, nullptr), "dfsph_get_rigid_body");
    // current center of the region; y is not followed (DFSPH_CUDA.cpp:4005-4006)
    const float* c = m_p.simulation.center;
    double d[3] = {b.position[0] - (c[0] + m_state.displacement[0]), 0.0,
                   b.position[2] - (c[2] + m_state.displacement[2])};
    const double len = std::sqrt(d[0] * d[0] + d[2] * d[2]);
    if (!(len > m_p.trigger)) return;
    const double sc = len > m_p.max_displacement ? m_p.max_displacement / len : 1.0;
    for (int a = 0; a < 3; ++a) out[a] = d[a] * sc;
}

bool DynamicWindow::apply()
{
    double d[3];
    proposedDisplacement(d);
    if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) return false;
    move(d);
    return true;
}

void DynamicWindow::move(const double d[3])
{
    // the existing fluid kept: inside the fluid interior before the move
    // (the layer along the old edge had the walls' arrangement) and more than
    // 1.5 h inside the moved region's edge (SA_keptExistingFluidArea,
    // DynamicWindow.cu:906-966)
    const dfsph_surface interior_before = fluidInteriorSurface();
    for (int a = 0; a < 3; ++a) m_state.displacement[a] += d[a];
    m_state.moves++;
    m_state.loaded = m_state.removed = 0;
    const dfsph_surface range_after = shrunk(simulationSurface(), 1.5 * m_hk);
    const dfsph_surface interior = fluidInteriorSurface(), inner = innerSurface(), fl = fluidSurface();

    // 0. the region's walls follow (displace_boundaries_and_surfaces, :912-928)
    if (m_p.wall_body >= 0) {
        dfsph_rigid_body w;
        wcheck(dfsph_get_rigid_body(m_h, (uint32_t)m_p.wall_body, &w, nullptr, nullptr), "dfsph_get_rigid_body");
        for (int a = 0; a < 3; ++a) w.position[a] += d[a];
        wcheck(dfsph_set_rigid_body(m_h, (uint32_t)m_p.wall_body, &w), "dfsph_set_rigid_body");
    }

    // 1. the existing fluid outside the kept area
    uint32_t n = 0;
    wcheck(dfsph_fluid_count(m_h, &n), "dfsph_fluid_count");
    if (n) {
        std::vector<float> x(3 * (size_t)n);
        wcheck(dfsph_download(m_h, DFSPH_FIELD_POS, x.data(), 3 * n), "dfsph_download");
        std::vector<uint8_t> keep = inside(interior_before, x);
        const std::vector<uint8_t> in_range = inside(range_after, x);
        for (uint32_t i = 0; i < n; ++i) keep[i] = keep[i] && in_range[i];
        uint32_t kept = n;
        wcheck(dfsph_remove_particles(m_h, keep.data(), &kept), "dfsph_remove_particles");
        m_state.removed = n - kept;
        n = kept;
    }

    // 2. background particles of the band inside the moved fluid interior,
    // under the surface (S_fluidInterior less S_bufferInterior, and S_fluid)
    const uint32_t nb = (uint32_t)(m_p.background.size() / 3);
    std::vector<float> bg(m_p.background);
    for (uint32_t i = 0; i < nb; ++i)
        for (int a = 0; a < 3; ++a) bg[3 * i + a] += (float)m_state.displacement[a];
    const std::vector<uint8_t> in_int = inside(interior, bg), in_inner = inside(inner, bg), in_fl = inside(fl, bg);
    std::vector<float> cand;
    for (uint32_t i = 0; i < nb; ++i)
        if (in_int[i] && !in_inner[i] && in_fl[i]) cand.insert(cand.end(), &bg[3 * i], &bg[3 * i + 3]);
    uint32_t added = 0;
    if (!cand.empty())
        wcheck(dfsph_load_fluid(m_h, cand.data(), (uint32_t)(cand.size() / 3), m_p.min_distance, m_p.max_density, 0,
                                &added),
               "dfsph_load_fluid");
    m_state.loaded = added;

    // 3. settle the loaded particles
    if (added) stabilize(n);
}

void DynamicWindow::stabilize(uint32_t first_loaded)
{
    uint32_t n = 0;
    wcheck(dfsph_fluid_count(m_h, &n), "dfsph_fluid_count");
    // the loaded particles are active; before every iteration their current
    // neighbors are tagged active neighbors again, from the neighbor lists of
    // an unrestricted search (the reference re-tags each iteration from its
    // backup, DynamicWindow.cu:3383-3400, restriction by tag: mode 1), so an
    // active particle never reads the density of a particle the restricted
    // solve leaves out.  The searches keep the order (no sort), so the active
    // particles stay at [first_loaded, n).
    std::vector<uint32_t> base(n, DFSPH_TAG_UNTAGGED);
    for (uint32_t i = first_loaded; i < n; ++i) base[i] = DFSPH_TAG_ACTIVE;
    const uint32_t na = n - first_loaded;
    m_state.active = na;
    // the step history and the rigid bodies are kept and restored afterwards:
    // the settling iterations are not simulated time
    dfsph_info info;
    wcheck(dfsph_get_info(m_h, &info), "dfsph_get_info");
    uint32_t nbody = 0;
    wcheck(dfsph_rigid_body_count(m_h, &nbody), "dfsph_rigid_body_count");
    std::vector<dfsph_rigid_body> bodies(nbody);
    for (uint32_t b = 0; b < nbody; ++b)
        wcheck(dfsph_get_rigid_body(m_h, b, &bodies[b], nullptr, nullptr), "dfsph_get_rigid_body");
    // the warm starts restart from zero (DynamicWindow.cu:3311-3313); the
    // stored density_adv of the particles left out is rho0, so the average
    // error is that of the restricted set
    std::vector<float> zero(n, 0.0f), rest(n, (float)m_rho0);
    wcheck(dfsph_upload_field(m_h, DFSPH_FIELD_KAPPA, zero.data(), n), "dfsph_upload_field");
    wcheck(dfsph_upload_field(m_h, DFSPH_FIELD_KAPPAV, zero.data(), n), "dfsph_upload_field");
    wcheck(dfsph_upload_field(m_h, DFSPH_FIELD_DENSITY_ADV, rest.data(), n), "dfsph_upload_field");
    wcheck(dfsph_set_time(m_h, m_p.stab_dt, m_p.stab_dt, info.time), "dfsph_set_time");
    const unsigned N = std::max(1u, m_p.stab_iters);
    const double decay = N > 1 ? std::pow(m_p.damping_last / m_p.damping_first, 1.0 / (N - 1)) : 1.0;
    double damping = m_p.damping_first;
    std::vector<float> v(3 * (size_t)n);
    for (unsigned it = 0; it < N; ++it) {
        wcheck(dfsph_set_restriction(m_h, 0, 0, 0, -1), "dfsph_set_restriction");
        wcheck(dfsph_upload_field(m_h, DFSPH_FIELD_TAG, base.data(), n), "dfsph_upload_field");
        wcheck(dfsph_neighbor_search(m_h, 0), "dfsph_neighbor_search");
        uint32_t nbrs = 0;
        wcheck(dfsph_tag_neighbors(m_h, DFSPH_TAG_ACTIVE, DFSPH_TAG_ACTIVE_NEIGHBORS, 1, &nbrs), "dfsph_tag_neighbors");
        m_state.active_neighbors = nbrs;
        wcheck(dfsph_set_restriction(m_h, 1, na, nbrs, (int64_t)na + nbrs), "dfsph_set_restriction");
        wcheck(dfsph_compute_density_factor(m_h), "dfsph_compute_density_factor");
        wcheck(dfsph_external_forces(m_h), "dfsph_external_forces");
        wcheck(dfsph_update_time_step(m_h, m_p.stab_dt), "dfsph_update_time_step");
        wcheck(dfsph_update_vel(m_h), "dfsph_update_vel");
        uint32_t iters = 0;
        double err = 0.0;
        wcheck(dfsph_pressure_solve(m_h, m_p.stab_max_iter, m_p.stab_max_error, &iters, &err), "dfsph_pressure_solve");
        // preUpdateVelocityDamping on the active particles (:3588-3593), reduced
        // toward damping_last (reduceDampingAndClamping)
        wcheck(dfsph_download(m_h, DFSPH_FIELD_VEL, v.data(), 3 * n), "dfsph_download");
        for (size_t k = 3 * (size_t)first_loaded; k < 3 * (size_t)n; ++k) v[k] = (float)(v[k] * damping);
        wcheck(dfsph_upload_field(m_h, DFSPH_FIELD_VEL, v.data(), 3 * n), "dfsph_upload_field");
        wcheck(dfsph_update_pos(m_h), "dfsph_update_pos");
        wcheck(dfsph_step_end(m_h), "dfsph_step_end");
        damping *= decay;
    }
    wcheck(dfsph_set_restriction(m_h, 0, 0, 0, -1), "dfsph_set_restriction");
    // the active particles leave at rest (:3752-3756); the warm starts are
    // cleared (clearWarmstartAfterStabilization, the default of
    // DynamicWindow.h:266): the active neighbors, frozen, accumulate pressure
    // over the settling solves that would kick them at the next step
    wcheck(dfsph_download(m_h, DFSPH_FIELD_VEL, v.data(), 3 * n), "dfsph_download");
    std::fill(v.begin() + 3 * (size_t)first_loaded, v.end(), 0.0f);
    wcheck(dfsph_upload_field(m_h, DFSPH_FIELD_VEL, v.data(), 3 * n), "dfsph_upload_field");
    wcheck(dfsph_upload_field(m_h, DFSPH_FIELD_KAPPA, zero.data(), n), "dfsph_upload_field");
    wcheck(dfsph_upload_field(m_h, DFSPH_FIELD_KAPPAV, zero.data(), n), "dfsph_upload_field");
    wcheck(dfsph_set_time(m_h, info.h, info.h_past, info.time), "dfsph_set_time");
    for (uint32_t b = 0; b < nbody; ++b) {
        if (bodies[b].type == DFSPH_BODY_SINE) continue;   // a function of the restored time
        wcheck(dfsph_set_rigid_body(m_h, b, &bodies[b]), "dfsph_set_rigid_body");
    }
    std::fill(base.begin(), base.end(), DFSPH_TAG_UNTAGGED);
    wcheck(dfsph_upload_field(m_h, DFSPH_FIELD_TAG, base.data(), n), "dfsph_upload_field");
}

}  // namespace SPH
