// RestFluidLoader.cpp -- see RestFluidLoader.h.
#include "RestFluidLoader.h"

#include <algorithm>
#include <cmath>
#include <string>
#include <unordered_map>

#include "DFSPH_CUDA.h"

namespace SPH {

namespace {

void rcheck(int rc, const char* what)
{
    if (rc != DFSPH_OK) throw DFSPHError(rc, std::string("rest fluid loader: ") + what + ": " + dfsph_last_error());
}

}  // namespace

RestFluidLoader::RestFluidLoader(void* handle, const RestFluidParams& p) : m_h(handle), m_p(p)
{
    if (p.background.size() % 3) throw DFSPHError(DFSPH_ERR_ARG, "rest fluid loader: background is xyz triples");
    if (!(p.step_density > 0.0f) || !(p.density_start >= p.density_end))
        throw DFSPHError(DFSPH_ERR_ARG, "rest fluid loader: needs step_density > 0 and density_start >= density_end");
    dfsph_params sp;
    rcheck(dfsph_get_params(m_h, &sp), "dfsph_get_params");
    m_rho0 = sp.density0;
    m_r = sp.particle_radius;
    if (m_p.boundary_limit <= 0.0f) m_p.boundary_limit = (float)(0.7 * m_rho0);
}

uint32_t RestFluidLoader::initializeFluidToSurface()
{
    m_state = RestFluidState{};
    // the background points inside the fluid surface are the candidates; those
    // outside it within one kernel radius of a candidate are air particles,
    // which give the candidates at the surface their full neighborhood and
    // are dropped at the end (TAG_AIR, RestFluidLoader.cu:2440-2520)
    const uint32_t nb = (uint32_t)(m_p.background.size() / 3);
    const float* B = m_p.background.data();
    std::vector<uint8_t> in(nb);
    if (nb) rcheck(dfsph_surface_inside(m_h, &m_p.fluid, B, nb, in.data()), "dfsph_surface_inside");
    const double hk = 4.0 * m_r;
    auto cell = [&](const float* x, int a) { return (int64_t)std::floor(x[a] / hk); };
    auto key = [](int64_t x, int64_t y, int64_t z) {
        return (uint64_t)((x & 0x1fffff) | ((y & 0x1fffff) << 21) | ((z & 0x1fffff) << 42));
    };
    std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
    for (uint32_t i = 0; i < nb; ++i)
        if (in[i]) grid[key(cell(B + 3 * i, 0), cell(B + 3 * i, 1), cell(B + 3 * i, 2))].push_back(i);
    std::vector<float> xyz;
    std::vector<uint32_t> tag;
    for (uint32_t i = 0; i < nb; ++i)
        if (in[i]) { xyz.insert(xyz.end(), B + 3 * i, B + 3 * i + 3); tag.push_back(DFSPH_TAG_UNTAGGED); }
    m_state.candidates = (uint32_t)tag.size();
    const double hk2 = hk * hk;
    for (uint32_t i = 0; i < nb; ++i) {
        if (in[i]) continue;
        const float* x = B + 3 * i;
        bool near = false;
        for (int dx = -1; dx <= 1 && !near; ++dx)
            for (int dy = -1; dy <= 1 && !near; ++dy)
                for (int dz = -1; dz <= 1 && !near; ++dz) {
                    auto it = grid.find(key(cell(x, 0) + dx, cell(x, 1) + dy, cell(x, 2) + dz));
                    if (it == grid.end()) continue;
                    for (uint32_t j : it->second) {
                        const float* y = B + 3 * j;
                        const double ex = x[0] - y[0], ey = x[1] - y[1], ez = x[2] - y[2];
                        if (ex * ex + ey * ey + ez * ez < hk2) { near = true; break; }
                    }
                }
        if (near) { xyz.insert(xyz.end(), x, x + 3); tag.push_back(DFSPH_TAG_AIR); }
    }
    const uint32_t n0 = (uint32_t)tag.size();
    // the upload restarts the step history: the solver's time goes on as it was
    dfsph_info info;
    rcheck(dfsph_get_info(m_h, &info), "dfsph_get_info");
    if (n0 > info.fluid_capacity && info.fluid_capacity > 0)
        rcheck(dfsph_change_max_particle_number(m_h, n0 + n0 / 4), "dfsph_change_max_particle_number");
    rcheck(dfsph_upload_fluid(m_h, xyz.data(), nullptr, nullptr, nullptr, n0), "dfsph_upload_fluid");
    rcheck(dfsph_set_time(m_h, info.h, info.h_past, info.time), "dfsph_set_time");
    if (n0) rcheck(dfsph_upload_field(m_h, DFSPH_FIELD_TAG, tag.data(), n0), "dfsph_upload_field");
    // extraction passes at a limit stepping down to density_end
    float limit = m_p.density_start;
    for (;;) {
        limit = std::max(limit - m_p.step_density, m_p.density_end);
        uint32_t removed = 0;
        do {
            if (m_state.passes >= m_p.max_passes) break;
            rcheck(dfsph_neighbor_search(m_h, 1), "dfsph_neighbor_search");
            rcheck(dfsph_remove_density_peaks(m_h, limit, m_p.boundary_limit, &removed, &m_state.avg_density),
                   "dfsph_remove_density_peaks");
            m_state.passes++;
        } while (removed);
        if (limit <= m_p.density_end || m_state.avg_density < m_p.density_end || m_state.passes >= m_p.max_passes)
            break;
    }
    // the air goes; the candidates left are the fluid
    uint32_t air = 0;
    rcheck(dfsph_remove_tagged(m_h, DFSPH_TAG_AIR, &air), "dfsph_remove_tagged");
    rcheck(dfsph_fluid_count(m_h, &m_state.loaded), "dfsph_fluid_count");
    std::vector<uint32_t> untagged(m_state.loaded, DFSPH_TAG_UNTAGGED);
    if (m_state.loaded)
        rcheck(dfsph_upload_field(m_h, DFSPH_FIELD_TAG, untagged.data(), m_state.loaded), "dfsph_upload_field");
    return m_state.loaded;
}

void RestFluidLoader::stabilizeFluid()
{
    uint32_t n = 0;
    rcheck(dfsph_fluid_count(m_h, &n), "dfsph_fluid_count");
    if (!n) return;
    dfsph_info info;
    rcheck(dfsph_get_info(m_h, &info), "dfsph_get_info");
    uint32_t nbody = 0;
    rcheck(dfsph_rigid_body_count(m_h, &nbody), "dfsph_rigid_body_count");
    std::vector<dfsph_rigid_body> bodies(nbody);
    for (uint32_t b = 0; b < nbody; ++b)
        rcheck(dfsph_get_rigid_body(m_h, b, &bodies[b], nullptr, nullptr), "dfsph_get_rigid_body");
    std::vector<float> zero(n, 0.0f);
    rcheck(dfsph_upload_field(m_h, DFSPH_FIELD_KAPPA, zero.data(), n), "dfsph_upload_field");
    rcheck(dfsph_upload_field(m_h, DFSPH_FIELD_KAPPAV, zero.data(), n), "dfsph_upload_field");
    rcheck(dfsph_set_restriction(m_h, 0, 0, 0, -1), "dfsph_set_restriction");
    rcheck(dfsph_set_time(m_h, m_p.stab_dt, m_p.stab_dt, info.time), "dfsph_set_time");
    const unsigned N = std::max(1u, m_p.stab_iters);
    const double decay = N > 1 ? std::pow(m_p.damping_last / m_p.damping_first, 1.0 / (N - 1)) : 1.0;
    double damping = m_p.damping_first;
    std::vector<float> v(3 * (size_t)n);
    for (unsigned it = 0; it < N; ++it) {
        rcheck(dfsph_neighbor_search(m_h, 1), "dfsph_neighbor_search");
        uint32_t iv = 0;
        double ev = 0.0;
        rcheck(dfsph_divergence_solve(m_h, m_p.stab_max_iter_v, m_p.stab_max_error_v, &iv, &ev),
               "dfsph_divergence_solve");
        rcheck(dfsph_external_forces(m_h), "dfsph_external_forces");
        rcheck(dfsph_update_time_step(m_h, m_p.stab_dt), "dfsph_update_time_step");
        rcheck(dfsph_update_vel(m_h), "dfsph_update_vel");
        uint32_t iters = 0;
        double err = 0.0;
        rcheck(dfsph_pressure_solve(m_h, m_p.stab_max_iter, m_p.stab_max_error, &iters, &err), "dfsph_pressure_solve");
        // preUpdateVelocityDamping on every particle, reduced toward damping_last
        rcheck(dfsph_download(m_h, DFSPH_FIELD_VEL, v.data(), 3 * n), "dfsph_download");
        for (float& x : v) x = (float)(x * damping);
        rcheck(dfsph_upload_field(m_h, DFSPH_FIELD_VEL, v.data(), 3 * n), "dfsph_upload_field");
        rcheck(dfsph_update_pos(m_h), "dfsph_update_pos");
        rcheck(dfsph_step_end(m_h), "dfsph_step_end");
        damping *= decay;
    }
    // at rest: velocities and warm starts zero (clearWarmstartAfterStabilization,
    // RestFluidLoader.cu:6520-6524), the time and the bodies as before
    std::fill(v.begin(), v.end(), 0.0f);
    rcheck(dfsph_upload_field(m_h, DFSPH_FIELD_VEL, v.data(), 3 * n), "dfsph_upload_field");
    rcheck(dfsph_upload_field(m_h, DFSPH_FIELD_KAPPA, zero.data(), n), "dfsph_upload_field");
    rcheck(dfsph_upload_field(m_h, DFSPH_FIELD_KAPPAV, zero.data(), n), "dfsph_upload_field");
    rcheck(dfsph_set_time(m_h, info.h, info.h_past, info.time), "dfsph_set_time");
    for (uint32_t b = 0; b < nbody; ++b) {
        if (bodies[b].type == DFSPH_BODY_SINE) continue;   // a function of the restored time
        rcheck(dfsph_set_rigid_body(m_h, b, &bodies[b]), "dfsph_set_rigid_body");
    }
}

}  // namespace SPH
