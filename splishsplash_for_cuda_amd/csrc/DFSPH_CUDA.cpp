// DFSPH_CUDA.cpp -- C++ host classes over the C ABI of libdfsph_hip.so.
// Built with g++ (no HIP header): every GPU action goes through dfsph_hip.h.
#include "DFSPH_CUDA.h"
#include "RestFluidLoader.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>

#include "SceneLoader.h"
#include "dfsph_host.h"

namespace SPH {

namespace {

void check(int rc, const char* what)
{
    if (rc != DFSPH_OK) throw DFSPHError(rc, std::string(what) + ": " + dfsph_last_error());
}

constexpr uint32_t kSnapMagic = 0x48505344u;   // "DSPH"
// version 2 adds the warm-start totals of the last step (the learned warm
// start's bound) and the rigid bodies' states; version 1 files still load
constexpr uint32_t kSnapVersion = 2;

}  // namespace

// ---------------------------------------------------------------- DFSPHCData

DFSPHCData::DFSPHCData(const dfsph_params& p, const ParticleArrays& boundary, const ParticleArrays& fluid)
    : params(p)
{
    check(dfsph_create(&params, &m_handle), "dfsph_create");
    try {
        check(dfsph_upload_boundary(m_handle, boundary.pos.data(), boundary.mass.empty() ? nullptr : boundary.mass.data(),
                                    boundary.size()),
              "dfsph_upload_boundary");
        reset(fluid);
    } catch (...) {
        dfsph_destroy(m_handle);
        m_handle = nullptr;
        throw;
    }
}

DFSPHCData::~DFSPHCData()
{
    if (m_handle) dfsph_destroy(m_handle);
}

void DFSPHCData::reset(const ParticleArrays& fluid)
{
    const uint32_t n = fluid.size();
    check(dfsph_upload_fluid(m_handle, fluid.pos.data(), fluid.vel.empty() ? nullptr : fluid.vel.data(),
                             fluid.mass.empty() ? nullptr : fluid.mass.data(), fluid.id.empty() ? nullptr : fluid.id.data(),
                             n),
          "dfsph_upload_fluid");
    h = h_future = h_past = params.dt;
    time = 0.0;
}

void DFSPHCData::updateTimeStep(double h_fut)
{
    check(dfsph_update_time_step(m_handle, h_fut), "dfsph_update_time_step");
    h_future = h_fut;
}

void DFSPHCData::onSimulationStepEnd()
{
    check(dfsph_step_end(m_handle), "dfsph_step_end");
    h_past = h;
    h = h_future;
    time += h;
}

int DFSPHCData::checkParticlesPositions(int mode, bool report)
{
    uint32_t count = 0;
    check(dfsph_check_particles_positions(m_handle, mode == 2 ? 2 : 0, &count), "dfsph_check_particles_positions");
    if (report && count)
        std::fprintf(stderr, "checkParticlesPositions: %u particles outside the bounding box%s\n", count,
                     mode == 2 ? " (removed)" : "");
    return (int)count;
}

void DFSPHCData::zeroFluidVelocities()
{
    const uint32_t n = (uint32_t)getFluidParticlesCount();
    std::vector<float> z(3 * (size_t)n, 0.0f);
    check(dfsph_upload_field(m_handle, DFSPH_FIELD_VEL, z.data(), 3 * n), "dfsph_upload_field");
}

int DFSPHCData::getFluidParticlesCount() const
{
    uint32_t n = 0;
    check(dfsph_fluid_count(m_handle, &n), "dfsph_fluid_count");
    return (int)n;
}

int DFSPHCData::getFluidParticlesCountMax() const
{
    dfsph_info info;
    check(dfsph_get_info(m_handle, &info), "dfsph_get_info");
    return (int)info.fluid_capacity;
}

std::vector<float> DFSPHCData::readField(int field) const
{
    const uint32_t n = (uint32_t)getFluidParticlesCount();
    const bool vec = field == DFSPH_FIELD_POS || field == DFSPH_FIELD_VEL || field == DFSPH_FIELD_ACC;
    std::vector<float> out((vec ? 3 : 1) * (size_t)n);
    check(dfsph_download(m_handle, field, out.data(), (uint32_t)out.size()), "dfsph_download");
    return out;
}

ParticleArrays DFSPHCData::readFluid() const
{
    const uint32_t n = (uint32_t)getFluidParticlesCount();
    std::vector<float> x = readField(DFSPH_FIELD_POS), v = readField(DFSPH_FIELD_VEL), m = readField(DFSPH_FIELD_MASS);
    std::vector<uint32_t> id(n);
    check(dfsph_download(m_handle, DFSPH_FIELD_ID, id.data(), n), "dfsph_download");
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return id[a] < id[b]; });
    ParticleArrays out;
    out.pos.resize(3 * (size_t)n);
    out.vel.resize(3 * (size_t)n);
    out.mass.resize(n);
    out.id.resize(n);
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t i = order[k];
        for (int a = 0; a < 3; ++a) {
            out.pos[3 * k + a] = x[3 * i + a];
            out.vel[3 * k + a] = v[3 * i + a];
        }
        out.mass[k] = m[i];
        out.id[k] = id[i];
    }
    return out;
}

namespace {
std::string in_folder(const std::string& folder, const char* name)
{
    if (folder.empty() || folder.back() == '/') return folder + name;
    return folder + "/" + name;
}
}  // namespace

void DFSPHCData::write_fluid_to_file(const std::string& folder) const
{
    const std::vector<float> x = readField(DFSPH_FIELD_POS), v = readField(DFSPH_FIELD_VEL),
                             m = readField(DFSPH_FIELD_MASS);
    // header flags of the fluid set (DFSPH_c_arrays_structure.cpp:911)
    writeParticleText(in_folder(folder, "fluid_file.txt"), x, v, m, 1, 1, 0);
}

void DFSPHCData::read_fluid_from_file(const std::string& folder, bool load_velocities)
{
    ParticleArrays p;
    readParticleText(in_folder(folder, "fluid_file.txt"), p.pos, p.vel, p.mass);
    if (!load_velocities) p.vel.clear();
    // only the fluid set is replaced: the device's time-step history carries over
    dfsph_info info;
    check(dfsph_get_info(m_handle, &info), "dfsph_get_info");
    reset(p);
    check(dfsph_set_time(m_handle, info.h, info.h_past, info.time), "dfsph_set_time");
    h = h_future = info.h;
    h_past = info.h_past;
    time = info.time;
}

void DFSPHCData::write_boundaries_to_file(const std::string& folder) const
{
    dfsph_info info;
    check(dfsph_get_info(m_handle, &info), "dfsph_get_info");
    const uint32_t n = info.n_boundary;
    std::vector<float> x(3 * (size_t)n), psi(n);
    check(dfsph_download(m_handle, DFSPH_FIELD_BOUNDARY_POS, x.data(), 3 * n), "dfsph_download");
    check(dfsph_download(m_handle, DFSPH_FIELD_BOUNDARY_PSI, psi.data(), n), "dfsph_download");
    // header flags of the boundary set (:893, boundary pressure off)
    writeParticleText(in_folder(folder, "boundaries_file.txt"), x, {}, psi, 0, 0, 0);
}

uint32_t DFSPHCData::addRigidBody(const dfsph_rigid_body& body, const std::vector<float>& xyz_body,
                                  const std::vector<float>& psi)
{
    uint32_t id = 0;
    check(dfsph_add_rigid_body(m_handle, &body, xyz_body.data(), psi.empty() ? nullptr : psi.data(),
                               (uint32_t)(xyz_body.size() / 3), &id),
          "dfsph_add_rigid_body");
    return id;
}

dfsph_rigid_body DFSPHCData::getRigidBody(uint32_t id, double* force, double* torque) const
{
    dfsph_rigid_body b;
    check(dfsph_get_rigid_body(m_handle, id, &b, force, torque), "dfsph_get_rigid_body");
    return b;
}

ParticleArrays DFSPHCData::read_boundaries_from_file(const std::string& folder, bool load_velocities)
{
    ParticleArrays p;
    readParticleText(in_folder(folder, "boundaries_file.txt"), p.pos, p.vel, p.mass);
    if (!load_velocities) p.vel.clear();
    return p;
}

// ------------------------------------------------------------------ DFSPHCUDA

DFSPHCUDA::DFSPHCUDA(const dfsph_params& p, const ParticleArrays& boundary, const ParticleArrays& fluid)
    : m_maxError(p.max_error), m_maxIterations(p.max_iter), m_maxErrorV(p.max_error_v),
      m_maxIterationsV(p.max_iter_v), m_data(p, boundary, fluid), m_initialFluid(fluid),
      m_enableDivergenceSolver(p.enable_divergence != 0)
{
}

DFSPHCUDA::~DFSPHCUDA() = default;

// The sequence of DFSPHCUDA::step (DFSPH_CUDA.cpp:197-456): neighbor search
// with sort, divergence solve, external forces, (CFL,) time-step update,
// velocity update, density solve, position update, step end, removal of the
// particles that left the bounding box.
void DFSPHCUDA::setOpenBoundaries(const OpenBoundaries& ob)
{
    uint32_t kept = 0;
    check(dfsph_open_boundary_init(m_data.handle(), &ob.boundary, &ob.interior, &ob.fluid_surface,
                                   ob.inflow_xyz.empty() ? nullptr : ob.inflow_xyz.data(),
                                   (uint32_t)(ob.inflow_xyz.size() / 3), &kept),
          "dfsph_open_boundary_init");
    m_ob = ob;
    m_openBoundaries = true;
}

void DFSPHCUDA::applyOpenBoundaries()
{
    const float dist = m_ob.allowed_new_distance > 0.0f ? m_ob.allowed_new_distance
                                                        : (float)(1.75 * m_data.params.particle_radius);
    // the open boundary follows the dynamic window (dynamicWindowTotalDisplacement)
    const double* disp = m_window ? m_window->state().displacement : nullptr;
    check(dfsph_open_boundary_apply(m_data.handle(), disp, dist, m_ob.allowed_new_density, m_ob.inflow ? 1 : 0,
                                    m_ob.outflow ? 1 : 0, &m_inflowAdded, &m_outflowRemoved),
          "dfsph_open_boundary_apply");
}

void DFSPHCUDA::setDynamicWindow(const DynamicWindowParams& p)
{
    m_window.reset(new DynamicWindow(m_data.handle(), p));
}

void DFSPHCUDA::applyDynamicWindow()
{
    if (m_window) m_window->apply();
}

void DFSPHCUDA::step()
{
    void* h = m_data.handle();
    m_inflowAdded = m_outflowRemoved = 0;
    if (m_window) applyDynamicWindow();              // DFSPH_CUDA.cpp:142-150
    if (m_openBoundaries) applyOpenBoundaries();     // DFSPH_CUDA.cpp:132-140
    check(dfsph_neighbor_search(h, 1), "dfsph_neighbor_search");
    if (m_enableDivergenceSolver) {
        check(dfsph_divergence_solve(h, m_maxIterationsV, m_maxErrorV, &m_iterationsV, &m_divergenceError),
              "dfsph_divergence_solve");
    } else {
        m_iterationsV = 0;
        m_divergenceError = 0.0;
        check(dfsph_compute_density_factor(h), "dfsph_compute_density_factor");
    }
    check(dfsph_external_forces(h), "dfsph_external_forces");
    if (m_data.params.use_cfl) {
        // the new step size stays on the device; the host copy is read after
        // the pressure solve from the state its loop already brought back
        check(dfsph_cfl(h, m_data.params.dt_min, m_data.params.cfl_factor, m_data.params.dt_max, nullptr), "dfsph_cfl");
    } else {
        m_data.updateTimeStep(m_data.h);
    }
    check(dfsph_update_vel(h), "dfsph_update_vel");
    check(dfsph_pressure_solve(h, m_maxIterations, m_maxError, &m_iterations, &m_densityError), "dfsph_pressure_solve");
    if (m_data.params.use_cfl) check(dfsph_get_time(h, nullptr, &m_data.h_future, nullptr), "dfsph_get_time");
    // the search's list statistics, brought back by the pressure solve's loop
    check(dfsph_neighbor_stats(h, &m_maxNeighborsSeen, &m_neighborOverflow, nullptr, nullptr),
          "dfsph_neighbor_stats");
    check(dfsph_update_pos(h), "dfsph_update_pos");
    m_data.onSimulationStepEnd();
    m_removedLastStep = 0;
    if (m_data.params.check_positions_mode >= 0)
        m_removedLastStep = m_data.checkParticlesPositions(m_data.params.check_positions_mode, false);
    count_steps++;
    // as dfsph_step: the step is complete, but some list was truncated
    if (m_neighborOverflow)
        throw DFSPHError(DFSPH_ERR_NEIGHBOR_OVERFLOW,
                         std::to_string(m_neighborOverflow) + " particles exceed max_neighbors");
}

void DFSPHCUDA::reset()
{
    m_iterations = 0;
    m_iterationsV = 0;
    count_steps = 0;
    m_data.reset(m_initialFluid);
}

void DFSPHCUDA::handleSimulationSave(const std::string& path) const
{
    void* h = m_data.handle();
    const uint32_t n = (uint32_t)m_data.getFluidParticlesCount();
    std::vector<float> x = m_data.readField(DFSPH_FIELD_POS), v = m_data.readField(DFSPH_FIELD_VEL),
                       m = m_data.readField(DFSPH_FIELD_MASS), k = m_data.readField(DFSPH_FIELD_KAPPA),
                       kv = m_data.readField(DFSPH_FIELD_KAPPAV);
    std::vector<uint32_t> id(n);
    check(dfsph_download(h, DFSPH_FIELD_ID, id.data(), n), "dfsph_download");
    dfsph_info info;
    check(dfsph_get_info(h, &info), "dfsph_get_info");
    std::ofstream f(path, std::ios::binary);
    if (!f) throw DFSPHError(DFSPH_ERR_ARG, "cannot open " + path);
    const uint32_t hdr[4] = {kSnapMagic, kSnapVersion, n, (uint32_t)count_steps};
    const double tm[3] = {info.h, info.h_past, info.time};
    f.write(reinterpret_cast<const char*>(hdr), sizeof hdr);
    f.write(reinterpret_cast<const char*>(tm), sizeof tm);
    f.write(reinterpret_cast<const char*>(x.data()), sizeof(float) * x.size());
    f.write(reinterpret_cast<const char*>(v.data()), sizeof(float) * v.size());
    f.write(reinterpret_cast<const char*>(m.data()), sizeof(float) * m.size());
    f.write(reinterpret_cast<const char*>(k.data()), sizeof(float) * k.size());
    f.write(reinterpret_cast<const char*>(kv.data()), sizeof(float) * kv.size());
    f.write(reinterpret_cast<const char*>(id.data()), sizeof(uint32_t) * id.size());
    const std::vector<float> kt = m_data.readField(DFSPH_FIELD_KAPPA_TOTAL),
                             kvt = m_data.readField(DFSPH_FIELD_KAPPAV_TOTAL);
    f.write(reinterpret_cast<const char*>(kt.data()), sizeof(float) * kt.size());
    f.write(reinterpret_cast<const char*>(kvt.data()), sizeof(float) * kvt.size());
    uint32_t nbody = 0;
    check(dfsph_rigid_body_count(h, &nbody), "dfsph_rigid_body_count");
    f.write(reinterpret_cast<const char*>(&nbody), sizeof nbody);
    for (uint32_t b = 0; b < nbody; ++b) {
        dfsph_rigid_body rb;
        check(dfsph_get_rigid_body(h, b, &rb, nullptr, nullptr), "dfsph_get_rigid_body");
        if (rb.type == DFSPH_BODY_SINE) {
            // dfsph_set_rigid_body takes a sinusoidal body's center at rest:
            // recovered from the current center and the time (to rounding)
            const double sn = std::sin(2.0 * M_PI * info.time / rb.period);
            for (int a = 0; a < 3; ++a) rb.position[a] -= rb.amplitude * sn * rb.axis[a];
        }
        f.write(reinterpret_cast<const char*>(&rb), sizeof rb);
    }
    if (!f) throw DFSPHError(DFSPH_ERR_ARG, "write failed: " + path);
}

void DFSPHCUDA::handleSimulationLoad(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw DFSPHError(DFSPH_ERR_ARG, "cannot open " + path);
    uint32_t hdr[4];
    double tm[3];
    f.read(reinterpret_cast<char*>(hdr), sizeof hdr);
    f.read(reinterpret_cast<char*>(tm), sizeof tm);
    if (!f || hdr[0] != kSnapMagic || hdr[1] < 1 || hdr[1] > kSnapVersion)
        throw DFSPHError(DFSPH_ERR_ARG, "not a snapshot: " + path);
    const uint32_t n = hdr[2];
    ParticleArrays p;
    p.pos.resize(3 * (size_t)n);
    p.vel.resize(3 * (size_t)n);
    p.mass.resize(n);
    p.id.resize(n);
    std::vector<float> k(n), kv(n);
    f.read(reinterpret_cast<char*>(p.pos.data()), sizeof(float) * p.pos.size());
    f.read(reinterpret_cast<char*>(p.vel.data()), sizeof(float) * p.vel.size());
    f.read(reinterpret_cast<char*>(p.mass.data()), sizeof(float) * p.mass.size());
    f.read(reinterpret_cast<char*>(k.data()), sizeof(float) * n);
    f.read(reinterpret_cast<char*>(kv.data()), sizeof(float) * n);
    f.read(reinterpret_cast<char*>(p.id.data()), sizeof(uint32_t) * n);
    std::vector<float> kt, kvt;
    std::vector<dfsph_rigid_body> bodies;
    if (hdr[1] >= 2) {
        kt.resize(n);
        kvt.resize(n);
        f.read(reinterpret_cast<char*>(kt.data()), sizeof(float) * n);
        f.read(reinterpret_cast<char*>(kvt.data()), sizeof(float) * n);
        uint32_t nbody = 0;
        f.read(reinterpret_cast<char*>(&nbody), sizeof nbody);
        bodies.resize(f ? nbody : 0);
        for (auto& rb : bodies) f.read(reinterpret_cast<char*>(&rb), sizeof rb);
    }
    if (!f) throw DFSPHError(DFSPH_ERR_ARG, "truncated snapshot: " + path);
    void* h = m_data.handle();
    uint32_t have = 0;
    check(dfsph_rigid_body_count(h, &have), "dfsph_rigid_body_count");
    if (!bodies.empty() && bodies.size() != have)
        throw DFSPHError(DFSPH_ERR_STATE, "the snapshot holds " + std::to_string(bodies.size()) +
                                              " rigid bodies, the solver " + std::to_string(have));
    m_data.reset(p);
    check(dfsph_upload_field(h, DFSPH_FIELD_KAPPA, k.data(), n), "dfsph_upload_field");
    check(dfsph_upload_field(h, DFSPH_FIELD_KAPPAV, kv.data(), n), "dfsph_upload_field");
    if (!kt.empty()) {
        check(dfsph_upload_field(h, DFSPH_FIELD_KAPPA_TOTAL, kt.data(), n), "dfsph_upload_field");
        check(dfsph_upload_field(h, DFSPH_FIELD_KAPPAV_TOTAL, kvt.data(), n), "dfsph_upload_field");
    }
    check(dfsph_set_time(h, tm[0], tm[1], tm[2]), "dfsph_set_time");
    m_data.h = m_data.h_future = tm[0];
    m_data.h_past = tm[1];
    m_data.time = tm[2];
    // bodies after the time: a sinusoidal body's phase is taken from it
    for (uint32_t b = 0; b < (uint32_t)bodies.size(); ++b)
        check(dfsph_set_rigid_body(h, b, &bodies[b]), "dfsph_set_rigid_body");
    count_steps = (int)hdr[3];
}

void DFSPHCUDA::handleSimulationSaveText(const std::string& folder, bool save_liquid, bool save_boundaries) const
{
    if (save_liquid) {
        m_data.write_fluid_to_file(folder);
        dfsph_info info;
        check(dfsph_get_info(m_data.handle(), &info), "dfsph_get_info");
        const std::string path = in_folder(folder, "general_data.txt");
        std::ofstream f(path);
        if (!f) throw DFSPHError(DFSPH_ERR_ARG, "cannot write " + path);
        f.precision(17);
        f << info.h;
        if (!f) throw DFSPHError(DFSPH_ERR_ARG, "write failed: " + path);
    }
    if (save_boundaries) m_data.write_boundaries_to_file(folder);
}

double DFSPHCUDA::readGeneralData(const std::string& folder)
{
    const std::string path = in_folder(folder, "general_data.txt");
    std::ifstream f(path);
    double h = 0.0;
    if (!(f >> h) || !(h > 0.0)) throw DFSPHError(DFSPH_ERR_ARG, "no time step in " + path);
    return h;
}

void DFSPHCUDA::handleSimulationLoadText(const std::string& folder, bool load_liquid_velocities)
{
    const double h = readGeneralData(folder);
    m_data.read_fluid_from_file(folder, load_liquid_velocities);
    check(dfsph_set_time(m_data.handle(), h, h, m_data.time), "dfsph_set_time");
    m_data.h = m_data.h_future = m_data.h_past = h;
}

}  // namespace SPH

// ---------------------------------------------------------------- C entry points

namespace {
thread_local std::string g_host_err;

SPH::DFSPHCUDA* S(void* s) { return static_cast<SPH::DFSPHCUDA*>(s); }

template <class F>
int guarded(F f)
{
    try {
        f();
        return DFSPH_OK;
    } catch (const SPH::DFSPHError& e) {
        g_host_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_host_err = e.what();
        return DFSPH_ERR_STATE;
    }
}
}  // namespace

extern "C" {

void* dfsph_cuda_new(const dfsph_params* p, const float* bxyz, uint32_t nb, const float* fxyz, const float* fvel,
                     uint32_t nf, char* err, uint32_t err_len)
{
    try {
        SPH::ParticleArrays b, f;
        b.pos.assign(bxyz, bxyz + 3 * (size_t)nb);
        f.pos.assign(fxyz, fxyz + 3 * (size_t)nf);
        if (fvel) f.vel.assign(fvel, fvel + 3 * (size_t)nf);
        return new SPH::DFSPHCUDA(*p, b, f);
    } catch (const std::exception& e) {
        g_host_err = e.what();
        if (err && err_len) {
            std::strncpy(err, e.what(), err_len - 1);
            err[err_len - 1] = 0;
        }
        return nullptr;
    }
}

void dfsph_cuda_delete(void* s) { delete S(s); }

int dfsph_cuda_step(void* s, dfsph_cuda_stats* st)
{
    return guarded([&] {
        SPH::DFSPHCUDA* d = S(s);
        auto fill = [&] {
            if (!st) return;
            st->iters_v = d->getIterationsV();
            st->iters = d->getIterations();
            st->err_v = d->getDivergenceError();
            st->err = d->getDensityError();
            st->h = d->data().h_past;
            st->time = d->data().time;
            st->n_fluid = (uint32_t)d->data().getFluidParticlesCount();
            st->removed = d->getRemovedLastStep();
            st->inflow_added = d->getInflowAddedLastStep();
            st->outflow_removed = d->getOutflowRemovedLastStep();
            st->max_neighbors_seen = d->getMaxNeighborsSeen();
            st->overflow = d->getNeighborOverflow();
        };
        try {
            d->step();
        } catch (const SPH::DFSPHError& e) {
            // a truncated neighbor list is reported after a whole step: the
            // statistics describe that step
            if (e.code == DFSPH_ERR_NEIGHBOR_OVERFLOW) fill();
            throw;
        }
        fill();
    });
}

int dfsph_cuda_reset(void* s) { return guarded([&] { S(s)->reset(); }); }

int dfsph_cuda_set_open_boundaries(void* s, const dfsph_surface* boundary, const dfsph_surface* interior,
                                   const dfsph_surface* fluid_surface, const float* inflow_xyz, uint32_t n,
                                   float allowed_new_distance, float allowed_new_density, int use_inflow,
                                   int use_outflow)
{
    return guarded([&] {
        if (!boundary || !interior || !fluid_surface || (n && !inflow_xyz))
            throw SPH::DFSPHError(DFSPH_ERR_ARG, "surfaces and inflow positions are required");
        SPH::DFSPHCUDA::OpenBoundaries ob;
        ob.boundary = *boundary;
        ob.interior = *interior;
        ob.fluid_surface = *fluid_surface;
        if (n) ob.inflow_xyz.assign(inflow_xyz, inflow_xyz + 3 * (size_t)n);
        ob.allowed_new_distance = allowed_new_distance;
        ob.allowed_new_density = allowed_new_density;
        ob.inflow = use_inflow != 0;
        ob.outflow = use_outflow != 0;
        S(s)->setOpenBoundaries(ob);
    });
}

int dfsph_cuda_set_dynamic_window(void* s, const dfsph_surface* simulation, const dfsph_surface* fluid,
                                  const float* background, uint32_t n, int tracked_body, int wall_body,
                                  double trigger, double max_displacement, double band)
{
    return guarded([&] {
        if (!simulation || !fluid || (n && !background))
            throw SPH::DFSPHError(DFSPH_ERR_ARG, "surfaces and the background buffer are required");
        SPH::DynamicWindowParams p;
        p.simulation = *simulation;
        p.fluid = *fluid;
        if (n) p.background.assign(background, background + 3 * (size_t)n);
        p.tracked_body = tracked_body;
        p.wall_body = wall_body;
        p.trigger = trigger;
        p.max_displacement = max_displacement;
        p.band = band;
        S(s)->setDynamicWindow(p);
    });
}

int dfsph_cuda_dynamic_window_state(void* s, double* displacement, uint32_t* moves, uint32_t* loaded,
                                    uint32_t* removed)
{
    return guarded([&] {
        const SPH::DynamicWindow* w = S(s)->dynamicWindow();
        if (!w) throw SPH::DFSPHError(DFSPH_ERR_STATE, "no dynamic window set");
        const SPH::DynamicWindowState& st = w->state();
        if (displacement)
            for (int a = 0; a < 3; ++a) displacement[a] = st.displacement[a];
        if (moves) *moves = st.moves;
        if (loaded) *loaded = st.loaded;
        if (removed) *removed = st.removed;
    });
}

int dfsph_cuda_load_rest_fluid(void* s, const dfsph_surface* fluid, const float* background, uint32_t n,
                               float density_start, float density_end, float step_density, uint32_t stab_iters,
                               uint32_t* loaded, uint32_t* passes, double* avg_density)
{
    return guarded([&] {
        if (!fluid || (n && !background))
            throw SPH::DFSPHError(DFSPH_ERR_ARG, "the fluid surface and the background buffer are required");
        SPH::RestFluidParams p;
        p.fluid = *fluid;
        if (n) p.background.assign(background, background + 3 * (size_t)n);
        if (density_start > 0.0f) p.density_start = density_start;
        if (density_end > 0.0f) p.density_end = density_end;
        if (step_density > 0.0f) p.step_density = step_density;
        if (stab_iters) p.stab_iters = stab_iters;
        SPH::RestFluidLoader loader(S(s)->data().handle(), p);
        loader.initializeFluidToSurface();
        loader.stabilizeFluid();
        if (loaded) *loaded = loader.state().loaded;
        if (passes) *passes = loader.state().passes;
        if (avg_density) *avg_density = loader.state().avg_density;
    });
}

void dfsph_cuda_set_divergence(void* s, int enable) { S(s)->setEnableDivergenceSolver(enable != 0); }

void* dfsph_cuda_handle(void* s) { return S(s)->data().handle(); }

int dfsph_cuda_save(void* s, const char* path) { return guarded([&] { S(s)->handleSimulationSave(path); }); }

int dfsph_cuda_load(void* s, const char* path) { return guarded([&] { S(s)->handleSimulationLoad(path); }); }

int dfsph_cuda_save_text(void* s, const char* folder, int save_liquid, int save_boundaries)
{
    return guarded([&] { S(s)->handleSimulationSaveText(folder, save_liquid != 0, save_boundaries != 0); });
}

int dfsph_cuda_load_text(void* s, const char* folder, int load_liquid_velocities)
{
    return guarded([&] { S(s)->handleSimulationLoadText(folder, load_liquid_velocities != 0); });
}

void* dfsph_cuda_new_from_text(const dfsph_params* p, const char* folder, int load_liquid_velocities)
{
    SPH::DFSPHCUDA* out = nullptr;
    guarded([&] {
        if (!p || !folder) throw SPH::DFSPHError(DFSPH_ERR_ARG, "params and folder are required");
        const SPH::ParticleArrays b = SPH::DFSPHCData::read_boundaries_from_file(folder, false);
        SPH::ParticleArrays f;
        SPH::readParticleText(std::string(folder) + "/fluid_file.txt", f.pos, f.vel, f.mass);
        if (!load_liquid_velocities) f.vel.clear();
        dfsph_params q = *p;
        q.dt = SPH::DFSPHCUDA::readGeneralData(folder);
        out = new SPH::DFSPHCUDA(q, b, f);
    });
    return out;
}

namespace {
struct SceneHandle {
    SPH::SceneLoader::Scene scene;
    std::vector<float> fluid_xyz, fluid_vel, boundary_xyz;
    std::vector<dfsph_rigid_body> bodies;
    std::vector<std::vector<float>> body_xyz;     // body frame
};

void to_float(const std::vector<SPH::Vec3>& v, std::vector<float>& out)
{
    out.resize(3 * v.size());
    for (size_t i = 0; i < v.size(); ++i)
        for (int a = 0; a < 3; ++a) out[3 * i + a] = (float)v[i][a];
}
}  // namespace

void* dfsph_scene_read(const char* path, const char* base_path)
{
    SceneHandle* out = nullptr;
    guarded([&] {
        if (!path) throw SPH::DFSPHError(DFSPH_ERR_ARG, "scene path is required");
        std::string base = base_path ? base_path : "";
        if (!base_path) {
            const std::string p(path);
            const size_t slash = p.find_last_of('/');
            base = slash == std::string::npos ? "." : p.substr(0, slash);
        }
        std::unique_ptr<SceneHandle> sc(new SceneHandle);
        SPH::SceneLoader::readScene(path, sc->scene);
        std::vector<SPH::Vec3> x, v, b;
        SPH::createFluidBlocks(sc->scene, x, v);
        SPH::sampleRigidBodies(sc->scene, base, b);
        to_float(x, sc->fluid_xyz);
        to_float(v, sc->fluid_vel);
        to_float(b, sc->boundary_xyz);
        for (const SPH::SceneBody& sb : SPH::dynamicRigidBodies(sc->scene, base)) {
            sc->bodies.push_back(sb.body);
            sc->body_xyz.emplace_back();
            to_float(sb.particles, sc->body_xyz.back());
        }
        out = sc.release();
    });
    return out;
}

void dfsph_scene_free(void* sc) { delete static_cast<SceneHandle*>(sc); }

void dfsph_scene_params(const void* sc, dfsph_params* p)
{
    dfsph_default_params(p);
    SPH::sceneParams(static_cast<const SceneHandle*>(sc)->scene, *p);
}

uint32_t dfsph_scene_fluid_count(const void* sc)
{
    return (uint32_t)(static_cast<const SceneHandle*>(sc)->fluid_xyz.size() / 3);
}

void dfsph_scene_fluid(const void* sc, float* xyz, float* vel)
{
    const SceneHandle* h = static_cast<const SceneHandle*>(sc);
    if (xyz) std::copy(h->fluid_xyz.begin(), h->fluid_xyz.end(), xyz);
    if (vel) std::copy(h->fluid_vel.begin(), h->fluid_vel.end(), vel);
}

uint32_t dfsph_scene_boundary_count(const void* sc)
{
    return (uint32_t)(static_cast<const SceneHandle*>(sc)->boundary_xyz.size() / 3);
}

void dfsph_scene_boundary(const void* sc, float* xyz)
{
    const SceneHandle* h = static_cast<const SceneHandle*>(sc);
    if (xyz) std::copy(h->boundary_xyz.begin(), h->boundary_xyz.end(), xyz);
}

uint32_t dfsph_scene_body_count(const void* sc)
{
    return (uint32_t)static_cast<const SceneHandle*>(sc)->bodies.size();
}

uint32_t dfsph_scene_body(const void* sc, uint32_t i, dfsph_rigid_body* body, float* xyz_body)
{
    const SceneHandle* h = static_cast<const SceneHandle*>(sc);
    if (i >= h->bodies.size()) return 0;
    if (body) *body = h->bodies[i];
    if (xyz_body) std::copy(h->body_xyz[i].begin(), h->body_xyz[i].end(), xyz_body);
    return (uint32_t)(h->body_xyz[i].size() / 3);
}

void* dfsph_cuda_new_from_scene(const void* sc, const dfsph_params* p)
{
    SPH::DFSPHCUDA* out = nullptr;
    guarded([&] {
        const SceneHandle* h = static_cast<const SceneHandle*>(sc);
        if (!h) throw SPH::DFSPHError(DFSPH_ERR_ARG, "scene is required");
        dfsph_params q;
        if (p) q = *p; else dfsph_scene_params(sc, &q);
        SPH::ParticleArrays b, f;
        b.pos = h->boundary_xyz;
        f.pos = h->fluid_xyz;
        f.vel = h->fluid_vel;
        std::unique_ptr<SPH::DFSPHCUDA> solver(new SPH::DFSPHCUDA(q, b, f));
        for (size_t k = 0; k < h->bodies.size(); ++k) solver->data().addRigidBody(h->bodies[k], h->body_xyz[k]);
        out = solver.release();
    });
    return out;
}

int dfsph_particles_write_text(const char* path, const float* xyz, const float* vel, const float* mass, uint32_t n,
                               int has_factor_computation, int velocity_impacted_by_fluid_solver)
{
    return guarded([&] {
        if (!path || (n && (!xyz || !mass))) throw SPH::DFSPHError(DFSPH_ERR_ARG, "path, xyz and mass are required");
        std::vector<float> x(xyz, xyz + 3 * (size_t)n), m(mass, mass + n), v;
        if (vel) v.assign(vel, vel + 3 * (size_t)n);
        SPH::writeParticleText(path, x, v, m, has_factor_computation != 0, velocity_impacted_by_fluid_solver != 0, 0);
    });
}

int64_t dfsph_particles_read_text(const char* path, float* xyz, float* vel, float* mass, uint32_t capacity)
{
    int64_t n = -1;
    guarded([&] {
        if (!path) throw SPH::DFSPHError(DFSPH_ERR_ARG, "path is required");
        std::vector<float> x, v, m;
        SPH::readParticleText(path, x, v, m);
        n = (int64_t)m.size();
        if ((uint64_t)n > capacity) return;
        if (xyz) std::copy(x.begin(), x.end(), xyz);
        if (vel) std::copy(v.begin(), v.end(), vel);
        if (mass) std::copy(m.begin(), m.end(), mass);
    });
    return n;
}

const char* dfsph_cuda_last_error(void) { return g_host_err.c_str(); }

}  // extern "C"
