// DFSPH_CUDA.h -- C++ host classes of the MI355X DFSPH solver.
//
// Keeps the host API of the reference's GPU path: SPH::DFSPHCData (time-step
// state and particle sets, DFSPH_c_arrays_structure.h:297-496) and
// SPH::DFSPHCUDA with step() / reset() and the iteration and tolerance
// members of TimeStep (DFSPH_CUDA.h:26-154).  The classes own a handle of
// libdfsph_hip.so and call the GPU only through its C ABI
// (include/dfsph_hip.h); this file and DFSPH_CUDA.cpp include no HIP
// header and build with a plain C++ compiler.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include <memory>

#include "DynamicWindow.h"
#include "dfsph_hip.h"

namespace SPH {

// Error raised when a C ABI call fails (the reference exits the process,
// DFSPH_macro_cuda.h:21-29).
class DFSPHError : public std::runtime_error {
public:
    DFSPHError(int code, const std::string& what) : std::runtime_error(what), code(code) {}
    int code;
};

// Host-side particle arrays in the interleaved xyz layout of the C ABI.
struct ParticleArrays {
    std::vector<float> pos;     // 3n
    std::vector<float> vel;     // 3n (may be empty: zero)
    std::vector<float> mass;    // n  (may be empty: 0.8 (2r)^3 rho0)
    std::vector<uint32_t> id;   // n  (may be empty: 0..n-1)
    uint32_t size() const { return (uint32_t)(pos.size() / 3); }
};

class DFSPHCData {
public:
    // DFSPHCData(FluidModel*) (DFSPH_c_arrays_structure.cpp:852-944): the
    // boundary set is uploaded first (it fixes the grid, initGridOffset
    // :1050-1098), then the fluid.
    DFSPHCData(const dfsph_params& params, const ParticleArrays& boundary, const ParticleArrays& fluid);
    ~DFSPHCData();
    DFSPHCData(const DFSPHCData&) = delete;
    DFSPHCData& operator=(const DFSPHCData&) = delete;

    // DFSPHCData::reset (:975-1048): reload the fluid, reset the step history.
    void reset(const ParticleArrays& fluid);
    // DFSPHCData::updateTimeStep / onSimulationStepEnd (:1100-1118)
    void updateTimeStep(double h_fut);
    void onSimulationStepEnd();
    // DFSPHCData::checkParticlesPositions (DFSPH_c_arrays_structure.h:486-492):
    // 0 report, 2 remove; returns the number of particles outside the box.
    int checkParticlesPositions(int mode = 0, bool report = true);
    // DFSPHCData::zeroFluidVelocities
    void zeroFluidVelocities();

    int getFluidParticlesCount() const;
    int getFluidParticlesCountMax() const;
    double get_current_timestep() const { return h; }
    double getSimulationTime() const { return time; }

    // read_UnifiedParticleSet_cuda: fluid state reordered by particle id
    ParticleArrays readFluid() const;
    std::vector<float> readField(int field) const;   // device order

    // UnifiedParticleSet text files in `folder` (DFSPH_c_arrays_structure.cpp:
    // 1196-1323): write_fluid_to_file writes fluid_file.txt, read_fluid_from_file
    // replaces the fluid with it (velocities zero unless load_velocities),
    // write_boundaries_to_file writes boundaries_file.txt (mass column = psi).
    void write_fluid_to_file(const std::string& folder) const;
    void read_fluid_from_file(const std::string& folder, bool load_velocities);
    void write_boundaries_to_file(const std::string& folder) const;
    // read_boundaries_from_file (:1294-1323) as arrays for the constructor:
    // the boundary set of a solver is static.
    static ParticleArrays read_boundaries_from_file(const std::string& folder, bool load_velocities = false);

    // The dynamic bodies (vector_dynamic_bodies_data, DFSPH_c_arrays_structure.cpp:
    // 897-905): add one from its body-frame particles (psi empty: computed),
    // read its state and the fluid force / torque of the last step
    // (compute_fluid_impact_on_dynamic_body_cuda).  step() moves them.
    uint32_t addRigidBody(const dfsph_rigid_body& body, const std::vector<float>& xyz_body,
                          const std::vector<float>& psi = {});
    dfsph_rigid_body getRigidBody(uint32_t id, double* force = nullptr, double* torque = nullptr) const;

    void* handle() const { return m_handle; }

    // mirrors of the device time-step state
    double h = 0.0, h_future = 0.0, h_past = 0.0, time = 0.0;
    dfsph_params params;

private:
    void* m_handle = nullptr;
};

class DFSPHCUDA {
public:
    DFSPHCUDA(const dfsph_params& params, const ParticleArrays& boundary, const ParticleArrays& fluid);
    virtual ~DFSPHCUDA();

    // DFSPHCUDA::step (DFSPH_CUDA.cpp:82-794)
    virtual void step();
    // DFSPHCUDA::reset (DFSPH_CUDA.cpp:796-817)
    virtual void reset();

    bool getEnableDivergenceSolver() const { return m_enableDivergenceSolver; }
    void setEnableDivergenceSolver(bool val) { m_enableDivergenceSolver = val; }

    unsigned int getIterations() const { return m_iterations; }
    unsigned int getIterationsV() const { return m_iterationsV; }
    double getDensityError() const { return m_densityError; }
    double getDivergenceError() const { return m_divergenceError; }
    int getRemovedLastStep() const { return m_removedLastStep; }
    // largest neighbor list of this step's search, and the particles whose
    // list was truncated at max_neighbors (step() throws when there are any)
    uint32_t getMaxNeighborsSeen() const { return m_maxNeighborsSeen; }
    uint32_t getNeighborOverflow() const { return m_neighborOverflow; }

    // DFSPHCUDA::zeroFluidVelocities
    void zeroFluidVelocities() { m_data.zeroFluidVelocities(); }
    // DFSPHCUDA::handleSimulationSave / handleSimulationLoad (DFSPH_CUDA.cpp:1887-1979):
    // binary snapshot with positions, velocities, masses, ids, kappa, kappaV and
    // the step history, so that a resumed run continues bit for bit.
    void handleSimulationSave(const std::string& path) const;
    void handleSimulationLoad(const std::string& path);
    // The reference's own handleSimulationSave(save_liquid, -, save_boundaries)
    // / handleSimulationLoad(load_liquid, load_liquid_velocities, ...) on the
    // text files of `folder`: fluid_file.txt + general_data.txt (h) and
    // boundaries_file.txt.  Loading sets h = h_past = h_future = the saved h
    // (updateTimeStep + onSimulationStepEnd twice, DFSPH_CUDA.cpp:1958-1962)
    // and keeps the simulated time; kappa and kappaV restart at zero.
    void handleSimulationSaveText(const std::string& folder, bool save_liquid = true, bool save_boundaries = true) const;
    void handleSimulationLoadText(const std::string& folder, bool load_liquid_velocities = true);
    // the time step saved in folder/general_data.txt
    static double readGeneralData(const std::string& folder);

    // OpenBoundariesSimple as DFSPHCUDA::applyOpenBoundaries drives it
    // (DFSPH_CUDA.cpp:3889-3924): the surfaces and inflow positions are set
    // once (OpenBoundariesSimple::init), then every step() starts with the
    // inflow / outflow pass (applyOpenBoundary with an allowed new distance of
    // 1.75 r and density 700 unless given).
    struct OpenBoundaries {
        dfsph_surface boundary, interior, fluid_surface;
        std::vector<float> inflow_xyz;      // 3 n
        float allowed_new_distance = 0.0f;  // 0: 1.75 r
        float allowed_new_density = 700.0f;
        bool inflow = true, outflow = true;
    };
    void setOpenBoundaries(const OpenBoundaries& ob);
    void applyOpenBoundaries();
    // DFSPHCUDA::applyDynamicWindow (DFSPH_CUDA.cpp:3988-4139): every step()
    // first moves the window when its tracked body asks for it (loading and
    // stabilization included); the open boundary follows the window.
    void setDynamicWindow(const DynamicWindowParams& p);
    void applyDynamicWindow();
    const DynamicWindow* dynamicWindow() const { return m_window.get(); }
    uint32_t getInflowAddedLastStep() const { return m_inflowAdded; }
    uint32_t getOutflowRemovedLastStep() const { return m_outflowRemoved; }

    DFSPHCData& data() { return m_data; }
    const DFSPHCData& data() const { return m_data; }

    // tolerances and caps (TimeStep members, DFSPH_CUDA.h:43-49)
    double m_maxError;
    unsigned int m_maxIterations;
    double m_maxErrorV;
    unsigned int m_maxIterationsV;

protected:
    DFSPHCData m_data;
    ParticleArrays m_initialFluid;
    unsigned int m_iterations = 0;
    unsigned int m_iterationsV = 0;
    double m_densityError = 0.0;
    double m_divergenceError = 0.0;
    int m_removedLastStep = 0;
    bool m_enableDivergenceSolver = true;
    int count_steps = 0;
    bool m_openBoundaries = false;
    OpenBoundaries m_ob;
    uint32_t m_inflowAdded = 0, m_outflowRemoved = 0;
    uint32_t m_maxNeighborsSeen = 0, m_neighborOverflow = 0;
    std::unique_ptr<DynamicWindow> m_window;
};

}  // namespace SPH
