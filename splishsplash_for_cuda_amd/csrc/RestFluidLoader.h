// RestFluidLoader.h -- the rest-fluid initializer of the reference
// (RestFLuidLoaderInterface, SPlisHSPlasH/DFSPH/RestFluidLoader.h:31-309,
// RestFluidLoader.cu:118-207 and :1505-7960).
//
// A volume of fluid at rest for any boundary: the solver's boundary is set
// (walls, bodies), a background buffer of fluid points larger than the volume
// is given, and
//   1. initializeFluidToSurface: the fluid becomes the background points
//      inside the fluid surface, with the background points outside it
//      within one kernel radius of them as air particles (density sources
//      only, dropped at the end); then density-extraction passes
//      (dfsph_remove_density_peaks) remove the points the boundary alone
//      makes denser than 0.7 rho0 and, at a density limit stepping down from
//      density_start to density_end, the locally densest points above it,
//      until a pass at density_end removes nothing or the mean density falls
//      below density_end (RestFLuidLoader::tagDataToSurface,
//      RestFluidLoader.cu:2700-3130);
//   2. stabilizeFluid: damped solver iterations on the whole fluid (method 0
//      of RestFLuidLoader::stabilizeFluid, :5772-6530, with the settings of
//      the reference's driver, DFSPH_CUDA.cpp:2123-2145: both solvers and
//      external forces at a fixed 3 ms, the velocities damped from 0.8 to
//      0.2), after which the fluid is at rest: velocities and warm starts
//      zero, the solver's time and rigid bodies as before.
//
// The reference steps its density limit once per pass and confirms removals
// in a separate kernel; here a limit is kept until a pass removes nothing,
// and one pass removes only local density maxima, so neighbours are not
// removed together.
//
// Host code over the C ABI only (no HIP header).
#pragma once

#include <cstdint>
#include <vector>

#include "dfsph_hip.h"

namespace SPH {

struct RestFluidParams {
    dfsph_surface fluid{};          // where the fluid is (e.g. a cuboid, or a plane: below a level)
    std::vector<float> background;  // 3 n: fluid points at rest (a lattice), covering the volume
    float density_start = 1900.0f;  // TaggingParameters defaults (RestFluidLoader.h:100-105)
    float density_end = 1000.0f;
    float step_density = 50.0f;
    float boundary_limit = 0.0f;    // 0: 0.7 rho0 (particle_selection_rule_1_kernel)
    unsigned max_passes = 1000;
    // stabilization (StabilizationParameters, method 0)
    unsigned stab_iters = 10;
    double stab_dt = 0.003;
    double damping_first = 0.8, damping_last = 0.2;
    double stab_max_error = 0.05;   // percent
    unsigned stab_max_iter = 100;
    double stab_max_error_v = 0.1;  // percent; the divergence solver runs (useDivergenceSolver,
    unsigned stab_max_iter_v = 100; // DFSPH_CUDA.cpp:2134)
};

struct RestFluidState {
    uint32_t candidates = 0;        // background points inside the fluid surface (air not counted)
    uint32_t loaded = 0;            // fluid particles after the extraction
    uint32_t passes = 0;            // extraction passes run
    double avg_density = 0.0;       // mean density after the last pass
};

class RestFluidLoader {
public:
    RestFluidLoader(void* handle, const RestFluidParams& p);
    // replaces the solver's fluid; returns the particle count
    uint32_t initializeFluidToSurface();
    void stabilizeFluid();
    const RestFluidState& state() const { return m_state; }

private:
    void* m_h;
    RestFluidParams m_p;
    RestFluidState m_state;
    double m_rho0 = 1000.0, m_r = 0.0;
};

}  // namespace SPH
