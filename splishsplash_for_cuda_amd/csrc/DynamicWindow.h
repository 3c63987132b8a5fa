// DynamicWindow.h -- the dynamic window of the reference (DynamicWindowInterface,
// SPlisHSPlasH/DFSPH/DynamicWindow.cu:150-237, driven by
// DFSPHCUDA::applyDynamicWindow, DFSPH_CUDA.cpp:3988-4139).
//
// The simulated region (a cuboid, cylinder or sphere) follows a point of
// interest, a rigid body, in x and z.  When the body is more than `trigger`
// from the region's center, the region moves toward it (at most
// `max_displacement` at a time) and:
//   0. the region's walls (a kinematic rigid body) move with it;
//   1. the existing fluid is kept inside the fluid interior (the region less
//      3 r) of before the move and more than 1.5 h inside the moved region,
//      the rest removed;
//   2. particles of a background buffer of fluid at rest, which moves with
//      the region, are loaded into the band (`band` wide) along the edge of
//      the moved fluid interior where no fluid is near
//      (DynamicWindow::loadDataToSimulation, :2624-3140, with the open
//      boundary's inflow rule: dfsph_load_fluid);
//   3. the loaded particles are settled by density-solver iterations
//      restricted to them and their neighbors (restriction by tag, the
//      neighbors re-tagged every iteration), with a decaying velocity damping
//      (DynamicWindow::stabilizeFluid method 0, :3160-3760, with the driver's
//      settings: 10 iterations of 3 ms, damping 0.8 -> 0.2, density
//      tolerance 0.05 %, no divergence solver), the other particles frozen;
//      the loaded particles then leave at rest and the warm starts are
//      cleared.
// The open boundary of the region's edge, when set, follows the region
// (dynamicWindowTotalDisplacement, OpenBoundariesSimple.cu:553-584).
//
// Host code over the C ABI only (no HIP header).
#pragma once

#include <cstdint>
#include <vector>

#include "dfsph_hip.h"

namespace SPH {

struct DynamicWindowParams {
    dfsph_surface simulation{};     // the simulated region at zero displacement (cuboid, cylinder or sphere)
    dfsph_surface fluid{};          // where the background is fluid (e.g. below the rest surface)
    std::vector<float> background;  // 3 n: the background buffer at rest, at zero displacement
    int tracked_body = -1;          // rigid body followed in x and z (-1: none, the window stays)
    int wall_body = -1;             // kinematic body of the region's walls, moved with the region
                                    // (the boundaries DynamicWindow displaces, :862-900); -1: none
    double trigger = 0.0;           // move once the body is this far from the center (0: 2.5 h)
    double max_displacement = 0.0;  // cap of one move (0: 3.5 h)
    double band = 0.0;              // width of the loading band inside the fluid interior (0: max_displacement + 1.5 h)
    float min_distance = 0.0f;      // loading: no fluid nearer than this (0: 1.75 r)
    float max_density = 0.0f;       // loading: at most this density (0: 0.7 rho0, the inflow's 700)
    // stabilization (DynamicWindow::stabilizeFluid, the driver's settings)
    unsigned stab_iters = 10;
    double stab_dt = 0.003;
    double damping_first = 0.8, damping_last = 0.2;
    double stab_max_error = 0.05;   // percent
    unsigned stab_max_iter = 100;
};

struct DynamicWindowState {
    double displacement[3] = {0.0, 0.0, 0.0};   // dynamicWindowTotalDisplacement
    uint32_t moves = 0;                         // moves so far
    uint32_t loaded = 0, removed = 0;           // particles of the last move
    uint32_t active = 0, active_neighbors = 0;  // the last stabilization's restriction
};

class DynamicWindow {
public:
    DynamicWindow(void* handle, const DynamicWindowParams& p);

    // the displacement the driver would apply now (x and z toward the tracked
    // body, capped), or zero when the body is within the trigger distance
    void proposedDisplacement(double out[3]) const;
    // move the region by d: remove, load, stabilize
    void move(const double d[3]);
    // applyDynamicWindow: move when the tracked body asks for it; true if moved
    bool apply();

    const DynamicWindowState& state() const { return m_state; }
    const DynamicWindowParams& params() const { return m_p; }
    // the surfaces at the current displacement
    dfsph_surface simulationSurface() const;
    dfsph_surface fluidInteriorSurface() const;   // the region less 3 r
    dfsph_surface innerSurface() const;           // the fluid interior less the loading band
    dfsph_surface fluidSurface() const;

private:
    void stabilize(uint32_t first_loaded);
    // the surface moved `by` inward (x and z faces of a cuboid, the radius otherwise)
    static dfsph_surface shrunk(const dfsph_surface& s, double by);
    std::vector<uint8_t> inside(const dfsph_surface& s, const std::vector<float>& xyz) const;

    void* m_h;
    DynamicWindowParams m_p;
    DynamicWindowState m_state;
    double m_hk = 0.0, m_r = 0.0, m_rho0 = 1000.0;
};

}  // namespace SPH
