// dfsph_launch.hpp -- host launch wrappers of dfsph_kernels.hip (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "dfsph_device.hpp"

namespace dfsph {

uint32_t grid_blocks_xcd(uint32_t n);

// Kernel timing bound to a dispatch: the next single-kernel sweep launch on
// this thread (the source, Jacobi, warm-start, XSPH, feature and position
// sweeps) carries the two events in its own dispatch (hipExtLaunchKernel), so
// their interval is that kernel's execution alone, without the gaps that
// separately recorded events would add around it.  The launch disarms them;
// sweep_events_armed() tells whether no such launch took place.
void arm_sweep_events(hipEvent_t start, hipEvent_t stop);
bool sweep_events_armed();
void disarm_sweep_events();
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, uint32_t n);
void launch_fill_f32(hipStream_t s, float* p, float v, uint32_t n);
size_t scan_temp_words(uint32_t n);
void launch_exclusive_scan(hipStream_t s, uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp,
                           bool clear_in = false);

void launch_nbr_ctl_init(hipStream_t s, DevCtl* c);
void launch_time_reset(hipStream_t s, DevTime* t, double h);
void launch_time_update(hipStream_t s, DevTime* t, double hf, const DevCtl* ctl, int mode,
                        double cfl_factor, double diameter, double dt_min, double dt_max, double seed);
void launch_time_step_end(hipStream_t s, DevTime* t, const Publish& pub = Publish{});
void launch_time_set(hipStream_t s, DevTime* t, double h, double hp, double time);

void launch_cell_hist(hipStream_t s, const SimConst& c, const float4* pos, uint32_t n, uint32_t* key,
                      uint32_t* slot, uint32_t* hist);
void launch_scatter_perm(hipStream_t s, const uint32_t* key, const uint32_t* slot, const uint32_t* cstart,
                         uint32_t n, uint32_t* perm);
void launch_cell_sort(hipStream_t s, const uint32_t* cstart, uint32_t ncell, uint32_t* perm);
void launch_gather_fluid(hipStream_t s, uint32_t n, const uint32_t* perm, const FluidPtrs& in, const FluidPtrs& out);
void launch_gather_f4(hipStream_t s, uint32_t n, const uint32_t* perm, const float4* in, float4* out);
void launch_neighbors(hipStream_t s, const SimConst& c, bool indirect, const float4* pos, const uint32_t* cs_f,
                      const uint32_t* perm, const float4* bpos, const uint32_t* cs_b, uint32_t n,
                      uint16_t* nbr16, uint32_t* nbrf, uint32_t* nbrb, uint32_t* cnt, uint32_t* stage_n,
                      uint32_t* stg, const Restrict& rs, uint4* part, DevCtl* ctl);
void launch_expand_lists(hipStream_t s, uint32_t n, uint32_t cap, const NbrLists& L, uint32_t* out);
void launch_boundary_psi(hipStream_t s, const SimConst& c, const float4* bpos, const uint32_t* cs_b,
                         uint32_t nb, float* psi);
void launch_set_w(hipStream_t s, float4* p, const float* w, uint32_t n);

// rigid bodies (dfsph_bodies.hip)
void launch_bodies_place(hipStream_t s, const DevBody* B, const float4* x0, const uint32_t* sbody, uint32_t ns,
                         float4* bpos, float4* bvel, unsigned long long* force);
void launch_bodies_integrate(hipStream_t s, DevBody* B, uint32_t nbody, const float4* spos,
                             const unsigned long long* force, double inv_fscale, const DevTime* t, const double* g);
void launch_solid_psi_take(hipStream_t s, const float4* bvel, const float* psi, uint32_t nb, float4* x0,
                           uint32_t first, uint32_t end);
void launch_wall_vel(hipStream_t s, float4* bvel, uint32_t n);

void launch_div_ws_init(hipStream_t s, int ws, const SimConst& c, const DevTime* t, uint32_t n,
                        const float4* pos, const float4* bpos, const NbrLists& L,
                        const float4* vel_i, float4* vel_o, const float* kws, const WsLearned& lw,
                        float* rho, float* alpha, float* radv, float* kvtot);
void launch_div_source(hipStream_t s, bool init, const SimConst& c, uint32_t n, const float4* pos,
                       const float4* bpos, const NbrLists& L, const float4* vel,
                       const float* alpha, float* radv, float* kk, float* kappaV, SolveCtl* ctl, double* partial,
                       const CtlInit& ci = CtlInit{});
// the one-block error reduction and loop decision after a loop-end sweep of n particles
void launch_solve_finalize(hipStream_t s, uint32_t n, SolveCtl* ctl, const double* partial, const Publish& pub);
void launch_jacobi(hipStream_t s, bool div, const SimConst& c, const DevTime* t, uint32_t n, const float4* pos,
                   const float4* bpos, const NbrLists& L, const float* kk, float* kappaV,
                   float4* vel, const SolveCtl* ctl);
void launch_xsph(hipStream_t s, const SimConst& c, const DevTime* t, uint32_t n, const float4* pos,
                 const NbrLists& L, const float4* vel, const float* rho, float4* acc,
                 DevCtl* ctl, float* partial);
void launch_update_vel(hipStream_t s, const DevTime* t, uint32_t n, const float4* acc, float4* vel, float* kappa,
                       const Restrict& rs);
void launch_pressure_ws(hipStream_t s, int ws, bool apply_acc, const SimConst& c, const DevTime* t, uint32_t n,
                        const float4* pos, const float4* bpos, const NbrLists& L,
                        const float4* acc, float4* vel, const float* kws, const WsLearned& lw, int rescale,
                        float* ktot);
void launch_density_source(hipStream_t s, bool init, const SimConst& c, const DevTime* t, uint32_t n,
                           const float4* pos, const float4* bpos, const NbrLists& L,
                           const float4* vel, const float* rho, float* alpha, float* radv, float* kk,
                           float* kappa, SolveCtl* ctl, double* partial, const CtlInit& ci = CtlInit{});
void launch_update_pos(hipStream_t s, const DevTime* t, uint32_t n, const float4* vel, float4* pos,
                       const float* bmin, const float* bmax, DevCtl* ctl, const Restrict& rs);
void launch_keep_inside(hipStream_t s, uint32_t n, const float4* pos, const float* bmin, const float* bmax,
                        uint32_t* keep);
void launch_keep_from_u8(hipStream_t s, uint32_t n, const uint8_t* k8, uint32_t* keep);
void launch_compact(hipStream_t s, uint32_t n, const uint32_t* keep, const uint32_t* rank, const FluidPtrs& in,
                    const FluidPtrs& out);
void launch_finish_totals(hipStream_t s, const DevTime* t, uint32_t n, int pressure, const float* inc, float* tot);
// features of the learned warm start; also copies the last step's totals
// (ktot, kvtot) into prev_p, prev_v for the sweeps' bound
void launch_features(hipStream_t s, const SimConst& c, const DevTime* t, uint32_t n, const float4* pos,
                     const float4* bpos, const NbrLists& L, const float4* vel, const float* kap,
                     const float* kapv, const float* ktot, const float* kvtot, float diameter, float* feat,
                     float* prev_p, float* prev_v);
void launch_harvest(hipStream_t s, const DevTime* t, uint32_t n, uint32_t stride, uint32_t offset, uint32_t cap,
                    const float* feat, const float* kapv, const float* kap, float diameter, float* ds_x,
                    float* ds_t);

// open boundary / dynamic window (dfsph_window.hip)
void launch_surface_keep(hipStream_t s, const float4* p, uint32_t n, const Surface& S, int want_inside, uint32_t* keep);
void launch_surface_keep_and(hipStream_t s, const float4* p, uint32_t n, const Surface& S, int want_inside,
                             uint32_t* keep);
void launch_compact_f4(hipStream_t s, uint32_t n, const uint32_t* keep, const uint32_t* rank, const float4* in,
                       float4* out);
void launch_outflow_keep(hipStream_t s, const float4* p, uint32_t n, const Surface& interior, const Surface& surface,
                         uint32_t* keep);
void launch_inflow(hipStream_t s, const SimConst& c, const float4* cand, uint32_t nc, const float4* pos,
                   const float4* vel, const uint32_t* cs_f, const uint32_t* perm, const float4* bpos,
                   const uint32_t* cs_b, const Surface& boundary, float min_dist, float max_density, float mass,
                   uint32_t* keep, float4* vel_out);
void launch_append_inflow(hipStream_t s, const float4* cand, uint32_t nc, const uint32_t* keep, const uint32_t* rank,
                          const float4* vnew, uint32_t n, float mass, uint32_t first_id, const FluidPtrs& f);
void launch_apply_delta(hipStream_t s, float4* p, uint32_t n, float dx, float dy, float dz);
void launch_count_tag(hipStream_t s, const uint32_t* tag, uint32_t n, uint32_t t, uint32_t* count);
void launch_tag_neighbors(hipStream_t s, const SimConst& c, uint32_t n, const NbrLists& L, const uint32_t* src,
                          uint32_t* dst, uint32_t tag_in, uint32_t tag_out, int untagged_only, uint32_t grid);
void launch_count_changed(hipStream_t s, const uint32_t* a, const uint32_t* b, uint32_t n, uint32_t* changed);
void launch_rest_extract(hipStream_t s, const SimConst& c, uint32_t n, const NbrLists& L, const float4* pos,
                         const float4* bpos, const float* rho, const uint32_t* tag, float limit, float boundary_limit,
                         uint32_t* keep, double* sum, uint32_t grid);
void launch_tag_classes(hipStream_t s, const uint32_t* tag, uint32_t n, uint32_t* f0, uint32_t* f1);
void launch_partition3(hipStream_t s, uint32_t n, const uint32_t* f0, const uint32_t* r0, const uint32_t* f1,
                       const uint32_t* r1, const FluidPtrs& in, const FluidPtrs& out);
void launch_keep_not_tag(hipStream_t s, const uint32_t* tag, uint32_t n, uint32_t t, uint32_t* keep);

}  // namespace dfsph
