// dfsph_capi.hip -- the C ABI of libdfsph_hip.so (declared in include/dfsph_hip.h).
//
// A handle owns every device buffer of one simulation (the role of
// SPH::DFSPHCData and its UnifiedParticleSet members, DFSPH_c_arrays_structure.h:96-496)
// and one HIP stream.  Convergence of the Jacobi loops is decided on the
// device (FP64 reductions, dfsph_kernels.hip); the host queues iterations in
// batches and reads the loop state back asynchronously, instead of the
// reference's blocking device-to-host copy per iteration (DFSPH_core_cuda.cu:779,1325).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dfsph_hip.h"
#include "dfsph_launch.hpp"
#include "dfsph_mlp_api.hpp"

using dfsph::DevCtl;
using dfsph::DevTime;
using dfsph::SimConst;
using dfsph::SolveCtl;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return fail(DFSPH_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CHK(x)                   \
    do {                         \
        int r_ = (x);            \
        if (r_ != DFSPH_OK) return r_; \
    } while (0)

constexpr double kCellMargin = 1.001;   // cell edge = 1.001 h_k (conservative binning)

struct Handle {
    dfsph_params p{};
    hipStream_t s = nullptr;          // the stream all work is queued on
    hipStream_t own_s = nullptr;      // the handle's own stream (s unless the caller set one)
    SimConst c{};
    bool grid_ready = false;
    uint32_t n = 0, nb = 0, cap = 0, tmp_cap = 0;
    uint32_t ncell_alloc = 0;

    // fluid (double-buffered persistent fields)
    float4* pos[2] = {nullptr, nullptr};
    float* kap[2] = {nullptr, nullptr};
    float* kapv[2] = {nullptr, nullptr};
    float* ktot[2] = {nullptr, nullptr};      // step totals (kappa h_f^2), learned warm start I/O
    float* kvtot[2] = {nullptr, nullptr};     // step totals (kappaV h)
    uint32_t* id[2] = {nullptr, nullptr};
    uint32_t* tag[2] = {nullptr, nullptr};    // dynamic-window / open-boundary tags
    // solver restriction (DFSPHCData::restriction_mode, count_active,
    // count_active_neighbors, true_particle_count)
    int rmode = 0;
    uint32_t r_active = 0, r_active_nbrs = 0;
    int64_t true_count = -1;
    // open boundary: surfaces, inflow positions, applied window displacement
    bool ob_ready = false;
    dfsph::Surface ob_boundary{}, ob_interior{}, ob_surface{};
    float4* ob_cand = nullptr;
    uint32_t ob_n = 0;
    uint32_t *ob_k = nullptr, *ob_r = nullptr, *ob_tmp = nullptr;   // inflow scratch, sized at init
    float4* ob_vnew = nullptr;
    double ob_disp[3] = {0.0, 0.0, 0.0};
    uint32_t next_id = 0;             // first id not yet used by any particle
    int pc = 0;                       // current buffer of pos / kappa / kappaV / id
    float4* vel[2] = {nullptr, nullptr};
    int vc = 0;                       // current velocity buffer
    float *rho = nullptr, *alpha = nullptr, *radv = nullptr, *kk = nullptr;
    float4* acc = nullptr;
    uint32_t* cnt = nullptr;
    uint16_t* nbr16 = nullptr;          // 16-bit fluid lists (blocks with a neighbor set)
    uint32_t *nbrf = nullptr, *nbrb = nullptr;   // 32-bit fluid / boundary lists
    uint32_t *stage_n = nullptr, *stg = nullptr;   // per-block staged sets
    // sort / scan scratch
    uint32_t *key = nullptr, *slot = nullptr, *perm = nullptr, *hist = nullptr, *cstart = nullptr;
    uint32_t *keep = nullptr, *rank = nullptr;
    uint8_t* keep8 = nullptr;
    uint32_t* scan_tmp = nullptr;
    size_t scan_words = 0;

    // boundary (walls cell-sorted once; with rigid bodies, walls + body
    // particles re-sorted at every neighbor search)
    float4* bpos = nullptr;
    uint32_t* bcstart = nullptr;
    // rigid bodies: body particles follow the walls in the boundary set
    uint32_t nbs = 0, ns = 0;                 // wall particles, body particles
    std::vector<dfsph::DevBody> bodies;       // host copies (ranges, motion parameters)
    dfsph::DevBody* dbody = nullptr;          // device state
    float4* x0 = nullptr;                     // body-frame positions + psi
    uint32_t* sbody = nullptr;                // body of each body particle
    float4* braw = nullptr;                   // walls (cell order), then body particles at their places
    float4* bvraw = nullptr;                  // velocity records of braw
    float4* bvel = nullptr;                   // velocity records in cell order (bpos order)
    unsigned long long* sforce = nullptr;     // reactions on the body particles, fixed point
    float fscale = 0.0f;                      // fixed-point units per newton
    uint32_t* tile_rank = nullptr;    // SimConst::tile_rank
    uint4* nbr_part = nullptr;        // per-block neighbor statistics of the search
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};

    // device-side control
    DevCtl* ctl = nullptr;
    DevTime* tm = nullptr;
    double* partial = nullptr;
    float* partial_f = nullptr;
    double* rest_sum = nullptr;       // dfsph_remove_density_peaks: density sum of the kept particles
    struct MeshRec {                  // dfsph_mesh_upload: triangles (3 float4 each) and bounding box
        float4* tri = nullptr;
        uint32_t ntri = 0;
        float bmin[3], bmax[3];
    };
    std::vector<MeshRec> meshes;
    uint32_t partial_n = 0;
    DevCtl* hctl = nullptr;           // pinned
    DevTime* htm = nullptr;           // pinned
    // loop state written by the device into pinned host memory (hrec; drec is
    // its device address), numbered by seq; last = the copy of the last one seen
    dfsph::HostRecord* hrec = nullptr;
    dfsph::HostRecord* drec = nullptr;
    uint32_t seq = 0;
    dfsph::HostRecord last{};
    // ABI calls on this handle (counted by H()); the op of the last record
    // wait and of the last step end, for the no-synchronisation fast paths
    uint64_t ops = 0, rec_op = 0, end_op = 0, pos_op = 0, nbr_op = 0;
    uint32_t end_seq = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};

    bool pend_vel = false;            // deferred cuda_update_vel (fused into the pressure warm start)
    bool lists_valid = false;
    int lists_rmode = 0;              // restriction mode the lists were built under
    bool acc_valid = false;
    bool pred_valid = false;          // features / predictions of this step computed
    // learned warm start gate: the warm-start mode of the current step (set by
    // its first solve after a search) and the last solves' iteration counts
    int ws_step = DFSPH_WARMSTART_REFERENCE;
    bool ws_step_set = false;
    uint32_t last_iters_v = UINT32_MAX, last_iters = UINT32_MAX;

    dfsph::MlpState* mlp = nullptr;   // learned warm start

    // per-kernel event timing (dfsph_profile_*)
    struct ProfRec {
        int kind;
        hipEvent_t a, b;
        bool bound;                   // events carried by the sweep's own dispatch
    };
    uint32_t prof_mask = 0;
    std::vector<ProfRec> prof_pending;
    std::vector<hipEvent_t> prof_pool;
    uint32_t prof_launches[DFSPH_KERNEL_COUNT] = {};
    double prof_ms[DFSPH_KERNEL_COUNT] = {};

    ~Handle();
};

// Kernel ids whose timed region is one sweep kernel: their events travel in
// that kernel's dispatch.  The others (sort, search + its finalize, the MLP)
// span several kernels and are bracketed by event records on the stream.
constexpr uint32_t kBoundKinds = (1u << DFSPH_KERNEL_DIV_WS_INIT) | (1u << DFSPH_KERNEL_DIV_INIT) |
                                 (1u << DFSPH_KERNEL_DIV_JACOBI) | (1u << DFSPH_KERNEL_DIV_LOOP_END) |
                                 (1u << DFSPH_KERNEL_XSPH) | (1u << DFSPH_KERNEL_PRESSURE_WS) |
                                 (1u << DFSPH_KERNEL_DENS_INIT) | (1u << DFSPH_KERNEL_DENS_JACOBI) |
                                 (1u << DFSPH_KERNEL_DENS_LOOP_END) | (1u << DFSPH_KERNEL_UPDATE_POS) |
                                 (1u << DFSPH_KERNEL_FEATURES);

// Event pair for one launch of kernel `kind` when it is being timed;
// returns the record index or -1.
int pbeg(Handle* h, int kind)
{
    if (!(h->prof_mask & (1u << kind))) return -1;
    hipEvent_t e[2];
    for (int k = 0; k < 2; ++k) {
        if (!h->prof_pool.empty()) {
            e[k] = h->prof_pool.back();
            h->prof_pool.pop_back();
        } else if (hipEventCreateWithFlags(&e[k], hipEventDisableSystemFence) != hipSuccess) {
            // timing events without the system-scope fence: that fence writes the
            // L2 back at every record and would slow the timed kernel itself
            return -1;
        }
    }
    const bool bound = (kBoundKinds >> kind) & 1u;
    if (bound)
        dfsph::arm_sweep_events(e[0], e[1]);
    else
        (void)hipEventRecord(e[0], h->s);
    h->prof_pending.push_back({kind, e[0], e[1], bound});
    return (int)h->prof_pending.size() - 1;
}

void pend(Handle* h, int idx)
{
    if (idx < 0) return;
    auto& r = h->prof_pending[idx];
    if (!r.bound) {
        (void)hipEventRecord(r.b, h->s);
    } else if (dfsph::sweep_events_armed()) {
        dfsph::disarm_sweep_events();   // nothing was launched (no particles): no record
        r.kind = -1;
    }
}

void pdrop(Handle* h, int idx)
{
    if (idx >= 0) h->prof_pending[idx].kind = -1;
}

// Turn the recorded pairs into durations: all of them after a stream
// synchronisation, else only the leading pairs that have completed.
void prof_collect(Handle* h, bool synced = true)
{
    size_t k = 0;
    for (; k < h->prof_pending.size(); ++k) {
        auto& r = h->prof_pending[k];
        if (!synced && hipEventQuery(r.b) != hipSuccess) break;
        if (r.kind >= 0) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
                h->prof_launches[r.kind]++;
                h->prof_ms[r.kind] += ms;
            }
        }
        h->prof_pool.push_back(r.a);
        h->prof_pool.push_back(r.b);
    }
    h->prof_pending.erase(h->prof_pending.begin(), h->prof_pending.begin() + (std::ptrdiff_t)k);
}

// Wait until the device has published record `seq`, then keep a copy.  The
// stream is queried now and then so that a failed launch ends the wait.
int wait_record(Handle* h, uint32_t seq)
{
    for (uint32_t k = 1;; ++k) {
        if (__atomic_load_n(&h->hrec->seq, __ATOMIC_ACQUIRE) == seq) break;
        if ((k & 1023u) == 0u) {
            const hipError_t e = hipStreamQuery(h->s);
            if (e == hipSuccess) {
                if (__atomic_load_n(&h->hrec->seq, __ATOMIC_ACQUIRE) == seq) break;
                return fail(DFSPH_ERR_STATE, "the device did not publish the loop state");
            }
            if (e != hipErrorNotReady) return fail(DFSPH_ERR_HIP, std::string("stream: ") + hipGetErrorString(e));
        }
        __builtin_ia32_pause();
    }
    std::memcpy(&h->last, (const void*)h->hrec, sizeof(dfsph::HostRecord));
    h->rec_op = h->ops;
    return DFSPH_OK;
}

template <class T>
void dfree(T*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

Handle::~Handle()
{
    if (s) (void)hipStreamSynchronize(s);
    for (int b = 0; b < 2; ++b) {
        dfree(pos[b]); dfree(kap[b]); dfree(kapv[b]); dfree(id[b]); dfree(tag[b]); dfree(vel[b]);
        dfree(ktot[b]); dfree(kvtot[b]);
        if (ev[b]) (void)hipEventDestroy(ev[b]);
    }
    dfree(rho); dfree(alpha); dfree(radv); dfree(kk); dfree(acc); dfree(cnt); dfree(nbrf); dfree(nbrb); dfree(nbr16); dfree(stage_n); dfree(stg); dfree(ob_cand);
    dfree(ob_k); dfree(ob_r); dfree(ob_tmp); dfree(ob_vnew);
    dfree(key); dfree(slot); dfree(perm); dfree(hist); dfree(cstart); dfree(keep); dfree(rank);
    dfree(keep8); dfree(scan_tmp); dfree(bpos); dfree(bcstart); dfree(tile_rank); dfree(nbr_part);
    dfree(dbody); dfree(x0); dfree(sbody); dfree(braw); dfree(bvraw); dfree(bvel); dfree(sforce);
    dfree(ctl); dfree(tm); dfree(partial); dfree(partial_f); dfree(rest_sum);
    for (auto& m : meshes) dfree(m.tri);
    if (hctl) (void)hipHostFree(hctl);
    if (htm) (void)hipHostFree(htm);
    if (hrec) (void)hipHostFree(hrec);
    if (mlp) dfsph::mlp_destroy(mlp);
    for (auto& r : prof_pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : prof_pool) (void)hipEventDestroy(e);
    if (own_s) (void)hipStreamDestroy(own_s);
}

template <class T>
int dalloc(T*& p, size_t count)
{
    dfree(p);
    if (count == 0) count = 1;
    HIPCHK(hipMalloc((void**)&p, count * sizeof(T)));
    return DFSPH_OK;
}

Handle* H(void* hv)
{
    Handle* h = static_cast<Handle*>(hv);
    ++h->ops;
    return h;
}

dfsph::Restrict restriction(const Handle* h)
{
    return dfsph::Restrict{h->rmode, h->r_active, h->r_active + h->r_active_nbrs, h->tag[h->pc]};
}

dfsph::NbrLists lists(const Handle* h)
{
    const dfsph::Solids sol{h->ns ? h->bvel : nullptr, h->sforce, h->fscale};
    return dfsph::NbrLists{h->nbr16, h->nbrf, h->nbrb, h->cnt, h->stage_n, h->stg, restriction(h), sol};
}

int ensure_scan(Handle* h, uint32_t n)
{
    const size_t w = dfsph::scan_temp_words(n);
    if (w > h->scan_words) {
        CHK(dalloc(h->scan_tmp, w));
        h->scan_words = w;
    }
    return DFSPH_OK;
}

int ensure_tmp(Handle* h, uint32_t n)
{
    if (n <= h->tmp_cap && h->key) return DFSPH_OK;
    CHK(dalloc(h->key, (size_t)n + 1));     // n + 1: also class flags / ranks with a sentinel
    CHK(dalloc(h->slot, (size_t)n + 1));
    CHK(dalloc(h->perm, n));
    CHK(dalloc(h->keep, (size_t)n + 1));
    CHK(dalloc(h->rank, (size_t)n + 1));
    CHK(dalloc(h->keep8, n));
    h->tmp_cap = n;
    CHK(ensure_scan(h, n + 1));
    return DFSPH_OK;
}

int ensure_partials(Handle* h, uint32_t n)
{
    const uint32_t g = dfsph::grid_blocks_xcd(n);
    if (g <= h->partial_n && h->partial) return DFSPH_OK;
    CHK(dalloc(h->partial, g));
    CHK(dalloc(h->partial_f, g));
    h->partial_n = g;
    return DFSPH_OK;
}

// counting sort of n points into the cell table cs (ncell+1 entries);
// perm[k] = index of the point at sorted slot k (stable)
int counting_sort(Handle* h, const float4* pts, uint32_t n, uint32_t* cs)
{
    const uint32_t nc = h->c.ncell;
    dfsph::launch_cell_hist(h->s, h->c, pts, n, h->key, h->slot, h->hist);
    dfsph::launch_exclusive_scan(h->s, h->hist, cs, nc + 1, h->scan_tmp, true);   // leaves hist zeroed
    dfsph::launch_scatter_perm(h->s, h->key, h->slot, cs, n, h->perm);
    dfsph::launch_cell_sort(h->s, cs, nc, h->perm);
    HIPCHK(hipGetLastError());
    return DFSPH_OK;
}

dfsph::FluidPtrs fluid_ptrs(Handle* h, int b, int vb)
{
    return dfsph::FluidPtrs{h->pos[b], h->vel[vb], h->kap[b], h->kapv[b], h->ktot[b], h->kvtot[b], h->id[b], h->tag[b]};
}

int flush_vel(Handle* h)
{
    if (!h->pend_vel) return DFSPH_OK;
    dfsph::launch_update_vel(h->s, h->tm, h->n, h->acc, h->vel[h->vc], h->kap[h->pc], restriction(h));
    h->pend_vel = false;
    HIPCHK(hipGetLastError());
    return DFSPH_OK;
}

int sync(Handle* h)
{
    HIPCHK(hipStreamSynchronize(h->s));
    return DFSPH_OK;
}

int read_ctl(Handle* h)
{
    HIPCHK(hipMemcpyAsync(h->hctl, h->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipMemcpyAsync(h->htm, h->tm, sizeof(DevTime), hipMemcpyDeviceToHost, h->s));
    return sync(h);
}

// Body particles at their bodies' state, then walls + body particles sorted
// into the boundary cell table (positions and velocity records in cell order).
int place_bodies(Handle* h)
{
    dfsph::launch_bodies_place(h->s, h->dbody, h->x0, h->sbody, h->ns, h->braw + h->nbs, h->bvraw + h->nbs,
                               h->sforce);
    CHK(counting_sort(h, h->braw, h->nb, h->bcstart));
    dfsph::launch_gather_f4(h->s, h->nb, h->perm, h->braw, h->bpos);
    dfsph::launch_gather_f4(h->s, h->nb, h->perm, h->bvraw, h->bvel);
    HIPCHK(hipGetLastError());
    return DFSPH_OK;
}

int neighbor_search(Handle* h, bool sort)
{
    if (!h->grid_ready) return fail(DFSPH_ERR_STATE, "no boundary set uploaded");
    CHK(flush_vel(h));
    h->pred_valid = false;
    if (h->ns) CHK(place_bodies(h));
    if (h->n == 0) {
        dfsph::launch_nbr_ctl_init(h->s, h->ctl);     // otherwise the search's finalize writes the statistics
        h->lists_valid = true;
        return DFSPH_OK;
    }
    const int ps = pbeg(h, DFSPH_KERNEL_SORT);
    CHK(counting_sort(h, h->pos[h->pc], h->n, h->cstart));
    if (sort) {
        const int o = h->pc ^ 1, vo = h->vc ^ 1;
        dfsph::launch_gather_fluid(h->s, h->n, h->perm, fluid_ptrs(h, h->pc, h->vc), fluid_ptrs(h, o, vo));
        h->pc = o;
        h->vc = vo;
    }
    pend(h, ps);
    const int pn = pbeg(h, DFSPH_KERNEL_NEIGHBORS);
    dfsph::launch_neighbors(h->s, h->c, !sort, h->pos[h->pc], h->cstart, h->perm, h->bpos, h->bcstart, h->n,
                            h->nbr16, h->nbrf, h->nbrb, h->cnt, h->stage_n, h->stg, restriction(h), h->nbr_part,
                            h->ctl);
    pend(h, pn);
    HIPCHK(hipGetLastError());
    h->lists_valid = true;
    h->lists_rmode = h->rmode;
    h->acc_valid = false;
    h->nbr_op = h->ops;
    h->ws_step_set = false;
    return DFSPH_OK;
}

// The warm-start mode of this step: the configured one, except that a gated
// learned warm start falls back to the reference warm start after a step whose
// solves both stopped at their minimum iterations (DFSPH_LEARNED_ABOVE_FLOORS).
int step_warm_start(Handle* h)
{
    if (!h->ws_step_set) {
        int ws = h->p.warm_start_mode;
        if (ws == DFSPH_WARMSTART_LEARNED && h->p.learned_gate == DFSPH_LEARNED_ABOVE_FLOORS &&
            h->last_iters_v != UINT32_MAX && h->last_iters != UINT32_MAX &&
            h->last_iters_v <= h->p.min_iter_v && h->last_iters <= h->p.min_iter)
            ws = DFSPH_WARMSTART_REFERENCE;
        h->ws_step = ws;
        h->ws_step_set = true;
    }
    return h->ws_step;
}

// Queue Jacobi iterations in batches; the device decides when the loop ends
// (kernels of later iterations exit at once).  Two readback slots keep one
// batch in flight while the previous batch's loop state is inspected.
// The loop control a solve's first sweep sets (block 0 of that sweep).
dfsph::CtlInit solve_ctl(Handle* h, bool div, uint32_t max_iter, double max_error)
{
    // error average over true_particle_count when set (DFSPH_core_cuda.cu:791-795, 1339-1343)
    const double n_div = h->true_count >= 0 ? (double)h->true_count : (double)h->n;
    return dfsph::CtlInit{div ? &h->ctl->div : &h->ctl->dens, h->tm, max_error * 0.01 * h->p.density0, n_div,
                          div ? 1 : 0, div ? h->p.min_iter_v : h->p.min_iter, max_iter};
}

// The loop control was set by the solve's first sweep (solve_ctl).
int jacobi_loop(Handle* h, bool div, uint32_t max_iter, double max_error, uint32_t* iters, double* err)
{
    SolveCtl* sc = div ? &h->ctl->div : &h->ctl->dens;
    const uint32_t min_iter = div ? h->p.min_iter_v : h->p.min_iter;
    (void)max_error;
    if (h->n == 0 || max_iter == 0) {
        if (iters) *iters = 0;
        if (err) *err = 0.0;
        return DFSPH_OK;
    }
    auto jacobi = [&]() -> int {
        const int pr = pbeg(h, div ? DFSPH_KERNEL_DIV_JACOBI : DFSPH_KERNEL_DENS_JACOBI);
        if (div)
            dfsph::launch_jacobi(h->s, true, h->c, h->tm, h->n, h->pos[h->pc], h->bpos, lists(h), h->kk,
                                 h->kapv[h->pc], h->vel[h->vc], sc);
        else
            dfsph::launch_jacobi(h->s, false, h->c, h->tm, h->n, h->pos[h->pc], h->bpos, lists(h), h->kk,
                                 nullptr, h->vel[h->vc], sc);
        pend(h, pr);
        return pr;
    };
    auto loop_end = [&]() {
        const int pr = pbeg(h, div ? DFSPH_KERNEL_DIV_LOOP_END : DFSPH_KERNEL_DENS_LOOP_END);
        if (div)
            dfsph::launch_div_source(h->s, false, h->c, h->n, h->pos[h->pc], h->bpos, lists(h), h->vel[h->vc],
                                     h->alpha, h->radv, h->kk, h->kapv[h->pc], sc, h->partial);
        else
            dfsph::launch_density_source(h->s, false, h->c, h->tm, h->n, h->pos[h->pc], h->bpos, lists(h),
                                         h->vel[h->vc], h->rho, h->alpha, h->radv, h->kk, h->kap[h->pc], sc,
                                         h->partial);
        pend(h, pr);                    // the timed region is the sweep alone
        dfsph::launch_solve_finalize(h->s, h->n, sc, h->partial, dfsph::Publish{h->ctl, h->tm, h->drec, ++h->seq});
    };
    // The minimum iterations are always run.  After that, the velocity sweep
    // of the next iteration is queued before the host waits for the loop
    // state of the current one (published by the finalize kernel into host
    // memory), so the wait hides behind it; the source sweep is queued only
    // once the loop is known to continue.  The iteration count is exact, and
    // at most one velocity sweep per solve exits at once.
    // with min_iter == 0 the reference loop does not start (its error starts at 0)
    const uint32_t first = std::min(min_iter, max_iter);
    if (first == 0) {
        if (iters) *iters = 0;
        if (err) *err = 0.0;
        return DFSPH_OK;
    }
    for (uint32_t k = 0; k < first; ++k) { jacobi(); loop_end(); }
    uint32_t queued = first;
    SolveCtl res{};
    for (;;) {
        const bool spec = queued < max_iter;
        const int pr = spec ? jacobi() : -1;
        HIPCHK(hipGetLastError());
        CHK(wait_record(h, h->seq));
        res = h->last.sc;
        if (res.done) pdrop(h, pr);   // that sweep exited at once
        if (res.done || !spec) break;
        loop_end();
        ++queued;
    }
    if (!res.done) return fail(DFSPH_ERR_STATE, "Jacobi loop ended without a decision");
    if (iters) *iters = res.iters;
    if (err) *err = res.err;
    return DFSPH_OK;
}

// Features of the learned warm start (once per step, after the neighbor
// search) and, in learned mode, the MLP's predictions of both warm starts.
// Features are also computed while a training set is being harvested.
int learned_predict(Handle* h)
{
    const bool learned = step_warm_start(h) == DFSPH_WARMSTART_LEARNED;
    const bool harvest = h->mlp && dfsph::mlp_harvesting(h->mlp);
    if (!learned && !harvest) return DFSPH_OK;
    if (h->pred_valid) return DFSPH_OK;
    if (learned && !(h->mlp && dfsph::mlp_loaded(h->mlp)))
        return fail(DFSPH_ERR_NO_MODEL, "learned warm start selected but no model loaded");
    CHK(dfsph::mlp_ensure_particles(h->mlp, h->cap));
    const int pf = pbeg(h, DFSPH_KERNEL_FEATURES);
    dfsph::launch_features(h->s, h->c, h->tm, h->n, h->pos[h->pc], h->bpos, lists(h), h->vel[h->vc],
                           h->kap[h->pc], h->kapv[h->pc], h->ktot[h->pc], h->kvtot[h->pc],
                           (float)(2.0 * h->p.particle_radius), dfsph::mlp_feat(h->mlp), dfsph::mlp_prev_p(h->mlp),
                           dfsph::mlp_prev_v(h->mlp));
    pend(h, pf);
    HIPCHK(hipGetLastError());
    if (learned) {
        const int pm = pbeg(h, DFSPH_KERNEL_MLP_INFER);
        const int rc = dfsph::mlp_infer_particles(h->mlp, h->s, h->n);
        pend(h, pm);
        CHK(rc);
    }
    h->pred_valid = true;
    return DFSPH_OK;
}

int divergence_solve(Handle* h, uint32_t max_iter, double max_error, uint32_t* iters, double* err)
{
    CHK(flush_vel(h));
    if (!h->lists_valid) return fail(DFSPH_ERR_STATE, "neighbor lists are stale: call dfsph_neighbor_search");
    const int ws = step_warm_start(h);
    CHK(learned_predict(h));
    const float dia = (float)(2.0 * h->p.particle_radius);
    const dfsph::WsLearned lw = ws == DFSPH_WARMSTART_LEARNED ? dfsph::mlp_ws_v(h->mlp, dia) : dfsph::WsLearned{};
    const int vo = h->vc ^ 1;
    int pr = pbeg(h, DFSPH_KERNEL_DIV_WS_INIT);
    dfsph::launch_div_ws_init(h->s, ws, h->c, h->tm, h->n, h->pos[h->pc], h->bpos, lists(h), h->vel[h->vc],
                              h->vel[vo], h->kapv[h->pc], lw, h->rho, h->alpha, h->radv, h->kvtot[h->pc]);
    pend(h, pr);
    h->vc = vo;
    pr = pbeg(h, DFSPH_KERNEL_DIV_INIT);
    dfsph::launch_div_source(h->s, true, h->c, h->n, h->pos[h->pc], h->bpos, lists(h), h->vel[h->vc],
                             h->alpha, h->radv, h->kk, h->kapv[h->pc], nullptr, nullptr,
                             solve_ctl(h, true, max_iter, max_error));
    pend(h, pr);
    HIPCHK(hipGetLastError());
    h->acc_valid = false;
    uint32_t it = 0;
    CHK(jacobi_loop(h, true, max_iter, max_error, &it, err));
    if (iters) *iters = it;
    h->last_iters_v = it;
    // the step's divergence total (applied warm start + increments) in kappaV h
    dfsph::launch_finish_totals(h->s, h->tm, h->n, 0, h->kapv[h->pc], h->kvtot[h->pc]);
    HIPCHK(hipGetLastError());
    return DFSPH_OK;
}

int density_only(Handle* h)
{
    // rho and alpha without the divergence solver (the reference leaves them stale)
    CHK(learned_predict(h));
    const int vo = h->vc ^ 1;
    dfsph::launch_div_ws_init(h->s, DFSPH_WARMSTART_NONE, h->c, h->tm, h->n, h->pos[h->pc], h->bpos, lists(h),
                              h->vel[h->vc], h->vel[vo], nullptr, dfsph::WsLearned{}, h->rho, h->alpha, h->radv,
                              h->kvtot[h->pc]);
    h->vc = vo;
    HIPCHK(hipGetLastError());
    return DFSPH_OK;
}

int external_forces(Handle* h)
{
    CHK(flush_vel(h));
    if (!h->lists_valid) return fail(DFSPH_ERR_STATE, "neighbor lists are stale: call dfsph_neighbor_search");
    const int pr = pbeg(h, DFSPH_KERNEL_XSPH);
    dfsph::launch_xsph(h->s, h->c, h->tm, h->n, h->pos[h->pc], lists(h), h->vel[h->vc], h->rho, h->acc,
                       h->ctl, h->partial_f);
    pend(h, pr);
    HIPCHK(hipGetLastError());
    h->acc_valid = true;
    return DFSPH_OK;
}

int pressure_solve(Handle* h, uint32_t max_iter, double max_error, uint32_t* iters, double* err)
{
    if (!h->lists_valid) return fail(DFSPH_ERR_STATE, "neighbor lists are stale: call dfsph_neighbor_search");
    const int ws = step_warm_start(h);
    dfsph::WsLearned lw{};
    if (ws == DFSPH_WARMSTART_LEARNED) {
        CHK(learned_predict(h));
        lw = dfsph::mlp_ws_p(h->mlp, (float)(2.0 * h->p.particle_radius));
    }
    if (ws == DFSPH_WARMSTART_NONE) {
        CHK(flush_vel(h));
        dfsph::launch_fill_f32(h->s, h->ktot[h->pc], 0.0f, h->n);
    } else {
        const bool fused = h->pend_vel;
        const int pw = pbeg(h, DFSPH_KERNEL_PRESSURE_WS);
        dfsph::launch_pressure_ws(h->s, ws, fused, h->c, h->tm, h->n, h->pos[h->pc], h->bpos, lists(h),
                                  h->acc, h->vel[h->vc], h->kap[h->pc], lw, fused ? 1 : 0, h->ktot[h->pc]);
        pend(h, pw);
        h->pend_vel = false;
    }
    const int pr = pbeg(h, DFSPH_KERNEL_DENS_INIT);
    dfsph::launch_density_source(h->s, true, h->c, h->tm, h->n, h->pos[h->pc], h->bpos, lists(h),
                                 h->vel[h->vc], h->rho, h->alpha, h->radv, h->kk, h->kap[h->pc], nullptr, nullptr,
                                 solve_ctl(h, false, max_iter, max_error));
    pend(h, pr);
    HIPCHK(hipGetLastError());
    h->acc_valid = false;
    uint32_t it = 0;
    CHK(jacobi_loop(h, false, max_iter, max_error, &it, err));
    if (iters) *iters = it;
    h->last_iters = it;
    if (!h->p.enable_divergence) h->last_iters_v = 0;
    // the step's density total (applied warm start + increments) in kappa h_f^2
    dfsph::launch_finish_totals(h->s, h->tm, h->n, 1, h->kap[h->pc], h->ktot[h->pc]);
    if (h->mlp && dfsph::mlp_harvesting(h->mlp)) {
        // targets: what the reference warm start left to the iterations
        if (ws != DFSPH_WARMSTART_REFERENCE)
            return fail(DFSPH_ERR_STATE, "a training set is harvested with the reference warm start");
        CHK(dfsph::mlp_harvest_step(h->mlp, h->s, h->tm, h->n, h->kapv[h->pc], h->kap[h->pc],
                                    (float)(2.0 * h->p.particle_radius)));
    }
    HIPCHK(hipGetLastError());
    return DFSPH_OK;
}

int compact(Handle* h, uint32_t* removed)
{
    const uint32_t n = h->n;
    dfsph::launch_exclusive_scan(h->s, h->keep, h->rank, n + 1, h->scan_tmp);
    uint32_t kept = 0;
    HIPCHK(hipMemcpyAsync(&h->hctl->pad, h->rank + n, sizeof(uint32_t), hipMemcpyDeviceToHost, h->s));
    CHK(sync(h));
    kept = h->hctl->pad;
    if (removed) *removed = n - kept;
    if (kept == n) return DFSPH_OK;
    const int o = h->pc ^ 1, vo = h->vc ^ 1;
    dfsph::launch_compact(h->s, n, h->keep, h->rank, fluid_ptrs(h, h->pc, h->vc), fluid_ptrs(h, o, vo));
    HIPCHK(hipGetLastError());
    h->pred_valid = false;
    h->pc = o;
    h->vc = vo;
    h->n = kept;
    h->lists_valid = false;
    h->acc_valid = false;
    return DFSPH_OK;
}

int check_positions(Handle* h, int mode, uint32_t* count)
{
    CHK(flush_vel(h));
    dfsph::launch_keep_inside(h->s, h->n, h->pos[h->pc], h->bmin, h->bmax, h->keep);
    if (mode == 2) return compact(h, count);
    dfsph::launch_exclusive_scan(h->s, h->keep, h->rank, h->n + 1, h->scan_tmp);
    HIPCHK(hipMemcpyAsync(&h->hctl->pad, h->rank + h->n, sizeof(uint32_t), hipMemcpyDeviceToHost, h->s));
    CHK(sync(h));
    if (count) *count = h->n - h->hctl->pad;
    return DFSPH_OK;
}

int alloc_fluid(Handle* h, uint32_t cap)
{
    for (int b = 0; b < 2; ++b) {
        CHK(dalloc(h->pos[b], cap));
        CHK(dalloc(h->vel[b], cap));
        CHK(dalloc(h->kap[b], cap));
        CHK(dalloc(h->kapv[b], cap));
        CHK(dalloc(h->ktot[b], cap));
        CHK(dalloc(h->kvtot[b], cap));
        CHK(dalloc(h->id[b], cap));
        CHK(dalloc(h->tag[b], cap));
        dfsph::launch_fill_u32(h->s, h->tag[b], dfsph::kTagUntagged, cap);
    }
    CHK(dalloc(h->rho, cap));
    CHK(dalloc(h->alpha, cap));
    CHK(dalloc(h->radv, cap));
    CHK(dalloc(h->kk, cap));
    CHK(dalloc(h->acc, cap));
    CHK(dalloc(h->cnt, cap));
    // blocked-interleaved lists: 16-byte groups, with room for the last batch
    // of 8 a sweep reads; zeroed so unread slots hold valid indices
    const uint32_t mx = h->p.max_neighbors;
    const size_t n16 = (size_t)cap * dfsph::list_groups16(mx) * dfsph::kGroup16;
    const size_t n32 = (size_t)cap * dfsph::list_groups32(mx) * dfsph::kGroup32;
    CHK(dalloc(h->nbr16, n16));
    CHK(dalloc(h->nbrf, n32));
    CHK(dalloc(h->nbrb, n32));
    HIPCHK(hipMemsetAsync(h->nbr16, 0, sizeof(uint16_t) * n16, h->s));
    HIPCHK(hipMemsetAsync(h->nbrf, 0, sizeof(uint32_t) * n32, h->s));
    HIPCHK(hipMemsetAsync(h->nbrb, 0, sizeof(uint32_t) * n32, h->s));
    // neighbor sets: one per block of the XCD-padded grid
    const uint32_t nblocks = dfsph::grid_blocks_xcd(cap);
    CHK(dalloc(h->stage_n, nblocks));
    CHK(dalloc(h->nbr_part, 2 * (size_t)nblocks));
    CHK(dalloc(h->stg, (size_t)nblocks * dfsph::kStageCap));
    CHK(ensure_tmp(h, std::max(cap, h->nb)));
    CHK(ensure_partials(h, cap));
    h->cap = cap;
    h->c.cap = cap;
    // zero-initialise the persistent fields beyond the active count
    for (int b = 0; b < 2; ++b) {
        HIPCHK(hipMemsetAsync(h->kap[b], 0, sizeof(float) * cap, h->s));
        HIPCHK(hipMemsetAsync(h->kapv[b], 0, sizeof(float) * cap, h->s));
        HIPCHK(hipMemsetAsync(h->ktot[b], 0, sizeof(float) * cap, h->s));
        HIPCHK(hipMemsetAsync(h->kvtot[b], 0, sizeof(float) * cap, h->s));
    }
    HIPCHK(hipMemsetAsync(h->cnt, 0, sizeof(uint32_t) * cap, h->s));
    // per-step fields start at zero: a restricted-out particle's stored
    // rho_adv enters the error sums before it was ever computed
    HIPCHK(hipMemsetAsync(h->rho, 0, sizeof(float) * cap, h->s));
    HIPCHK(hipMemsetAsync(h->alpha, 0, sizeof(float) * cap, h->s));
    HIPCHK(hipMemsetAsync(h->radv, 0, sizeof(float) * cap, h->s));
    HIPCHK(hipMemsetAsync(h->kk, 0, sizeof(float) * cap, h->s));
    HIPCHK(hipMemsetAsync(h->acc, 0, sizeof(float4) * cap, h->s));
    return DFSPH_OK;
}

// change_max_particle_number (DFSPH_memory_management_cuda.cu:962-1018): every
// per-particle buffer reallocated for new_cap particles, the active particles'
// persistent fields (position and mass, velocity, kappa, kappaV, the totals,
// id, tag) kept in their order; the neighbor lists are rebuilt by the next
// search.
int grow_capacity(Handle* h, uint32_t new_cap)
{
    if (new_cap <= h->cap) return DFSPH_OK;
    CHK(flush_vel(h));
    const uint32_t n = h->n;
    float4 *tp = nullptr, *tv = nullptr;
    float *tk = nullptr, *tkv = nullptr, *tkt = nullptr, *tkvt = nullptr;
    uint32_t *tid = nullptr, *ttag = nullptr;
    auto release = [&] { dfree(tp); dfree(tv); dfree(tk); dfree(tkv); dfree(tkt); dfree(tkvt); dfree(tid); dfree(ttag); };
    int rc = DFSPH_OK;
    auto keep = [&](auto*& tmp, auto* src) -> int {
        CHK(dalloc(tmp, n));
        if (n) HIPCHK(hipMemcpyAsync(tmp, src, sizeof(*src) * n, hipMemcpyDeviceToDevice, h->s));
        return DFSPH_OK;
    };
    if ((rc = keep(tp, h->pos[h->pc])) || (rc = keep(tv, h->vel[h->vc])) || (rc = keep(tk, h->kap[h->pc])) ||
        (rc = keep(tkv, h->kapv[h->pc])) || (rc = keep(tkt, h->ktot[h->pc])) || (rc = keep(tkvt, h->kvtot[h->pc])) ||
        (rc = keep(tid, h->id[h->pc])) || (rc = keep(ttag, h->tag[h->pc])) || (rc = sync(h)) ||
        (rc = alloc_fluid(h, new_cap))) {
        release();
        return rc;
    }
    h->pc = 0;
    h->vc = 0;
    auto back = [&](auto* dst, const auto* tmp) -> int {
        if (n) HIPCHK(hipMemcpyAsync(dst, tmp, sizeof(*tmp) * n, hipMemcpyDeviceToDevice, h->s));
        return DFSPH_OK;
    };
    if ((rc = back(h->pos[0], tp)) || (rc = back(h->vel[0], tv)) || (rc = back(h->kap[0], tk)) ||
        (rc = back(h->kapv[0], tkv)) || (rc = back(h->ktot[0], tkt)) || (rc = back(h->kvtot[0], tkvt)) ||
        (rc = back(h->id[0], tid)) || (rc = back(h->tag[0], ttag)) || (rc = sync(h))) {
        release();
        return rc;
    }
    release();
    // a capacity fixed at create follows the growth (a later upload keeps it)
    if (h->p.fluid_capacity) h->p.fluid_capacity = new_cap;
    h->lists_valid = false;
    h->acc_valid = false;
    h->pred_valid = false;
    return DFSPH_OK;
}

int write_fluid_range(Handle* h, uint32_t at, const float* xyz, const float* vel, const float* mass,
                      const uint32_t* id, uint32_t m)
{
    const double d = 2.0 * h->p.particle_radius;
    const float m0 = (float)(0.8 * d * d * d * h->p.density0);
    std::vector<float4> P(m), V(m);
    std::vector<uint32_t> I(m);
    std::vector<float> Z(m, 0.0f);
    for (uint32_t i = 0; i < m; ++i) {
        P[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], mass ? mass[i] : m0);
        V[i] = vel ? make_float4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.0f) : make_float4(0, 0, 0, 0);
        I[i] = id ? id[i] : at + i;
        h->next_id = std::max(h->next_id, I[i] + 1u);
    }
    // uniform fluid mass: the source sweeps then stage 24-byte records
    bool same = at == 0 || h->c.m_uniform > 0.0f;
    const float mref = at == 0 ? (m ? P[0].w : 0.0f) : h->c.m_uniform;
    for (uint32_t i = 0; i < m && same; ++i) same = P[i].w == mref;
    h->c.m_uniform = same && mref > 0.0f ? mref : 0.0f;
    HIPCHK(hipMemcpyAsync(h->pos[h->pc] + at, P.data(), sizeof(float4) * m, hipMemcpyHostToDevice, h->s));
    HIPCHK(hipMemcpyAsync(h->vel[h->vc] + at, V.data(), sizeof(float4) * m, hipMemcpyHostToDevice, h->s));
    HIPCHK(hipMemcpyAsync(h->id[h->pc] + at, I.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice, h->s));
    dfsph::launch_fill_u32(h->s, h->tag[h->pc] + at, dfsph::kTagUntagged, m);
    HIPCHK(hipMemcpyAsync(h->kap[h->pc] + at, Z.data(), sizeof(float) * m, hipMemcpyHostToDevice, h->s));
    HIPCHK(hipMemcpyAsync(h->kapv[h->pc] + at, Z.data(), sizeof(float) * m, hipMemcpyHostToDevice, h->s));
    HIPCHK(hipMemcpyAsync(h->ktot[h->pc] + at, Z.data(), sizeof(float) * m, hipMemcpyHostToDevice, h->s));
    HIPCHK(hipMemcpyAsync(h->kvtot[h->pc] + at, Z.data(), sizeof(float) * m, hipMemcpyHostToDevice, h->s));
    h->pred_valid = false;
    return sync(h);
}

}  // namespace

namespace dfsph {
// accessors used by the learned warm start (dfsph_mlp.hip)
MlpView handle_view(void* hv)
{
    Handle* h = H(hv);
    MlpView v;
    v.s = h->s;
    v.mlp = &h->mlp;
    return v;
}
void handle_set_error(const char* msg) { g_err = msg; }
}  // namespace dfsph

namespace {

// grow a device array to count elements, keeping its first `keep`
template <class T>
int grow(Handle* h, T*& p, size_t keep, size_t count)
{
    T* q = nullptr;
    CHK(dalloc(q, count));
    if (keep && p) HIPCHK(hipMemcpyAsync(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice, h->s));
    CHK(sync(h));
    dfree(p);
    p = q;
    return DFSPH_OK;
}

// dfsph_rigid_body -> device body (checked; quaternion and axis normalised);
// a sine body is placed at time t_now
int body_from_abi(const dfsph_rigid_body* b, double t_now, dfsph::DevBody& d)
{
    if (b->type < DFSPH_BODY_DYNAMIC || b->type > DFSPH_BODY_SINE) return fail(DFSPH_ERR_ARG, "unknown body type");
    const double qn = std::sqrt(b->rotation[0] * b->rotation[0] + b->rotation[1] * b->rotation[1] +
                                b->rotation[2] * b->rotation[2] + b->rotation[3] * b->rotation[3]);
    if (!(qn > 0.0)) return fail(DFSPH_ERR_ARG, "the body's rotation quaternion is zero");
    if (b->type == DFSPH_BODY_DYNAMIC) {
        if (!(b->mass > 0.0)) return fail(DFSPH_ERR_ARG, "a dynamic body needs a positive mass");
        for (int a = 0; a < 3; ++a)
            if (!(b->inertia[a] > 0.0)) return fail(DFSPH_ERR_ARG, "a dynamic body needs positive moments of inertia");
    }
    double an = 0.0;
    if (b->type == DFSPH_BODY_SINE) {
        an = std::sqrt(b->axis[0] * b->axis[0] + b->axis[1] * b->axis[1] + b->axis[2] * b->axis[2]);
        if (!(b->period > 0.0) || !(an > 0.0)) return fail(DFSPH_ERR_ARG, "a sine body needs a period and an axis");
    }
    d = dfsph::DevBody{};
    d.type = b->type;
    d.mass = b->mass;
    for (int a = 0; a < 3; ++a) {
        d.c[a] = b->position[a];
        d.v[a] = b->velocity[a];
        d.w[a] = b->angular_velocity[a];
        d.inertia[a] = b->inertia[a];
        d.c0[a] = b->position[a];
        d.axis[a] = an > 0.0 ? b->axis[a] / an : 0.0;
    }
    // a quaternion already of unit length is kept as it is, so that a state read
    // back by dfsph_get_rigid_body restores bit for bit
    const double qs = std::fabs(qn - 1.0) > 1e-12 ? 1.0 / qn : 1.0;
    for (int a = 0; a < 4; ++a) d.q[a] = b->rotation[a] * qs;
    d.clock = t_now;
    if (b->type == DFSPH_BODY_SINE) {
        d.amp = b->amplitude;
        d.omega = 2.0 * M_PI / b->period;
        const double sn = std::sin(d.omega * t_now), cs = std::cos(d.omega * t_now);
        for (int a = 0; a < 3; ++a) {
            d.c[a] = d.c0[a] + d.amp * sn * d.axis[a];
            d.v[a] = d.amp * d.omega * cs * d.axis[a];
            d.w[a] = 0.0;
        }
    }
    return DFSPH_OK;
}

}  // namespace

extern "C" {

void dfsph_default_params(dfsph_params* p)
{
    std::memset(p, 0, sizeof(*p));
    p->particle_radius = 0.025;
    p->density0 = 1000.0;
    p->viscosity = 0.01;
    p->gravity[1] = -9.81;
    p->dt = 0.001;
    p->max_iter_v = 100;
    p->max_iter = 100;
    p->min_iter_v = 3;
    p->min_iter = 2;
    p->max_error_v = 0.1;
    p->max_error = 0.01;
    p->use_cfl = 0;
    p->cfl_factor = 1.0;
    p->dt_min = 1.0e-4;
    p->dt_max = 5.0e-3;
    p->enable_divergence = 1;
    p->check_positions_mode = 2;
    p->max_neighbors = 90;
    p->fluid_capacity = 0;
    p->warm_start_mode = DFSPH_WARMSTART_REFERENCE;
    p->check_interval = 4;
    p->list_mode = DFSPH_LISTS_AUTO;
}

const char* dfsph_last_error(void) { return g_err.c_str(); }

int dfsph_create(const dfsph_params* p, void** out)
{
    if (!p || !out) return fail(DFSPH_ERR_ARG, "null argument");
    if (!(p->particle_radius > 0.0) || !(p->density0 > 0.0) || !(p->dt > 0.0))
        return fail(DFSPH_ERR_ARG, "particle_radius, density0 and dt must be positive");
    if (p->max_neighbors == 0 || p->max_neighbors > 255) return fail(DFSPH_ERR_ARG, "max_neighbors must be in [1, 255]");
    if (p->warm_start_mode < 0 || p->warm_start_mode > 2) return fail(DFSPH_ERR_ARG, "bad warm_start_mode");
    if (p->list_mode < 0 || p->list_mode > 2) return fail(DFSPH_ERR_ARG, "bad list_mode");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(DFSPH_ERR_HIP, "no HIP device");
    Handle* h = new Handle();
    h->p = *p;
    if (h->p.check_interval == 0) h->p.check_interval = 4;
    int rc = DFSPH_OK;
    do {
        if (hipStreamCreateWithFlags(&h->own_s, hipStreamNonBlocking) != hipSuccess) { rc = fail(DFSPH_ERR_HIP, "stream"); break; }
        h->s = h->own_s;
        if ((rc = dalloc(h->ctl, 1)) != DFSPH_OK) break;
        if ((rc = dalloc(h->tm, 1)) != DFSPH_OK) break;
        if (hipHostMalloc((void**)&h->hctl, sizeof(DevCtl)) != hipSuccess ||
            hipHostMalloc((void**)&h->htm, sizeof(DevTime)) != hipSuccess ||
            hipHostMalloc((void**)&h->hrec, sizeof(dfsph::HostRecord), hipHostMallocCoherent | hipHostMallocMapped) !=
                hipSuccess ||
            hipHostGetDevicePointer((void**)&h->drec, h->hrec, 0) != hipSuccess) {
            rc = fail(DFSPH_ERR_HIP, "pinned allocation");
            break;
        }
        std::memset(h->hrec, 0, sizeof(dfsph::HostRecord));
        for (int b = 0; b < 2; ++b)
            if (hipEventCreateWithFlags(&h->ev[b], hipEventDisableTiming) != hipSuccess) rc = fail(DFSPH_ERR_HIP, "event");
        if (rc) break;
        if (hipMemsetAsync(h->ctl, 0, sizeof(DevCtl), h->s) != hipSuccess) { rc = fail(DFSPH_ERR_HIP, "memset"); break; }
        dfsph::launch_time_reset(h->s, h->tm, h->p.dt);
        const double hk = 4.0 * h->p.particle_radius;
        const double kc = 8.0 / (M_PI * hk * hk * hk), lc = 48.0 / (M_PI * hk * hk * hk);
        SimConst& c = h->c;
        c.hk = (float)hk;
        c.hk2 = c.hk * c.hk;
        c.inv_hk = (float)(1.0 / hk);
        c.kcoef = (float)kc;
        c.lh2 = (float)(lc / (hk * hk));
        c.lh = (float)(lc / hk);
        c.W0 = (float)kc;
        c.hk2_4 = (float)(0.25 * hk * hk);
        c.gin_a = (float)(3.0 * lc / (hk * hk * hk));
        c.gin_b = (float)(-2.0 * lc / (hk * hk));
        c.gout = (float)(-lc / (hk * hk * hk));
        c.density0 = (float)h->p.density0;
        c.gx = (float)h->p.gravity[0];
        c.gy = (float)h->p.gravity[1];
        c.gz = (float)h->p.gravity[2];
        c.viscosity = (float)h->p.viscosity;
        c.max_nbr = h->p.max_neighbors;
        c.stage_lds = h->p.list_mode == DFSPH_LISTS_AUTO ? dfsph::kStageMax : 0u;
        c.stage_cap = h->p.list_mode == DFSPH_LISTS_GLOBAL ? 0u : dfsph::kStageCap;
        if (hipStreamSynchronize(h->s) != hipSuccess) { rc = fail(DFSPH_ERR_HIP, "sync"); break; }
    } while (0);
    if (rc != DFSPH_OK) { delete h; return rc; }
    *out = h;
    return DFSPH_OK;
}

int dfsph_destroy(void* hv)
{
    delete H(hv);
    return DFSPH_OK;
}

int dfsph_get_params(void* hv, dfsph_params* out)
{
    if (!out) return fail(DFSPH_ERR_ARG, "null params");
    *out = H(hv)->p;
    return DFSPH_OK;
}

int dfsph_set_stream(void* hv, void* stream)
{
    Handle* h = H(hv);
    CHK(flush_vel(h));
    // the work queued so far completes before work on the new stream starts
    CHK(sync(h));
    h->s = stream ? static_cast<hipStream_t>(stream) : h->own_s;
    return DFSPH_OK;
}

int dfsph_get_stream(void* hv, void** stream)
{
    if (!stream) return fail(DFSPH_ERR_ARG, "null stream");
    *stream = H(hv)->s;
    return DFSPH_OK;
}

int dfsph_upload_boundary(void* hv, const float* xyz, const float* psi, uint32_t n)
{
    Handle* h = H(hv);
    if (!h || !xyz || n == 0) return fail(DFSPH_ERR_ARG, "boundary needs at least one particle");
    if (h->grid_ready) return fail(DFSPH_ERR_STATE, "the boundary set is static and was already uploaded");
    // initGridOffset (DFSPH_c_arrays_structure.cpp:1050-1098) with a grid sized
    // to the boundary bounding box (no 256^3 cap)
    for (int a = 0; a < 3; ++a) { h->bmin[a] = xyz[a]; h->bmax[a] = xyz[a]; }
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            h->bmin[a] = std::min(h->bmin[a], xyz[3 * i + a]);
            h->bmax[a] = std::max(h->bmax[a], xyz[3 * i + a]);
        }
    const double cell = (double)h->c.hk * kCellMargin;
    // cell keys: 8x8x8-cell tiles, 512 keys per tile (the key table covers
    // the grid rounded up to whole tiles)
    uint64_t ncell = 512;
    for (int a = 0; a < 3; ++a) {
        const double org = (double)h->bmin[a] - 2.0 * cell;
        h->c.org[a] = (float)org;
        h->c.dim[a] = (int)std::floor(((double)h->bmax[a] - org) / cell) + 3;
        h->c.tiles[a] = (h->c.dim[a] + 7) / 8;
        ncell *= (uint64_t)h->c.tiles[a];
    }
    if (ncell >= (1ull << 31)) return fail(DFSPH_ERR_ARG, "domain too large for the cell grid");
    h->c.inv_cell = (float)(1.0 / cell);
    h->c.ncell = (uint32_t)ncell;
    {
        // tiles in Morton order of their coordinates (bits of x, y, z interleaved)
        const uint32_t tx = (uint32_t)h->c.tiles[0], ty = (uint32_t)h->c.tiles[1], tz = (uint32_t)h->c.tiles[2];
        const uint32_t nt = tx * ty * tz;
        auto spread = [](uint64_t v) {
            uint64_t r = 0;
            for (int bit = 0; bit < 21; ++bit) r |= ((v >> bit) & 1ull) << (3 * bit);
            return r;
        };
        std::vector<std::pair<uint64_t, uint32_t>> order(nt);
        for (uint32_t y = 0; y < ty; ++y)
            for (uint32_t z = 0; z < tz; ++z)
                for (uint32_t x = 0; x < tx; ++x) {
                    const uint32_t t = x + tx * (z + tz * y);
                    order[t] = {spread(x) | (spread(y) << 1) | (spread(z) << 2), t};
                }
        std::sort(order.begin(), order.end());
        std::vector<uint32_t> rank(nt);
        for (uint32_t k = 0; k < nt; ++k) rank[order[k].second] = k;
        CHK(dalloc(h->tile_rank, nt));
        HIPCHK(hipMemcpy(h->tile_rank, rank.data(), sizeof(uint32_t) * nt, hipMemcpyHostToDevice));
        h->c.tile_rank = h->tile_rank;
    }
    h->nb = n;
    CHK(dalloc(h->hist, ncell + 1));
    dfsph::launch_fill_u32(h->s, h->hist, 0u, (uint32_t)ncell + 1);   // each counting sort leaves it zeroed
    CHK(dalloc(h->cstart, ncell + 1));
    CHK(dalloc(h->bcstart, ncell + 1));
    CHK(ensure_scan(h, (uint32_t)ncell + 1));
    CHK(ensure_tmp(h, std::max(n, h->cap)));
    std::vector<float4> P(n);
    for (uint32_t i = 0; i < n; ++i) P[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], psi ? psi[i] : 0.0f);
    float4* raw = nullptr;
    CHK(dalloc(raw, n));
    CHK(dalloc(h->bpos, n));
    HIPCHK(hipMemcpyAsync(raw, P.data(), sizeof(float4) * n, hipMemcpyHostToDevice, h->s));
    int rc = counting_sort(h, raw, n, h->bcstart);
    if (rc == DFSPH_OK) {
        dfsph::launch_gather_f4(h->s, n, h->perm, raw, h->bpos);
        if (!psi) {
            float* tmp = nullptr;
            rc = dalloc(tmp, n);
            if (rc == DFSPH_OK) {
                dfsph::launch_boundary_psi(h->s, h->c, h->bpos, h->bcstart, n, tmp);
                dfsph::launch_set_w(h->s, h->bpos, tmp, n);
                rc = sync(h);
                dfree(tmp);
            }
        }
    }
    if (rc == DFSPH_OK) rc = sync(h);
    dfree(raw);
    if (rc != DFSPH_OK) return rc;
    h->nbs = n;
    h->grid_ready = true;
    h->ncell_alloc = (uint32_t)ncell;
    return DFSPH_OK;
}

int dfsph_upload_fluid(void* hv, const float* xyz, const float* vel, const float* mass, const uint32_t* id,
                       uint32_t n)
{
    Handle* h = H(hv);
    if (!h || (!xyz && n)) return fail(DFSPH_ERR_ARG, "null argument");
    if (!h->grid_ready) return fail(DFSPH_ERR_STATE, "upload the boundary set first");
    uint32_t cap = h->p.fluid_capacity ? h->p.fluid_capacity : (uint32_t)std::ceil(1.3 * (double)n);
    if (cap < n) return fail(DFSPH_ERR_CAPACITY, "fluid_capacity is smaller than the particle count");
    if (cap < 64) cap = 64;
    if (cap != h->cap) CHK(alloc_fluid(h, cap));
    h->pc = 0;
    h->vc = 0;
    h->pend_vel = false;
    h->n = n;
    if (n) CHK(write_fluid_range(h, 0, xyz, vel, mass, id, n));
    dfsph::launch_time_reset(h->s, h->tm, h->p.dt);
    h->lists_valid = false;
    h->acc_valid = false;
    return sync(h);
}

int dfsph_neighbor_search(void* hv, int sort) { return neighbor_search(H(hv), sort != 0); }

int dfsph_compute_density_factor(void* hv)
{
    Handle* h = H(hv);
    CHK(flush_vel(h));
    if (!h->lists_valid) return fail(DFSPH_ERR_STATE, "neighbor lists are stale: call dfsph_neighbor_search");
    return density_only(h);
}

int dfsph_divergence_solve(void* hv, uint32_t max_iter, double max_error, uint32_t* iters, double* err)
{
    return divergence_solve(H(hv), max_iter, max_error, iters, err);
}

int dfsph_external_forces(void* hv) { return external_forces(H(hv)); }

int dfsph_cfl(void* hv, double min_dt, double cfl_factor, double max_dt, double* dt_out)
{
    Handle* h = H(hv);
    if (!h->acc_valid) CHK(external_forces(h));
    dfsph::launch_time_update(h->s, h->tm, 0.0, h->ctl, 1, cfl_factor, 2.0 * h->p.particle_radius, min_dt, max_dt, 0.0);
    HIPCHK(hipGetLastError());
    if (!dt_out) return DFSPH_OK;       // stays on the device (dfsph_get_time reads it later)
    CHK(read_ctl(h));
    *dt_out = h->htm->hfd;
    return DFSPH_OK;
}

int dfsph_get_time(void* hv, double* h_now, double* h_future, double* time)
{
    Handle* h = H(hv);
    const dfsph::DevTime* t = nullptr;
    if (h->rec_op + 1 == h->ops && h->seq == h->last.seq) {
        t = &h->last.tm;                // the record the previous call waited for is the latest state
    } else {
        CHK(flush_vel(h));
        CHK(read_ctl(h));
        t = h->htm;
    }
    if (h_now) *h_now = t->hd;
    if (h_future) *h_future = t->hfd;
    if (time) *time = t->time;
    return DFSPH_OK;
}

int dfsph_fluid_count(void* hv, uint32_t* n)
{
    Handle* h = H(hv);
    if (!n) return fail(DFSPH_ERR_ARG, "null count");
    *n = h->n;
    return DFSPH_OK;
}

int dfsph_update_time_step(void* hv, double h_future)
{
    Handle* h = H(hv);
    if (!(h_future > 0.0)) return fail(DFSPH_ERR_ARG, "time step must be positive");
    CHK(flush_vel(h));
    dfsph::launch_time_update(h->s, h->tm, h_future, h->ctl, 0, 0, 0, 0, 0, 0);
    HIPCHK(hipGetLastError());
    return DFSPH_OK;
}

int dfsph_update_vel(void* hv)
{
    Handle* h = H(hv);
    CHK(flush_vel(h));
    h->pend_vel = true;    // applied by the next pressure warm start, or flushed
    h->acc_valid = false;
    return DFSPH_OK;
}

int dfsph_pressure_solve(void* hv, uint32_t max_iter, double max_error, uint32_t* iters, double* err)
{
    return pressure_solve(H(hv), max_iter, max_error, iters, err);
}

int dfsph_update_pos(void* hv)
{
    Handle* h = H(hv);
    CHK(flush_vel(h));
    const int pr = pbeg(h, DFSPH_KERNEL_UPDATE_POS);
    dfsph::launch_update_pos(h->s, h->tm, h->n, h->vel[h->vc], h->pos[h->pc], h->bmin, h->bmax, h->ctl,
                             restriction(h));
    pend(h, pr);
    // the bodies move with the fluid: forces of this step, then their step over h
    if (!h->bodies.empty())
        dfsph::launch_bodies_integrate(h->s, h->dbody, (uint32_t)h->bodies.size(), h->braw + h->nbs, h->sforce,
                                       1.0 / (double)h->fscale, h->tm, h->p.gravity);
    HIPCHK(hipGetLastError());
    h->lists_valid = false;
    h->acc_valid = false;
    h->pos_op = h->ops;
    return DFSPH_OK;
}

int dfsph_step_end(void* hv)
{
    Handle* h = H(hv);
    dfsph::launch_time_step_end(h->s, h->tm, dfsph::Publish{h->ctl, h->tm, h->drec, ++h->seq});
    HIPCHK(hipGetLastError());
    h->end_op = h->ops;
    h->end_seq = h->seq;
    return DFSPH_OK;
}

int dfsph_set_time(void* hv, double hh, double h_past, double time)
{
    Handle* h = H(hv);
    if (!(hh > 0.0) || !(h_past > 0.0)) return fail(DFSPH_ERR_ARG, "time steps must be positive");
    CHK(flush_vel(h));
    dfsph::launch_time_set(h->s, h->tm, hh, h_past, time);
    HIPCHK(hipGetLastError());
    return sync(h);
}

int dfsph_check_particles_positions(void* hv, int mode, uint32_t* count)
{
    Handle* h = H(hv);
    if (mode != 0 && mode != 2) return fail(DFSPH_ERR_ARG, "mode must be 0 (count) or 2 (remove)");
    if (h->pos_op + 1 == h->end_op && h->end_op + 1 == h->ops && h->end_seq == h->seq) {
        // right after dfsph_update_pos, dfsph_step_end: the step end's record
        // holds the update's count of particles outside the bounding box (the
        // same test on the same positions; an upper bound if the counter was
        // not cleared since an earlier update)
        CHK(wait_record(h, h->end_seq));
        if (h->last.escaped == 0) {
            if (count) *count = 0;
            return DFSPH_OK;
        }
    }
    return check_positions(h, mode, count);
}

int dfsph_remove_particles(void* hv, const uint8_t* keep, uint32_t* n_out)
{
    Handle* h = H(hv);
    if (!keep && h->n) return fail(DFSPH_ERR_ARG, "null keep array");
    CHK(flush_vel(h));
    if (h->n) {
        HIPCHK(hipMemcpyAsync(h->keep8, keep, h->n, hipMemcpyHostToDevice, h->s));
        dfsph::launch_keep_from_u8(h->s, h->n, h->keep8, h->keep);
        CHK(compact(h, nullptr));
    }
    if (n_out) *n_out = h->n;
    return DFSPH_OK;
}

int dfsph_add_particles(void* hv, const float* xyz, const float* vel, const float* mass, const uint32_t* id,
                        uint32_t m)
{
    Handle* h = H(hv);
    if (!xyz && m) return fail(DFSPH_ERR_ARG, "null positions");
    if (h->n + m > h->cap) return fail(DFSPH_ERR_CAPACITY, "adding particles beyond the fluid capacity");
    CHK(flush_vel(h));
    if (m) CHK(write_fluid_range(h, h->n, xyz, vel, mass, id, m));
    h->n += m;
    h->lists_valid = false;
    h->acc_valid = false;
    return DFSPH_OK;
}

int dfsph_step(void* hv, dfsph_step_stats* st)
{
    Handle* h = H(hv);
    if (!h->grid_ready) return fail(DFSPH_ERR_STATE, "no boundary set uploaded");
    dfsph_step_stats loc{};
    CHK(neighbor_search(h, h->rmode != 2));    // mode 2 needs the active prefix in place
    if (h->p.enable_divergence) CHK(divergence_solve(h, h->p.max_iter_v, h->p.max_error_v, &loc.iters_v, &loc.err_v));
    else CHK(density_only(h));
    CHK(external_forces(h));
    // h_future: CFL estimate, or the current step kept constant (the reference
    // takes TimeManager's step size, DFSPH_CUDA.cpp:299-306)
    dfsph::launch_time_update(h->s, h->tm, 0.0, h->ctl, h->p.use_cfl ? 1 : 2, h->p.cfl_factor,
                              2.0 * h->p.particle_radius, h->p.dt_min, h->p.dt_max, 0.0);
    h->pend_vel = true;
    const uint32_t seq0 = h->seq;
    CHK(pressure_solve(h, h->p.max_iter, h->p.max_error, &loc.iters, &loc.err));
    const bool published = h->seq != seq0;     // h->last is this pressure solve's record
    CHK(dfsph_update_pos(h));
    CHK(dfsph_step_end(h));
    uint32_t escaped = 0;
    if (h->p.check_positions_mode == 2) {
        // the step end's record: time after the step and the escaped count
        CHK(wait_record(h, h->end_seq));
        loc.dt = h->last.tm.hpd;
        loc.time = h->last.tm.time;
        escaped = h->last.escaped;
    } else if (published) {
        // the pressure solve's last record holds this step's statistics: no
        // wait, the next step queues behind this one (k_time_step_end:
        // h_past = h, time += h_future)
        loc.dt = h->last.tm.hd;
        loc.time = h->last.tm.time + h->last.tm.hfd;
    }
    if (h->p.check_positions_mode == 2 || published) {
        loc.max_neighbors_seen = h->last.max_seen;
        loc.overflow = h->last.overflow;
        loc.neighbors_fluid_total = h->last.nbr_fluid;
        loc.neighbors_boundary_total = h->last.nbr_boundary;
        prof_collect(h, false);
    } else {
        CHK(read_ctl(h));
        loc.dt = h->htm->hpd;
        loc.time = h->htm->time;
        loc.max_neighbors_seen = h->hctl->max_seen;
        loc.overflow = h->hctl->overflow;
        loc.neighbors_fluid_total = h->hctl->nbr_fluid;
        loc.neighbors_boundary_total = h->hctl->nbr_boundary;
        prof_collect(h);
    }
    if (h->p.check_positions_mode == 2 && escaped) {
        CHK(check_positions(h, 2, &loc.removed));
    } else if (h->p.check_positions_mode == 0) {
        loc.removed = 0;
    }
    loc.n_fluid = h->n;
    if (st) *st = loc;
    if (loc.overflow)
        return fail(DFSPH_ERR_NEIGHBOR_OVERFLOW, std::to_string(loc.overflow) + " particles exceed max_neighbors");
    return DFSPH_OK;
}

int dfsph_download(void* hv, int field, void* dst, uint32_t count)
{
    Handle* h = H(hv);
    if (!dst && count) return fail(DFSPH_ERR_ARG, "null destination");
    CHK(flush_vel(h));
    const uint32_t n = (field == DFSPH_FIELD_BOUNDARY_POS || field == DFSPH_FIELD_BOUNDARY_PSI) ? h->nb : h->n;
    const bool vec = field == DFSPH_FIELD_POS || field == DFSPH_FIELD_VEL || field == DFSPH_FIELD_ACC ||
                     field == DFSPH_FIELD_BOUNDARY_POS;
    if (count != (vec ? 3 * n : n)) return fail(DFSPH_ERR_ARG, "count does not match the field size");
    if (n == 0) return DFSPH_OK;
    CHK(sync(h));
    auto get_f4 = [&](const float4* src, int comp) -> int {
        std::vector<float4> tmp(n);
        HIPCHK(hipMemcpy(tmp.data(), src, sizeof(float4) * n, hipMemcpyDeviceToHost));
        float* o = static_cast<float*>(dst);
        for (uint32_t i = 0; i < n; ++i) {
            if (comp < 0) { o[3 * i] = tmp[i].x; o[3 * i + 1] = tmp[i].y; o[3 * i + 2] = tmp[i].z; }
            else o[i] = tmp[i].w;
        }
        return DFSPH_OK;
    };
    auto get_raw = [&](const void* src, size_t bytes) -> int {
        HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return DFSPH_OK;
    };
    switch (field) {
    case DFSPH_FIELD_POS: return get_f4(h->pos[h->pc], -1);
    case DFSPH_FIELD_VEL: return get_f4(h->vel[h->vc], -1);
    case DFSPH_FIELD_ACC: return get_f4(h->acc, -1);
    case DFSPH_FIELD_MASS: return get_f4(h->pos[h->pc], 3);
    case DFSPH_FIELD_DENSITY: return get_raw(h->rho, 4ull * n);
    case DFSPH_FIELD_FACTOR: return get_raw(h->alpha, 4ull * n);
    case DFSPH_FIELD_DENSITY_ADV: return get_raw(h->radv, 4ull * n);
    case DFSPH_FIELD_KAPPA: return get_raw(h->kap[h->pc], 4ull * n);
    case DFSPH_FIELD_KAPPAV: return get_raw(h->kapv[h->pc], 4ull * n);
    case DFSPH_FIELD_KAPPA_TOTAL: return get_raw(h->ktot[h->pc], 4ull * n);
    case DFSPH_FIELD_KAPPAV_TOTAL: return get_raw(h->kvtot[h->pc], 4ull * n);
    case DFSPH_FIELD_ID: return get_raw(h->id[h->pc], 4ull * n);
    case DFSPH_FIELD_TAG: return get_raw(h->tag[h->pc], 4ull * n);
    case DFSPH_FIELD_NBR_FLUID:
    case DFSPH_FIELD_NBR_BOUNDARY: {
        std::vector<uint32_t> c(n);
        HIPCHK(hipMemcpy(c.data(), h->cnt, 4ull * n, hipMemcpyDeviceToHost));
        uint32_t* o = static_cast<uint32_t*>(dst);
        for (uint32_t i = 0; i < n; ++i) o[i] = field == DFSPH_FIELD_NBR_FLUID ? (c[i] & 0xffffu) : (c[i] >> 16);
        return DFSPH_OK;
    }
    case DFSPH_FIELD_BOUNDARY_POS: return get_f4(h->bpos, -1);
    case DFSPH_FIELD_BOUNDARY_PSI: return get_f4(h->bpos, 3);
    default: return fail(DFSPH_ERR_ARG, "unknown field");
    }
}

int dfsph_upload_field(void* hv, int field, const void* src, uint32_t count)
{
    Handle* h = H(hv);
    CHK(flush_vel(h));
    const uint32_t n = h->n;
    const bool vec = field == DFSPH_FIELD_POS || field == DFSPH_FIELD_VEL;
    if (count != (vec ? 3 * n : n)) return fail(DFSPH_ERR_ARG, "count does not match the field size");
    if (n == 0) return DFSPH_OK;
    CHK(sync(h));
    const float* f = static_cast<const float*>(src);
    if (field == DFSPH_FIELD_POS || field == DFSPH_FIELD_VEL || field == DFSPH_FIELD_MASS) {
        float4* dptr = field == DFSPH_FIELD_VEL ? h->vel[h->vc] : h->pos[h->pc];
        std::vector<float4> tmp(n);
        HIPCHK(hipMemcpy(tmp.data(), dptr, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i) {
            if (field == DFSPH_FIELD_MASS) tmp[i].w = f[i];
            else { tmp[i].x = f[3 * i]; tmp[i].y = f[3 * i + 1]; tmp[i].z = f[3 * i + 2]; }
        }
        if (field == DFSPH_FIELD_MASS) {
            bool same = f[0] > 0.0f;
            for (uint32_t i = 1; i < n && same; ++i) same = f[i] == f[0];
            h->c.m_uniform = same ? f[0] : 0.0f;
        }
        HIPCHK(hipMemcpy(dptr, tmp.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
        if (field != DFSPH_FIELD_VEL) h->lists_valid = false;
        h->acc_valid = false;
        return DFSPH_OK;
    }
    if (field == DFSPH_FIELD_KAPPA) { HIPCHK(hipMemcpy(h->kap[h->pc], src, 4ull * n, hipMemcpyHostToDevice)); return DFSPH_OK; }
    if (field == DFSPH_FIELD_KAPPAV) { HIPCHK(hipMemcpy(h->kapv[h->pc], src, 4ull * n, hipMemcpyHostToDevice)); return DFSPH_OK; }
    if (field == DFSPH_FIELD_TAG) { HIPCHK(hipMemcpy(h->tag[h->pc], src, 4ull * n, hipMemcpyHostToDevice)); return DFSPH_OK; }
    // derived fields, read by restricted solves for the particles they leave out
    if (field == DFSPH_FIELD_DENSITY) { HIPCHK(hipMemcpy(h->rho, src, 4ull * n, hipMemcpyHostToDevice)); return DFSPH_OK; }
    if (field == DFSPH_FIELD_DENSITY_ADV) { HIPCHK(hipMemcpy(h->radv, src, 4ull * n, hipMemcpyHostToDevice)); return DFSPH_OK; }
    if (field == DFSPH_FIELD_KAPPA_TOTAL || field == DFSPH_FIELD_KAPPAV_TOTAL) {
        float* d = field == DFSPH_FIELD_KAPPA_TOTAL ? h->ktot[h->pc] : h->kvtot[h->pc];
        HIPCHK(hipMemcpy(d, src, 4ull * n, hipMemcpyHostToDevice));
        h->pred_valid = false;
        return DFSPH_OK;
    }
    return fail(DFSPH_ERR_ARG, "field cannot be uploaded");
}

int dfsph_get_neighbors(void* hv, uint32_t* n_fluid, uint32_t* n_boundary, uint32_t* list, uint32_t max)
{
    Handle* h = H(hv);
    if (!h->lists_valid) return fail(DFSPH_ERR_STATE, "neighbor lists are stale");
    CHK(sync(h));
    const uint32_t n = h->n, mx = h->p.max_neighbors, cap = h->cap;
    if (n == 0) return DFSPH_OK;
    std::vector<uint32_t> c(n), L((size_t)mx * cap);
    HIPCHK(hipMemcpy(c.data(), h->cnt, 4ull * n, hipMemcpyDeviceToHost));
    if (list) {
        // staged (16-bit) and unstaged lists as 32-bit global ids
        uint32_t* g = nullptr;
        CHK(dalloc(g, (size_t)mx * cap));
        dfsph::launch_expand_lists(h->s, n, cap, lists(h), g);
        int rc = sync(h);
        if (rc == DFSPH_OK && hipMemcpy(L.data(), g, 4ull * mx * cap, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(DFSPH_ERR_HIP, "hipMemcpy of the neighbor lists failed");
        dfree(g);
        if (rc != DFSPH_OK) return rc;
    }
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t f = c[i] & 0xffffu, b = c[i] >> 16;
        if (n_fluid) n_fluid[i] = f;
        if (n_boundary) n_boundary[i] = b;
        if (list)
            for (uint32_t k = 0; k < f + b && k < max; ++k) list[(size_t)i * max + k] = L[(size_t)k * cap + i];
    }
    return DFSPH_OK;
}

int dfsph_get_info(void* hv, dfsph_info* info)
{
    Handle* h = H(hv);
    if (!info) return fail(DFSPH_ERR_ARG, "null info");
    CHK(flush_vel(h));
    CHK(read_ctl(h));
    std::memset(info, 0, sizeof(*info));
    info->n_fluid = h->n;
    info->n_boundary = h->nb;
    info->fluid_capacity = h->cap;
    info->max_neighbors = h->p.max_neighbors;
    for (int a = 0; a < 3; ++a) {
        info->grid_dim[a] = (uint32_t)h->c.dim[a];
        info->grid_origin[a] = h->c.org[a];
        info->bbox_min[a] = h->bmin[a];
        info->bbox_max[a] = h->bmax[a];
    }
    info->cell_size = 1.0 / (double)h->c.inv_cell;
    info->h = h->htm->hd;
    info->h_future = h->htm->hfd;
    info->h_past = h->htm->hpd;
    info->time = h->htm->time;
    info->max_neighbors_seen = h->hctl->max_seen;
    info->overflow = h->hctl->overflow;
    info->escaped = h->hctl->escaped;
    info->neighbors_fluid_total = h->hctl->nbr_fluid;
    info->neighbors_boundary_total = h->hctl->nbr_boundary;
    info->unstaged_blocks = h->hctl->unstaged_blocks;
    info->neighbors_fluid_unstaged = h->hctl->nbr_fluid_unstaged;
    info->sweep_block = (uint32_t)dfsph::kBlock;
    return DFSPH_OK;
}

int dfsph_synchronize(void* hv) { return sync(H(hv)); }

int dfsph_neighbor_stats(void* hv, uint32_t* max_seen, uint32_t* overflow, uint64_t* fluid_total,
                         uint64_t* boundary_total)
{
    Handle* h = H(hv);
    uint32_t ms, of;
    uint64_t nf, nb;
    if (h->rec_op > h->nbr_op && h->last.seq == h->seq) {
        // the last record the host has seen was queued after the last search
        // and copied its statistics
        ms = h->last.max_seen;
        of = h->last.overflow;
        nf = h->last.nbr_fluid;
        nb = h->last.nbr_boundary;
    } else {
        CHK(read_ctl(h));
        ms = h->hctl->max_seen;
        of = h->hctl->overflow;
        nf = h->hctl->nbr_fluid;
        nb = h->hctl->nbr_boundary;
    }
    if (max_seen) *max_seen = ms;
    if (overflow) *overflow = of;
    if (fluid_total) *fluid_total = nf;
    if (boundary_total) *boundary_total = nb;
    return DFSPH_OK;
}

// ---- dynamic window: restriction and tags ----
int dfsph_set_restriction(void* hv, int mode, uint32_t count_active, uint32_t count_active_neighbors,
                          int64_t true_particle_count)
{
    Handle* h = H(hv);
    if (mode < 0 || mode > 2) return fail(DFSPH_ERR_ARG, "restriction mode must be 0, 1 or 2");
    if (mode == 2 && (uint64_t)count_active + count_active_neighbors > h->n)
        return fail(DFSPH_ERR_ARG, "restriction prefix beyond the active particles");
    CHK(flush_vel(h));
    h->rmode = mode;
    h->r_active = count_active;
    h->r_active_nbrs = count_active_neighbors;
    h->true_count = true_particle_count;
    // lists of a restricted search cover its restricted set only; those of an
    // unrestricted one cover every particle and serve any restriction
    if (h->lists_rmode != 0) h->lists_valid = false;
    return DFSPH_OK;
}

int dfsph_count_tagged(void* hv, uint32_t tag, uint32_t* count)
{
    Handle* h = H(hv);
    if (!count) return fail(DFSPH_ERR_ARG, "null count");
    CHK(flush_vel(h));
    dfsph::launch_fill_u32(h->s, &h->ctl->pad, 0u, 1);
    dfsph::launch_count_tag(h->s, h->tag[h->pc], h->n, tag, &h->ctl->pad);
    HIPCHK(hipMemcpyAsync(&h->hctl->pad, &h->ctl->pad, sizeof(uint32_t), hipMemcpyDeviceToHost, h->s));
    CHK(sync(h));
    *count = h->hctl->pad;
    return DFSPH_OK;
}

int dfsph_tag_neighbors(void* hv, uint32_t tag_in, uint32_t tag_out, int untagged_only, uint32_t* changed)
{
    Handle* h = H(hv);
    if (!h->lists_valid) return fail(DFSPH_ERR_STATE, "neighbor lists are stale: call dfsph_neighbor_search");
    CHK(flush_vel(h));
    const uint32_t n = h->n;
    if (n == 0) { if (changed) *changed = 0; return DFSPH_OK; }
    // snapshot of the tags in the spare buffer, then mark into the live array
    uint32_t* snap = h->tag[h->pc ^ 1];
    HIPCHK(hipMemcpyAsync(snap, h->tag[h->pc], 4ull * n, hipMemcpyDeviceToDevice, h->s));
    dfsph::launch_tag_neighbors(h->s, h->c, n, lists(h), snap, h->tag[h->pc], tag_in, tag_out, untagged_only,
                                dfsph::grid_blocks_xcd(n));
    dfsph::launch_fill_u32(h->s, &h->ctl->pad, 0u, 1);
    dfsph::launch_count_changed(h->s, snap, h->tag[h->pc], n, &h->ctl->pad);
    HIPCHK(hipMemcpyAsync(&h->hctl->pad, &h->ctl->pad, sizeof(uint32_t), hipMemcpyDeviceToHost, h->s));
    CHK(sync(h));
    if (changed) *changed = h->hctl->pad;
    return DFSPH_OK;
}

int dfsph_remove_density_peaks(void* hv, float limit, float boundary_limit, uint32_t* removed, double* avg_density)
{
    Handle* h = H(hv);
    CHK(flush_vel(h));
    if (!h->lists_valid) return fail(DFSPH_ERR_STATE, "neighbor lists are stale: call dfsph_neighbor_search");
    if (h->rmode != 0) return fail(DFSPH_ERR_STATE, "dfsph_remove_density_peaks runs without a restriction");
    const uint32_t n = h->n;
    if (n == 0) {
        if (removed) *removed = 0;
        if (avg_density) *avg_density = 0.0;
        return DFSPH_OK;
    }
    if (!h->rest_sum) CHK(dalloc(h->rest_sum, 2));
    CHK(density_only(h));
    dfsph::launch_fill_u32(h->s, reinterpret_cast<uint32_t*>(h->rest_sum), 0u, 4);   // 0.0, 0.0
    dfsph::launch_fill_u32(h->s, h->keep + n, 0u, 1);              // the compaction scan's sentinel
    dfsph::launch_rest_extract(h->s, h->c, n, lists(h), h->pos[h->pc], h->bpos, h->rho, h->tag[h->pc], limit,
                               boundary_limit, h->keep, h->rest_sum, dfsph::grid_blocks_xcd(n));
    HIPCHK(hipGetLastError());
    double sum[2] = {0.0, 0.0};
    HIPCHK(hipMemcpyAsync(sum, h->rest_sum, sizeof(sum), hipMemcpyDeviceToHost, h->s));
    uint32_t rm = 0;
    CHK(compact(h, &rm));                 // synchronises the stream
    if (removed) *removed = rm;
    if (avg_density) *avg_density = sum[1] > 0.0 ? sum[0] / sum[1] : 0.0;
    return DFSPH_OK;
}

int dfsph_order_active_first(void* hv, uint32_t* count_active, uint32_t* count_active_neighbors)
{
    Handle* h = H(hv);
    CHK(flush_vel(h));
    const uint32_t n = h->n;
    // f0 / r0 in keep / rank, f1 / r1 in the sort scratch (key / slot)
    dfsph::launch_tag_classes(h->s, h->tag[h->pc], n, h->keep, h->key);
    dfsph::launch_exclusive_scan(h->s, h->keep, h->rank, n + 1, h->scan_tmp);
    dfsph::launch_exclusive_scan(h->s, h->key, h->slot, n + 1, h->scan_tmp);
    const int o = h->pc ^ 1, vo = h->vc ^ 1;
    dfsph::launch_partition3(h->s, n, h->keep, h->rank, h->key, h->slot, fluid_ptrs(h, h->pc, h->vc),
                             fluid_ptrs(h, o, vo));
    uint32_t c2[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(&h->hctl->pad, h->rank + n, sizeof(uint32_t), hipMemcpyDeviceToHost, h->s));
    CHK(sync(h));
    c2[0] = h->hctl->pad;
    HIPCHK(hipMemcpyAsync(&h->hctl->pad, h->slot + n, sizeof(uint32_t), hipMemcpyDeviceToHost, h->s));
    CHK(sync(h));
    c2[1] = h->hctl->pad;
    h->pc = o;
    h->vc = vo;
    h->lists_valid = false;
    h->acc_valid = false;
    h->pred_valid = false;
    if (count_active) *count_active = c2[0];
    if (count_active_neighbors) *count_active_neighbors = c2[1];
    return DFSPH_OK;
}

int dfsph_remove_tagged(void* hv, uint32_t tag, uint32_t* removed)
{
    Handle* h = H(hv);
    CHK(flush_vel(h));
    dfsph::launch_keep_not_tag(h->s, h->tag[h->pc], h->n, tag, h->keep);
    return compact(h, removed);
}

// ---- open boundary ----
// (a surface checked by check_surface; a mesh id refers to an uploaded mesh)
static dfsph::Surface to_surface(const Handle* h, const dfsph_surface& a)
{
    dfsph::Surface S{};
    S.type = a.type;
    S.reversed = a.reversed;
    for (int k = 0; k < 3; ++k) {
        S.o[k] = a.o[k]; S.n[k] = a.n[k]; S.center[k] = a.center[k]; S.half[k] = a.half[k]; S.dir[k] = a.direction[k];
    }
    S.radius = a.radius;
    S.radius_internal = a.radius_internal;
    S.points = a.points;
    if (a.type == DFSPH_SURFACE_MESH) {
        const Handle::MeshRec& m = h->meshes[a.mesh - 1];
        S.tri = m.tri;
        S.ntri = m.ntri;
        for (int k = 0; k < 3; ++k) { S.bmin[k] = m.bmin[k]; S.bmax[k] = m.bmax[k]; }
    }
    return S;
}

static int check_surface(const Handle* h, const dfsph_surface* a)
{
    if (!a) return fail(DFSPH_ERR_ARG, "null surface");
    const int t = a->type;
    if (t != DFSPH_SURFACE_PLANE && t != DFSPH_SURFACE_CUBOID && t != DFSPH_SURFACE_CYLINDER &&
        t != DFSPH_SURFACE_SPHERE && t != DFSPH_SURFACE_STAR && t != DFSPH_SURFACE_MESH)
        return fail(DFSPH_ERR_ARG, "unknown surface type");
    if (t == DFSPH_SURFACE_STAR && a->points == 0) return fail(DFSPH_ERR_ARG, "star surface needs points > 0");
    if (t == DFSPH_SURFACE_MESH && (a->mesh == 0 || a->mesh > h->meshes.size()))
        return fail(DFSPH_ERR_ARG, "mesh surface: no mesh with this id (dfsph_mesh_upload)");
    return DFSPH_OK;
}

int dfsph_mesh_upload(void* hv, const float* vertices, uint32_t nv, const uint32_t* faces, uint32_t nf,
                      uint32_t* mesh)
{
    Handle* h = H(hv);
    if (!vertices || !faces || !mesh || nv < 3 || nf < 4)
        return fail(DFSPH_ERR_ARG, "a closed mesh needs vertices and at least 4 faces");
    Handle::MeshRec m;
    std::vector<float4> T(3 * (size_t)nf);
    for (int k = 0; k < 3; ++k) { m.bmin[k] = vertices[k]; m.bmax[k] = vertices[k]; }
    for (uint32_t v = 0; v < nv; ++v)
        for (int k = 0; k < 3; ++k) {
            m.bmin[k] = std::min(m.bmin[k], vertices[3 * v + k]);
            m.bmax[k] = std::max(m.bmax[k], vertices[3 * v + k]);
        }
    for (uint32_t f = 0; f < nf; ++f)
        for (int c = 0; c < 3; ++c) {
            const uint32_t v = faces[3 * f + c];
            if (v >= nv) return fail(DFSPH_ERR_ARG, "face index beyond the vertices");
            T[3 * (size_t)f + c] = make_float4(vertices[3 * v], vertices[3 * v + 1], vertices[3 * v + 2], 0.0f);
        }
    CHK(dalloc(m.tri, T.size()));
    m.ntri = nf;
    if (hipMemcpy(m.tri, T.data(), sizeof(float4) * T.size(), hipMemcpyHostToDevice) != hipSuccess) {
        dfree(m.tri);
        return fail(DFSPH_ERR_HIP, "hipMemcpy");
    }
    h->meshes.push_back(m);
    *mesh = (uint32_t)h->meshes.size();
    return DFSPH_OK;
}

int dfsph_surface_inside(void* hv, const dfsph_surface* s, const float* xyz, uint32_t n, uint8_t* inside)
{
    Handle* h = H(hv);
    CHK(check_surface(h, s));
    if (n == 0) return DFSPH_OK;
    if (!xyz || !inside) return fail(DFSPH_ERR_ARG, "null points");
    std::vector<float4> P(n);
    for (uint32_t i = 0; i < n; ++i) P[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.0f);
    float4* d = nullptr;
    uint32_t* k = nullptr;
    CHK(dalloc(d, n));
    int rc = dalloc(k, n + 1);
    std::vector<uint32_t> K(n);
    if (rc == DFSPH_OK) {
        if (hipMemcpyAsync(d, P.data(), sizeof(float4) * n, hipMemcpyHostToDevice, h->s) != hipSuccess)
            rc = fail(DFSPH_ERR_HIP, "hipMemcpy");
        if (rc == DFSPH_OK) {
            dfsph::launch_surface_keep(h->s, d, n, to_surface(h, *s), 1, k);
            if (hipMemcpyAsync(K.data(), k, 4ull * n, hipMemcpyDeviceToHost, h->s) != hipSuccess)
                rc = fail(DFSPH_ERR_HIP, "hipMemcpy");
        }
        if (rc == DFSPH_OK) rc = sync(h);
    }
    dfree(d);
    dfree(k);
    if (rc != DFSPH_OK) return rc;
    for (uint32_t i = 0; i < n; ++i) inside[i] = (uint8_t)K[i];
    return DFSPH_OK;
}

int dfsph_open_boundary_init(void* hv, const dfsph_surface* boundary, const dfsph_surface* interior,
                             const dfsph_surface* fluid_surface, const float* inflow_xyz, uint32_t n, uint32_t* kept)
{
    Handle* h = H(hv);
    CHK(check_surface(h, boundary));
    CHK(check_surface(h, interior));
    CHK(check_surface(h, fluid_surface));
    if (n && !inflow_xyz) return fail(DFSPH_ERR_ARG, "null inflow positions");
    h->ob_boundary = to_surface(h, *boundary);
    h->ob_interior = to_surface(h, *interior);
    h->ob_surface = to_surface(h, *fluid_surface);
    for (int a = 0; a < 3; ++a) h->ob_disp[a] = 0.0;
    h->ob_n = 0;
    dfree(h->ob_cand);
    if (n) {
        std::vector<float4> P(n);
        for (uint32_t i = 0; i < n; ++i) P[i] = make_float4(inflow_xyz[3 * i], inflow_xyz[3 * i + 1], inflow_xyz[3 * i + 2], 0.0f);
        float4* raw = nullptr;
        uint32_t *k = nullptr, *r = nullptr, *tmp = nullptr;
        CHK(dalloc(raw, n));
        CHK(dalloc(h->ob_cand, n));
        CHK(dalloc(k, n + 1));
        CHK(dalloc(r, n + 1));
        CHK(dalloc(tmp, dfsph::scan_temp_words(n + 1)));
        HIPCHK(hipMemcpyAsync(raw, P.data(), sizeof(float4) * n, hipMemcpyHostToDevice, h->s));
        // keep the positions outside the interior (the reference tags those
        // outside the reversed interior, OpenBoundariesSimple.cu:300-318) and
        // inside the fluid surface (:320-326), in their order
        dfsph::Surface rev = h->ob_interior;
        rev.reversed = !rev.reversed;
        dfsph::launch_surface_keep(h->s, raw, n, rev, 1, k);
        dfsph::launch_surface_keep_and(h->s, raw, n, h->ob_surface, 1, k);
        dfsph::launch_exclusive_scan(h->s, k, r, n + 1, tmp);
        dfsph::launch_compact_f4(h->s, n, k, r, raw, h->ob_cand);
        HIPCHK(hipMemcpyAsync(&h->hctl->pad, r + n, sizeof(uint32_t), hipMemcpyDeviceToHost, h->s));
        CHK(sync(h));
        h->ob_n = h->hctl->pad;
        dfree(raw);
        dfree(k);
        dfree(r);
        dfree(tmp);
    }
    if (h->ob_n) {
        CHK(dalloc(h->ob_k, h->ob_n + 1));
        CHK(dalloc(h->ob_r, h->ob_n + 1));
        CHK(dalloc(h->ob_tmp, dfsph::scan_temp_words(h->ob_n + 1)));
        CHK(dalloc(h->ob_vnew, h->ob_n));
    }
    h->ob_ready = true;
    if (kept) *kept = h->ob_n;
    return DFSPH_OK;
}

int dfsph_open_boundary_apply(void* hv, const double* window_displacement, float allowed_new_distance,
                              float allowed_new_density, int use_inflow, int use_outflow, uint32_t* added,
                              uint32_t* removed)
{
    Handle* h = H(hv);
    if (!h->ob_ready) return fail(DFSPH_ERR_STATE, "dfsph_open_boundary_init was not called");
    if (added) *added = 0;
    if (removed) *removed = 0;
    CHK(flush_vel(h));
    // follow the dynamic window (OpenBoundariesSimple.cu:553-584)
    if (window_displacement) {
        double d[3], sq = 0.0;
        for (int a = 0; a < 3; ++a) { d[a] = window_displacement[a] - h->ob_disp[a]; sq += d[a] * d[a]; }
        if (sq > 1e-6) {
            for (dfsph::Surface* S : {&h->ob_boundary, &h->ob_interior, &h->ob_surface})
                for (int a = 0; a < 3; ++a) { S->o[a] += (float)d[a]; S->center[a] += (float)d[a]; }
            dfsph::launch_apply_delta(h->s, h->ob_cand, h->ob_n, (float)d[0], (float)d[1], (float)d[2]);
            for (int a = 0; a < 3; ++a) h->ob_disp[a] = window_displacement[a];
        }
    }
    if (use_inflow) {
        if (!(allowed_new_distance > 0.0f) || !(allowed_new_density > 0.0f))
            return fail(DFSPH_ERR_ARG, "inflow needs a positive distance and density");
        // room for the inflow: the capacity doubles once more than 3/4 of it
        // is in use (OpenBoundariesSimple.cu:570-573)
        if (h->n > 0.75 * (double)h->cap) CHK(grow_capacity(h, 2 * h->cap));
        const uint32_t n = h->n, nc = h->ob_n;
        if (nc && n) {
            // cell table of the current positions without reordering
            CHK(counting_sort(h, h->pos[h->pc], n, h->cstart));
            uint32_t *k = h->ob_k, *r = h->ob_r, *tmp = h->ob_tmp;
            float4* vnew = h->ob_vnew;
            float4 p0;
            HIPCHK(hipMemcpyAsync(&p0, h->pos[h->pc], sizeof(float4), hipMemcpyDeviceToHost, h->s));
            CHK(sync(h));
            const float mass = p0.w;
            dfsph::launch_inflow(h->s, h->c, h->ob_cand, nc, h->pos[h->pc], h->vel[h->vc], h->cstart, h->perm,
                                 h->bpos, h->bcstart, h->ob_boundary, allowed_new_distance, allowed_new_density,
                                 mass, k, vnew);
            dfsph::launch_exclusive_scan(h->s, k, r, nc + 1, tmp);
            HIPCHK(hipMemcpyAsync(&h->hctl->pad, r + nc, sizeof(uint32_t), hipMemcpyDeviceToHost, h->s));
            CHK(sync(h));
            const uint32_t m = h->hctl->pad;
            int rc = DFSPH_OK;
            if (n + m > h->cap) {
                // more inflow than the doubled capacity holds: grow to fit (the
                // candidate flags, ranks and velocities are inflow scratch and
                // survive; the cell table is not used any more)
                rc = grow_capacity(h, std::max(2 * h->cap, n + m));
            }
            if (rc == DFSPH_OK && m) {
                dfsph::launch_append_inflow(h->s, h->ob_cand, nc, k, r, vnew, n, mass, h->next_id,
                                            fluid_ptrs(h, h->pc, h->vc));
                h->n = n + m;
                h->next_id += m;
                HIPCHK(hipGetLastError());
            }
            if (rc != DFSPH_OK) return rc;
            if (added) *added = m;
            h->lists_valid = false;
            h->acc_valid = false;
            h->pred_valid = false;
        }
    }
    if (use_outflow && h->n) {
        dfsph::launch_outflow_keep(h->s, h->pos[h->pc], h->n, h->ob_interior, h->ob_surface, h->keep);
        CHK(compact(h, removed));
    }
    return DFSPH_OK;
}

int dfsph_load_fluid(void* hv, const float* xyz, uint32_t nc, float min_distance, float max_density,
                     int interpolate_velocity, uint32_t* added)
{
    Handle* h = H(hv);
    if (added) *added = 0;
    if (!h->grid_ready || !h->pos[0]) return fail(DFSPH_ERR_STATE, "boundary and fluid must be uploaded first");
    if (nc && !xyz) return fail(DFSPH_ERR_ARG, "null candidates");
    if (!(min_distance > 0.0f) || !(max_density > 0.0f)) return fail(DFSPH_ERR_ARG, "positive distance and density");
    if (nc == 0 || h->n == 0) return DFSPH_OK;
    CHK(flush_vel(h));
    // candidates inside the grid only (the cell table covers the boundary's box)
    std::vector<float4> P;
    P.reserve(nc);
    for (uint32_t i = 0; i < nc; ++i) {
        const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        if (x >= h->bmin[0] && x <= h->bmax[0] && y >= h->bmin[1] && y <= h->bmax[1] && z >= h->bmin[2] &&
            z <= h->bmax[2])
            P.push_back(make_float4(x, y, z, 0.0f));
    }
    const uint32_t m = (uint32_t)P.size();
    if (m == 0) return DFSPH_OK;
    float4 *cand = nullptr, *vnew = nullptr;
    uint32_t *k = nullptr, *r = nullptr, *tmp = nullptr;
    auto release = [&] { dfree(cand); dfree(vnew); dfree(k); dfree(r); dfree(tmp); };
    int rc = DFSPH_OK;
    if ((rc = dalloc(cand, m)) || (rc = dalloc(vnew, m)) || (rc = dalloc(k, m + 1)) || (rc = dalloc(r, m + 1)) ||
        (rc = dalloc(tmp, dfsph::scan_temp_words(m + 1)))) {
        release();
        return rc;
    }
    auto body = [&]() -> int {
        HIPCHK(hipMemcpyAsync(cand, P.data(), sizeof(float4) * m, hipMemcpyHostToDevice, h->s));
        // rigid bodies at their current places in the boundary table (a body
        // moved since the last search counts where it is now)
        if (h->ns) CHK(place_bodies(h));
        const uint32_t n = h->n;
        CHK(counting_sort(h, h->pos[h->pc], n, h->cstart));
        float4 p0;
        HIPCHK(hipMemcpyAsync(&p0, h->pos[h->pc], sizeof(float4), hipMemcpyDeviceToHost, h->s));
        CHK(sync(h));
        // no direction rule: a plane surface with a zero normal admits every velocity
        dfsph::Surface any{};
        any.type = DFSPH_SURFACE_PLANE;
        dfsph::launch_inflow(h->s, h->c, cand, m, h->pos[h->pc], h->vel[h->vc], h->cstart, h->perm, h->bpos,
                             h->bcstart, any, min_distance, max_density, p0.w, k, vnew);
        if (!interpolate_velocity) dfsph::launch_fill_f32(h->s, reinterpret_cast<float*>(vnew), 0.0f, 4 * m);
        dfsph::launch_exclusive_scan(h->s, k, r, m + 1, tmp);
        HIPCHK(hipMemcpyAsync(&h->hctl->pad, r + m, sizeof(uint32_t), hipMemcpyDeviceToHost, h->s));
        CHK(sync(h));
        const uint32_t a = h->hctl->pad;
        if (n + a > h->cap) CHK(grow_capacity(h, std::max(2 * h->cap, n + a)));
        if (a) {
            dfsph::launch_append_inflow(h->s, cand, m, k, r, vnew, n, p0.w, h->next_id, fluid_ptrs(h, h->pc, h->vc));
            HIPCHK(hipGetLastError());
            h->n = n + a;
            h->next_id += a;
            h->lists_valid = false;
            h->acc_valid = false;
            h->pred_valid = false;
        }
        if (added) *added = a;
        return sync(h);
    };
    rc = body();
    release();
    return rc;
}

int dfsph_change_max_particle_number(void* hv, uint32_t max_particles)
{
    Handle* h = H(hv);
    if (max_particles < h->n) return fail(DFSPH_ERR_CAPACITY, "the new capacity is below the particle count");
    if (!h->pos[0]) return fail(DFSPH_ERR_STATE, "no fluid uploaded");
    return grow_capacity(h, max_particles);
}

int dfsph_profile_enable(void* hv, uint32_t mask)
{
    H(hv)->prof_mask = mask & ((1u << DFSPH_KERNEL_COUNT) - 1u);
    return DFSPH_OK;
}

int dfsph_profile_reset(void* hv)
{
    Handle* h = H(hv);
    CHK(sync(h));
    prof_collect(h);
    for (int k = 0; k < DFSPH_KERNEL_COUNT; ++k) { h->prof_launches[k] = 0; h->prof_ms[k] = 0.0; }
    return DFSPH_OK;
}

int dfsph_profile_read(void* hv, uint32_t kernel, uint32_t* launches, double* total_ms)
{
    Handle* h = H(hv);
    if (kernel >= DFSPH_KERNEL_COUNT) return fail(DFSPH_ERR_ARG, "unknown kernel id");
    CHK(sync(h));
    prof_collect(h);
    if (launches) *launches = h->prof_launches[kernel];
    if (total_ms) *total_ms = h->prof_ms[kernel];
    return DFSPH_OK;
}

int dfsph_add_rigid_body(void* hv, const dfsph_rigid_body* body, const float* xyz, const float* psi, uint32_t n,
                         uint32_t* id)
{
    Handle* h = H(hv);
    if (!body || !xyz || n == 0) return fail(DFSPH_ERR_ARG, "a body needs at least one particle");
    if (!h->grid_ready) return fail(DFSPH_ERR_STATE, "upload the boundary set first");
    CHK(flush_vel(h));
    CHK(read_ctl(h));                           // also waits for the stream
    dfsph::DevBody d;
    CHK(body_from_abi(body, h->htm->time, d));
    const uint32_t bid = (uint32_t)h->bodies.size();
    if (bid) HIPCHK(hipMemcpy(h->bodies.data(), h->dbody, sizeof(dfsph::DevBody) * bid, hipMemcpyDeviceToHost));
    d.first = h->ns;
    d.count = n;
    const uint32_t first = h->ns, ns = h->ns + n, nb = h->nbs + ns;
    std::vector<float4> P(n);
    for (uint32_t i = 0; i < n; ++i) P[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], psi ? psi[i] : 0.0f);
    std::vector<uint32_t> B(n, bid);
    CHK(grow(h, h->x0, first, ns));
    CHK(grow(h, h->sbody, first, ns));
    HIPCHK(hipMemcpy(h->x0 + first, P.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->sbody + first, B.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    // walls in cell order lead the raw boundary array (bpos holds only walls
    // before the first body)
    if (!h->braw) {
        CHK(dalloc(h->braw, nb));
        HIPCHK(hipMemcpyAsync(h->braw, h->bpos, sizeof(float4) * h->nbs, hipMemcpyDeviceToDevice, h->s));
    } else {
        CHK(grow(h, h->braw, h->nbs, nb));
    }
    CHK(dalloc(h->bvraw, nb));
    dfsph::launch_wall_vel(h->s, h->bvraw, h->nbs);
    CHK(sync(h));
    CHK(dalloc(h->bvel, nb));
    CHK(dalloc(h->bpos, nb));
    CHK(dalloc(h->sforce, 4ull * ns));
    h->bodies.push_back(d);
    CHK(dalloc(h->dbody, h->bodies.size()));
    HIPCHK(hipMemcpy(h->dbody, h->bodies.data(), sizeof(dfsph::DevBody) * h->bodies.size(), hipMemcpyHostToDevice));
    h->ns = ns;
    h->nb = nb;
    if (h->fscale == 0.0f) {
        // 2^20 units per weight of one fluid particle (0.8 (2r)^3 rho0 |g|)
        const double d2 = 2.0 * h->p.particle_radius;
        const double g = std::sqrt(h->p.gravity[0] * h->p.gravity[0] + h->p.gravity[1] * h->p.gravity[1] +
                                   h->p.gravity[2] * h->p.gravity[2]);
        h->fscale = (float)(1048576.0 / (0.8 * d2 * d2 * d2 * h->p.density0 * std::max(g, 1.0)));
    }
    CHK(ensure_tmp(h, std::max(nb, h->cap)));
    CHK(place_bodies(h));
    if (!psi) {
        float* tmp = nullptr;
        CHK(dalloc(tmp, nb));
        dfsph::launch_boundary_psi(h->s, h->c, h->bpos, h->bcstart, nb, tmp);
        dfsph::launch_solid_psi_take(h->s, h->bvel, tmp, nb, h->x0, first, ns);
        int rc = place_bodies(h);
        if (rc == DFSPH_OK) rc = sync(h);
        dfree(tmp);
        CHK(rc);
    }
    CHK(sync(h));
    h->lists_valid = false;
    if (id) *id = bid;
    return DFSPH_OK;
}

int dfsph_set_rigid_body(void* hv, uint32_t id, const dfsph_rigid_body* body)
{
    Handle* h = H(hv);
    if (!body || id >= h->bodies.size()) return fail(DFSPH_ERR_ARG, "no such body");
    CHK(flush_vel(h));
    CHK(read_ctl(h));
    dfsph::DevBody d;
    CHK(body_from_abi(body, h->htm->time, d));
    d.first = h->bodies[id].first;
    d.count = h->bodies[id].count;
    h->bodies[id] = d;
    HIPCHK(hipMemcpy(h->dbody + id, &d, sizeof(d), hipMemcpyHostToDevice));
    h->lists_valid = false;
    return DFSPH_OK;
}

int dfsph_get_rigid_body(void* hv, uint32_t id, dfsph_rigid_body* body, double* force, double* torque)
{
    Handle* h = H(hv);
    if (id >= h->bodies.size()) return fail(DFSPH_ERR_ARG, "no such body");
    CHK(flush_vel(h));
    CHK(sync(h));
    dfsph::DevBody d;
    HIPCHK(hipMemcpy(&d, h->dbody + id, sizeof(d), hipMemcpyDeviceToHost));
    if (body) {
        std::memset(body, 0, sizeof(*body));
        body->type = d.type;
        body->mass = d.mass;
        for (int a = 0; a < 3; ++a) {
            body->inertia[a] = d.inertia[a];
            body->position[a] = d.c[a];
            body->velocity[a] = d.v[a];
            body->angular_velocity[a] = d.w[a];
            body->axis[a] = d.axis[a];
        }
        for (int a = 0; a < 4; ++a) body->rotation[a] = d.q[a];
        body->amplitude = d.amp;
        body->period = d.omega > 0.0 ? 2.0 * M_PI / d.omega : 0.0;
    }
    for (int a = 0; a < 3; ++a) {
        if (force) force[a] = d.force[a];
        if (torque) torque[a] = d.torque[a];
    }
    return DFSPH_OK;
}

int dfsph_rigid_body_forces(void* hv, uint32_t id, float* f, uint32_t count)
{
    Handle* h = H(hv);
    if (id >= h->bodies.size()) return fail(DFSPH_ERR_ARG, "no such body");
    const dfsph::DevBody& d = h->bodies[id];
    if (!f || count != 3 * d.count) return fail(DFSPH_ERR_ARG, "count must be 3 x the body's particles");
    CHK(flush_vel(h));
    CHK(sync(h));
    std::vector<unsigned long long> raw(4ull * d.count);
    HIPCHK(hipMemcpy(raw.data(), h->sforce + 4ull * d.first, sizeof(unsigned long long) * raw.size(),
                     hipMemcpyDeviceToHost));
    const double inv = 1.0 / (double)h->fscale;
    for (uint32_t k = 0; k < d.count; ++k)
        for (int a = 0; a < 3; ++a) f[3 * k + a] = (float)((double)(long long)raw[4ull * k + a] * inv);
    return DFSPH_OK;
}

int dfsph_rigid_body_count(void* hv, uint32_t* n)
{
    Handle* h = H(hv);
    if (!n) return fail(DFSPH_ERR_ARG, "null count");
    *n = (uint32_t)h->bodies.size();
    return DFSPH_OK;
}

}  // extern "C"
