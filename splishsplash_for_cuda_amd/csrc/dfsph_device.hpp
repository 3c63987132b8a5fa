// dfsph_device.hpp -- device-side helpers shared by the DFSPH kernels (gfx950).
//
// Data layout in HBM (structure of arrays, particles in cell-sorted order):
//   pos   float4 (x, y, z, mass)        fluid; boundary: (x, y, z, psi)
//   vel   float4 (vx, vy, vz, 0)
//   rho, alpha, radv, k, kappa, kappaV  float
//   Neighbor lists are "blocked-interleaved": entry k of particle i lies in
//   group k / G at [(k / G) * capacity + i] (16 B per group), entry k % G
//   inside it, so a wavefront reads 64 consecutive 16-byte groups and a batch
//   of 8 neighbors is one or two vector loads per lane.
//   nbr16 uint16 groups of 8: fluid lists of blocks with a staged set
//         (indices into the block's set)
//   nbrf  uint32 groups of 4: fluid lists of blocks without one (global ids)
//   nbrb  uint32 groups of 4: boundary lists (boundary ids)
//   cnt   uint32  fluid count | boundary count << 16
//   stg   uint32 [blocks][kStageCap]  global ids of each block's staged set
//
// Cells are ordered in 8x8x8-cell tiles (tiles x-fastest, then z, then y),
// Morton order inside a tile, so a block of consecutive particles occupies a
// compact region: the neighbor set of a 256-particle block is about 5.2 times
// the block, that of a 512-particle block about 3.9 times (C2, at rest).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dfsph {

// Particles per block of every per-particle kernel; the sweep blocks own the
// neighbor sets, so this also sets the LDS stages below.  512 (8 wavefronts):
// a 512-particle set is about 3.9 times the block against 5.2 for 256, so the
// stages hold more particles per wavefront and the position + velocity sweeps
// run 16 wavefronts per CU instead of 12 (C4: 4.71 against 5.20 ms per step).
#ifndef DFSPH_BLOCK
#define DFSPH_BLOCK 512
#endif
constexpr int kBlock = DFSPH_BLOCK;
static_assert(kBlock == 256 || kBlock == 512, "256 or 512 threads: 4 or 8 wavefronts of 64");
constexpr int kNumXcd = 8;         // MI355X: 8 XCDs, blocks dispatched round robin
constexpr float kEps = 1.0e-5f;    // m_eps (DFSPH_define_cuda.h:7)

// Constants of one simulation, passed by value to every kernel.
struct SimConst {
    float hk, hk2, inv_hk;         // kernel support h_k = 4r, h_k^2, 1/h_k
    float kcoef;                   // 8 / (pi h_k^3)
    float lh2;                     // (48 / (pi h_k^3)) / h_k^2
    float lh;                      // (48 / (pi h_k^3)) / h_k
    float W0;                      // W(0)
    float hk2_4;                   // (h_k / 2)^2: the branch point q = 1/2
    float gin_a, gin_b;            // gradient, q <= 1/2: gin_a r + gin_b = l (3q - 2) / h^2
    float gout;                    // gradient, q > 1/2: gout (h_k - r)^2 / r = -l (1-q)^2 / (r h)
    float density0;
    float gx, gy, gz;
    float viscosity;
    float org[3];                  // grid origin
    float inv_cell;                // 1 / cell edge (cell edge = 1.001 h_k)
    int dim[3];                    // cells per axis
    int tiles[3];                  // 8-cell tiles per axis (dim rounded up / 8)
    const uint32_t* tile_rank;     // tile (x + tiles0 (z + tiles2 y)) -> its place in Morton order
    uint32_t ncell;                // cell keys: tiles0 tiles1 tiles2 512
    uint32_t cap;                  // neighbor list stride (fluid capacity)
    uint32_t max_nbr;
    uint32_t stage_lds;            // largest neighbor set staged in LDS (<= kStageMax)
    uint32_t stage_cap;            // largest neighbor set with 16-bit lists (<= kStageCap)
    float m_uniform;               // the mass of every fluid particle when they are all equal, else 0
};

// Time-step scalars live on the device so a whole step can be queued
// without host round trips (DFSPHCData h, h_future, h_past, ratios,
// DFSPH_c_arrays_structure.h:345-356).
struct DevTime {
    float h, h_future, h_past;
    float inv_h, inv_h_future;
    float ratio, ratio2;           // h_ratio_to_past, h_ratio_to_past2
    float pad;
    double hd, hfd, hpd;           // the same in FP64
    double time;                   // simulated time (FP64 accumulator)
};

// Device-side control of one Jacobi loop: a one-block kernel after each
// loop-end sweep turns the FP64 block partial sums into the average error and
// decides whether the loop continues, with the reference's condition
// while ((err > eta) || (it < min)) && (it < max)  (DFSPH_core_cuda.cu:828,1378).
struct SolveCtl {
    double sum, err, eta;
    double n_div;                  // particle count the error is averaged over
    uint32_t iters, done, min_iter, max_iter;
    uint32_t ticket, pad;
};

struct DevCtl {
    SolveCtl div, dens;
    float cfl_max;                 // max |v + a h|^2
    uint32_t cfl_ticket;
    uint32_t escaped;              // particles outside the bounding box
    uint32_t max_seen;             // largest neighbor count of the last search
    uint32_t overflow;             // particles whose list was truncated
    uint32_t n_active;             // active fluid particles (device copy)
    uint32_t unstaged_blocks;      // blocks of the last search that use global lists
    uint32_t pad;
    unsigned long long nbr_fluid;  // stored fluid neighbors of the last search
    unsigned long long nbr_boundary;
    unsigned long long nbr_fluid_unstaged;  // of which in 32-bit (unstaged) lists
};

// Loop state published to pinned, mapped host memory by the one-block
// finalize kernel of each Jacobi iteration: the host waits on seq (written
// last, after a system-scope fence) instead of a copy and an event.
struct HostRecord {
    SolveCtl sc;
    DevTime tm;
    uint32_t max_seen, overflow, unstaged_blocks, escaped;
    unsigned long long nbr_fluid, nbr_boundary, nbr_fluid_unstaged;
    uint32_t seq, pad[3];
};

// A Jacobi loop's control, set by the first sweep of its solve (block 0)
// (DFSPH_core_cuda.cu:791-795, 1339-1343: eta, the averaging count, limits).
struct CtlInit {
    SolveCtl* sc;                  // nullptr: nothing to set
    const DevTime* t;
    double eta_base, n_div;        // eta = eta_base (/ h when divide_by_h)
    int divide_by_h;
    uint32_t min_iter, max_iter;
};

__device__ __forceinline__ void ctl_init(const CtlInit& ci)
{
    SolveCtl* s = ci.sc;
    s->sum = 0.0;
    s->err = 0.0;
    s->eta = ci.divide_by_h ? ci.eta_base / ci.t->hd : ci.eta_base;
    s->n_div = ci.n_div;
    s->iters = 0;
    s->done = (ci.max_iter == 0) ? 1u : 0u;
    s->min_iter = ci.min_iter;
    s->max_iter = ci.max_iter;
    s->ticket = 0;
}

struct Publish {
    const DevCtl* dc;
    const DevTime* t;
    HostRecord* rec;               // device address of the host record (nullptr: none)
    uint32_t seq;
};

// ---- staged neighbor sets ----
// A sweep block (kBlock consecutive particles) owns the set of fluid particles in
// the cells adjacent to its particles' cells, cell by cell.  Up to
// kStageMax of them are loaded into LDS once per sweep (coalesced) and the
// block's fluid lists are 16-bit indices into the set; sets up to kStageCap
// keep the 16-bit lists and read the particles through the set's global ids;
// larger ones (sparse splashes, more than kAdjMax / 27 own cells, the
// unsorted search) use 32-bit global lists (stage_n = kUnstaged).
// LDS-staged particles per block.  A set read through its global ids instead
// costs about 1.5x per sweep, more than the occupancy a smaller stage would
// win, so the cap holds nearly every set.  256-particle blocks: 2048 (C2 sets:
// mean 1150-1400, p99 1780).  512-particle blocks: 3328, the most that keeps
// three position + k blocks (16 B records, 52 KB) and two position + velocity
// blocks (24 B, 78 KB) in a CU's 160 KB.
#ifndef DFSPH_STAGE_MAX
#if DFSPH_BLOCK == 512
#define DFSPH_STAGE_MAX 3328
#else
#define DFSPH_STAGE_MAX 2048
#endif
#endif
constexpr uint32_t kStageMax = DFSPH_STAGE_MAX;
// The sweeps that stage position and velocity (32 B per particle) may stage
// fewer (a smaller set fits more blocks per CU; larger sets are then read
// through the set's global ids, with the same lists and sums).
#ifndef DFSPH_PV_STAGE
#define DFSPH_PV_STAGE DFSPH_STAGE_MAX
#endif
constexpr uint32_t kStageMaxPV = DFSPH_PV_STAGE;
// The sweeps that stage seven floats per particle (position, velocity and one
// scalar: the XSPH sweep and the divergence warm start, 28 B): 2880 keeps two
// 512-particle blocks (79 KB) in a CU.
#ifndef DFSPH_X_STAGE
#if DFSPH_BLOCK == 512
#define DFSPH_X_STAGE 2880
#else
#define DFSPH_X_STAGE DFSPH_PV_STAGE
#endif
#endif
constexpr uint32_t kStageMaxX = DFSPH_X_STAGE;
#ifndef DFSPH_PV_WAVES
#if DFSPH_BLOCK == 512
#define DFSPH_PV_WAVES 4          // two 8-wave blocks per CU: at most 128 VGPRs
#else
#define DFSPH_PV_WAVES 1
#endif
#endif
constexpr uint32_t kStageCap = 4096;       // largest set with 16-bit lists
constexpr uint32_t kUnstaged = 0xffffffffu;
constexpr uint32_t kSetNoLds = 0x80000000u;  // stage_n flag: 16-bit lists, set not in LDS
constexpr uint32_t kAdjMax = 2048;         // 27 x unique own cells per block
constexpr uint32_t kKeyInvalid = 0xffffffffu;
constexpr uint32_t kGroup16 = 8;           // entries per 16-byte group
constexpr uint32_t kGroup32 = 4;

// list groups allocated per particle for a given max_neighbors: every batch
// of 8 that a sweep reads stays inside
__host__ __device__ inline uint32_t list_groups16(uint32_t max_nbr) { return max_nbr / kGroup16 + 2; }
__host__ __device__ inline uint32_t list_groups32(uint32_t max_nbr) { return max_nbr / kGroup32 + 4; }



// The persistent per-particle fields that move with a particle when the
// set is reordered or compacted.  kappa / kappaV are the reference's
// warm-start accumulators (this step's solver increments); ktot / kvtot are
// the step's totals (applied warm start + increments) in dt-invariant units
// (kappa h_f^2, kappaV h), the inputs and targets of the learned warm start.
struct FluidPtrs {
    float4* pos;
    float4* vel;
    float* kap;
    float* kapv;
    float* ktot;
    float* kvtot;
    uint32_t* id;
    uint32_t* tag;                 // dynamic-window / open-boundary tag
};

// Particle tags of the dynamic window and the open boundary, with the
// reference's values (cuda_compilation_center.cu:2-11).
constexpr uint32_t kTagActive = 10000000u;
constexpr uint32_t kTagActiveNeighbors = 20000000u;
constexpr uint32_t kTagUntagged = 40000000u;
constexpr uint32_t kTagAir = 50000000u;
constexpr uint32_t kTagRemoval = 310000000u;

// A BufferFluidSurface (BufferFluidSurface.cu:421-470) by value: the C ABI's
// dfsph_surface with the same fields.
// the direction of the inside test's ray for mesh surfaces: generic (no
// symmetry axis of a box or lattice), unit length
__device__ constexpr float kMeshRay[3] = {0.31190428f, 0.75361035f, 0.57860795f};
enum SurfaceType : int {
    kSurfPlane = 0, kSurfCuboid = 1, kSurfCylinder = 2, kSurfMesh = 3, kSurfSphere = 4, kSurfStar = 5
};
struct Surface {
    int type;
    int reversed;
    float o[3], n[3];              // plane: point and normal (inside: (p - o).n >= 0)
    float center[3], half[3];      // cuboid half lengths; cylinder / star: half[1] = half height; mesh: offset
    float radius, radius_internal; // cylinder, sphere; star: outer / inner radius
    uint32_t points;               // star
    float dir[3];                  // star: direction of the first point (unit, in x-z)
    const float4* tri;             // mesh: 3 vertices per triangle (mesh coordinates)
    uint32_t ntri;                 // mesh
    float bmin[3], bmax[3];        // mesh: bounding box (mesh coordinates)
};

// Restriction of the solver to part of the fluid (DFSPHCData::restriction_mode,
// count_active, count_active_neighbors, DFSPH_c_arrays_structure.h:410-418):
// mode 1 by tag, mode 2 by index prefix.  "Compute" kernels (velocity and
// position updates) run on the active particles only; "source" kernels
// (density, factor, divergence / density source, neighbor search) also on
// their active neighbors (DFSPH_core_cuda.cu:225-231, 500-506, 2545-2552).
struct Restrict {
    int mode;
    uint32_t n_active, n_source;   // mode 2: prefixes [0, n_active), [0, n_source)
    const uint32_t* tag;           // mode 1
};

__device__ __forceinline__ bool is_compute(const Restrict& r, uint32_t i)
{
    if (r.mode == 1) return r.tag[i] == kTagActive;
    if (r.mode == 2) return i < r.n_active;
    return true;
}

__device__ __forceinline__ bool is_source(const Restrict& r, uint32_t i)
{
    if (r.mode == 1) { const uint32_t t = r.tag[i]; return t == kTagActive || t == kTagActiveNeighbors; }
    if (r.mode == 2) return i < r.n_source;
    return true;
}

// Rigid-body particles (the dynamic bodies of DFSPHCData,
// vector_dynamic_bodies_data, DFSPH_c_arrays_structure.cpp:897-905) are
// moving boundary particles: they sit in the boundary set after the static
// walls, and while any exist the sweeps read the velocity of every boundary
// neighbor and push the reaction of the pressure updates onto the body
// particles (ITER_NEIGHBORS_SOLIDS, DFSPH_core_cuda.cu:467-477, 936-946).
// bvel[j] = (velocity, solid index as bits; kNoSolid for a wall particle).
// Reactions are summed in 64-bit fixed point, so integer atomics make the sum
// independent of their order (fscale units per newton).
constexpr uint32_t kNoSolid = 0xffffffffu;
struct Solids {
    const float4* bvel;            // null: no rigid bodies (walls at rest)
    unsigned long long* force;     // 4 words per solid particle (x, y, z, unused)
    float fscale;
};

struct NbrLists {
    const uint16_t* nbr16;         // 16-bit fluid lists (blocks with a staged set)
    const uint32_t* nbrf;          // 32-bit fluid lists (unstaged blocks)
    const uint32_t* nbrb;          // 32-bit boundary lists
    const uint32_t* cnt;
    const uint32_t* stage_n;       // per block: set size or kUnstaged
    const uint32_t* stg;           // per block: global ids of the set (kStageCap each)
    Restrict rs;                   // solver restriction (mode 0: none)
    Solids sol;
};

// boundary neighbor record of the sweeps: position + psi, velocity + solid index
struct BN {
    float4 p, v;
};

template <bool SOL>
__device__ __forceinline__ BN load_bnd(const float4* __restrict__ bpos, const Solids& S, uint32_t j)
{
    BN r;
    r.p = bpos[j];
    r.v = SOL ? S.bvel[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return r;
}

// The reaction f = -m_i k_i psi_j gradW_ij on body particle j of one velocity
// update (the reference's atomicAdd on body.F, DFSPH_core_cuda.cu:471-476).
template <bool SOL>
__device__ __forceinline__ void add_reaction(const Solids& S, const float4& vj, bool valid, float fx, float fy,
                                             float fz)
{
    if (!SOL || !valid) return;
    const uint32_t sid = __float_as_uint(vj.w);
    if (sid == kNoSolid) return;
    unsigned long long* f = S.force + 4ull * sid;
    atomicAdd(f + 0, (unsigned long long)__float2ll_rn(fx * S.fscale));
    atomicAdd(f + 1, (unsigned long long)__float2ll_rn(fy * S.fscale));
    atomicAdd(f + 2, (unsigned long long)__float2ll_rn(fz * S.fscale));
}

// One rigid body on the device (state in FP64).  Body particles are stored
// in the body frame relative to the center of mass c; the world position is
// R(q) x0 + c and the velocity w x (x - c) + v
// (DFSPH_updateDynamicObjectParticles_kernel, DFSPH_core_cuda.cu:3407-3420).
struct DevBody {
    double c[3], v[3], q[4], w[3];      // center of mass, velocity, orientation (x y z w), angular velocity
    double mass, inertia[3];            // body-frame principal moments
    double force[3], torque[3];         // from the fluid during the last step
    double c0[3], axis[3], amp, omega;  // sinusoidal translation: c = c0 + amp sin(omega t) axis
    double clock;                       // the body's time: advanced by the h of every step it takes
    int32_t type, pad;
    uint32_t first, count;              // particle range in the solid arrays
};
constexpr int kBodyDynamic = 0, kBodyKinematic = 1, kBodySine = 2;

// Blocks are dispatched round robin over the 8 XCDs, each with its own L2.
// Remapping gives every XCD a contiguous range of logical blocks, so
// neighbor gathers of cell-sorted particles stay in one L2.  gridDim.x must
// be a multiple of 8; logical ids >= the needed count are idle.
__device__ __forceinline__ uint32_t xcd_block()
{
    const uint32_t per = gridDim.x / kNumXcd;
    return (blockIdx.x % kNumXcd) * per + blockIdx.x / kNumXcd;
}

__device__ __forceinline__ uint32_t xcd_index()
{
    return xcd_block() * blockDim.x + threadIdx.x;
}

// The neighbor predicate: strict |xi - xj|^2 < h_k^2 in FP32
// (DFSPH_macro_cuda.h:124,149) with a fixed evaluation order
// d2 = fma(dz, dz, fma(dy, dy, dx*dx)); the oracle evaluates the same.
__device__ __forceinline__ float dist2(float dx, float dy, float dz)
{
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Cubic spline (CubicKernelPerso, SPH_kernels.h:42-94) evaluated from the
// squared distance.  q <= 1/2: W = k (6q^3 - 6q^2 + 1), grad = l (3q - 2) / h^2 r;
// q > 1/2: W = 2k (1-q)^3, grad = -l (1-q)^2 / (|r| h) r; zero for q > 1 and
// for |r| <= 1e-6 (gradient).
struct PairK {
    float W;    // kernel value
    float g;    // gradient coefficient: gradW = g * (xi - xj)
};

__device__ __forceinline__ float kernel_grad_coef(const SimConst& c, float r2)
{
    const float r2c = fmaxf(r2, 1.0e-24f);
    const float ir = __builtin_amdgcn_rsqf(r2c);
    const float r = r2c * ir;
    const float q = r * c.inv_hk;
    const float f = 1.0f - q;
    const float inner = c.lh2 * __builtin_fmaf(3.0f, q, -2.0f);
    const float outer = -c.lh * f * f * ir;
    float g = (q <= 0.5f) ? inner : outer;
    g = (q <= 1.0f && r > 1.0e-6f) ? g : 0.0f;
    return g;
}

__device__ __forceinline__ PairK kernel_pair(const SimConst& c, float r2)
{
    const float r2c = fmaxf(r2, 1.0e-24f);
    const float ir = __builtin_amdgcn_rsqf(r2c);
    const float r = r2c * ir;
    const float q = r * c.inv_hk;
    const float f = 1.0f - q;
    PairK o;
    const float w_in = c.kcoef * __builtin_fmaf(6.0f * q * q, q - 1.0f, 1.0f);
    const float w_out = 2.0f * c.kcoef * f * f * f;
    o.W = (q <= 0.5f) ? w_in : w_out;
    o.W = (q <= 1.0f) ? o.W : 0.0f;
    const float inner = c.lh2 * __builtin_fmaf(3.0f, q, -2.0f);
    const float outer = -c.lh * f * f * ir;
    float g = (q <= 0.5f) ? inner : outer;
    o.g = (q <= 1.0f && r > 1.0e-6f) ? g : 0.0f;
    return o;
}

__device__ __forceinline__ float kernel_W(const SimConst& c, float r2)
{
    return kernel_pair(c, r2).W;
}

// The same for a pair taken from a neighbor list, where r2 < h^2 holds by
// construction of the list (the same FP32 test on the same positions), so
// the support test is not needed, and with the constants folded.  ok is
// false for coincident particles (|r| <= 1e-6), whose gradient the reference
// zeroes; the caller folds it into its one final select.
struct PairL {
    float W, g;
    bool ok;
};

__device__ __forceinline__ PairL grad_listed(const SimConst& c, float r2)
{
    // r2 is clamped at 1e-12 (r = 1 um): a coincident pair then has a finite
    // coefficient and a zero offset, so it adds nothing to any sum (its
    // gradient is the coefficient times the offset) and needs no select
    const float ir = __builtin_amdgcn_rsqf(fmaxf(r2, 1.0e-12f));
    const float r = r2 * ir;
    const float d = c.hk - r;
    // both pieces are computed and one is selected: a branch per pair (the
    // select-to-branch transform) costs more than the other piece.  The
    // outer piece keeps the factored form gout (h - r)^2 / r: expanded, its
    // terms cancel near the support's edge.
    const float gi = __builtin_fmaf(c.gin_a, r, c.gin_b), go = (c.gout * ir) * (d * d);
    PairL o;
    o.W = 0.0f;
    o.g = __builtin_unpredictable(r2 <= c.hk2_4) ? gi : go;
    o.ok = true;
    return o;
}

__device__ __forceinline__ PairL pair_listed(const SimConst& c, float r2)
{
    const float r2c = fmaxf(r2, 1.0e-24f);     // W of a coincident pair is W(0)
    const float ir = __builtin_amdgcn_rsqf(r2c);
    const float r = r2c * ir;
    const float q = r * c.inv_hk;
    const float f = 1.0f - q;
    const bool in = __builtin_unpredictable(r2 <= c.hk2_4);
    const float wi = c.kcoef * __builtin_fmaf(6.0f * q * q, q - 1.0f, 1.0f), wo = 2.0f * c.kcoef * f * f * f;
    const float d = c.hk - r;
    const float gi = __builtin_fmaf(c.gin_a, r, c.gin_b), go = (c.gout * ir) * (d * d);
    PairL o;
    o.W = in ? wi : wo;
    o.g = in ? gi : go;
    o.ok = r2 > 1.0e-12f;
    return o;
}

__device__ __forceinline__ float W_listed(const SimConst& c, float r2)
{
    const float r2c = fmaxf(r2, 1.0e-24f);
    const float q = r2c * __builtin_amdgcn_rsqf(r2c) * c.inv_hk;
    const float f = 1.0f - q;
    const float wi = c.kcoef * __builtin_fmaf(6.0f * q * q, q - 1.0f, 1.0f), wo = 2.0f * c.kcoef * f * f * f;
    return __builtin_unpredictable(r2 <= c.hk2_4) ? wi : wo;
}

__device__ __forceinline__ int cell_coord(const SimConst& c, float v, int a)
{
    int k = (int)floorf((v - c.org[a]) * c.inv_cell);
    k = k < 0 ? 0 : k;
    k = k >= c.dim[a] ? c.dim[a] - 1 : k;
    return k;
}

// 3-bit value -> bits 0, 3, 6
__device__ __forceinline__ uint32_t spread3(uint32_t v)
{
    return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4);
}

// Cell key: the tile's place in a Morton order of the tiles * 512 + the Morton
// index of the cell in its tile.  Particles sorted by key fill blocks that
// are compact in 3D at both scales, so the blocks that share a neighbor set
// region run close together in time and on the same XCD (shared L2).
__device__ __forceinline__ uint32_t cell_key(const SimConst& c, int x, int y, int z)
{
    const uint32_t tile = (uint32_t)(x >> 3) +
                          (uint32_t)c.tiles[0] * ((uint32_t)(z >> 3) + (uint32_t)c.tiles[2] * (uint32_t)(y >> 3));
    const uint32_t local = spread3((uint32_t)x & 7u) | (spread3((uint32_t)y & 7u) << 1) | (spread3((uint32_t)z & 7u) << 2);
    return (c.tile_rank[tile] << 9) | local;
}

__device__ __forceinline__ uint32_t cell_id(const SimConst& c, float x, float y, float z)
{
    return cell_key(c, cell_coord(c, x, 0), cell_coord(c, y, 1), cell_coord(c, z, 2));
}

// The 27 cells around (cx, cy, cz) inside the grid, in a fixed order.
template <class F>
__device__ __forceinline__ void for_each_adjacent_cell(const SimConst& c, int cx, int cy, int cz, F&& f)
{
    for (int dy = -1; dy <= 1; ++dy) {
        const int y = cy + dy;
        if (y < 0 || y >= c.dim[1]) continue;
        for (int dz = -1; dz <= 1; ++dz) {
            const int z = cz + dz;
            if (z < 0 || z >= c.dim[2]) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = cx + dx;
                if (x < 0 || x >= c.dim[0]) continue;
                f(cell_key(c, x, y, z));
            }
        }
    }
}

// Neighbor loops with memory-level parallelism: a batch of 8 list entries is
// one 16-byte load (16-bit lists) or two (32-bit lists) per lane, then the 8
// neighbors' gathers can all be in flight before the first one is used.  The
// last batch reads the rest of its group(s), which the list allocation keeps
// in bounds; those entries are replaced by `pad` (a valid index) and passed
// with valid == false, so the callee zeroes their contributions.  Callees
// mask with keep() on values already computed, so the body has no branch the
// compiler could sink a gather into.
// Each batch first issues the loads of all 8 neighbors (load(entry) returns
// the neighbor's data), then runs body(data, valid) on them, so the LDS or
// L2 latency of one neighbor overlaps the others'; the next batch's list
// group is loaded while the current batch computes.
template <class Load, class Body>
__device__ __forceinline__ void for_each_nbr16(const uint16_t* __restrict__ lists, uint32_t cap, uint32_t i,
                                               uint32_t count, uint4 q, uint4 qn, Load&& load, Body&& body)
{
    // q, qn: the particle's groups 0 and 1 (loaded by the caller)
    using T = decltype(load(0u));
    const uint4* p = reinterpret_cast<const uint4*>(lists) + i;
    uint32_t k = 0;
    for (; k + 8 <= count; k += 8) {            // full batches: no masking
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
        q = qn;
        if (k + 16u < count) qn = p[(size_t)((k >> 3) + 2) * cap];   // two groups ahead, when a batch reads it
        T it[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) it[u] = load((w[u >> 1] >> ((u & 1) * 16)) & 0xffffu);
#pragma unroll
        for (int u = 0; u < 8; ++u) body(it[u], true);
    }
    if (k < count) {                            // the last, partial batch
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
        T it[8];
        bool v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            v[u] = k + u < count;
            it[u] = load(v[u] ? (w[u >> 1] >> ((u & 1) * 16)) & 0xffffu : 0u);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) body(it[u], v[u]);
    }
}

template <class Load, class Body>
__device__ __forceinline__ void for_each_nbr16(const uint16_t* __restrict__ lists, uint32_t cap, uint32_t i,
                                               uint32_t count, Load&& load, Body&& body)
{
    const uint4* p = reinterpret_cast<const uint4*>(lists) + i;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for_each_nbr16(lists, cap, i, count, count > 0u ? p[0] : z, count > 8u ? p[cap] : z, load, body);
}

template <class Load, class Body>
__device__ __forceinline__ void for_each_nbr32(const uint32_t* __restrict__ lists, uint32_t cap, uint32_t i,
                                               uint32_t count, uint32_t pad, uint4 a, uint4 b, Load&& load,
                                               Body&& body)
{
    // a, b: the particle's groups 0 and 1 (loaded by the caller)
    using T = decltype(load(0u));
    const uint4* p = reinterpret_cast<const uint4*>(lists) + i;
    uint32_t k = 0;
    for (; k + 8 <= count; k += 8) {
        const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        if (k + 8u < count) a = p[(size_t)((k >> 2) + 2) * cap];    // the next two groups, when read
        if (k + 12u < count) b = p[(size_t)((k >> 2) + 3) * cap];
        T it[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) it[u] = load(w[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u) body(it[u], true);
    }
    if (k < count) {
        const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        T it[8];
        bool v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            v[u] = k + u < count;
            it[u] = load(v[u] ? w[u] : pad);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) body(it[u], v[u]);
    }
}

template <class Load, class Body>
__device__ __forceinline__ void for_each_nbr32(const uint32_t* __restrict__ lists, uint32_t cap, uint32_t i,
                                               uint32_t count, uint32_t pad, Load&& load, Body&& body)
{
    const uint4* p = reinterpret_cast<const uint4*>(lists) + i;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for_each_nbr32(lists, cap, i, count, pad, count > 0u ? p[0] : z, count > 4u ? p[cap] : z, load, body);
}

// the plain forms: f(entry, valid)
template <class F>
__device__ __forceinline__ void for_each_nbr16(const uint16_t* __restrict__ lists, uint32_t cap, uint32_t i,
                                               uint32_t count, F&& f)
{
    for_each_nbr16(lists, cap, i, count, [](uint32_t e) { return e; }, f);
}

template <class F>
__device__ __forceinline__ void for_each_nbr32(const uint32_t* __restrict__ lists, uint32_t cap, uint32_t i,
                                               uint32_t count, uint32_t pad, F&& f)
{
    for_each_nbr32(lists, cap, i, count, pad, [](uint32_t e) { return e; }, f);
}

// Entries of the 16-bit lists are LDS byte offsets of float4 slots (16 x the
// index in the block's set), so a staged read needs no address arithmetic.
__device__ __forceinline__ const float4& lds_f4(const float4* base, uint32_t off)
{
    return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(base) + off);
}

__device__ __forceinline__ float lds_f1(const float* base, uint32_t off)   // base: one float per slot
{
    return base[off >> 4];
}

// Loads an LDS-staged set: wr(k, ld(global id)) for k < count, then a block
// barrier (reached by every thread).  Returns stage_n[b].  All of a thread's
// index loads are issued first, then all its record loads (ld takes any valid
// id: the lanes past the end load the set's last record), then the LDS
// writes, so the loads of one thread are in flight together.
template <uint32_t CAP = kStageMax, class Ld, class Wr>
__device__ __forceinline__ uint32_t stage_block(const NbrLists& L, uint32_t b, uint32_t sn, Ld&& ld, Wr&& wr)
{
    using T = decltype(ld(0u));
    constexpr uint32_t U = (CAP + kBlock - 1) / kBlock;
    if (sn <= CAP && sn > 0u) {
        const uint32_t* g = L.stg + (size_t)b * kStageCap;
        uint32_t gi[U];
        T d[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (u * kBlock < sn) gi[u] = g[min(threadIdx.x + u * kBlock, sn - 1u)];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (u * kBlock < sn) d[u] = ld(gi[u]);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t k = threadIdx.x + u * kBlock;
            if (k < sn) wr(k, d[u]);
        }
    }
    __syncthreads();
    return sn;
}

template <class Ld, class Wr>
__device__ __forceinline__ uint32_t stage_block(const NbrLists& L, uint32_t b, Ld&& ld, Wr&& wr)
{
    return stage_block(L, b, L.stage_n[b], ld, wr);
}

// The list counts and first two groups of particle i's fluid list (16-bit
// when its block's set has them, else 32-bit) and boundary list, loaded
// before the set is staged so that their latency hides behind it.
struct Head {
    uint32_t cnt;
    uint4 f0, f1, b0, b1;
};

__device__ __forceinline__ Head list_head(const NbrLists& L, uint32_t sn, uint32_t cap, uint32_t i)
{
    const uint4* f = (sn != kUnstaged ? reinterpret_cast<const uint4*>(L.nbr16) : reinterpret_cast<const uint4*>(L.nbrf)) + i;
    const uint4* bb = reinterpret_cast<const uint4*>(L.nbrb) + i;
    const uint32_t cn = L.cnt[i];
    const uint32_t nf = cn & 0xffffu, nb = cn >> 16;
    const uint32_t f1 = sn != kUnstaged ? 8u : 4u;   // entries in group 0
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    return Head{cn, nf > 0u ? f[0] : z, nf > f1 ? f[cap] : z, nb > 0u ? bb[0] : z, nb > 4u ? bb[cap] : z};
}

// The fluid neighbors of particle i (block b, sn = stage_n[b]) in whichever
// form its block's lists have: lds(LDS byte offset) loads a neighbor of an
// LDS-staged set, glob(global id) one from global memory; body(data, valid)
// consumes it.
template <uint32_t CAP = kStageMax, class FL, class FG, class Body>
__device__ __forceinline__ void fluid_nbrs(const NbrLists& L, uint32_t cap, uint32_t b, uint32_t i, uint32_t nf,
                                           uint32_t sn, FL&& lds, FG&& glob, Body&& body)
{
    if (sn <= CAP) {
        for_each_nbr16(L.nbr16, cap, i, nf, lds, body);
    } else if (sn != kUnstaged) {
        const uint32_t* g = L.stg + (size_t)b * kStageCap;
        for_each_nbr16(L.nbr16, cap, i, nf, [&](uint32_t e) { return glob(g[e >> 4]); }, body);
    } else {
        for_each_nbr32(L.nbrf, cap, i, nf, i, glob, body);
    }
}

// The same with the list head loaded beforehand (hd = list_head(L, sn, cap, i)).
template <uint32_t CAP = kStageMax, class FL, class FG, class Body>
__device__ __forceinline__ void fluid_nbrs(const NbrLists& L, uint32_t cap, uint32_t b, uint32_t i, uint32_t sn,
                                           const Head& hd, FL&& lds, FG&& glob, Body&& body)
{
    const uint32_t nf = hd.cnt & 0xffffu;
    if (sn <= CAP) {
        for_each_nbr16(L.nbr16, cap, i, nf, hd.f0, hd.f1, lds, body);
    } else if (sn != kUnstaged) {
        const uint32_t* g = L.stg + (size_t)b * kStageCap;
        for_each_nbr16(L.nbr16, cap, i, nf, hd.f0, hd.f1, [&](uint32_t e) { return glob(g[e >> 4]); }, body);
    } else {
        for_each_nbr32(L.nbrf, cap, i, nf, i, hd.f0, hd.f1, glob, body);
    }
}

// x where c holds, else 0 (a select of an already computed value)
__device__ __forceinline__ float keep(bool c, float x)
{
    return c ? x : 0.0f;
}

// ---- wavefront / block reductions (wave64) ----
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ uint32_t wave_max_u(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { const uint32_t w = __shfl_xor(v, o, 64); v = v > w ? v : w; }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum_u(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_incl_scan_u(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t w = __shfl_up(v, o, 64);
        if (lane >= o) v += w;
    }
    return v;
}

// Block-wide exclusive scan (kBlock threads); returns the block total to
// every thread.  s_w: kBlock / 64 words of LDS.  Contains two barriers.
__device__ __forceinline__ uint32_t block_excl_scan_u(uint32_t v, uint32_t* excl, uint32_t* s_w)
{
    const uint32_t incl = wave_incl_scan_u(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) s_w[w] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kBlock / 64; ++k) {
        const uint32_t s = s_w[k];
        off += k < w ? s : 0u;
        tot += s;
    }
    *excl = off + incl - v;
    __syncthreads();
    return tot;
}

// Block sum in a fixed order (wave tree, then waves 0..3): deterministic.
__device__ __forceinline__ double block_sum(double v, double* lds4)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds4[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        for (int k = 0; k < nw; ++k) t += lds4[k];
    }
    return t;   // valid in thread 0
}

// The learned warm start adds the network's correction to the reference warm
// start.  Features 4 and 5 of a particle hold the last solves' increments in
// units free of the time step and of the particle diameter d (kappa h_f^2 /
// d^2, kappaV h / d; kappa h_f^2 scales as d^2 and kappaV h as d for the same
// compression and velocities), from which the reference forms its warm starts
// (DFSPH_core_cuda.cu:240, :3063-3068); the network's targets are the
// increments the next solves still need beyond those warm starts, in the same
// units, so that a model trained at one resolution applies at another.
constexpr int kFeatKappaP = 4, kFeatKappaV = 5;

// The warm start a sweep applies in learned mode, in raw units: the
// reference's (ref, the value of the reference warm start), plus the
// network's correction dy, bounded between min(ref, prev) and max(ref, 0).
// prev is the particle's converged total of the last step in the current raw
// units (clamped at -0.5 like the reference): the whole total for the density
// solver, half of it for the divergence solver (a warm start of the whole
// divergence total set water at rest in motion).  It asks for no more
// pressure (a more negative kappa) than the larger of the reference warm
// start and prev, and for none where the reference asks for none.  A
// stronger start pushes the particles apart, which the solvers' errors, sums
// of compressions only, do not see.  dy = 0 gives ref exactly.
__host__ __device__ __forceinline__ float ws_learned(float ref, float prev, float dy)
{
    return fminf(fmaxf(ref + dy, fminf(ref, prev)), fmaxf(ref, 0.0f));
}

// The learned warm start's inputs to the sweeps: the network's corrections
// and the last step's converged totals (dt-invariant units), one float per
// particle each.  The network works in units free of the particle size too:
// its corrections are scaled by `scale` (the particle diameter d for the
// divergence solver, d^2 for the density solver) before use.
struct WsLearned {
    const float* dy;
    const float* prev;
    float scale;
};

}  // namespace dfsph
