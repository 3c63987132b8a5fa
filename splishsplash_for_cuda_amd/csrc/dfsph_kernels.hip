// dfsph_kernels.hip -- hand-written gfx950 kernels of the DFSPH time step.
//
// One thread per particle, particles in cell-sorted order, kBlock-thread
// blocks (512) with the XCD-aware block remap of dfsph_device.hpp.  Every kernel
// cites the reference kernel(s) it replaces (paths relative to the
// reference repository root).
#include <hip/hip_ext.h>

#include "dfsph_device.hpp"
#include "dfsph_launch.hpp"

namespace dfsph {

// ------------------------------------------------------------------------
// small utilities
// ------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f32(float* p, float v, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Loop-control init (the host-side locals of cuda_divergenceSolve /
// cuda_pressureSolve, DFSPH_core_cuda.cu:785-828, 1338-1378).
// eta = max_error * 0.01 * rho0 (/ h for the divergence solver).
__global__ void k_nbr_ctl_init(DevCtl* c)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->max_seen = 0;
        c->overflow = 0;
        c->nbr_fluid = 0;
        c->nbr_boundary = 0;
        c->nbr_fluid_unstaged = 0;
        c->unstaged_blocks = 0;
    }
}

static __device__ void time_fill(DevTime* t, double h, double hf, double hp)
{
    t->hd = h; t->hfd = hf; t->hpd = hp;
    t->h = (float)h; t->h_future = (float)hf; t->h_past = (float)hp;
    t->inv_h = (float)(1.0 / h);
    t->inv_h_future = (float)(1.0 / hf);
}

// Thread-level: copy the loop and step state into the host record, then
// release its sequence number at system scope.
__device__ void publish(const Publish& pub, const SolveCtl* sc)
{
    HostRecord* r = pub.rec;
    r->sc = *sc;
    r->tm = *pub.t;
    r->max_seen = pub.dc->max_seen;
    r->overflow = pub.dc->overflow;
    r->unstaged_blocks = pub.dc->unstaged_blocks;
    r->escaped = pub.dc->escaped;
    r->nbr_fluid = pub.dc->nbr_fluid;
    r->nbr_boundary = pub.dc->nbr_boundary;
    r->nbr_fluid_unstaged = pub.dc->nbr_fluid_unstaged;
    __threadfence_system();
    __hip_atomic_store(&r->seq, pub.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// DFSPHCData::reset time bookkeeping (DFSPH_c_arrays_structure.cpp:1000-1004)
__global__ void k_time_reset(DevTime* t, double h)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        time_fill(t, h, h, h);
        t->ratio = 1.0f; t->ratio2 = 1.0f;
        t->time = 0.0;
    }
}

// DFSPHCData::updateTimeStep (DFSPH_c_arrays_structure.cpp:1100-1106).  mode 0: the
// given h_future; mode 1: the CFL estimate of cuda_CFL (DFSPH_core_cuda.cu:3380-3387),
// h = cfl_factor * 0.4 * 2r / sqrt(max |v + a h|^2) clamped to [dt_min, dt_max];
// mode 2: keep the current h (constant step, DFSPH_CUDA.cpp:302-306).
__global__ void k_time_update(DevTime* t, double hf, const DevCtl* ctl, int mode,
                              double cfl_factor, double diameter, double dt_min, double dt_max,
                              double cfl_seed)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (mode == 2) hf = t->hd;
        if (mode == 1) {
            double m = (double)ctl->cfl_max;
            m = m > cfl_seed ? m : cfl_seed;
            double h = cfl_factor * 0.4 * (diameter / sqrt(m));
            if (!(h < dt_max)) h = dt_max;
            if (h < dt_min) h = dt_min;
            hf = h;
        }
        const double h = t->hd;
        time_fill(t, h, hf, t->hpd);
        t->ratio = (float)(h / hf);
        t->ratio2 = (float)((h * h) / (hf * hf));
    }
}

// resume from a snapshot: h, h_past, time; h_future = h and the ratios of the
// last updateTimeStep (h_past / h)
__global__ void k_time_set(DevTime* t, double h, double hp, double time)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        time_fill(t, h, h, hp);
        t->ratio = (float)(hp / h);
        t->ratio2 = (float)((hp * hp) / (h * h));
        t->time = time;
    }
}

// DFSPHCData::onSimulationStepEnd (DFSPH_c_arrays_structure.cpp:1108-1118) and the
// TimeManager advance time += h (DFSPH_CUDA.cpp:449-452, with the new h).
__global__ void k_time_step_end(DevTime* t, Publish pub)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const double hp = t->hd, h = t->hfd;
        time_fill(t, h, h, hp);
        t->time += h;
        if (pub.rec) publish(pub, &pub.dc->dens);
    }
}

// ------------------------------------------------------------------------
// exclusive scan of uint32 (replaces cub::DeviceScan::ExclusiveSum,
// DFSPH_core_cuda.cu:1920): tiles of 2048 in LDS, deterministic.
// ------------------------------------------------------------------------
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__global__ __launch_bounds__(kBlock) void k_scan_tiles(uint32_t* in, uint32_t* out, uint32_t n,
                                                        uint32_t* tile_sums, int clear_in)
{
    __shared__ uint32_t tile[kScanTile];
    __shared__ uint32_t wsum[kBlock / 64];
    const uint32_t base = blockIdx.x * kScanTile;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t g = base + k * kBlock + threadIdx.x;
        tile[k * kBlock + threadIdx.x] = g < n ? in[g] : 0u;
        if (clear_in && g < n) in[g] = 0u;     // a histogram left zeroed for its next use
    }
    __syncthreads();
    uint32_t loc[kScanItems];
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        loc[k] = run;
        run += tile[threadIdx.x * kScanItems + k];
    }
    const uint32_t incl = wave_incl_scan_u(run);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (int k = 0; k < w; ++k) woff += wsum[k];
    const uint32_t excl = woff + incl - run;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) tile[threadIdx.x * kScanItems + k] = excl + loc[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t g = base + k * kBlock + threadIdx.x;
        if (g < n) out[g] = tile[k * kBlock + threadIdx.x];
    }
    if (threadIdx.x == kBlock - 1) {
        uint32_t tot = 0;
        for (int k = 0; k < kBlock / 64; ++k) tot += wsum[k];
        tile_sums[blockIdx.x] = tot;
    }
}

// The tile's offset is the sum of the preceding tile sums, which each block
// adds up itself (grids of up to kScanDirect tiles): no scan of the sums.
constexpr uint32_t kScanDirect = 4096;
__global__ __launch_bounds__(kBlock) void k_scan_add_direct(uint32_t* out, uint32_t n, const uint32_t* tile_sums)
{
    __shared__ uint32_t ws[kBlock / 64];
    uint32_t acc = 0;
    for (uint32_t t = threadIdx.x; t < blockIdx.x; t += kBlock) acc += tile_sums[t];
    acc = wave_sum_u(acc);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
    __syncthreads();
    uint32_t off = 0;
#pragma unroll
    for (int k = 0; k < kBlock / 64; ++k) off += ws[k];
    const uint32_t base = blockIdx.x * kScanTile;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t g = base + k * kBlock + threadIdx.x;
        if (g < n) out[g] += off;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t* out, uint32_t n, const uint32_t* tile_off)
{
    const uint32_t off = tile_off[blockIdx.x];
    const uint32_t base = blockIdx.x * kScanTile;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t g = base + k * kBlock + threadIdx.x;
        if (g < n) out[g] += off;
    }
}

// ------------------------------------------------------------------------
// neighbor search (cuda_neighborsSearch, DFSPH_core_cuda.cu:1833-2129, 2219-2531)
// ------------------------------------------------------------------------

// DFSPH_computeGridIdx_kernel (DFSPH_core_cuda.cu:1833-1848): cell id and the
// particle's arrival slot in its cell.  Particles outside the grid are
// clamped into the border cells.
// Consecutive particles mostly share a cell (the previous step left them in
// cell order), so each run of equal keys in a wave takes its slots with one
// atomic: the run's first lane adds the run length and the others offset from
// the returned base.
__global__ __launch_bounds__(kBlock) void k_cell_hist(SimConst c, const float4* pos, uint32_t n, uint32_t* key,
                                                      uint32_t* slot, uint32_t* hist)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < n;
    uint32_t k = kKeyInvalid;
    if (act) {
        const float4 p = pos[i];
        k = cell_id(c, p.x, p.y, p.z);
    }
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t prev = __shfl_up(k, 1, 64);
    const bool head = lane == 0u || prev != k;
    const uint64_t heads = __ballot(head);
    const uint64_t upto = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);   // lanes 0..lane
    const uint32_t h = 63u - (uint32_t)__clzll((long long)(heads & upto));  // my run's first lane
    const uint64_t after = heads & ~upto;
    const uint32_t next = after ? (uint32_t)__ffsll((long long)after) - 1u : 64u;
    uint32_t base = 0;
    if (act && head) base = atomicAdd(hist + k, next - lane);
    base = __shfl(base, (int)h, 64);
    if (act) {
        key[i] = k;
        slot[i] = base + (lane - h);
    }
}

// DFSPH_CountingSortIds_kernel (DFSPH_core_cuda.cu:1851-1858)
__global__ void k_scatter_perm(const uint32_t* key, const uint32_t* slot, const uint32_t* cstart,
                               uint32_t n, uint32_t* perm)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    perm[cstart[key[i]] + slot[i]] = i;
}

// The arrival slots come from atomics, so the order inside a cell varies from
// run to run (DFSPH_core_cuda.cu:1845-1846).  Sorting each cell's few entries
// by original index makes the counting sort stable and the whole step
// bit-reproducible.
__global__ void k_cell_sort(const uint32_t* cstart, uint32_t ncell, uint32_t* perm)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const uint32_t s = cstart[c], e = cstart[c + 1];
    for (uint32_t a = s + 1; a < e; ++a) {
        const uint32_t v = perm[a];
        uint32_t b = a;
        while (b > s && perm[b - 1] > v) { perm[b] = perm[b - 1]; --b; }
        perm[b] = v;
    }
}

// cuda_sortData (DFSPH_core_cuda.cu:2057-2129): one gather of every persistent
// per-particle field into the alternate buffers (no device-to-device copies).
__device__ __forceinline__ void copy_fluid(const FluidPtrs& a, uint32_t from, const FluidPtrs& b, uint32_t to)
{
    b.pos[to] = a.pos[from];
    b.vel[to] = a.vel[from];
    b.kap[to] = a.kap[from];
    b.kapv[to] = a.kapv[from];
    b.ktot[to] = a.ktot[from];
    b.kvtot[to] = a.kvtot[from];
    b.id[to] = a.id[from];
    b.tag[to] = a.tag[from];
}

__global__ void k_gather_fluid(uint32_t n, const uint32_t* perm, FluidPtrs in, FluidPtrs out)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    copy_fluid(in, perm[k], out, k);
}

__global__ void k_gather_f4(uint32_t n, const uint32_t* perm, const float4* in, float4* out)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    out[k] = in[perm[k]];
}

// Appends list entries to a particle's blocked-interleaved list: entries
// shift into a 16-byte register group from the top (a funnel shift per word)
// and each full group is one store.
struct ListWriter16 {
    uint4* out;                    // group 0 of this particle
    uint32_t cap;
    uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
    __device__ void shift_in(uint32_t v)
    {
        w0 = __builtin_amdgcn_alignbit(w1, w0, 16);
        w1 = __builtin_amdgcn_alignbit(w2, w1, 16);
        w2 = __builtin_amdgcn_alignbit(w3, w2, 16);
        w3 = __builtin_amdgcn_alignbit(v, w3, 16);
    }
    __device__ void push(uint32_t n, uint32_t v)      // v = entry n (16 bits)
    {
        shift_in(v);
        if ((n & 7u) == 7u) out[(size_t)(n >> 3) * cap] = make_uint4(w0, w1, w2, w3);
    }
    __device__ void finish(uint32_t n)                // n entries pushed in total
    {
        if (n & 7u) {
            for (uint32_t k = n & 7u; k < 8u; ++k) shift_in(0u);
            out[(size_t)(n >> 3) * cap] = make_uint4(w0, w1, w2, w3);
        }
    }
};

struct ListWriter32 {
    uint4* out;
    uint32_t cap;
    uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
    __device__ void shift_in(uint32_t v)
    {
        w0 = w1;
        w1 = w2;
        w2 = w3;
        w3 = v;
    }
    __device__ void push(uint32_t n, uint32_t v)
    {
        shift_in(v);
        if ((n & 3u) == 3u) out[(size_t)(n >> 2) * cap] = make_uint4(w0, w1, w2, w3);
    }
    __device__ void finish(uint32_t n)
    {
        if (n & 3u) {
            for (uint32_t k = n & 3u; k < 4u; ++k) shift_in(0u);
            out[(size_t)(n >> 2) * cap] = make_uint4(w0, w1, w2, w3);
        }
    }
};

// Neighbor-search tables.  A block's set cells are the grid cells within one
// cell of its own cells; they are indexed through a dense table over the
// bounding box of those cells, x fastest, then z, then y, so the set lists
// them in (y, z, x) order and the three cells x-1..x+1 of any row (y, z) are
// consecutive: one contiguous range of the set per row.
// LDS of the search with 512-particle blocks: 2672 staged positions and 512
// set cells take 52 KB, three blocks per CU.  The search is bound by the
// latency of its per-block phases (set construction, staging, candidate
// tests between barriers), so the third block gains more than the larger
// stage of two blocks per CU would (C4: 0.62 against 0.76 ms per search);
// larger sets search through global memory.
#ifndef DFSPH_NBR_STAGE
#define DFSPH_NBR_STAGE (DFSPH_BLOCK == 512 ? 2672 : 2048)
#endif
constexpr uint32_t kNbrStage = DFSPH_NBR_STAGE;   // set positions staged in LDS by the search
constexpr uint32_t kDense = 8192;          // bounding-box cells of a tabled set
#ifndef DFSPH_SET_CELLS
#define DFSPH_SET_CELLS (DFSPH_BLOCK == 512 ? 512 : 384)
#endif
constexpr uint32_t kSetCells = DFSPH_SET_CELLS;   // set cells of a tabled set
constexpr uint32_t kNoRow = 0xffffffffu;
constexpr int kIntMax = 0x7fffffff;

__device__ __forceinline__ int wave_min_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ int wave_max_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// DFSPH_neighborsSearch_kernel<true> (DFSPH_core_cuda.cu:2219-2531,
// DFSPH_macro_cuda.h:136-155): every fluid particle within h (strict, FP32) of
// particle i except i itself, over the 27 cells around its cell, then the
// boundary particles the same way, at most max_neighbors in total (fluid
// first).  Cells are visited in for_each_adjacent_cell's order (y, then z,
// then x) and each cell's particles in index order, in every list form.
//
// Each block also builds its neighbor set: the set cells are marked in the
// dense bounding-box table, numbered by a block scan, and get their fluid
// ranges, set offsets and boundary ranges.  Every lane then keeps its 9 rows
// as (first, last) set cells in registers.  Sets up to kNbrStage have their
// positions staged in LDS for the candidate tests (over the table, which is
// dead by then); sets up to kStageCap store 16-bit lists; larger ones, blocks
// whose tables do not fit, and the INDIRECT search (fluid not reordered, the
// grid indexes it through perm: the need_sort == false path) store 32-bit
// global lists.
template <bool INDIRECT>
__global__ __launch_bounds__(kBlock) void k_neighbors(SimConst c, const float4* __restrict__ pos,
                                                      const uint32_t* __restrict__ cs_f,
                                                      const uint32_t* __restrict__ perm,
                                                      const float4* __restrict__ bpos,
                                                      const uint32_t* __restrict__ cs_b, uint32_t n,
                                                      uint16_t* __restrict__ nbr16, uint32_t* __restrict__ nbrf,
                                                      uint32_t* __restrict__ nbrb, uint32_t* __restrict__ cnt,
                                                      uint32_t* __restrict__ stage_n, uint32_t* __restrict__ stg,
                                                      Restrict rs, uint4* __restrict__ part)
{
    // the dense table shares one area with the staged positions
    __shared__ __align__(16) unsigned char s_raw[kNbrStage * sizeof(float4)];
    __shared__ uint32_t s_off[kSetCells + 1];  // set offsets of the set cells
    __shared__ uint32_t s_src[kSetCells];      // first fluid index of each set cell
    __shared__ uint32_t s_bs[kSetCells];       // boundary range of each set cell
    __shared__ uint32_t s_be[kSetCells];
    __shared__ uint32_t s_boff[kSetCells + 1]; // boundary set offsets of the set cells
    constexpr int NW = kBlock / 64;
    __shared__ int s_w[6 * NW];
    static_assert(kDense * sizeof(uint16_t) <= sizeof(s_raw), "dense table exceeds the stage");
    static_assert((kStageCap - 1) * 16 <= 0xffffu && kSetCells < 0xffffu, "16-bit byte offsets / cell ids");
    uint16_t* s_dense = reinterpret_cast<uint16_t*>(s_raw);
    float4* s_pos = reinterpret_cast<float4*>(s_raw);
    uint32_t* s_wu = reinterpret_cast<uint32_t*>(s_w);

    const uint32_t b = xcd_block();
    const uint32_t i = b * kBlock + threadIdx.x;
    const bool act = i < n;
    float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    int cx = 0, cy = 0, cz = 0;
    if (act) {
        p = pos[i];
        cx = cell_coord(c, p.x, 0);
        cy = cell_coord(c, p.y, 1);
        cz = cell_coord(c, p.z, 2);
    }
    uint32_t S = kUnstaged, U = 0, uown = 0, Sb = 0;
    uint32_t row[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) row[r] = kNoRow;
    bool tabled = false;                        // block uniform: row / s_off / s_src / s_b* valid
    if (!INDIRECT && n <= b * kBlock) S = 0;    // a block past the end: empty set
    if (!INDIRECT && n > b * kBlock) {
        // bounding box of the set cells, clamped to the grid
        const int w = threadIdx.x >> 6;
        const int lo_x = wave_min_i(act ? cx : kIntMax), hi_x = wave_max_i(act ? cx : -1);
        const int lo_y = wave_min_i(act ? cy : kIntMax), hi_y = wave_max_i(act ? cy : -1);
        const int lo_z = wave_min_i(act ? cz : kIntMax), hi_z = wave_max_i(act ? cz : -1);
        if ((threadIdx.x & 63) == 0) {
            s_w[w] = lo_x; s_w[NW + w] = lo_y; s_w[2 * NW + w] = lo_z;
            s_w[3 * NW + w] = hi_x; s_w[4 * NW + w] = hi_y; s_w[5 * NW + w] = hi_z;
        }
        __syncthreads();
        int lo[3] = {kIntMax, kIntMax, kIntMax}, hi[3] = {-1, -1, -1};
#pragma unroll
        for (int k = 0; k < NW; ++k) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = min(lo[a], s_w[NW * a + k]);
                hi[a] = max(hi[a], s_w[NW * (3 + a) + k]);
            }
        }
        __syncthreads();                        // s_w is reused by the scans
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = max(lo[a] - 1, 0);
            hi[a] = min(hi[a] + 1, c.dim[a] - 1);
        }
        const uint32_t W = (uint32_t)(hi[0] - lo[0] + 1), D = (uint32_t)(hi[2] - lo[2] + 1);
        const uint32_t V = W * D * (uint32_t)(hi[1] - lo[1] + 1);
        if (V <= kDense) {
            for (uint32_t k = threadIdx.x; k < V; k += kBlock) s_dense[k] = 0;
            __syncthreads();
            const int xl = max(cx - 1, 0), xh = min(cx + 1, c.dim[0] - 1);
            const uint32_t key = act ? cell_key(c, cx, cy, cz) : kKeyInvalid;
            const bool first = act && ((threadIdx.x & 63) == 0 || __shfl_up(key, 1, 64) != key);
            if (first) {                        // mark the 27 cells (one lane per run of a cell in a wave)
                for (int dy = -1; dy <= 1; ++dy) {
                    const int y = cy + dy;
                    if (y < 0 || y >= c.dim[1]) continue;
                    for (int dz = -1; dz <= 1; ++dz) {
                        const int z = cz + dz;
                        if (z < 0 || z >= c.dim[2]) continue;
                        const uint32_t base = W * ((uint32_t)(z - lo[2]) + D * (uint32_t)(y - lo[1])) - lo[0];
                        for (int x = xl; x <= xh; ++x) s_dense[base + x] = 1;
                    }
                }
            }
            __syncthreads();
            // number the marked cells in table order; thread t owns entries [tE, tE + E)
            const uint32_t E = (V + kBlock - 1) / kBlock;
            const uint32_t k0 = threadIdx.x * E, k1 = min(k0 + E, V);
            uint32_t occ = 0;
            for (uint32_t k = k0; k < k1; ++k) occ += s_dense[k];
            uint32_t ubase = 0;
            U = block_excl_scan_u(occ, &ubase, s_wu);
            tabled = U <= kSetCells;
            if (tabled) {
                for (uint32_t k = k0; k < k1; ++k) {
                    if (s_dense[k]) {
                        const uint32_t u = ubase++;
                        const uint32_t t = k / W;
                        const uint32_t key = cell_key(c, lo[0] + (int)(k - t * W), lo[1] + (int)(t / D),
                                                      lo[2] + (int)(t % D));
                        const uint32_t a = cs_f[key];
                        s_src[u] = a;
                        s_off[u] = cs_f[key + 1] - a;
                        s_bs[u] = cs_b[key];
                        s_be[u] = cs_b[key + 1];
                        s_dense[k] = (uint16_t)u;
                    }
                }
            }
            __syncthreads();
            if (tabled) {
                // fluid counts of the set cells -> set offsets
                const uint32_t E2 = (U + kBlock - 1) / kBlock;
                const uint32_t u0 = threadIdx.x * E2, u1 = min(u0 + E2, U);
                uint32_t sum = 0;
                for (uint32_t u = u0; u < u1; ++u) sum += s_off[u];
                uint32_t off = 0;
                const uint32_t tot = block_excl_scan_u(sum, &off, s_wu);
                for (uint32_t u = u0; u < u1; ++u) {
                    const uint32_t cu = s_off[u];
                    s_off[u] = off;
                    off += cu;
                }
                if (threadIdx.x == 0) s_off[U] = tot;
                S = tot <= c.stage_cap ? tot : kUnstaged;
                // boundary counts -> boundary set offsets (the set's boundary
                // particles follow its fluid ones when both fit in the stage)
                uint32_t bsum = 0;
                for (uint32_t u = u0; u < u1; ++u) bsum += s_be[u] - s_bs[u];
                uint32_t boff = 0;
                Sb = block_excl_scan_u(bsum, &boff, s_wu);
                for (uint32_t u = u0; u < u1; ++u) {
                    s_boff[u] = boff;
                    boff += s_be[u] - s_bs[u];
                }
                if (threadIdx.x == 0) s_boff[U] = Sb;
                // each lane's rows (dy, dz) as (first, last) set cell of x-1..x+1
                if (act) {
                    uown = s_dense[W * ((uint32_t)(cz - lo[2]) + D * (uint32_t)(cy - lo[1])) + (uint32_t)(cx - lo[0])];
#pragma unroll
                    for (int r = 0; r < 9; ++r) {
                        const int y = cy + r / 3 - 1, z = cz + r % 3 - 1;
                        if (y >= 0 && y < c.dim[1] && z >= 0 && z < c.dim[2]) {
                            const uint32_t base = W * ((uint32_t)(z - lo[2]) + D * (uint32_t)(y - lo[1])) - lo[0];
                            row[r] = (uint32_t)s_dense[base + xl] | ((uint32_t)s_dense[base + xh] << 16);
                        }
                    }
                }
                __syncthreads();                // the table area is staged over next
            }
        }
    }
    const bool lds_search = S <= kNbrStage && c.stage_lds > 0u;
    const bool lds_bnd = lds_search && S + Sb <= kNbrStage;   // boundary staged at [S, S + Sb)
    if (threadIdx.x == 0) stage_n[b] = S == kUnstaged || S <= c.stage_lds ? S : (S | kSetNoLds);
    if (S != kUnstaged) {
        uint32_t* g_out = stg + (size_t)b * kStageCap;
        for (uint32_t k = threadIdx.x; k < S; k += kBlock) {
            uint32_t lo = 0, hi = U;            // the set cell holding set index k
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_off[mid] <= k) lo = mid; else hi = mid;
            }
            const uint32_t g = s_src[lo] + (k - s_off[lo]);
            if (lds_search) s_pos[k] = pos[g];
            g_out[k] = g;
        }
        if (lds_bnd) {
            for (uint32_t k = threadIdx.x; k < Sb; k += kBlock) {
                uint32_t lo = 0, hi = U;        // the set cell holding boundary set index k
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_boff[mid] <= k) lo = mid; else hi = mid;
                }
                const uint32_t j = s_bs[lo] + (k - s_boff[lo]);
                const float4 q = bpos[j];
                s_pos[S + k] = make_float4(q.x, q.y, q.z, __uint_as_float(j));   // w: the global boundary id
            }
        }
        __syncthreads();
    }
    uint32_t total = 0, trunc = 0, stored_f = 0, stored_b = 0, stored_u32 = 0;
    if (act && !is_source(rs, i)) cnt[i] = 0;           // restricted out: no lists
    if (act && is_source(rs, i)) {
        const uint32_t maxn = c.max_nbr, cap = c.cap;
        const float hk2 = c.hk2;
        uint32_t nf = 0, nb = 0;
        if (S != kUnstaged) {
            ListWriter16 w{reinterpret_cast<uint4*>(nbr16) + i, cap};
            const uint32_t self = s_off[uown] + (i - s_src[uown]);
            if (lds_search) {
                // Each row range goes in chunks of 32 candidates: a branch-free
                // pass sets one bit per candidate within h, then the set bits
                // are appended in order (stored while below max_neighbors).
                uint32_t ntot = 0;
#pragma unroll
                for (int r = 0; r < 9; ++r) {
                    if (row[r] == kNoRow) continue;
                    const uint32_t e = s_off[(row[r] >> 16) + 1u];
                    for (uint32_t l0 = s_off[row[r] & 0xffffu]; l0 < e; l0 += 32u) {
                        const uint32_t cnt = min(e - l0, 32u);
                        uint32_t m = 0, t = 0;
                        for (; t + 2u <= cnt; t += 2u) {
                            const float4 qa = s_pos[l0 + t], qb = s_pos[l0 + t + 1u];
                            const uint32_t ia = dist2(p.x - qa.x, p.y - qa.y, p.z - qa.z) < hk2 ? 1u : 0u;
                            const uint32_t ib = dist2(p.x - qb.x, p.y - qb.y, p.z - qb.z) < hk2 ? 2u : 0u;
                            m |= (ia | ib) << t;
                        }
                        if (t < cnt) {
                            const float4 qa = s_pos[l0 + t];
                            m |= (dist2(p.x - qa.x, p.y - qa.y, p.z - qa.z) < hk2 ? 1u : 0u) << t;
                        }
                        if (self - l0 < cnt) m &= ~(1u << (self - l0));
                        ntot += (uint32_t)__popc(m);
                        while (m != 0u && nf < maxn) {
                            w.push(nf, (l0 + (uint32_t)__builtin_ctz(m)) << 4);   // LDS byte offset
                            ++nf;
                            m &= m - 1u;
                        }
                    }
                }
                nf = ntot;                      // stored: min(nf, max_neighbors), as below
            } else {
#pragma unroll
                for (int r = 0; r < 9; ++r) {
                    if (row[r] == kNoRow) continue;
                    for (uint32_t u = row[r] & 0xffffu; u <= (row[r] >> 16); ++u) {
                        const uint32_t e = s_off[u + 1u], src = s_src[u] - s_off[u];
                        for (uint32_t l = s_off[u]; l < e; ++l) {
                            const float4 qp = pos[src + l];
                            const bool in = dist2(p.x - qp.x, p.y - qp.y, p.z - qp.z) < hk2 && l != self;
                            if (in) {
                                if (nf < maxn) w.push(nf, l << 4);
                                ++nf;
                            }
                        }
                    }
                }
            }
            w.finish(nf < maxn ? nf : maxn);
        } else {
            ListWriter32 w{reinterpret_cast<uint4*>(nbrf) + i, cap};
            auto scan_range = [&](uint32_t s0, uint32_t e) {
                for (uint32_t q = s0; q < e; ++q) {
                    const uint32_t j = INDIRECT ? perm[q] : q;
                    const float4 pj = pos[j];
                    if (dist2(p.x - pj.x, p.y - pj.y, p.z - pj.z) < hk2 && j != i) {
                        if (nf < maxn) w.push(nf, j);
                        ++nf;
                    }
                }
            };
            if (tabled) {
#pragma unroll
                for (int r = 0; r < 9; ++r) {
                    if (row[r] == kNoRow) continue;
                    for (uint32_t u = row[r] & 0xffffu; u <= (row[r] >> 16); ++u)
                        scan_range(s_src[u], s_src[u] + (s_off[u + 1u] - s_off[u]));
                }
            } else {
                for_each_adjacent_cell(c, cx, cy, cz, [&](uint32_t nk) { scan_range(cs_f[nk], cs_f[nk + 1]); });
            }
            w.finish(nf < maxn ? nf : maxn);
        }
        const uint32_t sf = nf < maxn ? nf : maxn;
        ListWriter32 wb{reinterpret_cast<uint4*>(nbrb) + i, cap};
        auto scan_boundary = [&](uint32_t s0, uint32_t e) {
            for (uint32_t j = s0; j < e; ++j) {
                const float4 pj = bpos[j];
                if (dist2(p.x - pj.x, p.y - pj.y, p.z - pj.z) < hk2) {
                    if (sf + nb < maxn) wb.push(nb, j);
                    ++nb;
                }
            }
        };
        if (lds_bnd) {
            // the row's boundary particles are one contiguous staged range;
            // the same chunked bitmask pass as the fluid search
            uint32_t ntot = 0;
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                if (row[r] == kNoRow) continue;
                const uint32_t e = S + s_boff[(row[r] >> 16) + 1u];
                for (uint32_t l0 = S + s_boff[row[r] & 0xffffu]; l0 < e; l0 += 32u) {
                    const uint32_t cntc = min(e - l0, 32u);
                    uint32_t m = 0, t = 0;
                    for (; t + 2u <= cntc; t += 2u) {
                        const float4 qa = s_pos[l0 + t], qb = s_pos[l0 + t + 1u];
                        const uint32_t ia = dist2(p.x - qa.x, p.y - qa.y, p.z - qa.z) < hk2 ? 1u : 0u;
                        const uint32_t ib = dist2(p.x - qb.x, p.y - qb.y, p.z - qb.z) < hk2 ? 2u : 0u;
                        m |= (ia | ib) << t;
                    }
                    if (t < cntc) {
                        const float4 qa = s_pos[l0 + t];
                        m |= (dist2(p.x - qa.x, p.y - qa.y, p.z - qa.z) < hk2 ? 1u : 0u) << t;
                    }
                    ntot += (uint32_t)__popc(m);
                    while (m != 0u && sf + nb < maxn) {
                        wb.push(nb, __float_as_uint(s_pos[l0 + (uint32_t)__builtin_ctz(m)].w));
                        ++nb;
                        m &= m - 1u;
                    }
                }
            }
            nb = ntot;
        } else if (tabled) {
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                if (row[r] == kNoRow) continue;
                for (uint32_t u = row[r] & 0xffffu; u <= (row[r] >> 16); ++u) scan_boundary(s_bs[u], s_be[u]);
            }
        } else {
            for_each_adjacent_cell(c, cx, cy, cz, [&](uint32_t nk) { scan_boundary(cs_b[nk], cs_b[nk + 1]); });
        }
        const uint32_t sb = nb < maxn - sf ? nb : maxn - sf;
        wb.finish(sb);
        total = nf + nb;
        trunc = total > maxn ? 1u : 0u;
        cnt[i] = sf | (sb << 16);
        stored_f = sf;
        stored_b = sb;
        stored_u32 = S == kUnstaged ? sf : 0u;
    }
    // the block's statistics go to part[2b], part[2b + 1] (k_nbr_finalize
    // sums them): per-wave atomics on one control line would serialise
    const uint32_t wv[5] = {wave_max_u(total), wave_sum_u(trunc), wave_sum_u(stored_f), wave_sum_u(stored_b),
                            wave_sum_u(stored_u32)};
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 5; ++k) s_w[5 * w + k] = (int)wv[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a[5] = {0u, 0u, 0u, 0u, 0u};
        for (int v = 0; v < NW; ++v)
            for (int k = 0; k < 5; ++k) {
                const uint32_t x = (uint32_t)s_w[5 * v + k];
                a[k] = k == 0 ? max(a[0], x) : a[k] + x;
            }
        const uint32_t unstaged = S > c.stage_lds && n > b * kBlock ? 1u : 0u;
        part[2 * b] = make_uint4(a[0], a[1], a[2], a[3]);
        part[2 * b + 1] = make_uint4(a[4], unstaged, 0u, 0u);
    }
}

// The neighbor statistics of the last search from the block records of
// k_neighbors (maximum, sums in a fixed order).
__global__ __launch_bounds__(kBlock) void k_nbr_finalize(const uint4* __restrict__ part, uint32_t nblocks, DevCtl* ctl)
{
    __shared__ unsigned long long s[6][kBlock];
    unsigned long long a[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < nblocks; b += kBlock) {
        const uint4 p = part[2 * b], q = part[2 * b + 1];
        a[0] = a[0] > p.x ? a[0] : p.x;
        a[1] += p.y; a[2] += p.z; a[3] += p.w; a[4] += q.x; a[5] += q.y;
    }
    for (int k = 0; k < 6; ++k) s[k][threadIdx.x] = a[k];
    __syncthreads();
    for (uint32_t h = kBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int k = 0; k < 6; ++k) {
                const unsigned long long o = s[k][threadIdx.x + h];
                s[k][threadIdx.x] = k == 0 ? (s[0][threadIdx.x] > o ? s[0][threadIdx.x] : o) : s[k][threadIdx.x] + o;
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ctl->max_seen = (uint32_t)s[0][0];
        ctl->overflow = (uint32_t)s[1][0];
        ctl->nbr_fluid = s[2][0];
        ctl->nbr_boundary = s[3][0];
        ctl->nbr_fluid_unstaged = s[4][0];
        ctl->unstaged_blocks = (uint32_t)s[5][0];
    }
}

// Every list of the last search as 32-bit global ids, fluid then boundary
// (the reference's plain interleaved layout, slot k at k * cap + i), for readback.
__global__ __launch_bounds__(kBlock) void k_expand_lists(uint32_t n, uint32_t cap, NbrLists L, uint32_t* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = i / kBlock;
    const uint32_t cn = L.cnt[i];
    const uint32_t nf = cn & 0xffffu, nb = cn >> 16;
    const uint32_t sn = L.stage_n[b];
    for (uint32_t k = 0; k < nf; ++k) {
        uint32_t j;
        if (sn != kUnstaged) {
            const uint32_t l = L.nbr16[(((size_t)(k >> 3) * cap + i) << 3) + (k & 7u)] >> 4;
            j = L.stg[(size_t)b * kStageCap + l];
        } else {
            j = L.nbrf[(((size_t)(k >> 2) * cap + i) << 2) + (k & 3u)];
        }
        out[(size_t)k * cap + i] = j;
    }
    for (uint32_t k = 0; k < nb; ++k)
        out[(size_t)(nf + k) * cap + i] = L.nbrb[(((size_t)(k >> 2) * cap + i) << 2) + (k & 3u)];
}

// FluidModel::computeBoundaryPsi (FluidModel.cpp:233-258) on the device:
// psi_i = rho0 / (W0 + sum_{j != i, |xij| < h} W(xij)) over the boundary set.
__global__ void k_boundary_psi(SimConst c, const float4* bpos, const uint32_t* cs_b, uint32_t nb, float* psi)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const float4 p = bpos[i];
    const int cx = cell_coord(c, p.x, 0), cy = cell_coord(c, p.y, 1), cz = cell_coord(c, p.z, 2);
    float delta = c.W0;
    for_each_adjacent_cell(c, cx, cy, cz, [&](uint32_t nk) {
        for (uint32_t j = cs_b[nk]; j < cs_b[nk + 1]; ++j) {
            if (j == i) continue;
            const float4 q = bpos[j];
            const float r2 = dist2(p.x - q.x, p.y - q.y, p.z - q.z);
            if (r2 < c.hk2) delta += kernel_W(c, r2);
        }
    });
    psi[i] = c.density0 / delta;
}

__global__ void k_set_w(float4* p, const float* w, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i].w = w[i];
}

// ------------------------------------------------------------------------
// solver sweeps
// ------------------------------------------------------------------------

// neighbor records loaded in the first phase of a batch
struct PV {
    float4 p, v;                   // position + mass, velocity (+ a per-kernel scalar in w)
};

// LDS stage of (position, velocity) records.  UM (every fluid particle has the
// mass c.m_uniform): 24 B per particle, (x y z vx) and (vy vz), so that two
// 512-particle blocks fit a CU (78 KB each) instead of one; the mass is put
// back from the constant, the same value bit for bit.  Entries of the 16-bit
// lists are byte offsets of 16-byte slots; the 8-byte slots are at half.
template <bool UM>
struct PVStage {
    float4 p[kStageMaxPV], v[kStageMaxPV];
    __device__ __forceinline__ void put(uint32_t k, const PV& d) { p[k] = d.p; v[k] = d.v; }
    __device__ __forceinline__ PV get(uint32_t off, float) const { return PV{lds_f4(p, off), lds_f4(v, off)}; }
};
template <>
struct PVStage<true> {
    float4 a[kStageMaxPV];
    float2 b[kStageMaxPV];
    __device__ __forceinline__ void put(uint32_t k, const PV& d)
    {
        a[k] = make_float4(d.p.x, d.p.y, d.p.z, d.v.x);
        b[k] = make_float2(d.v.y, d.v.z);
    }
    __device__ __forceinline__ PV get(uint32_t off, float m) const
    {
        const float4 u = lds_f4(a, off);
        const float2 w = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(b) + (off >> 1));
        return PV{make_float4(u.x, u.y, u.z, m), make_float4(u.w, w.x, w.y, 0.0f)};
    }
};
// LDS stage of (position, velocity, scalar) records.  UM: 28 B per particle,
// (x y z vx), (vy vz) and the scalar, the mass put back from the constant;
// otherwise 32 B, (x y z m) and (vx vy vz scalar).
template <bool UM>
struct PVSStage {
    float4 p[kStageMaxX], v[kStageMaxX];
    __device__ __forceinline__ void put(uint32_t k, const PV& d) { p[k] = d.p; v[k] = d.v; }
    __device__ __forceinline__ PV get(uint32_t off, float) const { return PV{lds_f4(p, off), lds_f4(v, off)}; }
};
template <>
struct PVSStage<true> {
    float4 a[kStageMaxX];
    float2 b[kStageMaxX];
    float s[kStageMaxX];
    __device__ __forceinline__ void put(uint32_t k, const PV& d)
    {
        a[k] = make_float4(d.p.x, d.p.y, d.p.z, d.v.x);
        b[k] = make_float2(d.v.y, d.v.z);
        s[k] = d.v.w;
    }
    __device__ __forceinline__ PV get(uint32_t off, float m) const
    {
        const float4 u = lds_f4(a, off);
        const float2 w = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(b) + (off >> 1));
        const float q = lds_f1(s, off);
        return PV{make_float4(u.x, u.y, u.z, m), make_float4(u.w, w.x, w.y, q)};
    }
};

struct PK {
    float4 p;                      // position + mass
    float k;                       // k_j of the Jacobi sweep / the warm start
};

// LDS stage of (position, k) records: 20 B per particle, or 16 B (x y z k)
// with a uniform fluid mass (UM; three 512-particle blocks per CU instead of
// two).
template <bool UM>
struct PKStage {
    float4 p[kStageMax];
    float k[kStageMax];
    __device__ __forceinline__ void put(uint32_t i, const PK& d) { p[i] = d.p; k[i] = d.k; }
    __device__ __forceinline__ PK get(uint32_t off, float) const { return PK{lds_f4(p, off), lds_f1(k, off)}; }
};
template <>
struct PKStage<true> {
    float4 a[kStageMax];
    __device__ __forceinline__ void put(uint32_t i, const PK& d) { a[i] = make_float4(d.p.x, d.p.y, d.p.z, d.k); }
    __device__ __forceinline__ PK get(uint32_t off, float m) const
    {
        const float4 u = lds_f4(a, off);
        return PK{make_float4(u.x, u.y, u.z, m), u.w};
    }
};

// Reducing sweeps write one FP64 partial per block with a plain store; the
// separate one-block kernel below sums them.  (A "last block finishes"
// reduction inside the sweep needs a device-scope release/acquire per block,
// which on MI355X writes back and invalidates the XCD's L2 every time and
// throws away the neighbor data the other blocks are gathering.)
__device__ __forceinline__ void store_partial(double* partial, double bsum)
{
    if (threadIdx.x == 0) partial[blockIdx.x] = bsum;
}

// Sums the block partials in a fixed order (deterministic), forms
// err = sum / N and evaluates the loop condition of cuda_divergenceSolve /
// cuda_pressureSolve: continue while ((err > eta) || (it < min)) && (it < max).
__global__ __launch_bounds__(kBlock) void k_solve_finalize(SolveCtl* ctl, const double* partial, uint32_t nparts,
                                                          Publish pub)
{
    __shared__ double lds4[kBlock / 64];
    if (!ctl->done) {
        // each thread sums its strided share in order; the loads of 8 terms
        // are issued together (adding the 0.0 of a missing term is exact, so
        // the sum is the one-at-a-time sum, bit for bit)
        constexpr uint32_t kBatch = 8;
        double acc = 0.0;
        for (uint32_t k0 = threadIdx.x; k0 < nparts; k0 += kBatch * kBlock) {
            double v[kBatch];
#pragma unroll
            for (uint32_t u = 0; u < kBatch; ++u) {
                const uint32_t k = k0 + u * kBlock;
                v[u] = k < nparts ? partial[k] : 0.0;
            }
#pragma unroll
            for (uint32_t u = 0; u < kBatch; ++u) acc += v[u];
        }
        const double tot = block_sum(acc, lds4);
        if (threadIdx.x == 0) {
            const double err = ctl->n_div > 0.0 ? tot / ctl->n_div : 0.0;
            const uint32_t it = ctl->iters + 1;
            ctl->sum = tot;
            ctl->err = err;
            ctl->iters = it;
            const bool cont = ((err > ctl->eta) || (it < ctl->min_iter)) && (it < ctl->max_iter);
            ctl->done = cont ? 0u : 1u;
        }
    }
    if (threadIdx.x == 0 && pub.rec) publish(pub, ctl);
}

// Fused density + factor + divergence source, then the divergence warm start:
// DFSPH_divergence_warmstart_init_kernel (DFSPH_core_cuda.cu:220-376) followed by
// DFSPH_divergence_compute_kernel<true> (:495-517, divergenceSolveParticle :400-481).
// rho_i = m_i W0 + sum m_j W_ij; alpha_i = -(1/h) / max(sum |m_j gradW|^2 + |sum m_j gradW|^2, eps);
// rho_adv_i = 0 if fewer than 20 neighbors else max(sum m_j (v_i - v_j).gradW, 0).
// Warm start (only where rho_adv_i > 0): v_i += h sum (k_i + k_j) m_j gradW with
// k = max(kappaV h_ratio / 2, -0.5) (WS == 0), that value plus the learned
// correction, bounded by ws_learned with half the last converged total
// (WS == 1), or nothing (WS == 2).  kws holds
// the last solve's kappaV increments.  Velocities are read from vel_i and
// written to vel_o.  The applied k_i starts the particle's total of the step
// in kvtot (when given).
template <int WS, bool SOL, bool UM>
__global__ __launch_bounds__(kBlock, UM ? DFSPH_PV_WAVES : 1) void k_div_ws_init(SimConst c, const DevTime* t, uint32_t n,
                                                        const float4* __restrict__ pos,
                                                        const float4* __restrict__ bpos, NbrLists L,
                                                        const float4* __restrict__ vel_i,
                                                        float4* __restrict__ vel_o,
                                                        const float* __restrict__ kws, WsLearned lw,
                                                        float* __restrict__ rho, float* __restrict__ alpha,
                                                        float* __restrict__ radv, float* __restrict__ kvtot)
{
    __shared__ PVSStage<UM> st;                  // (x, v, warm-start k) of the staged set
    const float ratio = t->ratio, h = t->h, inv_h = t->inv_h;
    auto ws_of = [&](uint32_t j) -> float {
        if (WS == 2) return 0.0f;
        const float ref = fmaxf(kws[j] * ratio / 2.0f, -0.5f);
        if (WS == 0) return ref;
        return ws_learned(ref, fmaxf(0.5f * lw.prev[j] * inv_h, -0.5f), lw.dy[j] * lw.scale * inv_h);
    };
    const uint32_t b = xcd_block();
    const uint32_t i = b * kBlock + threadIdx.x;
    const uint32_t ic = i < n ? i : 0u;         // own loads before staging, so they overlap it
    const uint32_t sn0 = L.stage_n[b];
    const Head hd = list_head(L, sn0, c.cap, ic);
    const float4 vi = vel_i[ic];
    const float4 pi = pos[ic];
    const float ki = ws_of(ic);
    const uint32_t sn = stage_block<kStageMaxX>(L, b, sn0, [&](uint32_t g) {
        const float4 v = vel_i[g];
        return PV{pos[g], make_float4(v.x, v.y, v.z, ws_of(g))};
    }, [&](uint32_t k, const PV& d) { st.put(k, d); });
    if (i >= n) return;
    if (kvtot) kvtot[i] = ki;
    if (!is_source(L.rs, i)) { vel_o[i] = vi; return; }
    const uint32_t cn = hd.cnt;
    const uint32_t nf = cn & 0xffffu, nb = cn >> 16;
    const uint32_t stride = c.cap;
    float dens = pi.w * c.W0, sg2 = 0.0f, dr = 0.0f;
    float gx = 0.0f, gy = 0.0f, gz = 0.0f, ax = 0.0f, ay = 0.0f, az = 0.0f;
    auto fluid = [&](const float4& pj, const float4& vj, bool valid) {
        const float m = keep(valid, pj.w);
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const PairL kp = pair_listed(c, dist2(dx, dy, dz));
        dens += m * kp.W;
        const float mg = keep(kp.ok, m * kp.g);
        const float px = mg * dx, py = mg * dy, pz = mg * dz;
        sg2 += px * px + py * py + pz * pz;
        gx += px; gy += py; gz += pz;
        dr += (vi.x - vj.x) * px + (vi.y - vj.y) * py + (vi.z - vj.z) * pz;
        if (WS != 2) {
            const float ks = ki + vj.w;
            const float s = keep(fabsf(ks) > kEps, ks);
            ax += s * px; ay += s * py; az += s * pz;
        }
    };
    fluid_nbrs<kStageMaxX>(L, stride, b, i, sn, hd, [&](uint32_t e) { return st.get(e, c.m_uniform); },
               [&](uint32_t j) {
                   const float4 v = vel_i[j];
                   return PV{pos[j], make_float4(v.x, v.y, v.z, ws_of(j))};
               },
               [&](const PV& d, bool valid) { fluid(d.p, d.v, valid); });
    (void)nf;
    float bx = 0.0f, by = 0.0f, bz = 0.0f;
    for_each_nbr32(L.nbrb, stride, i, nb, 0u, hd.b0, hd.b1, [&](uint32_t j) { return load_bnd<SOL>(bpos, L.sol, j); },
        [&](const BN& bj, bool valid) {
        const float4& pj = bj.p;
        const float m = keep(valid, pj.w);
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const PairL kp = pair_listed(c, dist2(dx, dy, dz));
        dens += m * kp.W;
        const float mg = keep(kp.ok, m * kp.g);
        const float px = mg * dx, py = mg * dy, pz = mg * dz;
        sg2 += px * px + py * py + pz * pz;
        gx += px; gy += py; gz += pz;
        dr += (vi.x - bj.v.x) * px + (vi.y - bj.v.y) * py + (vi.z - bj.v.z) * pz;
        bx += px; by += py; bz += pz;
    });
    sg2 += gx * gx + gy * gy + gz * gz;
    const float a = -inv_h / fmaxf(sg2, kEps);
    const float r = (nf + nb < 20u) ? 0.0f : fmaxf(dr, 0.0f);
    rho[i] = dens;
    alpha[i] = a;
    radv[i] = r;
    float4 vo = vi;
    if (WS != 2 && r > 0.0f && is_compute(L.rs, i)) {
        if (fabsf(ki) > kEps) {
            ax += ki * bx; ay += ki * by; az += ki * bz;
            if (SOL) {
                // reactions on the body particles: a second pass, since the
                // gate (a positive source) is known only after the first
                const float mk = -pi.w * ki;
                for_each_nbr32(L.nbrb, stride, i, nb, 0u, [&](uint32_t j) { return load_bnd<SOL>(bpos, L.sol, j); },
                    [&](const BN& bj, bool valid) {
                    const float dx = pi.x - bj.p.x, dy = pi.y - bj.p.y, dz = pi.z - bj.p.z;
                    const PairL kp = grad_listed(c, dist2(dx, dy, dz));
                    const float f = keep(valid && kp.ok, mk * bj.p.w * kp.g);
                    add_reaction<SOL>(L.sol, bj.v, valid, f * dx, f * dy, f * dz);
                });
            }
        }
        vo.x += h * ax; vo.y += h * ay; vo.z += h * az;
    }
    vel_o[i] = vo;
}

// Divergence source from the current velocities: computeDensityChange
// (DFSPH_core_cuda.cu:546-614) in DFSPH_divergence_init_kernel (:618-709, INIT:
// kappaV = 0) and DFSPH_divergence_loop_end_kernel (:732-748) + the FP64 error
// reduction that replaces cub::DeviceReduce::Sum + cudaMemcpy (:774-781).
// Also stores k_i = rho_adv_i * alpha_i for the next Jacobi sweep.
template <bool INIT, bool SOL, bool UM>
__global__ __launch_bounds__(kBlock, UM ? DFSPH_PV_WAVES : 1) void k_div_source(SimConst c, uint32_t n, const float4* __restrict__ pos,
                                                       const float4* __restrict__ bpos, NbrLists L,
                                                       const float4* __restrict__ vel,
                                                       const float* __restrict__ alpha,
                                                       float* __restrict__ radv, float* __restrict__ kk,
                                                       float* __restrict__ kappaV, SolveCtl* ctl,
                                                       double* partial, CtlInit ci)
{
    if (INIT && ci.sc && blockIdx.x == 0 && threadIdx.x == 0) ctl_init(ci);
    __shared__ double lds4[kBlock / 64];
    __shared__ PVStage<UM> st;
    if (!INIT && ctl->done) return;
    const uint32_t b = xcd_block();
    const uint32_t i = b * kBlock + threadIdx.x;
    const uint32_t ic = i < n ? i : 0u;         // own loads before staging, so they overlap it
    const uint32_t sn0 = L.stage_n[b];
    const Head hd = list_head(L, sn0, c.cap, ic);
    const float4 pi = pos[ic];
    const float4 vi = vel[ic];
    const float ai = alpha[ic];
    const uint32_t sn = stage_block<kStageMaxPV>(L, b, sn0, [&](uint32_t g) { return PV{pos[g], vel[g]}; },
                                    [&](uint32_t k, const PV& d) { st.put(k, d); });
    float r = 0.0f;
    if (i < n && !is_source(L.rs, i)) r = radv[i];      // restricted out: stored value in the error sum
    if (i < n && is_source(L.rs, i)) {
        const uint32_t cn = hd.cnt;
        const uint32_t nf = cn & 0xffffu, nb = cn >> 16;
        // with a uniform mass the fluid sum is taken without it and scaled once
        float dr = 0.0f;
        auto fluid = [&](const float4& pj, const float4& vj, bool valid) {
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const PairL kp = grad_listed(c, dist2(dx, dy, dz));
            const float mg = keep(valid && kp.ok, UM ? kp.g : pj.w * kp.g);
            dr += mg * ((vi.x - vj.x) * dx + (vi.y - vj.y) * dy + (vi.z - vj.z) * dz);
        };
        fluid_nbrs<kStageMaxPV>(L, c.cap, b, i, sn, hd, [&](uint32_t e) { return st.get(e, c.m_uniform); },
                   [&](uint32_t j) { return PV{pos[j], vel[j]}; },
                   [&](const PV& d, bool valid) { fluid(d.p, d.v, valid); });
        if (UM) dr *= c.m_uniform;
        for_each_nbr32(L.nbrb, c.cap, i, nb, 0u, hd.b0, hd.b1,
            [&](uint32_t j) { return load_bnd<SOL>(bpos, L.sol, j); }, [&](const BN& bj, bool valid) {
            const float4& pj = bj.p;
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const PairL kp = grad_listed(c, dist2(dx, dy, dz));
            const float mg = keep(valid && kp.ok, pj.w * kp.g);
            dr += mg * ((vi.x - bj.v.x) * dx + (vi.y - bj.v.y) * dy + (vi.z - bj.v.z) * dz);
        });
        r = (nf + nb < 20u) ? 0.0f : fmaxf(dr, 0.0f);
        radv[i] = r;
        kk[i] = r * ai;
        if (INIT) kappaV[i] = 0.0f;
    }
    if (!INIT) {
        const double bs = block_sum((double)r, lds4);
        store_partial(partial, bs);
    }
}

// Jacobi velocity update: divergenceSolveParticle<false> (DFSPH_core_cuda.cu:400-481,
// DIV: kappaV += k_i, step h) and pressureSolveParticle<false> (:867-951, step h_future):
// v_i += step (sum_f [|k_i + k_j| > eps] (k_i + k_j) m_j gradW + [|k_i| > eps] sum_b k_i psi_j gradW)
template <bool DIV, bool SOL, bool UM>
__global__ __launch_bounds__(kBlock) void k_jacobi(SimConst c, const DevTime* t, uint32_t n,
                                                   const float4* __restrict__ pos,
                                                   const float4* __restrict__ bpos, NbrLists L,
                                                   const float* __restrict__ kk, float* __restrict__ kappaV,
                                                   float4* __restrict__ vel, const SolveCtl* ctl)
{
    __shared__ PKStage<UM> st;
    if (ctl->done) return;
    const uint32_t b = xcd_block();
    const uint32_t i = b * kBlock + threadIdx.x;
    const uint32_t ic = i < n ? i : 0u;         // own loads before staging, so they overlap it
    const uint32_t sn0 = L.stage_n[b];
    const float ki = kk[ic];
    const float kv = DIV ? kappaV[ic] : 0.0f;
    const float4 pi = pos[ic];
    float4 v = vel[ic];
    const uint32_t sn = stage_block(L, b, sn0, [&](uint32_t g) { return PK{pos[g], kk[g]}; },
                                    [&](uint32_t k, const PK& d) { st.put(k, d); });
    if (i >= n || !is_compute(L.rs, i)) return;
    if (DIV) kappaV[i] = kv + ki;
    const uint32_t cn = L.cnt[i];
    const uint32_t nf = cn & 0xffffu, nb = cn >> 16;
    // with a uniform mass the fluid sum is taken without it and scaled once
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    auto fluid = [&](const float4& pj, float kj, bool valid) {
        const float ks = ki + kj;
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const PairL kp = grad_listed(c, dist2(dx, dy, dz));
        const float s = keep(valid && kp.ok && fabsf(ks) > kEps, UM ? ks * kp.g : ks * pj.w * kp.g);
        ax += s * dx; ay += s * dy; az += s * dz;
    };
    fluid_nbrs(L, c.cap, b, i, nf, sn, [&](uint32_t e) { return st.get(e, c.m_uniform); },
               [&](uint32_t j) { return PK{pos[j], kk[j]}; },
               [&](const PK& d, bool valid) { fluid(d.p, d.k, valid); });
    if (UM) { ax *= c.m_uniform; ay *= c.m_uniform; az *= c.m_uniform; }
    if (fabsf(ki) > kEps) {
        const float mi = pi.w;
        for_each_nbr32(L.nbrb, c.cap, i, nb, 0u, [&](uint32_t j) { return load_bnd<SOL>(bpos, L.sol, j); },
            [&](const BN& bj, bool valid) {
            const float4& pj = bj.p;
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const PairL kp = grad_listed(c, dist2(dx, dy, dz));
            const float s = keep(valid && kp.ok, ki * pj.w * kp.g);
            ax += s * dx; ay += s * dy; az += s * dz;
            add_reaction<SOL>(L.sol, bj.v, valid, -mi * s * dx, -mi * s * dy, -mi * s * dz);
        });
    }
    const float step = DIV ? t->h : t->h_future;
    v.x += step * ax; v.y += step * ay; v.z += step * az;
    vel[i] = v;
}

// DFSPH_viscosityXSPH_kernel (DFSPH_core_cuda.cu:1423-1510):
// a_i = g - sum_f (1/h) nu (m_j / rho_j) (v_i - v_j) W_ij, fused with the CFL
// maximum of |v + a h|^2 (DFSPH_CFLVelSquaredNorm_kernel, :3273-3278).
__global__ __launch_bounds__(kBlock, DFSPH_PV_WAVES) void k_xsph(SimConst c, const DevTime* t, uint32_t n,
                                                 const float4* __restrict__ pos, NbrLists L,
                                                 const float4* __restrict__ vel,
                                                 const float* __restrict__ rho, float4* __restrict__ acc,
                                                 DevCtl* ctl, float* partial)
{
    __shared__ float lds4[kBlock / 64];
    __shared__ PVSStage<true> st;                // (x, v, m / rho) of the staged set (the mass is not used)
    const uint32_t b = xcd_block();
    const uint32_t i = b * kBlock + threadIdx.x;
    const uint32_t ic = i < n ? i : 0u;         // own loads before staging, so they overlap it
    const uint32_t sn0 = L.stage_n[b];
    const Head hd = list_head(L, sn0, c.cap, ic);
    const float4 pi = pos[ic];
    const float4 vi = vel[ic];
    const uint32_t sn = stage_block<kStageMaxX>(L, b, sn0, [&](uint32_t g) {
        const float4 p = pos[g];
        const float4 v = vel[g];
        return PV{p, make_float4(v.x, v.y, v.z, p.w / rho[g])};
    }, [&](uint32_t k, const PV& d) { st.put(k, d); });
    float m2 = 0.0f;
    if (i < n && !is_compute(L.rs, i)) {
        // restricted out: its stored acceleration still enters the CFL bound
        // (the reference's maximum runs over all particles, DFSPH_core_cuda.cu:3340-3355)
        const float4 a = acc[i];
        const float h = t->h;
        const float wx = vi.x + a.x * h, wy = vi.y + a.y * h, wz = vi.z + a.z * h;
        m2 = wx * wx + wy * wy + wz * wz;
    }
    if (i < n && is_compute(L.rs, i)) {
        const float coef = t->inv_h * c.viscosity;
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        auto fluid = [&](const float4& pj, const float4& vj, bool valid) {
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float w = W_listed(c, dist2(dx, dy, dz));
            const float s = keep(valid, coef * vj.w * w);
            ax -= s * (vi.x - vj.x); ay -= s * (vi.y - vj.y); az -= s * (vi.z - vj.z);
        };
        fluid_nbrs<kStageMaxX>(L, c.cap, b, i, sn, hd, [&](uint32_t e) { return st.get(e, 0.0f); },
                   [&](uint32_t j) {
                       const float4 p = pos[j];
                       const float4 v = vel[j];
                       return PV{p, make_float4(v.x, v.y, v.z, p.w / rho[j])};
                   },
                   [&](const PV& d, bool valid) { fluid(d.p, d.v, valid); });
        const float4 a = make_float4(c.gx + ax, c.gy + ay, c.gz + az, 0.0f);
        acc[i] = a;
        const float h = t->h;
        const float wx = vi.x + a.x * h, wy = vi.y + a.y * h, wz = vi.z + a.z * h;
        m2 = wx * wx + wy * wy + wz * wz;
    }
    float m = wave_max(m2);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds4[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float bm = lds4[0];
        for (int k = 1; k < kBlock / 64; ++k) bm = fmaxf(bm, lds4[k]);
        partial[blockIdx.x] = bm;
    }
}

// maximum of the XSPH kernel's block maxima (the CFL velocity bound)
__global__ __launch_bounds__(kBlock) void k_cfl_finalize(DevCtl* ctl, const float* partial, uint32_t nparts)
{
    __shared__ float lds4[kBlock / 64];
    constexpr uint32_t kBatch = 8;
    float mm = 0.0f;
    for (uint32_t k0 = threadIdx.x; k0 < nparts; k0 += kBatch * kBlock) {
        float v[kBatch];
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
            const uint32_t k = k0 + u * kBlock;
            v[u] = k < nparts ? partial[k] : 0.0f;
        }
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) mm = fmaxf(mm, v[u]);
    }
    mm = wave_max(mm);
    if ((threadIdx.x & 63) == 0) lds4[threadIdx.x >> 6] = mm;
    __syncthreads();
    if (threadIdx.x == 0) {
        float bm = lds4[0];
        for (int k = 1; k < kBlock / 64; ++k) bm = fmaxf(bm, lds4[k]);
        ctl->cfl_max = bm;
        ctl->escaped = 0;                      // counted afresh by the step's update_pos
    }
}

// DFSPH_update_vel_kernel (DFSPH_core_cuda.cu:3047-3070): v += h a and the
// pressure warm-start rescale kappa = max(kappa h^2 / h_f^2, -0.5).
__global__ void k_update_vel(const DevTime* t, uint32_t n, const float4* acc, float4* vel, float* kappa, Restrict rs)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !is_compute(rs, i)) return;
    const float h = t->h;
    const float4 a = acc[i];
    float4 v = vel[i];
    v.x += h * a.x; v.y += h * a.y; v.z += h * a.z;
    vel[i] = v;
    kappa[i] = fmaxf(kappa[i] * t->ratio2, -0.5f);
}

// Pressure warm start cuda_pressure_compute<true> (DFSPH_core_cuda.cu:965-1000,
// pressureSolveParticle<true> :867-951), optionally fused with the velocity
// update of DFSPH_update_vel_kernel (APPLY_ACC):
// v_i += h a_i;  v_i += h_f (sum_f [|k_i+k_j|>eps] (k_i+k_j) m_j gradW + [|k_i|>eps] sum_b k_i psi_j gradW)
// with k = max(kappa ratio2, -0.5) (WS == 0; ratio2 = 1 when already applied),
// or that value plus the learned correction, bounded by ws_learned (WS == 1).
// kws holds the last solve's kappa increments.  No gate (unlike the divergence
// warm start).  The applied k_i starts the particle's total of the step in
// ktot (when given).
template <int WS, bool APPLY_ACC, bool SOL>
__global__ __launch_bounds__(kBlock) void k_pressure_ws(SimConst c, const DevTime* t, uint32_t n,
                                                        const float4* __restrict__ pos,
                                                        const float4* __restrict__ bpos, NbrLists L,
                                                        const float4* __restrict__ acc,
                                                        float4* __restrict__ vel,
                                                        const float* __restrict__ kws, WsLearned lw, int rescale,
                                                        float* __restrict__ ktot)
{
    __shared__ float4 sP[kStageMax];
    __shared__ float sK[kStageMax];
    const float ratio2 = rescale ? t->ratio2 : 1.0f;
    const float ihf2 = t->inv_h_future * t->inv_h_future;
    auto ws_of = [&](uint32_t j) -> float {
        const float ref = fmaxf(kws[j] * ratio2, -0.5f);
        if (WS == 0) return ref;
        return ws_learned(ref, fmaxf(lw.prev[j] * ihf2, -0.5f), lw.dy[j] * lw.scale * ihf2);
    };
    const uint32_t b = xcd_block();
    const uint32_t i = b * kBlock + threadIdx.x;
    const uint32_t ic = i < n ? i : 0u;         // own loads before staging, so they overlap it
    const uint32_t sn0 = L.stage_n[b];
    const float ki = ws_of(ic);
    const float4 pi = pos[ic];
    float4 v = vel[ic];
    const float4 aown = APPLY_ACC ? acc[ic] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint32_t sn = stage_block(L, b, sn0, [&](uint32_t g) { return PK{pos[g], ws_of(g)}; },
                                    [&](uint32_t k, const PK& d) { sP[k] = d.p; sK[k] = d.k; });
    if (i >= n || !is_compute(L.rs, i)) return;
    if (ktot) ktot[i] = ki;
    const uint32_t cn = L.cnt[i];
    const uint32_t nf = cn & 0xffffu, nb = cn >> 16;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    auto fluid = [&](const float4& pj, float kj, bool valid) {
        const float ks = ki + kj;
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const PairL kp = grad_listed(c, dist2(dx, dy, dz));
        const float s = keep(valid && kp.ok && fabsf(ks) > kEps, ks * pj.w * kp.g);
        ax += s * dx; ay += s * dy; az += s * dz;
    };
    fluid_nbrs(L, c.cap, b, i, nf, sn, [&](uint32_t e) { return PK{lds_f4(sP, e), lds_f1(sK, e)}; },
               [&](uint32_t j) { return PK{pos[j], ws_of(j)}; },
               [&](const PK& d, bool valid) { fluid(d.p, d.k, valid); });
    if (fabsf(ki) > kEps) {
        const float mi = pi.w;
        for_each_nbr32(L.nbrb, c.cap, i, nb, 0u, [&](uint32_t j) { return load_bnd<SOL>(bpos, L.sol, j); },
            [&](const BN& bj, bool valid) {
            const float4& pj = bj.p;
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const PairL kp = grad_listed(c, dist2(dx, dy, dz));
            const float s = keep(valid && kp.ok, ki * pj.w * kp.g);
            ax += s * dx; ay += s * dy; az += s * dz;
            add_reaction<SOL>(L.sol, bj.v, valid, -mi * s * dx, -mi * s * dy, -mi * s * dz);
        });
    }
    if (APPLY_ACC) {
        const float h = t->h;
        const float4 a = aown;
        v.x += h * a.x; v.y += h * a.y; v.z += h * a.z;
    }
    const float hf = t->h_future;
    v.x += hf * ax; v.y += hf * ay; v.z += hf * az;
    vel[i] = v;
}

// Density source computeDensityAdv (DFSPH_core_cuda.cu:1165-1215):
// rho_adv = max(rho + h_f sum m_j (v_i - v_j).gradW - rho0, 0); kappa += rho_adv alpha.
// INIT: DFSPH_pressure_init_kernel (:1217-1262): alpha *= 1/h_f, kappa = 0 first.
// Otherwise DFSPH_pressure_loop_end_kernel (:1264-1284) + FP64 error reduction
// (replaces cub Sum + cudaMemcpy, :1321-1325).
template <bool INIT, bool SOL, bool UM>
__global__ __launch_bounds__(kBlock, UM ? DFSPH_PV_WAVES : 1) void k_density_source(SimConst c, const DevTime* t, uint32_t n,
                                                           const float4* __restrict__ pos,
                                                           const float4* __restrict__ bpos, NbrLists L,
                                                           const float4* __restrict__ vel,
                                                           const float* __restrict__ rho,
                                                           float* __restrict__ alpha, float* __restrict__ radv,
                                                           float* __restrict__ kk, float* __restrict__ kappa,
                                                           SolveCtl* ctl, double* partial, CtlInit ci)
{
    if (INIT && ci.sc && blockIdx.x == 0 && threadIdx.x == 0) ctl_init(ci);
    __shared__ double lds4[kBlock / 64];
    __shared__ PVStage<UM> st;
    if (!INIT && ctl->done) return;
    const uint32_t b = xcd_block();
    const uint32_t i = b * kBlock + threadIdx.x;
    const uint32_t ic = i < n ? i : 0u;         // own loads before staging, so they overlap it
    const uint32_t sn0 = L.stage_n[b];
    const Head hd = list_head(L, sn0, c.cap, ic);
    const float4 pi = pos[ic];
    const float4 vi = vel[ic];
    const float a0 = alpha[ic], rho_i = rho[ic];
    const float kap0 = INIT ? 0.0f : kappa[ic];
    const uint32_t sn = stage_block<kStageMaxPV>(L, b, sn0, [&](uint32_t g) { return PV{pos[g], vel[g]}; },
                                    [&](uint32_t k, const PV& d) { st.put(k, d); });
    float r = 0.0f;
    if (i < n && !is_source(L.rs, i)) r = radv[i];      // restricted out: stored value in the error sum
    if (i < n && is_source(L.rs, i)) {
        const float hf = t->h_future;
        const uint32_t cn = hd.cnt;
        const uint32_t nf = cn & 0xffffu, nb = cn >> 16;
        // with a uniform mass the fluid sum is taken without it and scaled once
        float dr = 0.0f;
        auto fluid = [&](const float4& pj, const float4& vj, bool valid) {
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const PairL kp = grad_listed(c, dist2(dx, dy, dz));
            const float mg = keep(valid && kp.ok, UM ? kp.g : pj.w * kp.g);
            dr += mg * ((vi.x - vj.x) * dx + (vi.y - vj.y) * dy + (vi.z - vj.z) * dz);
        };
        fluid_nbrs<kStageMaxPV>(L, c.cap, b, i, sn, hd, [&](uint32_t e) { return st.get(e, c.m_uniform); },
                   [&](uint32_t j) { return PV{pos[j], vel[j]}; },
                   [&](const PV& d, bool valid) { fluid(d.p, d.v, valid); });
        if (UM) dr *= c.m_uniform;
        (void)nf;
        for_each_nbr32(L.nbrb, c.cap, i, nb, 0u, hd.b0, hd.b1,
            [&](uint32_t j) { return load_bnd<SOL>(bpos, L.sol, j); }, [&](const BN& bj, bool valid) {
            const float4& pj = bj.p;
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const PairL kp = grad_listed(c, dist2(dx, dy, dz));
            const float mg = keep(valid && kp.ok, pj.w * kp.g);
            dr += mg * ((vi.x - bj.v.x) * dx + (vi.y - bj.v.y) * dy + (vi.z - bj.v.z) * dz);
        });
        float a = a0;
        if (INIT) { a *= t->inv_h_future; alpha[i] = a; }
        r = fmaxf(rho_i + hf * dr - c.density0, 0.0f);
        radv[i] = r;
        const float ka = r * a;
        kk[i] = ka;
        kappa[i] = INIT ? ka : kap0 + ka;
    }
    if (!INIT) {
        const double bs = block_sum((double)r, lds4);
        store_partial(partial, bs);
    }
}

// DFSPH_update_pos_kernel (DFSPH_core_cuda.cu:3127-3214): x += h v, fused with the
// escape count of tag_particles_outside_bounding_box_kernel
// (SPH_other_systems_cuda.cu:1244-1281): outside if any coordinate leaves the box.
__global__ void k_update_pos(const DevTime* t, uint32_t n, const float4* vel, float4* pos,
                             float bx0, float by0, float bz0, float bx1, float by1, float bz1, DevCtl* ctl,
                             Restrict rs)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t out = 0;
    if (i < n) {
        const float h = t->h;
        const float4 v = vel[i];
        float4 x = pos[i];
        if (is_compute(rs, i)) {
            x.x += h * v.x; x.y += h * v.y; x.z += h * v.z;
            pos[i] = x;
        }
        out = (x.x > bx1 || x.y > by1 || x.z > bz1 || x.x < bx0 || x.y < by0 || x.z < bz0) ? 1u : 0u;
    }
    const uint32_t s = wave_sum_u(out);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(&ctl->escaped, s);
}

// keep flags for the compaction (tag_particles_outside_bounding_box_kernel)
__global__ void k_keep_inside(uint32_t n, const float4* pos, float bx0, float by0, float bz0,
                              float bx1, float by1, float bz1, uint32_t* keep)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { keep[n] = 0; return; }
    const float4 x = pos[i];
    const bool out = x.x > bx1 || x.y > by1 || x.z > bz1 || x.x < bx0 || x.y < by0 || x.z < bz0;
    keep[i] = out ? 0u : 1u;
}

__global__ void k_keep_from_u8(uint32_t n, const uint8_t* k8, uint32_t* keep)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    keep[i] = (i < n && k8[i]) ? 1u : 0u;
}

// Stable compaction: the kept particles move to their exclusive-scan rank
// (replaces the radix sort by tag + cuda_sortData of remove_tagged_particles,
// SPH_other_systems_cuda.cu:1606-1641).
__global__ void k_compact(uint32_t n, const uint32_t* keep, const uint32_t* rank, FluidPtrs in, FluidPtrs out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    copy_fluid(in, i, out, rank[i]);
}

// End of a solve: the step's total = applied warm start + the solver's
// accumulated increments, in dt-invariant units (scale h for kappaV, h_f^2 for kappa).
__global__ void k_finish_totals(const DevTime* t, uint32_t n, int pressure, const float* inc, float* tot)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float s = pressure ? t->h_future * t->h_future : t->h;
    tot[i] = (tot[i] + inc[i]) * s;
}

// Features of the learned warm start, one neighbor sweep over the current state
// (before the divergence warm start):
//  0 rho / rho0 - 1            1 divergence source h / rho0 (unclamped)
//  2 (n_fluid + n_boundary) / 40   3 |v| h / (2r)
//  4 last density increments kappa h_f^2 / d^2   5 last divergence increments kappaV h / d
//  6 boundary fraction n_b / n     7-9 mean neighbor offset / h_k
// 10 mean neighbor distance / h_k 11 fluid share of the density
template <bool SOL, bool UM>
__global__ __launch_bounds__(kBlock, UM ? DFSPH_PV_WAVES : 1) void k_features(SimConst c, const DevTime* t, uint32_t n,
                                                     const float4* __restrict__ pos,
                                                     const float4* __restrict__ bpos, NbrLists L,
                                                     const float4* __restrict__ vel,
                                                     const float* __restrict__ kap,
                                                     const float* __restrict__ kapv,
                                                     const float* __restrict__ ktot,
                                                     const float* __restrict__ kvtot, float diameter,
                                                     float* __restrict__ feat, float* __restrict__ prev_p,
                                                     float* __restrict__ prev_v)
{
    __shared__ PVStage<UM> st;
    const uint32_t b = xcd_block();
    const uint32_t i = b * kBlock + threadIdx.x;
    const uint32_t ic = i < n ? i : 0u;         // own loads before staging, so they overlap it
    const uint32_t sn0 = L.stage_n[b];
    const Head hd = list_head(L, sn0, c.cap, ic);
    const float4 pi = pos[ic];
    const float4 vi = vel[ic];
    const uint32_t sn = stage_block<kStageMaxPV>(L, b, sn0, [&](uint32_t g) { return PV{pos[g], vel[g]}; },
                                    [&](uint32_t k, const PV& d) { st.put(k, d); });
    if (i >= n) return;
    const uint32_t cn = hd.cnt;
    const uint32_t nf = cn & 0xffffu, nb = cn >> 16;
    float dens = pi.w * c.W0, dfl = 0.0f, dr = 0.0f;
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, od = 0.0f;
    auto fluid = [&](const float4& pj, const float4& vj, bool valid) {
        const float m = keep(valid, 1.0f);
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float r2 = dist2(dx, dy, dz);
        const PairL kp = pair_listed(c, r2);
        const float mj = m * pj.w;
        dfl += mj * kp.W;
        dr += keep(kp.ok, mj * kp.g) * ((vi.x - vj.x) * dx + (vi.y - vj.y) * dy + (vi.z - vj.z) * dz);
        ox -= m * dx; oy -= m * dy; oz -= m * dz;
        od += m * __builtin_sqrtf(r2);
    };
    fluid_nbrs<kStageMaxPV>(L, c.cap, b, i, sn, hd, [&](uint32_t e) { return st.get(e, c.m_uniform); },
               [&](uint32_t j) { return PV{pos[j], vel[j]}; },
               [&](const PV& d, bool valid) { fluid(d.p, d.v, valid); });
    float dbd = 0.0f;
    for_each_nbr32(L.nbrb, c.cap, i, nb, 0u, hd.b0, hd.b1, [&](uint32_t j) { return load_bnd<SOL>(bpos, L.sol, j); },
        [&](const BN& bj, bool valid) {
        const float4& pj = bj.p;
        const float m = keep(valid, 1.0f);
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float r2 = dist2(dx, dy, dz);
        const PairL kp = pair_listed(c, r2);
        const float mj = m * pj.w;
        dbd += mj * kp.W;
        dr += keep(kp.ok, mj * kp.g) * ((vi.x - bj.v.x) * dx + (vi.y - bj.v.y) * dy + (vi.z - bj.v.z) * dz);
        ox -= m * dx; oy -= m * dy; oz -= m * dz;
        od += m * __builtin_sqrtf(r2);
    });
    dens += dfl + dbd;
    const float nn = (float)(nf + nb);
    const float inv_n = nn > 0.0f ? 1.0f / nn : 0.0f;
    const float s = c.inv_hk * inv_n;
    float* f = feat + (size_t)i * 12;
    f[0] = dens / c.density0 - 1.0f;
    f[1] = dr * t->h / c.density0;
    f[2] = nn * (1.0f / 40.0f);
    f[3] = __builtin_sqrtf(vi.x * vi.x + vi.y * vi.y + vi.z * vi.z) * t->h / diameter;
    const float inv_d = 1.0f / diameter;
    f[kFeatKappaP] = kap[i] * (t->h * t->h) * (inv_d * inv_d);
    f[kFeatKappaV] = kapv[i] * t->h_past * inv_d;
    f[6] = (float)nb * inv_n;
    f[7] = ox * s;
    f[8] = oy * s;
    f[9] = oz * s;
    f[10] = od * s;
    f[11] = dfl / dens;
    prev_p[i] = ktot[i];
    prev_v[i] = kvtot[i];
}

// training set, after a step's solves with the reference warm start: append
// the step's features and the solves' increments in dt-invariant units
// (kappaV h, kappa h_f^2), what the reference warm start left to the Jacobi
// iterations: the correction the network learns
__global__ void k_harvest(const DevTime* t, uint32_t n, uint32_t stride, uint32_t offset, uint32_t cap,
                          const float* feat, const float* kapv, const float* kap, float inv_d, float* ds_x,
                          float* ds_t)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = k * stride;
    if (i >= n || offset + k >= cap) return;
    const size_t o = offset + k;
    const float* f = feat + (size_t)i * 12;
    for (int a = 0; a < 12; ++a) ds_x[o * 12 + a] = f[a];
    ds_t[o * 2 + 0] = kapv[i] * t->h * inv_d;
    ds_t[o * 2 + 1] = kap[i] * (t->h_future * t->h_future) * (inv_d * inv_d);
}

// ------------------------------------------------------------------------
// host launch wrappers
// ------------------------------------------------------------------------
static inline uint32_t nblk(uint32_t n) { return (n + kBlock - 1) / kBlock; }
static inline uint32_t nblk_xcd(uint32_t n)
{
    const uint32_t b = nblk(n);
    return ((b + kNumXcd - 1) / kNumXcd) * kNumXcd;
}

uint32_t grid_blocks_xcd(uint32_t n) { return nblk_xcd(n ? n : 1); }

static thread_local hipEvent_t tl_ev_start = nullptr, tl_ev_stop = nullptr;

void arm_sweep_events(hipEvent_t start, hipEvent_t stop)
{
    tl_ev_start = start;
    tl_ev_stop = stop;
}
bool sweep_events_armed() { return tl_ev_start != nullptr; }
void disarm_sweep_events() { tl_ev_start = tl_ev_stop = nullptr; }

// A sweep launch of g blocks of kBlock threads; it takes the armed timing
// events, if any, into its dispatch.
template <class... P, class... A>
static void sweep_launch(void (*k)(P...), uint32_t g, hipStream_t s, A... a)
{
    if (tl_ev_start) {
        hipExtLaunchKernelGGL(k, dim3(g), dim3(kBlock), 0u, s, tl_ev_start, tl_ev_stop, 0u, a...);
        disarm_sweep_events();
    } else {
        k<<<g, kBlock, 0, s>>>(a...);
    }
}

void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, uint32_t n)
{
    if (n) k_fill_u32<<<nblk(n), kBlock, 0, s>>>(p, v, n);
}

void launch_fill_f32(hipStream_t s, float* p, float v, uint32_t n)
{
    if (n) k_fill_f32<<<nblk(n), kBlock, 0, s>>>(p, v, n);
}

size_t scan_temp_words(uint32_t n)
{
    size_t words = 0;
    uint32_t m = n;
    while (m > 1) {
        const uint32_t tiles = (m + kScanTile - 1) / kScanTile;
        words += 2 * (size_t)tiles + 2;
        m = tiles;
    }
    return words + 4;
}

void launch_exclusive_scan(hipStream_t s, uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp, bool clear_in)
{
    if (n == 0) return;
    const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    uint32_t* sums = tmp;
    uint32_t* offs = tmp + tiles + 1;
    k_scan_tiles<<<tiles, kBlock, 0, s>>>(in, out, n, sums, clear_in ? 1 : 0);
    if (tiles > 1 && tiles <= kScanDirect) {
        k_scan_add_direct<<<tiles, kBlock, 0, s>>>(out, n, sums);
    } else if (tiles > 1) {
        launch_exclusive_scan(s, sums, offs, tiles, tmp + 2 * (size_t)tiles + 2, false);
        k_scan_add<<<tiles, kBlock, 0, s>>>(out, n, offs);
    }
}

void launch_nbr_ctl_init(hipStream_t s, DevCtl* c) { k_nbr_ctl_init<<<1, 64, 0, s>>>(c); }

void launch_time_reset(hipStream_t s, DevTime* t, double h) { k_time_reset<<<1, 64, 0, s>>>(t, h); }

void launch_time_update(hipStream_t s, DevTime* t, double hf, const DevCtl* ctl, int mode,
                        double cfl_factor, double diameter, double dt_min, double dt_max, double seed)
{
    k_time_update<<<1, 64, 0, s>>>(t, hf, ctl, mode, cfl_factor, diameter, dt_min, dt_max, seed);
}

void launch_time_step_end(hipStream_t s, DevTime* t, const Publish& pub) { k_time_step_end<<<1, 64, 0, s>>>(t, pub); }

void launch_time_set(hipStream_t s, DevTime* t, double h, double hp, double time)
{
    k_time_set<<<1, 64, 0, s>>>(t, h, hp, time);
}

void launch_cell_hist(hipStream_t s, const SimConst& c, const float4* pos, uint32_t n, uint32_t* key,
                      uint32_t* slot, uint32_t* hist)
{
    if (n) k_cell_hist<<<nblk(n), kBlock, 0, s>>>(c, pos, n, key, slot, hist);
}

void launch_scatter_perm(hipStream_t s, const uint32_t* key, const uint32_t* slot, const uint32_t* cstart,
                         uint32_t n, uint32_t* perm)
{
    if (n) k_scatter_perm<<<nblk(n), kBlock, 0, s>>>(key, slot, cstart, n, perm);
}

void launch_cell_sort(hipStream_t s, const uint32_t* cstart, uint32_t ncell, uint32_t* perm)
{
    k_cell_sort<<<nblk(ncell), kBlock, 0, s>>>(cstart, ncell, perm);
}

void launch_gather_fluid(hipStream_t s, uint32_t n, const uint32_t* perm, const FluidPtrs& in, const FluidPtrs& out)
{
    if (n) k_gather_fluid<<<nblk(n), kBlock, 0, s>>>(n, perm, in, out);
}

void launch_gather_f4(hipStream_t s, uint32_t n, const uint32_t* perm, const float4* in, float4* out)
{
    if (n) k_gather_f4<<<nblk(n), kBlock, 0, s>>>(n, perm, in, out);
}

void launch_neighbors(hipStream_t s, const SimConst& c, bool indirect, const float4* pos, const uint32_t* cs_f,
                      const uint32_t* perm, const float4* bpos, const uint32_t* cs_b, uint32_t n, uint16_t* nbr16,
                      uint32_t* nbrf, uint32_t* nbrb, uint32_t* cnt, uint32_t* stage_n, uint32_t* stg,
                      const Restrict& rs, uint4* part, DevCtl* ctl)
{
    if (!n) return;
    const uint32_t g = nblk_xcd(n);
    if (indirect)
        k_neighbors<true><<<g, kBlock, 0, s>>>(c, pos, cs_f, perm, bpos, cs_b, n, nbr16, nbrf, nbrb, cnt, stage_n, stg,
                                               rs, part);
    else
        k_neighbors<false><<<g, kBlock, 0, s>>>(c, pos, cs_f, perm, bpos, cs_b, n, nbr16, nbrf, nbrb, cnt, stage_n, stg,
                                                rs, part);
    k_nbr_finalize<<<1, kBlock, 0, s>>>(part, g, ctl);
}

void launch_expand_lists(hipStream_t s, uint32_t n, uint32_t cap, const NbrLists& L, uint32_t* out)
{
    if (n) k_expand_lists<<<nblk(n), kBlock, 0, s>>>(n, cap, L, out);
}

void launch_boundary_psi(hipStream_t s, const SimConst& c, const float4* bpos, const uint32_t* cs_b,
                         uint32_t nb, float* psi)
{
    if (nb) k_boundary_psi<<<nblk(nb), kBlock, 0, s>>>(c, bpos, cs_b, nb, psi);
}

void launch_set_w(hipStream_t s, float4* p, const float* w, uint32_t n)
{
    if (n) k_set_w<<<nblk(n), kBlock, 0, s>>>(p, w, n);
}

void launch_div_ws_init(hipStream_t s, int ws, const SimConst& c, const DevTime* t, uint32_t n,
                        const float4* pos, const float4* bpos, const NbrLists& L,
                        const float4* vel_i, float4* vel_o, const float* kws, const WsLearned& lw,
                        float* rho, float* alpha, float* radv, float* kvtot)
{
    if (!n) return;
    const uint32_t g = nblk_xcd(n);
    const bool sol = L.sol.bvel != nullptr;
    const bool um = c.m_uniform > 0.0f;
#define DWU(W, S, U) sweep_launch(k_div_ws_init<W, S, U>, g, s, c, t, n, pos, bpos, L, vel_i, vel_o, kws, lw, rho, alpha, radv, kvtot)
#define DWS(W, S) do { if (um) DWU(W, S, true); else DWU(W, S, false); } while (0)
    if (ws == 0) { if (sol) DWS(0, true); else DWS(0, false); }
    else if (ws == 1) { if (sol) DWS(1, true); else DWS(1, false); }
    else { if (sol) DWS(2, true); else DWS(2, false); }
#undef DWS
#undef DWU
}

void launch_div_source(hipStream_t s, bool init, const SimConst& c, uint32_t n, const float4* pos,
                       const float4* bpos, const NbrLists& L, const float4* vel,
                       const float* alpha, float* radv, float* kk, float* kappaV, SolveCtl* ctl, double* partial,
                       const CtlInit& ci)
{
    const uint32_t g = nblk_xcd(n ? n : 1);
    const bool sol = L.sol.bvel != nullptr;
    const bool um = c.m_uniform > 0.0f;
#define DS(I, S, U, CI) sweep_launch(k_div_source<I, S, U>, g, s, c, n, pos, bpos, L, vel, alpha, radv, kk, kappaV, ctl, partial, CI)
#define DSU(I, S, CI) do { if (um) DS(I, S, true, CI); else DS(I, S, false, CI); } while (0)
    if (init) { if (sol) DSU(true, true, ci); else DSU(true, false, ci); }
    else { if (sol) DSU(false, true, CtlInit{}); else DSU(false, false, CtlInit{}); }
#undef DSU
#undef DS
}

void launch_solve_finalize(hipStream_t s, uint32_t n, SolveCtl* ctl, const double* partial, const Publish& pub)
{
    k_solve_finalize<<<1, kBlock, 0, s>>>(ctl, partial, nblk_xcd(n ? n : 1), pub);
}

void launch_jacobi(hipStream_t s, bool div, const SimConst& c, const DevTime* t, uint32_t n, const float4* pos,
                   const float4* bpos, const NbrLists& L, const float* kk, float* kappaV,
                   float4* vel, const SolveCtl* ctl)
{
    if (!n) return;
    const uint32_t g = nblk_xcd(n);
    const bool sol = L.sol.bvel != nullptr;
    const bool um = c.m_uniform > 0.0f;
#define JA(D, S, U) sweep_launch(k_jacobi<D, S, U>, g, s, c, t, n, pos, bpos, L, kk, kappaV, vel, ctl)
#define JAU(D, S) do { if (um) JA(D, S, true); else JA(D, S, false); } while (0)
    if (div) { if (sol) JAU(true, true); else JAU(true, false); }
    else { if (sol) JAU(false, true); else JAU(false, false); }
#undef JAU
#undef JA
}

void launch_xsph(hipStream_t s, const SimConst& c, const DevTime* t, uint32_t n, const float4* pos,
                 const NbrLists& L, const float4* vel, const float* rho, float4* acc,
                 DevCtl* ctl, float* partial)
{
    const uint32_t g = nblk_xcd(n ? n : 1);
    sweep_launch(k_xsph, g, s, c, t, n, pos, L, vel, rho, acc, ctl, partial);
    k_cfl_finalize<<<1, kBlock, 0, s>>>(ctl, partial, g);
}

void launch_update_vel(hipStream_t s, const DevTime* t, uint32_t n, const float4* acc, float4* vel, float* kappa,
                       const Restrict& rs)
{
    if (n) k_update_vel<<<nblk(n), kBlock, 0, s>>>(t, n, acc, vel, kappa, rs);
}

void launch_pressure_ws(hipStream_t s, int ws, bool apply_acc, const SimConst& c, const DevTime* t, uint32_t n,
                        const float4* pos, const float4* bpos, const NbrLists& L,
                        const float4* acc, float4* vel, const float* kws, const WsLearned& lw, int rescale,
                        float* ktot)
{
    if (!n) return;
    const uint32_t g = nblk_xcd(n);
    const bool sol = L.sol.bvel != nullptr;
#define PW(W, A, S) sweep_launch(k_pressure_ws<W, A, S>, g, s, c, t, n, pos, bpos, L, acc, vel, kws, lw, rescale, ktot)
    if (ws == 1) {
        if (apply_acc) { if (sol) PW(1, true, true); else PW(1, true, false); }
        else { if (sol) PW(1, false, true); else PW(1, false, false); }
    } else {
        if (apply_acc) { if (sol) PW(0, true, true); else PW(0, true, false); }
        else { if (sol) PW(0, false, true); else PW(0, false, false); }
    }
#undef PW
}

void launch_density_source(hipStream_t s, bool init, const SimConst& c, const DevTime* t, uint32_t n,
                           const float4* pos, const float4* bpos, const NbrLists& L,
                           const float4* vel, const float* rho, float* alpha, float* radv, float* kk,
                           float* kappa, SolveCtl* ctl, double* partial, const CtlInit& ci)
{
    const uint32_t g = nblk_xcd(n ? n : 1);
    const bool sol = L.sol.bvel != nullptr;
    const bool um = c.m_uniform > 0.0f;
#define DN(I, S, U, CI) sweep_launch(k_density_source<I, S, U>, g, s, c, t, n, pos, bpos, L, vel, rho, alpha, radv, kk, kappa, ctl, partial, CI)
#define DNU(I, S, CI) do { if (um) DN(I, S, true, CI); else DN(I, S, false, CI); } while (0)
    if (init) { if (sol) DNU(true, true, ci); else DNU(true, false, ci); }
    else { if (sol) DNU(false, true, CtlInit{}); else DNU(false, false, CtlInit{}); }
#undef DNU
#undef DN
}

void launch_update_pos(hipStream_t s, const DevTime* t, uint32_t n, const float4* vel, float4* pos,
                       const float* bmin, const float* bmax, DevCtl* ctl, const Restrict& rs)
{
    if (n)
        sweep_launch(k_update_pos, nblk(n), s, t, n, vel, pos, bmin[0], bmin[1], bmin[2], bmax[0], bmax[1], bmax[2], ctl,
                     rs);
}

void launch_keep_inside(hipStream_t s, uint32_t n, const float4* pos, const float* bmin, const float* bmax,
                        uint32_t* keep)
{
    k_keep_inside<<<nblk(n + 1), kBlock, 0, s>>>(n, pos, bmin[0], bmin[1], bmin[2], bmax[0], bmax[1], bmax[2], keep);
}

void launch_keep_from_u8(hipStream_t s, uint32_t n, const uint8_t* k8, uint32_t* keep)
{
    k_keep_from_u8<<<nblk(n + 1), kBlock, 0, s>>>(n, k8, keep);
}

void launch_compact(hipStream_t s, uint32_t n, const uint32_t* keep, const uint32_t* rank, const FluidPtrs& in,
                    const FluidPtrs& out)
{
    if (n) k_compact<<<nblk(n), kBlock, 0, s>>>(n, keep, rank, in, out);
}

void launch_finish_totals(hipStream_t s, const DevTime* t, uint32_t n, int pressure, const float* inc, float* tot)
{
    if (n) k_finish_totals<<<nblk(n), kBlock, 0, s>>>(t, n, pressure, inc, tot);
}

void launch_features(hipStream_t s, const SimConst& c, const DevTime* t, uint32_t n, const float4* pos,
                     const float4* bpos, const NbrLists& L, const float4* vel, const float* kap,
                     const float* kapv, const float* ktot, const float* kvtot, float diameter, float* feat,
                     float* prev_p, float* prev_v)
{
    if (!n) return;
    const bool um = c.m_uniform > 0.0f;
#define KF(S, U) sweep_launch(k_features<S, U>, nblk_xcd(n), s, c, t, n, pos, bpos, L, vel, kap, kapv, ktot, kvtot, \
                                  diameter, feat, prev_p, prev_v)
    if (L.sol.bvel) { if (um) KF(true, true); else KF(true, false); }
    else { if (um) KF(false, true); else KF(false, false); }
#undef KF
}

void launch_harvest(hipStream_t s, const DevTime* t, uint32_t n, uint32_t stride, uint32_t offset, uint32_t cap,
                    const float* feat, const float* kapv, const float* kap, float diameter, float* ds_x,
                    float* ds_t)
{
    const uint32_t m = (n + stride - 1) / stride;
    if (m) k_harvest<<<nblk(m), kBlock, 0, s>>>(t, n, stride, offset, cap, feat, kapv, kap, 1.0f / diameter, ds_x,
                                                ds_t);
}

}  // namespace dfsph
