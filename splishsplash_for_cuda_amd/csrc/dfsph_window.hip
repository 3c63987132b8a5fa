// dfsph_window.hip -- open boundary and dynamic-window kernels (gfx950).
//
// The reference's open boundary (OpenBoundariesSimple.cu:117-621) adds fluid
// at predefined inflow positions where there is room and removes fluid that
// leaves through the open side; its dynamic window (DynamicWindow.cu) tags
// the particles the solver keeps working on and restricts the solver to them
// (restriction_mode).  Surfaces follow BufferFluidSurface
// (BufferFluidSurface.cu:421-1070): plane, cuboid, cylinder, sphere and the
// linear-sided star prism, each optionally reversed.
#include "dfsph_device.hpp"
#include "dfsph_launch.hpp"

namespace dfsph {

// ---------------------------------------------------------------- surfaces
__device__ __forceinline__ float norm3(float x, float y, float z)
{
    return __builtin_sqrtf(x * x + y * y + z * z);
}

// Whether the ray q + t u, t > 0, crosses the triangle (v0, v1, v2): the
// Moeller-Trumbore test (Moeller and Trumbore, "Fast, minimum storage
// ray-triangle intersection", 1997): barycentric coordinates (a, b) and the
// ray parameter t from Cramer's rule on [-u, e1, e2].
__device__ __forceinline__ bool ray_crosses(const float q[3], const float u[3], const float4& v0, const float4& v1,
                                            const float4& v2)
{
    const float e1[3] = {v1.x - v0.x, v1.y - v0.y, v1.z - v0.z};
    const float e2[3] = {v2.x - v0.x, v2.y - v0.y, v2.z - v0.z};
    const float p[3] = {u[1] * e2[2] - u[2] * e2[1], u[2] * e2[0] - u[0] * e2[2], u[0] * e2[1] - u[1] * e2[0]};
    const float det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (fabsf(det) < 1e-12f) return false;                  // the ray runs along the triangle's plane
    const float inv = 1.0f / det;
    const float s[3] = {q[0] - v0.x, q[1] - v0.y, q[2] - v0.z};
    const float a = inv * (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]);
    if (a < 0.0f || a > 1.0f) return false;
    const float w[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
    const float b = inv * (u[0] * w[0] + u[1] * w[1] + u[2] * w[2]);
    if (b < 0.0f || a + b > 1.0f) return false;
    return inv * (e2[0] * w[0] + e2[1] * w[1] + e2[2] * w[2]) > 0.0f;
}

// MeshPerso::isInside (BufferFluidSurface.cu:224-270): parity of the
// crossings of one ray of a fixed, generic direction (no mesh edge or vertex
// lies along it for the meshes of this library's tests)
__device__ bool mesh_inside(const Surface& S, float px, float py, float pz)
{
    const float q[3] = {px - S.center[0], py - S.center[1], pz - S.center[2]};
    for (int a = 0; a < 3; ++a)
        if (q[a] < S.bmin[a] || q[a] > S.bmax[a]) return false;
    const float u[3] = {kMeshRay[0], kMeshRay[1], kMeshRay[2]};
    uint32_t crossings = 0;
    for (uint32_t t = 0; t < S.ntri; ++t)
        crossings += ray_crosses(q, u, S.tri[3 * t], S.tri[3 * t + 1], S.tri[3 * t + 2]) ? 1u : 0u;
    return (crossings & 1u) != 0u;
}

// BufferFluidSurface::isinside (BufferFluidSurface.cu:775-880)
__device__ bool surface_inside(const Surface& S, float px, float py, float pz)
{
    bool r = false;
    switch (S.type) {
    case kSurfMesh: {
        r = mesh_inside(S, px, py, pz);
        break;
    }
    case kSurfPlane: {
        const float d = (px - S.o[0]) * S.n[0] + (py - S.o[1]) * S.n[1] + (pz - S.o[2]) * S.n[2];
        r = !(d < 0.0f);
        break;
    }
    case kSurfCuboid: {
        const float vx = fabsf(px - S.center[0]), vy = fabsf(py - S.center[1]), vz = fabsf(pz - S.center[2]);
        r = (vx < S.half[0]) && (vy < S.half[1]) && (vz < S.half[2]);
        break;
    }
    case kSurfCylinder: {
        const float vx = px - S.center[0], vy = py - S.center[1], vz = pz - S.center[2];
        if (fabsf(vy) > S.half[1]) { r = false; break; }
        r = norm3(vx, 0.0f, vz) < S.radius;
        break;
    }
    case kSurfSphere: {
        r = norm3(px - S.center[0], py - S.center[1], pz - S.center[2]) < S.radius;
        break;
    }
    case kSurfStar: {
        const float vx = px - S.center[0], vy = py - S.center[1], vz = pz - S.center[2];
        if (fabsf(vy) > S.half[1]) { r = false; break; }
        const float len = norm3(vx, 0.0f, vz);
        if (len < S.radius_internal) { r = true; break; }
        if (len > S.radius) { r = false; break; }
        // polar angle against the direction of the first point, folded into
        // one star period; the border there is the segment between the outer
        // and the inner vertex
        const float dot = (vx * S.dir[0] + vz * S.dir[2]) / len;
        const float alpha = acosf(dot);
        const float alpha0 = 2.0f * 3.14159265358979f / (float)S.points;
        float k = alpha / alpha0;
        while (k > 1.0f) k -= 1.0f;
        const float x = k * alpha0;
        float a1 = 0.0f, r1 = S.radius, a2 = alpha0 / 2.0f, r2 = S.radius_internal;
        if (x > alpha0 / 2.0f) { r1 = S.radius_internal; r2 = S.radius; a1 = a2; a2 = a1 + alpha0 / 2.0f; }
        const float x1 = r1 * cosf(a1), y1 = r1 * sinf(a1), x2 = r2 * cosf(a2), y2 = r2 * sinf(a2);
        const float m = (y2 - y1) / (x2 - x1);
        const float d = (-x1 * m + y1) / (-m * cosf(x) + sinf(x));
        r = len <= d;
        break;
    }
    default: r = false;
    }
    return S.reversed ? !r : r;
}

// BufferFluidSurface::getNormal (BufferFluidSurface.cu:730-772): towards the
// centre for the radial shapes, zero for planes, cuboids and meshes; unit length.
__device__ void surface_normal(const Surface& S, float px, float py, float pz, float* nx, float* ny, float* nz)
{
    float x = 0.0f, y = 0.0f, z = 0.0f;
    if (S.type == kSurfCylinder || S.type == kSurfSphere || S.type == kSurfStar) {
        x = S.center[0] - px; y = S.center[1] - py; z = S.center[2] - pz;
    }
    if (S.reversed) { x = -x; y = -y; z = -z; }
    const float l = norm3(x, y, z);
    if (l != 0.0f) { x /= l; y /= l; z /= l; }
    *nx = x; *ny = y; *nz = z;
}

// keep[i] = (inside S) == want_inside; keep[n] = 0 (the compaction scan's sentinel)
__global__ void k_surface_keep(const float4* __restrict__ p, uint32_t n, Surface S, int want_inside,
                               uint32_t* __restrict__ keep)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { keep[n] = 0; return; }
    const float4 q = p[i];
    keep[i] = surface_inside(S, q.x, q.y, q.z) == (want_inside != 0) ? 1u : 0u;
}

// keep[i] &= (inside S) == want_inside (after k_surface_keep)
__global__ void k_surface_keep_and(const float4* __restrict__ p, uint32_t n, Surface S, int want_inside,
                                   uint32_t* __restrict__ keep)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 q = p[i];
    if (surface_inside(S, q.x, q.y, q.z) != (want_inside != 0)) keep[i] = 0u;
}

// out[rank[i]] = in[i] for the kept entries (order kept)
__global__ void k_compact_f4(uint32_t n, const uint32_t* __restrict__ keep, const uint32_t* __restrict__ rank,
                             const float4* __restrict__ in, float4* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && keep[i]) out[rank[i]] = in[i];
}

// outflow_basic_kernel (OpenBoundariesSimple.cu:517-534): a particle outside
// both the interior and the fluid-surface shape leaves
__global__ void k_outflow_keep(const float4* __restrict__ p, uint32_t n, Surface interior, Surface surface,
                               uint32_t* __restrict__ keep)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { keep[n] = 0; return; }
    const float4 q = p[i];
    const bool out = !surface_inside(interior, q.x, q.y, q.z) && !surface_inside(surface, q.x, q.y, q.z);
    keep[i] = out ? 0u : 1u;
}

// inflow_with_predefined_positions_kernel (OpenBoundariesSimple.cu:360-508):
// a candidate position is filled when the nearest fluid particle within h is
// farther than min_dist; the new particle's velocity is the (m W)-weighted
// mean over the fluid (the boundary adds weight, not velocity); it is
// rejected when that density W0 m + sum m_j W_ij + sum psi_j W_ij exceeds
// max_density or the velocity points against the boundary's normal.
// Fluid through the cell table of the current order (perm), boundary through
// its own table.  keep[c] = 1 and vel_out[c] = v for accepted candidates.
__global__ __launch_bounds__(kBlock) void k_inflow(SimConst c, const float4* __restrict__ cand, uint32_t nc,
                                                   const float4* __restrict__ pos,
                                                   const float4* __restrict__ vel,
                                                   const uint32_t* __restrict__ cs_f,
                                                   const uint32_t* __restrict__ perm,
                                                   const float4* __restrict__ bpos,
                                                   const uint32_t* __restrict__ cs_b, Surface boundary,
                                                   float min_dist, float max_density, float mass,
                                                   uint32_t* __restrict__ keep, float4* __restrict__ vel_out)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nc) return;
    if (k == nc) { keep[nc] = 0; return; }
    const float4 p = cand[k];
    const int cx = cell_coord(c, p.x, 0), cy = cell_coord(c, p.y, 1), cz = cell_coord(c, p.z, 2);
    // nearest fluid particle within h; the scan stops once one is closer
    // than min_dist (the candidate is then rejected whatever the minimum)
    float dmin = 100.0f;
    for_each_adjacent_cell(c, cx, cy, cz, [&](uint32_t nk) {
        if (dmin <= min_dist) return;
        for (uint32_t q = cs_f[nk]; q < cs_f[nk + 1]; ++q) {
            const float4 pj = pos[perm[q]];
            const float r2 = dist2(p.x - pj.x, p.y - pj.y, p.z - pj.z);
            if (r2 < c.hk2) dmin = fminf(dmin, __builtin_sqrtf(r2));
        }
    });
    uint32_t ok = 0;
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (dmin > min_dist) {
        float dens = c.W0 * mass, sw = 0.0f, vx = 0.0f, vy = 0.0f, vz = 0.0f;
        for_each_adjacent_cell(c, cx, cy, cz, [&](uint32_t nk) {
            for (uint32_t q = cs_f[nk]; q < cs_f[nk + 1]; ++q) {
                const uint32_t j = perm[q];
                const float4 pj = pos[j];
                const float r2 = dist2(p.x - pj.x, p.y - pj.y, p.z - pj.z);
                if (r2 < c.hk2) {
                    const float w = kernel_W(c, r2) * pj.w;
                    const float4 vj = vel[j];
                    dens += w;
                    vx += vj.x * w; vy += vj.y * w; vz += vj.z * w;
                    sw += w;
                }
            }
        });
        for_each_adjacent_cell(c, cx, cy, cz, [&](uint32_t nk) {
            for (uint32_t j = cs_b[nk]; j < cs_b[nk + 1]; ++j) {
                const float4 pj = bpos[j];
                const float r2 = dist2(p.x - pj.x, p.y - pj.y, p.z - pj.z);
                if (r2 < c.hk2) {
                    const float w = kernel_W(c, r2) * pj.w;
                    sw += w;
                    dens += w;
                }
            }
        });
        if (sw > 0.0f) { vx /= sw; vy /= sw; vz /= sw; }
        float nx, ny, nz;
        surface_normal(boundary, p.x, p.y, p.z, &nx, &ny, &nz);
        ok = (dens > max_density || vx * nx + vy * ny + vz * nz < 0.0f) ? 0u : 1u;
        v = make_float4(vx, vy, vz, 0.0f);
    }
    keep[k] = ok;
    vel_out[k] = v;
}

// The accepted candidates, in candidate order, appended after the n current
// particles (kappa = kappaV = 0, mass of fluid particle 0, fresh ids).
__global__ void k_append_inflow(const float4* __restrict__ cand, uint32_t nc, const uint32_t* __restrict__ keep,
                                const uint32_t* __restrict__ rank, const float4* __restrict__ vnew, uint32_t n,
                                float mass, uint32_t first_id, FluidPtrs f)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc || !keep[k]) return;
    const uint32_t o = n + rank[k];
    const float4 p = cand[k];
    f.pos[o] = make_float4(p.x, p.y, p.z, mass);
    f.vel[o] = vnew[k];
    f.kap[o] = 0.0f;
    f.kapv[o] = 0.0f;
    f.ktot[o] = 0.0f;
    f.kvtot[o] = 0.0f;
    f.id[o] = first_id + rank[k];
    f.tag[o] = kTagUntagged;
}

// apply_delta_to_buffer_kernel (basic_kernels_cuda.cu:104-110)
__global__ void k_apply_delta(float4* __restrict__ p, uint32_t n, float dx, float dy, float dz)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 q = p[i];
    q.x += dx; q.y += dy; q.z += dz;
    p[i] = q;
}

// ---------------------------------------------------------------- tags
// count_particle_with_tag_kernel (RestFluidLoader.cu:2238-2252)
__global__ void k_count_tag(const uint32_t* __restrict__ tag, uint32_t n, uint32_t t, uint32_t* count)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t hit = (i < n && tag[i] == t) ? 1u : 0u;
    const uint32_t s = wave_sum_u(hit);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(count, s);
}

// tag_neighbors_of_tagged_kernel<untagged_only, true> (RestFluidLoader.cu:2070-2101):
// the fluid neighbors of every particle tagged tag_in get tag_out (only the
// untagged ones if untagged_only).  Reads the tags of a snapshot (src) and
// writes the live array (dst), so the result does not depend on the order
// the particles run in.
__global__ __launch_bounds__(kBlock) void k_tag_neighbors(SimConst c, uint32_t n, NbrLists L,
                                                          const uint32_t* __restrict__ src,
                                                          uint32_t* __restrict__ dst, uint32_t tag_in,
                                                          uint32_t tag_out, int untagged_only)
{
    const uint32_t b = xcd_block();
    const uint32_t i = b * kBlock + threadIdx.x;
    if (i >= n || src[i] != tag_in) return;
    const uint32_t nf = L.cnt[i] & 0xffffu;
    const uint32_t sn = L.stage_n[b];
    auto mark = [&](uint32_t j, bool valid) {
        if (valid && j != i && (!untagged_only || src[j] == kTagUntagged)) dst[j] = tag_out;
    };
    if (sn != kUnstaged) {
        const uint32_t* g = L.stg + (size_t)b * kStageCap;
        for_each_nbr16(L.nbr16, c.cap, i, nf, [&](uint32_t e, bool v) { mark(g[e >> 4], v); });
    } else {
        for_each_nbr32(L.nbrf, c.cap, i, nf, i, mark);
    }
}

// One extraction pass of the rest-fluid loader (RestFLuidLoader::tagDataToSurface,
// RestFluidLoader.cu:2700-3130, its selection particle_selection_rule_1_kernel,
// :773-890).  Particles tagged air (kTagAir) only add density: they stay and
// are not counted.  A candidate (any other tag) leaves when the boundary
// alone gives it a density above boundary_limit (the reference's 0.7 rho0),
// or when its density rho[i] (of the last density computation, boundary and
// air included) exceeds limit and no listed candidate above the limit is
// denser; of equal densities the lower index leaves.  So of two crowded
// neighbors one leaves per pass.  keep[i] = 0 / 1; sum[0] += rho and sum[1]
// += 1 over the kept candidates.  keep[n] is set by the caller.
__global__ __launch_bounds__(kBlock) void k_rest_extract(SimConst c, uint32_t n, NbrLists L,
                                                         const float4* __restrict__ pos,
                                                         const float4* __restrict__ bpos,
                                                         const float* __restrict__ rho,
                                                         const uint32_t* __restrict__ tag, float limit,
                                                         float boundary_limit, uint32_t* __restrict__ keep,
                                                         double* __restrict__ sum)
{
    const uint32_t b = xcd_block();
    const uint32_t i = b * kBlock + threadIdx.x;
    double kept_rho = 0.0, kept = 0.0;
    if (i < n && tag[i] == kTagAir) {
        keep[i] = 1u;
    } else if (i < n) {
        const uint32_t cn = L.cnt[i];
        const uint32_t nf = cn & 0xffffu, nbn = cn >> 16;
        const float4 pi = pos[i];
        float rb = 0.0f;
        for_each_nbr32(L.nbrb, c.cap, i, nbn, 0u, [&](uint32_t j, bool valid) {
            const float4 pj = bpos[j];
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            rb += valid ? pj.w * kernel_W(c, dist2(dx, dy, dz)) : 0.0f;
        });
        const float ri = rho[i];
        bool leave = rb > boundary_limit;
        if (!leave && ri > limit) {
            bool densest = true;
            auto look = [&](uint32_t j, bool valid) {
                if (!valid || j == i || tag[j] == kTagAir) return;
                const float rj = rho[j];
                if (rj > limit && (rj > ri || (rj == ri && j < i))) densest = false;
            };
            const uint32_t sn = L.stage_n[b];
            if (sn != kUnstaged) {
                const uint32_t* g = L.stg + (size_t)b * kStageCap;
                for_each_nbr16(L.nbr16, c.cap, i, nf, [&](uint32_t e, bool v) { look(g[e >> 4], v); });
            } else {
                for_each_nbr32(L.nbrf, c.cap, i, nf, i, look);
            }
            leave = densest;
        }
        keep[i] = leave ? 0u : 1u;
        kept_rho = leave ? 0.0 : (double)ri;
        kept = leave ? 0.0 : 1.0;
    }
    const double ws = wave_sum(kept_rho), wk = wave_sum(kept);
    if ((threadIdx.x & 63) == 0 && wk != 0.0) {
        atomicAdd(sum, ws);
        atomicAdd(sum + 1, wk);
    }
}

// changed[0] += number of i with a[i] != b[i]
__global__ void k_count_changed(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint32_t n,
                                uint32_t* changed)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t d = (i < n && a[i] != b[i]) ? 1u : 0u;
    const uint32_t s = wave_sum_u(d);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(changed, s);
}

// class flags of the active-first order: f0 = active, f1 = active neighbor
// (entries n of both are the scan sentinels)
__global__ void k_tag_classes(const uint32_t* __restrict__ tag, uint32_t n, uint32_t* __restrict__ f0,
                              uint32_t* __restrict__ f1)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { f0[n] = 0; f1[n] = 0; return; }
    const uint32_t t = tag[i];
    f0[i] = t == kTagActive ? 1u : 0u;
    f1[i] = t == kTagActiveNeighbors ? 1u : 0u;
}

// Stable three-way partition: active, then active neighbors, then the rest,
// each in its current order (the prefix order restriction mode 2 needs; the
// reference gets it from a radix sort by tag, DynamicWindow.cu:2990-2998).
__global__ void k_partition3(uint32_t n, const uint32_t* __restrict__ f0, const uint32_t* __restrict__ r0,
                             const uint32_t* __restrict__ f1, const uint32_t* __restrict__ r1, FluidPtrs in,
                             FluidPtrs out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t n0 = r0[n], n1 = r1[n];
    uint32_t o;
    if (f0[i]) o = r0[i];
    else if (f1[i]) o = n0 + r1[i];
    else o = n0 + n1 + (i - r0[i] - r1[i]);
    out.pos[o] = in.pos[i];
    out.vel[o] = in.vel[i];
    out.kap[o] = in.kap[i];
    out.kapv[o] = in.kapv[i];
    out.ktot[o] = in.ktot[i];
    out.kvtot[o] = in.kvtot[i];
    out.id[o] = in.id[i];
    out.tag[o] = in.tag[i];
}

// keep[i] = tag[i] != t; keep[n] = 0
__global__ void k_keep_not_tag(const uint32_t* __restrict__ tag, uint32_t n, uint32_t t, uint32_t* __restrict__ keep)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    keep[i] = (i < n && tag[i] != t) ? 1u : 0u;
}

// ---------------------------------------------------------------- launchers
static inline uint32_t nb_(uint32_t n) { return (n + kBlock - 1) / kBlock; }

void launch_surface_keep(hipStream_t s, const float4* p, uint32_t n, const Surface& S, int want_inside, uint32_t* keep)
{
    k_surface_keep<<<nb_(n + 1), kBlock, 0, s>>>(p, n, S, want_inside, keep);
}

void launch_surface_keep_and(hipStream_t s, const float4* p, uint32_t n, const Surface& S, int want_inside,
                             uint32_t* keep)
{
    if (n) k_surface_keep_and<<<nb_(n), kBlock, 0, s>>>(p, n, S, want_inside, keep);
}

void launch_compact_f4(hipStream_t s, uint32_t n, const uint32_t* keep, const uint32_t* rank, const float4* in,
                       float4* out)
{
    if (n) k_compact_f4<<<nb_(n), kBlock, 0, s>>>(n, keep, rank, in, out);
}

void launch_outflow_keep(hipStream_t s, const float4* p, uint32_t n, const Surface& interior, const Surface& surface,
                         uint32_t* keep)
{
    k_outflow_keep<<<nb_(n + 1), kBlock, 0, s>>>(p, n, interior, surface, keep);
}

void launch_inflow(hipStream_t s, const SimConst& c, const float4* cand, uint32_t nc, const float4* pos,
                   const float4* vel, const uint32_t* cs_f, const uint32_t* perm, const float4* bpos,
                   const uint32_t* cs_b, const Surface& boundary, float min_dist, float max_density, float mass,
                   uint32_t* keep, float4* vel_out)
{
    k_inflow<<<nb_(nc + 1), kBlock, 0, s>>>(c, cand, nc, pos, vel, cs_f, perm, bpos, cs_b, boundary, min_dist,
                                            max_density, mass, keep, vel_out);
}

void launch_append_inflow(hipStream_t s, const float4* cand, uint32_t nc, const uint32_t* keep, const uint32_t* rank,
                          const float4* vnew, uint32_t n, float mass, uint32_t first_id, const FluidPtrs& f)
{
    if (nc) k_append_inflow<<<nb_(nc), kBlock, 0, s>>>(cand, nc, keep, rank, vnew, n, mass, first_id, f);
}

void launch_apply_delta(hipStream_t s, float4* p, uint32_t n, float dx, float dy, float dz)
{
    if (n) k_apply_delta<<<nb_(n), kBlock, 0, s>>>(p, n, dx, dy, dz);
}

void launch_count_tag(hipStream_t s, const uint32_t* tag, uint32_t n, uint32_t t, uint32_t* count)
{
    if (n) k_count_tag<<<nb_(n), kBlock, 0, s>>>(tag, n, t, count);
}

void launch_tag_neighbors(hipStream_t s, const SimConst& c, uint32_t n, const NbrLists& L, const uint32_t* src,
                          uint32_t* dst, uint32_t tag_in, uint32_t tag_out, int untagged_only, uint32_t grid)
{
    if (n) k_tag_neighbors<<<grid, kBlock, 0, s>>>(c, n, L, src, dst, tag_in, tag_out, untagged_only);
}

void launch_rest_extract(hipStream_t s, const SimConst& c, uint32_t n, const NbrLists& L, const float4* pos,
                         const float4* bpos, const float* rho, const uint32_t* tag, float limit, float boundary_limit,
                         uint32_t* keep, double* sum, uint32_t grid)
{
    if (n) k_rest_extract<<<grid, kBlock, 0, s>>>(c, n, L, pos, bpos, rho, tag, limit, boundary_limit, keep, sum);
}

void launch_count_changed(hipStream_t s, const uint32_t* a, const uint32_t* b, uint32_t n, uint32_t* changed)
{
    if (n) k_count_changed<<<nb_(n), kBlock, 0, s>>>(a, b, n, changed);
}

void launch_tag_classes(hipStream_t s, const uint32_t* tag, uint32_t n, uint32_t* f0, uint32_t* f1)
{
    k_tag_classes<<<nb_(n + 1), kBlock, 0, s>>>(tag, n, f0, f1);
}

void launch_partition3(hipStream_t s, uint32_t n, const uint32_t* f0, const uint32_t* r0, const uint32_t* f1,
                       const uint32_t* r1, const FluidPtrs& in, const FluidPtrs& out)
{
    if (n) k_partition3<<<nb_(n), kBlock, 0, s>>>(n, f0, r0, f1, r1, in, out);
}

void launch_keep_not_tag(hipStream_t s, const uint32_t* tag, uint32_t n, uint32_t t, uint32_t* keep)
{
    k_keep_not_tag<<<nb_(n + 1), kBlock, 0, s>>>(tag, n, t, keep);
}

}  // namespace dfsph
