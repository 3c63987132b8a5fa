// dfsph_bodies.hip -- rigid bodies as moving boundary particles (gfx950).
//
// The reference keeps each dynamic body as its own particle set with a
// third neighbor list (vector_dynamic_bodies_data, ITER_NEIGHBORS_SOLIDS,
// DFSPH_macro_cuda.h:265-274).  Here body particles join the boundary set:
// every step they are placed from the body state and sorted into the cell
// table together with the static walls, so the neighbor search and the
// sweeps need no third list; the sweeps read the boundary neighbor's
// velocity and add the reaction of each pressure update to the body particle
// (64-bit fixed point, see Solids).  At the position update the reactions
// are summed into the body's force and torque and the body is integrated,
// which replaces the external PositionBasedDynamics step of the reference's
// demo (Demos/DynamicBoundaryDemo/PositionBasedDynamicsWrapper).
#include <hip/hip_runtime.h>

#include "dfsph_launch.hpp"

namespace dfsph {

namespace {

inline uint32_t nblk(uint32_t n) { return (n + kBlock - 1) / kBlock; }

// R(q) for a unit quaternion (x, y, z, w)
__device__ __forceinline__ void quat_mat(const double* q, double R[3][3])
{
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    R[0][0] = 1.0 - 2.0 * (y * y + z * z);
    R[0][1] = 2.0 * (x * y - z * w);
    R[0][2] = 2.0 * (x * z + y * w);
    R[1][0] = 2.0 * (x * y + z * w);
    R[1][1] = 1.0 - 2.0 * (x * x + z * z);
    R[1][2] = 2.0 * (y * z - x * w);
    R[2][0] = 2.0 * (x * z - y * w);
    R[2][1] = 2.0 * (y * z + x * w);
    R[2][2] = 1.0 - 2.0 * (x * x + y * y);
}

}  // namespace

// Body particles at the body state: x = R(q) x0 + c, v = w x (x - c) + v_c, and
// the reaction accumulators cleared (DFSPH_updateDynamicObjectParticles_kernel,
// DFSPH_core_cuda.cu:3407-3420).  FP64 without contraction, so the CPU check
// reproduces the same floats.
__global__ void k_bodies_place(const DevBody* __restrict__ B, const float4* __restrict__ x0,
                               const uint32_t* __restrict__ sbody, uint32_t ns, float4* __restrict__ bpos,
                               float4* __restrict__ bvel, unsigned long long* __restrict__ force)
{
#pragma clang fp contract(off)
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const DevBody& b = B[sbody[s]];
    double R[3][3];
    quat_mat(b.q, R);
    const float4 p0 = x0[s];
    const double l[3] = {p0.x, p0.y, p0.z};
    double x[3], r[3];
    for (int a = 0; a < 3; ++a) {
        r[a] = R[a][0] * l[0] + R[a][1] * l[1] + R[a][2] * l[2];
        x[a] = r[a] + b.c[a];
    }
    const double vx = b.w[1] * r[2] - b.w[2] * r[1] + b.v[0];
    const double vy = b.w[2] * r[0] - b.w[0] * r[2] + b.v[1];
    const double vz = b.w[0] * r[1] - b.w[1] * r[0] + b.v[2];
    bpos[s] = make_float4((float)x[0], (float)x[1], (float)x[2], p0.w);
    bvel[s] = make_float4((float)vx, (float)vy, (float)vz, __uint_as_float(s));
    force[4ull * s] = 0ull;
    force[4ull * s + 1] = 0ull;
    force[4ull * s + 2] = 0ull;
}

// One block per body: the fluid force (sum of the particle reactions, exact in
// fixed point) and torque about the center of mass (compute_fluid_impact_on_
// dynamic_body_kernel, DFSPH_memory_management_cuda.cu:688-705), then one
// step of the body over h:
//   dynamic    semi-implicit Euler: v += h (F / m + g); w += h I^-1 (T - w x I w)
//              with I = R diag(I_body) R^T; c += h v; q += h/2 (w, 0) q, normalised
//   sine       c = c0 + amp sin(omega t) axis, v = amp omega cos(omega t) axis at
//              the body's time after the step (its clock sums the steps' h; the
//              solver's time adds each step's successor instead,
//              DFSPH_CUDA.cpp:441-452)
//   kinematic  state left to the caller
__global__ __launch_bounds__(kBlock) void k_bodies_integrate(DevBody* __restrict__ B,
                                                             const float4* __restrict__ spos,
                                                             const unsigned long long* __restrict__ force,
                                                             double inv_fscale, const DevTime* __restrict__ t,
                                                             double gx, double gy, double gz)
{
#pragma clang fp contract(off)
    __shared__ long long sF[3][kBlock];
    __shared__ double sT[3][kBlock];
    DevBody& b = B[blockIdx.x];
    long long fx = 0, fy = 0, fz = 0;
    double tx = 0.0, ty = 0.0, tz = 0.0;
    for (uint32_t k = threadIdx.x; k < b.count; k += kBlock) {
        const uint32_t s = b.first + k;
        const long long ix = (long long)force[4ull * s], iy = (long long)force[4ull * s + 1],
                        iz = (long long)force[4ull * s + 2];
        fx += ix; fy += iy; fz += iz;
        const float4 p = spos[s];
        const double rx = (double)p.x - b.c[0], ry = (double)p.y - b.c[1], rz = (double)p.z - b.c[2];
        const double ex = (double)ix * inv_fscale, ey = (double)iy * inv_fscale, ez = (double)iz * inv_fscale;
        tx += ry * ez - rz * ey;
        ty += rz * ex - rx * ez;
        tz += rx * ey - ry * ex;
    }
    sF[0][threadIdx.x] = fx; sF[1][threadIdx.x] = fy; sF[2][threadIdx.x] = fz;
    sT[0][threadIdx.x] = tx; sT[1][threadIdx.x] = ty; sT[2][threadIdx.x] = tz;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int a = 0; a < 3; ++a) {
                sF[a][threadIdx.x] += sF[a][threadIdx.x + w];
                sT[a][threadIdx.x] += sT[a][threadIdx.x + w];
            }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double F[3] = {(double)sF[0][0] * inv_fscale, (double)sF[1][0] * inv_fscale, (double)sF[2][0] * inv_fscale};
    const double T[3] = {sT[0][0], sT[1][0], sT[2][0]};
    for (int a = 0; a < 3; ++a) { b.force[a] = F[a]; b.torque[a] = T[a]; }
    const double h = t->hd;
    b.clock += h;
    if (b.type == kBodySine) {
        const double tn = b.clock;
        const double sn = sin(b.omega * tn), cs = cos(b.omega * tn);
        for (int a = 0; a < 3; ++a) {
            b.c[a] = b.c0[a] + b.amp * sn * b.axis[a];
            b.v[a] = b.amp * b.omega * cs * b.axis[a];
            b.w[a] = 0.0;
        }
        return;
    }
    if (b.type != kBodyDynamic) return;
    const double g[3] = {gx, gy, gz};
    for (int a = 0; a < 3; ++a) b.v[a] += h * (F[a] / b.mass + g[a]);
    double R[3][3];
    quat_mat(b.q, R);
    // body-frame angular velocity and torque: wb = R^T w
    double wb[3], Tb[3];
    for (int a = 0; a < 3; ++a) {
        wb[a] = R[0][a] * b.w[0] + R[1][a] * b.w[1] + R[2][a] * b.w[2];
        Tb[a] = R[0][a] * T[0] + R[1][a] * T[1] + R[2][a] * T[2];
    }
    const double Lb[3] = {b.inertia[0] * wb[0], b.inertia[1] * wb[1], b.inertia[2] * wb[2]};
    const double gyro[3] = {wb[1] * Lb[2] - wb[2] * Lb[1], wb[2] * Lb[0] - wb[0] * Lb[2], wb[0] * Lb[1] - wb[1] * Lb[0]};
    for (int a = 0; a < 3; ++a) wb[a] += h * (Tb[a] - gyro[a]) / b.inertia[a];
    for (int a = 0; a < 3; ++a) b.w[a] = R[a][0] * wb[0] + R[a][1] * wb[1] + R[a][2] * wb[2];
    for (int a = 0; a < 3; ++a) b.c[a] += h * b.v[a];
    const double* q = b.q;
    const double* w = b.w;
    // dq = (w, 0) q: scalar -w.qv, vector qw w + w x qv
    const double dx = q[3] * w[0] + (w[1] * q[2] - w[2] * q[1]);
    const double dy = q[3] * w[1] + (w[2] * q[0] - w[0] * q[2]);
    const double dz = q[3] * w[2] + (w[0] * q[1] - w[1] * q[0]);
    const double dw = -(w[0] * q[0] + w[1] * q[1] + w[2] * q[2]);
    double nq[4] = {q[0] + 0.5 * h * dx, q[1] + 0.5 * h * dy, q[2] + 0.5 * h * dz, q[3] + 0.5 * h * dw};
    const double nn = sqrt(nq[0] * nq[0] + nq[1] * nq[1] + nq[2] * nq[2] + nq[3] * nq[3]);
    for (int a = 0; a < 4; ++a) b.q[a] = nq[a] / nn;
}

// psi of the body particles in [first, end) from a psi computed over the whole
// boundary set in cell order (bvel.w names each slot's solid index)
__global__ void k_solid_psi_take(const float4* __restrict__ bvel, const float* __restrict__ psi, uint32_t nb,
                                 float4* __restrict__ x0, uint32_t first, uint32_t end)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const uint32_t s = __float_as_uint(bvel[k].w);
    if (s != kNoSolid && s >= first && s < end) x0[s].w = psi[k];
}

// velocity records of the wall particles: at rest, no solid index
__global__ void k_wall_vel(float4* bvel, uint32_t n)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) bvel[k] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(kNoSolid));
}

void launch_bodies_place(hipStream_t s, const DevBody* B, const float4* x0, const uint32_t* sbody, uint32_t ns,
                         float4* bpos, float4* bvel, unsigned long long* force)
{
    if (ns) k_bodies_place<<<nblk(ns), kBlock, 0, s>>>(B, x0, sbody, ns, bpos, bvel, force);
}

void launch_bodies_integrate(hipStream_t s, DevBody* B, uint32_t nbody, const float4* spos,
                             const unsigned long long* force, double inv_fscale, const DevTime* t, const double* g)
{
    if (nbody) k_bodies_integrate<<<nbody, kBlock, 0, s>>>(B, spos, force, inv_fscale, t, g[0], g[1], g[2]);
}

void launch_solid_psi_take(hipStream_t s, const float4* bvel, const float* psi, uint32_t nb, float4* x0,
                           uint32_t first, uint32_t end)
{
    if (nb) k_solid_psi_take<<<nblk(nb), kBlock, 0, s>>>(bvel, psi, nb, x0, first, end);
}

void launch_wall_vel(hipStream_t s, float4* bvel, uint32_t n)
{
    if (n) k_wall_vel<<<nblk(n), kBlock, 0, s>>>(bvel, n);
}

}  // namespace dfsph
