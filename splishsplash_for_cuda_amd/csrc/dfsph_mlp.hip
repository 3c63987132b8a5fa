// dfsph_mlp.hip -- learned warm start of the two pressure solvers (no
// reference counterpart; the warm starts it replaces are the kappaV rescale of
// DFSPH_core_cuda.cu:240 and the kappa rescale of :3063-3068).
//
// A per-particle MLP 12 -> 64 -> 64 -> 64 -> 2 with SiLU activations predicts
// the converged (kappaV h, kappa h_f^2) of a step, the dt-invariant units of
// the reference CPU path (TimeStepDFSPH.cpp:458-459, 659-660), as a correction
// to the reference warm start formed from the previous step's totals
// (ws_base): the network's outputs are the totals less that warm start, and
// the solver's prediction adds it back.  Weights are
// BF16 on the matrix cores (v_mfma_f32_16x16x16_bf16, FP32 accumulate);
// biases, activations between layers and all reductions are FP32.
//
// Wave-level tiling: one wavefront owns 64 particles (four 16-row tiles).
// MFMA 16x16x16 operand layout (lane l = 0..63, q = 0..3):
//   A[i][k]: lane l holds A[l % 16][4 (l / 16) + q]
//   B[k][n]: lane l holds B[4 (l / 16) + q][l % 16]
//   D[i][n]: lane l holds D[4 (l / 16) + q][l % 16]
// Activations pass between layers through LDS (bf16), which turns the D
// layout of one layer into the A layout of the next.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dfsph_hip.h"
#include "dfsph_launch.hpp"
#include "dfsph_mlp_api.hpp"

namespace dfsph {

constexpr int kIn = 12, kH = 64, kOut = 2;
constexpr int kInPad = 16, kOutPad = 16;
// parameter offsets of the flat FP32 layout shared with the oracle:
// W1[64x12] b1[64] W2[64x64] b2[64] W3[64x64] b3[64] W4[2x64] b4[2]
constexpr int oW1 = 0, ob1 = oW1 + kH * kIn, oW2 = ob1 + kH, ob2 = oW2 + kH * kH, oW3 = ob2 + kH,
              ob3 = oW3 + kH * kH, oW4 = ob3 + kH, ob4 = oW4 + kOut * kH, kNParam = ob4 + kOut;
constexpr int kNNorm = 2 * kIn + 2 * kOut;   // in_mu, in_sig, out_mu, out_sig

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f)
{
    // round to nearest even (finite inputs)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float(((uint32_t)b) << 16); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }

__device__ __forceinline__ float dsilu(float z)
{
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// General strided operand access: X(r, c) = base[r * rs + c * cs] (bf16 in LDS or global)
struct Mat {
    const unsigned short* p;
    int rs, cs;
    __device__ unsigned short at(int r, int c) const { return p[r * rs + c * cs]; }
};

// One 16x16 output tile C[mt*16.., nt*16..] = sum_k A(i, k) B(k, n) over K (multiple of 16).
__device__ __forceinline__ f32x4 gemm_tile(const Mat& A, const Mat& B, int mt, int nt, int K, f32x4 acc)
{
    const int l = threadIdx.x & 63;
    const int i = mt * 16 + (l & 15), n = nt * 16 + (l & 15);
    for (int k0 = 0; k0 < K; k0 += 16) {
        const int kb = k0 + 4 * (l >> 4);
        bf16x4 a, b;
        a[0] = (short)A.at(i, kb); a[1] = (short)A.at(i, kb + 1); a[2] = (short)A.at(i, kb + 2); a[3] = (short)A.at(i, kb + 3);
        b[0] = (short)B.at(kb, n); b[1] = (short)B.at(kb + 1, n); b[2] = (short)B.at(kb + 2, n); b[3] = (short)B.at(kb + 3, n);
        acc = mfma(a, b, acc);
    }
    return acc;
}

// Device copy of the model: padded BF16 weight matrices for the matrix cores
// and FP32 biases / normalisation.
struct DevModel {
    unsigned short W1[kH * kInPad];      // [64][16], columns 12..15 zero
    unsigned short W2[kH * kH];
    unsigned short W3[kH * kH];
    unsigned short W4[kOutPad * kH];     // [16][64], rows 2..15 zero
    float b1[kH], b2[kH], b3[kH], b4[kOutPad];
    float norm[kNNorm];
};

// Forward pass of one wavefront over 64 rows of X0 (bf16, [64][16] in LDS).
// Writes pre-activations Z1..Z3 (FP32, may be null) and activations A1..A3
// (bf16) to LDS, returns the two outputs of the lane's rows in y (D layout:
// rows 4 (l/16) + q of the 16-row tiles, column l % 16 < 2 valid).
struct WaveBufs {
    unsigned short* X0;   // [64][16]
    unsigned short* A1;   // [64][64]
    unsigned short* A2;
    unsigned short* A3;
    float* Z1;            // [64][64] or null (inference)
    float* Z2;
    float* Z3;
};

__device__ void layer(const Mat& Ain, int K, const unsigned short* W, int ldw, const float* bias, float* Zout,
                      unsigned short* Aout)
{
    const int l = threadIdx.x & 63;
    const Mat B{W, 1, ldw};   // B(k, n) = W[n][k]
    for (int mt = 0; mt < 4; ++mt)
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = gemm_tile(Ain, B, mt, nt, K, acc);
            const int n = nt * 16 + (l & 15);
            for (int q = 0; q < 4; ++q) {
                const int r = mt * 16 + 4 * (l >> 4) + q;
                const float z = acc[q] + bias[n];
                if (Zout) Zout[r * kH + n] = z;
                Aout[r * kH + n] = f2bf(silu(z));
            }
        }
}

__device__ void forward_wave(const DevModel& M, const WaveBufs& w, float out[4][4])
{
    const int l = threadIdx.x & 63;
    layer(Mat{w.X0, kInPad, 1}, kInPad, M.W1, kInPad, M.b1, w.Z1, w.A1);
    __syncthreads();
    layer(Mat{w.A1, kH, 1}, kH, M.W2, kH, M.b2, w.Z2, w.A2);
    __syncthreads();
    layer(Mat{w.A2, kH, 1}, kH, M.W3, kH, M.b3, w.Z3, w.A3);
    __syncthreads();
    const Mat B{M.W4, 1, kH};
    for (int mt = 0; mt < 4; ++mt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = gemm_tile(Mat{w.A3, kH, 1}, B, mt, 0, kH, acc);
        for (int q = 0; q < 4; ++q) out[mt][q] = acc[q] + M.b4[l & 15];
    }
}

// ---------------------------------------------------------------- inference
// x: [n][12] raw features (FP32); output o of row r goes to y[r * ystride + o * ocol]
// (denormalised when normalise is set: the solver's corrections).
//
// Register-resident form: each layer is computed transposed, H_out^T = W H_in^T,
// with W as the MFMA A operand and the particles as the N dimension.  The
// 16x16 output tile a lane holds (out features 4 (l/16) + q of 16 particles)
// is exactly the B operand layout of the next layer's matching K block, so
// activations never leave registers.  The weights' A fragments are loaded
// once into registers and each wave runs over 16-particle tiles; biases come
// from LDS as accumulator seeds.  Launch bounds ask for 3 waves per SIMD (168
// VGPRs, no spills; the default allocation held 2, and 4 spill).
constexpr int kInferBlock = 256;

__device__ __forceinline__ bf16x4 frag8(const unsigned short* p)     // 4 consecutive bf16 (8-byte aligned)
{
    const uint2 u = *reinterpret_cast<const uint2*>(p);
    bf16x4 r;
    r[0] = (short)(u.x & 0xffffu); r[1] = (short)(u.x >> 16);
    r[2] = (short)(u.y & 0xffffu); r[3] = (short)(u.y >> 16);
    return r;
}

// Inference activation: SiLU with the hardware reciprocal (z / (1 + e^-z) to
// about 1 ulp; the IEEE division took three quarters of the kernel's vector
// instructions), rounded to bf16 two at a time by v_cvt_pk_bf16_f32 (round
// to nearest even, as f2bf).
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

__device__ __forceinline__ bf16x4 pack_bf16(float a, float b, float c, float d)
{
    struct {
        bf16x2 lo, hi;
    } p{__builtin_convertvector(f32x2{a, b}, bf16x2), __builtin_convertvector(f32x2{c, d}, bf16x2)};
    return __builtin_bit_cast(bf16x4, p);
}

__device__ __forceinline__ bf16x4 act_frag(const f32x4& z)
{
    return pack_bf16(silu_fast(z[0]), silu_fast(z[1]), silu_fast(z[2]), silu_fast(z[3]));
}

__global__ __launch_bounds__(kInferBlock, 3) void k_mlp_infer(const DevModel* __restrict__ Mg, const float* __restrict__ x,
                                                           uint32_t n, float* __restrict__ y, int normalise,
                                                           int ystride, int ocol)
{
    __shared__ __align__(16) float sb[3 * kH + kOutPad];
    const DevModel& M = *Mg;
    for (int k = threadIdx.x; k < 3 * kH + kOutPad; k += blockDim.x)
        sb[k] = k < kH ? M.b1[k] : k < 2 * kH ? M.b2[k - kH] : k < 3 * kH ? M.b3[k - 2 * kH] : M.b4[k - 3 * kH];
    __syncthreads();
    const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;   // particle column, 4-row group
    bf16x4 w1[4], w2[4][4], w3[4][4], w4[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        w1[mt] = frag8(M.W1 + (mt * 16 + r) * kInPad + 4 * g);
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            w2[mt][kt] = frag8(M.W2 + (mt * 16 + r) * kH + kt * 16 + 4 * g);
            w3[mt][kt] = frag8(M.W3 + (mt * 16 + r) * kH + kt * 16 + 4 * g);
        }
        w4[mt] = frag8(M.W4 + r * kH + mt * 16 + 4 * g);      // K block mt of the (padded) output layer
    }
    float mu[4], isg[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int f = 4 * g + j;
        mu[j] = (normalise && f < kIn) ? M.norm[f] : 0.0f;
        isg[j] = (normalise && f < kIn) ? 1.0f / M.norm[kIn + f] : 1.0f;
    }
    const uint32_t tiles = (n + 15) / 16;
    const uint32_t waves = gridDim.x * (blockDim.x / 64);
    for (uint32_t t = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); t < tiles; t += waves) {
        const uint32_t p = t * 16 + r;
        // layer 1 input: features 4g..4g+3 of particle p (12..15 zero)
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (p < n && g < 3) v = *reinterpret_cast<const float4*>(x + (size_t)p * kIn + 4 * g);
        bf16x4 h[4];
        {
            bf16x4 xb = pack_bf16((v.x - mu[0]) * isg[0], (v.y - mu[1]) * isg[1], (v.z - mu[2]) * isg[2],
                                  (v.w - mu[3]) * isg[3]);
            if (g == 3) xb = bf16x4{0, 0, 0, 0};
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                const float4 b = *reinterpret_cast<const float4*>(sb + mt * 16 + 4 * g);
                h[mt] = act_frag(mfma(w1[mt], xb, f32x4{b.x, b.y, b.z, b.w}));
            }
        }
#pragma unroll
        for (int layer = 0; layer < 2; ++layer) {
            bf16x4 o[4];
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                const float4 b = *reinterpret_cast<const float4*>(sb + (layer + 1) * kH + mt * 16 + 4 * g);
                f32x4 acc{b.x, b.y, b.z, b.w};
#pragma unroll
                for (int kt = 0; kt < 4; ++kt) acc = mfma(layer == 0 ? w2[mt][kt] : w3[mt][kt], h[kt], acc);
                o[mt] = act_frag(acc);
            }
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) h[mt] = o[mt];
        }
        const float4 b = *reinterpret_cast<const float4*>(sb + 3 * kH + 4 * g);
        f32x4 acc{b.x, b.y, b.z, b.w};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) acc = mfma(w4[kt], h[kt], acc);
        if (g == 0 && p < n) {                  // outputs 0 and 1 of particle p
#pragma unroll
            for (int o = 0; o < kOut; ++o) {
                float val = acc[o];
                if (normalise) val = val * M.norm[2 * kIn + kOut + o] + M.norm[2 * kIn + o];
                y[(size_t)p * ystride + (size_t)o * ocol] = val;
            }
        }
    }
}

static uint32_t infer_grid(uint32_t n)
{
    const uint32_t tiles = (n + 15) / 16, per_block = kInferBlock / 64;
    const uint32_t need = (tiles + per_block - 1) / per_block;
    return std::max(1u, std::min(need, 2048u));
}

// ---------------------------------------------------------------- training
// One wavefront per 64 samples: forward, loss 0.5 |y - t|^2 / n_total, backward
// through the four layers; the per-wave parameter gradient (flat FP32 layout)
// goes to partial[blockIdx.x][kNParam] and the wave's loss to loss[blockIdx.x].
__global__ __launch_bounds__(64) void k_mlp_grad(const DevModel* __restrict__ Mg, const float* __restrict__ x,
                                                  const float* __restrict__ t, const uint32_t* __restrict__ idx,
                                                  uint32_t n, float inv_total, int normalise,
                                                  float* __restrict__ partial, float* __restrict__ loss)
{
    __shared__ unsigned short X0[64 * kInPad];
    __shared__ unsigned short A1[64 * kH], A2[64 * kH], A3[64 * kH];
    __shared__ float Z1[64 * kH], Z2[64 * kH], Z3[64 * kH];
    __shared__ unsigned short D[64 * kH];      // dZ (bf16) of the current layer
    __shared__ unsigned short DY[64 * kOutPad];
    __shared__ float lsum[64];
    const DevModel& M = *Mg;
    const int l = threadIdx.x;
    const uint32_t base = blockIdx.x * 64u;
    const uint32_t row = base + l;
    const bool valid = row < n;
    const uint32_t src = valid ? (idx ? idx[row] : row) : 0u;
    for (int k = 0; k < kInPad; ++k) {
        float v = 0.0f;
        if (k < kIn && valid) {
            v = x[(size_t)src * kIn + k];
            if (normalise) v = (v - M.norm[k]) / M.norm[kIn + k];
        }
        X0[l * kInPad + k] = f2bf(v);
    }
    __syncthreads();
    WaveBufs w{X0, A1, A2, A3, Z1, Z2, Z3};
    float out[4][4];
    forward_wave(M, w, out);
    // dY = (y - t) / n_total for the valid rows (columns 0, 1), zero elsewhere
    const int o = l & 15;
    float lacc = 0.0f;
    for (int mt = 0; mt < 4; ++mt)
        for (int q = 0; q < 4; ++q) {
            const int r = mt * 16 + 4 * (l >> 4) + q;
            float d = 0.0f;
            if (o < kOut && base + r < n) {
                const uint32_t s = idx ? idx[base + r] : base + r;
                float tv = t[(size_t)s * kOut + o];
                if (normalise) tv = (tv - M.norm[2 * kIn + o]) / M.norm[2 * kIn + kOut + o];
                const float e = out[mt][q] - tv;
                lacc += 0.5f * e * e * inv_total;
                d = e * inv_total;
            }
            DY[r * kOutPad + o] = f2bf(d);
        }
    lsum[l] = lacc;
    __syncthreads();
    if (l == 0) {
        float s = 0.0f;
        for (int k = 0; k < 64; ++k) s += lsum[k];
        loss[blockIdx.x] = s;
    }
    float* g = partial + (size_t)blockIdx.x * kNParam;
    // bias gradients are column sums of dZ: done in FP32 from the bf16 dZ rows
    auto col_sum_store = [&](const unsigned short* Dm, int ld, int ncol, float* dst) {
        for (int c = l; c < ncol; c += 64) {
            float s = 0.0f;
            for (int r = 0; r < 64; ++r) s += bf2f(Dm[r * ld + c]);
            dst[c] = s;
        }
    };
    // weight gradient dW[o][k] = sum_s dZ[s][o] Aprev[s][k]: M = out, N = in, K = samples
    auto weight_grad = [&](const unsigned short* Dm, int ldd, int nout_tiles, int nout, const unsigned short* Ap,
                           int lda, int nin_tiles, int nin, float* dst) {
        const Mat At{Dm, 1, ldd};     // A(i = out, k = sample) = dZ[sample][out]
        const Mat Bs{Ap, lda, 1};     // B(k = sample, n = in) = Aprev[sample][in]
        for (int mt = 0; mt < nout_tiles; ++mt)
            for (int nt = 0; nt < nin_tiles; ++nt) {
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                acc = gemm_tile(At, Bs, mt, nt, 64, acc);
                const int ci = nt * 16 + (l & 15);
                for (int q = 0; q < 4; ++q) {
                    const int ro = mt * 16 + 4 * (l >> 4) + q;
                    if (ro < nout && ci < nin) dst[ro * nin + ci] = acc[q];
                }
            }
    };
    // input gradient dA[s][k] = sum_o dZ[s][o] W[o][k]: M = samples, N = in, K = out
    auto input_grad_to_dz = [&](const unsigned short* Dm, int ldd, int kout, const unsigned short* W, int ldw,
                                const float* Zprev, unsigned short* Dout) {
        const Mat A{Dm, ldd, 1};      // A(i = sample, k = out)
        const Mat B{W, ldw, 1};       // B(k = out, n = in) = W[out][in]
        for (int mt = 0; mt < 4; ++mt)
            for (int nt = 0; nt < 4; ++nt) {
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                acc = gemm_tile(A, B, mt, nt, kout, acc);
                const int c = nt * 16 + (l & 15);
                for (int q = 0; q < 4; ++q) {
                    const int r = mt * 16 + 4 * (l >> 4) + q;
                    Dout[r * kH + c] = f2bf(acc[q] * dsilu(Zprev[r * kH + c]));
                }
            }
    };
    // layer 4
    weight_grad(DY, kOutPad, 1, kOut, A3, kH, 4, kH, g + oW4);
    if (l < kOut) {
        float s = 0.0f;
        for (int r = 0; r < 64; ++r) s += bf2f(DY[r * kOutPad + l]);
        g[ob4 + l] = s;
    }
    input_grad_to_dz(DY, kOutPad, kOutPad, M.W4, kH, Z3, D);
    __syncthreads();
    // layer 3
    weight_grad(D, kH, 4, kH, A2, kH, 4, kH, g + oW3);
    col_sum_store(D, kH, kH, g + ob3);
    __syncthreads();
    {
        // dZ2 needs D (= dZ3) as input while D is overwritten: stage through A3
        for (int e = l; e < 64 * kH; e += 64) A3[e] = D[e];
        __syncthreads();
        input_grad_to_dz(A3, kH, kH, M.W3, kH, Z2, D);
        __syncthreads();
    }
    // layer 2
    weight_grad(D, kH, 4, kH, A1, kH, 4, kH, g + oW2);
    col_sum_store(D, kH, kH, g + ob2);
    __syncthreads();
    {
        for (int e = l; e < 64 * kH; e += 64) A3[e] = D[e];
        __syncthreads();
        input_grad_to_dz(A3, kH, kH, M.W2, kH, Z1, D);
        __syncthreads();
    }
    // layer 1 (inputs: 12 used columns of X0)
    weight_grad(D, kH, 4, kH, X0, kInPad, 1, kIn, g + oW1);
    col_sum_store(D, kH, kH, g + ob1);
}

// fixed-order sum of the per-wave partial gradients and losses
__global__ void k_mlp_reduce(const float* __restrict__ partial, const float* __restrict__ loss, uint32_t nparts,
                             float* __restrict__ grad, float* __restrict__ loss_out)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < (uint32_t)kNParam) {
        float s = 0.0f;
        for (uint32_t b = 0; b < nparts; ++b) s += partial[(size_t)b * kNParam + p];
        grad[p] = s;
    }
    if (p == 0) {
        float s = 0.0f;
        for (uint32_t b = 0; b < nparts; ++b) s += loss[b];
        *loss_out = s;
    }
}

// Adam (Kingma & Ba 2015) on the FP32 master weights, then refresh the BF16 copies
__global__ void k_mlp_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)kNParam) return;
    const float gi = g[p];
    const float mi = b1 * m[p] + (1.0f - b1) * gi;
    const float vi = b2 * v[p] + (1.0f - b2) * gi * gi;
    m[p] = mi;
    v[p] = vi;
    w[p] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

__global__ void k_mlp_pack(const float* __restrict__ w, const float* __restrict__ norm, DevModel* __restrict__ M)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (uint32_t)(kH * kInPad)) {
        const int o = t / kInPad, k = t % kInPad;
        M->W1[t] = k < kIn ? f2bf(w[oW1 + o * kIn + k]) : 0;
    }
    if (t < (uint32_t)(kH * kH)) {
        M->W2[t] = f2bf(w[oW2 + t]);
        M->W3[t] = f2bf(w[oW3 + t]);
    }
    if (t < (uint32_t)(kOutPad * kH)) {
        const int o = t / kH;
        M->W4[t] = o < kOut ? f2bf(w[oW4 + t]) : 0;
    }
    if (t < (uint32_t)kH) {
        M->b1[t] = w[ob1 + t];
        M->b2[t] = w[ob2 + t];
        M->b3[t] = w[ob3 + t];
    }
    if (t < (uint32_t)kOutPad) M->b4[t] = t < (uint32_t)kOut ? w[ob4 + t] : 0.0f;
    if (t < (uint32_t)kNNorm) M->norm[t] = norm[t];
}

// ------------------------------------------------------------- model state

struct MlpState {
    DevModel* model = nullptr;
    float *w = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *norm = nullptr;
    float* partial = nullptr;
    float* lossp = nullptr;
    float* loss = nullptr;
    uint32_t partial_cap = 0;
    uint32_t adam_t = 0;
    bool loaded = false;
    // per-particle buffers of the solver integration
    float* feat = nullptr;      // [cap][12]
    float* pred = nullptr;      // [4][cap]: corrections (kappaV h, kappa h_f^2), then the last
                                // step's converged totals in the same units (the sweeps' bound)
    float* kvws = nullptr;      // predicted kappaV warm start (GPU units)
    uint32_t cap = 0;
    // training set
    float* ds_x = nullptr;
    float* ds_t = nullptr;
    uint32_t ds_cap = 0, ds_n = 0, ds_stride = 1;
    uint32_t* idx = nullptr;
    uint32_t idx_cap = 0;
    ~MlpState()
    {
        void* ps[] = {model, w, m, v, grad, norm, partial, lossp, loss, feat, pred, kvws, ds_x, ds_t, idx};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }
};

void mlp_destroy(MlpState* m) { delete m; }

}  // namespace dfsph

// ------------------------------------------------------------- C ABI part
namespace {

using dfsph::MlpState;

int mfail(int code, const std::string& msg)
{
    dfsph::handle_set_error(msg.c_str());
    return code;
}

#define MHIP(x)                                                                                       \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) return mfail(DFSPH_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int malloc_dev(T*& p, size_t n)
{
    if (p) (void)hipFree(p);
    p = nullptr;
    MHIP(hipMalloc((void**)&p, (n ? n : 1) * sizeof(T)));
    return DFSPH_OK;
}

int ensure_state(void* handle, MlpState** out)
{
    dfsph::MlpView v = dfsph::handle_view(handle);
    if (!*v.mlp) {
        MlpState* m = new MlpState();
        int rc = DFSPH_OK;
        if ((rc = malloc_dev(m->model, 1)) || (rc = malloc_dev(m->w, dfsph::kNParam)) ||
            (rc = malloc_dev(m->m, dfsph::kNParam)) || (rc = malloc_dev(m->v, dfsph::kNParam)) ||
            (rc = malloc_dev(m->grad, dfsph::kNParam)) || (rc = malloc_dev(m->norm, dfsph::kNNorm)) ||
            (rc = malloc_dev(m->loss, 1))) {
            delete m;
            return rc;
        }
        *v.mlp = m;
    }
    *out = *v.mlp;
    return DFSPH_OK;
}

int pack(MlpState* m, hipStream_t s)
{
    dfsph::k_mlp_pack<<<(dfsph::kH * dfsph::kH + 255) / 256, 256, 0, s>>>(m->w, m->norm, m->model);
    MHIP(hipGetLastError());
    return DFSPH_OK;
}

}  // namespace

// Staging of dfsph_mlp_forward_noalloc: features then outputs of up to
// DFSPH_MLP_NOALLOC_MAX samples, one copy per device for the life of the
// process.  Host code uses the device address that hipGetSymbolAddress
// returns for the current device, never the array's name.
static __device__ __align__(256) float scratch[(size_t)DFSPH_MLP_NOALLOC_MAX * (dfsph::kIn + dfsph::kOut)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// the scratch address on the current device, looked up on its first use
int scratch_address(float** out)
{
    int dev = 0;
    MHIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return mfail(DFSPH_ERR_STATE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        MHIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return DFSPH_OK;
}
}  // namespace

extern "C" {

int dfsph_mlp_set_params(void* handle, const float* params, const float* norm)
{
    if (!params) return mfail(DFSPH_ERR_ARG, "null parameters");
    MlpState* m = nullptr;
    if (int rc = ensure_state(handle, &m)) return rc;
    hipStream_t s = dfsph::handle_view(handle).s;
    std::vector<float> nrm(dfsph::kNNorm);
    if (norm) std::memcpy(nrm.data(), norm, sizeof(float) * dfsph::kNNorm);
    else {
        for (int k = 0; k < dfsph::kIn; ++k) { nrm[k] = 0.0f; nrm[dfsph::kIn + k] = 1.0f; }
        for (int k = 0; k < dfsph::kOut; ++k) { nrm[2 * dfsph::kIn + k] = 0.0f; nrm[2 * dfsph::kIn + dfsph::kOut + k] = 1.0f; }
    }
    MHIP(hipMemcpyAsync(m->w, params, sizeof(float) * dfsph::kNParam, hipMemcpyHostToDevice, s));
    MHIP(hipMemcpyAsync(m->norm, nrm.data(), sizeof(float) * dfsph::kNNorm, hipMemcpyHostToDevice, s));
    MHIP(hipMemsetAsync(m->m, 0, sizeof(float) * dfsph::kNParam, s));
    MHIP(hipMemsetAsync(m->v, 0, sizeof(float) * dfsph::kNParam, s));
    m->adam_t = 0;
    if (int rc = pack(m, s)) return rc;
    MHIP(hipStreamSynchronize(s));
    m->loaded = true;
    return DFSPH_OK;
}

int dfsph_mlp_get_params(void* handle, float* params, float* norm)
{
    MlpState* m = nullptr;
    if (int rc = ensure_state(handle, &m)) return rc;
    if (!m->loaded) return mfail(DFSPH_ERR_NO_MODEL, "no model parameters set");
    hipStream_t s = dfsph::handle_view(handle).s;
    MHIP(hipStreamSynchronize(s));
    if (params) MHIP(hipMemcpy(params, m->w, sizeof(float) * dfsph::kNParam, hipMemcpyDeviceToHost));
    if (norm) MHIP(hipMemcpy(norm, m->norm, sizeof(float) * dfsph::kNNorm, hipMemcpyDeviceToHost));
    return DFSPH_OK;
}

int dfsph_mlp_forward(void* handle, const float* x, uint32_t n, float* y, int normalise)
{
    MlpState* m = nullptr;
    if (int rc = ensure_state(handle, &m)) return rc;
    if (!m->loaded) return mfail(DFSPH_ERR_NO_MODEL, "no model parameters set");
    if (n == 0) return DFSPH_OK;
    hipStream_t s = dfsph::handle_view(handle).s;
    float *dx = nullptr, *dy = nullptr;
    int rc = malloc_dev(dx, (size_t)n * dfsph::kIn);
    if (!rc) rc = malloc_dev(dy, (size_t)n * dfsph::kOut);
    if (!rc) {
        (void)hipMemcpyAsync(dx, x, sizeof(float) * n * dfsph::kIn, hipMemcpyHostToDevice, s);
        dfsph::k_mlp_infer<<<dfsph::infer_grid(n), dfsph::kInferBlock, 0, s>>>(m->model, dx, n, dy, normalise,
                                                                                  dfsph::kOut, 1);
        (void)hipMemcpyAsync(y, dy, sizeof(float) * n * dfsph::kOut, hipMemcpyDeviceToHost, s);
        if (hipStreamSynchronize(s) != hipSuccess) rc = mfail(DFSPH_ERR_HIP, "mlp forward");
    }
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    return rc;
}

int dfsph_mlp_forward_noalloc(void* handle, const float* x, uint32_t n, float* y, int normalise)
{
    if (n > DFSPH_MLP_NOALLOC_MAX)
        return mfail(DFSPH_ERR_ARG, "dfsph_mlp_forward_noalloc: n exceeds DFSPH_MLP_NOALLOC_MAX (" +
                                        std::to_string(DFSPH_MLP_NOALLOC_MAX) + ")");
    if (n && (!x || !y)) return mfail(DFSPH_ERR_ARG, "null input or output");
    const dfsph::MlpView v = dfsph::handle_view(handle);
    MlpState* m = *v.mlp;            // never created here: that would allocate
    if (!m || !m->loaded) return mfail(DFSPH_ERR_NO_MODEL, "no model parameters set");
    if (n == 0) return DFSPH_OK;
    float* base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    float* dx = base;
    float* dy = base + (size_t)DFSPH_MLP_NOALLOC_MAX * dfsph::kIn;
    MHIP(hipMemcpyAsync(dx, x, sizeof(float) * n * dfsph::kIn, hipMemcpyHostToDevice, v.s));
    dfsph::k_mlp_infer<<<dfsph::infer_grid(n), dfsph::kInferBlock, 0, v.s>>>(m->model, dx, n, dy, normalise,
                                                                              dfsph::kOut, 1);
    MHIP(hipGetLastError());
    MHIP(hipMemcpyAsync(y, dy, sizeof(float) * n * dfsph::kOut, hipMemcpyDeviceToHost, v.s));
    MHIP(hipStreamSynchronize(v.s));
    return DFSPH_OK;
}

int dfsph_mlp_loss_grad(void* handle, const float* x, const float* t, uint32_t n, float* grad, double* loss,
                        int normalise)
{
    MlpState* m = nullptr;
    if (int rc = ensure_state(handle, &m)) return rc;
    if (!m->loaded) return mfail(DFSPH_ERR_NO_MODEL, "no model parameters set");
    if (n == 0) return mfail(DFSPH_ERR_ARG, "no samples");
    hipStream_t s = dfsph::handle_view(handle).s;
    const uint32_t nb = (n + 63) / 64;
    float *dx = nullptr, *dt = nullptr, *part = nullptr, *lp = nullptr;
    int rc = malloc_dev(dx, (size_t)n * dfsph::kIn);
    if (!rc) rc = malloc_dev(dt, (size_t)n * dfsph::kOut);
    if (!rc) rc = malloc_dev(part, (size_t)nb * dfsph::kNParam);
    if (!rc) rc = malloc_dev(lp, nb);
    if (!rc) {
        (void)hipMemcpyAsync(dx, x, sizeof(float) * n * dfsph::kIn, hipMemcpyHostToDevice, s);
        (void)hipMemcpyAsync(dt, t, sizeof(float) * n * dfsph::kOut, hipMemcpyHostToDevice, s);
        dfsph::k_mlp_grad<<<nb, 64, 0, s>>>(m->model, dx, dt, nullptr, n, 1.0f / (float)n, normalise, part, lp);
        dfsph::k_mlp_reduce<<<(dfsph::kNParam + 255) / 256, 256, 0, s>>>(part, lp, nb, m->grad, m->loss);
        float l = 0.0f;
        (void)hipMemcpyAsync(grad, m->grad, sizeof(float) * dfsph::kNParam, hipMemcpyDeviceToHost, s);
        (void)hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, s);
        if (hipStreamSynchronize(s) != hipSuccess) rc = mfail(DFSPH_ERR_HIP, "mlp grad");
        if (loss) *loss = l;
    }
    void* ps[] = {dx, dt, part, lp};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    return rc;
}

}  // extern "C"

namespace dfsph {

// ------------------------------------------------------- solver integration

bool mlp_loaded(const MlpState* m) { return m && m->loaded; }
bool mlp_harvesting(const MlpState* m) { return m && m->ds_cap > 0 && m->ds_n < m->ds_cap; }
float* mlp_feat(MlpState* m) { return m->feat; }
WsLearned mlp_ws_v(const MlpState* m, float diameter)
{
    return WsLearned{m->pred, m->pred + 2 * (size_t)m->cap, diameter};
}
WsLearned mlp_ws_p(const MlpState* m, float diameter)
{
    return WsLearned{m->pred + m->cap, m->pred + 3 * (size_t)m->cap, diameter * diameter};
}
float* mlp_prev_p(MlpState* m) { return m->pred + 3 * (size_t)m->cap; }
float* mlp_prev_v(MlpState* m) { return m->pred + 2 * (size_t)m->cap; }

int mlp_ensure_particles(MlpState* m, uint32_t cap)
{
    if (!m) return mfail(DFSPH_ERR_STATE, "no learned warm start state");
    if (cap <= m->cap && m->feat) return DFSPH_OK;
    if (int rc = malloc_dev(m->feat, (size_t)cap * kIn)) return rc;
    if (int rc = malloc_dev(m->pred, (size_t)cap * 4)) return rc;
    m->cap = cap;
    return DFSPH_OK;
}

int mlp_infer_particles(MlpState* m, hipStream_t s, uint32_t n)
{
    if (n == 0) return DFSPH_OK;
    k_mlp_infer<<<infer_grid(n), kInferBlock, 0, s>>>(m->model, m->feat, n, m->pred, 1, 1, (int)m->cap);
    MHIP(hipGetLastError());
    return DFSPH_OK;
}

int mlp_harvest_step(MlpState* m, hipStream_t s, const DevTime* t, uint32_t n, const float* kapv, const float* kap,
                     float diameter)
{
    if (!mlp_harvesting(m) || n == 0) return DFSPH_OK;
    const uint32_t stride = m->ds_stride ? m->ds_stride : 1;
    launch_harvest(s, t, n, stride, m->ds_n, m->ds_cap, m->feat, kapv, kap, diameter, m->ds_x, m->ds_t);
    MHIP(hipGetLastError());
    const uint32_t add = (n + stride - 1) / stride;
    m->ds_n = m->ds_n + add > m->ds_cap ? m->ds_cap : m->ds_n + add;
    return DFSPH_OK;
}

// sample indices of one training batch: a counter-based hash (splitmix64)
__global__ void k_batch_indices(uint32_t* idx, uint32_t batch, uint32_t n, uint64_t seed)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    uint64_t z = seed + 0x9e3779b97f4a7c15ull * (uint64_t)(b + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    idx[b] = (uint32_t)(z % n);
}

}  // namespace dfsph

extern "C" {

int dfsph_mlp_harvest(void* handle, uint32_t capacity, uint32_t stride)
{
    MlpState* m = nullptr;
    if (int rc = ensure_state(handle, &m)) return rc;
    if (capacity != m->ds_cap) {
        if (int rc = malloc_dev(m->ds_x, (size_t)capacity * dfsph::kIn)) return rc;
        if (int rc = malloc_dev(m->ds_t, (size_t)capacity * dfsph::kOut)) return rc;
    }
    m->ds_cap = capacity;
    m->ds_n = 0;
    m->ds_stride = stride ? stride : 1;
    return DFSPH_OK;
}

int dfsph_mlp_dataset_size(void* handle, uint32_t* n)
{
    MlpState* m = nullptr;
    if (int rc = ensure_state(handle, &m)) return rc;
    if (n) *n = m->ds_n;
    return DFSPH_OK;
}

int dfsph_mlp_dataset_download(void* handle, float* x, float* t, uint32_t n)
{
    MlpState* m = nullptr;
    if (int rc = ensure_state(handle, &m)) return rc;
    if (n > m->ds_n) return mfail(DFSPH_ERR_ARG, "more samples requested than harvested");
    MHIP(hipStreamSynchronize(dfsph::handle_view(handle).s));
    if (x && n) MHIP(hipMemcpy(x, m->ds_x, sizeof(float) * n * dfsph::kIn, hipMemcpyDeviceToHost));
    if (t && n) MHIP(hipMemcpy(t, m->ds_t, sizeof(float) * n * dfsph::kOut, hipMemcpyDeviceToHost));
    return DFSPH_OK;
}

int dfsph_mlp_dataset_upload(void* handle, const float* x, const float* t, uint32_t n)
{
    MlpState* m = nullptr;
    if (int rc = ensure_state(handle, &m)) return rc;
    if (int rc = dfsph_mlp_harvest(handle, n, 1)) return rc;
    hipStream_t s = dfsph::handle_view(handle).s;
    if (n) {
        MHIP(hipMemcpyAsync(m->ds_x, x, sizeof(float) * n * dfsph::kIn, hipMemcpyHostToDevice, s));
        MHIP(hipMemcpyAsync(m->ds_t, t, sizeof(float) * n * dfsph::kOut, hipMemcpyHostToDevice, s));
        MHIP(hipStreamSynchronize(s));
    }
    m->ds_n = n;
    m->ds_cap = n;   // full: no further harvesting into it
    return DFSPH_OK;
}

// Mean / standard deviation of the training set's features and targets (FP64
// on the host) become the model's normalisation.
int dfsph_mlp_fit_normalization(void* handle)
{
    MlpState* m = nullptr;
    if (int rc = ensure_state(handle, &m)) return rc;
    if (!m->loaded) return mfail(DFSPH_ERR_NO_MODEL, "set initial parameters first");
    const uint32_t n = m->ds_n;
    if (n < 2) return mfail(DFSPH_ERR_STATE, "training set is empty");
    std::vector<float> x((size_t)n * dfsph::kIn), t((size_t)n * dfsph::kOut);
    if (int rc = dfsph_mlp_dataset_download(handle, x.data(), t.data(), n)) return rc;
    std::vector<float> nrm(dfsph::kNNorm);
    auto stats = [&](const std::vector<float>& a, int w, int col, float* mu, float* sd) {
        double s = 0.0, s2 = 0.0;
        for (uint32_t i = 0; i < n; ++i) {
            const double v = a[(size_t)i * w + col];
            s += v;
            s2 += v * v;
        }
        // no absolute floor on the spread: the dt-invariant totals are O(1e-9)
        // in SI units; only a constant column falls back to its magnitude or 1
        const double mean = s / n, var = std::max(s2 / n - mean * mean, 0.0);
        double dev = std::sqrt(var);
        if (!(dev > 1e-7 * std::fabs(mean)) || !(dev > 1e-30)) dev = std::max(std::fabs(mean), 1.0);
        *mu = (float)mean;
        *sd = (float)dev;
    };
    for (int k = 0; k < dfsph::kIn; ++k) stats(x, dfsph::kIn, k, &nrm[k], &nrm[dfsph::kIn + k]);
    for (int k = 0; k < dfsph::kOut; ++k)
        stats(t, dfsph::kOut, k, &nrm[2 * dfsph::kIn + k], &nrm[2 * dfsph::kIn + dfsph::kOut + k]);
    hipStream_t s = dfsph::handle_view(handle).s;
    MHIP(hipMemcpyAsync(m->norm, nrm.data(), sizeof(float) * dfsph::kNNorm, hipMemcpyHostToDevice, s));
    if (int rc = pack(m, s)) return rc;
    MHIP(hipStreamSynchronize(s));
    return DFSPH_OK;
}

// Mini-batch Adam on the training set with the normalised loss; the learning
// rate decays linearly to 10 % over the run.  loss_first / loss_last: the
// batch losses of the first and last step.
int dfsph_mlp_train(void* handle, uint32_t steps, uint32_t batch, float lr, uint32_t seed, float* loss_first,
                    float* loss_last)
{
    MlpState* m = nullptr;
    if (int rc = ensure_state(handle, &m)) return rc;
    if (!m->loaded) return mfail(DFSPH_ERR_NO_MODEL, "set initial parameters first");
    if (m->ds_n == 0) return mfail(DFSPH_ERR_STATE, "training set is empty");
    if (batch == 0 || batch % 64) return mfail(DFSPH_ERR_ARG, "batch must be a positive multiple of 64");
    hipStream_t s = dfsph::handle_view(handle).s;
    const uint32_t nb = batch / 64;
    if (nb > m->partial_cap) {
        if (int rc = malloc_dev(m->partial, (size_t)nb * dfsph::kNParam)) return rc;
        if (int rc = malloc_dev(m->lossp, nb)) return rc;
        m->partial_cap = nb;
    }
    if (batch > m->idx_cap) {
        if (int rc = malloc_dev(m->idx, batch)) return rc;
        m->idx_cap = batch;
    }
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (uint32_t it = 0; it < steps; ++it) {
        dfsph::k_batch_indices<<<(batch + 255) / 256, 256, 0, s>>>(m->idx, batch, m->ds_n,
                                                                   ((uint64_t)seed << 32) ^ (uint64_t)it * 0x632be59bd9b4e019ull);
        dfsph::k_mlp_grad<<<nb, 64, 0, s>>>(m->model, m->ds_x, m->ds_t, m->idx, batch, 1.0f / (float)batch, 1,
                                            m->partial, m->lossp);
        dfsph::k_mlp_reduce<<<(dfsph::kNParam + 255) / 256, 256, 0, s>>>(m->partial, m->lossp, nb, m->grad, m->loss);
        m->adam_t++;
        const float c1 = 1.0f / (1.0f - std::pow(b1, (float)m->adam_t));
        const float c2 = 1.0f / (1.0f - std::pow(b2, (float)m->adam_t));
        const float lr_t = lr * (1.0f - 0.9f * (float)it / (float)(steps > 1 ? steps - 1 : 1));
        dfsph::k_mlp_adam<<<(dfsph::kNParam + 255) / 256, 256, 0, s>>>(m->w, m->m, m->v, m->grad, lr_t, b1, b2, eps,
                                                                       c1, c2);
        if (int rc = pack(m, s)) return rc;
        if (it == 0 && loss_first) MHIP(hipMemcpyAsync(loss_first, m->loss, sizeof(float), hipMemcpyDeviceToHost, s));
        if (it + 1 == steps && loss_last) MHIP(hipMemcpyAsync(loss_last, m->loss, sizeof(float), hipMemcpyDeviceToHost, s));
    }
    MHIP(hipGetLastError());
    MHIP(hipStreamSynchronize(s));
    return DFSPH_OK;
}

// features and predictions of the last learned / harvesting step (device order)
int dfsph_mlp_step_features(void* handle, float* feat, float* pred_v, float* pred_p, uint32_t n)
{
    MlpState* m = nullptr;
    if (int rc = ensure_state(handle, &m)) return rc;
    if (!m->feat || n > m->cap) return mfail(DFSPH_ERR_STATE, "no features computed");
    MHIP(hipStreamSynchronize(dfsph::handle_view(handle).s));
    if (feat) MHIP(hipMemcpy(feat, m->feat, sizeof(float) * n * dfsph::kIn, hipMemcpyDeviceToHost));
    if (pred_v) MHIP(hipMemcpy(pred_v, m->pred, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (pred_p) MHIP(hipMemcpy(pred_p, m->pred + m->cap, sizeof(float) * n, hipMemcpyDeviceToHost));
    return DFSPH_OK;
}

}  // extern "C"
