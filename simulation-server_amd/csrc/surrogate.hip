// surrogate.hip -- the learned per-cell spring surrogate: an MLP
// 294 -> 256 -> 256 -> 256 -> 126 (GELU) in BF16 on gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, and its training.
//
// It stands in for the spring part of the cell-force kernel
// (gatherForcesKernel, src/objects/blood_cells.cu:66-120;
// calculateParticlesSpringForceComponent, src/simulation/physics.cuh:53-78)
// for 42-particle cells.  One row per cell:
//   input   126 centroid-relative positions / pos_scale, 126 velocities
//           relative to the cell's mean velocity / vel_scale, 42 contact
//           flags of the previous frame (1 = the particle touched another)
//   output  126 spring displacements d = (dt^2 / 4) * S (the spring force's
//           share of the frame's Heun position update) / out_scale
//
// Kernels (all hand-written):
//   k_gemm_nt    C = A * B^T (+ bias, GELU): A is M x K, B is N x K, both
//                bf16 row-major with K contiguous; 128 x 128 output tiles per
//                256-thread block, each wave 64 x 64 as 4 x 4 MFMA 16x16x32
//                tiles, A and B staged through LDS 32 columns at a time
//   k_transpose  bf16/fp32 -> bf16 transposes (the backward GEMMs are NT
//                products of transposed operands)
//   k_loss_grad, k_gelu_back, k_colsum, k_adam, k_to_bf16, k_gather_rows
// Training keeps fp32 master weights and Adam moments; the GEMMs read bf16
// copies.
#include "surrogate.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace bcs {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ uint16_t to_bf16(float f) { return bf16_bits(f); }
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU in its tanh form (Hendrycks & Gimpel 2016): 0.5 x (1 + tanh(z / 2))
// = x / (1 + exp(-z)), z = 2 sqrt(2/pi) (x + 0.044715 x^3); and its
// derivative s + x s (1 - s) dz/dx with s = 1 / (1 + exp(-z)).  exp through
// the hardware base-2 exponential and 1/(1 + e) through its reciprocal: a
// few instructions per element, against a result rounded to bf16.
constexpr float kGeluC = 1.5957691216057308f;  // 2 sqrt(2/pi)
constexpr float kGeluA = 0.044715f;
__device__ __forceinline__ float gelu_sig(float x)
{
    const float z = kGeluC * (x + kGeluA * x * x * x);
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_sig(x); }
__device__ __forceinline__ float gelu_grad(float x)
{
    const float sg = gelu_sig(x);
    return sg + x * sg * (1.0f - sg) * kGeluC * (1.0f + 3.0f * kGeluA * x * x);
}

constexpr int kBM = 128, kBN = 128, kBK = 32;
constexpr int kLdk = kBK + 8;  // LDS row pitch in bf16 (80 bytes: 16-byte aligned rows)

struct GemmArgs {
    const uint16_t* A;
    int64_t lda;
    const uint16_t* B;
    int64_t ldb;
    int32_t M, N, K;  // K % 32 == 0
    const float* bias;  // [N] or null
    float* Cf;          // fp32 result (before the activation), or null
    int64_t ldcf;
    uint16_t* Cb;       // bf16 result (after the activation), or null
    int64_t ldcb;
    int act;            // 1: GELU
};

// Operand layout of v_mfma_f32_16x16x32_bf16: lane l holds A[m = l % 16]
// [k = 8 (l / 16) .. + 7] and B[k = 8 (l / 16) .. + 7][n = l % 16]; its four
// results are D[m = 4 (l / 16) + r][n = l % 16], r = 0..3.
__global__ void __launch_bounds__(256) k_gemm_nt(GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) uint16_t sA[kBM * kLdk];
    __shared__ __attribute__((aligned(16))) uint16_t sB[kBN * kLdk];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t m0 = (int64_t)blockIdx.x * kBM;
    const int n0 = blockIdx.y * kBN;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += kBK) {
        // 128 rows x 32 columns of A and of B: 512 chunks of 8 bf16 each
#pragma unroll
        for (int c = tid; c < 512; c += 256) {
            const int r = c >> 2, q = c & 3;
            uint4 va = make_uint4(0u, 0u, 0u, 0u), vb = va;
            if (m0 + r < g.M) va = *reinterpret_cast<const uint4*>(g.A + (m0 + r) * g.lda + k0 + 8 * q);
            if (n0 + r < g.N) vb = *reinterpret_cast<const uint4*>(g.B + (int64_t)(n0 + r) * g.ldb + k0 + 8 * q);
            *reinterpret_cast<uint4*>(sA + r * kLdk + 8 * q) = va;
            *reinterpret_cast<uint4*>(sB + r * kLdk + 8 * q) = vb;
        }
        __syncthreads();
        bf16x8 a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            a[i] = *reinterpret_cast<const bf16x8*>(sA + (wm + 16 * i + (lane & 15)) * kLdk + 8 * (lane >> 4));
#pragma unroll
        for (int j = 0; j < 4; ++j)
            b[j] = *reinterpret_cast<const bf16x8*>(sB + (wn + 16 * j + (lane & 15)) * kLdk + 8 * (lane >> 4));
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + wn + 16 * j + (lane & 15);
            if (n >= g.N) continue;
            const float bn = g.bias ? g.bias[n] : 0.0f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                if (m >= g.M) continue;
                float v = acc[i][j][r] + bn;
                if (g.Cf) g.Cf[m * g.ldcf + n] = v;
                if (g.act) v = gelu(v);
                if (g.Cb) g.Cb[m * g.ldcb + n] = to_bf16(v);
            }
        }
}

// dst[c][r] = src[r][c] for r < R, c < C (src fp32 or bf16, dst bf16).
template <bool SRC_F32>
__global__ void __launch_bounds__(256) k_transpose(const void* __restrict__ src, int64_t lds, int32_t R, int32_t C,
                                                   uint16_t* __restrict__ dst, int64_t ldd)
{
    __shared__ uint16_t t[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const int64_t r0 = (int64_t)blockIdx.x * 32;
    const int c0 = blockIdx.y * 32;
    for (int k = ty; k < 32; k += 8) {
        const int64_t r = r0 + k;
        const int c = c0 + tx;
        uint16_t v = 0;
        if (r < R && c < C) {
            if (SRC_F32) v = to_bf16(static_cast<const float*>(src)[r * lds + c]);
            else v = static_cast<const uint16_t*>(src)[r * lds + c];
        }
        t[k][tx] = v;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k;
        const int64_t r = r0 + tx;
        if (r < R && c < C) dst[(int64_t)c * ldd + r] = t[tx][k];
    }
}

// Mean squared error over the kSurrOut real outputs: dZ = 2 (o - y) / (B * 126)
// (zero on the padding columns), fp32 and bf16; loss partial sums (report only).
__global__ void __launch_bounds__(256) k_loss_grad(const float* __restrict__ out, const float* __restrict__ y,
                                                   int32_t B, float* __restrict__ dz, uint16_t* __restrict__ dzb,
                                                   float* __restrict__ loss)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float sq = 0.f;
    if (e < (int64_t)B * kSurrOutPad) {
        const int64_t m = e / kSurrOutPad;
        const int n = (int)(e % kSurrOutPad);
        float gv = 0.f;
        if (n < kSurrOut) {
            const float d = out[e] - y[m * kSurrOut + n];
            sq = d * d;
            gv = 2.0f * d / ((float)B * (float)kSurrOut);
        }
        dz[e] = gv;
        dzb[e] = to_bf16(gv);
    }
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_down(sq, o, 64);
    if ((threadIdx.x & 63) == 0 && sq != 0.f) atomicAdd(loss, sq);
}

// dZ = dY * GELU'(Z) (fp32 and bf16), n elements.
__global__ void __launch_bounds__(256) k_gelu_back(const float* __restrict__ dy, const float* __restrict__ z, int64_t n,
                                                   float* __restrict__ dz, uint16_t* __restrict__ dzb)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float v = dy[e] * gelu_grad(z[e]);
    dz[e] = v;
    dzb[e] = to_bf16(v);
}

// db[n] = sum over the B rows of dz[m][n] (row stride ld), in two
// deterministic passes: partial sums of kColRows rows per (chunk, column),
// then the chunks in order.
constexpr int kColRows = 64;
__global__ void __launch_bounds__(256) k_colsum_part(const float* __restrict__ dz, int32_t B, int32_t N, int64_t ld,
                                                     float* __restrict__ part)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (n >= N) return;
    const int32_t m0 = c * kColRows, m1 = min(B, m0 + kColRows);
    float s = 0.f;
    for (int32_t m = m0; m < m1; ++m) s += dz[(int64_t)m * ld + n];
    part[(int64_t)c * N + n] = s;
}
__global__ void __launch_bounds__(256) k_colsum(const float* __restrict__ part, int32_t chunks, int32_t N,
                                                float* __restrict__ db)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int32_t c = 0; c < chunks; ++c) s += part[(int64_t)c * N + n];
    db[n] = s;
}

// Adam (Kingma & Ba 2015) over a flat parameter vector; bias-corrected step.
__global__ void __launch_bounds__(256) k_adam(float* __restrict__ p, const float* __restrict__ grad,
                                              float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                                              float b1, float b2, float eps, float c1, float c2)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gv = grad[e];
    const float mm = b1 * m[e] + (1.0f - b1) * gv;
    const float vv = b2 * v[e] + (1.0f - b2) * gv * gv;
    m[e] = mm;
    v[e] = vv;
    p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

__global__ void __launch_bounds__(256) k_to_bf16(const float* __restrict__ src, int64_t n, uint16_t* __restrict__ dst)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) dst[e] = to_bf16(src[e]);
}

// Batch rows idx[0..B) of the training set: features (bf16, ld kSurrInPad)
// and targets (fp32, ld kSurrOut).
__global__ void __launch_bounds__(256) k_gather_rows(const uint16_t* __restrict__ X, const float* __restrict__ Y,
                                                     const int32_t* __restrict__ idx, int32_t B,
                                                     uint16_t* __restrict__ xb, float* __restrict__ yb)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nx = (int64_t)B * kSurrInPad;
    if (e < nx) {
        const int64_t m = e / kSurrInPad;
        xb[e] = X[(int64_t)idx[m] * kSurrInPad + e % kSurrInPad];
    } else if (e < nx + (int64_t)B * kSurrOut) {
        const int64_t f = e - nx;
        const int64_t m = f / kSurrOut;
        yb[f] = Y[(int64_t)idx[m] * kSurrOut + f % kSurrOut];
    }
}

// ---- fused inference: all four layers in one kernel ------------------------
// A block takes kFM rows (cells).  Their activations stay in LDS (bf16, row
// pitch kHPitch) from the input to the last hidden layer; each weight matrix
// streams through one LDS buffer 32 columns at a time, fetched into registers
// two chunks ahead.  With 62 KB of LDS and at most 128 registers per lane two
// blocks share a CU, so that one block's waits (its barriers, the weight
// stores) overlap the other's MFMA chain.  The products are formed
// transposed, W x H^T, so that each lane ends with 4 consecutive output
// features of one cell: the bias, GELU and bf16 rounding of a hidden layer go
// back to LDS as one 8-byte store per 4 values.  The 4 waves split a layer's
// output features 4 ways: 64 cells x 64 features per wave for the hidden
// layers (4 x 4 MFMA tiles), 64 x 32 for the 128-feature output layer.  Only
// the input rows and the output rows touch HBM.
constexpr int kFM = 64;
constexpr int kHPitch = kSurrInPad + 8;  // 328 bf16: rows 16 B apart modulo the LDS banks
constexpr int kWPitch = kBK + 8;         // 40 bf16
constexpr int kWBuf = kSurrHidden * kWPitch;
constexpr int kAhead = 2;                // weight chunks in flight in registers

__device__ __forceinline__ uint16_t bf16_rn(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

// Columns [k0, k0 + 32) of the first `rows` rows of W (ld K): 4 16-byte pieces
// per row, piece c = tid + 256 q of the block.
__device__ __forceinline__ void w_chunk_load(const uint16_t* __restrict__ W, int K, int rows, int k0, int tid,
                                             uint4 (&r)[4])
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = tid + 256 * q, row = c >> 2, part = c & 3;
        r[q] = row < rows ? *reinterpret_cast<const uint4*>(W + (size_t)row * K + k0 + 8 * part)
                          : make_uint4(0u, 0u, 0u, 0u);
    }
}
__device__ __forceinline__ void w_chunk_store(uint16_t* buf, int rows, int tid, const uint4 (&r)[4])
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = tid + 256 * q, row = c >> 2, part = c & 3;
        if (row < rows) *reinterpret_cast<uint4*>(buf + row * kWPitch + 8 * part) = r[q];
    }
}

// acc[i][j] = W[wn + 16 j ..][0 .. K) x H[16 i ..][0 .. K)^T: lane l holds
// features wn + 16 j + 4 (l / 16) + r of cell 16 i + l % 16.  W has ROWS
// rows.  Ends with a barrier: every wave is done with H.
template <int NT, int ROWS, int K>
__device__ __forceinline__ void mlp_layer(const uint16_t* __restrict__ W, const uint16_t* H, uint16_t* wbuf,
                                          f32x4 (&acc)[4][NT], int wn, int lane, int tid)
{
    constexpr int nks = K / kBK;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    uint4 r[kAhead][4];
#pragma unroll
    for (int a = 0; a < kAhead; ++a)
        if (a < nks) w_chunk_load(W, K, ROWS, a * kBK, tid, r[a]);
    w_chunk_store(wbuf, ROWS, tid, r[0]);
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < nks; ++ks) {
        // chunk ks + kAhead into the registers of chunk ks (in LDS since the
        // last step); it is stored kAhead - 1 steps from now
        if (ks + kAhead < nks) w_chunk_load(W, K, ROWS, (ks + kAhead) * kBK, tid, r[ks % kAhead]);
        bf16x8 a[4], b[NT];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            a[i] = *reinterpret_cast<const bf16x8*>(H + (16 * i + (lane & 15)) * kHPitch + ks * kBK + 8 * (lane >> 4));
#pragma unroll
        for (int j = 0; j < NT; ++j)
            b[j] = *reinterpret_cast<const bf16x8*>(wbuf + (wn + 16 * j + (lane & 15)) * kWPitch + 8 * (lane >> 4));
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < NT; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j], a[i], acc[i][j], 0, 0, 0);
        __syncthreads();  // every wave is done with this chunk (and, at the end, with H)
        if (ks + 1 < nks) {
            w_chunk_store(wbuf, ROWS, tid, r[(ks + 1) % kAhead]);
            __syncthreads();
        }
    }
}

__global__ void __launch_bounds__(256, 2) k_mlp_infer(SurrNet net, const uint16_t* __restrict__ X, int32_t n,
                                                      float* __restrict__ out)
{
    __shared__ __attribute__((aligned(16))) uint16_t H[kFM * kHPitch];
    __shared__ __attribute__((aligned(16))) uint16_t wbuf[kWBuf];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t m0 = (int64_t)blockIdx.x * kFM;
    // input rows (zeros past the last row)
    for (int c = tid; c < kFM * (kSurrInPad / 8); c += 256) {
        const int row = c / (kSurrInPad / 8), part = c % (kSurrInPad / 8);
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (m0 + row < n) v = *reinterpret_cast<const uint4*>(X + (m0 + row) * kSurrInPad + 8 * part);
        *reinterpret_cast<uint4*>(H + row * kHPitch + 8 * part) = v;
    }
    __syncthreads();
    const int cm = lane & 15, fq = 4 * (lane >> 4);  // the lane's cell within a tile, first feature of its 4
    {
        f32x4 acc[4][4];
        const int wn = wave * 64;
#pragma unroll 1
        for (int l = 0; l < 3; ++l) {
            if (l == 0) mlp_layer<4, kSurrHidden, kSurrInPad>(net.W[0], H, wbuf, acc, wn, lane, tid);
            else mlp_layer<4, kSurrHidden, kSurrHidden>(net.W[l], H, wbuf, acc, wn, lane, tid);
            // bias, GELU, bf16: the next layer's input, in place
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int nn = wn + 16 * j + fq;
                const float4 bn = *reinterpret_cast<const float4*>(net.b[l] + nn);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const f32x4 v = acc[i][j];
                    const uint32_t lo = (uint32_t)bf16_rn(gelu(v[0] + bn.x)) | ((uint32_t)bf16_rn(gelu(v[1] + bn.y)) << 16);
                    const uint32_t hi = (uint32_t)bf16_rn(gelu(v[2] + bn.z)) | ((uint32_t)bf16_rn(gelu(v[3] + bn.w)) << 16);
                    *reinterpret_cast<uint2*>(H + (16 * i + cm) * kHPitch + nn) = make_uint2(lo, hi);
                }
            }
            __syncthreads();
        }
    }
    f32x4 acc[4][2];
    const int wn = wave * 32;
    mlp_layer<2, kSurrOutPad, kSurrHidden>(net.W[3], H, wbuf, acc, wn, lane, tid);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int nn = wn + 16 * j + fq;
        const float4 bn = *reinterpret_cast<const float4*>(net.b[3] + nn);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t m = m0 + 16 * i + cm;
            const f32x4 v = acc[i][j];
            if (m < n)
                *reinterpret_cast<float4*>(out + m * kSurrOutPad + nn) =
                    make_float4(v[0] + bn.x, v[1] + bn.y, v[2] + bn.z, v[3] + bn.w);
        }
    }
}

void gemm(const uint16_t* A, int64_t lda, const uint16_t* B, int64_t ldb, int32_t M, int32_t N, int32_t K,
          const float* bias, float* Cf, int64_t ldcf, uint16_t* Cb, int64_t ldcb, int act, hipStream_t st)
{
    if (M <= 0 || N <= 0) return;
    const GemmArgs g{A, lda, B, ldb, M, N, K, bias, Cf, ldcf, Cb, ldcb, act};
    hipLaunchKernelGGL(k_gemm_nt, dim3((unsigned)((M + kBM - 1) / kBM), (unsigned)((N + kBN - 1) / kBN)), dim3(256), 0,
                       st, g);
}

template <bool F32>
void transpose(const void* src, int64_t lds, int32_t R, int32_t C, uint16_t* dst, int64_t ldd, hipStream_t st)
{
    hipLaunchKernelGGL(k_transpose<F32>, dim3((unsigned)((R + 31) / 32), (unsigned)((C + 31) / 32)), dim3(256), 0, st,
                       src, lds, R, C, dst, ldd);
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

// layer l: in -> out, weights [out][in] (in padded for layer 0, out padded for layer 3)
constexpr int kLayerIn[4] = {kSurrInPad, kSurrHidden, kSurrHidden, kSurrHidden};
constexpr int kLayerOut[4] = {kSurrHidden, kSurrHidden, kSurrHidden, kSurrOutPad};
constexpr int kLayerInReal[4] = {kSurrIn, kSurrHidden, kSurrHidden, kSurrHidden};
constexpr int kLayerOutReal[4] = {kSurrHidden, kSurrHidden, kSurrHidden, kSurrOut};

#define SHIP(call)                                                                                           \
    do {                                                                                                     \
        hipError_t e_ = (call);                                                                              \
        if (e_ != hipSuccess) return fail(BCS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

void surrogate_infer(const SurrNet& net, const uint16_t* X, int32_t n, uint16_t* act0, uint16_t* act1, float* out,
                     hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + kFM - 1) / kFM)), dim3(256), 0, st, net, X, n, out);
}

}  // namespace bcs

using namespace bcs;

struct bcs_surrogate {
    int device = 0;
    hipStream_t stream = nullptr;
    // flat fp32 parameters: W0 b0 W1 b1 W2 b2 W3 b3 (W [out][in] padded)
    int64_t off_w[4]{}, off_b[4]{}, nparam = 0;
    int64_t off_wb[4]{}, nwb = 0;  // bf16 weight copies
    float *p = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    uint16_t* wb = nullptr;
    float pos_scale = 1.0f, vel_scale = 1.0f, out_scale = 1.0f;
    int64_t adam_t = 0;
    std::vector<void*> allocs;
    ~bcs_surrogate()
    {
        for (void* q : allocs) (void)hipFree(q);
        if (stream) (void)hipStreamDestroy(stream);
    }
    template <class T>
    int dalloc(T** q, size_t count)
    {
        hipError_t e = hipMalloc((void**)q, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) return fail(BCS_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back((void*)*q);
        return BCS_OK;
    }
    SurrNet net() const
    {
        SurrNet n{};
        for (int l = 0; l < 4; ++l) {
            n.W[l] = wb + off_wb[l];
            n.b[l] = p + off_b[l];
        }
        n.pos_scale = pos_scale;
        n.vel_scale = vel_scale;
        n.out_scale = out_scale;
        return n;
    }
    void refresh_bf16()
    {
        for (int l = 0; l < 4; ++l) {
            const int64_t n = (int64_t)kLayerOut[l] * kLayerIn[l];
            hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(n)), dim3(256), 0, stream, p + off_w[l], n, wb + off_wb[l]);
        }
    }
};

#define RCS(x)                \
    do {                      \
        int rc_ = (x);        \
        if (rc_) return rc_;  \
    } while (0)

int bcs_surrogate_net(bcs_surrogate* s, SurrNet* out, int* device)
{
    if (!s || !out) return fail(BCS_ERR_INVALID_ARGUMENT, "null surrogate");
    *out = s->net();
    if (device) *device = s->device;
    return BCS_OK;
}

extern "C" int bcs_surrogate_create(int32_t device, uint64_t seed, bcs_surrogate** out)
{
    if (!out) return fail(BCS_ERR_INVALID_ARGUMENT, "null output");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(BCS_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return fail(BCS_ERR_NO_DEVICE, "device index out of range");
    std::unique_ptr<bcs_surrogate> s(new bcs_surrogate());
    s->device = device;
    SHIP(hipSetDevice(device));
    SHIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    int64_t o = 0, ob = 0;
    for (int l = 0; l < 4; ++l) {
        s->off_w[l] = o;
        o += (int64_t)kLayerOut[l] * kLayerIn[l];
        s->off_b[l] = o;
        o += kLayerOut[l];
        s->off_wb[l] = ob;
        ob += (int64_t)kLayerOut[l] * kLayerIn[l];
    }
    s->nparam = o;
    s->nwb = ob;
    RCS(s->dalloc(&s->p, o));
    RCS(s->dalloc(&s->grad, o));
    RCS(s->dalloc(&s->m1, o));
    RCS(s->dalloc(&s->m2, o));
    RCS(s->dalloc(&s->wb, ob));
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases on the real
    // entries (SplitMix64 draws), zeros on the padding
    std::vector<float> h((size_t)o, 0.0f);
    uint64_t st = seed ^ 0x5EED5EEDull;
    auto u01 = [&st]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * 0x1p-53;
    };
    for (int l = 0; l < 4; ++l) {
        const double a = 1.0 / std::sqrt((double)kLayerInReal[l]);
        for (int r = 0; r < kLayerOutReal[l]; ++r) {
            for (int c = 0; c < kLayerInReal[l]; ++c)
                h[(size_t)(s->off_w[l] + (int64_t)r * kLayerIn[l] + c)] = (float)((2.0 * u01() - 1.0) * a);
            h[(size_t)(s->off_b[l] + r)] = (float)((2.0 * u01() - 1.0) * a);
        }
    }
    SHIP(hipMemcpyAsync(s->p, h.data(), (size_t)o * 4, hipMemcpyHostToDevice, s->stream));
    SHIP(hipMemsetAsync(s->m1, 0, (size_t)o * 4, s->stream));
    SHIP(hipMemsetAsync(s->m2, 0, (size_t)o * 4, s->stream));
    s->refresh_bf16();
    SHIP(hipStreamSynchronize(s->stream));
    *out = s.release();
    return BCS_OK;
}

extern "C" int bcs_surrogate_destroy(bcs_surrogate* s)
{
    if (!s) return BCS_OK;
    (void)hipSetDevice(s->device);
    (void)hipStreamSynchronize(s->stream);
    delete s;
    return BCS_OK;
}

// weights[l]: [out][in] row-major fp32 of the real sizes (294x256 ... 256x126
// as out x in), biases[l]: out entries.
extern "C" int bcs_surrogate_set_weights(bcs_surrogate* s, const float* const* weights, const float* const* biases)
{
    if (!s || !weights || !biases) return fail(BCS_ERR_INVALID_ARGUMENT, "null argument");
    SHIP(hipSetDevice(s->device));
    std::vector<float> h((size_t)s->nparam, 0.0f);
    for (int l = 0; l < 4; ++l) {
        if (!weights[l] || !biases[l]) return fail(BCS_ERR_INVALID_ARGUMENT, "null layer");
        for (int r = 0; r < kLayerOutReal[l]; ++r) {
            for (int c = 0; c < kLayerInReal[l]; ++c)
                h[(size_t)(s->off_w[l] + (int64_t)r * kLayerIn[l] + c)] = weights[l][(size_t)r * kLayerInReal[l] + c];
            h[(size_t)(s->off_b[l] + r)] = biases[l][r];
        }
    }
    SHIP(hipMemcpyAsync(s->p, h.data(), (size_t)s->nparam * 4, hipMemcpyHostToDevice, s->stream));
    s->refresh_bf16();
    SHIP(hipStreamSynchronize(s->stream));
    return BCS_OK;
}

extern "C" int bcs_surrogate_get_weights(bcs_surrogate* s, float* const* weights, float* const* biases)
{
    if (!s || !weights || !biases) return fail(BCS_ERR_INVALID_ARGUMENT, "null argument");
    SHIP(hipSetDevice(s->device));
    std::vector<float> h((size_t)s->nparam);
    SHIP(hipMemcpyAsync(h.data(), s->p, (size_t)s->nparam * 4, hipMemcpyDeviceToHost, s->stream));
    SHIP(hipStreamSynchronize(s->stream));
    for (int l = 0; l < 4; ++l) {
        for (int r = 0; r < kLayerOutReal[l]; ++r) {
            if (weights[l])
                for (int c = 0; c < kLayerInReal[l]; ++c)
                    weights[l][(size_t)r * kLayerInReal[l] + c] = h[(size_t)(s->off_w[l] + (int64_t)r * kLayerIn[l] + c)];
            if (biases[l]) biases[l][r] = h[(size_t)(s->off_b[l] + r)];
        }
    }
    return BCS_OK;
}

extern "C" int bcs_surrogate_set_scales(bcs_surrogate* s, float pos_scale, float vel_scale, float out_scale)
{
    if (!s || !(pos_scale > 0) || !(vel_scale > 0) || !(out_scale > 0))
        return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_surrogate_set_scales: positive scales");
    s->pos_scale = pos_scale;
    s->vel_scale = vel_scale;
    s->out_scale = out_scale;
    return BCS_OK;
}

extern "C" int bcs_surrogate_get_scales(bcs_surrogate* s, float* pos_scale, float* vel_scale, float* out_scale)
{
    if (!s) return fail(BCS_ERR_INVALID_ARGUMENT, "null surrogate");
    if (pos_scale) *pos_scale = s->pos_scale;
    if (vel_scale) *vel_scale = s->vel_scale;
    if (out_scale) *out_scale = s->out_scale;
    return BCS_OK;
}

namespace {

// Host rows (n x 294, physical units) -> device bf16 (n x 320, scaled).
int upload_features(bcs_surrogate* s, const float* feat, int32_t n, uint16_t* dst)
{
    std::vector<uint16_t> h((size_t)n * kSurrInPad, 0);
    for (int32_t m = 0; m < n; ++m)
        for (int c = 0; c < kSurrIn; ++c) {
            const float sc = c < 126 ? s->pos_scale : (c < 252 ? s->vel_scale : 1.0f);
            h[(size_t)m * kSurrInPad + c] = to_bf16(feat[(size_t)m * kSurrIn + c] / sc);
        }
    SHIP(hipMemcpyAsync(dst, h.data(), h.size() * 2, hipMemcpyHostToDevice, s->stream));
    SHIP(hipStreamSynchronize(s->stream));
    return BCS_OK;
}

}  // namespace

// Spring displacements (n x 126, physical units) of n feature rows (n x 294,
// physical units).
extern "C" int bcs_surrogate_predict(bcs_surrogate* s, const float* features, int32_t n, float* out)
{
    if (!s || n < 0 || (n > 0 && (!features || !out))) return fail(BCS_ERR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return BCS_OK;
    SHIP(hipSetDevice(s->device));
    uint16_t *X = nullptr, *a0 = nullptr, *a1 = nullptr;
    float* o = nullptr;
    SHIP(hipMalloc(&X, (size_t)n * kSurrInPad * 2));
    SHIP(hipMalloc(&a0, (size_t)n * kSurrHidden * 2));
    SHIP(hipMalloc(&a1, (size_t)n * kSurrHidden * 2));
    SHIP(hipMalloc(&o, (size_t)n * kSurrOutPad * 4));
    int rc = upload_features(s, features, n, X);
    if (!rc) {
        surrogate_infer(s->net(), X, n, a0, a1, o, s->stream);
        std::vector<float> h((size_t)n * kSurrOutPad);
        hipError_t e = hipMemcpyAsync(h.data(), o, h.size() * 4, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) rc = fail(BCS_ERR_HIP, std::string("bcs_surrogate_predict: ") + hipGetErrorString(e));
        for (int32_t m = 0; m < n && !rc; ++m)
            for (int c = 0; c < kSurrOut; ++c) out[(size_t)m * kSurrOut + c] = h[(size_t)m * kSurrOutPad + c] * s->out_scale;
    }
    (void)hipFree(X);
    (void)hipFree(a0);
    (void)hipFree(a1);
    (void)hipFree(o);
    return rc;
}

// Mean-squared-error training with Adam (lr, betas 0.9 / 0.999, eps 1e-8)
// on n samples (features n x 294 and targets n x 126, physical units; the
// scales set before apply), `epochs` passes in a seeded random order over
// batches of `batch` rows (a multiple of 128; the last partial batch is
// dropped).  loss_per_epoch (optional) receives the mean squared error in
// scaled units of each epoch.
extern "C" int bcs_surrogate_train(bcs_surrogate* s, const float* features, const float* targets, int32_t n,
                                   int32_t epochs, int32_t batch, float lr, uint64_t seed, float* loss_per_epoch)
{
    if (!s || !features || !targets || n < batch || epochs < 1 || batch < 128 || batch % 128 || !(lr > 0))
        return fail(BCS_ERR_INVALID_ARGUMENT,
                    "bcs_surrogate_train: n >= batch, batch a positive multiple of 128, epochs >= 1, lr > 0");
    SHIP(hipSetDevice(s->device));
    hipStream_t st = s->stream;
    std::vector<void*> tmp;
    auto talloc = [&](void** q, size_t bytes) -> int {
        hipError_t e = hipMalloc(q, std::max<size_t>(bytes, 1));
        if (e != hipSuccess) return fail(BCS_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
        tmp.push_back(*q);
        return BCS_OK;
    };
    auto cleanup = [&]() {
        for (void* q : tmp) (void)hipFree(q);
    };
    const int32_t B = batch;
    uint16_t *X = nullptr, *xb = nullptr, *wt = nullptr, *ta = nullptr, *tb = nullptr;
    float *Y = nullptr, *yb = nullptr, *loss = nullptr;
    int32_t* idx = nullptr;
    uint16_t* act[4]{};  // bf16 input of layer l (act[0] = xb)
    float* z[3]{};       // pre-activations of the hidden layers
    float *out = nullptr, *dyf = nullptr, *dzf = nullptr;
    uint16_t* dzb = nullptr;
    int rc = BCS_OK;
    rc = rc ? rc : talloc((void**)&X, (size_t)n * kSurrInPad * 2);
    rc = rc ? rc : talloc((void**)&Y, (size_t)n * kSurrOut * 4);
    rc = rc ? rc : talloc((void**)&xb, (size_t)B * kSurrInPad * 2);
    rc = rc ? rc : talloc((void**)&yb, (size_t)B * kSurrOut * 4);
    rc = rc ? rc : talloc((void**)&idx, (size_t)B * 4);
    rc = rc ? rc : talloc((void**)&loss, 4);
    for (int l = 1; l < 4 && !rc; ++l) rc = talloc((void**)&act[l], (size_t)B * kSurrHidden * 2);
    for (int l = 0; l < 3 && !rc; ++l) rc = talloc((void**)&z[l], (size_t)B * kSurrHidden * 4);
    rc = rc ? rc : talloc((void**)&out, (size_t)B * kSurrOutPad * 4);
    rc = rc ? rc : talloc((void**)&dyf, (size_t)B * kSurrHidden * 4);
    rc = rc ? rc : talloc((void**)&dzf, (size_t)B * kSurrHidden * 4);
    rc = rc ? rc : talloc((void**)&dzb, (size_t)B * kSurrHidden * 2);
    rc = rc ? rc : talloc((void**)&wt, (size_t)kSurrInPad * kSurrHidden * 2);
    rc = rc ? rc : talloc((void**)&ta, (size_t)kSurrHidden * B * 2);
    rc = rc ? rc : talloc((void**)&tb, (size_t)kSurrInPad * B * 2);
    const int32_t colchunks = (B + kColRows - 1) / kColRows;
    float* colpart = nullptr;  // per row chunk: the partial bias gradients
    rc = rc ? rc : talloc((void**)&colpart, (size_t)colchunks * kSurrHidden * 4);
    if (rc) {
        cleanup();
        return rc;
    }
    act[0] = xb;
    rc = upload_features(s, features, n, X);
    if (!rc) {
        std::vector<float> hy((size_t)n * kSurrOut);
        for (size_t e = 0; e < hy.size(); ++e) hy[e] = targets[e] / s->out_scale;
        hipError_t e = hipMemcpyAsync(Y, hy.data(), hy.size() * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = fail(BCS_ERR_HIP, std::string("bcs_surrogate_train: ") + hipGetErrorString(e));
    }
    std::vector<int32_t> order((size_t)n);
    for (int32_t k = 0; k < n; ++k) order[k] = k;
    uint64_t rs = seed ^ 0x7A11u;
    auto rnd = [&rs]() {
        uint64_t q = (rs += 0x9E3779B97F4A7C15ull);
        q = (q ^ (q >> 30)) * 0xBF58476D1CE4E5B9ull;
        q = (q ^ (q >> 27)) * 0x94D049BB133111EBull;
        return q ^ (q >> 31);
    };
    const int32_t nb = n / B;
    for (int32_t ep = 0; ep < epochs && !rc; ++ep) {
        for (int32_t k = n - 1; k > 0; --k) std::swap(order[k], order[rnd() % (uint64_t)(k + 1)]);
        (void)hipMemsetAsync(loss, 0, 4, st);
        for (int32_t bi = 0; bi < nb && !rc; ++bi) {
            (void)hipMemcpyAsync(idx, order.data() + (size_t)bi * B, (size_t)B * 4, hipMemcpyHostToDevice, st);
            const int64_t ng = (int64_t)B * (kSurrInPad + kSurrOut);
            hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for(ng)), dim3(256), 0, st, X, Y, idx, B, xb, yb);
            // forward
            const SurrNet net = s->net();
            for (int l = 0; l < 3; ++l)
                gemm(act[l], kLayerIn[l], net.W[l], kLayerIn[l], B, kSurrHidden, kLayerIn[l], net.b[l], z[l],
                     kSurrHidden, act[l + 1], kSurrHidden, 1, st);
            gemm(act[3], kSurrHidden, net.W[3], kSurrHidden, B, kSurrOutPad, kSurrHidden, net.b[3], out, kSurrOutPad,
                 nullptr, 0, 0, st);
            hipLaunchKernelGGL(k_loss_grad, dim3(blocks_for((int64_t)B * kSurrOutPad)), dim3(256), 0, st, out, yb, B,
                               dzf, dzb, loss);
            // backward, layer 3 down to 0; dzf / dzb hold dZ of layer l (B x out_l)
            for (int l = 3; l >= 0; --l) {
                const int in = kLayerIn[l], outn = kLayerOut[l];
                // dW_l = dZ^T X_l : A = dZ^T (out x B), B = X_l^T (in x B), K = B
                transpose<false>(dzb, outn, B, outn, ta, B, st);
                transpose<false>(act[l], in, B, in, tb, B, st);
                gemm(ta, B, tb, B, outn, in, B, nullptr, s->grad + s->off_w[l], in, nullptr, 0, 0, st);
                hipLaunchKernelGGL(k_colsum_part, dim3(blocks_for(outn), (unsigned)colchunks), dim3(256), 0, st, dzf, B,
                                   outn, (int64_t)outn, colpart);
                hipLaunchKernelGGL(k_colsum, dim3(blocks_for(outn)), dim3(256), 0, st, colpart, colchunks, outn,
                                   s->grad + s->off_b[l]);
                if (l == 0) break;
                // dY_{l-1} = dZ W_l : A = dZ (B x out), B = W_l^T (in x out), K = out
                transpose<false>(net.W[l], in, outn, in, wt, outn, st);
                gemm(dzb, outn, wt, outn, B, in, outn, nullptr, dyf, in, nullptr, 0, 0, st);
                hipLaunchKernelGGL(k_gelu_back, dim3(blocks_for((int64_t)B * in)), dim3(256), 0, st, dyf, z[l - 1],
                                   (int64_t)B * in, dzf, dzb);
            }
            s->adam_t += 1;
            const float c1 = 1.0f - (float)std::pow(0.9, (double)s->adam_t);
            const float c2 = 1.0f - (float)std::pow(0.999, (double)s->adam_t);
            hipLaunchKernelGGL(k_adam, dim3(blocks_for(s->nparam)), dim3(256), 0, st, s->p, s->grad, s->m1, s->m2,
                               s->nparam, lr, 0.9f, 0.999f, 1e-8f, c1, c2);
            s->refresh_bf16();
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) rc = fail(BCS_ERR_HIP, std::string("bcs_surrogate_train: ") + hipGetErrorString(e));
        }
        float lsum = 0.f;
        hipError_t e = hipMemcpyAsync(&lsum, loss, 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess && !rc) rc = fail(BCS_ERR_HIP, std::string("bcs_surrogate_train: ") + hipGetErrorString(e));
        if (loss_per_epoch) loss_per_epoch[ep] = lsum / ((float)nb * (float)B * (float)kSurrOut);
    }
    (void)hipStreamSynchronize(st);
    cleanup();
    return rc;
}

// Times `reps` inference passes over n_cells rows of (random) features on the
// device: ms per pass.
extern "C" int bcs_surrogate_time_inference(bcs_surrogate* s, int32_t n_cells, int32_t reps, float* ms_per_pass)
{
    if (!s || n_cells < 1 || reps < 1 || !ms_per_pass) return fail(BCS_ERR_INVALID_ARGUMENT, "bad arguments");
    SHIP(hipSetDevice(s->device));
    uint16_t *X = nullptr, *a0 = nullptr, *a1 = nullptr;
    float* o = nullptr;
    SHIP(hipMalloc(&X, (size_t)n_cells * kSurrInPad * 2));
    SHIP(hipMalloc(&a0, (size_t)n_cells * kSurrHidden * 2));
    SHIP(hipMalloc(&a1, (size_t)n_cells * kSurrHidden * 2));
    SHIP(hipMalloc(&o, (size_t)n_cells * kSurrOutPad * 4));
    SHIP(hipMemsetAsync(X, 0x3C, (size_t)n_cells * kSurrInPad * 2, s->stream));  // bf16 ~0.0115 everywhere
    const SurrNet net = s->net();
    surrogate_infer(net, X, n_cells, a0, a1, o, s->stream);  // warm-up
    hipEvent_t e0, e1;
    SHIP(hipEventCreate(&e0));
    SHIP(hipEventCreate(&e1));
    SHIP(hipEventRecord(e0, s->stream));
    for (int r = 0; r < reps; ++r) surrogate_infer(net, X, n_cells, a0, a1, o, s->stream);
    SHIP(hipEventRecord(e1, s->stream));
    SHIP(hipEventSynchronize(e1));
    float ms = 0.f;
    SHIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(X);
    (void)hipFree(a0);
    (void)hipFree(a1);
    (void)hipFree(o);
    *ms_per_pass = ms / (float)reps;
    return BCS_OK;
}
