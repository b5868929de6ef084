/*
 * spmd_mlp.hip -- the closure network: MLP 24-128-128-128-6 with SiLU,
 * BF16 operands and FP32 accumulation on the gfx950 matrix cores
 * (v_mfma_f32_16x16x32_bf16), forward, backward and Adam, behind a C ABI.
 *
 * Parameters (FP32 master copy, this order): W0[128][24] b0[128]
 * W1[128][128] b1[128] W2[128][128] b2[128] W3[6][128] b3[6]; y = W x + b.
 * The kernels read BF16 copies padded to the matrix-core shapes (inputs to
 * 32 columns, the last layer to 16 outputs) and, for the backward data
 * products, transposed copies.  Activations live in a per-handle workspace
 * sized for the largest batch seen, rows padded to a multiple of 64.
 *
 * Output: mean (3) and log-variance (3) of a diagonal Gaussian; the
 * training loss is its negative log-likelihood
 *   L = 1/M sum_m sum_c 1/2 [ (y - mu)^2 exp(-s) + s ],   s clamped to [-12, 12].
 * Inputs are standardised with per-feature mean/std held by the handle;
 * targets with a per-component std (mu and exp(s/2) are in those units).
 */
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/spmd.h"
#include "spmd_math.h"

namespace mlpk {

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kIn = 24, kInP = 32, kH = 128, kOut = 6, kOutP = 16, kOutK = 32;
constexpr int kParams = kH * kIn + kH + 2 * (kH * kH + kH) + kOut * kH + kOut;  /* 36,998 */
constexpr int kRowTile = 64;    /* rows per workgroup of the row GEMMs */
constexpr int kRedChunk = 512;  /* rows per workgroup of the weight-gradient products */

/* offsets of the FP32 parameter blocks */
constexpr int oW0 = 0, oB0 = oW0 + kH * kIn, oW1 = oB0 + kH, oB1 = oW1 + kH * kH, oW2 = oB1 + kH,
              oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kOut * kH;

__device__ __forceinline__ uint16_t f2bf(float f) {  /* round to nearest even */
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16 | ((u & 0xFFFF) ? 0x40 : 0));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

/* ---------------------------------------------------------------- kernels */

/* features (M x 24 FP32) -> standardised BF16 rows of 32 (zero padded);
 * rows M..Mp-1 are zero */
__global__ __launch_bounds__(256) void k_in(const float* __restrict__ x, const float* __restrict__ mu,
                                            const float* __restrict__ sd, uint16_t* __restrict__ x0, uint64_t M,
                                            uint64_t Mp) {
    const uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= Mp * kInP) return;
    const uint64_t m = e / kInP;
    const int c = (int)(e % kInP);
    float v = 0.0f;
    if (m < M && c < kIn) v = (x[m * kIn + c] - mu[c]) / sd[c];
    x0[e] = f2bf(v);
}

enum { EPI_SILU = 0, EPI_LINEAR = 1, EPI_F32 = 2 };

/* C[Mp x N] = A[Mp x K] B, with B given as its transpose Bt[N x K] (both
 * BF16 row-major).  One wave: 16 rows x N columns; a workgroup of four
 * waves: 64 rows.  Matrix-core operand layout (16x16x32): lane l supplies
 * A[l%16][8(l/16) .. +7] and B[8(l/16) .. +7][l%16] and holds
 * D[4(l/16) + r][l%16], r = 0..3.
 *   EPI_SILU   : z = C + b  -> Z (BF16), silu(z) -> A_out (BF16)
 *   EPI_LINEAR : C + b      -> F (FP32, ld N)
 *   EPI_F32    : C          -> F (FP32, ld N) */
template <int K, int N, int EPI>
__global__ __launch_bounds__(256) void k_gemm_nt(const uint16_t* __restrict__ A, const uint16_t* __restrict__ Bt,
                                                 const float* __restrict__ bias, uint16_t* __restrict__ Zo,
                                                 uint16_t* __restrict__ Ao, float* __restrict__ F) {
    constexpr int NT = N / 16;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t row0 = (uint64_t)blockIdx.x * kRowTile + (uint64_t)w * 16;
    const int li = lane & 15, lk = (lane >> 4) * 8;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < K; kk += 32) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(A + (row0 + li) * K + kk + lk);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(Bt + (size_t)(t * 16 + li) * K + kk + lk);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int col = t * 16 + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint64_t row = row0 + (lane >> 4) * 4 + r;
            float v = acc[t][r];
            if (EPI != EPI_F32) v += bias[col];
            if (EPI == EPI_SILU) {
                Zo[row * N + col] = f2bf(v);
                Ao[row * N + col] = f2bf(silu(v));
            } else {
                F[row * N + col] = v;
            }
        }
    }
}

/* Gaussian NLL of the output (Mp x 16 FP32, columns 0..2 mean, 3..5
 * log-variance) against y (M x 3, standardised by sd_y): gradient dZ
 * (BF16, Mp x 32, zero beyond column 5 and row M) and per-block loss sums */
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ O, const float* __restrict__ y,
                                              const float* __restrict__ sdy, uint16_t* __restrict__ dZ,
                                              float* __restrict__ lpart, uint64_t M, uint64_t Mp) {
    __shared__ float sred[256];
    const uint64_t m = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    float l = 0.0f;
    if (m < Mp) {
        float g[kOutK];
#pragma unroll
        for (int c = 0; c < kOutK; ++c) g[c] = 0.0f;
        if (m < M) {
            const float inv = 1.0f / (float)M;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float mu = O[m * kOutP + c];
                float s = O[m * kOutP + 3 + c];
                const bool clamped = s < -12.0f || s > 12.0f;
                s = fminf(fmaxf(s, -12.0f), 12.0f);
                const float r = y[m * 3 + c] / sdy[c] - mu;
                const float e = __expf(-s);
                l += 0.5f * (r * r * e + s);
                g[c] = -r * e * inv;
                g[3 + c] = clamped ? 0.0f : 0.5f * (1.0f - r * r * e) * inv;
            }
        }
#pragma unroll
        for (int c = 0; c < kOutK; ++c) dZ[m * kOutK + c] = f2bf(g[c]);
    }
    sred[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sred[threadIdx.x] += sred[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = sred[0];
}

/* dZ = dA * silu'(Z) (BF16) */
__global__ __launch_bounds__(256) void k_dsilu(const float* __restrict__ dA, const uint16_t* __restrict__ Z,
                                               uint16_t* __restrict__ dZ, uint64_t n) {
    const uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < n) dZ[e] = f2bf(dA[e] * dsilu(bf2f(Z[e])));
}

/* Partial weight and bias gradients of one workgroup's kRedChunk rows:
 * P[n][k] = sum_m dZ[m][n] X[m][k] (dZ: Mp x LDZ, first N columns used;
 * X: Mp x K), Pb[n] = sum_m dZ[m][n].  32-row blocks of both are staged in
 * LDS and read transposed into matrix-core operands. */
template <int N, int K, int LDZ>
__global__ __launch_bounds__(256) void k_wgrad(const uint16_t* __restrict__ dZ, const uint16_t* __restrict__ X,
                                               float* __restrict__ P, float* __restrict__ Pb) {
    constexpr int TN = N / 16, TK = K / 16, TILES = TN * TK, TPW = TILES / 4;
    static_assert(TILES % 4 == 0, "tiles must split evenly over the four waves");
    __shared__ uint16_t sZ[32][N + 2];
    __shared__ uint16_t sX[32][K + 2];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, lk = (lane >> 4) * 8;
    const uint64_t m0 = (uint64_t)blockIdx.x * kRedChunk;
    f32x4 acc[TPW];
#pragma unroll
    for (int q = 0; q < TPW; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.0f;
    for (int mb = 0; mb < kRedChunk; mb += 32) {
        __syncthreads();
        for (int e = tid; e < 32 * N; e += 256) sZ[e / N][e % N] = dZ[(m0 + mb + e / N) * LDZ + e % N];
        for (int e = tid; e < 32 * K; e += 256) sX[e / K][e % K] = X[(m0 + mb + e / K) * K + e % K];
        __syncthreads();
        if (tid < N) {
            for (int r = 0; r < 32; ++r) bsum += bf2f(sZ[r][tid]);
        }
#pragma unroll
        for (int q = 0; q < TPW; ++q) {
            const int tile = w * TPW + q;
            const int tn = tile / TK, tk = tile % TK;
            bf16x8 a, b;
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                a[v] = (short)sZ[lk + v][tn * 16 + li];
                b[v] = (short)sX[lk + v][tk * 16 + li];
            }
            acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[q], 0, 0, 0);
        }
    }
    float* Pw = P + (size_t)blockIdx.x * N * K;
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
        const int tile = w * TPW + q;
        const int tn = tile / TK, tk = tile % TK;
#pragma unroll
        for (int r = 0; r < 4; ++r) Pw[(size_t)(tn * 16 + (lane >> 4) * 4 + r) * K + tk * 16 + li] = acc[q][r];
    }
    if (tid < N) Pb[(size_t)blockIdx.x * N + tid] = bsum;
}

/* sum the partials of nb workgroups into the FP32 gradient block
 * g[n][k] (n < nv, k < kv of the padded N x K partials), bias after it */
__global__ __launch_bounds__(256) void k_reduce(const float* __restrict__ P, const float* __restrict__ Pb, int nb,
                                                int N, int K, int nv, int kv, float* __restrict__ gw,
                                                float* __restrict__ gb) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e < nv * kv) {
        const int n = e / kv, k = e % kv;
        float s = 0.0f;
        for (int b = 0; b < nb; ++b) s += P[(size_t)b * N * K + (size_t)n * K + k];
        gw[e] = s;
    } else if (e < nv * kv + nv) {
        const int n = e - nv * kv;
        float s = 0.0f;
        for (int b = 0; b < nb; ++b) s += Pb[(size_t)b * N + n];
        gb[n] = s;
    }
}

/* minibatch of B rows drawn uniformly with replacement (Philox, counter =
 * (row, step, 7, 0), key = seed) from X (N x 24) and Y (N x 3) */
__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ X, const float* __restrict__ Y, uint64_t N,
                                                float* __restrict__ Xb, float* __restrict__ Yb, uint32_t B,
                                                uint64_t seed, uint32_t step) {
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b >= B) return;
    uint32_t w[4];
    spmd::philox10(b, step, 7u, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), w);
    const uint64_t r = ((((uint64_t)w[0] << 32) | w[1]) % N);
    for (int c = 0; c < kIn; ++c) Xb[(uint64_t)b * kIn + c] = X[r * kIn + c];
    for (int c = 0; c < 3; ++c) Yb[(uint64_t)b * 3 + c] = Y[r * 3 + c];
}

/* Adam on the FP32 master parameters */
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                              const float* __restrict__ g, int n, float lr, float b1, float b2,
                                              float c1, float c2, float eps) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const float gg = g[e];
    const float mm = b1 * m[e] + (1.0f - b1) * gg;
    const float vv = b2 * v[e] + (1.0f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

/* BF16 copies of one layer's FP32 weights W[nv][kv]: padded W (N x K) and
 * its transpose Wt (KT x NT, with NT >= N rows of zero padding) */
__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ W, int nv, int kv, int N, int K,
                                              uint16_t* __restrict__ Wp, uint16_t* __restrict__ Wt, int NTP) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e < N * K) {
        const int n = e / K, k = e % K;
        Wp[e] = f2bf(n < nv && k < kv ? W[n * kv + k] : 0.0f);
    }
    if (Wt && e < K * NTP) {
        const int k = e / NTP, n = e % NTP;
        Wt[e] = f2bf(n < nv && k < kv ? W[n * kv + k] : 0.0f);
    }
}

/* ---------------------------------------------- fused inference pass */

/* The inference pass in one kernel: the transposed products D = W X^T, with
 * the weights as the matrix-core A operand (16 output neurons x 32 inputs)
 * and 16 rows as B.  Lane l then holds neurons 4(l/16)+r (r < 4) of row
 * l%16 in each 16-neuron tile, which is, up to an order of the inputs, the
 * B operand the next layer needs from that lane: inputs 8g..8g+7 of
 * 32-input block kb (g = l/16) are taken as neurons kperm(kb, g, 0..7),
 * and the next layer's weight columns are packed in that order.  The
 * activations never leave registers; the weights (BF16, 77,824 bytes)
 * and biases sit in LDS in the order the lanes read them. */
constexpr int fL0 = 0;                          /* 8 tiles x 64 lanes x 8 */
constexpr int fL1 = fL0 + 8 * 64 * 8;           /* 8 tiles x 4 blocks x 64 x 8 */
constexpr int fL2 = fL1 + 8 * 4 * 64 * 8;
constexpr int fL3 = fL2 + 8 * 4 * 64 * 8;       /* 4 blocks x 64 x 8 */
constexpr int fEnd = fL3 + 4 * 64 * 8;          /* 38,912 BF16 */
constexpr int fBias = 3 * kH + 16;              /* b0, b1, b2, b3 (padded) */
constexpr int kFwdThreads = 256;
constexpr size_t kFwdLds = (size_t)fEnd * 2 + (size_t)fBias * 4;

__host__ __device__ __forceinline__ int kperm(int kb, int g, int i) {
    return 32 * kb + (i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4));
}

/* FP32 parameters -> the lane-ready BF16 weights of k_forward_fused */
__global__ __launch_bounds__(256) void k_pack_fused(const float* __restrict__ p, uint16_t* __restrict__ wf) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= fEnd) return;
    const int i = e & 7, lane = (e >> 3) & 63, n16 = lane & 15, g = lane >> 4;
    float v = 0.0f;
    if (e < fL1) {
        const int n = (e >> 9) * 16 + n16, k = 8 * g + i;
        v = k < kIn ? p[oW0 + n * kIn + k] : 0.0f;
    } else if (e < fL3) {
        const int q = e - (e < fL2 ? fL1 : fL2), off = e < fL2 ? oW1 : oW2;
        const int n = (q >> 11) * 16 + n16, kb = (q >> 9) & 3;
        v = p[off + n * kH + kperm(kb, g, i)];
    } else {
        const int kb = (e - fL3) >> 9;
        v = n16 < kOut ? p[oW3 + n16 * kH + kperm(kb, g, i)] : 0.0f;
    }
    wf[e] = f2bf(v);
}

/* BF16 by the gfx950 conversion instruction (round to nearest even, as
 * f2bf for finite values) and SiLU with the hardware reciprocal: the
 * result is rounded to BF16 at once, far coarser than either step's error */
__device__ __forceinline__ short bf16_rne(float f) { return __builtin_bit_cast(short, (__bf16)f); }
__device__ __forceinline__ float silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

/* z + b, SiLU, BF16 into the next layer's B operands */
template <int NS>
__device__ __forceinline__ void fwd_epilogue(const f32x4 (&acc)[8][NS], const float* __restrict__ bias, int g,
                                             bf16x8 (&bh)[4][NS]) {
#pragma unroll
    for (int mt = 0; mt < 8; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float b = bias[mt * 16 + 4 * g + r];
#pragma unroll
            for (int ns = 0; ns < NS; ++ns) bh[mt >> 1][ns][(mt & 1) * 4 + r] = bf16_rne(silu_fast(acc[mt][ns][r] + b));
        }
}

/* M feature rows (FP32, 24 each) -> M x 6 outputs (FP32).  Each wave takes
 * 16 NS rows at a time, wave-strided over the grid (no synchronisation
 * after the weights are staged). */
template <int NS>
__global__ __launch_bounds__(kFwdThreads, 2) void k_forward_fused(const float* __restrict__ x,
                                                                   const float* __restrict__ mu,
                                                                   const float* __restrict__ sd,
                                                                   const uint16_t* __restrict__ wf,
                                                                   const float* __restrict__ p, float* __restrict__ out,
                                                                   uint64_t M) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint16_t* sw = reinterpret_cast<uint16_t*>(smem);
    float* sb = reinterpret_cast<float*>(smem + (size_t)fEnd * 2);
    const int tid = threadIdx.x;
    {
        const uint4* src = reinterpret_cast<const uint4*>(wf);
        uint4* dst = reinterpret_cast<uint4*>(sw);
        for (int e = tid; e < fEnd / 8; e += kFwdThreads) dst[e] = src[e];
        for (int e = tid; e < fBias; e += kFwdThreads) {
            float v = 0.0f;
            if (e < kH) v = p[oB0 + e];
            else if (e < 2 * kH) v = p[oB1 + e - kH];
            else if (e < 3 * kH) v = p[oB2 + e - 2 * kH];
            else if (e - 3 * kH < kOut) v = p[oB3 + e - 3 * kH];
            sb[e] = v;
        }
    }
    __syncthreads();
    const int lane = tid & 63, w = tid >> 6, g = lane >> 4, s = lane & 15;
    float m8[8], d8[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int c = 8 * g + i;
        m8[i] = c < kIn ? mu[c] : 0.0f;
        d8[i] = c < kIn ? sd[c] : 1.0f;
    }
    const uint64_t tiles = (M + 16 * NS - 1) / (16 * NS);
    const uint64_t stride = (uint64_t)gridDim.x * (kFwdThreads / 64);
    for (uint64_t t = (uint64_t)blockIdx.x * (kFwdThreads / 64) + w; t < tiles; t += stride) {
        const uint64_t row0 = t * 16 * NS;
        /* the weight reads are the same in every pass: an opaque base keeps
         * the compiler from hoisting all 76 of them out of the loop (304
         * registers, spilled to scratch) */
        uint32_t wo = 0;  /* an offset, not a pointer: the reads stay LDS reads */
        asm volatile("" : "+v"(wo));
        bf16x8 bx[NS];
#pragma unroll
        for (int ns = 0; ns < NS; ++ns) {
            const uint64_t row = row0 + ns * 16 + s;
            float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (row < M && g < 3) {
                const float4* xr = reinterpret_cast<const float4*>(x + row * kIn + 8 * g);
                const float4 u0 = xr[0], u1 = xr[1];
                v[0] = u0.x; v[1] = u0.y; v[2] = u0.z; v[3] = u0.w;
                v[4] = u1.x; v[5] = u1.y; v[6] = u1.z; v[7] = u1.w;
#pragma unroll
                for (int i = 0; i < 8; ++i) v[i] = (v[i] - m8[i]) / d8[i];
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) bx[ns][i] = bf16_rne(v[i]);
        }
        f32x4 acc[8][NS];
        bf16x8 bh[4][NS];
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(sw + wo + fL0 + (mt * 64 + lane) * 8);
#pragma unroll
            for (int ns = 0; ns < NS; ++ns)
                acc[mt][ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bx[ns], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        }
        fwd_epilogue<NS>(acc, sb, g, bh);
#pragma unroll
        for (int layer = 0; layer < 2; ++layer) {
#pragma unroll
            for (int mt = 0; mt < 8; ++mt)
#pragma unroll
                for (int ns = 0; ns < NS; ++ns) acc[mt][ns] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) {
                /* one 32-input block at a time: 8 weight operands live */
                uint32_t ko = (layer ? fL2 : fL1) + wo;
                asm volatile("" : "+v"(ko));
                const uint16_t* wl = sw + ko;
#pragma unroll
                for (int mt = 0; mt < 8; ++mt) {
                    const bf16x8 a = *reinterpret_cast<const bf16x8*>(wl + ((mt * 4 + kb) * 64 + lane) * 8);
#pragma unroll
                    for (int ns = 0; ns < NS; ++ns)
                        acc[mt][ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh[kb][ns], acc[mt][ns], 0, 0, 0);
                }
            }
            fwd_epilogue<NS>(acc, sb + (layer + 1) * kH, g, bh);
        }
        f32x4 o[NS];
#pragma unroll
        for (int ns = 0; ns < NS; ++ns) o[ns] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(sw + wo + fL3 + (kb * 64 + lane) * 8);
#pragma unroll
            for (int ns = 0; ns < NS; ++ns) o[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh[kb][ns], o[ns], 0, 0, 0);
        }
        if (g < 2) {
#pragma unroll
            for (int ns = 0; ns < NS; ++ns) {
                const uint64_t row = row0 + ns * 16 + s;
                if (row < M)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int c = 4 * g + r;
                        if (c < kOut) out[row * kOut + c] = o[ns][r] + sb[3 * kH + c];
                    }
            }
        }
    }
}

}  // namespace mlpk

using namespace mlpk;

struct spmd_mlp {
    std::string err;
    float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
    float *mu = nullptr, *sd = nullptr, *sdy = nullptr;
    uint16_t *W0 = nullptr, *W1 = nullptr, *W2 = nullptr, *W3 = nullptr;
    uint16_t *W1t = nullptr, *W2t = nullptr, *W3t = nullptr;
    uint16_t* Wf = nullptr;  /* lane-ready weights of the fused inference pass */
    int n_cu = 0;
    /* workspace */
    uint64_t cap = 0;
    uint16_t *X0 = nullptr, *Z1 = nullptr, *A1 = nullptr, *Z2 = nullptr, *A2 = nullptr, *Z3 = nullptr, *A3 = nullptr;
    uint16_t *dZ4 = nullptr, *dZ = nullptr;
    float *O = nullptr, *dA = nullptr, *P = nullptr, *Pb = nullptr, *lpart = nullptr;
    float *Xb = nullptr, *Yb = nullptr;  /* minibatch of spmd_mlp_fit */
    uint32_t bcap = 0;
    hipStream_t st = nullptr;
    long step = 0;
};

namespace {

int mlp_fail(spmd_mlp* h, int code, const std::string& m) {
    if (h) h->err = m;
    return code;
}

#define MCHK(h, x)                                                                  \
    do {                                                                            \
        const hipError_t e_ = (x);                                                  \
        if (e_ != hipSuccess) return mlp_fail((h), SPMD_ERR_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class T>
hipError_t dal(T** p, size_t n) {
    return hipMalloc((void**)p, n * sizeof(T) + 64);
}
template <class T>
void dfr(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

uint64_t pad64(uint64_t M) { return (M + kRedChunk - 1) / kRedChunk * kRedChunk; }

int ensure(spmd_mlp* h, uint64_t M) {
    const uint64_t Mp = pad64(M);
    if (Mp <= h->cap) return SPMD_OK;
    dfr(h->X0); dfr(h->Z1); dfr(h->A1); dfr(h->Z2); dfr(h->A2); dfr(h->Z3); dfr(h->A3);
    dfr(h->dZ4); dfr(h->dZ); dfr(h->O); dfr(h->dA); dfr(h->P); dfr(h->Pb); dfr(h->lpart);
    MCHK(h, dal(&h->X0, Mp * kInP));
    MCHK(h, dal(&h->Z1, Mp * kH)); MCHK(h, dal(&h->A1, Mp * kH));
    MCHK(h, dal(&h->Z2, Mp * kH)); MCHK(h, dal(&h->A2, Mp * kH));
    MCHK(h, dal(&h->Z3, Mp * kH)); MCHK(h, dal(&h->A3, Mp * kH));
    MCHK(h, dal(&h->dZ4, Mp * kOutK)); MCHK(h, dal(&h->dZ, Mp * kH));
    MCHK(h, dal(&h->O, Mp * kOutP)); MCHK(h, dal(&h->dA, Mp * kH));
    const uint64_t nb = Mp / kRedChunk;
    MCHK(h, dal(&h->P, nb * kH * kH)); MCHK(h, dal(&h->Pb, nb * kH));
    MCHK(h, dal(&h->lpart, Mp / 256 + 1));
    h->cap = Mp;
    return SPMD_OK;
}

int pack_all(spmd_mlp* h) {
    const float* p = h->p;
    auto g = [](int n) { return dim3((n + 255) / 256); };
    hipLaunchKernelGGL(k_pack, g(kH * kInP), dim3(256), 0, h->st, p + oW0, kH, kIn, kH, kInP, h->W0,
                       (uint16_t*)nullptr, 0);
    hipLaunchKernelGGL(k_pack, g(kH * kH), dim3(256), 0, h->st, p + oW1, kH, kH, kH, kH, h->W1, h->W1t, kH);
    hipLaunchKernelGGL(k_pack, g(kH * kH), dim3(256), 0, h->st, p + oW2, kH, kH, kH, kH, h->W2, h->W2t, kH);
    hipLaunchKernelGGL(k_pack, g(kH * kOutK), dim3(256), 0, h->st, p + oW3, kOut, kH, kOutP, kH, h->W3, h->W3t,
                       kOutK);
    hipLaunchKernelGGL(k_pack_fused, g(fEnd), dim3(256), 0, h->st, p, h->Wf);
    MCHK(h, hipGetLastError());
    return SPMD_OK;
}

/* the fused inference pass of M rows on stream st: x (M x 24) -> out (M x 6) */
constexpr int kFwdRows = 2;  /* 16-row tiles per wave and pass (168 VGPRs, no spills) */
int forward_fused(spmd_mlp* h, const float* x, float* out, uint64_t M, hipStream_t st) {
    static bool attr = false;
    if (!attr) {
        MCHK(h, hipFuncSetAttribute((const void*)k_forward_fused<kFwdRows>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFwdLds));
        attr = true;
    }
    const uint64_t tiles = (M + 16 * kFwdRows - 1) / (16 * kFwdRows);
    const uint64_t want = (tiles + kFwdThreads / 64 - 1) / (kFwdThreads / 64);
    const uint64_t cap = 2ull * (uint64_t)(h->n_cu > 0 ? h->n_cu : 256);
    const unsigned nb = (unsigned)(want < cap ? want : cap);
    hipLaunchKernelGGL(k_forward_fused<kFwdRows>, dim3(nb), dim3(kFwdThreads), kFwdLds, st, x, h->mu, h->sd, h->Wf,
                       h->p, out, M);
    MCHK(h, hipGetLastError());
    return SPMD_OK;
}

/* forward of M rows already converted into X0 */
int forward_rows(spmd_mlp* h, uint64_t Mp) {
    const dim3 g((unsigned)(Mp / kRowTile)), b(256);
    const float* p = h->p;
    hipLaunchKernelGGL((k_gemm_nt<kInP, kH, EPI_SILU>), g, b, 0, h->st, h->X0, h->W0, p + oB0, h->Z1, h->A1,
                       (float*)nullptr);
    hipLaunchKernelGGL((k_gemm_nt<kH, kH, EPI_SILU>), g, b, 0, h->st, h->A1, h->W1, p + oB1, h->Z2, h->A2,
                       (float*)nullptr);
    hipLaunchKernelGGL((k_gemm_nt<kH, kH, EPI_SILU>), g, b, 0, h->st, h->A2, h->W2, p + oB2, h->Z3, h->A3,
                       (float*)nullptr);
    hipLaunchKernelGGL((k_gemm_nt<kH, kOutP, EPI_LINEAR>), g, b, 0, h->st, h->A3, h->W3, p + oB3,
                       (uint16_t*)nullptr, (uint16_t*)nullptr, h->O);
    MCHK(h, hipGetLastError());
    return SPMD_OK;
}

/* gradient of the mean loss into h->g; loss sum into *loss */
int backward_rows(spmd_mlp* h, const float* y, uint64_t M, uint64_t Mp, double* loss) {
    const unsigned nl = (unsigned)((Mp + 255) / 256);
    hipLaunchKernelGGL(k_loss, dim3(nl), dim3(256), 0, h->st, h->O, y, h->sdy, h->dZ4, h->lpart, M, Mp);
    const dim3 gr((unsigned)(Mp / kRowTile)), b(256);
    const unsigned nb = (unsigned)(Mp / kRedChunk);
    float* g = h->g;
    auto red = [&](int N, int K, int nv, int kv, float* gw, float* gb) {
        const int n = nv * kv + nv;
        hipLaunchKernelGGL(k_reduce, dim3((n + 255) / 256), dim3(256), 0, h->st, h->P, h->Pb, (int)nb, N, K, nv,
                           kv, gw, gb);
    };
    const uint64_t nh = Mp * kH;
    const unsigned ne = (unsigned)((nh + 255) / 256);
    /* layer 3 (linear): dW3, dA3 = dZ4 W3 */
    hipLaunchKernelGGL((k_wgrad<kOutP, kH, kOutK>), dim3(nb), b, 0, h->st, h->dZ4, h->A3, h->P, h->Pb);
    red(kOutP, kH, kOut, kH, g + oW3, g + oB3);
    hipLaunchKernelGGL((k_gemm_nt<kOutK, kH, EPI_F32>), gr, b, 0, h->st, h->dZ4, h->W3t, (const float*)nullptr,
                       (uint16_t*)nullptr, (uint16_t*)nullptr, h->dA);
    /* layer 2 */
    hipLaunchKernelGGL(k_dsilu, dim3(ne), b, 0, h->st, h->dA, h->Z3, h->dZ, nh);
    hipLaunchKernelGGL((k_wgrad<kH, kH, kH>), dim3(nb), b, 0, h->st, h->dZ, h->A2, h->P, h->Pb);
    red(kH, kH, kH, kH, g + oW2, g + oB2);
    hipLaunchKernelGGL((k_gemm_nt<kH, kH, EPI_F32>), gr, b, 0, h->st, h->dZ, h->W2t, (const float*)nullptr,
                       (uint16_t*)nullptr, (uint16_t*)nullptr, h->dA);
    /* layer 1 */
    hipLaunchKernelGGL(k_dsilu, dim3(ne), b, 0, h->st, h->dA, h->Z2, h->dZ, nh);
    hipLaunchKernelGGL((k_wgrad<kH, kH, kH>), dim3(nb), b, 0, h->st, h->dZ, h->A1, h->P, h->Pb);
    red(kH, kH, kH, kH, g + oW1, g + oB1);
    hipLaunchKernelGGL((k_gemm_nt<kH, kH, EPI_F32>), gr, b, 0, h->st, h->dZ, h->W1t, (const float*)nullptr,
                       (uint16_t*)nullptr, (uint16_t*)nullptr, h->dA);
    /* layer 0 */
    hipLaunchKernelGGL(k_dsilu, dim3(ne), b, 0, h->st, h->dA, h->Z1, h->dZ, nh);
    hipLaunchKernelGGL((k_wgrad<kH, kInP, kH>), dim3(nb), b, 0, h->st, h->dZ, h->X0, h->P, h->Pb);
    red(kH, kInP, kH, kIn, g + oW0, g + oB0);
    MCHK(h, hipGetLastError());
    if (loss) {
        std::vector<float> lp(nl);
        MCHK(h, hipMemcpyAsync(lp.data(), h->lpart, nl * 4ull, hipMemcpyDeviceToHost, h->st));
        MCHK(h, hipStreamSynchronize(h->st));
        double s = 0.0;
        for (float v : lp) s += v;
        *loss = s / (double)M;
    }
    return SPMD_OK;
}

int to_rows(spmd_mlp* h, const float* x, uint64_t M, uint64_t* Mp_out) {
    if (int rc = ensure(h, M)) return rc;
    const uint64_t Mp = pad64(M);
    const uint64_t n = Mp * kInP;
    hipLaunchKernelGGL(k_in, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->st, x, h->mu, h->sd, h->X0, M, Mp);
    MCHK(h, hipGetLastError());
    *Mp_out = Mp;
    return SPMD_OK;
}

}  // namespace

extern "C" {

int spmd_mlp_create(uint64_t seed, spmd_mlp** out) {
    if (!out) return SPMD_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SPMD_ERR_NO_DEVICE;
    spmd_mlp* h = new spmd_mlp();
    auto bad = [&](hipError_t e) { return e != hipSuccess; };
    int dev = 0;
    if (bad(hipGetDevice(&dev)) || bad(hipDeviceGetAttribute(&h->n_cu, hipDeviceAttributeMultiprocessorCount, dev))) {
        delete h;
        return SPMD_ERR_HIP;
    }
    if (bad(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking)) || bad(dal(&h->p, kParams)) ||
        bad(dal(&h->m, kParams)) || bad(dal(&h->v, kParams)) || bad(dal(&h->g, kParams)) || bad(dal(&h->mu, kIn)) ||
        bad(dal(&h->sd, kIn)) || bad(dal(&h->sdy, 3)) || bad(dal(&h->W0, kH * kInP)) || bad(dal(&h->W1, kH * kH)) ||
        bad(dal(&h->W2, kH * kH)) || bad(dal(&h->W3, kOutP * kH)) || bad(dal(&h->W1t, kH * kH)) ||
        bad(dal(&h->W2t, kH * kH)) || bad(dal(&h->W3t, kH * kOutK)) || bad(dal(&h->Wf, fEnd))) {
        spmd_mlp_destroy(h);
        return SPMD_ERR_OOM;
    }
    /* Kaiming-uniform weights (bound sqrt(6/fan_in)), zero biases, from a
     * 64-bit LCG (Knuth's MMIX constants) seeded by `seed` */
    std::vector<float> p(kParams, 0.0f);
    uint64_t s = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto uni = [&]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(s >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int n, int fan_in) {
        const double bnd = sqrt(6.0 / fan_in);
        for (int i = 0; i < n; ++i) p[off + i] = (float)((2.0 * uni() - 1.0) * bnd);
    };
    fill(oW0, kH * kIn, kIn);
    fill(oW1, kH * kH, kH);
    fill(oW2, kH * kH, kH);
    fill(oW3, kOut * kH, kH);
    for (int i = 0; i < kOut * kH; ++i) p[oW3 + i] *= 0.1f;
    std::vector<float> one(kIn, 1.0f), zero(kIn, 0.0f);
    const bool ok = hipMemcpy(h->p, p.data(), kParams * 4ull, hipMemcpyHostToDevice) == hipSuccess &&
                    hipMemset(h->m, 0, kParams * 4ull) == hipSuccess &&
                    hipMemset(h->v, 0, kParams * 4ull) == hipSuccess &&
                    hipMemcpy(h->mu, zero.data(), kIn * 4ull, hipMemcpyHostToDevice) == hipSuccess &&
                    hipMemcpy(h->sd, one.data(), kIn * 4ull, hipMemcpyHostToDevice) == hipSuccess &&
                    hipMemcpy(h->sdy, one.data(), 3 * 4ull, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok || pack_all(h) || hipStreamSynchronize(h->st) != hipSuccess) {
        spmd_mlp_destroy(h);
        return SPMD_ERR_HIP;
    }
    *out = h;
    return SPMD_OK;
}

int spmd_mlp_destroy(spmd_mlp* h) {
    if (!h) return SPMD_OK;
    if (h->st) (void)hipStreamSynchronize(h->st);
    dfr(h->p); dfr(h->m); dfr(h->v); dfr(h->g); dfr(h->mu); dfr(h->sd); dfr(h->sdy);
    dfr(h->W0); dfr(h->W1); dfr(h->W2); dfr(h->W3); dfr(h->W1t); dfr(h->W2t); dfr(h->W3t); dfr(h->Wf);
    dfr(h->X0); dfr(h->Z1); dfr(h->A1); dfr(h->Z2); dfr(h->A2); dfr(h->Z3); dfr(h->A3);
    dfr(h->dZ4); dfr(h->dZ); dfr(h->O); dfr(h->dA); dfr(h->P); dfr(h->Pb); dfr(h->lpart);
    dfr(h->Xb); dfr(h->Yb);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
    return SPMD_OK;
}

const char* spmd_mlp_last_error(const spmd_mlp* h) { return h ? h->err.c_str() : "null handle"; }

uint32_t spmd_mlp_num_params(void) { return (uint32_t)kParams; }

int spmd_mlp_get_params(spmd_mlp* h, float* host) {
    if (!h || !host) return SPMD_ERR_ARG;
    MCHK(h, hipStreamSynchronize(h->st));
    MCHK(h, hipMemcpy(host, h->p, kParams * 4ull, hipMemcpyDeviceToHost));
    return SPMD_OK;
}

int spmd_mlp_set_params(spmd_mlp* h, const float* host) {
    if (!h || !host) return SPMD_ERR_ARG;
    MCHK(h, hipStreamSynchronize(h->st));
    MCHK(h, hipMemcpy(h->p, host, kParams * 4ull, hipMemcpyHostToDevice));
    if (int rc = pack_all(h)) return rc;
    MCHK(h, hipStreamSynchronize(h->st));
    return SPMD_OK;
}

int spmd_mlp_set_norm(spmd_mlp* h, const float* mu_x, const float* sd_x, const float* sd_y) {
    if (!h || !mu_x || !sd_x || !sd_y) return SPMD_ERR_ARG;
    for (int i = 0; i < kIn; ++i)
        if (!(sd_x[i] > 0.0f)) return mlp_fail(h, SPMD_ERR_ARG, "feature std must be positive");
    for (int i = 0; i < 3; ++i)
        if (!(sd_y[i] > 0.0f)) return mlp_fail(h, SPMD_ERR_ARG, "target std must be positive");
    MCHK(h, hipStreamSynchronize(h->st));
    MCHK(h, hipMemcpy(h->mu, mu_x, kIn * 4ull, hipMemcpyHostToDevice));
    MCHK(h, hipMemcpy(h->sd, sd_x, kIn * 4ull, hipMemcpyHostToDevice));
    MCHK(h, hipMemcpy(h->sdy, sd_y, 3 * 4ull, hipMemcpyHostToDevice));
    return SPMD_OK;
}

int spmd_mlp_get_norm(const spmd_mlp* h, float* mu_x, float* sd_x, float* sd_y) {
    if (!h) return SPMD_ERR_ARG;
    if (hipStreamSynchronize(h->st) != hipSuccess) return SPMD_ERR_HIP;
    if (mu_x && hipMemcpy(mu_x, h->mu, kIn * 4ull, hipMemcpyDeviceToHost) != hipSuccess) return SPMD_ERR_HIP;
    if (sd_x && hipMemcpy(sd_x, h->sd, kIn * 4ull, hipMemcpyDeviceToHost) != hipSuccess) return SPMD_ERR_HIP;
    if (sd_y && hipMemcpy(sd_y, h->sdy, 3 * 4ull, hipMemcpyDeviceToHost) != hipSuccess) return SPMD_ERR_HIP;
    return SPMD_OK;
}

}  // extern "C"

/* The forward pass enqueued on the caller's stream (no synchronisation):
 * used by the closure loop, which runs features -> network -> kick back to
 * back on its own stream.  The workspace is grown first (host call). */
int spmd_mlp_forward_on(spmd_mlp* h, const float* x_dev, float* out_dev, uint64_t M, hipStream_t st) {
    if (!h || !x_dev || !out_dev) return SPMD_ERR_ARG;
    if (M == 0) return SPMD_OK;
    MCHK(h, hipStreamSynchronize(h->st));  /* parameters packed by training are complete */
    return forward_fused(h, x_dev, out_dev, M, st);
}

extern "C" {

int spmd_mlp_forward(spmd_mlp* h, const float* x_dev, float* out_dev, uint64_t M) {
    if (!h || !x_dev || !out_dev) return SPMD_ERR_ARG;
    if (M == 0) return SPMD_OK;
    if (int rc = forward_fused(h, x_dev, out_dev, M, h->st)) return rc;
    MCHK(h, hipStreamSynchronize(h->st));
    return SPMD_OK;
}

/* the training-path forward (four row-tiled products through the
 * workspace, the activations kept for the backward pass) */
int spmd_mlp_forward_unfused(spmd_mlp* h, const float* x_dev, float* out_dev, uint64_t M) {
    if (!h || !x_dev || !out_dev) return SPMD_ERR_ARG;
    if (M == 0) return SPMD_OK;
    uint64_t Mp;
    if (int rc = to_rows(h, x_dev, M, &Mp)) return rc;
    if (int rc = forward_rows(h, Mp)) return rc;
    MCHK(h, hipMemcpy2DAsync(out_dev, kOut * 4, h->O, kOutP * 4, kOut * 4, M, hipMemcpyDeviceToDevice, h->st));
    MCHK(h, hipStreamSynchronize(h->st));
    return SPMD_OK;
}

int spmd_mlp_gradient(spmd_mlp* h, const float* x_dev, const float* y_dev, uint64_t M, float* grad_host,
                      double* loss) {
    if (!h || !x_dev || !y_dev || M == 0) return SPMD_ERR_ARG;
    uint64_t Mp;
    if (int rc = to_rows(h, x_dev, M, &Mp)) return rc;
    if (int rc = forward_rows(h, Mp)) return rc;
    if (int rc = backward_rows(h, y_dev, M, Mp, loss)) return rc;
    if (grad_host) MCHK(h, hipMemcpyAsync(grad_host, h->g, kParams * 4ull, hipMemcpyDeviceToHost, h->st));
    MCHK(h, hipStreamSynchronize(h->st));
    return SPMD_OK;
}

int spmd_mlp_train_step(spmd_mlp* h, const float* x_dev, const float* y_dev, uint64_t M, double lr, double* loss) {
    if (!h || !x_dev || !y_dev || M == 0 || !(lr > 0)) return SPMD_ERR_ARG;
    uint64_t Mp;
    if (int rc = to_rows(h, x_dev, M, &Mp)) return rc;
    if (int rc = forward_rows(h, Mp)) return rc;
    if (int rc = backward_rows(h, y_dev, M, Mp, loss)) return rc;
    h->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)h->step), c2 = 1.0f - powf(b2, (float)h->step);
    hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, h->st, h->p, h->m, h->v, h->g, kParams,
                       (float)lr, b1, b2, c1, c2, 1e-8f);
    MCHK(h, hipGetLastError());
    if (int rc = pack_all(h)) return rc;
    if (!loss) MCHK(h, hipStreamSynchronize(h->st));
    return SPMD_OK;
}

int spmd_mlp_fit(spmd_mlp* h, const float* x_dev, const float* y_dev, uint64_t N, uint32_t batch, uint32_t steps,
                 double lr, uint64_t seed, double* loss_hist) {
    if (!h || !x_dev || !y_dev || N == 0 || batch == 0 || !(lr > 0)) return SPMD_ERR_ARG;
    if (batch > h->bcap) {
        dfr(h->Xb); dfr(h->Yb);
        MCHK(h, dal(&h->Xb, (size_t)batch * kIn));
        MCHK(h, dal(&h->Yb, (size_t)batch * 3));
        h->bcap = batch;
    }
    for (uint32_t it = 0; it < steps; ++it) {
        hipLaunchKernelGGL(k_gather, dim3((batch + 255) / 256), dim3(256), 0, h->st, x_dev, y_dev, N, h->Xb, h->Yb,
                           batch, seed, (uint32_t)h->step);
        MCHK(h, hipGetLastError());
        double l = 0.0;
        if (int rc = spmd_mlp_train_step(h, h->Xb, h->Yb, batch, lr, loss_hist ? &l : nullptr)) return rc;
        if (loss_hist) loss_hist[it] = l;
    }
    MCHK(h, hipStreamSynchronize(h->st));
    return SPMD_OK;
}

/* device buffers for callers without a device allocator of their own */
int spmd_dev_alloc(uint64_t bytes, void** p) {
    if (!p) return SPMD_ERR_ARG;
    *p = nullptr;
    return hipMalloc(p, bytes ? bytes : 1) == hipSuccess ? SPMD_OK : SPMD_ERR_OOM;
}
int spmd_dev_free(void* p) { return !p || hipFree(p) == hipSuccess ? SPMD_OK : SPMD_ERR_HIP; }
int spmd_dev_upload(void* dst, const void* src, uint64_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? SPMD_OK : SPMD_ERR_HIP;
}
int spmd_dev_download(void* dst, const void* src, uint64_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? SPMD_OK : SPMD_ERR_HIP;
}

}  // extern "C"
