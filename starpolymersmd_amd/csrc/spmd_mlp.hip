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
constexpr int kFwdThreads = 512;  /* two workgroups per CU (LDS): four waves per SIMD */
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
__global__ __launch_bounds__(kFwdThreads, 4) void k_forward_fused(const float* __restrict__ x,
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

/* ---------------------------------------------- fused training step */

/* Transposed BF16 weights for the backward data products dA = W^T dZ, laid
 * out like the forward pack with rows and columns swapped: tile mt holds
 * the layer's inputs 16mt..16mt+15 as the matrix-core A operand's rows and
 * block kb its outputs kperm(kb, g, 0..7) as the K values, so the output
 * gradients in the order the previous product leaves them (lane l: outputs
 * 4(l/16)+r of row l%16 in each 16-output tile) are the B operand as they
 * stand. */
constexpr int bL3 = 0;                      /* W3^T: 8 tiles x 64 lanes x 8 (outputs padded to 32) */
constexpr int bL2 = bL3 + 8 * 64 * 8;       /* W2^T: 8 tiles x 4 blocks x 64 x 8 */
constexpr int bL1 = bL2 + 8 * 4 * 64 * 8;   /* W1^T */
constexpr int bEnd = bL1 + 8 * 4 * 64 * 8;  /* 36,864 BF16 */

__global__ __launch_bounds__(256) void k_pack_bwd(const float* __restrict__ p, uint16_t* __restrict__ wb) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= bEnd) return;
    const int i = e & 7, lane = (e >> 3) & 63, s = lane & 15, g = lane >> 4;
    float v;
    if (e < bL2) {
        const int mt = e >> 9, o = kperm(0, g, i);
        v = o < kOut ? p[oW3 + o * kH + 16 * mt + s] : 0.0f;
    } else {
        const int q = e - (e < bL1 ? bL2 : bL1), off = e < bL1 ? oW2 : oW1;
        const int mt = q >> 11, kb = (q >> 9) & 3;
        v = p[off + kperm(kb, g, i) * kH + 16 * mt + s];
    }
    wb[e] = f2bf(v);
}

/* the block, lane group and position with kperm(kb, g, i) = k */
__device__ __forceinline__ void kperm_inv(int k, int& kb, int& g, int& i) {
    kb = k >> 5;
    const int j = k & 31;
    g = (j & 15) >> 2;
    i = (j & 3) + (j >= 16 ? 4 : 0);
}

/* where weight parameter e sits in the forward pack (pf) and in the
 * transposed pack (pb); -1 where it has no place (biases, W0 in pb) */
__device__ __forceinline__ void pack_pos(int e, int& pf, int& pb) {
    pf = -1;
    pb = -1;
    int kb, g, i;
    if (e < oB0) {
        const int n = e / kIn, k = e % kIn;
        pf = fL0 + ((n >> 4) * 64 + 16 * (k >> 3) + (n & 15)) * 8 + (k & 7);
    } else if ((e >= oW1 && e < oB1) || (e >= oW2 && e < oB2)) {
        const bool l1 = e < oB1;
        const int q = e - (l1 ? oW1 : oW2), n = q >> 7, k = q & (kH - 1);
        kperm_inv(k, kb, g, i);
        pf = (l1 ? fL1 : fL2) + (((n >> 4) * 4 + kb) * 64 + 16 * g + (n & 15)) * 8 + i;
        kperm_inv(n, kb, g, i);
        pb = (l1 ? bL1 : bL2) + (((k >> 4) * 4 + kb) * 64 + 16 * g + (k & 15)) * 8 + i;
    } else if (e >= oW3 && e < oB3) {
        const int q = e - oW3, o = q >> 7, k = q & (kH - 1);
        kperm_inv(k, kb, g, i);
        pf = fL3 + (kb * 64 + 16 * g + o) * 8 + i;
        kperm_inv(o, kb, g, i);  /* kb = 0: o < 6 */
        pb = bL3 + ((k >> 4) * 64 + 16 * g + (k & 15)) * 8 + i;
    }
}

/* One workgroup of eight waves (two per SIMD) takes 64-row chunks and
 * keeps the weight gradients of all its chunks in registers.  Waves 0-3
 * work on rows (16 each), waves 4-7 on the weight gradients; the two roles
 * run their own loops over the same chunks and meet at the same eight
 * barriers per chunk, so each role's registers stay within half a SIMD's
 * and one role's products run while the other waits on memory.
 *   rows:  forward as k_forward_fused (weights in LDS), SiLU' kept per
 *          value, every layer's input written transposed (feature-major) to
 *          the staging area; the Gaussian NLL and its gradient dZ4 (lanes
 *          exchange the mean and log-variance of a row); backward dA =
 *          W^T dZ with the transposed pack read from L2, dZ = dA SiLU'(z),
 *          each dZ^T staged over the activation it replaces once the
 *          gradient waves have read that
 *   dW:    dW += dZ^T a over the chunk's rows (rows as the products' K
 *          dimension, both operands read from the staging); gradient wave q
 *          owns output tiles 2q, 2q+1 of each layer (input tiles 2q, 2q+1
 *          of dW3); bias sums from the staged dZ^T
 * At the end each workgroup writes its partial gradient (parameter order)
 * and its loss sum; k_train_finish adds them up. */
constexpr int kTrThreads = 512;
constexpr int kTrRows = 64;
constexpr int kTS = kTrRows + 8;  /* staging row stride (BF16): 144 bytes, 16-byte reads conflict-free */
constexpr int sA0o = 0;                      /* inputs^T (32 features) */
constexpr int sA1o = sA0o + kInP * kTS;      /* a1^T, then dZ1^T */
constexpr int sA2o = sA1o + kH * kTS;        /* a2^T, then dZ2^T */
constexpr int sA3o = sA2o + kH * kTS;        /* a3^T, then dZ3^T */
constexpr int sD4o = sA3o + kH * kTS;        /* dZ4^T (16 outputs) */
constexpr int sEnd = sD4o + kOutP * kTS;
/* LDS: the staging first (its constant offsets then fit the 16-bit offset
 * field of the LDS instructions), the forward weights, the biases */
constexpr size_t kTrLdsS = (size_t)sEnd * 2;
constexpr size_t kTrLds = kTrLdsS + (size_t)fEnd * 2 + (size_t)fBias * 4 + (4 + 2 * kInP) * sizeof(float);
static_assert(kTrLds <= 160 * 1024, "LDS of one CU");

struct TrainArgs {
    const float* x;   /* rows of 24 features */
    const float* y;   /* rows of 3 targets */
    uint64_t n_src;   /* > 0: row b of the batch is source row philox(b, step) mod n_src */
    uint64_t M;       /* batch rows */
    uint64_t seed;
    uint32_t step;
    const float* mu;
    const float* sd;
    const float* sdy;
    const uint16_t* wf;
    const uint16_t* wb;
    const float* p;
    float* part;      /* [gridDim.x][kParams] */
    float* lpart;     /* [gridDim.x] */
};

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) { return *reinterpret_cast<const bf16x8*>(p); }

/* z + b: SiLU into the next product's B operands and to the staging,
 * SiLU'(z) kept (BF16) for the backward pass */
__device__ __forceinline__ void tr_fwd_epilogue(const f32x4 (&acc)[8], const float* __restrict__ bias, int g,
                                                uint16_t* sT, int rw, bf16x8 (&bh)[4], bf16x8 (&ds)[4]) {
#pragma unroll
    for (int mt = 0; mt < 8; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int n = mt * 16 + 4 * g + r;
            const float z = acc[mt][r] + bias[n];
            const float sg = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
            const float a = z * sg;  /* SiLU; its derivative sg (1 + z (1 - sg)) = sg + a (1 - sg) */
            const short av = bf16_rne(a);
            bh[mt >> 1][(mt & 1) * 4 + r] = av;
            ds[mt >> 1][(mt & 1) * 4 + r] = bf16_rne(__builtin_fmaf(a, 1.0f - sg, sg));
            sT[n * kTS + rw] = (uint16_t)av;
        }
}

/* dZ = dA SiLU'(z) (BF16), the next backward product's B operands */
__device__ __forceinline__ void tr_bwd_epilogue(const f32x4 (&acc)[8], const bf16x8 (&ds)[4], bf16x8 (&bd)[4]) {
#pragma unroll
    for (int mt = 0; mt < 8; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int q = (mt & 1) * 4 + r;
            bd[mt >> 1][q] = bf16_rne(acc[mt][r] * bf2f((uint16_t)ds[mt >> 1][q]));
        }
}

/* dZ^T to the staging from the B operands */
__device__ __forceinline__ void tr_stage(const bf16x8 (&v)[4], int g, uint16_t* sT, int rw) {
#pragma unroll
    for (int mt = 0; mt < 8; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) sT[(mt * 16 + 4 * g + r) * kTS + rw] = (uint16_t)v[mt >> 1][(mt & 1) * 4 + r];
}

/* acc[8 jn + tk] += dZ^T (features 16(2w + jn) + .) a^T (features 16 tk + .)
 * over the chunk's 64 rows */
__device__ __forceinline__ void tr_wgrad(f32x4 (&acc)[16], const uint16_t* sD, const uint16_t* sA, int w, int g,
                                         int s) {
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        const int ro = 32 * ks + 8 * g;
        const bf16x8 a0 = ld8(sD + (32 * w + s) * kTS + ro), a1 = ld8(sD + (32 * w + 16 + s) * kTS + ro);
#pragma unroll
        for (int tk = 0; tk < 8; ++tk) {
            const bf16x8 b = ld8(sA + (16 * tk + s) * kTS + ro);
            acc[tk] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b, acc[tk], 0, 0, 0);
            acc[8 + tk] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b, acc[8 + tk], 0, 0, 0);
        }
    }
}

/* half the chunk's rows (32) of staged feature n */
__device__ __forceinline__ float tr_row_sum(const uint16_t* sT, int n, int h) {
    float t = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const bf16x8 v = ld8(sT + n * kTS + 32 * h + 8 * q);
#pragma unroll
        for (int i = 0; i < 8; ++i) t += bf2f((uint16_t)v[i]);
    }
    return t;
}

/* The row waves' loop (16 rows each; w = 0..3). */
__device__ __forceinline__ void tr_rows(const TrainArgs& a, uint16_t* st, const uint16_t* sw, const float* sb,
                                        float* sred, const float* snorm, int w, int lane) {
    const float sy[3] = {a.sdy[0], a.sdy[1], a.sdy[2]};
    const float inv = 1.0f / (float)a.M;
    float lsum = 0.0f;
    const uint64_t chunks = (a.M + kTrRows - 1) / kTrRows;
    for (uint64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        /* opaque bases keep the weight reads and the staging addresses
         * inside the loop (hoisted, each base + constant would take a
         * register for the whole loop instead of an offset field) */
        uint32_t wo = 0, so = 0;
        int ln = lane;
        asm volatile("" : "+v"(wo));
        asm volatile("" : "+v"(so));
        asm volatile("" : "+v"(ln));
        uint16_t* sl = st + so;
        const float* sbl = sb + so;
        const int g = ln >> 4, s = ln & 15, rw = 16 * w + s;
        const uint64_t row = c * kTrRows + rw;
        const bool live = row < a.M;
        uint64_t src = row;
        if (a.n_src && live) {
            uint32_t q[4];
            spmd::philox10((uint32_t)row, a.step, 7u, 0u, (uint32_t)a.seed, (uint32_t)(a.seed >> 32), q);
            src = ((((uint64_t)q[0] << 32) | q[1]) % a.n_src);
        }
        float yv[3] = {0.0f, 0.0f, 0.0f};
        if (live)
#pragma unroll
            for (int k = 0; k < 3; ++k) yv[k] = a.y[src * 3 + k] / sy[k];
        bf16x8 bx;
        {
            float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (live && g < 3) {
                const float4* xr = reinterpret_cast<const float4*>(a.x + src * kIn + 8 * g);
                const float4 u0 = xr[0], u1 = xr[1];
                v[0] = u0.x; v[1] = u0.y; v[2] = u0.z; v[3] = u0.w;
                v[4] = u1.x; v[5] = u1.y; v[6] = u1.z; v[7] = u1.w;
#pragma unroll
                for (int i = 0; i < 8; ++i) v[i] = (v[i] - snorm[so + 8 * g + i]) / snorm[so + kInP + 8 * g + i];
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                bx[i] = bf16_rne(v[i]);
                sl[sA0o + (8 * g + i) * kTS + rw] = (uint16_t)bx[i];
            }
        }
        /* forward */
        f32x4 acc[8];
        bf16x8 bh[4], ds1[4], ds2[4], ds3[4];
#pragma unroll
        for (int mt = 0; mt < 8; ++mt)
            acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld8(sw + wo + fL0 + (mt * 64 + ln) * 8), bx,
                                                              f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        tr_fwd_epilogue(acc, sbl, g, sl + sA1o, rw, bh, ds1);
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            uint32_t ko = fL1 + wo;
            asm volatile("" : "+v"(ko));
#pragma unroll
            for (int mt = 0; mt < 8; ++mt)
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld8(sw + ko + ((mt * 4 + kb) * 64 + ln) * 8),
                                                                  bh[kb], acc[mt], 0, 0, 0);
        }
        tr_fwd_epilogue(acc, sbl + kH, g, sl + sA2o, rw, bh, ds2);
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            uint32_t ko = fL2 + wo;
            asm volatile("" : "+v"(ko));
#pragma unroll
            for (int mt = 0; mt < 8; ++mt)
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld8(sw + ko + ((mt * 4 + kb) * 64 + ln) * 8),
                                                                  bh[kb], acc[mt], 0, 0, 0);
        }
        tr_fwd_epilogue(acc, sbl + 2 * kH, g, sl + sA3o, rw, bh, ds3);
        f32x4 o = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
            o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld8(sw + wo + fL3 + (kb * 64 + ln) * 8), bh[kb], o, 0, 0, 0);
        /* W3^T for the first backward product */
        bf16x8 wt[8];
        uint32_t bo = bL3 + wo;
        asm volatile("" : "+v"(bo));
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) wt[mt] = ld8(a.wb + bo + (mt * 64 + ln) * 8);

        /* loss (k_loss's expressions): lane group 0 holds outputs 0..3 of
         * its row, group 1 outputs 4..7 */
        float ov[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) ov[r] = o[r] + sbl[3 * kH + 4 * g + r];
        const float mu[3] = {__shfl(ov[0], s), __shfl(ov[1], s), __shfl(ov[2], s)};
        const float lv[3] = {__shfl(ov[3], s), __shfl(ov[0], 16 + s), __shfl(ov[1], 16 + s)};
        float gm[3] = {0.0f, 0.0f, 0.0f}, gs[3] = {0.0f, 0.0f, 0.0f};
        if (live) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                float sv = lv[k];
                const bool clamped = sv < -12.0f || sv > 12.0f;
                sv = fminf(fmaxf(sv, -12.0f), 12.0f);
                const float r = yv[k] - mu[k];
                const float e = __expf(-sv);
                if (g == 0) lsum += 0.5f * (r * r * e + sv);
                gm[k] = -r * e * inv;
                gs[k] = clamped ? 0.0f : 0.5f * (1.0f - r * r * e) * inv;
            }
        }
        /* this lane's outputs 4g + r: g = 0 -> mean 0..2, log-variance 0;
         * g = 1 -> log-variance 1, 2; the rest padding */
        const float d4[4] = {g == 0 ? gm[0] : (g == 1 ? gs[1] : 0.0f), g == 0 ? gm[1] : (g == 1 ? gs[2] : 0.0f),
                             g == 0 ? gm[2] : 0.0f, g == 0 ? gs[0] : 0.0f};
        bf16x8 bd4;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            bd4[i] = bf16_rne(d4[i]);
            bd4[4 + i] = 0;
            sl[sD4o + (4 * g + i) * kTS + rw] = (uint16_t)bd4[i];
        }
        __syncthreads();  /* B1: inputs^T, a1^T..a3^T and dZ4^T staged */
        /* dA3 = W3^T dZ4, dZ3 = dA3 SiLU'(z3) */
#pragma unroll
        for (int mt = 0; mt < 8; ++mt)
            acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wt[mt], bd4, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        bf16x8 bd[4];
        tr_bwd_epilogue(acc, ds3, bd);
        /* W2^T, fetched once dZ3 is formed (a data dependence orders it) */
        bf16x8 wl[4][8];
        bo = bL2 + wo;
        asm volatile("" : "+v"(bo) : "v"(bd[3]));
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
#pragma unroll
            for (int mt = 0; mt < 8; ++mt) wl[kb][mt] = ld8(a.wb + bo + ((mt * 4 + kb) * 64 + ln) * 8);
        __syncthreads();  /* B2: the gradient waves have read a3^T */
        tr_stage(bd, g, sl + sA3o, rw);
        __syncthreads();  /* B3: dZ3^T staged */
        /* dA2 = W2^T dZ3, dZ2 */
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
#pragma unroll
            for (int mt = 0; mt < 8; ++mt)
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl[kb][mt], bd[kb], acc[mt], 0, 0, 0);
        tr_bwd_epilogue(acc, ds2, bd);
        bo = bL1 + wo;
        asm volatile("" : "+v"(bo) : "v"(bd[3]));
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
#pragma unroll
            for (int mt = 0; mt < 8; ++mt) wl[kb][mt] = ld8(a.wb + bo + ((mt * 4 + kb) * 64 + ln) * 8);
        __syncthreads();  /* B4: the gradient waves have read a2^T */
        tr_stage(bd, g, sl + sA2o, rw);
        __syncthreads();  /* B5: dZ2^T staged */
        /* dA1 = W1^T dZ2, dZ1 */
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
#pragma unroll
            for (int mt = 0; mt < 8; ++mt)
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl[kb][mt], bd[kb], acc[mt], 0, 0, 0);
        tr_bwd_epilogue(acc, ds1, bd);
        __syncthreads();  /* B6: the gradient waves have read a1^T */
        tr_stage(bd, g, sl + sA1o, rw);
        __syncthreads();  /* B7: dZ1^T staged */
        __syncthreads();  /* B8: dW0 has read the staging */
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
    if (lane == 0) sred[w] = lsum;
}

/* The gradient waves' loop (q = 0..3; t = 0..255 their thread index). */
__device__ __forceinline__ void tr_grads(const TrainArgs& a, const uint16_t* st, int q, int lane, int t) {
    const int bn = t >> 1, bhf = t & 1;  /* bias sums: feature and half of the rows */
    f32x4 gw0[4], gw1[16], gw2[16], gw3[2];
#pragma unroll
    for (int j = 0; j < 16; ++j) gw1[j] = gw2[j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < 4; ++j) gw0[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    gw3[0] = gw3[1] = f32x4{0.f, 0.f, 0.f, 0.f};
    float db0 = 0.0f, db1 = 0.0f, db2 = 0.0f, db3 = 0.0f;
    const uint64_t chunks = (a.M + kTrRows - 1) / kTrRows;
    for (uint64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        uint32_t so = 0;
        int ln = lane;
        asm volatile("" : "+v"(so));
        asm volatile("" : "+v"(ln));
        const uint16_t* sl = st + so;
        const int g = ln >> 4, s = ln & 15;
        __syncthreads();  /* B1 */
        /* dW3 (outputs 0..15 x inputs 16(2q + j)..), b3 */
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int ro = 32 * ks + 8 * g;
            const bf16x8 av = ld8(sl + sD4o + s * kTS + ro);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                gw3[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, ld8(sl + sA3o + (16 * (2 * q + j) + s) * kTS + ro),
                                                                 gw3[j], 0, 0, 0);
        }
        if (t < 2 * kOutP) db3 += tr_row_sum(sl + sD4o, bn, bhf);
        __syncthreads();  /* B2 */
        __syncthreads();  /* B3 */
        tr_wgrad(gw2, sl + sA3o, sl + sA2o, q, g, s);
        db2 += tr_row_sum(sl + sA3o, bn, bhf);
        __syncthreads();  /* B4 */
        __syncthreads();  /* B5 */
        tr_wgrad(gw1, sl + sA2o, sl + sA1o, q, g, s);
        db1 += tr_row_sum(sl + sA2o, bn, bhf);
        __syncthreads();  /* B6 */
        __syncthreads();  /* B7 */
        /* dW0 = dZ1^T inputs, b0 */
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int ro = 32 * ks + 8 * g;
#pragma unroll
            for (int jn = 0; jn < 2; ++jn) {
                const bf16x8 av = ld8(sl + sA1o + (32 * q + 16 * jn + s) * kTS + ro);
#pragma unroll
                for (int tk = 0; tk < 2; ++tk)
                    gw0[2 * jn + tk] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        av, ld8(sl + sA0o + (16 * tk + s) * kTS + ro), gw0[2 * jn + tk], 0, 0, 0);
            }
        }
        db0 += tr_row_sum(sl + sA1o, bn, bhf);
        __syncthreads();  /* B8 */
    }
    /* this workgroup's partial gradient, parameter order */
    const int g = lane >> 4, s = lane & 15;
    float* P = a.part + (size_t)blockIdx.x * kParams;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = 16 * (2 * q + (j >> 1)) + 4 * g + r, k = 16 * (j & 1) + s;
            if (k < kIn) P[oW0 + n * kIn + k] = gw0[j][r];
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int n = 16 * (2 * q + (j >> 3)) + 4 * g + r, k = 16 * (j & 7) + s;
            P[oW1 + n * kH + k] = gw1[j][r];
            P[oW2 + n * kH + k] = gw2[j][r];
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int o = 4 * g + r, k = 16 * (2 * q + j) + s;
            if (o < kOut) P[oW3 + o * kH + k] = gw3[j][r];
        }
    }
    db0 += __shfl_xor(db0, 1);
    db1 += __shfl_xor(db1, 1);
    db2 += __shfl_xor(db2, 1);
    db3 += __shfl_xor(db3, 1);
    if (bhf == 0) {
        P[oB0 + bn] = db0;
        P[oB1 + bn] = db1;
        P[oB2 + bn] = db2;
        if (bn < kOut) P[oB3 + bn] = db3;
    }
}

__global__ __launch_bounds__(kTrThreads, 2) void k_train_fused(TrainArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint16_t* st = reinterpret_cast<uint16_t*>(smem);
    uint16_t* sw = reinterpret_cast<uint16_t*>(smem + kTrLdsS);
    float* sb = reinterpret_cast<float*>(smem + kTrLdsS + (size_t)fEnd * 2);
    float* sred = sb + fBias;
    float* snorm = sred + 4;  /* feature means (32), then standard deviations (32) */
    const int tid = threadIdx.x;
    {
        const uint4* src = reinterpret_cast<const uint4*>(a.wf);
        uint4* dst = reinterpret_cast<uint4*>(sw);
        for (int e = tid; e < fEnd / 8; e += kTrThreads) dst[e] = src[e];
        if (tid < 2 * kInP) {
            const int c = tid & (kInP - 1);
            snorm[tid] = tid < kInP ? (c < kIn ? a.mu[c] : 0.0f) : (c < kIn ? a.sd[c] : 1.0f);
        }
        for (int e = tid; e < fBias; e += kTrThreads) {
            float v = 0.0f;
            if (e < kH) v = a.p[oB0 + e];
            else if (e < 2 * kH) v = a.p[oB1 + e - kH];
            else if (e < 3 * kH) v = a.p[oB2 + e - 2 * kH];
            else if (e - 3 * kH < kOut) v = a.p[oB3 + e - 3 * kH];
            sb[e] = v;
        }
    }
    __syncthreads();
    const int lane = tid & 63;
    const int wv = (int)__builtin_amdgcn_readfirstlane((uint32_t)(tid >> 6));  /* uniform: a scalar branch */
    if (wv < 4) tr_rows(a, st, sw, sb, sred, snorm, wv, lane);
    else tr_grads(a, st, wv - 4, lane, tid - 256);
    __syncthreads();  /* the row waves' loss sums */
    if (tid == 0) a.lpart[blockIdx.x] = ((sred[0] + sred[1]) + sred[2]) + sred[3];
}

/* The gradient: the nb workgroups' partials summed in a fixed order (64
 * parameters per workgroup, each wave a quarter of the partials) and the
 * mean loss; with lr > 0 the Adam step on the FP32 master copy, and the two
 * BF16 packs updated in place. */
__global__ __launch_bounds__(256) void k_train_finish(const float* __restrict__ part, int nb,
                                                      const float* __restrict__ lpart, float inv_m,
                                                      float* __restrict__ g, float* __restrict__ loss_out,
                                                      float* __restrict__ p, float* __restrict__ m,
                                                      float* __restrict__ v, uint16_t* __restrict__ wf,
                                                      uint16_t* __restrict__ wb, float lr, float c1, float c2) {
    __shared__ float sp[4][64];
    const int t = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int e = blockIdx.x * 64 + t;
    float acc = 0.0f;
    if (e < kParams) {
#pragma unroll 8
        for (int b = q; b < nb; b += 4) acc += part[(size_t)b * kParams + e];
    }
    sp[q][t] = acc;
    __syncthreads();
    if (q == 0 && e < kParams) {
        const float gg = ((sp[0][t] + sp[1][t]) + sp[2][t]) + sp[3][t];
        g[e] = gg;
        if (lr > 0.0f) {
            const float b1 = 0.9f, b2 = 0.999f;
            const float mm = b1 * m[e] + (1.0f - b1) * gg;
            const float vv = b2 * v[e] + (1.0f - b2) * gg * gg;
            m[e] = mm;
            v[e] = vv;
            const float pn = p[e] - lr * (mm / c1) / (sqrtf(vv / c2) + 1e-8f);
            p[e] = pn;
            int pf, pb;
            pack_pos(e, pf, pb);
            if (pf >= 0) wf[pf] = f2bf(pn);
            if (pb >= 0) wb[pb] = f2bf(pn);
        }
    }
    if (blockIdx.x == 0 && q == 1 && loss_out) {
        float l = 0.0f;
        for (int b = t; b < nb; b += 64) l += lpart[b];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
        if (t == 0) *loss_out = l * inv_m;
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
    uint16_t* Wb = nullptr;  /* lane-ready transposed weights of the fused training pass */
    int n_cu = 0;
    /* workspace */
    uint64_t cap = 0;
    uint16_t *X0 = nullptr, *Z1 = nullptr, *A1 = nullptr, *Z2 = nullptr, *A2 = nullptr, *Z3 = nullptr, *A3 = nullptr;
    uint16_t *dZ4 = nullptr, *dZ = nullptr;
    float *O = nullptr, *dA = nullptr, *P = nullptr, *Pb = nullptr, *lpart = nullptr;
    /* fused training: per-workgroup partial gradients and loss sums, loss per step */
    float *tpart = nullptr, *tloss = nullptr, *lossd = nullptr;
    uint32_t tcap = 0, lcap = 0;
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

/* the BF16 copies of the per-layer products (the unfused path): the fused
 * training step does not keep them current, so the unfused entry points
 * pack them from the FP32 parameters on every call */
int pack_unfused(spmd_mlp* h) {
    const float* p = h->p;
    auto g = [](int n) { return dim3((n + 255) / 256); };
    hipLaunchKernelGGL(k_pack, g(kH * kInP), dim3(256), 0, h->st, p + oW0, kH, kIn, kH, kInP, h->W0,
                       (uint16_t*)nullptr, 0);
    hipLaunchKernelGGL(k_pack, g(kH * kH), dim3(256), 0, h->st, p + oW1, kH, kH, kH, kH, h->W1, h->W1t, kH);
    hipLaunchKernelGGL(k_pack, g(kH * kH), dim3(256), 0, h->st, p + oW2, kH, kH, kH, kH, h->W2, h->W2t, kH);
    hipLaunchKernelGGL(k_pack, g(kH * kOutK), dim3(256), 0, h->st, p + oW3, kOut, kH, kOutP, kH, h->W3, h->W3t,
                       kOutK);
    MCHK(h, hipGetLastError());
    return SPMD_OK;
}

/* every BF16 copy from the FP32 parameters (after they were set) */
int pack_all(spmd_mlp* h) {
    auto g = [](int n) { return dim3((n + 255) / 256); };
    if (int rc = pack_unfused(h)) return rc;
    hipLaunchKernelGGL(k_pack_fused, g(fEnd), dim3(256), 0, h->st, h->p, h->Wf);
    hipLaunchKernelGGL(k_pack_bwd, g(bEnd), dim3(256), 0, h->st, h->p, h->Wb);
    MCHK(h, hipGetLastError());
    return SPMD_OK;
}

int ensure_loss(spmd_mlp* h, uint32_t n) {
    if (n <= h->lcap) return SPMD_OK;
    dfr(h->lossd);
    MCHK(h, dal(&h->lossd, n));
    h->lcap = n;
    return SPMD_OK;
}

/* One fused training pass over M rows (x, y; with n_src > 0 the rows are
 * drawn from n_src source rows by Philox(seed, step)): the gradient into
 * h->g, the mean loss into loss_dev (device, optional), and with lr > 0
 * the Adam step with bias corrections c1, c2.  Two launches on h->st. */
int train_fused(spmd_mlp* h, const float* x, const float* y, uint64_t n_src, uint64_t M, uint64_t seed,
                uint32_t step, float lr, float c1, float c2, float* loss_dev) {
    static bool attr = false;
    if (!attr) {
        MCHK(h, hipFuncSetAttribute((const void*)k_train_fused, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)kTrLds));
        attr = true;
    }
    const uint64_t chunks = (M + kTrRows - 1) / kTrRows;
    const uint64_t cap = (uint64_t)(h->n_cu > 0 ? h->n_cu : 256);
    const unsigned nb = (unsigned)(chunks < cap ? chunks : cap);
    if (nb > h->tcap) {
        MCHK(h, hipStreamSynchronize(h->st));
        dfr(h->tpart);
        dfr(h->tloss);
        MCHK(h, dal(&h->tpart, (size_t)nb * kParams));
        MCHK(h, dal(&h->tloss, nb));
        h->tcap = nb;
    }
    TrainArgs a{x, y, n_src, M, seed, step, h->mu, h->sd, h->sdy, h->Wf, h->Wb, h->p, h->tpart, h->tloss};
    hipLaunchKernelGGL(k_train_fused, dim3(nb), dim3(kTrThreads), kTrLds, h->st, a);
    hipLaunchKernelGGL(k_train_finish, dim3((kParams + 63) / 64), dim3(256), 0, h->st, h->tpart, (int)nb, h->tloss,
                       1.0f / (float)M, h->g, loss_dev, h->p, h->m, h->v, h->Wf, h->Wb, lr, c1, c2);
    MCHK(h, hipGetLastError());
    return SPMD_OK;
}

/* the fused inference pass of M rows on stream st: x (M x 24) -> out (M x 6) */
constexpr int kFwdRows = 1;  /* 16-row tiles per wave and pass (128 VGPRs: four waves per SIMD) */
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
        bad(dal(&h->W2t, kH * kH)) || bad(dal(&h->W3t, kH * kOutK)) || bad(dal(&h->Wf, fEnd)) || bad(dal(&h->Wb, bEnd))) {
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
    dfr(h->W0); dfr(h->W1); dfr(h->W2); dfr(h->W3); dfr(h->W1t); dfr(h->W2t); dfr(h->W3t); dfr(h->Wf); dfr(h->Wb);
    dfr(h->X0); dfr(h->Z1); dfr(h->A1); dfr(h->Z2); dfr(h->A2); dfr(h->Z3); dfr(h->A3);
    dfr(h->dZ4); dfr(h->dZ); dfr(h->O); dfr(h->dA); dfr(h->P); dfr(h->Pb); dfr(h->lpart);
    dfr(h->tpart); dfr(h->tloss); dfr(h->lossd);
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

/* the per-layer path's forward (four row-tiled products through the
 * workspace, the activations kept for a backward pass) */
int spmd_mlp_forward_unfused(spmd_mlp* h, const float* x_dev, float* out_dev, uint64_t M) {
    if (!h || !x_dev || !out_dev) return SPMD_ERR_ARG;
    if (M == 0) return SPMD_OK;
    uint64_t Mp;
    if (int rc = pack_unfused(h)) return rc;
    if (int rc = to_rows(h, x_dev, M, &Mp)) return rc;
    if (int rc = forward_rows(h, Mp)) return rc;
    MCHK(h, hipMemcpy2DAsync(out_dev, kOut * 4, h->O, kOutP * 4, kOut * 4, M, hipMemcpyDeviceToDevice, h->st));
    MCHK(h, hipStreamSynchronize(h->st));
    return SPMD_OK;
}

/* the fused training pass without the Adam step */
int spmd_mlp_gradient(spmd_mlp* h, const float* x_dev, const float* y_dev, uint64_t M, float* grad_host,
                      double* loss) {
    if (!h || !x_dev || !y_dev || M == 0) return SPMD_ERR_ARG;
    if (int rc = ensure_loss(h, 1)) return rc;
    if (int rc = train_fused(h, x_dev, y_dev, 0, M, 0, 0, 0.0f, 1.0f, 1.0f, h->lossd)) return rc;
    if (grad_host) MCHK(h, hipMemcpyAsync(grad_host, h->g, kParams * 4ull, hipMemcpyDeviceToHost, h->st));
    float l = 0.0f;
    MCHK(h, hipMemcpyAsync(&l, h->lossd, 4, hipMemcpyDeviceToHost, h->st));
    MCHK(h, hipStreamSynchronize(h->st));
    if (loss) *loss = l;
    return SPMD_OK;
}

/* the same gradient through the per-layer products (loss, data and weight
 * products, SiLU' and the partial sums as separate kernels) */
int spmd_mlp_gradient_unfused(spmd_mlp* h, const float* x_dev, const float* y_dev, uint64_t M, float* grad_host,
                              double* loss) {
    if (!h || !x_dev || !y_dev || M == 0) return SPMD_ERR_ARG;
    uint64_t Mp;
    if (int rc = pack_unfused(h)) return rc;
    if (int rc = to_rows(h, x_dev, M, &Mp)) return rc;
    if (int rc = forward_rows(h, Mp)) return rc;
    if (int rc = backward_rows(h, y_dev, M, Mp, loss)) return rc;
    if (grad_host) MCHK(h, hipMemcpyAsync(grad_host, h->g, kParams * 4ull, hipMemcpyDeviceToHost, h->st));
    MCHK(h, hipStreamSynchronize(h->st));
    return SPMD_OK;
}

int spmd_mlp_train_step(spmd_mlp* h, const float* x_dev, const float* y_dev, uint64_t M, double lr, double* loss) {
    if (!h || !x_dev || !y_dev || M == 0 || !(lr > 0)) return SPMD_ERR_ARG;
    if (int rc = ensure_loss(h, 1)) return rc;
    h->step += 1;
    const float c1 = 1.0f - powf(0.9f, (float)h->step), c2 = 1.0f - powf(0.999f, (float)h->step);
    if (int rc = train_fused(h, x_dev, y_dev, 0, M, 0, 0, (float)lr, c1, c2, loss ? h->lossd : nullptr)) return rc;
    float l = 0.0f;
    if (loss) MCHK(h, hipMemcpyAsync(&l, h->lossd, 4, hipMemcpyDeviceToHost, h->st));
    MCHK(h, hipStreamSynchronize(h->st));
    if (loss) *loss = l;
    return SPMD_OK;
}

/* Each step: one fused pass over `batch` rows drawn on the device (Philox
 * counter (row, step, 7, 0), key seed) and the Adam step; the losses stay
 * on the device until the end. */
int spmd_mlp_fit(spmd_mlp* h, const float* x_dev, const float* y_dev, uint64_t N, uint32_t batch, uint32_t steps,
                 double lr, uint64_t seed, double* loss_hist) {
    if (!h || !x_dev || !y_dev || N == 0 || batch == 0 || !(lr > 0)) return SPMD_ERR_ARG;
    if (int rc = ensure_loss(h, loss_hist && steps ? steps : 1)) return rc;
    for (uint32_t it = 0; it < steps; ++it) {
        const uint32_t draw = (uint32_t)h->step;
        h->step += 1;
        const float c1 = 1.0f - powf(0.9f, (float)h->step), c2 = 1.0f - powf(0.999f, (float)h->step);
        if (int rc = train_fused(h, x_dev, y_dev, N, batch, seed, draw, (float)lr, c1, c2,
                                 loss_hist ? h->lossd + it : nullptr))
            return rc;
    }
    if (loss_hist && steps) {
        std::vector<float> l(steps);
        MCHK(h, hipMemcpyAsync(l.data(), h->lossd, steps * 4ull, hipMemcpyDeviceToHost, h->st));
        MCHK(h, hipStreamSynchronize(h->st));
        for (uint32_t it = 0; it < steps; ++it) loss_hist[it] = l[it];
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
