// Collision surrogate: a 64 -> 128 -> 128 -> 128 -> 1 ReLU MLP with BF16
// operands and FP32 accumulation on gfx950 matrix cores, trained on labels
// from the exact checker.  It has no counterpart in the reference; it only
// orders the waypoint checks of link() so that likely collisions are tested
// first — every verdict still comes from the exact checker.
//
// Matrix products use v_mfma_f32_16x16x32_bf16, one wave per 16x16 output
// tile.  Operand layout (checked on the device): lane l supplies row l%16 of
// A and column l%16 of B, k = 8*(l/16) .. +7; it receives column l%16 and
// rows 4*(l/16) .. +3 of D.  Every product is written C = A * Bt^T with A and
// Bt row-major, so both operands are contiguous 16-byte loads.  Epilogues:
//   FWD  : + bias, ReLU, store H (BF16) and H^T (BF16, for the weight grads)
//   DX   : zero where the layer input was 0 (ReLU'), store dH and dH^T (BF16)
//   F32  : store FP32 (weight gradients)
// Rounding points are those of the CPU oracle (oracle/orc_mlp.cpp).
//
// Parameter layout (FP32 master copy, same as the oracle):
//   W1[64][128] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] w4[128] b4
#include <algorithm>
#include <cmath>
#include <vector>

#include "bisect_dev.hpp"

namespace mptg {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int IN = 64, HID = 128;
constexpr int64_t OFF_W1 = 0, OFF_B1 = OFF_W1 + IN * HID, OFF_W2 = OFF_B1 + HID, OFF_B2 = OFF_W2 + HID * HID,
                  OFF_W3 = OFF_B2 + HID, OFF_B3 = OFF_W3 + HID * HID, OFF_W4 = OFF_B3 + HID,
                  OFF_B4 = OFF_W4 + HID, NPARAM = OFF_B4 + 1;

enum { EPI_FWD = 0, EPI_DX = 1, EPI_F32 = 2 };

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}

struct Epi {
    const float* bias;     // FWD
    const uint16_t* mask;  // DX: layer input, [M][N]
    uint16_t* out;         // FWD/DX: [M][N] BF16
    uint16_t* out_t;       // FWD/DX: [N][M] BF16 (may be null)
    float* out_f;          // F32: [M][N]
};

// C[M][N] = A[M][K] * Bt[N][K]^T; M, N multiples of 16, K multiple of 32.
template <int MODE>
__global__ __launch_bounds__(256) void gemm_nt(int M, int N, int K, const uint16_t* __restrict__ A,
                                               const uint16_t* __restrict__ Bt, Epi ep) {
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int tiles_n = N / 16;
    const int tm = tile / tiles_n, tn = tile % tiles_n;
    if (tm * 16 >= M) return;
    const int r = lane & 15, kq = lane >> 4;
    const uint16_t* pa = A + (int64_t)(tm * 16 + r) * K + 8 * kq;
    const uint16_t* pb = Bt + (int64_t)(tn * 16 + r) * K + 8 * kq;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(load8(pa + k0), load8(pb + k0), acc, 0, 0, 0);
    const int n = tn * 16 + r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = tm * 16 + 4 * kq + i;
        float v = acc[i];
        if constexpr (MODE == EPI_F32) {
            ep.out_f[(int64_t)m * N + n] = v;
        } else {
            if constexpr (MODE == EPI_FWD) {
                v = v + ep.bias[n];
                v = v > 0.f ? v : 0.f;
            } else {
                if (ep.mask[(int64_t)m * N + n] == 0) v = 0.f;  // ReLU output is 0 or positive
            }
            uint16_t h = float_to_bf16(v);
            ep.out[(int64_t)m * N + n] = h;
            if (ep.out_t) ep.out_t[(int64_t)n * M + m] = h;
        }
    }
}

// Output layer and loss: z = H3 . w4 + b4; dz = (sigmoid(z) - y) / N;
// dH3 = bf16(dz * w4 where H3 > 0).  A block takes 64 rows, each wave 16 of
// them in turn: lanes hold 2 of the 128 hidden units (coalesced row loads), a
// wave reduction gives the logit, and the gradient row goes to dH3 directly
// and to dH3^T through an LDS tile so both stores are coalesced.
constexpr int kHeadRows = 64;

__global__ __launch_bounds__(256) void head_kernel(int M, const uint16_t* __restrict__ H3,
                                                   const uint16_t* __restrict__ w4, const float* __restrict__ b4,
                                                   const float* __restrict__ y, float inv_n, float* __restrict__ z,
                                                   float* __restrict__ dz, uint16_t* __restrict__ dH3,
                                                   uint16_t* __restrict__ dH3t, float* __restrict__ loss) {
    __shared__ uint16_t tile[HID][kHeadRows + 2];
    __shared__ float wl[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int r0 = blockIdx.x * kHeadRows;
    const float wa = bf16_to_float(w4[2 * lane]), wb = bf16_to_float(w4[2 * lane + 1]);
    float lsum = 0.f;
    // all 16 row loads of this wave in flight before any is used
    constexpr int kRowsPerWave = kHeadRows / 4;
    uint32_t pairs[kRowsPerWave];
    float ys[kRowsPerWave];
#pragma unroll
    for (int j = 0; j < kRowsPerWave; ++j) {
        const int r = r0 + wid + 4 * j;
        pairs[j] = r < M ? reinterpret_cast<const uint32_t*>(H3 + (int64_t)r * HID)[lane] : 0u;
        ys[j] = (y && r < M) ? y[r] : 0.f;
    }
#pragma unroll
    for (int j = 0; j < kRowsPerWave; ++j) {
        const int i = wid + 4 * j;
        const int r = r0 + i;
        uint16_t g0 = 0, g1 = 0;
        if (r < M) {  // wave-uniform
            const uint16_t h0 = (uint16_t)(pairs[j] & 0xffffu), h1 = (uint16_t)(pairs[j] >> 16);
            float acc = bf16_to_float(h0) * wa + bf16_to_float(h1) * wb;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
            const float zz = acc + b4[0];
            if (z && lane == 0) z[r] = zz;
            if (y) {
                const float yy = ys[j];
                const float d = (1.f / (1.f + expf(-zz)) - yy) * inv_n;
                if (lane == 0) {
                    lsum += fmaxf(zz, 0.f) - zz * yy + log1pf(expf(-fabsf(zz)));
                    dz[r] = d;
                }
                g0 = float_to_bf16(h0 != 0 ? d * wa : 0.f);
                g1 = float_to_bf16(h1 != 0 ? d * wb : 0.f);
                reinterpret_cast<uint32_t*>(dH3 + (int64_t)r * HID)[lane] = (uint32_t)g0 | ((uint32_t)g1 << 16);
            }
        }
        tile[2 * lane][i] = g0;
        tile[2 * lane + 1][i] = g1;
    }
    if (!y) return;  // block-uniform
    if (lane == 0) wl[wid] = lsum;
    __syncthreads();
    // dH3^T rows k, columns r0 .. r0 + 63: 64 consecutive rows per k
    for (int idx = threadIdx.x; idx < HID * kHeadRows; idx += 256) {
        const int k = idx / kHeadRows, i = idx % kHeadRows;
        if (r0 + i < M) dH3t[(int64_t)k * M + r0 + i] = tile[k][i];
    }
    if (threadIdx.x == 0) atomicAdd(loss, (wl[0] + wl[1] + wl[2] + wl[3]) * inv_n);
}

// out[row] = sum_j X[row][j] * (w ? w[j] : 1), X BF16 [rows][len]: bias
// gradients from dH^T rows, and dw4 from H3^T rows weighted by dz.
__global__ __launch_bounds__(256) void row_sums(const uint16_t* __restrict__ X, int len, const float* __restrict__ w,
                                                float* __restrict__ out) {
    const uint16_t* x = X + (int64_t)blockIdx.x * len;
    float acc = 0.f;
    for (int j = threadIdx.x; j < len; j += blockDim.x) acc += bf16_to_float(x[j]) * (w ? w[j] : 1.f);
    __shared__ float red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// Transpose of a BF16 matrix [rows][cols] -> [cols][rows].
__global__ void transpose_bf16(const uint16_t* __restrict__ X, int rows, int cols, uint16_t* __restrict__ Xt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rows * cols) return;
    int r = (int)(i / cols), c = (int)(i % cols);
    Xt[(int64_t)c * rows + r] = X[i];
}

// BF16 operand copies of the FP32 master weights: W[in][out] and W^T[out][in].
__global__ void cast_weights(const float* __restrict__ p, uint16_t* __restrict__ wb, uint16_t* __restrict__ wtb) {
    // wb/wtb hold W1, W2, W3, w4 back to back (W4 has no transpose copy)
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n1 = IN * HID, n2 = HID * HID;
    if (i < n1) {
        int a = (int)(i / HID), o = (int)(i % HID);
        uint16_t h = float_to_bf16(p[OFF_W1 + i]);
        wb[i] = h;
        wtb[(int64_t)o * IN + a] = h;
    } else if (i < n1 + 2 * n2) {
        int64_t j = i - n1;
        int layer = (int)(j / n2);
        int64_t k = j % n2;
        int a = (int)(k / HID), o = (int)(k % HID);
        uint16_t h = float_to_bf16(p[(layer == 0 ? OFF_W2 : OFF_W3) + k]);
        wb[n1 + layer * n2 + k] = h;
        wtb[n1 + layer * n2 + (int64_t)o * HID + a] = h;
    } else if (i < n1 + 2 * n2 + HID) {
        int64_t k = i - n1 - 2 * n2;
        wb[n1 + 2 * n2 + k] = float_to_bf16(p[OFF_W4 + k]);
    }
}

__global__ void adam_kernel(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, int64_t n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void init_params(float* __restrict__ p, uint64_t seed) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    float fan_in = 0.f;
    if (i < OFF_B1) fan_in = IN;
    else if ((i >= OFF_W2 && i < OFF_B2) || (i >= OFF_W3 && i < OFF_B3) || (i >= OFF_W4 && i < OFF_B4)) fan_in = HID;
    if (fan_in == 0.f) {
        p[i] = 0.f;  // biases
        return;
    }
    double u = uniform_at(seed, (uint64_t)i, 0x4d4c5030u, 0);  // He-uniform
    float lim = sqrtf(6.f / fan_in);
    p[i] = (float)(u * 2.0 - 1.0) * lim;
}

// Surrogate features of an arm state (FP32): joint positions p_1..p_20
// scaled by 1/total length (40 values), then cos/sin of the first 12
// cumulative angles (24 values).  Hardware sine/cosine: the features only
// steer the check order, never a verdict.
template <int D>
__device__ __forceinline__ void arm_features(const double* q, const double* lengths, float inv_total, uint16_t* out) {
    float x = 0.f, y = 0.f, ang = 0.f;
    int k = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        ang += (float)q[i];
        float s, c;
        __sincosf(ang, &s, &c);
        x += (float)lengths[i] * c;
        y += (float)lengths[i] * s;
        if (2 * i + 1 < 40) {
            out[2 * i] = float_to_bf16(x * inv_total);
            out[2 * i + 1] = float_to_bf16(y * inv_total);
        }
        if (i < 12) {
            out[40 + 2 * i] = float_to_bf16(c);
            out[40 + 2 * i + 1] = float_to_bf16(s);
        }
        ++k;
    }
    for (int j = 2 * D; j < 40; ++j) out[j] = 0;
    for (int j = 40 + 2 * (D < 12 ? D : 12); j < 64; ++j) out[j] = 0;
}

}  // namespace
}  // namespace mptg

namespace mptg {
namespace {

__global__ void sum_f32(const float* __restrict__ x, int n, float* __restrict__ out) {
    float acc = 0.f;
    for (int j = threadIdx.x; j < n; j += blockDim.x) acc += x[j];
    __shared__ float red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

template <int D>
__global__ void features_kernel(const double* __restrict__ q, int64_t n, const double* __restrict__ lengths,
                                float inv_total, uint16_t* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s[D];
#pragma unroll
    for (int j = 0; j < D; ++j) s[j] = q[i * D + j];
    uint16_t f[IN];
    arm_features<D>(s, lengths, inv_total, f);
#pragma unroll
    for (int j = 0; j < IN; ++j) out[i * IN + j] = f[j];
}

}  // namespace

// ---- fused inference over the waypoints of arm edges ----------------------
// One wave (= one block) scores 64 waypoints of one edge.  The features come
// from an FP32 interpolation of the waypoint (its dyadic position along the
// edge) rather than the reference's FP64 midpoint recursion: scores only
// order the exact checks, so a 1e-7 relative difference in a feature is
// immaterial.  Each hidden layer is computed transposed, H^T = W . X^T, with
// the weights as the MFMA A operand (16 output units x 32 inputs) and the
// activations as the B operand (32 inputs x 16 waypoints): a lane's output
// fragment is then 4 consecutive units of one waypoint, stored as one 8-byte
// LDS write after a hardware BF16 conversion, and read back by the next
// layer as 16-byte B fragments.  The output unit's dot product is folded
// into the last layer's epilogue.  Weights come from the BF16 W^T copies
// (L2/L1 resident, 80 KiB).
constexpr int kLdsRow = HID + 8;  // padded row (BF16 elements)

__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b) {
    const __bf16 x = (__bf16)a, y = (__bf16)b;
    return (uint32_t)__builtin_bit_cast(uint16_t, x) | ((uint32_t)__builtin_bit_cast(uint16_t, y) << 16);
}

// Orders a wave's LDS accesses: a block barrier in one-wave blocks, a wave
// barrier with a workgroup fence where waves of a block run apart.
template <bool WAVE>
__device__ __forceinline__ void buf_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    } else {
        __syncthreads();
    }
}

// buf[16 NST waypoints][K] (BF16, row stride kLdsRow) -> buf[16 NST][128] =
// relu(X W^T + b); W^T rows WS elements apart (K in global memory, padded in
// LDS).  A row's result does not depend on NST or on the other rows.
template <int K, int WS = K, bool WAVE = false, int NST = 4>
__device__ __forceinline__ void wave_layer_t(uint16_t* __restrict__ buf, const uint16_t* __restrict__ wt,
                                             const float* __restrict__ bias, int lane) {
    constexpr int KS = K / 32;
    const int r = lane & 15, kq = lane >> 4;
    bf16x8 xf[NST][KS];
#pragma unroll
    for (int st = 0; st < NST; ++st)
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) xf[st][kk] = load8(buf + (st * 16 + r) * kLdsRow + kk * 32 + 8 * kq);
    buf_sync<WAVE>();  // every lane holds its fragments before the buffer is overwritten
#pragma unroll 2
    for (int ot = 0; ot < HID / 16; ++ot) {
        bf16x8 wf[KS];
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) wf[kk] = load8(wt + (ot * 16 + r) * WS + kk * 32 + 8 * kq);
        const float4 bb = *reinterpret_cast<const float4*>(bias + ot * 16 + 4 * kq);
#pragma unroll
        for (int st = 0; st < NST; ++st) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < KS; ++kk) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[kk], xf[st][kk], acc, 0, 0, 0);
            const float v0 = fmaxf(acc[0] + bb.x, 0.f), v1 = fmaxf(acc[1] + bb.y, 0.f);
            const float v2 = fmaxf(acc[2] + bb.z, 0.f), v3 = fmaxf(acc[3] + bb.w, 0.f);
            uint2 pk;
            pk.x = pack_bf16x2(v0, v1);
            pk.y = pack_bf16x2(v2, v3);
            *reinterpret_cast<uint2*>(buf + (st * 16 + r) * kLdsRow + ot * 16 + 4 * kq) = pk;
        }
    }
    buf_sync<WAVE>();
}

// Last hidden layer and the output unit: logit of waypoint st * 16 + (lane & 15)
// in out[st] (all lanes), H3 rounded to BF16 as the training forward does.
template <int WS = HID, int NST = 4>
__device__ __forceinline__ void wave_layer_logit(const uint16_t* __restrict__ buf, const uint16_t* __restrict__ wt,
                                                 const float* __restrict__ bias, const uint16_t* __restrict__ w4,
                                                 int lane, float out[NST]) {
    constexpr int KS = HID / 32;
    const int r = lane & 15, kq = lane >> 4;
    bf16x8 xf[NST][KS];
#pragma unroll
    for (int st = 0; st < NST; ++st)
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) xf[st][kk] = load8(buf + (st * 16 + r) * kLdsRow + kk * 32 + 8 * kq);
    float part[NST];
#pragma unroll
    for (int st = 0; st < NST; ++st) part[st] = 0.f;
#pragma unroll 2
    for (int ot = 0; ot < HID / 16; ++ot) {
        bf16x8 wf[KS];
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) wf[kk] = load8(wt + (ot * 16 + r) * WS + kk * 32 + 8 * kq);
        const float4 bb = *reinterpret_cast<const float4*>(bias + ot * 16 + 4 * kq);
        const uint2 wq = *reinterpret_cast<const uint2*>(w4 + ot * 16 + 4 * kq);
        const float w0 = bf16_to_float((uint16_t)(wq.x & 0xffffu)), w1 = bf16_to_float((uint16_t)(wq.x >> 16));
        const float w2 = bf16_to_float((uint16_t)(wq.y & 0xffffu)), w3 = bf16_to_float((uint16_t)(wq.y >> 16));
#pragma unroll
        for (int st = 0; st < NST; ++st) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < KS; ++kk) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[kk], xf[st][kk], acc, 0, 0, 0);
            const float h0 = (float)(__bf16)fmaxf(acc[0] + bb.x, 0.f), h1 = (float)(__bf16)fmaxf(acc[1] + bb.y, 0.f);
            const float h2 = (float)(__bf16)fmaxf(acc[2] + bb.z, 0.f), h3 = (float)(__bf16)fmaxf(acc[3] + bb.w, 0.f);
            part[st] += ((h0 * w0 + h1 * w1) + h2 * w2) + h3 * w3;
        }
    }
#pragma unroll
    for (int st = 0; st < NST; ++st) {
        float v = part[st];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        out[st] = v;
    }
}

// FP32 features of a waypoint given as FP32 coordinates (see arm_features).
template <int D>
__device__ __forceinline__ void arm_features_f(const float* q, const double* lengths, float inv_total,
                                               uint16_t* out) {
    float x = 0.f, y = 0.f, ang = 0.f;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        ang += q[i];
        float s, c;
        __sincosf(ang, &s, &c);
        x += (float)lengths[i] * c;
        y += (float)lengths[i] * s;
        if (2 * i + 1 < 40) {
            out[2 * i] = __builtin_bit_cast(uint16_t, (__bf16)(x * inv_total));
            out[2 * i + 1] = __builtin_bit_cast(uint16_t, (__bf16)(y * inv_total));
        }
        if (i < 12) {
            out[40 + 2 * i] = __builtin_bit_cast(uint16_t, (__bf16)c);
            out[40 + 2 * i + 1] = __builtin_bit_cast(uint16_t, (__bf16)s);
        }
    }
#pragma unroll
    for (int j = 2 * D; j < 40; ++j) out[j] = 0;
#pragma unroll
    for (int j = 40 + 2 * (D < 12 ? D : 12); j < 64; ++j) out[j] = 0;
}

// FP32 features of the waypoint at (path, depth) of link(a, b): the dyadic
// position t of the node's midpoint along the edge, the FP32 interpolation
// a + t (b - a), then arm_features_f.  Every scoring kernel computes a
// waypoint's features here, so a waypoint gets the same score in each.
template <int D>
__device__ __forceinline__ void node_features(const double* a, const double* b, uint32_t path, int depth,
                                              const double* lengths, float inv_total, uint16_t* f) {
    float t = 0.f, step = 0.5f;
#pragma unroll 1
    for (int k = 0; k < depth; ++k, step *= 0.5f)
        if ((path >> k) & 1u) t += step;
    t += step;
    float q[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        const float ai = (float)a[i], bi = (float)b[i];
        q[i] = ai + t * (bi - ai);
    }
    arm_features_f<D>(q, lengths, inv_total, f);
}

// Lane feature row -> the wave's LDS activation buffer (row `lane`).
__device__ __forceinline__ void store_row(uint16_t* buf, int lane, const uint16_t (&f)[IN]) {
#pragma unroll
    for (int j = 0; j < IN; j += 8) {
        uint4 v;
        v.x = (uint32_t)f[j] | ((uint32_t)f[j + 1] << 16);
        v.y = (uint32_t)f[j + 2] | ((uint32_t)f[j + 3] << 16);
        v.z = (uint32_t)f[j + 4] | ((uint32_t)f[j + 5] << 16);
        v.w = (uint32_t)f[j + 6] | ((uint32_t)f[j + 7] << 16);
        *reinterpret_cast<uint4*>(buf + lane * kLdsRow + j) = v;
    }
}

// Scores of every waypoint of edges e < n with mask[e] set (or all): one
// wave per tile of 64 reference ordinals from c0, written to
// scores[e * kScoreStride + ordinal].
template <int D>
__global__ __launch_bounds__(64) void score_waypoints(const double* __restrict__ from, const double* __restrict__ to,
                                                      int64_t n, const uint8_t* __restrict__ mask, double tol,
                                                      const double* __restrict__ lengths, float inv_total,
                                                      const uint16_t* __restrict__ wt, const uint16_t* __restrict__ wb,
                                                      const float* __restrict__ p, float* __restrict__ scores,
                                                      int tiles) {
    __shared__ __attribute__((aligned(16))) uint16_t buf[64 * kLdsRow];
    const int lane = threadIdx.x;
    const int64_t e = blockIdx.x / tiles;
    const uint32_t c0 = (uint32_t)(blockIdx.x % tiles) * 64;
    if (e >= n || (mask && !mask[e])) return;
    const double* a = from + e * D;
    const double* b = to + e * D;
    int h = 0;
    bool regular = false;
    if (!bisect_shape<D>(a, b, tol, &h, &regular)) return;
    const uint32_t W = (1u << h) - 1u;
    if (c0 >= W || W > (uint32_t)kScoreStride) return;
    const uint32_t o = c0 + (uint32_t)lane;
    uint16_t f[IN];
    if (o < W) {
        uint32_t path;
        int depth;
        preorder_path(h, o, &path, &depth);
        node_features<D>(a, b, path, depth, lengths, inv_total, f);
    } else {
#pragma unroll
        for (int j = 0; j < IN; ++j) f[j] = 0;
    }
    store_row(buf, lane, f);
    __syncthreads();
    wave_layer_t<IN>(buf, wt, p + OFF_B1, lane);
    wave_layer_t<HID>(buf, wt + IN * HID, p + OFF_B2, lane);
    float logit[4];
    wave_layer_logit(buf, wt + IN * HID + HID * HID, p + OFF_B3, wb + IN * HID + 2 * HID * HID, lane, logit);
    const float b4 = p[OFF_B4];
    // lanes 0..15 write waypoints st * 16 + lane
    if (lane < 16) {
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            const uint32_t oo = c0 + (uint32_t)(st * 16 + lane);
            if (oo < W) scores[e * kScoreStride + oo] = logit[st] + b4;
        }
    }
}

// The surrogate's picks of edges e < n with mask[e] set (or all) whose tree
// is regular with at most kMaxPickDepth levels: the reference ordinals of
// each edge's kPicks best-scored top-tree nodes -- descending score, ties by
// ordinal -- at picks[e * kPicks + r], in that order (fewer on trees of fewer
// nodes).
//
// Persistent blocks of kPickWaves waves (two per SIMD) with the weights
// staged once into LDS (W^T rows padded by 8 elements against bank
// conflicts).  Block b takes the edge slots [b S, (b + 1) S), kPickChunk at a
// time: the block lists the chunk's live edges in LDS, then its waves take
// tiles of kPickEdges listed edges in turn.  A tile has 16 rows per edge
// (edge slot j = lane / 16, node lane % 16 of the top tree in breadth-first
// order, 15 at most).  A row's arithmetic is that of score_waypoints, so a
// waypoint gets the same score.
constexpr int kPickWaves = 8, kPickEdges = 2, kPickChunk = 64 * kPickWaves;
constexpr int kW1S = IN + 8, kWS = HID + 8;
struct PickLds {
    uint16_t w1[HID * kW1S], w2[HID * kWS], w3[HID * kWS];
    float b1[HID], b2[HID], b3[HID];
    uint16_t w4[HID];
    uint16_t act[kPickWaves][16 * kPickEdges * kLdsRow];
    uint32_t list[kPickChunk];  // (slot offset in the chunk << 6) | h
    uint32_t wsum[kPickWaves];
};

template <int D>
__global__ __launch_bounds__(64 * kPickWaves) void score_picks(const double* __restrict__ from,
                                                               const double* __restrict__ to, int64_t n,
                                                               const uint8_t* __restrict__ mask, double tol,
                                                               int64_t per_block, const double* __restrict__ lengths,
                                                               float inv_total, const uint16_t* __restrict__ wt,
                                                               const uint16_t* __restrict__ wb,
                                                               const float* __restrict__ p,
                                                               uint16_t* __restrict__ picks) {
    __shared__ __attribute__((aligned(16))) PickLds L;
    const int t = threadIdx.x;
    // stage W1^T [128][64], W2^T, W3^T [128][128] in 16-byte chunks: every
    // load of a thread issued before its stores
    {
        constexpr int T = 64 * kPickWaves, N1 = HID * IN / 8 / T, N2 = HID * HID / 8 / T;
        uint4 v1[N1], v2[N2], v3[N2];
#pragma unroll
        for (int k = 0; k < N1; ++k) v1[k] = reinterpret_cast<const uint4*>(wt)[t + k * T];
#pragma unroll
        for (int k = 0; k < N2; ++k) {
            v2[k] = reinterpret_cast<const uint4*>(wt + IN * HID)[t + k * T];
            v3[k] = reinterpret_cast<const uint4*>(wt + IN * HID + HID * HID)[t + k * T];
        }
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const int i = t + k * T, row = i / (IN / 8), c = i % (IN / 8);
            *reinterpret_cast<uint4*>(L.w1 + row * kW1S + c * 8) = v1[k];
        }
#pragma unroll
        for (int k = 0; k < N2; ++k) {
            const int i = t + k * T, row = i / (HID / 8), c = i % (HID / 8);
            *reinterpret_cast<uint4*>(L.w2 + row * kWS + c * 8) = v2[k];
            *reinterpret_cast<uint4*>(L.w3 + row * kWS + c * 8) = v3[k];
        }
    }
    for (int i = t; i < HID; i += 64 * kPickWaves) {
        L.b1[i] = p[OFF_B1 + i];
        L.b2[i] = p[OFF_B2 + i];
        L.b3[i] = p[OFF_B3 + i];
        L.w4[i] = wb[IN * HID + 2 * HID * HID + i];
    }
    const float b4 = p[OFF_B4];
    const int wave = t >> 6, lane = t & 63;
    const int slot = lane >> 4, node = lane & 15;
    uint16_t* buf = L.act[wave];
    const int64_t s0 = (int64_t)blockIdx.x * per_block;
    const int64_t s1 = s0 + per_block < n ? s0 + per_block : n;
    // block-uniform chunk loop: every wave reaches each barrier
    for (int64_t c0 = s0; c0 < s1; c0 += kPickChunk) {
        // the chunk's live edges, listed in slot order
        const int64_t ei = c0 + t;
        bool live = ei < s1 && (!mask || mask[ei]);
        int hh = 0;
        if (live) {
            bool regular = false;
            live = bisect_shape<D>(from + ei * D, to + ei * D, tol, &hh, &regular) && regular && hh <= kMaxPickDepth;
        }
        const uint64_t lm = __ballot(live);
        __syncthreads();  // the previous chunk's tiles are done with the list (and, once, the weights are staged)
        if (lane == 0) L.wsum[wave] = (uint32_t)__popcll(lm);
        __syncthreads();
        uint32_t before = 0, cnt = 0;
#pragma unroll
        for (int w = 0; w < kPickWaves; ++w) {
            before += w < wave ? L.wsum[w] : 0u;
            cnt += L.wsum[w];
        }
        if (live) L.list[before + __popcll(lm & (lane ? (~0ull >> (64 - lane)) : 0ull))] = ((uint32_t)t << 6) | (uint32_t)hh;
        __syncthreads();
        const int tiles = (int)((cnt + kPickEdges - 1) / kPickEdges);
        for (int tile = wave; tile < tiles; tile += kPickWaves) {
            // lanes of slots >= kPickEdges hold no row (they load and
            // multiply fragments like the others)
            const int j = tile * kPickEdges + slot;
            const bool have = slot < kPickEdges && (uint32_t)j < cnt;
            const uint32_t ent = have ? L.list[j] : 0u;
            const int64_t e = c0 + (int64_t)(ent >> 6);
            const int h = (int)(ent & 63u);
            const uint32_t T = have ? (1u << (h < kPickLevels ? h : kPickLevels)) - 1u : 0u;
            const bool row = (uint32_t)node < T;
            uint16_t f[IN];
            uint32_t ord = 0;
            if (row) {
                uint32_t path;
                int depth;
                bfs_path((uint32_t)node, &path, &depth);
                ord = path_preorder(h, path, depth);
                node_features<D>(from + e * D, to + e * D, path, depth, lengths, inv_total, f);
            } else {
#pragma unroll
                for (int q = 0; q < IN; ++q) f[q] = 0;
            }
            if (slot < kPickEdges) store_row(buf, lane, f);
            buf_sync<true>();
            wave_layer_t<IN, kW1S, true, kPickEdges>(buf, L.w1, L.b1, lane);
            wave_layer_t<HID, kWS, true, kPickEdges>(buf, L.w2, L.b2, lane);
            float logit[kPickEdges];
            wave_layer_logit<kWS, kPickEdges>(buf, L.w3, L.b3, L.w4, lane, logit);
            // every lane holds the logits of rows st * 16 + (lane & 15): its
            // own row is st = slot
            const float sc = (slot == 0 ? logit[0] : logit[kPickEdges - 1]) + b4;
            uint64_t key = row ? score_key(sc, ord) : ~0ull;
            // kPicks rounds of a minimum over the edge's 16 lanes; the lane
            // holding it (keys are distinct: ordinals differ) writes its
            // ordinal and drops out
#pragma unroll
            for (int r = 0; r < kPicks; ++r) {
                uint64_t m = key;
#pragma unroll
                for (int o = 8; o > 0; o >>= 1) {
                    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)m, o);
                    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(m >> 32), o);
                    const uint64_t x = ((uint64_t)hi << 32) | lo;
                    m = x < m ? x : m;
                }
                if (m != ~0ull && key == m) {
                    picks[e * kPicks + r] = (uint16_t)ord;
                    key = ~0ull;
                }
            }
            buf_sync<true>();  // the next tile's rows overwrite the buffer
        }
    }
}

// Training samples: one waypoint per edge (ordinal uniform over the nominal
// bisection tree, Philox(seed, edge)), its features and the exact label
// (1 = the waypoint collides).
template <int D>
__global__ __launch_bounds__(256) void sample_waypoints(const double* __restrict__ from, const double* __restrict__ to,
                                                        int64_t n, double tol, SceneArgs sc, float inv_total,
                                                        uint64_t seed, uint16_t* __restrict__ feat,
                                                        float* __restrict__ label) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double* a = from + e * D;
    const double* b = to + e * D;
    int h = 0;
    bool regular = false;
    double q[D];
    if (!bisect_shape<D>(a, b, tol, &h, &regular) || h == 0) {
#pragma unroll
        for (int i = 0; i < D; ++i) q[i] = b[i];
    } else {
        const uint32_t W = (1u << h) - 1u;
        uint32_t o = (uint32_t)(uniform_at(seed, (uint64_t)e, 0x53414d57u, 0) * (double)W);
        if (o >= W) o = W - 1;
        uint32_t path;
        int depth;
        preorder_path(h, o, &path, &depth);
#pragma unroll
        for (int i = 0; i < D; ++i) q[i] = coord_mid(a[i], b[i], path, depth);
    }
    uint16_t f[IN];
    arm_features<D>(q, sc.lengths, inv_total, f);
#pragma unroll
    for (int j = 0; j < IN; ++j) feat[e * IN + j] = f[j];
    label[e] = arm_ok<D>(sc, q) ? 0.f : 1.f;
}

// Device state of the surrogate (owned by the context).
struct Surrogate {
    bool init = false;
    int64_t cap = 0;  // rows the batch buffers hold
    int32_t adam_t = 0;
    DevBuf params, m, v, grad, wb, wtb, loss;
    DevBuf x, xt, y, h1, h1t, h2, h2t, h3, h3t, z, dz, dh3, dh3t, dh2, dh2t, dh1, dh1t;
};

Surrogate& surrogate(mptg_ctx* ctx) {
    if (!ctx->surrogate) ctx->surrogate = new Surrogate();
    return *ctx->surrogate;
}

void surrogate_free(mptg_ctx* ctx) {
    if (!ctx->surrogate) return;
    Surrogate& S = *ctx->surrogate;
    for (DevBuf* b : {&S.params, &S.m, &S.v, &S.grad, &S.wb, &S.wtb, &S.loss, &S.x, &S.xt, &S.y, &S.h1, &S.h1t,
                      &S.h2, &S.h2t, &S.h3, &S.h3t, &S.z, &S.dz, &S.dh3, &S.dh3t, &S.dh2, &S.dh2t, &S.dh1, &S.dh1t})
        (void)hipFree(b->p);
    delete ctx->surrogate;
    ctx->surrogate = nullptr;
}

static constexpr int64_t kWb = IN * HID + 2 * HID * HID + HID;  // BF16 weight copies

static int sur_reserve(mptg_ctx* ctx, Surrogate& S, int64_t rows) {
    if (rows <= S.cap) return MPTG_OK;
    int rc;
    const size_t bh = sizeof(uint16_t) * HID * rows;
    if ((rc = ensure(ctx, S.x, sizeof(uint16_t) * IN * rows))) return rc;
    if ((rc = ensure(ctx, S.xt, sizeof(uint16_t) * IN * rows))) return rc;
    if ((rc = ensure(ctx, S.y, sizeof(float) * rows))) return rc;
    for (DevBuf* b : {&S.h1, &S.h1t, &S.h2, &S.h2t, &S.h3, &S.h3t, &S.dh3, &S.dh3t, &S.dh2, &S.dh2t, &S.dh1, &S.dh1t})
        if ((rc = ensure(ctx, *b, bh))) return rc;
    if ((rc = ensure(ctx, S.z, sizeof(float) * rows))) return rc;
    if ((rc = ensure(ctx, S.dz, sizeof(float) * rows))) return rc;
    S.cap = rows;
    return MPTG_OK;
}

static int sur_alloc(mptg_ctx* ctx, Surrogate& S) {
    int rc;
    for (DevBuf* b : {&S.params, &S.m, &S.v, &S.grad})
        if ((rc = ensure(ctx, *b, sizeof(float) * NPARAM))) return rc;
    if ((rc = ensure(ctx, S.wb, sizeof(uint16_t) * kWb))) return rc;
    if ((rc = ensure(ctx, S.wtb, sizeof(uint16_t) * kWb))) return rc;
    if ((rc = ensure(ctx, S.loss, sizeof(float) * 4))) return rc;
    return MPTG_OK;
}

static int sur_cast(mptg_ctx* ctx, Surrogate& S) {
    hipLaunchKernelGGL(cast_weights, dim3((unsigned)((kWb + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const float*)S.params.p, (uint16_t*)S.wb.p, (uint16_t*)S.wtb.p);
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

static void gemm(mptg_ctx* ctx, int mode, int M, int N, int K, const void* A, const void* Bt, Epi ep) {
    unsigned blocks = (unsigned)(((int64_t)(M / 16) * (N / 16) + 3) / 4);
    if (mode == EPI_FWD)
        hipLaunchKernelGGL(gemm_nt<EPI_FWD>, dim3(blocks), dim3(256), 0, ctx->stream, M, N, K, (const uint16_t*)A,
                           (const uint16_t*)Bt, ep);
    else if (mode == EPI_DX)
        hipLaunchKernelGGL(gemm_nt<EPI_DX>, dim3(blocks), dim3(256), 0, ctx->stream, M, N, K, (const uint16_t*)A,
                           (const uint16_t*)Bt, ep);
    else
        hipLaunchKernelGGL(gemm_nt<EPI_F32>, dim3(blocks), dim3(256), 0, ctx->stream, M, N, K, (const uint16_t*)A,
                           (const uint16_t*)Bt, ep);
}

// Forward over M rows of S.x (M multiple of 16); logits to S.z.  With
// labels (S.y) also the loss, dz and dH3.
static void sur_forward(mptg_ctx* ctx, Surrogate& S, int M, bool with_loss) {
    const uint16_t* wt = (const uint16_t*)S.wtb.p;
    const float* p = (const float*)S.params.p;
    gemm(ctx, EPI_FWD, M, HID, IN, S.x.p, wt, Epi{p + OFF_B1, nullptr, (uint16_t*)S.h1.p, (uint16_t*)S.h1t.p, nullptr});
    gemm(ctx, EPI_FWD, M, HID, HID, S.h1.p, wt + IN * HID,
         Epi{p + OFF_B2, nullptr, (uint16_t*)S.h2.p, (uint16_t*)S.h2t.p, nullptr});
    gemm(ctx, EPI_FWD, M, HID, HID, S.h2.p, wt + IN * HID + HID * HID,
         Epi{p + OFF_B3, nullptr, (uint16_t*)S.h3.p, (uint16_t*)S.h3t.p, nullptr});
    hipLaunchKernelGGL(head_kernel, dim3((unsigned)((M + kHeadRows - 1) / kHeadRows)), dim3(256), 0, ctx->stream, M,
                       (const uint16_t*)S.h3.p, (const uint16_t*)S.wb.p + IN * HID + 2 * HID * HID, p + OFF_B4,
                       with_loss ? (const float*)S.y.p : nullptr, 1.0f / (float)M, (float*)S.z.p, (float*)S.dz.p,
                       (uint16_t*)S.dh3.p, (uint16_t*)S.dh3t.p, (float*)S.loss.p);
}

// Gradient of the mean loss over the M rows already forwarded; into S.grad.
static void sur_backward(mptg_ctx* ctx, Surrogate& S, int M) {
    float* g = (float*)S.grad.p;
    const uint16_t* wb = (const uint16_t*)S.wb.p;
    hipLaunchKernelGGL(row_sums, dim3(HID), dim3(256), 0, ctx->stream, (const uint16_t*)S.h3t.p, M,
                       (const float*)S.dz.p, g + OFF_W4);
    hipLaunchKernelGGL(sum_f32, dim3(1), dim3(256), 0, ctx->stream, (const float*)S.dz.p, M, g + OFF_B4);
    // layer 3
    gemm(ctx, EPI_F32, HID, HID, M, S.h2t.p, S.dh3t.p, Epi{nullptr, nullptr, nullptr, nullptr, g + OFF_W3});
    hipLaunchKernelGGL(row_sums, dim3(HID), dim3(256), 0, ctx->stream, (const uint16_t*)S.dh3t.p, M,
                       (const float*)nullptr, g + OFF_B3);
    gemm(ctx, EPI_DX, M, HID, HID, S.dh3.p, wb + IN * HID + HID * HID,
         Epi{nullptr, (const uint16_t*)S.h2.p, (uint16_t*)S.dh2.p, (uint16_t*)S.dh2t.p, nullptr});
    // layer 2
    gemm(ctx, EPI_F32, HID, HID, M, S.h1t.p, S.dh2t.p, Epi{nullptr, nullptr, nullptr, nullptr, g + OFF_W2});
    hipLaunchKernelGGL(row_sums, dim3(HID), dim3(256), 0, ctx->stream, (const uint16_t*)S.dh2t.p, M,
                       (const float*)nullptr, g + OFF_B2);
    gemm(ctx, EPI_DX, M, HID, HID, S.dh2.p, wb + IN * HID,
         Epi{nullptr, (const uint16_t*)S.h1.p, (uint16_t*)S.dh1.p, (uint16_t*)S.dh1t.p, nullptr});
    // layer 1
    hipLaunchKernelGGL(transpose_bf16, dim3((unsigned)(((int64_t)M * IN + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const uint16_t*)S.x.p, M, IN, (uint16_t*)S.xt.p);
    gemm(ctx, EPI_F32, IN, HID, M, S.xt.p, S.dh1t.p, Epi{nullptr, nullptr, nullptr, nullptr, g + OFF_W1});
    hipLaunchKernelGGL(row_sums, dim3(HID), dim3(256), 0, ctx->stream, (const uint16_t*)S.dh1t.p, M,
                       (const float*)nullptr, g + OFF_B1);
}

static int sur_check(mptg_ctx* ctx, bool need_init) {
    Surrogate& S = surrogate(ctx);
    if (need_init && !S.init) return fail(ctx, MPTG_E_STATE, "surrogate not initialised (mptg_surrogate_init)");
    return MPTG_OK;
}

static int score_check(mptg_ctx* ctx) {
    Surrogate& S = surrogate(ctx);
    if (!S.init) return fail(ctx, MPTG_E_STATE, "surrogate order needs an initialised surrogate");
    if (ctx->scene.kind != MPTG_SCENE_ARM || ctx->dim != ctx->scene.nlinks)
        return fail(ctx, MPTG_E_STATE, "surrogate order is defined for the arm scene");
    if (ctx->dim != 5 && ctx->dim != 10 && ctx->dim != 20)
        return fail(ctx, MPTG_E_STATE, "surrogate order: arms of 5, 10 or 20 links");
    return MPTG_OK;
}

#define MPTG_SCORE_ARGS()                                                                                        \
    const Surrogate& S = surrogate(ctx);                                                                         \
    const float inv_total = (float)(1.0 / ctx->scene.total_length);                                              \
    const uint16_t* wt = (const uint16_t*)S.wtb.p;                                                               \
    const uint16_t* wb = (const uint16_t*)S.wb.p;                                                                \
    const float* p = (const float*)S.params.p;                                                                   \
    const double* L = (const double*)ctx->scene.lengths

int surrogate_score_edges(mptg_ctx* ctx, const double* from, const double* to, int64_t n, const uint8_t* mask,
                          float* scores) {
    int rc;
    if ((rc = score_check(ctx))) return rc;
    if (n <= 0) return MPTG_OK;
    MPTG_SCORE_ARGS();
    const int tiles = kScoreStride / 64;  // 64-waypoint tiles per edge
    const unsigned blocks = (unsigned)(n * tiles);
    switch (ctx->dim) {
        case 5: hipLaunchKernelGGL(score_waypoints<5>, dim3(blocks), dim3(64), 0, ctx->stream, from, to, n, mask, ctx->link_param, L, inv_total, wt, wb, p, scores, tiles); break;
        case 10: hipLaunchKernelGGL(score_waypoints<10>, dim3(blocks), dim3(64), 0, ctx->stream, from, to, n, mask, ctx->link_param, L, inv_total, wt, wb, p, scores, tiles); break;
        default: hipLaunchKernelGGL(score_waypoints<20>, dim3(blocks), dim3(64), 0, ctx->stream, from, to, n, mask, ctx->link_param, L, inv_total, wt, wb, p, scores, tiles); break;
    }
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

int surrogate_picks(mptg_ctx* ctx, const double* from, const double* to, int64_t n, const uint8_t* mask,
                    uint16_t* picks) {
    int rc;
    if ((rc = score_check(ctx))) return rc;
    if (n <= 0) return MPTG_OK;
    MPTG_SCORE_ARGS();
    // one block per CU at most (the LDS holds one), a few tiles per wave at
    // the least
    const int64_t blocks = std::max<int64_t>(
        1, std::min<int64_t>(ctx->cus, (n + 2 * kPickEdges * kPickWaves - 1) / (2 * kPickEdges * kPickWaves)));
    const int64_t per_block = (n + blocks - 1) / blocks;
    const dim3 g((unsigned)blocks), blk(64 * kPickWaves);
    switch (ctx->dim) {
        case 5: hipLaunchKernelGGL(score_picks<5>, g, blk, 0, ctx->stream, from, to, n, mask, ctx->link_param, per_block, L, inv_total, wt, wb, p, picks); break;
        case 10: hipLaunchKernelGGL(score_picks<10>, g, blk, 0, ctx->stream, from, to, n, mask, ctx->link_param, per_block, L, inv_total, wt, wb, p, picks); break;
        default: hipLaunchKernelGGL(score_picks<20>, g, blk, 0, ctx->stream, from, to, n, mask, ctx->link_param, per_block, L, inv_total, wt, wb, p, picks); break;
    }
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}
#undef MPTG_SCORE_ARGS

}  // namespace mptg

using namespace mptg;

extern "C" {

int64_t mptg_surrogate_nparams(void) { return NPARAM; }

// One training sample per edge: a waypoint of link(from_i, to_i), its
// features and its exact label (device buffers).
int mptg_surrogate_samples(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, uint64_t seed,
                           uint16_t* d_feat, float* d_label) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (ctx->scene.kind != MPTG_SCENE_ARM || ctx->dim != ctx->scene.nlinks || ctx->link_mode != MPTG_LINK_BISECT)
        return fail(ctx, MPTG_E_STATE, "surrogate samples are defined for the arm scene with bisection links");
    if (n < 0 || (n > 0 && (!d_from || !d_to || !d_feat || !d_label))) return fail(ctx, MPTG_E_ARG, "bad buffers");
    if (n == 0) return MPTG_OK;
    const float inv_total = (float)(1.0 / ctx->scene.total_length);
    const unsigned blocks = (unsigned)((n + 255) / 256);
    const SceneArgs sa = scene_args(ctx);
    switch (ctx->dim) {
        case 5: hipLaunchKernelGGL(sample_waypoints<5>, dim3(blocks), dim3(256), 0, ctx->stream, d_from, d_to, n, ctx->link_param, sa, inv_total, seed, d_feat, d_label); break;
        case 10: hipLaunchKernelGGL(sample_waypoints<10>, dim3(blocks), dim3(256), 0, ctx->stream, d_from, d_to, n, ctx->link_param, sa, inv_total, seed, d_feat, d_label); break;
        case 20: hipLaunchKernelGGL(sample_waypoints<20>, dim3(blocks), dim3(256), 0, ctx->stream, d_from, d_to, n, ctx->link_param, sa, inv_total, seed, d_feat, d_label); break;
        default: return fail(ctx, MPTG_E_ARG, "surrogate samples: arms of 5, 10 or 20 links");
    }
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

// Surrogate scores of every waypoint of every edge: d_scores[i * 512 + o]
// for reference ordinal o (slots beyond the edge's waypoint count untouched).
int mptg_surrogate_score_edges(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n,
                               float* d_scores) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!d_from || !d_to || !d_scores))) return fail(ctx, MPTG_E_ARG, "bad buffers");
    if (n == 0) return MPTG_OK;
    return surrogate_score_edges(ctx, d_from, d_to, n, nullptr, d_scores);
}

int mptg_surrogate_init(mptg_ctx* ctx, uint64_t seed) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    Surrogate& S = surrogate(ctx);
    if ((rc = sur_alloc(ctx, S))) return rc;
    hipLaunchKernelGGL(init_params, dim3((unsigned)((NPARAM + 255) / 256)), dim3(256), 0, ctx->stream,
                       (float*)S.params.p, seed);
    MPTG_HIP(ctx, hipMemsetAsync(S.m.p, 0, sizeof(float) * NPARAM, ctx->stream));
    MPTG_HIP(ctx, hipMemsetAsync(S.v.p, 0, sizeof(float) * NPARAM, ctx->stream));
    if ((rc = sur_cast(ctx, S))) return rc;
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    S.adam_t = 0;
    S.init = true;
    return MPTG_OK;
}

int mptg_surrogate_get(mptg_ctx* ctx, float* h_params, int64_t n) {
    int rc = check_device(ctx);
    if (rc || (rc = sur_check(ctx, true))) return rc;
    if (n != NPARAM || !h_params) return fail(ctx, MPTG_E_ARG, "parameter buffer must hold mptg_surrogate_nparams()");
    MPTG_HIP(ctx, hipMemcpyAsync(h_params, surrogate(ctx).params.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost,
                                 ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

int mptg_surrogate_set(mptg_ctx* ctx, const float* h_params, int64_t n) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (n != NPARAM || !h_params) return fail(ctx, MPTG_E_ARG, "parameter buffer must hold mptg_surrogate_nparams()");
    Surrogate& S = surrogate(ctx);
    if ((rc = sur_alloc(ctx, S))) return rc;
    MPTG_HIP(ctx, hipMemcpyAsync(S.params.p, h_params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, ctx->stream));
    MPTG_HIP(ctx, hipMemsetAsync(S.m.p, 0, sizeof(float) * NPARAM, ctx->stream));
    MPTG_HIP(ctx, hipMemsetAsync(S.v.p, 0, sizeof(float) * NPARAM, ctx->stream));
    if ((rc = sur_cast(ctx, S))) return rc;
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    S.adam_t = 0;
    S.init = true;
    return MPTG_OK;
}

int mptg_surrogate_features(mptg_ctx* ctx, const double* d_states, int64_t n, uint16_t* d_feat) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (ctx->scene.kind != MPTG_SCENE_ARM || ctx->dim != ctx->scene.nlinks)
        return fail(ctx, MPTG_E_STATE, "surrogate features are defined for the arm scene");
    if (n < 0 || (n > 0 && (!d_states || !d_feat))) return fail(ctx, MPTG_E_ARG, "bad buffers");
    if (n == 0) return MPTG_OK;
    float inv_total = (float)(1.0 / ctx->scene.total_length);
    unsigned blocks = (unsigned)((n + 255) / 256);
    switch (ctx->dim) {
        case 5: hipLaunchKernelGGL(features_kernel<5>, dim3(blocks), dim3(256), 0, ctx->stream, d_states, n, (const double*)ctx->scene.lengths, inv_total, d_feat); break;
        case 10: hipLaunchKernelGGL(features_kernel<10>, dim3(blocks), dim3(256), 0, ctx->stream, d_states, n, (const double*)ctx->scene.lengths, inv_total, d_feat); break;
        case 20: hipLaunchKernelGGL(features_kernel<20>, dim3(blocks), dim3(256), 0, ctx->stream, d_states, n, (const double*)ctx->scene.lengths, inv_total, d_feat); break;
        default: return fail(ctx, MPTG_E_ARG, "surrogate features: arms of 5, 10 or 20 links");
    }
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

// Logits of n feature rows (device buffers).
int mptg_surrogate_forward(mptg_ctx* ctx, const uint16_t* d_feat, int64_t n, float* d_logits) {
    int rc = check_device(ctx);
    if (rc || (rc = sur_check(ctx, true))) return rc;
    if (n < 0 || (n > 0 && (!d_feat || !d_logits))) return fail(ctx, MPTG_E_ARG, "bad buffers");
    Surrogate& S = surrogate(ctx);
    const int64_t chunk = 1 << 16;
    for (int64_t r0 = 0; r0 < n; r0 += chunk) {
        int64_t m = std::min(chunk, n - r0);
        int M = (int)((m + 15) / 16 * 16);
        if ((rc = sur_reserve(ctx, S, M))) return rc;
        MPTG_HIP(ctx, hipMemsetAsync(S.x.p, 0, sizeof(uint16_t) * IN * M, ctx->stream));
        MPTG_HIP(ctx, hipMemcpyAsync(S.x.p, d_feat + r0 * IN, sizeof(uint16_t) * IN * m, hipMemcpyDeviceToDevice,
                                     ctx->stream));
        sur_forward(ctx, S, M, false);
        MPTG_HIP(ctx, hipMemcpyAsync(d_logits + r0, S.z.p, sizeof(float) * m, hipMemcpyDeviceToDevice, ctx->stream));
    }
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

// Mean loss and its gradient over n rows (one batch; n multiple of 16).
int mptg_surrogate_loss_grad(mptg_ctx* ctx, const uint16_t* d_feat, const float* d_label, int64_t n, float* h_loss,
                             float* h_grad) {
    int rc = check_device_ready(ctx);
    if (rc || (rc = sur_check(ctx, true))) return rc;
    if (n <= 0 || n % 16 || n > (1 << 20) || !d_feat || !d_label)
        return fail(ctx, MPTG_E_ARG, "batch rows must be a positive multiple of 16 (at most 2^20)");
    Surrogate& S = surrogate(ctx);
    if ((rc = sur_reserve(ctx, S, n))) return rc;
    MPTG_HIP(ctx, hipMemcpyAsync(S.x.p, d_feat, sizeof(uint16_t) * IN * n, hipMemcpyDeviceToDevice, ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(S.y.p, d_label, sizeof(float) * n, hipMemcpyDeviceToDevice, ctx->stream));
    MPTG_HIP(ctx, hipMemsetAsync(S.loss.p, 0, sizeof(float), ctx->stream));
    sur_forward(ctx, S, (int)n, true);
    sur_backward(ctx, S, (int)n);
    MPTG_HIP(ctx, hipGetLastError());
    if (h_loss) MPTG_HIP(ctx, hipMemcpyAsync(h_loss, S.loss.p, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    if (h_grad)
        MPTG_HIP(ctx, hipMemcpyAsync(h_grad, S.grad.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

// Adam training over n rows (device features/labels), in order, `epochs`
// passes of batches of `batch` rows (a multiple of 16; a short last batch
// is dropped).  h_losses (optional) receives the mean loss of each epoch.
int mptg_surrogate_train(mptg_ctx* ctx, const uint16_t* d_feat, const float* d_label, int64_t n, int epochs,
                         int batch, float lr, float* h_losses) {
    int rc = check_device_ready(ctx);
    if (rc || (rc = sur_check(ctx, true))) return rc;
    if (batch <= 0 || batch % 16 || batch > (1 << 16)) return fail(ctx, MPTG_E_ARG, "batch must be a multiple of 16");
    if (n < batch || epochs < 1 || !(lr > 0) || !d_feat || !d_label) return fail(ctx, MPTG_E_ARG, "bad training args");
    Surrogate& S = surrogate(ctx);
    if ((rc = sur_reserve(ctx, S, batch))) return rc;
    const int64_t nb = n / batch;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
        MPTG_HIP(ctx, hipMemsetAsync(S.loss.p, 0, sizeof(float) * 2, ctx->stream));
        for (int64_t bi = 0; bi < nb; ++bi) {
            const int64_t r0 = bi * batch;
            MPTG_HIP(ctx, hipMemcpyAsync(S.x.p, d_feat + r0 * IN, sizeof(uint16_t) * IN * batch,
                                         hipMemcpyDeviceToDevice, ctx->stream));
            MPTG_HIP(ctx, hipMemcpyAsync(S.y.p, d_label + r0, sizeof(float) * batch, hipMemcpyDeviceToDevice,
                                         ctx->stream));
            sur_forward(ctx, S, batch, true);
            sur_backward(ctx, S, batch);
            S.adam_t += 1;
            float c1 = 1.f - std::pow(b1, (float)S.adam_t), c2 = 1.f - std::pow(b2, (float)S.adam_t);
            hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((NPARAM + 255) / 256)), dim3(256), 0, ctx->stream,
                               (float*)S.params.p, (float*)S.m.p, (float*)S.v.p, (const float*)S.grad.p,
                               (int64_t)NPARAM, lr, b1, b2, eps, c1, c2);
            if ((rc = sur_cast(ctx, S))) return rc;
        }
        MPTG_HIP(ctx, hipGetLastError());
        if (h_losses) {
            float tot = 0.f;
            MPTG_HIP(ctx, hipMemcpyAsync(&tot, S.loss.p, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
            MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
            h_losses[ep] = tot / (float)nb;
        }
    }
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

}  // extern "C"
