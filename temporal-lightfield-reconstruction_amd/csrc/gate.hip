// Learned gate MLP (24-64-64-1) on MFMA: BF16 operands, FP32 accumulation.
//
// One wavefront processes 32 feature rows per iteration:
//   layer 1  H1 = relu(X W1 + b1)        2x4 MFMA 16x16x32 (K = 24 padded to 32)
//   layer 2  H2 = relu(H1 W2 + b2)       2x4x2 MFMA
//   layer 3  z  = H2 w3 + b3             FP32 dot products
// Backward (training):
//   dz = (sigmoid(z) - y) * weight(y)
//   dW2 += H1^T dH2                      4x4 MFMA, K = 32 rows
//   dH1  = dH2 W2^T (masked)             2x4x2 MFMA
//   dW1 += X^T dH1                       2x4 MFMA, K = 32 rows
// MFMA fragment layout (v_mfma_f32_16x16x32_bf16, 64 lanes): lane l holds
// A[m = l%16][k = 8*(l/16) .. +7], B[k = 8*(l/16) .. +7][n = l%16] and
// D[m = 4*(l/16) + j][n = l%16], j = 0..3.
// Operands are rounded to BF16 (round to nearest even) when fragments are
// built from FP32 LDS tiles; the host check rounds at the same places.
#include "gate.hpp"

namespace tlr {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kRows = 32;      // feature rows per wave iteration
constexpr int kXld = 32;       // padded input width

__device__ __forceinline__ __bf16 to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    const uint16_t h = (uint16_t)(u >> 16);
    return __builtin_bit_cast(__bf16, h);
}
__device__ __forceinline__ float bf16_round(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xFFFF0000u);
}

// A fragment of a row-major tile: element (m, k) = src[m*ld + k]
__device__ __forceinline__ bf16x8 frag_a(const float* src, int ld, int lane) {
    const int m = lane & 15, kb = (lane >> 4) * 8;
    bf16x8 a;
#pragma unroll
    for (int j = 0; j < 8; j++) a[j] = to_bf16(src[m * ld + kb + j]);
    return a;
}
// A fragment of a transposed tile: element (m, k) = src[k*ld + m]
__device__ __forceinline__ bf16x8 frag_at(const float* src, int ld, int lane) {
    const int m = lane & 15, kb = (lane >> 4) * 8;
    bf16x8 a;
#pragma unroll
    for (int j = 0; j < 8; j++) a[j] = to_bf16(src[(kb + j) * ld + m]);
    return a;
}
// B fragment of a row-major K x N tile: element (k, n) = src[k*ld + n]
__device__ __forceinline__ bf16x8 frag_b(const float* src, int ld, int lane) {
    const int n = lane & 15, kb = (lane >> 4) * 8;
    bf16x8 b;
#pragma unroll
    for (int j = 0; j < 8; j++) b[j] = to_bf16(src[(kb + j) * ld + n]);
    return b;
}
// B fragment of a transposed tile: element (k, n) = src[n*ld + k]
__device__ __forceinline__ bf16x8 frag_bt(const float* src, int ld, int lane) {
    const int n = lane & 15, kb = (lane >> 4) * 8;
    bf16x8 b;
#pragma unroll
    for (int j = 0; j < 8; j++) b[j] = to_bf16(src[n * ld + kb + j]);
    return b;
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

struct GateLds {
    float W1[kXld * kGateHid];        // rows 24..31 zero
    float W2[kGateHid * kGateHid];
    float b1[kGateHid], b2[kGateHid], w3[kGateHid];
    float b3;
    float X[kRows * kXld];
    float H1[kRows * kGateHid];
    float H2[kRows * kGateHid];       // H2, then dH2
    float D1[kRows * kGateHid];       // dH1
    float z[kRows];
};

__device__ void load_weights(GateLds& L, const float* __restrict__ P, int lane) {
    for (int i = lane; i < kXld * kGateHid; i += 64) {
        const int k = i / kGateHid;
        L.W1[i] = k < kGateIn ? P[kGateOffW1 + i] : 0.0f;
    }
    for (int i = lane; i < kGateHid * kGateHid; i += 64) L.W2[i] = P[kGateOffW2 + i];
    L.b1[lane] = P[kGateOffB1 + lane];
    L.b2[lane] = P[kGateOffB2 + lane];
    L.w3[lane] = bf16_round(P[kGateOffW3 + lane]);
    if (lane == 0) L.b3 = P[kGateOffB3];
    __syncthreads();
}

__device__ void load_rows(GateLds& L, const float* __restrict__ X, const int* __restrict__ idx, int base, int n,
                          int lane) {
    for (int i = lane; i < kRows * kXld; i += 64) {
        const int r = i / kXld, c = i % kXld;
        const int row = base + r;
        float v = 0.0f;
        if (row < n && c < kGateIn) {
            const int src = idx ? idx[row] : row;
            v = X[(size_t)src * kGateIn + c];
        }
        L.X[i] = v;
    }
    __syncthreads();
}

// Forward pass of the 32 rows in L.X; leaves H1, H2 and z in LDS.
__device__ void forward(GateLds& L, int lane) {
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
        const bf16x8 a = frag_a(L.X + mt * 16 * kXld, kXld, lane);
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(a, frag_b(L.W1 + nt * 16, kGateHid, lane), acc);
            const int col = nt * 16 + (lane & 15);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int row = mt * 16 + 4 * (lane >> 4) + j;
                const float v = acc[j] + L.b1[col];
                L.H1[row * kGateHid + col] = v > 0.0f ? v : 0.0f;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
        const bf16x8 a0 = frag_a(L.H1 + mt * 16 * kGateHid, kGateHid, lane);
        const bf16x8 a1 = frag_a(L.H1 + mt * 16 * kGateHid + 32, kGateHid, lane);
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(a0, frag_b(L.W2 + nt * 16, kGateHid, lane), acc);
            acc = mfma(a1, frag_b(L.W2 + 32 * kGateHid + nt * 16, kGateHid, lane), acc);
            const int col = nt * 16 + (lane & 15);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int row = mt * 16 + 4 * (lane >> 4) + j;
                const float v = acc[j] + L.b2[col];
                L.H2[row * kGateHid + col] = v > 0.0f ? v : 0.0f;
            }
        }
    }
    __syncthreads();
    if (lane < kRows) {
        float z = 0.0f;
        for (int c = 0; c < kGateHid; c++) z += L.H2[lane * kGateHid + c] * L.w3[c];
        L.z[lane] = z + L.b3;
    }
    __syncthreads();
}

// Inference: the weights live in registers as MFMA B fragments (loaded once
// per wave), each wave takes 16 feature rows at a time, and only the layer-1
// activations pass through LDS (to become layer 2's A operand).  Layer 3 is
// reduced from layer 2's accumulators across the 16 lanes of each row group.
constexpr int kFwdBlock = 256;
constexpr int kFwdWaves = kFwdBlock / 64;

__global__ __launch_bounds__(kFwdBlock) void k_gate_forward(const float* __restrict__ X, int n,
                                                            const unsigned* __restrict__ n_dev,
                                                            const float* __restrict__ P, float* __restrict__ logits) {
    if (n_dev) n = (int)*n_dev;
    __shared__ float H1s[kFwdWaves][16 * kGateHid];
    const int lane = threadIdx.x & 63;
    float* H1 = H1s[threadIdx.x >> 6];
    const int col = lane & 15, kb = (lane >> 4) * 8;
    bf16x8 w1f[4], w2f[2][4];
    float b1c[4], b2c[4], w3c[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        const int c = nt * 16 + col;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = kb + j;
            w1f[nt][j] = to_bf16(k < kGateIn ? P[kGateOffW1 + k * kGateHid + c] : 0.0f);
            w2f[0][nt][j] = to_bf16(P[kGateOffW2 + k * kGateHid + c]);
            w2f[1][nt][j] = to_bf16(P[kGateOffW2 + (32 + k) * kGateHid + c]);
        }
        b1c[nt] = P[kGateOffB1 + c];
        b2c[nt] = P[kGateOffB2 + c];
        w3c[nt] = bf16_round(P[kGateOffW3 + c]);
    }
    const float b3 = P[kGateOffB3];
    const int gw = (gridDim.x * kFwdBlock) >> 6;
    for (int base = ((blockIdx.x * kFwdBlock + threadIdx.x) >> 6) * 16; base < n; base += gw * 16) {
        // A fragment of X: row base + col, k = kb .. kb+7 (k >= 24 is zero)
        bf16x8 a;
        const int row = base + col;
        if (row < n && kb < kGateIn) {
            const float4* src = reinterpret_cast<const float4*>(X + (size_t)row * kGateIn + kb);
            const float4 x0 = src[0], x1 = src[1];
            a[0] = to_bf16(x0.x); a[1] = to_bf16(x0.y); a[2] = to_bf16(x0.z); a[3] = to_bf16(x0.w);
            a[4] = to_bf16(x1.x); a[5] = to_bf16(x1.y); a[6] = to_bf16(x1.z); a[7] = to_bf16(x1.w);
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++) a[j] = to_bf16(0.0f);
        }
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(a, w1f[nt], acc);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float v = acc[j] + b1c[nt];
                H1[(4 * (lane >> 4) + j) * kGateHid + nt * 16 + col] = v > 0.0f ? v : 0.0f;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const bf16x8 a0 = frag_a(H1, kGateHid, lane);
        const bf16x8 a1 = frag_a(H1 + 32, kGateHid, lane);
        float zp[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(a0, w2f[0][nt], acc);
            acc = mfma(a1, w2f[1][nt], acc);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float v = acc[j] + b2c[nt];
                zp[j] += (v > 0.0f ? v : 0.0f) * w3c[nt];
            }
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
#pragma unroll
            for (int d = 1; d < 16; d <<= 1) zp[j] += __shfl_xor(zp[j], d, 64);
        }
        if (col == 0) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int r = base + 4 * (lane >> 4) + j;
                if (r < n) logits[r] = zp[j] + b3;
            }
        }
        // H1 is rewritten by the next tile
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ __launch_bounds__(64) void k_gate_grad(const float* __restrict__ X, const float* __restrict__ y,
                                                  const int* __restrict__ idx, int n, const float* __restrict__ P,
                                                  float* __restrict__ grad, float* __restrict__ loss, float pos_w) {
    __shared__ GateLds L;
    const int lane = threadIdx.x;
    load_weights(L, P, lane);
    f32x4 gW2[4][4], gW1[2][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) gW2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) gW1[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, lsum = 0.f;

    for (int base = blockIdx.x * kRows; base < n; base += gridDim.x * kRows) {
        load_rows(L, X, idx, base, n, lane);
        forward(L, lane);
        // dz, loss
        if (lane < kRows) {
            const int row = base + lane;
            float dz = 0.0f;
            if (row < n) {
                const float t = y[idx ? idx[row] : row];
                const float z = L.z[lane];
                const float w = t > 0.5f ? pos_w : 1.0f;
                const float sg = 1.0f / (1.0f + expf(-z));
                dz = (sg - t) * w;
                // softplus(z) - t*z, computed stably
                const float sp = z > 0.0f ? z + log1pf(expf(-z)) : log1pf(expf(z));
                lsum += w * (sp - t * z);
            }
            L.z[lane] = dz;
        }
        __syncthreads();
        // dw3, db3, dH2 (in place of H2), db2 : lane = column
        {
            const int c = lane;
            float s3 = 0.0f, s2 = 0.0f;
            for (int r = 0; r < kRows; r++) {
                const float h = L.H2[r * kGateHid + c];
                const float dz = L.z[r];
                s3 += dz * h;
                const float d = h > 0.0f ? dz * L.w3[c] : 0.0f;
                L.H2[r * kGateHid + c] = d;
                s2 += d;
            }
            gw3 += s3;
            gb2 += s2;
            if (lane == 0)
                for (int r = 0; r < kRows; r++) gb3 += L.z[r];
        }
        __syncthreads();
        // dW2 += H1^T dH2
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            const bf16x8 a = frag_at(L.H1 + mt * 16, kGateHid, lane);
#pragma unroll
            for (int nt = 0; nt < 4; nt++) gW2[mt][nt] = mfma(a, frag_b(L.H2 + nt * 16, kGateHid, lane), gW2[mt][nt]);
        }
        // dH1 = (dH2 W2^T) * [H1 > 0]
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            const bf16x8 a0 = frag_a(L.H2 + mt * 16 * kGateHid, kGateHid, lane);
            const bf16x8 a1 = frag_a(L.H2 + mt * 16 * kGateHid + 32, kGateHid, lane);
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mfma(a0, frag_bt(L.W2 + nt * 16 * kGateHid, kGateHid, lane), acc);
                acc = mfma(a1, frag_bt(L.W2 + nt * 16 * kGateHid + 32, kGateHid, lane), acc);
                const int col = nt * 16 + (lane & 15);
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int row = mt * 16 + 4 * (lane >> 4) + j;
                    L.D1[row * kGateHid + col] = L.H1[row * kGateHid + col] > 0.0f ? acc[j] : 0.0f;
                }
            }
        }
        __syncthreads();
        {
            float s1 = 0.0f;
            for (int r = 0; r < kRows; r++) s1 += L.D1[r * kGateHid + lane];
            gb1 += s1;
        }
        // dW1 += X^T dH1
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            const bf16x8 a = frag_at(L.X + mt * 16, kXld, lane);
#pragma unroll
            for (int nt = 0; nt < 4; nt++) gW1[mt][nt] = mfma(a, frag_b(L.D1 + nt * 16, kGateHid, lane), gW1[mt][nt]);
        }
        __syncthreads();
    }
    // flush
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int k = mt * 16 + 4 * (lane >> 4) + j, c = nt * 16 + (lane & 15);
                atomicAdd(&grad[kGateOffW2 + k * kGateHid + c], gW2[mt][nt][j]);
            }
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int f = mt * 16 + 4 * (lane >> 4) + j, c = nt * 16 + (lane & 15);
                if (f < kGateIn) atomicAdd(&grad[kGateOffW1 + f * kGateHid + c], gW1[mt][nt][j]);
            }
    atomicAdd(&grad[kGateOffB1 + lane], gb1);
    atomicAdd(&grad[kGateOffB2 + lane], gb2);
    atomicAdd(&grad[kGateOffW3 + lane], gw3);
    if (lane == 0) atomicAdd(&grad[kGateOffB3], gb3);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) lsum += __shfl_xor(lsum, d, 64);
    if (lane == 0) atomicAdd(loss, lsum);
}

__global__ __launch_bounds__(256) void k_gate_adam(float* __restrict__ P, const float* __restrict__ G,
                                                   float* __restrict__ m, float* __restrict__ v, float lr,
                                                   float b1, float b2, float eps, float c1, float c2, float gs) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kGateParams) return;
    const float g = G[i] * gs;
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace

hipError_t launch_gate_forward(const float* X, int n, const float* params, float* logits, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int blocks = (n + 16 * kFwdWaves - 1) / (16 * kFwdWaves);
    if (blocks > 4096) blocks = 4096;
    k_gate_forward<<<blocks, kFwdBlock, 0, s>>>(X, n, nullptr, params, logits);
    return hipGetLastError();
}

hipError_t launch_gate_forward_dev(const float* X, const unsigned* n_dev, size_t max_n, const float* params,
                                   float* logits, hipStream_t s) {
    if (max_n == 0) return hipSuccess;
    size_t blocks = (max_n + 16 * kFwdWaves - 1) / (16 * kFwdWaves);
    if (blocks > 4096) blocks = 4096;
    k_gate_forward<<<(unsigned)blocks, kFwdBlock, 0, s>>>(X, 0, n_dev, params, logits);
    return hipGetLastError();
}

hipError_t launch_gate_grad(const float* X, const float* y, const int* idx, int n, const float* params,
                            float* grad, float* loss, float pos_w, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int blocks = (n + kRows - 1) / kRows;
    if (blocks > 2048) blocks = 2048;
    k_gate_grad<<<blocks, 64, 0, s>>>(X, y, idx, n, params, grad, loss, pos_w);
    return hipGetLastError();
}

hipError_t launch_gate_adam(float* params, const float* grad, float* m, float* v, float lr, float beta1,
                            float beta2, float eps, int step, float gscale, hipStream_t s) {
    const float c1 = 1.0f - powf(beta1, (float)step);
    const float c2 = 1.0f - powf(beta2, (float)step);
    k_gate_adam<<<(kGateParams + 255) / 256, 256, 0, s>>>(params, grad, m, v, lr, beta1, beta2, eps, c1, c2, gscale);
    return hipGetLastError();
}

}  // namespace tlr
