// Shadow rays and shading of reconstructDofMotionShadows
// (FilterTask::process2, Reconstruction.cpp:722-833).
#include "shadow.hpp"

namespace tlr {
namespace {

// RANROT-A (src/framework/base/Random.cpp:52-92) seeded per pixel: the two
// first draws of Random(hashBits(x, y)).getF32() (Random.hpp:57) are the
// light-sample offset of a pixel, taken as x then y.
__device__ __forceinline__ void pixel_uv_offset(uint32_t seed, float& ox, float& oy) {
    uint32_t buf[11];
    if (seed == 0) seed--;
    for (int i = 0; i < 11; i++) {
        seed ^= seed << 13;
        seed ^= seed >> 17;
        seed ^= seed << 5;
        buf[i] = seed;
    }
    int p1 = 0, p2 = 7;
    uint32_t x = 0;
    for (int i = 0; i < 13; i++) {
        x = buf[p1] + buf[p2];
        x = (x << 13) | (x >> 19);
        buf[p1] = x;
        p1 = (p1 == 0) ? 10 : p1 - 1;
        p2 = (p2 == 0) ? 10 : p2 - 1;
        if (i == 11) ox = (float)x * (1.0f / 4294967296.0f);
    }
    oy = (float)x * (1.0f / 4294967296.0f);
}

constexpr int kRayBlock = 256;

// One output sample per lane: the camera result's point back at the (u,v)
// centre (Sample::reprojectToUVCenter, Reconstruction.hpp:87-91), into the
// light's window (c2l), and the light-space query with its position on the
// square light (Reconstruction.cpp:766-793).
__global__ __launch_bounds__(kRayBlock) void k_shadow_rays(ShadowRays R) {
    const unsigned q = blockIdx.x * kRayBlock + threadIdx.x;
    bool query = false;
    if (q < R.nq) {
        const unsigned pix = q / (unsigned)R.M;
        const int i = (int)(q - pix * (unsigned)R.M);
        const int x = (int)(pix % (unsigned)R.W), y = (int)(pix / (unsigned)R.W);
        const float4 c = R.qcol[q];
        int f = 0;
        if (c.x == 0.0f && c.y == 0.0f && c.z == 0.0f && c.w == 0.0f) {
            f = 2;
        } else {
            const float4 e0 = R.qext[2 * (size_t)q];
            const float w = e0.z;
            const float camFarW = R.S.cam_far - 1e-2f;
            if (w <= camFarW) {
                const int pat = (int)(hash_bits((uint32_t)x, (uint32_t)y) % 64u);
                const float* tp = R.pattern + 5 * ((size_t)pat * R.M + i);
                const float coc = R.S.cam_coc[0] / w + R.S.cam_coc[1];
                const float cx = e0.x - coc * tp[2];
                const float cy = e0.y - coc * tp[3];
                const float v4[4] = {cx * w, cy * w, 1.0f, w};
                R.xyw[q] = make_float4(v4[0], v4[1], v4[2], v4[3]);
                float lp[4];
                mat4_vec(R.S.c2l, v4, lp);
                const float lc = rcp(lp[3]);
                const float lx = lp[0] * lc, ly = lp[1] * lc, lw = lp[3];
                if (lx >= 0 && ly >= 0 && lx <= (float)R.S.cw && ly <= (float)R.S.ch && lw > 0) {
                    float ox, oy;
                    pixel_uv_offset(hash_bits((uint32_t)x, (uint32_t)y), ox, oy);
                    float uu = R.light_uv[2 * i] + ox, vv = R.light_uv[2 * i + 1] + oy;
                    if (uu >= 1.f) uu -= 1.f;
                    if (vv >= 1.f) vv -= 1.f;
                    const float lu = 2.f * uu - 1.f, lv = 2.f * vv - 1.f;
                    const float lcoc = R.S.light_coc[0] / lw + R.S.light_coc[1];
                    R.lq[2 * (size_t)q] = make_float4(lx + lu * lcoc, ly + lv * lcoc, lu, lv);
                    R.lq[2 * (size_t)q + 1] = make_float4(tp[4], lw, R.qext[2 * (size_t)q].w, 0.0f);
                    f = 1;
                    query = true;
                }
            }
        }
        R.flag[q] = f;
    }
    // append the light queries, one atomic per block (atomics on the one
    // counter serialise): the waves' counts in LDS, the block's base from one
    // atomic, each wave's base past the earlier waves' entries
    __shared__ unsigned wcount[kRayBlock / 64];
    __shared__ unsigned bbase;
    const uint64_t m = __ballot(query);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wcount[w] = (unsigned)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = 0;
        for (int k = 0; k < kRayBlock / 64; k++) t += wcount[k];
        bbase = t ? atomicAdd(R.count, t) : 0u;
    }
    __syncthreads();
    if (query) {
        unsigned base = bbase;
        for (int k = 0; k < w; k++) base += wcount[k];
        R.list[base + (unsigned)__popcll(m & ((1ull << lane) - 1ull))] = q;
    }
}

constexpr int kShadeBlock = 128;

// One output sample's shaded colour: the light pass's colour scaled by the
// Lambert term (light-space dot of the normal from the camera result's
// w-gradients), plus the ambient term from the world-space normal
// (Reconstruction.cpp:795-833); f: its flag (2: no surface).
__device__ __forceinline__ void shade_sample(const ShadowShade& D, size_t q, int& f, float col[4]) {
    const float LIGHTING_SCALE = 1.5f, AMBIENT_SCALE = 0.5f;  // Reconstruction.cpp:34-35
    f = D.flag[q];
    col[0] = col[1] = col[2] = col[3] = 0.0f;
    if (f == 2) return;
    const float4 c = D.qcol[q];
    col[0] = c.x; col[1] = c.y; col[2] = c.z; col[3] = c.w;
    if (f == 1) {
        const float4 sc = D.scol[q];
        col[0] = sc.x; col[1] = sc.y; col[2] = sc.z; col[3] = sc.w;
        const float4 e1 = D.qext[2 * q + 1];
        const float4 v = D.xyw[q];
        const float v4[4] = {v.x, v.y, v.z, v.w};
        float n[3];
        const float ldotn = light_space_dot(v4, e1.x, e1.y, D.S, n);
        const float n4[4] = {n[0], n[1], n[2], 0.0f};
        float wn[4];
        mat4_vec(D.S.world_inv_t, n4, wn);
        const float ambient = wn[1] * 0.5f + 0.5f;
        const float cosine = ldotn > 0.0f ? ldotn : 0.0f;
        const float s3 = LIGHTING_SCALE * cosine, s4 = LIGHTING_SCALE * 1.0f;
        col[0] *= s3; col[1] *= s3; col[2] *= s3; col[3] *= s4;
        const float a = ambient * AMBIENT_SCALE;
        col[0] += a * c.x; col[1] += a * c.y; col[2] += a * c.z; col[3] += a * c.w;
        col[3] = 1.0f;
    }
}

// Per pixel, its output samples' shaded colours summed in order; a sample
// with no surface makes the pixel red (Reconstruction.cpp:762-766).  When M
// divides 64, one thread per output sample: a wave shades 64 consecutive
// output samples (64 / M pixels) with coalesced loads and every lane sums its
// pixel's samples, in order, from the lanes that hold them; otherwise one
// thread per pixel.
__global__ __launch_bounds__(kShadeBlock) void k_shadow_shade(ShadowShade D) {
    const int npix = D.W * D.H, M = D.M;
    float pc[4] = {0.f, 0.f, 0.f, 0.f};
    bool red = false;
    int p;
    if (M <= 64 && (64 % M) == 0) {
        const size_t k = (size_t)blockIdx.x * kShadeBlock + threadIdx.x;
        const bool in = k < (size_t)npix * (size_t)M;
        int f = 0;
        float col[4] = {0.f, 0.f, 0.f, 0.f};
        if (in) shade_sample(D, k, f, col);
        const int lane = threadIdx.x & 63, first = lane & ~(M - 1);
        for (int i = 0; i < M; i++) {
            const int fi = __shfl(f, first + i, 64);
            float ci[4];
            for (int c = 0; c < 4; c++) ci[c] = __shfl(col[c], first + i, 64);
            if (!red) {
                if (fi == 2) {
                    red = true;
                } else {
                    for (int c = 0; c < 4; c++) pc[c] += ci[c];
                }
            }
        }
        if (!in || lane != first) return;
        p = (int)(k / (size_t)M);
    } else {
        p = blockIdx.x * kShadeBlock + threadIdx.x;
        if (p >= npix) return;
        for (int i = 0; i < M && !red; i++) {
            int f;
            float col[4];
            shade_sample(D, (size_t)p * M + i, f, col);
            if (f == 2) {
                red = true;
            } else {
                for (int c = 0; c < 4; c++) pc[c] += col[c];
            }
        }
    }
    if (red || pc[3] == 0) {
        pc[0] = 1.f; pc[1] = 0.f; pc[2] = 0.f; pc[3] = 1.f;
    }
    const float rw = rcp(pc[3]);
    float4* o = reinterpret_cast<float4*>(D.out);
    o[p] = make_float4(pc[0] * rw, pc[1] * rw, pc[2] * rw, pc[3] * rw);
}

__global__ __launch_bounds__(256) void k_gather_ext(const int* __restrict__ cand, int nb,
                                                    const float4* __restrict__ in_ext, float4* __restrict__ out) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < nb) out[k] = in_ext ? in_ext[cand[k]] : make_float4(0.0f, 0.0f, 1.0f, 0.0f);
}

}  // namespace

hipError_t launch_shadow_rays(const ShadowRays& r, hipStream_t s) {
    if (r.nq == 0) return hipSuccess;
    k_shadow_rays<<<(r.nq + kRayBlock - 1) / kRayBlock, kRayBlock, 0, s>>>(r);
    return hipGetLastError();
}

hipError_t launch_shadow_shade(const ShadowShade& sh, hipStream_t s) {
    const int n = sh.W * sh.H;
    if (n == 0) return hipSuccess;
    const size_t threads = (sh.M <= 64 && (64 % sh.M) == 0) ? (size_t)n * sh.M : (size_t)n;
    k_shadow_shade<<<(unsigned)((threads + kShadeBlock - 1) / kShadeBlock), kShadeBlock, 0, s>>>(sh);
    return hipGetLastError();
}

hipError_t launch_gather_ext(const int* cand, int nb, const float4* in_ext, float4* out, hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    k_gather_ext<<<(nb + 255) / 256, 256, 0, s>>>(cand, nb, in_ext, out);
    return hipGetLastError();
}

}  // namespace tlr
