// Shadow reconstruction (TreeGather::reconstructDofMotionShadows,
// Reconstruction.cpp:355-433, FilterTask::process2 :722-833): the camera's
// output samples, each one's shadow "ray" into the light's buffer, shading.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tlr_math.hpp"

namespace tlr {

// Row-major 4x4 matrices (m[4*i+j] = row i, column j) and camera data, as
// FilterTask::initShadow receives them (Reconstruction.cpp:404) plus the two
// matrix products lightSpaceDot forms first.
struct ShadowMats {
    float c2l[16];              // ws * cameraToShadow * invws * cameraProjectedZfromW
    float zfromw_ndc[16];       // invws * cameraProjectedZfromW
    float inv_projection[16];   // invCameraProjection
    float shadow2_inv_proj[16]; // cameraToShadow2 * invCameraProjection
    float shadow2_inv_t[16];    // cameraToShadow2.inverted().transposed()
    float world_inv_t[16];      // cameraToWorld.inverted().transposed()
    float cam_coc[2], light_coc[2];
    float cam_far;
    int cw, ch;                 // camera buffer size (the light's is the same)
};

// Matrix * Vector summed from zero: src/framework/base/Math.hpp:989-1000
TLR_HD void mat4_vec(const float* a, const float v[4], float out[4]) {
    float r[4];
    for (int i = 0; i < 4; i++) {
        float rr = 0.0f;
        for (int j = 0; j < 4; j++) rr += a[4 * i + j] * v[j];
        r[i] = rr;
    }
    for (int i = 0; i < 4; i++) out[i] = r[i];
}

// normalized(): v * (1 * rcp(length())) (Math.hpp:139-141)
TLR_HD void normalize3(const float v[3], float out[3]) {
    float l2 = 0.0f;
    for (int k = 0; k < 3; k++) l2 += v[k] * v[k];
    const float f = 1.0f * rcp(sqrtf(l2));
    for (int k = 0; k < 3; k++) out[k] = v[k] * f;
}

// viewSpaceNormalFromWGradients: common/Util.cpp:172-198 (p = clip x, y, w)
TLR_HD void normal_from_w_gradients(const float p[3], float dwdu, float dwdv, int vw, int vh, const float* inv_proj,
                                    float n[3]) {
    const float x = p[0], y = p[1], w = p[2];
    const float u = x / w, v = y / w;
    dwdu *= (float)vw / 2.0f;
    dwdv *= (float)vh / 2.0f;
    float b1[4] = {w + u * dwdu, v * dwdu, dwdu, 0.0f};
    float b2[4] = {u * dwdv, w + v * dwdv, dwdv, 0.0f};
    float ip[16];
    for (int k = 0; k < 8; k++) ip[k] = inv_proj[k];
    ip[8] = 0.0f; ip[9] = 0.0f; ip[10] = -1.0f; ip[11] = 0.0f;
    ip[12] = 0.0f; ip[13] = 0.0f; ip[14] = 0.0f; ip[15] = 1.0f;
    mat4_vec(ip, b1, b1);
    mat4_vec(ip, b2, b2);
    const float c[3] = {b1[1] * b2[2] - b1[2] * b2[1], b1[2] * b2[0] - b1[0] * b2[2], b1[0] * b2[1] - b1[1] * b2[0]};
    normalize3(c, n);
}

// lightSpaceDot: common/Util.cpp:202-224; n receives the view-space normal
TLR_HD float light_space_dot(const float xyw[4], float wgx, float wgy, const ShadowMats& S, float n[3]) {
    float xyzw[4];
    mat4_vec(S.zfromw_ndc, xyw, xyzw);
    const float p3[3] = {xyzw[0], xyzw[1], xyzw[3]};
    normal_from_w_gradients(p3, wgx, wgy, S.cw, S.ch, S.inv_projection, n);
    const float n4[4] = {n[0], n[1], n[2], 0.0f};
    float nl4[4], pl[4];
    mat4_vec(S.shadow2_inv_t, n4, nl4);
    mat4_vec(S.shadow2_inv_proj, xyzw, pl);
    const float c = rcp(pl[3]);
    const float plight[3] = {pl[0] * c, pl[1] * c, pl[2] * c};
    const float nl[3] = {nl4[0], nl4[1], nl4[2]};
    float a[3], b[3];
    normalize3(plight, a);
    normalize3(nl, b);
    float d = 0.0f;
    for (int k = 0; k < 3; k++) d += a[k] * b[k];
    return -d;
}

struct ShadowRays {
    int W, H, M;
    unsigned nq;
    const float* pattern;      // 64*M templates (x, y, u, v, t)
    const float* light_uv;     // M pairs (sobol(3, i+1), sobol(4, i+1))
    const float4* qcol;        // camera pass colours
    const float4* qext;        // camera pass extended results (2 per query)
    ShadowMats S;
    float4* lq;                // light queries, 2 per output sample
    float4* xyw;               // (x w, y w, 1, w) at the (u,v) centre
    int* flag;                 // 0 no light query, 1 light query, 2 no surface (red pixel)
    unsigned* list;            // output samples with a light query
    unsigned* count;
};

struct ShadowShade {
    int W, H, M;
    const float4* qcol;
    const float4* qext;
    const float4* scol;        // light pass colours (flag 1)
    const float4* xyw;
    const int* flag;
    ShadowMats S;
    float* out;                // W*H*4
};

hipError_t launch_shadow_rays(const ShadowRays& r, hipStream_t s);
hipError_t launch_shadow_shade(const ShadowShade& sh, hipStream_t s);
// Tree-order extras (wgx, wgy, density, 0) from the input-order ones
// (in_ext null: 0, 0, 1, 0 for every sample).
hipError_t launch_gather_ext(const int* cand, int nb, const float4* in_ext, float4* out, hipStream_t s);

}  // namespace tlr
