// The matrices TreeGather::reconstructDofMotionShadows builds from the camera
// and the light (Reconstruction.cpp:365-388), in FW's float arithmetic
// (src/framework/base/Math.hpp): row-major 4x4, products and matrix-vector
// sums started at zero, inverted() by cofactors.  Host only.
#include <cstring>

#include "../../include/tlr.h"

namespace {

struct Mat {
    float m[16];
    float& at(int i, int j) { return m[4 * i + j]; }
    float at(int i, int j) const { return m[4 * i + j]; }
};

Mat identity() {
    Mat r;
    for (int k = 0; k < 16; k++) r.m[k] = (k % 5 == 0) ? 1.0f : 0.0f;
    return r;
}

Mat load(const float* p) {
    Mat r;
    std::memcpy(r.m, p, sizeof(r.m));
    return r;
}

// Math.hpp:1024-1037
Mat mul(const Mat& a, const Mat& b) {
    Mat r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float acc = 0.0f;
            for (int k = 0; k < 4; k++) acc += a.at(i, k) * b.at(k, j);
            r.at(i, j) = acc;
        }
    return r;
}

// 3x3 determinant, Math.hpp:928-933
float det3(const float v[3][3]) {
    return v[0][0] * v[1][1] * v[2][2] - v[0][0] * v[1][2] * v[2][1] + v[1][0] * v[2][1] * v[0][2] -
           v[1][0] * v[2][2] * v[0][1] + v[2][0] * v[0][1] * v[1][2] - v[2][0] * v[0][2] * v[1][1];
}

// inverted(), Math.hpp:955-976: the cofactor matrix, d summed over every
// (i, j) (four times the determinant), scaled by rcp(d) * 4
Mat inverse(const Mat& a) {
    Mat r;
    float d = 0.0f, si = 1.0f;
    for (int i = 0; i < 4; i++) {
        float sj = si;
        for (int j = 0; j < 4; j++) {
            float sub[3][3];
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) sub[k][l] = a.at(k < j ? k : k + 1, l < i ? l : l + 1);
            const float dd = det3(sub) * sj;
            r.at(i, j) = dd;
            d += dd * a.at(j, i);
            sj = -sj;
        }
        si = -si;
    }
    const float rd = (d != 0.0f) ? 1.0f / d : 0.0f;
    for (int k = 0; k < 16; k++) r.m[k] = r.m[k] * rd * 4.0f;
    return r;
}

Mat transpose(const Mat& a) {
    Mat r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) r.at(i, j) = a.at(j, i);
    return r;
}

Mat scale(float x, float y, float z) {
    Mat r = identity();
    r.at(0, 0) = x;
    r.at(1, 1) = y;
    r.at(2, 2) = z;
    return r;
}

Mat translate(float x, float y, float z) {
    Mat r = identity();
    r.at(0, 3) = x;
    r.at(1, 3) = y;
    r.at(2, 3) = z;
    return r;
}

}  // namespace

extern "C" tlr_status tlr_shadow_setup(const float* cam_proj, const float* cam_to_world, const float* light_proj,
                                       const float* light_world_to_cam, int w, int h, const float* cam_coc,
                                       const float* light_coc, float cam_far, tlr_shadow_params* out) {
    if (!cam_proj || !cam_to_world || !light_proj || !light_world_to_cam || !cam_coc || !light_coc || !out ||
        w <= 0 || h <= 0)
        return TLR_E_INVALID;
    const Mat P = load(cam_proj), toWorld = load(cam_to_world), lightP = load(light_proj),
              toLight = load(light_world_to_cam);
    const Mat invP = inverse(P);
    const Mat cameraToShadow = mul(mul(mul(lightP, toLight), toWorld), invP);
    const Mat cameraToShadow2 = mul(toLight, toWorld);
    const Mat ws = mul(scale(0.5f * w, 0.5f * h, 0.5f), translate(1.0f, 1.0f, 1.0f));
    const Mat invws = mul(translate(-1.0f, -1.0f, -1.0f), scale(2.f / w, 2.f / h, 2.f));
    // (x w, y w, 1, w) -> (x w, y w, z w, w), z still in viewport units
    const Mat wsP = mul(ws, P);
    Mat zFromW = identity();
    zFromW.at(2, 2) = wsP.at(2, 3);
    zFromW.at(2, 3) = -wsP.at(2, 2);
    const Mat c2l = mul(mul(mul(ws, cameraToShadow), invws), zFromW);
    out->width = w;
    out->height = h;
    out->camera_coc[0] = cam_coc[0];
    out->camera_coc[1] = cam_coc[1];
    out->light_coc[0] = light_coc[0];
    out->light_coc[1] = light_coc[1];
    out->camera_far = cam_far;
    std::memcpy(out->c2l, c2l.m, 64);
    std::memcpy(out->zfromw_ndc, mul(invws, zFromW).m, 64);
    std::memcpy(out->inv_projection, invP.m, 64);
    std::memcpy(out->shadow2_inv_proj, mul(cameraToShadow2, invP).m, 64);
    std::memcpy(out->shadow2_inv_t, transpose(inverse(cameraToShadow2)).m, 64);
    std::memcpy(out->world_inv_t, transpose(inverse(toWorld)).m, 64);
    return TLR_OK;
}
