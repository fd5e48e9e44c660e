// C++ facade with the reference's reconstruction_lib interface, implemented
// on the libtlr C ABI (include/tlr.h).  A caller of the reference
//
//     FW::UVTSampleBuffer sbuf("scene.bin");
//     FW::CameraParams params;
//     FW::TreeGather tg(sbuf, params);
//     FW::Image image(sbuf.getWidth(), sbuf.getHeight());
//     tg.reconstructDofMotion(image);
//
// compiles against this header unchanged (src/reconstruction_lib/
// reconstruction/Reconstruction.hpp:42-48, common/SampleBuffer.hpp:138-183,
// common/CameraParams.hpp:92-107).  Differences: fatal conditions throw
// FW::ReconstructionError instead of FW::fail's print-and-exit
// (src/framework/base/Defs.cpp:304-331); the GPU is always used
// (params.enableCuda is ignored); overrideRefocusDistance needs the camera's
// projection, windowWidth and apertureRadius set (CameraParams below).
#pragma once
#include <cfloat>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "tlr.h"

namespace FW {

struct Vec2i {
    int x, y;
    Vec2i(int a = 0, int b = 0) : x(a), y(b) {}
};
struct Vec2f {
    float x, y;
    Vec2f(float a = 0, float b = 0) : x(a), y(b) {}
    float operator[](int i) const { return i ? y : x; }
};
struct Vec3f {
    float x, y, z;
    Vec3f(float a = 0) : x(a), y(a), z(a) {}
    Vec3f(float a, float b, float c) : x(a), y(b), z(c) {}
    bool operator!=(const Vec3f& o) const { return x != o.x || y != o.y || z != o.z; }
};
struct Vec4f {
    float x, y, z, w;
    Vec4f(float a = 0, float b = 0, float c = 0, float d = 0) : x(a), y(b), z(c), w(d) {}
};

class ReconstructionError : public std::runtime_error {
  public:
    ReconstructionError(tlr_status s, const std::string& m) : std::runtime_error(m), status(s) {}
    tlr_status status;
};

enum LensFilter { LENS_BOX = 0, LENS_GAUSSIAN = 1 };
enum TimeFilter { TIME_BOX = 0, TIME_GAUSSIAN = 1 };

// RGBA float image (the RGBA_Vec4f Image the reference writes into).
class Image {
  public:
    Image(int w, int h) : size_(w, h), px_((size_t)w * h * 4, 0.0f) {}
    Vec2i getSize() const { return size_; }
    Vec4f getVec4f(const Vec2i& p) const {
        const float* q = &px_[4 * ((size_t)p.y * size_.x + p.x)];
        return Vec4f(q[0], q[1], q[2], q[3]);
    }
    void setVec4f(const Vec2i& p, const Vec4f& c) {
        float* q = &px_[4 * ((size_t)p.y * size_.x + p.x)];
        q[0] = c.x; q[1] = c.y; q[2] = c.z; q[3] = c.w;
    }
    float* getMutablePtr() { return px_.data(); }
    const float* getPtr() const { return px_.data(); }

  private:
    Vec2i size_;
    std::vector<float> px_;
};

// Sample buffer: v1.3 records (x,y,z/w,w,u,v,t,r,g,b,a,mv_x,mv_y,mv_w,dwdx,dwdy).
class UVTSampleBuffer {
  public:
    explicit UVTSampleBuffer(const char* filename) {
        float* rec = nullptr;
        int64_t n = 0;
        tlr_status s = tlr_load_sample_file(filename, &h_, &rec, &n);
        if (s != TLR_OK) throw ReconstructionError(s, std::string("File not found or unreadable: ") + filename);
        rec_.assign(rec, rec + n * 16);
        tlr_free(rec);
    }
    UVTSampleBuffer(int w, int h, int spp, const Vec2f& coc, std::vector<float> records) : rec_(std::move(records)) {
        h_.width = w; h_.height = h; h_.spp = spp;
        h_.coc_c0 = coc.x; h_.coc_c1 = coc.y;
    }
    int getWidth() const { return h_.width; }
    int getHeight() const { return h_.height; }
    int getNumSamples() const { return h_.spp; }
    int getNumSamples(int, int) const { return h_.spp; }
    Vec2f getCocCoeffs() const { return Vec2f(h_.coc_c0, h_.coc_c1); }
    void setCocCoeffs(const Vec2f& c) { h_.coc_c0 = c.x; h_.coc_c1 = c.y; }
    int getIndex(int x, int y, int i) const { return (y * h_.width + x) * h_.spp + i; }
    Vec2f getSampleXY(int x, int y, int i) const { const float* r = rec(x, y, i); return Vec2f(r[0], r[1]); }
    Vec2f getSampleUV(int x, int y, int i) const { const float* r = rec(x, y, i); return Vec2f(r[4], r[5]); }
    float getSampleT(int x, int y, int i) const { return rec(x, y, i)[6]; }
    float getSampleW(int x, int y, int i) const { return rec(x, y, i)[3]; }
    Vec4f getSampleColor(int x, int y, int i) const { const float* r = rec(x, y, i); return Vec4f(r[7], r[8], r[9], 1); }
    Vec3f getSampleMV(int x, int y, int i) const { const float* r = rec(x, y, i); return Vec3f(r[11], r[12], r[13]); }
    void serialize(const char* filename, bool separateHeader = false, bool binary = false) const {
        tlr_status s = tlr_save_sample_file(filename, &h_, rec_.data(), binary, separateHeader);
        if (s != TLR_OK) throw ReconstructionError(s, std::string("cannot write ") + filename);
    }
    const tlr_header& header() const { return h_; }
    const std::vector<float>& records() const { return rec_; }

  private:
    const float* rec(int x, int y, int i) const { return &rec_[16 * (size_t)getIndex(x, y, i)]; }
    tlr_header h_{};
    std::vector<float> rec_;
};

// The CameraParams fields the reconstruction reads.  projection (row-major,
// projection[4*row+col]), windowWidth and apertureRadius are needed only with
// overrideRefocusDistance.
class CameraParams {
  public:
    LensFilter lensFilter = LENS_BOX;
    TimeFilter timeFilter = TIME_BOX;
    Vec3f overrideUVT = Vec3f(FLT_MAX);
    float overrideRefocusDistance = FLT_MAX;
    bool enableCuda = true;
    float projection[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    int windowWidth = 0;
    float apertureRadius = 0;

    // getCocCoeffs at focus distance f (CameraParams.hpp:97-118): m =
    // projection * getDofMatrix((0.5, 0)) - projection, products summed over k
    // from 0 as the framework's Mat4f does (Math.hpp:1024-1035); coefficients
    // (-m03, m02) * window width.
    void cocCoeffsAt(float f, float out[2]) const {
        float dof[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        dof[0 * 4 + 2] = (0.5f * apertureRadius) / f;
        dof[1 * 4 + 2] = (0.0f * apertureRadius) / f;
        dof[0 * 4 + 3] = 0.5f * apertureRadius;
        dof[1 * 4 + 3] = 0.0f * apertureRadius;
        float m[16];
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                float rr = 0.0f;
                for (int k = 0; k < 4; k++) rr += projection[4 * i + k] * dof[4 * k + j];
                m[4 * i + j] = rr - projection[4 * i + j];
            }
        out[0] = -m[0 * 4 + 3] * (float)windowWidth;
        out[1] = m[0 * 4 + 2] * (float)windowWidth;
    }
};

class TreeGather {
  public:
    enum { NUM_OUTPUT_SAMPLES = 128, NUM_OUTPUT_SAMPLES_OVERRIDE = 1, NUM_PATTERNS = 64, MAX_LEAF_SIZE = 48 };

    TreeGather(const UVTSampleBuffer& sbuf, const CameraParams& params, float apertureAdjust = 1.f,
               float focalDistanceAdjust = 1.f, int device = 0)
        : sbuf_(&sbuf), params_(&params) {
        check(tlr_create(&ctx_, device), "tlr_create");
        const tlr_header& h = sbuf.header();
        check(tlr_set_samples(ctx_, &h, sbuf.records().data(), (int64_t)h.width * h.height * h.spp), "set samples");
        if (params.overrideRefocusDistance != FLT_MAX) {
            // Reconstruction.cpp:88-95: the refocused camera's coefficients
            float cc[2];
            params.cocCoeffsAt(params.overrideRefocusDistance, cc);
            check(tlr_build_ex(ctx_, apertureAdjust, focalDistanceAdjust, cc), "build");
        } else {
            check(tlr_build(ctx_, apertureAdjust, focalDistanceAdjust), "build");
        }
    }
    ~TreeGather() { tlr_destroy(ctx_); }
    TreeGather(const TreeGather&) = delete;
    TreeGather& operator=(const TreeGather&) = delete;

    // The reference hard-wires 128 output samples per pixel
    // (Reconstruction.hpp:67); this sets the count for later calls.
    void setOutputSpp(int spp) { out_spp_ = spp; }

    void reconstructDofMotion(Image& image, Image* debugImage = nullptr) {
        if (image.getSize().x < sbuf_->getWidth() || image.getSize().y < sbuf_->getHeight())
            throw ReconstructionError(TLR_E_IMAGE_TOO_SMALL,
                                      "TreeGather::reconstructDofMotion image smaller than < sample buffer");
        tlr_recon_params p{};
        p.out_spp = out_spp_;
        const CameraParams& c = *params_;
        p.override_uvt = c.overrideUVT != Vec3f(FLT_MAX);
        p.uvt[0] = c.overrideUVT.x; p.uvt[1] = c.overrideUVT.y; p.uvt[2] = c.overrideUVT.z;
        p.lens_gaussian = c.lensFilter == LENS_GAUSSIAN;
        p.time_gaussian = c.timeFilter == TIME_GAUSSIAN;
        p.debug_image = debugImage != nullptr;
        const int W = sbuf_->getWidth(), H = sbuf_->getHeight();
        std::vector<float> rgba((size_t)W * H * 4);
        check(tlr_reconstruct(ctx_, &p, rgba.data(), 0, &stats_), "reconstruct");
        if (debugImage) {
            // (#surfaces, #samples within 1R) of each pixel's last output
            // sample (ReconstructionOur.cpp:312-314, Reconstruction.cpp:276)
            std::vector<float> dbg((size_t)W * H * 4);
            check(tlr_copy_debug_image(ctx_, dbg.data()), "debug image");
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++) {
                    const float* q = &dbg[4 * ((size_t)y * W + x)];
                    debugImage->setVec4f(Vec2i(x, y), Vec4f(q[0], q[1], q[2], q[3]));
                }
        }
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                const float* q = &rgba[4 * ((size_t)y * W + x)];
                image.setVec4f(Vec2i(x, y), Vec4f(q[0], q[1], q[2], q[3]));
            }
    }

    const tlr_stats& lastStats() const { return stats_; }

  private:
    void check(tlr_status s, const char* what) {
        if (s != TLR_OK)
            throw ReconstructionError(s, std::string(what) + ": " + (ctx_ ? tlr_last_error(ctx_) : "no device"));
    }
    const UVTSampleBuffer* sbuf_;
    const CameraParams* params_;
    tlr_ctx* ctx_ = nullptr;
    int out_spp_ = NUM_OUTPUT_SAMPLES;
    tlr_stats stats_{};
};

}  // namespace FW
