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
    Vec2f getSampleWG(int x, int y, int i) const { const float* r = rec(x, y, i); return Vec2f(r[14], r[15]); }
    void setSampleColor(int x, int y, int i, const Vec4f& c) {
        float* r = &rec_[16 * (size_t)getIndex(x, y, i)];
        r[7] = c.x; r[8] = c.y; r[9] = c.z; r[10] = c.w;
    }
    // Extra per-sample float channels by name (SampleBuffer.hpp:104-112; the
    // shadow path reads "LIGHTSPACE_DENSITY").  getSampleExtra of channel -1
    // is 0, as in the reference.
    int reserveChannel(const std::string& name) {
        int id = getChannelID(name);
        if (id >= 0) return id;
        names_.push_back(name);
        channels_.emplace_back(rec_.size() / 16, 0.0f);
        return (int)names_.size() - 1;
    }
    int getChannelID(const std::string& name) const {
        for (size_t k = 0; k < names_.size(); k++)
            if (names_[k] == name) return (int)k;
        return -1;
    }
    template <class T>
    T getSampleExtra(int cid, int x, int y, int i) const {
        return cid < 0 ? T(0) : T(channels_[cid][getIndex(x, y, i)]);
    }
    template <class T>
    void setSampleExtra(int cid, int x, int y, int i, T v) {
        channels_[cid][getIndex(x, y, i)] = (float)v;
    }
    const std::vector<float>* channel(const std::string& name) const {
        const int id = getChannelID(name);
        return id < 0 ? nullptr : &channels_[id];
    }
    void serialize(const char* filename, bool separateHeader = false, bool binary = false) const {
        tlr_status s = tlr_save_sample_file(filename, &h_, rec_.data(), binary, separateHeader);
        if (s != TLR_OK) throw ReconstructionError(s, std::string("cannot write ") + filename);
    }
    const tlr_header& header() const { return h_; }
    const std::vector<float>& records() const { return rec_; }

  private:
    const float* rec(int x, int y, int i) const { return &rec_[16 * (size_t)getIndex(x, y, i)]; }
    std::vector<std::string> names_;
    std::vector<std::vector<float>> channels_;
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
    float focusDistance = 1;
    // The camera data the shadow path reads from the reference's
    // CameraControls (getCameraToWorld, getWorldToCamera, getFar), row-major.
    float cameraToWorld[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float worldToCamera[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float farPlane = 100.0f;

    // getCocCoeffs (CameraParams.hpp:97-118) at the focus distance
    void getCocCoeffs(float out[2]) const { cocCoeffsAt(focusDistance, out); }

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

    // Shadow-only reconstruction (Reconstruction.cpp:302-347): every sample
    // of qbuf is a light-space query (its xy, uv, t, w and LIGHTSPACE_DENSITY
    // channel -- 0 without the channel, as the reference reads it); its
    // colour is replaced by the result.  debugImage is not filled.
    void reconstructShadows(UVTSampleBuffer* qbuf, Image* debugImage = nullptr) {
        (void)debugImage;
        uploadShadowInputs();
        const int W = qbuf->getWidth(), H = qbuf->getHeight(), S = qbuf->getNumSamples();
        const int cid = qbuf->getChannelID("LIGHTSPACE_DENSITY");
        std::vector<float> q((size_t)W * H * S * 8);
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++)
                for (int i = 0; i < S; i++) {
                    float* o = &q[8 * (size_t)qbuf->getIndex(x, y, i)];
                    const Vec2f xy = qbuf->getSampleXY(x, y, i), uv = qbuf->getSampleUV(x, y, i);
                    o[0] = xy.x; o[1] = xy.y; o[2] = uv.x; o[3] = uv.y;
                    o[4] = qbuf->getSampleT(x, y, i);
                    o[5] = qbuf->getSampleW(x, y, i);
                    o[6] = qbuf->getSampleExtra<float>(cid, x, y, i);
                    o[7] = 0.0f;
                }
        std::vector<float> rgba((size_t)W * H * S * 4);
        check(tlr_reconstruct_queries(ctx_, nullptr, q.data(), (int64_t)W * H * S, 1, rgba.data(), nullptr, &stats_),
              "shadow queries");
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++)
                for (int i = 0; i < S; i++) {
                    const float* c = &rgba[4 * (size_t)qbuf->getIndex(x, y, i)];
                    qbuf->setSampleColor(x, y, i, Vec4f(c[0], c[1], c[2], c[3]));
                }
    }

    // Depth of field and motion with shadows from shadowTG, the light's
    // TreeGather of the same size (Reconstruction.cpp:355-433); the matrices
    // come from both CameraParams.  The reference also writes a screenshot
    // file; this does not.
    void reconstructDofMotionShadows(Image& image, const TreeGather& shadowTG) {
        const int W = sbuf_->getWidth(), H = sbuf_->getHeight();
        if (image.getSize().x < W || image.getSize().y < H)
            throw ReconstructionError(TLR_E_IMAGE_TOO_SMALL,
                                      "TreeGather::reconstructDofMotionShadows image smaller than < sample buffer");
        uploadShadowInputs();
        const_cast<TreeGather&>(shadowTG).uploadShadowInputs();
        const CameraParams& c = *params_;
        const CameraParams& l = *shadowTG.params_;
        float cc[2], lc[2];
        c.getCocCoeffs(cc);
        l.getCocCoeffs(lc);
        tlr_shadow_params sp;
        check(tlr_shadow_setup(c.projection, c.cameraToWorld, l.projection, l.worldToCamera, W, H, cc, lc,
                               c.farPlane, &sp),
              "shadow setup");
        tlr_recon_params p{};
        p.out_spp = out_spp_;
        p.override_uvt = c.overrideUVT != Vec3f(FLT_MAX);
        p.uvt[0] = c.overrideUVT.x; p.uvt[1] = c.overrideUVT.y; p.uvt[2] = c.overrideUVT.z;
        p.lens_gaussian = c.lensFilter == LENS_GAUSSIAN;
        p.time_gaussian = c.timeFilter == TIME_GAUSSIAN;
        std::vector<float> rgba((size_t)W * H * 4);
        check(tlr_reconstruct_dof_motion_shadows(ctx_, shadowTG.ctx_, &p, &sp, rgba.data(), 0, &stats_),
              "dof motion shadows");
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                const float* q = &rgba[4 * ((size_t)y * W + x)];
                image.setVec4f(Vec2i(x, y), Vec4f(q[0], q[1], q[2], q[3]));
            }
    }

    const tlr_stats& lastStats() const { return stats_; }

  private:
    // w-gradients (records' floats 14, 15) and the LIGHTSPACE_DENSITY
    // channel (1 without it) of the input buffer, once
    void uploadShadowInputs() {
        if (shadow_inputs_) return;
        const std::vector<float>& r = sbuf_->records();
        const size_t n = r.size() / 16;
        std::vector<float> wg(2 * n);
        for (size_t k = 0; k < n; k++) {
            wg[2 * k] = r[16 * k + 14];
            wg[2 * k + 1] = r[16 * k + 15];
        }
        const std::vector<float>* d = sbuf_->channel("LIGHTSPACE_DENSITY");
        check(tlr_set_shadow_inputs(ctx_, wg.data(), d ? d->data() : nullptr, (int64_t)n), "shadow inputs");
        shadow_inputs_ = true;
    }

    void check(tlr_status s, const char* what) {
        if (s != TLR_OK)
            throw ReconstructionError(s, std::string(what) + ": " + (ctx_ ? tlr_last_error(ctx_) : "no device"));
    }
    const UVTSampleBuffer* sbuf_;
    const CameraParams* params_;
    tlr_ctx* ctx_ = nullptr;
    int out_spp_ = NUM_OUTPUT_SAMPLES;
    bool shadow_inputs_ = false;
    tlr_stats stats_{};
};

}  // namespace FW
