// libtlr host side: context, device memory, build and reconstruct pipelines
// behind the C ABI of include/tlr.h.
//
// TreeGather (Reconstruction.cpp:43-295) does its build eagerly in the
// constructor and filters in reconstructDofMotion; tlr_build and
// tlr_reconstruct are those two halves, running entirely on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <numeric>
#include <random>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tlr.h"
#include "build.hpp"
#include "gate.hpp"
#include "noalloc.hpp"
#include "pattern.hpp"
#include "prims.hpp"
#include "query.hpp"
#include "shadow.hpp"

using namespace tlr;

namespace {

constexpr int kMaxGridSide = 4100;   // the bucketing grid is clamped to 4096 (+2)

int bits_for(uint64_t n) {
    int b = 0;
    while (b < 32 && ((uint64_t)1 << b) < n) b++;
    return b;
}

// One buffer of a batch: its device-resident input and, after a build, the
// products the query kernels read plus its own output buffers.
struct BatchSlot {
    tlr_header h{};
    int N = 0;
    float* d_in = nullptr;   // 12 fields x N, like tlr_ctx::d_in
    size_t d_in_cap = 0;
    DeviceArena prod;
    Points pts{};
    LeafRec* leaves = nullptr;
    Bounds* leaf_b = nullptr;
    TravNode* trav = nullptr;
    TravPair* pairs = nullptr;
    float* pattern = nullptr;
    float4* qcol = nullptr;
    float* img = nullptr;
    int M = 0;
    size_t nq = 0;
    int gate = 0;            // this batch: the learned gate on
    QueryParams q{};
};

// Learned-gate buffers of a batch's chained form, shared by its gated items
// (they run one after another), and the per-item completion stamps.  Without
// `gate` every item runs ungated whatever its flag (the autotuner's timing
// runs have no gate buffers).
struct BatchExtras {
    bool gate = false;
    float* features = nullptr;   // kGateFeatures per hand-off entry
    float* logits = nullptr;
    unsigned* accepted = nullptr;  // one word per item
    float gate_thr = 0.0f;
    const float* gate_p = nullptr;
    unsigned long long* stamps = nullptr;  // per item, after its image
};

}  // namespace

struct tlr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    hipDeviceProp_t prop;

    // input
    tlr_header h{};
    int N = 0;
    bool has_samples = false;
    float* d_in = nullptr;          // 12 fields x N (x,y,w,u,v,t,r,g,b,mvx,mvy,mvw)
    size_t d_in_cap = 0;

    // build products (live until the next build)
    DeviceArena arena;
    bool built = false;
    GridInfo* d_gi = nullptr;
    GridInfo gi{};
    int spp_pow2 = 1;
    float cc0 = 0, cc1 = 0;
    int nb = 0, nil = 0, L = 0, nv = 0;
    float *rx = nullptr, *ry = nullptr, *rw = nullptr, *rt = nullptr;
    uint32_t* cand = nullptr;
    int4* ileaf_rect = nullptr;
    int2* ileaf_oc = nullptr;
    Points pts{};
    LeafRec* leaves = nullptr;
    Bounds* leaf_b = nullptr;
    // hierarchy: slot -> leaf id, internal children (leaf slot k is n-1+k), bounds
    uint32_t* tree_ids = nullptr;
    int* tree_c0 = nullptr;
    int* tree_c1 = nullptr;
    Bounds* tree_b = nullptr;
    TravNode* trav = nullptr;
    TravPair* pairs = nullptr;

    // reconstruct
    DeviceArena qarena;
    float* d_pattern = nullptr;
    float4* qcol = nullptr;
    size_t qcol_n = 0;
    float* d_img = nullptr;
    unsigned long long* d_stats = nullptr;
    unsigned* d_counter = nullptr;
    unsigned* d_fail_count = nullptr;
    unsigned* fail_list = nullptr;
    // second 1R pass (wide leaf lists) over fail_list: its work counter, its
    // hand-off list and count (what the general kernel then takes)
    unsigned* d_wide_next = nullptr;
    unsigned* d_fail_count2 = nullptr;
    unsigned* fail_list2 = nullptr;
    int wide_blocks = 0;
    float* scratch = nullptr;
    int num_waves = 0;
    int general_wpc = 0;  // resident general-kernel waves per CU
    int last_M = 0;
    float* features = nullptr;      // gate: per query features
    float* logits = nullptr;
    float* labels = nullptr;
    float4* prov = nullptr;         // gate: provisional colour per hand-off entry
    unsigned* d_accept_count = nullptr;
    float2* qdebug = nullptr;       // debug image values per output sample
    float* d_debug = nullptr;       // debug image (W*H*4) of the last reconstruct with debug_image
    bool has_debug = false;
    // reference GPU-variant method (path 2)
    DeviceArena rarena;             // lists, counters, stage codes, first 2R lists
    DeviceArena rwide;              // lists of the overflow launches
    int* ref_stage = nullptr;       // per query stage of the last path-2 call

    // learned gate
    float* gate_p = nullptr;        // fp32 master parameters
    float* gate_m = nullptr;
    float* gate_v = nullptr;
    float* gate_g = nullptr;
    float* gate_loss = nullptr;
    bool has_gate = false;
    // shadow inputs: per input sample (wgx, wgy, density, 0) (null: 0, 0, 1),
    // and the same per tree sample, gathered after each build on demand
    float4* d_wgd = nullptr;
    int wgd_n = 0;
    float4* pext = nullptr;
    size_t pext_cap = 0;
    bool pext_valid = false;
    DeviceArena sarena;             // explicit queries and shadow passes
    unsigned long long* d_stamps = nullptr;  // batch completion stamps (n + 1)
    int stamps_cap = 0;
    float gate_thr = 0.0f;
    int gate_step = 0;
    float* train_X = nullptr;
    float* train_y = nullptr;
    int* train_idx = nullptr;
    int64_t train_n = 0, train_cap = 0;

    hipEvent_t ev[9];
    bool events = false;

    // batched reconstruction
    std::vector<BatchSlot*> slots;
    DeviceArena barena;                  // items, block table, hand-off list, scratch
    hipGraphExec_t bgraph = nullptr;     // captured query phase of the last batch shape
    std::vector<uint64_t> bkey;          // that shape
    int tuned_variant = -1;              // autotuned fast-kernel variant
    int tuned_general_blocks = 0;        // autotuned general-kernel grid
    int tuned_chain = 0;                 // 1: per-buffer launches chained in the graph
    int batch_mode = -1;                 // tlr_batch_set_mode: -1 autotune, 0 merged, 1 chained

    // tlr_reconstruct_noalloc: this device's fixed scratch array and the
    // host copy of the output pattern last uploaded into it
    char* noalloc = nullptr;
    std::vector<float> noalloc_pat;

    tlr_status fail(tlr_status s, const char* fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        err = buf;
        return s;
    }
};

#define HIPCHK(ctx, call)                                                                              \
    do {                                                                                               \
        hipError_t _e = (call);                                                                        \
        if (_e != hipSuccess)                                                                          \
            return (ctx)->fail(TLR_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, \
                               __LINE__);                                                              \
    } while (0)

extern "C" {

tlr_status tlr_create(tlr_ctx** out, int device) {
    if (!out) return TLR_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return TLR_E_HIP;
    if (device < 0 || device >= n) return TLR_E_INVALID;
    tlr_ctx* c = new tlr_ctx;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&c->prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return TLR_E_HIP;
    }
    for (auto& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return TLR_E_HIP;
        }
    c->events = true;
    if (noalloc_scratch(device, &c->noalloc) != hipSuccess) {
        delete c;
        return TLR_E_HIP;
    }
    if (hipMalloc(&c->d_gi, sizeof(GridInfo)) != hipSuccess ||
        hipMalloc(&c->d_stats, sizeof(unsigned long long) * kNumStats) != hipSuccess ||
        hipMalloc(&c->d_counter, sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&c->d_fail_count, sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&c->d_wide_next, sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&c->d_fail_count2, sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&c->d_accept_count, sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&c->gate_p, sizeof(float) * kGateParams) != hipSuccess ||
        hipMalloc(&c->gate_m, sizeof(float) * kGateParams) != hipSuccess ||
        hipMalloc(&c->gate_v, sizeof(float) * kGateParams) != hipSuccess ||
        hipMalloc(&c->gate_g, sizeof(float) * kGateParams) != hipSuccess ||
        hipMalloc(&c->gate_loss, sizeof(float)) != hipSuccess) {
        delete c;
        return TLR_E_HIP;
    }
    *out = c;
    return TLR_OK;
}

void tlr_destroy(tlr_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->arena.release();
    c->qarena.release();
    c->rarena.release();
    c->rwide.release();
    if (c->d_in) (void)hipFree(c->d_in);
    if (c->d_gi) (void)hipFree(c->d_gi);
    if (c->d_stats) (void)hipFree(c->d_stats);
    if (c->d_counter) (void)hipFree(c->d_counter);
    if (c->d_wgd) (void)hipFree(c->d_wgd);
    if (c->pext) (void)hipFree(c->pext);
    if (c->d_stamps) (void)hipFree(c->d_stamps);
    if (c->d_fail_count) (void)hipFree(c->d_fail_count);
    if (c->d_wide_next) (void)hipFree(c->d_wide_next);
    if (c->d_fail_count2) (void)hipFree(c->d_fail_count2);
    if (c->d_accept_count) (void)hipFree(c->d_accept_count);
    for (float* p : {c->gate_p, c->gate_m, c->gate_v, c->gate_g, c->gate_loss, c->train_X, c->train_y})
        if (p) (void)hipFree(p);
    if (c->train_idx) (void)hipFree(c->train_idx);
    if (c->events)
        for (auto& e : c->ev) (void)hipEventDestroy(e);
    if (c->bgraph) (void)hipGraphExecDestroy(c->bgraph);
    for (BatchSlot* b : c->slots) {
        if (b->d_in) (void)hipFree(b->d_in);
        b->prod.release();
        delete b;
    }
    c->barena.release();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* tlr_last_error(const tlr_ctx* c) { return c ? c->err.c_str() : "null context"; }

// spp 0 (an irregular buffer, Reconstruction.cpp:65-71) only where the
// caller gives the sample count.
static tlr_status check_header(tlr_ctx* c, const tlr_header* h, bool irregular_ok = false) {
    if (!h || h->width <= 0 || h->height <= 0 || h->spp < 0 || (h->spp == 0 && !irregular_ok))
        return c->fail(TLR_E_INVALID, "invalid header");
    const int64_t n = (int64_t)h->width * h->height * h->spp;
    if (n >= (int64_t)1 << 31) return c->fail(TLR_E_INVALID, "sample count %lld exceeds 2^31", (long long)n);
    return TLR_OK;
}

static tlr_status reserve_input(tlr_ctx* c, int N) {
    const size_t need = (size_t)12 * N * sizeof(float);
    if (need > c->d_in_cap) {
        if (c->d_in) HIPCHK(c, hipFree(c->d_in));
        c->d_in = nullptr;
        c->d_in_cap = 0;
        HIPCHK(c, hipMalloc(&c->d_in, need));
        c->d_in_cap = need;
    }
    return TLR_OK;
}

// AoS records -> device SoA: the transpose runs on the host once per upload.
tlr_status tlr_set_samples(tlr_ctx* c, const tlr_header* h, const float* rec, int64_t n) {
    if (!c) return TLR_E_INVALID;
    tlr_status s = check_header(c, h, true);
    if (s != TLR_OK) return s;
    if (h->spp == 0 ? (!rec || n < 1 || n >= ((int64_t)1 << 31)) : (!rec || n != (int64_t)h->width * h->height * h->spp))
        return c->fail(TLR_E_INVALID, "record count does not match the header");
    HIPCHK(c, hipSetDevice(c->device));
    const int N = (int)n;
    s = reserve_input(c, N);
    if (s != TLR_OK) return s;
    static const int field[12] = {0, 1, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13};  // x,y,w,u,v,t,r,g,b,mv
    std::vector<float> tmp((size_t)N);
    for (int f = 0; f < 12; f++) {
        const int src = field[f];
        for (int k = 0; k < N; k++) tmp[k] = rec[(size_t)16 * k + src];
        HIPCHK(c, hipMemcpy(c->d_in + (size_t)f * N, tmp.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    }
    c->h = *h;
    c->N = N;
    c->has_samples = true;
    c->built = false;
    c->wgd_n = 0;
    return TLR_OK;
}

tlr_status tlr_set_samples_soa(tlr_ctx* c, const tlr_header* h, const tlr_soa* soa, int on_device) {
    if (!c) return TLR_E_INVALID;
    tlr_status s = check_header(c, h);
    if (s != TLR_OK) return s;
    if (!soa) return c->fail(TLR_E_INVALID, "null soa");
    HIPCHK(c, hipSetDevice(c->device));
    const int N = h->width * h->height * h->spp;
    s = reserve_input(c, N);
    if (s != TLR_OK) return s;
    const float* src[12] = {soa->x, soa->y, soa->w, soa->u, soa->v, soa->t,
                            soa->r, soa->g, soa->b, soa->mv_x, soa->mv_y, soa->mv_w};
    for (int f = 0; f < 12; f++) {
        if (!src[f]) return c->fail(TLR_E_INVALID, "null soa field %d", f);
        HIPCHK(c, hipMemcpyAsync(c->d_in + (size_t)f * N, src[f], sizeof(float) * N,
                                 on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->h = *h;
    c->N = N;
    c->has_samples = true;
    c->built = false;
    c->wgd_n = 0;
    return TLR_OK;
}

static InSoA in_view(tlr_ctx* c) {
    const float* b = c->d_in;
    const size_t N = (size_t)c->N;
    InSoA in;
    in.x = b; in.y = b + N; in.w = b + 2 * N; in.u = b + 3 * N; in.v = b + 4 * N; in.t = b + 5 * N;
    in.r = b + 6 * N; in.g = b + 7 * N; in.b = b + 8 * N;
    in.mvx = b + 9 * N; in.mvy = b + 10 * N; in.mvw = b + 11 * N;
    return in;
}

static uint32_t read_u32(tlr_ctx* c, const uint32_t* d) {
    uint32_t v = 0;
    (void)hipMemcpyAsync(&v, d, sizeof(v), hipMemcpyDeviceToHost, c->stream);
    (void)hipStreamSynchronize(c->stream);
    return v;
}

static tlr_status build_impl(tlr_ctx* c, float aperture_adjust, float focal_adjust, const float* coc_override);

tlr_status tlr_build(tlr_ctx* c, float aperture_adjust, float focal_adjust) {
    return build_impl(c, aperture_adjust, focal_adjust, nullptr);
}

tlr_status tlr_build_ex(tlr_ctx* c, float aperture_adjust, float focal_adjust, const float* coc_override) {
    if (coc_override && !(std::isfinite(coc_override[0]) && std::isfinite(coc_override[1])))
        return c ? c->fail(TLR_E_INVALID, "non-finite CoC coefficients") : TLR_E_INVALID;
    return build_impl(c, aperture_adjust, focal_adjust, coc_override);
}

static tlr_status build_impl(tlr_ctx* c, float aperture_adjust, float focal_adjust, const float* coc_override) {
    if (!c) return TLR_E_INVALID;
    if (!c->has_samples) return c->fail(TLR_E_STATE, "tlr_build before tlr_set_samples");
    HIPCHK(c, hipSetDevice(c->device));
    c->built = false;
    hipStream_t s = c->stream;
    const int N = c->N, W = c->h.width, H = c->h.height;

    // TreeGather::init preamble: Reconstruction.cpp:63-87; an irregular
    // buffer (spp 0) averages its per-pixel counts, integer division (:65-71)
    {
        uint32_t v = (uint32_t)(c->h.spp ? c->h.spp : N / (W * H));
        v--;
        v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
        v++;
        c->spp_pow2 = std::max(1, (int)v);
        switch (c->spp_pow2) {
            case 1: case 2: case 4: case 8: case 16: case 32: case 64: case 128: case 256: break;
            default: return c->fail(TLR_E_UNSUPPORTED_SPP, "unsupported spp %d", c->h.spp);
        }
        float cx = c->h.coc_c0, cy = c->h.coc_c1;
        float a = -cx;
        float f = -cx * rcp(cy);
        a *= aperture_adjust;
        f *= focal_adjust;
        c->cc0 = -a;
        c->cc1 = a * rcp(f);
        // overrideRefocusDistance (Reconstruction.cpp:88-95): the new
        // coefficients replace the adjusted ones; the reprojection to the
        // (u,v,t) centre keeps the header's (Reconstruction.cpp:519)
        if (coc_override) {
            c->cc0 = coc_override[0];
            c->cc1 = coc_override[1];
        }
    }

    // ---- device memory for this build ----
    const size_t cells_max = (size_t)kMaxGridSide * kMaxGridSide;
    const size_t hist_words = radix_hist_words((size_t)N);
    const size_t scan_words = scan_tmp_words(std::max<size_t>(std::max<size_t>((size_t)N, cells_max), hist_words));
    size_t words = 0;
    words += 4 * (size_t)N;           // rx ry rw rt
    words += 12 * (size_t)N;          // packed reprojected records
    words += 4 * (size_t)N;           // radix ping-pong
    words += 3 * (size_t)N;           // cand ct cw
    words += 12 * (size_t)N + 8 * kPointPad;  // points: pab (2 x float4, padded), pc
    words += 16 * (size_t)N;          // per-sample bounds
    words += 8 * (size_t)N;           // boxes
    words += (size_t)N + 1;           // leaf starts
    words += 2 * (size_t)N;           // start flags, leaf pos
    words += 8 * (size_t)N;           // leaf records
    words += 16 * (size_t)N;          // leaf bounds
    words += 2 * (size_t)N;           // valid flag / pos
    words += 4 * (size_t)N;           // hierarchy keys / values ping-pong
    words += 30 * (size_t)N;          // leaf features (8), leaf of valid id, split info (4), extents (16), slot ids
    words += 7 * (size_t)N;           // child0/1, parent(2N), range(2N), split owner
    words += 16 * (size_t)N + (size_t)N;  // node bounds, flags
    words += 32 * (size_t)N;          // traversal nodes (2N x 16 words)
    words += 32 * (size_t)N;          // child-pair records (N x 32 words)
    words += (size_t)W * H;           // hits
    words += 7 * cells_max + 2 * (size_t)kMaxGridSide + 16;  // counts, cell start, sat, owner(4), flags, y1
    words += (size_t)N;               // initial-leaf idx (<= cells)
    words += 6 * (size_t)N;           // initial leaf rect / offcnt
    words += hist_words + scan_words + 4096;
    const size_t bytes = words * 4 + 64 * 256;
    HIPCHK(c, c->arena.reserve(bytes));
    DeviceArena& A = c->arena;
    A.reset();

    c->rx = A.alloc<float>(N); c->ry = A.alloc<float>(N); c->rw = A.alloc<float>(N); c->rt = A.alloc<float>(N);
    float4* packed = A.alloc<float4>(3 * (size_t)std::max(N, 1));
    uint32_t* k0 = A.alloc<uint32_t>(N);
    uint32_t* v0 = A.alloc<uint32_t>(N);
    uint32_t* k1 = A.alloc<uint32_t>(N);
    uint32_t* v1 = A.alloc<uint32_t>(N);
    int* hits = A.alloc<int>((size_t)W * H);
    uint32_t* hist = A.alloc<uint32_t>(hist_words);
    uint32_t* stmp = A.alloc<uint32_t>(scan_words + 64);
    if (A.failed) return c->fail(TLR_E_HIP, "arena exhausted");

    // ---- reprojection + grid ----
    HIPCHK(c, hipEventRecord(c->ev[0], s));
    GridInfo g0;
    memset(&g0, 0, sizeof(g0));
    g0.bbmin_key[0] = g0.bbmin_key[1] = 0x7FFFFFFF;
    g0.bbmax_key[0] = g0.bbmax_key[1] = (int)0x80000000;
    HIPCHK(c, hipMemcpyAsync(c->d_gi, &g0, sizeof(g0), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemsetAsync(hits, 0, sizeof(int) * (size_t)W * H, s));
    const InSoA in = in_view(c);
    launch_reproject(in, N, W, H, c->h.coc_c0, c->h.coc_c1, c->rx, c->ry, c->rw, c->rt, packed, hits, c->d_gi, s);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(&c->gi, c->d_gi, sizeof(GridInfo), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (c->gi.error == 1) return c->fail(TLR_E_INVALID, "%u samples reproject to non-finite positions", c->gi.num_bad);
    if (c->gi.error == 3)
        return c->fail(TLR_E_ALL_DISCARDED, "All samples were discarded in reprojection to (u,v,t)=0");
    const int gw = c->gi.gw, gh = c->gi.gh;
    if (gw <= 0 || gh <= 0 || gw > kMaxGridSide || gh > kMaxGridSide)
        return c->fail(TLR_E_INVALID, "degenerate bucketing grid %dx%d", gw, gh);
    const size_t cells = (size_t)gw * gh;

    uint32_t* cell_end = A.alloc<uint32_t>(cells);
    uint32_t* cell_start = A.alloc<uint32_t>(cells + 1);
    uint32_t* sat = A.alloc<uint32_t>((size_t)(gw + 1) * (gh + 1));
    int4* owner = A.alloc<int4>(cells);
    uint32_t* own_flag = A.alloc<uint32_t>(cells);
    int* own_y1 = A.alloc<int>(cells);
    uint32_t* own_idx = A.alloc<uint32_t>(cells + 1);
    if (A.failed) return c->fail(TLR_E_HIP, "arena exhausted (grid)");

    // ---- bucketing: stable sort of (cell, input index); each cell's run ----
    launch_cell_keys(c->rx, c->ry, c->rw, N, c->d_gi, k0, v0, s);
    HIPCHK(c, hipGetLastError());
    bool first = true;
    HIPCHK(c, radix_sort_pairs(k0, v0, k1, v1, (size_t)N, bits_for(cells + 1), hist, stmp, &first, s));
    const uint32_t* skeys = first ? k0 : k1;
    const uint32_t* svals = first ? v0 : v1;
    HIPCHK(c, hipMemsetAsync(cell_start, 0, sizeof(uint32_t) * cells, s));
    HIPCHK(c, hipMemsetAsync(cell_end, 0, sizeof(uint32_t) * cells, s));
    launch_cell_runs(skeys, N, (uint32_t)cells, cell_start, cell_end, s);
    HIPCHK(c, hipGetLastError());
    const int nb = (int)read_u32(c, &c->d_gi->num_bucketed);
    c->nb = nb;

    // ---- initial kd split ----
    launch_sat(cell_start, cell_end, gw, gh, sat, s);
    launch_kd(sat, gw, gh, owner, own_flag, own_y1, s);
    HIPCHK(c, exclusive_scan_u32(own_flag, own_idx, own_idx + cells, cells, stmp, s));
    const int nil = (int)read_u32(c, own_idx + cells);
    c->nil = nil;
    c->ileaf_rect = A.alloc<int4>(std::max(nil, 1));
    c->ileaf_oc = A.alloc<int2>(std::max(nil, 1));
    c->cand = A.alloc<uint32_t>(std::max(nb, 1));
    float* ct = A.alloc<float>(std::max(nb, 1));
    float* cw = A.alloc<float>(std::max(nb, 1));
    if (A.failed) return c->fail(TLR_E_HIP, "arena exhausted (candidates)");
    launch_emit_initial(owner, own_flag, own_y1, own_idx, sat, gw, gh, c->ileaf_rect, c->ileaf_oc, s);
    launch_place(skeys, svals, nb, gw, owner, sat, cell_start, packed, c->cand, ct, cw, s);
    launch_leaf_sort(c->ileaf_oc, nil, c->cand, ct, cw, s);
    HIPCHK(c, hipGetLastError());

    // ---- per-sample data in tree order ----
    const int nb1 = std::max(nb, 1);
    Points P;
    P.pab = A.alloc<float4>(2 * ((size_t)nb1 + kPointPad));
    P.pc = A.alloc<float4>(nb1);
    c->pts = P;
    Bounds* sb = A.alloc<Bounds>(nb1);
    float4* boxes = A.alloc<float4>(2 * (size_t)nb1);
    uint32_t* sflag = A.alloc<uint32_t>(nb1 + 1);
    uint32_t* lpos = A.alloc<uint32_t>(nb1 + 1);
    if (A.failed) return c->fail(TLR_E_HIP, "arena exhausted (points)");
    launch_points(c->cand, nb, packed, (float)W, (float)H, c->cc0, c->cc1, P, sb, boxes, s);
    HIPCHK(c, hipMemsetAsync(sflag, 0, sizeof(uint32_t) * (nb1 + 1), s));
    HIPCHK(c, hipMemsetAsync(P.pab + 2 * (size_t)nb, 0, sizeof(float4) * 2 * kPointPad, s));
    launch_surface_scan(c->ileaf_oc, nil, P.pab, boxes, sflag, s);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, exclusive_scan_u32(sflag, lpos, lpos + nb, (size_t)nb, stmp, s));
    const int L = nb > 0 ? (int)read_u32(c, lpos + nb) : 0;
    c->L = L;
    const int L1 = std::max(L, 1);
    c->leaves = A.alloc<LeafRec>(L1);
    c->leaf_b = A.alloc<Bounds>(L1);
    uint32_t* vflag = A.alloc<uint32_t>(L1 + 1);
    uint32_t* vpos = A.alloc<uint32_t>(L1 + 1);
    int* lstart = A.alloc<int>((size_t)L1 + 1);
    if (A.failed) return c->fail(TLR_E_HIP, "arena exhausted (leaves)");
    launch_leaf_emit(sflag, lpos, nb, L, lstart, sb, P, c->leaves, c->leaf_b, vflag, s);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, exclusive_scan_u32(vflag, vpos, vpos + L, (size_t)L, stmp, s));
    const int nv = L > 0 ? (int)read_u32(c, vpos + L) : 0;
    c->nv = nv;

    // ---- median-split hierarchy over the valid leaves ----
    const int nv1 = std::max(nv, 1);
    float* feat = A.alloc<float>(8 * (size_t)nv1);
    uint32_t* vid_leaf = A.alloc<uint32_t>(nv1);
    uint32_t* tk0 = A.alloc<uint32_t>(nv1);
    uint32_t* tv0 = A.alloc<uint32_t>(nv1);
    uint32_t* tk1 = A.alloc<uint32_t>(nv1);
    uint32_t* tv1 = A.alloc<uint32_t>(nv1);
    float4* kinfo = A.alloc<float4>(nv1);
    uint32_t* kmn = A.alloc<uint32_t>(8 * (size_t)nv1);
    uint32_t* kmx = A.alloc<uint32_t>(8 * (size_t)nv1);
    c->tree_ids = A.alloc<uint32_t>(nv1);
    c->tree_c0 = A.alloc<int>(nv1);
    c->tree_c1 = A.alloc<int>(nv1);
    int* parent = A.alloc<int>(2 * (size_t)nv1);
    int2* range = A.alloc<int2>(nv1);
    int* split_owner = A.alloc<int>(nv1);
    c->tree_b = A.alloc<Bounds>(nv1);
    c->trav = A.alloc<TravNode>(2 * (size_t)nv1);
    c->pairs = A.alloc<TravPair>(nv1);
    if (A.failed) return c->fail(TLR_E_HIP, "arena exhausted (hierarchy)");
    if (nv > 0) {
        HIPCHK(c, hipMemsetAsync(parent, 0xFF, sizeof(int) * (2 * (size_t)nv), s));
        HIPCHK(c, launch_kdtree(vflag, vpos, L, c->leaf_b, nv, feat, vid_leaf, tk0, tv0, tk1, tv1, kinfo, kmn, kmx, hist, stmp,
                                c->tree_ids, c->tree_c0, c->tree_c1, parent, range, split_owner, c->tree_b, s));
        launch_tree_trav(c->tree_ids, nv, c->leaf_b, c->tree_b, c->tree_c0, c->tree_c1, range, split_owner,
                         c->leaves, c->trav, c->pairs, s);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipEventRecord(c->ev[1], s));
    HIPCHK(c, hipStreamSynchronize(s));
    HIPCHK(c, hipMemcpy(&c->gi, c->d_gi, sizeof(GridInfo), hipMemcpyDeviceToHost));
    c->built = true;
    c->pext_valid = false;
    return TLR_OK;
}

// Allocates the per-reconstruction buffers and fills the query parameters.
static tlr_status prepare_queries(tlr_ctx* c, const tlr_recon_params* p, bool gate_buffers, QueryParams& Q) {
    const bool debug = p->debug_image != 0;
    if (!c->built) return c->fail(TLR_E_STATE, "reconstruct before tlr_build");
    if (p->out_spp <= 0 && !p->override_uvt) return c->fail(TLR_E_INVALID, "out_spp must be positive");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    std::vector<float> pat;
    const int M = make_output_pattern(p->out_spp, p->override_uvt != 0, p->uvt, pat);
    const int W = c->h.width, H = c->h.height;
    const size_t nq = (size_t)W * H * M;
    if (nq >= ((size_t)1 << 31)) return c->fail(TLR_E_INVALID, "too many output samples");
    // persistent general kernel: as many waves as can be resident
    if (c->general_wpc == 0) c->general_wpc = general_waves_per_cu();
    const int waves = std::max(4, c->prop.multiProcessorCount * c->general_wpc);
    size_t qbytes = pat.size() * 4 + nq * sizeof(float4) + (size_t)W * H * 16 + 2 * nq * 4 +
                    (size_t)waves * kWaveScratchWords * 4 + 16 * 256;
    if (gate_buffers) qbytes += nq * 4 * ((size_t)kGateFeatures + 2) + nq * sizeof(float4) + 5 * 256;
    if (debug) qbytes += nq * sizeof(float2) + (size_t)W * H * 16 + 2 * 256;
    HIPCHK(c, c->qarena.reserve(qbytes));
    c->qarena.reset();
    c->d_pattern = c->qarena.alloc<float>(pat.size());
    c->qcol = c->qarena.alloc<float4>(nq);
    c->d_img = c->qarena.alloc<float>((size_t)W * H * 4);
    c->fail_list = c->qarena.alloc<unsigned>(nq);
    c->fail_list2 = c->qarena.alloc<unsigned>(nq);
    c->scratch = c->qarena.alloc<float>((size_t)waves * kWaveScratchWords);
    c->features = c->logits = c->labels = nullptr;
    c->prov = nullptr;
    c->qdebug = nullptr;
    c->d_debug = nullptr;
    c->has_debug = false;
    if (debug) {
        c->qdebug = c->qarena.alloc<float2>(nq);
        c->d_debug = c->qarena.alloc<float>((size_t)W * H * 4);
    }
    if (gate_buffers) {
        c->features = c->qarena.alloc<float>(nq * kGateFeatures);
        c->logits = c->qarena.alloc<float>(nq);
        c->labels = c->qarena.alloc<float>(nq);
        c->prov = c->qarena.alloc<float4>(nq);
    }
    if (c->qarena.failed) return c->fail(TLR_E_HIP, "query arena exhausted");
    c->qcol_n = nq;
    c->num_waves = waves;
    c->last_M = M;
    HIPCHK(c, hipMemcpyAsync(c->d_pattern, pat.data(), pat.size() * 4, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemsetAsync(c->d_stats, 0, sizeof(unsigned long long) * kNumStats, s));
    HIPCHK(c, hipMemsetAsync(c->d_counter, 0, sizeof(unsigned), s));
    HIPCHK(c, hipMemsetAsync(c->d_fail_count, 0, sizeof(unsigned), s));
    HIPCHK(c, hipMemsetAsync(c->d_accept_count, 0, sizeof(unsigned), s));
    HIPCHK(c, hipMemsetAsync(c->d_wide_next, 0, sizeof(unsigned), s));
    HIPCHK(c, hipMemsetAsync(c->d_fail_count2, 0, sizeof(unsigned), s));
    if (c->wide_blocks == 0) c->wide_blocks = c->prop.multiProcessorCount * fast_wide_blocks_per_cu();

    memset(&Q, 0, sizeof(Q));
    Q.W = W; Q.H = H; Q.M = M;
    Q.spp_pow2 = c->spp_pow2;
    Q.cc0 = c->cc0; Q.cc1 = c->cc1;
    Q.lens_gauss = p->lens_gaussian;
    Q.time_gauss = p->time_gaussian;
    Q.collect_stats = p->collect_stats;
    Q.n_nodes = c->nv > 0 ? 2 * c->nv - 1 : 0;
    Q.root = c->nv > 1 ? c->nv / 2 - 1 : 0;
    Q.num_queries = (unsigned)nq;
    Q.pattern = c->d_pattern;
    Q.trav = c->trav;
    Q.pairs = c->pairs;
    Q.leaves = c->leaves;
    Q.leaf_b = c->leaf_b;
    Q.pts = c->pts;
    Q.qcol = c->qcol;
    Q.stats = c->d_stats;
    Q.next_query = c->d_counter;
    Q.scratch = c->scratch;
    Q.num_waves = waves;
    Q.fail_list = c->fail_list;
    Q.fail_count = c->d_fail_count;
    Q.qdebug = c->qdebug;
    // tile order pads every 8x8 tile to whole waves; the fast kernel's lane
    // index is 32-bit, so very narrow images fall back to query order
    Q.tile_order = TLR_TILE_ORDER;
    if (fast_slots(Q) >= ((size_t)1 << 32)) Q.tile_order = 0;
    return TLR_OK;
}

// The second 1R lane pass (wide leaf lists) over the context's hand-off
// list, into its second list; event 7 marks its start.  Entries marked for
// 2R, and entries the gate accepted (kSkipEntry has that bit set), pass
// through unchanged.
static tlr_status run_wide(tlr_ctx* c, const QueryParams& Q0, bool stats_mode) {
    hipStream_t s = c->stream;
    QueryParams V = Q0;
    V.fast_list = 1;
    V.qlist = c->fail_list;
    V.qcount = c->d_fail_count;
    V.next_query = c->d_wide_next;
    V.fail_list = c->fail_list2;
    V.fail_count = c->d_fail_count2;
    V.features = nullptr;
    HIPCHK(c, hipEventRecord(c->ev[7], s));
    HIPCHK(c, launch_query_fast_wide(V, stats_mode, c->wide_blocks, s));
    return TLR_OK;
}

// Modes 1 and 2 (extended results, shadow queries): the 1R lane kernel over
// F's queries (F.fast_list: over F.qlist), the second lane pass over its
// hand-offs, the general kernel over what is left.  list1/cnt1 and
// list2/cnt2 are the two hand-off lists (counts zeroed here), wide_next the
// second pass's work counter, gen_next the general kernel's.
static tlr_status run_mode_passes(tlr_ctx* c, const QueryParams& F0, unsigned* list1, unsigned* cnt1,
                                  unsigned* list2, unsigned* cnt2, unsigned* wide_next, unsigned* gen_next,
                                  bool stats_mode) {
    hipStream_t s = c->stream;
    if (c->wide_blocks == 0) c->wide_blocks = c->prop.multiProcessorCount * fast_wide_blocks_per_cu();
    HIPCHK(c, hipMemsetAsync(cnt1, 0, sizeof(unsigned), s));
    HIPCHK(c, hipMemsetAsync(cnt2, 0, sizeof(unsigned), s));
    HIPCHK(c, hipMemsetAsync(wide_next, 0, sizeof(unsigned), s));
    HIPCHK(c, hipMemsetAsync(gen_next, 0, sizeof(unsigned), s));
    QueryParams F = F0;
    F.fail_list = list1;
    F.fail_count = cnt1;
    HIPCHK(c, launch_query_fast_mode(F, stats_mode, s));
    QueryParams V = F0;
    V.fast_list = 1;
    V.qlist = list1;
    V.qcount = cnt1;
    V.next_query = wide_next;
    V.fail_list = list2;
    V.fail_count = cnt2;
    HIPCHK(c, launch_query_fast_wide(V, stats_mode, c->wide_blocks, s));
    QueryParams G = F0;
    G.fast_list = 0;
    G.qlist = list2;
    G.qcount = cnt2;
    G.next_query = gen_next;
    HIPCHK(c, launch_query_general(G, s));
    return TLR_OK;
}

// Plain method over all queries (or over Q.qlist when list != nullptr):
// fast 1R lane kernel, then (wide) the second 1R lane pass with the
// context's wide lists over its hand-offs, then the general kernel for what
// is left.
static tlr_status run_plain(tlr_ctx* c, const QueryParams& Q0, const unsigned* list, const unsigned* count,
                            bool stats_mode, int path, bool wide = false) {
    hipStream_t s = c->stream;
    if (path == 1 && !list) {
        HIPCHK(c, launch_query_general(Q0, s));
        HIPCHK(c, hipEventRecord(c->ev[5], s));
        return TLR_OK;
    }
    QueryParams F = Q0;
    if (list) {
        F.fast_list = 1;
        F.qlist = list;
        F.qcount = count;
    }
    HIPCHK(c, launch_query_fast(F, stats_mode, s));
    // the general kernel takes the hand-offs from the list the fast kernel
    // just filled (Q0's, which need not be the context's own)
    QueryParams G = Q0;
    G.qlist = Q0.fail_list;
    G.qcount = Q0.fail_count;
    if (wide) {
        // (the context's own lists: Q0's are those)
        const tlr_status st = run_wide(c, Q0, stats_mode);
        if (st != TLR_OK) return st;
        G.qlist = c->fail_list2;
        G.qcount = c->d_fail_count2;
    }
    HIPCHK(c, hipEventRecord(c->ev[5], s));
    HIPCHK(c, launch_query_general(G, s));
    return TLR_OK;
}

// The reference's GPU-variant method (ReconstructionCuda.cpp:327-465): 1R
// over every output sample, 2R over the ones with several surfaces, nearest
// sample for the rest; samples beyond a launch's per-lane lists are run again
// with lists sized for them.  *n2r: output samples that went to 2R.
static tlr_status run_refgpu(tlr_ctx* c, const QueryParams& Q, int list_cap, unsigned* n2r) {
    hipStream_t s = c->stream;
    const size_t nq = Q.num_queries;
    const int cus = c->prop.multiProcessorCount;
    const int blocks = ref_stage_blocks(cus, kRefWavesPerCU);
    const size_t waves = (size_t)blocks * 4;
    const size_t fail_bytes = waves * kRefWaveScratchWords * 4;
    enum { N_FAIL, N_NEAR, N_OVER1, N_OVER2, N_OVER3, N_OVER4, NX0, NX1, NX2, NX3, NX4, NEED1, NEED1B, NEED2,
           NEED2B, NEED3, NEED3B, NEED4, NEED4B, NUM_COUNTERS };
    HIPCHK(c, c->rarena.reserve(5 * nq * 4 + fail_bytes + 16 * 256));
    c->rarena.reset();
    unsigned* cnt = c->rarena.alloc<unsigned>(NUM_COUNTERS);
    unsigned* fail = c->rarena.alloc<unsigned>(nq);
    unsigned* near = c->rarena.alloc<unsigned>(nq);
    unsigned* over1 = c->rarena.alloc<unsigned>(nq);
    unsigned* over2 = c->rarena.alloc<unsigned>(nq);
    c->ref_stage = c->rarena.alloc<int>(nq);
    float* wscratch = c->rarena.alloc<float>(waves * kRefWaveScratchWords);
    if (c->rarena.failed) return c->fail(TLR_E_HIP, "reference-method arena exhausted");
    HIPCHK(c, hipMemsetAsync(cnt, 0, NUM_COUNTERS * sizeof(unsigned), s));

    RefLaunch R;
    memset(&R, 0, sizeof(R));
    R.q = Q;
    R.to_fail = fail;
    R.n_fail = cnt + N_FAIL;
    R.to_near = near;
    R.n_near = cnt + N_NEAR;
    R.stage = c->ref_stage;

    // Overflow launch: lists sized by need[] for the samples in `list`.
    auto wide = [&](int stage, unsigned* list, unsigned* count, unsigned n, const unsigned* need, unsigned* over,
                    unsigned* n_over, unsigned* need_out, unsigned* next) -> tlr_status {
        const size_t leaf_cap = std::max(1u, need[0]), point_cap = stage == 1 ? std::max(1u, need[1]) : 0;
        const size_t per_wave = 64 * (leaf_cap * sizeof(int) + point_cap * 8);
        const size_t budget = (size_t)1 << 30;
        size_t waves = std::min<size_t>((n + 63) / 64, (size_t)cus * kRefWavesPerCU);
        waves = std::min(waves, std::max<size_t>(1, budget / per_wave));
        waves = (waves + 3) / 4 * 4;
        if (per_wave > ((size_t)4 << 30))
            return c->fail(TLR_E_CAPACITY, "an output sample needs %zu leaves and %zu points", leaf_cap, point_cap);
        HIPCHK(c, c->rwide.reserve(waves * per_wave + 2 * 256));
        c->rwide.reset();
        RefLaunch W = R;
        W.list = list;
        W.count = count;
        W.next = next;
        W.to_over = over;
        W.n_over = n_over;
        W.need = need_out;
        W.leaf_cap = (int)leaf_cap;
        W.point_cap = (int)point_cap;
        W.lists = c->rwide.alloc<int>(waves * 64 * leaf_cap);
        W.xy = point_cap ? c->rwide.alloc<float2>(waves * 64 * point_cap) : nullptr;
        if (c->rwide.failed) return c->fail(TLR_E_HIP, "reference-method overflow arena exhausted");
        HIPCHK(c, launch_ref_stage(W, stage, (int)(waves / 4), s));
        return TLR_OK;
    };

    // 1R over every output sample
    const int cap1 = list_cap > 0 ? std::min(list_cap, kRefFirstCap) : kRefFirstCap;
    RefLaunch F = R;
    F.leaf_cap = cap1;
    F.to_over = over1;
    F.n_over = cnt + N_OVER1;
    F.need = cnt + NEED1;
    HIPCHK(c, launch_ref_first(F, s));
    HIPCHK(c, hipEventRecord(c->ev[5], s));
    unsigned h[NUM_COUNTERS];
    HIPCHK(c, hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (h[N_OVER1]) {
        tlr_status st = wide(0, over1, cnt + N_OVER1, h[N_OVER1], h + NEED1, over2, cnt + N_OVER3, cnt + NEED3,
                             cnt + NX1);
        if (st != TLR_OK) return st;
    }
    // 2R over the samples with several surfaces at 1R
    RefLaunch T = R;
    T.list = fail;
    T.count = cnt + N_FAIL;
    T.next = cnt + NX0;
    T.to_over = over1;
    T.n_over = cnt + N_OVER2;
    T.need = cnt + NEED2;
    T.wscratch = wscratch;
    T.leaf_cap = list_cap > 0 ? std::min(list_cap, kRefWaveLeaves) : kRefWaveLeaves;
    T.point_cap = list_cap > 0 ? std::min(list_cap, kRefWavePoints) : kRefWavePoints;
    HIPCHK(c, launch_ref_fail_wave(T, blocks, s));
    HIPCHK(c, hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (h[N_OVER3])
        return c->fail(TLR_E_CAPACITY, "%u output samples exceeded the sized 1R lists", h[N_OVER3]);
    if (h[N_OVER2]) {
        tlr_status st = wide(1, over1, cnt + N_OVER2, h[N_OVER2], h + NEED2, over2, cnt + N_OVER4, cnt + NEED4,
                             cnt + NX2);
        if (st != TLR_OK) return st;
    }
    // nearest sample for the rest
    RefLaunch N = R;
    N.list = near;
    N.count = cnt + N_NEAR;
    N.next = cnt + NX3;
    HIPCHK(c, launch_ref_nearest(N, blocks, s));
    HIPCHK(c, hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (h[N_OVER4])
        return c->fail(TLR_E_CAPACITY, "%u output samples exceeded the sized 2R lists", h[N_OVER4]);
    *n2r = h[N_FAIL];
    return TLR_OK;
}

tlr_status tlr_reconstruct(tlr_ctx* c, const tlr_recon_params* p, float* out, int out_on_device, tlr_stats* stats) {
    if (!c || !p) return TLR_E_INVALID;
    if (out_on_device && !out) return c->fail(TLR_E_INVALID, "out_on_device set with a NULL rgba_out");
    if (p->gate && !c->has_gate) return c->fail(TLR_E_STATE, "gate requested but no gate parameters set");
    if (p->path < 0 || p->path > 2) return c->fail(TLR_E_INVALID, "unknown path %d", p->path);
    if (p->path == 2 && p->gate) return c->fail(TLR_E_INVALID, "the gate applies to the plain method only");
    if (p->debug_image && (p->gate || p->path == 2))
        return c->fail(TLR_E_INVALID, "the debug image comes from the plain method (path 0 or 1, no gate)");
    c->ref_stage = nullptr;
    QueryParams Q;
    tlr_status st = prepare_queries(c, p, p->gate != 0, Q);
    if (st != TLR_OK) return st;
    hipStream_t s = c->stream;
    const int W = c->h.width, H = c->h.height, M = c->last_M;
    const size_t nq = c->qcol_n;
    unsigned n2r = 0;

    HIPCHK(c, hipEventRecord(c->ev[2], s));
    if (p->gate) {
        // the 1R lane kernel with provisional first-surface colours and
        // features for its hand-offs; the gate scores the hand-offs; the
        // general kernel runs only for the ones it does not accept
        QueryParams F = Q;
        F.features = c->features;
        HIPCHK(c, launch_query_fast_gate(F, s));
        HIPCHK(c, hipEventRecord(c->ev[5], s));
        HIPCHK(c, launch_gate_forward_dev(c->features, c->d_fail_count, nq, c->gate_p, c->logits, s));
        HIPCHK(c, launch_gate_accept(c->logits, c->features, c->fail_list, c->d_fail_count, nq, c->gate_thr,
                                     c->d_accept_count, s));
        HIPCHK(c, hipEventRecord(c->ev[6], s));
        // the second lane pass over what the gate left, then the general kernel
        st = run_wide(c, Q, false);
        if (st != TLR_OK) return st;
        HIPCHK(c, hipEventRecord(c->ev[8], s));
        QueryParams G = Q;
        G.qlist = c->fail_list2;
        G.qcount = c->d_fail_count2;
        HIPCHK(c, launch_query_general(G, s));
    } else if (p->path == 2) {
        st = run_refgpu(c, Q, p->list_cap, &n2r);
    } else {
        st = run_plain(c, Q, nullptr, nullptr, p->collect_stats != 0 || p->debug_image != 0, p->path, true);
    }
    if (st != TLR_OK) return st;
    if (p->debug_image) {
        HIPCHK(c, launch_debug_image(c->qdebug, W, H, M, c->d_debug, s));
        c->has_debug = true;
    }
    HIPCHK(c, hipEventRecord(c->ev[3], s));
    float* dst = out_on_device ? out : c->d_img;
    HIPCHK(c, launch_combine(c->qcol, W, H, M, dst, s));
    HIPCHK(c, hipEventRecord(c->ev[4], s));
    if (!out_on_device && out)
        HIPCHK(c, hipMemcpyAsync(out, c->d_img, (size_t)W * H * 16, hipMemcpyDeviceToHost, s));
    unsigned long long hs[kNumStats];
    unsigned nfail = 0, naccept = 0;
    HIPCHK(c, hipMemcpyAsync(hs, c->d_stats, sizeof(hs), hipMemcpyDeviceToHost, s));
    const bool wide = p->path == 0;
    unsigned nwide = 0;
    HIPCHK(c, hipMemcpyAsync(&nfail, wide ? c->d_fail_count2 : c->d_fail_count, sizeof(nfail),
                             hipMemcpyDeviceToHost, s));
    if (wide) HIPCHK(c, hipMemcpyAsync(&nwide, c->d_fail_count, sizeof(nwide), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(&naccept, c->d_accept_count, sizeof(naccept), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->leaf_nodes = (double)hs[ST_LEAF_NODES];
        stats->samples_in_leaf_nodes = (double)hs[ST_SAMPLES_IN_LEAVES];
        stats->samples_within_1r = (double)hs[ST_WITHIN_1R];
        stats->samples_fetched = (double)hs[ST_FETCHED];
        stats->surfaces = (double)hs[ST_SURFACES];
        stats->num_2r = (double)hs[ST_2R];
        stats->surfaces_merged = (double)hs[ST_MERGED];
        stats->at_least_one = (double)hs[ST_AT_LEAST_ONE];
        stats->count = (double)hs[ST_COUNT];
        stats->overflow = (double)hs[ST_OVERFLOW];
        stats->fast_leaf_nodes = (double)hs[ST_FAST_LEAVES];
        stats->fast_samples_in_leaf_nodes = (double)hs[ST_FAST_IN_LEAVES];
        stats->fast_samples_within_1r = (double)hs[ST_FAST_KEPT];
        stats->wide_leaf_nodes = (double)hs[ST_WIDE_LEAVES];
        stats->wide_samples_in_leaf_nodes = (double)hs[ST_WIDE_IN_LEAVES];
        stats->wide_samples_within_1r = (double)hs[ST_WIDE_KEPT];
        float ms = 0;
        if (wide) {
            // entries the first pass had marked for 2R are relayed, not run
            stats->wide_queries = (double)nwide;
            if (hipEventElapsedTime(&ms, c->ev[7], c->ev[p->gate ? 8 : 5]) == hipSuccess) stats->wide_ms = ms;
        }
        if (hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) stats->build_ms = ms;
        if (hipEventElapsedTime(&ms, c->ev[2], c->ev[3]) == hipSuccess) stats->query_ms = ms;
        if (hipEventElapsedTime(&ms, c->ev[3], c->ev[4]) == hipSuccess) stats->combine_ms = ms;
        if (hipEventElapsedTime(&ms, c->ev[2], c->ev[5]) == hipSuccess) stats->fast_ms = ms;
        if (p->gate) {
            if (hipEventElapsedTime(&ms, c->ev[5], c->ev[6]) == hipSuccess) stats->gate_ms = ms;
            if (hipEventElapsedTime(&ms, c->ev[8], c->ev[3]) == hipSuccess) stats->general_ms = ms;
            // the gate scored the first lane pass's hand-offs; accepted
            // entries are relayed through the second pass and skipped
            stats->gate_candidates = (double)nwide;
            stats->gate_rejected = (double)(nwide - naccept);
        } else if (hipEventElapsedTime(&ms, c->ev[5], c->ev[3]) == hipSuccess) {
            stats->general_ms = ms;
        }
        stats->general_queries = (p->path == 1 && !p->gate) ? (double)nq : (double)(p->gate ? nfail - naccept : nfail);
        if (p->gate) stats->fast_ms += stats->wide_ms;
        if (p->path == 2) {
            stats->num_2r = stats->general_queries = (double)n2r;
            stats->count = (double)nq;
        }
        stats->total_ms = stats->build_ms + stats->query_ms + stats->combine_ms;
        stats->num_leaves = c->L;
        stats->num_samples = c->nb;
    }
    if (hs[ST_OVERFLOW])
        return c->fail(TLR_E_CAPACITY, "%llu output samples exceeded the per-query capacity",
                       (unsigned long long)hs[ST_OVERFLOW]);
    return TLR_OK;
}

// Allocation-free variant of tlr_reconstruct (plain method): every buffer of
// the query phase lives in this device's fixed scratch array, and the image
// is reconstructed in row bands of at most kNoAllocBandQueries output samples
// (fast kernel, general kernel over its hand-offs, combine into the caller's
// rows).  No allocation: counters are cleared by a fill kernel, the pattern
// is copied from a host vector the context keeps.
tlr_status tlr_reconstruct_noalloc(tlr_ctx* c, const tlr_recon_params* p, float* out, tlr_stats* stats) {
    if (!c || !p) return TLR_E_INVALID;
    if (!out) return c->fail(TLR_E_INVALID, "tlr_reconstruct_noalloc needs a device destination");
    if (!c->built) return c->fail(TLR_E_STATE, "reconstruct before tlr_build");
    if (p->gate || p->path == 2 || p->path < 0 || p->path > 2)
        return c->fail(TLR_E_INVALID, "tlr_reconstruct_noalloc runs the plain method only (path 0 or 1, no gate)");
    if (p->out_spp <= 0 && !p->override_uvt) return c->fail(TLR_E_INVALID, "out_spp must be positive");
    if (!p->override_uvt && p->out_spp > kNoAllocMaxM)
        return c->fail(TLR_E_INVALID, "tlr_reconstruct_noalloc supports at most %d output samples per pixel",
                       kNoAllocMaxM);
    const int W = c->h.width, H = c->h.height;
    const int M = p->override_uvt ? 1 : p->out_spp;
    if ((size_t)W * M > kNoAllocBandQueries)
        return c->fail(TLR_E_INVALID, "an image row of %zu output samples exceeds the %zu of one band",
                       (size_t)W * M, kNoAllocBandQueries);
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    c->noalloc_pat.clear();
    make_output_pattern(p->out_spp, p->override_uvt != 0, p->uvt, c->noalloc_pat);
    char* base = c->noalloc;
    unsigned* counters = reinterpret_cast<unsigned*>(base + kNoAllocCounters);
    unsigned long long* st = reinterpret_cast<unsigned long long*>(base + kNoAllocCounters + 32);
    float* pattern = reinterpret_cast<float*>(base + kNoAllocPattern);
    float4* qcol = reinterpret_cast<float4*>(base + kNoAllocQcol);
    unsigned* fail = reinterpret_cast<unsigned*>(base + kNoAllocFail);
    float* wscratch = reinterpret_cast<float*>(base + kNoAllocWave);
    HIPCHK(c, hipMemcpyAsync(pattern, c->noalloc_pat.data(), c->noalloc_pat.size() * 4, hipMemcpyHostToDevice, s));
    HIPCHK(c, launch_fill_u32(counters, 64, 0u, s));
    // rows per band: a multiple of 8 (whole 8x8 tiles) when the band holds 8
    int rows = (int)(kNoAllocBandQueries / ((size_t)W * M));
    if (rows >= 8) rows &= ~7;
    rows = std::min(rows, H);
    HIPCHK(c, hipEventRecord(c->ev[2], s));
    for (int y0 = 0; y0 < H; y0 += rows) {
        const int bh = std::min(rows, H - y0);
        QueryParams Q;
        memset(&Q, 0, sizeof(Q));
        Q.W = W; Q.H = bh; Q.M = M; Q.y0 = y0;
        Q.spp_pow2 = c->spp_pow2;
        Q.cc0 = c->cc0; Q.cc1 = c->cc1;
        Q.lens_gauss = p->lens_gaussian;
        Q.time_gauss = p->time_gaussian;
        Q.collect_stats = p->collect_stats;
        Q.n_nodes = c->nv > 0 ? 2 * c->nv - 1 : 0;
        Q.root = c->nv > 1 ? c->nv / 2 - 1 : 0;
        Q.num_queries = (unsigned)((size_t)W * bh * M);
        Q.pattern = pattern;
        Q.trav = c->trav;
        Q.pairs = c->pairs;
        Q.leaves = c->leaves;
        Q.leaf_b = c->leaf_b;
        Q.pts = c->pts;
        Q.qcol = qcol;
        Q.stats = st;
        Q.next_query = counters;
        Q.fail_list = fail;
        Q.fail_count = counters + 1;
        Q.scratch = wscratch;
        Q.num_waves = kNoAllocWaves;
        Q.tile_order = 1;
        if (y0 > 0) HIPCHK(c, launch_fill_u32(counters, 2, 0u, s));
        tlr_status r = run_plain(c, Q, nullptr, nullptr, p->collect_stats != 0, p->path);
        if (r != TLR_OK) return r;
        HIPCHK(c, launch_combine(qcol, W, bh, M, out + (size_t)y0 * W * 4, s));
    }
    HIPCHK(c, hipEventRecord(c->ev[3], s));
    unsigned long long hs[kNumStats];
    HIPCHK(c, hipMemcpyAsync(hs, st, sizeof(hs), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->leaf_nodes = (double)hs[ST_LEAF_NODES];
        stats->samples_in_leaf_nodes = (double)hs[ST_SAMPLES_IN_LEAVES];
        stats->samples_within_1r = (double)hs[ST_WITHIN_1R];
        stats->samples_fetched = (double)hs[ST_FETCHED];
        stats->surfaces = (double)hs[ST_SURFACES];
        stats->num_2r = (double)hs[ST_2R];
        stats->surfaces_merged = (double)hs[ST_MERGED];
        stats->at_least_one = (double)hs[ST_AT_LEAST_ONE];
        stats->count = (double)hs[ST_COUNT];
        stats->overflow = (double)hs[ST_OVERFLOW];
        stats->fast_leaf_nodes = (double)hs[ST_FAST_LEAVES];
        stats->fast_samples_in_leaf_nodes = (double)hs[ST_FAST_IN_LEAVES];
        stats->fast_samples_within_1r = (double)hs[ST_FAST_KEPT];
        float ms = 0;
        if (hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) stats->build_ms = ms;
        if (hipEventElapsedTime(&ms, c->ev[2], c->ev[3]) == hipSuccess) stats->query_ms = ms;
        stats->total_ms = stats->build_ms + stats->query_ms;
        stats->num_leaves = c->L;
        stats->num_samples = c->nb;
    }
    if (hs[ST_OVERFLOW])
        return c->fail(TLR_E_CAPACITY, "%llu output samples exceeded the per-query capacity",
                       (unsigned long long)hs[ST_OVERFLOW]);
    return TLR_OK;
}

tlr_status tlr_sync(tlr_ctx* c) {
    if (!c) return TLR_E_INVALID;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TLR_OK;
}

tlr_status tlr_get_build_info(tlr_ctx* c, tlr_build_info* o) {
    if (!c || !o) return TLR_E_INVALID;
    if (!c->built) return c->fail(TLR_E_STATE, "not built");
    memset(o, 0, sizeof(*o));
    o->num_input = c->N;
    o->num_kept = (int)c->gi.num_kept;
    o->num_bucketed = c->nb;
    o->spp_pow2 = c->spp_pow2;
    o->coc_adj[0] = c->cc0;
    o->coc_adj[1] = c->cc1;
    o->bbmin[0] = c->gi.bbmin[0]; o->bbmin[1] = c->gi.bbmin[1];
    o->bbmax[0] = c->gi.bbmax[0]; o->bbmax[1] = c->gi.bbmax[1];
    o->scale = c->gi.scale;
    o->bbmin_int[0] = c->gi.bbmin_i[0]; o->bbmin_int[1] = c->gi.bbmin_i[1];
    o->grid_w = c->gi.gw;
    o->grid_h = c->gi.gh;
    o->max_hits = c->gi.max_hits;
    o->num_initial_leaves = c->nil;
    o->num_leaves = c->L;
    o->num_valid_leaves = c->nv;
    return TLR_OK;
}

tlr_status tlr_copy_reprojected(tlr_ctx* c, float* x, float* y, float* w, float* t) {
    if (!c || !c->built) return TLR_E_STATE;
    const size_t n = (size_t)c->N * 4;
    if (x) HIPCHK(c, hipMemcpy(x, c->rx, n, hipMemcpyDeviceToHost));
    if (y) HIPCHK(c, hipMemcpy(y, c->ry, n, hipMemcpyDeviceToHost));
    if (w) HIPCHK(c, hipMemcpy(w, c->rw, n, hipMemcpyDeviceToHost));
    if (t) HIPCHK(c, hipMemcpy(t, c->rt, n, hipMemcpyDeviceToHost));
    return TLR_OK;
}

tlr_status tlr_copy_layout(tlr_ctx* c, int32_t* src) {
    if (!c || !c->built || !src) return TLR_E_STATE;
    if (c->nb) HIPCHK(c, hipMemcpy(src, c->cand, (size_t)c->nb * 4, hipMemcpyDeviceToHost));
    return TLR_OK;
}

tlr_status tlr_copy_initial_leaves(tlr_ctx* c, int32_t* out6) {
    if (!c || !c->built || !out6) return TLR_E_STATE;
    std::vector<int4> r(c->nil);
    std::vector<int2> oc(c->nil);
    if (c->nil) {
        HIPCHK(c, hipMemcpy(r.data(), c->ileaf_rect, sizeof(int4) * c->nil, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(oc.data(), c->ileaf_oc, sizeof(int2) * c->nil, hipMemcpyDeviceToHost));
    }
    for (int k = 0; k < c->nil; k++) {
        out6[6 * k + 0] = r[k].x; out6[6 * k + 1] = r[k].y;
        out6[6 * k + 2] = r[k].z; out6[6 * k + 3] = r[k].w;
        out6[6 * k + 4] = oc[k].x; out6[6 * k + 5] = oc[k].y;
    }
    return TLR_OK;
}

tlr_status tlr_copy_hierarchy(tlr_ctx* c, int32_t* order, int32_t* ch2, float* b14) {
    if (!c || !c->built) return TLR_E_STATE;
    const int n = c->nv;
    if (n == 0) return TLR_OK;
    std::vector<uint32_t> ids(n);
    HIPCHK(c, hipMemcpy(ids.data(), c->tree_ids, 4 * (size_t)n, hipMemcpyDeviceToHost));
    if (order)
        for (int k = 0; k < n; k++) order[k] = (int32_t)ids[k];
    if (n < 2) return TLR_OK;
    std::vector<int> c0(n - 1), c1(n - 1);
    std::vector<Bounds> nb(n - 1);
    HIPCHK(c, hipMemcpy(c0.data(), c->tree_c0, 4 * (size_t)(n - 1), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(c1.data(), c->tree_c1, 4 * (size_t)(n - 1), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(nb.data(), c->tree_b, sizeof(Bounds) * (size_t)(n - 1), hipMemcpyDeviceToHost));
    auto enc = [&](int x) { return x >= n - 1 ? (int32_t)(ids[x - (n - 1)] | 0x80000000u) : (int32_t)x; };
    for (int g = 0; g < n - 1; g++) {
        if (ch2) {
            ch2[2 * g] = enc(c0[g]);
            ch2[2 * g + 1] = enc(c1[g]);
        }
        if (b14) {
            memcpy(b14 + 14 * (size_t)g, nb[g].p, 12 * 4);
            b14[14 * (size_t)g + 12] = nb[g].ts0;
            b14[14 * (size_t)g + 13] = nb[g].ts1;
        }
    }
    return TLR_OK;
}

tlr_status tlr_copy_leaves(tlr_ctx* c, int32_t* s01, float* b14) {
    if (!c || !c->built) return TLR_E_STATE;
    std::vector<LeafRec> lr(c->L);
    std::vector<Bounds> lb(c->L);
    if (c->L) {
        HIPCHK(c, hipMemcpy(lr.data(), c->leaves, sizeof(LeafRec) * c->L, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(lb.data(), c->leaf_b, sizeof(Bounds) * c->L, hipMemcpyDeviceToHost));
    }
    for (int k = 0; k < c->L; k++) {
        if (s01) {
            s01[2 * k] = lr[k].s0;
            s01[2 * k + 1] = lr[k].s1;
        }
        if (b14) {
            memcpy(b14 + 14 * (size_t)k, lb[k].p, 12 * 4);
            b14[14 * (size_t)k + 12] = lb[k].ts0;
            b14[14 * (size_t)k + 13] = lb[k].ts1;
        }
    }
    return TLR_OK;
}

tlr_status tlr_copy_query_colors(tlr_ctx* c, float* out4) {
    if (!c || !c->qcol || !out4) return TLR_E_STATE;
    HIPCHK(c, hipMemcpy(out4, c->qcol, c->qcol_n * sizeof(float4), hipMemcpyDeviceToHost));
    return TLR_OK;
}

// ---- Shadows ----------------------------------------------------------------

tlr_status tlr_set_shadow_inputs(tlr_ctx* c, const float* wg, const float* density, int64_t n) {
    if (!c) return TLR_E_INVALID;
    if (!c->has_samples) return c->fail(TLR_E_STATE, "no samples");
    if (n != c->N) return c->fail(TLR_E_INVALID, "shadow inputs for %lld samples, the buffer has %d", (long long)n, c->N);
    HIPCHK(c, hipSetDevice(c->device));
    if (c->d_wgd) HIPCHK(c, hipFree(c->d_wgd));
    c->d_wgd = nullptr;
    c->wgd_n = 0;
    std::vector<float4> tmp((size_t)std::max(n, (int64_t)1));
    for (int64_t k = 0; k < n; k++)
        tmp[k] = make_float4(wg ? wg[2 * k] : 0.0f, wg ? wg[2 * k + 1] : 0.0f, density ? density[k] : 1.0f, 0.0f);
    HIPCHK(c, hipMalloc(&c->d_wgd, sizeof(float4) * tmp.size()));
    HIPCHK(c, hipMemcpy(c->d_wgd, tmp.data(), sizeof(float4) * tmp.size(), hipMemcpyHostToDevice));
    c->wgd_n = (int)n;
    c->pext_valid = false;
    return TLR_OK;
}

// Tree-order (wgx, wgy, density, 0) of the current build.
static tlr_status ensure_pext(tlr_ctx* c) {
    if (c->pext_valid) return TLR_OK;
    const size_t need = (size_t)std::max(c->nb, 1);
    if (c->pext_cap < need) {
        if (c->pext) HIPCHK(c, hipFree(c->pext));
        c->pext = nullptr;
        c->pext_cap = 0;
        HIPCHK(c, hipMalloc(&c->pext, sizeof(float4) * need));
        c->pext_cap = need;
    }
    const float4* in = (c->wgd_n == c->N && c->d_wgd) ? c->d_wgd : nullptr;
    HIPCHK(c, launch_gather_ext(reinterpret_cast<const int*>(c->cand), c->nb, in, c->pext, c->stream));
    c->pext_valid = true;
    return TLR_OK;
}

static void stats_from(const unsigned long long* hs, tlr_stats* stats) {
    memset(stats, 0, sizeof(*stats));
    stats->leaf_nodes = (double)hs[ST_LEAF_NODES];
    stats->samples_in_leaf_nodes = (double)hs[ST_SAMPLES_IN_LEAVES];
    stats->samples_within_1r = (double)hs[ST_WITHIN_1R];
    stats->samples_fetched = (double)hs[ST_FETCHED];
    stats->surfaces = (double)hs[ST_SURFACES];
    stats->num_2r = (double)hs[ST_2R];
    stats->surfaces_merged = (double)hs[ST_MERGED];
    stats->at_least_one = (double)hs[ST_AT_LEAST_ONE];
    stats->count = (double)hs[ST_COUNT];
    stats->overflow = (double)hs[ST_OVERFLOW];
}

// QueryParams of a built context's tree for the general kernel over an
// explicit list (or the output pattern), with the context's wave scratch.
static void tree_params(const tlr_ctx* c, QueryParams& Q) {
    Q.spp_pow2 = c->spp_pow2;
    Q.cc0 = c->cc0; Q.cc1 = c->cc1;
    Q.n_nodes = c->nv > 0 ? 2 * c->nv - 1 : 0;
    Q.root = c->nv > 1 ? c->nv / 2 - 1 : 0;
    Q.trav = c->trav;
    Q.pairs = c->pairs;
    Q.leaves = c->leaves;
    Q.leaf_b = c->leaf_b;
    Q.pts = c->pts;
    Q.pext = c->pext;
}

tlr_status tlr_reconstruct_queries(tlr_ctx* c, const tlr_recon_params* p, const float* queries, int64_t n, int shadow,
                                   float* rgba_out, float* ext_out, tlr_stats* stats) {
    if (!c) return TLR_E_INVALID;
    if (!c->built) return c->fail(TLR_E_STATE, "not built");
    if (n < 0 || n >= ((int64_t)1 << 31)) return c->fail(TLR_E_INVALID, "query count %lld", (long long)n);
    if (n > 0 && (!queries || !rgba_out)) return c->fail(TLR_E_INVALID, "null queries or output");
    if (n == 0) return TLR_OK;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    tlr_status st = ensure_pext(c);
    if (st != TLR_OK) return st;
    if (c->general_wpc == 0) c->general_wpc = general_waves_per_cu();
    const int waves = std::max(4, c->prop.multiProcessorCount * c->general_wpc);
    const size_t nq = (size_t)n;
    HIPCHK(c, c->sarena.reserve(nq * (5 * sizeof(float4) + 8) + (size_t)waves * kWaveScratchWords * 4 + 9 * 256));
    c->sarena.reset();
    float4* qin = c->sarena.alloc<float4>(2 * nq);
    float4* qcol = c->sarena.alloc<float4>(nq);
    float4* qext = c->sarena.alloc<float4>(2 * nq);
    unsigned* hand = c->sarena.alloc<unsigned>(nq);
    unsigned* hand2 = c->sarena.alloc<unsigned>(nq);
    float* scratch = c->sarena.alloc<float>((size_t)waves * kWaveScratchWords);
    if (c->sarena.failed) return c->fail(TLR_E_HIP, "query arena exhausted");
    HIPCHK(c, hipMemcpyAsync(qin, queries, nq * 2 * sizeof(float4), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemsetAsync(c->d_stats, 0, sizeof(unsigned long long) * kNumStats, s));
    QueryParams Q;
    memset(&Q, 0, sizeof(Q));
    tree_params(c, Q);
    Q.W = c->h.width; Q.H = c->h.height; Q.M = 1;
    Q.lens_gauss = p ? p->lens_gaussian : 0;
    Q.time_gauss = p ? p->time_gaussian : 0;
    Q.collect_stats = 1;
    Q.num_queries = (unsigned)nq;
    Q.qcol = qcol;
    Q.stats = c->d_stats;
    Q.next_query = c->d_counter;
    Q.scratch = scratch;
    Q.num_waves = waves;
    Q.mode = shadow ? 2 : 1;
    Q.qin = qin;
    Q.qext = qext;
    HIPCHK(c, hipEventRecord(c->ev[2], s));
    // the 1R lane passes complete what they can; the general kernel (one
    // wave per query) takes the rest
    st = run_mode_passes(c, Q, hand, c->d_fail_count, hand2, c->d_fail_count2, c->d_wide_next, c->d_counter, true);
    if (st != TLR_OK) return st;
    HIPCHK(c, hipEventRecord(c->ev[3], s));
    HIPCHK(c, hipMemcpyAsync(rgba_out, qcol, nq * sizeof(float4), hipMemcpyDeviceToHost, s));
    if (ext_out) HIPCHK(c, hipMemcpyAsync(ext_out, qext, nq * 2 * sizeof(float4), hipMemcpyDeviceToHost, s));
    unsigned long long hs[kNumStats];
    HIPCHK(c, hipMemcpyAsync(hs, c->d_stats, sizeof(hs), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (stats) {
        stats_from(hs, stats);
        float ms = 0;
        if (hipEventElapsedTime(&ms, c->ev[2], c->ev[3]) == hipSuccess) stats->query_ms = stats->general_ms = ms;
        stats->general_queries = (double)nq;
        stats->total_ms = stats->query_ms;
    }
    if (hs[ST_OVERFLOW])
        return c->fail(TLR_E_CAPACITY, "%llu queries exceeded the per-query capacity", (unsigned long long)hs[ST_OVERFLOW]);
    return TLR_OK;
}

tlr_status tlr_reconstruct_dof_motion_shadows(tlr_ctx* c, tlr_ctx* light, const tlr_recon_params* p,
                                              const tlr_shadow_params* sp, float* out, int out_on_device,
                                              tlr_stats* stats) {
    if (!c || !light || !p || !sp) return TLR_E_INVALID;
    if (out_on_device && !out) return c->fail(TLR_E_INVALID, "out_on_device set with a NULL rgba_out");
    if (!c->built || !light->built) return c->fail(TLR_E_STATE, "camera or light not built");
    if (c->device != light->device) return c->fail(TLR_E_INVALID, "camera and light contexts on different GPUs");
    if (c->h.width != light->h.width || c->h.height != light->h.height || sp->width != c->h.width ||
        sp->height != c->h.height)
        return c->fail(TLR_E_INVALID, "camera, light and shadow parameters differ in size");
    HIPCHK(c, hipSetDevice(c->device));
    // the light's tree (and its tree-order extras) come from its own stream
    tlr_status st = ensure_pext(light);
    if (st != TLR_OK) return st;
    HIPCHK(c, hipStreamSynchronize(light->stream));
    st = ensure_pext(c);
    if (st != TLR_OK) return st;
    tlr_recon_params pp = *p;
    pp.gate = 0;
    pp.path = 1;
    pp.debug_image = 0;
    QueryParams Q;
    st = prepare_queries(c, &pp, false, Q);
    if (st != TLR_OK) return st;
    hipStream_t s = c->stream;
    const int W = c->h.width, H = c->h.height, M = c->last_M;
    const size_t nq = c->qcol_n;
    std::vector<float> luv;
    light_sample_positions(M, luv);
    HIPCHK(c, c->sarena.reserve(nq * (6 * sizeof(float4) + 8) + luv.size() * 4 + 8 * 256));
    c->sarena.reset();
    float4* qext = c->sarena.alloc<float4>(2 * nq);
    float4* lq = c->sarena.alloc<float4>(2 * nq);
    float4* xyw = c->sarena.alloc<float4>(nq);
    float4* scol = c->sarena.alloc<float4>(nq);
    int* flag = c->sarena.alloc<int>(nq);
    unsigned* list = c->sarena.alloc<unsigned>(nq);
    float* d_luv = c->sarena.alloc<float>(luv.size());
    unsigned* cnt = c->sarena.alloc<unsigned>(64);
    if (c->sarena.failed) return c->fail(TLR_E_HIP, "shadow arena exhausted");
    HIPCHK(c, hipMemcpyAsync(d_luv, luv.data(), luv.size() * 4, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemsetAsync(cnt, 0, 64 * sizeof(unsigned), s));
    ShadowMats S;
    memcpy(S.c2l, sp->c2l, 64);
    memcpy(S.zfromw_ndc, sp->zfromw_ndc, 64);
    memcpy(S.inv_projection, sp->inv_projection, 64);
    memcpy(S.shadow2_inv_proj, sp->shadow2_inv_proj, 64);
    memcpy(S.shadow2_inv_t, sp->shadow2_inv_t, 64);
    memcpy(S.world_inv_t, sp->world_inv_t, 64);
    S.cam_coc[0] = sp->camera_coc[0]; S.cam_coc[1] = sp->camera_coc[1];
    S.light_coc[0] = sp->light_coc[0]; S.light_coc[1] = sp->light_coc[1];
    S.cam_far = sp->camera_far;
    S.cw = W; S.ch = H;
    HIPCHK(c, hipEventRecord(c->ev[2], s));
    // 1. the camera's output samples, with the nearest sample's xy, w, wg
    //    and density (the 1R lane kernel over the pattern, the general kernel
    //    over its hand-offs)
    Q.mode = 1;
    Q.pext = c->pext;
    Q.qext = qext;
    st = run_mode_passes(c, Q, c->fail_list, c->d_fail_count, c->fail_list2, c->d_fail_count2, c->d_wide_next,
                         c->d_counter, Q.collect_stats != 0);
    if (st != TLR_OK) return st;
    HIPCHK(c, hipMemcpyAsync(cnt + 40, c->d_fail_count, sizeof(unsigned), hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipMemcpyAsync(cnt + 41, c->d_fail_count2, sizeof(unsigned), hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipEventRecord(c->ev[5], s));
    // 2. shadow rays into the light's window
    ShadowRays R;
    R.W = W; R.H = H; R.M = M;
    R.nq = (unsigned)nq;
    R.pattern = c->d_pattern;
    R.light_uv = d_luv;
    R.qcol = c->qcol;
    R.qext = qext;
    R.S = S;
    R.lq = lq;
    R.xyw = xyw;
    R.flag = flag;
    R.list = list;
    R.count = cnt;
    HIPCHK(c, launch_shadow_rays(R, s));
    // 3. the light queries on the light's tree (shadow mode)
    QueryParams L;
    memset(&L, 0, sizeof(L));
    tree_params(light, L);
    L.W = W; L.H = H; L.M = M;
    L.collect_stats = Q.collect_stats;
    L.num_queries = (unsigned)nq;
    L.qcol = scol;
    L.stats = c->d_stats;
    L.next_query = cnt + 32;
    L.scratch = Q.scratch;
    L.num_waves = Q.num_waves;
    L.mode = 2;
    L.qin = lq;
    // the listed light queries through the same passes (the camera pass's
    // hand-off lists are free again: the stream has finished with them)
    L.fast_list = 1;
    L.qlist = list;
    L.qcount = cnt;
    st = run_mode_passes(c, L, c->fail_list, c->d_fail_count, c->fail_list2, c->d_fail_count2, c->d_wide_next,
                         cnt + 32, L.collect_stats != 0);
    if (st != TLR_OK) return st;
    HIPCHK(c, hipMemcpyAsync(cnt + 42, c->d_fail_count, sizeof(unsigned), hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipMemcpyAsync(cnt + 43, c->d_fail_count2, sizeof(unsigned), hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipEventRecord(c->ev[6], s));
    // 4. shading, per pixel in output-sample order
    ShadowShade D;
    D.W = W; D.H = H; D.M = M;
    D.qcol = c->qcol;
    D.qext = qext;
    D.scol = scol;
    D.xyw = xyw;
    D.flag = flag;
    D.S = S;
    D.out = out_on_device ? out : c->d_img;
    HIPCHK(c, hipEventRecord(c->ev[3], s));
    HIPCHK(c, launch_shadow_shade(D, s));
    HIPCHK(c, hipEventRecord(c->ev[4], s));
    if (!out_on_device && out) HIPCHK(c, hipMemcpyAsync(out, c->d_img, (size_t)W * H * 16, hipMemcpyDeviceToHost, s));
    unsigned long long hs[kNumStats];
    unsigned nl = 0, hand[4] = {0, 0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(hs, c->d_stats, sizeof(hs), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(&nl, cnt, sizeof(nl), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(hand, cnt + 40, sizeof(hand), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (stats) {
        stats_from(hs, stats);
        float ms = 0;
        if (hipEventElapsedTime(&ms, c->ev[2], c->ev[3]) == hipSuccess) stats->query_ms = ms;
        if (hipEventElapsedTime(&ms, c->ev[3], c->ev[4]) == hipSuccess) stats->combine_ms = ms;
        if (hipEventElapsedTime(&ms, c->ev[2], c->ev[5]) == hipSuccess) stats->general_ms = ms;
        if (hipEventElapsedTime(&ms, c->ev[5], c->ev[6]) == hipSuccess) stats->fast_ms = ms;
        stats->light_queries = (double)nl;
        stats->wide_queries = (double)hand[0] + hand[2];
        stats->general_queries = (double)hand[1] + hand[3];
        stats->total_ms = stats->query_ms + stats->combine_ms;
    }
    if (hs[ST_OVERFLOW])
        return c->fail(TLR_E_CAPACITY, "%llu output samples exceeded the per-query capacity",
                       (unsigned long long)hs[ST_OVERFLOW]);
    return TLR_OK;
}

tlr_status tlr_copy_debug_image(tlr_ctx* c, float* out4) {
    if (!c || !out4) return TLR_E_INVALID;
    if (!c->has_debug) return c->fail(TLR_E_STATE, "the last reconstruction did not make a debug image");
    HIPCHK(c, hipMemcpy(out4, c->d_debug, (size_t)c->h.width * c->h.height * 16, hipMemcpyDeviceToHost));
    return TLR_OK;
}

tlr_status tlr_copy_query_gather(tlr_ctx* c, const float* o5, int n, float radius, int leaf_cap, int32_t* leaves,
                                 int32_t* n_leaves, int sample_cap, int32_t* samples, int32_t* surface_of,
                                 int32_t* n_samples) {
    if (!c || (!o5 && n > 0) || n < 0 || leaf_cap < 0 || sample_cap < 0 || !n_leaves || !n_samples)
        return TLR_E_INVALID;
    if ((leaf_cap > 0 && !leaves) || (sample_cap > 0 && (!samples || !surface_of))) return TLR_E_INVALID;
    if (!(radius > 0.0f)) return c->fail(TLR_E_INVALID, "gather radius must be positive");
    if (!c->built) return c->fail(TLR_E_STATE, "gather before tlr_build");
    if (n == 0) return TLR_OK;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int waves = std::min(n, 256);
    const int wv = (waves + 3) / 4 * 4;
    char* buf = nullptr;
    const size_t b_o5 = (size_t)n * 5 * 4, b_l = (size_t)n * leaf_cap * 4, b_s = (size_t)n * sample_cap * 4,
                 b_c = (size_t)n * 4, b_w = (size_t)wv * kWaveScratchWords * 4;
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t total = up(b_o5) + up(b_l) + 2 * up(b_s) + 2 * up(b_c) + up(b_w);
    HIPCHK(c, hipMalloc(&buf, total));
    char* q = buf;
    auto take = [&](size_t b) { char* r = q; q += up(b); return r; };
    float* d_o5 = reinterpret_cast<float*>(take(b_o5));
    int* d_l = reinterpret_cast<int*>(take(b_l));
    int* d_s = reinterpret_cast<int*>(take(b_s));
    int* d_f = reinterpret_cast<int*>(take(b_s));
    int* d_nl = reinterpret_cast<int*>(take(b_c));
    int* d_ns = reinterpret_cast<int*>(take(b_c));
    float* d_w = reinterpret_cast<float*>(take(b_w));
    QueryParams Q;
    memset(&Q, 0, sizeof(Q));
    Q.W = c->h.width; Q.H = c->h.height; Q.M = 1;
    Q.spp_pow2 = c->spp_pow2;
    Q.cc0 = c->cc0; Q.cc1 = c->cc1;
    Q.n_nodes = c->nv > 0 ? 2 * c->nv - 1 : 0;
    Q.root = c->nv > 1 ? c->nv / 2 - 1 : 0;
    Q.trav = c->trav;
    Q.pairs = c->pairs;
    Q.leaves = c->leaves;
    Q.leaf_b = c->leaf_b;
    Q.pts = c->pts;
    Q.scratch = d_w;
    Q.num_waves = wv;
    hipError_t e = hipMemcpyAsync(d_o5, o5, b_o5, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = launch_query_gather(Q, d_o5, n, radius, leaf_cap, d_l, d_nl, sample_cap, d_s, d_f, d_ns, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipMemcpy(n_leaves, d_nl, b_c, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(n_samples, d_ns, b_c, hipMemcpyDeviceToHost);
    if (e == hipSuccess && leaf_cap) e = hipMemcpy(leaves, d_l, b_l, hipMemcpyDeviceToHost);
    if (e == hipSuccess && sample_cap) e = hipMemcpy(samples, d_s, b_s, hipMemcpyDeviceToHost);
    if (e == hipSuccess && sample_cap) e = hipMemcpy(surface_of, d_f, b_s, hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    if (e != hipSuccess) return c->fail(TLR_E_HIP, "query gather: %s", hipGetErrorString(e));
    return TLR_OK;
}

tlr_status tlr_copy_query_stages(tlr_ctx* c, int32_t* out) {
    if (!c || !out) return TLR_E_INVALID;
    if (!c->ref_stage) return c->fail(TLR_E_STATE, "the last reconstruction did not use path 2");
    HIPCHK(c, hipMemcpy(out, c->ref_stage, c->qcol_n * sizeof(int32_t), hipMemcpyDeviceToHost));
    return TLR_OK;
}

tlr_status tlr_calibrate_traffic(tlr_ctx* c, int kind, int64_t bytes, int64_t* read, int64_t* written) {
    if (!c) return TLR_E_INVALID;
    if (kind < 0 || kind > 2) return c->fail(TLR_E_INVALID, "calibration kind %d: expected 0, 1 or 2", kind);
    if (bytes < ((int64_t)1 << 20)) return c->fail(TLR_E_INVALID, "calibration size below 1 MiB");
    HIPCHK(c, hipSetDevice(c->device));
    int64_t b = (int64_t)1 << 20;
    while (b * 2 <= bytes) b *= 2;
    const int rec = kind == 0 ? 1 : (kind == 1 ? 2 : 4);     // float4 per record
    const size_t n = (size_t)b / (16 * (size_t)rec);          // records (a power of two)
    float4 *src = nullptr, *dst = nullptr;
    HIPCHK(c, hipMalloc(&src, (size_t)b));
    hipError_t e = hipMalloc(&dst, n * sizeof(float4));
    if (e == hipSuccess) e = hipMemsetAsync(src, 0, (size_t)b, c->stream);
    if (e == hipSuccess) e = launch_calibrate(kind, src, dst, n, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (e != hipSuccess) return c->fail(TLR_E_HIP, "calibration: %s", hipGetErrorString(e));
    if (read) *read = b;
    if (written) *written = (int64_t)(n * sizeof(float4));
    return TLR_OK;
}

int tlr_output_pattern(int out_spp, int override_uvt, const float* uvt, float* out5) {
    std::vector<float> pat;
    const float zero[3] = {0, 0, 0};
    const int M = make_output_pattern(out_spp, override_uvt != 0, uvt ? uvt : zero, pat);
    if (out5) memcpy(out5, pat.data(), pat.size() * 4);
    return M;
}


// ---------------------------------------------------------------------------
// Learned gate
// ---------------------------------------------------------------------------
tlr_status tlr_gate_set_params(tlr_ctx* c, const float* params) {
    if (!c || !params) return TLR_E_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpy(c->gate_p, params, sizeof(float) * kGateParams, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(c->gate_m, 0, sizeof(float) * kGateParams));
    HIPCHK(c, hipMemset(c->gate_v, 0, sizeof(float) * kGateParams));
    c->gate_step = 0;
    c->has_gate = true;
    return TLR_OK;
}

tlr_status tlr_gate_get_params(tlr_ctx* c, float* params) {
    if (!c || !params) return TLR_E_INVALID;
    if (!c->has_gate) return c->fail(TLR_E_STATE, "no gate parameters");
    HIPCHK(c, hipMemcpy(params, c->gate_p, sizeof(float) * kGateParams, hipMemcpyDeviceToHost));
    return TLR_OK;
}

tlr_status tlr_gate_set_threshold(tlr_ctx* c, float thr) {
    if (!c) return TLR_E_INVALID;
    c->gate_thr = thr;
    return TLR_OK;
}

tlr_status tlr_gate_forward(tlr_ctx* c, const float* X, int64_t n, float* logits) {
    if (!c || !X || !logits || n < 0) return TLR_E_INVALID;
    if (!c->has_gate) return c->fail(TLR_E_STATE, "no gate parameters");
    HIPCHK(c, hipSetDevice(c->device));
    float *dX = nullptr, *dz = nullptr;
    HIPCHK(c, hipMalloc(&dX, sizeof(float) * kGateIn * std::max<int64_t>(n, 1)));
    HIPCHK(c, hipMalloc(&dz, sizeof(float) * std::max<int64_t>(n, 1)));
    hipError_t e = hipMemcpy(dX, X, sizeof(float) * kGateIn * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_gate_forward(dX, (int)n, c->gate_p, dz, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(logits, dz, sizeof(float) * n, hipMemcpyDeviceToHost);
    (void)hipFree(dX);
    (void)hipFree(dz);
    if (e != hipSuccess) return c->fail(TLR_E_HIP, "gate forward: %s", hipGetErrorString(e));
    return TLR_OK;
}

tlr_status tlr_gate_grad(tlr_ctx* c, const float* X, const float* y, int64_t n, float pos_w, float* grad,
                         float* loss) {
    if (!c || !X || !y || !grad || n <= 0) return TLR_E_INVALID;
    if (!c->has_gate) return c->fail(TLR_E_STATE, "no gate parameters");
    HIPCHK(c, hipSetDevice(c->device));
    float *dX = nullptr, *dy = nullptr;
    HIPCHK(c, hipMalloc(&dX, sizeof(float) * kGateIn * n));
    HIPCHK(c, hipMalloc(&dy, sizeof(float) * n));
    hipError_t e = hipMemcpy(dX, X, sizeof(float) * kGateIn * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dy, y, sizeof(float) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(c->gate_g, 0, sizeof(float) * kGateParams);
    if (e == hipSuccess) e = hipMemset(c->gate_loss, 0, sizeof(float));
    if (e == hipSuccess) e = launch_gate_grad(dX, dy, nullptr, (int)n, c->gate_p, c->gate_g, c->gate_loss, pos_w,
                                              c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(grad, c->gate_g, sizeof(float) * kGateParams, hipMemcpyDeviceToHost);
    if (e == hipSuccess && loss) e = hipMemcpy(loss, c->gate_loss, sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(dX);
    (void)hipFree(dy);
    if (e != hipSuccess) return c->fail(TLR_E_HIP, "gate grad: %s", hipGetErrorString(e));
    return TLR_OK;
}

// Training data of the gate, over the current build: the gated 1R lane
// kernel (provisional first-surface colours and features of its hand-offs),
// the provisional colours kept aside, the general kernel over every hand-off
// (the plain method's result), then each entry's label: 1 when its
// provisional colour is within kGateLabelTau of the plain one in every
// channel.  *n_entries: hand-off entries (rows).
constexpr float kGateLabelTau = 1e-3f;

static tlr_status gate_pass(tlr_ctx* c, const tlr_recon_params* p, unsigned* n_entries) {
    if (p->path != 0) return c->fail(TLR_E_INVALID, "gate data come from the plain method (path 0)");
    tlr_recon_params pp = *p;
    pp.collect_stats = 0;
    pp.gate = 0;
    QueryParams Q;
    tlr_status st = prepare_queries(c, &pp, true, Q);
    if (st != TLR_OK) return st;
    hipStream_t s = c->stream;
    const size_t nq = c->qcol_n;
    QueryParams F = Q;
    F.features = c->features;
    HIPCHK(c, launch_query_fast_gate(F, s));
    HIPCHK(c, launch_gate_prov(c->qcol, c->fail_list, c->d_fail_count, nq, c->prov, s));
    QueryParams G = Q;
    G.qlist = c->fail_list;
    G.qcount = c->d_fail_count;
    HIPCHK(c, launch_query_general(G, s));
    HIPCHK(c, launch_gate_label(c->qcol, c->fail_list, c->d_fail_count, nq, c->features, c->prov, kGateLabelTau,
                                c->labels, s));
    *n_entries = read_u32(c, c->d_fail_count);
    unsigned long long ovf = 0;
    HIPCHK(c, hipMemcpy(&ovf, c->d_stats + ST_OVERFLOW, sizeof(ovf), hipMemcpyDeviceToHost));
    if (ovf) return c->fail(TLR_E_CAPACITY, "%llu output samples exceeded the per-query capacity", ovf);
    return TLR_OK;
}

tlr_status tlr_gate_collect(tlr_ctx* c, const tlr_recon_params* p, int stride, int64_t cap, float* X, float* y,
                            float* prov, float* plain, uint32_t* query, int64_t* n) {
    if (!c || !p || stride <= 0 || cap < 0) return TLR_E_INVALID;
    unsigned ne = 0;
    tlr_status st = gate_pass(c, p, &ne);
    if (st != TLR_OK) return st;
    hipStream_t s = c->stream;
    const size_t rows_all = ((size_t)ne + stride - 1) / stride;
    const size_t rows = std::min(rows_all, (size_t)cap);
    if (n) *n = (int64_t)rows_all;
    if (rows == 0) return TLR_OK;
    hipError_t e = hipSuccess;
    // every stride-th entry, compacted on the device, then one copy per
    // array (a pitched copy of millions of short rows is slow); the plain
    // colours are gathered per entry from qcol first
    void* tmp = nullptr;
    float4* plain_e = nullptr;
    e = hipMalloc(&tmp, rows * sizeof(float) * kGateFeatures);
    if (e == hipSuccess && plain) e = hipMalloc(&plain_e, (size_t)ne * sizeof(float4));
    if (e == hipSuccess && plain) e = launch_gate_prov(c->qcol, c->fail_list, c->d_fail_count, ne, plain_e, s);
    auto take = [&](const void* src, void* dst, int words) {
        if (!dst || e != hipSuccess) return;
        e = launch_take_rows(src, tmp, rows, words, stride, s);
        if (e == hipSuccess) e = hipMemcpyAsync(dst, tmp, rows * 4 * (size_t)words, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    };
    take(c->features, X, kGateFeatures);
    take(c->labels, y, 1);
    take(c->prov, prov, 4);
    take(plain_e, plain, 4);
    take(c->fail_list, query, 1);
    if (tmp) (void)hipFree(tmp);
    if (plain_e) (void)hipFree(plain_e);
    if (e != hipSuccess) return c->fail(TLR_E_HIP, "gate collect: %s", hipGetErrorString(e));
    if (query)
        for (size_t r = 0; r < rows; r++) query[r] &= ~kNeeds2R;
    return TLR_OK;
}

tlr_status tlr_gate_append_training(tlr_ctx* c, const tlr_recon_params* p, int stride) {
    if (!c || !p || stride <= 0) return TLR_E_INVALID;
    unsigned ne = 0;
    tlr_status st = gate_pass(c, p, &ne);
    if (st != TLR_OK) return st;
    const int64_t rows = (int64_t)(((size_t)ne + stride - 1) / stride);
    if (rows == 0) return TLR_OK;
    if (c->train_n + rows > c->train_cap) {
        const int64_t cap = std::max<int64_t>(c->train_n + rows, c->train_cap * 2);
        float *nX = nullptr, *ny = nullptr;
        HIPCHK(c, hipMalloc(&nX, sizeof(float) * kGateFeatures * cap));
        HIPCHK(c, hipMalloc(&ny, sizeof(float) * cap));
        if (c->train_n) {
            HIPCHK(c, hipMemcpyAsync(nX, c->train_X, sizeof(float) * kGateFeatures * c->train_n,
                                     hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(ny, c->train_y, sizeof(float) * c->train_n, hipMemcpyDeviceToDevice, c->stream));
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->train_X) (void)hipFree(c->train_X);
        if (c->train_y) (void)hipFree(c->train_y);
        c->train_X = nX;
        c->train_y = ny;
        c->train_cap = cap;
    }
    hipStream_t s = c->stream;
    HIPCHK(c, launch_take_rows(c->features, c->train_X + kGateFeatures * c->train_n, (size_t)rows, kGateFeatures,
                               stride, s));
    HIPCHK(c, launch_take_rows(c->labels, c->train_y + c->train_n, (size_t)rows, 1, stride, s));
    HIPCHK(c, hipStreamSynchronize(s));
    c->train_n += rows;
    return TLR_OK;
}

tlr_status tlr_gate_clear_training(tlr_ctx* c) {
    if (!c) return TLR_E_INVALID;
    c->train_n = 0;
    return TLR_OK;
}

int64_t tlr_gate_training_size(tlr_ctx* c) { return c ? c->train_n : 0; }

tlr_status tlr_gate_train(tlr_ctx* c, int epochs, int batch, float lr, float pos_w, uint64_t seed,
                          float* epoch_loss) {
    if (!c || epochs < 0 || batch <= 0) return TLR_E_INVALID;
    if (!c->has_gate) return c->fail(TLR_E_STATE, "initialise the gate parameters first");
    if (c->train_n == 0) return c->fail(TLR_E_STATE, "empty training set");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t n = c->train_n;
    if (c->train_idx) (void)hipFree(c->train_idx);
    HIPCHK(c, hipMalloc(&c->train_idx, sizeof(int) * n));
    std::vector<int> perm((size_t)n);
    std::mt19937_64 rng(seed);
    for (int e = 0; e < epochs; e++) {
        std::iota(perm.begin(), perm.end(), 0);
        std::shuffle(perm.begin(), perm.end(), rng);
        HIPCHK(c, hipMemcpyAsync(c->train_idx, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        HIPCHK(c, hipMemsetAsync(c->gate_loss, 0, sizeof(float), s));
        for (int64_t b0 = 0; b0 < n; b0 += batch) {
            const int nb = (int)std::min<int64_t>(batch, n - b0);
            HIPCHK(c, hipMemsetAsync(c->gate_g, 0, sizeof(float) * kGateParams, s));
            HIPCHK(c, launch_gate_grad(c->train_X, c->train_y, c->train_idx + b0, nb, c->gate_p, c->gate_g,
                                       c->gate_loss, pos_w, s));
            c->gate_step++;
            HIPCHK(c, launch_gate_adam(c->gate_p, c->gate_g, c->gate_m, c->gate_v, lr, 0.9f, 0.999f, 1e-8f,
                                       c->gate_step, 1.0f / (float)nb, s));
        }
        float l = 0;
        HIPCHK(c, hipMemcpyAsync(&l, c->gate_loss, sizeof(float), hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        if (epoch_loss) epoch_loss[e] = l / (float)n;
    }
    return TLR_OK;
}


// ---------------------------------------------------------------------------
// Batched reconstruction.  No reference counterpart: the reference builds and
// filters one TreeGather at a time.  Buffers are uploaded once into slots;
// each call builds every item in turn (the build needs a few host reads of
// device counts), keeps each item's query products in its slot, then runs the
// query phase for all items as one captured graph: a counter clear, one
// persistent launch of the fast kernel over every item's queries, one
// persistent launch of the general kernel over the shared hand-off list, and
// the per-item combines.  The graph is replayed while the batch's shape stays
// the same.  The first call (per context) times the fast-kernel variants and
// the general-kernel grid on the actual batch and keeps the fastest.
// ---------------------------------------------------------------------------

tlr_status tlr_batch_set_samples(tlr_ctx* c, int slot, const tlr_header* h, const float* rec, int64_t n) {
    if (!c) return TLR_E_INVALID;
    if (slot < 0 || slot > 4096) return c->fail(TLR_E_INVALID, "slot %d out of range", slot);
    tlr_status st = check_header(c, h);
    if (st != TLR_OK) return st;
    if (!rec || n != (int64_t)h->width * h->height * h->spp)
        return c->fail(TLR_E_INVALID, "record count does not match the header");
    HIPCHK(c, hipSetDevice(c->device));
    while ((int)c->slots.size() <= slot) c->slots.push_back(new BatchSlot);
    BatchSlot* b = c->slots[slot];
    const int N = (int)n;
    const size_t need = (size_t)12 * N * sizeof(float);
    if (need > b->d_in_cap) {
        if (b->d_in) HIPCHK(c, hipFree(b->d_in));
        b->d_in = nullptr;
        b->d_in_cap = 0;
        HIPCHK(c, hipMalloc(&b->d_in, need));
        b->d_in_cap = need;
    }
    static const int field[12] = {0, 1, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13};  // x,y,w,u,v,t,r,g,b,mv
    std::vector<float> tmp((size_t)N);
    for (int f = 0; f < 12; f++) {
        const int src = field[f];
        for (int k = 0; k < N; k++) tmp[k] = rec[(size_t)16 * k + src];
        HIPCHK(c, hipMemcpy(b->d_in + (size_t)f * N, tmp.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    }
    b->h = *h;
    b->N = N;
    return TLR_OK;
}

// Builds slot b with the context's build pipeline and keeps its products.
static tlr_status batch_build(tlr_ctx* c, BatchSlot* b, const tlr_batch_item& it, float* build_ms) {
    // the build reads the context's input: lend it the slot's
    std::swap(c->d_in, b->d_in);
    std::swap(c->d_in_cap, b->d_in_cap);
    const tlr_header h0 = c->h;
    const int N0 = c->N;
    const bool has0 = c->has_samples;
    c->h = b->h;
    c->N = b->N;
    c->has_samples = true;
    tlr_status st = tlr_build(c, it.aperture_adjust, it.focal_adjust);
    std::swap(c->d_in, b->d_in);
    std::swap(c->d_in_cap, b->d_in_cap);
    c->h = h0;
    c->N = N0;
    c->has_samples = has0;
    c->built = false;  // the context's own input was not what was built
    if (st != TLR_OK) return st;
    float ms = 0;
    if (build_ms && hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) *build_ms = ms;

    std::vector<float> pat;
    const int M = make_output_pattern(it.out_spp, false, nullptr, pat);
    const int W = b->h.width, H = b->h.height;
    const size_t nq = (size_t)W * H * M;
    if (nq >= ((size_t)1 << 31)) return c->fail(TLR_E_INVALID, "too many output samples");
    const size_t nb = (size_t)c->nb, L = (size_t)c->L, nv = (size_t)c->nv;
    const size_t nt = nv > 0 ? 2 * nv - 1 : 0;
    const size_t bytes = sizeof(float4) * (2 * (nb + kPointPad) + nb) + sizeof(LeafRec) * L + sizeof(Bounds) * L +
                         sizeof(TravNode) * nt + sizeof(TravPair) * nv + 4 * pat.size() + sizeof(float4) * nq +
                         16 * (size_t)W * H + 17 * 256;
    HIPCHK(c, b->prod.reserve(bytes));
    b->prod.reset();
    b->pts.pab = b->prod.alloc<float4>(2 * (nb + kPointPad));
    b->pts.pc = b->prod.alloc<float4>(std::max<size_t>(nb, 1));
    b->leaves = b->prod.alloc<LeafRec>(std::max<size_t>(L, 1));
    b->leaf_b = b->prod.alloc<Bounds>(std::max<size_t>(L, 1));
    b->trav = b->prod.alloc<TravNode>(std::max<size_t>(nt, 1));
    b->pairs = b->prod.alloc<TravPair>(std::max<size_t>(nv, 1));
    b->pattern = b->prod.alloc<float>(pat.size());
    b->qcol = b->prod.alloc<float4>(nq);
    b->img = b->prod.alloc<float>((size_t)W * H * 4);
    if (b->prod.failed) return c->fail(TLR_E_HIP, "batch slot arena exhausted");
    hipStream_t s = c->stream;
    HIPCHK(c, hipMemcpyAsync(b->pts.pab, c->pts.pab, sizeof(float4) * 2 * (nb + kPointPad), hipMemcpyDeviceToDevice, s));
    if (nb) HIPCHK(c, hipMemcpyAsync(b->pts.pc, c->pts.pc, sizeof(float4) * nb, hipMemcpyDeviceToDevice, s));
    if (L) {
        HIPCHK(c, hipMemcpyAsync(b->leaves, c->leaves, sizeof(LeafRec) * L, hipMemcpyDeviceToDevice, s));
        HIPCHK(c, hipMemcpyAsync(b->leaf_b, c->leaf_b, sizeof(Bounds) * L, hipMemcpyDeviceToDevice, s));
    }
    if (nt) HIPCHK(c, hipMemcpyAsync(b->trav, c->trav, sizeof(TravNode) * nt, hipMemcpyDeviceToDevice, s));
    if (nv > 1)
        HIPCHK(c, hipMemcpyAsync(b->pairs, c->pairs, sizeof(TravPair) * (nv - 1), hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipMemcpyAsync(b->pattern, pat.data(), 4 * pat.size(), hipMemcpyHostToDevice, s));
    b->M = M;
    b->nq = nq;
    QueryParams& Q = b->q;
    memset(&Q, 0, sizeof(Q));
    Q.W = W; Q.H = H; Q.M = M;
    Q.spp_pow2 = c->spp_pow2;
    Q.cc0 = c->cc0; Q.cc1 = c->cc1;
    Q.n_nodes = (int)nt;
    Q.root = nv > 1 ? (int)nv / 2 - 1 : 0;
    Q.num_queries = (unsigned)nq;
    Q.pattern = b->pattern;
    Q.trav = b->trav;
    Q.pairs = b->pairs;
    Q.leaves = b->leaves;
    Q.leaf_b = b->leaf_b;
    Q.pts = b->pts;
    Q.qcol = b->qcol;
    Q.tile_order = 1;
    if (fast_slots(Q) >= ((size_t)1 << 32)) Q.tile_order = 0;
    // the pattern copy reads host memory that goes away with `pat`
    HIPCHK(c, hipStreamSynchronize(s));
    return TLR_OK;
}

// Batch counters, in 32-bit words: block_next, list_count (fail_count),
// list_next, one word of padding so that the 64-bit stats start 8-byte
// aligned, the stats (2 * kNumStats words), then a (next, fail_count) pair per
// item for the chained form.
constexpr int kBatchStatsWord = 4;
constexpr int kBatchChainWord = kBatchStatsWord + 2 * kNumStats;

// Query phase of a batch, merged form: one launch of each query kernel over
// every buffer, with a given fast variant and general grid.
static tlr_status batch_queries(tlr_ctx* c, const BatchQuery& B, int variant, int general_blocks,
                                const std::vector<BatchSlot*>& its, unsigned* counters,
                                unsigned long long* stamps) {
    hipStream_t s = c->stream;
    HIPCHK(c, launch_fill_u32(counters, kBatchChainWord, 0u, s));
    HIPCHK(c, launch_query_fast_batch(B, variant, s));
    HIPCHK(c, launch_query_general_batch(B, general_blocks, s));
    for (size_t i = 0; i < its.size(); i++) {
        BatchSlot* b = its[i];
        HIPCHK(c, launch_combine(b->qcol, b->h.width, b->h.height, b->M, b->img, s));
        if (stamps) HIPCHK(c, launch_stamp(stamps + i, s));
    }
    return TLR_OK;
}

// Chained form: the single-buffer launches (fast kernel in tile order,
// general kernel over its hand-offs, combine) buffer after buffer, sharing
// the hand-off list and the general kernel's scratch.  No host work between
// buffers, so in a graph it runs back to back.
static tlr_status batch_chain(tlr_ctx* c, const std::vector<BatchSlot*>& its, unsigned* list, unsigned* counters,
                              float* scratch, int waves, const BatchExtras& x) {
    hipStream_t s = c->stream;
    const int n = (int)its.size();
    HIPCHK(c, launch_fill_u32(counters, kBatchChainWord + 2 * n, 0u, s));
    if (x.gate) HIPCHK(c, launch_fill_u32(x.accepted, n, 0u, s));
    for (int i = 0; i < n; i++) {
        BatchSlot* b = its[i];
        QueryParams Q = b->q;
        unsigned* pair = counters + kBatchChainWord + 2 * i;
        Q.next_query = pair;
        Q.fail_list = list;
        Q.fail_count = pair + 1;
        Q.fail_item = nullptr;
        Q.stats = reinterpret_cast<unsigned long long*>(counters + kBatchStatsWord);
        Q.scratch = scratch;
        Q.num_waves = waves;
        if (b->gate && x.gate) {
            // as tlr_reconstruct with the gate: provisional colours and
            // features for the hand-offs, the gate's scores, the general
            // kernel over the entries it does not accept
            QueryParams F = Q;
            F.features = x.features;
            HIPCHK(c, launch_query_fast_gate(F, s));
            HIPCHK(c, launch_gate_forward_dev(x.features, pair + 1, b->nq, x.gate_p, x.logits, s));
            HIPCHK(c, launch_gate_accept(x.logits, x.features, list, pair + 1, b->nq, x.gate_thr, x.accepted + i, s));
        } else {
            HIPCHK(c, launch_query_fast(Q, false, s));
        }
        Q.qlist = list;
        Q.qcount = pair + 1;
        HIPCHK(c, launch_query_general(Q, s));
        HIPCHK(c, launch_combine(b->qcol, b->h.width, b->h.height, b->M, b->img, s));
        if (x.stamps) HIPCHK(c, launch_stamp(x.stamps + i, s));
    }
    return TLR_OK;
}

tlr_status tlr_reconstruct_batch(tlr_ctx* c, const tlr_batch_item* items, int n, tlr_batch_stats* bs) {
    if (!c) return TLR_E_INVALID;
    if (n <= 0 || !items) return c->fail(TLR_E_INVALID, "empty batch");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    std::vector<BatchSlot*> its((size_t)n);
    double build_ms = 0;
    bool any_gate = false;
    for (int i = 0; i < n; i++) {
        const tlr_batch_item& it = items[i];
        if (it.slot < 0 || it.slot >= (int)c->slots.size() || c->slots[it.slot]->N == 0)
            return c->fail(TLR_E_STATE, "batch item %d: slot %d has no samples", i, it.slot);
        if (it.out_spp <= 0) return c->fail(TLR_E_INVALID, "batch item %d: out_spp must be positive", i);
        for (int j = 0; j < i; j++)
            if (items[j].slot == it.slot) return c->fail(TLR_E_INVALID, "slot %d appears twice in the batch", it.slot);
        any_gate = any_gate || it.gate != 0;
    }
    if (any_gate && !c->has_gate) return c->fail(TLR_E_STATE, "gate requested but no gate parameters set");
    // completion stamps: [n] = the batch's start, [i] = item i's image done
    if (c->stamps_cap < n + 1) {
        if (c->d_stamps) HIPCHK(c, hipFree(c->d_stamps));
        c->d_stamps = nullptr;
        c->stamps_cap = 0;
        HIPCHK(c, hipMalloc(&c->d_stamps, sizeof(unsigned long long) * (size_t)(n + 1)));
        c->stamps_cap = n + 1;
    }
    HIPCHK(c, launch_stamp(c->d_stamps + n, s));
    for (int i = 0; i < n; i++) {
        const tlr_batch_item& it = items[i];
        its[i] = c->slots[it.slot];
        its[i]->gate = it.gate != 0;
        float ms = 0;
        tlr_status st = batch_build(c, its[i], it, &ms);
        if (st != TLR_OK) return st;
        if (it.build_ms) *it.build_ms = ms;
        build_ms += ms;
    }
    // shared buffers
    std::vector<unsigned> prefix((size_t)n + 1, 0);
    size_t total_q = 0;
    if (c->tuned_variant < 0) c->tuned_variant = 0;
    if (c->general_wpc == 0) c->general_wpc = general_waves_per_cu();
    const int cus = c->prop.multiProcessorCount;
    const int waves = std::max(4, cus * c->general_wpc);
    size_t gate_nq = 0;  // largest gated item
    for (int i = 0; i < n; i++) {
        total_q += its[i]->nq;
        if (its[i]->gate) gate_nq = std::max(gate_nq, its[i]->nq);
    }
    if (total_q >= ((size_t)1 << 31)) return c->fail(TLR_E_INVALID, "batch has too many output samples");
    size_t bbytes = sizeof(QueryParams) * n + 4 * (size_t)(n + 1) * kNumFastVariants + 8 * total_q +
                    4 * ((size_t)kBatchChainWord + 2 * (size_t)n) + (size_t)waves * kWaveScratchWords * 4 + 16 * 256;
    if (any_gate) bbytes += 4 * gate_nq * ((size_t)kGateFeatures + 1) + 4 * (size_t)n + 3 * 256;
    HIPCHK(c, c->barena.reserve(bbytes));
    c->barena.reset();
    QueryParams* d_items = c->barena.alloc<QueryParams>(n);
    unsigned* d_prefix = c->barena.alloc<unsigned>((size_t)(n + 1) * kNumFastVariants);
    unsigned* d_list = c->barena.alloc<unsigned>(total_q);
    unsigned* d_item = c->barena.alloc<unsigned>(total_q);
    unsigned* d_cnt = c->barena.alloc<unsigned>((size_t)kBatchChainWord + 2 * (size_t)n);
    float* d_scratch = c->barena.alloc<float>((size_t)waves * kWaveScratchWords);
    BatchExtras xs;
    xs.stamps = c->d_stamps;
    if (any_gate) {
        xs.gate = true;
        xs.features = c->barena.alloc<float>(gate_nq * kGateFeatures);
        xs.logits = c->barena.alloc<float>(gate_nq);
        xs.accepted = c->barena.alloc<unsigned>((size_t)n);
        xs.gate_thr = c->gate_thr;
        xs.gate_p = c->gate_p;
    }
    BatchExtras xs_plain;  // what the autotuner runs: no gate, no stamps
    if (c->barena.failed) return c->fail(TLR_E_HIP, "batch arena exhausted");
    std::vector<QueryParams> qp((size_t)n);
    for (int i = 0; i < n; i++) {
        qp[i] = its[i]->q;
        qp[i].fail_list = d_list;
        qp[i].fail_count = d_cnt + 1;
        qp[i].fail_item = d_item;
        qp[i].stats = reinterpret_cast<unsigned long long*>(d_cnt + kBatchStatsWord);
    }
    // block tables, one per variant (logical blocks depend on the block size)
    std::vector<unsigned> tables((size_t)(n + 1) * kNumFastVariants);
    for (int v = 0; v < kNumFastVariants; v++) {
        const unsigned bs_v = (unsigned)fast_variant_block(v);
        unsigned acc = 0;
        for (int i = 0; i < n; i++) {
            tables[(size_t)v * (n + 1) + i] = acc;
            acc += (unsigned)((fast_slots(its[i]->q) + bs_v - 1) / bs_v);
        }
        tables[(size_t)v * (n + 1) + n] = acc;
    }
    HIPCHK(c, hipMemcpyAsync(d_items, qp.data(), sizeof(QueryParams) * n, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(d_prefix, tables.data(), 4 * tables.size(), hipMemcpyHostToDevice, s));
    auto make_bq = [&](int v) {
        BatchQuery B;
        memset(&B, 0, sizeof(B));
        B.items = d_items;
        B.n_items = n;
        B.block_prefix = d_prefix + (size_t)v * (n + 1);
        B.total_blocks = tables[(size_t)v * (n + 1) + n];
        B.block_next = d_cnt;
        B.list = d_list;
        B.list_item = d_item;
        B.list_count = d_cnt + 1;
        B.list_next = d_cnt + 2;
        B.scratch = d_scratch;
        B.stats = reinterpret_cast<unsigned long long*>(d_cnt + kBatchStatsWord);
        return B;
    };
    const int general_full = waves / 4;  // blocks of four waves
    // times one query phase on the stream
    auto timed = [&](auto&& run, float* ms) -> tlr_status {
        HIPCHK(c, hipEventRecord(c->ev[2], s));
        tlr_status st = run();
        if (st != TLR_OK) return st;
        HIPCHK(c, hipEventRecord(c->ev[3], s));
        HIPCHK(c, hipEventSynchronize(c->ev[3]));
        HIPCHK(c, hipEventElapsedTime(ms, c->ev[2], c->ev[3]));
        return TLR_OK;
    };
    // autotune once per context: fast variants at the full general grid,
    // then the general grid (full or three quarters) with the best variant,
    // then the chained form against the best merged one
    bool tuned_now = false;
    if (c->tuned_general_blocks == 0) {
        float best = 1e30f, ms = 0;
        int best_v = 0, best_g = general_full, chain = c->batch_mode == 1 ? 1 : 0;
        if (c->batch_mode != 1) {
            for (int v = 0; v < kNumFastVariants; v++) {
                tlr_status st =
                    timed([&] { return batch_queries(c, make_bq(v), v, general_full, its, d_cnt, nullptr); }, &ms);
                if (st != TLR_OK) return st;
                if (ms < best) {
                    best = ms;
                    best_v = v;
                }
            }
            const int g34 = std::max(1, general_full * 3 / 4);
            tlr_status st =
                timed([&] { return batch_queries(c, make_bq(best_v), best_v, g34, its, d_cnt, nullptr); }, &ms);
            if (st != TLR_OK) return st;
            if (ms < best) {
                best = ms;
                best_g = g34;
            }
        }
        if (c->batch_mode == -1) {
            tlr_status st = timed([&] { return batch_chain(c, its, d_list, d_cnt, d_scratch, waves, xs_plain); }, &ms);
            if (st != TLR_OK) return st;
            if (ms < best) chain = 1;
        }
        c->tuned_variant = best_v;
        c->tuned_general_blocks = best_g;
        c->tuned_chain = chain;
        tuned_now = true;
    }
    const int v = c->tuned_variant;
    const BatchQuery B = make_bq(v);
    // the gate runs in the chained form (the single-buffer gated kernels)
    const int chain = any_gate ? 1 : c->tuned_chain;
    // graph of the query phase, keyed by everything its launches capture:
    // the chained form passes each item's QueryParams by value (CoC
    // coefficients, spp, tree size and root, every device pointer), so all
    // of its bytes are part of the key, as are the shared buffers' addresses
    uint32_t thr_bits = 0;
    memcpy(&thr_bits, &xs.gate_thr, sizeof(thr_bits));
    std::vector<uint64_t> key = {(uint64_t)n, (uint64_t)v, (uint64_t)c->tuned_general_blocks,
                                 (uint64_t)chain, (uint64_t)(uintptr_t)c->barena.base,
                                 (uint64_t)B.total_blocks, (uint64_t)(uintptr_t)d_items,
                                 (uint64_t)(uintptr_t)d_prefix, (uint64_t)(uintptr_t)d_cnt,
                                 (uint64_t)(uintptr_t)d_scratch, (uint64_t)waves,
                                 (uint64_t)(uintptr_t)xs.stamps, (uint64_t)(uintptr_t)xs.features,
                                 (uint64_t)(uintptr_t)xs.logits, (uint64_t)(uintptr_t)xs.accepted,
                                 (uint64_t)(uintptr_t)xs.gate_p, (uint64_t)thr_bits};
    static_assert(sizeof(QueryParams) % sizeof(uint64_t) == 0, "QueryParams is keyed as 64-bit words");
    for (BatchSlot* b : its) {
        key.push_back((uint64_t)(uintptr_t)b->img);
        key.push_back(((uint64_t)b->h.width << 40) | ((uint64_t)b->h.height << 20) | (uint64_t)b->M);
        key.push_back((uint64_t)b->gate);
        uint64_t words[sizeof(QueryParams) / sizeof(uint64_t)];
        memcpy(words, &b->q, sizeof(QueryParams));
        key.insert(key.end(), words, words + sizeof(QueryParams) / sizeof(uint64_t));
    }
    bool reused = c->bgraph && key == c->bkey;
    if (!reused) {
        if (c->bgraph) HIPCHK(c, hipGraphExecDestroy(c->bgraph));
        c->bgraph = nullptr;
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        tlr_status st = chain ? batch_chain(c, its, d_list, d_cnt, d_scratch, waves, xs)
                              : batch_queries(c, B, v, c->tuned_general_blocks, its, d_cnt, xs.stamps);
        hipError_t e = hipStreamEndCapture(s, &g);
        if (st != TLR_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        HIPCHK(c, e);
        e = hipGraphInstantiate(&c->bgraph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(c, e);
        c->bkey = key;
    }
    HIPCHK(c, hipEventRecord(c->ev[2], s));
    HIPCHK(c, hipGraphLaunch(c->bgraph, s));
    HIPCHK(c, hipEventRecord(c->ev[3], s));
    for (int i = 0; i < n; i++) {
        const tlr_batch_item& it = items[i];
        if (!it.rgba_out) continue;
        const size_t bytes = (size_t)its[i]->h.width * its[i]->h.height * 16;
        HIPCHK(c, hipMemcpyAsync(it.rgba_out, its[i]->img, bytes,
                                 it.out_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    }
    std::vector<unsigned> cnt((size_t)kBatchChainWord + 2 * (size_t)n);
    std::vector<unsigned long long> stamps((size_t)n + 1);
    std::vector<unsigned> acc((size_t)n, 0u);
    HIPCHK(c, hipMemcpyAsync(cnt.data(), d_cnt, 4 * cnt.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(stamps.data(), c->d_stamps, 8 * stamps.size(), hipMemcpyDeviceToHost, s));
    if (xs.accepted) HIPCHK(c, hipMemcpyAsync(acc.data(), xs.accepted, 4 * acc.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    unsigned long long st64[kNumStats];
    memcpy(st64, cnt.data() + kBatchStatsWord, sizeof(st64));
    double handed = cnt[1];
    if (chain)
        for (int i = 0; i < n; i++) handed += cnt[(size_t)kBatchChainWord + 2 * (size_t)i + 1];
    double accepted = 0;
    for (unsigned a : acc) accepted += a;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || khz <= 0) khz = 0;
    for (int i = 0; i < n; i++)
        if (items[i].done_ms)
            *items[i].done_ms = khz ? (float)((double)(stamps[i] - stamps[n]) / (double)khz) : 0.0f;
    if (bs) {
        memset(bs, 0, sizeof(*bs));
        float ms = 0;
        if (hipEventElapsedTime(&ms, c->ev[2], c->ev[3]) == hipSuccess) bs->query_ms = ms;
        bs->build_ms = build_ms;
        bs->total_ms = bs->build_ms + bs->query_ms;
        bs->general_queries = handed;
        bs->output_samples = (double)total_q;
        bs->fast_leaf_cap = chain ? kFastLeafCap : fast_variant_cap(v);
        bs->fast_block = chain ? kFastBlock : fast_variant_block(v);
        bs->general_blocks = chain ? general_full : c->tuned_general_blocks;
        bs->graph_reused = reused ? 1 : 0;
        bs->autotuned = tuned_now ? 1 : 0;
        bs->chained = chain;
        bs->gate_accepted = accepted;
    }
    if (st64[ST_OVERFLOW])
        return c->fail(TLR_E_CAPACITY, "%llu output samples exceeded the per-query capacity",
                       (unsigned long long)st64[ST_OVERFLOW]);
    return TLR_OK;
}

tlr_status tlr_use_slot(tlr_ctx* c, int slot) {
    if (!c) return TLR_E_INVALID;
    if (slot < 0 || slot >= (int)c->slots.size() || c->slots[slot]->N == 0)
        return c->fail(TLR_E_STATE, "slot %d has no samples", slot);
    HIPCHK(c, hipSetDevice(c->device));
    BatchSlot* b = c->slots[slot];
    tlr_status st = reserve_input(c, b->N);
    if (st != TLR_OK) return st;
    HIPCHK(c, hipMemcpyAsync(c->d_in, b->d_in, sizeof(float) * 12 * (size_t)b->N, hipMemcpyDeviceToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->h = b->h;
    c->N = b->N;
    c->has_samples = true;
    c->built = false;
    c->wgd_n = 0;
    return TLR_OK;
}

tlr_status tlr_batch_retune(tlr_ctx* c) {
    if (!c) return TLR_E_INVALID;
    c->tuned_variant = -1;
    c->tuned_general_blocks = 0;
    c->tuned_chain = 0;
    return TLR_OK;
}

tlr_status tlr_batch_set_mode(tlr_ctx* c, int mode) {
    if (!c) return TLR_E_INVALID;
    if (mode < -1 || mode > 1) return c->fail(TLR_E_INVALID, "batch mode %d: expected -1, 0 or 1", mode);
    c->batch_mode = mode;
    return tlr_batch_retune(c);
}

}  // extern "C"
