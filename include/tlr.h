/*
 * libtlr: MI355X-native temporal light field reconstruction (C ABI).
 *
 * Drop-in for the reconstruction slot of the reference's TreeGather
 * (src/reconstruction_lib/reconstruction/Reconstruction.hpp:42-48): build the
 * sample tree from a UVTSampleBuffer, then reconstruct an RGBA image.  Plain
 * pointers and sizes only; the C++ facade in
 * temporal-lightfield-reconstruction_amd/csrc/facade mirrors the reference
 * classes on top of it.
 *
 * Threading: a context owns one HIP stream and its device buffers; use one
 * context per host thread.  Calls on one context are serialized by the
 * caller.  Every call returns a tlr_status; tlr_last_error() describes the
 * last failure (the reference aborts through FW::fail at the same sites).
 */
#ifndef TLR_H
#define TLR_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tlr_status {
    TLR_OK = 0,
    TLR_E_INVALID = 1,          /* bad argument or non-finite sample data               */
    TLR_E_IO = 2,               /* sample file missing/unreadable (SampleBuffer.cpp:233) */
    TLR_E_ALL_DISCARDED = 3,    /* every sample discarded (Reconstruction.cpp:587,636)   */
    TLR_E_UNSUPPORTED_SPP = 4,  /* spp rounds above 256 (ReconstructionOur.cpp:66)       */
    TLR_E_IMAGE_TOO_SMALL = 5,  /* output smaller than the buffer (Reconstruction.cpp:223)*/
    TLR_E_HIP = 6,              /* HIP runtime error                                     */
    TLR_E_CAPACITY = 7,         /* a query exceeded the per-query scratch capacity       */
    TLR_E_STATE = 8             /* call order violated (e.g. reconstruct before build)   */
} tlr_status;

typedef struct tlr_ctx tlr_ctx;

/* Sample-buffer header: UVTSampleBuffer width/height/spp/CoC coefficients
 * (SampleBuffer.hpp:138-183, v1.3 header SampleBuffer.cpp:244-268). */
typedef struct tlr_header {
    int   width, height, spp;      /* spp 0: an irregular buffer (tlr_set_samples only) */
    float coc_c0, coc_c1;      /* CoC radius = C0/w + C1 (CameraParams.hpp:72) */
} tlr_header;

/* SoA view of the per-sample fields the path reads, indexed (y*W+x)*spp+i
 * (UVTSampleBuffer getters SampleBuffer.hpp:147-162).  Colour alpha is 1. */
typedef struct tlr_soa {
    const float *x, *y, *w, *u, *v, *t, *r, *g, *b, *mv_x, *mv_y, *mv_w;
} tlr_soa;

typedef struct tlr_recon_params {
    int   out_spp;             /* output samples per pixel (reference: 128)        */
    int   override_uvt;        /* constant (u,v,t): CameraParams::overrideUVT       */
    float uvt[3];
    int   lens_gaussian;       /* CameraParams::lensFilter == LENS_GAUSSIAN         */
    int   time_gaussian;       /* CameraParams::timeFilter == TIME_GAUSSIAN         */
    int   collect_stats;       /* fill the Stats counters (slower)                  */
    int   gate;                /* 0: plain method; 1: learned gate (path 0 only)    */
    int   path;                /* 0: fast 1R kernel + general kernel for the rest;
                                  1: general kernel for every output sample;
                                  2: the reference's GPU-variant method
                                     (ReconstructionCudaKernels.cu:242-868:
                                     1R first surface, 2R surfaces, nearest
                                     sample); lens/time filters do not apply;
                                     stats: leaf_nodes, samples_in_leaf_nodes,
                                     samples_within_1r count the first 1R
                                     launch (leaves, samples tested, colours
                                     read), num_2r the samples sent to 2R   */
    int   list_cap;            /* path 2, if > 0: per-lane list capacity of the first
                                  1R and 2R launches (output samples beyond it are
                                  run again with lists sized for them; results do
                                  not change)                                       */
    int   debug_image;         /* also make the reference's debug image (#surfaces,
                                  #samples within 1R of each pixel's last output
                                  sample; ReconstructionOur.cpp:312-314), read with
                                  tlr_copy_debug_image; path 0 or 1, no gate     */
    int   reserved;
} tlr_recon_params;

/* TreeGather::Stats (Reconstruction.hpp:134-147) as sums over output samples,
 * plus device timings of the last build/reconstruct (milliseconds). */
typedef struct tlr_stats {
    double leaf_nodes, samples_in_leaf_nodes, samples_within_1r, samples_fetched;
    double surfaces, num_2r, surfaces_merged, at_least_one, count, overflow;
    double build_ms, query_ms, combine_ms, total_ms;
    double num_leaves, num_samples;
    double fast_ms, general_ms;  /* split of query_ms: 1R lane kernel, general kernel */
    double general_queries;      /* output samples handled by the general kernel     */
    double gate_candidates;      /* gated path: hand-offs of the 1R lane kernel      */
    double gate_ms;              /* gated path: MLP + routing                        */
    double gate_rejected;        /* hand-offs the gate sent to the general kernel    */
    /* the 1R lane kernel's share of leaf_nodes, samples_in_leaf_nodes and
     * samples_within_1r (its algorithmic bytes) */
    double fast_leaf_nodes, fast_samples_in_leaf_nodes, fast_samples_within_1r;
    /* the plain path's second 1R lane pass, over the first one's hand-offs
     * that held more 1R leaves than its per-lane lists: its time (part of
     * fast_ms), the output samples it ran, and its share of leaf_nodes,
     * samples_in_leaf_nodes and samples_within_1r */
    double wide_ms, wide_queries;
    double wide_leaf_nodes, wide_samples_in_leaf_nodes, wide_samples_within_1r;
    /* tlr_reconstruct_dof_motion_shadows: the light-space queries its shadow
     * rays made (its wide_queries and general_queries count both passes) */
    double light_queries;
} tlr_stats;

typedef struct tlr_build_info {
    int   num_input, num_kept, num_bucketed;
    int   spp_pow2;
    float coc_adj[2];
    float bbmin[2], bbmax[2];
    float scale;
    int   bbmin_int[2];
    int   grid_w, grid_h;
    int   max_hits;
    int   num_initial_leaves;
    int   num_leaves, num_valid_leaves;
} tlr_build_info;

/* Context lifetime.  device: HIP device ordinal. */
tlr_status  tlr_create(tlr_ctx** out, int device);
void        tlr_destroy(tlr_ctx* ctx);
const char* tlr_last_error(const tlr_ctx* ctx);

/* UVTSampleBuffer(const char*) (SampleBuffer.cpp:230-387): v1.3 file, text or
 * binary, header inline or in "<file>.header".  Fills the header and returns
 * records (16 floats each, alpha forced to 1) in a malloc'd buffer that the
 * caller frees with tlr_free(). */
tlr_status  tlr_load_sample_file(const char* path, tlr_header* h, float** records, int64_t* n);
/* UVTSampleBuffer::serialize (SampleBuffer.cpp:404-564), v1.3. */
tlr_status  tlr_save_sample_file(const char* path, const tlr_header* h, const float* records, int binary,
                                 int separate_header);
void        tlr_free(void* p);

/* Input upload.  AoS records (16 floats, v1.3 order) from host memory, or an
 * SoA view of host (on_device=0) or device (on_device=1) arrays.  The buffers
 * are borrowed for the duration of the call.  With h->spp 0 (an irregular
 * buffer, TreeGather::init's hook Reconstruction.cpp:65-71), tlr_set_samples
 * takes n records in any order and the build's dispersion uses spp =
 * n / (width * height), rounded up to a power of two as for regular buffers;
 * tlr_set_samples_soa and the batch slots take regular buffers only. */
tlr_status  tlr_set_samples(tlr_ctx* ctx, const tlr_header* h, const float* records, int64_t n);
tlr_status  tlr_set_samples_soa(tlr_ctx* ctx, const tlr_header* h, const tlr_soa* soa, int on_device);

/* TreeGather ctor (Reconstruction.cpp:43-215): spp rounding, CoC adjust,
 * reprojection, bucketing, leaves and hierarchy, on the device. */
tlr_status  tlr_build(tlr_ctx* ctx, float aperture_adjust, float focal_adjust);
/* The same with CameraParams::overrideRefocusDistance
 * (Reconstruction.cpp:88-95): coc_override (2 floats, C0 and C1 of the
 * refocused camera, CameraParams::getCocCoeffs) replaces the adjusted
 * coefficients for the tree bounds and the filter; the reprojection to the
 * (u,v,t) centre keeps the header's.  NULL: as tlr_build. */
tlr_status  tlr_build_ex(tlr_ctx* ctx, float aperture_adjust, float focal_adjust, const float* coc_override);

/* reconstructDofMotion (Reconstruction.cpp:221-295): rgba_out receives
 * width*height*4 floats (row-major, the buffer's y order); on_device selects
 * a device or host destination.  rgba_out NULL (with out_on_device 0): the
 * image stays in the context's device buffer; NULL with out_on_device 1 is
 * TLR_E_INVALID.  stats may be NULL. */
tlr_status  tlr_reconstruct(tlr_ctx* ctx, const tlr_recon_params* p, float* rgba_out, int out_on_device,
                            tlr_stats* stats);

/* Allocation-free variant of tlr_reconstruct.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The plain method only (path 0 or 1, no gate); rgba_out is device memory
 * receiving width*height*4 floats.  Every query-phase buffer (pattern,
 * per-output-sample colours, hand-off list, counters, the general kernel's
 * per-wave scratch) lives in one fixed device array of 73.7 MB per GPU, held
 * until the process ends, and the image is processed in row bands of at most
 * 2^20 output samples; out_spp above 256, or an image row of more than 2^20
 * output samples, is TLR_E_INVALID.  The general kernel runs 64 waves, so
 * hand-off-heavy frames take longer than through tlr_reconstruct.  Images and
 * stats equal tlr_reconstruct's.  All calls on one GPU share the array (it
 * keeps the last call's contents; each call clears what it needs): calls that
 * overlap in time on one GPU, from any context or thread, race. */
tlr_status  tlr_reconstruct_noalloc(tlr_ctx* ctx, const tlr_recon_params* p, float* rgba_out, tlr_stats* stats);

/* ---- Shadows --------------------------------------------------------------
 * TreeGather::reconstructShadows and reconstructDofMotionShadows
 * (src/reconstruction_lib/reconstruction/Reconstruction.cpp:302-433, declared
 * Reconstruction.hpp:47-48), and the filter's shadow mode
 * (ReconstructionOur.cpp:78-80, 283-291, 325-347). */

/* Per-sample inputs of shadow reconstruction, input order, n = the buffer's
 * sample count: wg = n x (dw/dx, dw/dy) (UVTSampleBuffer::getSampleWG,
 * SampleBuffer.hpp:156; the records' floats 14, 15), density = n floats (the
 * LIGHTSPACE_DENSITY channel, Reconstruction.cpp:547).  NULL: 0 and 1, the
 * reference's defaults.  Set after tlr_set_samples*; new samples reset them. */
tlr_status  tlr_set_shadow_inputs(tlr_ctx* ctx, const float* wg, const float* density, int64_t n);

/* Explicit output samples through the filter (FilterTask::process with a
 * query buffer, Reconstruction.cpp:648-719, as reconstructShadows drives it):
 * queries = n x (x, y, u, v, t, w, density, 0), host memory.  shadow != 0:
 * Filterer::reconstruct with reconstructShadow -- the dispersion scaled by
 * sqrt(density), a gathered sample's colour counted only when it lights the
 * query (o.w - 0.01 <= its w at t + dot(o.xy - its xy, its w-gradient)),
 * black with its alpha otherwise, no lens/time weighting.  shadow == 0: the
 * plain filter (p's lens/time filters; w and density unused).  rgba_out: n x
 * 4 floats; ext_out (optional): n x 8 floats (xy of the nearest gathered
 * sample at the output's (u,v,t), its w there, its density, its w-gradient,
 * 0, 0): Filterer::Result's xy, w, density, wg.  p may be NULL. */
tlr_status  tlr_reconstruct_queries(tlr_ctx* ctx, const tlr_recon_params* p, const float* queries, int64_t n,
                                    int shadow, float* rgba_out, float* ext_out, tlr_stats* stats);

/* Matrices and camera data of reconstructDofMotionShadows, row-major
 * (m[4*i+j] = row i, column j). */
typedef struct tlr_shadow_params {
    int   width, height;          /* camera buffer size (= the light's)                 */
    float camera_coc[2];          /* the camera's CameraParams::getCocCoeffs()          */
    float light_coc[2];           /* the light's                                        */
    float camera_far;             /* the camera's far plane (camera.getFar())           */
    float c2l[16];                /* ws * cameraToShadow * invws * cameraProjectedZfromW */
    float zfromw_ndc[16];         /* invws * cameraProjectedZfromW                      */
    float inv_projection[16];     /* invCameraProjection                                */
    float shadow2_inv_proj[16];   /* cameraToShadow2 * invCameraProjection              */
    float shadow2_inv_t[16];      /* cameraToShadow2.inverted().transposed()            */
    float world_inv_t[16];        /* cameraToWorld.inverted().transposed()              */
} tlr_shadow_params;

/* The matrices of Reconstruction.cpp:365-388 from the two cameras, with FW's
 * float arithmetic (Mat4f products, inverted(), transposed()): the camera's
 * projection and camera-to-world, the light's projection and world-to-camera,
 * all row-major.  Host only (no GPU needed). */
tlr_status  tlr_shadow_setup(const float* cam_proj, const float* cam_to_world, const float* light_proj,
                             const float* light_world_to_cam, int width, int height, const float* cam_coc,
                             const float* light_coc, float cam_far, tlr_shadow_params* out);

/* TreeGather::reconstructDofMotionShadows (FilterTask::process2,
 * Reconstruction.cpp:722-833): camera and light are built contexts of the
 * same size on one GPU; p: out_spp, override_uvt and the lens/time filters of
 * the camera pass (gate, path and debug_image are ignored).  Each output
 * sample's result is taken back to the (u,v) centre, mapped into the light's
 * window (c2l), queried there in shadow mode at a position on the square
 * light (Sobol dimensions 3 and 4 plus a per-pixel offset: the first and
 * second draws of Random(hashBits(x, y)) as x and y), and shaded: 1.5 x the
 * light-space cosine of the normal from its w-gradients, plus 0.5 x the
 * ambient term from the world-space normal; a sample with no surface makes
 * its pixel red. */
tlr_status  tlr_reconstruct_dof_motion_shadows(tlr_ctx* camera, tlr_ctx* light, const tlr_recon_params* p,
                                               const tlr_shadow_params* sp, float* rgba_out, int out_on_device,
                                               tlr_stats* stats);

/* Blocks until the context's stream is idle. */
tlr_status  tlr_sync(tlr_ctx* ctx);

/* ---- Batched reconstruction (no reference counterpart: the reference
 * builds and filters one TreeGather at a time) ------------------------------
 * Many sample buffers of uneven size go through one build each and then one
 * query phase for all of them, captured in a hipGraph that is replayed while
 * the batch's shape (items, sizes, output spp) stays the same.  The query
 * phase takes one of two forms: merged (a single persistent launch of each
 * query kernel over every buffer) or chained (each buffer's own launches,
 * back to back).  The first batch of a context times the merged form's
 * launch variants and the chained form on that batch and keeps the fastest
 * (tlr_batch_retune starts over; tlr_batch_set_mode fixes the form).
 * Buffers are uploaded once into numbered slots; an item names its slot.  One
 * context runs one batch at a time. */
typedef struct tlr_batch_item {
    int    slot;                 /* filled by tlr_batch_set_samples           */
    float  aperture_adjust;      /* TreeGather's apertureAdjust (1 = none)    */
    float  focal_adjust;         /* TreeGather's focalDistanceAdjust          */
    int    out_spp;              /* output samples per pixel                  */
    float* rgba_out;             /* W*H*4 floats, or NULL to keep on device   */
    int    out_on_device;        /* rgba_out is device memory                 */
    float* build_ms;             /* optional: this item's build time          */
    int    gate;                 /* 1: learned gate (the context's parameters
                                    and threshold, as tlr_recon_params.gate);
                                    a batch with a gated item runs the chained
                                    form                                      */
    float* done_ms;              /* optional: time from the batch's start (its
                                    first build) to this item's image being
                                    complete on the device                    */
} tlr_batch_item;

typedef struct tlr_batch_stats {
    double build_ms;             /* sum of the items' builds                  */
    double query_ms;             /* the captured query phase, all items       */
    double total_ms;
    double output_samples;       /* over all items                            */
    double general_queries;      /* handed to the general kernel              */
    int    fast_leaf_cap, fast_block, general_blocks;  /* tuned launch shape  */
    int    graph_reused;         /* 1: the previous call's graph was replayed */
    int    autotuned;            /* 1: this call ran the autotuner            */
    int    chained;              /* 1: chained form, 0: merged form           */
    double gate_accepted;        /* hand-offs the gate accepted, all items    */
} tlr_batch_stats;

/* Uploads buffer `slot` (AoS records as for tlr_set_samples) to device
 * memory, where it stays until overwritten or the context is destroyed. */
tlr_status  tlr_batch_set_samples(tlr_ctx* ctx, int slot, const tlr_header* h, const float* records, int64_t n);
/* Builds and reconstructs every item; images go to rgba_out. */
tlr_status  tlr_reconstruct_batch(tlr_ctx* ctx, const tlr_batch_item* items, int n, tlr_batch_stats* stats);
/* Makes slot `slot` the context's input for tlr_build / tlr_reconstruct (a
 * device-to-device copy), e.g. to run a batch's buffers one at a time. */
tlr_status  tlr_use_slot(tlr_ctx* ctx, int slot);
/* Forgets the tuned launch shape; the next batch tunes again. */
tlr_status  tlr_batch_retune(tlr_ctx* ctx);
/* Query-phase form of later batches: -1 the autotuner chooses (default),
 * 0 merged, 1 chained.  Retunes.  TLR_E_INVALID for other values. */
tlr_status  tlr_batch_set_mode(tlr_ctx* ctx, int mode);

/* ---- Introspection for parity checks (no reference counterpart) -------- */
tlr_status  tlr_get_build_info(tlr_ctx* ctx, tlr_build_info* out);
/* reprojected x,y,w,t per input sample (w = 0 if discarded) */
tlr_status  tlr_copy_reprojected(tlr_ctx* ctx, float* x, float* y, float* w, float* t);
/* tree sample order: input index of each sample (num_bucketed) */
tlr_status  tlr_copy_layout(tlr_ctx* ctx, int32_t* src);
/* initial leaves: x0,y0,x1,y1,offset,count (num_initial_leaves, row-major owner order) */
tlr_status  tlr_copy_initial_leaves(tlr_ctx* ctx, int32_t* out6);
/* leaves: s0,s1 and 14 bound floats (12 plane coefficients, tspan) */
tlr_status  tlr_copy_leaves(tlr_ctx* ctx, int32_t* s01, float* bounds14);
/* hierarchy over the valid leaves (num_valid = n): leaf id of each slot (n),
 * internal node children (2 per node, n-1 nodes; a leaf child is its leaf id
 * with bit 31 set) and internal node bounds (14 floats each).  Root: internal
 * node n/2-1 (n >= 2) or the single leaf. */
tlr_status  tlr_copy_hierarchy(tlr_ctx* ctx, int32_t* leaf_order, int32_t* children2, float* bounds14);
/* per output sample colour of the last reconstruct, (y,x,i) order, 4 floats */
tlr_status  tlr_copy_query_colors(tlr_ctx* ctx, float* out4);
/* The debug image of the last reconstruct made with debug_image set:
 * width*height*4 floats, (#surfaces, #samples within 1R, 0, 0) per pixel
 * (TreeGather::reconstructDofMotion's debugImage, Reconstruction.cpp:276). */
tlr_status  tlr_copy_debug_image(tlr_ctx* ctx, float* rgba_out);
/* Gathered sets of n output samples o5 (x,y,u,v,t each, host) at `radius`
 * over the current build, as the plain method's general kernel gathers them
 * (collectInputSamples2, ReconstructionOur.cpp:412-511): per output sample j,
 * leaves[j*leaf_cap ..] its leaf ids front to back (by key, then id),
 * n_leaves[j] their count; samples[j*sample_cap ..] the tree-layout index of
 * each gathered sample in gather order, surface_of[..] its surface, and
 * n_samples[j] the count (-1 for both counts past the kernel's capacity).
 * Rows are cut at leaf_cap / sample_cap; the counts are not. */
tlr_status  tlr_copy_query_gather(tlr_ctx* ctx, const float* o5, int n, float radius, int leaf_cap, int32_t* leaves,
                                  int32_t* n_leaves, int sample_cap, int32_t* samples, int32_t* surface_of,
                                  int32_t* n_samples);
/* After a path-2 reconstruction: per output sample, the stage that produced
 * its colour (0: recKernel 1R, 1: failKernel 2R, 2: nearestKernel),
 * ReconstructionCudaKernels.cu:242,476,634. */
tlr_status  tlr_copy_query_stages(tlr_ctx* ctx, int32_t* out);
/* output pattern: 64*M templates x,y,u,v,t (host computation) */
int         tlr_output_pattern(int out_spp, int override_uvt, const float* uvt, float* out5);

/* ---- Profiler counter calibration (no reference counterpart) ------------
 * One kernel launch (on the context's stream, then a sync) that moves a known
 * number of bytes, so that rocprofv3's FETCH_SIZE / WRITE_SIZE readings of
 * the query kernels can be scaled to bytes with the same access pattern:
 *   kind 0: streaming copy, 16 bytes per lane, `bytes` read and written;
 *   kind 1: gather of 32-byte records (the query kernels' per-sample record)
 *           in a fixed pseudo-random order, `bytes` read, bytes/2 written;
 *   kind 2: gather of 64-byte records (a hierarchy node), `bytes` read,
 *           bytes/4 written.
 * `bytes` is rounded down to a power of two of at least 1 MiB.  *read and
 * *written (optional) receive the bytes the kernel moves. */
tlr_status  tlr_calibrate_traffic(tlr_ctx* ctx, int kind, int64_t bytes, int64_t* read, int64_t* written);

/* ---- Learned gate (no reference counterpart) ------------------------------
 * The plain method's 1R lane kernel finishes most output samples and hands
 * the rest (a 2R fetch, no surface, too many leaves) to the general kernel.
 * With the gate on, the lane kernel also filters the first 1R surface of each
 * hand-off that has one (a provisional colour) and writes 24 features per
 * hand-off; a 24-64-64-1 MLP (BF16 operands, FP32 accumulation on MFMA)
 * scores them, and hand-offs scoring at least the threshold keep the
 * provisional colour instead of running the general kernel.  Parameter
 * layout: W1[24][64], b1[64], W2[64][64], b2[64], w3[64], b3
 * (TLR_GATE_PARAMS floats, row-major with the input index first). */
#define TLR_GATE_FEATURES 24
#define TLR_GATE_PARAMS 5825
tlr_status  tlr_gate_set_params(tlr_ctx* ctx, const float* params);
tlr_status  tlr_gate_get_params(tlr_ctx* ctx, float* params);
/* hand-offs with logit >= threshold keep their provisional colour */
tlr_status  tlr_gate_set_threshold(tlr_ctx* ctx, float logit_threshold);
/* host arrays: logits of n feature rows */
tlr_status  tlr_gate_forward(tlr_ctx* ctx, const float* features, int64_t n, float* logits);
/* host arrays: summed gradient of the weighted BCE loss over n rows */
tlr_status  tlr_gate_grad(tlr_ctx* ctx, const float* features, const float* labels, int64_t n, float pos_weight,
                          float* grad, float* loss);
/* Over the current build (plain method, path 0): every stride-th hand-off
 * entry's features, label (1 = the provisional colour is within 0.001 of the
 * plain method's in every channel), provisional colour, plain colour and
 * output-sample index ((y*W+x)*M+i).  At most `cap` rows are written; *n
 * receives the row count (which may exceed cap).  Buffers may be NULL. */
tlr_status  tlr_gate_collect(tlr_ctx* ctx, const tlr_recon_params* p, int stride, int64_t cap, float* features,
                             float* labels, float* prov_rgba, float* plain_rgba, uint32_t* query, int64_t* n);
/* Same rows appended to a device-resident training set. */
tlr_status  tlr_gate_append_training(tlr_ctx* ctx, const tlr_recon_params* p, int stride);
tlr_status  tlr_gate_clear_training(tlr_ctx* ctx);
int64_t     tlr_gate_training_size(tlr_ctx* ctx);
/* Adam on the device training set; epoch_loss (optional) gets the mean loss
 * of each epoch. */
tlr_status  tlr_gate_train(tlr_ctx* ctx, int epochs, int batch, float lr, float pos_weight, uint64_t seed,
                           float* epoch_loss);

#ifdef __cplusplus
}
#endif
#endif
