/*
 * CPU oracle for the temporal light field reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C++17 restatement of the
 * reference CPU path (TreeGather build + Filterer::reconstruct) used as the
 * correctness checker for the HIP path and as the timed CPU baseline.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library never links or calls it.
 *
 * Every stage cites the reference file:line it restates (paths relative to the
 * reference repository root).  Arithmetic is IEEE FP32 with no contraction
 * (built with -ffp-contract=off, no fast-math), in the reference's operation
 * order; the triangle edge functions are evaluated in FP64 as the reference
 * does.
 *
 * Where the reference is ill-defined or hierarchy-dependent the oracle fixes a
 * rule, and the HIP path follows the same rule:
 *   - query-side leaf order: sorted by (key, leaf id) (the reference's
 *     unstable FW::sort tie order depends on its traversal order);
 *   - leaves whose time span is empty are never gathered;
 *   - box values of samples with an empty time span are zero (the reference
 *     reads stale scratch contents there);
 *   - the hierarchy above the leaves is a balanced median-split tree (the
 *     reference's is an agglomerative tree; the leaf set a query finds does
 *     not depend on it).
 *
 * Parity status of each function is listed in the header comment of
 * oracle.cpp.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* Records: N x 16 floats in the v1.3 order
 * x,y,z/w,w,u,v,t,r,g,b,a,mv_x,mv_y,mv_w,dwdx,dwdy, indexed (y*W+x)*spp+i. */
orc_ctx* orc_create(int W, int H, int spp, float C0, float C1, const float* records);
void     orc_destroy(orc_ctx* c);
const char* orc_error(const orc_ctx* c);

/* Build: spp rounding, CoC adjust, reprojection, grid, leaves, hierarchy.
 * threads: worker threads for the leaf stage (<=0: hardware concurrency).
 * Returns 0 on success, else a TLR_E_* style code (see include/tlr.h). */
int orc_build(orc_ctx* c, float aperture_adjust, float focal_adjust, int threads);

/* Scalars of the build. */
typedef struct orc_build_info {
    int   num_input, num_kept, num_bucketed;
    int   spp_pow2;
    float coc_adj[2];
    float bbmin[2], bbmax[2];       /* reprojected xy bbox (unscaled) */
    float scale;
    int   bbmin_int[2];
    int   grid_w, grid_h;
    int   max_hits;
    int   num_initial_nodes, num_initial_leaves;
    int   num_leaves, num_internal;
    int   root;                     /* hierarchy root (internal id, or leaf id | 0x80000000) */
} orc_build_info;
void orc_get_build_info(const orc_ctx* c, orc_build_info* out);

/* Reprojected samples in input order (only the kept ones), 10 floats each:
 * x,y,w,t,mvx,mvy,mvw,r,g,b ; src[k] = input record index. */
void orc_get_reprojected(const orc_ctx* c, float* out10, int32_t* src);
/* Cell id (row-major in the grid) of each kept sample, -1 if discarded. */
void orc_get_cells(const orc_ctx* c, int32_t* cell);
/* Initial leaves in DFS order: x0,y0,x1,y1,offset,count. */
void orc_get_initial_leaves(const orc_ctx* c, int32_t* out6);
/* Layout (tree sample order): input record index of each sample. */
void orc_get_layout(const orc_ctx* c, int32_t* src);
/* Leaves in layout order: s0,s1 and 14 floats (12 plane coeffs, tspan). */
void orc_get_leaves(const orc_ctx* c, int32_t* s01, float* bounds14);
/* Hierarchy: sorted leaf order (leaf ids), internal nodes: child0, child1 (leaf ids
 * tagged with 0x80000000) and 14 floats of bounds. */
void orc_get_hierarchy(const orc_ctx* c, int32_t* leaf_order, int32_t* children2, float* bounds14);

/* Reconstruction parameters. */
typedef struct orc_recon_params {
    int   out_spp;          /* output samples per pixel (reference default 128) */
    int   override_uvt;     /* 0/1: constant (u,v,t) for every output sample   */
    float uvt[3];
    int   lens_gaussian;    /* 0: box lens filter (default), 1: Gaussian        */
    int   time_gaussian;    /* 0: box time filter (default), 1: Gaussian        */
    int   threads;          /* <=0: hardware concurrency                        */
    int   row_stride;       /* reconstruct every k-th scanline (1 = all)        */
} orc_recon_params;

/* Stats counters (sum over output samples), the reference's Stats layout:
 * numLeafNodes, numSamplesInLeafNodes, numSamplesWithin1R, numSamplesFetched,
 * numSurfaces, num2R, numSurfacesMerged, numAtLeastOne; plus count. */
typedef struct orc_stats {
    double leaf_nodes, samples_in_leaf_nodes, samples_within_1r, samples_fetched;
    double surfaces, num_2r, surfaces_merged, at_least_one;
    double count;
} orc_stats;

/* Output pattern: 64 * spp' templates, 5 floats each (x,y,u,v,t) + key. */
int orc_output_pattern(int out_spp, int override_uvt, const float* uvt, float* out6);

/* Full reconstruction: image W*H*4 floats (row-major, y down as in the
 * sample buffer).  query_colors (optional): per output sample result colour,
 * W*H*out_spp*4 floats in (y, x, i) order. */
int orc_reconstruct(orc_ctx* c, const orc_recon_params* p, float* image,
                    float* query_colors, orc_stats* stats);
/* orc_reconstruct plus the debug image (optional, W*H*4): per pixel, the last
 * output sample's (#surfaces of its last gather, #samples within 1R, 0, 0)
 * (ReconstructionOur.cpp:312-314, Reconstruction.cpp:276). */
int orc_reconstruct_debug(orc_ctx* c, const orc_recon_params* p, float* image, float* query_colors,
                          orc_stats* stats, float* debug4);
/* overrideRefocusDistance (Reconstruction.cpp:88-95): later builds use cc
 * (C0, C1) in place of the adjusted coefficients; NULL clears it. */
void orc_set_coc_override(orc_ctx* c, const float* cc);

/* The reference's GPU-variant algorithm (recKernel, failKernel and
 * nearestKernel of ReconstructionCudaKernels.cu) over the same hierarchy.
 * image and query_colors as for orc_reconstruct; stage (optional): per
 * output sample, the pass that produced its colour (0: first surface at 1R,
 * 1: surfaces at 2R, 2: nearest-sample search). */
int orc_reconstruct_refgpu(orc_ctx* c, const orc_recon_params* p, float* image,
                           float* query_colors, int32_t* stage);

/* Pieces of the GPU-variant method, for checks: one output sample (o5 =
 * x,y,u,v,t) through all stages; the nearest search alone (returns the
 * tree-layout index of the sample found or -1); checkTriangles on n points
 * (x,y pairs relative to the output sample); the variant's reprojection of
 * every tree sample to o5 (dx, dy, dist^2 per sample, layout order). */
int orc_refgpu_query(orc_ctx* c, const float* o5, float* col4, int32_t* stage);
int orc_refgpu_nearest(orc_ctx* c, const float* o5, float* col4);
int orc_refgpu_failed(orc_ctx* c, const float* o5, float* col4);  /* 1: colour found */
int orc_refgpu_triangles(const float* xy, int n, float disp);
int orc_refgpu_reproject(orc_ctx* c, const float* o5, float* out3);

/* Debug gather for one output sample at radius R: leaf ids (sorted order),
 * gathered sample layout indices and their surface ids.  Returns the number
 * of gathered samples; *num_leaves receives the leaf count.  Buffers may be
 * NULL to query sizes. */
int orc_gather(orc_ctx* c, const float* o5 /* x,y,u,v,t */, float R,
               int32_t* leaves, int max_leaves, int* num_leaves,
               int32_t* samples, int32_t* surface_of, int max_samples);

/* Per-query decision record for one output sample (pixel x,y, index i):
 * fetch2R, number of surfaces, samples within 1R, final dispersion, colour. */
typedef struct orc_query_info {
    int   fetch2r, num_surfaces_1r, num_samples_1r, num_samples_final;
    int   emergency_passes, found_surface, used_nearest;
    float dispersion;
    float color[4];
} orc_query_info;
int orc_query(orc_ctx* c, const orc_recon_params* p, int x, int y, int i, orc_query_info* out);

/* Brute-force leaf set (every valid leaf whose own bounds intersect), for
 * checking that the hierarchy traversal loses nothing. */
int orc_leaves_bruteforce(orc_ctx* c, const float* o5, float R, int32_t* leaves, int max_leaves);

/* Instrumentation: hierarchy nodes tested by orc_gather/orc_reconstruct so
 * far (single-threaded use); reset != 0 clears the counter. */
long long orc_nodes_tested(int reset);

/* Primitive restatements exposed for known-answer tests. */
uint32_t orc_hash_bits(uint32_t a, uint32_t b);
void     orc_ranrot(uint32_t seed, int n, uint32_t* out);
float    orc_sobol(int dim, int i);
void     orc_sobol2d(int i, float* out2);
float    orc_halton(int base, int i);
float    orc_lp(int i);
void     orc_to_unit_disk(float x, float y, float* out2);
/* computeTUXPlanes: pos/motion homogeneous (3 floats), screen W,H, coc (2).
 * out: 12 plane floats + 2 tspan; returns 1 if valid. */
int      orc_tux_planes(const float* pos, const float* motion, float W, float H,
                        const float* coc, float* out14);
/* TimeLensBounds(tb0, tb1) merge: 14 + 14 -> 14. */
void     orc_merge_bounds(const float* a14, const float* b14, float* out14);
/* FW::sort replica over floats with ties: sorts (key, payload) pairs by key
 * with the reference's unstable quicksort; returns the payload permutation. */
void     orc_fw_sort(int n, float* keys, int32_t* payload);
/* Edge-function point-in-triangle as the reference evaluates it. */
int      orc_triangle_covers(const float* r0, const float* r1, const float* r2, const float* o);
/* sameSurface on two 14-float bounds at output (u,v,t). */
int      orc_same_surface(const float* n0, const float* n1, const float* o5);
/* Leaf surface scan of n candidates sorted by (t, w), given each one's time
 * and its screen boxes (x, y pairs): corners at u=-1 / u=+1 at t=0 (ba, bb)
 * and t=1 (ba1, bb1) (ReconstructionTreeBuilder.cpp:222-340).  surf[k]
 * receives candidate k's surface id. */
int orc_surface_scan(int n, const float* t, const float* ba, const float* bb, const float* ba1, const float* bb1,
                     int32_t* surf);

/* An irregular buffer (Reconstruction.cpp:65-71): n records in any order,
 * spp 0; the dispersion takes spp = n / (W*H) (integer division). */
orc_ctx* orc_create_irregular(int W, int H, float C0, float C1, const float* records, int n);

/* ---- Shadows (TreeGather::reconstructShadows / reconstructDofMotionShadows,
 * Reconstruction.cpp:302-433; the filter's shadow test
 * ReconstructionOur.cpp:78-80, 283-291, 325-347) ---- */

/* Matrices and camera data of reconstructDofMotionShadows, row-major
 * (m[4*i+j] = row i, column j). */
typedef struct orc_shadow_params {
    int   width, height;          /* camera buffer size (= the light's)            */
    float camera_coc[2];          /* the camera's CameraParams::getCocCoeffs()     */
    float light_coc[2];           /* the light's                                   */
    float camera_far;             /* the camera's far plane                        */
    float c2l[16];                /* ws * cameraToShadow * invws * cameraProjectedZfromW */
    float zfromw_ndc[16];         /* invws * cameraProjectedZfromW                 */
    float inv_projection[16];     /* invCameraProjection                           */
    float shadow2_inv_proj[16];   /* cameraToShadow2 * invCameraProjection         */
    float shadow2_inv_t[16];      /* cameraToShadow2.inverted().transposed()       */
    float world_inv_t[16];        /* cameraToWorld.inverted().transposed()         */
} orc_shadow_params;

/* Per-sample LIGHTSPACE_DENSITY channel, input order (N floats; NULL: 1).
 * The w-gradients are the records' dwdx, dwdy (floats 14, 15).  Rebuild after. */
void  orc_set_density(orc_ctx* c, const float* density);
/* Explicit output samples: q8 = n x (x, y, u, v, t, w, density, 0).
 * shadow != 0: Filterer::reconstruct with reconstructShadow (w and density
 * are the query's own).  rgba: n x 4; ext8 (optional): n x (xy of the
 * nearest gathered sample at the output's (u,v,t), its w there, its density,
 * its w-gradient, 0, 0) -- the Result's xy, w, density, wg. */
int   orc_reconstruct_queries(orc_ctx* c, const orc_recon_params* p, const float* q8, int n, int shadow,
                              float* rgba, float* ext8, orc_stats* stats);
/* The matrices from the cameras (Reconstruction.cpp:365-388): projections,
 * the camera's camera-to-world and the light's world-to-camera, all
 * row-major, with FW's float arithmetic. */
void  orc_shadow_setup(const float* cam_proj, const float* cam_to_world, const float* light_proj,
                       const float* light_world_to_cam, int w, int h, const float* cam_coc,
                       const float* light_coc, float cam_far, orc_shadow_params* out);
/* FW's Mat4f::inverted (Math.hpp:955-976). */
void  orc_mat4_inverse(const float* m16, float* out16);
/* lightSpaceDot (common/Util.cpp:202-224); n3 receives the view-space normal. */
float orc_light_space_dot(const float* xyw4, const float* wg2, const orc_shadow_params* sp, float* n3);
/* FilterTask::process2 over the whole image (cam and light built, same size). */
int   orc_reconstruct_dof_motion_shadows(orc_ctx* cam, orc_ctx* light, const orc_recon_params* p,
                                         const orc_shadow_params* sp, float* image);

#ifdef __cplusplus
}
#endif
