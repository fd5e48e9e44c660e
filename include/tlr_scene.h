/*
 * Seeded synthetic sample-buffer generator (host only, C ABI).
 *
 * The reference ships no sample buffers (its only one is missing from the
 * repository), so every workload is synthesized: textured screen-aligned
 * quads under a thin-lens camera with homogeneous affine motion, rendered by
 * exact first-hit point sampling at stratified-jittered (x,y,u,v,t).
 *
 * Output records follow the v1.3 sample-buffer layout documented in the
 * reference README.txt:138-156: 16 floats
 *   x, y, z/w, w, u, v, t, r, g, b, a, mv_x, mv_y, mv_w, dwdx, dwdy
 * indexed (y*W+x)*spp+i, with mv the homogeneous motion of the centre-lens
 * position over the shutter: (xy*w)(T=0) = xy(T=t)*w(T=t) - t*mv.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tlr_scene_desc {
    int      width, height, spp;
    float    coc_c0, coc_c1;        /* CoC radius = C0/w + C1 (pixels)          */
    int      num_quads;             /* large textured quads                     */
    float    quad_min_frac;         /* side length range, fraction of min(W,H)  */
    float    quad_max_frac;
    int      num_occluders;         /* small quads                              */
    float    occluder_min_px;       /* side length range in pixels              */
    float    occluder_max_px;
    float    w_min, w_max;          /* depth range                              */
    int      w_log_uniform;         /* 1: log-uniform depths, 0: uniform         */
    float    mv_xy_px;              /* screen motion bound: mv_xy ~ U(+-m)*w    */
    float    mv_w_frac;             /* mv_w ~ U(+-f)*w                          */
    float    static_frac;           /* fraction of surfaces without motion      */
    float    checker_period_px;     /* checker albedo period in pixels          */
    uint64_t seed;
    int      threads;               /* <=0: hardware concurrency                */
} tlr_scene_desc;

/* Presets of the five workloads (1..5; 4 and 5 use the config-2 scene type
 * at the given size).  Fills *out; returns 0 or -1 for an unknown preset. */
int tlr_scene_preset(int preset, int width, int height, int spp, uint64_t seed, tlr_scene_desc* out);

/* Renders width*height*spp records into `records` (16 floats each).
 * Returns 0 on success. */
int tlr_scene_generate(const tlr_scene_desc* d, float* records);

#ifdef __cplusplus
}
#endif
