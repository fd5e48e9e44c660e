// Launchers of the build kernels (build.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "tlr_device.hpp"

namespace tlr {

// Reprojection to the (u,v,t) centre: per-field arrays rx..rt (w = 0 marks a
// discarded sample) and, for kept samples, packed 48-byte records
// (x, y, w, t | mv_x, mv_y, mv_w, r | g, b, 0, 0) for the tree-order gathers.
void launch_reproject(const InSoA& in, int n, int W, int H, float C0, float C1, float* rx, float* ry, float* rw,
                      float* rt, float4* packed, int* hits, GridInfo* gi, hipStream_t s);
// Bucket index (cell) of each kept sample, 0xFFFFFFFF outside the grid, and
// the identity as values for the stable sort by cell.
void launch_cell_keys(const float* rx, const float* ry, const float* rw, int n, GridInfo* gi, uint32_t* keys,
                      uint32_t* vals, hipStream_t s);
// Each cell's run [start, end) in the sorted keys (zero-filled beforehand).
void launch_cell_runs(const uint32_t* skeys, int n, uint32_t cells, uint32_t* start, uint32_t* end, hipStream_t s);
// Summed-area table of the cell counts end - start.
void launch_sat(const uint32_t* cstart, const uint32_t* cend, int gw, int gh, uint32_t* sat, hipStream_t s);
void launch_kd(const uint32_t* sat, int gw, int gh, int4* owner, uint32_t* own_flag, int* own_y1, hipStream_t s);
void launch_emit_initial(const int4* owner, const uint32_t* own_flag, const int* own_y1, const uint32_t* idx,
                         const uint32_t* sat, int gw, int gh, int4* rect, int2* offcnt, hipStream_t s);
void launch_place(const uint32_t* skeys, const uint32_t* svals, int nb, int gw, const int4* owner,
                  const uint32_t* sat, const uint32_t* cell_start, const float4* packed, uint32_t* cand, float* ct,
                  float* cw, hipStream_t s);
void launch_leaf_sort(const int2* offcnt, int nil, uint32_t* cand, float* ct, float* cw, hipStream_t s);
void launch_points(const uint32_t* cand, int nb, const float4* packed, float W, float H, float cc0, float cc1,
                   const Points& pts, Bounds* bounds, float4* boxes, hipStream_t s);
void launch_surface_scan(const int2* offcnt, int nil, const float4* pab, const float4* boxes, uint32_t* flags,
                         hipStream_t s);
// Leaves from the surface-scan start flags (L leaves; leaf_start: L+1 ints of
// scratch): records, and bounds folded in sample order.
void launch_leaf_emit(const uint32_t* flags, const uint32_t* leaf_pos, int nb, int L, int* leaf_start,
                      const Bounds* bounds, const Points& pts, LeafRec* leaves, Bounds* leaf_b,
                      uint32_t* valid_flag, hipStream_t s);
// Median-split tree over the nv valid leaves: ids[k] = leaf id of slot k,
// internal nodes g in [0, nv-1) with children / parent / range / split owner.
hipError_t launch_kdtree(const uint32_t* valid_flag, const uint32_t* valid_pos, int L, const Bounds* leaf_b,
                         int nv, float* feat, uint32_t* vid_leaf, uint32_t* k0, uint32_t* v0, uint32_t* k1,
                         uint32_t* v1, float4* info, uint32_t* emn, uint32_t* emx, uint32_t* hist, uint32_t* tmp, uint32_t* ids, int* c0, int* c1,
                         int* parent, int2* range, int* split_owner, Bounds* node_b, hipStream_t s);
// Traversal records (2n - 1) and the internal nodes' child-pair records (n - 1).
void launch_tree_trav(const uint32_t* ids, int n, const Bounds* leaf_b, const Bounds* node_b, const int* c0,
                      const int* c1, const int2* range, const int* split_owner, const LeafRec* leaves,
                      TravNode* trav, TravPair* pairs, hipStream_t s);

}  // namespace tlr
