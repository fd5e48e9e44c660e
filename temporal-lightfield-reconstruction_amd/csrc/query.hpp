// Per-output-sample reconstruction kernels (query.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tlr_device.hpp"

namespace tlr {

// Capacities of one wave's scratch (per query).  A query that exceeds them
// is reported as an overflow and the call fails with TLR_E_CAPACITY.  (The
// leaf capacity can be set at compile time; a test build sets it tiny so
// that the capacity error is reached on ordinary scenes.)
#ifndef TLR_LEAF_CAP
#define TLR_LEAF_CAP 4096
#endif
constexpr int kLeafCap = TLR_LEAF_CAP;
constexpr int kSampleCap = 32768;   // gathered samples + merged-surface copies
constexpr int kStackCap = 8192;

// Marks an entry of the fast kernel's hand-off list whose 1R pass already
// decided on 2R (the rest of the entry is the query id).
constexpr unsigned kNeeds2R = 0x80000000u;
// A hand-off entry the learned gate accepted (the general kernel skips it);
// query ids stay below 2^31 - 1, so it never names a query.
constexpr unsigned kSkipEntry = 0xFFFFFFFFu;

constexpr int kNumStats = 16;
enum StatIdx {
    ST_LEAF_NODES = 0, ST_SAMPLES_IN_LEAVES, ST_WITHIN_1R, ST_FETCHED, ST_SURFACES, ST_2R, ST_MERGED,
    ST_AT_LEAST_ONE, ST_COUNT, ST_OVERFLOW,
    // the fast kernel's share of the first three (its 1R passes)
    ST_FAST_LEAVES, ST_FAST_IN_LEAVES, ST_FAST_KEPT,
    // the second (wide-list) 1R pass's share
    ST_WIDE_LEAVES, ST_WIDE_IN_LEAVES, ST_WIDE_KEPT
};

// Per-wave scratch words (int/float) for the general query kernel.
constexpr size_t kWaveScratchWords =
    (size_t)kStackCap + 3 * (size_t)kLeafCap + 5 * (size_t)kSampleCap + 5 * (size_t)kLeafCap;

// Tile order the single-buffer reconstruction uses (QueryParams::tile_order).
#ifndef TLR_TILE_ORDER
#define TLR_TILE_ORDER 2
#endif

struct QueryParams {
    int W, H, M;              // image size, output samples per pixel
    int spp_pow2;             // input spp rounded up to a power of two
    float cc0, cc1;           // adjusted CoC coefficients
    int lens_gauss, time_gauss;
    int collect_stats;
    int n_nodes;              // traversal nodes (0: empty tree)
    int root;                 // traversal index of the root node
    unsigned num_queries;
    const float* pattern;     // 64*M templates: x,y,u,v,t
    const TravNode* trav;
    const TravPair* pairs;    // child pairs of the internal nodes (n_nodes / 2 records)
    const LeafRec* leaves;
    const Bounds* leaf_b;
    Points pts;
    float4* qcol;             // per query colour
    unsigned long long* stats;
    unsigned* next_query;     // work counter (zeroed before launch)
    float* scratch;           // kWaveScratchWords per resident wave
    int num_waves;
    // general kernel: optional list of query ids (qlist[0 .. *qcount))
    const unsigned* qlist;
    const unsigned* qcount;
    // fast kernel: queries it hands to the general kernel
    unsigned* fail_list;
    unsigned* fail_count;
    // fast kernel over a list: qlist/qcount above when fast_list != 0
    int fast_list;
    // batched launches: the buffer index of each fail_list entry (else null)
    unsigned* fail_item;
    // gated fast kernel: kGateFeatures floats per hand-off entry (row = the
    // entry's position in fail_list)
    float* features;
    // fast kernel over all queries: 0 = lanes in query order (pixel-major);
    // 1 = each wave takes one output-sample index over an 8x8 pixel tile;
    // 2 = tiles as in 1, each tile row cut into 8 column segments, segment
    //     (x + row) % 8 run by the workgroups of XCD x (tile columns padded
    //     to a multiple of 16)
    int tile_order;
    // first image row of this launch: its W x H pixels are rows [y0, y0+H)
    // of the image (0 for a whole image); query ids and qcol are local
    int y0;
    // optional, with stats: per query (#surfaces of the last gather, #samples
    // within 1R), the values of the reference's debug image
    float2* qdebug;
    // general kernel: 0 = the plain method; 1 = also the extended result of
    // each query (qext); 2 = shadow queries (ReconstructionOur.cpp:78-80,
    // 283-291, 336-347) with extended results.  qin (modes 1, 2; optional in
    // 1): explicit queries, 2 float4 each: (x, y, u, v), (t, w, density, 0)
    // in place of the output pattern.  pext: per tree sample (wgx, wgy,
    // density, 0).  qext: 2 float4 per query: (x, y, w, density), (wgx, wgy,
    // 0, 0) of the nearest gathered sample (Filterer::Result's xy, w,
    // density, wg).
    int mode;
    const float4* qin;
    const float4* pext;
    float4* qext;
};

// Lanes the fast kernel is launched with over all queries of Q (not a list):
// with tile order every 8x8 pixel tile takes M whole waves, edge tiles
// included, so the count rounds up past num_queries.
inline size_t fast_slots(const QueryParams& Q) {
    if (!Q.tile_order) return Q.num_queries;
    size_t tiles_x = (size_t)((Q.W + 7) / 8);
    if (Q.tile_order == 2) tiles_x = (tiles_x + 15) & ~(size_t)15;
    return tiles_x * (size_t)((Q.H + 7) / 8) * (size_t)Q.M * 64;
}

// A batch of built buffers queried by one launch of each kernel.
struct BatchQuery {
    const QueryParams* items;     // per buffer (device array)
    int n_items;
    // fast kernel: blocks of the variant's size; buffer b owns blocks
    // [block_prefix[b], block_prefix[b+1])
    const unsigned* block_prefix;
    unsigned total_blocks;
    unsigned* block_next;         // unused by the current kernels (kept zeroed)
    // general kernel: the shared hand-off list and its buffer indices
    const unsigned* list;
    const unsigned* list_item;
    const unsigned* list_count;
    unsigned* list_next;
    float* scratch;               // kWaveScratchWords per resident wave
    unsigned long long* stats;    // aggregate counters (overflow)
};

// Capacity of the per-lane sorted leaf list of the fast kernel; queries with
// more 1R leaves go to the general kernel.  22 ids x 128 lanes = 11 KB of LDS
// per block, so the 7 waves per SIMD its registers allow (14 blocks) fit the
// CU's 160 KB (24 would hold 13 blocks).
#ifndef TLR_FAST_LEAF_CAP
#define TLR_FAST_LEAF_CAP 22
#endif
constexpr int kFastLeafCap = TLR_FAST_LEAF_CAP;
constexpr int kFastBlock = 128;   // block size of the single-buffer fast kernel
static_assert(kFastLeafCap <= 64, "the fast kernel keeps a 64-bit mask over its leaves");
constexpr int kGateFeatures = 24;
// Pending hit children a 1R lane walk holds (a deeper one is handed over).
constexpr int kWalkStack = 8;
// Per-lane leaf list of the second 1R pass (k_query_fast_wide) over the
// first pass's hand-offs that had more than kFastLeafCap leaves.
#ifndef TLR_WIDE_LEAF_CAP
#define TLR_WIDE_LEAF_CAP 40
#endif
constexpr int kWideLeafCap = TLR_WIDE_LEAF_CAP;
static_assert(kWideLeafCap > kFastLeafCap && kWideLeafCap <= 64, "the wide pass keeps a 64-bit leaf mask");

// Batched launches (tlr_reconstruct_batch).  The fast kernel comes in
// kNumFastVariants (leaf cap, block size) variants, chosen by the autotuner.
constexpr int kNumFastVariants = 4;
constexpr int kFastBatchBlockMax = 256;   // largest variant block (logical block size is per variant)
int fast_variant_cap(int v);
int fast_variant_block(int v);
hipError_t launch_fill_u32(unsigned* p, int n, unsigned v, hipStream_t s);
// dst row r = src row r*stride (rows of width_words 32-bit words), on the device.
hipError_t launch_take_rows(const void* src, void* dst, size_t rows, int width_words, int stride, hipStream_t s);
hipError_t launch_query_fast_batch(const BatchQuery& b, int variant, hipStream_t s);
hipError_t launch_query_general_batch(const BatchQuery& b, int blocks, hipStream_t s);

// Resident waves of the general kernel per CU (its occupancy limit).
int general_waves_per_cu();
hipError_t launch_query_general(const QueryParams& p, hipStream_t s);
// One lane per output sample, 1R only; completes every query whose result
// is decided by the first surface without a 2R fetch
// (ReconstructionOur.cpp:99,204: single surface, or a first surface covering
// all four quadrants) and appends the others to fail_list.
hipError_t launch_query_fast(const QueryParams& p, bool stats, hipStream_t s);
// Gated fast kernel: launch_query_fast (no stats) plus, for every hand-off
// with a first surface at 1R, its filtered colour in qcol (provisional) and,
// for every hand-off entry, its gate features in p.features.
hipError_t launch_query_fast_gate(const QueryParams& p, hipStream_t s);
// The 1R lane kernel in the general kernel's mode p.mode (1: extended
// results; 2: shadow queries): completes what it can, writing qcol and qext,
// and appends the rest to p.fail_list for the general kernel in that mode.
hipError_t launch_query_fast_mode(const QueryParams& p, bool stats, hipStream_t s);
// Second 1R pass (kWideLeafCap leaves per lane, persistent 64-lane blocks)
// over p.qlist[0 .. *p.qcount), the first pass's hand-offs; p.next_query is
// its work counter (zeroed before launch); its hand-offs, with the entries
// already marked kNeeds2R, go to p.fail_list.
hipError_t launch_query_fast_wide(const QueryParams& p, bool stats, int blocks, hipStream_t s);
int fast_wide_blocks_per_cu();
hipError_t launch_combine(const float4* qcol, int W, int H, int M, float* out_rgba, hipStream_t s);
// Writes the device's constant-rate wall clock (hipDeviceAttributeWallClockRate
// kHz) to *dst when the stream reaches it (also inside a captured graph).
hipError_t launch_stamp(unsigned long long* dst, hipStream_t s);
// Debug image from QueryParams::qdebug: pixel = (qdebug of its last output
// sample, 0, 0).
hipError_t launch_debug_image(const float2* qdebug, int W, int H, int M, float* out_rgba, hipStream_t s);
// The general kernel's gather for n listed output samples (o5: x,y,u,v,t
// each) at radius R; p.num_waves waves with p.scratch; see k_query_gather.
hipError_t launch_query_gather(const QueryParams& p, const float* o5, int n, float R, int leaf_cap, int* leaves,
                               int* n_leaves, int sample_cap, int* samples, int* surface_of, int* n_samples,
                               hipStream_t s);
// Counter calibration: kind 0 streams n float4 from src to dst; kinds 1 / 2
// gather n records of 2 / 4 float4 from src in the order (i * odd) mod n (n
// a power of two) and write one float4 per record to dst.
hipError_t launch_calibrate(int kind, const float4* src, float4* dst, size_t n, hipStream_t s);
// The gate's decision over the hand-off entries (*n_entries on the device,
// at most max_n): entries whose logit reaches the threshold and that have a
// provisional colour become kSkipEntry in place; *accepted counts them.
hipError_t launch_gate_accept(const float* logits, const float* features, unsigned* entries,
                              const unsigned* n_entries, size_t max_n, float threshold, unsigned* accepted,
                              hipStream_t s);
// Provisional colour of each hand-off entry (copied before the general
// kernel overwrites qcol).
hipError_t launch_gate_prov(const float4* qcol, const unsigned* entries, const unsigned* n_entries, size_t max_n,
                            float4* prov, hipStream_t s);
// Label of each entry: 1 if it has a provisional colour within tau (every
// channel) of the plain method's final colour.
hipError_t launch_gate_label(const float4* qcol, const unsigned* entries, const unsigned* n_entries, size_t max_n,
                             const float* features, const float4* prov, float tau, float* labels, hipStream_t s);

// The reference's GPU-variant method (ReconstructionCudaKernels.cu:242-868):
// stage launches over output samples and lists of them.
constexpr int kRefFirstCap = 22;        // 1R leaves per lane in LDS (first launch): 7 waves per SIMD fit
constexpr int kRefWavesPerCU = 8;       // resident waves of the persistent stage launches
// 2R stage with one wave per output sample: per-wave lists in global scratch
constexpr int kRefWaveStack = 4096;
constexpr int kRefWaveLeaves = 2048;
constexpr int kRefWavePoints = 4096;
constexpr size_t kRefWaveScratchWords = (size_t)kRefWaveStack + 3 * (size_t)kRefWaveLeaves +
                                        4 * (size_t)kRefWavePoints;
struct RefLaunch {
    QueryParams q;             // geometry, tree, points, qcol
    const unsigned* list;      // stage input (query ids) and its length
    const unsigned* count;
    unsigned* next;            // work counter (zeroed before launch)
    unsigned* to_fail;         // output lists: 2R, nearest, beyond this launch's lists
    unsigned* n_fail;
    unsigned* to_near;
    unsigned* n_near;
    unsigned* to_over;
    unsigned* n_over;
    unsigned* need;            // [0] leaves, [1] points: maxima over the to_over entries
    int* stage;                // optional: per query 0 (1R), 1 (2R), 2 (nearest)
    int* lists;                // per lane leaf_cap leaf ids, lane-interleaved per wave
    float2* xy;                // per lane point_cap points
    int leaf_cap, point_cap;   // per-lane list capacities (first launch: leaf_cap <= kRefFirstCap)
    float* wscratch;           // 2R wave launch: kRefWaveScratchWords per resident wave
};
int ref_stage_blocks(int cus, int waves_per_cu);
// first stage over every output sample (1R, LDS leaf lists)
hipError_t launch_ref_first(const RefLaunch& r, hipStream_t s);
// stage 0: 1R over r.list; stage 1: 2R over r.list (blocks of 4 waves)
hipError_t launch_ref_stage(const RefLaunch& r, int stage, int blocks, hipStream_t s);
// 2R stage, one wave per output sample of r.list (blocks of 4 waves); samples
// beyond r.leaf_cap leaves or r.point_cap points (<= the kRefWave* caps) go
// to r.to_over with their list sizes in r.need
hipError_t launch_ref_fail_wave(const RefLaunch& r, int blocks, hipStream_t s);
hipError_t launch_ref_nearest(const RefLaunch& r, int blocks, hipStream_t s);

}  // namespace tlr
