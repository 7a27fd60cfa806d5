// fgs_internal.h -- argument records and launchers shared between the
// kernel translation units and the C-ABI layer.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fgs {

// Per-launch geometry arguments of the preprocess kernels (forward and
// backward).  All pointers are device pointers.
struct PreArgs {
    int P, D, M, W, H, gx, gy;
    float tanx, tany, fx, fy, mod;
    float log_lambda, energy_floor, ratio_min;
    int variant;      // 0 wodilate, 1 dilate
    int use_filter;
    int prefiltered;
    int grad_mode;    // 0 reference, 1 exact
    const float* view;
    const float* proj;
    const float* campos;
    const float* means3D;
    const float* scales;
    const float* rotations;
    const float* opacities;
    const float* shs;
    const float* colors_precomp;
    const float* cov3D_precomp;
    // FilterGS rows (launch_preprocess_fwd_rows): Gaussian i is row
    // rows_index[i] of the raw parameters -- means3D, scales = log scales,
    // rotations = raw quaternions, opacities = logits, rows_dc / rows_rest
    // the colour coefficients -- activated on load.
    const int64_t* rows_index;
    int64_t rows_n_src;
    const float* rows_dc;
    const float* rows_rest;
    int rows_rest_floats, rows_deg;
    const float* rows_campos;
};

// Outputs of the forward preprocess (SoA).
struct PreOut {
    int* radii;
    float2* means2D;
    float* depths;
    float4* conic_opacity;
    float* rgb;
    uint32_t* tiles;
    uint4* recs;                        // packed tile rectangle (x0|y0<<16, x1|y1<<16), tiles, 0
    uint64_t* vkeys;                    // (depth_bits << 32 | i), culled: (~0 << 32 | i)
    int* error_flag;                    // bit 0: prefiltered point culled
    uint32_t* seg_cnt;                  // batch: keys each wave appended to its segment
    uint64_t* vis_masks;                // batch: per (view, 64-group) visibility ballot
    uint32_t* gid;                      // batch: Gaussian index of each element
    // batch: tiles of each (view, 64-group); recs[e].w of an element is its
    // group's tiles before it (the offset of its slots in the group's run)
    uint32_t* grp_tiles;
    // batch: the elements of more than kGatherSolo tiles, appended in any
    // order (big_count cleared by the caller)
    uint32_t* big_list;
    unsigned long long* big_count;
};

struct PreBwdIn {
    const int* radii;
    const float* dL_dmean2D;   // [P,3] (x, y used); a batch: per element
    const float* dL_dconic;    // [P,3]; a batch: per element
    const float* dL_dcolors;   // [P,3]; a batch: per element
    const uint64_t* vis_masks; // batch: the element layout
    // The forward's colours, [P,3] or per element: a channel the SH clamp
    // zeroed holds -0, so the backward reads the clamp instead of
    // evaluating the SH again.
    const float* rgb;
};
struct PreBwdOut {
    float* dL_dopacity;        // [P], modified in exact mode
    float* dL_dmeans3D;        // [P,3]
    float* dL_dcov3D;          // [P,6]
    float* dL_dsh;             // [P,M,3] or null
    float* dL_dscales;         // [P,3] or null
    float* dL_drot;            // [P,4] or null
};

// One view of a batch, as the batched kernels read it from device memory.
// Matrices are column-major float[16] like fgs_raster_settings'.
struct BatchView {
    float view[16];
    float proj[16];
    float campos[3];
    float bg[3];
    float tanx, tany, fx, fy;
    int W, H, gx, gy;
    uint32_t tile_base;   // first tile of the view in the batch's tile numbering
    uint32_t pad;
    uint64_t pix_base;    // first pixel of the view in the concatenated per-pixel arrays
};

__device__ inline void bind_view(PreArgs& a, const BatchView& v)
{
    a.view = v.view;
    a.proj = v.proj;
    a.campos = v.campos;
    a.tanx = v.tanx;
    a.tany = v.tany;
    a.fx = v.fx;
    a.fy = v.fy;
    a.W = v.W;
    a.H = v.H;
    a.gx = v.gx;
    a.gy = v.gy;
}

struct BlendArgs {
    const uint2* ranges;
    // the Gaussian (element) of sorted kv pair k: pair_ids[k << pair_shift]
    // (the low words of the plain path's u64 pairs, shift 1; the batch's
    // value array, shift 0)
    const uint32_t* pair_ids;
    uint32_t pair_shift;
    // Written by the forward, read by the backward: per tile, the kv values
    // that blended into at least one pixel, in blend order, from the tile's
    // first kv position (vis_ids[ranges[t].x + i]); the strips of the tile
    // each of them blended into (kv_strips, same positions); their count
    // (tile_vis[t]).
    uint32_t* vis_ids;
    uint8_t* kv_strips;
    uint32_t* tile_vis;
    // batch: the Gaussian index i of element g (gid[g]; null: g itself),
    // reported in point ids and used for point_weight[id_base + i]
    const uint32_t* gid;
    // Backward: the nine gradient sums of a (tile, Gaussian) pair are stored
    // at part[9 p ..] with mark[p] = 1, p = the Gaussian's first slot plus
    // the row-major offset of the tile in its rectangle; p < part_cap.  The
    // first slot is recs[g].w (one view: the Gaussian's first pair before
    // the tile sort), plus slot_base[g / 64] when slot_base is given (batch:
    // the slots run per (view, 64-group) of the element layout).  A Gaussian
    // of more than kGatherSolo tiles adds its sums into the part entries of
    // its first slot with float atomics instead.
    const uint4* recs;
    const uint32_t* slot_base;
    float* part;
    uint8_t* mark;
    size_t part_cap;
    // Forward statistics (RenderStats, device layout of kStats*): null = off
    double* stats;
    float stats_thresh;        // KPC used threshold
    int W, H, gx, gy;
    const float2* means2D;
    const float* colors;       // [P,3]
    const float4* conic_opacity;
    const float* bg;           // device [3]
    // work shape: Gaussians staged in LDS per batch (forward 128 or 256,
    // backward 64, 128 or 256)
    int batch, bwd_batch;
    int walk;                  // forward walk: 0 adaptive, 1 wave-uniform, 2 divergent
    int bwd_layout;            // backward: 0 four waves per tile (one strip each), 1 one wave per tile
    int xcd_chunk;             // batched one-wave blend: runs of tiles per XCD (0: launch order)
    int fwd_layout;            // forward without statistics: 0 four waves per tile, 1 one wave per tile
    // batched path: kv values are elements (see batch_elem); point ids
    // report i, point weights go to point_weight[v * P + i]
    uint32_t id_stride;        // P of a batch, 0 for one view
    uint32_t id_base;          // v * P (set per view)
    // Batched one-wave blend as a persistent launch: runs of queue_chunk
    // consecutive tiles, in the order of queue_order (run = low 32 bits),
    // dealt round robin to eight queues, one per XCD, each claimed tile by
    // tile through its counter in queue_head[8]; a workgroup drains the
    // queue of its XCD, then the others.  Null queue_order: one workgroup per
    // tile.
    const uint64_t* queue_order;
    uint32_t* queue_head;
    uint32_t queue_chunk;
};

// The work queue of a persistent batched blend launch over T tiles in runs
// of C consecutive tiles: with by_cost, keys (255 - cost class) << 32 | run
// in keysA and one stable radix pass on the class into keysB, the heaviest
// runs first; without, the runs in launch order in keysB.  The eight claim
// counters in head are cleared.  The cost of a tile is its visible-list
// length when tile_vis is given (backward), else its kv count (forward);
// classes are quarter octaves of a run's summed cost.
hipError_t build_tile_queue(const uint2* ranges, const uint32_t* tile_vis, uint32_t T, uint32_t C, int by_cost,
                            uint64_t* keysA, uint64_t* keysB, uint32_t* hist, uint32_t* partials,
                            uint32_t* head, hipStream_t s);

// launchers (return hipError_t of the launch)
hipError_t launch_preprocess_fwd(const PreArgs& a, const PreOut& o, hipStream_t s);
hipError_t launch_preprocess_bwd(const PreArgs& a, const PreBwdIn& in, const PreBwdOut& o,
                                 hipStream_t s);
// error_flag bit 1: an index outside [0, rows_n_src)
hipError_t launch_preprocess_fwd_rows(const PreArgs& a, const PreOut& o, hipStream_t s);
hipError_t launch_fill_u32(void* p, uint32_t v, size_t n_words, hipStream_t s);
hipError_t launch_preprocess_fwd_batch(const PreArgs& a, const BatchView* views, int V,
                                       const PreOut& o, hipStream_t s);
// Zeroes the blend gradients of a batch's visible elements (groups = V *
// P64 / 64 visibility ballots): a3, b3, c3 hold three floats per element.
hipError_t launch_clear_elements(const uint64_t* vis_masks, size_t groups, float* a3, float* b3, float* c3,
                                 float* d1, hipStream_t s);
hipError_t launch_preprocess_bwd_batch(const PreArgs& a, const BatchView* views, int V,
                                       const PreBwdIn& in, const float* dL_dopacity_view,
                                       const PreBwdOut& o, hipStream_t s);

// scans / sort / binning
size_t scan_partials_words(size_t n);
hipError_t launch_sorted_records(const uint64_t* vkeys, const uint4* recs, int P,
                                 uint32_t* tiles_s, uint2* rects_s, hipStream_t s);
hipError_t launch_scan_tiles(const uint32_t* tiles_s, uint32_t* offs, int P, uint32_t* partials,
                             unsigned long long* counters, hipStream_t s);
size_t radix_hist_words(size_t n);
hipError_t radix_sort_u64(uint64_t* a, uint64_t* b, size_t n, int shift_lo, int n_passes,
                          uint32_t* hist, uint32_t* partials, hipStream_t s,
                          const unsigned long long* n_dev = nullptr);
// Batched depth sort over the per-wave segments the batched preprocess wrote
// (segment w: seg_cnt[w] keys from w * 64 * V): the first pass reads the
// segments, the later ones the dense result, whose size the first pass's
// scan stores in *n_dev.  The sorted keys end in b (5 passes).
hipError_t radix_sort_u64_segments(uint64_t* a, uint64_t* b, const uint32_t* seg_cnt, int P, int V,
                                   int shift_lo, int n_passes, uint32_t* hist, uint32_t* partials,
                                   unsigned long long* n_dev, hipStream_t s);
// duplicate also records in recs[g].w the position of the Gaussian's first pair
hipError_t launch_duplicate(const uint64_t* vkeys, const uint32_t* offs, const uint2* rects_s,
                            int P, int gx, uint64_t* pairs, uint4* recs, hipStream_t s);
hipError_t launch_ranges(const uint64_t* sorted, size_t K, uint2* ranges, int T, hipStream_t s);

// batched binning (depth key (v << 58) | (depth_bits << 26) | i)
// Grid of the grid-stride kernels that run over a device-side count: enough
// 256-lane blocks to fill the chip several times over.
constexpr size_t kStrideBlocks = 4096;
constexpr int kBatchViewShift = 58;
constexpr int kBatchIdxBits = 26;
constexpr uint64_t kBatchIdxMask = (1ull << kBatchIdxBits) - 1;
constexpr int kBatchMaxViews = 64;
// Element layout of a batch's per-view records and blend gradients: Gaussian
// i visible in view v is element v * P64 + 64 * (i / 64) + r, with r the rank
// of i among the Gaussians of its 64-group that view v sees (bit i % 64 of
// vis_masks[v * P64 / 64 + i / 64]) and P64 = P rounded up to 64.  The
// visible elements of a 64-group are packed at the front of its 64 slots, so
// the preprocess writes whole runs and culled elements take no stores.  The
// low 26 bits of a depth key hold 64 * (i / 64) + r.
__host__ __device__ inline uint32_t batch_p64(int P) { return ((uint32_t)P + 63u) & ~63u; }
__device__ __forceinline__ uint32_t batch_elem(uint64_t key, uint32_t P64)
{
    return (uint32_t)(key >> kBatchViewShift) * P64 + (uint32_t)(key & kBatchIdxMask);
}
__device__ __forceinline__ uint32_t batch_slot(const uint64_t* __restrict__ vis_masks, uint32_t P64, int v, int i)
{
    const uint32_t l = (uint32_t)i & 63u;
    const uint64_t m = vis_masks[(size_t)v * (P64 >> 6) + ((uint32_t)i >> 6)];
    return (uint32_t)v * P64 + ((uint32_t)i & ~63u) + (uint32_t)__popcll(m & ((1ull << l) - 1ull));
}
// The batched binning stages run over the first *n_dev of n slots (the
// visible (view, Gaussian) keys the depth sort left).  rects_s: each
// rectangle as (first tile, width | the view's tiles per row << 16).
hipError_t launch_sorted_records_batch(const uint64_t* vkeys, const uint4* recs, size_t n,
                                       const unsigned long long* n_dev, int P, const BatchView* views,
                                       uint32_t* tiles_s, uint2* rects_s, hipStream_t s);
// out[i] = in[0] + ... + in[i - 1] for i < n (u32).
hipError_t launch_scan_u32(const uint32_t* in, size_t n, uint32_t* out, uint32_t* partials, hipStream_t s);
// The visible elements of a batch in element order: grp_start[g] = visible
// elements of the (view, 64-group)s before g, list[grp_start[g] + r] =
// 64 g + r for r below group g's visible count.
hipError_t launch_visible_list(const uint64_t* vis_masks, size_t n_groups, uint32_t* grp_start, uint32_t* list,
                               uint32_t* partials, hipStream_t s);
hipError_t launch_scan_tiles_batch(const uint32_t* tiles_s, uint32_t* offs, size_t n,
                                   const unsigned long long* n_dev, uint32_t* partials,
                                   unsigned long long* counters, hipStream_t s);
// kv pairs of a batch are two u32 arrays: the tile key (view-local tile <<
// kBatchKeyViewBits) | view, and the element v * P + i.  A stable sort on the
// tile bits (from bit kBatchKeyViewBits) leaves each tile's pairs in view
// order, each view's in depth order, and a (tile, view) run is a run of equal
// keys.
constexpr int kBatchKeyViewBits = 6;
constexpr uint32_t kBatchMaxViewTiles = 1u << (32 - kBatchKeyViewBits);
// The duplication of the first min(n, *n_dev) depth-ordered keys into their
// min(cap, *K_dev) kv pairs, fused with the first tile-sort pass: the pairs
// land in (keys, vals) ordered by the db-bit digit at bit shift, without
// being written in duplication order.  chunk_first: scratch of cap / 4096 + 1
// words; hist / partials: the sort's histogram and scan scratch.
hipError_t launch_duplicate_sort_batch(const uint64_t* vkeys, const uint32_t* offs, const uint2* rects_s,
                                       const uint32_t* tiles_s, size_t n, const unsigned long long* n_dev,
                                       const unsigned long long* K_dev, int P, size_t cap, uint32_t* chunk_first,
                                       int shift, int db, uint32_t* keys, uint32_t* vals, uint32_t* hist,
                                       uint32_t* partials, hipStream_t s);
// Stable LSD sort of n (key, value) u32 pairs on n_passes digits of db <= 8
// bits of the key from bit shift_lo (only the first min(n, *n_dev) with
// n_dev); the result is in (ka, va) if n_passes is even, else in (kb, vb).
hipError_t radix_sort_kv32(uint32_t* ka, uint32_t* va, uint32_t* kb, uint32_t* vb, size_t n, int shift_lo,
                           int n_passes, int db, uint32_t* hist, uint32_t* partials, hipStream_t s,
                           const unsigned long long* n_dev);
hipError_t launch_ranges_batch(const uint32_t* sorted_keys, size_t cap, const unsigned long long* K,
                               const BatchView* views, uint2* ranges, int T, hipStream_t s);

// blend
// vis_last (per pixel): 1 + the index in the tile's visible list of the last
// Gaussian blended into the pixel (written by the forward, read by the backward)
hipError_t launch_blend_fwd(const BlendArgs& a, float* final_T, uint32_t* n_contrib, uint32_t* vis_last,
                            float* out_color, int* out_point_id, float* out_weight_pixel,
                            float* point_weight, float* kv_sum, uint32_t* kv_cnt, hipStream_t s);
// The per-Gaussian sums of the backward's per-pair partials are gathered
// from each Gaussian's run of slots (GradGather).  Gaussians of up to
// kGatherSolo tiles store one slot per pair, larger ones add their per-tile
// sums with global float atomics (at 16 the atomics of the close orbit views
// cost more than the longer per-lane gather at 128 does).
constexpr uint32_t kGatherSolo = 128;

// RenderStats on the device (rasterizer.h:22-36, packed as
// rasterize_points.cu:138-177): a head of 9 + 2 * kStatsMaxBins doubles
// (bar_g, K/P, K/T, K, T, P, used kv, unused kv, bin count, bin edges, bin
// counts), then tile_ratio[T] and tile_total[T]; after the packed part, the
// scratch the blend fills: tile_avg[T] and 2 + kKpcBins u64 counters (used,
// unused, bins).
constexpr int kStatsMaxBins = 64;
constexpr int kKpcBins = 11;
constexpr size_t kStatsHead = 9 + 2 * kStatsMaxBins;
inline size_t stats_words(size_t T) { return kStatsHead + 3 * T + 2 + kKpcBins; }
hipError_t launch_stats_finalize(double* stats, int T, int P, int64_t K, hipStream_t s);
// One view: slot s of the depth order is Gaussian vkeys[s] (low 32 bits)
// with tiles_s[s] tiles and first pair offs[s].  Batch (slot_base given):
// item j is element e = elems[j] (the visible elements in element order),
// with recs[e].z tiles and first slot slot_base[e / 64] + recs[e].w; the
// gather stores every item's gradients (zero where no pair blended), and the
// elements of more than kGatherSolo tiles are big[0 .. *n_big).  The first
// *n_dev (null: n) of the n items are in use.
struct GradGather {
    const uint64_t* vkeys;
    const uint32_t* tiles_s;
    const uint32_t* offs;
    size_t n;
    const unsigned long long* n_dev;
    const uint32_t* elems;
    const uint4* recs;
    const uint32_t* slot_base;
    const uint32_t* big;
    const unsigned long long* n_big;
};
hipError_t launch_blend_bwd(const BlendArgs& a, const GradGather& gg, const float* final_T,
                            const uint32_t* vis_last, const float* dL_dpix, float* dL_dmean2D,
                            float* dL_dconic, float* dL_dopacity, float* dL_dcolors, int exact,
                            hipStream_t s);
// batched: common.ranges / pairs / kv_strips / means2D / colors / conic_opacity
// cover the whole batch; per-pixel arrays are the views' blocks concatenated
hipError_t launch_blend_fwd_batch(const BlendArgs& common, const BatchView* views, int V, int T_total,
                                  float* final_T, uint32_t* n_contrib, uint32_t* vis_last,
                                  float* out_color,
                                  int* out_point_id, float* out_weight_pixel, float* point_weight,
                                  hipStream_t s);
hipError_t launch_blend_bwd_batch(const BlendArgs& common, const GradGather& gg, const BatchView* views,
                                  int V, int T_total, const float* final_T, const uint32_t* vis_last,
                                  const float* dL_dpix,
                                  float* dL_dmean2D, float* dL_dconic, float* dL_dopacity,
                                  float* dL_dcolors, int exact, hipStream_t s);

// training-side neighbours (train.hip)
struct SparseAdamArgs {
    int64_t n, N;              // selected rows, rows of the parameter
    int D;                     // floats per row
    const int64_t* index;      // [n] parameter row of selected row j
    const uint8_t* flag;       // [n] row j takes the step (null: all)
    float* param;              // [N, D]
    const float* param_sel;    // [n, D] the gathered rows the gradient belongs to
    const float* grad;         // [n, D]
    float* exp_avg;            // [N, D]
    float* exp_avg_sq;         // [N, D]
    float* max_exp_avg_sq;     // [N, D] or null (no amsgrad)
    float beta1, beta2, one_m_beta1, one_m_beta2, eps, bc2_sqrt, neg_step_size;
    int* bad;                  // bit 0: a row index outside [0, N)
};
hipError_t launch_sparse_adam(const SparseAdamArgs& a, hipStream_t s);

struct CounterArgs {
    int64_t n_vis, n_pid, N;
    const int64_t* visible_index;   // [n_vis] point of visible row j
    const int32_t* radii;           // [n_vis]
    const float* point_weight;      // [n_vis]
    const int64_t* point_id;        // [n_pid] visible rows that are the max at some pixel
    const int64_t* point_count;     // [n_pid] their pixel counts
    const float* grad;              // [n_vis, 3] view-space gradient, or null
    int32_t* create_steps;
    int16_t* visible_count;
    float* weights_max;
    float* weights_sum;
    int16_t* radii_max;
    int32_t* radii_max_max;
    int32_t* area_sum;
    float* grad_sum;
    int* bad;                       // bit 0: point index outside [0, N); bit 1: point id outside the rows
};
hipError_t launch_counter_update(const CounterArgs& a, hipStream_t s);

// Optional timing of the blend kernels alone (a no-op unless profiling is on).
void prof_kernel(bool backward, bool begin, hipStream_t s);

// TensorTree.traverse (traverse.hip)
constexpr int kTravGrid = 1024;    // blocks of the per-level count / scatter launches
constexpr int kTravBlock = 256;
constexpr int kTravCtl = 8;        // control words at the head of the workspace
constexpr int kTravFrontierA = 0, kTravFrontierB = 1, kTravOut = 2, kTravErr = 3;
struct TravArgs {
    int32_t N, C, n_nodes;
    const float* xyz;              // [N, 3]
    const float* scaling;          // [N, 3] raw (log) scales
    const float* rotation;         // [N, 4] raw quaternions
    const int32_t* node_index;     // [N] row of the node in `tree`, -1 for a leaf
    const int32_t* tree;           // [n_nodes, C] child ids, -1 for an empty slot
    const float* proj;
    const float* view;
    float fx, fy, tanx, tany, min_px;
    int32_t cap_frontier, cap_out; // capacities of the frontier and output arrays
    int32_t* ctl;                  // [kTravCtl] sizes and error bits (bit 0 bad index, bit 1 overflow)
    uint32_t* counts;              // [2 * kTravGrid] per-block keep / next counts
    uint32_t* offs;                // [2 * kTravGrid] their exclusive scans
    uint8_t* flags;                // per-candidate decision of the level
};
hipError_t launch_traverse(const TravArgs& a, const int64_t* roots, int32_t n_roots,
                           int32_t* frontier_a, int32_t* frontier_b, int32_t* out, int levels,
                           bool append, hipStream_t s);
hipError_t launch_traverse_emit(const int32_t* ids, int64_t n, int64_t* out, hipStream_t s);

// LoD cut and per-point utilities
struct LodArgs {
    int N, L;
    const float* means3d;
    const float* log_scales;
    const float* rotations;
    const float* proj;
    const float* view;
    const int32_t* node_index;
    const int32_t* ancestor_path;
    // instead of ancestor rows: walk the parent links (index_parent, -1 at a
    // root) of a leaf, at most kLodMaxWalk steps
    const int32_t* index_parent;
    float fx, fy, tanx, tany, padding, radius_max, radius_filter_threshold;
    uint8_t* mask;
    float* radius2d;
    float* depth;
    uint8_t* cull_flag;        // scratch [N]: internal & visible & 0 < r <= radius_max
};
hipError_t launch_lod_cut(const LodArgs& a, hipStream_t s);
constexpr int kLodMaxWalk = 255;
// ancestor rows [N, L] from parent links and depths (apps/ancestor.py:21-81);
// *bad |= 1 when a chain is longer than L columns
hipError_t launch_rebuild_ancestor(int N, int L, const int32_t* index_parent, const int32_t* depth,
                                   int32_t* out, int* bad, hipStream_t s);

// gather + activation of FilterGS.get_all
struct ActivateArgs {
    int64_t N, n_src;
    const int64_t* index;
    const float* xyz;            // [n_src,3]
    const float* log_scale;      // [n_src,3]
    const float* rot;            // [n_src,4] raw
    const float* opacity_logit;  // [n_src]
    const float* dc;             // [n_src,3] SH DC coefficient
    const float* rest;           // [n_src, rest_floats] coefficients 1.. (or null when deg == 0)
    int rest_floats;
    int deg;                     // active SH degree, 0..4
    const float* campos;         // device [3]
    float* xyz_o;
    float* scale_o;
    float* rot_o;
    float* opacity_o;
    float* colors_o;
    int* bad_index;
};
hipError_t launch_activate_gather(const ActivateArgs& a, hipStream_t s);
struct ActivateBwdArgs {
    int64_t N;
    const float* xyz;            // [N,3] (SH direction, deg > 0)
    const float* log_scale;      // [N,3]
    const float* rot;            // [N,4] raw
    const float* opacity_logit;  // [N]
    int rest_floats;             // floats per row of d_rest
    int deg;
    const float* campos;
    const float* g_scale;        // [N,3] gradients of the activated rows (null: zero)
    const float* g_rot;          // [N,4]
    const float* g_opacity;      // [N]
    const float* g_colors;       // [N,3]
    float* d_log_scale;          // outputs (null: not wanted)
    float* d_rot;
    float* d_opacity_logit;
    float* d_dc;
    float* d_rest;               // [N, rest_floats]
};
hipError_t launch_activate_bwd(const ActivateBwdArgs& a, hipStream_t s);
hipError_t launch_radius_v2(int P, const float* means3d, const float* scales,
                            const float* rotations, const float* proj, const float* view,
                            float fx, float fy, float tanx, float tany, float* radii,
                            hipStream_t s);
hipError_t launch_radius_aa(int P, const float* means3d, const float* scales,
                            const float* rotations, const float* view, const float* proj,
                            float tanx, float tany, int H, int W, float* radii, hipStream_t s);
hipError_t launch_mark_visible(int P, const float* means3d, const float* view, uint8_t* present,
                               hipStream_t s);
hipError_t launch_visible_flag(int B, int N, const float* xyz, const float* proj, float padding,
                               uint8_t* valid, float* depth, float* p_proj, hipStream_t s);
hipError_t launch_gather(int64_t N, const int64_t* index, int64_t n_src, const float* xyz,
                         const float* scaling, const float* rotation, const float* opacity,
                         const float* colors, const float* shs, int sh_floats, float* xyz_o,
                         float* scaling_o, float* rotation_o, float* opacity_o, float* colors_o,
                         float* shs_o, int* bad_index, hipStream_t s);

}  // namespace fgs
