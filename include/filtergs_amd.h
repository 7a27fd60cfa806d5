/*
 * filtergs_amd.h -- C ABI of the MI355X (gfx950) FilterGS rasterizer and
 * LoD-cut kernels (libfiltergs_amd.so).
 *
 * Plain C: device pointers, sizes and a hipStream_t passed as void*.  No
 * allocation happens inside any entry point: callers size their buffers with
 * the *_bytes() queries and own all memory.  Every entry point returns 0 on
 * success and a nonzero FGS_E_* code otherwise; fgs_last_error() gives the
 * message of the calling thread's last failure.  Launches are asynchronous on
 * the given stream; no entry point synchronises the host.
 *
 * Each function names the reference interface it replaces (paths relative to
 * the FilterGS repository).
 */
#ifndef FILTERGS_AMD_H
#define FILTERGS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FGS_ABI_VERSION 1

enum {
    FGS_OK = 0,
    FGS_E_INVALID = 1,   /* bad argument (shape, size, null pointer)  */
    FGS_E_LAUNCH = 2,    /* HIP launch / runtime failure              */
    FGS_E_CAPACITY = 3,  /* caller buffer too small                    */
};

enum { FGS_VARIANT_WODILATE = 0, FGS_VARIANT_DILATE = 1 };
enum { FGS_GRAD_REFERENCE = 0, FGS_GRAD_EXACT = 1 };

/* GaussianRasterizationSettings plus the per-call OGFS overrides
 * (antialias/diff_gaussian_rasterization_wodilate/__init__.py:193-210,
 *  :245-268).  viewmatrix/projmatrix are device float[16] in column-major
 * order (the transposed row-vector matrices of LoG/dataset/base.py:20-55);
 * bg and campos are device float[3]. */
typedef struct {
    int32_t image_height;
    int32_t image_width;
    float tanfovx;
    float tanfovy;
    const float* bg;
    float scale_modifier;
    const float* viewmatrix;
    const float* projmatrix;
    int32_t sh_degree;
    const float* campos;
    int32_t prefiltered;
    int32_t debug;              /* nonzero: synchronise and check after each stage */
    int32_t variant;            /* FGS_VARIANT_*                                    */
    int32_t use_filter;         /* wodilate: max(cov2D_ii, 0.3) when nonzero        */
    float ogfs_epsilon_scale;
    float ogfs_epsilon_max;
    float ogfs_energy_floor;
    float ogfs_radius_ratio_min;
    int32_t ogfs_enable_stats;
    int32_t grad_mode;          /* FGS_GRAD_*                                       */
} fgs_raster_settings;

/* Per-Gaussian inputs (device, contiguous float32).  Exactly one of shs /
 * colors_precomp and one of (scales, rotations) / cov3D_precomp is non-null
 * (the checks of wodilate/__init__.py:270-274). */
typedef struct {
    int32_t P;                    /* number of Gaussians                          */
    int32_t M;                    /* SH coefficients per Gaussian (shs [P,M,3])   */
    const float* means3D;         /* [P,3]                                        */
    const float* shs;             /* [P,M,3] or NULL                              */
    const float* colors_precomp;  /* [P,3] or NULL                                */
    const float* opacities;       /* [P]                                          */
    const float* scales;          /* [P,3] or NULL                                */
    const float* rotations;       /* [P,4] (w,x,y,z) or NULL                      */
    const float* cov3D_precomp;   /* [P,6] or NULL                                */
} fgs_gaussians;

/* ---- rasterizer: replaces _C.rasterize_gaussians / RasterizeGaussiansCUDA
 * (antialias/rasterize_points.cu:36-179, ext.cpp:19-47) and
 * CudaRasterizer::Rasterizer::forward (cuda_rasterizer/rasterizer_impl.cu:254-563).
 * The reference sizes its binning buffer after a blocking read of the kv-pair
 * count; here the forward is split at that point into two calls. */

/* Bytes of the per-Gaussian state (the reference's geomBuffer). */
size_t fgs_geom_bytes(int32_t P);
/* Bytes of the per-pixel state (imgBuffer): final T, n_contrib, tile ranges. */
size_t fgs_image_bytes(int32_t image_width, int32_t image_height);
/* Bytes of the per-kv-pair state (binningBuffer) for K pairs. */
size_t fgs_binning_bytes(int64_t K, int32_t image_width, int32_t image_height);

/* Phase 1 -- preprocessCUDA (forward.cu:266-401) + per-Gaussian depth sort.
 * Writes radii[P] (int32, device) and counts[4] (uint64, device):
 * counts[0] = visible Gaussians, counts[1] = K (kv pairs), counts[2] =
 * error bits (bit 0: a point was culled although prefiltered was set). */
int fgs_forward_geometry(const fgs_raster_settings* s, const fgs_gaussians* g, void* geom,
                         int32_t* radii, uint64_t* counts, void* stream);

/* Phase 2 -- duplicateWithKeys, stable tile sort, identifyTileRanges and
 * renderCUDA (forward.cu:406-567).  K is counts[1] of phase 1.  Outputs:
 * out_color [3,H,W], out_point_id [H,W], out_weight_pixel [H,W],
 * out_point_weight [P] (must be zero on entry).  kv_alpha_sum [K] /
 * kv_alpha_count [K] may be null; when given they must be zero on entry and
 * receive the per-kv-pair sum of T*alpha and contribution counts
 * (the OGFS statistics inputs of rasterizer_impl.cu:450-458, 532-540). */
int fgs_forward_render(const fgs_raster_settings* s, const fgs_gaussians* g, const int32_t* radii,
                       void* geom, int64_t K, void* binning, void* image, float* out_color,
                       int32_t* out_point_id, float* out_weight_pixel, float* out_point_weight,
                       float* kv_alpha_sum, uint32_t* kv_alpha_count, void* stream);

/* fgs_forward_render plus RenderStats computed on the device
 * (rasterizer.h:22-36 packed as rasterize_points.cu:138-177; the host loops
 * of rasterizer_impl.cu:479-556 fused into the blend).  stats: device
 * doubles, fgs_stats_words(W, H) of them, any content on entry; on return the
 * first 9 + 2*64 + 2*T hold the packed head (bar_g_tile, K/P, K/T, K, T, P,
 * used kv, unused kv, bin count, 64 bin edges, 64 bin counts), tile_ratio[T]
 * and tile_total_kv[T]; the rest is scratch.  kpc_used_thresh is the
 * KPC_USED_THRESH of rasterizer_impl.cu:44-55 (> 0).  stats may be null. */
int fgs_forward_render_stats(const fgs_raster_settings* s, const fgs_gaussians* g, const int32_t* radii,
                             void* geom, int64_t K, void* binning, void* image, float* out_color,
                             int32_t* out_point_id, float* out_weight_pixel, float* out_point_weight,
                             float* kv_alpha_sum, uint32_t* kv_alpha_count, double* stats,
                             float kpc_used_thresh, void* stream);
/* Doubles of the stats buffer of fgs_forward_render_stats at W x H. */
size_t fgs_stats_words(int32_t image_width, int32_t image_height);

/* The LoD cut with the tree's parent links (index_parent int32 [N], -1 at a
 * root) in place of the ancestor rows: a leaf is culled when an ancestor on
 * its parent chain culls its subtree.  Same mask as fgs_compute_visibility_
 * radius with the rows fgs_rebuild_ancestor_path builds from those links, and
 * no [N, L] array (FilterGS/cuda/compute_radius_kernel_v2.cu:231-353 reads
 * the rows; this reads 4 bytes per node). */
int fgs_compute_visibility_radius_parent(int32_t N, const float* means3D, const float* log_scales,
                                         const float* rotations, const float* projmatrix,
                                         const float* viewmatrix, const int32_t* tree_node_index,
                                         const int32_t* index_parent, float focal_x, float focal_y,
                                         float tan_fovx, float tan_fovy, float padding, float radius_max,
                                         float radius_filter_threshold, uint8_t* mask, float* radius2d,
                                         float* depth, void* ws, void* stream);

/* rebuild_ancestor_path (apps/ancestor.py:21-81) on the device: row i of
 * ancestor_path [N, L] int32 gets the ancestor of node i at depth d in column
 * d (d < depth[i]) and -1 elsewhere; index_parent and depth are int32 [N].
 * An orphan (parent -1 below the root level) keeps a row of -1.  *bad
 * (device int32, zero on entry) gets bit 0 when a depth exceeds L. */
int fgs_rebuild_ancestor_path(int32_t N, int32_t L, const int32_t* index_parent, const int32_t* depth,
                              int32_t* ancestor_path, int32_t* bad, void* stream);

/* TensorTree.traverse (LoG/model/tensor_tree.py:133-188, with the radius of
 * Gaussian.compute_radius, LoG/model/level_of_gaussian.py:66-88, and
 * LoG/cuda/compute_radius_kernel.cu:107-156) on the device, without a host
 * wait per level.
 *   xyz [N,3], scaling [N,3] (raw, exp-activated here), rotation [N,4] (raw,
 *   normalised here): the Gaussian parameters, float;
 *   node_index [N] int32: row of node i in tree, -1 for a leaf;
 *   tree [n_nodes, C] int32: child ids, -1 for an empty slot;
 *   root_index [n_roots] int64: the roots to start from;
 *   min_resolution_pixel: a node whose radius is below it is kept whole;
 *   max_depth: the reference's argument (levels past min(30, max_depth) are
 *   not expanded; the frontier there is returned as is);
 *   level_hint: levels to run at most (>= the tree's depth below the roots),
 *   or -1 for min(30, max_depth); a hint below the depth returns an error.
 * ws: fgs_tree_traverse_ws_bytes(N, C, n_roots) device bytes.  Waits for the
 * stream once and stores the number of ids in *count (host); the ids stay in
 * ws until fgs_tree_traverse_result copies them out, in the reference's order
 * (kept roots, then the kept nodes of level 1, 2, ..., then the frontier at
 * the level limit).  Errors: ids out of range (FGS_E_INVALID), more than
 * N + n_roots ids (FGS_E_CAPACITY: duplicate roots or a cyclic tree). */
size_t fgs_tree_traverse_ws_bytes(int32_t N, int32_t C, int32_t n_roots);
int fgs_tree_traverse(int32_t N, const float* xyz, const float* scaling, const float* rotation,
                      const int32_t* node_index, const int32_t* tree, int32_t n_nodes, int32_t C,
                      const int64_t* root_index, int32_t n_roots, const float* projmatrix,
                      const float* viewmatrix, float focal_x, float focal_y, float tan_fovx,
                      float tan_fovy, float min_resolution_pixel, int32_t max_depth,
                      int32_t level_hint, void* ws, int64_t* count, void* stream);
/* The count ids of the last fgs_tree_traverse on ws as int64 into out [count]
 * (the reference returns an int64 index tensor). */
int fgs_tree_traverse_result(int32_t N, int32_t C, int32_t n_roots, const void* ws, int64_t count,
                             int64_t* out, void* stream);

/* One step of LoG's SparseOptimizer for one parameter
 * (LoG/model/sparse_optimizer.py:41-78, 163-195): for each selected row j < n
 * with flag[j] (flag may be null: all rows), parameter row i = index[j] gets
 * the Adam update of the gradient row grad[j] from the gathered row
 * param_sel[j], and exp_avg / exp_avg_sq (/ max_exp_avg_sq with amsgrad,
 * else null) rows i are updated in place -- the reference's gather of the
 * state, update and scatter back, without the copies.  Arrays are float32:
 * param, exp_avg, exp_avg_sq [N, D]; param_sel, grad [n, D].  step is the
 * global step (>= 1); *bad (device, zero on entry) gets bit 0 for a row
 * index outside [0, N).  Rows must be distinct. */
int fgs_sparse_adam(int64_t n, int32_t D, const int64_t* index, const uint8_t* flag, int64_t N, float* param,
                    const float* param_sel, const float* grad, float* exp_avg, float* exp_avg_sq,
                    float* max_exp_avg_sq, double lr, double beta1, double beta2, double eps, double step,
                    int32_t* bad, void* stream);

/* Counter.update_by_output for one render (LoG/model/counter.py:36-83):
 * visible row j (point visible_index[j]) with radii[j] > 0 adds to
 * create_steps, visible_count, weights_sum and maxes weights_max, radii_max;
 * max-weight point k (row point_id[k], point_count[k] pixels) adds its area
 * to area_sum, |grad[j, :2]| * area to grad_sum (grad [n_vis, 3], may be
 * null) and maxes radii_max_max.  Counters are the Counter buffers over N
 * points; visible rows and point ids must be distinct.  *bad (device, zero
 * on entry): bit 0 point index outside [0, N), bit 1 point id outside the
 * rows. */
int fgs_counter_update(int64_t n_vis, const int64_t* visible_index, const int32_t* radii,
                       const float* point_weight, int64_t n_pid, const int64_t* point_id,
                       const int64_t* point_count, const float* grad, int64_t N, int32_t* create_steps,
                       int16_t* visible_count, float* weights_max, float* weights_sum, int16_t* radii_max,
                       int32_t* radii_max_max, int32_t* area_sum, float* grad_sum, int32_t* bad, void* stream);

/* FilterGS render rows (FilterGS/filtergs.py:98-223: get_all, then render):
 * the rasterizer draws row index[i] of the raw parameter arrays for i < N,
 * activated while it is loaded -- scale exp(scaling), opacity
 * sigmoid(opacity), rotation normalised, colour SH2RGB(colors) +
 * eval_sh_wobase(dir, shs) with no clamp (LoG/model/activation.py:36-43,
 * LoG/model/sh_utils.py:31-68) used as colors_precomp.  A row with a
 * non-finite value, a zero scale or a zero opacity is culled, as render()
 * drops it (:150-167).  No gathered copy of the rows is made. */
typedef struct {
    int32_t N;                 /* rows drawn */
    const int64_t* index;      /* [N] rows of the arrays below */
    int64_t n_src;             /* rows of the arrays */
    const float* xyz;          /* [n_src,3] */
    const float* scaling;      /* [n_src,3] log scales */
    const float* rotation;     /* [n_src,4] quaternions, not normalised */
    const float* opacity;      /* [n_src] logits */
    const float* colors;       /* [n_src,3] DC coefficients */
    const float* shs;          /* [n_src,rest_coeffs,3] higher coefficients, or null */
    int32_t rest_coeffs;
    int32_t sh_degree;         /* active degree, 0..4 */
    const float* campos;       /* device [3], camera centre (sh_degree > 0) */
} fgs_filtergs_rows;

/* Phase 1 for FilterGS rows; replaces gather_attributes
 * (FilterGS/cuda/gather_attributes_kernel.cu:6-145), the activation passes and
 * the preprocess of the rasterizer call.  Buffers and counts as
 * fgs_forward_geometry with P = N; counts[2] bit 1 flags an index outside
 * [0, n_src) (that row is culled). */
int fgs_filtergs_geometry(const fgs_raster_settings* s, const fgs_filtergs_rows* f, void* geom,
                          int32_t* radii, uint64_t* counts, void* stream);
/* Phase 2 for FilterGS rows: fgs_forward_render_stats with the colours of the
 * activation (stats may be null). */
int fgs_filtergs_render(const fgs_raster_settings* s, const fgs_filtergs_rows* f, const int32_t* radii,
                        void* geom, int64_t K, void* binning, void* image, float* out_color,
                        int32_t* out_point_id, float* out_weight_pixel, float* out_point_weight,
                        double* stats, float kpc_used_thresh, void* stream);

/* Sorted kv list of phase 2, for inspection: keys[k] = tile << 32 |
 * float_bits(depth of vals[k]) exactly as the reference's point_list_keys,
 * vals[k] = point_list.  Either output may be null. */
int fgs_read_sorted(const void* geom, int32_t P, int64_t K, const void* binning, uint64_t* keys,
                    uint32_t* vals, void* stream);
/* Byte offsets inside a geometry buffer of P Gaussians of: means2D (float2),
 * depths (float), conic_opacity (float4), rgb (float[3]), tiles_touched
 * (uint32) -- the preprocess outputs, for inspection. */
void fgs_geom_offsets(int32_t P, size_t* out5);
/* Byte offsets inside an image buffer of: final T (float [H*W]), n_contrib
 * (uint32 [H*W]), tile ranges (uint32 pairs [T,2]) -- the per-pixel state
 * renderCUDA leaves for the backward (forward.cu:552-553), for inspection. */
void fgs_image_offsets(int32_t image_width, int32_t image_height, size_t* out3);
/* Device pointer to the tile ranges (uint32 pairs [T,2]) inside an image buffer. */
const uint32_t* fgs_image_ranges(const void* image, int32_t image_width, int32_t image_height);

/* Backward -- replaces _C.rasterize_gaussians_backward /
 * RasterizeGaussiansBackwardCUDA (rasterize_points.cu:181-260) and
 * Rasterizer::backward (rasterizer_impl.cu:582-676): renderCUDA backward
 * (backward.cu:399-557), computeCov2DCUDA (:144-274) and preprocessCUDA
 * (:346-396).  dL_dpix [3,H,W].  dL_dmeans2D [P,3], dL_dcolors [P,3],
 * dL_dopacity [P] must be zero on entry (they are accumulated);
 * dL_dmeans3D [P,3], dL_dcov3D [P,6], dL_dsh [P,M,3], dL_dscales [P,3],
 * dL_drot [P,4] are fully written.  ws: fgs_backward_ws_bytes(P) bytes. */
size_t fgs_backward_ws_bytes(int32_t P);
int fgs_backward(const fgs_raster_settings* s, const fgs_gaussians* g, const int32_t* radii,
                 const void* geom, int64_t K, const void* binning, const void* image,
                 const float* dL_dpix, float* dL_dmeans2D, float* dL_dcolors, float* dL_dopacity,
                 float* dL_dmeans3D, float* dL_dcov3D, float* dL_dsh, float* dL_dscales,
                 float* dL_drot, void* ws, void* stream);

/* fgs_backward without the workspace argument.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The workspace (dL/dconic, 12 bytes per Gaussian) is a static
 * device array of the library, 48 MiB on every device the library runs on,
 * held until the process ends; calls with more than
 * FGS_NOALLOC_MAX_GAUSSIANS Gaussians fail with FGS_E_CAPACITY.  The array
 * is cleared by every call before use.  Every call on one device shares
 * that device's array: calls that overlap on one device (concurrent
 * streams, threads) race and must be serialised by the caller.  The stream
 * must belong to the current device.  Allocates nothing. */
#define FGS_NOALLOC_MAX_GAUSSIANS 4194304
int fgs_backward_noalloc(const fgs_raster_settings* s, const fgs_gaussians* g, const int32_t* radii,
                         const void* geom, int64_t K, const void* binning, const void* image,
                         const float* dL_dpix, float* dL_dmeans2D, float* dL_dcolors,
                         float* dL_dopacity, float* dL_dmeans3D, float* dL_dcov3D, float* dL_dsh,
                         float* dL_dscales, float* dL_drot, void* stream);

/* ---- batched multi-view path --------------------------------------------
 * No single counterpart in the reference: it replaces V sequential
 * _C.rasterize_gaussians / _C.rasterize_gaussians_backward calls
 * (rasterize_points.cu:36-260) over the same Gaussians, e.g. the per-camera
 * loop of a training step.  One launch per stage covers every view:
 *   - preprocess reads each Gaussian once and projects it into all views;
 *   - one depth sort orders all views ((view, depth, index) keys);
 *   - duplicate / tile sort / ranges run over all views at once; pairs keep
 *     their view's own tile numbers, so the stable tile sort needs only the
 *     bits of the largest view's tile count and leaves each (view, tile)
 *     range contiguous; ranges are stored in the batch's tile numbering
 *     (view v owns tiles [tile_base_v, tile_base_v + T_v));
 *   - one blend launch renders every tile of every view;
 *   - the preprocess backward folds the per-view gradients into one set of
 *     parameter gradients.
 * Per view the results equal those of the plain path: identical images,
 * radii, n_contrib and kv order; gradients up to float summation order.
 * The kv-pair buffer has a caller-chosen capacity K_cap, so a step issues no
 * host read and can be captured in a graph; counts[1] reports the true K,
 * and a step with K > K_cap must be redone with a larger buffer. */

/* One camera of a batch (host memory).  Matrices are column-major as in
 * fgs_raster_settings; campos and bg by value. */
typedef struct {
    int32_t image_height;
    int32_t image_width;
    float tanfovx;
    float tanfovy;
    float viewmatrix[16];
    float projmatrix[16];
    float campos[3];
    float bg[3];
} fgs_batch_view;

typedef struct {
    int32_t views;
    int32_t P;
    int64_t tiles;            /* tiles over all views                        */
    int64_t max_view_tiles;   /* tiles of the largest view                   */
    int64_t pixels;           /* pixels over all views                       */
    size_t plan_bytes;        /* device bytes of the plan (per-view records) */
    size_t geom_bytes;        /* device bytes of the batch geometry buffer   */
    size_t image_bytes;       /* device bytes of the batch image buffer      */
    size_t backward_ws_bytes; /* device bytes of the backward workspace      */
} fgs_batch_sizes;

/* Sizes for V views of P Gaussians (V <= 64, P <= 2^26, V*P64 < 2^31 with
 * P64 = P rounded up to 64, at most 2^26 tiles per view).
 * Element layout of the per-view records and screen-space gradients of a
 * batch: Gaussian i seen by view v is element v*P64 + 64*(i/64) + r, r = the
 * rank of i among the Gaussians of its 64-group that view v sees (radius > 0
 * in radii[v*P + i]); culled (view, Gaussian) pairs have no element. */
int fgs_batch_sizes_query(const fgs_batch_view* views, int32_t V, int32_t P, fgs_batch_sizes* out);
size_t fgs_batch_binning_bytes(int64_t K_cap);
/* Uploads the per-view records into plan (plan_bytes, device).  Synchronises
 * the stream: call it when the cameras change, outside graph capture. */
int fgs_batch_plan(const fgs_batch_view* views, int32_t V, void* plan, void* stream);
/* Forward of every view.  s supplies the settings shared by the views
 * (scale_modifier, sh_degree, prefiltered, debug, variant, use_filter, OGFS
 * fields, grad_mode); its image size, tan-fov, matrices, bg and campos are
 * ignored.  g->colors_precomp, when given, is per view: [V,P,3].
 * Outputs, views concatenated in plan order: radii [V,P]; out_color the
 * views' [3,H,W] blocks; out_point_id / out_weight_pixel the views' [H,W]
 * blocks; out_point_weight [V,P] (zero on entry).  counts[4] (device u64):
 * visible (Gaussian, view) pairs, K, error bits, 0. */
int fgs_batch_forward(const fgs_raster_settings* s, const fgs_gaussians* g, const fgs_batch_sizes* sz,
                      const void* plan, void* geom, void* image, void* binning, int64_t K_cap,
                      int32_t* radii, float* out_color, int32_t* out_point_id, float* out_weight_pixel,
                      float* out_point_weight, uint64_t* counts, void* stream);
/* Backward of every view.  dL_dpix: the views' [3,H,W] blocks.  Per view,
 * in the element layout (V*P64*3 floats each, written for every element,
 * the slots of culled pairs left as they were): dL_dmeans2D, dL_dcolors.
 * Summed over views (fully written): dL_dopacity [P], dL_dmeans3D [P,3],
 * dL_dcov3D [P,6], dL_dsh [P,M,3], dL_dscales [P,3], dL_drot [P,4].
 * ws: backward_ws_bytes. */
int fgs_batch_backward(const fgs_raster_settings* s, const fgs_gaussians* g, const fgs_batch_sizes* sz,
                       const void* plan, const int32_t* radii, const void* geom, const void* image,
                       const void* binning, int64_t K_cap, const float* dL_dpix, float* dL_dmeans2D,
                       float* dL_dcolors, float* dL_dopacity, float* dL_dmeans3D, float* dL_dcov3D,
                       float* dL_dsh, float* dL_dscales, float* dL_drot, void* ws, void* stream);

/* Byte offsets inside the batch buffers, for inspection (tests, tools):
 * out4[0] the sorted kv pairs in the binning buffer of capacity K_cap: K_cap
 * u32 tile keys ((view-local tile << 6) | v), then K_cap u32 elements
 * (v * P + i), min(K, K_cap) of each valid;
 * out4[1] the tile ranges in the image buffer (uint2 per batch tile, views
 * in plan order); out4[2] final_T and out4[3] n_contrib there (per pixel,
 * views concatenated).  No counterpart in the reference: its per-view
 * equivalents are the buffers RasterizeGaussiansCUDA returns
 * (rasterize_points.cu:36-179). */
int fgs_batch_offsets(const fgs_batch_sizes* sz, int64_t K_cap, size_t* out4);

/* ---- per-point utilities ------------------------------------------------ */
/* _C.mark_visible / markVisible (rasterize_points.cu:299-318,
 * rasterizer_impl.cu:97-109): present[i] = z_view > 0.2. */
int fgs_mark_visible(int32_t P, const float* means3D, const float* viewmatrix,
                     const float* projmatrix, uint8_t* present, void* stream);

/* _C.compute_radius (rasterize_points.cu:263-296, forward.cu:187-263). */
int fgs_compute_radius(int32_t P, const float* means3D, const float* scales,
                       const float* rotations, const float* viewmatrix, const float* projmatrix,
                       float tanfovx, float tanfovy, int32_t image_height, int32_t image_width,
                       float* radii, void* stream);

/* compute_radius_module.compute_radius of FilterGS
 * (FilterGS/cuda/compute_radius_kernel_v2.cu:180-229, :355-380). */
int fgs_compute_radius_v2(int32_t P, const float* means3D, const float* scales,
                          const float* rotations, const float* projmatrix,
                          const float* viewmatrix, float focal_x, float focal_y, float tan_fovx,
                          float tan_fovy, float* radii, void* stream);

/* compute_visibility_radius -- the LoD cut
 * (FilterGS/cuda/compute_radius_kernel_v2.cu:231-353, :382-430;
 *  FilterGS/cuda/compute_radius.py:23-45).  ancestor_path is int32 [N, L].
 * mask is bool/uint8 [N]; ws: fgs_lod_ws_bytes(N) bytes. */
size_t fgs_lod_ws_bytes(int32_t N);
int fgs_compute_visibility_radius(int32_t N, const float* means3D, const float* log_scales,
                                  const float* rotations, const float* projmatrix,
                                  const float* viewmatrix, const int32_t* tree_node_index,
                                  const int32_t* ancestor_path, int32_t L, float focal_x,
                                  float focal_y, float tan_fovx, float tan_fovy, float padding,
                                  float radius_max, float radius_filter_threshold, uint8_t* mask,
                                  float* radius2d, float* depth, void* ws, void* stream);

/* visible_flag_cuda / visible_flag_cuda_batch
 * (FilterGS/cuda/visible_flag_kernel.cu:19-81, :165-287): B batches of N
 * points sharing one projection matrix (B = 1 for the unbatched call). */
int fgs_visible_flag(int32_t B, int32_t N, const float* xyz, const float* proj_matrix,
                     float padding, uint8_t* valid, float* depth, float* p_proj, void* stream);

/* gather_attributes_cuda (FilterGS/cuda/gather_attributes_kernel.cu:70-145).
 * shs / shs_out may be null; sh_floats = coefficients*channels per row.
 * bad_index (device int, zero on entry) is set when an index is outside
 * [0, n_src). */
int fgs_gather_attributes(int64_t N, const int64_t* index, int64_t n_src, const float* xyz,
                          const float* scaling, const float* rotation, const float* opacity,
                          const float* colors, const float* shs, int32_t sh_floats,
                          float* xyz_out, float* scaling_out, float* rotation_out,
                          float* opacity_out, float* colors_out, float* shs_out, int32_t* bad_index,
                          void* stream);

/* FilterGS.get_all (FilterGS/filtergs.py:98-134): gather_attributes of the
 * rows in index, then Activation.activate_root_return
 * (LoG/model/activation.py:36-43) -- scaling = exp, opacity = sigmoid,
 * rotation = q / max(|q|, 1e-12), colors = SH2RGB(dc) + eval_sh_wobase
 * (LoG/model/sh_utils.py:31-68, no clamp) toward campos (device float[3])
 * for active_sh_degree > 0 -- fused into one pass.  features_rest is
 * [n_src, rest_coeffs, 3] (coefficients 1..); bad_index (device int, zero on
 * entry) is set when an index is outside [0, n_src). */
int fgs_activate_gather(int64_t N, const int64_t* index, int64_t n_src, const float* xyz,
                        const float* log_scaling, const float* rotation, const float* opacity_logit,
                        const float* features_dc, const float* features_rest, int32_t rest_coeffs,
                        int32_t active_sh_degree, const float* campos, float* xyz_out,
                        float* scaling_out, float* rotation_out, float* opacity_out, float* colors_out,
                        int32_t* bad_index, void* stream);
/* index may be null: row j is source row j (N <= n_src), the in-place rows of
 * LoG.get_all's gathered leaves (LoG/model/level_of_gaussian.py:260-294);
 * xyz_out may be null (no copy of the means). */

/* Backward of Activation.activate_root_return (LoG/model/activation.py:27-43)
 * for rows given in place, as LoG.get_all's autograd graph computes it
 * (level_of_gaussian.py:260-294): from the gradients of the activated
 * scaling [N,3], rotation [N,4], opacity [N] and colors [N,3] (any may be
 * null: zero), the gradients of the raw log-scales, quaternions, opacity
 * logits, SH DC [N,3] and SH rest [N, rest_coeffs, 3] (any output may be
 * null: not written).  Rotation: (g - y (y.g)) / |q| with y = q / |q|, or
 * g / 1e-12 below F.normalize's eps; rest coefficients past the active degree
 * get 0; the SH direction is the detached mean minus campos, so no mean
 * gradient arises. */
int fgs_activate_backward(int64_t N, const float* xyz, const float* log_scaling, const float* rotation,
                          const float* opacity_logit, int32_t rest_coeffs, int32_t active_sh_degree,
                          const float* campos, const float* dL_dscaling, const float* dL_drotation,
                          const float* dL_dopacity, const float* dL_dcolors, float* dL_dlog_scaling,
                          float* dL_drotation_raw, float* dL_dopacity_logit, float* dL_dfeatures_dc,
                          float* dL_dfeatures_rest, void* stream);

/* ---- misc ---------------------------------------------------------------- */
/* Per-stage timing with HIP events recorded on the launch stream (off by
 * default; not for use under graph capture).  Stages, in order: preprocess,
 * depth sort, tile offsets, duplicate, tile sort, ranges, blend forward,
 * blend backward, preprocess backward.  fgs_profile_read fills up to n
 * accumulated milliseconds / call counts and returns the number of stages. */
int fgs_profile_enable(int32_t on);
int fgs_profile_read(double* ms, uint64_t* calls, int32_t n, int32_t reset);
/* Blend work shape, process-wide: Gaussians staged in LDS per batch by the
 * forward blend (128 or 256; default 256) and the backward blend (64, 128
 * or 256; default 128).  Images and integer outputs do not depend on it (the
 * 16x16 binning tile is fixed); gradients change only in float summation
 * order.  No counterpart in the reference, whose renderCUDA
 * (forward.cu:406, backward.cu:399) fixes BLOCK_SIZE batches; the
 * autotuner picks it.  fgs_get_work_shape fills {fwd_batch, bwd_batch}. */
int fgs_set_work_shape(int32_t fwd_batch, int32_t bwd_batch);
int fgs_get_work_shape(int32_t* out2);
/* Forward blend walk, process-wide: 0 adaptive (default: each wave starts
 * with the wave-uniform walk and its DPP maximum, and switches to the
 * divergent walk with per-lane LDS maxima when few of its lanes blend per
 * entry), 1 wave-uniform only, 2 divergent only.  All three render the same
 * bits; only the speed differs.  No counterpart in the reference. */
int fgs_set_blend_walk(int32_t mode);
/* Backward tile layout (process-wide; no counterpart in the reference):
 * 0 -- four waves per 16x16 tile, each wave one 16x4 strip, the waves'
 *      per-Gaussian sums met with LDS float atomics;
 * 1 -- one wave per tile, four pixels per lane (one per strip): the strips'
 *      terms are added in registers and reduced across the wave once per
 *      (tile, Gaussian), in a fixed order (the default).
 * Both compute the same per-pixel terms; gradients differ by float
 * summation order only. */
int fgs_set_backward_layout(int32_t layout);
int fgs_get_backward_layout(void);
/* Forward tile layout without statistics (process-wide; no counterpart in
 * the reference): 0 -- four waves per tile with per-wave compacted lists;
 * 1 -- one wave per tile, four pixels per lane; 2 -- layout 1 for launches
 * of at least 12288 tiles, else 0 (the default: single waves per tile need
 * that many tiles to fill the GPU).  Identical outputs, bit for bit; the
 * forward with statistics keeps layout 0. */
int fgs_set_forward_layout(int32_t layout);
int fgs_get_forward_layout(void);
/* Tile order of the batched one-wave blend kernels (process-wide; no
 * counterpart in the reference): 0 launch order; C > 0 gives each XCD runs
 * of C consecutive tiles, so that neighbouring tiles share an L2 (default
 * 32).  Outputs do not depend on it. */
int fgs_set_xcd_chunk(int32_t chunk);
/* Batched one-wave blends as persistent launches (process-wide; no
 * counterpart in the reference, which renders one view per launch,
 * LoG/render/renderer.py:234): with chunk C > 0 the forward and the
 * backward each run one grid of as many one-wave workgroups as the GPU
 * holds at once, which claim the batch's (view, tile) items tile by tile
 * from a work queue of runs of C consecutive tiles, dealt round robin to
 * one queue per XCD.  by_cost = 1 orders the runs by estimated cost,
 * heaviest first (kv count for the forward, visible-list length for the
 * backward); 0 keeps the launch order.  chunk 0 (the default) launches one
 * workgroup per tile.  Outputs do not depend on it. */
int fgs_set_tile_queue(int32_t chunk, int32_t by_cost);
int fgs_get_tile_queue(void);
const char* fgs_last_error(void);
int fgs_abi_version(void);
/* OGFS log(lambda_G) for the given epsilons (forward.cu:346-351). */
float fgs_ogfs_log_lambda(float ogfs_epsilon_scale, float ogfs_epsilon_max);

#ifdef __cplusplus
}
#endif
#endif /* FILTERGS_AMD_H */
