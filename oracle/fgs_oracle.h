/*
 * fgs_oracle.h -- CPU oracle for the FilterGS rasterizer / LoD-cut hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (filtergs_amd) never links, loads or calls it.
 *
 * It restates the reference equations in plain C, in two arithmetic modes:
 *   *_f32  FP32 emulation: the same operation order as the GPU kernels, no FMA
 *          contraction (built with -ffp-contract=off), correctly rounded
 *          division/sqrt, and transcendental functions evaluated as the
 *          correctly rounded float of the double-precision libm result.  This
 *          is the bit-exact reference for every integer-producing stage
 *          (radius, tile rectangle, depth key, sort order, tile ranges, LoD
 *          mask).
 *   *_f64  FP64: the same equations in double precision; the correctness
 *          reference for images and gradients.  Thresholds use the float
 *          constants of the reference (widened), so both modes compute the
 *          same function, the f64 one with less rounding error.
 *
 * Reference files followed (paths relative to the reference repository):
 *   antialias/cuda_rasterizer/forward.cu, backward.cu, auxiliary.h, config.h,
 *   rasterizer_impl.cu; antialias/rasterize_points.cu;
 *   FilterGS/cuda/compute_radius_kernel_v2.cu, visible_flag_kernel.cu,
 *   gather_attributes_kernel.cu.  Function-level citations are in fgs_oracle.c
 *   and orc_impl.h.
 *
 * Pinning status of each function is listed in the header of orc_impl.h.
 */
#ifndef FGS_ORACLE_H
#define FGS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Camera and rasterizer settings (mirrors GaussianRasterizationSettings,
 * antialias/diff_gaussian_rasterization_wodilate/__init__.py:193-210). */
typedef struct {
    int32_t image_height, image_width;
    float tanfovx, tanfovy;
    float bg[3];
    float scale_modifier;
    float viewmatrix[16];   /* column-major V, translation at [12..14] */
    float projmatrix[16];   /* column-major P*V (full projection)      */
    float campos[3];
    int32_t sh_degree;
    int32_t prefiltered;
    /* variant: 0 = wodilate (max(.,0.3) when use_filter, OGFS radius),
     *          1 = dilate (+0.3, 3-sigma radius, raw opacity). */
    int32_t variant;
    int32_t use_filter;
    float ogfs_epsilon_scale, ogfs_epsilon_max, ogfs_energy_floor, ogfs_radius_ratio_min;
    /* grad_mode: 0 = reference equations (backward.cu), 1 = exact derivative
     * of the forward actually computed. */
    int32_t grad_mode;
} orc_settings;

/* OGFS per-launch constant log(lambda_G) (forward.cu:346-351), float. */
float orc_ogfs_log_lambda(float eps_scale, float eps_max);
/* Correctly rounded float log/exp via double libm (helpers, exported for tests). */
float orc_logf_cr(float x);
float orc_expf_cr(float x);

/* ---- per-Gaussian preprocess (forward.cu:266-401) ------------------------ */
/* Inputs: means3D[P,3], scales[P,3] or NULL, rotations[P,4] or NULL,
 * opacities[P], shs[P,M,3] or NULL, colors_precomp[P,3] or NULL,
 * cov3D_precomp[P,6] or NULL.  Outputs (all P-sized, zero for culled):
 * radii, means2D[2P], depths, cov3D[6P], rgb[3P], conic_opacity[4P],
 * tiles_touched, clamped[3P], rect[4P] (xmin,ymin,xmax,ymax). */
void orc_preprocess_f32(const orc_settings* s, int32_t P, int32_t M,
                        const float* means3D, const float* scales, const float* rotations,
                        const float* opacities, const float* shs, const float* colors_precomp,
                        const float* cov3D_precomp,
                        int32_t* radii, float* means2D, float* depths, float* cov3D, float* rgb,
                        float* conic_opacity, uint32_t* tiles_touched, uint8_t* clamped,
                        int32_t* rect, int32_t nthreads);
void orc_preprocess_f64(const orc_settings* s, int32_t P, int32_t M,
                        const double* means3D, const double* scales, const double* rotations,
                        const double* opacities, const double* shs, const double* colors_precomp,
                        const double* cov3D_precomp,
                        int32_t* radii, double* means2D, double* depths, double* cov3D, double* rgb,
                        double* conic_opacity, uint32_t* tiles_touched, uint8_t* clamped,
                        int32_t* rect, int32_t nthreads);

/* ---- binning: duplicateWithKeys + stable radix sort + identifyTileRanges --
 * (rasterizer_impl.cu:113-147, 427-445, 172-194).  Integer stage, one mode.
 * Gaussian i contributes iff radii[i] > 0; its tile rectangle is rect[4i..].
 * keys/vals may be NULL to query K only.  ranges has 2*T entries. */
int64_t orc_num_rendered(int32_t P, const int32_t* radii, const int32_t* rect);
void orc_bin(int32_t P, int32_t grid_x, int32_t grid_y, const int32_t* radii, const int32_t* rect,
             const float* depths, uint64_t* keys, uint32_t* vals, uint32_t* ranges);
/* getRect (auxiliary.h:46-56) on one point, for tests. */
void orc_get_rect(float px, float py, int32_t radius, int32_t grid_x, int32_t grid_y, int32_t* out4);
/* getHigherMsb (rasterizer_impl.cu:78-93) == bit_length for 1 <= n < 2^21. */
uint32_t orc_higher_msb(uint32_t n);

/* ---- per-tile alpha blending forward (forward.cu:406-567) ---------------- */
/* colors[3P] (rgb or colors_precomp), ranges[2T], point_list[K].
 * Outputs: out_color[3,H,W], final_T[H*W], n_contrib[H*W], point_id[H*W],
 * weight_pixel[H*W], point_weight[P] (exact max of T*alpha; caller zeroes),
 * kv_alpha_sum[K], kv_alpha_count[K] (optional, caller zeroes),
 * decision_hash[H*W] (optional; a checksum of the per-pixel blend decisions). */
void orc_blend_fwd_f32(const orc_settings* s, int32_t P, const uint32_t* ranges,
                       const uint32_t* point_list, const float* means2D, const float* colors,
                       const float* conic_opacity, float* out_color, float* final_T,
                       uint32_t* n_contrib, int32_t* point_id, float* weight_pixel,
                       float* point_weight, float* kv_alpha_sum, uint32_t* kv_alpha_count,
                       uint64_t* decision_hash, int32_t nthreads);
void orc_blend_fwd_f64(const orc_settings* s, int32_t P, const uint32_t* ranges,
                       const uint32_t* point_list, const double* means2D, const double* colors,
                       const double* conic_opacity, double* out_color, double* final_T,
                       uint32_t* n_contrib, int32_t* point_id, double* weight_pixel,
                       double* point_weight, double* kv_alpha_sum, uint32_t* kv_alpha_count,
                       uint64_t* decision_hash, int32_t nthreads);

/* ---- blend backward (backward.cu:399-557) -------------------------------- */
/* dL_dpix[3,H,W].  Accumulates (caller zeroes): dL_dmean2D[3P] (x,y used),
 * dL_dconic[3P] (a, b/2-convention, c as in backward.cu:549-551),
 * dL_dopacity[P], dL_dcolors[3P].  nthreads > 1 splits tiles over threads
 * with per-thread accumulators (10 values for each Gaussian of the list)
 * summed in thread order. */
void orc_blend_bwd_f32(const orc_settings* s, int32_t P, const uint32_t* ranges,
                       const uint32_t* point_list, const float* means2D, const float* colors,
                       const float* conic_opacity, const float* final_T, const uint32_t* n_contrib,
                       const float* dL_dpix, float* dL_dmean2D, float* dL_dconic,
                       float* dL_dopacity, float* dL_dcolors, int32_t nthreads);
void orc_blend_bwd_f64(const orc_settings* s, int32_t P, const uint32_t* ranges,
                       const uint32_t* point_list, const double* means2D, const double* colors,
                       const double* conic_opacity, const double* final_T, const uint32_t* n_contrib,
                       const double* dL_dpix, double* dL_dmean2D, double* dL_dconic,
                       double* dL_dopacity, double* dL_dcolors, int32_t nthreads);

/* ---- preprocess backward (backward.cu:144-396) --------------------------- */
/* Inputs as preprocess plus radii (the mask), cov3D (computed or precomputed),
 * the blend-backward accumulators, opacities (exact mode only).
 * Outputs (caller zeroes): dL_dmeans3D[3P], dL_dcov3D[6P], dL_dsh[P*M*3],
 * dL_dscales[3P], dL_drot[4P]; dL_dcolors is read (and, for SH input, its
 * clamped channels are not modified); dL_dopacity is modified in exact mode. */
void orc_preprocess_bwd_f32(const orc_settings* s, int32_t P, int32_t M, const float* means3D,
                            const float* scales, const float* rotations, const float* opacities,
                            const float* shs, const float* cov3D, const int32_t* radii,
                            const float* dL_dmean2D, const float* dL_dconic, float* dL_dopacity,
                            const float* dL_dcolors, float* dL_dmeans3D, float* dL_dcov3D,
                            float* dL_dsh, float* dL_dscales, float* dL_drot, int32_t nthreads);
void orc_preprocess_bwd_f64(const orc_settings* s, int32_t P, int32_t M, const double* means3D,
                            const double* scales, const double* rotations, const double* opacities,
                            const double* shs, const double* cov3D, const int32_t* radii,
                            const double* dL_dmean2D, const double* dL_dconic, double* dL_dopacity,
                            const double* dL_dcolors, double* dL_dmeans3D, double* dL_dcov3D,
                            double* dL_dsh, double* dL_dscales, double* dL_drot, int32_t nthreads);

/* ---- LoD cut and neighbours (FilterGS/cuda, antialias compute_radius) ---- */
/* compute_visibility_radius (compute_radius_kernel_v2.cu:109-169, 231-353). */
void orc_lod_cut(int32_t N, const float* means3d, const float* log_scales, const float* rotations,
                 const float* projmatrix, const float* viewmatrix,
                 const int32_t* tree_node_index, const int32_t* ancestor_path, int32_t L,
                 float focal_x, float focal_y, float tanx, float tany, float padding,
                 float radius_max, float radius_filter_threshold,
                 uint8_t* mask, float* radius2d, float* depth, int32_t nthreads);
/* FilterGS compute_radius (compute_radius_kernel_v2.cu:180-229), raw inputs. */
void orc_radius_v2(int32_t P, const float* means3d, const float* scales, const float* rotations,
                   const float* projmatrix, const float* viewmatrix, float focal_x, float focal_y,
                   float tanx, float tany, float* radii);
/* antialias compute_radius (forward.cu:187-263). */
void orc_radius_aa(int32_t P, const float* means3d, const float* scales, const float* rotations,
                   const float* viewmatrix, const float* projmatrix, float tanx, float tany,
                   int32_t H, int32_t W, float* radii);
/* markVisible (rasterizer_impl.cu:97-109). */
void orc_mark_visible(int32_t P, const float* means3d, const float* viewmatrix,
                      const float* projmatrix, uint8_t* present);
/* visible_flag_kernel (visible_flag_kernel.cu:19-81). */
void orc_visible_flag(int32_t N, const float* xyz, const float* proj, float padding,
                      uint8_t* valid, float* depth, float* p_proj);

#ifdef __cplusplus
}
#endif
#endif
