/*
 * fgs_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see fgs_oracle.h).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  The FP32
 * emulation relies on -ffp-contract=off and on SSE single-precision
 * arithmetic (FLT_EVAL_METHOD == 0 on x86-64).
 */
#include "fgs_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#if FLT_EVAL_METHOD != 0
#error "the FP32 emulation needs FLT_EVAL_METHOD == 0"
#endif

/* Saturating float->int truncation (the GPU's v_cvt_i32_f32 behaviour made
 * explicit; NaN and huge values map to values that the [0, grid] clamp sends
 * to the same tile as the reference kernel would). */
static inline int orc_sat_int(double v)
{
    if (!(v > -1.0e9)) return -1000000000;
    if (v > 1.0e9) return 1000000000;
    return (int)v;
}
static inline int orc_clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

float orc_logf_cr(float x) { return (float)log((double)x); }
float orc_expf_cr(float x) { return (float)exp((double)x); }

/* OGFS lambda_G (forward.cu:346-351): a per-launch constant. */
float orc_ogfs_log_lambda(float eps_scale, float eps_max)
{
    const float epsilon_pix = 1.0f / 255.0f;
    const float eps_base = fmaxf(0.0f, 1.0f - epsilon_pix);
    const float lambda_raw = 1.0f - (float)pow((double)eps_base, 256.0);
    float lambda_g = epsilon_pix + eps_scale * (lambda_raw - epsilon_pix);
    lambda_g = fminf(fmaxf(lambda_g, epsilon_pix), eps_max);
    return orc_logf_cr(lambda_g);
}

/* ---- FP32 emulation instantiation -------------------------------------- */
#define R float
#define SFX(n) n##_f32
#define RMIN fminf
#define RMAX fmaxf
#define RSQRT sqrtf
#define RCEIL ceilf
#define RLOG orc_logf_cr
#define REXP expf
#include "orc_impl.h"
#undef R
#undef SFX
#undef RMIN
#undef RMAX
#undef RSQRT
#undef RCEIL
#undef RLOG
#undef REXP

/* ---- FP64 instantiation ------------------------------------------------- */
#define R double
#define SFX(n) n##_f64
#define RMIN fmin
#define RMAX fmax
#define RSQRT sqrt
#define RCEIL ceil
#define RLOG log
#define REXP exp
#include "orc_impl.h"
#undef R
#undef SFX
#undef RMIN
#undef RMAX
#undef RSQRT
#undef RCEIL
#undef RLOG
#undef REXP

/* ---- integer stage: binning ------------------------------------------- */
void orc_get_rect(float px, float py, int32_t radius, int32_t grid_x, int32_t grid_y, int32_t* out4)
{
    get_rect_f32(px, py, radius, grid_x, grid_y, out4);
}

/* getHigherMsb (rasterizer_impl.cu:78-93), restated as written. */
uint32_t orc_higher_msb(uint32_t n)
{
    uint32_t msb = sizeof(n) * 4;
    uint32_t step = msb;
    while (step > 1) {
        step /= 2;
        if (n >> msb) msb += step;
        else msb -= step;
    }
    if (n >> msb) msb++;
    return msb;
}

int64_t orc_num_rendered(int32_t P, const int32_t* radii, const int32_t* rect)
{
    int64_t K = 0;
    for (int i = 0; i < P; ++i)
        if (radii[i] > 0)
            K += (int64_t)(rect[4 * i + 2] - rect[4 * i]) * (rect[4 * i + 3] - rect[4 * i + 1]);
    return K;
}

typedef struct { uint64_t key; uint64_t seq; uint32_t val; } orc_kv;
static int orc_kv_cmp(const void* a, const void* b)
{
    const orc_kv* x = (const orc_kv*)a;
    const orc_kv* y = (const orc_kv*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->seq < y->seq ? -1 : (x->seq > y->seq ? 1 : 0);
}

/* duplicateWithKeys (rasterizer_impl.cu:113-147) emits, per Gaussian in
 * index order, one (tile << 32 | float_bits(depth), idx) pair per tile of its
 * rectangle (y-major, x-minor).  SortPairs (rasterizer_impl.cu:430-435) is a
 * stable LSD radix sort, i.e. the result is ordered by (key, emission order).
 * identifyTileRanges (rasterizer_impl.cu:172-194) marks [first, last+1). */
void orc_bin(int32_t P, int32_t grid_x, int32_t grid_y, const int32_t* radii, const int32_t* rect,
             const float* depths, uint64_t* keys, uint32_t* vals, uint32_t* ranges)
{
    const int64_t K = orc_num_rendered(P, radii, rect);
    orc_kv* kv = (orc_kv*)malloc(sizeof(orc_kv) * (size_t)(K > 0 ? K : 1));
    int64_t off = 0;
    for (int i = 0; i < P; ++i) {
        if (radii[i] <= 0) continue;
        uint32_t dbits;
        memcpy(&dbits, &depths[i], 4);
        for (int y = rect[4 * i + 1]; y < rect[4 * i + 3]; ++y)
            for (int x = rect[4 * i]; x < rect[4 * i + 2]; ++x) {
                kv[off].key = ((uint64_t)(uint32_t)(y * grid_x + x) << 32) | dbits;
                kv[off].seq = (uint64_t)off;
                kv[off].val = (uint32_t)i;
                off++;
            }
    }
    qsort(kv, (size_t)K, sizeof(orc_kv), orc_kv_cmp);
    const int T = grid_x * grid_y;
    for (int t = 0; t < 2 * T; ++t) ranges[t] = 0;
    for (int64_t k = 0; k < K; ++k) {
        if (keys) keys[k] = kv[k].key;
        if (vals) vals[k] = kv[k].val;
        const uint32_t tile = (uint32_t)(kv[k].key >> 32);
        if (k == 0 || (uint32_t)(kv[k - 1].key >> 32) != tile) ranges[2 * tile] = (uint32_t)k;
        if (k == K - 1 || (uint32_t)(kv[k + 1].key >> 32) != tile) ranges[2 * tile + 1] = (uint32_t)(k + 1);
    }
    free(kv);
}

/* ---- LoD cut (FilterGS/cuda/compute_radius_kernel_v2.cu) ----------------- */
/* compute_visibility_radius_point (:109-169), FP32 emulation. */
static int orc_vis_radius_point(int idx, const float* pts, const float* sl, const float* rot,
                                const float* proj, const float* view, float fx, float fy,
                                float tanx, float tany, float padding, float* depth_out,
                                float* radius_out)
{
    const float px = pts[3 * idx], py = pts[3 * idx + 1], pz = pts[3 * idx + 2];
    float hom[4];
    xf4_f32(proj, px, py, pz, hom);
    const float p_w = 1.0f / (hom[3] + 0.0000001f);
    const float qx = hom[0] * p_w, qy = hom[1] * p_w, qz = hom[2] * p_w;
    *depth_out = qz;
    const int visible = (qz > 0.0f) && (qz < 1.0f) && (qx > -1.0f - padding) &&
                        (qx < 1.0f + padding) && (qy > -1.0f - padding) && (qy < 1.0f + padding);
    if (!visible) { *radius_out = 0.0f; return 0; }
    const float sx = orc_expf_cr(sl[3 * idx]), sy = orc_expf_cr(sl[3 * idx + 1]),
                sz = orc_expf_cr(sl[3 * idx + 2]);
    float r = rot[4 * idx], x = rot[4 * idx + 1], y = rot[4 * idx + 2], z = rot[4 * idx + 3];
    const float n = sqrtf(r * r + x * x + y * y + z * z);
    if (n > 0.f) { r /= n; x /= n; y /= n; z /= n; }
    else { r = 1.f; x = 0.f; y = 0.f; z = 0.f; }
    float c3[6];
    cov3d_f32(sx, sy, sz, r, x, y, z, c3);
    cov2d_t_f32 cv;
    cov2d_f32(view, px, py, pz, fx, fy, tanx, tany, c3, &cv);
    const float a = fmaxf(cv.a, 0.3f), b = cv.b, c = fmaxf(cv.c, 0.3f);
    const float det = a * c - b * b;
    if (det <= 0.0f) { *radius_out = 0.0f; return 1; }
    const float mid = 0.5f * (a + c);
    const float root = sqrtf(fmaxf(0.1f, mid * mid - det));
    const float l1 = mid + root, l2 = mid - root;
    *radius_out = 3.f * sqrtf(fmaxf(l1, l2));
    return 1;
}

/* visibility_radius_cuda (:231-353).  The lazy CAS publication of ancestor
 * radii computes the same deterministic value as the ancestor's own thread,
 * so the mask equals: internal node -> r <= radius_max; leaf -> no ancestor
 * a (row order, stop at -1, skip a == i, skip leaves) with 0 < r_a <= radius_max;
 * then the radius_filter_threshold rule. */
void orc_lod_cut(int32_t N, const float* means3d, const float* log_scales, const float* rotations,
                 const float* projmatrix, const float* viewmatrix,
                 const int32_t* tree_node_index, const int32_t* ancestor_path, int32_t L,
                 float focal_x, float focal_y, float tanx, float tany, float padding,
                 float radius_max, float radius_filter_threshold,
                 uint8_t* mask, float* radius2d, float* depth, int32_t nthreads)
{
    uint8_t* vis = (uint8_t*)malloc((size_t)(N > 0 ? N : 1));
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int i = 0; i < N; ++i) {
        float d, r;
        vis[i] = (uint8_t)orc_vis_radius_point(i, means3d, log_scales, rotations, projmatrix,
                                               viewmatrix, focal_x, focal_y, tanx, tany, padding,
                                               &d, &r);
        depth[i] = d;
        radius2d[i] = r;
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int i = 0; i < N; ++i) {
        if (!vis[i]) { mask[i] = 0; continue; }
        const float rv = radius2d[i];
        int keep = 1;
        if (tree_node_index[i] != -1) {
            keep = rv <= radius_max;
        } else {
            const int32_t* row = ancestor_path + (size_t)i * L;
            for (int lv = 0; lv < L; ++lv) {
                const int32_t a = row[lv];
                if (a < 0) break;
                if (a == i) continue;
                if (tree_node_index[a] == -1) continue;
                const float ra = radius2d[a];
                if (ra > 0.f && ra <= radius_max) { keep = 0; break; }
            }
        }
        if (keep && radius_filter_threshold > 0.0f && rv > radius_filter_threshold) keep = 0;
        mask[i] = (uint8_t)keep;
    }
    free(vis);
}

/* compute_radius_cuda of FilterGS (compute_radius_kernel_v2.cu:180-229; the
 * same kernel as LoG/cuda/compute_radius_kernel.cu:107-156) for one point:
 * scale and quaternion used as given. */
static float orc_radius_v2_point(const float* projmatrix, const float* viewmatrix, float focal_x,
                                 float focal_y, float tanx, float tany, float px, float py,
                                 float pz, float sx, float sy, float sz, float qr, float qx,
                                 float qy, float qz)
{
    float hom[4];
    xf4_f32(projmatrix, px, py, pz, hom);
    const float p_w = 1.0f / (hom[3] + 0.0000001f);
    const float nx = hom[0] * p_w, ny = hom[1] * p_w;
    if ((double)nx < -1.3 || (double)nx > 1.3 || (double)ny < -1.3 || (double)ny > 1.3) return 0.f;
    float c3[6];
    cov3d_f32(sx, sy, sz, qr, qx, qy, qz, c3);
    cov2d_t_f32 cv;
    cov2d_f32(viewmatrix, px, py, pz, focal_x, focal_y, tanx, tany, c3, &cv);
    const float a = fmaxf(cv.a, 0.3f), b = cv.b, c = fmaxf(cv.c, 0.3f);
    const float det = a * c - b * b;
    if (det == 0.0f) return 0.f;
    const float mid = 0.5f * (a + c);
    const float l1 = mid + sqrtf(fmaxf(0.1f, mid * mid - det));
    const float l2 = mid - sqrtf(fmaxf(0.1f, mid * mid - det));
    return 3.f * sqrtf(fmaxf(l1, l2));
}

void orc_radius_v2(int32_t P, const float* means3d, const float* scales, const float* rotations,
                   const float* projmatrix, const float* viewmatrix, float focal_x, float focal_y,
                   float tanx, float tany, float* radii)
{
    for (int i = 0; i < P; ++i)
        radii[i] = orc_radius_v2_point(projmatrix, viewmatrix, focal_x, focal_y, tanx, tany,
                                       means3d[3 * i], means3d[3 * i + 1], means3d[3 * i + 2],
                                       scales[3 * i], scales[3 * i + 1], scales[3 * i + 2],
                                       rotations[4 * i], rotations[4 * i + 1], rotations[4 * i + 2],
                                       rotations[4 * i + 3]);
}

/* LoG's Gaussian.compute_radius (level_of_gaussian.py:66-88) of the rows ids:
 * scaling = exp(raw) (correctly rounded), rotation = raw / max(|raw|, 1e-12)
 * (F.normalize, squares summed in index order), then compute_radius_cuda. */
void orc_log_radius(int64_t n, const int64_t* ids, const float* xyz, const float* scaling,
                    const float* rotation, const float* projmatrix, const float* viewmatrix,
                    float focal_x, float focal_y, float tanx, float tany, float* radii)
{
    for (int64_t k = 0; k < n; ++k) {
        const int64_t i = ids[k];
        const float* q = rotation + 4 * i;
        const float qn = fmaxf(sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]), 1e-12f);
        radii[k] = orc_radius_v2_point(projmatrix, viewmatrix, focal_x, focal_y, tanx, tany,
                                       xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2],
                                       orc_expf_cr(scaling[3 * i]), orc_expf_cr(scaling[3 * i + 1]),
                                       orc_expf_cr(scaling[3 * i + 2]), q[0] / qn, q[1] / qn,
                                       q[2] / qn, q[3] / qn);
    }
}

/* antialias compute_radius_cuda (forward.cu:187-263): near/1.3-NDC guard, no
 * dilation, mod 1, focal from the image size. */
void orc_radius_aa(int32_t P, const float* means3d, const float* scales, const float* rotations,
                   const float* viewmatrix, const float* projmatrix, float tanx, float tany,
                   int32_t H, int32_t W, float* radii)
{
    const float focal_y = (float)H / (2.0f * tany);
    const float focal_x = (float)W / (2.0f * tanx);
    for (int i = 0; i < P; ++i) {
        radii[i] = 0;
        const float px = means3d[3 * i], py = means3d[3 * i + 1], pz = means3d[3 * i + 2];
        float hom[4], view[3];
        xf4_f32(projmatrix, px, py, pz, hom);
        const float p_w = 1.0f / (hom[3] + 0.0000001f);
        const float qx = hom[0] * p_w, qy = hom[1] * p_w;
        xf3_f32(viewmatrix, px, py, pz, view);
        if (view[2] <= 0.2f || (double)qx < -1.3 || (double)qx > 1.3 || (double)qy < -1.3 ||
            (double)qy > 1.3)
            continue;
        float c3[6];
        cov3d_f32(scales[3 * i], scales[3 * i + 1], scales[3 * i + 2], rotations[4 * i],
                  rotations[4 * i + 1], rotations[4 * i + 2], rotations[4 * i + 3], c3);
        cov2d_t_f32 cv;
        cov2d_f32(viewmatrix, px, py, pz, focal_x, focal_y, tanx, tany, c3, &cv);
        const float a = cv.a, b = cv.b, c = cv.c;
        const float det = a * c - b * b;
        if (det == 0.0f) continue;
        const float mid = 0.5f * (a + c);
        const float l1 = mid + sqrtf(fmaxf(0.1f, mid * mid - det));
        const float l2 = mid - sqrtf(fmaxf(0.1f, mid * mid - det));
        radii[i] = 3.f * sqrtf(fmaxf(l1, l2));
    }
}

/* checkFrustum / in_frustum (rasterizer_impl.cu:97-109, auxiliary.h:139-164). */
void orc_mark_visible(int32_t P, const float* means3d, const float* viewmatrix,
                      const float* projmatrix, uint8_t* present)
{
    (void)projmatrix;
    for (int i = 0; i < P; ++i) {
        float view[3];
        xf3_f32(viewmatrix, means3d[3 * i], means3d[3 * i + 1], means3d[3 * i + 2], view);
        present[i] = !(view[2] <= 0.2f);
    }
}

/* visible_flag_kernel (visible_flag_kernel.cu:19-81), exact division. */
void orc_visible_flag(int32_t N, const float* xyz, const float* proj, float padding,
                      uint8_t* valid, float* depth, float* p_proj)
{
    for (int i = 0; i < N; ++i) {
        float hom[4];
        xf4_f32(proj, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], hom);
        const float w = hom[3];
        const float inv_w = 1.0f / (w + 1e-7f);
        const float x = hom[0] * inv_w, y = hom[1] * inv_w, z = hom[2] * inv_w;
        valid[i] = (w > 0.0f) && (w < 100.0f) && (x > -1.0f - padding) && (x < 1.0f + padding) &&
                   (y > -1.0f - padding) && (y < 1.0f + padding);
        depth[i] = w;
        p_proj[3 * i] = x;
        p_proj[3 * i + 1] = y;
        p_proj[3 * i + 2] = z;
    }
}
