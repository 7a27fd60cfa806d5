// preprocess.hip -- per-Gaussian forward and backward preprocessing.
//
// Forward: antialias/cuda_rasterizer/forward.cu:266-401 (preprocessCUDA)
//   with computeCov3D (:151-185), computeCov2D (:102-146),
//   computeColorFromSH (:48-99), in_frustum (auxiliary.h:139-164),
//   ndc2Pix (auxiliary.h:41-44) and getRect (auxiliary.h:46-56).
// Backward: antialias/cuda_rasterizer/backward.cu:144-274 (computeCov2DCUDA)
//   and :346-396 (preprocessCUDA), fused into one kernel.
//
// This file is compiled with -ffp-contract=off: the radius, rectangle and
// depth key are integer results of FP32 arithmetic and must not depend on
// FMA contraction.
#include "fgs_device.h"
#include "fgs_internal.h"
#include "fgs_activation.h"

namespace fgs {

__constant__ float kSH_C0 = 0.28209479177387814f;
__constant__ float kSH_C1 = 0.4886025119029199f;
__constant__ float kSH_C2[5] = { 1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                                 -1.0925484305920792f, 0.5462742152960396f };
__constant__ float kSH_C3[7] = { -0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                                 0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                                 -0.5900435899266435f };

// SH colour of one channel, before the +0.5 shift.  sh(f) returns float f of
// the Gaussian's coefficient row (coefficient k, channel c at f = 3k + c).
// The colour is no integer's input: its products and sums may fuse (as the
// reference's compiler fuses them), unlike the rest of this file.
template <typename Sh>
__device__ __forceinline__ float sh_eval(int deg, const Sh& sh, int c, float x, float y, float z)
{
#pragma clang fp contract(fast)
#define SH(k) sh((k) * 3 + c)
    float res = kSH_C0 * SH(0);
    if (deg > 0) {
        res = res - kSH_C1 * y * SH(1) + kSH_C1 * z * SH(2) - kSH_C1 * x * SH(3);
        if (deg > 1) {
            const float xx = x * x, yy = y * y, zz = z * z;
            const float xy = x * y, yz = y * z, xz = x * z;
            res = res + kSH_C2[0] * xy * SH(4) + kSH_C2[1] * yz * SH(5) +
                  kSH_C2[2] * (2.0f * zz - xx - yy) * SH(6) + kSH_C2[3] * xz * SH(7) +
                  kSH_C2[4] * (xx - yy) * SH(8);
            if (deg > 2) {
                res = res + kSH_C3[0] * y * (3.0f * xx - yy) * SH(9) + kSH_C3[1] * xy * z * SH(10) +
                      kSH_C3[2] * y * (4.0f * zz - xx - yy) * SH(11) +
                      kSH_C3[3] * z * (2.0f * zz - 3.0f * xx - 3.0f * yy) * SH(12) +
                      kSH_C3[4] * x * (4.0f * zz - xx - yy) * SH(13) +
                      kSH_C3[5] * z * (xx - yy) * SH(14) + kSH_C3[6] * x * (xx - 3.0f * yy) * SH(15);
            }
        }
    }
#undef SH
    return res;
}

// SH rows of a 256-Gaussian block staged in LDS, transposed (float f of the
// block's row t at s[f * kPitch + t]; the odd pitch keeps both the staging
// writes and the per-thread reads free of bank conflicts).  Rows are read
// from HBM as one contiguous, fully coalesced stream.
constexpr int kPreThreads = 256;
constexpr int kPitch = kPreThreads + 1;
constexpr int kMaxShFloats = 48;

struct ShLds {
    const float* s;
    int t;
    __device__ float operator()(int f) const { return s[f * kPitch + t]; }
};

// An SH row held in the thread's registers (float f at r[f]).
struct ShReg {
    const float* r;
    __device__ float operator()(int f) const { return r[f]; }
};

template <int NU>
__device__ __forceinline__ void stage_rows_fixed(const float* __restrict__ src, int rs, int nrows, float* s)
{
    const int total = nrows * NU;
    for (int k = threadIdx.x; k < total; k += kPreThreads) {
        const int t = k / NU;
        const int f = k - t * NU;
        s[f * kPitch + t] = src[(size_t)t * rs + f];
    }
}

// Contiguous rows (nu == rs, 16-byte aligned block): float4 loads.
template <int NU>
__device__ __forceinline__ void stage_rows_vec4(const float* __restrict__ src, int nrows, float* s)
{
    static_assert(NU % 4 == 0, "row length must be a multiple of 4 floats");
    const float4* src4 = reinterpret_cast<const float4*>(src);
    const int total4 = nrows * (NU / 4);
    for (int k = threadIdx.x; k < total4; k += kPreThreads) {
        const float4 v = src4[k];
        const int t = k / (NU / 4);
        const int f = (k - t * (NU / 4)) * 4;
        s[(f + 0) * kPitch + t] = v.x;
        s[(f + 1) * kPitch + t] = v.y;
        s[(f + 2) * kPitch + t] = v.z;
        s[(f + 3) * kPitch + t] = v.w;
    }
}

__device__ __forceinline__ void stage_rows(const float* __restrict__ src, int rs, int nu, int nrows, float* s)
{
    if (nu == 48 && rs == 48 && ((uintptr_t)src & 15) == 0) { stage_rows_vec4<48>(src, nrows, s); return; }
    if (nu == 12 && rs == 12 && ((uintptr_t)src & 15) == 0) { stage_rows_vec4<12>(src, nrows, s); return; }
    if (nu == 48) { stage_rows_fixed<48>(src, rs, nrows, s); return; }
    if (nu == 27) { stage_rows_fixed<27>(src, rs, nrows, s); return; }
    if (nu == 12) { stage_rows_fixed<12>(src, rs, nrows, s); return; }
    if (nu == 3) { stage_rows_fixed<3>(src, rs, nrows, s); return; }
    const int total = nrows * nu;
    for (int k = threadIdx.x; k < total; k += kPreThreads) {
        const int t = k / nu;
        const int f = k - t * nu;
        s[f * kPitch + t] = src[(size_t)t * rs + f];
    }
}

// Inverse of stage_rows for the first nu floats of each row: LDS -> HBM.
template <int NU>
__device__ __forceinline__ void unstage_rows_fixed(const float* s, int rs, int nrows, float* __restrict__ dst)
{
    const int total = nrows * NU;
    for (int k = threadIdx.x; k < total; k += kPreThreads) {
        const int t = k / NU;
        const int f = k - t * NU;
        dst[(size_t)t * rs + f] = s[f * kPitch + t];
    }
}

template <int NU>
__device__ __forceinline__ void unstage_rows_vec4(const float* s, int nrows, float* __restrict__ dst)
{
    float4* dst4 = reinterpret_cast<float4*>(dst);
    const int total4 = nrows * (NU / 4);
    for (int k = threadIdx.x; k < total4; k += kPreThreads) {
        const int t = k / (NU / 4);
        const int f = (k - t * (NU / 4)) * 4;
        dst4[k] = make_float4(s[(f + 0) * kPitch + t], s[(f + 1) * kPitch + t],
                              s[(f + 2) * kPitch + t], s[(f + 3) * kPitch + t]);
    }
}

__device__ __forceinline__ void unstage_rows(const float* s, int rs, int nu, int nrows, float* __restrict__ dst)
{
    if (nu == 48 && rs == 48 && ((uintptr_t)dst & 15) == 0) { unstage_rows_vec4<48>(s, nrows, dst); return; }
    if (nu == 48) { unstage_rows_fixed<48>(s, rs, nrows, dst); return; }
    const int total = nrows * nu;
    for (int k = threadIdx.x; k < total; k += kPreThreads) {
        const int t = k / nu;
        const int f = k - t * nu;
        dst[(size_t)t * rs + f] = s[f * kPitch + t];
    }
}

__device__ __forceinline__ void view_dir(const float* campos, float px, float py, float pz,
                                         float* d, float* dor)
{
    dor[0] = px - campos[0];
    dor[1] = py - campos[1];
    dor[2] = pz - campos[2];
    const float len = sqrtf(dor[0] * dor[0] + dor[1] * dor[1] + dor[2] * dor[2]);
    d[0] = dor[0] / len;
    d[1] = dor[1] / len;
    d[2] = dor[2] / len;
}

// Forward preprocess of one Gaussian in one view.  (px, py, pz) is its mean,
// c3 its 3D covariance (upper triangle) and opacity_in its stored opacity;
// the view, the image and the OGFS settings come from a.
struct PreRes {
    int radius;
    uint32_t tiles;
    uint4 rec;                  // packed tile rectangle, tiles, 0
    float2 mean2d;
    float depth;
    float4 con;
    float col[3];
};

// The view-independent log of the opacity the OGFS radius uses (variant 0).
__device__ __forceinline__ float pre_log_opacity(const PreArgs& a, float opacity_in)
{
    return a.variant == 1 ? 0.f : logf_cr(fmaxf(opacity_in, 1e-6f));
}

// log_opacity: pre_log_opacity(a, opacity_in).
template <typename Sh>
__device__ __forceinline__ void preprocess_view(const PreArgs& a, float px, float py, float pz,
                                                const float* c3, float opacity_in, float log_opacity,
                                                const Sh& sh, int* error_flag, PreRes& r)
{
    int radius_px = 0;
    uint32_t tiles = 0;
    float mx = 0.f, my = 0.f, depth = 0.f;
    float4 con = make_float4(0.f, 0.f, 0.f, 0.f);
    float col[3] = { 0.f, 0.f, 0.f };
    int rc[4] = { 0, 0, 0, 0 };

    float hom[4], view[3];
    xf4(a.proj, px, py, pz, hom);
    const float p_w = 1.0f / (hom[3] + 0.0000001f);
    const float projx = hom[0] * p_w, projy = hom[1] * p_w;
    xf3(a.view, px, py, pz, view);
    bool alive = !(view[2] <= 0.2f);
    if (!alive && a.prefiltered) atomicOr(error_flag, 1);
    float ca = 0.f, cb = 0.f, cc = 0.f, det = 0.f;
    if (alive) {
        Cov2D cv;
        cov2d(a.view, px, py, pz, a.fx, a.fy, a.tanx, a.tany, c3, cv);
        ca = cv.a; cb = cv.b; cc = cv.c;
        if (a.variant == 1) {
            ca = (float)((double)ca + 0.3);
            cc = (float)((double)cc + 0.3);
        } else if (a.use_filter) {
            ca = fmaxf(ca, 0.3f);
            cc = fmaxf(cc, 0.3f);
        }
        det = ca * cc - cb * cb;
        alive = det != 0.0f;
    }
    float opacity = 0.f;
    if (alive) {
        const float det_inv = 1.0f / det;
        con.x = cc * det_inv;
        con.y = -cb * det_inv;
        con.z = ca * det_inv;
        if (a.variant == 1) {
            opacity = opacity_in;
            const float mid = 0.5f * (ca + cc);
            const float l1 = mid + sqrtf(fmaxf(0.1f, mid * mid - det));
            const float l2 = mid - sqrtf(fmaxf(0.1f, mid * mid - det));
            radius_px = sat_int(ceilf(3.0f * sqrtf(fmaxf(l1, l2))));
            if (radius_px <= 0) alive = false;
        } else {
            opacity = fmaxf(opacity_in, 1e-6f);
            const float energy_clamped = fmaxf(log_opacity - a.log_lambda, a.energy_floor);
            const float mid = 0.5f * (ca + cc);
            const float disc = fmaxf(0.0f, mid * mid - det);
            const float sqrt_disc = sqrtf(disc);
            const float l1 = mid + sqrt_disc, l2 = mid - sqrt_disc;
            const float max_lambda = fmaxf(fmaxf(l1, l2), 1e-6f);
            const float radius_base = 3.0f * sqrtf(max_lambda);
            float radius_major = sqrtf(2.0f * energy_clamped * max_lambda);
            radius_major = fminf(radius_major, radius_base);
            radius_major = fmaxf(radius_major, a.ratio_min * radius_base);
            const int rp = sat_int(ceilf(radius_major + 1.0f));
            radius_px = rp > 1 ? rp : 1;
        }
        con.w = opacity;
    }
    if (alive) {
        mx = (float)((((double)projx + 1.0) * (double)a.W - 1.0) * 0.5);
        my = (float)((((double)projy + 1.0) * (double)a.H - 1.0) * 0.5);
        get_rect(mx, my, radius_px, a.gx, a.gy, rc);
        tiles = (uint32_t)((rc[2] - rc[0]) * (rc[3] - rc[1]));
        alive = tiles != 0;
    }
    if (alive) {
        depth = view[2];
        if (!a.colors_precomp) {
            float d[3], dor[3];
            view_dir(a.campos, px, py, pz, d, dor);
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                const float v = sh_eval(a.D, sh, ch, d[0], d[1], d[2]) + 0.5f;
                // clamped: -0, which blends like 0 and tells the backward
                col[ch] = v < 0.0f ? -0.0f : v;
            }
        }
    } else {
        radius_px = 0;
        tiles = 0;
        mx = my = 0.f;
        con = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    r.radius = radius_px;
    r.tiles = tiles;
    r.rec = tiles ? make_uint4((uint32_t)rc[0] | ((uint32_t)rc[1] << 16),
                               (uint32_t)rc[2] | ((uint32_t)rc[3] << 16), tiles, 0u)
                  : make_uint4(0u, 0u, 0u, 0u);
    r.mean2d = make_float2(mx, my);
    r.depth = depth;
    r.con = con;
    r.col[0] = col[0];
    r.col[1] = col[1];
    r.col[2] = col[2];
}

// View-independent inputs of Gaussian i: mean and 3D covariance.
__device__ __forceinline__ void gaussian_inputs(const PreArgs& a, int i, float* p, float* c3)
{
    p[0] = a.means3D[3 * i];
    p[1] = a.means3D[3 * i + 1];
    p[2] = a.means3D[3 * i + 2];
    if (a.cov3D_precomp) {
#pragma unroll
        for (int k = 0; k < 6; ++k) c3[k] = a.cov3D_precomp[6 * i + k];
    } else {
        cov3d(a.mod * a.scales[3 * i], a.mod * a.scales[3 * i + 1], a.mod * a.scales[3 * i + 2],
              a.rotations[4 * i], a.rotations[4 * i + 1], a.rotations[4 * i + 2],
              a.rotations[4 * i + 3], c3);
    }
}

__device__ __forceinline__ void write_pre(const PreOut& o, size_t e, const PreRes& r, bool rgb)
{
    o.radii[e] = r.radius;
    o.tiles[e] = r.tiles;
    o.recs[e] = r.rec;
    o.means2D[e] = r.mean2d;
    o.depths[e] = r.depth;
    o.conic_opacity[e] = r.con;
    if (rgb) {
        o.rgb[3 * e + 0] = r.col[0];
        o.rgb[3 * e + 1] = r.col[1];
        o.rgb[3 * e + 2] = r.col[2];
    }
}

__global__ __launch_bounds__(kPreThreads) void k_preprocess_fwd(PreArgs a, PreOut o)
{
    __shared__ float s_sh[kMaxShFloats * kPitch];
    const int base = blockIdx.x * kPreThreads;
    const int i = base + threadIdx.x;
    const int nrows = min(kPreThreads, a.P - base);
    const int nu = (a.D + 1) * (a.D + 1) * 3;
    if (!a.colors_precomp) {
        stage_rows(a.shs + (size_t)base * a.M * 3, a.M * 3, nu, nrows, s_sh);
        __syncthreads();
    }
    if (i >= a.P) return;
    float p[3], c3[6];
    gaussian_inputs(a, i, p, c3);
    PreRes r;
    preprocess_view(a, p[0], p[1], p[2], c3, a.opacities[i], pre_log_opacity(a, a.opacities[i]),
                    ShLds{ s_sh, (int)threadIdx.x }, o.error_flag, r);
    write_pre(o, (size_t)i, r, !a.colors_precomp);
    const uint32_t hi = r.tiles ? __float_as_uint(r.depth) : 0xFFFFFFFFu;
    o.vkeys[i] = ((uint64_t)hi << 32) | (uint32_t)i;
}

// FilterGS render (FilterGS/filtergs.py:98-223 in one pass): Gaussian i is
// row rows_index[i] of the raw parameters, activated as FilterGS.get_all does
// (scale exp, opacity sigmoid, quaternion normalised, colour SH2RGB(dc) +
// eval_sh_wobase, no clamp), and used as the rasterizer's colours_precomp
// input.  A row with a non-finite value, a zero scale or a zero opacity is
// culled, as FilterGS.render drops it before rasterising (:150-167).
__global__ __launch_bounds__(kPreThreads) void k_preprocess_fwd_rows(PreArgs a, PreOut o)
{
    const int i = blockIdx.x * kPreThreads + threadIdx.x;
    if (i >= a.P) return;
    const int64_t src = a.rows_index[i];
    PreRes r{};
    bool keep = src >= 0 && src < a.rows_n_src;
    if (!keep) atomicOr(o.error_flag, 2);
    if (keep) {
        ActRow ar;
        activate_row(src, a.means3D, a.scales, a.rotations, a.opacities, a.rows_dc, a.rows_rest,
                     a.rows_rest_floats, a.rows_deg, a.rows_campos, ar);
        bool fin = isfinite(ar.o) && ar.o > 0.f;
#pragma unroll
        for (int k = 0; k < 3; ++k)
            fin = fin && isfinite(ar.p[k]) && isfinite(ar.s[k]) && ar.s[k] > 0.f && isfinite(ar.c[k]);
#pragma unroll
        for (int k = 0; k < 4; ++k) fin = fin && isfinite(ar.q[k]);
        if (fin) {
            float c3[6];
            cov3d(a.mod * ar.s[0], a.mod * ar.s[1], a.mod * ar.s[2], ar.q[0], ar.q[1], ar.q[2], ar.q[3], c3);
            PreArgs ac = a;
            ac.colors_precomp = a.rows_dc;   // colours come from the activation: no SH pass
            preprocess_view(ac, ar.p[0], ar.p[1], ar.p[2], c3, ar.o, pre_log_opacity(ac, ar.o), ShLds{ nullptr, 0 },
                            o.error_flag, r);
            r.col[0] = ar.c[0];
            r.col[1] = ar.c[1];
            r.col[2] = ar.c[2];
        }
    }
    write_pre(o, (size_t)i, r, true);
    const uint32_t hi = r.tiles ? __float_as_uint(r.depth) : 0xFFFFFFFFu;
    o.vkeys[i] = ((uint64_t)hi << 32) | (uint32_t)i;
}

hipError_t launch_preprocess_fwd_rows(const PreArgs& a, const PreOut& o, hipStream_t s)
{
    if (a.P == 0) return hipSuccess;
    hipLaunchKernelGGL(k_preprocess_fwd_rows, dim3((a.P + kPreThreads - 1) / kPreThreads), dim3(kPreThreads),
                       0, s, a, o);
    return hipGetLastError();
}

// All views of a batch in one pass: the Gaussian's inputs (mean, covariance,
// opacity, SH row) are read once and reused for every view.  radii[v * P + i]
// is written for every (view, Gaussian); the records of a visible one (mean,
// conic and opacity, colour, tile rectangle, Gaussian index) go to its
// element (batch_elem / batch_slot: packed per 64-group and view, so each
// wave stores one run per view), with the group's visibility ballot in
// vis_masks.  The depth key is (v << 58) | (depth_bits << 26) | j with j the
// element's offset in its view, so one sort on bits 26..63 orders every view
// by depth (ties by Gaussian index) and keeps the views apart.  colors_precomp
// of a batch is per view ([V, P, 3]) and is copied to the elements.
__global__ __launch_bounds__(kPreThreads) void k_preprocess_fwd_batch(PreArgs a, const BatchView* views,
                                                                      int V, PreOut o)
{
    const int i = blockIdx.x * kPreThreads + threadIdx.x;
    if (i >= a.P) return;
    const int nu = (a.D + 1) * (a.D + 1) * 3;
    // The Gaussian's SH row in registers, read once for all views (no LDS:
    // the block's rows would cap the kernel at three waves per SIMD).
    float shr[kMaxShFloats];
    if (!a.colors_precomp) {
        const int rs = a.M * 3;
        const float* row = a.shs + (size_t)i * rs;
        if (nu == 48 && rs == 48 && ((uintptr_t)a.shs & 15) == 0) {
#pragma unroll
            for (int q = 0; q < 12; ++q) {
                const float4 v4 = reinterpret_cast<const float4*>(row)[q];
                shr[4 * q] = v4.x;
                shr[4 * q + 1] = v4.y;
                shr[4 * q + 2] = v4.z;
                shr[4 * q + 3] = v4.w;
            }
        } else {
#pragma unroll
            for (int f = 0; f < kMaxShFloats; ++f) shr[f] = f < nu ? row[f] : 0.f;
        }
    }
    float p[3], c3[6];
    gaussian_inputs(a, i, p, c3);
    const float opacity_in = a.opacities[i];
    const float log_opacity = pre_log_opacity(a, opacity_in);   // once for all views
    const ShReg sh{ shr };
    const uint32_t P64 = batch_p64(a.P);
    // Only visible (view, Gaussian) keys go to the sort: each wave appends
    // its own to a segment of 64 * V slots (ballot ranks, no atomics) and
    // records how many; the key order inside a segment does not matter, the
    // keys are unique.
    const uint32_t w64 = (uint32_t)i & ~63u;   // the wave's 64-group (i / 64) * 64
    uint64_t* seg = o.vkeys + (size_t)(i / FGS_WAVE) * FGS_WAVE * V;
    const int lane = lane_id();
    const uint64_t below = lane ? (~0ull >> (FGS_WAVE - lane)) : 0ull;
    uint32_t appended = 0;
    for (int v = 0; v < V; ++v) {
        PreArgs av = a;
        bind_view(av, views[v]);
        PreRes r;
        preprocess_view(av, p[0], p[1], p[2], c3, opacity_in, log_opacity, sh, o.error_flag, r);
        o.radii[(size_t)v * a.P + i] = r.radius;
        const bool vis = r.tiles != 0;
        const uint64_t m = __ballot(vis);
        const uint32_t rank = (uint32_t)__popcll(m & below);
        // the group's tiles before this lane (the lanes past P have exited:
        // they are all above the live ones, which read only lanes below)
        uint32_t tincl = vis ? r.tiles : 0u;
        if (m != 0u) tincl = wave_incl_sum_u32(tincl);
        if (vis) {
            // no depth: the depth travels in the key; the record holds the
            // tile count and the element's slot offset in its group's run
            const size_t e = (size_t)v * P64 + w64 + rank;
            uint4 rec = r.rec;
            rec.w = tincl - r.tiles;
            o.recs[e] = rec;
            if (r.tiles > kGatherSolo) o.big_list[atomicAdd(o.big_count, 1ull)] = (uint32_t)e;
            o.means2D[e] = r.mean2d;
            o.conic_opacity[e] = r.con;
            o.gid[e] = (uint32_t)i;
            float c0 = r.col[0], c1 = r.col[1], c2 = r.col[2];
            if (a.colors_precomp) {
                const size_t q = 3 * ((size_t)v * a.P + i);
                c0 = a.colors_precomp[q];
                c1 = a.colors_precomp[q + 1];
                c2 = a.colors_precomp[q + 2];
            }
            o.rgb[3 * e + 0] = c0;
            o.rgb[3 * e + 1] = c1;
            o.rgb[3 * e + 2] = c2;
            seg[appended + rank] =
                ((uint64_t)v << 58) | ((uint64_t)__float_as_uint(r.depth) << 26) | (w64 + rank);
        }
        const uint32_t gtot = __builtin_amdgcn_readlane(tincl, 63 - __clzll(__ballot(true)));
        if (lane == 0) {
            o.vis_masks[(size_t)v * (P64 >> 6) + (i >> 6)] = m;
            o.grp_tiles[(size_t)v * (P64 >> 6) + (i >> 6)] = gtot;
        }
        appended += (uint32_t)__popcll(m);
    }
    if (lane == 0) o.seg_cnt[i / FGS_WAVE] = appended;
}

// ---------------------------------------------------------------------------
// Backward: backward.cu:144-396 fused (cov2D chain, projection chain, SH, cov3D).
// Every output entry of every Gaussian is written (zeros where radius == 0).
// ---------------------------------------------------------------------------

// Backward of one Gaussian in one view, accumulated into dm (dL/dmean3D),
// dcov (dL/dcov3D) and the Gaussian's dL/dsh column s_out[f * kPitch + t].
// (mx, my, mz) is the mean, c3 the 3D covariance, shr the SH row (first nu
// floats); the per-view gradients of the blend are read at element e.
__device__ __forceinline__ void bwd_view_geom(const PreArgs& a, const float* dconic, const float* dmean2D,
                                              float mx, float my, float mz, const float* c3,
                                              float* dmv, float* dcov)
{
    const bool exact = a.grad_mode == 1;
    Cov2D cv;
    cov2d(a.view, mx, my, mz, a.fx, a.fy, a.tanx, a.tany, c3, cv);
    const float x_grad_mul = (cv.txtz < -cv.limx || cv.txtz > cv.limx) ? 0.f : 1.f;
    const float y_grad_mul = (cv.tytz < -cv.limy || cv.tytz > cv.limy) ? 0.f : 1.f;
    float ca = cv.a, cb = cv.b, cc = cv.c;
    float mask_a = 1.f, mask_c = 1.f;
    if (!exact || a.variant == 1) {
        ca = (float)((double)ca + 0.3);   // backward.cu:197-199
        cc = (float)((double)cc + 0.3);
    } else if (a.use_filter) {
        mask_a = ca > 0.3f ? 1.f : 0.f;
        mask_c = cc > 0.3f ? 1.f : 0.f;
        ca = fmaxf(ca, 0.3f);
        cc = fmaxf(cc, 0.3f);
    }
    const float dcx = dconic[0], dcy = dconic[1], dcz = dconic[2];
    const float denom = ca * cc - cb * cb;
    const float denom2inv = exact ? 1.0f / (denom * denom) : 1.0f / ((denom * denom) + 0.0000001f);
    float dL_da = 0.f, dL_db = 0.f, dL_dc = 0.f;
    const float (*T)[3] = cv.T;
    if (denom2inv != 0.f) {
        dL_da = denom2inv * (-cc * cc * dcx + 2.f * cb * cc * dcy + (denom - ca * cc) * dcz);
        dL_dc = denom2inv * (-ca * ca * dcz + 2.f * ca * cb * dcy + (denom - ca * cc) * dcx);
        dL_db = denom2inv * 2.f * (cb * cc * dcx - (denom + 2.f * cb * cb) * dcy + ca * cb * dcz);
        dL_da *= mask_a;
        dL_dc *= mask_c;
        dcov[0] += T[0][0] * T[0][0] * dL_da + T[0][0] * T[1][0] * dL_db + T[1][0] * T[1][0] * dL_dc;
        dcov[3] += T[0][1] * T[0][1] * dL_da + T[0][1] * T[1][1] * dL_db + T[1][1] * T[1][1] * dL_dc;
        dcov[5] += T[0][2] * T[0][2] * dL_da + T[0][2] * T[1][2] * dL_db + T[1][2] * T[1][2] * dL_dc;
        dcov[1] += 2.f * T[0][0] * T[0][1] * dL_da + (T[0][0] * T[1][1] + T[0][1] * T[1][0]) * dL_db + 2.f * T[1][0] * T[1][1] * dL_dc;
        dcov[2] += 2.f * T[0][0] * T[0][2] * dL_da + (T[0][0] * T[1][2] + T[0][2] * T[1][0]) * dL_db + 2.f * T[1][0] * T[1][2] * dL_dc;
        dcov[4] += 2.f * T[0][2] * T[0][1] * dL_da + (T[0][1] * T[1][2] + T[0][2] * T[1][1]) * dL_db + 2.f * T[1][1] * T[1][2] * dL_dc;
    }
    const float Vk[3][3] = { { c3[0], c3[1], c3[2] }, { c3[1], c3[3], c3[4] }, { c3[2], c3[4], c3[5] } };
    float dT[2][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float t0 = T[0][0] * Vk[k][0] + T[0][1] * Vk[k][1] + T[0][2] * Vk[k][2];
        const float t1 = T[1][0] * Vk[k][0] + T[1][1] * Vk[k][1] + T[1][2] * Vk[k][2];
        dT[0][k] = 2.f * t0 * dL_da + t1 * dL_db;
        dT[1][k] = 2.f * t1 * dL_dc + t0 * dL_db;
    }
    const float (*W)[3] = cv.W;
    const float dL_dJ00 = W[0][0] * dT[0][0] + W[0][1] * dT[0][1] + W[0][2] * dT[0][2];
    const float dL_dJ02 = W[2][0] * dT[0][0] + W[2][1] * dT[0][1] + W[2][2] * dT[0][2];
    const float dL_dJ11 = W[1][0] * dT[1][0] + W[1][1] * dT[1][1] + W[1][2] * dT[1][2];
    const float dL_dJ12 = W[2][0] * dT[1][0] + W[2][1] * dT[1][1] + W[2][2] * dT[1][2];
    const float tz = 1.f / cv.t[2];
    const float tz2 = tz * tz, tz3 = tz2 * tz;
    const float dL_dtx = x_grad_mul * -a.fx * tz2 * dL_dJ02;
    const float dL_dty = y_grad_mul * -a.fy * tz2 * dL_dJ12;
    const float dL_dtz = -a.fx * tz2 * dL_dJ00 - a.fy * tz2 * dL_dJ11 + (2.f * a.fx * cv.t[0]) * tz3 * dL_dJ02 +
                         (2.f * a.fy * cv.t[1]) * tz3 * dL_dJ12;
    const float* V = a.view;
    dmv[0] = V[0] * dL_dtx + V[1] * dL_dty + V[2] * dL_dtz;
    dmv[1] = V[4] * dL_dtx + V[5] * dL_dty + V[6] * dL_dtz;
    dmv[2] = V[8] * dL_dtx + V[9] * dL_dty + V[10] * dL_dtz;

    // projection of the mean (backward.cu:370-387)
    const float* Pm = a.proj;
    float hom[4];
    xf4(Pm, mx, my, mz, hom);
    const float m_w = 1.0f / (hom[3] + 0.0000001f);
    const float mul1 = hom[0] * m_w * m_w;
    const float mul2 = hom[1] * m_w * m_w;
    const float g2x = dmean2D[0], g2y = dmean2D[1];
    dmv[0] += (Pm[0] * m_w - Pm[3] * mul1) * g2x + (Pm[1] * m_w - Pm[3] * mul2) * g2y;
    dmv[1] += (Pm[4] * m_w - Pm[7] * mul1) * g2x + (Pm[5] * m_w - Pm[7] * mul2) * g2y;
    dmv[2] += (Pm[8] * m_w - Pm[11] * mul1) * g2x + (Pm[9] * m_w - Pm[11] * mul2) * g2y;
}

// SH backward (backward.cu:20-139) of one Gaussian in one view: dL/dsh into
// the LDS column s_out[f * kPitch + t], the view direction's term into dmv.
// rgb: the forward's colour of the Gaussian in the view (a clamped channel
// is -0 and passes no gradient).
__device__ __forceinline__ void bwd_view_sh(const PreArgs& a, const float* campos, const float* dcolors,
                                            const float* rgb, float mx, float my, float mz, const float* shr,
                                            float* dmv, float* s_out, int t)
{
    {
        const int deg = a.D;
        float d[3], dor[3];
        view_dir(campos, mx, my, mz, d, dor);
        const float x = d[0], y = d[1], z = d[2];
        float dRGB[3];
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) dRGB[ch] = signbit(rgb[ch]) ? 0.f : dcolors[ch];
        float ddx[3] = { 0.f, 0.f, 0.f }, ddy[3] = { 0.f, 0.f, 0.f }, ddz[3] = { 0.f, 0.f, 0.f };
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
#define SHv(k) shr[(k) * 3 + ch]
#define DSH(k) s_out[((k) * 3 + ch) * kPitch + t]
            const float g = dRGB[ch];
            DSH(0) += kSH_C0 * g;
            if (deg > 0) {
                const float C1 = kSH_C1;
                DSH(1) += -C1 * y * g;
                DSH(2) += C1 * z * g;
                DSH(3) += -C1 * x * g;
                ddx[ch] = -C1 * SHv(3);
                ddy[ch] = -C1 * SHv(1);
                ddz[ch] = C1 * SHv(2);
                if (deg > 1) {
                    const float xx = x * x, yy = y * y, zz = z * z;
                    const float xy = x * y, yz = y * z, xz = x * z;
                    DSH(4) += kSH_C2[0] * xy * g;
                    DSH(5) += kSH_C2[1] * yz * g;
                    DSH(6) += kSH_C2[2] * (2.f * zz - xx - yy) * g;
                    DSH(7) += kSH_C2[3] * xz * g;
                    DSH(8) += kSH_C2[4] * (xx - yy) * g;
                    ddx[ch] += kSH_C2[0] * y * SHv(4) + kSH_C2[2] * 2.f * -x * SHv(6) + kSH_C2[3] * z * SHv(7) + kSH_C2[4] * 2.f * x * SHv(8);
                    ddy[ch] += kSH_C2[0] * x * SHv(4) + kSH_C2[1] * z * SHv(5) + kSH_C2[2] * 2.f * -y * SHv(6) + kSH_C2[4] * 2.f * -y * SHv(8);
                    ddz[ch] += kSH_C2[1] * y * SHv(5) + kSH_C2[2] * 2.f * 2.f * z * SHv(6) + kSH_C2[3] * x * SHv(7);
                    if (deg > 2) {
                        DSH(9) += kSH_C3[0] * y * (3.f * xx - yy) * g;
                        DSH(10) += kSH_C3[1] * xy * z * g;
                        DSH(11) += kSH_C3[2] * y * (4.f * zz - xx - yy) * g;
                        DSH(12) += kSH_C3[3] * z * (2.f * zz - 3.f * xx - 3.f * yy) * g;
                        DSH(13) += kSH_C3[4] * x * (4.f * zz - xx - yy) * g;
                        DSH(14) += kSH_C3[5] * z * (xx - yy) * g;
                        DSH(15) += kSH_C3[6] * x * (xx - 3.f * yy) * g;
                        ddx[ch] += (kSH_C3[0] * SHv(9) * 3.f * 2.f * xy + kSH_C3[1] * SHv(10) * yz +
                                    kSH_C3[2] * SHv(11) * -2.f * xy + kSH_C3[3] * SHv(12) * -3.f * 2.f * xz +
                                    kSH_C3[4] * SHv(13) * (-3.f * xx + 4.f * zz - yy) +
                                    kSH_C3[5] * SHv(14) * 2.f * xz + kSH_C3[6] * SHv(15) * 3.f * (xx - yy));
                        ddy[ch] += (kSH_C3[0] * SHv(9) * 3.f * (xx - yy) + kSH_C3[1] * SHv(10) * xz +
                                    kSH_C3[2] * SHv(11) * (-3.f * yy + 4.f * zz - xx) +
                                    kSH_C3[3] * SHv(12) * -3.f * 2.f * yz + kSH_C3[4] * SHv(13) * -2.f * xy +
                                    kSH_C3[5] * SHv(14) * -2.f * yz + kSH_C3[6] * SHv(15) * -3.f * 2.f * xy);
                        ddz[ch] += (kSH_C3[1] * SHv(10) * xy + kSH_C3[2] * SHv(11) * 4.f * 2.f * yz +
                                    kSH_C3[3] * SHv(12) * 3.f * (2.f * zz - xx - yy) +
                                    kSH_C3[4] * SHv(13) * 4.f * 2.f * xz + kSH_C3[5] * SHv(14) * (xx - yy));
                    }
                }
            }
#undef SHv
#undef DSH
        }
        const float ddir0 = ddx[0] * dRGB[0] + ddx[1] * dRGB[1] + ddx[2] * dRGB[2];
        const float ddir1 = ddy[0] * dRGB[0] + ddy[1] * dRGB[1] + ddy[2] * dRGB[2];
        const float ddir2 = ddz[0] * dRGB[0] + ddz[1] * dRGB[1] + ddz[2] * dRGB[2];
        const float sum2 = dor[0] * dor[0] + dor[1] * dor[1] + dor[2] * dor[2];
        const float invsum32 = 1.0f / sqrtf(sum2 * sum2 * sum2);
        dmv[0] += ((sum2 - dor[0] * dor[0]) * ddir0 - dor[1] * dor[0] * ddir1 - dor[2] * dor[0] * ddir2) * invsum32;
        dmv[1] += (-dor[0] * dor[1] * ddir0 + (sum2 - dor[1] * dor[1]) * ddir1 - dor[2] * dor[1] * ddir2) * invsum32;
        dmv[2] += (-dor[0] * dor[2] * ddir0 - dor[1] * dor[2] * ddir1 + (sum2 - dor[2] * dor[2]) * ddir2) * invsum32;
    }
}

__device__ __forceinline__ void preprocess_bwd_view(const PreArgs& a, const PreBwdIn& in, size_t e,
                                                    float mx, float my, float mz, const float* c3,
                                                    const float* shr, float* dm, float* dcov,
                                                    float* s_out, int t)
{
    float dmv[3];
    bwd_view_geom(a, in.dL_dconic + 3 * e, in.dL_dmean2D + 3 * e, mx, my, mz, c3, dmv, dcov);
    if (a.shs) bwd_view_sh(a, a.campos, in.dL_dcolors + 3 * e, in.rgb + 3 * e, mx, my, mz, shr, dmv, s_out, t);
#pragma unroll
    for (int k = 0; k < 3; ++k) dm[k] += dmv[k];
}

// Writes dL/dmean3D and dL/dcov3D of Gaussian i and backpropagates dcov to
// scale and rotation (backward.cu:278-341: dL/dscale is not multiplied by the
// modifier in reference mode; the quaternion is used unnormalised).
// A Gaussian rendered in no view gets zeros (alive = false).
__device__ __forceinline__ void preprocess_bwd_finish(const PreArgs& a, const PreBwdOut& o, int i,
                                                      bool alive, const float* dm, const float* dcov)
{
    const bool exact = a.grad_mode == 1;
#pragma unroll
    for (int k = 0; k < 3; ++k) o.dL_dmeans3D[3 * i + k] = alive ? dm[k] : 0.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) o.dL_dcov3D[6 * i + k] = alive ? dcov[k] : 0.f;
    if (alive && !a.cov3D_precomp && o.dL_dscales) {
        float sv[3];
        sv[0] = a.mod * a.scales[3 * i];
        sv[1] = a.mod * a.scales[3 * i + 1];
        sv[2] = a.mod * a.scales[3 * i + 2];
        const float qr = a.rotations[4 * i], qx = a.rotations[4 * i + 1];
        const float qy = a.rotations[4 * i + 2], qz = a.rotations[4 * i + 3];
        float R[3][3];
        quat_rot(qr, qx, qy, qz, R);
        float Mm[3][3];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int q = 0; q < 3; ++q) Mm[j][q] = sv[q] * R[j][q];
        const float dS[3][3] = { { dcov[0], 0.5f * dcov[1], 0.5f * dcov[2] },
                                 { 0.5f * dcov[1], dcov[3], 0.5f * dcov[4] },
                                 { 0.5f * dcov[2], 0.5f * dcov[4], dcov[5] } };
        float dM[3][3];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int q = 0; q < 3; ++q)
                dM[j][q] = (2.f * Mm[0][q]) * dS[j][0] + (2.f * Mm[1][q]) * dS[j][1] + (2.f * Mm[2][q]) * dS[j][2];
        const float smul = exact ? a.mod : 1.f;
#pragma unroll
        for (int q = 0; q < 3; ++q)
            o.dL_dscales[3 * i + q] = (R[0][q] * dM[0][q] + R[1][q] * dM[1][q] + R[2][q] * dM[2][q]) * smul;
        float dMt[3][3];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int q = 0; q < 3; ++q) dMt[j][q] = dM[q][j] * sv[j];
        const float r = qr, x = qx, y = qy, z = qz;
        float4 dq;
        dq.x = 2.f * z * (dMt[0][1] - dMt[1][0]) + 2.f * y * (dMt[2][0] - dMt[0][2]) + 2.f * x * (dMt[1][2] - dMt[2][1]);
        dq.y = 2.f * y * (dMt[1][0] + dMt[0][1]) + 2.f * z * (dMt[2][0] + dMt[0][2]) + 2.f * r * (dMt[1][2] - dMt[2][1]) - 4.f * x * (dMt[2][2] + dMt[1][1]);
        dq.z = 2.f * x * (dMt[1][0] + dMt[0][1]) + 2.f * r * (dMt[2][0] - dMt[0][2]) + 2.f * z * (dMt[1][2] + dMt[2][1]) - 4.f * y * (dMt[2][2] + dMt[0][0]);
        dq.w = 2.f * r * (dMt[0][1] - dMt[1][0]) + 2.f * x * (dMt[2][0] + dMt[0][2]) + 2.f * y * (dMt[1][2] + dMt[2][1]) - 4.f * z * (dMt[1][1] + dMt[0][0]);
        o.dL_drot[4 * i + 0] = dq.x;
        o.dL_drot[4 * i + 1] = dq.y;
        o.dL_drot[4 * i + 2] = dq.z;
        o.dL_drot[4 * i + 3] = dq.w;
    } else {
        if (o.dL_dscales)
            for (int k = 0; k < 3; ++k) o.dL_dscales[3 * i + k] = 0.f;
        if (o.dL_drot)
            for (int k = 0; k < 4; ++k) o.dL_drot[4 * i + k] = 0.f;
    }
}

// SH rows of the block into registers (shr) and the LDS buffer cleared for
// the dL/dsh accumulation.
__device__ __forceinline__ void bwd_stage_sh(const PreArgs& a, int base, int nrows, int nu, float* shr,
                                             float* s_sh, int t)
{
    const int rs = a.M * 3;
#pragma unroll
    for (int f = 0; f < kMaxShFloats; ++f) shr[f] = 0.f;
    stage_rows(a.shs + (size_t)base * rs, rs, nu, nrows, s_sh);
    __syncthreads();
#pragma unroll
    for (int f = 0; f < kMaxShFloats; ++f)
        if (f < nu) shr[f] = s_sh[f * kPitch + t];
    __syncthreads();
#pragma unroll
    for (int f = 0; f < kMaxShFloats; ++f)
        if (f < nu) s_sh[f * kPitch + t] = 0.f;   // the buffer now collects dL/dsh
}

__device__ __forceinline__ void bwd_store_sh(const PreArgs& a, const PreBwdOut& o, int base, int nrows,
                                             int nu, const float* s_sh, int t)
{
    const int rs = a.M * 3;
    __syncthreads();
    float* dst = o.dL_dsh + (size_t)base * rs;
    unstage_rows(s_sh, rs, nu, nrows, dst);
    if (nu < rs) {
        // coefficients above the active degree get zero gradient
        const int tail = rs - nu;
        for (int k = t; k < nrows * tail; k += kPreThreads) {
            const int r = k / tail;
            dst[(size_t)r * rs + nu + (k - r * tail)] = 0.f;
        }
    }
}

__global__ __launch_bounds__(kPreThreads) void k_preprocess_bwd(PreArgs a, PreBwdIn in, PreBwdOut o)
{
    __shared__ float s_sh[kMaxShFloats * kPitch];
    const int t = threadIdx.x;
    const int base = blockIdx.x * kPreThreads;
    const int i = base + t;
    const int nrows = min(kPreThreads, a.P - base);
    const bool has_sh = a.shs != nullptr;
    const int nu = (a.D + 1) * (a.D + 1) * 3;
    float shr[kMaxShFloats];
    if (has_sh) bwd_stage_sh(a, base, nrows, nu, shr, s_sh, t);
    if (i < a.P) {
        float dm[3] = { 0.f, 0.f, 0.f }, dcov[6] = { 0.f, 0.f, 0.f, 0.f, 0.f, 0.f };
        const bool alive = in.radii[i] > 0;
        if (alive) {
            float m[3], c3[6];
            gaussian_inputs(a, i, m, c3);
            preprocess_bwd_view(a, in, (size_t)i, m[0], m[1], m[2], c3, shr, dm, dcov, s_sh, t);
        }
        preprocess_bwd_finish(a, o, i, alive, dm, dcov);
        if (alive && a.grad_mode == 1 && a.variant == 0 && a.opacities[i] < 1e-6f) o.dL_dopacity[i] = 0.f;
    }
    if (has_sh) bwd_store_sh(a, o, base, nrows, nu, s_sh, t);
}

// Bit v of the result: Gaussian i is visible in view v (radius > 0).
// The views that see Gaussian i, from the batch's visibility ballots (one
// wave-uniform load per view: a wave's Gaussians share their 64-group).
__device__ __forceinline__ uint64_t batch_visible_views(const uint64_t* __restrict__ vis_masks, uint32_t P64,
                                                        int V, int i)
{
    const uint32_t nw = P64 >> 6;
    const uint32_t grp = __builtin_amdgcn_readfirstlane((uint32_t)i >> 6);
    const uint32_t l = (uint32_t)i & 63u;
    uint64_t vis = 0ull;
    for (int v = 0; v < V; ++v) vis |= ((vis_masks[(size_t)v * nw + grp] >> l) & 1ull) << v;
    return vis;
}

// All views of a batch, in two passes over the Gaussians so that neither
// holds both the SH row and the covariance chain in registers.  The SH pass
// sums dL/dsh over the visible views and leaves the view-direction term of
// dL/dmean3D in o.dL_dmeans3D; the geometry pass adds the cov2D and
// projection terms to it, sums dL/dcov3D over the views before the scale /
// rotation backward (linear in it) and sums the per-view dL/dopacity of the
// blend into o.dL_dopacity.  Per-view gradients are read at the Gaussian's
// element of the view (batch_slot).
// Both passes keep the views' constants in LDS and load the next visible
// view's blend gradients before working the current one.  The SH pass is held
// to 3 waves per SIMD (its LDS column of dL/dsh allows no more), the geometry
// pass to 5 (89 VGPRs, no spills).
__global__ __launch_bounds__(kPreThreads) __attribute__((amdgpu_waves_per_eu(3)))
void k_preprocess_bwd_batch_sh(PreArgs a, const BatchView* views, int V, PreBwdIn in, PreBwdOut o)
{
    __shared__ float s_sh[kMaxShFloats * kPitch];
    __shared__ float s_cam[3 * kBatchMaxViews];
    const int t = threadIdx.x;
    const int base = blockIdx.x * kPreThreads;
    const int i = base + t;
    const int nrows = min(kPreThreads, a.P - base);
    const int nu = (a.D + 1) * (a.D + 1) * 3;
    for (int k = t; k < 3 * V; k += kPreThreads) s_cam[k] = views[k / 3].campos[k % 3];
    float shr[kMaxShFloats];
    bwd_stage_sh(a, base, nrows, nu, shr, s_sh, t);   // its barrier also publishes s_cam
    if (i < a.P) {
        const float mx = a.means3D[3 * i], my = a.means3D[3 * i + 1], mz = a.means3D[3 * i + 2];
        float dm[3] = { 0.f, 0.f, 0.f };
        const uint32_t P64 = batch_p64(a.P);
        uint64_t vis = batch_visible_views(in.vis_masks, P64, V, i);
        int v = vis ? __builtin_ctzll(vis) : 0;
        float gc[3] = { 0.f, 0.f, 0.f }, rc[3] = { 0.f, 0.f, 0.f };
        if (vis) {
            const size_t e = batch_slot(in.vis_masks, P64, v, i);
            const float* src = in.dL_dcolors + 3 * e;
            gc[0] = src[0]; gc[1] = src[1]; gc[2] = src[2];
            rc[0] = in.rgb[3 * e]; rc[1] = in.rgb[3 * e + 1]; rc[2] = in.rgb[3 * e + 2];
        }
        while (vis) {
            // Opaque to the optimiser: products of the SH row with constants
            // are not hoisted out of the loop as a second set of 48 registers
            // (168 VGPRs with spills in the loop, 136 without).
#pragma unroll
            for (int f = 0; f < kMaxShFloats; ++f) asm volatile("" : "+v"(shr[f]));
            vis &= vis - 1ull;
            const int vn = vis ? __builtin_ctzll(vis) : v;
            const size_t en = batch_slot(in.vis_masks, P64, vn, i);
            const float* nsrc = in.dL_dcolors + 3 * en;
            const float gn[3] = { nsrc[0], nsrc[1], nsrc[2] };
            const float rn[3] = { in.rgb[3 * en], in.rgb[3 * en + 1], in.rgb[3 * en + 2] };
            bwd_view_sh(a, s_cam + 3 * v, gc, rc, mx, my, mz, shr, dm, s_sh, t);
            v = vn;
            gc[0] = gn[0]; gc[1] = gn[1]; gc[2] = gn[2];
            rc[0] = rn[0]; rc[1] = rn[1]; rc[2] = rn[2];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) o.dL_dmeans3D[3 * i + k] = dm[k];
    }
    bwd_store_sh(a, o, base, nrows, nu, s_sh, t);
}

__global__ __launch_bounds__(kPreThreads) __attribute__((amdgpu_waves_per_eu(5)))
void k_preprocess_bwd_batch_geom(PreArgs a, const BatchView* views, int V, PreBwdIn in,
                                 const float* dL_dopacity_view, PreBwdOut o)
{
    __shared__ float s_mat[kBatchMaxViews][32];   // view, then projection matrix
    __shared__ float4 s_cam[kBatchMaxViews];      // tanx, tany, fx, fy
    for (int k = threadIdx.x; k < 32 * V; k += kPreThreads)
        s_mat[k >> 5][k & 31] = (k & 31) < 16 ? views[k >> 5].view[k & 15] : views[k >> 5].proj[k & 15];
    for (int k = threadIdx.x; k < V; k += kPreThreads)
        s_cam[k] = make_float4(views[k].tanx, views[k].tany, views[k].fx, views[k].fy);
    __syncthreads();
    const int i = blockIdx.x * kPreThreads + threadIdx.x;
    if (i >= a.P) return;
    float dm[3] = { 0.f, 0.f, 0.f }, dcov[6] = { 0.f, 0.f, 0.f, 0.f, 0.f, 0.f };
    if (a.shs) {
#pragma unroll
        for (int k = 0; k < 3; ++k) dm[k] = o.dL_dmeans3D[3 * i + k];
    }
    float m[3], c3[6];
    gaussian_inputs(a, i, m, c3);
    float dop = 0.f;
    const uint32_t P64 = batch_p64(a.P);
    uint64_t vis = batch_visible_views(in.vis_masks, P64, V, i);
    const bool any = vis != 0ull;
    int v = vis ? __builtin_ctzll(vis) : 0;
    float gcur[6] = { 0.f, 0.f, 0.f, 0.f, 0.f, 0.f };   // dconic[3], dmean2D[2], dopacity
    auto load6 = [&](int w, float* g) {
        const size_t e = batch_slot(in.vis_masks, P64, w, i);
        g[0] = in.dL_dconic[3 * e]; g[1] = in.dL_dconic[3 * e + 1]; g[2] = in.dL_dconic[3 * e + 2];
        g[3] = in.dL_dmean2D[3 * e]; g[4] = in.dL_dmean2D[3 * e + 1];
        g[5] = dL_dopacity_view[e];
    };
    if (vis) load6(v, gcur);
    while (vis) {
        vis &= vis - 1ull;
        const int vn = vis ? __builtin_ctzll(vis) : v;
        float gn[6];
        load6(vn, gn);
        PreArgs av = a;
        av.view = s_mat[v];
        av.proj = s_mat[v] + 16;
        const float4 cam = s_cam[v];
        av.tanx = cam.x; av.tany = cam.y; av.fx = cam.z; av.fy = cam.w;
        float dmv[3];
        bwd_view_geom(av, gcur, gcur + 3, m[0], m[1], m[2], c3, dmv, dcov);
#pragma unroll
        for (int k = 0; k < 3; ++k) dm[k] += dmv[k];
        dop += gcur[5];
        v = vn;
#pragma unroll
        for (int k = 0; k < 6; ++k) gcur[k] = gn[k];
    }
    preprocess_bwd_finish(a, o, i, any, dm, dcov);
    if (any && a.grad_mode == 1 && a.variant == 0 && a.opacities[i] < 1e-6f) dop = 0.f;
    o.dL_dopacity[i] = dop;
}

// Zeroes the per-element blend gradients (three [.,3] arrays and one [.])
// of every visible element of a batch, before the blend adds into them: one
// wave per (view, 64-group), its visible elements being the run at the front
// of the group's 64 slots.  Culled slots are never read and stay as they are.
__global__ __launch_bounds__(256) void k_clear_elements(const uint64_t* __restrict__ vis_masks, size_t groups,
                                                        float* __restrict__ a3, float* __restrict__ b3,
                                                        float* __restrict__ c3, float* __restrict__ d1)
{
    const size_t grp = (size_t)blockIdx.x * (256 / FGS_WAVE) + threadIdx.x / FGS_WAVE;
    if (grp >= groups) return;
    const uint32_t lane = (uint32_t)lane_id();
    if (lane >= (uint32_t)__popcll(vis_masks[grp])) return;
    const size_t e = grp * FGS_WAVE + lane;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        a3[3 * e + k] = 0.f;
        b3[3 * e + k] = 0.f;
        c3[3 * e + k] = 0.f;
    }
    d1[e] = 0.f;
}

hipError_t launch_clear_elements(const uint64_t* vis_masks, size_t groups, float* a3, float* b3, float* c3,
                                 float* d1, hipStream_t s)
{
    if (groups == 0) return hipSuccess;
    const size_t per = 256 / FGS_WAVE;
    hipLaunchKernelGGL(k_clear_elements, dim3((unsigned)((groups + per - 1) / per)), dim3(256), 0, s, vis_masks,
                       groups, a3, b3, c3, d1);
    return hipGetLastError();
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

hipError_t launch_fill_u32(void* p, uint32_t v, size_t n_words, hipStream_t s)
{
    if (n_words == 0) return hipSuccess;
    size_t blocks = (n_words + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t*)p, v, n_words);
    return hipGetLastError();
}

hipError_t launch_preprocess_fwd(const PreArgs& a, const PreOut& o, hipStream_t s)
{
    if (a.P == 0) return hipSuccess;
    hipLaunchKernelGGL(k_preprocess_fwd, dim3((a.P + kPreThreads - 1) / kPreThreads), dim3(kPreThreads), 0, s, a, o);
    return hipGetLastError();
}

hipError_t launch_preprocess_fwd_batch(const PreArgs& a, const BatchView* views, int V,
                                       const PreOut& o, hipStream_t s)
{
    if (a.P == 0 || V == 0) return hipSuccess;
    hipLaunchKernelGGL(k_preprocess_fwd_batch, dim3((a.P + kPreThreads - 1) / kPreThreads),
                       dim3(kPreThreads), 0, s, a, views, V, o);
    return hipGetLastError();
}

hipError_t launch_preprocess_bwd_batch(const PreArgs& a, const BatchView* views, int V,
                                       const PreBwdIn& in, const float* dL_dopacity_view,
                                       const PreBwdOut& o, hipStream_t s)
{
    if (a.P == 0) return hipSuccess;
    const dim3 grid((a.P + kPreThreads - 1) / kPreThreads);
    if (a.shs) hipLaunchKernelGGL(k_preprocess_bwd_batch_sh, grid, dim3(kPreThreads), 0, s, a, views, V, in, o);
    hipLaunchKernelGGL(k_preprocess_bwd_batch_geom, grid, dim3(kPreThreads), 0, s, a, views, V, in,
                       dL_dopacity_view, o);
    return hipGetLastError();
}

hipError_t launch_preprocess_bwd(const PreArgs& a, const PreBwdIn& in, const PreBwdOut& o,
                                 hipStream_t s)
{
    if (a.P == 0) return hipSuccess;
    hipLaunchKernelGGL(k_preprocess_bwd, dim3((a.P + kPreThreads - 1) / kPreThreads), dim3(kPreThreads), 0, s, a, in, o);
    return hipGetLastError();
}

}  // namespace fgs
