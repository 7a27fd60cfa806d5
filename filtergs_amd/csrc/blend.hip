// blend.hip -- per-tile alpha blending, forward and backward.
//
// Forward: antialias/cuda_rasterizer/forward.cu:406-567 (renderCUDA).
// Backward: antialias/cuda_rasterizer/backward.cu:399-557 (renderCUDA).
//
// One workgroup of 256 lanes (four wave64s) per 16x16 tile, one pixel per
// lane; wave w owns the 16x4 strip of rows 4w..4w+3.  Gaussians of the tile's
// range are staged in LDS in batches of BATCH, and each wave walks a compacted
// list of the batch slots that concern its strip (ballot + prefix count, order
// preserved), so the inner loop carries no per-Gaussian cull test.
//
// Forward: a slot enters strip w's list when a conservative box of the region
// where alpha >= 1/255 (the ellipse q <= 2 ln(255 o) of the conic, slightly
// inflated) touches the strip.  Per slot and strip the largest T*alpha is
// kept in LDS; after the batch it gives point_weight (one float-as-uint
// atomicMax per (tile, Gaussian): exact and deterministic) and the pair's
// strip mask -- the strips where the pair actually blended -- stored next to
// the sorted pairs.
// Backward: the lists come from those masks, so every listed slot blends into
// at least one pixel of the strip and pairs that blended nowhere are never
// loaded.  The nine per-pixel gradient terms of each Gaussian are reduced
// across the wave (a DPP reduce-scatter inside 16-lane rows, then two lane
// swaps), added over the waves with LDS float atomics, and published with one
// global atomic add per term per (tile, Gaussian).
#include "fgs_device.h"
#include "fgs_internal.h"

namespace fgs {

constexpr float kLog2e = 1.4426950408889634f;
constexpr int kStrips = FGS_BLOCK / FGS_WAVE;   // 4 waves, one 16x4 strip each

// Minimum over the rectangle [dx0, dx1] x [dy0, dy1] (offsets from the mean)
// of q(d) = A dx^2 + 2 B dx dy + C dy^2 for a positive definite conic: zero
// when the mean lies inside, else attained on an edge, where q is a 1-D
// convex quadratic minimised at its clamped stationary point.
__device__ __forceinline__ float rect_min_q(float A, float B, float C, float rA, float rC, float dx0,
                                            float dx1, float dy0, float dy1)
{
    if (dx0 <= 0.f && dx1 >= 0.f && dy0 <= 0.f && dy1 >= 0.f) return 0.f;
    float qmin = INFINITY;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const float dx = e ? dx1 : dx0;
        const float dy = fminf(fmaxf(-B * dx * rC, dy0), dy1);
        qmin = fminf(qmin, A * dx * dx + 2.f * B * dx * dy + C * dy * dy);
    }
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const float dy = e ? dy1 : dy0;
        const float dx = fminf(fmaxf(-B * dy * rA, dx0), dx1);
        qmin = fminf(qmin, A * dx * dx + 2.f * B * dx * dy + C * dy * dy);
    }
    return qmin;
}

// Strips of the tile whose top-left pixel is (tx0, ty0) that the region where
// alpha >= 1/255 -- q(p - mean) <= 2 ln(255 o), slightly inflated -- can
// reach: the exact minimum of q over each strip's pixel rectangle.  A conic
// that is not positive definite keeps every strip.
__device__ __forceinline__ uint32_t strip_bits_ellipse(float2 m, float4 co, float tx0, float ty0)
{
    const float A = co.x, B = co.y, C = co.z;
    const float det = A * C - B * B;
    if (!(det > 0.f) || !(A > 0.f)) return (1u << kStrips) - 1u;
    const float l = 255.f * co.w;
    if (!(l >= 1.f)) return 0u;
    const float tau = 2.f * __logf(l) * 1.001f + 1e-4f;
    const float rA = 1.f / A, rC = 1.f / C;
    const float dx0 = tx0 - m.x, dx1 = tx0 + (FGS_TILE - 1) - m.x;
    uint32_t bits = 0u;
#pragma unroll
    for (int w = 0; w < kStrips; ++w) {
        const float dy0 = ty0 + 4.f * w - m.y;
        if (rect_min_q(A, B, C, rA, rC, dx0, dx1, dy0, dy0 + 3.f) <= tau) bits |= 1u << w;
    }
    return bits;
}

// Per-wave compacted lists of batch positions.  Every lane of the block calls
// it (q = threadIdx.x is the batch slot it staged; bits its strip mask, zero
// for empty slots).  Two barriers; on return s_list[w][0 .. count(w)) holds
// the slots whose bit w is set, in increasing order, and the count for the
// calling wave is returned.
template <int BATCH>
__device__ __forceinline__ int build_lists(uint32_t bits, uint16_t (*s_list)[BATCH],
                                           uint32_t (*s_wcnt)[kStrips])
{
    const int tid = threadIdx.x;
    const int wave = tid / FGS_WAVE;
    const int lane = lane_id();
    int rank[kStrips];
#pragma unroll
    for (int w = 0; w < kStrips; ++w) {
        const uint64_t bal = __ballot((bits >> w) & 1u);
        rank[w] = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
        if (lane == 0) s_wcnt[wave][w] = (uint32_t)__popcll(bal);
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kStrips; ++w) {
        if ((bits >> w) & 1u) {
            int off = 0;
            for (int sw = 0; sw < wave; ++sw) off += (int)s_wcnt[sw][w];
            s_list[w][off + rank[w]] = (uint16_t)tid;
        }
    }
    int total = 0;
#pragma unroll
    for (int sw = 0; sw < kStrips; ++sw) total += (int)s_wcnt[sw][wave];
    __syncthreads();
    return total;
}

// DPP read of another lane's value inside a 16-lane row; written as
// update_dpp with a zero old value so the compiler folds it into the
// consuming add (v_add_f32_dpp).
template <int CTRL>
__device__ __forceinline__ float dpp(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

// x[l] + x[l ^ 32] and x[l] + x[l ^ 16] in every lane with the gfx950 lane
// swaps (v_permlane32_swap / v_permlane16_swap): VALU only, no LDS round trip.
// Both halves add the pair in the same order, so every lane gets the same bits.
__device__ __forceinline__ float sum_xor32(float x)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false,
                                                    false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float sum_xor16(float x)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false,
                                                    false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
constexpr int kDppRowMirror = 0x140;      // lane r <-> 15 - r inside each 16-lane row
constexpr int kDppRowHalfMirror = 0x141;  // lane r <-> 7 - r inside each 8-lane half row
constexpr int kDppXor2 = 0x4E;            // quad_perm [2,3,0,1]
constexpr int kDppXor1 = 0xB1;            // quad_perm [1,0,3,2]

// Wave total of one value (every lane receives it).
__device__ __forceinline__ float wave_total(float x)
{
    x += dpp<0x128>(x);   // row_ror 8
    x += dpp<0x124>(x);   // row_ror 4
    x += dpp<0x122>(x);   // row_ror 2
    x += dpp<0x121>(x);   // row_ror 1
    return sum_xor32(sum_xor16(x));
}

// Exponent of the Gaussian in log2 units at offset (dx, dy) from its mean,
// from the staged terms (a', b') = g0.zw and c' = g1.x.  Shared by the
// forward and the backward so that both take the same alpha decisions.
__device__ __forceinline__ float exponent2(float4 g0, float2 g1, float dx, float dy)
{
#pragma clang fp contract(off)
    return fmaf(g0.z * dx, dx, fmaf(g1.x * dy, dy, (g0.w * dx) * dy));
}

// ---------------------------------------------------------------------------
// Forward
// ---------------------------------------------------------------------------
// Forward of one tile (one workgroup of FGS_BLOCK lanes).
template <int BATCH>
__device__ __forceinline__ void blend_fwd_tile(const BlendArgs& a, int tile, float* __restrict__ final_T,
                                               uint32_t* __restrict__ n_contrib,
                                               float* __restrict__ out_color,
                                               int* __restrict__ out_point_id,
                                               float* __restrict__ out_weight_pixel,
                                               float* __restrict__ point_weight,
                                               float* __restrict__ kv_sum, uint32_t* __restrict__ kv_cnt)
{
    // Every product and sum of the per-pixel chain is written out (explicit
    // fmaf, no contraction by the compiler), so all work shapes round alike
    // and render bit-identical images.
#pragma clang fp contract(off)
    static_assert(BATCH <= FGS_BLOCK, "one staging lane per batch slot");
    const int tx = tile % a.gx, ty = tile / a.gx;
    const int tid = threadIdx.x;
    const int wave = tid / FGS_WAVE;
    const int px = tx * FGS_TILE + (tid & (FGS_TILE - 1));
    const int py = ty * FGS_TILE + (tid / FGS_TILE);
    const bool inside = px < a.W && py < a.H;
    const float pfx = (float)px, pfy = (float)py;
    const float tx0 = (float)(tx * FGS_TILE), ty0 = (float)(ty * FGS_TILE);
    bool done = !inside;

    const uint2 range = a.ranges[tile];
    const int n = (int)(range.y - range.x);
    const int rounds = (n + BATCH - 1) / BATCH;

    __shared__ float4 s_g0[BATCH];   // x, y, a', b'  (power in log2 units)
    __shared__ float2 s_g1[BATCH];   // c', opacity
    __shared__ float4 s_col[BATCH];  // r, g, b, id bits
    __shared__ uint32_t s_w[kStrips][BATCH];   // per strip: max T*alpha (float bits)
    __shared__ float s_kv[BATCH];
    __shared__ uint32_t s_kc[BATCH];
    __shared__ uint16_t s_list[kStrips][BATCH];
    __shared__ uint32_t s_wcnt[kStrips][kStrips];

    float T = 1.0f;
    float C0 = 0.f, C1 = 0.f, C2 = 0.f;
    uint32_t last = 0;
    float wmax = 0.f;
    int max_id = -1;

    for (int r = 0; r < rounds; ++r) {
        const int num_done = __syncthreads_count(done);
        if (num_done == FGS_BLOCK) break;
        const int q = tid;
        const int k = r * BATCH + q;
        uint32_t bits = 0u;
        if (q < BATCH) {
            if (k < n) {
                const uint32_t g = (uint32_t)a.pairs[range.x + k];
                const float2 m = a.means2D[g];
                const float4 co = a.conic_opacity[g];
                s_g0[q] = make_float4(m.x, m.y, -0.5f * kLog2e * co.x, -kLog2e * co.y);
                s_g1[q] = make_float2(-0.5f * kLog2e * co.z, co.w);
                s_col[q] = make_float4(a.colors[3 * g + 0], a.colors[3 * g + 1], a.colors[3 * g + 2],
                                       __uint_as_float(g));
                bits = strip_bits_ellipse(m, co, tx0, ty0);
            }
#pragma unroll
            for (int w = 0; w < kStrips; ++w) s_w[w][q] = 0u;
            s_kv[q] = 0.f;
            s_kc[q] = 0u;
        }
        const int cnt = build_lists<BATCH>(bits, s_list, s_wcnt);
        for (int i = 0; !done && i < cnt; ++i) {
            const int j = s_list[wave][i];
            const float4 g0 = s_g0[j];
            const float2 g1 = s_g1[j];
            const float dx = g0.x - pfx, dy = g0.y - pfy;
            const float p2 = exponent2(g0, g1, dx, dy);
            const float alpha = fminf(0.99f, g1.y * __builtin_amdgcn_exp2f(p2));
            const float test_T = T * (1.0f - alpha);
            const bool pass = !(p2 > 0.0f) && !(alpha < 1.0f / 255.0f);
            const bool stop = pass && test_T < 0.0001f;
            done = done || stop;
            if (pass && !stop) {
                const float w = alpha * T;
                const float4 col = s_col[j];
                C0 = fmaf(col.x, w, C0);
                C1 = fmaf(col.y, w, C1);
                C2 = fmaf(col.z, w, C2);
                if (w > wmax) {
                    wmax = w;
                    max_id = (int)__float_as_uint(col.w);
                }
                atomicMax(&s_w[wave][j], __float_as_uint(w));
                if (kv_sum) {
                    atomicAdd(&s_kv[j], w);
                    atomicAdd(&s_kc[j], 1u);
                }
                T = test_T;
                last = (uint32_t)(r * BATCH + j + 1);
            }
        }
        __syncthreads();
        if (q < BATCH && k < n) {
            const uint32_t g = __float_as_uint(s_col[q].w);
            uint32_t m = 0u, used = 0u;
#pragma unroll
            for (int w = 0; w < kStrips; ++w) {
                m = s_w[w][q] > m ? s_w[w][q] : m;
                used |= (s_w[w][q] != 0u ? 1u : 0u) << w;
            }
            if (m != 0u) atomicMax((uint32_t*)&point_weight[g], m);
            a.kv_strips[range.x + k] = (uint8_t)used;
            if (kv_sum) {
                // each kv pair belongs to exactly one tile: plain stores
                kv_sum[range.x + k] = s_kv[q];
                kv_cnt[range.x + k] = s_kc[q];
            }
        }
    }

    if (inside) {
        const int pix = a.W * py + px;
        const int HW = a.W * a.H;
        final_T[pix] = T;
        n_contrib[pix] = last;
        out_color[pix] = fmaf(T, a.bg[0], C0);
        out_color[HW + pix] = fmaf(T, a.bg[1], C1);
        out_color[2 * HW + pix] = fmaf(T, a.bg[2], C2);
        out_point_id[pix] = max_id >= 0 ? (int)((uint32_t)max_id - a.id_base) : -1;
        out_weight_pixel[pix] = wmax;
    }
}

template <int BATCH>
__global__ __launch_bounds__(FGS_BLOCK) void k_blend_fwd(BlendArgs a, float* final_T, uint32_t* n_contrib,
                                                         float* out_color, int* out_point_id,
                                                         float* out_weight_pixel, float* point_weight,
                                                         float* kv_sum, uint32_t* kv_cnt)
{
    blend_fwd_tile<BATCH>(a, blockIdx.x, final_T, n_contrib, out_color, out_point_id, out_weight_pixel,
                          point_weight, kv_sum, kv_cnt);
}

// View of a batch that owns workgroup b (views sorted by tile_base).
__device__ __forceinline__ int find_view(const BatchView* views, int V, uint32_t b)
{
    int v = 0;
    while (v + 1 < V && views[v + 1].tile_base <= b) ++v;
    return v;
}

// The view-dependent fields of a BlendArgs; per-pixel pointers are offset by
// the caller with pix_base.
__device__ __forceinline__ BlendArgs bind_blend(const BlendArgs& common, const BatchView& vw, int v)
{
    BlendArgs a = common;
    a.id_base = (uint32_t)v * common.id_stride;
    a.ranges = common.ranges + vw.tile_base;
    a.W = vw.W;
    a.H = vw.H;
    a.gx = vw.gx;
    a.gy = vw.gy;
    a.bg = vw.bg;
    return a;
}

// Every tile of every view of a batch in one launch: workgroup b renders tile
// b - tile_base of its view.  Per-pixel outputs are the views' [3,H,W] / [H,W]
// blocks concatenated (view v starts at pixel pix_base).
template <int BATCH>
__global__ __launch_bounds__(FGS_BLOCK) void k_blend_fwd_batch(BlendArgs common, const BatchView* views,
                                                               int V, float* final_T,
                                                               uint32_t* n_contrib, float* out_color,
                                                               int* out_point_id,
                                                               float* out_weight_pixel,
                                                               float* point_weight)
{
    const int v = find_view(views, V, blockIdx.x);
    const BatchView& vw = views[v];
    const BlendArgs a = bind_blend(common, vw, v);
    const size_t pb = vw.pix_base;
    blend_fwd_tile<BATCH>(a, (int)(blockIdx.x - vw.tile_base), final_T + pb, n_contrib + pb,
                          out_color + 3 * pb, out_point_id + pb, out_weight_pixel + pb, point_weight,
                          nullptr, nullptr);
}

hipError_t launch_blend_fwd_batch(const BlendArgs& common, const BatchView* views, int V, int T_total,
                                  float* final_T, uint32_t* n_contrib, float* out_color,
                                  int* out_point_id, float* out_weight_pixel, float* point_weight,
                                  hipStream_t s)
{
    if (T_total == 0) return hipSuccess;
    if (common.batch == 128)
        hipLaunchKernelGGL(k_blend_fwd_batch<128>, dim3(T_total), dim3(FGS_BLOCK), 0, s, common, views, V,
                           final_T, n_contrib, out_color, out_point_id, out_weight_pixel, point_weight);
    else
        hipLaunchKernelGGL(k_blend_fwd_batch<256>, dim3(T_total), dim3(FGS_BLOCK), 0, s, common, views, V,
                           final_T, n_contrib, out_color, out_point_id, out_weight_pixel, point_weight);
    return hipGetLastError();
}

hipError_t launch_blend_fwd(const BlendArgs& a, float* final_T, uint32_t* n_contrib,
                            float* out_color, int* out_point_id, float* out_weight_pixel,
                            float* point_weight, float* kv_sum, uint32_t* kv_cnt, hipStream_t s)
{
    const int T = a.gx * a.gy;
    if (T == 0) return hipSuccess;
    if (a.batch == 128)
        hipLaunchKernelGGL(k_blend_fwd<128>, dim3(T), dim3(FGS_BLOCK), 0, s, a, final_T, n_contrib,
                           out_color, out_point_id, out_weight_pixel, point_weight, kv_sum, kv_cnt);
    else
        hipLaunchKernelGGL(k_blend_fwd<256>, dim3(T), dim3(FGS_BLOCK), 0, s, a, final_T, n_contrib,
                           out_color, out_point_id, out_weight_pixel, point_weight, kv_sum, kv_cnt);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Backward
// ---------------------------------------------------------------------------

// Reduce-scatter of 8 values across the wave.  Inside every 16-lane row the
// halving steps pair lanes with DPP permutations (mirror: bit 3, half mirror:
// bit 2, quad_perm: bits 1 and 0), leaving in lane r the row sum of value
// 4*bit3 + 2*bit2 + bit1; two lane swaps then add the four rows.  On return
// lane l < 16 holds the wave total of value l >> 1.
__device__ __forceinline__ float wave_reduce_scatter8(const float v[8])
{
    const int lane = lane_id();
    const bool h3 = (lane >> 3) & 1, h2 = (lane >> 2) & 1, h1 = (lane >> 1) & 1;
    float w4[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float send = h3 ? v[k] : v[k + 4];
        const float keep = h3 ? v[k + 4] : v[k];
        w4[k] = keep + dpp<kDppRowMirror>(send);
    }
    float w2[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const float send = h2 ? w4[k] : w4[k + 2];
        const float keep = h2 ? w4[k + 2] : w4[k];
        w2[k] = keep + dpp<kDppRowHalfMirror>(send);
    }
    float z;
    {
        const float send = h1 ? w2[0] : w2[1];
        const float keep = h1 ? w2[1] : w2[0];
        z = keep + dpp<kDppXor2>(send);
    }
    z += dpp<kDppXor1>(z);
    return sum_xor32(sum_xor16(z));
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(v, off, FGS_WAVE);
        v = o > v ? o : v;
    }
    return v;
}

constexpr int kGradTerms = 9;   // dcolor x3, dmean2D x2, dconic x3, dopacity

// Backward of one tile (one workgroup of FGS_BLOCK lanes).
template <int BATCH>
__device__ __forceinline__ void blend_bwd_tile(const BlendArgs& a, int tile, const float* __restrict__ final_Ts,
                                               const uint32_t* __restrict__ n_contrib,
                                               const float* __restrict__ dL_dpix,
                                               float* __restrict__ dL_dmean2D,
                                               float* __restrict__ dL_dconic,
                                               float* __restrict__ dL_dopacity,
                                               float* __restrict__ dL_dcolors, int exact)
{
    static_assert(BATCH <= FGS_BLOCK, "one staging lane per batch slot");
    const int tx = tile % a.gx, ty = tile / a.gx;
    const int tid = threadIdx.x;
    const int wave = tid / FGS_WAVE;
    const int lane = lane_id();
    const int px = tx * FGS_TILE + (tid & (FGS_TILE - 1));
    const int py = ty * FGS_TILE + (tid / FGS_TILE);
    const bool inside = px < a.W && py < a.H;
    const float pfx = (float)px, pfy = (float)py;
    const int HW = a.W * a.H;
    const int pix = a.W * py + px;

    const uint2 range = a.ranges[tile];

    __shared__ float4 s_g0[BATCH];   // x, y, a', b'  (power in log2 units, as the forward)
    __shared__ float2 s_g1[BATCH];   // c', opacity
    __shared__ float4 s_con[BATCH];  // conic A, B, C
    __shared__ float4 s_col[BATCH];  // r, g, b, id bits
    __shared__ float s_g[kGradTerms][BATCH];
    __shared__ uint32_t s_wmax[kStrips];
    __shared__ uint16_t s_list[kStrips][BATCH];
    __shared__ uint32_t s_wcnt[kStrips][kStrips];

    const float T_final = inside ? final_Ts[pix] : 0.f;
    float T = T_final;
    const uint32_t last = inside ? n_contrib[pix] : 0u;
    const float dp0 = inside ? dL_dpix[pix] : 0.f;
    const float dp1 = inside ? dL_dpix[HW + pix] : 0.f;
    const float dp2 = inside ? dL_dpix[2 * HW + pix] : 0.f;
    const float bg_dot = a.bg[0] * dp0 + a.bg[1] * dp1 + a.bg[2] * dp2;
    const float tfb = -T_final * bg_dot;   // dL/dalpha gets -T_final / (1 - alpha) * (bg . dL/dpix)
    float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f;
    float last_alpha = 0.f, lc0 = 0.f, lc1 = 0.f, lc2 = 0.f;

    // positions at or past the largest n_contrib of the tile blend into no pixel
    const uint32_t wmax = wave_max_u32(last);
    if (lane == 0) s_wmax[wave] = wmax;
    __syncthreads();
    uint32_t bmax = 0;
#pragma unroll
    for (int w = 0; w < kStrips; ++w) bmax = s_wmax[w] > bmax ? s_wmax[w] : bmax;
    const int n = (int)min(bmax, range.y - range.x);
    const int rounds = (n + BATCH - 1) / BATCH;
    const float ddelx_dx = 0.5f * (float)a.W;
    const float ddely_dy = 0.5f * (float)a.H;

    for (int r = 0; r < rounds; ++r) {
        __syncthreads();
        const int q = tid;
        const int k = r * BATCH + q;
        uint32_t bits = 0u;
        if (q < BATCH) {
            // back to front: position n-1-k of the tile's list.  Only the
            // strips where the forward blended this pair take part, and a
            // pair that blended nowhere is not even loaded.
            if (k < n) bits = a.kv_strips[range.x + n - 1 - k];
            if (bits != 0u) {
                const uint32_t g = (uint32_t)a.pairs[range.x + n - 1 - k];
                const float2 m = a.means2D[g];
                const float4 co = a.conic_opacity[g];
                s_g0[q] = make_float4(m.x, m.y, -0.5f * kLog2e * co.x, -kLog2e * co.y);
                s_g1[q] = make_float2(-0.5f * kLog2e * co.z, co.w);
                s_con[q] = co;
                s_col[q] = make_float4(a.colors[3 * g + 0], a.colors[3 * g + 1], a.colors[3 * g + 2],
                                       __uint_as_float(g));
            }
#pragma unroll
            for (int t = 0; t < kGradTerms; ++t) s_g[t][q] = 0.f;
        }
        const int cnt = build_lists<BATCH>(bits, s_list, s_wcnt);
        for (int i = 0; i < cnt; ++i) {
            const int j = s_list[wave][i];
            const uint32_t position = (uint32_t)(n - 1 - (r * BATCH + j));
            // alpha exactly as the forward computed it (same staged terms,
            // same operations), so both passes take the same decisions
            const float4 g0 = s_g0[j];
            const float2 g1 = s_g1[j];
            const float dx = g0.x - pfx, dy = g0.y - pfy;
            const float p2 = exponent2(g0, g1, dx, dy);
            const float G = __builtin_amdgcn_exp2f(fminf(p2, 0.0f));
            const float alpha_raw = g1.y * G;
            const float alpha = fminf(0.99f, alpha_raw);
            const bool act = position < last && !(p2 > 0.0f) && !(alpha < 1.0f / 255.0f);
            if (!__any(act)) continue;   // wave-uniform
            // every lane computes; inactive lanes contribute zeros and keep
            // their state (all terms stay finite: G <= 1, alpha <= 0.99)
            const float4 col = s_col[j];
            const float inv = __builtin_amdgcn_rcpf(1.f - alpha);
            const float Tn = T * inv;
            const float a0 = last_alpha * lc0 + (1.f - last_alpha) * acc0;
            const float a1 = last_alpha * lc1 + (1.f - last_alpha) * acc1;
            const float a2 = last_alpha * lc2 + (1.f - last_alpha) * acc2;
            const float dLa = ((col.x - a0) * dp0 + (col.y - a1) * dp1 + (col.z - a2) * dp2) * Tn +
                              inv * tfb;
            const bool flows = act && !(exact && alpha_raw > 0.99f);
            const float dch = act ? alpha * Tn : 0.f;
            float v[8];
            v[0] = dch * dp0;
            v[1] = dch * dp1;
            v[2] = dch * dp2;
            // h = dL/dG * G.  dL/dmean2D = -(W/2, H/2) * (A Sx + B Sy, B Sx + C Sy)
            // and dL/dconic = -1/2 (Sxx, Sxy, Syy) with Sx = sum h dx, ...:
            // the conic and the constants are applied once per (tile,
            // Gaussian) when the sums are published
            const float vop = flows ? G * dLa : 0.f;
            const float h = g1.y * vop;
            const float hdx = h * dx, hdy = h * dy;
            v[3] = hdx;
            v[4] = hdy;
            v[5] = hdx * dx;
            v[6] = hdx * dy;
            v[7] = hdy * dy;
            if (act) {
                T = Tn;
                acc0 = a0;
                acc1 = a1;
                acc2 = a2;
                lc0 = col.x;
                lc1 = col.y;
                lc2 = col.z;
                last_alpha = alpha;
            }
            // the waves meet on the same slot: LDS float atomics
            const float z = wave_reduce_scatter8(v);
            const float zop = wave_total(vop);
            if (lane < 16 && (lane & 1) == 0) atomicAdd(&s_g[lane >> 1][j], z);
            if (lane == 16) atomicAdd(&s_g[8][j], zop);
        }
        __syncthreads();
        if (bits != 0u) {
            const uint32_t g = __float_as_uint(s_col[q].w);
            float t[kGradTerms];
#pragma unroll
            for (int e = 0; e < kGradTerms; ++e) t[e] = s_g[e][q];
            if (t[0] != 0.f) atomic_add_f32(&dL_dcolors[3 * g + 0], t[0]);
            if (t[1] != 0.f) atomic_add_f32(&dL_dcolors[3 * g + 1], t[1]);
            if (t[2] != 0.f) atomic_add_f32(&dL_dcolors[3 * g + 2], t[2]);
            if (t[3] != 0.f || t[4] != 0.f) {
                const float4 con = s_con[q];
                atomic_add_f32(&dL_dmean2D[3 * g + 0], -ddelx_dx * (con.x * t[3] + con.y * t[4]));
                atomic_add_f32(&dL_dmean2D[3 * g + 1], -ddely_dy * (con.y * t[3] + con.z * t[4]));
            }
            if (t[5] != 0.f) atomic_add_f32(&dL_dconic[3 * g + 0], -0.5f * t[5]);
            if (t[6] != 0.f) atomic_add_f32(&dL_dconic[3 * g + 1], -0.5f * t[6]);
            if (t[7] != 0.f) atomic_add_f32(&dL_dconic[3 * g + 2], -0.5f * t[7]);
            if (t[8] != 0.f) atomic_add_f32(&dL_dopacity[g], t[8]);
        }
    }
}

template <int BATCH>
__global__ __launch_bounds__(FGS_BLOCK) void k_blend_bwd(BlendArgs a, const float* final_Ts,
                                                         const uint32_t* n_contrib, const float* dL_dpix,
                                                         float* dL_dmean2D, float* dL_dconic,
                                                         float* dL_dopacity, float* dL_dcolors, int exact)
{
    blend_bwd_tile<BATCH>(a, blockIdx.x, final_Ts, n_contrib, dL_dpix, dL_dmean2D, dL_dconic, dL_dopacity,
                          dL_dcolors, exact);
}

template <int BATCH>
__global__ __launch_bounds__(FGS_BLOCK) void k_blend_bwd_batch(BlendArgs common, const BatchView* views,
                                                               int V, const float* final_Ts,
                                                               const uint32_t* n_contrib,
                                                               const float* dL_dpix, float* dL_dmean2D,
                                                               float* dL_dconic, float* dL_dopacity,
                                                               float* dL_dcolors, int exact)
{
    const int v = find_view(views, V, blockIdx.x);
    const BatchView& vw = views[v];
    const BlendArgs a = bind_blend(common, vw, v);
    const size_t pb = vw.pix_base;
    blend_bwd_tile<BATCH>(a, (int)(blockIdx.x - vw.tile_base), final_Ts + pb, n_contrib + pb,
                          dL_dpix + 3 * pb, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors, exact);
}

hipError_t launch_blend_bwd_batch(const BlendArgs& common, const BatchView* views, int V, int T_total,
                                  const float* final_T, const uint32_t* n_contrib, const float* dL_dpix,
                                  float* dL_dmean2D, float* dL_dconic, float* dL_dopacity,
                                  float* dL_dcolors, int exact, hipStream_t s)
{
    if (T_total == 0) return hipSuccess;
#define FGS_BWD(BATCH)                                                                             \
    hipLaunchKernelGGL(k_blend_bwd_batch<BATCH>, dim3(T_total), dim3(FGS_BLOCK), 0, s, common, views, V, \
                       final_T, n_contrib, dL_dpix, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors, exact)
    if (common.bwd_batch == 256) FGS_BWD(256);
    else if (common.bwd_batch == 64) FGS_BWD(64);
    else FGS_BWD(128);
#undef FGS_BWD
    return hipGetLastError();
}

hipError_t launch_blend_bwd(const BlendArgs& a, const float* final_T, const uint32_t* n_contrib,
                            const float* dL_dpix, float* dL_dmean2D, float* dL_dconic,
                            float* dL_dopacity, float* dL_dcolors, int exact, hipStream_t s)
{
    const int T = a.gx * a.gy;
    if (T == 0) return hipSuccess;
#define FGS_BWD(BATCH)                                                                            \
    hipLaunchKernelGGL(k_blend_bwd<BATCH>, dim3(T), dim3(FGS_BLOCK), 0, s, a, final_T, n_contrib,  \
                       dL_dpix, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors, exact)
    if (a.bwd_batch == 256) FGS_BWD(256);
    else if (a.bwd_batch == 64) FGS_BWD(64);
    else FGS_BWD(128);
#undef FGS_BWD
    return hipGetLastError();
}

}  // namespace fgs
