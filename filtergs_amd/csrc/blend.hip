// blend.hip -- per-tile alpha blending, forward and backward.
//
// Forward: antialias/cuda_rasterizer/forward.cu:406-567 (renderCUDA).
// Backward: antialias/cuda_rasterizer/backward.cu:399-557 (renderCUDA).
//
// One workgroup per 16x16 tile, one pixel per lane: 256 lanes = four wave64s,
// each wave covering a 16x4 pixel strip.  Gaussians of the tile's range are
// staged in LDS in batches of 256 together with a conservative bounding box
// of the region where alpha >= 1/255 (the ellipse q <= 2 ln(255 o) of the
// conic, slightly inflated); a wave whose strip misses that box skips the
// Gaussian with one uniform branch -- no lane of the strip could blend it.
// point_weight (the maximum T*alpha of each Gaussian over all pixels) is
// reduced in LDS per batch and published with one float-as-uint atomicMax
// per (tile, Gaussian), so it is exact and deterministic.
// The backward visits only positions below the tile's largest n_contrib,
// pre-reduces the nine per-pixel gradient terms of each Gaussian across the
// wave with a reduce-scatter butterfly (ten shuffles for eight values, six for
// the ninth), sums the four waves in LDS and issues one atomic add per term
// per (tile, Gaussian).
#include "fgs_device.h"
#include "fgs_internal.h"

namespace fgs {

constexpr float kLog2e = 1.4426950408889634f;

// Conservative half extents of {alpha >= 1/255} for conic (A, B, C) and
// opacity o, centred on the mean; negative when the Gaussian can never reach
// alpha >= 1/255; +inf when the conic is not positive definite (no culling).
__device__ __forceinline__ float2 alpha_extent(float A, float B, float C, float o)
{
    const float det = A * C - B * B;
    if (!(det > 0.f) || !(A > 0.f)) return make_float2(INFINITY, INFINITY);
    const float l = 255.f * o;
    if (!(l >= 1.f)) return make_float2(-1.f, -1.f);
    const float tau = 2.f * __logf(l) * 1.001f + 1e-4f;
    const float ex = sqrtf(tau * C / det) * 1.001f + 0.01f;
    const float ey = sqrtf(tau * A / det) * 1.001f + 0.01f;
    return make_float2(ex, ey);
}

__device__ __forceinline__ float4 alpha_box(float2 m, float4 co)
{
    const float2 e = alpha_extent(co.x, co.y, co.z, co.w);
    if (e.x < 0.f) return make_float4(INFINITY, -INFINITY, INFINITY, -INFINITY);
    return make_float4(m.x - e.x, m.x + e.x, m.y - e.y, m.y + e.y);
}

__device__ __forceinline__ bool strip_misses(float4 bb, float x0, float x1, float y0, float y1)
{
    return bb.x > x1 || bb.y < x0 || bb.z > y1 || bb.w < y0;
}

__global__ __launch_bounds__(FGS_BLOCK) void k_blend_fwd(BlendArgs a, float* __restrict__ final_T,
                                                         uint32_t* __restrict__ n_contrib,
                                                         float* __restrict__ out_color,
                                                         int* __restrict__ out_point_id,
                                                         float* __restrict__ out_weight_pixel,
                                                         float* __restrict__ point_weight,
                                                         float* __restrict__ kv_sum,
                                                         uint32_t* __restrict__ kv_cnt)
{
    const int tile = blockIdx.x;
    const int tx = tile % a.gx, ty = tile / a.gx;
    const int tid = threadIdx.x;
    const int wave = tid / FGS_WAVE;
    const int px = tx * FGS_TILE + (tid & (FGS_TILE - 1));
    const int py = ty * FGS_TILE + (tid / FGS_TILE);
    const bool inside = px < a.W && py < a.H;
    const float pfx = (float)px, pfy = (float)py;
    const float sx0 = (float)(tx * FGS_TILE), sx1 = sx0 + (FGS_TILE - 1);
    const float sy0 = (float)(ty * FGS_TILE + wave * 4), sy1 = sy0 + 3.f;
    bool done = !inside;

    const uint2 range = a.ranges[tile];
    const int n = (int)(range.y - range.x);
    const int rounds = (n + FGS_BLOCK - 1) / FGS_BLOCK;

    __shared__ float4 s_g0[FGS_BLOCK];   // x, y, a', b'  (power in log2 units)
    __shared__ float4 s_g1[FGS_BLOCK];   // c', opacity
    __shared__ float4 s_col[FGS_BLOCK];  // r, g, b, id bits
    __shared__ float4 s_bb[FGS_BLOCK];   // alpha >= 1/255 box
    __shared__ uint32_t s_w[FGS_BLOCK];
    __shared__ float s_kv[FGS_BLOCK];
    __shared__ uint32_t s_kc[FGS_BLOCK];

    float T = 1.0f;
    float C0 = 0.f, C1 = 0.f, C2 = 0.f;
    uint32_t contributor = 0, last_contributor = 0;
    float weight_max = 0.f;
    int max_id = -1;
    int todo = n;

    for (int r = 0; r < rounds; ++r, todo -= FGS_BLOCK) {
        const int num_done = __syncthreads_count(done);
        if (num_done == FGS_BLOCK) break;
        const int k = r * FGS_BLOCK + tid;
        if (k < n) {
            const uint32_t g = (uint32_t)a.pairs[range.x + k];
            const float2 m = a.means2D[g];
            const float4 co = a.conic_opacity[g];
            s_g0[tid] = make_float4(m.x, m.y, -0.5f * kLog2e * co.x, -kLog2e * co.y);
            s_g1[tid] = make_float4(-0.5f * kLog2e * co.z, co.w, 0.f, 0.f);
            s_col[tid] = make_float4(a.colors[3 * g + 0], a.colors[3 * g + 1], a.colors[3 * g + 2],
                                     __uint_as_float(g));
            s_bb[tid] = alpha_box(m, co);
        }
        s_w[tid] = 0u;
        s_kv[tid] = 0.f;
        s_kc[tid] = 0u;
        __syncthreads();
        const int cnt = todo < FGS_BLOCK ? todo : FGS_BLOCK;
        for (int j = 0; !done && j < cnt; ++j) {
            contributor++;
            if (strip_misses(s_bb[j], sx0, sx1, sy0, sy1)) continue;
            const float4 g0 = s_g0[j];
            const float4 g1 = s_g1[j];
            const float dx = g0.x - pfx, dy = g0.y - pfy;
            const float p2 = g0.z * dx * dx + g1.x * dy * dy + g0.w * dx * dy;
            if (p2 > 0.0f) continue;
            const float alpha = fminf(0.99f, g1.y * __builtin_amdgcn_exp2f(p2));
            if (alpha < 1.0f / 255.0f) continue;
            const float test_T = T * (1.0f - alpha);
            if (test_T < 0.0001f) {
                done = true;
                continue;
            }
            const float w = alpha * T;
            const float4 col = s_col[j];
            C0 += col.x * w;
            C1 += col.y * w;
            C2 += col.z * w;
            if (w > weight_max) {
                weight_max = w;
                max_id = (int)__float_as_uint(col.w);
            }
            atomicMax(&s_w[j], __float_as_uint(w));
            if (kv_sum) {
                atomicAdd(&s_kv[j], w);
                atomicAdd(&s_kc[j], 1u);
            }
            T = test_T;
            last_contributor = contributor;
        }
        __syncthreads();
        if (k < n) {
            const uint32_t g = __float_as_uint(s_col[tid].w);
            if (s_w[tid] != 0u) atomicMax((uint32_t*)&point_weight[g], s_w[tid]);
            if (kv_sum) {
                // each kv pair belongs to exactly one tile: plain stores
                kv_sum[range.x + k] = s_kv[tid];
                kv_cnt[range.x + k] = s_kc[tid];
            }
        }
    }

    if (inside) {
        const int pix = a.W * py + px;
        const int HW = a.W * a.H;
        final_T[pix] = T;
        n_contrib[pix] = last_contributor;
        out_color[pix] = C0 + T * a.bg[0];
        out_color[HW + pix] = C1 + T * a.bg[1];
        out_color[2 * HW + pix] = C2 + T * a.bg[2];
        out_point_id[pix] = max_id;
        out_weight_pixel[pix] = weight_max;
    }
}

hipError_t launch_blend_fwd(const BlendArgs& a, float* final_T, uint32_t* n_contrib,
                            float* out_color, int* out_point_id, float* out_weight_pixel,
                            float* point_weight, float* kv_sum, uint32_t* kv_cnt, hipStream_t s)
{
    const int T = a.gx * a.gy;
    if (T == 0) return hipSuccess;
    hipLaunchKernelGGL(k_blend_fwd, dim3(T), dim3(FGS_BLOCK), 0, s, a, final_T, n_contrib, out_color,
                       out_point_id, out_weight_pixel, point_weight, kv_sum, kv_cnt);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Backward.
// ---------------------------------------------------------------------------

// Reduce-scatter of 8 values across the wave: on return, lanes whose low
// three bits are zero hold the wave total of value (lane >> 3).
__device__ __forceinline__ float wave_reduce_scatter8(const float v[8])
{
    const int lane = lane_id();
    float w4[4];
    {
        const bool h = (lane >> 5) & 1;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float send = h ? v[k] : v[k + 4];
            const float keep = h ? v[k + 4] : v[k];
            w4[k] = keep + __shfl_xor(send, 32, FGS_WAVE);
        }
    }
    float w2[2];
    {
        const bool h = (lane >> 4) & 1;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const float send = h ? w4[k] : w4[k + 2];
            const float keep = h ? w4[k + 2] : w4[k];
            w2[k] = keep + __shfl_xor(send, 16, FGS_WAVE);
        }
    }
    float z;
    {
        const bool h = (lane >> 3) & 1;
        const float send = h ? w2[0] : w2[1];
        const float keep = h ? w2[1] : w2[0];
        z = keep + __shfl_xor(send, 8, FGS_WAVE);
    }
    z += __shfl_xor(z, 4, FGS_WAVE);
    z += __shfl_xor(z, 2, FGS_WAVE);
    z += __shfl_xor(z, 1, FGS_WAVE);
    return z;
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

__global__ __launch_bounds__(FGS_BLOCK) void k_blend_bwd(BlendArgs a, const float* __restrict__ final_Ts,
                                                         const uint32_t* __restrict__ n_contrib,
                                                         const float* __restrict__ dL_dpix,
                                                         float* __restrict__ dL_dmean2D,
                                                         float* __restrict__ dL_dconic,
                                                         float* __restrict__ dL_dopacity,
                                                         float* __restrict__ dL_dcolors, int exact)
{
    const int tile = blockIdx.x;
    const int tx = tile % a.gx, ty = tile / a.gx;
    const int tid = threadIdx.x;
    const int wave = tid / FGS_WAVE;
    const int lane = lane_id();
    const int px = tx * FGS_TILE + (tid & (FGS_TILE - 1));
    const int py = ty * FGS_TILE + (tid / FGS_TILE);
    const bool inside = px < a.W && py < a.H;
    const float pfx = (float)px, pfy = (float)py;
    const float sx0 = (float)(tx * FGS_TILE), sx1 = sx0 + (FGS_TILE - 1);
    const float sy0 = (float)(ty * FGS_TILE + wave * 4), sy1 = sy0 + 3.f;
    const int pix = a.W * py + px;
    const int HW = a.W * a.H;

    const uint2 range = a.ranges[tile];

    __shared__ float4 s_g0[FGS_BLOCK];   // x, y, A, B
    __shared__ float4 s_g1[FGS_BLOCK];   // C, opacity
    __shared__ float4 s_col[FGS_BLOCK];  // r, g, b, id bits
    __shared__ float4 s_bb[FGS_BLOCK];
    __shared__ float s_g[FGS_BLOCK / FGS_WAVE][kGradTerms][FGS_BLOCK];
    __shared__ uint32_t s_wmax[FGS_BLOCK / FGS_WAVE];

    const float T_final = inside ? final_Ts[pix] : 0.f;
    float T = T_final;
    const uint32_t last_contributor = inside ? n_contrib[pix] : 0u;
    // positions >= the largest n_contrib of the tile contribute to no pixel
    const uint32_t wmax = wave_max_u32(last_contributor);
    if (lane == 0) s_wmax[wave] = wmax;
    __syncthreads();
    uint32_t bmax = 0;
#pragma unroll
    for (int w = 0; w < FGS_BLOCK / FGS_WAVE; ++w) bmax = s_wmax[w] > bmax ? s_wmax[w] : bmax;
    const int n = (int)min(bmax, range.y - range.x);
    const int rounds = (n + FGS_BLOCK - 1) / FGS_BLOCK;

    float dpix0 = 0.f, dpix1 = 0.f, dpix2 = 0.f;
    if (inside) {
        dpix0 = dL_dpix[pix];
        dpix1 = dL_dpix[HW + pix];
        dpix2 = dL_dpix[2 * HW + pix];
    }
    const float bg_dot = a.bg[0] * dpix0 + a.bg[1] * dpix1 + a.bg[2] * dpix2;
    float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f;
    float last_alpha = 0.f;
    float lc0 = 0.f, lc1 = 0.f, lc2 = 0.f;
    const float ddelx_dx = 0.5f * (float)a.W;
    const float ddely_dy = 0.5f * (float)a.H;
    int todo = n;

    for (int r = 0; r < rounds; ++r, todo -= FGS_BLOCK) {
        __syncthreads();
        const int k = r * FGS_BLOCK + tid;
        if (k < n) {
            // back to front: position n-1-k of the tile's list
            const uint32_t g = (uint32_t)a.pairs[range.x + n - 1 - k];
            const float2 m = a.means2D[g];
            const float4 co = a.conic_opacity[g];
            s_g0[tid] = make_float4(m.x, m.y, co.x, co.y);
            s_g1[tid] = make_float4(co.z, co.w, 0.f, 0.f);
            s_col[tid] = make_float4(a.colors[3 * g + 0], a.colors[3 * g + 1], a.colors[3 * g + 2],
                                     __uint_as_float(g));
            s_bb[tid] = alpha_box(m, co);
        }
#pragma unroll
        for (int w = 0; w < FGS_BLOCK / FGS_WAVE; ++w)
#pragma unroll
            for (int t = 0; t < kGradTerms; ++t) s_g[w][t][tid] = 0.f;
        __syncthreads();
        const int cnt = todo < FGS_BLOCK ? todo : FGS_BLOCK;
        for (int j = 0; j < cnt; ++j) {
            const uint32_t position = (uint32_t)(n - 1 - (r * FGS_BLOCK + j));
            if (position >= wmax) continue;                           // wave-uniform
            if (strip_misses(s_bb[j], sx0, sx1, sy0, sy1)) continue;  // wave-uniform
            float v[8] = { 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f };
            float vop = 0.f;
            bool act = false;
            if (position < last_contributor) {
                const float4 g0 = s_g0[j];
                const float4 g1 = s_g1[j];
                const float dx = g0.x - pfx, dy = g0.y - pfy;
                const float power = -0.5f * (g0.z * dx * dx + g1.x * dy * dy) - g0.w * dx * dy;
                if (power <= 0.0f) {
                    const float G = __builtin_amdgcn_exp2f(power * kLog2e);
                    const float alpha_raw = g1.y * G;
                    const float alpha = fminf(0.99f, alpha_raw);
                    if (alpha >= 1.0f / 255.0f) {
                        act = true;
                        const float inv = __builtin_amdgcn_rcpf(1.f - alpha);
                        T = T * inv;
                        const float dch = alpha * T;
                        const float4 col = s_col[j];
                        acc0 = last_alpha * lc0 + (1.f - last_alpha) * acc0;
                        acc1 = last_alpha * lc1 + (1.f - last_alpha) * acc1;
                        acc2 = last_alpha * lc2 + (1.f - last_alpha) * acc2;
                        lc0 = col.x;
                        lc1 = col.y;
                        lc2 = col.z;
                        float dL_dalpha = (col.x - acc0) * dpix0 + (col.y - acc1) * dpix1 + (col.z - acc2) * dpix2;
                        v[0] = dch * dpix0;
                        v[1] = dch * dpix1;
                        v[2] = dch * dpix2;
                        dL_dalpha *= T;
                        last_alpha = alpha;
                        dL_dalpha += (-T_final * inv) * bg_dot;
                        const float slope = (exact && alpha_raw > 0.99f) ? 0.f : 1.f;
                        const float dL_dG = g1.y * dL_dalpha * slope;
                        const float gdx = G * dx, gdy = G * dy;
                        const float dG_ddelx = -gdx * g0.z - gdy * g0.w;
                        const float dG_ddely = -gdy * g1.x - gdx * g0.w;
                        v[3] = dL_dG * dG_ddelx * ddelx_dx;
                        v[4] = dL_dG * dG_ddely * ddely_dy;
                        v[5] = -0.5f * gdx * dx * dL_dG;
                        v[6] = -0.5f * gdx * dy * dL_dG;
                        v[7] = -0.5f * gdy * dy * dL_dG;
                        vop = G * dL_dalpha * slope;
                    }
                }
            }
            if (__any(act)) {
                const float z = wave_reduce_scatter8(v);
                const float zop = wave_sum(vop);
                if ((lane & 7) == 0) s_g[wave][lane >> 3][j] = z;
                if (lane == 0) s_g[wave][8][j] = zop;
            }
        }
        __syncthreads();
        if (k < n) {
            const uint32_t g = __float_as_uint(s_col[tid].w);
            float t[kGradTerms];
#pragma unroll
            for (int q = 0; q < kGradTerms; ++q)
                t[q] = s_g[0][q][tid] + s_g[1][q][tid] + s_g[2][q][tid] + s_g[3][q][tid];
            if (t[0] != 0.f) atomic_add_f32(&dL_dcolors[3 * g + 0], t[0]);
            if (t[1] != 0.f) atomic_add_f32(&dL_dcolors[3 * g + 1], t[1]);
            if (t[2] != 0.f) atomic_add_f32(&dL_dcolors[3 * g + 2], t[2]);
            if (t[3] != 0.f) atomic_add_f32(&dL_dmean2D[3 * g + 0], t[3]);
            if (t[4] != 0.f) atomic_add_f32(&dL_dmean2D[3 * g + 1], t[4]);
            if (t[5] != 0.f) atomic_add_f32(&dL_dconic[3 * g + 0], t[5]);
            if (t[6] != 0.f) atomic_add_f32(&dL_dconic[3 * g + 1], t[6]);
            if (t[7] != 0.f) atomic_add_f32(&dL_dconic[3 * g + 2], t[7]);
            if (t[8] != 0.f) atomic_add_f32(&dL_dopacity[g], t[8]);
        }
    }
}

hipError_t launch_blend_bwd(const BlendArgs& a, const float* final_T, const uint32_t* n_contrib,
                            const float* dL_dpix, float* dL_dmean2D, float* dL_dconic,
                            float* dL_dopacity, float* dL_dcolors, int exact, hipStream_t s)
{
    const int T = a.gx * a.gy;
    if (T == 0) return hipSuccess;
    hipLaunchKernelGGL(k_blend_bwd, dim3(T), dim3(FGS_BLOCK), 0, s, a, final_T, n_contrib, dL_dpix,
                       dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors, exact);
    return hipGetLastError();
}

}  // namespace fgs
