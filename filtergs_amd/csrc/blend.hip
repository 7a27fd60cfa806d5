// blend.hip -- per-tile alpha blending, forward and backward.
//
// Forward: antialias/cuda_rasterizer/forward.cu:406-567 (renderCUDA).
// Backward: antialias/cuda_rasterizer/backward.cu:399-557 (renderCUDA).
//
// One workgroup per 16x16 tile, one pixel per lane: 256 lanes = four wave64s,
// each wave covering four pixel rows.  Gaussians of the tile's range are
// staged in LDS in batches of 256 (position, conic+opacity, colour, id).
// point_weight (the maximum T*alpha of each Gaussian over all pixels) is
// reduced in LDS per batch and published with one float-as-uint atomicMax
// per (tile, Gaussian), so it is exact and deterministic.
// The backward pre-reduces the nine per-pixel gradient terms of each
// Gaussian across the wave with a reduce-scatter butterfly (ten shuffles for
// eight values, six for the ninth), sums the four waves in LDS and issues one
// atomic add per term per (tile, Gaussian).
#include "fgs_device.h"
#include "fgs_internal.h"

namespace fgs {

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
    const int px = tx * FGS_TILE + (tid & (FGS_TILE - 1));
    const int py = ty * FGS_TILE + (tid / FGS_TILE);
    const bool inside = px < a.W && py < a.H;
    const float pfx = (float)px, pfy = (float)py;
    bool done = !inside;

    const uint2 range = a.ranges[tile];
    const int n = (int)(range.y - range.x);
    const int rounds = (n + FGS_BLOCK - 1) / FGS_BLOCK;

    __shared__ float2 s_xy[FGS_BLOCK];
    __shared__ float4 s_co[FGS_BLOCK];
    __shared__ float s_rgb[3][FGS_BLOCK];
    __shared__ uint32_t s_id[FGS_BLOCK];
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
            s_id[tid] = g;
            s_xy[tid] = a.means2D[g];
            s_co[tid] = a.conic_opacity[g];
            s_rgb[0][tid] = a.colors[3 * g + 0];
            s_rgb[1][tid] = a.colors[3 * g + 1];
            s_rgb[2][tid] = a.colors[3 * g + 2];
        }
        s_w[tid] = 0u;
        s_kv[tid] = 0.f;
        s_kc[tid] = 0u;
        __syncthreads();
        const int cnt = todo < FGS_BLOCK ? todo : FGS_BLOCK;
        for (int j = 0; !done && j < cnt; ++j) {
            contributor++;
            const float2 xy = s_xy[j];
            const float dx = xy.x - pfx, dy = xy.y - pfy;
            const float4 co = s_co[j];
            const float power = -0.5f * (co.x * dx * dx + co.z * dy * dy) - co.y * dx * dy;
            if (power > 0.0f) continue;
            const float alpha = fminf(0.99f, co.w * __expf(power));
            if (alpha < 1.0f / 255.0f) continue;
            const float test_T = T * (1.0f - alpha);
            if (test_T < 0.0001f) {
                done = true;
                continue;
            }
            const float w = alpha * T;
            C0 += s_rgb[0][j] * w;
            C1 += s_rgb[1][j] * w;
            C2 += s_rgb[2][j] * w;
            if (w > weight_max) {
                weight_max = w;
                max_id = (int)s_id[j];
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
            if (s_w[tid] != 0u) atomicMax((uint32_t*)&point_weight[s_id[tid]], s_w[tid]);
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
    const int pix = a.W * py + px;
    const int HW = a.W * a.H;

    const uint2 range = a.ranges[tile];
    const int n = (int)(range.y - range.x);
    const int rounds = (n + FGS_BLOCK - 1) / FGS_BLOCK;

    __shared__ float2 s_xy[FGS_BLOCK];
    __shared__ float4 s_co[FGS_BLOCK];
    __shared__ float s_rgb[3][FGS_BLOCK];
    __shared__ uint32_t s_id[FGS_BLOCK];
    __shared__ float s_g[FGS_BLOCK / FGS_WAVE][kGradTerms][FGS_BLOCK];

    const float T_final = inside ? final_Ts[pix] : 0.f;
    float T = T_final;
    const uint32_t last_contributor = inside ? n_contrib[pix] : 0u;
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
    uint32_t contributor = (uint32_t)n;
    bool done = !inside;
    int todo = n;

    for (int r = 0; r < rounds; ++r, todo -= FGS_BLOCK) {
        __syncthreads();
        const int k = r * FGS_BLOCK + tid;
        if (k < n) {
            const uint32_t g = (uint32_t)a.pairs[range.y - k - 1];
            s_id[tid] = g;
            s_xy[tid] = a.means2D[g];
            s_co[tid] = a.conic_opacity[g];
            s_rgb[0][tid] = a.colors[3 * g + 0];
            s_rgb[1][tid] = a.colors[3 * g + 1];
            s_rgb[2][tid] = a.colors[3 * g + 2];
        }
#pragma unroll
        for (int t = 0; t < kGradTerms; ++t) s_g[wave][t][tid] = 0.f;
        __syncthreads();
        const int cnt = todo < FGS_BLOCK ? todo : FGS_BLOCK;
        for (int j = 0; j < cnt; ++j) {
            // wave-uniform early exit: every lane of this wave is past its range
            if (__all(done)) break;
            float v[8] = { 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f };
            float vop = 0.f;
            bool act = false;
            if (!done) {
                contributor--;
                if (contributor < last_contributor) {
                    const float2 xy = s_xy[j];
                    const float dx = xy.x - pfx, dy = xy.y - pfy;
                    const float4 co = s_co[j];
                    const float power = -0.5f * (co.x * dx * dx + co.z * dy * dy) - co.y * dx * dy;
                    if (power <= 0.0f) {
                        const float G = __expf(power);
                        const float alpha_raw = co.w * G;
                        const float alpha = fminf(0.99f, alpha_raw);
                        if (alpha >= 1.0f / 255.0f) {
                            act = true;
                            T = T / (1.f - alpha);
                            const float dch = alpha * T;
                            const float c0 = s_rgb[0][j], c1 = s_rgb[1][j], c2 = s_rgb[2][j];
                            acc0 = last_alpha * lc0 + (1.f - last_alpha) * acc0;
                            acc1 = last_alpha * lc1 + (1.f - last_alpha) * acc1;
                            acc2 = last_alpha * lc2 + (1.f - last_alpha) * acc2;
                            lc0 = c0; lc1 = c1; lc2 = c2;
                            float dL_dalpha = (c0 - acc0) * dpix0 + (c1 - acc1) * dpix1 + (c2 - acc2) * dpix2;
                            v[0] = dch * dpix0;
                            v[1] = dch * dpix1;
                            v[2] = dch * dpix2;
                            dL_dalpha *= T;
                            last_alpha = alpha;
                            dL_dalpha += (-T_final / (1.f - alpha)) * bg_dot;
                            const float slope = (exact && alpha_raw > 0.99f) ? 0.f : 1.f;
                            const float dL_dG = co.w * dL_dalpha * slope;
                            const float gdx = G * dx, gdy = G * dy;
                            const float dG_ddelx = -gdx * co.x - gdy * co.y;
                            const float dG_ddely = -gdy * co.z - gdx * co.y;
                            v[3] = dL_dG * dG_ddelx * ddelx_dx;
                            v[4] = dL_dG * dG_ddely * ddely_dy;
                            v[5] = -0.5f * gdx * dx * dL_dG;
                            v[6] = -0.5f * gdx * dy * dL_dG;
                            v[7] = -0.5f * gdy * dy * dL_dG;
                            vop = G * dL_dalpha * slope;
                        }
                    }
                }
                if (contributor == 0) done = true;
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
            const uint32_t g = s_id[tid];
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
