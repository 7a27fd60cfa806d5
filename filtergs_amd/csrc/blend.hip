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
// atomicMax per (tile, Gaussian): exact and deterministic) and the strip mask
// of the pair -- the strips where it actually blended.  The pairs that
// blended anywhere are appended to the tile's visible list (id + mask).
// Backward: walks the visible list back to front; the wave lists come from
// the masks, so every listed slot blends into at least one pixel of the
// strip.  The nine per-pixel gradient terms of each Gaussian are reduced
// across the wave (lane swaps across halves and rows, then DPP inside rows),
// added over the waves with LDS float atomics, and published per (tile,
// Gaussian): as partial sums gathered per Gaussian afterwards for Gaussians
// of at most kGatherSolo tiles, with global float atomics for larger ones.
#include "fgs_device.h"
#include "fgs_internal.h"

namespace fgs {

constexpr float kLog2e = 1.4426950408889634f;

// The Gaussian (element) of sorted kv pair k.
__device__ __forceinline__ uint32_t pair_id(const BlendArgs& a, uint32_t k)
{
    return a.pair_ids[(size_t)k << a.pair_shift];
}

// The Gaussian index of element g (a batch's elements carry it in gid).
__device__ __forceinline__ uint32_t gauss_id(const BlendArgs& a, uint32_t g)
{
    return a.gid ? a.gid[g] : g;
}
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
// the slots whose bit w is set (as slot << SHIFT), in increasing order, and
// the count for the calling wave is returned (wave-uniform, in an SGPR).
template <int BATCH, int SHIFT, typename L>
__device__ __forceinline__ int build_lists(uint32_t bits, L (*s_list)[BATCH],
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
            s_list[w][off + rank[w]] = (L)(tid << SHIFT);
        }
    }
    int total = 0;
#pragma unroll
    for (int sw = 0; sw < kStrips; ++sw) total += (int)s_wcnt[sw][wave];
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(total);
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

// Maximum of two non-negative floats on their bit patterns (the order of
// non-negative IEEE floats is that of their bits): one integer max, no NaN
// quieting.
__device__ __forceinline__ float max_nonneg(float a, float b)
{
    return __uint_as_float(max(__float_as_uint(a), __float_as_uint(b)));
}

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true);
}

// Wave maximum of a non-negative value (every lane receives it).
__device__ __forceinline__ float wave_maxf(float xf)
{
    uint32_t x = __float_as_uint(xf);
    x = max(x, dpp_u<0x128>(x));
    x = max(x, dpp_u<0x124>(x));
    x = max(x, dpp_u<0x122>(x));
    x = max(x, dpp_u<0x121>(x));
    {
        const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
        x = max((uint32_t)r[0], (uint32_t)r[1]);
    }
    {
        const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
        x = max((uint32_t)r[0], (uint32_t)r[1]);
    }
    return __uint_as_float(x);
}

// Wave reductions into lane 63 only (the other lanes hold partial results):
// a prefix inside each 16-lane row by DPP row shifts (lanes shifted in from
// outside the row read 0, the identity of both reductions here), then the
// row broadcasts of lane 15 into rows 1 and 3 and of lane 31 into rows 2
// and 3.  Six DPP operations, against ten for the every-lane forms above
// (whose two lane swaps each need a copy).
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_rows_u(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_rows(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

// Maximum of a non-negative value over the wave, in lane 63.
__device__ __forceinline__ float wave_maxf_last(float xf)
{
    uint32_t x = __float_as_uint(xf);
    x = max(x, dpp_u<kDppRowShr1>(x));
    x = max(x, dpp_u<kDppRowShr2>(x));
    x = max(x, dpp_u<kDppRowShr4>(x));
    x = max(x, dpp_u<kDppRowShr8>(x));
    x = max(x, dpp_rows_u<kDppRowBcast15, 0xA>(x));
    x = max(x, dpp_rows_u<kDppRowBcast31, 0xC>(x));
    return __uint_as_float(x);
}

// Sum over the wave, in lane 63.
__device__ __forceinline__ float wave_total_last(float x)
{
    x += dpp<kDppRowShr1>(x);
    x += dpp<kDppRowShr2>(x);
    x += dpp<kDppRowShr4>(x);
    x += dpp<kDppRowShr8>(x);
    x += dpp_rows<kDppRowBcast15, 0xA>(x);
    x += dpp_rows<kDppRowBcast31, 0xC>(x);
    return x;
}

// Exponent of the Gaussian in log2 units at offset (dx, dy) from its mean,
// from the staged terms (a', b') = g0.zw and c' = g1.x:
// dy (c' dy + b' dx) + a' dx^2, two fused steps per pixel row once the
// column terms b' dx and a' dx^2 are formed (the one-wave kernels form them
// once per Gaussian and lane: exponent2_col / exponent2_row give the same
// bits).  Shared by the forward and the backward so that both take the same
// alpha decisions.
struct ExpCol {
    float bdx;    // b' dx
    float adx2;   // a' dx^2
};
__device__ __forceinline__ ExpCol exponent2_col(float4 g0, float dx)
{
#pragma clang fp contract(off)
    return ExpCol{ g0.w * dx, (g0.z * dx) * dx };
}
__device__ __forceinline__ float exponent2_row(ExpCol c, float2 g1, float dy)
{
    return fmaf(dy, fmaf(g1.x, dy, c.bdx), c.adx2);
}
__device__ __forceinline__ float exponent2(float4 g0, float2 g1, float dx, float dy)
{
    return exponent2_row(exponent2_col(g0, dx), g1, dy);
}

// ---------------------------------------------------------------------------
// Forward
// ---------------------------------------------------------------------------
// One staged Gaussian of the forward, 64 bytes.  The per-wave lists hold slot
// byte offsets, so the loop reads a slot at fixed offsets from one address.
struct __align__(16) FwdSlot {
    float4 g0;         // x, y, a', b'  (power in log2 units)
    float4 col;        // r, g, b, id bits
    float2 g1;         // c', opacity
    uint32_t w[kStrips];   // per strip: max T*alpha (float bits)
    float kv;          // sum of T*alpha (statistics)
    uint32_t kc;       // contributions (statistics)
};
static_assert(sizeof(FwdSlot) == 64, "forward slot is 64 bytes");
constexpr int kFwdSlotShift = 6;

// KPC histogram bin edges (rasterizer_impl.cu:60-72): bin b holds sums in
// (edge[b-1], edge[b]]; sums above the last edge are not counted.
__constant__ float kKpcEdges[kKpcBins] = { 0.01f, 0.05f, 0.10f, 0.15f, 0.20f, 0.25f,
                                           0.30f, 0.35f, 0.40f, 0.45f, 0.50f };

// Forward of one tile (one workgroup of FGS_BLOCK lanes).  STATS: also the
// tile's share of RenderStats (the host loops of rasterizer_impl.cu:479-556):
// per kv pair of the tile its T*alpha sum and contribution count, binned and
// summed without leaving the chip.
template <int BATCH, bool STATS>
__device__ __forceinline__ void blend_fwd_tile(const BlendArgs& a, int tile, float* __restrict__ final_T,
                                               uint32_t* __restrict__ n_contrib,
                                               uint32_t* __restrict__ vis_last,
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
    float pfx = (float)px, pfy = (float)py;
    // keep the pixel position in registers (the compiler would otherwise
    // convert it again in every iteration)
    asm volatile("" : "+v"(pfx), "+v"(pfy));
    const float tx0 = (float)(tx * FGS_TILE), ty0 = (float)(ty * FGS_TILE);
    bool done = !inside;

    const uint2 range = a.ranges[tile];
    const int n = (int)(range.y - range.x);
    const int rounds = (n + BATCH - 1) / BATCH;

    __shared__ FwdSlot s_slot[BATCH];
    __shared__ uint32_t s_list[kStrips][BATCH];
    __shared__ uint32_t s_wcnt[kStrips][kStrips];
    __shared__ uint32_t s_vcnt[kStrips];   // visible slots per wave in this round
    __shared__ uint32_t s_vidx[BATCH];     // visible-list index of each visible slot

    float T = 1.0f;
    float C0 = 0.f, C1 = 0.f, C2 = 0.f;
    uint32_t last = 0;
    float wmax = 0.f;
    int max_id = -1;
    // The tile's visible list (Gaussians that blend into at least one pixel)
    // is appended round by round; the pixel remembers the slot of its last
    // blend in the round and turns it into a list index after the round.
    uint32_t vis_n = 0;        // entries appended so far (block-uniform)
    uint32_t vlast = 0;        // 1 + list index of the pixel's last blend
    int last_slot = -1;
    // statistics: per-lane sum of the pair sums (reduced in a fixed order at
    // the end), used / unused counts and the histogram in LDS
    double st_sum = 0.0;
    int r_end = rounds;        // rounds staged (the tile stops early when saturated)
    // walk choice of the wave: dense lists (many lanes blend per entry) use
    // the wave-uniform walk with a DPP maximum, sparse ones the divergent
    // walk with per-lane LDS maxima; decided from the previous round
    bool dense_walk = a.walk != 2;
    __shared__ uint32_t s_st[2 + kKpcBins];   // used in tile, unused, bins
    if (STATS) {
        if (tid < 2 + kKpcBins) s_st[tid] = 0u;
    }

    for (int r = 0; r < rounds; ++r) {
        const int num_done = __syncthreads_count(done);
        if (last_slot >= 0) {
            vlast = s_vidx[last_slot] + 1u;
            last_slot = -1;
        }
        if (num_done == FGS_BLOCK) {
            r_end = r;
            break;
        }
        const int q = tid;
        const int k = r * BATCH + q;
        uint32_t bits = 0u;
        uint32_t stage_g = 0u;   // the element this lane staged
        if (q < BATCH) {
            if (k < n) {
                const uint32_t g = pair_id(a, range.x + k);
                stage_g = g;
                const float2 m = a.means2D[g];
                const float4 co = a.conic_opacity[g];
                s_slot[q].g0 = make_float4(m.x, m.y, -0.5f * kLog2e * co.x, -kLog2e * co.y);
                s_slot[q].g1 = make_float2(-0.5f * kLog2e * co.z, co.w);
                s_slot[q].col = make_float4(a.colors[3 * g + 0], a.colors[3 * g + 1], a.colors[3 * g + 2],
                                            __uint_as_float(gauss_id(a, g)));
                bits = strip_bits_ellipse(m, co, tx0, ty0);
            }
#pragma unroll
            for (int w = 0; w < kStrips; ++w) s_slot[q].w[w] = 0u;
            s_slot[q].kv = 0.f;
            s_slot[q].kc = 0u;
        }
        const int cnt = build_lists<BATCH, kFwdSlotShift>(bits, s_list, s_wcnt);
        char* const slots = reinterpret_cast<char*>(s_slot);
        // the next entry of the list is read one iteration ahead
        uint32_t next = s_list[wave][0];
        if constexpr (STATS) {
            // The whole wave walks the list (until all of its pixels are
            // done), so the per-slot statistics are wave sums: one LDS atomic
            // per wave instead of one per blending lane.
            for (int i = 0; i < cnt; ++i) {
                if (__all(done)) break;
                FwdSlot* const sp = reinterpret_cast<FwdSlot*>(slots + next);
                next = s_list[wave][min(i + 1, BATCH - 1)];
                const float4 g0 = sp->g0;
                const float2 g1 = sp->g1;
                const float4 col = sp->col;
                const float dx = g0.x - pfx, dy = g0.y - pfy;
                const float p2 = exponent2(g0, g1, dx, dy);
                const float alpha = fminf(0.99f, g1.y * __builtin_amdgcn_exp2f(p2));
                const float test_T = T * (1.0f - alpha);
                const bool pass = !done && !(p2 > 0.0f) && !(alpha < 1.0f / 255.0f);
                const bool stop = pass && test_T < 0.0001f;
                const bool blend = pass && !stop;
                done = done || stop;
                const float w = blend ? alpha * T : 0.f;
                const uint64_t bl = __ballot(blend);
                const bool many = __popcll(bl) > 12;   // wave-uniform
                if (bl) {
                    const float ws = wave_total(w);
                    const float wm = many ? wave_maxf(w) : 0.f;
                    if (lane_id() == 0) {
                        atomicAdd(&sp->kv, ws);
                        atomicAdd(&sp->kc, (uint32_t)__popcll(bl));
                        if (many) sp->w[wave] = __float_as_uint(wm);
                    }
                }
                if (blend) {
                    C0 = fmaf(col.x, w, C0);
                    C1 = fmaf(col.y, w, C1);
                    C2 = fmaf(col.z, w, C2);
                    if (w > wmax) {
                        wmax = w;
                        max_id = (int)__float_as_uint(col.w);
                    }
                    if (!many) atomicMax(&sp->w[wave], __float_as_uint(w));
                    T = test_T;
                    last_slot = (int)(sp - s_slot);
                    last = (uint32_t)(r * BATCH + last_slot + 1);
                }
            }
        } else {
            // Dense walk: wave-uniform, the per-strip maximum weight of the
            // slot is a DPP wave maximum stored by one lane (each slot is
            // visited once by each wave) instead of an LDS atomic from every
            // blending lane, which serialises when most lanes blend.  Sparse
            // walk: divergent, per-lane LDS atomics.
            uint32_t act_sum = 0, act_iters = 0;
            if (dense_walk) {
                for (int i = 0; i < cnt; ++i) {
                    if (__all(done)) break;
                    FwdSlot* const sp = reinterpret_cast<FwdSlot*>(slots + next);
                    next = s_list[wave][min(i + 1, BATCH - 1)];
                    const float4 g0 = sp->g0;
                    const float2 g1 = sp->g1;
                    const float4 col = sp->col;
                    const float dx = g0.x - pfx, dy = g0.y - pfy;
                    const float p2 = exponent2(g0, g1, dx, dy);
                    const float alpha = fminf(0.99f, g1.y * __builtin_amdgcn_exp2f(p2));
                    const float test_T = T * (1.0f - alpha);
                    const bool pass = !done && !(p2 > 0.0f) && !(alpha < 1.0f / 255.0f);
                    const bool stop = pass && test_T < 0.0001f;
                    const bool blend = pass && !stop;
                    done = done || stop;
                    const float w = blend ? alpha * T : 0.f;
                    const uint64_t blm = __ballot(blend);
                    act_sum += (uint32_t)__popcll(blm);
                    act_iters += 1u;
                    if (blm) {   // wave-uniform
                        const float wm = wave_maxf(w);
                        if (lane_id() == 0) sp->w[wave] = __float_as_uint(wm);
                        if (kv_sum) {
                            if (blend) {
                                atomicAdd(&sp->kv, w);
                                atomicAdd(&sp->kc, 1u);
                            }
                        }
                    }
                    if (blend) {
                        C0 = fmaf(col.x, w, C0);
                        C1 = fmaf(col.y, w, C1);
                        C2 = fmaf(col.z, w, C2);
                        if (w > wmax) {
                            wmax = w;
                            max_id = (int)__float_as_uint(col.w);
                        }
                        T = test_T;
                        last_slot = (int)(sp - s_slot);
                        last = (uint32_t)(r * BATCH + last_slot + 1);
                    }
                }
            } else {
                for (int i = 0; !done && i < cnt; ++i) {
                    FwdSlot* const sp = reinterpret_cast<FwdSlot*>(slots + next);
                    next = s_list[wave][min(i + 1, BATCH - 1)];
                    const float4 g0 = sp->g0;
                    const float2 g1 = sp->g1;
                    const float4 col = sp->col;
                    const float dx = g0.x - pfx, dy = g0.y - pfy;
                    const float p2 = exponent2(g0, g1, dx, dy);
                    const float alpha = fminf(0.99f, g1.y * __builtin_amdgcn_exp2f(p2));
                    const float test_T = T * (1.0f - alpha);
                    const bool pass = !(p2 > 0.0f) && !(alpha < 1.0f / 255.0f);
                    const bool stop = pass && test_T < 0.0001f;
                    done = done || stop;
                    if (pass && !stop) {
                        const float w = alpha * T;
                        C0 = fmaf(col.x, w, C0);
                        C1 = fmaf(col.y, w, C1);
                        C2 = fmaf(col.z, w, C2);
                        if (w > wmax) {
                            wmax = w;
                            max_id = (int)__float_as_uint(col.w);
                        }
                        atomicMax(&sp->w[wave], __float_as_uint(w));
                        if (kv_sum) {
                            atomicAdd(&sp->kv, w);
                            atomicAdd(&sp->kc, 1u);
                        }
                        T = test_T;
                        last_slot = (int)(sp - s_slot);
                        last = (uint32_t)(r * BATCH + last_slot + 1);
                    }
                }
            }
            if (a.walk == 0 && dense_walk && act_iters > 0u) dense_walk = act_sum > 24u * act_iters;
        }
        __syncthreads();
        uint32_t used = 0u, g = 0u;
        if (q < BATCH && k < n) {
            g = stage_g;
            uint32_t m = 0u;
#pragma unroll
            for (int w = 0; w < kStrips; ++w) {
                const uint32_t sw = s_slot[q].w[w];
                m = sw > m ? sw : m;
                used |= (sw != 0u ? 1u : 0u) << w;
            }
            if (m != 0u) atomicMax((uint32_t*)&point_weight[a.id_base + __float_as_uint(s_slot[q].col.w)], m);
            if (kv_sum) {
                // each kv pair belongs to exactly one tile: plain stores
                kv_sum[range.x + k] = s_slot[q].kv;
                kv_cnt[range.x + k] = s_slot[q].kc;
            }
            if (STATS) {
                const float kv = s_slot[q].kv;
                const uint32_t kc = s_slot[q].kc;
                st_sum += (double)kv;
                if (kc > 0u && kv >= a.stats_thresh) atomicAdd(&s_st[0], 1u);
                if (kc == 0u) {
                    atomicAdd(&s_st[1], 1u);
                } else {
                    int b = 0;
#pragma unroll
                    for (int e = 0; e < kKpcBins; ++e) b += kv > kKpcEdges[e] ? 1 : 0;
                    if (b < kKpcBins) atomicAdd(&s_st[2 + b], 1u);
                }
            }
        }
        // append the round's visible slots in slot order
        const uint64_t vb = __ballot(used != 0u);
        if (lane_id() == 0) s_vcnt[wave] = (uint32_t)__popcll(vb);
        __syncthreads();
        uint32_t off = vis_n;
#pragma unroll
        for (int w = 0; w < kStrips; ++w) {
            const uint32_t c = s_vcnt[w];
            off += w < wave ? c : 0u;
            vis_n += c;
        }
        if (used != 0u) {
            const uint32_t idx = off + __builtin_amdgcn_mbcnt_hi((uint32_t)(vb >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)vb, 0u));
            a.vis_ids[range.x + idx] = g;
            a.kv_strips[range.x + idx] = (uint8_t)used;
            s_vidx[q] = idx;
        }
    }
    __syncthreads();
    if (last_slot >= 0) vlast = s_vidx[last_slot] + 1u;
    if (tid == 0) a.tile_vis[tile] = vis_n;
    if (STATS) {
        // fixed-order reduction of the tile sum (the same bits on every run)
        __shared__ double s_red[FGS_BLOCK];
        s_red[tid] = st_sum;
        __syncthreads();
#pragma unroll
        for (int h = FGS_BLOCK / 2; h > 0; h >>= 1) {
            if (tid < h) s_red[tid] += s_red[tid + h];
            __syncthreads();
        }
        if (tid == 0) {
            const int T = a.gx * a.gy;
            double* st = a.stats;
            const double nk = (double)n;
            // pairs of rounds never staged (all pixels saturated) blended nowhere
            const uint32_t unstaged = (uint32_t)max(0, n - r_end * BATCH);
            st[kStatsHead + tile] = n > 0 ? (double)s_st[0] / nk : 0.0;           // tile_ratio
            st[kStatsHead + T + tile] = (double)n;                                 // tile_total
            st[kStatsHead + 2 * T + tile] = n > 0 ? s_red[0] / nk : 0.0;          // tile_avg
            unsigned long long* cnt = reinterpret_cast<unsigned long long*>(st + kStatsHead + 3 * T);
            uint32_t used = 0;
#pragma unroll
            for (int b = 0; b < kKpcBins; ++b) {
                used += s_st[2 + b];
                if (s_st[2 + b]) atomicAdd(&cnt[2 + b], (unsigned long long)s_st[2 + b]);
            }
            if (used) atomicAdd(&cnt[0], (unsigned long long)used);
            if (s_st[1] + unstaged) atomicAdd(&cnt[1], (unsigned long long)(s_st[1] + unstaged));
        }
    }

    if (inside) {
        const int pix = a.W * py + px;
        const int HW = a.W * a.H;
        final_T[pix] = T;
        n_contrib[pix] = last;
        vis_last[pix] = vlast;
        out_color[pix] = fmaf(T, a.bg[0], C0);
        out_color[HW + pix] = fmaf(T, a.bg[1], C1);
        out_color[2 * HW + pix] = fmaf(T, a.bg[2], C2);
        out_point_id[pix] = max_id;
        out_weight_pixel[pix] = wmax;
    }
}

// The record of one Gaussian a staging lane loads: mean, conic and opacity,
// colour, id.
struct StageRec {
    float2 m;
    float4 co;
    float c0, c1, c2;
    uint32_t g;    // element
    uint32_t id;   // Gaussian index (point ids and weights)
};

__device__ __forceinline__ StageRec load_stage(const BlendArgs& a, uint32_t g)
{
    StageRec r;
    r.g = g;
    r.id = gauss_id(a, g);
    r.m = a.means2D[g];
    r.co = a.conic_opacity[g];
    r.c0 = a.colors[3 * g + 0];
    r.c1 = a.colors[3 * g + 1];
    r.c2 = a.colors[3 * g + 2];
    return r;
}

// Forward of one tile by a single wave (no statistics): lane l owns column
// l % 16 and, in each of the four strips, row l / 16 of the strip (4 pixels
// per lane).  Batches of 64 Gaussians of the tile's range are staged in LDS,
// one per lane, with the strip mask of the conservative alpha >= 1/255 box;
// each staged Gaussian is worked strip by strip (wave-uniform branches on the
// mask and on the strips' saturation).  A pixel sees the same Gaussians in
// the same order with the same operations as in blend_fwd_tile, so the
// images, n_contrib, point ids and weights are the same bits.  point_weight
// takes one DPP wave maximum per (tile, Gaussian) over the four strips.
constexpr int kFwdWaveBatch = FGS_WAVE;
constexpr float kFarRow = 1e30f;
constexpr int kLiveEvery = 8;

// Strips of the wave with a pixel that has not finished (a bit per strip).
__device__ __forceinline__ uint32_t live_strips(const float* pfy)
{
    uint32_t lv = 0u;
#pragma unroll
    for (int k = 0; k < kStrips; ++k)
        lv |= (__builtin_amdgcn_ballot_w64(pfy[k] != kFarRow) != 0u ? 1u : 0u) << k;
    return lv;
}

__device__ __forceinline__ void blend_fwd_tile_wave(const BlendArgs& a, int tile, float* __restrict__ final_T,
                                                    uint32_t* __restrict__ n_contrib,
                                                    uint32_t* __restrict__ vis_last,
                                                    float* __restrict__ out_color,
                                                    int* __restrict__ out_point_id,
                                                    float* __restrict__ out_weight_pixel,
                                                    float* __restrict__ point_weight,
                                                    float* __restrict__ kv_sum, uint32_t* __restrict__ kv_cnt)
{
#pragma clang fp contract(off)
    const int tx = tile % a.gx, ty = tile / a.gx;
    const int lane = lane_id();
    const int px = tx * FGS_TILE + (lane & (FGS_TILE - 1));
    const int py0 = ty * FGS_TILE + (lane / FGS_TILE);
    float pfx = (float)px;
    asm volatile("" : "+v"(pfx));
    const float tx0 = (float)(tx * FGS_TILE), ty0 = (float)(ty * FGS_TILE);
    const uint2 range = a.ranges[tile];
    const int n = __builtin_amdgcn_readfirstlane((int)(range.y - range.x));

    __shared__ FwdSlot s_slot[kFwdWaveBatch];
    __shared__ uint32_t s_vidx[kFwdWaveBatch];

    // A finished pixel (saturated, or outside the image) moves its row to
    // kFarRow: dy is then of order -1e30 and the exponent -inf (or +inf for
    // a conic that is not positive definite), so alpha < 1/255 or power > 0
    // and the pixel neither blends nor stops again.
    float pfy[kStrips], T[kStrips], C0[kStrips], C1[kStrips], C2[kStrips], wmax[kStrips];
    int max_id[kStrips];
    uint32_t last[kStrips], vlast[kStrips];
    uint32_t live = 0u;   // strips with a pixel not yet saturated (wave-uniform)
#pragma unroll
    for (int k = 0; k < kStrips; ++k) {
        const int py = py0 + 4 * k;
        pfy[k] = (px < a.W && py < a.H) ? (float)py : kFarRow;
        asm volatile("" : "+v"(pfy[k]));
        T[k] = 1.0f;
        C0[k] = C1[k] = C2[k] = 0.f;
        wmax[k] = 0.f;
        max_id[k] = -1;
        last[k] = 0u;
        vlast[k] = 0u;
    }
    live = live_strips(pfy);
    uint32_t vis_n = 0;   // visible-list entries so far (wave-uniform)

    // Staging is software-pipelined: while round r is worked, the records of
    // round r + 1 are in flight into registers and the ids of round r + 2
    // into another (the id loads and the record loads they feed are one
    // round apart, so neither waits on the other).
    StageRec cur, nxt;
    uint32_t id_next = 0u;
    if (lane < n) cur = load_stage(a, pair_id(a, range.x + lane));
    if (kFwdWaveBatch + lane < n) id_next = pair_id(a, range.x + kFwdWaveBatch + lane);

    for (int r = 0; r * kFwdWaveBatch < n && live != 0u; ++r) {
        const int k0 = r * kFwdWaveBatch;
        const int cnt = __builtin_amdgcn_readfirstlane(min(kFwdWaveBatch, n - k0));
        uint32_t bits = 0u;
        __syncthreads();
        if (lane < cnt) {
            s_slot[lane].g0 = make_float4(cur.m.x, cur.m.y, -0.5f * kLog2e * cur.co.x, -kLog2e * cur.co.y);
            s_slot[lane].g1 = make_float2(-0.5f * kLog2e * cur.co.z, cur.co.w);
            s_slot[lane].col = make_float4(cur.c0, cur.c1, cur.c2, __uint_as_float(cur.id));
            bits = strip_bits_ellipse(cur.m, cur.co, tx0, ty0);
        }
        if (k0 + kFwdWaveBatch + lane < n) nxt = load_stage(a, id_next);
        if (k0 + 2 * kFwdWaveBatch + lane < n) id_next = pair_id(a, range.x + k0 + 2 * kFwdWaveBatch + lane);
        s_slot[lane].w[0] = 0u;   // point weight of the slot (max T*alpha, float bits)
        s_slot[lane].w[1] = 0u;   // strips it blended into
        s_slot[lane].kv = 0.f;
        s_slot[lane].kc = 0u;
        __syncthreads();
        for (int j = 0; j < cnt; ++j) {
            // strips whose pixels have all finished drop out of the walk,
            // checked every kLiveEvery Gaussians and after each round
            if (j != 0 && (j & (kLiveEvery - 1)) == 0) live = live_strips(pfy);   // wave-uniform
            const uint32_t sb = __builtin_amdgcn_readlane(bits, j) & live;
            if (sb == 0u) continue;   // wave-uniform
            FwdSlot* const sp = &s_slot[j];
            const float4 g0 = sp->g0;
            const float2 g1 = sp->g1;
            const ExpCol ec = exponent2_col(g0, g0.x - pfx);
            float wl = 0.f, wsum = 0.f;
            uint32_t used = 0u, nbl = 0u;
#pragma unroll
            for (int k = 0; k < kStrips; ++k) {
                if (!((sb >> k) & 1u)) continue;   // wave-uniform
                const float dy = g0.y - pfy[k];
                const float p2 = exponent2_row(ec, g1, dy);
                const float alpha = fminf(0.99f, g1.y * __builtin_amdgcn_exp2f(p2));
                const float test_T = T[k] * (1.0f - alpha);
                const bool hit = !(p2 > 0.0f) && !(alpha < 1.0f / 255.0f);
                const bool keep = test_T >= 0.0001f;   // test_T is never NaN: !keep is test_T < 1e-4
                const bool blend = hit && keep;
                // the ballot of blend from those of its compares (the ballot of
                // a compare is its lane mask; of an "and", a materialised 0/1)
                const uint64_t bl = __builtin_amdgcn_ballot_w64(!(p2 > 0.0f)) &
                                    __builtin_amdgcn_ballot_w64(!(alpha < 1.0f / 255.0f)) &
                                    __builtin_amdgcn_ballot_w64(keep);
                pfy[k] = (hit && !keep) ? kFarRow : pfy[k];   // saturated: finished
                used |= (bl != 0u ? 1u : 0u) << k;
                nbl += (uint32_t)__popcll(bl);
                if (blend) {   // skipped by the whole wave when no lane blends
                    const float4 col = sp->col;
                    const float w = alpha * T[k];
                    C0[k] = fmaf(col.x, w, C0[k]);
                    C1[k] = fmaf(col.y, w, C1[k]);
                    C2[k] = fmaf(col.z, w, C2[k]);
                    if (w > wmax[k]) {
                        wmax[k] = w;
                        max_id[k] = (int)__float_as_uint(col.w);
                    }
                    wl = max_nonneg(wl, w);
                    wsum += w;
                    T[k] = test_T;
                    last[k] = (uint32_t)(k0 + j + 1);
                }
            }
            if (used != 0u) {   // wave-uniform
                const float pw = wave_maxf_last(wl);   // in lane 63
                if (kv_sum) {
                    const float ws = wave_total(wsum);
                    if (lane == 0) {
                        sp->kv = ws;
                        sp->kc = nbl;
                    }
                }
                if (lane == FGS_WAVE - 1) {
                    sp->w[0] = __float_as_uint(pw);
                    sp->w[1] = used;
                }
            }
        }
        live = live_strips(pfy);
        __syncthreads();
        uint32_t used = 0u, g = 0u;
        if (lane < cnt) {
            g = cur.g;
            used = s_slot[lane].w[1];
            const uint32_t m = s_slot[lane].w[0];
            if (m != 0u) atomicMax((uint32_t*)&point_weight[a.id_base + cur.id], m);
            if (kv_sum) {
                kv_sum[range.x + k0 + lane] = s_slot[lane].kv;
                kv_cnt[range.x + k0 + lane] = s_slot[lane].kc;
            }
        }
        // append the round's visible slots in slot order
        const uint64_t vb = __ballot(used != 0u);
        if (used != 0u) {
            const uint32_t idx = vis_n + __builtin_amdgcn_mbcnt_hi((uint32_t)(vb >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)vb, 0u));
            a.vis_ids[range.x + idx] = g;
            a.kv_strips[range.x + idx] = (uint8_t)used;
            s_vidx[lane] = idx;
        }
        vis_n += (uint32_t)__popcll(vb);
        __syncthreads();
        // a pixel that blended in this round last did so at slot last - 1 - k0
#pragma unroll
        for (int k = 0; k < kStrips; ++k) {
            if (last[k] > (uint32_t)k0) vlast[k] = s_vidx[last[k] - 1u - (uint32_t)k0] + 1u;
        }
        cur = nxt;
    }
    if (lane == 0) a.tile_vis[tile] = vis_n;
    const int HW = a.W * a.H;
#pragma unroll
    for (int k = 0; k < kStrips; ++k) {
        const int py = py0 + 4 * k;
        if (px < a.W && py < a.H) {
            const int pix = a.W * py + px;
            final_T[pix] = T[k];
            n_contrib[pix] = last[k];
            vis_last[pix] = vlast[k];
            out_color[pix] = fmaf(T[k], a.bg[0], C0[k]);
            out_color[HW + pix] = fmaf(T[k], a.bg[1], C1[k]);
            out_color[2 * HW + pix] = fmaf(T[k], a.bg[2], C2[k]);
            out_point_id[pix] = max_id[k];
            out_weight_pixel[pix] = wmax[k];
        }
    }
}

__global__ __launch_bounds__(FGS_WAVE) void k_blend_fwd_wave(BlendArgs a, float* final_T, uint32_t* n_contrib,
                                                             uint32_t* vis_last, float* out_color,
                                                             int* out_point_id, float* out_weight_pixel,
                                                             float* point_weight, float* kv_sum,
                                                             uint32_t* kv_cnt)
{
    blend_fwd_tile_wave(a, blockIdx.x, final_T, n_contrib, vis_last, out_color, out_point_id, out_weight_pixel,
                        point_weight, kv_sum, kv_cnt);
}

template <int BATCH, bool STATS>
__global__ __launch_bounds__(FGS_BLOCK) void k_blend_fwd(BlendArgs a, float* final_T, uint32_t* n_contrib,
                                                         uint32_t* vis_last, float* out_color,
                                                         int* out_point_id, float* out_weight_pixel,
                                                         float* point_weight, float* kv_sum,
                                                         uint32_t* kv_cnt)
{
    blend_fwd_tile<BATCH, STATS>(a, blockIdx.x, final_T, n_contrib, vis_last, out_color, out_point_id,
                                 out_weight_pixel, point_weight, kv_sum, kv_cnt);
}

// Workgroup -> batch tile.  Workgroups go to the eight XCDs round robin
// (workgroup b to XCD b % 8), each with its own L2.  With chunk C > 0, XCD x
// takes runs of C consecutive tiles (its k-th workgroup works tile
// ((k / C) * 8 + x) * C + k % C), so neighbouring tiles, which stage many of
// the same Gaussians, share an L2; the runs interleave over the XCDs, which
// keeps their loads even.  Tiles past the last whole 8 C group map to
// themselves.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n, uint32_t C)
{
    if (C == 0u) return b;
    const uint32_t n8 = n / (8u * C) * (8u * C);
    if (b >= n8) return b;
    const uint32_t x = b % 8u, k = b / 8u;
    return ((k / C) * 8u + x) * C + k % C;
}

// View of a batch that owns workgroup b (views sorted by tile_base).
__device__ __forceinline__ int find_view(const BatchView* views, int V, uint32_t b)
{
    int v = 0;
    while (v + 1 < V && views[v + 1].tile_base <= b) ++v;
    return v;
}

// ---- persistent launches over a cost-ordered tile queue -------------------
//
// Queue items are runs of C consecutive tiles of the batch numbering (the
// tiles of one run are neighbours in one view and stage many of the same
// Gaussians).  The runs, heaviest first, are dealt round robin to eight
// queues: queue q holds the runs at positions q, q + 8, q + 16, ... of the
// order, and its k-th tile is tile k % C of the run at position
// 8 (k / C) + q.  Workgroup b starts on queue b % 8, the queue of its XCD
// (workgroups go to the XCDs round robin), so each XCD works its runs
// heaviest first with their tiles sharing its L2; when its queue is drained
// it claims from the other seven in turn.
__device__ __forceinline__ uint32_t queue_len(uint32_t runs, uint32_t C, uint32_t q)
{
    return (runs / 8u + (q < runs % 8u ? 1u : 0u)) * C;
}

// Lane 0 claims the next item of the queue whose counter is h; the value is
// lane 0's (the caller reads it with readfirstlane).  (Claiming four items
// per atomic measured slower: 20.9 against 20.6 ms for the 16-view step.)
__device__ __forceinline__ uint32_t queue_claim(uint32_t* h)
{
    uint32_t k = 0u;
    if (lane_id() == 0) k = atomicAdd(h, 1u);
    return k;
}

// Runs work(tile) for every tile the calling wave claims, until every queue
// is drained.  The claim of the next item is issued before the current tile
// is worked, so its latency overlaps the tile's first loads.  All control
// flow here is wave-uniform.
template <typename Work>
__device__ __forceinline__ void drain_queue(const BlendArgs& c, uint32_t n, Work&& work)
{
    const uint32_t C = c.queue_chunk;
    const uint32_t runs = (n + C - 1u) / C;
    const uint32_t q0 = blockIdx.x & 7u;
    uint32_t qs = 0u, q = q0;
    uint32_t len = queue_len(runs, C, q);
    uint32_t k = queue_claim(c.queue_head + q);
    for (;;) {
        k = __builtin_amdgcn_readfirstlane(k);
        if (k >= len) {
            if (++qs == 8u) return;
            q = (q0 + qs) & 7u;
            len = queue_len(runs, C, q);
            k = queue_claim(c.queue_head + q);
            continue;
        }
        const uint32_t t = (uint32_t)c.queue_order[(k / C) * 8u + q] * C + k % C;
        const uint32_t k_next = queue_claim(c.queue_head + q);
        if (t < n) work(t);   // the last run may be short
        k = k_next;
    }
}

// One lane per run: its cost (the sum of its tiles' visible-list lengths
// when tile_vis is given, else of their kv counts) in quarter-octave classes,
// 4 * bit_length(c) + the two bits below the leading one (0 for c = 0, at
// most 131).  by_cost: key (255 - class) << 32 | run into keys, so a stable
// sort on the class puts the heaviest runs first; else the run itself into
// order (launch order).  The first eight lanes clear the claim counters.
__global__ __launch_bounds__(256) void k_tile_cost_keys(const uint2* __restrict__ ranges,
                                                        const uint32_t* __restrict__ tile_vis, uint32_t T,
                                                        uint32_t C, int by_cost, uint64_t* __restrict__ keys,
                                                        uint64_t* __restrict__ order, uint32_t* __restrict__ head)
{
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    const uint32_t runs = (T + C - 1u) / C;
    if (r < 8u) head[r] = 0u;
    if (r >= runs) return;
    if (!by_cost) {
        order[r] = r;
        return;
    }
    uint64_t c64 = 0u;
    const uint32_t t1 = min(T, (r + 1u) * C);
    for (uint32_t t = r * C; t < t1; ++t) {
        if (tile_vis) {
            c64 += tile_vis[t];
        } else {
            const uint2 rg = ranges[t];
            c64 += rg.y - rg.x;
        }
    }
    const uint32_t c = c64 > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)c64;
    uint32_t cls = 0u;
    if (c != 0u) {
        const uint32_t bl = 32u - (uint32_t)__clz(c);
        const uint32_t mant = bl >= 3u ? (c >> (bl - 3u)) & 3u : (c << (3u - bl)) & 3u;
        cls = 4u * bl + mant;
    }
    keys[r] = ((uint64_t)(255u - cls) << 32) | r;
}

hipError_t build_tile_queue(const uint2* ranges, const uint32_t* tile_vis, uint32_t T, uint32_t C, int by_cost,
                            uint64_t* keysA, uint64_t* keysB, uint32_t* hist, uint32_t* partials,
                            uint32_t* head, hipStream_t s)
{
    if (C == 0u) return hipErrorInvalidValue;
    const uint32_t runs = (T + C - 1u) / C;
    const uint32_t blocks = (runs + 8u + 255u) / 256u;
    hipLaunchKernelGGL(k_tile_cost_keys, dim3(blocks), dim3(256), 0, s, ranges, tile_vis, T, C, by_cost, keysA,
                       keysB, head);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || runs == 0 || !by_cost) return e;
    return radix_sort_u64(keysA, keysB, runs, 32, 1, hist, partials, s);
}

// Workgroups of a persistent launch: as many one-wave workgroups as the
// device holds at once (occupancy x CUs, a multiple of 8), at most one per
// tile rounded up to 8.  Cached per device and kernel.
static uint32_t persistent_grid(const void* fn, int slot, uint32_t T)
{
    constexpr int kDevs = 64;
    static uint32_t cache[3][kDevs];   // forward, backward, exact-mode backward
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kDevs) dev = 0;
    uint32_t g = cache[slot][dev];
    if (g == 0u) {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, FGS_WAVE, 0) != hipSuccess || per_cu < 1)
            per_cu = 8;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
        g = (uint32_t)per_cu * (uint32_t)cus / 8u * 8u;
        if (g < 8u) g = 8u;
        cache[slot][dev] = g;
    }
    const uint32_t cap = (T + 7u) / 8u * 8u;
    return g < cap ? g : cap;
}

// The view-dependent fields of a BlendArgs; per-pixel pointers are offset by
// the caller with pix_base.
__device__ __forceinline__ BlendArgs bind_blend(const BlendArgs& common, const BatchView& vw, int v)
{
    BlendArgs a = common;
    a.id_base = (uint32_t)v * common.id_stride;
    a.ranges = common.ranges + vw.tile_base;
    a.tile_vis = common.tile_vis + vw.tile_base;
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
                                                               uint32_t* n_contrib, uint32_t* vis_last,
                                                               float* out_color, int* out_point_id,
                                                               float* out_weight_pixel,
                                                               float* point_weight)
{
    const int v = find_view(views, V, blockIdx.x);
    const BatchView& vw = views[v];
    const BlendArgs a = bind_blend(common, vw, v);
    const size_t pb = vw.pix_base;
    blend_fwd_tile<BATCH, false>(a, (int)(blockIdx.x - vw.tile_base), final_T + pb, n_contrib + pb,
                          vis_last + pb, out_color + 3 * pb, out_point_id + pb, out_weight_pixel + pb,
                          point_weight, nullptr, nullptr);
}

__global__ __launch_bounds__(FGS_WAVE) void k_blend_fwd_batch_wave(BlendArgs common, const BatchView* views,
                                                                   int V, float* final_T, uint32_t* n_contrib,
                                                                   uint32_t* vis_last, float* out_color,
                                                                   int* out_point_id, float* out_weight_pixel,
                                                                   float* point_weight)
{
    const uint32_t b = xcd_tile(blockIdx.x, gridDim.x, (uint32_t)common.xcd_chunk);
    const int v = find_view(views, V, b);
    const BatchView& vw = views[v];
    const BlendArgs a = bind_blend(common, vw, v);
    const size_t pb = vw.pix_base;
    blend_fwd_tile_wave(a, (int)(b - vw.tile_base), final_T + pb, n_contrib + pb, vis_last + pb,
                        out_color + 3 * pb, out_point_id + pb, out_weight_pixel + pb, point_weight, nullptr,
                        nullptr);
}

// The same tiles as k_blend_fwd_batch_wave, claimed from the cost-ordered
// queue by a persistent grid.  Each tile is worked exactly as there, so the
// outputs are the same bits.
__global__ __launch_bounds__(FGS_WAVE) __attribute__((amdgpu_waves_per_eu(5))) void k_blend_fwd_batch_queue(BlendArgs common, const BatchView* views,
                                                                    int V, uint32_t T_total, float* final_T,
                                                                    uint32_t* n_contrib, uint32_t* vis_last,
                                                                    float* out_color, int* out_point_id,
                                                                    float* out_weight_pixel,
                                                                    float* point_weight)
{
    drain_queue(common, T_total, [&](uint32_t b) {
        const int v = find_view(views, V, b);
        const BatchView& vw = views[v];
        const BlendArgs a = bind_blend(common, vw, v);
        const size_t pb = vw.pix_base;
        blend_fwd_tile_wave(a, (int)(b - vw.tile_base), final_T + pb, n_contrib + pb, vis_last + pb,
                            out_color + 3 * pb, out_point_id + pb, out_weight_pixel + pb, point_weight,
                            nullptr, nullptr);
    });
}

hipError_t launch_blend_fwd_batch(const BlendArgs& common, const BatchView* views, int V, int T_total,
                                  float* final_T, uint32_t* n_contrib, uint32_t* vis_last,
                                  float* out_color, int* out_point_id, float* out_weight_pixel,
                                  float* point_weight, hipStream_t s)
{
    if (T_total == 0) return hipSuccess;
    prof_kernel(false, true, s);
    if (common.fwd_layout == 1 && common.queue_order) {
        const uint32_t grid =
            persistent_grid(reinterpret_cast<const void*>(&k_blend_fwd_batch_queue), 0, (uint32_t)T_total);
        hipLaunchKernelGGL(k_blend_fwd_batch_queue, dim3(grid), dim3(FGS_WAVE), 0, s, common, views, V,
                           (uint32_t)T_total, final_T, n_contrib, vis_last, out_color, out_point_id,
                           out_weight_pixel, point_weight);
    } else if (common.fwd_layout == 1)
        hipLaunchKernelGGL(k_blend_fwd_batch_wave, dim3(T_total), dim3(FGS_WAVE), 0, s, common, views, V,
                           final_T, n_contrib, vis_last, out_color, out_point_id, out_weight_pixel,
                           point_weight);
    else if (common.batch == 128)
        hipLaunchKernelGGL(k_blend_fwd_batch<128>, dim3(T_total), dim3(FGS_BLOCK), 0, s, common, views, V,
                           final_T, n_contrib, vis_last, out_color, out_point_id, out_weight_pixel,
                           point_weight);
    else
        hipLaunchKernelGGL(k_blend_fwd_batch<256>, dim3(T_total), dim3(FGS_BLOCK), 0, s, common, views, V,
                           final_T, n_contrib, vis_last, out_color, out_point_id, out_weight_pixel,
                           point_weight);
    prof_kernel(false, false, s);
    return hipGetLastError();
}

hipError_t launch_blend_fwd(const BlendArgs& a, float* final_T, uint32_t* n_contrib, uint32_t* vis_last,
                            float* out_color, int* out_point_id, float* out_weight_pixel,
                            float* point_weight, float* kv_sum, uint32_t* kv_cnt, hipStream_t s)
{
    const int T = a.gx * a.gy;
    if (T == 0) return hipSuccess;
#define FGS_FWD(BATCH, STATS)                                                                            \
    hipLaunchKernelGGL((k_blend_fwd<BATCH, STATS>), dim3(T), dim3(FGS_BLOCK), 0, s, a, final_T, n_contrib, \
                       vis_last, out_color, out_point_id, out_weight_pixel, point_weight, kv_sum, kv_cnt)
    prof_kernel(false, true, s);
    if (a.fwd_layout == 1 && !a.stats) {
        hipLaunchKernelGGL(k_blend_fwd_wave, dim3(T), dim3(FGS_WAVE), 0, s, a, final_T, n_contrib, vis_last,
                           out_color, out_point_id, out_weight_pixel, point_weight, kv_sum, kv_cnt);
    } else if (a.stats) {
        if (a.batch == 128) FGS_FWD(128, true);
        else FGS_FWD(256, true);
    } else {
        if (a.batch == 128) FGS_FWD(128, false);
        else FGS_FWD(256, false);
    }
#undef FGS_FWD
    prof_kernel(false, false, s);
    return hipGetLastError();
}

// The packed head of RenderStats from the per-tile values and counters the
// blend left in the buffer; bar_g is summed over the tiles in a fixed order.
__global__ __launch_bounds__(256) void k_stats_finalize(double* st, int T, int P, int64_t K)
{
    __shared__ double s_red[256];
    const int tid = threadIdx.x;
    double acc = 0.0;
    for (int t = tid; t < T; t += 256) acc += st[kStatsHead + 2 * (size_t)T + t];
    s_red[tid] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) s_red[tid] += s_red[tid + h];
        __syncthreads();
    }
    const unsigned long long* cnt =
        reinterpret_cast<const unsigned long long*>(st + kStatsHead + 3 * (size_t)T);
    if (tid < 9 + 2 * kStatsMaxBins) {
        double v = 0.0;
        if (tid == 0) v = T > 0 ? s_red[0] / (double)T : 0.0;
        else if (tid == 1) v = P > 0 ? (double)K / (double)P : 0.0;
        else if (tid == 2) v = T > 0 ? (double)K / (double)T : 0.0;
        else if (tid == 3) v = (double)K;
        else if (tid == 4) v = (double)T;
        else if (tid == 5) v = (double)P;
        else if (tid == 6) v = (double)cnt[0];
        else if (tid == 7) v = (double)cnt[1];
        else if (tid == 8) v = (double)kKpcBins;
        else if (tid < 9 + kKpcBins) v = (double)kKpcEdges[tid - 9];
        else if (tid >= 9 + kStatsMaxBins && tid < 9 + kStatsMaxBins + kKpcBins)
            v = (double)cnt[2 + (tid - 9 - kStatsMaxBins)];
        st[tid] = v;
    }
}

hipError_t launch_stats_finalize(double* stats, int T, int P, int64_t K, hipStream_t s)
{
    hipLaunchKernelGGL(k_stats_finalize, dim3(1), dim3(256), 0, s, stats, T, P, K);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Backward
// ---------------------------------------------------------------------------

// Lane swaps of the gfx950 (v_permlane32_swap / v_permlane16_swap) on a
// pair of registers: one half (row) of a is exchanged with the other half
// (row) of b, so a + b afterwards holds a's pair sums in one half (every
// other row) and b's in the other.
__device__ __forceinline__ void swap_halves(uint32_t& a, uint32_t& b)
{
    const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
    a = r[0];
    b = r[1];
}
__device__ __forceinline__ void swap_rows(uint32_t& a, uint32_t& b)
{
    const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
    a = r[0];
    b = r[1];
}
__device__ __forceinline__ float fold_halves(float a, float b)
{
    uint32_t x = __float_as_uint(a), y = __float_as_uint(b);
    swap_halves(x, y);
    return __uint_as_float(x) + __uint_as_float(y);
}
__device__ __forceinline__ float fold_rows(float a, float b)
{
    uint32_t x = __float_as_uint(a), y = __float_as_uint(b);
    swap_rows(x, y);
    return __uint_as_float(x) + __uint_as_float(y);
}

// Reduce-scatter of 8 values across the wave: four pairs folded across the
// wave halves, two across row pairs, then inside each 16-lane row the two
// remaining values split by lane bit 3 (mirror) and three DPP steps add the
// 8-lane groups.  Every lane ends with the wave total of one value; which
// value is fixed per lane and given by scatter_term().
__device__ __forceinline__ float wave_reduce_scatter8(const float v[8], bool h3)
{
    float u[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) u[k] = fold_halves(v[k], v[k + 4]);
    const float w0 = fold_rows(u[0], u[2]);
    const float w1 = fold_rows(u[1], u[3]);
    float z = (h3 ? w1 : w0) + dpp<kDppRowMirror>(h3 ? w0 : w1);
    z += dpp<kDppRowHalfMirror>(z);
    z += dpp<kDppXor2>(z);
    z += dpp<kDppXor1>(z);
    return z;
}

// The value index whose total wave_reduce_scatter8 leaves in the calling
// lane: the same exchanges applied to the indices, keeping the first operand
// of every fold.
__device__ __forceinline__ uint32_t scatter_term(bool h3)
{
    uint32_t id[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) id[k] = (uint32_t)k;
#pragma unroll
    for (int k = 0; k < 4; ++k) swap_halves(id[k], id[k + 4]);
    swap_rows(id[0], id[2]);
    swap_rows(id[1], id[3]);
    return h3 ? id[1] : id[0];
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

// Per-slot sums of the backward: dcolor x3, (Sx, Sy) for dmean2D,
// (Sxx, Sxy, Syy) for dconic, dopacity.
constexpr int kGradTerms = 9;

// One staged Gaussian of the backward, 64 bytes.
struct __align__(16) BwdSlot {
    float4 g0;    // x, y, a', b'  (power in log2 units, as the forward)
    float4 col;   // r, g, b, id bits
    float4 con;   // conic A, B, C
    float2 g1;    // c', opacity
    float2 pad;
};
static_assert(sizeof(BwdSlot) == 64, "backward slot is 64 bytes");

// A Gaussian of more than kGatherSolo tiles owns at least 9 * (kGatherSolo + 1)
// floats of partial slots and uses none of them for partials: its sums accumulate
// there in kBigCopies copies (chosen by tile) of nine terms, each term on
// its own 128-byte line, so the float atomics of its tiles spread over 36
// lines instead of meeting on one.
constexpr int kBigLine = 32;                        // floats per 128-byte line
constexpr int kBigCopies = 4;
constexpr int kBigCopyStride = 9 * kBigLine;        // floats per copy
static_assert(kBigCopies * kBigCopyStride <= 9 * (kGatherSolo + 1), "copies fit the slots of a large Gaussian");
__device__ __forceinline__ int big_copy(int tx, int ty) { return (tx + 2 * ty) & (kBigCopies - 1); }

// Publishes the nine sums of a (tile, Gaussian) pair: the opacity, the conic
// and the screen-space constants applied, then stored at the pair's pre-sort
// position for the gather (rectangles of at most kGatherSolo tiles), or, for
// a larger rectangle, added with global float atomics into the partial slot
// of the Gaussian's first pair (its other slots stay unused); either way the
// gather writes the Gaussian's totals.  The first slot is recs[g].w, plus
// slot_base[g / 64] in a batch.
__device__ __forceinline__ void publish_pair(const BlendArgs& a, int tx, int ty, uint32_t g, float o,
                                             float4 con, const float* t)
{
    const float ddelx_dx = 0.5f * (float)a.W;
    const float ddely_dy = 0.5f * (float)a.H;
    const float sx = o * t[3], sy = o * t[4];
    const float ho = -0.5f * o;
    float d[kGradTerms];
    d[0] = t[0];
    d[1] = t[1];
    d[2] = t[2];
    d[3] = -ddelx_dx * (con.x * sx + con.y * sy);
    d[4] = -ddely_dy * (con.y * sx + con.z * sy);
    d[5] = ho * t[5];
    d[6] = ho * t[6];
    d[7] = ho * t[7];
    d[8] = t[8];
    const uint4 rc = a.recs[g];
    const size_t first = (size_t)rc.w + (a.slot_base ? (size_t)a.slot_base[g >> 6] : 0u);
    if (rc.z <= kGatherSolo) {
        const uint32_t x0 = rc.x & 0xFFFFu, y0 = rc.x >> 16, w = (rc.y & 0xFFFFu) - x0;
        const size_t p = first + (size_t)((uint32_t)ty - y0) * w + ((uint32_t)tx - x0);
        if (p < a.part_cap) {
            float* dst = a.part + (size_t)kGradTerms * p;
#pragma unroll
            for (int e = 0; e < kGradTerms; ++e) dst[e] = d[e];
            a.mark[p] = 1;
        }
    } else if (first + kGatherSolo < a.part_cap) {   // its slots lie inside the buffer
        float* acc = a.part + (size_t)kGradTerms * first + big_copy(tx, ty) * kBigCopyStride;
#pragma unroll
        for (int e = 0; e < kGradTerms; ++e)
            if (d[e] != 0.f) atomic_add_f32(&acc[e * kBigLine], d[e]);
    }
}

// Backward of one tile (one workgroup of FGS_BLOCK lanes).
//
// Per pixel the walk goes back to front from T_final like the reference
// (antialias/cuda_rasterizer/backward.cu:441-557): T is recovered by dividing by
// (1 - alpha), the colour accumulated behind the current Gaussian is carried
// already dotted with dL/dpix (the reference's per-channel sum
// (c - accum_rec) . dL/dpix is linear in accum_rec), and per (pixel,
// Gaussian) the nine sums below are formed:
//   dcolor_c += alpha T dL/dpix_c
//   v = G dL/dalpha  ->  dopacity += v,  Sx += v dx, Sy += v dy,
//                        Sxx += v dx dx, Sxy += v dx dy, Syy += v dy dy
// The opacity, the conic, -(W/2, H/2) and -1/2 of
//   dmean2D = -(W/2, H/2) o (A Sx + B Sy, B Sx + C Sy),  dconic = -1/2 o (Sxx, Sxy, Syy)
// are applied once per (tile, Gaussian) when the sums are published.
template <int BATCH>
__device__ __forceinline__ void blend_bwd_tile(const BlendArgs& a, int tile, const float* __restrict__ final_Ts,
                                               const uint32_t* __restrict__ vis_last,
                                               const float* __restrict__ dL_dpix, int exact)
{
    static_assert(BATCH <= FGS_BLOCK, "one staging lane per batch slot");
    const int tx = tile % a.gx, ty = tile / a.gx;
    const int tid = threadIdx.x;
    const int wave = tid / FGS_WAVE;
    const int lane = lane_id();
    const int px = tx * FGS_TILE + (tid & (FGS_TILE - 1));
    const int py = ty * FGS_TILE + (tid / FGS_TILE);
    const bool inside = px < a.W && py < a.H;
    float pfx = (float)px, pfy = (float)py;
    asm volatile("" : "+v"(pfx), "+v"(pfy));
    const int HW = a.W * a.H;
    const int pix = a.W * py + px;

    const uint2 range = a.ranges[tile];

    __shared__ BwdSlot s_slot[BATCH];
    __shared__ float s_g[BATCH][kGradTerms];
    __shared__ uint32_t s_wmax[kStrips];
    __shared__ uint32_t s_list[kStrips][BATCH];
    __shared__ uint32_t s_wcnt[kStrips][kStrips];

    const float T_final = inside ? final_Ts[pix] : 0.f;
    float T = T_final;
    const uint32_t last = inside ? vis_last[pix] : 0u;   // visible-list positions below it
    const float dp0 = inside ? dL_dpix[pix] : 0.f;
    const float dp1 = inside ? dL_dpix[HW + pix] : 0.f;
    const float dp2 = inside ? dL_dpix[2 * HW + pix] : 0.f;
    const float bg_dot = a.bg[0] * dp0 + a.bg[1] * dp1 + a.bg[2] * dp2;
    const float tfb = -T_final * bg_dot;   // dL/dalpha gets -T_final / (1 - alpha) * (bg . dL/dpix)
    // colour accumulated behind the current Gaussian and the colour of the
    // last blended one, both dotted with dL/dpix
    float acc_d = 0.f, last_d = 0.f, last_alpha = 0.f;

    // positions at or past the largest vis_last of the tile blend into no pixel
    const uint32_t wmax = wave_max_u32(last);
    if (lane == 0) s_wmax[wave] = wmax;
    __syncthreads();
    uint32_t bmax = 0;
#pragma unroll
    for (int w = 0; w < kStrips; ++w) bmax = s_wmax[w] > bmax ? s_wmax[w] : bmax;
    const int n = (int)min(bmax, a.tile_vis[tile]);
    const int rounds = (n + BATCH - 1) / BATCH;

    // lane roles in the reduction: lanes 0, 8, ..., 56 publish the eight
    // scattered sums, lane 1 the dopacity total
    const bool h3 = (lane >> 3) & 1;
    const bool pub_op = lane == 1;
    const bool pub = (lane & 7) == 0 || pub_op;
    const uint32_t term = pub_op ? 8u : scatter_term(h3);

    for (int r = 0; r < rounds; ++r) {
        __syncthreads();
        const int q = tid;
        const int k = r * BATCH + q;
        uint32_t bits = 0u;
        if (q < BATCH) {
            // back to front: position n-1-k of the tile's visible list (the
            // Gaussians the forward blended somewhere in the tile).  Only the
            // strips where it blended take part.
            if (k < n) {
                bits = a.kv_strips[range.x + n - 1 - k];
                const uint32_t g = a.vis_ids[range.x + n - 1 - k];
                const float2 m = a.means2D[g];
                const float4 co = a.conic_opacity[g];
                s_slot[q].g0 = make_float4(m.x, m.y, -0.5f * kLog2e * co.x, -kLog2e * co.y);
                s_slot[q].g1 = make_float2(-0.5f * kLog2e * co.z, co.w);
                s_slot[q].con = co;
                s_slot[q].col = make_float4(a.colors[3 * g + 0], a.colors[3 * g + 1],
                                            a.colors[3 * g + 2], __uint_as_float(g));
            }
#pragma unroll
            for (int t = 0; t < kGradTerms; ++t) s_g[q][t] = 0.f;
        }
        const int cnt = build_lists<BATCH, 0>(bits, s_list, s_wcnt);
        // slot j holds visible position n-1-(r*BATCH+j), which blended into
        // this pixel iff it lies below vis_last
        const int thr = n - 1 - r * BATCH - (int)last;
        for (int i = 0; i < cnt; ++i) {
            const int j = (int)s_list[wave][i];
            const BwdSlot* const sp = &s_slot[j];
            // alpha exactly as the forward computed it (same staged terms,
            // same operations), so both passes take the same decisions
            const float4 g0 = sp->g0;
            const float2 g1 = sp->g1;
            const float dx = g0.x - pfx, dy = g0.y - pfy;
            const float p2 = exponent2(g0, g1, dx, dy);
            const float G = __builtin_amdgcn_exp2f(fminf(p2, 0.0f));
            const float alpha_raw = g1.y * G;
            const float alpha = fminf(0.99f, alpha_raw);
            const bool act = j > thr && !(p2 > 0.0f) && !(alpha < 1.0f / 255.0f);
            if (!__any(act)) continue;   // wave-uniform
            // every lane computes; inactive lanes contribute zeros and keep
            // their state (all terms stay finite: G <= 1, alpha <= 0.99)
            const float4 col = sp->col;
            const float inv = __builtin_amdgcn_rcpf(1.f - alpha);
            const float Tn = T * inv;
            const float c_d = col.x * dp0 + col.y * dp1 + col.z * dp2;
            const float acc_n = acc_d + last_alpha * (last_d - acc_d);
            const float dLa = (c_d - acc_n) * Tn + inv * tfb;
            const bool flows = act && !(exact && alpha_raw > 0.99f);
            const float dch = act ? alpha * Tn : 0.f;
            const float vop = flows ? G * dLa : 0.f;
            float v[8];
            v[0] = dch * dp0;
            v[1] = dch * dp1;
            v[2] = dch * dp2;
            const float vdx = vop * dx, vdy = vop * dy;
            v[3] = vdx;
            v[4] = vdy;
            v[5] = vdx * dx;
            v[6] = vdx * dy;
            v[7] = vdy * dy;
            if (act) {
                T = Tn;
                acc_d = acc_n;
                last_d = c_d;
                last_alpha = alpha;
            }
            // the waves meet on the same slot: LDS float atomics
            const float z = wave_reduce_scatter8(v, h3);
            const float zop = wave_total(vop);
            if (pub) atomicAdd(&s_g[j][term], pub_op ? zop : z);
        }
        __syncthreads();
        if (bits != 0u) {
            float t[kGradTerms];
#pragma unroll
            for (int e = 0; e < kGradTerms; ++e) t[e] = s_g[q][e];
            publish_pair(a, tx, ty, __float_as_uint(s_slot[q].col.w), s_slot[q].g1.y, s_slot[q].con, t);
        }
    }
}

// Backward of one tile by a single wave: lane l owns column l % 16 and, in
// each of the four strips, row l / 16 of the strip (4 pixels per lane, one
// per strip).  Per staged Gaussian the wave works the strips its mask names
// (a wave-uniform branch each), adds the strips' terms in registers and
// reduces the nine sums across the wave once; the sums go straight to the
// slot's LDS row (one wave: no atomics, and a fixed summation order).  The
// per-pixel arithmetic is that of blend_bwd_tile, term for term.
constexpr int kWaveBatch = FGS_WAVE;

template <bool EXACT>
__device__ __forceinline__ void blend_bwd_tile_wave(const BlendArgs& a, int tile,
                                                      const float* __restrict__ final_Ts,
                                                      const uint32_t* __restrict__ vis_last,
                                                      const float* __restrict__ dL_dpix)
{
    const int tx = tile % a.gx, ty = tile / a.gx;
    const int lane = lane_id();
    const int px = tx * FGS_TILE + (lane & (FGS_TILE - 1));
    const int py0 = ty * FGS_TILE + (lane / FGS_TILE);
    float pfx = (float)px;
    asm volatile("" : "+v"(pfx));
    const int HW = a.W * a.H;

    __shared__ BwdSlot s_slot[kWaveBatch];
    __shared__ float s_g[kWaveBatch][kGradTerms];

    float pfy[kStrips], T[kStrips], dp0[kStrips], dp1[kStrips], dp2[kStrips], tfb[kStrips];
    // acc: the colour of the Gaussians already walked (those in front of
    // the next one), dotted with dL/dpix -- the reference's accum_rec,
    // updated as soon as a Gaussian is walked rather than at the next one
    float acc[kStrips];
    uint32_t last[kStrips];
    uint32_t lmax = 0u;
#pragma unroll
    for (int k = 0; k < kStrips; ++k) {
        const int py = py0 + 4 * k;
        pfy[k] = (float)py;
        asm volatile("" : "+v"(pfy[k]));
        const bool inside = px < a.W && py < a.H;
        const int pix = a.W * py + px;
        const float T_final = inside ? final_Ts[pix] : 0.f;
        T[k] = T_final;
        last[k] = inside ? vis_last[pix] : 0u;
        dp0[k] = inside ? dL_dpix[pix] : 0.f;
        dp1[k] = inside ? dL_dpix[HW + pix] : 0.f;
        dp2[k] = inside ? dL_dpix[2 * HW + pix] : 0.f;
        const float bg_dot = a.bg[0] * dp0[k] + a.bg[1] * dp1[k] + a.bg[2] * dp2[k];
        tfb[k] = -T_final * bg_dot;
        acc[k] = 0.f;
        lmax = last[k] > lmax ? last[k] : lmax;
    }
    const uint32_t wmax = __builtin_amdgcn_readfirstlane(wave_max_u32(lmax));
    const int n = __builtin_amdgcn_readfirstlane((int)min(wmax, a.tile_vis[tile]));
    const uint2 range = a.ranges[tile];
    const bool h3 = (lane >> 3) & 1;
    const bool pub_op = lane == FGS_WAVE - 1;   // the opacity total lands in lane 63
    const bool pub = (lane & 7) == 0 || pub_op;
    const uint32_t term = pub_op ? 8u : scatter_term(h3);

    for (int r = 0; r * kWaveBatch < n; ++r) {
        const int k0 = r * kWaveBatch;
        const int cnt = __builtin_amdgcn_readfirstlane(min(kWaveBatch, n - k0));
        uint32_t bits = 0u;
        __syncthreads();
        if (lane < cnt) {
            const int pos = n - 1 - (k0 + lane);   // back to front
            bits = a.kv_strips[range.x + pos];
            const uint32_t g = a.vis_ids[range.x + pos];
            const float2 m = a.means2D[g];
            const float4 co = a.conic_opacity[g];
            s_slot[lane].g0 = make_float4(m.x, m.y, -0.5f * kLog2e * co.x, -kLog2e * co.y);
            s_slot[lane].g1 = make_float2(-0.5f * kLog2e * co.z, co.w);
            s_slot[lane].con = co;
            s_slot[lane].col = make_float4(a.colors[3 * g + 0], a.colors[3 * g + 1], a.colors[3 * g + 2],
                                           __uint_as_float(g));
        }
#pragma unroll
        for (int t = 0; t < kGradTerms; ++t) s_g[lane][t] = 0.f;
        __syncthreads();
        // slot j holds position n-1-(k0+j), which blended into a pixel iff it
        // lies below that pixel's vis_last
        int thr[kStrips];
#pragma unroll
        for (int k = 0; k < kStrips; ++k) thr[k] = n - 1 - k0 - (int)last[k];
        for (int j = 0; j < cnt; ++j) {
            const uint32_t sb = __builtin_amdgcn_readlane(bits, j);
            const BwdSlot* const sp = &s_slot[j];
            const float4 g0 = sp->g0;
            const float2 g1 = sp->g1;
            const float dx = g0.x - pfx;
            const ExpCol ec = exponent2_col(g0, dx);
            // per lane the strips share the column, so dx: the strips add
            // v, v dy and v dy^2, and v dx, v dx^2, v dx dy follow per entry
            float v[8] = { 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f };
            float vsum = 0.f;
            bool any = false;
#pragma unroll
            for (int k = 0; k < kStrips; ++k) {
                if (!((sb >> k) & 1u)) continue;   // wave-uniform
                const float dy = g0.y - pfy[k];
                const float p2 = exponent2_row(ec, g1, dy);
                // G is only used where the pair acts (power <= 0); elsewhere
                // it may be inf, and every use of it there is a select
                const float G = __builtin_amdgcn_exp2f(p2);
                const float alpha_raw = g1.y * G;
                const float alpha = fminf(0.99f, alpha_raw);
                const bool act = j > thr[k] && !(p2 > 0.0f) && !(alpha < 1.0f / 255.0f);
                // the ballot of act from those of its compares (lane masks)
                if ((__builtin_amdgcn_ballot_w64(j > thr[k]) & __builtin_amdgcn_ballot_w64(!(p2 > 0.0f)) &
                     __builtin_amdgcn_ballot_w64(!(alpha < 1.0f / 255.0f))) == 0u)
                    continue;   // wave-uniform
                any = true;
                const float4 col = sp->col;
                const float inv = __builtin_amdgcn_rcpf(1.f - alpha);
                const float Tn = T[k] * inv;
                const float c_d = col.x * dp0[k] + col.y * dp1[k] + col.z * dp2[k];
                const float cd_acc = c_d - acc[k];
                const float dLa = cd_acc * Tn + inv * tfb[k];
                if (act) {   // the sums and the state of the acting lanes only
                    const float dch = alpha * Tn;
                    v[0] += dch * dp0[k];
                    v[1] += dch * dp1[k];
                    v[2] += dch * dp2[k];
                    if (!(EXACT && alpha_raw > 0.99f)) {   // exact mode: no flow through the clamp
                        const float vop = G * dLa;
                        const float vdy = vop * dy;
                        v[4] += vdy;
                        v[7] += vdy * dy;
                        vsum += vop;
                    }
                    acc[k] = acc[k] + alpha * cd_acc;
                    T[k] = Tn;
                }
            }
            if (!any) continue;   // wave-uniform
            v[3] = vsum * dx;
            v[5] = v[3] * dx;
            v[6] = v[4] * dx;
            const float z = wave_reduce_scatter8(v, h3);
            const float zop = wave_total_last(vsum);
            if (pub) s_g[j][term] = pub_op ? zop : z;
        }
        __syncthreads();
        if (bits != 0u) {
            float t[kGradTerms];
#pragma unroll
            for (int e = 0; e < kGradTerms; ++e) t[e] = s_g[lane][e];
            publish_pair(a, tx, ty, __float_as_uint(s_slot[lane].col.w), s_slot[lane].g1.y, s_slot[lane].con, t);
        }
    }
}

// Gather of the per-pair gradients: each Gaussian (one view: slot s of the
// depth order; batch: element e) sums the marked part entries of its slots
// [off, off + n) and stores the nine totals (plain stores: every Gaussian
// has one item).

// whole: also the third (zero) component of dL_dmean2D.
__device__ __forceinline__ void store_grads(uint32_t g, const float* t, float* __restrict__ dL_dmean2D,
                                            float* __restrict__ dL_dconic, float* __restrict__ dL_dopacity,
                                            float* __restrict__ dL_dcolors, bool whole)
{
    dL_dcolors[3 * (size_t)g + 0] = t[0];
    dL_dcolors[3 * (size_t)g + 1] = t[1];
    dL_dcolors[3 * (size_t)g + 2] = t[2];
    dL_dmean2D[3 * (size_t)g + 0] = t[3];
    dL_dmean2D[3 * (size_t)g + 1] = t[4];
    if (whole) dL_dmean2D[3 * (size_t)g + 2] = 0.f;
    dL_dconic[3 * (size_t)g + 0] = t[5];
    dL_dconic[3 * (size_t)g + 1] = t[6];
    dL_dconic[3 * (size_t)g + 2] = t[7];
    dL_dopacity[g] = t[8];
}

// Segments of up to kGatherLane pairs are summed by their own lane; longer
// ones (up to kGatherSolo) by the whole wave, one at a time, lanes striding
// over the pairs and the nine sums reduced across the wave, so a long
// segment does not hold up its wave for a hundred dependent loads.  Larger
// rectangles read the sums the blend added into their first pair's slot.
// One view: Gaussians of at most kGatherSolo tiles none of whose pairs
// blended store nothing (the outputs start at zero); a batch stores every
// visible element's gradients, whole, in element order.
constexpr uint32_t kGatherLane = 16;

__device__ __forceinline__ size_t gather_live(const GradGather& gg)
{
    return gg.n_dev && (size_t)*gg.n_dev < gg.n ? (size_t)*gg.n_dev : gg.n;
}

// Item i of the gather: the Gaussian g with n tiles (0: none) whose slots
// start at off.  A batch's items are its visible elements in element order
// (neighbouring lanes read neighbouring records and slot runs and store to
// neighbouring outputs).
__device__ __forceinline__ void gather_item(const GradGather& gg, size_t i, size_t live, uint32_t& n, uint32_t& g,
                                            uint32_t& off)
{
    n = 0u;
    g = 0u;
    off = 0u;
    if (i >= live) return;
    if (gg.slot_base) {
        const uint32_t e = gg.elems[i];
        const uint4 rc = gg.recs[e];
        n = rc.z;
        g = e;
        off = gg.slot_base[e >> 6] + rc.w;
    } else {
        n = gg.tiles_s[i];
        if (n > 0u) {
            g = (uint32_t)gg.vkeys[i];
            off = gg.offs[i];
        }
    }
}

__global__ __launch_bounds__(256) void k_grad_gather(GradGather gg, const float* __restrict__ part,
                                                     const uint8_t* __restrict__ mark, size_t cap,
                                                     float* __restrict__ dL_dmean2D,
                                                     float* __restrict__ dL_dconic,
                                                     float* __restrict__ dL_dopacity,
                                                     float* __restrict__ dL_dcolors)
{
    // grid-stride in whole blocks (the wave steps below need every lane):
    // with a device count the grid is sized for the capacity
    const size_t live = gather_live(gg);
    const int lane = lane_id();
    const bool all = gg.slot_base != nullptr;
    for (size_t b0 = (size_t)blockIdx.x * blockDim.x; b0 < live; b0 += (size_t)gridDim.x * blockDim.x) {
        uint32_t n, g, off;
        gather_item(gg, b0 + threadIdx.x, live, n, g, off);
        const bool gathered = n > 0u && n <= kGatherSolo;
        if (gathered && n <= kGatherLane) {
            float t[kGradTerms] = {};
            bool any = false;
            for (uint32_t k = 0; k < n; ++k) {
                const size_t p = (size_t)off + k;
                if (p < cap && mark[p]) {
                    any = true;
#pragma unroll
                    for (int e = 0; e < kGradTerms; ++e) t[e] += part[(size_t)kGradTerms * p + e];
                }
            }
            if (any || all) store_grads(g, t, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors, all);
        } else if (n > kGatherSolo) {
            float t[kGradTerms] = {};
            if ((size_t)off + kGatherSolo < cap) {
                const float* acc = part + (size_t)kGradTerms * off;
#pragma unroll
                for (int c = 0; c < kBigCopies; ++c)
#pragma unroll
                    for (int e = 0; e < kGradTerms; ++e) t[e] += acc[c * kBigCopyStride + e * kBigLine];
            }
            store_grads(g, t, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors, all);
        }
        uint64_t big = __ballot(gathered && n > kGatherLane);
        while (big) {
            const int src = __ffsll((long long)big) - 1;
            big &= big - 1;
            const uint32_t bn = __builtin_amdgcn_readlane(n, src);
            const uint32_t boff = __builtin_amdgcn_readlane(off, src);
            const uint32_t bg = __builtin_amdgcn_readlane(g, src);
            float t[kGradTerms] = {};
            bool mine = false;
            for (uint32_t k = lane; k < bn; k += FGS_WAVE) {
                const size_t p = (size_t)boff + k;
                if (p < cap && mark[p]) {
                    mine = true;
#pragma unroll
                    for (int e = 0; e < kGradTerms; ++e) t[e] += part[(size_t)kGradTerms * p + e];
                }
            }
            if (__any(mine) || all) {   // wave-uniform
#pragma unroll
                for (int e = 0; e < kGradTerms; ++e) t[e] = wave_total(t[e]);
                if (lane == src) store_grads(bg, t, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors, all);
            }
        }
    }
}

// Segmented inclusive wave sum of N values per lane: head marks the first
// lane of a segment (lane 0 starts one in any case).  Kogge-Stone steps of
// 1, 2, 4 and 8 lanes inside each 16-lane row (DPP row shifts), then the
// rows take the running sum of the row before them (lanes 15, 31 and 47)
// up to their first head.  The order of the additions depends on the lane
// positions only.
template <int N>
__device__ __forceinline__ void wave_segmented_sum(float* x, bool head)
{
    const int lane = lane_id();
    const int col = lane & 15;
    uint32_t f = head || lane == 0 ? 1u : 0u;
#define FGS_SEG_STEP(D)                                                                              \
    {                                                                                                \
        const uint32_t fu = dpp_u<0x110 + D>(f);                                                     \
        float xu[N];                                                                                 \
        _Pragma("unroll") for (int e = 0; e < N; ++e) xu[e] = dpp<0x110 + D>(x[e]);                  \
        if (col >= D) {                                                                              \
            if (!f) {                                                                                \
                _Pragma("unroll") for (int e = 0; e < N; ++e) x[e] += xu[e];                         \
            }                                                                                        \
            f |= fu;                                                                                 \
        }                                                                                            \
    }
    FGS_SEG_STEP(1)
    FGS_SEG_STEP(2)
    FGS_SEG_STEP(4)
    FGS_SEG_STEP(8)
#undef FGS_SEG_STEP
    // rows 1 and 3 take lanes 15 and 47, then rows 2 and 3 lane 31
    const int row = lane >> 4;
    {
        const uint32_t f15 = __builtin_amdgcn_readlane(f, 15), f47 = __builtin_amdgcn_readlane(f, 47);
        float xu[N];
#pragma unroll
        for (int e = 0; e < N; ++e) {
            const float a = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(x[e]), 15));
            const float b = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(x[e]), 47));
            xu[e] = row == 1 ? a : b;
        }
        if (row & 1) {
            if (!f) {
#pragma unroll
                for (int e = 0; e < N; ++e) x[e] += xu[e];
            }
            f |= row == 1 ? f15 : f47;
        }
    }
    {
        const uint32_t f31 = __builtin_amdgcn_readlane(f, 31);
        float xu[N];
#pragma unroll
        for (int e = 0; e < N; ++e) xu[e] = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(x[e]), 31));
        if (row >= 2) {
            if (!f) {
#pragma unroll
                for (int e = 0; e < N; ++e) x[e] += xu[e];
            }
            f |= f31;
        }
    }
}

// Batch gather: one wave per 64 consecutive visible elements of the list.
// Their slot runs follow each other (the list is in element order and the
// runs are numbered in that order), so the wave walks the slots of its
// elements of up to kGatherSolo tiles together, 64 a step, one per lane: the
// marks and sums load coalesced, a segmented wave sum adds up each
// element's slots of the step, and the last lane of each segment adds it
// into the element's totals in LDS.  Larger elements read the copies the
// blend accumulated into.
__global__ __launch_bounds__(FGS_WAVE) void k_grad_gather_elems(GradGather gg, const float* __restrict__ part,
                                                                const uint8_t* __restrict__ mark, size_t cap,
                                                                float* __restrict__ dL_dmean2D,
                                                                float* __restrict__ dL_dconic,
                                                                float* __restrict__ dL_dopacity,
                                                                float* __restrict__ dL_dcolors)
{
    __shared__ uint32_t s_pre[FGS_WAVE];   // slots of the wave's solo elements before the lane's
    __shared__ uint32_t s_off[FGS_WAVE];   // the lane's first slot
    __shared__ uint32_t s_start[FGS_WAVE]; // step: 1 + the lane whose run starts at the position
    __shared__ float s_acc[kGradTerms][FGS_WAVE];
    const int lane = lane_id();
    const size_t live = gather_live(gg);
    for (size_t j0 = (size_t)blockIdx.x * FGS_WAVE; j0 < live; j0 += (size_t)gridDim.x * FGS_WAVE) {
        uint32_t n, g, off;
        gather_item(gg, j0 + lane, live, n, g, off);
        const uint32_t ns = n <= kGatherSolo ? n : 0u;
        const uint32_t incl = wave_incl_sum_u32(ns);
        const uint32_t pre = incl - ns;
        const uint32_t total = __builtin_amdgcn_readlane(incl, FGS_WAVE - 1);
        __syncthreads();   // the previous elements' totals are read
        s_pre[lane] = pre;
        s_off[lane] = off;
#pragma unroll
        for (int e = 0; e < kGradTerms; ++e) s_acc[e][lane] = 0.f;
        uint32_t carry = 0u;   // the lane owning the previous step's last slot
        for (uint32_t r0 = 0; r0 < total; r0 += FGS_WAVE) {
            s_start[lane] = 0u;
            __syncthreads();
            if (ns > 0u && pre >= r0 && pre < r0 + FGS_WAVE) s_start[pre - r0] = (uint32_t)lane + 1u;
            __syncthreads();
            const uint32_t st = s_start[lane];
            const uint32_t ow = wave_incl_max_u32(st);
            const uint32_t owner = ow ? ow - 1u : carry;
            carry = __builtin_amdgcn_readlane(owner, FGS_WAVE - 1);
            const uint32_t idx = r0 + (uint32_t)lane;
            const bool act = idx < total;
            const size_t p = (size_t)s_off[owner] + (idx - s_pre[owner]);
            float x[kGradTerms];
#pragma unroll
            for (int e = 0; e < kGradTerms; ++e) x[e] = 0.f;
            if (act && p < cap && mark[p]) {
#pragma unroll
                for (int e = 0; e < kGradTerms; ++e) x[e] = part[(size_t)kGradTerms * p + e];
            }
            wave_segmented_sum<kGradTerms>(x, st != 0u);
            // lane + 1's start flag (DPP wave_shl:1; lane 63 reads 0 and is a
            // tail anyway)
            const bool next_head = __builtin_amdgcn_update_dpp(0, (int)st, 0x130, 0xF, 0xF, true) != 0;
            const bool tail = act && (lane == FGS_WAVE - 1 || idx + 1 == total || next_head);
            if (tail) {
#pragma unroll
                for (int e = 0; e < kGradTerms; ++e) s_acc[e][owner] += x[e];
            }
            __syncthreads();
        }
        if (n > 0u && n <= kGatherSolo) {
            float t[kGradTerms];
#pragma unroll
            for (int e = 0; e < kGradTerms; ++e) t[e] = s_acc[e][lane];
            store_grads(g, t, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors, true);
        } else if (n > kGatherSolo) {
            float t[kGradTerms] = {};
            if ((size_t)off + kGatherSolo < cap) {
                const float* acc = part + (size_t)kGradTerms * off;
#pragma unroll
                for (int c = 0; c < kBigCopies; ++c)
#pragma unroll
                    for (int e = 0; e < kGradTerms; ++e) t[e] += acc[c * kBigCopyStride + e * kBigLine];
            }
            store_grads(g, t, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors, true);
        }
    }
}

// Clears the accumulation slot (first slot's partials) of every Gaussian of
// more than kGatherSolo tiles before the blend adds into it (a batch: those
// of the big list).
__global__ __launch_bounds__(256) void k_clear_big(GradGather gg, float* __restrict__ part, size_t cap)
{
    const size_t live = gg.big ? (size_t)*gg.n_big : gather_live(gg);
    for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < live; s += (size_t)gridDim.x * blockDim.x) {
        uint32_t n, o;
        if (gg.big) {
            const uint32_t e = gg.big[s];
            const uint4 rc = gg.recs[e];
            n = rc.z;
            o = gg.slot_base[e >> 6] + rc.w;
        } else {
            n = gg.tiles_s[s];
            o = n > kGatherSolo ? gg.offs[s] : 0u;
        }
        if (n > kGatherSolo) {
            const size_t off = o;
            if (off + kGatherSolo < cap) {
                float* acc = part + (size_t)kGradTerms * off;
#pragma unroll
                for (int c = 0; c < kBigCopies; ++c)
#pragma unroll
                    for (int e = 0; e < kGradTerms; ++e) acc[c * kBigCopyStride + e * kBigLine] = 0.f;
            }
        }
    }
}


// Clear the marks and the large Gaussians' slots, blend, gather into the outputs.
template <typename Launch>
hipError_t blend_bwd_and_gather(const BlendArgs& a, const GradGather& gg, float* dL_dmean2D,
                                float* dL_dconic, float* dL_dopacity, float* dL_dcolors, hipStream_t s,
                                Launch&& launch)
{
    hipError_t e = launch_fill_u32(a.mark, 0u, (a.part_cap + 3) / 4, s);
    if (e != hipSuccess) return e;
    if (gg.n > 0) {
        size_t blocks = (gg.n + 255) / 256;
        if (blocks > (gg.big ? 256u : kStrideBlocks)) blocks = gg.big ? 256u : kStrideBlocks;
        hipLaunchKernelGGL(k_clear_big, dim3((unsigned)blocks), dim3(256), 0, s, gg, a.part, a.part_cap);
    }
    prof_kernel(true, true, s);
    launch(a);
    prof_kernel(true, false, s);
    if ((e = hipGetLastError()) != hipSuccess || gg.n == 0) return e;
    if (gg.slot_base) {   // batch: one wave per 64 visible elements
        size_t blocks = (gg.n + FGS_WAVE - 1) / FGS_WAVE;
        if (blocks > 16 * kStrideBlocks) blocks = 16 * kStrideBlocks;
        hipLaunchKernelGGL(k_grad_gather_elems, dim3((unsigned)blocks), dim3(FGS_WAVE), 0, s, gg, a.part, a.mark,
                           a.part_cap, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors);
        return hipGetLastError();
    }
    size_t blocks = (gg.n + 255) / 256;
    if (gg.n_dev && blocks > kStrideBlocks) blocks = kStrideBlocks;
    hipLaunchKernelGGL(k_grad_gather, dim3((unsigned)blocks), dim3(256), 0, s, gg, a.part,
                       a.mark, a.part_cap, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors);
    return hipGetLastError();
}

// The one-wave backward kernels come in one instance per gradient mode
// (EXACT): one inlined copy each keeps their register count, and so their
// occupancy, that of a single path.
template <bool EXACT>
__global__ __launch_bounds__(FGS_WAVE) void k_blend_bwd_wave(BlendArgs a, const float* final_Ts,
                                                             const uint32_t* vis_last, const float* dL_dpix)
{
    blend_bwd_tile_wave<EXACT>(a, blockIdx.x, final_Ts, vis_last, dL_dpix);
}

template <bool EXACT>
__global__ __launch_bounds__(FGS_WAVE) void k_blend_bwd_batch_wave(BlendArgs common, const BatchView* views,
                                                                   int V, const float* final_Ts,
                                                                   const uint32_t* vis_last,
                                                                   const float* dL_dpix)
{
    const uint32_t b = xcd_tile(blockIdx.x, gridDim.x, (uint32_t)common.xcd_chunk);
    const int v = find_view(views, V, b);
    const BatchView& vw = views[v];
    const BlendArgs a = bind_blend(common, vw, v);
    const size_t pb = vw.pix_base;
    blend_bwd_tile_wave<EXACT>(a, (int)(b - vw.tile_base), final_Ts + pb, vis_last + pb, dL_dpix + 3 * pb);
}

// The same tiles as k_blend_bwd_batch_wave, claimed from the cost-ordered
// queue by a persistent grid.
template <bool EXACT>
__global__ __launch_bounds__(FGS_WAVE) void k_blend_bwd_batch_queue(BlendArgs common, const BatchView* views,
                                                                    int V, uint32_t T_total,
                                                                    const float* final_Ts,
                                                                    const uint32_t* vis_last,
                                                                    const float* dL_dpix)
{
    drain_queue(common, T_total, [&](uint32_t b) {
        const int v = find_view(views, V, b);
        const BatchView& vw = views[v];
        const BlendArgs a = bind_blend(common, vw, v);
        const size_t pb = vw.pix_base;
        blend_bwd_tile_wave<EXACT>(a, (int)(b - vw.tile_base), final_Ts + pb, vis_last + pb, dL_dpix + 3 * pb);
    });
}

template <int BATCH>
__global__ __launch_bounds__(FGS_BLOCK) void k_blend_bwd(BlendArgs a, const float* final_Ts,
                                                         const uint32_t* vis_last, const float* dL_dpix,
                                                         int exact)
{
    blend_bwd_tile<BATCH>(a, blockIdx.x, final_Ts, vis_last, dL_dpix, exact);
}

template <int BATCH>
__global__ __launch_bounds__(FGS_BLOCK) void k_blend_bwd_batch(BlendArgs common, const BatchView* views,
                                                               int V, const float* final_Ts,
                                                               const uint32_t* vis_last,
                                                               const float* dL_dpix, int exact)
{
    const int v = find_view(views, V, blockIdx.x);
    const BatchView& vw = views[v];
    const BlendArgs a = bind_blend(common, vw, v);
    const size_t pb = vw.pix_base;
    blend_bwd_tile<BATCH>(a, (int)(blockIdx.x - vw.tile_base), final_Ts + pb, vis_last + pb,
                          dL_dpix + 3 * pb, exact);
}

hipError_t launch_blend_bwd_batch(const BlendArgs& common, const GradGather& gg, const BatchView* views,
                                  int V, int T_total, const float* final_T, const uint32_t* vis_last,
                                  const float* dL_dpix, float* dL_dmean2D, float* dL_dconic,
                                  float* dL_dopacity, float* dL_dcolors, int exact, hipStream_t s)
{
    if (T_total == 0) return hipSuccess;
    return blend_bwd_and_gather(common, gg, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors, s,
                                [&](const BlendArgs& b) {
#define FGS_BWD(BATCH)                                                                               \
    hipLaunchKernelGGL(k_blend_bwd_batch<BATCH>, dim3(T_total), dim3(FGS_BLOCK), 0, s, b, views, V, \
                       final_T, vis_last, dL_dpix, exact)
        if (b.bwd_layout == 1 && b.queue_order) {
            const void* k = exact ? reinterpret_cast<const void*>(&k_blend_bwd_batch_queue<true>)
                                  : reinterpret_cast<const void*>(&k_blend_bwd_batch_queue<false>);
            const uint32_t grid = persistent_grid(k, exact ? 2 : 1, (uint32_t)T_total);
            if (exact)
                hipLaunchKernelGGL(k_blend_bwd_batch_queue<true>, dim3(grid), dim3(FGS_WAVE), 0, s, b, views, V,
                                   (uint32_t)T_total, final_T, vis_last, dL_dpix);
            else
                hipLaunchKernelGGL(k_blend_bwd_batch_queue<false>, dim3(grid), dim3(FGS_WAVE), 0, s, b, views, V,
                                   (uint32_t)T_total, final_T, vis_last, dL_dpix);
        } else if (b.bwd_layout == 1 && exact)
            hipLaunchKernelGGL(k_blend_bwd_batch_wave<true>, dim3(T_total), dim3(FGS_WAVE), 0, s, b, views, V,
                               final_T, vis_last, dL_dpix);
        else if (b.bwd_layout == 1)
            hipLaunchKernelGGL(k_blend_bwd_batch_wave<false>, dim3(T_total), dim3(FGS_WAVE), 0, s, b, views, V,
                               final_T, vis_last, dL_dpix);
        else if (b.bwd_batch == 256) FGS_BWD(256);
        else if (b.bwd_batch == 64) FGS_BWD(64);
        else FGS_BWD(128);
#undef FGS_BWD
    });
}

hipError_t launch_blend_bwd(const BlendArgs& a, const GradGather& gg, const float* final_T,
                            const uint32_t* vis_last, const float* dL_dpix, float* dL_dmean2D,
                            float* dL_dconic, float* dL_dopacity, float* dL_dcolors, int exact,
                            hipStream_t s)
{
    const int T = a.gx * a.gy;
    if (T == 0) return hipSuccess;
    return blend_bwd_and_gather(a, gg, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors, s,
                                [&](const BlendArgs& b) {
#define FGS_BWD(BATCH)                                                                              \
    hipLaunchKernelGGL(k_blend_bwd<BATCH>, dim3(T), dim3(FGS_BLOCK), 0, s, b, final_T, vis_last, dL_dpix, \
                       exact)
        if (b.bwd_layout == 1 && exact)
            hipLaunchKernelGGL(k_blend_bwd_wave<true>, dim3(T), dim3(FGS_WAVE), 0, s, b, final_T, vis_last,
                               dL_dpix);
        else if (b.bwd_layout == 1)
            hipLaunchKernelGGL(k_blend_bwd_wave<false>, dim3(T), dim3(FGS_WAVE), 0, s, b, final_T, vis_last,
                               dL_dpix);
        else if (b.bwd_batch == 256) FGS_BWD(256);
        else if (b.bwd_batch == 64) FGS_BWD(64);
        else FGS_BWD(128);
#undef FGS_BWD
    });
}

}  // namespace fgs
