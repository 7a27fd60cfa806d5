// sort.hip -- tile binning: scans, wave64 LSD radix sort, duplicate-with-keys
// and tile ranges.
//
// The reference (rasterizer_impl.cu:113-147, :363, :427-445, :172-194) sorts
// K = sum(tiles_touched) 64-bit keys (tile << 32 | depth bits) on 32 +
// bit_length(T) bits.  The result is the order (tile, depth bits, Gaussian
// index).  Here the same order is produced in two smaller sorts:
//   1. the P per-Gaussian keys (depth bits << 32 | index) are sorted on their
//      upper 32 bits (culled Gaussians carry 0xFFFFFFFF and land last);
//   2. duplicate_keys emits (tile << 32 | index) in that depth order;
//   3. a stable LSD sort on the tile bits only (bit_length(T) bits) keeps the
//      depth order inside every tile.
// The tile rectangle of each Gaussian is packed by the preprocess and moved
// into depth order once, so the offset scan and the duplication read
// coalesced streams.
// Stable LSD passes use 8-bit digits: a per-block digit histogram, a
// device-wide exclusive scan of the digit-major histogram table, and a
// scatter that ranks equal digits inside each 256-element slab with wave64
// ballots (match masks) and an LDS prefix over the four waves.
#include "fgs_device.h"
#include "fgs_internal.h"

namespace fgs {

constexpr int kThreads = 256;
constexpr int kIpt = 16;                      // items per thread
constexpr int kChunk = kThreads * kIpt;       // 4096 items per block
constexpr int kRadix = 256;

static inline size_t div_up(size_t a, size_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// Block-wide exclusive scan of one u32 per thread (256 threads = 4 waves).
// Returns the exclusive prefix; *total receives the block sum.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t* total)
{
    const int lane = lane_id();
    const int wave = threadIdx.x / FGS_WAVE;
    const uint32_t incl = wave_incl_sum_u32(v);
    if (lane == FGS_WAVE - 1) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wave_off = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < kThreads / FGS_WAVE; ++w) {
        const uint32_t t = s_wave[w];
        if (w < wave) wave_off += t;
        sum += t;
    }
    __syncthreads();
    *total = sum;
    return wave_off + incl - v;
}

// Loaders of the scans: v[0 .. kIpt) = the values at base .. base + kIpt
// (zero at or past n); the plain one reads 16 bytes at a time.
struct PlainLoader {
    const uint32_t* in;
    __device__ void load(size_t base, size_t n, uint32_t* v) const
    {
        if (base + kIpt <= n && ((uintptr_t)in & 15u) == 0) {
            const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
            for (int q = 0; q < kIpt / 4; ++q) {
                const uint4 x = p[q];
                v[4 * q] = x.x;
                v[4 * q + 1] = x.y;
                v[4 * q + 2] = x.z;
                v[4 * q + 3] = x.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kIpt; ++j) v[j] = base + j < n ? in[base + j] : 0u;
        }
    }
};

// n_dev (may be null): only the first min(n, *n_dev) elements are scanned;
// the grid still covers n.
__device__ __forceinline__ size_t live_count(size_t n, const unsigned long long* n_dev)
{
    if (!n_dev) return n;
    const size_t d = (size_t)*n_dev;
    return d < n ? d : n;
}

template <typename Loader>
__global__ __launch_bounds__(kThreads) void k_scan_reduce(Loader ld, size_t n, uint32_t* partials,
                                                          unsigned long long* nz_count,
                                                          const unsigned long long* n_dev)
{
    __shared__ uint32_t s_wave[kThreads / FGS_WAVE];
    n = live_count(n, n_dev);
    if ((size_t)blockIdx.x * kChunk >= n) {   // past the live values: nothing to add
        if (threadIdx.x == 0) partials[blockIdx.x] = 0u;
        return;
    }
    const size_t base = (size_t)blockIdx.x * kChunk + (size_t)threadIdx.x * kIpt;
    uint32_t v[kIpt];
    ld.load(base, n, v);
    uint32_t acc = 0, nz = 0;
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
        acc += v[j];
        nz += v[j] != 0u;
    }
    uint32_t total;
    block_exclusive_scan(acc, s_wave, &total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
    if (nz_count) {
        uint32_t nz_total;
        block_exclusive_scan(nz, s_wave, &nz_total);
        if (threadIdx.x == 0 && nz_total) atomicAdd(nz_count, (unsigned long long)nz_total);
    }
}

// Single workgroup: exclusive scan of partials[0..nb) in place, kIpt per
// thread per step; the grand total (64-bit) goes to *total when given, and
// *copy_src to *copy_dst when both are.
__global__ __launch_bounds__(kThreads) void k_scan_partials(uint32_t* partials, size_t nb,
                                                            unsigned long long* total_out,
                                                            const unsigned long long* copy_src,
                                                            unsigned long long* copy_dst)
{
    if (copy_src && copy_dst && threadIdx.x == 0) *copy_dst = *copy_src;
    __shared__ uint32_t s_wave[kThreads / FGS_WAVE];
    uint32_t carry = 0;
    unsigned long long carry64 = 0;
    for (size_t b0 = 0; b0 < nb; b0 += kChunk) {
        const size_t base = b0 + (size_t)threadIdx.x * kIpt;
        uint32_t v[kIpt];
        PlainLoader{ partials }.load(base, nb, v);
        uint32_t acc = 0;
#pragma unroll
        for (int j = 0; j < kIpt; ++j) acc += v[j];
        uint32_t total;
        uint32_t run = carry + block_exclusive_scan(acc, s_wave, &total);
#pragma unroll
        for (int j = 0; j < kIpt; ++j) {
            if (base + j < nb) partials[base + j] = run;
            run += v[j];
        }
        carry += total;
        carry64 += total;
    }
    if (total_out && threadIdx.x == 0) *total_out = carry64;
}

template <typename Loader>
__global__ __launch_bounds__(kThreads) void k_scan_down(Loader ld, size_t n, const uint32_t* partials,
                                                        uint32_t* out, const unsigned long long* n_dev)
{
    __shared__ uint32_t s_wave[kThreads / FGS_WAVE];
    n = live_count(n, n_dev);
    if ((size_t)blockIdx.x * kChunk >= n) return;
    const size_t base = (size_t)blockIdx.x * kChunk + (size_t)threadIdx.x * kIpt;
    uint32_t v[kIpt];
    ld.load(base, n, v);
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < kIpt; ++j) acc += v[j];
    uint32_t total;
    uint32_t run = block_exclusive_scan(acc, s_wave, &total) + partials[blockIdx.x];
    if (base + kIpt <= n && ((uintptr_t)out & 15u) == 0) {
        uint4* o = reinterpret_cast<uint4*>(out + base);
#pragma unroll
        for (int q = 0; q < kIpt / 4; ++q) {
            uint4 x;
            x.x = run;
            run += v[4 * q];
            x.y = run;
            run += v[4 * q + 1];
            x.z = run;
            run += v[4 * q + 2];
            x.w = run;
            run += v[4 * q + 3];
            o[q] = x;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kIpt; ++j) {
            if (base + j < n) out[base + j] = run;
            run += v[j];
        }
    }
}

size_t scan_partials_words(size_t n) { return div_up(n > 0 ? n : 1, kChunk); }

// nz_live: every live value is nonzero, so *nz_count is the live count
// (copied, not counted).
template <typename Loader>
static hipError_t exclusive_scan(Loader ld, size_t n, uint32_t* out, uint32_t* partials,
                                 hipStream_t s, unsigned long long* total = nullptr,
                                 unsigned long long* nz_count = nullptr,
                                 const unsigned long long* n_dev = nullptr, bool nz_live = false)
{
    if (n == 0) return hipSuccess;
    const size_t nb = div_up(n, kChunk);
    hipLaunchKernelGGL(k_scan_reduce<Loader>, dim3((unsigned)nb), dim3(kThreads), 0, s, ld, n, partials,
                       nz_live ? nullptr : nz_count, n_dev);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kThreads), 0, s, partials, nb, total,
                       nz_live ? n_dev : nullptr, nz_live ? nz_count : nullptr);
    hipLaunchKernelGGL(k_scan_down<Loader>, dim3((unsigned)nb), dim3(kThreads), 0, s, ld, n, partials, out,
                       n_dev);
    return hipGetLastError();
}

// Per sorted slot: tiles touched and packed tile rectangle of the Gaussian
// (one 16-byte random read per Gaussian; everything after is coalesced).
__global__ __launch_bounds__(256) void k_sorted_records(const uint64_t* __restrict__ vkeys,
                                                        const uint4* __restrict__ recs, int P,
                                                        uint32_t* __restrict__ tiles_s,
                                                        uint2* __restrict__ rects_s)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P) return;
    const uint4 r = recs[(uint32_t)vkeys[s]];
    tiles_s[s] = r.z;
    rects_s[s] = make_uint2(r.x, r.y);
}

hipError_t launch_sorted_records(const uint64_t* vkeys, const uint4* recs, int P,
                                 uint32_t* tiles_s, uint2* rects_s, hipStream_t s)
{
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sorted_records, dim3((P + 255) / 256), dim3(256), 0, s, vkeys, recs, P,
                       tiles_s, rects_s);
    return hipGetLastError();
}

hipError_t launch_scan_tiles(const uint32_t* tiles_s, uint32_t* offs, int P, uint32_t* partials,
                             unsigned long long* counters, hipStream_t s)
{
    // counters[0] += Gaussians with tiles > 0, counters[1] = sum of tiles (K)
    return exclusive_scan(PlainLoader{ tiles_s }, (size_t)P, offs, partials, s, counters + 1,
                          counters);
}

// ---------------------------------------------------------------------------
// LSD radix sort of u64 keys on 8-bit digits.
// ---------------------------------------------------------------------------
// Adds the 8-bit digits of keys src[0 .. c) to the per-wave LDS histograms.
__device__ __forceinline__ void rs_count_run(const uint64_t* __restrict__ src, size_t c, int shift,
                                             uint32_t (*s_h)[kRadix])
{
    const int wave = threadIdx.x / FGS_WAVE;
    for (size_t e = threadIdx.x; e < c; e += kThreads) {
        const uint32_t d = (uint32_t)(src[e] >> shift) & (kRadix - 1);
        atomicAdd(&s_h[wave][d], 1u);
    }
}

__device__ __forceinline__ void rs_count_store(uint32_t (*s_h)[kRadix], uint32_t* __restrict__ hist, size_t nb)
{
    __syncthreads();
    const uint32_t c = s_h[0][threadIdx.x] + s_h[1][threadIdx.x] + s_h[2][threadIdx.x] + s_h[3][threadIdx.x];
    hist[(size_t)threadIdx.x * nb + blockIdx.x] = c;
}

__global__ __launch_bounds__(kThreads) void k_rs_count(const uint64_t* __restrict__ in, size_t n,
                                                       int shift, uint32_t* __restrict__ hist,
                                                       size_t nb, const unsigned long long* n_dev)
{
    __shared__ uint32_t s_h[kThreads / FGS_WAVE][kRadix];
    for (int w = 0; w < kThreads / FGS_WAVE; ++w) s_h[w][threadIdx.x] = 0;
    __syncthreads();
    n = live_count(n, n_dev);
    const size_t base = (size_t)blockIdx.x * kChunk;
    if (base < n) rs_count_run(in + base, n - base < (size_t)kChunk ? n - base : (size_t)kChunk, shift, s_h);
    rs_count_store(s_h, hist, nb);
}

// The ranked write of one block of up to kChunk u64 keys: key[j] is item j
// of the thread, block position per_wave * wave + 64 j + lane (per_wave a
// multiple of 64, at most 1024; valid below live and inside the wave's
// span), s_glob holds the global start of the block's run of each digit and
// s_cnt zeros.  A wave's span is a run of positions, so ranks follow the
// positions (the sort is stable).  s_keys may hold the items on entry: they
// are read before it is written.
__device__ __forceinline__ void rs_u64_rank_write(const uint64_t* key, uint32_t live, uint32_t per_wave, int shift,
                                                  uint64_t* __restrict__ out, uint32_t (*s_cnt)[kRadix],
                                                  uint32_t* s_start, const uint32_t* s_glob, uint32_t* s_wave,
                                                  uint64_t* s_keys)
{
    constexpr int kWaves = kThreads / FGS_WAVE;
    const int t = threadIdx.x;
    const int wave = t / FGS_WAVE;
    const int lane = lane_id();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (FGS_WAVE - lane));
    const uint32_t wbase = (uint32_t)wave * per_wave;
    uint32_t rank[kIpt];
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
        if ((uint32_t)j * FGS_WAVE >= per_wave) break;   // wave-uniform
        const bool valid = wbase + (uint32_t)j * FGS_WAVE + (uint32_t)lane < live;
        const uint32_t d = (uint32_t)(key[j] >> shift) & (kRadix - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t below = __popcll(peers & lt_mask);
        const uint32_t before = valid ? s_cnt[wave][d] : 0u;
        rank[j] = before + below;
        if (valid && below == 0) s_cnt[wave][d] = before + __popcll(peers);
    }
    __syncthreads();
    {
        uint32_t c[kWaves], tot = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            c[w] = s_cnt[w][t];
            tot += c[w];
        }
        uint32_t all;
        uint32_t run = block_exclusive_scan(tot, s_wave, &all);
        s_start[t] = run;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            s_cnt[w][t] = run;
            run += c[w];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
        if ((uint32_t)j * FGS_WAVE >= per_wave) break;
        if (wbase + (uint32_t)j * FGS_WAVE + (uint32_t)lane < live) {
            const uint32_t d = (uint32_t)(key[j] >> shift) & (kRadix - 1);
            s_keys[s_cnt[wave][d] + rank[j]] = key[j];
        }
    }
    __syncthreads();
    for (uint32_t p = t; p < live; p += kThreads) {
        const uint64_t k = s_keys[p];
        const uint32_t d = (uint32_t)(k >> shift) & (kRadix - 1);
        out[s_glob[d] + (p - s_start[d])] = k;
    }
}

__global__ __launch_bounds__(kThreads) void k_rs_scatter(const uint64_t* __restrict__ in,
                                                         uint64_t* __restrict__ out, size_t n,
                                                         int shift, const uint32_t* __restrict__ offs,
                                                         size_t nb, const unsigned long long* n_dev)
{
    constexpr int kWaves = kThreads / FGS_WAVE;
    constexpr int kPerWave = kChunk / kWaves;
    __shared__ uint32_t s_cnt[kWaves][kRadix];   // per-wave digit counts, then local positions
    __shared__ uint32_t s_start[kRadix];         // local start of each digit's run
    __shared__ uint32_t s_glob[kRadix];          // global start of the block's run of each digit
    __shared__ uint32_t s_wave[kWaves];
    __shared__ uint64_t s_keys[kChunk];
    n = live_count(n, n_dev);
    const size_t base = (size_t)blockIdx.x * kChunk;
    if (base >= n) return;
    const int t = threadIdx.x;
    const int wave = t / FGS_WAVE;
    const int lane = lane_id();
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s_cnt[w][t] = 0;
    s_glob[t] = offs[(size_t)t * nb + blockIdx.x];
    __syncthreads();
    const size_t wbase = base + (size_t)wave * kPerWave;
    uint64_t key[kIpt];
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
        const size_t e = wbase + (size_t)j * FGS_WAVE + lane;
        key[j] = e < n ? in[e] : 0ull;
    }
    const uint32_t live = (uint32_t)(n - base < (size_t)kChunk ? n - base : (size_t)kChunk);
    rs_u64_rank_write(key, live, kPerWave, shift, out, s_cnt, s_start, s_glob, s_wave, s_keys);
}

// Segmented input of the batched preprocess: segment w holds seg_cnt[w]
// keys from w * seg_len; block b owns segments [b * spb, (b + 1) * spb).
__global__ __launch_bounds__(kThreads) void k_rs_count_seg(const uint64_t* __restrict__ in,
                                                           const uint32_t* __restrict__ seg_cnt,
                                                           size_t n_seg, size_t seg_len, int spb,
                                                           int shift, uint32_t* __restrict__ hist,
                                                           size_t nb)
{
    __shared__ uint32_t s_h[kThreads / FGS_WAVE][kRadix];
    for (int w = 0; w < kThreads / FGS_WAVE; ++w) s_h[w][threadIdx.x] = 0;
    __syncthreads();
    for (int q = 0; q < spb; ++q) {
        const size_t w = (size_t)blockIdx.x * spb + q;
        if (w >= n_seg) break;
        const size_t c = seg_cnt[w] < seg_len ? seg_cnt[w] : seg_len;
        rs_count_run(in + w * seg_len, c, shift, s_h);
    }
    rs_count_store(s_h, hist, nb);
}

// The segments' keys are first copied into one dense run in LDS, then
// ranked and written like a dense block (spb * seg_len <= kChunk).
__global__ __launch_bounds__(kThreads) void k_rs_scatter_seg(const uint64_t* __restrict__ in,
                                                             uint64_t* __restrict__ out,
                                                             const uint32_t* __restrict__ seg_cnt,
                                                             size_t n_seg, size_t seg_len, int spb,
                                                             int shift, const uint32_t* __restrict__ offs,
                                                             size_t nb)
{
    constexpr int kWaves = kThreads / FGS_WAVE;
    __shared__ uint32_t s_cnt[kWaves][kRadix];
    __shared__ uint32_t s_start[kRadix];
    __shared__ uint32_t s_glob[kRadix];
    __shared__ uint32_t s_wave[kWaves];
    __shared__ uint32_t s_pre[kThreads + 1];   // the segments' keys before each segment
    __shared__ uint64_t s_keys[kChunk];
    const int t = threadIdx.x;
    const int wave = t / FGS_WAVE;
    const int lane = lane_id();
    const size_t w0 = (size_t)blockIdx.x * spb;
    uint32_t c = 0u;
    if (t < spb && w0 + t < n_seg) c = seg_cnt[w0 + t] < seg_len ? seg_cnt[w0 + t] : (uint32_t)seg_len;
    uint32_t live;
    const uint32_t pre = block_exclusive_scan(c, s_wave, &live);
    s_pre[t] = pre;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s_cnt[w][t] = 0;
    s_glob[t] = offs[(size_t)t * nb + blockIdx.x];
    __syncthreads();
    for (int q = 0; q < spb; ++q) {
        const uint32_t cq = (q + 1 < spb ? s_pre[q + 1] : live) - s_pre[q];
        const uint64_t* src = in + (w0 + q) * seg_len;
        for (uint32_t k = t; k < cq; k += kThreads) s_keys[s_pre[q] + k] = src[k];
    }
    __syncthreads();
    // the waves split the live keys into four runs of whole 64-key steps
    const uint32_t per_wave = (live + kWaves * FGS_WAVE - 1) / (kWaves * FGS_WAVE) * FGS_WAVE;
    const uint32_t wbase = (uint32_t)wave * per_wave;
    uint64_t key[kIpt];
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
        const uint32_t off = (uint32_t)j * FGS_WAVE + (uint32_t)lane;
        key[j] = off < per_wave && wbase + off < live ? s_keys[wbase + off] : 0ull;
    }
    rs_u64_rank_write(key, live, per_wave, shift, out, s_cnt, s_start, s_glob, s_wave, s_keys);
}

size_t radix_hist_words(size_t n) { return (size_t)kRadix * div_up(n > 0 ? n : 1, kChunk); }

// Sorts n keys on n_passes 8-bit digits starting at bit shift_lo; with
// n_dev only the first min(n, *n_dev) (read on the device).  The input is in
// a; after the call the sorted keys are in a if n_passes is even and in b
// otherwise.
hipError_t radix_sort_u64(uint64_t* a, uint64_t* b, size_t n, int shift_lo, int n_passes,
                          uint32_t* hist, uint32_t* partials, hipStream_t s,
                          const unsigned long long* n_dev)
{
    if (n == 0) return hipSuccess;
    const size_t nb = div_up(n, kChunk);
    uint64_t* src = a;
    uint64_t* dst = b;
    for (int p = 0; p < n_passes; ++p) {
        const int shift = shift_lo + 8 * p;
        hipLaunchKernelGGL(k_rs_count, dim3((unsigned)nb), dim3(kThreads), 0, s, src, n, shift, hist, nb, n_dev);
        hipError_t e = exclusive_scan(PlainLoader{ hist }, (size_t)kRadix * nb, hist, partials, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)nb), dim3(kThreads), 0, s, src, dst, n, shift, hist, nb,
                           n_dev);
        uint64_t* t = src;
        src = dst;
        dst = t;
    }
    return hipGetLastError();
}

// Segment length 64 * V (one preprocess wave's keys); spb = kChunk / seg_len
// segments per block (at least one: V <= 64), so a block's keys fit one
// dense block; the first pass takes at most twice the histogram columns of a
// dense pass over V * P keys.
hipError_t radix_sort_u64_segments(uint64_t* a, uint64_t* b, const uint32_t* seg_cnt, int P, int V,
                                   int shift_lo, int n_passes, uint32_t* hist, uint32_t* partials,
                                   unsigned long long* n_dev, hipStream_t s)
{
    const size_t n = (size_t)P * (size_t)V;
    if (n == 0 || n_passes < 1) return hipSuccess;
    const size_t seg_len = (size_t)FGS_WAVE * (size_t)V;
    const size_t n_seg = div_up((size_t)P, FGS_WAVE);
    const int spb = (int)((size_t)kChunk / seg_len);
    const size_t nb_seg = div_up(n_seg, (size_t)spb);
    hipLaunchKernelGGL(k_rs_count_seg, dim3((unsigned)nb_seg), dim3(kThreads), 0, s, a, seg_cnt, n_seg,
                       seg_len, spb, shift_lo, hist, nb_seg);
    hipError_t e = exclusive_scan(PlainLoader{ hist }, (size_t)kRadix * nb_seg, hist, partials, s, n_dev);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rs_scatter_seg, dim3((unsigned)nb_seg), dim3(kThreads), 0, s, a, b, seg_cnt, n_seg,
                       seg_len, spb, shift_lo, hist, nb_seg);
    // the rest: dense over the first *n_dev slots; ping-pong so the result ends in b
    const size_t nb = div_up(n, kChunk);
    uint64_t* src = b;
    uint64_t* dst = a;
    for (int p = 1; p < n_passes; ++p) {
        const int shift = shift_lo + 8 * p;
        hipLaunchKernelGGL(k_rs_count, dim3((unsigned)nb), dim3(kThreads), 0, s, src, n, shift, hist, nb,
                           (const unsigned long long*)n_dev);
        e = exclusive_scan(PlainLoader{ hist }, (size_t)kRadix * nb, hist, partials, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)nb), dim3(kThreads), 0, s, src, dst, n, shift, hist, nb,
                           (const unsigned long long*)n_dev);
        uint64_t* t = src;
        src = dst;
        dst = t;
    }
    if (src != b) return hipErrorInvalidValue;   // needs an odd number of passes
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// duplicateWithKeys in depth order (rasterizer_impl.cu:113-147).
// ---------------------------------------------------------------------------
// Rectangles of up to kDupSolo tiles are written by their own lane; larger
// ones by the whole wave, one at a time, lanes striding over the tiles so the
// stores coalesce and one huge splat does not hold up its wave.
constexpr uint32_t kDupSolo = 16;

// kv pair stores: the plain path's u64 (tile << 32) | val, and a batch's
// two u32 arrays, key (view-local tile << kBatchKeyViewBits) | view and val.
struct PairStore64 {
    uint64_t* pairs;
    __device__ void operator()(size_t p, uint32_t tile, uint32_t tag, uint32_t val) const
    {
        (void)tag;
        pairs[p] = ((uint64_t)tile << 32) | val;
    }
};
struct PairStoreBatch {
    uint32_t* keys;
    uint32_t* vals;
    __device__ void operator()(size_t p, uint32_t tile, uint32_t view, uint32_t val) const
    {
        keys[p] = (tile << kBatchKeyViewBits) | view;
        vals[p] = val;
    }
};

// Writes the kv pairs of tiles tile_base + y * gx + x of rectangle rc
// (packed x0 | y0 << 16, x1 | y1 << 16) at off through st (with tag, the
// batch's view); every argument is the calling lane's own.  Every lane of
// the wave calls it (valid = false for lanes without an element); pairs at
// or past cap are dropped.
template <typename Store>
__device__ __forceinline__ void dup_rect(bool valid, uint2 rc, size_t off, uint32_t val, uint32_t gx,
                                         uint32_t tile_base, uint32_t tag, Store st, size_t cap)
{
    const uint32_t x0 = rc.x & 0xFFFFu, y0 = rc.x >> 16, x1 = rc.y & 0xFFFFu, y1 = rc.y >> 16;
    const uint32_t w = (valid && x1 > x0 && y1 > y0) ? x1 - x0 : 0u;
    const uint32_t n = w ? w * (y1 - y0) : 0u;
    if (n > 0 && n <= kDupSolo) {
        for (uint32_t y = y0; y < y1; ++y)
            for (uint32_t x = x0; x < x1; ++x, ++off)
                if (off < cap) st(off, tile_base + y * gx + x, tag, val);
    }
    uint64_t big = __ballot(n > kDupSolo);
    const uint32_t lane = (uint32_t)lane_id();
    while (big) {
        const int src = __ffsll((long long)big) - 1;
        big &= big - 1;
        const uint32_t bx0 = __builtin_amdgcn_readlane(x0, src);
        const uint32_t by0 = __builtin_amdgcn_readlane(y0, src);
        const uint32_t bw = __builtin_amdgcn_readlane(w, src);
        const uint32_t bn = __builtin_amdgcn_readlane(n, src);
        const uint32_t bval = __builtin_amdgcn_readlane(val, src);
        const uint32_t bgx = __builtin_amdgcn_readlane(gx, src);
        const uint32_t btb = __builtin_amdgcn_readlane(tile_base, src);
        const uint32_t btag = __builtin_amdgcn_readlane(tag, src);
        const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)off, src);
        const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(off >> 32), src);
        const size_t boff = ((size_t)hi << 32) | lo;
        for (uint32_t k = lane; k < bn; k += FGS_WAVE) {
            const uint32_t dy = k / bw;
            const uint32_t x = bx0 + (k - dy * bw), y = by0 + dy;
            if (boff + k < cap) st(boff + k, btb + y * bgx + x, btag, bval);
        }
    }
}

__global__ __launch_bounds__(256) void k_duplicate(const uint64_t* __restrict__ vkeys,
                                                   const uint32_t* __restrict__ offs,
                                                   const uint2* __restrict__ rects_s, int P, int gx,
                                                   uint64_t* __restrict__ pairs, uint4* __restrict__ recs)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = s < P;
    const uint2 rc = valid ? rects_s[s] : make_uint2(0u, 0u);
    const uint32_t g = valid ? (uint32_t)vkeys[s] : 0u;
    const size_t off = valid ? offs[s] : 0u;
    if (valid) recs[g].w = (uint32_t)off;
    dup_rect(valid, rc, off, g, (uint32_t)gx, 0u, 0u, PairStore64{ pairs }, ~(size_t)0);
}

hipError_t launch_duplicate(const uint64_t* vkeys, const uint32_t* offs, const uint2* rects_s,
                            int P, int gx, uint64_t* pairs, uint4* recs, hipStream_t s)
{
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(k_duplicate, dim3((P + 255) / 256), dim3(256), 0, s, vkeys, offs, rects_s, P,
                       gx, pairs, recs);
    return hipGetLastError();
}

// identifyTileRanges (rasterizer_impl.cu:172-194); ranges zeroed beforehand.
__global__ __launch_bounds__(256) void k_ranges(const uint64_t* __restrict__ sorted, size_t K,
                                                uint2* __restrict__ ranges)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const uint32_t t = (uint32_t)(sorted[k] >> 32);
    if (k == 0 || (uint32_t)(sorted[k - 1] >> 32) != t) ranges[t].x = (uint32_t)k;
    if (k == K - 1 || (uint32_t)(sorted[k + 1] >> 32) != t) ranges[t].y = (uint32_t)(k + 1);
}

hipError_t launch_ranges(const uint64_t* sorted, size_t K, uint2* ranges, int T, hipStream_t s)
{
    hipError_t e = launch_fill_u32(ranges, 0u, (size_t)T * 2, s);
    if (e != hipSuccess || K == 0) return e;
    hipLaunchKernelGGL(k_ranges, dim3((unsigned)div_up(K, 256)), dim3(256), 0, s, sorted, K, ranges);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched binning.  Depth keys of a batch are (v << 58) | (depth_bits << 26) | j
// (element batch_elem; j orders a view's Gaussians like their index), so one
// sort orders every view by depth with the views in sequence.  kv pairs carry
// the view's own tile number and the element: a stable sort on the tile bits
// alone then leaves, for every tile number t, the pairs of view 0, view 1,
// ... one after the other, each in depth order, so (view, t) ranges stay
// contiguous and the sort needs only the bits of the largest view's tile
// count.
// ---------------------------------------------------------------------------
constexpr int kRecPerThread = 4;
__global__ __launch_bounds__(256) void k_sorted_records_batch(const uint64_t* __restrict__ vkeys,
                                                              const uint4* __restrict__ recs, size_t n,
                                                              const unsigned long long* __restrict__ n_dev,
                                                              int P, const BatchView* __restrict__ views,
                                                              uint32_t* __restrict__ tiles_s,
                                                              uint2* __restrict__ rects_s)
{
    // kRecPerThread keys per thread, a block's 256 apart (coalesced key loads
    // and stores): their record gathers are independent, so each wave has
    // that many random loads in flight instead of one
    const size_t live = live_count(n, n_dev);
    const size_t s0 = (size_t)blockIdx.x * (kRecPerThread * 256) + threadIdx.x;
    if (s0 >= live) return;
    const uint32_t P64 = batch_p64(P);
    uint64_t key[kRecPerThread];
    uint4 r[kRecPerThread];
#pragma unroll
    for (int i = 0; i < kRecPerThread; ++i) {
        const size_t s = s0 + (size_t)i * 256;
        key[i] = s < live ? vkeys[s] : 0ull;   // only visible (view, Gaussian) keys reach the sort
    }
#pragma unroll
    for (int i = 0; i < kRecPerThread; ++i)
        if (s0 + (size_t)i * 256 < live) r[i] = recs[batch_elem(key[i], P64)];
#pragma unroll
    for (int i = 0; i < kRecPerThread; ++i) {
        const size_t s = s0 + (size_t)i * 256;
        if (s >= live) break;
        const uint32_t gx = (uint32_t)views[(uint32_t)(key[i] >> kBatchViewShift)].gx;
        const uint32_t x0 = r[i].x & 0xFFFFu, y0 = r[i].x >> 16;
        tiles_s[s] = r[i].z;
        // the rectangle as its first tile and row width | view row stride << 16
        rects_s[s] = make_uint2(y0 * gx + x0, ((r[i].y & 0xFFFFu) - x0) | (gx << 16));
    }
}

hipError_t launch_sorted_records_batch(const uint64_t* vkeys, const uint4* recs, size_t n,
                                       const unsigned long long* n_dev, int P, const BatchView* views,
                                       uint32_t* tiles_s, uint2* rects_s, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sorted_records_batch, dim3((unsigned)div_up(n, (size_t)kRecPerThread * 256)), dim3(256), 0, s,
                       vkeys,
                       recs, n, n_dev, P, views, tiles_s, rects_s);
    return hipGetLastError();
}

hipError_t launch_scan_u32(const uint32_t* in, size_t n, uint32_t* out, uint32_t* partials, hipStream_t s)
{
    return exclusive_scan(PlainLoader{ in }, n, out, partials, s);
}

struct PopLoader {
    const uint64_t* masks;
    __device__ void load(size_t base, size_t n, uint32_t* v) const
    {
#pragma unroll
        for (int j = 0; j < kIpt; ++j) v[j] = base + j < n ? (uint32_t)__popcll(masks[base + j]) : 0u;
    }
};

// One wave per 64 groups: lane l loads group l's mask and start, then the
// wave writes the groups' visible elements one group after the other.
__global__ __launch_bounds__(256) void k_visible_list(const uint64_t* __restrict__ vis_masks, size_t n_groups,
                                                      const uint32_t* __restrict__ grp_start,
                                                      uint32_t* __restrict__ list)
{
    const size_t g0 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) & ~(size_t)(FGS_WAVE - 1);
    if (g0 >= n_groups) return;
    const uint32_t lane = (uint32_t)lane_id();
    const size_t gl = g0 + lane;
    const uint32_t cnt = gl < n_groups ? (uint32_t)__popcll(vis_masks[gl]) : 0u;
    const uint32_t start = gl < n_groups ? grp_start[gl] : 0u;
    const uint32_t ng = n_groups - g0 < (size_t)FGS_WAVE ? (uint32_t)(n_groups - g0) : (uint32_t)FGS_WAVE;
    for (uint32_t k = 0; k < ng; ++k) {
        const uint32_t c = __builtin_amdgcn_readlane(cnt, k);
        if (lane < c) list[__builtin_amdgcn_readlane(start, k) + lane] = (uint32_t)((g0 + k) * FGS_WAVE) + lane;
    }
}

hipError_t launch_visible_list(const uint64_t* vis_masks, size_t n_groups, uint32_t* grp_start, uint32_t* list,
                               uint32_t* partials, hipStream_t s)
{
    if (n_groups == 0) return hipSuccess;
    hipError_t e = exclusive_scan(PopLoader{ vis_masks }, n_groups, grp_start, partials, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_visible_list, dim3((unsigned)div_up(n_groups, 256)), dim3(256), 0, s, vis_masks, n_groups,
                       grp_start, list);
    return hipGetLastError();
}

hipError_t launch_scan_tiles_batch(const uint32_t* tiles_s, uint32_t* offs, size_t n,
                                   const unsigned long long* n_dev, uint32_t* partials,
                                   unsigned long long* counters, hipStream_t s)
{
    // the slots hold the visible keys only: every one has tiles
    return exclusive_scan(PlainLoader{ tiles_s }, n, offs, partials, s, counters + 1, counters, n_dev,
                          n_dev != nullptr);
}

// Block-wide exclusive maximum of one u32 per thread (256 threads = 4 waves).
__device__ __forceinline__ uint32_t block_exclusive_max(uint32_t v, uint32_t* s_wave)
{
    const int lane = lane_id();
    const int wave = threadIdx.x / FGS_WAVE;
    const uint32_t incl = wave_incl_max_u32(v);
    uint32_t excl = __shfl_up(incl, 1, FGS_WAVE);
    if (lane == 0) excl = 0u;
    if (lane == FGS_WAVE - 1) s_wave[wave] = incl;
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kThreads / FGS_WAVE; ++w)
        if (w < wave) excl = max(excl, s_wave[w]);
    __syncthreads();
    return excl;
}

// The batched duplication works in chunks of kDupChunk output positions.
// k_dup_chunk_owners records, per chunk, the element (depth-order slot)
// whose pairs hold the chunk's first position; the elements' pair runs
// [offs[s], offs[s] + tiles_s[s]) tile [0, K) without gaps.
constexpr uint32_t kDupChunk = 4096;
constexpr int kDupPer = kDupChunk / 256;   // positions per thread
static_assert(kDupChunk == (uint32_t)kChunk, "a duplication chunk is one sort block");
// The owner table in LDS holds position p at own_slot(p): one pad word per
// 16 positions, so the scan's threads (16 consecutive positions each) read
// distinct banks.
constexpr uint32_t kDupOwnWords = kDupChunk + kDupChunk / kDupPer;
__device__ __forceinline__ uint32_t own_slot(uint32_t p) { return p + p / kDupPer; }
static_assert(kDupOwnWords <= 2u * (uint32_t)kChunk, "the owner table fits the sort block's LDS");

__global__ __launch_bounds__(256) void k_dup_chunk_owners(const uint32_t* __restrict__ offs,
                                                          const uint32_t* __restrict__ tiles_s, size_t n,
                                                          const unsigned long long* __restrict__ n_dev, size_t cap,
                                                          uint32_t* __restrict__ chunk_first)
{
    const size_t live = live_count(n, n_dev);
    for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < live; s += (size_t)gridDim.x * blockDim.x) {
        const uint64_t off = offs[s], end = off + tiles_s[s];
        for (uint64_t c = (off + kDupChunk - 1) / kDupChunk; c * kDupChunk < end && c * kDupChunk < cap; ++c)
            chunk_first[c] = (uint32_t)s;
    }
}

// ---------------------------------------------------------------------------
// LSD radix sort of (u32 key, u32 value) pairs on 8-bit digits (the batched
// tile sort): the passes of radix_sort_u64 with the keys and values in two
// arrays, so the counting pass reads 4 bytes per pair.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_rs_count32(const uint32_t* __restrict__ in, size_t n, int shift,
                                                         int db, uint32_t* __restrict__ hist, size_t nb,
                                                         const unsigned long long* n_dev)
{
    const uint32_t dm = (1u << db) - 1u;
    __shared__ uint32_t s_h[kThreads / FGS_WAVE][kRadix];
    for (int w = 0; w < kThreads / FGS_WAVE; ++w) s_h[w][threadIdx.x] = 0;
    __syncthreads();
    n = live_count(n, n_dev);
    const size_t base = (size_t)blockIdx.x * kChunk;
    if (base < n) {
        const size_t c = n - base < (size_t)kChunk ? n - base : (size_t)kChunk;
        const int wave = threadIdx.x / FGS_WAVE;
        const uint32_t* src = in + base;
        // four keys per 16-byte load (base is a multiple of kChunk)
        const size_t c4 = c / 4;
        for (size_t q = threadIdx.x; q < c4; q += kThreads) {
            const uint4 k4 = reinterpret_cast<const uint4*>(src)[q];
            atomicAdd(&s_h[wave][(k4.x >> shift) & dm], 1u);
            atomicAdd(&s_h[wave][(k4.y >> shift) & dm], 1u);
            atomicAdd(&s_h[wave][(k4.z >> shift) & dm], 1u);
            atomicAdd(&s_h[wave][(k4.w >> shift) & dm], 1u);
        }
        for (size_t e = 4 * c4 + threadIdx.x; e < c; e += kThreads)
            atomicAdd(&s_h[wave][(src[e] >> shift) & dm], 1u);
    }
    rs_count_store(s_h, hist, nb);
}

// The ranked write of one 4096-pair block of a (key, value) pass (the
// dense scatter of k_rs_scatter): key[j] / val[j] are item j of the thread
// (block position 1024 * wave + 64 j + lane; valid below live), s_glob holds
// the global start of the block's run of each digit and s_cnt zeros.
__device__ __forceinline__ void rs_kv32_rank_write(const uint32_t* key, const uint32_t* val, uint32_t live,
                                                   int shift, int db, uint32_t* __restrict__ kout,
                                                   uint32_t* __restrict__ vout,
                                                   uint32_t (*s_cnt)[kRadix], uint32_t* s_start,
                                                   const uint32_t* s_glob, uint32_t* s_wave, uint32_t* s_keys,
                                                   uint32_t* s_vals)
{
    constexpr int kWaves = kThreads / FGS_WAVE;
    constexpr int kPerWave = kChunk / kWaves;
    const int t = threadIdx.x;
    const int wave = t / FGS_WAVE;
    const int lane = lane_id();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (FGS_WAVE - lane));
    const uint32_t wbase = (uint32_t)wave * kPerWave;
    const uint32_t dm = (1u << db) - 1u;
    uint32_t rank[kIpt];
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
        const bool valid = wbase + (uint32_t)j * FGS_WAVE + (uint32_t)lane < live;
        const uint32_t d = (key[j] >> shift) & dm;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            if (b >= db) break;
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t below = __popcll(peers & lt_mask);
        const uint32_t before = valid ? s_cnt[wave][d] : 0u;
        rank[j] = before + below;
        if (valid && below == 0) s_cnt[wave][d] = before + __popcll(peers);
    }
    __syncthreads();
    {
        uint32_t c[kWaves], tot = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            c[w] = s_cnt[w][t];
            tot += c[w];
        }
        uint32_t all;
        uint32_t run = block_exclusive_scan(tot, s_wave, &all);
        s_start[t] = run;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            s_cnt[w][t] = run;
            run += c[w];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
        if (wbase + (uint32_t)j * FGS_WAVE + (uint32_t)lane < live) {
            const uint32_t d = (key[j] >> shift) & dm;
            const uint32_t q = s_cnt[wave][d] + rank[j];
            s_keys[q] = key[j];
            s_vals[q] = val[j];
        }
    }
    __syncthreads();
    for (uint32_t p = t; p < live; p += kThreads) {
        const uint32_t k = s_keys[p];
        const uint32_t d = (k >> shift) & dm;
        const size_t o = s_glob[d] + (p - s_start[d]);
        kout[o] = k;
        vout[o] = s_vals[p];
    }
}

__global__ __launch_bounds__(kThreads) void k_rs_scatter_kv32(const uint32_t* __restrict__ kin,
                                                              const uint32_t* __restrict__ vin,
                                                              uint32_t* __restrict__ kout,
                                                              uint32_t* __restrict__ vout, size_t n, int shift,
                                                              int db, const uint32_t* __restrict__ offs, size_t nb,
                                                              const unsigned long long* n_dev)
{
    constexpr int kWaves = kThreads / FGS_WAVE;
    constexpr int kPerWave = kChunk / kWaves;
    __shared__ uint32_t s_cnt[kWaves][kRadix];
    __shared__ uint32_t s_start[kRadix];
    __shared__ uint32_t s_glob[kRadix];
    __shared__ uint32_t s_wave[kWaves];
    __shared__ uint32_t s_keys[kChunk];
    __shared__ uint32_t s_vals[kChunk];
    n = live_count(n, n_dev);
    const size_t base = (size_t)blockIdx.x * kChunk;
    if (base >= n) return;
    const int t = threadIdx.x;
    const int wave = t / FGS_WAVE;
    const int lane = lane_id();
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s_cnt[w][t] = 0;
    s_glob[t] = offs[(size_t)t * nb + blockIdx.x];
    __syncthreads();
    const size_t wbase = base + (size_t)wave * kPerWave;
    uint32_t key[kIpt], val[kIpt];
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
        const size_t e = wbase + (size_t)j * FGS_WAVE + lane;
        key[j] = e < n ? kin[e] : 0u;
        val[j] = e < n ? vin[e] : 0u;
    }
    const uint32_t live = (uint32_t)(n - base < (size_t)kChunk ? n - base : (size_t)kChunk);
    rs_kv32_rank_write(key, val, live, shift, db, kout, vout, s_cnt, s_start, s_glob, s_wave, s_keys, s_vals);
}

// Pairs past the capacity are dropped (the caller compares K with it).
// Chunk b holds output positions [b * kDupChunk, (b + 1) * kDupChunk): the
// elements starting inside the chunk mark their first position in LDS and
// a block maximum-scan turns the marks into the owner of every position
// (s_own[p] = owner - chunk_first[b]).  The tile of a position is then the
// pair's index in its owner's row-major rectangle, so no lane walks a
// rectangle of its own and large rectangles do not hold up a wave.
struct DupChunk {
    size_t base, K;
    uint32_t s_lo;
};

__device__ __forceinline__ DupChunk dup_chunk_owners(const uint32_t* __restrict__ offs, size_t n,
                                                     const unsigned long long* __restrict__ n_dev, size_t K,
                                                     const uint32_t* __restrict__ chunk_first, uint32_t* s_own,
                                                     uint32_t* s_wave)
{
    const int t = threadIdx.x;
    DupChunk c;
    c.base = (size_t)blockIdx.x * kDupChunk;
    c.K = K;
    c.s_lo = chunk_first[blockIdx.x];
    // the elements overlapping the chunk: s_lo up to the next chunk's owner
    const size_t s_end = c.base + kDupChunk < K ? (size_t)chunk_first[blockIdx.x + 1] + 1 : live_count(n, n_dev);
#pragma unroll
    for (int q = 0; q < kDupPer; ++q) s_own[own_slot(t + 256 * q)] = 0u;
    __syncthreads();
    // two elements per thread in flight (a chunk overlaps a few hundred)
    for (size_t s = (size_t)c.s_lo + (size_t)t; s < s_end; s += 512) {
        const bool has2 = s + 256 < s_end;
        const uint32_t off0 = offs[s];
        const uint32_t off1 = has2 ? offs[s + 256] : 0u;
        if ((size_t)off0 >= c.base && (size_t)off0 < c.base + kDupChunk && s > c.s_lo)
            s_own[own_slot(off0 - (uint32_t)c.base)] = (uint32_t)(s - c.s_lo);
        if (has2 && (size_t)off1 >= c.base && (size_t)off1 < c.base + kDupChunk)
            s_own[own_slot(off1 - (uint32_t)c.base)] = (uint32_t)(s + 256 - c.s_lo);
    }
    __syncthreads();
    // inclusive maximum over the chunk (owners grow with the position)
    uint32_t m = 0u;
#pragma unroll
    for (int i = 0; i < kDupPer; ++i) {
        m = max(m, s_own[(kDupPer + 1) * t + i]);
        s_own[(kDupPer + 1) * t + i] = m;
    }
    const uint32_t carry = block_exclusive_max(m, s_wave);
#pragma unroll
    for (int i = 0; i < kDupPer; ++i) s_own[(kDupPer + 1) * t + i] = max(s_own[(kDupPer + 1) * t + i], carry);
    __syncthreads();
    return c;
}

// q = k / w and r = k % w for w >= 1, k < 2^26 and q < 2^20: the float
// estimate is within 3 of q (float(k) within 2 of k, the reciprocal within
// an ulp) and the fix loops step it to q.
__device__ __forceinline__ uint32_t div_rows(uint32_t k, uint32_t w, uint32_t& r)
{
    uint32_t q = (uint32_t)((float)k * __builtin_amdgcn_rcpf((float)w));
    int32_t rr = (int32_t)k - (int32_t)(q * w);
    while (rr < 0) {
        --q;
        rr += (int32_t)w;
    }
    while (rr >= (int32_t)w) {
        ++q;
        rr -= (int32_t)w;
    }
    r = (uint32_t)rr;
    return q;
}

// The (tile << kBatchKeyViewBits | view, element) pairs at chunk positions
// p0 + i * stride, i < N; the owner's rectangle is (first tile, width | row
// stride << 16).  All owner reads, then all record loads, are issued before
// any is used.  Positions past K hold the chunk's last owner (the scan
// carries it), so their loads stay in bounds; their pairs are not used.
template <int N>
__device__ __forceinline__ void dup_pairs(const DupChunk& c, uint32_t p0, uint32_t stride, const uint32_t* s_own,
                                          const uint64_t* __restrict__ vkeys, const uint32_t* __restrict__ offs,
                                          const uint2* __restrict__ rects_s, uint32_t P64, uint32_t* key,
                                          uint32_t* val)
{
    uint32_t own[N];
#pragma unroll
    for (int i = 0; i < N; ++i) own[i] = s_own[own_slot(p0 + (uint32_t)i * stride)];
    uint2 rc[N];
    uint64_t vk[N];
    uint32_t of[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const size_t s = (size_t)c.s_lo + own[i];
        rc[i] = rects_s[s];
        vk[i] = vkeys[s];
        of[i] = offs[s];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const uint32_t k = (uint32_t)(c.base + p0 + (uint32_t)i * stride - of[i]);
        uint32_t dx;
        const uint32_t dy = div_rows(k, max(rc[i].y & 0xFFFFu, 1u), dx);
        const uint32_t tile = rc[i].x + dy * (rc[i].y >> 16) + dx;
        key[i] = (tile << kBatchKeyViewBits) | (uint32_t)(vk[i] >> kBatchViewShift);
        val[i] = batch_elem(vk[i], P64);
    }
}

// Duplication fused with the first tile-sort pass.  Counting half: the
// chunk's pairs are generated in registers and their digits counted into
// the pass histogram (chunks past K count nothing).
__global__ __launch_bounds__(256) void k_dup_count_batch(const uint64_t* __restrict__ vkeys,
                                                         const uint32_t* __restrict__ offs,
                                                         const uint2* __restrict__ rects_s, size_t n,
                                                         const unsigned long long* __restrict__ n_dev,
                                                         const unsigned long long* __restrict__ K_dev,
                                                         const uint32_t* __restrict__ chunk_first, int P,
                                                         size_t cap, int shift, int db,
                                                         uint32_t* __restrict__ hist, size_t nb)
{
    __shared__ uint32_t s_h[kThreads / FGS_WAVE][kRadix];
    __shared__ uint32_t s_own[kDupOwnWords];
    __shared__ uint32_t s_wave[kThreads / FGS_WAVE];
    const int t = threadIdx.x;
    const int wave = t / FGS_WAVE;
#pragma unroll
    for (int w = 0; w < kThreads / FGS_WAVE; ++w) s_h[w][t] = 0;
    const size_t K = live_count(cap, K_dev);
    if ((size_t)blockIdx.x * kDupChunk < K) {
        const uint32_t P64 = batch_p64(P);
        const DupChunk c = dup_chunk_owners(offs, n, n_dev, K, chunk_first, s_own, s_wave);
        constexpr int kRound = 8;
#pragma unroll
        for (int r = 0; r < kDupPer / kRound; ++r) {
            const uint32_t p0 = (uint32_t)t + 256u * kRound * (uint32_t)r;
            uint32_t key[kRound], val[kRound];
            dup_pairs<kRound>(c, p0, 256u, s_own, vkeys, offs, rects_s, P64, key, val);
#pragma unroll
            for (int i = 0; i < kRound; ++i)
                if (c.base + p0 + 256u * (uint32_t)i < K)
                    atomicAdd(&s_h[wave][(key[i] >> shift) & ((1u << db) - 1u)], 1u);
        }
    }
    rs_count_store(s_h, hist, nb);
}

// Scattering half: the chunk's pairs are generated again, in the item
// layout of k_rs_scatter_kv32, and written ranked by digit.  The owner
// table shares its LDS with the ranked keys.
__global__ __launch_bounds__(256) void k_dup_scatter_batch(const uint64_t* __restrict__ vkeys,
                                                           const uint32_t* __restrict__ offs,
                                                           const uint2* __restrict__ rects_s, size_t n,
                                                           const unsigned long long* __restrict__ n_dev,
                                                           const unsigned long long* __restrict__ K_dev,
                                                           const uint32_t* __restrict__ chunk_first, int P,
                                                           size_t cap, int shift, int db, const uint32_t* __restrict__ hist, size_t nb,
                                                           uint32_t* __restrict__ kout, uint32_t* __restrict__ vout)
{
    constexpr int kWaves = kThreads / FGS_WAVE;
    constexpr int kPerWave = kChunk / kWaves;
    __shared__ uint32_t s_cnt[kWaves][kRadix];
    __shared__ uint32_t s_start[kRadix];
    __shared__ uint32_t s_glob[kRadix];
    __shared__ uint32_t s_wave[kWaves];
    __shared__ uint32_t s_kv[2 * kChunk];   // ranked keys, then values; first the owner table
    const size_t K = live_count(cap, K_dev);
    if ((size_t)blockIdx.x * kDupChunk >= K) return;
    const int t = threadIdx.x;
    const int wave = t / FGS_WAVE;
    const int lane = lane_id();
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s_cnt[w][t] = 0;
    s_glob[t] = hist[(size_t)t * nb + blockIdx.x];
    const uint32_t P64 = batch_p64(P);
    const DupChunk c = dup_chunk_owners(offs, n, n_dev, K, chunk_first, s_kv, s_wave);
    const uint32_t live = (uint32_t)(K - c.base < (size_t)kChunk ? K - c.base : (size_t)kChunk);
    // item j of the thread is position 1024 * wave + 64 j + lane
    uint32_t key[kIpt], val[kIpt];
    constexpr int kRound = 8;
#pragma unroll
    for (int r = 0; r < kIpt / kRound; ++r)
        dup_pairs<kRound>(c, (uint32_t)wave * kPerWave + (uint32_t)lane + FGS_WAVE * kRound * (uint32_t)r, FGS_WAVE,
                          s_kv, vkeys, offs, rects_s, P64, key + kRound * r, val + kRound * r);
    rs_kv32_rank_write(key, val, live, shift, db, kout, vout, s_cnt, s_start, s_glob, s_wave, s_kv, s_kv + kChunk);
}

hipError_t radix_sort_kv32(uint32_t* ka, uint32_t* va, uint32_t* kb, uint32_t* vb, size_t n, int shift_lo,
                           int n_passes, int db, uint32_t* hist, uint32_t* partials, hipStream_t s,
                           const unsigned long long* n_dev)
{
    if (n == 0) return hipSuccess;
    const size_t nb = div_up(n, kChunk);
    uint32_t *ks = ka, *vs = va, *kd = kb, *vd = vb;
    for (int p = 0; p < n_passes; ++p) {
        const int shift = shift_lo + db * p;
        hipLaunchKernelGGL(k_rs_count32, dim3((unsigned)nb), dim3(kThreads), 0, s, ks, n, shift, db, hist, nb, n_dev);
        hipError_t e = exclusive_scan(PlainLoader{ hist }, (size_t)kRadix * nb, hist, partials, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_rs_scatter_kv32, dim3((unsigned)nb), dim3(kThreads), 0, s, ks, vs, kd, vd, n, shift,
                           db, hist, nb, n_dev);
        uint32_t* t = ks;
        ks = kd;
        kd = t;
        t = vs;
        vs = vd;
        vd = t;
    }
    return hipGetLastError();
}

hipError_t launch_duplicate_sort_batch(const uint64_t* vkeys, const uint32_t* offs, const uint2* rects_s,
                                       const uint32_t* tiles_s, size_t n, const unsigned long long* n_dev,
                                       const unsigned long long* K_dev, int P, size_t cap, uint32_t* chunk_first,
                                       int shift, int db, uint32_t* keys, uint32_t* vals, uint32_t* hist,
                                       uint32_t* partials, hipStream_t s)
{
    if (cap == 0 || n == 0) return hipSuccess;
    size_t blocks = div_up(n, 256);
    if (blocks > kStrideBlocks) blocks = kStrideBlocks;
    hipLaunchKernelGGL(k_dup_chunk_owners, dim3((unsigned)blocks), dim3(256), 0, s, offs, tiles_s, n, n_dev, cap,
                       chunk_first);
    const size_t nb = div_up(cap, kChunk);
    hipLaunchKernelGGL(k_dup_count_batch, dim3((unsigned)nb), dim3(kThreads), 0, s, vkeys, offs, rects_s, n, n_dev,
                       K_dev, chunk_first, P, cap, shift, db, hist, nb);
    hipError_t e = exclusive_scan(PlainLoader{ hist }, (size_t)kRadix * nb, hist, partials, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_dup_scatter_batch, dim3((unsigned)nb), dim3(kThreads), 0, s, vkeys, offs, rects_s, n,
                       n_dev, K_dev, chunk_first, P, cap, shift, db, hist, nb, keys, vals);
    return hipGetLastError();
}

// identifyTileRanges over the batch: a (tile, view) run is a run of equal
// tile keys; it is stored at the view's tile_base + tile.  Only the first
// min(cap, *K) sorted pairs exist.  Each thread takes four consecutive keys
// (one 16-byte load) and their two neighbours.
constexpr int kRangeKeys = 8;   // a multiple of 4
__global__ __launch_bounds__(256) void k_ranges_batch(const uint32_t* __restrict__ sorted, size_t cap,
                                                      const unsigned long long* __restrict__ K,
                                                      const BatchView* __restrict__ views,
                                                      uint2* __restrict__ ranges)
{
    const size_t n = live_count(cap, K);
    const size_t k0 = (size_t)kRangeKeys * ((size_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (k0 >= n) return;
    // kRangeKeys keys per thread in 16-byte loads issued together (more bytes
    // in flight per wave)
    uint32_t c[kRangeKeys];
    if (k0 + kRangeKeys <= n) {
        uint4 q[kRangeKeys / 4];
#pragma unroll
        for (int h = 0; h < kRangeKeys / 4; ++h) q[h] = reinterpret_cast<const uint4*>(sorted + k0)[h];
#pragma unroll
        for (int h = 0; h < kRangeKeys / 4; ++h) {
            c[4 * h] = q[h].x;
            c[4 * h + 1] = q[h].y;
            c[4 * h + 2] = q[h].z;
            c[4 * h + 3] = q[h].w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < kRangeKeys; ++i) c[i] = k0 + i < n ? sorted[k0 + i] : 0u;
    }
    uint32_t prev = k0 > 0 ? sorted[k0 - 1] : 0u;
#pragma unroll
    for (int i = 0; i < kRangeKeys; ++i) {
        const size_t k = k0 + i;
        if (k >= n) break;
        const uint32_t key = c[i];
        const uint32_t next = i < kRangeKeys - 1 ? c[i + 1] : (k + 1 < n ? sorted[k + 1] : 0u);
        const bool first = k == 0 || prev != key;
        const bool last = k == n - 1 || next != key;
        prev = key;
        if (!first && !last) continue;
        const uint32_t v = key & ((1u << kBatchKeyViewBits) - 1u);
        const uint32_t gt = views[v].tile_base + (key >> kBatchKeyViewBits);
        if (first) ranges[gt].x = (uint32_t)k;
        if (last) ranges[gt].y = (uint32_t)(k + 1);
    }
}

hipError_t launch_ranges_batch(const uint32_t* sorted_keys, size_t cap, const unsigned long long* K,
                               const BatchView* views, uint2* ranges, int T, hipStream_t s)
{
    hipError_t e = launch_fill_u32(ranges, 0u, (size_t)T * 2, s);
    if (e != hipSuccess || cap == 0) return e;
    hipLaunchKernelGGL(k_ranges_batch, dim3((unsigned)div_up(cap, (size_t)kRangeKeys * 256)), dim3(256), 0, s,
                       sorted_keys, cap, K,
                       views, ranges);
    return hipGetLastError();
}

}  // namespace fgs
