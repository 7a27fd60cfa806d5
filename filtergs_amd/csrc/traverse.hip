// traverse.hip -- TensorTree.traverse of LoG (LoG/model/tensor_tree.py:133-188)
// as device work.  One level of the breadth-first cut is three launches on a
// fixed grid that read the level's size from device memory, so the host
// never waits between levels; the only synchronisation is the final read of
// the result size.
//
// Level 0 tests the given roots.  Level l >= 1 expands the frontier (the
// nodes that went on at level l-1) into its children in (frontier position,
// child slot) order, skipping empty slots (-1).  A candidate is kept (goes to
// the output) when its screen radius is below min_resolution_pixel or it has
// no children (node_index == -1); otherwise it joins the next frontier.  At
// the level limit (min(30, max_depth) in the reference) the remaining
// frontier itself is appended.  Stable compaction keeps the reference's
// order: kept roots, then the kept children of level 1, level 2, ..., then
// the frontier at the limit.
//
// The radius is LoG's Gaussian.compute_radius (level_of_gaussian.py:66-88):
// scaling = exp(raw), rotation = raw / max(|raw|, 1e-12) (F.normalize), then
// compute_radius_cuda (LoG/cuda/compute_radius_kernel.cu:107-156).  exp is
// rounded through double and the norm summed in index order, compiled without
// contraction, so the kept set is reproducible bit for bit.
//
// Per level and candidate: 4 B frontier id + 4 B node_index + 4 B child id +
// 4 B node_index of the child + 40 B parameters + 1 B flag written and read
// back + 4 B output id; 64 B of HBM traffic at most.
#include "fgs_device.h"
#include "fgs_internal.h"

namespace fgs {

namespace {

struct TravLevel {
    const int64_t* roots;      // level 0: candidate j is roots[j]
    int32_t n_roots;
    const int32_t* frontier;   // level >= 1: candidates are children of frontier[j / C]
    const int32_t* n_frontier;
    int32_t* next;
    int32_t* n_next;
    int32_t* out;
    int32_t* n_out;
};

__device__ __forceinline__ int64_t trav_size(const TravArgs& a, const TravLevel& L)
{
    return L.roots ? (int64_t)L.n_roots : (int64_t)(*L.n_frontier) * a.C;
}

// Range of candidates block b owns: contiguous, a multiple of the block size.
__device__ __forceinline__ void trav_chunk(int64_t m, int64_t& beg, int64_t& end)
{
    const int64_t per = (int64_t)kTravGrid * kTravBlock;
    const int64_t chunk = (m + per - 1) / per * kTravBlock;
    beg = (int64_t)blockIdx.x * chunk;
    end = beg + chunk < m ? beg + chunk : m;
}

// Node id of candidate j, -1 for an empty slot, -2 for a bad index (an id
// outside [0, N) or a node row outside [0, n_nodes)).
__device__ __forceinline__ int32_t trav_candidate(const TravArgs& a, const TravLevel& L, int64_t j)
{
    if (L.roots) {
        const int64_t r = L.roots[j];
        return (r >= 0 && r < a.N) ? (int32_t)r : -2;
    }
    const int32_t p = L.frontier[j / a.C];
    const int32_t nd = a.node_index[p];
    if (nd < 0 || nd >= a.n_nodes) return -2;
    const int32_t c = a.tree[(int64_t)nd * a.C + (j % a.C)];
    if (c == -1) return -1;
    return (c >= 0 && c < a.N) ? c : -2;
}

// 1 = keep, 2 = next, 0 = empty slot.
__device__ __forceinline__ uint8_t trav_decide(const TravArgs& a, int32_t id)
{
    const int32_t ni = a.node_index[id];
    if (ni == -1) return 1;
    const float* s = a.scaling + 3 * (int64_t)id;
    const float* q = a.rotation + 4 * (int64_t)id;
    const float* p = a.xyz + 3 * (int64_t)id;
    const float q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    const float qn = fmaxf(sqrtf(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3), 1e-12f);
    const float r = radius_v2_point(a.proj, a.view, a.fx, a.fy, a.tanx, a.tany, p[0], p[1], p[2],
                                    expf_cr(s[0]), expf_cr(s[1]), expf_cr(s[2]), q0 / qn, q1 / qn,
                                    q2 / qn, q3 / qn);
    return r < a.min_px ? 1 : 2;
}

__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, uint32_t* red)
{
    v = wave_sum_u32(v);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) red[w] = v;
    __syncthreads();
    uint32_t t = 0;
#pragma unroll
    for (int k = 0; k < kTravBlock / 64; ++k) t += red[k];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(kTravBlock) void k_trav_count(TravArgs a, TravLevel L)
{
    __shared__ uint32_t red[kTravBlock / 64];
    int64_t beg, end;
    trav_chunk(trav_size(a, L), beg, end);
    uint32_t keep = 0, nxt = 0;
    for (int64_t j = beg + threadIdx.x; j < end; j += kTravBlock) {
        const int32_t id = trav_candidate(a, L, j);
        uint8_t f = 0;
        if (id >= 0) {
            f = trav_decide(a, id);
        } else if (id == -2) {
            atomicOr(a.ctl + kTravErr, 1);
        }
        a.flags[j] = f;
        keep += f == 1;
        nxt += f == 2;
    }
    keep = block_sum_u32(keep, red);
    nxt = block_sum_u32(nxt, red);
    if (threadIdx.x == 0) {
        a.counts[blockIdx.x] = keep;
        a.counts[kTravGrid + blockIdx.x] = nxt;
    }
}

// One block of kTravGrid threads: exclusive scans of the per-block counts,
// then the new output and frontier sizes.  Sizes past the capacities are
// clamped (the scatter checks every position) and flagged.
__global__ __launch_bounds__(kTravGrid) void k_trav_scan(TravArgs a, TravLevel L)
{
    __shared__ uint32_t sk[kTravGrid], sn[kTravGrid];
    const int t = threadIdx.x;
    const uint32_t base = (uint32_t)*L.n_out;
    const uint32_t k = a.counts[t], n = a.counts[kTravGrid + t];
    sk[t] = k;
    sn[t] = n;
    __syncthreads();
    for (int d = 1; d < kTravGrid; d <<= 1) {
        const uint32_t vk = t >= d ? sk[t - d] : 0u, vn = t >= d ? sn[t - d] : 0u;
        __syncthreads();
        sk[t] += vk;
        sn[t] += vn;
        __syncthreads();
    }
    a.offs[t] = base + sk[t] - k;
    a.offs[kTravGrid + t] = sn[t] - n;
    if (t == kTravGrid - 1) {
        uint64_t tot_out = (uint64_t)base + sk[t];
        uint64_t tot_next = sn[t];
        if (tot_out > (uint64_t)a.cap_out) { tot_out = a.cap_out; atomicOr(a.ctl + kTravErr, 2); }
        if (tot_next > (uint64_t)a.cap_frontier) {
            tot_next = a.cap_frontier;
            atomicOr(a.ctl + kTravErr, 2);
        }
        *L.n_out = (int32_t)tot_out;
        *L.n_next = (int32_t)tot_next;
    }
}

__global__ __launch_bounds__(kTravBlock) void k_trav_scatter(TravArgs a, TravLevel L)
{
    __shared__ uint32_t wk[kTravBlock / 64], wn[kTravBlock / 64];
    int64_t beg, end;
    trav_chunk(trav_size(a, L), beg, end);
    uint32_t ko = a.offs[blockIdx.x], no = a.offs[kTravGrid + blockIdx.x];
    const int w = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane_id()) - 1ull;
    for (int64_t b0 = beg; b0 < end; b0 += kTravBlock) {
        const int64_t j = b0 + threadIdx.x;
        const uint8_t f = j < end ? a.flags[j] : 0;
        const uint64_t mk = __ballot(f == 1), mn = __ballot(f == 2);
        if (lane_id() == 0) {
            wk[w] = __popcll(mk);
            wn[w] = __popcll(mn);
        }
        __syncthreads();
        uint32_t pk = 0, pn = 0, tk = 0, tn = 0;
#pragma unroll
        for (int k = 0; k < kTravBlock / 64; ++k) {
            if (k < w) { pk += wk[k]; pn += wn[k]; }
            tk += wk[k];
            tn += wn[k];
        }
        __syncthreads();
        if (f) {
            const int32_t id = trav_candidate(a, L, j);
            if (f == 1) {
                const uint32_t pos = ko + pk + __popcll(mk & below);
                if (pos < (uint32_t)a.cap_out) L.out[pos] = id;
            } else {
                const uint32_t pos = no + pn + __popcll(mn & below);
                if (pos < (uint32_t)a.cap_frontier) L.next[pos] = id;
            }
        }
        ko += tk;
        no += tn;
    }
}

// The frontier at the level limit goes to the output after the kept nodes.
__global__ __launch_bounds__(kTravBlock) void k_trav_append(TravArgs a, const int32_t* frontier,
                                                            const int32_t* n_frontier,
                                                            const int32_t* n_out, int32_t* out)
{
    const int64_t n = *n_frontier, base = *n_out;
    for (int64_t i = (int64_t)blockIdx.x * kTravBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kTravBlock) {
        if (base + i < a.cap_out) out[base + i] = frontier[i];
    }
}

__global__ void k_trav_clear(int32_t* ctl)
{
    if (threadIdx.x < kTravCtl) ctl[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kTravBlock) void k_trav_emit(const int32_t* ids, int64_t n, int64_t* out)
{
    const int64_t i = (int64_t)blockIdx.x * kTravBlock + threadIdx.x;
    if (i < n) out[i] = ids[i];
}

}  // namespace

hipError_t launch_traverse(const TravArgs& a, const int64_t* roots, int32_t n_roots,
                           int32_t* frontier_a, int32_t* frontier_b, int32_t* out, int levels,
                           bool append, hipStream_t s)
{
    hipLaunchKernelGGL(k_trav_clear, dim3(1), dim3(64), 0, s, a.ctl);
    int32_t* fr[2] = { frontier_a, frontier_b };
    int32_t* nf[2] = { a.ctl + kTravFrontierA, a.ctl + kTravFrontierB };
    for (int l = 0; l <= levels; ++l) {
        TravLevel L{};
        if (l == 0) {
            L.roots = roots;
            L.n_roots = n_roots;
        } else {
            L.frontier = fr[(l - 1) & 1];
            L.n_frontier = nf[(l - 1) & 1];
        }
        L.next = fr[l & 1];
        L.n_next = nf[l & 1];
        L.out = out;
        L.n_out = a.ctl + kTravOut;
        hipLaunchKernelGGL(k_trav_count, dim3(kTravGrid), dim3(kTravBlock), 0, s, a, L);
        hipLaunchKernelGGL(k_trav_scan, dim3(1), dim3(kTravGrid), 0, s, a, L);
        hipLaunchKernelGGL(k_trav_scatter, dim3(kTravGrid), dim3(kTravBlock), 0, s, a, L);
    }
    if (append) {
        hipLaunchKernelGGL(k_trav_append, dim3(kTravGrid), dim3(kTravBlock), 0, s, a, fr[levels & 1],
                           nf[levels & 1], a.ctl + kTravOut, out);
    }
    return hipGetLastError();
}

hipError_t launch_traverse_emit(const int32_t* ids, int64_t n, int64_t* out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_trav_emit, dim3((unsigned)((n + kTravBlock - 1) / kTravBlock)),
                       dim3(kTravBlock), 0, s, ids, n, out);
    return hipGetLastError();
}

}  // namespace fgs
