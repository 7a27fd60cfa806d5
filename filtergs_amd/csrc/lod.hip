// lod.hip -- LoD-tree cut and per-point utilities of FilterGS/cuda and the
// antialias extension.
//
//   lod_cut        FilterGS/cuda/compute_radius_kernel_v2.cu:109-169, 231-353
//   radius_v2      FilterGS/cuda/compute_radius_kernel_v2.cu:180-229
//   radius_aa      antialias/cuda_rasterizer/forward.cu:187-263
//   mark_visible   antialias/cuda_rasterizer/rasterizer_impl.cu:97-109
//   visible_flag   FilterGS/cuda/visible_flag_kernel.cu:19-81 (exact division)
//   gather         FilterGS/cuda/gather_attributes_kernel.cu:6-66
//
// The LoD cut runs in two streaming phases instead of the reference's lazy
// per-leaf ancestor recomputation with CAS publication: phase 1 computes every
// node's visibility, NDC depth and 3-sigma radius and marks internal nodes
// that would cull their descendants (visible, 0 < r <= radius_max); phase 2
// decides each node, leaves by scanning their ancestor row against those
// one-byte marks.  The result equals the reference's (the lazily computed
// radius is the same deterministic value).  Compiled with -ffp-contract=off
// so the mask is reproducible bit for bit.
#include "fgs_device.h"
#include "fgs_internal.h"

namespace fgs {

__device__ __forceinline__ bool vis_radius(int idx, const LodArgs& a, float* depth_out,
                                           float* radius_out)
{
    const float px = a.means3d[3 * idx], py = a.means3d[3 * idx + 1], pz = a.means3d[3 * idx + 2];
    float hom[4];
    xf4(a.proj, px, py, pz, hom);
    const float p_w = 1.0f / (hom[3] + 0.0000001f);
    const float qx = hom[0] * p_w, qy = hom[1] * p_w, qz = hom[2] * p_w;
    *depth_out = qz;
    const bool visible = (qz > 0.0f) && (qz < 1.0f) && (qx > -1.0f - a.padding) &&
                         (qx < 1.0f + a.padding) && (qy > -1.0f - a.padding) &&
                         (qy < 1.0f + a.padding);
    if (!visible) {
        *radius_out = 0.0f;
        return false;
    }
    const float sx = expf_cr(a.log_scales[3 * idx]);
    const float sy = expf_cr(a.log_scales[3 * idx + 1]);
    const float sz = expf_cr(a.log_scales[3 * idx + 2]);
    float r = a.rotations[4 * idx], x = a.rotations[4 * idx + 1], y = a.rotations[4 * idx + 2],
          z = a.rotations[4 * idx + 3];
    const float nrm = sqrtf(r * r + x * x + y * y + z * z);
    if (nrm > 0.f) {
        r /= nrm; x /= nrm; y /= nrm; z /= nrm;
    } else {
        r = 1.f; x = 0.f; y = 0.f; z = 0.f;
    }
    float c3[6];
    cov3d(sx, sy, sz, r, x, y, z, c3);
    Cov2D cv;
    cov2d(a.view, px, py, pz, a.fx, a.fy, a.tanx, a.tany, c3, cv);
    const float ca = fmaxf(cv.a, 0.3f), cb = cv.b, cc = fmaxf(cv.c, 0.3f);
    const float det = ca * cc - cb * cb;
    if (det <= 0.0f) {
        *radius_out = 0.0f;
        return true;
    }
    const float mid = 0.5f * (ca + cc);
    const float root = sqrtf(fmaxf(0.1f, mid * mid - det));
    *radius_out = 3.f * sqrtf(fmaxf(mid + root, mid - root));
    return true;
}

__global__ __launch_bounds__(256) void k_lod_phase1(LodArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    float d, r;
    const bool vis = vis_radius(i, a, &d, &r);
    a.depth[i] = d;
    a.radius2d[i] = r;
    const bool internal = a.node_index[i] != -1;
    // bit 0: visible; bit 1: internal node that culls its descendant leaves
    a.cull_flag[i] = (uint8_t)((vis ? 1 : 0) | ((vis && internal && r > 0.f && r <= a.radius_max) ? 2 : 0));
}

__global__ __launch_bounds__(256) void k_lod_phase2(LodArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    const uint8_t f = a.cull_flag[i];
    if (!(f & 1)) {
        a.mask[i] = 0;
        return;
    }
    const float rv = a.radius2d[i];
    bool keep = true;
    if (a.node_index[i] != -1) {
        keep = rv <= a.radius_max;
    } else if (a.index_parent) {
        // the same set of ancestors as the rows rebuilt from these links
        int32_t p = a.index_parent[i];
        for (int step = 0; p >= 0 && step < kLodMaxWalk; ++step) {
            if (p != i && (a.cull_flag[p] & 2)) {
                keep = false;
                break;
            }
            p = a.index_parent[p];
        }
    } else {
        const int32_t* row = a.ancestor_path + (size_t)i * a.L;
        for (int lv = 0; lv < a.L; ++lv) {
            const int32_t anc = row[lv];
            if (anc < 0) break;
            if (anc == i) continue;
            if (a.cull_flag[anc] & 2) {
                keep = false;
                break;
            }
        }
    }
    if (keep && a.radius_filter_threshold > 0.0f && rv > a.radius_filter_threshold) keep = false;
    a.mask[i] = keep ? 1 : 0;
}

// Ancestor row of node i: column d holds its ancestor at depth d (d <
// depth[i]), -1 elsewhere.  The reference fills rows level by level, copying
// the parent's row (apps/ancestor.py:55-73); walking the parent links gives
// the same rows when every parent is one level up.  A node whose parent is
// -1 below the root level is an orphan: its row stays -1, and the rows of
// its descendants stop at it.
__global__ __launch_bounds__(256) void k_rebuild_ancestor(int N, int L, const int32_t* __restrict__ index_parent,
                                                          const int32_t* __restrict__ depth,
                                                          int32_t* __restrict__ out, int* bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int32_t* row = out + (size_t)i * L;
    for (int c = 0; c < L; ++c) row[c] = -1;
    int lvl = depth[i];
    int cur = i;
    while (lvl >= 1) {
        const int32_t p = index_parent[cur];
        if (p < 0 || p >= N) break;
        if (lvl > L) {
            atomicOr(bad, 1);
            break;
        }
        row[lvl - 1] = p;
        cur = p;
        --lvl;
    }
}

hipError_t launch_rebuild_ancestor(int N, int L, const int32_t* index_parent, const int32_t* depth,
                                   int32_t* out, int* bad, hipStream_t s)
{
    if (N == 0 || L == 0) return hipSuccess;
    hipLaunchKernelGGL(k_rebuild_ancestor, dim3((N + 255) / 256), dim3(256), 0, s, N, L, index_parent, depth,
                       out, bad);
    return hipGetLastError();
}

hipError_t launch_lod_cut(const LodArgs& a, hipStream_t s)
{
    if (a.N == 0) return hipSuccess;
    const dim3 grid((a.N + 255) / 256);
    hipLaunchKernelGGL(k_lod_phase1, grid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_lod_phase2, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_radius_v2(int P, const float* means3d, const float* scales,
                                                   const float* rotations, const float* proj,
                                                   const float* view, float fx, float fy, float tanx,
                                                   float tany, float* radii)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const float out = radius_v2_point(proj, view, fx, fy, tanx, tany, means3d[3 * i],
                                      means3d[3 * i + 1], means3d[3 * i + 2], scales[3 * i],
                                      scales[3 * i + 1], scales[3 * i + 2], rotations[4 * i],
                                      rotations[4 * i + 1], rotations[4 * i + 2],
                                      rotations[4 * i + 3]);
    radii[i] = out;
}

hipError_t launch_radius_v2(int P, const float* means3d, const float* scales,
                            const float* rotations, const float* proj, const float* view,
                            float fx, float fy, float tanx, float tany, float* radii, hipStream_t s)
{
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(k_radius_v2, dim3((P + 255) / 256), dim3(256), 0, s, P, means3d, scales,
                       rotations, proj, view, fx, fy, tanx, tany, radii);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_radius_aa(int P, const float* means3d, const float* scales,
                                                   const float* rotations, const float* view,
                                                   const float* proj, float tanx, float tany,
                                                   float fx, float fy, float* radii)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    float out = 0.f;
    const float px = means3d[3 * i], py = means3d[3 * i + 1], pz = means3d[3 * i + 2];
    float hom[4], v[3];
    xf4(proj, px, py, pz, hom);
    const float p_w = 1.0f / (hom[3] + 0.0000001f);
    const float qx = hom[0] * p_w, qy = hom[1] * p_w;
    xf3(view, px, py, pz, v);
    if (!(v[2] <= 0.2f || (double)qx < -1.3 || (double)qx > 1.3 || (double)qy < -1.3 ||
          (double)qy > 1.3)) {
        float c3[6];
        cov3d(scales[3 * i], scales[3 * i + 1], scales[3 * i + 2], rotations[4 * i],
              rotations[4 * i + 1], rotations[4 * i + 2], rotations[4 * i + 3], c3);
        Cov2D cv;
        cov2d(view, px, py, pz, fx, fy, tanx, tany, c3, cv);
        const float det = cv.a * cv.c - cv.b * cv.b;
        if (det != 0.0f) {
            const float mid = 0.5f * (cv.a + cv.c);
            const float l1 = mid + sqrtf(fmaxf(0.1f, mid * mid - det));
            const float l2 = mid - sqrtf(fmaxf(0.1f, mid * mid - det));
            out = 3.f * sqrtf(fmaxf(l1, l2));
        }
    }
    radii[i] = out;
}

hipError_t launch_radius_aa(int P, const float* means3d, const float* scales,
                            const float* rotations, const float* view, const float* proj,
                            float tanx, float tany, int H, int W, float* radii, hipStream_t s)
{
    if (P == 0) return hipSuccess;
    const float fy = (float)H / (2.0f * tany);
    const float fx = (float)W / (2.0f * tanx);
    hipLaunchKernelGGL(k_radius_aa, dim3((P + 255) / 256), dim3(256), 0, s, P, means3d, scales,
                       rotations, view, proj, tanx, tany, fx, fy, radii);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_mark_visible(int P, const float* means3d, const float* view,
                                                      uint8_t* present)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    float v[3];
    xf3(view, means3d[3 * i], means3d[3 * i + 1], means3d[3 * i + 2], v);
    present[i] = !(v[2] <= 0.2f);
}

hipError_t launch_mark_visible(int P, const float* means3d, const float* view, uint8_t* present,
                               hipStream_t s)
{
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mark_visible, dim3((P + 255) / 256), dim3(256), 0, s, P, means3d, view, present);
    return hipGetLastError();
}

// B views of N points each sharing one matrix (visible_flag_cuda_batch runs B
// launches; here one 2-D grid).
__global__ __launch_bounds__(256) void k_visible_flag(int N, const float* xyz, const float* proj,
                                                      float padding, uint8_t* valid, float* depth,
                                                      float* p_proj)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t b = blockIdx.y;
    const size_t e = b * (size_t)N + i;
    float hom[4];
    xf4(proj, xyz[3 * e], xyz[3 * e + 1], xyz[3 * e + 2], hom);
    const float w = hom[3];
    const float inv_w = 1.0f / (w + 1e-7f);
    const float x = hom[0] * inv_w, y = hom[1] * inv_w, z = hom[2] * inv_w;
    valid[e] = (w > 0.0f) && (w < 100.0f) && (x > -1.0f - padding) && (x < 1.0f + padding) &&
               (y > -1.0f - padding) && (y < 1.0f + padding);
    depth[e] = w;
    p_proj[3 * e] = x;
    p_proj[3 * e + 1] = y;
    p_proj[3 * e + 2] = z;
}

hipError_t launch_visible_flag(int B, int N, const float* xyz, const float* proj, float padding,
                               uint8_t* valid, float* depth, float* p_proj, hipStream_t s)
{
    if (N == 0 || B == 0) return hipSuccess;
    hipLaunchKernelGGL(k_visible_flag, dim3((N + 255) / 256, B), dim3(256), 0, s, N, xyz, proj,
                       padding, valid, depth, p_proj);
    return hipGetLastError();
}

// Row gather; out-of-range indices raise the bad_index flag and yield zeros.
__global__ __launch_bounds__(256) void k_gather(int64_t N, const int64_t* index, int64_t n_src,
                                                const float* xyz, const float* scaling,
                                                const float* rotation, const float* opacity,
                                                const float* colors, const float* shs, int sh_floats,
                                                float* xyz_o, float* scaling_o, float* rotation_o,
                                                float* opacity_o, float* colors_o, float* shs_o,
                                                int* bad_index)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const int64_t src = index[j];
    const bool ok = src >= 0 && src < n_src;
    if (!ok) atomicOr(bad_index, 1);
    const int64_t sidx = ok ? src : 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        xyz_o[3 * j + k] = ok ? xyz[3 * sidx + k] : 0.f;
        scaling_o[3 * j + k] = ok ? scaling[3 * sidx + k] : 0.f;
        colors_o[3 * j + k] = ok ? colors[3 * sidx + k] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) rotation_o[4 * j + k] = ok ? rotation[4 * sidx + k] : 0.f;
    opacity_o[j] = ok ? opacity[sidx] : 0.f;
    if (shs && shs_o)
        for (int k = 0; k < sh_floats; ++k)
            shs_o[j * sh_floats + k] = ok ? shs[sidx * sh_floats + k] : 0.f;
}

hipError_t launch_gather(int64_t N, const int64_t* index, int64_t n_src, const float* xyz,
                         const float* scaling, const float* rotation, const float* opacity,
                         const float* colors, const float* shs, int sh_floats, float* xyz_o,
                         float* scaling_o, float* rotation_o, float* opacity_o, float* colors_o,
                         float* shs_o, int* bad_index, hipStream_t s)
{
    if (N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, N, index, n_src,
                       xyz, scaling, rotation, opacity, colors, shs, sh_floats, xyz_o, scaling_o,
                       rotation_o, opacity_o, colors_o, shs_o, bad_index);
    return hipGetLastError();
}

}  // namespace fgs
