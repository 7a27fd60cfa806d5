// fgs_device.h -- device helpers shared by the gfx950 kernels.
//
// The per-Gaussian geometry helpers are written in one fixed operation order
// (glm column-major products, antialias/cuda_rasterizer/forward.cu:102-185)
// and the files that produce integer results from them (radius, tile
// rectangle, depth key, LoD mask) are compiled with -ffp-contract=off, so that
// the FP32 results are reproducible bit for bit on the CPU.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FGS_TILE 16
#define FGS_BLOCK (FGS_TILE * FGS_TILE)
#define FGS_WAVE 64

namespace fgs {

struct Mat16 { float m[16]; };

// transformPoint4x4 / 4x3 (auxiliary.h:58-77)
__device__ __forceinline__ void xf4(const float* m, float px, float py, float pz, float* o)
{
    o[0] = m[0] * px + m[4] * py + m[8] * pz + m[12];
    o[1] = m[1] * px + m[5] * py + m[9] * pz + m[13];
    o[2] = m[2] * px + m[6] * py + m[10] * pz + m[14];
    o[3] = m[3] * px + m[7] * py + m[11] * pz + m[15];
}
__device__ __forceinline__ void xf3(const float* m, float px, float py, float pz, float* o)
{
    o[0] = m[0] * px + m[4] * py + m[8] * pz + m[12];
    o[1] = m[1] * px + m[5] * py + m[9] * pz + m[13];
    o[2] = m[2] * px + m[6] * py + m[10] * pz + m[14];
}

// Rotation of the unnormalised quaternion (r, x, y, z), glm column-major
// R[col][row] (forward.cu:167-171).
__device__ __forceinline__ void quat_rot(float r, float x, float y, float z, float R[3][3])
{
    R[0][0] = 1.f - 2.f * (y * y + z * z);
    R[0][1] = 2.f * (x * y - r * z);
    R[0][2] = 2.f * (x * z + r * y);
    R[1][0] = 2.f * (x * y + r * z);
    R[1][1] = 1.f - 2.f * (x * x + z * z);
    R[1][2] = 2.f * (y * z - r * x);
    R[2][0] = 2.f * (x * z - r * y);
    R[2][1] = 2.f * (y * z + r * x);
    R[2][2] = 1.f - 2.f * (x * x + y * y);
}

// Sigma = (S R)^T (S R), upper triangle (forward.cu:151-185).
__device__ __forceinline__ void cov3d(float sx, float sy, float sz, float r, float x, float y,
                                      float z, float* cov)
{
    float R[3][3], M[3][3];
    quat_rot(r, x, y, z, R);
    const float s[3] = { sx, sy, sz };
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) M[j][i] = s[i] * R[j][i];
#define FGS_SIG(j, i) (M[i][0] * M[j][0] + M[i][1] * M[j][1] + M[i][2] * M[j][2])
    cov[0] = FGS_SIG(0, 0);
    cov[1] = FGS_SIG(0, 1);
    cov[2] = FGS_SIG(0, 2);
    cov[3] = FGS_SIG(1, 1);
    cov[4] = FGS_SIG(1, 2);
    cov[5] = FGS_SIG(2, 2);
#undef FGS_SIG
}

// EWA 2D covariance before dilation (forward.cu:102-146).  Keeps the
// intermediates the backward pass needs.
struct Cov2D {
    float t[3];
    float txtz, tytz, limx, limy;
    float J00, J02, J11, J12;
    float W[3][3];
    float T[2][3];
    float a, b, c;
};

__device__ __forceinline__ void cov2d(const float* V, float px, float py, float pz, float fx,
                                      float fy, float tanx, float tany, const float* c3, Cov2D& o)
{
    xf3(V, px, py, pz, o.t);
    o.limx = 1.3f * tanx;
    o.limy = 1.3f * tany;
    o.txtz = o.t[0] / o.t[2];
    o.tytz = o.t[1] / o.t[2];
    o.t[0] = fminf(o.limx, fmaxf(-o.limx, o.txtz)) * o.t[2];
    o.t[1] = fminf(o.limy, fmaxf(-o.limy, o.tytz)) * o.t[2];
    const float tz = o.t[2];
    o.J00 = fx / tz;
    o.J02 = -(fx * o.t[0]) / (tz * tz);
    o.J11 = fy / tz;
    o.J12 = -(fy * o.t[1]) / (tz * tz);
    o.W[0][0] = V[0]; o.W[0][1] = V[4]; o.W[0][2] = V[8];
    o.W[1][0] = V[1]; o.W[1][1] = V[5]; o.W[1][2] = V[9];
    o.W[2][0] = V[2]; o.W[2][1] = V[6]; o.W[2][2] = V[10];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        o.T[0][i] = o.W[0][i] * o.J00 + o.W[2][i] * o.J02;
        o.T[1][i] = o.W[1][i] * o.J11 + o.W[2][i] * o.J12;
    }
    const float Vk[3][3] = { { c3[0], c3[1], c3[2] }, { c3[1], c3[3], c3[4] }, { c3[2], c3[4], c3[5] } };
    float A[3][2];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 2; ++i)
            A[j][i] = o.T[i][0] * Vk[j][0] + o.T[i][1] * Vk[j][1] + o.T[i][2] * Vk[j][2];
    o.a = A[0][0] * o.T[0][0] + A[1][0] * o.T[0][1] + A[2][0] * o.T[0][2];
    o.b = A[0][1] * o.T[0][0] + A[1][1] * o.T[0][1] + A[2][1] * o.T[0][2];
    o.c = A[0][1] * o.T[1][0] + A[1][1] * o.T[1][1] + A[2][1] * o.T[1][2];
}

// Saturating truncation used by the tile rectangle.
__device__ __forceinline__ int sat_int(float v)
{
    return (int)fminf(fmaxf(v, -1.0e9f), 1.0e9f);
}
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// getRect (auxiliary.h:46-56).
__device__ __forceinline__ void get_rect(float px, float py, int r, int gx, int gy, int* out)
{
    const float rr = (float)r;
    out[0] = clampi(sat_int((px - rr) / 16.0f), 0, gx);
    out[1] = clampi(sat_int((py - rr) / 16.0f), 0, gy);
    out[2] = clampi(sat_int((px + rr + 16.0f - 1.0f) / 16.0f), 0, gx);
    out[3] = clampi(sat_int((py + rr + 16.0f - 1.0f) / 16.0f), 0, gy);
}

// Correctly rounded float log / exp through double precision.
__device__ __forceinline__ float logf_cr(float x) { return (float)log((double)x); }
__device__ __forceinline__ float expf_cr(float x) { return (float)exp((double)x); }

// 3-sigma screen radius of compute_radius_cuda
// (FilterGS/cuda/compute_radius_kernel_v2.cu:180-229, the same kernel as
// LoG/cuda/compute_radius_kernel.cu:107-156): scale and rotation used as given,
// 0 outside |NDC x|, |NDC y| <= 1.3 (compared in double, as the reference's
// double literal promotes them), 0.3 px^2 floor on the diagonal of the 2D
// covariance, 0 for a zero determinant.
__device__ __forceinline__ float radius_v2_point(const float* proj, const float* view, float fx,
                                                 float fy, float tanx, float tany, float px,
                                                 float py, float pz, float sx, float sy, float sz,
                                                 float qr, float qx, float qy, float qz)
{
    float hom[4];
    xf4(proj, px, py, pz, hom);
    const float p_w = 1.0f / (hom[3] + 0.0000001f);
    const float nx = hom[0] * p_w, ny = hom[1] * p_w;
    if ((double)nx < -1.3 || (double)nx > 1.3 || (double)ny < -1.3 || (double)ny > 1.3) return 0.f;
    float c3[6];
    cov3d(sx, sy, sz, qr, qx, qy, qz, c3);
    Cov2D cv;
    cov2d(view, px, py, pz, fx, fy, tanx, tany, c3, cv);
    const float ca = fmaxf(cv.a, 0.3f), cb = cv.b, cc = fmaxf(cv.c, 0.3f);
    const float det = ca * cc - cb * cb;
    if (det == 0.0f) return 0.f;
    const float mid = 0.5f * (ca + cc);
    const float l1 = mid + sqrtf(fmaxf(0.1f, mid * mid - det));
    const float l2 = mid - sqrtf(fmaxf(0.1f, mid * mid - det));
    return 3.f * sqrtf(fmaxf(l1, l2));
}

// Wave64 helpers.
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, FGS_WAVE);
    return v;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, FGS_WAVE);
    return v;
}

// Inclusive scans of a u32 across the wave with DPP (no LDS round trips): a
// prefix inside each 16-lane row by row shifts 1, 2, 4, 8 (lanes shifted in
// from outside the row read 0, the identity of both scans), then lane 15's
// value into rows 1 and 3 and lane 31's into rows 2 and 3 (the row
// broadcasts of the wave64 DPP).  A lane reads only lanes below it, so lanes
// that have exited must lie above every live lane.
template <int CTRL, int ROWS, bool BOUND>
__device__ __forceinline__ uint32_t dpp_from_below(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, BOUND);
}
__device__ __forceinline__ uint32_t wave_incl_sum_u32(uint32_t x)
{
    x += dpp_from_below<0x111, 0xF, true>(x);   // row_shr:1
    x += dpp_from_below<0x112, 0xF, true>(x);   // row_shr:2
    x += dpp_from_below<0x114, 0xF, true>(x);   // row_shr:4
    x += dpp_from_below<0x118, 0xF, true>(x);   // row_shr:8
    x += dpp_from_below<0x142, 0xA, false>(x);  // row_bcast:15
    x += dpp_from_below<0x143, 0xC, false>(x);  // row_bcast:31
    return x;
}
__device__ __forceinline__ uint32_t wave_incl_max_u32(uint32_t x)
{
    x = max(x, dpp_from_below<0x111, 0xF, true>(x));
    x = max(x, dpp_from_below<0x112, 0xF, true>(x));
    x = max(x, dpp_from_below<0x114, 0xF, true>(x));
    x = max(x, dpp_from_below<0x118, 0xF, true>(x));
    x = max(x, dpp_from_below<0x142, 0xA, false>(x));
    x = max(x, dpp_from_below<0x143, 0xC, false>(x));
    return x;
}

__device__ __forceinline__ void atomic_add_f32(float* p, float v)
{
    unsafeAtomicAdd(p, v);
}

}  // namespace fgs
