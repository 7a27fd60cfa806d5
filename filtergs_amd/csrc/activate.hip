// activate.hip -- gather + activation of the visible Gaussians of a LoD cut.
//
// Replaces FilterGS.get_all (FilterGS/filtergs.py:98-134): gather_attributes
// (FilterGS/cuda/gather_attributes_kernel.cu:6-145) followed by
// Activation.activate_root_return (LoG/model/activation.py:36-43):
//   scaling  = exp(log_scale)
//   opacity  = sigmoid(logit)
//   rotation = q / max(|q|, 1e-12)                       (F.normalize)
//   colors   = SH2RGB(dc) + eval_sh_wobase(dir, rest)    (no clamp; the SH
//              term only for active degree > 0, dir = normalised xyz - campos)
// in one pass over the index: each output row reads its source row once, the
// five torch kernels and the materialised copy of the raw rows disappear.
#include "fgs_device.h"
#include "fgs_internal.h"

namespace fgs {

// Real spherical-harmonics constants of LoG/model/sh_utils.py:1-28.
__constant__ float kA_C0 = 0.28209479177387814f;
__constant__ float kA_C1 = 0.4886025119029199f;
__constant__ float kA_C2[5] = { 1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                                -1.0925484305920792f, 0.5462742152960396f };
__constant__ float kA_C3[7] = { -0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                                0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                                -0.5900435899266435f };
__constant__ float kA_C4[9] = { 2.5033429417967046f, -1.7701307697799304f, 0.9461746957575601f,
                                -0.6690465435572892f, 0.10578554691520431f, -0.6690465435572892f,
                                0.47308734787878004f, -1.7701307697799304f, 0.6258357354491761f };

// eval_sh_wobase (sh_utils.py:31-68) of one channel c: basis functions of
// degree 1..deg against the coefficient row r (coefficient k >= 1 of channel c
// at r[(k - 1) * 3 + c]).
__device__ __forceinline__ float sh_wobase(int deg, const float* __restrict__ r, int c, float x,
                                           float y, float z)
{
#define R(k) r[((k) - 1) * 3 + c]
    float res = -kA_C1 * y * R(1) + kA_C1 * z * R(2) - kA_C1 * x * R(3);
    if (deg > 1) {
        const float xx = x * x, yy = y * y, zz = z * z;
        const float xy = x * y, yz = y * z, xz = x * z;
        res = res + kA_C2[0] * xy * R(4) + kA_C2[1] * yz * R(5) + kA_C2[2] * (2.0f * zz - xx - yy) * R(6) +
              kA_C2[3] * xz * R(7) + kA_C2[4] * (xx - yy) * R(8);
        if (deg > 2) {
            res = res + kA_C3[0] * y * (3.f * xx - yy) * R(9) + kA_C3[1] * xy * z * R(10) +
                  kA_C3[2] * y * (4.f * zz - xx - yy) * R(11) +
                  kA_C3[3] * z * (2.f * zz - 3.f * xx - 3.f * yy) * R(12) +
                  kA_C3[4] * x * (4.f * zz - xx - yy) * R(13) + kA_C3[5] * z * (xx - yy) * R(14) +
                  kA_C3[6] * x * (xx - 3.f * yy) * R(15);
            if (deg > 3) {
                res = res + kA_C4[0] * xy * (xx - yy) * R(16) + kA_C4[1] * yz * (3.f * xx - yy) * R(17) +
                      kA_C4[2] * xy * (7.f * zz - 1.f) * R(18) + kA_C4[3] * yz * (7.f * zz - 3.f) * R(19) +
                      kA_C4[4] * (zz * (35.f * zz - 30.f) + 3.f) * R(20) +
                      kA_C4[5] * xz * (7.f * zz - 3.f) * R(21) +
                      kA_C4[6] * (xx - yy) * (7.f * zz - 1.f) * R(22) +
                      kA_C4[7] * xz * (xx - 3.f * yy) * R(23) +
                      kA_C4[8] * (xx * (xx - 3.f * yy) - yy * (3.f * xx - yy)) * R(24);
            }
        }
    }
#undef R
    return res;
}

__global__ __launch_bounds__(256) void k_activate_gather(ActivateArgs a)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.N) return;
    const int64_t src = a.index[j];
    const bool ok = src >= 0 && src < a.n_src;
    if (!ok) {
        atomicOr(a.bad_index, 1);
        return;
    }
    const float px = a.xyz[3 * src], py = a.xyz[3 * src + 1], pz = a.xyz[3 * src + 2];
    a.xyz_o[3 * j] = px;
    a.xyz_o[3 * j + 1] = py;
    a.xyz_o[3 * j + 2] = pz;
#pragma unroll
    for (int k = 0; k < 3; ++k) a.scale_o[3 * j + k] = expf(a.log_scale[3 * src + k]);
    const float q0 = a.rot[4 * src], q1 = a.rot[4 * src + 1], q2 = a.rot[4 * src + 2], q3 = a.rot[4 * src + 3];
    const float qn = fmaxf(sqrtf(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3), 1e-12f);
    a.rot_o[4 * j] = q0 / qn;
    a.rot_o[4 * j + 1] = q1 / qn;
    a.rot_o[4 * j + 2] = q2 / qn;
    a.rot_o[4 * j + 3] = q3 / qn;
    a.opacity_o[j] = 1.0f / (1.0f + expf(-a.opacity_logit[src]));
    float col[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) col[c] = a.dc[3 * src + c] * kA_C0 + 0.5f;
    if (a.deg > 0) {
        const float* row = a.rest + src * (int64_t)a.rest_floats;
        float dx = px - a.campos[0], dy = py - a.campos[1], dz = pz - a.campos[2];
        const float dn = sqrtf(dx * dx + dy * dy + dz * dz);
        dx /= dn;
        dy /= dn;
        dz /= dn;
#pragma unroll
        for (int c = 0; c < 3; ++c) col[c] += sh_wobase(a.deg, row, c, dx, dy, dz);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) a.colors_o[3 * j + c] = col[c];
}

hipError_t launch_activate_gather(const ActivateArgs& a, hipStream_t s)
{
    if (a.N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_activate_gather, dim3((unsigned)((a.N + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace fgs
