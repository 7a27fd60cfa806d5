// fgs_activation.h -- the activation of FilterGS rows (device functions
// shared by the gather kernel and the preprocess of the fused render).
//
// FilterGS.get_all (FilterGS/filtergs.py:98-134) with
// Activation.activate_root_return (LoG/model/activation.py:36-43):
//   scaling  = exp(log_scale)
//   opacity  = sigmoid(logit)
//   rotation = q / max(|q|, 1e-12)                       (F.normalize)
//   colors   = SH2RGB(dc) + eval_sh_wobase(dir, rest)    (no clamp; the SH
//              term only for active degree > 0, dir = normalised xyz - campos)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fgs {

// Real spherical-harmonics constants of LoG/model/sh_utils.py:1-28.
constexpr float kA_C0 = 0.28209479177387814f;
constexpr float kA_C1 = 0.4886025119029199f;
constexpr float kA_C2[5] = { 1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                             -1.0925484305920792f, 0.5462742152960396f };
constexpr float kA_C3[7] = { -0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                             0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                             -0.5900435899266435f };
constexpr float kA_C4[9] = { 2.5033429417967046f, -1.7701307697799304f, 0.9461746957575601f,
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

// The basis functions of eval_sh_wobase themselves: b[k - 1] multiplies
// coefficient k (k = 1 .. (deg+1)^2 - 1), each formed as the factor the
// reference multiplies the coefficient by.
__device__ __forceinline__ void sh_basis(int deg, float x, float y, float z, float* __restrict__ b)
{
    b[0] = -kA_C1 * y;
    b[1] = kA_C1 * z;
    b[2] = -kA_C1 * x;
    if (deg > 1) {
        const float xx = x * x, yy = y * y, zz = z * z;
        const float xy = x * y, yz = y * z, xz = x * z;
        b[3] = kA_C2[0] * xy;
        b[4] = kA_C2[1] * yz;
        b[5] = kA_C2[2] * (2.0f * zz - xx - yy);
        b[6] = kA_C2[3] * xz;
        b[7] = kA_C2[4] * (xx - yy);
        if (deg > 2) {
            b[8] = kA_C3[0] * y * (3.f * xx - yy);
            b[9] = kA_C3[1] * xy * z;
            b[10] = kA_C3[2] * y * (4.f * zz - xx - yy);
            b[11] = kA_C3[3] * z * (2.f * zz - 3.f * xx - 3.f * yy);
            b[12] = kA_C3[4] * x * (4.f * zz - xx - yy);
            b[13] = kA_C3[5] * z * (xx - yy);
            b[14] = kA_C3[6] * x * (xx - 3.f * yy);
            if (deg > 3) {
                b[15] = kA_C4[0] * xy * (xx - yy);
                b[16] = kA_C4[1] * yz * (3.f * xx - yy);
                b[17] = kA_C4[2] * xy * (7.f * zz - 1.f);
                b[18] = kA_C4[3] * yz * (7.f * zz - 3.f);
                b[19] = kA_C4[4] * (zz * (35.f * zz - 30.f) + 3.f);
                b[20] = kA_C4[5] * xz * (7.f * zz - 3.f);
                b[21] = kA_C4[6] * (xx - yy) * (7.f * zz - 1.f);
                b[22] = kA_C4[7] * xz * (xx - 3.f * yy);
                b[23] = kA_C4[8] * (xx * (xx - 3.f * yy) - yy * (3.f * xx - yy));
            }
        }
    }
}

// One activated FilterGS row: mean, scale, unit quaternion, opacity and
// colour of row src (rest_floats = 3 * rest coefficients per row).
struct ActRow {
    float p[3], s[3], q[4], o, c[3];
};

__device__ __forceinline__ void activate_row(int64_t src, const float* __restrict__ xyz,
                                             const float* __restrict__ log_scale,
                                             const float* __restrict__ rot,
                                             const float* __restrict__ opacity_logit,
                                             const float* __restrict__ dc,
                                             const float* __restrict__ rest, int rest_floats, int deg,
                                             const float* __restrict__ campos, ActRow& r)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) r.p[k] = xyz[3 * src + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) r.s[k] = expf(log_scale[3 * src + k]);
    const float q0 = rot[4 * src], q1 = rot[4 * src + 1], q2 = rot[4 * src + 2], q3 = rot[4 * src + 3];
    const float qn = fmaxf(sqrtf(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3), 1e-12f);
    r.q[0] = q0 / qn;
    r.q[1] = q1 / qn;
    r.q[2] = q2 / qn;
    r.q[3] = q3 / qn;
    r.o = 1.0f / (1.0f + expf(-opacity_logit[src]));
#pragma unroll
    for (int c = 0; c < 3; ++c) r.c[c] = dc[3 * src + c] * kA_C0 + 0.5f;
    if (deg > 0) {
        const float* row = rest + src * (int64_t)rest_floats;
        float dx = r.p[0] - campos[0], dy = r.p[1] - campos[1], dz = r.p[2] - campos[2];
        const float dn = sqrtf(dx * dx + dy * dy + dz * dz);
        dx /= dn;
        dy /= dn;
        dz /= dn;
#pragma unroll
        for (int c = 0; c < 3; ++c) r.c[c] += sh_wobase(deg, row, c, dx, dy, dz);
    }
}

}  // namespace fgs
