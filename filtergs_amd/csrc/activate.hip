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
#include "fgs_activation.h"

namespace fgs {

__global__ __launch_bounds__(256) void k_activate_gather(ActivateArgs a)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.N) return;
    const int64_t src = a.index ? a.index[j] : j;
    const bool ok = src >= 0 && src < a.n_src;
    if (!ok) {
        atomicOr(a.bad_index, 1);
        return;
    }
    ActRow r;
    activate_row(src, a.xyz, a.log_scale, a.rot, a.opacity_logit, a.dc, a.rest, a.rest_floats, a.deg,
                 a.campos, r);
    if (a.xyz_o) {
#pragma unroll
        for (int k = 0; k < 3; ++k) a.xyz_o[3 * j + k] = r.p[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) a.scale_o[3 * j + k] = r.s[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) a.rot_o[4 * j + k] = r.q[k];
    a.opacity_o[j] = r.o;
#pragma unroll
    for (int c = 0; c < 3; ++c) a.colors_o[3 * j + c] = r.c[c];
}

hipError_t launch_activate_gather(const ActivateArgs& a, hipStream_t s)
{
    if (a.N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_activate_gather, dim3((unsigned)((a.N + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Backward of the activation for rows given in place (row j of every input is
// source row j): the gradients of the raw rows from the gradients of the
// activated ones.  The mean passes through unchanged and the SH direction uses
// the detached mean (LoG/model/activation.py:27-35), so no mean gradient is
// formed here.  Per row:
//   d log_scale = g_s * exp(log_scale)
//   d rot       = (g - y (y . g)) / |q|  with y = q / |q|, for |q| >= 1e-12;
//                 g / 1e-12 below (the clamp of F.normalize stops the norm term)
//   d logit     = g_o * (1 - o) * o, o = sigmoid(logit)
//   d dc        = g_c * C0
//   d rest_k    = g_c * basis_k(dir) for k < (deg+1)^2, 0 for the others
// A null gradient input counts as zero; a null output is not written.
__global__ __launch_bounds__(256) void k_activate_bwd(ActivateBwdArgs a)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.N) return;
    if (a.d_log_scale) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float g = a.g_scale ? a.g_scale[3 * j + k] : 0.f;
            a.d_log_scale[3 * j + k] = g * expf(a.log_scale[3 * j + k]);
        }
    }
    if (a.d_rot) {
        const float* q = a.rot + 4 * j;
        const float q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
        float g[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) g[k] = a.g_rot ? a.g_rot[4 * j + k] : 0.f;
        const float n = sqrtf(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
        if (n >= 1e-12f) {
            const float y[4] = { q0 / n, q1 / n, q2 / n, q3 / n };
            const float yg = y[0] * g[0] + y[1] * g[1] + y[2] * g[2] + y[3] * g[3];
#pragma unroll
            for (int k = 0; k < 4; ++k) a.d_rot[4 * j + k] = (g[k] - y[k] * yg) / n;
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) a.d_rot[4 * j + k] = g[k] / 1e-12f;
        }
    }
    if (a.d_opacity_logit) {
        const float o = 1.0f / (1.0f + expf(-a.opacity_logit[j]));
        const float g = a.g_opacity ? a.g_opacity[j] : 0.f;
        a.d_opacity_logit[j] = g * (1.0f - o) * o;
    }
    float gc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) gc[c] = a.g_colors ? a.g_colors[3 * j + c] : 0.f;
    if (a.d_dc) {
#pragma unroll
        for (int c = 0; c < 3; ++c) a.d_dc[3 * j + c] = gc[c] * kA_C0;
    }
    if (a.d_rest) {
        float* out = a.d_rest + j * (int64_t)a.rest_floats;
        const int used = a.deg > 0 ? (a.deg + 1) * (a.deg + 1) - 1 : 0;
        int k0 = 0;
        if (used > 0) {
            float dx = a.xyz[3 * j] - a.campos[0], dy = a.xyz[3 * j + 1] - a.campos[1],
                  dz = a.xyz[3 * j + 2] - a.campos[2];
            const float dn = sqrtf(dx * dx + dy * dy + dz * dz);
            dx /= dn;
            dy /= dn;
            dz /= dn;
            float b[24];
            sh_basis(a.deg, dx, dy, dz, b);
#pragma unroll
            for (int k = 0; k < 24; ++k) {
                if (k < used) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) out[3 * k + c] = gc[c] * b[k];
                }
            }
            k0 = 3 * used;
        }
        for (int f = k0; f < a.rest_floats; ++f) out[f] = 0.f;
    }
}

hipError_t launch_activate_bwd(const ActivateBwdArgs& a, hipStream_t s)
{
    if (a.N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_activate_bwd, dim3((unsigned)((a.N + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace fgs
