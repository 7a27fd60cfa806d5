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
    const int64_t src = a.index[j];
    const bool ok = src >= 0 && src < a.n_src;
    if (!ok) {
        atomicOr(a.bad_index, 1);
        return;
    }
    ActRow r;
    activate_row(src, a.xyz, a.log_scale, a.rot, a.opacity_logit, a.dc, a.rest, a.rest_floats, a.deg,
                 a.campos, r);
#pragma unroll
    for (int k = 0; k < 3; ++k) a.xyz_o[3 * j + k] = r.p[k];
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

}  // namespace fgs
