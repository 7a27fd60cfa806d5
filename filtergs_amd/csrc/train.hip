// train.hip -- training-side neighbours of the LoG render path.
//
//   sparse_adam     LoG/model/sparse_optimizer.py:41-78, 163-195: the Adam step
//                   of the visible rows, with the gather of the optimiser
//                   state, the update and the scatter back in one pass
//   counter_update  LoG/model/counter.py:36-83: the per-point statistics of
//                   one render, in one pass over the visible rows and one over
//                   the max-weight point ids
//
// Compiled with -ffp-contract=off: every product and sum rounds as the
// reference's separate torch operations do.
#include "fgs_device.h"
#include "fgs_internal.h"

namespace fgs {

__global__ __launch_bounds__(256) void k_sparse_adam(SparseAdamArgs a)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.n * a.D) return;
    const int64_t j = e / a.D;
    const int d = (int)(e - j * a.D);
    if (a.flag && !a.flag[j]) return;
    const int64_t i = a.index[j];
    if (i < 0 || i >= a.N) {
        atomicOr(a.bad, 1);
        return;
    }
    const size_t src = (size_t)j * a.D + d, dst = (size_t)i * a.D + d;
    const float g = a.grad[src];
    // exp_avg.mul_(beta1).add_(grad, alpha=1-beta1)
    const float m = a.exp_avg[dst] * a.beta1 + a.one_m_beta1 * g;
    // exp_avg_sq.mul_(beta2).addcmul_(grad, grad, value=1-beta2)
    const float v = a.exp_avg_sq[dst] * a.beta2 + a.one_m_beta2 * (g * g);
    float vd = v;
    if (a.max_exp_avg_sq) {
        vd = fmaxf(a.max_exp_avg_sq[dst], v);
        a.max_exp_avg_sq[dst] = vd;
    }
    // denom = sqrt(v) / sqrt(bias_correction2) + eps; p += -step_size * (m / denom)
    const float denom = sqrtf(vd) / a.bc2_sqrt + a.eps;
    a.param[dst] = a.param_sel[src] + a.neg_step_size * (m / denom);
    a.exp_avg[dst] = m;
    a.exp_avg_sq[dst] = v;
}

hipError_t launch_sparse_adam(const SparseAdamArgs& a, hipStream_t s)
{
    const int64_t n = a.n * a.D;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sparse_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Visible rows with a positive radius: counts, weights, radius maximum.
__global__ __launch_bounds__(256) void k_counter_visible(CounterArgs a)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.n_vis || a.radii[j] <= 0) return;
    const int64_t i = a.visible_index[j];
    if (i < 0 || i >= a.N) {
        atomicOr(a.bad, 1);
        return;
    }
    a.create_steps[i] += 1;
    a.visible_count[i] = (int16_t)(a.visible_count[i] + 1);
    const float w = a.point_weight[j];
    a.weights_max[i] = fmaxf(a.weights_max[i], w);
    a.weights_sum[i] += w;
    const int16_t r = (int16_t)a.radii[j];
    if (r > a.radii_max[i]) a.radii_max[i] = r;
}

// Max-weight points of the render (unique ids of the point-id image and
// their pixel counts): area, view-space gradient norm times area, area max.
__global__ __launch_bounds__(256) void k_counter_points(CounterArgs a)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n_pid) return;
    const int64_t j = a.point_id[k];
    if (j < 0 || j >= a.n_vis) {
        atomicOr(a.bad, 2);
        return;
    }
    const int64_t i = a.visible_index[j];
    if (i < 0 || i >= a.N) {
        atomicOr(a.bad, 1);
        return;
    }
    const int64_t c = a.point_count[k];
    a.area_sum[i] += (int32_t)c;
    if (a.grad) {
        const float gx = a.grad[3 * j], gy = a.grad[3 * j + 1];
        a.grad_sum[i] += sqrtf(gx * gx + gy * gy) * (float)c;
    }
    const int32_t ci = (int32_t)c;
    if (ci > a.radii_max_max[i]) a.radii_max_max[i] = ci;
}

hipError_t launch_counter_update(const CounterArgs& a, hipStream_t s)
{
    if (a.n_vis > 0)
        hipLaunchKernelGGL(k_counter_visible, dim3((unsigned)((a.n_vis + 255) / 256)), dim3(256), 0, s, a);
    if (a.n_pid > 0)
        hipLaunchKernelGGL(k_counter_points, dim3((unsigned)((a.n_pid + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace fgs
