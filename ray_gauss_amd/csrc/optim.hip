// Optimiser step of the training path: Adam over every per-Gaussian parameter tensor in one
// launch, and the post-step projections of the training loop.
//
// Reference: each parameter tensor has its own torch.optim.Adam(eps=1e-15) (classes/optimizer.py:
// 203-218, utils/optim_utils.py:4-83); train.py steps them one after the other (scripts/train.py:
// 121-124) and then projects the parameters (scripts/train.py:168-183).  Here all tensors form one
// grid: each work item is one float4 (or tail float) of one tensor, found from a prefix table of
// tensor sizes, so the step streams parameter, gradient and both moments once (28 B per element:
// 16 read, 12 written) in a single HBM-bound launch.
//
// The arithmetic follows torch's single-tensor CPU Adam step term by term, with contraction off
// so every product and sum is rounded as there:
//   m <- fma(1-b1, g - m, m)                    (lerp with weight 1-b1 < 0.5; torch's vectorised
//                                                lerp is a fused multiply-add)
//   v <- v*b2 + ((1-b2) g) g                    (mul_ then addcmul_)
//   p <- p + ((-lr/(1-b1^t)) m) / (sqrt(v)/sqrt(1-b2^t) + eps)     (addcdiv_)
// where the scalars are formed in double precision on the host and rounded to float once.

#include <hip/hip_runtime.h>

#include "rg_internal.h"

#pragma clang fp contract(off)

namespace rg {
namespace {

constexpr int kOptBlock = 256;

__device__ __forceinline__ float adam_one(float& p, float g, float& m, float& v, const AdamScalars& c) {
    m = fmaf(c.w1, g - m, m);
    v = v * c.b2 + (c.w2 * g) * g;
    const float denom = sqrtf(v) / c.bc2_sqrt + c.eps;
    p = p + (c.neg_step * m) / denom;
    return p;
}

__global__ void __launch_bounds__(kOptBlock) k_adam(AdamBatch b, int64_t total_items) {
    const int64_t it = (int64_t)blockIdx.x * kOptBlock + threadIdx.x;
    if (it >= total_items) return;
    int k = 0;
    while (k + 1 < b.count && it >= b.item_start[k + 1]) ++k;
    const AdamTensor& t = b.t[k];
    const AdamScalars c = b.s[k];
    const int64_t i = (it - b.item_start[k]) * 4;  // first element of this item
    if (i + 4 <= t.n && t.aligned) {
        float4 p = reinterpret_cast<float4*>(t.param)[i >> 2];
        const float4 g = reinterpret_cast<const float4*>(t.grad)[i >> 2];
        float4 m = reinterpret_cast<float4*>(t.exp_avg)[i >> 2];
        float4 v = reinterpret_cast<float4*>(t.exp_avg_sq)[i >> 2];
        adam_one(p.x, g.x, m.x, v.x, c);
        adam_one(p.y, g.y, m.y, v.y, c);
        adam_one(p.z, g.z, m.z, v.z, c);
        adam_one(p.w, g.w, m.w, v.w, c);
        reinterpret_cast<float4*>(t.param)[i >> 2] = p;
        reinterpret_cast<float4*>(t.exp_avg)[i >> 2] = m;
        reinterpret_cast<float4*>(t.exp_avg_sq)[i >> 2] = v;
    } else {
        for (int64_t j = i; j < i + 4 && j < t.n; ++j) {
            float p = t.param[j], m = t.exp_avg[j], v = t.exp_avg_sq[j];
            adam_one(p, t.grad[j], m, v, c);
            t.param[j] = p;
            t.exp_avg[j] = m;
            t.exp_avg_sq[j] = v;
        }
    }
}

// Post-step projections, one thread per Gaussian (scripts/train.py:168-183):
//   densities < 0 -> 0; quaternions normalised, a zero quaternion becoming (1,0,0,0)
//   (classes/point_cloud.py:703-706); lobe axes normalised, a zero axis becoming (0,0,1)
//   (:721-724); log-scales clamped below by the per-Gaussian filter_3D and above by max_scale.
__global__ void __launch_bounds__(kOptBlock) k_project(rg_projection pr) {
    const int64_t g = (int64_t)blockIdx.x * kOptBlock + threadIdx.x;
    if (g >= pr.n) return;
    if (pr.densities) {
        const float d = pr.densities[g];
        if (d < 0.0f) pr.densities[g] = 0.0f;
    }
    if (pr.quats && pr.normalize_quaternion) {
        float* q = pr.quats + 4 * g;
        float q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
        float nq = sqrtf(((q0 * q0 + q1 * q1) + q2 * q2) + q3 * q3);
        if (nq == 0.0f) {
            q0 = 1.0f;
            q1 = q2 = q3 = 0.0f;
            nq = 1.0f;
        }
        q[0] = q0 / nq;
        q[1] = q1 / nq;
        q[2] = q2 / nq;
        q[3] = q3 / nq;
    }
    if (pr.log_scales) {
        const float lo = pr.filter_3d ? pr.filter_3d[g] : -INFINITY;
        for (int k = 0; k < 3; ++k) {
            float s = pr.log_scales[3 * g + k];
            if (s < lo) s = lo;
            if (s > pr.max_scale) s = pr.max_scale;
            pr.log_scales[3 * g + k] = s;
        }
    }
    if (pr.lobe_axis && pr.normalize_lobe_axis) {
        for (int l = 0; l < pr.num_sg; ++l) {
            float* a = pr.lobe_axis + (g * pr.num_sg + l) * 3;
            float ax = a[0], ay = a[1], az = a[2];
            float na = sqrtf((ax * ax + ay * ay) + az * az);
            if (na == 0.0f) {
                ax = 0.0f;
                ay = 0.0f;
                az = 1.0f;
                na = 1.0f;
            }
            a[0] = ax / na;
            a[1] = ay / na;
            a[2] = az / na;
        }
    }
}

}  // namespace

hipError_t launch_adam(const AdamBatch& b, int64_t total_items, hipStream_t s) {
    if (total_items <= 0) return hipSuccess;
    const int64_t nb = (total_items + kOptBlock - 1) / kOptBlock;
    hipLaunchKernelGGL(k_adam, dim3((unsigned)nb), dim3(kOptBlock), 0, s, b, total_items);
    return hipGetLastError();
}

hipError_t launch_project(const rg_projection& pr, hipStream_t s) {
    if (pr.n <= 0) return hipSuccess;
    const int64_t nb = (pr.n + kOptBlock - 1) / kOptBlock;
    hipLaunchKernelGGL(k_project, dim3((unsigned)nb), dim3(kOptBlock), 0, s, pr);
    return hipGetLastError();
}

}  // namespace rg
