// Photometric loss of the training loop and its gradient w.r.t. the rendered ray colours:
//   L = (1 - lambda) * mean|x - y| + lambda * (1 - mean SSIM(x, y))
// with SSIM over an 11x11 Gaussian window (sigma 1.5), zero padding, C1 = 0.01^2, C2 = 0.03^2
// (reference utils/loss_utils.py:17-64, used at scripts/train.py:110-112 with lambda = 0.2), and
// the supersampling reduction in front of it (utils/general_utils.py:489-502): the image pixel
// (x, y) is the mean of the fx*fy rays (x*fx + i, y*fy + j) of a row-major (H*fy) x (W*fx) grid.
//
// SSIM gradient in closed form.  Per pixel q, with mu1 = w*x, mu2 = w*y, E11 = w*x^2,
// E22 = w*y^2, E12 = w*xy (w the window, * the zero-padded "same" correlation):
//   A1 = 2 mu1 mu2 + C1, A2 = 2 (E12 - mu1 mu2) + C2, B1 = mu1^2 + mu2^2 + C1,
//   B2 = (E11 - mu1^2) + (E22 - mu2^2) + C2, s = A1 A2 / (B1 B2)
//   Gm  = ds/dmu1 = (2 mu2 (A2 - A1) - 2 mu1 s (B2 - B1)) / (B1 B2)
//   G11 = ds/dE11 = -s / B2,   G12 = ds/dE12 = 2 A1 / (B1 B2)
// and, the window being symmetric, d(mean s)/dx(p) = (w*Gm + 2 x w*G11 + y w*G12)(p) / N.
// Every convolution is separable (the window is the outer product of the 1-D Gaussian), done as
// a horizontal and a vertical 11-tap pass over channel-planar images.  All passes are HBM-bound
// streaming kernels (one thread per pixel and plane).

#include <hip/hip_runtime.h>

#include "rg_internal.h"

namespace rg {
namespace {

constexpr int kLB = 256;
constexpr int kHalf = 5;

struct Win {
    float w[2 * kHalf + 1];
};

// Supersampling reduction into channel-planar x, and the target (3,H,W) copied alongside.
__global__ void __launch_bounds__(kLB) k_pack(const float* __restrict__ rays, int W, int H, int fx, int fy,
                                              float* __restrict__ xp) {
    const int64_t p = (int64_t)blockIdx.x * kLB + threadIdx.x;
    if (p >= (int64_t)W * H) return;
    const int x = (int)(p % W), y = (int)(p / W);
    const int64_t RW = (int64_t)W * fx;
    float s[3] = {0.0f, 0.0f, 0.0f};
    for (int j = 0; j < fy; ++j)
        for (int i = 0; i < fx; ++i) {
            const int64_t r = ((int64_t)y * fy + j) * RW + (int64_t)x * fx + i;
            s[0] += rays[3 * r];
            s[1] += rays[3 * r + 1];
            s[2] += rays[3 * r + 2];
        }
    const float inv = 1.0f / (float)(fx * fy);
    const int64_t HW = (int64_t)W * H;
    for (int c = 0; c < 3; ++c) xp[c * HW + p] = s[c] * inv;
}

// Horizontal pass of the five SSIM moments: out[k][c][y][x] for k = x, y, xx, yy, xy.
__global__ void __launch_bounds__(kLB) k_h_moments(const float* __restrict__ xp, const float* __restrict__ yp, int W,
                                                   int H, Win win, float* __restrict__ out) {
    const int64_t HW = (int64_t)W * H;
    const int64_t t = (int64_t)blockIdx.x * kLB + threadIdx.x;
    if (t >= 3 * HW) return;
    const int64_t c = t / HW, p = t % HW;
    const int x = (int)(p % W);
    const float* xr = xp + c * HW + (p - x);
    const float* yr = yp + c * HW + (p - x);
    float m1 = 0.0f, m2 = 0.0f, e11 = 0.0f, e22 = 0.0f, e12 = 0.0f;
#pragma unroll
    for (int k = -kHalf; k <= kHalf; ++k) {
        const int xx = x + k;
        if (xx < 0 || xx >= W) continue;
        const float a = xr[xx], b = yr[xx], wk = win.w[k + kHalf];
        m1 += wk * a;
        m2 += wk * b;
        e11 += wk * (a * a);
        e22 += wk * (b * b);
        e12 += wk * (a * b);
    }
    const int64_t o = c * HW + p, S = 3 * HW;
    out[o] = m1;
    out[S + o] = m2;
    out[2 * S + o] = e11;
    out[3 * S + o] = e22;
    out[4 * S + o] = e12;
}

// Vertical pass of the moments, SSIM map, its three gradient maps and per-block partial sums of
// SSIM and |x - y| (double, so the mean does not depend on the block count's rounding).
__global__ void __launch_bounds__(kLB) k_v_ssim(const float* __restrict__ hm, const float* __restrict__ xp,
                                                const float* __restrict__ yp, int W, int H, Win win,
                                                float* __restrict__ gmap, double* __restrict__ partial) {
    __shared__ double red[2][kLB];
    const int64_t HW = (int64_t)W * H, S = 3 * HW;
    const int64_t t = (int64_t)blockIdx.x * kLB + threadIdx.x;
    double s_sum = 0.0, l1_sum = 0.0;
    if (t < S) {
        const int64_t c = t / HW, p = t % HW;
        const int y = (int)(p / W), x = (int)(p % W);
        float mu1 = 0.0f, mu2 = 0.0f, e11 = 0.0f, e22 = 0.0f, e12 = 0.0f;
#pragma unroll
        for (int k = -kHalf; k <= kHalf; ++k) {
            const int yy = y + k;
            if (yy < 0 || yy >= H) continue;
            const int64_t q = c * HW + (int64_t)yy * W + x;
            const float wk = win.w[k + kHalf];
            mu1 += wk * hm[q];
            mu2 += wk * hm[S + q];
            e11 += wk * hm[2 * S + q];
            e22 += wk * hm[3 * S + q];
            e12 += wk * hm[4 * S + q];
        }
        const float C1 = 0.0001f, C2 = 0.0009f;
        const float mu12 = mu1 * mu2;
        const float A1 = 2.0f * mu12 + C1;
        const float A2 = 2.0f * (e12 - mu12) + C2;
        const float B1 = mu1 * mu1 + mu2 * mu2 + C1;
        const float B2 = (e11 - mu1 * mu1) + (e22 - mu2 * mu2) + C2;
        const float den = B1 * B2;
        const float s = (A1 * A2) / den;
        const int64_t o = c * HW + p;
        gmap[o] = (2.0f * mu2 * (A2 - A1) - 2.0f * mu1 * s * (B2 - B1)) / den;
        gmap[S + o] = -s / B2;
        gmap[2 * S + o] = 2.0f * A1 / den;
        s_sum = (double)s;
        l1_sum = (double)fabsf(xp[o] - yp[o]);
    }
    red[0][threadIdx.x] = s_sum;
    red[1][threadIdx.x] = l1_sum;
    __syncthreads();
    for (int off = kLB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            red[0][threadIdx.x] += red[0][threadIdx.x + off];
            red[1][threadIdx.x] += red[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = red[0][0];
        partial[2 * blockIdx.x + 1] = red[1][0];
    }
}

// Horizontal pass of the three gradient maps.
__global__ void __launch_bounds__(kLB) k_h3(const float* __restrict__ gmap, int W, int H, Win win,
                                            float* __restrict__ out) {
    const int64_t HW = (int64_t)W * H, S = 3 * HW;
    const int64_t t = (int64_t)blockIdx.x * kLB + threadIdx.x;
    if (t >= S) return;
    const int64_t p = t % HW;
    const int x = (int)(p % W);
    const int64_t row = t - x;
    float a = 0.0f, b = 0.0f, d = 0.0f;
#pragma unroll
    for (int k = -kHalf; k <= kHalf; ++k) {
        const int xx = x + k;
        if (xx < 0 || xx >= W) continue;
        const float wk = win.w[k + kHalf];
        a += wk * gmap[row + xx];
        b += wk * gmap[S + row + xx];
        d += wk * gmap[2 * S + row + xx];
    }
    out[t] = a;
    out[S + t] = b;
    out[2 * S + t] = d;
}

// Vertical pass, gradient assembly and scatter to the supersampled rays.
__global__ void __launch_bounds__(kLB) k_v_grad(const float* __restrict__ h3, const float* __restrict__ xp,
                                                const float* __restrict__ yp, int W, int H, int fx, int fy, Win win,
                                                float lambda, float* __restrict__ dl_rays) {
    const int64_t HW = (int64_t)W * H, S = 3 * HW;
    const int64_t t = (int64_t)blockIdx.x * kLB + threadIdx.x;
    if (t >= S) return;
    const int64_t c = t / HW, p = t % HW;
    const int y = (int)(p / W), x = (int)(p % W);
    float a = 0.0f, b = 0.0f, d = 0.0f;
#pragma unroll
    for (int k = -kHalf; k <= kHalf; ++k) {
        const int yy = y + k;
        if (yy < 0 || yy >= H) continue;
        const int64_t q = c * HW + (int64_t)yy * W + x;
        const float wk = win.w[k + kHalf];
        a += wk * h3[q];
        b += wk * h3[S + q];
        d += wk * h3[2 * S + q];
    }
    const float xv = xp[t], yv = yp[t];
    const float invN = 1.0f / (float)S;
    const float ds = (a + 2.0f * xv * b + yv * d) * invN;
    const float e = xv - yv;
    const float dl1 = (e > 0.0f ? invN : (e < 0.0f ? -invN : 0.0f));
    const float g = (1.0f - lambda) * dl1 - lambda * ds;
    const float gr = g / (float)(fx * fy);
    const int64_t RW = (int64_t)W * fx;
    for (int j = 0; j < fy; ++j)
        for (int i = 0; i < fx; ++i) {
            const int64_t r = ((int64_t)y * fy + j) * RW + (int64_t)x * fx + i;
            dl_rays[3 * r + c] = gr;
        }
}

__global__ void k_loss_final(const double* __restrict__ partial, int nb, int64_t N, float lambda,
                             float* __restrict__ loss, float* __restrict__ parts) {
    __shared__ double red[2][kLB];
    double s = 0.0, l = 0.0;
    for (int i = threadIdx.x; i < nb; i += kLB) {
        s += partial[2 * i];
        l += partial[2 * i + 1];
    }
    red[0][threadIdx.x] = s;
    red[1][threadIdx.x] = l;
    __syncthreads();
    for (int off = kLB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            red[0][threadIdx.x] += red[0][threadIdx.x + off];
            red[1][threadIdx.x] += red[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double ssim = red[0][0] / (double)N, l1 = red[1][0] / (double)N;
        loss[0] = (float)((1.0 - lambda) * l1 + lambda * (1.0 - ssim));
        if (parts) {
            parts[0] = (float)l1;
            parts[1] = (float)ssim;
        }
    }
}

}  // namespace

int64_t loss_workspace_bytes(int W, int H) {
    const int64_t S = 3 * (int64_t)W * H;
    const int64_t nb = (S + kLB - 1) / kLB;
    // x planes, moments (5), gradient maps (3), their horizontal pass (3), block partials
    return (int64_t)sizeof(float) * S * (1 + 5 + 3 + 3) + (int64_t)sizeof(double) * 2 * nb + 256;
}

hipError_t launch_loss_l1_ssim(const float* rays, const float* target, int W, int H, int fx, int fy,
                               const float* window11, float lambda, float* loss, float* parts, float* dl_rays,
                               void* ws, hipStream_t s) {
    const int64_t HW = (int64_t)W * H, S = 3 * HW;
    Win win;
    for (int k = 0; k < 2 * kHalf + 1; ++k) win.w[k] = window11[k];
    float* xp = static_cast<float*>(ws);
    float* hm = xp + S;
    float* gm = hm + 5 * S;
    float* h3 = gm + 3 * S;
    uintptr_t pa = reinterpret_cast<uintptr_t>(h3 + 3 * S);
    pa = (pa + 255) & ~(uintptr_t)255;
    double* partial = reinterpret_cast<double*>(pa);
    const int64_t nbp = (HW + kLB - 1) / kLB, nbs = (S + kLB - 1) / kLB;
    hipLaunchKernelGGL(k_pack, dim3((unsigned)nbp), dim3(kLB), 0, s, rays, W, H, fx, fy, xp);
    hipLaunchKernelGGL(k_h_moments, dim3((unsigned)nbs), dim3(kLB), 0, s, xp, target, W, H, win, hm);
    hipLaunchKernelGGL(k_v_ssim, dim3((unsigned)nbs), dim3(kLB), 0, s, hm, xp, target, W, H, win, gm, partial);
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(kLB), 0, s, partial, (int)nbs, S, lambda, loss, parts);
    if (dl_rays) {
        hipLaunchKernelGGL(k_h3, dim3((unsigned)nbs), dim3(kLB), 0, s, gm, W, H, win, h3);
        hipLaunchKernelGGL(k_v_grad, dim3((unsigned)nbs), dim3(kLB), 0, s, h3, xp, target, W, H, fx, fy, win, lambda,
                           dl_rays);
    }
    return hipGetLastError();
}

}  // namespace rg
