// Per-Gaussian kernels: ellipsoid preprocessing + scene box, per-view SH/SG colour, and the
// gradient finish (SH/SG backward, view-direction position term, un-permute).
//
// These stream the boundary tensors (N x attribute arrays) with coalesced loads, one thread per
// Gaussian in original order, and scatter the derived 64-B records into Morton slots.  They are
// HBM-bound; nothing here is a contraction.
#include <hip/hip_runtime.h>
#include <math.h>

#include "rg_device.h"
#include "rg_internal.h"

namespace rg {
namespace {

constexpr int kBlock = 256;

// Ordered-int encoding so float min/max can use integer atomics.
__device__ __forceinline__ int f2o(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float o2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

__global__ void k_init_box(float* box) {
    int* b = reinterpret_cast<int*>(box);
    const int t = threadIdx.x;
    if (t < 3) b[t] = f2o(INFINITY);
    else if (t < 6) b[t] = f2o(-INFINITY);
}

__global__ void k_decode_box(float* box) {
    int* b = reinterpret_cast<int*>(box);
    const int t = threadIdx.x;
    if (t < 6) box[t] = o2f(b[t]);
}

// quaternion_to_matrix (reference cuda_train/forward/gaussians_aabb.cu:296-310) and the
// ellipsoid AABB of compute_ellipsoids_bbox (optix_utils.py:54-66).
__global__ void __launch_bounds__(kBlock) k_preprocess(int n, const float* __restrict__ pos,
                                                       const float* __restrict__ scl, const float* __restrict__ quat,
                                                       const float* __restrict__ dens, const int* __restrict__ iperm,
                                                       float thr, float4* __restrict__ geo, float4* __restrict__ brec,
                                                       float4* __restrict__ leafbox, float* box) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (g < n) {
        const float mx = pos[3 * g], my = pos[3 * g + 1], mz = pos[3 * g + 2];
        const float sx = scl[3 * g], sy = scl[3 * g + 1], sz = scl[3 * g + 2];
        const float4 q = reinterpret_cast<const float4*>(quat)[g];
        const float sigma = dens[g];
        const float r = q.x, i = q.y, j = q.z, k = q.w;
        // columns of R
        const float c0x = 1.0f - 2.0f * (j * j + k * k), c0y = 2.0f * (i * j + r * k), c0z = 2.0f * (i * k - r * j);
        const float c1x = 2.0f * (i * j - r * k), c1y = 1.0f - 2.0f * (i * i + k * k), c1z = 2.0f * (j * k + r * i);
        const float c2x = 2.0f * (i * k + r * j), c2y = 2.0f * (j * k - r * i), c2z = 1.0f - 2.0f * (i * i + j * j);
        const float ratio = sigma / thr;
        const float l2 = logf(ratio * ratio);
        const bool valid = l2 > 0.0f && l2 < INFINITY;
        const float delta = valid ? sqrtf(l2) : 0.0f;
        const float h = valid ? 0.5f * l2 : 0.0f;
        // M' rows = col_k(R) / (s_k * delta)   (reference: scales*sqrtf(logf(ratio^2)), inverse)
        const float ispx = valid ? 1.0f / (sx * delta) : 0.0f;
        const float ispy = valid ? 1.0f / (sy * delta) : 0.0f;
        const float ispz = valid ? 1.0f / (sz * delta) : 0.0f;
        const int slot = iperm[g];
        geo[4 * slot + 0] = make_float4(mx, my, mz, sigma);
        geo[4 * slot + 1] = make_float4(c0x * ispx, c0y * ispx, c0z * ispx, h);
        geo[4 * slot + 2] = make_float4(c1x * ispy, c1y * ispy, c1z * ispy, 0.0f);
        geo[4 * slot + 3] = make_float4(c2x * ispz, c2y * ispz, c2z * ispz, 0.0f);
        brec[2 * slot + 0] = make_float4(sx, sy, sz, delta);
        brec[2 * slot + 1] = q;
        // half extent on world axis a = || (R_a0 s'_0, R_a1 s'_1, R_a2 s'_2) ||, s' = delta * s
        // (reference uses max(0, ln((sigma/thr)^2)) so invalid Gaussians are a point box at mu)
        const float spx = delta * sx, spy = delta * sy, spz = delta * sz;
        const float hx = sqrtf((c0x * spx) * (c0x * spx) + (c1x * spy) * (c1x * spy) + (c2x * spz) * (c2x * spz));
        const float hy = sqrtf((c0y * spx) * (c0y * spx) + (c1y * spy) * (c1y * spy) + (c2y * spz) * (c2y * spz));
        const float hz = sqrtf((c0z * spx) * (c0z * spx) + (c1z * spy) * (c1z * spy) + (c2z * spz) * (c2z * spz));
        lo[0] = mx - hx;
        lo[1] = my - hy;
        lo[2] = mz - hz;
        hi[0] = mx + hx;
        hi[1] = my + hy;
        hi[2] = mz + hz;
        if (valid) {
            // Pad so that FP32 rounding in the predicate and in the box test can never cull a
            // Gaussian the predicate accepts: relative on the extent, absolute on the magnitude.
            const float mag = fmaxf(fabsf(mx), fmaxf(fabsf(my), fabsf(mz))) + fmaxf(hx, fmaxf(hy, hz));
            const float pa = 3e-5f * (1.0f + mag);
            leafbox[2 * slot + 0] =
                make_float4(mx - (hx * 1.001f + pa), my - (hy * 1.001f + pa), mz - (hz * 1.001f + pa), 0.0f);
            leafbox[2 * slot + 1] =
                make_float4(mx + (hx * 1.001f + pa), my + (hy * 1.001f + pa), mz + (hz * 1.001f + pa), 0.0f);
        } else {
            // never hit: an empty box that unions to nothing
            leafbox[2 * slot + 0] = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
            leafbox[2 * slot + 1] = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
        }
    }
    // scene box (render_optix.py:105-106): block reduction then one atomic per block
    __shared__ float red[6][kBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    float v[6] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
    for (int a = 0; a < 6; ++a) {
        float x = v[a];
        for (int o = 32; o > 0; o >>= 1) {
            const float y = __shfl_xor(x, o, 64);
            x = a < 3 ? fminf(x, y) : fmaxf(x, y);
        }
        if (lane == 0) red[a][w] = x;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        float x = red[a][0];
        for (int i = 1; i < kBlock / 64; ++i) x = a < 3 ? fminf(x, red[a][i]) : fmaxf(x, red[a][i]);
        int* b = reinterpret_cast<int*>(box);
        if (a < 3) atomicMin(&b[a], f2o(x));
        else atomicMax(&b[a], f2o(x));
    }
}

// ------------------------------------------------------------------ SH / SG colour -----------

__constant__ float cC0 = 0.28209479177387814f;
__constant__ float cC1 = 0.4886025119029199f;
__constant__ float cC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f, -1.0925484305920792f,
                             0.5462742152960396f};
__constant__ float cC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f, 0.3731763325901154f,
                             -0.4570457994644658f, 1.445305721320277f, -0.5900435899266435f};

// Real SH basis, reference order and signs (utils/sh_utils.cu:60-117).
__device__ __forceinline__ void sh_basis(float x, float y, float z, int deg, float* Y) {
    Y[0] = cC0;
    if (deg < 1) return;
    Y[1] = -cC1 * y;
    Y[2] = cC1 * z;
    Y[3] = -cC1 * x;
    if (deg < 2) return;
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    Y[4] = cC2[0] * xy;
    Y[5] = cC2[1] * yz;
    Y[6] = cC2[2] * (2.0f * zz - xx - yy);
    Y[7] = cC2[3] * xz;
    Y[8] = cC2[4] * (xx - yy);
    if (deg < 3) return;
    Y[9] = cC3[0] * y * (3.0f * xx - yy);
    Y[10] = cC3[1] * xy * z;
    Y[11] = cC3[2] * y * (4.0f * zz - xx - yy);
    Y[12] = cC3[3] * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
    Y[13] = cC3[4] * x * (4.0f * zz - xx - yy);
    Y[14] = cC3[5] * z * (xx - yy);
    Y[15] = cC3[6] * x * (xx - 3.0f * yy);
}

// forward_sh (utils/sh_utils.cu:25-119): rgb = 0.5 + sum_l a_l exp(lam_l (nu_l . v - 1)) + sum_k
// sh_k Y_k(v), v = normalize(mu - eye).  All views of one Gaussian in one thread.
__global__ void __launch_bounds__(kBlock) k_sh_forward(int n, int deg, int L, const float* __restrict__ pos,
                                                       const float* __restrict__ sh, const float* __restrict__ sg,
                                                       const float* __restrict__ lam, const float* __restrict__ axis,
                                                       const int* __restrict__ iperm, const rg_camera* __restrict__ cams,
                                                       int num_views, float4* __restrict__ rgb) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= n) return;
    const int K = (deg + 1) * (deg + 1);
    float c[3][16];
    for (int ch = 0; ch < 3; ++ch)
        for (int k = 0; k < 16; ++k) c[ch][k] = k < K ? sh[(size_t)g * 3 * K + ch * K + k] : 0.0f;
    const float mx = pos[3 * g], my = pos[3 * g + 1], mz = pos[3 * g + 2];
    const int slot = iperm[g];
    for (int v = 0; v < num_views; ++v) {
        const rg_camera& cam = cams[v];
        float dx = mx - cam.eye[0], dy = my - cam.eye[1], dz = mz - cam.eye[2];
        const float len = sqrtf(dx * dx + dy * dy + dz * dz);
        dx /= len;
        dy /= len;
        dz /= len;
        float Y[16];
        sh_basis(dx, dy, dz, deg, Y);
        float out[3];
        for (int ch = 0; ch < 3; ++ch) {
            float acc = 0.0f;
            for (int l = 0; l < L; ++l) {
                const float a = sg[(size_t)g * 3 * L + ch * L + l];
                const float lm = lam[(size_t)g * L + l];
                const float* ax = axis + (size_t)g * 3 * L + 3 * l;
                const float dp = dx * ax[0] + dy * ax[1] + dz * ax[2];
                acc += a * expf(lm * (dp - 1.0f));
            }
            for (int k = 0; k < K; ++k) acc += c[ch][k] * Y[k];
            out[ch] = acc + 0.5f;
        }
        rgb[(size_t)v * n + slot] = make_float4(out[0], out[1], out[2], 0.0f);
    }
}

// backward_sh (utils/sh_utils.cu:121-315) summed over views, the position term through the view
// direction (render_optix.py:170-175), and the un-permute of the ray-caster gradients.  Clears
// the accumulators it reads.
__global__ void __launch_bounds__(kBlock) k_finish(int n, int deg, int L, const float* __restrict__ pos,
                                                   const float* __restrict__ sh, const float* __restrict__ sg,
                                                   const float* __restrict__ lam, const float* __restrict__ axis,
                                                   const int* __restrict__ iperm, const rg_camera* __restrict__ cams,
                                                   int num_views, DevGrad gr, int exact, rg_grads out) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= n) return;
    const int K = (deg + 1) * (deg + 1);
    const int slot = iperm[g];
    float dsh[3][16];
    for (int ch = 0; ch < 3; ++ch)
        for (int k = 0; k < 16; ++k) dsh[ch][k] = 0.0f;
    float dsg[3][8], dlam[8], dax[8][3];
    const int Lc = L < 8 ? L : 8;
    for (int l = 0; l < 8; ++l) {
        dlam[l] = 0.0f;
        for (int k = 0; k < 3; ++k) {
            dsg[k][l] = 0.0f;
            dax[l][k] = 0.0f;
        }
    }
    const float mx = pos[3 * g], my = pos[3 * g + 1], mz = pos[3 * g + 2];
    float dpx = 0.0f, dpy = 0.0f, dpz = 0.0f;
    for (int v = 0; v < num_views; ++v) {
        float4* p = gr.drgb + (size_t)v * n + slot;
        const float4 d4 = *p;
        *p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float dL[3] = {d4.x, d4.y, d4.z};
        if (dL[0] == 0.0f && dL[1] == 0.0f && dL[2] == 0.0f) continue;  // sh_utils.cu:129-131
        const rg_camera& cam = cams[v];
        const float d0x = mx - cam.eye[0], d0y = my - cam.eye[1], d0z = mz - cam.eye[2];
        const float len = sqrtf(d0x * d0x + d0y * d0y + d0z * d0z);
        const float x = d0x / len, y = d0y / len, z = d0z / len;
        float Y[16];
        sh_basis(x, y, z, deg, Y);
        for (int ch = 0; ch < 3; ++ch)
            for (int k = 0; k < K; ++k) dsh[ch][k] += Y[k] * dL[ch];
        // dRGB/d dir per channel
        float gx[3] = {0, 0, 0}, gy[3] = {0, 0, 0}, gz[3] = {0, 0, 0};
        if (deg > 0) {
            for (int ch = 0; ch < 3; ++ch) {
                const float* s = sh + (size_t)g * 3 * K + ch * K;
                gx[ch] = -cC1 * s[3];
                gy[ch] = -cC1 * s[1];
                gz[ch] = cC1 * s[2];
                if (deg > 1) {
                    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
                    gx[ch] += cC2[0] * y * s[4] + cC2[2] * 2.0f * -x * s[6] + cC2[3] * z * s[7] + cC2[4] * 2.0f * x * s[8];
                    gy[ch] += cC2[0] * x * s[4] + cC2[1] * z * s[5] + cC2[2] * 2.0f * -y * s[6] + cC2[4] * 2.0f * -y * s[8];
                    gz[ch] += cC2[1] * y * s[5] + cC2[2] * 2.0f * 2.0f * z * s[6] + cC2[3] * x * s[7];
                    if (deg > 2) {
                        const float y9 = exact ? 3.0f * (xx - yy) : (3.0f * xx - yy);
                        gx[ch] += cC3[0] * s[9] * 3.0f * 2.0f * xy + cC3[1] * s[10] * yz + cC3[2] * s[11] * -2.0f * xy +
                                  cC3[3] * s[12] * -3.0f * 2.0f * xz + cC3[4] * s[13] * (-3.0f * xx + 4.0f * zz - yy) +
                                  cC3[5] * s[14] * 2.0f * xz + cC3[6] * s[15] * 3.0f * (xx - yy);
                        gy[ch] += cC3[0] * s[9] * y9 + cC3[1] * s[10] * xz + cC3[2] * s[11] * (-3.0f * yy + 4.0f * zz - xx) +
                                  cC3[3] * s[12] * -3.0f * 2.0f * yz + cC3[4] * s[13] * -2.0f * xy +
                                  cC3[5] * s[14] * -2.0f * yz + cC3[6] * s[15] * -3.0f * 2.0f * xy;
                        gz[ch] += cC3[1] * s[10] * xy + cC3[2] * s[11] * 4.0f * 2.0f * yz +
                                  cC3[3] * s[12] * 3.0f * (2.0f * zz - xx - yy) + cC3[4] * s[13] * 4.0f * 2.0f * xz +
                                  cC3[5] * s[14] * (xx - yy);
                    }
                }
            }
        }
        float ddx = dL[0] * gx[0] + dL[1] * gx[1] + dL[2] * gx[2];
        float ddy = dL[0] * gy[0] + dL[1] * gy[1] + dL[2] * gy[2];
        float ddz = dL[0] * gz[0] + dL[1] * gz[1] + dL[2] * gz[2];
        for (int l = 0; l < Lc; ++l) {
            const float* ax = axis + (size_t)g * 3 * L + 3 * l;
            const float lm = lam[(size_t)g * L + l];
            const float dm1 = x * ax[0] + y * ax[1] + z * ax[2] - 1.0f;
            const float w = expf(dm1 * lm);
            float dLdw = 0.0f;
            for (int ch = 0; ch < 3; ++ch) {
                dsg[ch][l] += dL[ch] * w;
                dLdw += dL[ch] * sg[(size_t)g * 3 * L + ch * L + l];
            }
            dax[l][0] += dLdw * lm * x * w;
            dax[l][1] += dLdw * lm * y * w;
            dax[l][2] += dLdw * lm * z * w;
            dlam[l] += dLdw * dm1 * w;
            ddx += dLdw * lm * w * ax[0];
            ddy += dLdw * lm * w * ax[1];
            ddz += dLdw * lm * w * ax[2];
        }
        // d normalize(d0)/d d0 = (|d0|^2 I - d0 d0^T) / |d0|^3
        const float s2 = d0x * d0x + d0y * d0y + d0z * d0z;
        const float inv = 1.0f / sqrtf(s2 * s2 * s2);
        dpx += ((s2 - d0x * d0x) * ddx - d0y * d0x * ddy - d0z * d0x * ddz) * inv;
        dpy += (-d0x * d0y * ddx + (s2 - d0y * d0y) * ddy - d0z * d0y * ddz) * inv;
        dpz += (-d0x * d0z * ddx - d0y * d0z * ddy + (s2 - d0z * d0z) * ddz) * inv;
    }
    for (int ch = 0; ch < 3; ++ch)
        for (int k = 0; k < K; ++k) out.sh[(size_t)g * 3 * K + ch * K + k] = dsh[ch][k];
    if (L > 0 && out.sg) {
        for (int l = 0; l < L; ++l) {
            const bool inr = l < 8;
            for (int ch = 0; ch < 3; ++ch) out.sg[(size_t)g * 3 * L + ch * L + l] = inr ? dsg[ch][l] : 0.0f;
            out.sharpness[(size_t)g * L + l] = inr ? dlam[l] : 0.0f;
            for (int k = 0; k < 3; ++k) out.lobe_axis[(size_t)g * 3 * L + 3 * l + k] = inr ? dax[l][k] : 0.0f;
        }
    }
    const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4* a = reinterpret_cast<float4*>(gr.acc + (size_t)kAccStride * slot);
    const float4 a0 = a[0], a1 = a[1], a2 = a[2];
    a[0] = z4;
    a[1] = z4;
    a[2] = z4;
    out.positions[3 * g + 0] = a0.x + dpx;
    out.positions[3 * g + 1] = a0.y + dpy;
    out.positions[3 * g + 2] = a0.z + dpz;
    out.densities[g] = a0.w;
    out.scales[3 * g + 0] = a1.x;
    out.scales[3 * g + 1] = a1.y;
    out.scales[3 * g + 2] = a1.z;
    reinterpret_cast<float4*>(out.quats)[g] = a2;
}

// Caller-supplied colour table of one view: rgb (n,3) in original order -> slot order.
__global__ void __launch_bounds__(kBlock) k_set_colors(int n, const float* __restrict__ rgb,
                                                       const int* __restrict__ iperm, float4* __restrict__ table) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= n) return;
    table[iperm[g]] = make_float4(rgb[3 * g], rgb[3 * g + 1], rgb[3 * g + 2], 0.0f);
}

// make_color of the viewer (cuda_gui/helpers.h:35-57, the sRGB step commented out there):
// clamp to [0, 1] as max(0, min(x, 1)) (so NaN -> 1), quantise x -> min(floor(256 x), 255),
// alpha 255.
__global__ void __launch_bounds__(kBlock) k_rgba8(int64_t n, const float* __restrict__ rgb,
                                                  uchar4* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    unsigned char q[3];
    for (int c = 0; c < 3; ++c) {
        const float x = fmaxf(0.0f, fminf(rgb[3 * i + c], 1.0f));
        q[c] = (unsigned char)min((unsigned int)(x * 256.0f), 255u);
    }
    out[i] = make_uchar4(q[0], q[1], q[2], 255);
}

__global__ void k_zero(float4* p, int64_t count4) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count4) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ void k_export_pred(int n, const int* __restrict__ iperm, const float4* __restrict__ geo, float* mu,
                              float* Mp, int* valid, float* hlog) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int s = iperm[g];
    const float4 g0 = geo[4 * s], g1 = geo[4 * s + 1], g2 = geo[4 * s + 2], g3 = geo[4 * s + 3];
    mu[3 * g] = g0.x;
    mu[3 * g + 1] = g0.y;
    mu[3 * g + 2] = g0.z;
    const float m[9] = {g1.x, g1.y, g1.z, g2.x, g2.y, g2.z, g3.x, g3.y, g3.z};
    for (int k = 0; k < 9; ++k) Mp[9 * g + k] = m[k];
    valid[g] = g1.w > 0.0f ? 1 : 0;
    if (hlog) hlog[g] = g1.w;
}

inline int blocks(int64_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

hipError_t launch_init_box(float* box, hipStream_t s) {
    hipLaunchKernelGGL(k_init_box, dim3(1), dim3(64), 0, s, box);
    return hipGetLastError();
}

hipError_t launch_preprocess(const rg_gaussians& g, const int* iperm, float sigma_thr, float4* geo, float4* brec,
                             float4* leafbox, float* box, hipStream_t s) {
    hipLaunchKernelGGL(k_init_box, dim3(1), dim3(64), 0, s, box);
    if (g.n > 0)
        hipLaunchKernelGGL(k_preprocess, dim3(blocks(g.n, kBlock)), dim3(kBlock), 0, s, (int)g.n, g.positions, g.scales,
                           g.quats, g.densities, iperm, sigma_thr, geo, brec, leafbox, box);
    hipLaunchKernelGGL(k_decode_box, dim3(1), dim3(64), 0, s, box);
    return hipGetLastError();
}

hipError_t launch_sh_forward(const rg_gaussians& g, const int* iperm, const rg_camera* cams_dev, int num_views,
                             float4* rgb, hipStream_t s) {
    if (g.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sh_forward, dim3(blocks(g.n, kBlock)), dim3(kBlock), 0, s, (int)g.n, g.sh_degree, g.num_sg,
                       g.positions, g.sh, g.sg, g.sharpness, g.lobe_axis, iperm, cams_dev, num_views, rgb);
    return hipGetLastError();
}

hipError_t launch_finish_grads(const rg_gaussians& g, const int* iperm, const rg_camera* cams_dev, int num_views,
                               const DevGrad& gr, int exact, const rg_grads& out, hipStream_t s) {
    if (g.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_finish, dim3(blocks(g.n, kBlock)), dim3(kBlock), 0, s, (int)g.n, g.sh_degree, g.num_sg,
                       g.positions, g.sh, g.sg, g.sharpness, g.lobe_axis, iperm, cams_dev, num_views, gr, exact,
                       out);
    return hipGetLastError();
}

hipError_t launch_set_colors(int n, const float* rgb, const int* iperm, float4* table, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_colors, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, rgb, iperm, table);
    return hipGetLastError();
}

hipError_t launch_rgba8(int64_t n, const float* rgb, uchar4* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rgba8, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, n, rgb, out);
    return hipGetLastError();
}

hipError_t launch_zero(float* p, int64_t count, hipStream_t s) {
    const int64_t c4 = (count + 3) / 4;
    if (c4 == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero, dim3(blocks(c4, 256)), dim3(256), 0, s, reinterpret_cast<float4*>(p), c4);
    return hipGetLastError();
}

hipError_t launch_export_predicate(int n, const int* iperm, const float4* geo, float* mu, float* Mp, int* valid,
                                   float* hlog,
                                   hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_export_pred, dim3(blocks(n, 256)), dim3(256), 0, s, n, iperm, geo, mu, Mp, valid, hlog);
    return hipGetLastError();
}

}  // namespace rg
