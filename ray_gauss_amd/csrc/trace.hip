// Ray kernels of the gfx950 ray caster: forward render, recompute-based backward, fused
// train step, and the hit-set export used for verification.
//
// Per ray (one lane), the reference's per-slab loop (cuda_train/forward/gaussians_aabb.cu:521-612,
// backward/gaussians_aabb.cu:505-775) is kept exactly -- slab grid tb accumulated in FP32 from
// tenter, 8-sample buffers, thresholded density sums, per-slab transmittance exit -- but the hit
// sets are found differently: instead of one BVH traversal per slab, one traversal gathers every
// bounding-ellipsoid interval [r1, r2] overlapping a SEGMENT of m slabs into a per-lane list in
// LDS (capacity hit_depth); each slab's hit set is then {entries: max(tmin,r1) <= min(tmax,r2)},
// the reference's own predicate.  m adapts to the list fill (doubling across empty space,
// halving on overflow); a slab whose own hit set exceeds the list is processed by a per-slab
// traversal that accumulates on the fly (no storage).  Fill order is the traversal's DFS order,
// and forward and backward make identical segment decisions, so both sum each slab in the same
// order.
//
// Sampling uses the closed form of the Gaussian along the ray: with the predicate's quantities
// (t_c = b l, P = O' + b D^, D' = M' d, h = ln|sigma/thr|),
//     0.5 |M (o + t d - mu)|^2 = h |P|^2 + h |D'|^2 (t - t_c)^2,
// so each (sample, hit) costs one FMA pair and one exp.  The backward accumulates, per
// (slab, hit), four scalars (sum w phi, sum w phi tau, sum w phi tau^2, sum w B) and expands
// them into the 14 parameter gradients of backward/gaussians_aabb.cu:700-763 once.
#include <hip/hip_runtime.h>

#include "rg_device.h"
#include "rg_internal.h"

namespace rg {
namespace {

constexpr int kWave = 64;
constexpr int kMaxSeg = 4096;

enum CounterIdx { C_K = 0, C_P = 1, C_OVF = 2, C_FALLBACK = 3, C_GATHER = 4, C_LOSS = 6, C_QUEUE0 = 8 };

struct RayOut {
    float C[3];
    float depth;
    float T;
};

// Per-lane interval list in LDS, [entry][lane] so lanes access consecutive words.
template <int CAP>
struct IList {
    int g[CAP][kWave];
    float r1[CAP][kWave];
    float r2[CAP][kWave];
};

struct Counters {
    unsigned int K = 0, P = 0, ovf = 0, fallback = 0, gathers = 0;
};

// XCD (accelerator complex die) the wave runs on: hardware register XCC_ID, low 3 bits.
__device__ __forceinline__ int xcc_id() { return (int)(__builtin_amdgcn_s_getreg((20) | (0 << 6) | (3 << 11)) & 7); }

// FP32 gradient add: workgroup scope keeps the atomic in the XCD's L2 (per-XCD copies);
// device scope otherwise.
__device__ __forceinline__ void gadd(float* p, float v, bool local) {
    if (local) __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else atomicAdd(p, v);
}

__device__ __forceinline__ void load_rec(const float4* __restrict__ geo, int s, float4& g0, float4& g1, float4& g2,
                                         float4& g3) {
    g0 = geo[4 * s];
    g1 = geo[4 * s + 1];
    g2 = geo[4 * s + 2];
    g3 = geo[4 * s + 3];
}

// Accumulate one hit's contribution to the slab's sample buffers (computeBufferForward,
// forward/gaussians_aabb.cu:437-498).
template <int SPB>
__device__ __forceinline__ void accum_hit(const Isect& I, float sigma, float4 col, float tb, const DevParams& p,
                                          float* sig, float* cr, float* cg, float* cb) {
#pragma unroll
    for (int i = 0; i < SPB; ++i) {
        if (SPB > 8 && i >= p.spb) break;
        const float t = sample_t(tb, i, p.dt);
        const float tau = t - I.tc;
        const float pw = I.A + I.Cq * (tau * tau);
        const float sw = sigma * __expf(-pw);
        if (sw > p.sigma_thr) {
            sig[i] += sw;
            cr[i] += sw * col.x;
            cg[i] += sw * col.y;
            cb[i] += sw * col.z;
        }
    }
}

// colorBlending (forward/gaussians_aabb.cu:500-519) + depth (cuda_gui/gaussians_aabb.cu:415).
template <int SPB>
__device__ __forceinline__ void composite(const float* sig, const float* cr, const float* cg, const float* cb, float tb,
                                          const DevParams& p, RayOut& o) {
#pragma unroll
    for (int i = 0; i < SPB; ++i) {
        if (SPB > 8 && i >= p.spb) break;
        const float s = sig[i];
        const float alpha = 1.0f - expf(-s * p.dt);
        float r = cr[i], g = cg[i], b = cb[i];
        if (s > 0.0f) {
            r /= s;
            g /= s;
            b /= s;
        }
        const float ta = o.T * alpha;
        o.C[0] += ta * r;
        o.C[1] += ta * g;
        o.C[2] += ta * b;
        o.depth += ta * sample_t(tb, i, p.dt);
        o.T *= 1.0f - alpha;
    }
}

// Per-sample backward coefficients of one slab, from its buffers: overwrites (sig, cr, cg) with
// (A_i, Bq_i, E_i) and advances T and the colour remainder Crem past the slab.
//   A_i  = dt * dLdC . (T_i c^_i - Crem_i)      (dcolor_dsigma_aux, backward :706)
//   Bq_i = T_i alpha_i / sigma_i                (0 when sigma_i == 0)
//   E_i  = dLdC . c^_i
template <int SPB>
__device__ __forceinline__ void slab_coeffs(float* sig, float* cr, float* cg, float* cb, const DevParams& p,
                                            const float* dL, float& T, float* Crem) {
#pragma unroll
    for (int i = 0; i < SPB; ++i) {
        if (SPB > 8 && i >= p.spb) break;
        const float s = sig[i];
        const float alpha = 1.0f - expf(-s * p.dt);
        float r = cr[i], g = cg[i], b = cb[i];
        if (s > 0.0f) {
            r /= s;
            g /= s;
            b /= s;
        }
        const float A = p.dt * (dL[0] * (T * r - Crem[0]) + dL[1] * (T * g - Crem[1]) + dL[2] * (T * b - Crem[2]));
        const float Bq = s > 0.0f ? T * alpha / s : 0.0f;
        const float E = dL[0] * r + dL[1] * g + dL[2] * b;
        const float ta = T * alpha;
        Crem[0] -= ta * r;
        Crem[1] -= ta * g;
        Crem[2] -= ta * b;
        T *= 1.0f - alpha;
        sig[i] = A;
        cr[i] = Bq;
        cg[i] = E;
    }
}

// Gradient of one (slab, hit): four sufficient statistics over the slab's samples, expanded
// into d mu, d s, d q, d sigma, d rgb and scattered with FP32 atomics.
template <int SPB>
__device__ __forceinline__ void grad_hit(const Isect& I, int slot, float4 g0, float4 g1, float4 g2, float4 g3,
                                         float4 col, float tb, const DevParams& p, const float* Aq, const float* Bq,
                                         const float* Eq, const float* dL, const DevScene& sc, const DevGrad& gr,
                                         int view) {
    const float sigma = g0.w;
    const float F = dL[0] * col.x + dL[1] * col.y + dL[2] * col.z;
    float Ss = 0.0f, S1 = 0.0f, S2 = 0.0f, Sc = 0.0f;
#pragma unroll
    for (int i = 0; i < SPB; ++i) {
        if (SPB > 8 && i >= p.spb) break;
        const float t = sample_t(tb, i, p.dt);
        const float tau = t - I.tc;
        const float pw = I.A + I.Cq * (tau * tau);
        const float w = __expf(-pw);
        if (sigma * w > p.sigma_thr) {
            const float phi = Aq[i] + Bq[i] * (F - Eq[i]);
            const float wp = w * phi;
            Ss += wp;
            S1 += wp * tau;
            S2 += wp * (tau * tau);
            Sc += w * Bq[i];
        }
    }
    const float4 b0 = sc.brec[2 * slot], q = sc.brec[2 * slot + 1];
    const float d = b0.w;
    const float S0 = sigma * Ss, T1 = sigma * S1, T2 = sigma * S2;
    // y(t) = yc + tau v in the unenlarged ellipsoid frame
    const float yc[3] = {d * I.Px, d * I.Py, d * I.Pz};
    const float v[3] = {d * I.Dx, d * I.Dy, d * I.Dz};
    float Gy[3], Y[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        Gy[a] = yc[a] * S0 + v[a] * T1;
#pragma unroll
        for (int b = 0; b < 3; ++b) Y[a][b] = yc[a] * yc[b] * S0 + (yc[a] * v[b] + v[a] * yc[b]) * T1 + v[a] * v[b] * T2;
    }
    const float m[3][3] = {{g1.x, g1.y, g1.z}, {g2.x, g2.y, g2.z}, {g3.x, g3.y, g3.z}};  // rows of M'
    const float s[3] = {b0.x, b0.y, b0.z};
    float dmu[3], dcol[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) dmu[c] = d * (m[0][c] * Gy[0] + m[1][c] * Gy[1] + m[2][c] * Gy[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float f = -d / s[k];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            dcol[k][c] = f * (s[0] * s[0] * Y[0][k] * m[0][c] + s[1] * s[1] * Y[1][k] * m[1][c] +
                              s[2] * s[2] * Y[2][k] * m[2][c]);
    }
    // quaternion chain (backward/gaussians_aabb.cu:747-763), q = (r, i, j, k) = (w, x, y, z)
    const float r = q.x, qi = q.y, qj = q.z, qk = q.w;
    const float* c0 = dcol[0];
    const float* c1 = dcol[1];
    const float* c2 = dcol[2];
    const float dqr = 2.0f * (qk * c0[1] - qj * c0[2] - qk * c1[0] + qi * c1[2] + qj * c2[0] - qi * c2[1]);
    const float dqi = 2.0f * (qj * c0[1] + qk * c0[2] + qj * c1[0] + r * c1[2] + qk * c2[0] - r * c2[1]) -
                      4.0f * qi * (c1[1] + c2[2]);
    const float dqj = 2.0f * (qi * c0[1] - r * c0[2] + qi * c1[0] + qk * c1[2] + r * c2[0] + qk * c2[1]) -
                      4.0f * qj * (c0[0] + c2[2]);
    const float dqk = 2.0f * (r * c0[1] + qi * c0[2] - r * c1[0] + qj * c1[2] + qi * c2[0] + qj * c2[1]) -
                      4.0f * qk * (c0[0] + c1[1]);
    const bool local = gr.copies > 1;
    const int copy = local ? xcc_id() : 0;
    float* acc = gr.acc + (size_t)copy * gr.acc_stride + (size_t)kAccStride * slot;
    gadd(acc + 0, dmu[0], local);
    gadd(acc + 1, dmu[1], local);
    gadd(acc + 2, dmu[2], local);
    gadd(acc + 3, Ss, local);
    gadd(acc + 4, Y[0][0] / s[0], local);
    gadd(acc + 5, Y[1][1] / s[1], local);
    gadd(acc + 6, Y[2][2] / s[2], local);
    gadd(acc + 8, dqr, local);
    gadd(acc + 9, dqi, local);
    gadd(acc + 10, dqj, local);
    gadd(acc + 11, dqk, local);
    float* dr = reinterpret_cast<float*>(gr.drgb + (size_t)copy * gr.drgb_stride + (size_t)view * sc.n + slot);
    const float sc_ = sigma * Sc;
    gadd(dr + 0, dL[0] * sc_, local);
    gadd(dr + 1, dL[1] * sc_, local);
    gadd(dr + 2, dL[2] * sc_, local);
}

__device__ __forceinline__ int wave_lane() { return threadIdx.x & (kWave - 1); }

// Gathers the intervals overlapping [ta, tb] into the lane's list; returns the count, or
// CAP + 1 on overflow.
template <int CAP>
__device__ __forceinline__ int gather(const DevScene& sc, const Ray& ray, float ta, float tb, IList<CAP>& L,
                                      int lane) {
    int cnt = 0;
    traverse(sc.nodes, sc.total_nodes, ray, ta, tb, [&](int slot) -> bool {
        float4 g0, g1, g2, g3;
        load_rec(sc.geo, slot, g0, g1, g2, g3);
        Isect I;
        if (!intersect(g0, g1, g2, g3, ray, I)) return true;
        if (!overlaps(ta, tb, I.r1, I.r2)) return true;
        if (cnt >= CAP) {
            cnt = CAP + 1;
            return false;
        }
        L.g[cnt][lane] = slot;
        L.r1[cnt][lane] = I.r1;
        L.r2[cnt][lane] = I.r2;
        ++cnt;
        return true;
    });
    return cnt;
}

// Shared segment driver.  Calls slab_list(tb, tmin, tmax, cnt) for slabs covered by the list
// and slab_direct(tb, tmin, tmax) for slabs whose hit set exceeds the list.  Returns when the ray
// leaves the scene or `alive()` turns false.
template <int CAP, typename FL, typename FD, typename AL>
__device__ __forceinline__ void march(const DevScene& sc, const DevParams& p, const Ray& ray, IList<CAP>& L, int lane,
                                      Counters& ctr, FL&& slab_list, FD&& slab_direct, AL&& alive) {
    const float slab = p.dt * (float)p.spb;
    float tb = ray.tenter;
    int m = p.seg_slabs;
    if (!(add_exact(tb, slab) > tb)) return;  // slab below half an ulp of t: no progress possible
    while (tb < ray.texit && alive()) {
        const float seg_hi = fminf(ray.texit, tb + (float)m * slab);
        const int cnt = gather<CAP>(sc, ray, tb, seg_hi, L, lane);
        ctr.gathers++;
        if (cnt > CAP) {
            if (m > 1) {
                m = m >> 1;
                continue;
            }
            // one slab with more hits than the list holds: per-slab traversal
            ctr.fallback++;
            const float tmin = fmaxf(ray.tenter, tb);
            const float tmax = fminf(ray.texit, add_exact(tb, slab));
            if (tmax > ray.tenter) slab_direct(tb, tmin, tmax);
            const float nt = add_exact(tb, slab);
            if (!(nt > tb)) return;
            tb = nt;
            continue;
        }
        // slabs fully inside the gathered segment
        while (tb < ray.texit && alive()) {
            const float tmax = fminf(ray.texit, add_exact(tb, slab));
            if (tmax > seg_hi) break;
            const float tmin = fmaxf(ray.tenter, tb);
            if (cnt > 0 && tmax > ray.tenter) slab_list(tb, tmin, tmax, cnt);
            const float nt = add_exact(tb, slab);
            if (!(nt > tb)) return;
            tb = nt;
        }
        if (cnt == 0 || 4 * cnt < CAP) m = min(m * 2, kMaxSeg);
        else if (4 * cnt > 3 * CAP) m = max(m >> 1, 1);
    }
}

struct RayRef {
    int view;
    int pixel;
    int64_t out;  // output row
    bool active;
};

// Maps a work index to (view, pixel, output row).  Single-view implicit mode walks the image in
// 8x8 pixel tiles so a wavefront's rays are spatially coherent.
__device__ __forceinline__ RayRef ray_ref(const RayJob& job, int64_t r) {
    RayRef q;
    q.active = false;
    q.view = job.view;
    q.pixel = 0;
    q.out = r;
    if (r >= job.num_rays) return q;
    if (job.ray_view) {
        q.view = job.ray_view[r];
        q.pixel = job.ray_pixel[r];
        q.active = true;
        return q;
    }
    if (job.pixels) {
        q.pixel = job.pixels[r];
        q.active = true;
        return q;
    }
    q.pixel = (int)r;
    q.active = true;
    return q;
}

// Tile-ordered enumeration for the implicit full-view mode: work item w -> pixel.
__device__ __forceinline__ bool tile_pixel(int64_t w, int width, int height, int& pixel) {
    const int64_t tile = w >> 6;
    const int lane = (int)(w & 63);
    const int tiles_x = (width + 7) >> 3;
    const int tx = (int)(tile % tiles_x), ty = (int)(tile / tiles_x);
    const int px = tx * 8 + (lane & 7), py = ty * 8 + (lane >> 3);
    if (px >= width || py >= height) return false;
    pixel = py * width + px;
    return true;
}

__device__ __forceinline__ void flush_counters(const Counters& c, DevStats st) {
    // wave-level reduction, one atomic per counter per wave
    unsigned int v[5] = {c.K, c.P, c.ovf, c.fallback, c.gathers};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        unsigned int x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        v[k] = x;
    }
    if (wave_lane() == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k)
            if (v[k]) atomicAdd(&st.c[k], (unsigned long long)v[k]);
    }
}

// ------------------------------------------------------------------ per-ray bodies -----------

template <int SPB, int CAP>
__device__ void forward_ray(const DevScene& sc, const DevParams& p, const Ray& ray, const float4* __restrict__ rgbv,
                            IList<CAP>& L, int lane, Counters& ctr, RayOut& o) {
    o.C[0] = o.C[1] = o.C[2] = 0.0f;
    o.depth = 0.0f;
    o.T = 1.0f;
    if (!(ray.tenter < ray.texit)) return;
    float first_tb = ray.tenter;
    auto eval_hit = [&](int slot, const Isect& I, float tb, float tmin, float* sig, float* cr, float* cg, float* cb) {
        const float4 g0 = sc.geo[4 * slot];
        const float4 col = rgbv[slot];
        accum_hit<SPB>(I, g0.w, col, tb, p, sig, cr, cg, cb);
        if (tb == first_tb || I.r1 > tmin) ctr.P++;
    };
    auto finish_slab = [&](int nh, float tb, const float* sig, const float* cr, const float* cg, const float* cb) {
        ctr.K += nh;
        if (p.max_hits > 0 && nh > p.max_hits) ctr.ovf++;
        if (nh > 0) composite<SPB>(sig, cr, cg, cb, tb, p, o);
    };
    auto slab_list = [&](float tb, float tmin, float tmax, int cnt) {
        float sig[SPB], cr[SPB], cg[SPB], cb[SPB];
#pragma unroll
        for (int i = 0; i < SPB; ++i) sig[i] = cr[i] = cg[i] = cb[i] = 0.0f;
        int nh = 0;
        for (int e = 0; e < cnt; ++e) {
            const float r1 = L.r1[e][lane], r2 = L.r2[e][lane];
            if (!overlaps(tmin, tmax, r1, r2)) continue;
            const int slot = L.g[e][lane];
            float4 g0, g1, g2, g3;
            load_rec(sc.geo, slot, g0, g1, g2, g3);
            Isect I;
            intersect(g0, g1, g2, g3, ray, I);
            refine_sampling(g0, g1, g2, g3, ray, I);
            eval_hit(slot, I, tb, tmin, sig, cr, cg, cb);
            ++nh;
        }
        finish_slab(nh, tb, sig, cr, cg, cb);
    };
    auto slab_direct = [&](float tb, float tmin, float tmax) {
        float sig[SPB], cr[SPB], cg[SPB], cb[SPB];
#pragma unroll
        for (int i = 0; i < SPB; ++i) sig[i] = cr[i] = cg[i] = cb[i] = 0.0f;
        int nh = 0;
        traverse(sc.nodes, sc.total_nodes, ray, tmin, tmax, [&](int slot) -> bool {
            float4 g0, g1, g2, g3;
            load_rec(sc.geo, slot, g0, g1, g2, g3);
            Isect I;
            if (!intersect(g0, g1, g2, g3, ray, I)) return true;
            if (!overlaps(tmin, tmax, I.r1, I.r2)) return true;
            refine_sampling(g0, g1, g2, g3, ray, I);
            eval_hit(slot, I, tb, tmin, sig, cr, cg, cb);
            ++nh;
            return true;
        });
        finish_slab(nh, tb, sig, cr, cg, cb);
    };
    march<CAP>(sc, p, ray, L, lane, ctr, slab_list, slab_direct, [&]() { return o.T > p.t_eps; });
}

template <int SPB, int CAP>
__device__ void backward_ray(const DevScene& sc, const DevParams& p, const Ray& ray, const float4* __restrict__ rgbv,
                             const float* Cfwd, const float* dL, IList<CAP>& L, int lane, Counters& ctr,
                             const DevGrad& gr, int view) {
    if (!(ray.tenter < ray.texit)) return;
    float T = 1.0f;
    float Crem[3] = {Cfwd[0], Cfwd[1], Cfwd[2]};
    auto slab_list = [&](float tb, float tmin, float tmax, int cnt) {
        float sig[SPB], cr[SPB], cg[SPB], cb[SPB];
#pragma unroll
        for (int i = 0; i < SPB; ++i) sig[i] = cr[i] = cg[i] = cb[i] = 0.0f;
        int nh = 0;
        for (int e = 0; e < cnt; ++e) {
            const float r1 = L.r1[e][lane], r2 = L.r2[e][lane];
            if (!overlaps(tmin, tmax, r1, r2)) continue;
            const int slot = L.g[e][lane];
            float4 g0, g1, g2, g3;
            load_rec(sc.geo, slot, g0, g1, g2, g3);
            Isect I;
            intersect(g0, g1, g2, g3, ray, I);
            refine_sampling(g0, g1, g2, g3, ray, I);
            accum_hit<SPB>(I, g0.w, rgbv[slot], tb, p, sig, cr, cg, cb);
            ++nh;
        }
        if (nh == 0) return;
        slab_coeffs<SPB>(sig, cr, cg, cb, p, dL, T, Crem);
        for (int e = 0; e < cnt; ++e) {
            const float r1 = L.r1[e][lane], r2 = L.r2[e][lane];
            if (!overlaps(tmin, tmax, r1, r2)) continue;
            const int slot = L.g[e][lane];
            float4 g0, g1, g2, g3;
            load_rec(sc.geo, slot, g0, g1, g2, g3);
            Isect I;
            intersect(g0, g1, g2, g3, ray, I);
            refine_sampling(g0, g1, g2, g3, ray, I);
            grad_hit<SPB>(I, slot, g0, g1, g2, g3, rgbv[slot], tb, p, sig, cr, cg, dL, sc, gr, view);
        }
    };
    auto slab_direct = [&](float tb, float tmin, float tmax) {
        float sig[SPB], cr[SPB], cg[SPB], cb[SPB];
#pragma unroll
        for (int i = 0; i < SPB; ++i) sig[i] = cr[i] = cg[i] = cb[i] = 0.0f;
        int nh = 0;
        traverse(sc.nodes, sc.total_nodes, ray, tmin, tmax, [&](int slot) -> bool {
            float4 g0, g1, g2, g3;
            load_rec(sc.geo, slot, g0, g1, g2, g3);
            Isect I;
            if (!intersect(g0, g1, g2, g3, ray, I)) return true;
            if (!overlaps(tmin, tmax, I.r1, I.r2)) return true;
            refine_sampling(g0, g1, g2, g3, ray, I);
            accum_hit<SPB>(I, g0.w, rgbv[slot], tb, p, sig, cr, cg, cb);
            ++nh;
            return true;
        });
        if (nh == 0) return;
        slab_coeffs<SPB>(sig, cr, cg, cb, p, dL, T, Crem);
        traverse(sc.nodes, sc.total_nodes, ray, tmin, tmax, [&](int slot) -> bool {
            float4 g0, g1, g2, g3;
            load_rec(sc.geo, slot, g0, g1, g2, g3);
            Isect I;
            if (!intersect(g0, g1, g2, g3, ray, I)) return true;
            if (!overlaps(tmin, tmax, I.r1, I.r2)) return true;
            refine_sampling(g0, g1, g2, g3, ray, I);
            grad_hit<SPB>(I, slot, g0, g1, g2, g3, rgbv[slot], tb, p, sig, cr, cg, dL, sc, gr, view);
            return true;
        });
    };
    march<CAP>(sc, p, ray, L, lane, ctr, slab_list, slab_direct, [&]() { return T > p.t_eps; });
}

// ------------------------------------------------------------------ kernels ------------------

// Work item -> (view, pixel, output row); for the implicit single-view mode the items are
// enumerated in 8x8 tiles (num_items is then the padded tile count * 64).
__device__ __forceinline__ bool resolve(const RayJob& job, int64_t w, RayRef& q) {
    if (job.ray_view || job.pixels) {
        q = ray_ref(job, w);
        return q.active;
    }
    const rg_camera& cam = job.cams[job.view];
    int pixel;
    if (!tile_pixel(w, cam.width, cam.height, pixel)) return false;
    q.view = job.view;
    q.pixel = pixel;
    q.out = pixel;
    q.active = true;
    return true;
}

// XCD-aware block order: consecutive work blocks land on the same XCD's L2.
__device__ __forceinline__ int64_t block_item(int64_t b, int64_t nblocks) {
    const int64_t per = (nblocks + 7) >> 3;
    const int64_t xcd = b & 7, k = b >> 3;
    const int64_t logical = xcd * per + k;
    return logical;
}

template <int SPB, int CAP>
__global__ void __launch_bounds__(kWave) k_forward(DevScene sc, DevParams p, RayJob job, int64_t num_items,
                                                   float* __restrict__ out_rgb, float* __restrict__ out_depth,
                                                   float* __restrict__ out_opac, DevStats st) {
    __shared__ IList<CAP> L;
    const int lane = threadIdx.x;
    const int64_t nb = gridDim.x;
    const int64_t item = block_item(blockIdx.x, nb) * kWave + lane;
    Counters ctr;
    RayRef q;
    if (item < num_items && resolve(job, item, q)) {
        const rg_camera cam = job.cams[q.view];
        const Ray ray = make_ray(cam, q.pixel, sc.box, p.t_near, p.t_far);
        RayOut o;
        forward_ray<SPB, CAP>(sc, p, ray, sc.rgb + (size_t)q.view * sc.n, L, lane, ctr, o);
        out_rgb[3 * q.out + 0] = o.C[0] + o.T * p.bg;
        out_rgb[3 * q.out + 1] = o.C[1] + o.T * p.bg;
        out_rgb[3 * q.out + 2] = o.C[2] + o.T * p.bg;
        if (out_depth) out_depth[q.out] = o.depth;
        if (out_opac) out_opac[q.out] = 1.0f - o.T;
    }
    flush_counters(ctr, st);
}

template <int SPB, int CAP>
__global__ void __launch_bounds__(kWave) k_backward(DevScene sc, DevParams p, RayJob job, int64_t num_items,
                                                    const float* __restrict__ ray_colors, const float* __restrict__ dLdC,
                                                    DevGrad gr, DevStats st) {
    __shared__ IList<CAP> L;
    const int lane = threadIdx.x;
    const int64_t nb = gridDim.x;
    const int64_t item = block_item(blockIdx.x, nb) * kWave + lane;
    Counters ctr;
    RayRef q;
    if (item < num_items && resolve(job, item, q)) {
        const float dL[3] = {dLdC[3 * q.out], dLdC[3 * q.out + 1], dLdC[3 * q.out + 2]};
        if (dL[0] != 0.0f || dL[1] != 0.0f || dL[2] != 0.0f) {
            const float Cf[3] = {ray_colors[3 * q.out], ray_colors[3 * q.out + 1], ray_colors[3 * q.out + 2]};
            const rg_camera cam = job.cams[q.view];
            const Ray ray = make_ray(cam, q.pixel, sc.box, p.t_near, p.t_far);
            backward_ray<SPB, CAP>(sc, p, ray, sc.rgb + (size_t)q.view * sc.n, Cf, dL, L, lane, ctr, gr, q.view);
        }
    }
    flush_counters(ctr, st);
}

// Fused train step, persistent: waves pull chunks of `chunk` rays from a device queue; each ray
// is rendered, its per-pixel loss gradient formed, and back-propagated in the same lane.
template <int SPB, int CAP>
__global__ void __launch_bounds__(kWave) k_train(DevScene sc, DevParams p, RayJob job, int64_t num_items, int chunk,
                                                 const float* __restrict__ target, int loss_kind,
                                                 float* __restrict__ out_rgb, DevGrad gr, DevStats st) {
    __shared__ IList<CAP> L;
    const int lane = threadIdx.x;
    Counters ctr;
    float loss = 0.0f;
    // The work items are split into 8 contiguous ranges with one queue each; a wave drains the
    // queue of the XCD it runs on first (that XCD's L2 then serves a compact part of the scene)
    // and then steals from the others.  Every queue is drained, so the XCD id only affects speed.
    const int xcc = xcc_id();
    const int64_t per = (num_items + 7) / 8;
    for (int k = 0; k < 8; ++k) {
      const int qid = (xcc + k) & 7;
      const int64_t q0 = (int64_t)qid * per;
      const int64_t q1 = min(num_items, q0 + per);
      while (q0 < q1) {
        int64_t base = 0;
        if (lane == 0) base = (int64_t)atomicAdd(&st.c[C_QUEUE0 + qid], (unsigned long long)chunk);
        base = __shfl(base, 0, 64);
        if (q0 + base >= q1) break;
        for (int j = 0; j < chunk; j += kWave) {
            const int64_t item = q0 + base + j + lane;
            RayRef q;
            if (item < q1 && resolve(job, item, q)) {
                const rg_camera cam = job.cams[q.view];
                const Ray ray = make_ray(cam, q.pixel, sc.box, p.t_near, p.t_far);
                const float4* rgbv = sc.rgb + (size_t)q.view * sc.n;
                RayOut o;
                forward_ray<SPB, CAP>(sc, p, ray, rgbv, L, lane, ctr, o);
                float C[3] = {o.C[0] + o.T * p.bg, o.C[1] + o.T * p.bg, o.C[2] + o.T * p.bg};
                if (out_rgb) {
                    out_rgb[3 * q.out + 0] = C[0];
                    out_rgb[3 * q.out + 1] = C[1];
                    out_rgb[3 * q.out + 2] = C[2];
                }
                float dL[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float e = C[k] - target[3 * q.out + k];
                    if (loss_kind == 1) {
                        loss += fabsf(e);
                        dL[k] = e > 0.0f ? 1.0f : (e < 0.0f ? -1.0f : 0.0f);
                    } else {
                        loss += e * e;
                        dL[k] = 2.0f * e;
                    }
                }
                if (dL[0] != 0.0f || dL[1] != 0.0f || dL[2] != 0.0f) {
                    Counters dummy;
                    backward_ray<SPB, CAP>(sc, p, ray, rgbv, C, dL, L, lane, dummy, gr, q.view);
                }
            }
        }
      }
    }
    flush_counters(ctr, st);
    const float ls = wave_sum(loss);
    if (lane == 0 && ls != 0.0f) atomicAdd(reinterpret_cast<float*>(&st.c[C_LOSS]), ls);
}

// Verification: per-slab hit sets from the same traversal + predicate (every slab to texit).
__global__ void __launch_bounds__(kWave) k_debug_hits(DevScene sc, DevParams p, RayJob job, int max_slabs,
                                                      int max_pairs, int32_t* pairs, int64_t* counts, float* ro,
                                                      float* rd, float* te, float* tx) {
    const int64_t r = (int64_t)blockIdx.x * kWave + threadIdx.x;
    if (r >= job.num_rays) return;
    const RayRef q = ray_ref(job, r);
    const rg_camera cam = job.cams[q.view];
    const Ray ray = make_ray(cam, q.pixel, sc.box, p.t_near, p.t_far);
    ro[3 * r] = ray.ox;
    ro[3 * r + 1] = ray.oy;
    ro[3 * r + 2] = ray.oz;
    rd[3 * r] = ray.dx;
    rd[3 * r + 1] = ray.dy;
    rd[3 * r + 2] = ray.dz;
    te[r] = ray.tenter;
    tx[r] = ray.texit;
    int64_t n = 0;
    if (ray.tenter < ray.texit) {
        const float slab = p.dt * (float)p.spb;
        float tb = ray.tenter;
        for (int s = 0; s < max_slabs && tb < ray.texit; ++s) {
            const float tmin = fmaxf(ray.tenter, tb);
            const float tmax = fminf(ray.texit, add_exact(tb, slab));
            if (tmax > ray.tenter) {
                traverse(sc.nodes, sc.total_nodes, ray, tmin, tmax, [&](int slot) -> bool {
                    float4 g0, g1, g2, g3;
                    load_rec(sc.geo, slot, g0, g1, g2, g3);
                    Isect I;
                    if (intersect(g0, g1, g2, g3, ray, I) && overlaps(tmin, tmax, I.r1, I.r2)) {
                        if (n < max_pairs) {
                            pairs[(r * max_pairs + n) * 2] = s;
                            pairs[(r * max_pairs + n) * 2 + 1] = sc.perm[slot];
                        }
                        ++n;
                    }
                    return true;
                });
            }
            tb = add_exact(tb, slab);
        }
    }
    counts[r] = n;
}

__global__ void k_reset_counters(unsigned long long* c) {
    if (threadIdx.x < 16) c[threadIdx.x] = 0;
}

// ------------------------------------------------------------------ dispatch -----------------

}  // namespace

// In the implicit single-view mode RayJob.num_rays carries the padded 8x8-tile item count.
#define RG_DISPATCH(SPBV, CAPV, KERNEL, ...)                                                                         \
    do {                                                                                                             \
        if (SPBV == 8) {                                                                                             \
            if (CAPV <= 8) hipLaunchKernelGGL((KERNEL<8, 8>), __VA_ARGS__);                                          \
            else if (CAPV <= 16) hipLaunchKernelGGL((KERNEL<8, 16>), __VA_ARGS__);                                   \
            else hipLaunchKernelGGL((KERNEL<8, 32>), __VA_ARGS__);                                                   \
        } else {                                                                                                     \
            if (CAPV <= 8) hipLaunchKernelGGL((KERNEL<16, 8>), __VA_ARGS__);                                         \
            else if (CAPV <= 16) hipLaunchKernelGGL((KERNEL<16, 16>), __VA_ARGS__);                                  \
            else hipLaunchKernelGGL((KERNEL<16, 32>), __VA_ARGS__);                                                  \
        }                                                                                                            \
    } while (0)

hipError_t launch_reset_counters(DevStats st, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, st.c);
    return hipGetLastError();
}

hipError_t launch_forward(const DevScene& sc, const DevParams& p, const RayJob& job, int hit_depth, float* out_rgb,
                          float* out_depth, float* out_opacity, DevStats st, hipStream_t s) {
    const int64_t items = job.num_rays;  // already the padded item count in implicit mode
    if (items <= 0) return hipSuccess;
    const int64_t nb = (items + kWave - 1) / kWave;
    const int64_t nbp = ((nb + 7) / 8) * 8;
    RG_DISPATCH(p.spb <= 8 ? 8 : 16, hit_depth, k_forward, dim3((unsigned)nbp), dim3(kWave), 0, s, sc, p, job, items,
                out_rgb, out_depth, out_opacity, st);
    return hipGetLastError();
}

hipError_t launch_backward(const DevScene& sc, const DevParams& p, const RayJob& job, int hit_depth,
                           const float* ray_colors, const float* dLdC, DevGrad gr, DevStats st, hipStream_t s) {
    const int64_t items = job.num_rays;
    if (items <= 0) return hipSuccess;
    const int64_t nb = (items + kWave - 1) / kWave;
    const int64_t nbp = ((nb + 7) / 8) * 8;
    RG_DISPATCH(p.spb <= 8 ? 8 : 16, hit_depth, k_backward, dim3((unsigned)nbp), dim3(kWave), 0, s, sc, p, job,
                items, ray_colors, dLdC, gr, st);
    return hipGetLastError();
}

hipError_t launch_train(const DevScene& sc, const DevParams& p, const RayJob& job, int hit_depth, int chunk,
                        const float* target, int loss_kind, float* out_rgb, float* out_loss, DevGrad gr, DevStats st,
                        hipStream_t s) {
    (void)out_loss;
    const int64_t items = job.num_rays;
    if (items <= 0) return hipSuccess;
    int dev = 0;
    hipGetDevice(&dev);
    int cus = 0;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (cus <= 0) cus = 256;
    const int64_t want = (items + chunk - 1) / chunk;
    int64_t nb = (int64_t)cus * 16;
    if (nb > want) nb = want;
    RG_DISPATCH(p.spb <= 8 ? 8 : 16, hit_depth, k_train, dim3((unsigned)nb), dim3(kWave), 0, s, sc, p, job, items,
                chunk, target, loss_kind, out_rgb, gr, st);
    return hipGetLastError();
}

hipError_t launch_debug_hits(const DevScene& sc, const DevParams& p, const RayJob& job, int max_slabs, int max_pairs,
                             int32_t* pairs, int64_t* counts, float* ro, float* rd, float* te, float* tx,
                             hipStream_t s) {
    if (job.num_rays <= 0) return hipSuccess;
    const int64_t nb = (job.num_rays + kWave - 1) / kWave;
    hipLaunchKernelGGL(k_debug_hits, dim3((unsigned)nb), dim3(kWave), 0, s, sc, p, job, max_slabs, max_pairs, pairs,
                       counts, ro, rd, te, tx);
    return hipGetLastError();
}

}  // namespace rg
