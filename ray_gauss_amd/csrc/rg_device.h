// Device-side building blocks of the ray caster: FP32 ray setup, the bounding-ellipsoid
// predicate, box tests and the stackless (threaded) LBVH traversal.
//
// Functions marked "exact FP32" are evaluated with contraction off so that every operation
// rounds once; they define the slab grid and the hit sets, which must be reproducible bit for
// bit by an FP32 restatement on the host.
#pragma once

#include <hip/hip_runtime.h>

#include "rg_internal.h"

namespace rg {

struct Ray {
    float ox, oy, oz;
    float dx, dy, dz;
    float tenter, texit;
};

// The reference's random generator (forward/gaussians_aabb.h:63-94): tea<4> (four cycles of the
// TEA cipher with a fixed key) seeds a ray from (iteration, pixel); a 32-bit LCG steps it and a
// draw is its low 24 bits / 2^24.
__device__ __forceinline__ unsigned int tea4(unsigned int v0, unsigned int v1) {
    unsigned int s0 = 0;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

__device__ __forceinline__ float rnd01(unsigned int& st) {
    st = 1664525u * st + 1013904223u;
    return (float)(st & 0x00FFFFFFu) / 16777216.0f;
}

// Generator state after n draws.
__device__ __forceinline__ unsigned int skip_draws(unsigned int st, int n) {
    for (int i = 0; i < n; ++i) st = 1664525u * st + 1013904223u;
    return st;
}

// computeRay (reference cuda_train/forward/gaussians_aabb.cu:405-434) followed by the scene-box
// slab test (:530-541) clipped to [t_near, t_far].  Exact FP32.  With jitter the pixel position
// is x + 1/3 + u/3 (two draws from the ray's generator, :409-418), else x + 1/2.
__device__ __forceinline__ Ray make_ray(const rg_camera& c, int pixel, const float* box, float t_near, float t_far,
                                        int jitter = 0, unsigned int* seed = nullptr) {
#pragma clang fp contract(off)
    Ray r;
    const int ix = pixel % c.width;
    const int iy = pixel / c.width;
    float fx, fy;
    if (jitter) {
        const float jx = rnd01(*seed) * 0.333333f;
        const float jy = rnd01(*seed) * 0.333333f;
        fx = (((float)ix + 0.333333f) + jx) / (float)c.width;
        fy = (((float)iy + 0.333333f) + jy) / (float)c.height;
    } else {
        fx = ((float)ix + 0.5f) / (float)c.width;
        fy = ((float)iy + 0.5f) / (float)c.height;
    }
    const float px = 2.0f * fx - 1.0f;
    const float py = 2.0f * fy - 1.0f;
    const float ddx = px * c.tanx;
    const float ddy = py * c.tany;
    float v0 = (ddx * c.u[0] + ddy * c.v[0]) + c.w[0];
    float v1 = (ddx * c.u[1] + ddy * c.v[1]) + c.w[1];
    float v2 = (ddx * c.u[2] + ddy * c.v[2]) + c.w[2];
    const float len2 = (v0 * v0 + v1 * v1) + v2 * v2;
    const float inv = 1.0f / __builtin_sqrtf(len2);
    r.ox = c.eye[0];
    r.oy = c.eye[1];
    r.oz = c.eye[2];
    r.dx = v0 * inv;
    r.dy = v1 * inv;
    r.dz = v2 * inv;
    const float t0x = (box[0] - r.ox) / r.dx, t1x = (box[3] - r.ox) / r.dx;
    const float t0y = (box[1] - r.oy) / r.dy, t1y = (box[4] - r.oy) / r.dy;
    const float t0z = (box[2] - r.oz) / r.dz, t1z = (box[5] - r.oz) / r.dz;
    const float mnx = fminf(t0x, t1x), mxx = fmaxf(t0x, t1x);
    const float mny = fminf(t0y, t1y), mxy = fmaxf(t0y, t1y);
    const float mnz = fminf(t0z, t1z), mxz = fmaxf(t0z, t1z);
    r.tenter = fmaxf(t_near, fmaxf(mnx, fmaxf(mny, mnz)));
    r.texit = fminf(t_far, fminf(mxx, fminf(mxy, mxz)));
    return r;
}

// Exact FP32: next slab start (reference accumulates tbuffer += slab_spacing, :596).
__device__ __forceinline__ float add_exact(float a, float b) {
#pragma clang fp contract(off)
    return a + b;
}

// Exact FP32: t of sample i in the slab starting at tb (reference :474), with the rnd_sample
// offsets of :473-482 -- mode 1: + u dt, u the next draw of st (each hit of a slab restarts from
// the slab's state); mode 2: + dt/2.
__device__ __forceinline__ float sample_t(float tb, int i, float dt, int mode, unsigned int& st) {
#pragma clang fp contract(off)
    const float off = (float)i * dt;
    const float base = tb + off;
    if (mode == 1) return base + rnd01(st) * dt;
    if (mode == 2) return base + 0.5f * dt;
    return base;
}

// Per (ray, Gaussian) result of the bounding-ellipsoid test.
struct Isect {
    float r1, r2;  // predicate interval (root1, root2)
    float tc;      // t of closest approach in ellipsoid metric (b * l)
    float A;       // h * |P|^2          -> weight exponent p(t) = A + Cq (t - tc)^2
    float Cq;      // h * |D'|^2
    float Px, Py, Pz;     // O' + b D^   (ellipsoid-space offset at closest approach)
    float Dx, Dy, Dz;     // D' = M' d
    float l;              // 1 / |D'|
};

// Sampling terms of the bounding-ellipsoid test (everything but the roots), exact FP32 with the
// operation order of the host FP32 checker.  Returns disc = 1 - |P|^2.
__device__ __forceinline__ float isect_terms(const float4 g0, const float4 g1, const float4 g2, const float4 g3,
                                             const Ray& r, Isect& I, float& b, float& c) {
#pragma clang fp contract(off)
    const float Ox = r.ox - g0.x, Oy = r.oy - g0.y, Oz = r.oz - g0.z;
    const float Op0 = (g1.x * Ox + g1.y * Oy) + g1.z * Oz;
    const float Op1 = (g2.x * Ox + g2.y * Oy) + g2.z * Oz;
    const float Op2 = (g3.x * Ox + g3.y * Oy) + g3.z * Oz;
    const float Dp0 = (g1.x * r.dx + g1.y * r.dy) + g1.z * r.dz;
    const float Dp1 = (g2.x * r.dx + g2.y * r.dy) + g2.z * r.dz;
    const float Dp2 = (g3.x * r.dx + g3.y * r.dy) + g3.z * r.dz;
    const float ddp = (Dp0 * Dp0 + Dp1 * Dp1) + Dp2 * Dp2;
    const float l = 1.0f / __builtin_sqrtf(ddp);
    const float Dh0 = Dp0 * l, Dh1 = Dp1 * l, Dh2 = Dp2 * l;
    b = -((Op0 * Dh0 + Op1 * Dh1) + Op2 * Dh2);
    c = ((Op0 * Op0 + Op1 * Op1) + Op2 * Op2) - 1.0f;
    const float P0 = Op0 + b * Dh0, P1 = Op1 + b * Dh1, P2 = Op2 + b * Dh2;
    const float pp = (P0 * P0 + P1 * P1) + P2 * P2;
    I.tc = b * l;
    I.A = g1.w * pp;
    I.Cq = g1.w * ddp;
    I.Px = P0;
    I.Py = P1;
    I.Pz = P2;
    I.Dx = Dp0;
    I.Dy = Dp1;
    I.Dz = Dp2;
    I.l = l;
    return 1.0f - pp;
}

// __intersection__gaussian (reference cuda_train/forward/gaussians_aabb.cu:324-402), exact FP32
// for the decision part.  Returns true when disc > 1e-7; the hit test on [tmin, tmax] is then
// fmaxf(tmin, r1) <= fminf(tmax, r2).  The operation order is shared with the host FP32 checker.
__device__ __forceinline__ bool intersect(const float4 g0, const float4 g1, const float4 g2, const float4 g3,
                                          const Ray& r, Isect& I) {
#pragma clang fp contract(off)
    // The terms are computed before the validity test (h <= 0 or NaN: no bounding ellipsoid,
    // |sigma| <= thr; the terms of such a record are then unused), so that the record's four
    // loads are issued together instead of waiting for h before loading the rest.
    float b, c;
    const float disc = isect_terms(g0, g1, g2, g3, r, I, b, c);
    if (!(g1.w > 0.0f)) return false;
    if (!(disc > 1e-7f)) return false;
    const float sd = __builtin_sqrtf(disc);
    const float q = (b > 0.0f) ? b + sd : b - sd;
    I.r1 = (c / q) * I.l;
    I.r2 = q * I.l;
    return true;
}

// Sampling set-up re-centred at the closest approach.  b = -O'.D^ is a difference of
// ellipsoid-space quantities of size |o - mu| / s', so t_c = b l and P = O' + b D^ carry an
// absolute error that grows with the distance to the camera.  Evaluating O' again at
// x0 = o + t_c d (a point inside the ellipsoid) leaves only the rounding of x0 itself, like the
// reference's per-sample x = o + t d.  Does not touch the hit decision (r1, r2).
__device__ __forceinline__ void refine_sampling(const float4 g0, const float4 g1, const float4 g2, const float4 g3,
                                                const Ray& r, Isect& I) {
#pragma clang fp contract(off)
    // contraction off: the result must not depend on the inlining context, because the gather
    // caches it and the backward recomputes it
    const float t0 = I.tc;
    const float X0 = (r.ox + t0 * r.dx) - g0.x;
    const float X1 = (r.oy + t0 * r.dy) - g0.y;
    const float X2 = (r.oz + t0 * r.dz) - g0.z;
    const float O0 = g1.x * X0 + g1.y * X1 + g1.z * X2;
    const float O1 = g2.x * X0 + g2.y * X1 + g2.z * X2;
    const float O2 = g3.x * X0 + g3.y * X1 + g3.z * X2;
    const float Dh0 = I.Dx * I.l, Dh1 = I.Dy * I.l, Dh2 = I.Dz * I.l;
    const float b = -(O0 * Dh0 + O1 * Dh1 + O2 * Dh2);
    const float P0 = O0 + b * Dh0, P1 = O1 + b * Dh1, P2 = O2 + b * Dh2;
    I.tc = t0 + b * I.l;
    I.A = g1.w * (P0 * P0 + P1 * P1 + P2 * P2);
    I.Px = P0;
    I.Py = P1;
    I.Pz = P2;
}

__device__ __forceinline__ bool overlaps(float tmin, float tmax, float r1, float r2) {
    return fmaxf(tmin, r1) <= fminf(tmax, r2);
}

// Conservative ray-segment / box test (boxes are padded at build time).
__device__ __forceinline__ bool box_hit(const float4 lo, const float4 hi, const Ray& r, float ix, float iy, float iz,
                                        float ta, float tb) {
    // one FMA per plane (-o*inv is loop-invariant); the rounding differs from (lo - o) * inv by a
    // few ulp of t, far inside the box padding.  A zero direction component gives NaN planes,
    // which fminf/fmaxf ignore: the test then only widens.
    const float ax = fmaf(lo.x, ix, -r.ox * ix), bx = fmaf(hi.x, ix, -r.ox * ix);
    const float ay = fmaf(lo.y, iy, -r.oy * iy), by = fmaf(hi.y, iy, -r.oy * iy);
    const float az = fmaf(lo.z, iz, -r.oz * iz), bz = fmaf(hi.z, iz, -r.oz * iz);
    const float tn = fmaxf(fmaxf(ta, fminf(ax, bx)), fmaxf(fminf(ay, by), fminf(az, bz)));
    const float tf = fminf(fminf(tb, fmaxf(ax, bx)), fminf(fmaxf(ay, by), fmaxf(az, bz)));
    return tn <= tf;
}

// Tree nodes are read through the constant address space: the traversal kernels never write
// them, and this lets a wave-uniform node index become a scalar load (through the scalar cache)
// even in kernels that store or add to global memory, where the compiler cannot otherwise prove
// the node array unchanged and falls back to per-lane vector loads.
typedef float node_f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(4))) const node_f32x4 const_node_f32x4;

__device__ __forceinline__ float4 ld_node(const float4* __restrict__ nodes, int i) {
    const node_f32x4 v = ((const const_node_f32x4*)nodes)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}

// Stackless traversal of the threaded LBVH: visits, in DFS preorder, every leaf whose padded box
// meets the ray segment [ta, tb].  on_leaf(slot) returns false to stop the traversal.
template <typename F>
__device__ __forceinline__ void traverse(const float4* __restrict__ nodes, int total, const Ray& r, float ta, float tb,
                                         F&& on_leaf) {
    const float ix = 1.0f / r.dx, iy = 1.0f / r.dy, iz = 1.0f / r.dz;
    int idx = 0;
    while (idx < total) {
        const float4 lo = ld_node(nodes, 2 * idx);
        const float4 hi = ld_node(nodes, 2 * idx + 1);
        const int miss = __float_as_int(lo.w);
        const int leaf = __float_as_int(hi.w);
        if (miss <= idx) return;  // malformed link: never loop (idx strictly increases)
        if (box_hit(lo, hi, r, ix, iy, iz, ta, tb)) {
            if (leaf >= 0) {
                if (leaf >= ((total + 1) >> 1)) return;  // malformed leaf index: stop
                if (!on_leaf(leaf)) return;
                idx = miss;
            } else {
                idx = idx + 1;
            }
        } else {
            idx = miss;
        }
    }
}

// Packet form of traverse(): the lanes executing the call walk ONE node sequence together (the
// node index is wave-uniform, so node loads are uniform), each lane testing the boxes against its
// own ray segment; the wave descends into a node when any lane's segment meets its box, and at a
// leaf every lane whose segment meets the leaf box calls on_leaf.  Each lane therefore sees
// exactly the leaves the per-lane traversal would give it, in the same DFS order.  For the rays
// of an 8x8 pixel tile the union of visited nodes is close to one ray's set.
template <typename F>
__device__ __forceinline__ void traverse_packet(const float4* __restrict__ nodes, int total, bool active, const Ray& r,
                                                float ta, float tb, F&& on_leaf) {
    const float ix = 1.0f / r.dx, iy = 1.0f / r.dy, iz = 1.0f / r.dz;
    bool live = active;
    // the live lanes as a scalar mask, so that the ballot of a box test (a compare result) needs
    // one scalar AND; a ballot of the loop-carried lane flag itself costs two vector instructions
    unsigned long long livem = __builtin_amdgcn_ballot_w64(live);
    int idx = 0;
    while (idx < total && livem != 0ULL) {
        idx = __builtin_amdgcn_readfirstlane(idx);
        const float4 lo = ld_node(nodes, 2 * idx);
        const float4 hi = ld_node(nodes, 2 * idx + 1);
        const int miss = __float_as_int(lo.w);
        const int leaf = __float_as_int(hi.w);
        if (miss <= idx) return;  // malformed link: never loop (idx strictly increases)
        const bool b = box_hit(lo, hi, r, ix, iy, iz, ta, tb);
        if ((__builtin_amdgcn_ballot_w64(b) & livem) != 0ULL) {
            if (leaf >= 0) {
                if (leaf >= ((total + 1) >> 1)) return;  // malformed leaf index: stop
                if (live && b && !on_leaf(leaf)) live = false;
                livem = __builtin_amdgcn_ballot_w64(live);
                idx = miss;
            } else {
                idx = idx + 1;
            }
        } else {
            idx = miss;
        }
    }
}

// ---------------------------------------------------------------- cross-lane operations -------
// Wavefront reductions without LDS.  Within a row of 16 lanes, DPP moves (quad permutes, the
// half-row mirror, the row rotation by 8) feed the ALU directly; across rows, gfx950's
// v_permlane16_swap / v_permlane32_swap exchange row pairs and wave halves in the VALU.  (A
// __shfl_xor compiles to ds_bpermute_b32, one LDS round trip per exchange.)  Every lane of the
// wavefront must execute these calls.

constexpr int kDppXor1 = 0xB1;         // quad_perm [1,0,3,2]: lane l reads lane l ^ 1
constexpr int kDppXor2 = 0x4E;         // quad_perm [2,3,0,1]: lane l reads lane l ^ 2
constexpr int kDppHalfMirror = 0x141;  // row_half_mirror: lane l reads lane l ^ 7 (groups of 8)
constexpr int kDppRor8 = 0x128;        // row_ror:8: lane l reads lane l ^ 8 (rows of 16)

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}

// v_permlane16_swap(a, b): the two results hold, in rows 2r, a of row 2r and a of row 2r+1, and
// in rows 2r+1, b of row 2r and b of row 2r+1 (same column).
struct Pair {
    unsigned int x, y;
};
__device__ __forceinline__ Pair swap16(unsigned int a, unsigned int b) {
    const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
    return Pair{r[0], r[1]};
}
// v_permlane32_swap(a, b): in lanes l < 32, a[l] and a[l + 32]; in lanes l >= 32, b[l - 32], b[l].
__device__ __forceinline__ Pair swap32(unsigned int a, unsigned int b) {
    const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
    return Pair{r[0], r[1]};
}

// Reduce-scatter steps: rows 2r get a[l] + a[l ^ 16], rows 2r+1 get b[l ^ 16] + b[l]; and
// lanes below 32 get a[l] + a[l ^ 32], the others b[l ^ 32] + b[l].
__device__ __forceinline__ float pair_add16(float a, float b) {
    const Pair r = swap16(__float_as_uint(a), __float_as_uint(b));
    return __uint_as_float(r.x) + __uint_as_float(r.y);
}
__device__ __forceinline__ float pair_add32(float a, float b) {
    const Pair r = swap32(__float_as_uint(a), __float_as_uint(b));
    return __uint_as_float(r.x) + __uint_as_float(r.y);
}

__device__ __forceinline__ float wave_sum64(float v) {
    v += dpp_f<kDppXor1>(v);
    v += dpp_f<kDppXor2>(v);
    v += dpp_f<kDppHalfMirror>(v);
    v += dpp_f<kDppRor8>(v);
    v = pair_add16(v, v);
    return pair_add32(v, v);
}

__device__ __forceinline__ int wave_min64(int v) {
    v = min(v, dpp_i<kDppXor1>(v));
    v = min(v, dpp_i<kDppXor2>(v));
    v = min(v, dpp_i<kDppHalfMirror>(v));
    v = min(v, dpp_i<kDppRor8>(v));
    Pair r = swap16((unsigned int)v, (unsigned int)v);
    v = min((int)r.x, (int)r.y);
    r = swap32((unsigned int)v, (unsigned int)v);
    return min((int)r.x, (int)r.y);
}

__device__ __forceinline__ float wave_min64f(float v) {
    v = fminf(v, dpp_f<kDppXor1>(v));
    v = fminf(v, dpp_f<kDppXor2>(v));
    v = fminf(v, dpp_f<kDppHalfMirror>(v));
    v = fminf(v, dpp_f<kDppRor8>(v));
    Pair r = swap16(__float_as_uint(v), __float_as_uint(v));
    v = fminf(__uint_as_float(r.x), __uint_as_float(r.y));
    r = swap32(__float_as_uint(v), __float_as_uint(v));
    return fminf(__uint_as_float(r.x), __uint_as_float(r.y));
}

}  // namespace rg
