// TEST INFRASTRUCTURE ONLY -- never linked into, loaded by, or called from the product path.
//
// CPU FP64 restatement of the RayGauss training ray caster (hugobl1/ray_gauss), used as the
// correctness oracle for the HIP implementation in ray_gauss_amd/.  Only tests/, the smoke()
// check and bench.py's cpu_baseline leg may load this library.
//
// What is FP32 and what is FP64:
//   * The ray origin/direction, the scene-box slab test (tenter/texit) and the slab/sample grid
//     (tb accumulated by repeated FP32 addition, t_i = tb + i*dt) are computed in FP32 exactly as
//     the reference's raygen does (forward/gaussians_aabb.cu:405-434, :530-541, :547-548, :596,
//     :474).  These are discrete positions chosen by the reference, not approximations of a real
//     number, so the oracle reproduces them bit for bit; the HIP path computes them with the
//     same FP32 operation sequence (contraction off).
//   * Everything evaluated at those positions -- bounding ellipsoids, the hit predicate, Gaussian
//     weights, compositing, gradients, SH/SG colour -- is FP64.
//
// Reference functions restated here (paths relative to the reference repository):
//   quaternion_to_matrix        optix_raycasting/cuda_train/forward/gaussians_aabb.cu:296-310
//   __intersection__gaussian    optix_raycasting/cuda_train/forward/gaussians_aabb.cu:324-402
//   computeRay                  optix_raycasting/cuda_train/forward/gaussians_aabb.cu:405-434
//   computeBufferForward        optix_raycasting/cuda_train/forward/gaussians_aabb.cu:437-498
//   colorBlending               optix_raycasting/cuda_train/forward/gaussians_aabb.cu:500-519
//   __raygen__rg (forward)      optix_raycasting/cuda_train/forward/gaussians_aabb.cu:521-612
//   __raygen__rg (backward)     optix_raycasting/cuda_train/backward/gaussians_aabb.cu:505-775
//   forward_sh / backward_sh    utils/sh_utils.cu:25-119 / :121-315
//   compute_ellipsoids_bbox     optix_raycasting/optix_utils.py:54-66
//   depth (viewer only)         optix_raycasting/cuda_gui/gaussians_aabb.cu:397-417
//   tea<4>, lcg, rnd            optix_raycasting/cuda_train/forward/gaussians_aabb.h:63-94
//
// Random numbers (jitter, rnd_sample): a ray's generator is seeded with tea<4>(iteration,
// pixel) (:526); jitter draws two numbers in computeRay (:413-414); with rnd_sample = 1 every
// slab that has hits draws one number per sample, shared by all its hits, and the generator
// then stands after those draws (:440-496; backward :573-611, :768); slabs without hits draw
// nothing.
//
// Runtime parameters replace the reference's compile-time constants (gaussians_aabb.h:2-5):
// dt (DT), samples per slab (BUFFER_SIZE), t_eps (TRANSMITTANCE_EPSILON), sigma_thr
// (SIGMA_THRESHOLD).  near/far clip the scene-box interval; near=0, far=+inf is the reference.
//
// Hit sets: the reference lets OptiX find, per slab, every Gaussian whose bounding ellipsoid
// satisfies the intersection predicate on [tmin_slab, tmax_slab].  The predicate's interval
// [root1, root2] does not depend on the slab, so this oracle computes it once per
// (ray, Gaussian) and derives each slab's hit set from the same inequality
// max(tmin, root1) <= min(tmax, root2).  orc_render_literal() evaluates the predicate for every
// Gaussian in every slab, literally as the reference does, and pins the interval form.
// The reference drops hits beyond max_prim_slice (1024) in traversal order; that order is
// OptiX-internal and cannot be reproduced, so no hits are dropped here (the per-slab maximum is
// reported so tests can assert the 1024 limit is never reached).

//
// FP32 CPU baseline build (ORC_BASELINE, liboracle_rg32.so): the same source with `real` = float
// and single-precision constants, minus the fragility bookkeeping only the oracle needs (near
// misses, rounding margins, near-threshold counts).  It exports orc_baseline_step() alone and is
// a speed reference for the same algorithm, not a checker.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;

#ifdef ORC_BASELINE
constexpr bool kOracle = false;
#else
constexpr bool kOracle = true;
#endif

extern "C" {

struct orc_camera {
    float eye[3];
    float u[3];
    float v[3];
    float w[3];
    float tanx;   // tan(FoVx/2), FP32 as in optix_utils.py:181
    float tany;
    int32_t width;
    int32_t height;
};

struct orc_params {
    float dt;           // sample spacing (reference DT)
    int32_t spb;        // samples per slab (reference BUFFER_SIZE = 8)
    float t_eps;        // transmittance cut-off (reference TRANSMITTANCE_EPSILON)
    float sigma_thr;    // density threshold (reference SIGMA_THRESHOLD = 0.1)
    float t_near;
    float t_far;
    int32_t white_bg;
    int32_t exact_sh_grad;  // 0: reference backward_sh (utils/sh_utils.cu:254) ; 1: exact derivative
    int32_t jitter;         // 1: sub-pixel jitter of the ray (forward/gaussians_aabb.cu:409-418)
    int32_t rnd_sample;     // 0: t_i = tb + i dt; 1: + u_i dt, u_i random; 2: + dt/2 (:473-482)
    uint32_t iteration;     // seeds the per-ray generator with tea<4>(iteration, pixel) (:526)
    float eps_w;            // fragility margin on sigma*w / thr (<= 0: the default kEpsW)
};

struct orc_scene {
    int64_t n;
    int32_t deg;     // SH degree (0..3)
    int32_t L;       // SG lobes
    const float* pos;     // (n,3)
    const float* scale;   // (n,3) activated
    const float* quat;    // (n,4) w,x,y,z as given (not normalised here, like the reference)
    const float* dens;    // (n,)
    const float* sh;      // (n,3,K)
    const float* sg;      // (n,3,L)
    const float* lam;     // (n,L)
    const float* axis;    // (n,L,3)
};

// The HIP path's FP32 per-Gaussian records (rg_export_predicate), original order.  Given to
// orc_render / orc_backward, they switch the oracle to decision-matched mode: every discrete
// decision (slab hit sets, sample thresholds sigma*w > thr) is taken in FP32 from these records
// with the HIP path's operation sequence, while every value (weights, compositing, gradients)
// stays FP64.  The records' FP32 predicate is the one the hit-set test checks bit for bit.
struct orc_record {
    const float* mu;       // (n,3)
    const float* Mp;       // (n,9) rows of the bounding-ellipsoid transform
    const int32_t* valid;  // (n)
    const float* h;        // (n) ln|sigma/thr|
};

}  // extern "C"

namespace {

// ---------------------------------------------------------------- FP32 ray / grid ----------
// All FP32 arithmetic below is compiled with -ffp-contract=off so that each operation rounds
// once, exactly as the HIP path does it.

struct Ray32 {
    float o[3];
    float d[3];
};

// The reference's generator (forward/gaussians_aabb.h:63-94): TEA with 2N rounds of the
// Wheeler-Needham cipher and the fixed key (0xa341316c, 0xc8013ea4, 0xad90777d, 0x7e95761e)
// seeds; a 32-bit LCG (1664525, 1013904223) steps; rnd = low 24 bits / 2^24.
uint32_t tea_keyed(uint32_t v0, uint32_t v1, int cycles, const uint32_t k[4]) {
    uint32_t s0 = 0;
    for (int n = 0; n < cycles; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + k[0]) ^ (v1 + s0) ^ ((v1 >> 5) + k[1]);
        v1 += ((v0 << 4) + k[2]) ^ (v0 + s0) ^ ((v0 >> 5) + k[3]);
    }
    return v0;
}
uint32_t tea4(uint32_t v0, uint32_t v1) {
    static const uint32_t k[4] = {0xa341316cu, 0xc8013ea4u, 0xad90777du, 0x7e95761eu};
    return tea_keyed(v0, v1, 4, k);
}
inline uint32_t lcg(uint32_t& s) {
    s = 1664525u * s + 1013904223u;
    return s & 0x00FFFFFFu;
}
inline float rnd(uint32_t& s) { return (float)lcg(s) / (float)0x01000000; }

// computeRay (forward/gaussians_aabb.cu:405-434); with jitter the two draws advance seed.
Ray32 make_ray32(const orc_camera& c, int64_t pixel, int jitter = 0, uint32_t* seed = nullptr) {
    Ray32 r;
    const int64_t ix = pixel % c.width;
    const int64_t iy = pixel / c.width;
    float fx, fy;
    if (jitter) {
        const float jx = rnd(*seed) * 0.333333f;
        const float jy = rnd(*seed) * 0.333333f;
        fx = (((float)ix + 0.333333f) + jx) / (float)c.width;
        fy = (((float)iy + 0.333333f) + jy) / (float)c.height;
    } else {
        fx = ((float)ix + 0.5f) / (float)c.width;
        fy = ((float)iy + 0.5f) / (float)c.height;
    }
    const float px = 2.0f * fx - 1.0f;
    const float py = 2.0f * fy - 1.0f;
    const float dx = px * c.tanx;
    const float dy = py * c.tany;
    float v[3];
    for (int k = 0; k < 3; ++k) {
        const float a = dx * c.u[k];
        const float b = dy * c.v[k];
        v[k] = (a + b) + c.w[k];
    }
    const float len2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    const float inv = 1.0f / std::sqrt(len2);
    for (int k = 0; k < 3; ++k) {
        r.o[k] = c.eye[k];
        r.d[k] = v[k] * inv;
    }
    return r;
}

// Scene-box slab test (forward/gaussians_aabb.cu:530-541) with near/far clipping.
void scene_interval32(const Ray32& r, const float* bmin, const float* bmax, float t_near, float t_far,
                      float* tenter, float* texit) {
    float tmn[3], tmx[3];
    for (int k = 0; k < 3; ++k) {
        const float t0 = (bmin[k] - r.o[k]) / r.d[k];
        const float t1 = (bmax[k] - r.o[k]) / r.d[k];
        tmn[k] = std::fmin(t0, t1);
        tmx[k] = std::fmax(t0, t1);
    }
    *tenter = std::fmax(t_near, std::fmax(tmn[0], std::fmax(tmn[1], tmn[2])));
    *texit = std::fmin(t_far, std::fmin(tmx[0], std::fmin(tmx[1], tmx[2])));
}

inline float sample_t32(float tb, int i, float dt) {
    const float off = (float)i * dt;
    return tb + off;
}

// Sample positions of one slab (FP32, :473-482): t_i = (tb + i dt) [+ u_i dt | + dt/2]; with
// rnd_sample = 1, u_i is the (i+1)-th draw from the slab's generator state `seed`.
struct SlabSamples {
    float t[16];
    SlabSamples(const orc_params& p, float tb, uint32_t seed) {
        for (int i = 0; i < p.spb; ++i) {
            const float base = sample_t32(tb, i, p.dt);
            if (p.rnd_sample == 1) t[i] = base + rnd(seed) * p.dt;
            else if (p.rnd_sample == 2) t[i] = base + 0.5f * p.dt;
            else t[i] = base;
        }
    }
};

// The generator after a slab with hits (spb draws when rnd_sample = 1).
inline uint32_t after_slab(const orc_params& p, uint32_t seed) {
    if (p.rnd_sample == 1)
        for (int i = 0; i < p.spb; ++i) lcg(seed);
    return seed;
}

// ---------------------------------------------------------------- FP64 geometry -------------

struct Geo {
    real mu[3];
    real R[3][3];     // R[row][col]; columns are the rotated axes (quaternion_to_matrix)
    real s[3];
    real sigma;
    real M[3][3];     // rows: col_k(R)/s_k       (unenlarged, sampling)
    real Mp[3][3];    // rows: col_k(R)/(s_k*delta) (bounding ellipsoid, predicate)
    bool valid;         // bounding ellipsoid exists: ln((sigma/thr)^2) > 0
};

// quaternion_to_matrix: q = (r,i,j,k) = (w,x,y,z), not normalised.
void quat_to_R(const float* q, real R[3][3]) {
    const real r = q[0], i = q[1], j = q[2], k = q[3];
    // column 0
    R[0][0] = 1.0 - 2.0 * (j * j + k * k);
    R[1][0] = 2.0 * (i * j + r * k);
    R[2][0] = 2.0 * (i * k - r * j);
    // column 1
    R[0][1] = 2.0 * (i * j - r * k);
    R[1][1] = 1.0 - 2.0 * (i * i + k * k);
    R[2][1] = 2.0 * (j * k + r * i);
    // column 2
    R[0][2] = 2.0 * (i * k + r * j);
    R[1][2] = 2.0 * (j * k - r * i);
    R[2][2] = 1.0 - 2.0 * (i * i + j * j);
}

Geo make_geo(const orc_scene& S, int64_t g, real thr) {
    Geo G;
    for (int k = 0; k < 3; ++k) {
        G.mu[k] = S.pos[3 * g + k];
        G.s[k] = S.scale[3 * g + k];
    }
    G.sigma = S.dens[g];
    quat_to_R(S.quat + 4 * g, G.R);
    const real ratio = G.sigma / thr;
    const real l2 = std::log(ratio * ratio);
    G.valid = l2 > 0.0 && std::isfinite(l2);
    const real delta = G.valid ? std::sqrt(l2) : 0.0;
    for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c) {
            G.M[k][c] = G.R[c][k] / G.s[k];
            G.Mp[k][c] = G.valid ? G.R[c][k] / (G.s[k] * delta) : 0.0;
        }
    return G;
}

struct Interval {
    real r1, r2;
    real disc, b, c;
    bool valid;
};

// __intersection__gaussian (forward/gaussians_aabb.cu:324-402) in FP64; the slab-independent
// part.  valid <=> disc > 1e-7 (the reference's only pre-test); hit on [tmin,tmax] <=>
// valid && max(tmin, r1) <= min(tmax, r2).
Interval intersect64(const Geo& G, const real o[3], const real d[3]) {
    Interval I;
    I.valid = false;
    I.r1 = I.r2 = 0.0;
    I.disc = I.b = I.c = 0.0;
    if (!G.valid) return I;
    real O[3], Op[3], Dp[3];
    for (int k = 0; k < 3; ++k) O[k] = o[k] - G.mu[k];
    for (int k = 0; k < 3; ++k) {
        Op[k] = G.Mp[k][0] * O[0] + G.Mp[k][1] * O[1] + G.Mp[k][2] * O[2];
        Dp[k] = G.Mp[k][0] * d[0] + G.Mp[k][1] * d[1] + G.Mp[k][2] * d[2];
    }
    const real l = 1.0 / std::sqrt(Dp[0] * Dp[0] + Dp[1] * Dp[1] + Dp[2] * Dp[2]);
    real Dh[3];
    for (int k = 0; k < 3; ++k) Dh[k] = Dp[k] * l;
    const real b = -(Op[0] * Dh[0] + Op[1] * Dh[1] + Op[2] * Dh[2]);
    const real c = Op[0] * Op[0] + Op[1] * Op[1] + Op[2] * Op[2] - 1.0;
    real pp = 0.0;
    for (int k = 0; k < 3; ++k) {
        const real p = Op[k] + b * Dh[k];
        pp += p * p;
    }
    const real disc = 1.0 - pp;
    I.disc = disc;
    I.b = b;
    I.c = c;
    if (disc > 1e-7) {
        const real sd = std::sqrt(disc);
        const real q = (b > 0.0) ? b + sd : b - sd;
        I.r1 = (c / q) * l;
        I.r2 = q * l;
        I.valid = true;
    }
    return I;
}

inline bool overlap(real tmin, real tmax, real r1, real r2) {
    return std::fmax(tmin, r1) <= std::fmin(tmax, r2);
}

// ---------------------------------------------------------------- CPU BVH (oracle only) -----

struct BNode {
    real lo[3], hi[3];
    int64_t left, right;   // >=0 internal; leaf when left < 0: prims [first, first+count)
    int64_t first, count;
};

struct CpuBVH {
    std::vector<BNode> nodes;
    std::vector<int64_t> prims;
};

void bvh_build_rec(CpuBVH& B, const std::vector<real>& box, std::vector<int64_t>& idx, int64_t lo,
                   int64_t hi, int64_t node) {
    BNode nd;
    for (int k = 0; k < 3; ++k) {
        nd.lo[k] = INFINITY;
        nd.hi[k] = -INFINITY;
    }
    for (int64_t i = lo; i < hi; ++i) {
        const real* b = &box[6 * idx[i]];
        for (int k = 0; k < 3; ++k) {
            nd.lo[k] = std::min(nd.lo[k], b[k]);
            nd.hi[k] = std::max(nd.hi[k], b[3 + k]);
        }
    }
    if (hi - lo <= 4) {
        nd.left = nd.right = -1;
        nd.first = lo;
        nd.count = hi - lo;
        B.nodes[node] = nd;
        return;
    }
    real clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = lo; i < hi; ++i) {
        const real* b = &box[6 * idx[i]];
        for (int k = 0; k < 3; ++k) {
            const real cc = 0.5 * (b[k] + b[3 + k]);
            clo[k] = std::min(clo[k], cc);
            chi[k] = std::max(chi[k], cc);
        }
    }
    int axis = 0;
    for (int k = 1; k < 3; ++k)
        if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
    const int64_t mid = (lo + hi) / 2;
    std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi, [&](int64_t a, int64_t b) {
        return box[6 * a + axis] + box[6 * a + 3 + axis] < box[6 * b + axis] + box[6 * b + 3 + axis];
    });
    const int64_t l = (int64_t)B.nodes.size();
    B.nodes.emplace_back();
    const int64_t r = (int64_t)B.nodes.size();
    B.nodes.emplace_back();
    nd.left = l;
    nd.right = r;
    nd.first = nd.count = 0;
    B.nodes[node] = nd;
    bvh_build_rec(B, box, idx, lo, mid, l);
    bvh_build_rec(B, box, idx, mid, hi, r);
}

// Conservative FP64 AABB of the bounding ellipsoid, padded so the FP64 predicate can never be
// true outside it.  Invalid Gaussians get no box (they can never be hit).
CpuBVH bvh_build(const std::vector<Geo>& geo) {
    const int64_t n = (int64_t)geo.size();
    std::vector<real> box(6 * n);
    std::vector<int64_t> idx;
    idx.reserve(n);
    for (int64_t g = 0; g < n; ++g) {
        const Geo& G = geo[g];
        if (!G.valid) continue;
        for (int k = 0; k < 3; ++k) {
            // half extent along world axis k = || row k of (R * S') ||, S' = S*delta.
            // Mp rows are col_j(R)/s'_j, so s'_j = 1/||Mp row j||.
            real h2 = 0.0;
            for (int j = 0; j < 3; ++j) {
                const real nrm = std::sqrt(G.Mp[j][0] * G.Mp[j][0] + G.Mp[j][1] * G.Mp[j][1] +
                                             G.Mp[j][2] * G.Mp[j][2]);
                const real sp = 1.0 / nrm;
                const real e = G.R[k][j] * sp;
                h2 += e * e;
            }
            const real pad = kOracle ? 1e-9 : 1e-5;  // relative to the type's rounding
            const real h = std::sqrt(h2) * (1.0 + pad) + pad * (1.0 + std::fabs(G.mu[k]));
            box[6 * g + k] = G.mu[k] - h;
            box[6 * g + 3 + k] = G.mu[k] + h;
        }
        idx.push_back(g);
    }
    CpuBVH B;
    B.nodes.reserve(2 * idx.size() + 2);
    B.nodes.emplace_back();
    if (idx.empty()) {
        BNode& r = B.nodes[0];
        for (int k = 0; k < 3; ++k) {
            r.lo[k] = INFINITY;
            r.hi[k] = -INFINITY;
        }
        r.left = r.right = -1;
        r.first = r.count = 0;
        return B;
    }
    bvh_build_rec(B, box, idx, 0, (int64_t)idx.size(), 0);
    B.prims = idx;
    return B;
}

bool box_seg(const BNode& nd, const real o[3], const real invd[3], real t0, real t1) {
    real tn = t0, tf = t1;
    for (int k = 0; k < 3; ++k) {
        real a = (nd.lo[k] - o[k]) * invd[k];
        real b = (nd.hi[k] - o[k]) * invd[k];
        if (std::isnan(a) || std::isnan(b)) {
            // ray parallel to the slab with origin on its boundary: inside iff lo<=o<=hi
            if (o[k] < nd.lo[k] || o[k] > nd.hi[k]) return false;
            continue;
        }
        if (a > b) std::swap(a, b);
        tn = std::max(tn, a);
        tf = std::min(tf, b);
    }
    const real tol = kOracle ? 1e-12 : 1e-6;
    return tn <= tf * (1.0 + tol) + tol;
}

void bvh_query(const CpuBVH& B, const real o[3], const real d[3], real t0, real t1,
               std::vector<int64_t>& out) {
    out.clear();
    if (B.prims.empty()) return;
    real invd[3];
    for (int k = 0; k < 3; ++k) invd[k] = 1.0 / d[k];
    int64_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const BNode& nd = B.nodes[stack[--sp]];
        if (!box_seg(nd, o, invd, t0, t1)) continue;
        if (nd.left < 0) {
            for (int64_t i = 0; i < nd.count; ++i) out.push_back(B.prims[nd.first + i]);
        } else {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
        }
    }
}

// ---------------------------------------------------------------- SH / SG ------------------

const real kC0 = 0.28209479177387814;
const real kC1 = 0.4886025119029199;
const real kC2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792,
                       0.5462742152960396};
const real kC3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
                       -0.4570457994644658, 1.445305721320277, -0.5900435899266435};

// Real SH basis in the reference's order and signs (utils/sh_utils.cu:60-117).
void sh_basis(real x, real y, real z, int deg, real* Y) {
    Y[0] = kC0;
    if (deg < 1) return;
    Y[1] = -kC1 * y;
    Y[2] = kC1 * z;
    Y[3] = -kC1 * x;
    if (deg < 2) return;
    const real xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    Y[4] = kC2[0] * xy;
    Y[5] = kC2[1] * yz;
    Y[6] = kC2[2] * (2.0 * zz - xx - yy);
    Y[7] = kC2[3] * xz;
    Y[8] = kC2[4] * (xx - yy);
    if (deg < 3) return;
    Y[9] = kC3[0] * y * (3.0 * xx - yy);
    Y[10] = kC3[1] * xy * z;
    Y[11] = kC3[2] * y * (4.0 * zz - xx - yy);
    Y[12] = kC3[3] * z * (2.0 * zz - 3.0 * xx - 3.0 * yy);
    Y[13] = kC3[4] * x * (4.0 * zz - xx - yy);
    Y[14] = kC3[5] * z * (xx - yy);
    Y[15] = kC3[6] * x * (xx - 3.0 * yy);
}

// dY/dx, dY/dy, dY/dz as used by backward_sh (utils/sh_utils.cu:176-303).  With exact=0 the
// d/dy of Y9 is C3[0]*(3xx - yy), as the reference writes it (sh_utils.cu:254); the exact
// derivative is C3[0]*3*(xx - yy).
void sh_basis_grad(real x, real y, real z, int deg, int exact, real* dX, real* dY, real* dZ) {
    const int K = (deg + 1) * (deg + 1);
    for (int k = 0; k < K; ++k) dX[k] = dY[k] = dZ[k] = 0.0;
    if (deg < 1) return;
    dY[1] = -kC1;
    dZ[2] = kC1;
    dX[3] = -kC1;
    if (deg < 2) return;
    const real xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    dX[4] = kC2[0] * y;
    dY[4] = kC2[0] * x;
    dY[5] = kC2[1] * z;
    dZ[5] = kC2[1] * y;
    dX[6] = kC2[2] * 2.0 * -x;
    dY[6] = kC2[2] * 2.0 * -y;
    dZ[6] = kC2[2] * 2.0 * 2.0 * z;
    dX[7] = kC2[3] * z;
    dZ[7] = kC2[3] * x;
    dX[8] = kC2[4] * 2.0 * x;
    dY[8] = kC2[4] * 2.0 * -y;
    if (deg < 3) return;
    dX[9] = kC3[0] * 3.0 * 2.0 * xy;
    dY[9] = exact ? kC3[0] * 3.0 * (xx - yy) : kC3[0] * (3.0 * xx - yy);
    dX[10] = kC3[1] * yz;
    dY[10] = kC3[1] * xz;
    dZ[10] = kC3[1] * xy;
    dX[11] = kC3[2] * -2.0 * xy;
    dY[11] = kC3[2] * (-3.0 * yy + 4.0 * zz - xx);
    dZ[11] = kC3[2] * 4.0 * 2.0 * yz;
    dX[12] = kC3[3] * -3.0 * 2.0 * xz;
    dY[12] = kC3[3] * -3.0 * 2.0 * yz;
    dZ[12] = kC3[3] * 3.0 * (2.0 * zz - xx - yy);
    dX[13] = kC3[4] * (-3.0 * xx + 4.0 * zz - yy);
    dY[13] = kC3[4] * -2.0 * xy;
    dZ[13] = kC3[4] * 4.0 * 2.0 * xz;
    dX[14] = kC3[5] * 2.0 * xz;
    dY[14] = kC3[5] * -2.0 * yz;
    dZ[14] = kC3[5] * (xx - yy);
    dX[15] = kC3[6] * 3.0 * (xx - yy);
    dY[15] = kC3[6] * -3.0 * 2.0 * xy;
}

// ---------------------------------------------------------------- per-ray march -------------

// ---- FP32 evaluations of sigma*w, for measuring the rounding margin of the threshold test ----

void quat_cols32(const float* q, float c0[3], float c1[3], float c2[3]) {
    const float r = q[0], i = q[1], j = q[2], k = q[3];
    c0[0] = 1.0f - 2.0f * (j * j + k * k), c0[1] = 2.f * (i * j + r * k), c0[2] = 2.f * (i * k - r * j);
    c1[0] = 2.f * (i * j - r * k), c1[1] = 1.0f - 2.0f * (i * i + k * k), c1[2] = 2.f * (j * k + r * i);
    c2[0] = 2.f * (i * k + r * j), c2[1] = 2.f * (j * k - r * i), c2[2] = 1.0f - 2.0f * (i * i + j * j);
}

// sigma*w as the reference evaluates it in FP32 (computeBufferForward,
// forward/gaussians_aabb.cu:437-498): M = (R S^-1)^T from the float quaternion and scales,
// x = o + d t, power = -0.5 |M (x - mu)|^2, w = expf(power).
float sigw_ref32(const orc_scene& S, int64_t g, const Ray32& r, float t) {
    float c0[3], c1[3], c2[3];
    quat_cols32(S.quat + 4 * g, c0, c1, c2);
    const float* s = S.scale + 3 * g;
    const float is0 = 1.0f / s[0], is1 = 1.0f / s[1], is2 = 1.0f / s[2];
    float U[3], V[3], W[3];
    for (int k = 0; k < 3; ++k) {
        U[k] = c0[k] * is0;
        V[k] = c1[k] * is1;
        W[k] = c2[k] * is2;
    }
    float X[3];
    for (int k = 0; k < 3; ++k) {
        const float x = r.o[k] + r.d[k] * t;
        X[k] = x - S.pos[3 * g + k];
    }
    const float m0 = (X[0] * U[0] + X[1] * U[1]) + X[2] * U[2];
    const float m1 = (X[0] * V[0] + X[1] * V[1]) + X[2] * V[2];
    const float m2 = (X[0] * W[0] + X[1] * W[1]) + X[2] * W[2];
    const float power = -0.5f * ((m0 * m0 + m1 * m1) + m2 * m2);
    return S.dens[g] * std::exp(power);
}

// sigma*w as the HIP path evaluates it in FP32: the preprocessed record (rows of
// M' = col_k(R) / (s_k delta), h = ln|sigma/thr|), the predicate's terms at the closest approach
// t_c re-centred at x0 = o + t_c d, and w = exp2(-log2(e) (A + Cq (t - t_c)^2)).  Returns a
// negative value for a Gaussian without a bounding ellipsoid (never sampled there).
float sigw_closed32(const orc_scene& S, int64_t g, float thr, const Ray32& r, float t) {
    const float sigma = S.dens[g];
    const float ratio = sigma / thr;
    const float l2 = std::log(ratio * ratio);
    if (!(l2 > 0.0f && l2 < INFINITY)) return -1.0f;
    const float delta = std::sqrt(l2), h = 0.5f * l2;
    float c0[3], c1[3], c2[3];
    quat_cols32(S.quat + 4 * g, c0, c1, c2);
    const float* s = S.scale + 3 * g;
    const float i0 = 1.0f / (s[0] * delta), i1 = 1.0f / (s[1] * delta), i2 = 1.0f / (s[2] * delta);
    float M[3][3];
    for (int k = 0; k < 3; ++k) {
        M[0][k] = c0[k] * i0;
        M[1][k] = c1[k] * i1;
        M[2][k] = c2[k] * i2;
    }
    const float* mu = S.pos + 3 * g;
    const float Ox = r.o[0] - mu[0], Oy = r.o[1] - mu[1], Oz = r.o[2] - mu[2];
    float Op[3], Dp[3];
    for (int k = 0; k < 3; ++k) {
        Op[k] = (M[k][0] * Ox + M[k][1] * Oy) + M[k][2] * Oz;
        Dp[k] = (M[k][0] * r.d[0] + M[k][1] * r.d[1]) + M[k][2] * r.d[2];
    }
    const float ddp = (Dp[0] * Dp[0] + Dp[1] * Dp[1]) + Dp[2] * Dp[2];
    const float l = 1.0f / std::sqrt(ddp);
    float Dh[3];
    for (int k = 0; k < 3; ++k) Dh[k] = Dp[k] * l;
    const float b = -((Op[0] * Dh[0] + Op[1] * Dh[1]) + Op[2] * Dh[2]);
    const float t0 = b * l;
    const float Cq = h * ddp;
    float X0[3], O[3];
    for (int k = 0; k < 3; ++k) X0[k] = (r.o[k] + t0 * r.d[k]) - mu[k];
    for (int k = 0; k < 3; ++k) O[k] = M[k][0] * X0[0] + M[k][1] * X0[1] + M[k][2] * X0[2];
    const float b2 = -(O[0] * Dh[0] + O[1] * Dh[1] + O[2] * Dh[2]);
    float P[3];
    for (int k = 0; k < 3; ++k) P[k] = O[k] + b2 * Dh[k];
    const float tc = t0 + b2 * l;
    const float A = h * (P[0] * P[0] + P[1] * P[1] + P[2] * P[2]);
    constexpr float kLog2e = 1.4426950408889634f;
    const float a = -A * kLog2e, c = -Cq * kLog2e;
    const float tau = t - tc;
    return sigma * std::exp2(std::fma(c, tau * tau, a));
}

// FP32 bounding-ellipsoid predicate on a record, operation for operation as the HIP path:
// returns whether disc > 1e-7, with the interval [r1, r2].
bool pred32(const float* mu, const float* m, const Ray32& r, float* r1, float* r2) {
    const float Ox = r.o[0] - mu[0], Oy = r.o[1] - mu[1], Oz = r.o[2] - mu[2];
    float Op[3], Dp[3];
    for (int k = 0; k < 3; ++k) {
        const float a0 = m[3 * k + 0] * Ox;
        const float a1 = m[3 * k + 1] * Oy;
        const float a2 = m[3 * k + 2] * Oz;
        Op[k] = (a0 + a1) + a2;
        const float b0 = m[3 * k + 0] * r.d[0];
        const float b1 = m[3 * k + 1] * r.d[1];
        const float b2 = m[3 * k + 2] * r.d[2];
        Dp[k] = (b0 + b1) + b2;
    }
    const float ddp = (Dp[0] * Dp[0] + Dp[1] * Dp[1]) + Dp[2] * Dp[2];
    const float l = 1.0f / std::sqrt(ddp);
    float Dh[3];
    for (int k = 0; k < 3; ++k) Dh[k] = Dp[k] * l;
    const float b = -((Op[0] * Dh[0] + Op[1] * Dh[1]) + Op[2] * Dh[2]);
    const float c = ((Op[0] * Op[0] + Op[1] * Op[1]) + Op[2] * Op[2]) - 1.0f;
    float P[3];
    for (int k = 0; k < 3; ++k) {
        const float e = b * Dh[k];
        P[k] = Op[k] + e;
    }
    const float pp = (P[0] * P[0] + P[1] * P[1]) + P[2] * P[2];
    const float disc = 1.0f - pp;
    if (!(disc > 1e-7f)) return false;
    const float sd = std::sqrt(disc);
    const float q = (b > 0.0f) ? b + sd : b - sd;
    *r1 = (c / q) * l;
    *r2 = q * l;
    return true;
}

struct Dec32 {
    orc_record rec;
    const float* sigma;  // densities as given
    float thr;
};

// sigma*w in FP32 from a record, as the HIP path samples a hit: the predicate's terms at the
// closest approach re-centred at x0 = o + t_c d, and w = exp2(-log2(e) (A + Cq (t - t_c)^2)).
float sigw32(const Dec32& D, int64_t g, const Ray32& r, float t) {
    const float* mu = D.rec.mu + 3 * g;
    const float* m = D.rec.Mp + 9 * g;
    const float h = D.rec.h[g];
    const float Ox = r.o[0] - mu[0], Oy = r.o[1] - mu[1], Oz = r.o[2] - mu[2];
    float Op[3], Dp[3];
    for (int k = 0; k < 3; ++k) {
        Op[k] = (m[3 * k] * Ox + m[3 * k + 1] * Oy) + m[3 * k + 2] * Oz;
        Dp[k] = (m[3 * k] * r.d[0] + m[3 * k + 1] * r.d[1]) + m[3 * k + 2] * r.d[2];
    }
    const float ddp = (Dp[0] * Dp[0] + Dp[1] * Dp[1]) + Dp[2] * Dp[2];
    const float l = 1.0f / std::sqrt(ddp);
    float Dh[3];
    for (int k = 0; k < 3; ++k) Dh[k] = Dp[k] * l;
    const float b = -((Op[0] * Dh[0] + Op[1] * Dh[1]) + Op[2] * Dh[2]);
    const float t0 = b * l;
    const float Cq = h * ddp;
    float X0[3], O[3];
    for (int k = 0; k < 3; ++k) X0[k] = (r.o[k] + t0 * r.d[k]) - mu[k];
    for (int k = 0; k < 3; ++k) O[k] = (m[3 * k] * X0[0] + m[3 * k + 1] * X0[1]) + m[3 * k + 2] * X0[2];
    const float b2 = -((O[0] * Dh[0] + O[1] * Dh[1]) + O[2] * Dh[2]);
    float P[3];
    for (int k = 0; k < 3; ++k) P[k] = O[k] + b2 * Dh[k];
    const float tc = t0 + b2 * l;
    const float A = h * ((P[0] * P[0] + P[1] * P[1]) + P[2] * P[2]);
    constexpr float kLog2e = 1.4426950408889634f;
    const float a = -A * kLog2e, c = -Cq * kLog2e;
    const float tau = t - tc;
    return D.sigma[g] * std::exp2(std::fma(c, tau * tau, a));
}

// Decision-matched mode: a sample decision within this relative distance of the threshold may be
// taken differently by the GPU's native exp2 (about one ulp from the host's).
const real kEpsDec = 1e-6;

// Fragility kinds (RayOut.fragile bits): near-threshold sample decisions, a transmittance check
// within rounding of T_eps, the origin-inside branch decided by a b within rounding of zero.
enum { kFragSample = 1, kFragExit = 2, kFragQuirk = 4 };

struct RayOut {
    real C[3];
    real depth;
    real T;
    int fragile;      // kFrag* bits
    int64_t nfrag;    // sample decisions (hit or near Gaussians) with |sigma*w/thr - 1| <= eps_w
    real cmax;      // largest |colour component| among the ray's hit Gaussians and the background
    int64_t K;        // slab interactions (sum over slabs of |hit set|)
    int64_t P;        // distinct (ray, Gaussian) pairs among hit sets
    int64_t maxhits;  // largest per-slab hit set
    int64_t slabs;    // slabs visited (loop iterations)
    int64_t nc;       // contributing (sample, hit) events: sigma*w > thr
};

inline bool near_rel(real a, real b, real rel) {
    return std::fabs(a - b) <= rel * std::max(std::fabs(a), std::fabs(b));
}

// Fragility: a ray is flagged when some threshold decision of the reference's algorithm lies
// within FP32 rounding of its boundary, so that an FP32 evaluation may legitimately decide it
// the other way.  The decisions that change the image are the sample tests sigma*w > thr; a
// hit-set decision (disc vs 1e-7, interval end vs slab end) only matters through the samples of
// that Gaussian in that slab, and every such sample then has sigma*w at the threshold.  So the
// oracle also evaluates, without accumulating, the samples of "near" Gaussians -- predicate
// just failing (disc within kEpsD below 1e-7) or interval just missing the slab (within
// kEpsT) -- and counts the samples of hit or near Gaussians with |sigma*w/thr - 1| <= eps_w.
// eps_w is set from a measurement: orc_sigmaw_fp32_error() reports the largest relative FP32
// error of sigma*w within 1 % of the threshold, for the reference's FP32 evaluation
// (computeBufferForward, forward/gaussians_aabb.cu:437-498) and for the HIP path's closed form,
// and the tests assert that eps_w is at least twice that error.
//
// Each counted decision bounds what a flip can do to the colour: the sample's density changes by
// d <= thr (1 + eps_w), which moves C by at most d dt (3 cmax + max(cmax, bg)) (the sample's
// opacity, its mixed colour and the transmittance of everything behind it), i.e. at most
// 4 thr (1 + eps_w) dt max(cmax, 1) per decision.
const real kEpsW = 1e-4;  // default relative margin on sigma*w
const real kEpsD = 5e-5;  // absolute, on disc
inline real eps_t(real t) { return 1e-6 * std::max(1.0, std::fabs(t)); }

struct Hit {
    int64_t g;
    real r1, r2;
    bool near;  // predicate fails by less than kEpsD: evaluated for fragility only
};

// Collects every Gaussian whose predicate interval overlaps [tenter, texit] (and the near
// misses), sorted by (r1, g).
void gather(const CpuBVH& B, const std::vector<Geo>& geo, const real o[3], const real d[3], real t0,
            real t1, std::vector<Hit>& hits, std::vector<int64_t>& scratch, int* fragile) {
    // fragile: kFragQuirk is or-ed in
    hits.clear();
    bvh_query(B, o, d, t0 - eps_t(t0), t1 + eps_t(t1), scratch);
    for (int64_t g : scratch) {
        const Interval I = intersect64(geo[g], o, d);
        if (!kOracle) {  // baseline: hits only, no margins
            if (I.valid && overlap(t0, t1, I.r1, I.r2)) hits.push_back(Hit{g, I.r1, I.r2, false});
            continue;
        }
        // the b > 0 branch decides hit / no hit when the origin is inside the ellipsoid
        if (I.c < 0.0 && std::fabs(I.b) < 1e-5) *fragile |= kFragQuirk;
        Hit h{g, I.r1, I.r2, false};
        if (!I.valid) {
            if (!(geo[g].valid && I.disc > 1e-7 - kEpsD && I.b > 0.0)) continue;
            // near miss: interval of the ellipsoid it grazes
            const real sd = std::sqrt(std::max(I.disc, 0.0));
            const real q = I.b + sd;
            real l = 0.0;
            {
                real Dp[3];
                for (int k = 0; k < 3; ++k)
                    Dp[k] = geo[g].Mp[k][0] * d[0] + geo[g].Mp[k][1] * d[1] + geo[g].Mp[k][2] * d[2];
                l = 1.0 / std::sqrt(Dp[0] * Dp[0] + Dp[1] * Dp[1] + Dp[2] * Dp[2]);
            }
            h.r1 = (I.b - sd) * l;
            h.r2 = q * l;
            h.near = true;
        }
        if (!overlap(t0 - eps_t(t0), t1 + eps_t(t1), h.r1, h.r2)) continue;
        hits.push_back(h);
    }
    std::sort(hits.begin(), hits.end(), [](const Hit& a, const Hit& b) {
        return a.r1 < b.r1 || (a.r1 == b.r1 && a.g < b.g);
    });
}

// Slab iterator shared by forward and backward: produces, per slab, the hit set (indices into
// `hits`) in ascending Gaussian order, and the near set (entries that miss the slab's hit set
// by less than the rounding margins).
struct SlabWalker {
    const std::vector<Hit>* hits;
    size_t next = 0;
    std::vector<int> active;  // indices into hits
    std::vector<int> cur;
    std::vector<int> near;
    void step(real tmin, real tmax, int* fragile) {
        (void)fragile;
        const std::vector<Hit>& H = *hits;
        const real emin = eps_t(tmin), emax = eps_t(tmax);
        while (next < H.size() && H[next].r1 <= tmax + emax) active.push_back((int)next++);
        size_t w = 0;
        for (size_t a = 0; a < active.size(); ++a)
            if (H[active[a]].r2 >= tmin - emin) active[w++] = active[a];
        active.resize(w);
        cur.clear();
        near.clear();
        for (int a : active) {
            if (!H[a].near && overlap(tmin, tmax, H[a].r1, H[a].r2)) cur.push_back(a);
            else if (overlap(tmin - emin, tmax + emax, H[a].r1, H[a].r2)) near.push_back(a);
        }
        std::sort(cur.begin(), cur.end(), [&](int a, int b) { return H[a].g < H[b].g; });
    }
    bool exhausted() const { return next >= hits->size() && active.empty(); }
};

struct SampleBuf {
    std::vector<real> sig, c0, c1, c2;
    void reset(int spb) {
        sig.assign(spb, 0.0);
        c0.assign(spb, 0.0);
        c1.assign(spb, 0.0);
        c2.assign(spb, 0.0);
    }
};

inline real gauss_w(const Geo& G, const real x[3], real y[3]) {
    real X[3];
    for (int k = 0; k < 3; ++k) X[k] = x[k] - G.mu[k];
    for (int k = 0; k < 3; ++k) y[k] = G.M[k][0] * X[0] + G.M[k][1] * X[1] + G.M[k][2] * X[2];
    return std::exp(-0.5 * (y[0] * y[0] + y[1] * y[1] + y[2] * y[2]));
}

// computeBufferForward (forward/gaussians_aabb.cu:437-498) for one slab, FP64.
void accumulate_slab(const orc_params& p, const std::vector<Geo>& geo, const real* rgb,
                     const std::vector<Hit>& H, const std::vector<int>& cur, const Ray32& r, const SlabSamples& ts,
                     SampleBuf& buf, int64_t* nfrag, int64_t* nc = nullptr, const std::vector<int>* near = nullptr,
                     const Dec32* dec = nullptr) {
    const real thr = p.sigma_thr;
    const real epsw = p.eps_w > 0.0f ? (real)p.eps_w : kEpsW;
    if (near)
        for (int a : *near) {
            const Geo& G = geo[H[a].g];
            for (int i = 0; i < p.spb; ++i) {
                const real t = ts.t[i];
                real x[3], y[3];
                for (int k = 0; k < 3; ++k) x[k] = (real)r.o[k] + t * (real)r.d[k];
                if (G.sigma * gauss_w(G, x, y) >= thr * (1.0 - epsw)) ++*nfrag;
            }
        }
    for (int a : cur) {
        const Geo& G = geo[H[a].g];
        for (int i = 0; i < p.spb; ++i) {
            const real t = ts.t[i];
            real x[3], y[3];
            for (int k = 0; k < 3; ++k) x[k] = (real)r.o[k] + t * (real)r.d[k];
            const real w = gauss_w(G, x, y);
            const real sw = G.sigma * w;
            bool take;
            if (dec) {  // the HIP path's FP32 decision, the FP64 value
                const float v = sigw32(*dec, H[a].g, r, ts.t[i]);
                if (std::fabs((real)v / (real)dec->thr - 1.0) <= kEpsDec) ++*nfrag;
                take = v > dec->thr;
            } else {
                if (kOracle && std::fabs(sw / thr - 1.0) <= epsw) ++*nfrag;
                take = sw > thr;
            }
            if (take) {
                if (nc) ++*nc;
                buf.sig[i] += sw;
                buf.c0[i] += sw * rgb[3 * H[a].g + 0];
                buf.c1[i] += sw * rgb[3 * H[a].g + 1];
                buf.c2[i] += sw * rgb[3 * H[a].g + 2];
            }
        }
    }
}

struct RayCtx {
    Ray32 r;
    real o[3], d[3];
    float tenter, texit;
    uint32_t seed;  // generator state after ray generation
};

RayCtx make_ctx(const orc_params& p, const orc_camera& cam, int64_t pixel, const float* bmin, const float* bmax) {
    RayCtx c;
    c.seed = tea4(p.iteration, (uint32_t)pixel);
    c.r = make_ray32(cam, pixel, p.jitter, &c.seed);
    for (int k = 0; k < 3; ++k) {
        c.o[k] = c.r.o[k];
        c.d[k] = c.r.d[k];
    }
    scene_interval32(c.r, bmin, bmax, p.t_near, p.t_far, &c.tenter, &c.texit);
    return c;
}

// Slab hit set of the decision-matched mode: the gathered hit and near entries that the FP32
// record predicate accepts on [tmin, tmax] (the near entries cover every FP32-only hit: the FP32
// predicate differs from the FP64 one far inside the near margins), ascending by Gaussian.
const std::vector<int>& dec_members(const Dec32& D, const std::vector<Hit>& H, const SlabWalker& W, const Ray32& r,
                                    float tmin, float tmax, std::vector<int>& out) {
    out.clear();
    for (const std::vector<int>* L : {&W.cur, &W.near})
        for (int a : *L) {
            const int64_t g = H[a].g;
            float r1, r2;
            if (D.rec.valid[g] && pred32(D.rec.mu + 3 * g, D.rec.Mp + 9 * g, r, &r1, &r2) &&
                std::fmax(tmin, r1) <= std::fmin(tmax, r2))
                out.push_back(a);
        }
    std::sort(out.begin(), out.end(), [&](int a, int b) { return H[a].g < H[b].g; });
    return out;
}

// __raygen__rg forward (forward/gaussians_aabb.cu:521-612) + colorBlending (:500-519), FP64.
RayOut march_forward(const orc_params& p, const std::vector<Geo>& geo, const CpuBVH& B, const real* rgb,
                     const RayCtx& c, std::vector<Hit>& hits, std::vector<int64_t>& scratch,
                     const Dec32* dec = nullptr) {
    RayOut out;
    std::memset(&out, 0, sizeof(out));
    const real bg = p.white_bg ? 1.0 : 0.0;
    real T = 1.0;
    if (!(c.tenter < c.texit)) {
        for (int k = 0; k < 3; ++k) out.C[k] = bg;
        out.T = 1.0;
        return out;
    }
    gather(B, geo, c.o, c.d, c.tenter, c.texit, hits, scratch, &out.fragile);
    out.cmax = bg;
    for (const Hit& h : hits)
        if (!h.near)
            for (int k = 0; k < 3; ++k) out.cmax = std::max(out.cmax, std::fabs(rgb[3 * h.g + k]));
    SlabWalker W;
    W.hits = &hits;
    SampleBuf buf;
    const float slab = p.dt * (float)p.spb;
    const real dt = p.dt;
    const real eps = p.t_eps;
    float tb = c.tenter;
    uint32_t seed = c.seed;
    std::vector<char> seen(hits.size(), 0);
    std::vector<int> memv;
    while (tb < c.texit && T > eps) {
        if (near_rel(T, eps, 1e-3)) out.fragile |= kFragExit;
        const float tmin = std::fmax(c.tenter, tb);
        const float tmax = std::fmin(c.texit, tb + slab);
        out.slabs++;
        if (tmax > c.tenter) {
            W.step(tmin, tmax, &out.fragile);
            const SlabSamples ts(p, tb, seed);
            const std::vector<int>& cur = dec ? dec_members(*dec, hits, W, c.r, tmin, tmax, memv) : W.cur;
            if (!dec && W.cur.empty() && !W.near.empty()) {
                buf.reset(p.spb);
                const std::vector<int> none;
                accumulate_slab(p, geo, rgb, hits, none, c.r, ts, buf, &out.nfrag, nullptr, &W.near);
            }
            if (!cur.empty()) {
                out.K += (int64_t)cur.size();
                out.maxhits = std::max<int64_t>(out.maxhits, (int64_t)cur.size());
                for (int a : cur)
                    if (!seen[a]) {
                        seen[a] = 1;
                        out.P++;
                    }
                buf.reset(p.spb);
                accumulate_slab(p, geo, rgb, hits, cur, c.r, ts, buf, &out.nfrag, &out.nc, dec ? nullptr : &W.near,
                                dec);
                for (int i = 0; i < p.spb; ++i) {
                    const real s = buf.sig[i];
                    const real alpha = 1.0 - std::exp(-s * dt);
                    real ch[3] = {buf.c0[i], buf.c1[i], buf.c2[i]};
                    if (s > 0.0)
                        for (int k = 0; k < 3; ++k) ch[k] /= s;
                    for (int k = 0; k < 3; ++k) out.C[k] += T * alpha * ch[k];
                    out.depth += T * alpha * (real)ts.t[i];
                    T *= 1.0 - alpha;
                }
                seed = after_slab(p, seed);
            } else if (W.exhausted()) {
                // nothing left on this ray: the remaining slabs are empty; skip them (T unchanged)
                float t = tb;
                while (t < c.texit) t = t + slab;
                tb = t;
                continue;
            }
        }
        tb = tb + slab;
    }
    for (int k = 0; k < 3; ++k) out.C[k] += T * bg;
    out.T = T;
    if (out.nfrag > 0) out.fragile |= kFragSample;
    return out;
}

// Literal restatement: every Gaussian's predicate is evaluated in every slab (no interval
// reuse, no BVH).  Quadratic cost; for pinning march_forward on small scenes only.
RayOut march_forward_literal(const orc_params& p, const std::vector<Geo>& geo, const real* rgb, const RayCtx& c) {
    RayOut out;
    std::memset(&out, 0, sizeof(out));
    const real bg = p.white_bg ? 1.0 : 0.0;
    real T = 1.0;
    if (!(c.tenter < c.texit)) {
        for (int k = 0; k < 3; ++k) out.C[k] = bg;
        out.T = 1.0;
        return out;
    }
    const float slab = p.dt * (float)p.spb;
    float tb = c.tenter;
    uint32_t seed = c.seed;
    std::vector<Hit> hits;
    std::vector<int> cur;
    SampleBuf buf;
    while (tb < c.texit && T > (real)p.t_eps) {
        const float tmin = std::fmax(c.tenter, tb);
        const float tmax = std::fmin(c.texit, tb + slab);
        out.slabs++;
        if (tmax > c.tenter) {
            hits.clear();
            for (int64_t g = 0; g < (int64_t)geo.size(); ++g) {
                const Interval I = intersect64(geo[g], c.o, c.d);
                if (I.valid && overlap(tmin, tmax, I.r1, I.r2)) hits.push_back({g, I.r1, I.r2});
            }
            if (!hits.empty()) {
                cur.resize(hits.size());
                for (size_t i = 0; i < hits.size(); ++i) cur[i] = (int)i;
                out.K += (int64_t)hits.size();
                buf.reset(p.spb);
                int64_t dummy = 0;
                const SlabSamples ts(p, tb, seed);
                accumulate_slab(p, geo, rgb, hits, cur, c.r, ts, buf, &dummy);
                for (int i = 0; i < p.spb; ++i) {
                    const real s = buf.sig[i];
                    const real alpha = 1.0 - std::exp(-s * (real)p.dt);
                    real ch[3] = {buf.c0[i], buf.c1[i], buf.c2[i]};
                    if (s > 0.0)
                        for (int k = 0; k < 3; ++k) ch[k] /= s;
                    for (int k = 0; k < 3; ++k) out.C[k] += T * alpha * ch[k];
                    out.depth += T * alpha * (real)ts.t[i];
                    T *= 1.0 - alpha;
                }
                seed = after_slab(p, seed);
            }
        }
        tb = tb + slab;
    }
    for (int k = 0; k < 3; ++k) out.C[k] += T * bg;
    out.T = T;
    return out;
}

inline void atomic_add(real* p, real v) {
#pragma omp atomic
    *p += v;
}

// __raygen__rg backward (backward/gaussians_aabb.cu:505-775), FP64, literal per-hit formulas.
void march_backward(const orc_params& p, const std::vector<Geo>& geo, const CpuBVH& B, const real* rgb,
                    const RayCtx& c, const real* Cfwd, const real* dLdC, real* dpos, real* dscale,
                    real* dquat, real* ddens, real* drgb, const orc_scene& S, std::vector<Hit>& hits,
                    std::vector<int64_t>& scratch, const Dec32* dec = nullptr) {
    if (!(c.tenter < c.texit)) return;
    int fragile = 0;
    gather(B, geo, c.o, c.d, c.tenter, c.texit, hits, scratch, &fragile);
    SlabWalker W;
    W.hits = &hits;
    SampleBuf buf;
    const float slab = p.dt * (float)p.spb;
    const real dt = p.dt;
    const real thr = p.sigma_thr;
    real T = 1.0;
    real Crem[3] = {Cfwd[0], Cfwd[1], Cfwd[2]};
    float tb = c.tenter;
    uint32_t seed = c.seed;
    const int spb = p.spb;
    std::vector<real> alpha(spb), chat(3 * spb);
    std::vector<int> memv;
    while (tb < c.texit && T > (real)p.t_eps) {
        const float tmin = std::fmax(c.tenter, tb);
        const float tmax = std::fmin(c.texit, tb + slab);
        if (tmax > c.tenter) {
            W.step(tmin, tmax, &fragile);
            const std::vector<int>& cur = dec ? dec_members(*dec, hits, W, c.r, tmin, tmax, memv) : W.cur;
            if (!cur.empty()) {
                buf.reset(spb);
                const SlabSamples ts(p, tb, seed);
                seed = after_slab(p, seed);
                int64_t nfrag = 0;
                accumulate_slab(p, geo, rgb, hits, cur, c.r, ts, buf, &nfrag, nullptr, nullptr, dec);
                for (int i = 0; i < spb; ++i) {
                    const real s = buf.sig[i];
                    real ch[3] = {buf.c0[i], buf.c1[i], buf.c2[i]};
                    if (s > 0.0)
                        for (int k = 0; k < 3; ++k) ch[k] /= s;
                    for (int k = 0; k < 3; ++k) chat[3 * i + k] = ch[k];
                    alpha[i] = 1.0 - std::exp(-s * dt);
                }
                real Tn = T, Cn[3] = {Crem[0], Crem[1], Crem[2]};
                for (int a : cur) {
                    const int64_t g = hits[a].g;
                    const Geo& G = geo[g];
                    real Ta = T;
                    real Ca[3] = {Crem[0], Crem[1], Crem[2]};
                    real grad_color = 0.0;
                    real dcds[3] = {0, 0, 0};
                    real dsc[3] = {0, 0, 0}, dps[3] = {0, 0, 0};
                    real dr[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // dr[k] = dL/dcol_k
                    const real* col = rgb + 3 * g;
                    for (int i = 0; i < spb; ++i) {
                        const real cd = buf.sig[i];
                        const real* cc = &chat[3 * i];
                        const real al = alpha[i];
                        const real t = ts.t[i];
                        real x[3], y[3];
                        for (int k = 0; k < 3; ++k) x[k] = (real)c.r.o[k] + t * (real)c.r.d[k];
                        const real w = gauss_w(G, x, y);
                        if (dec ? sigw32(*dec, g, c.r, ts.t[i]) > dec->thr : G.sigma * w > thr) {
                            real wn = w * G.sigma;
                            if (cd > 0.0) wn /= cd;
                            grad_color += wn * al * Ta;
                            real aux[3], dcdw[3];
                            for (int k = 0; k < 3; ++k) aux[k] = (Ta * cc[k] - Ca[k]) * dt;
                            for (int k = 0; k < 3; ++k) {
                                dcds[k] += aux[k] * w;
                                dcdw[k] = aux[k] * G.sigma;
                            }
                            if (cd > 0.0)
                                for (int k = 0; k < 3; ++k) {
                                    dcdw[k] += Ta * al * ((col[k] - cc[k]) / cd) * G.sigma;
                                    dcds[k] += Ta * al * ((col[k] - cc[k]) / cd) * w;
                                }
                            const real dLdw = dLdC[0] * dcdw[0] + dLdC[1] * dcdw[1] + dLdC[2] * dcdw[2];
                            real X[3];
                            for (int k = 0; k < 3; ++k) X[k] = x[k] - G.mu[k];
                            for (int k = 0; k < 3; ++k) dsc[k] += dLdw * w * (y[k] * y[k]) / G.s[k];
                            for (int k = 0; k < 3; ++k) {
                                // dot(M_k, y) with M_k = column k of M (coefficients of X[k])
                                const real m = G.M[0][k] * y[0] + G.M[1][k] * y[1] + G.M[2][k] * y[2];
                                dps[k] += dLdw * w * m;
                            }
                            for (int kk = 0; kk < 3; ++kk)
                                for (int k = 0; k < 3; ++k) dr[kk][k] -= dLdw * w * X[k] * (y[kk] / G.s[kk]);
                        }
                        for (int k = 0; k < 3; ++k) Ca[k] -= Ta * al * cc[k];
                        Ta *= 1.0 - al;
                    }
                    for (int k = 0; k < 3; ++k) atomic_add(&drgb[3 * g + k], dLdC[k] * grad_color);
                    atomic_add(&ddens[g], dLdC[0] * dcds[0] + dLdC[1] * dcds[1] + dLdC[2] * dcds[2]);
                    for (int k = 0; k < 3; ++k) atomic_add(&dscale[3 * g + k], dsc[k]);
                    for (int k = 0; k < 3; ++k) atomic_add(&dpos[3 * g + k], dps[k]);
                    // quaternion chain (backward/gaussians_aabb.cu:747-763), q = (r,i,j,k)
                    const float* q = S.quat + 4 * g;
                    const real r = q[0], qi = q[1], qj = q[2], qk = q[3];
                    const real* g0 = dr[0];
                    const real* g1 = dr[1];
                    const real* g2 = dr[2];
                    const real dq_r = 2 * qk * g0[1] - 2 * qj * g0[2] - 2 * qk * g1[0] + 2 * qi * g1[2] +
                                        2 * qj * g2[0] - 2 * qi * g2[1];
                    const real dq_i = 2 * qj * g0[1] + 2 * qk * g0[2] + 2 * qj * g1[0] - 4 * qi * g1[1] +
                                        2 * r * g1[2] + 2 * qk * g2[0] - 2 * r * g2[1] - 4 * qi * g2[2];
                    const real dq_j = -4 * qj * g0[0] + 2 * qi * g0[1] - 2 * r * g0[2] + 2 * qi * g1[0] +
                                        2 * qk * g1[2] + 2 * r * g2[0] + 2 * qk * g2[1] - 4 * qj * g2[2];
                    const real dq_k = -4 * qk * g0[0] + 2 * r * g0[1] + 2 * qi * g0[2] - 2 * r * g1[0] -
                                        4 * qk * g1[1] + 2 * qj * g1[2] + 2 * qi * g2[0] + 2 * qj * g2[1];
                    atomic_add(&dquat[4 * g + 0], dq_r);
                    atomic_add(&dquat[4 * g + 1], dq_i);
                    atomic_add(&dquat[4 * g + 2], dq_j);
                    atomic_add(&dquat[4 * g + 3], dq_k);
                    Tn = Ta;
                    for (int k = 0; k < 3; ++k) Cn[k] = Ca[k];
                }
                T = Tn;
                for (int k = 0; k < 3; ++k) Crem[k] = Cn[k];
            } else if (W.exhausted()) {
                return;
            }
        }
        tb = tb + slab;
    }
}

std::vector<Geo> make_geos(const orc_scene& S, real thr, int nthreads) {
    std::vector<Geo> geo(S.n);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t g = 0; g < S.n; ++g) geo[g] = make_geo(S, g, thr);
    return geo;
}

}  // namespace

extern "C" {

#ifndef ORC_BASELINE

int orc_version() { return 1; }

// Ray generation (FP32, bit-exact with the HIP path).  pixels may be null: ray r -> pixel r.
void orc_rays(const orc_camera* cam, int64_t nrays, const int32_t* pixels, float* o, float* d) {
    for (int64_t r = 0; r < nrays; ++r) {
        const Ray32 R = make_ray32(*cam, pixels ? pixels[r] : r);
        for (int k = 0; k < 3; ++k) {
            o[3 * r + k] = R.o[k];
            d[3 * r + k] = R.d[k];
        }
    }
}

// Ray generation with the reference's jitter (generator seeded with tea<4>(iteration, pixel)).
void orc_rays_rng(const orc_camera* cam, int64_t nrays, const int32_t* pixels, int32_t jitter, uint32_t iteration,
                  float* o, float* d) {
    for (int64_t r = 0; r < nrays; ++r) {
        const int64_t px = pixels ? pixels[r] : r;
        uint32_t seed = tea4(iteration, (uint32_t)px);
        const Ray32 R = make_ray32(*cam, px, jitter, &seed);
        for (int k = 0; k < 3; ++k) {
            o[3 * r + k] = R.o[k];
            d[3 * r + k] = R.d[k];
        }
    }
}

// The generator's pieces, for pinning: TEA with an explicit key and cycle count, and the first
// n LCG outputs (24-bit integers and rnd floats) from a seed.
uint32_t orc_tea(uint32_t v0, uint32_t v1, int32_t cycles, const uint32_t* key) {
    return tea_keyed(v0, v1, cycles, key);
}
uint32_t orc_tea4(uint32_t v0, uint32_t v1) { return tea4(v0, v1); }
void orc_lcg(uint32_t seed, int32_t n, uint32_t* ints, float* floats) {
    uint32_t s = seed, t = seed;
    for (int i = 0; i < n; ++i) {
        ints[i] = lcg(s);
        floats[i] = rnd(t);
    }
}

// Scene interval (FP32) per ray, for checking the HIP ray setup.
void orc_ray_interval(const orc_camera* cam, int64_t nrays, const int32_t* pixels, const float* bmin,
                      const float* bmax, float t_near, float t_far, float* tenter, float* texit) {
    for (int64_t r = 0; r < nrays; ++r) {
        const Ray32 R = make_ray32(*cam, pixels ? pixels[r] : r);
        scene_interval32(R, bmin, bmax, t_near, t_far, &tenter[r], &texit[r]);
    }
}

// forward_sh (utils/sh_utils.cu:25-119), FP64.  rgb: (n,3).
void orc_sh_forward(const orc_scene* S, const float* eye, real* rgb, int nthreads) {
    const int K = (S->deg + 1) * (S->deg + 1);
    const int L = S->L;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t g = 0; g < S->n; ++g) {
        real dir[3];
        for (int k = 0; k < 3; ++k) dir[k] = (real)S->pos[3 * g + k] - (real)eye[k];
        const real len = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        for (int k = 0; k < 3; ++k) dir[k] /= len;
        real Y[16];
        sh_basis(dir[0], dir[1], dir[2], S->deg, Y);
        for (int ch = 0; ch < 3; ++ch) {
            real acc = 0.0;
            for (int l = 0; l < L; ++l) {
                const real a = S->sg[g * 3 * L + ch * L + l];
                const real lam = S->lam[g * L + l];
                const float* ax = S->axis + g * 3 * L + 3 * l;
                const real dp = dir[0] * ax[0] + dir[1] * ax[1] + dir[2] * ax[2];
                acc += a * std::exp(lam * (dp - 1.0));
            }
            for (int k = 0; k < K; ++k) acc += S->sh[g * 3 * K + ch * K + k] * Y[k];
            rgb[3 * g + ch] = acc + 0.5;
        }
    }
}

// backward_sh (utils/sh_utils.cu:121-315), FP64.  Outputs are overwritten.
void orc_sh_backward(const orc_scene* S, const float* eye, const real* dLdrgb, int exact, real* dsh,
                     real* dsg, real* dlam, real* daxis, real* dpos_dir, int nthreads) {
    const int K = (S->deg + 1) * (S->deg + 1);
    const int L = S->L;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t g = 0; g < S->n; ++g) {
        const real* dL = dLdrgb + 3 * g;
        real d0[3];
        for (int k = 0; k < 3; ++k) d0[k] = (real)S->pos[3 * g + k] - (real)eye[k];
        const real len = std::sqrt(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]);
        real dir[3] = {d0[0] / len, d0[1] / len, d0[2] / len};
        real Y[16], dX[16], dY[16], dZ[16];
        sh_basis(dir[0], dir[1], dir[2], S->deg, Y);
        sh_basis_grad(dir[0], dir[1], dir[2], S->deg, exact, dX, dY, dZ);
        real ddir[3] = {0, 0, 0};
        for (int ch = 0; ch < 3; ++ch)
            for (int k = 0; k < K; ++k) {
                dsh[g * 3 * K + ch * K + k] = Y[k] * dL[ch];
                const real c = S->sh[g * 3 * K + ch * K + k];
                ddir[0] += dL[ch] * c * dX[k];
                ddir[1] += dL[ch] * c * dY[k];
                ddir[2] += dL[ch] * c * dZ[k];
            }
        for (int l = 0; l < L; ++l) {
            const float* ax = S->axis + g * 3 * L + 3 * l;
            const real lam = S->lam[g * L + l];
            const real dm1 = dir[0] * ax[0] + dir[1] * ax[1] + dir[2] * ax[2] - 1.0;
            const real wsg = std::exp(dm1 * lam);
            real dLdw = 0.0;
            for (int ch = 0; ch < 3; ++ch) {
                dsg[g * 3 * L + ch * L + l] = dL[ch] * wsg;
                dLdw += dL[ch] * S->sg[g * 3 * L + ch * L + l];
            }
            for (int k = 0; k < 3; ++k) daxis[g * 3 * L + 3 * l + k] = dLdw * lam * dir[k] * wsg;
            dlam[g * L + l] = dLdw * dm1 * wsg;
            for (int k = 0; k < 3; ++k) ddir[k] += dLdw * lam * wsg * ax[k];
        }
        // d normalize(d0) / d d0 = (I - dir dir^T)/|d0|
        const real s2 = d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2];
        const real inv = 1.0 / (std::sqrt(s2) * s2);
        for (int k = 0; k < 3; ++k) {
            real acc = s2 * ddir[k];
            for (int j = 0; j < 3; ++j) acc -= d0[j] * d0[k] * ddir[j];
            dpos_dir[3 * g + k] = acc * inv;
        }
    }
}

// compute_ellipsoids_bbox (optix_utils.py:54-66), FP64: (n,6) = min xyz, max xyz.
void orc_aabb(const orc_scene* S, float sigma_thr, real* box, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t g = 0; g < S->n; ++g) {
        real R[3][3];
        quat_to_R(S->quat + 4 * g, R);
        const real ratio = (real)S->dens[g] / (real)sigma_thr;
        real del = std::log(ratio * ratio);
        if (del < 0.0) del = 0.0;
        del = std::sqrt(del);
        for (int k = 0; k < 3; ++k) {
            // L_k row of R, scaled component-wise by delta*s
            real h2 = 0.0;
            for (int j = 0; j < 3; ++j) {
                const real e = del * (real)S->scale[3 * g + j] * R[k][j];
                h2 += e * e;
            }
            const real h = std::sqrt(h2);
            box[6 * g + k] = (real)S->pos[3 * g + k] - h;
            box[6 * g + 3 + k] = (real)S->pos[3 * g + k] + h;
        }
    }
}

// The FP64 predicate of every (ray, Gaussian) pair: out (nrays, n, 5) = r1, r2, disc, b, c and
// valid (nrays, n).  For pinning the predicate on small inputs.
void orc_intersect(const orc_scene* S, float sigma_thr, int64_t nrays, const real* o, const real* d, real* out,
                   int32_t* valid) {
    const std::vector<Geo> geo = make_geos(*S, sigma_thr, 1);
    for (int64_t r = 0; r < nrays; ++r)
        for (int64_t g = 0; g < S->n; ++g) {
            const Interval I = intersect64(geo[g], o + 3 * r, d + 3 * r);
            real* q = out + 5 * (r * S->n + g);
            q[0] = I.r1;
            q[1] = I.r2;
            q[2] = I.disc;
            q[3] = I.b;
            q[4] = I.c;
            valid[r * S->n + g] = I.valid ? 1 : 0;
        }
}

// Forward render.  rgb: per-Gaussian colour (n,3) for this view (orc_sh_forward).
// pixels may be null (ray r -> pixel r).  Outputs per ray; any output pointer may be null.
int orc_render(const orc_params* p, const orc_scene* S, const real* rgb, const float* bmin, const float* bmax,
               const orc_camera* cam, int64_t nrays, const int32_t* pixels, real* out_rgb, real* out_depth,
               real* out_opacity, int32_t* fragile, int64_t* K, int64_t* P, int64_t* maxhits, int64_t* nc,
               int64_t* nfrag, real* cmax, const orc_record* rec, int nthreads) {
    const std::vector<Geo> geo = make_geos(*S, p->sigma_thr, nthreads);
    const CpuBVH B = bvh_build(geo);
    Dec32 dec{};
    if (rec) dec = Dec32{*rec, S->dens, p->sigma_thr};
#pragma omp parallel num_threads(nthreads)
    {
        std::vector<Hit> hits;
        std::vector<int64_t> scratch;
#pragma omp for schedule(dynamic, 16)
        for (int64_t r = 0; r < nrays; ++r) {
            const RayCtx c = make_ctx(*p, *cam, pixels ? pixels[r] : r, bmin, bmax);
            const RayOut o = march_forward(*p, geo, B, rgb, c, hits, scratch, rec ? &dec : nullptr);
            if (out_rgb)
                for (int k = 0; k < 3; ++k) out_rgb[3 * r + k] = o.C[k];
            if (out_depth) out_depth[r] = o.depth;
            if (out_opacity) out_opacity[r] = 1.0 - o.T;
            if (fragile) fragile[r] = o.fragile;
            if (nfrag) nfrag[r] = o.nfrag;
            if (cmax) cmax[r] = o.cmax;
            if (K) K[r] = o.K;
            if (P) P[r] = o.P;
            if (maxhits) maxhits[r] = o.maxhits;
            if (nc) nc[r] = o.nc;
        }
    }
    return 0;
}

// Largest relative FP32 error of sigma*w, against FP64, over the samples of the rays' hit
// Gaussians whose FP64 sigma*w lies within `band` (relative) of the threshold, walking every slab
// from tenter to texit: out[0] for the reference's FP32 evaluation, out[1] for the HIP path's
// closed form, out[2] the number of such samples.
void orc_sigmaw_fp32_error(const orc_params* p, const orc_scene* S, const float* bmin, const float* bmax,
                           const orc_camera* cam, int64_t nrays, const int32_t* pixels, real band, real* out,
                           int nthreads) {
    const std::vector<Geo> geo = make_geos(*S, p->sigma_thr, nthreads);
    const CpuBVH B = bvh_build(geo);
    const real thr = p->sigma_thr;
    real e_ref = 0.0, e_closed = 0.0, count = 0.0;
#pragma omp parallel num_threads(nthreads) reduction(max : e_ref, e_closed) reduction(+ : count)
    {
        std::vector<Hit> hits;
        std::vector<int64_t> scratch;
#pragma omp for schedule(dynamic, 16)
        for (int64_t r = 0; r < nrays; ++r) {
            const RayCtx c = make_ctx(*p, *cam, pixels ? pixels[r] : r, bmin, bmax);
            if (!(c.tenter < c.texit)) continue;
            int quirk = 0;
            gather(B, geo, c.o, c.d, c.tenter, c.texit, hits, scratch, &quirk);
            SlabWalker W;
            W.hits = &hits;
            const float slab = p->dt * (float)p->spb;
            uint32_t seed = c.seed;
            for (float tb = c.tenter; tb < c.texit; tb = tb + slab) {
                const float tmin = std::fmax(c.tenter, tb);
                const float tmax = std::fmin(c.texit, tb + slab);
                W.step(tmin, tmax, &quirk);
                if (W.cur.empty()) continue;
                const SlabSamples ts(*p, tb, seed);
                seed = after_slab(*p, seed);
                for (int a : W.cur) {
                    const int64_t g = hits[a].g;
                    for (int i = 0; i < p->spb; ++i) {
                        real x[3], y[3];
                        for (int k = 0; k < 3; ++k) x[k] = (real)c.r.o[k] + (real)ts.t[i] * (real)c.r.d[k];
                        const real sw = geo[g].sigma * gauss_w(geo[g], x, y);
                        if (!(std::fabs(sw / thr - 1.0) < band)) continue;
                        count += 1.0;
                        e_ref = std::max(e_ref, std::fabs((real)sigw_ref32(*S, g, c.r, ts.t[i]) / sw - 1.0));
                        const float scl = sigw_closed32(*S, g, p->sigma_thr, c.r, ts.t[i]);
                        if (scl >= 0.0f) e_closed = std::max(e_closed, std::fabs((real)scl / sw - 1.0));
                    }
                }
            }
        }
    }
    out[0] = e_ref;
    out[1] = e_closed;
    out[2] = count;
}

// Literal per-slab restatement (no BVH, no interval reuse): O(slabs * n) per ray.
int orc_render_literal(const orc_params* p, const orc_scene* S, const real* rgb, const float* bmin,
                       const float* bmax, const orc_camera* cam, int64_t nrays, const int32_t* pixels,
                       real* out_rgb, real* out_depth, int64_t* K, int nthreads) {
    const std::vector<Geo> geo = make_geos(*S, p->sigma_thr, nthreads);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
    for (int64_t r = 0; r < nrays; ++r) {
        const RayCtx c = make_ctx(*p, *cam, pixels ? pixels[r] : r, bmin, bmax);
        const RayOut o = march_forward_literal(*p, geo, rgb, c);
        if (out_rgb)
            for (int k = 0; k < 3; ++k) out_rgb[3 * r + k] = o.C[k];
        if (out_depth) out_depth[r] = o.depth;
        if (K) K[r] = o.K;
    }
    return 0;
}

// Backward of the ray caster for one view: accumulates (+=) into the gradient arrays
// dpos (n,3), dscale (n,3), dquat (n,4), ddens (n), drgb (n,3).  ray_colors are the forward
// outputs (including T*bg), dLdC the incoming gradient, both (nrays,3).
int orc_backward(const orc_params* p, const orc_scene* S, const real* rgb, const float* bmin, const float* bmax,
                 const orc_camera* cam, int64_t nrays, const int32_t* pixels, const real* ray_colors,
                 const real* dLdC, real* dpos, real* dscale, real* dquat, real* ddens, real* drgb,
                 const orc_record* rec, int nthreads) {
    const std::vector<Geo> geo = make_geos(*S, p->sigma_thr, nthreads);
    const CpuBVH B = bvh_build(geo);
    Dec32 dec{};
    if (rec) dec = Dec32{*rec, S->dens, p->sigma_thr};
#pragma omp parallel num_threads(nthreads)
    {
        std::vector<Hit> hits;
        std::vector<int64_t> scratch;
#pragma omp for schedule(dynamic, 16)
        for (int64_t r = 0; r < nrays; ++r) {
            const real* dl = dLdC + 3 * r;
            if (dl[0] == 0.0 && dl[1] == 0.0 && dl[2] == 0.0) continue;
            const RayCtx c = make_ctx(*p, *cam, pixels ? pixels[r] : r, bmin, bmax);
            march_backward(*p, geo, B, rgb, c, ray_colors + 3 * r, dl, dpos, dscale, dquat, ddens, drgb, *S, hits,
                           scratch, rec ? &dec : nullptr);
        }
    }
    return 0;
}

// FP32 brute-force hit sets, for bit-exact comparison with the HIP traversal.
// Per-Gaussian predicate parameters are given in FP32 exactly as the HIP preprocess stores them:
// mu (n,3), Mp (n,9) rows of the bounding-ellipsoid transform, valid (n).  Rays: o, d (nrays,3),
// tenter/texit (nrays).  For every ray, slabs are enumerated from tenter with FP32 accumulation
// until texit or max_slabs (no transmittance exit).  Output: for each ray, pairs (slab, g)
// sorted by (slab, g), at most max_pairs per ray; counts[r] is the true number of pairs.
// The FP32 operation order is the one the HIP predicate uses (contraction off).
void orc_hits_fp32(int64_t n, const float* mu, const float* Mp, const int32_t* valid, int64_t nrays,
                   const float* o, const float* d, const float* tenter, const float* texit, float dt, int32_t spb,
                   int32_t max_slabs, int32_t max_pairs, int32_t* pairs, int64_t* counts, int nthreads) {
    const float slab = dt * (float)spb;
#pragma omp parallel num_threads(nthreads)
    {
        std::vector<float> r1v(n), r2v(n);
        std::vector<int32_t> cand;
        std::vector<std::pair<int32_t, int32_t>> out;
#pragma omp for schedule(dynamic, 4)
        for (int64_t r = 0; r < nrays; ++r) {
            cand.clear();
            out.clear();
            const float* oo = o + 3 * r;
            const float* dd = d + 3 * r;
            for (int64_t g = 0; g < n; ++g) {
                if (!valid[g]) continue;
                const float* m = Mp + 9 * g;
                const float* u = mu + 3 * g;
                const float Ox = oo[0] - u[0], Oy = oo[1] - u[1], Oz = oo[2] - u[2];
                float Op[3], Dp[3];
                for (int k = 0; k < 3; ++k) {
                    const float a0 = m[3 * k + 0] * Ox;
                    const float a1 = m[3 * k + 1] * Oy;
                    const float a2 = m[3 * k + 2] * Oz;
                    Op[k] = (a0 + a1) + a2;
                    const float b0 = m[3 * k + 0] * dd[0];
                    const float b1 = m[3 * k + 1] * dd[1];
                    const float b2 = m[3 * k + 2] * dd[2];
                    Dp[k] = (b0 + b1) + b2;
                }
                const float ddp = (Dp[0] * Dp[0] + Dp[1] * Dp[1]) + Dp[2] * Dp[2];
                const float l = 1.0f / std::sqrt(ddp);
                float Dh[3];
                for (int k = 0; k < 3; ++k) Dh[k] = Dp[k] * l;
                const float b = -((Op[0] * Dh[0] + Op[1] * Dh[1]) + Op[2] * Dh[2]);
                const float c = ((Op[0] * Op[0] + Op[1] * Op[1]) + Op[2] * Op[2]) - 1.0f;
                float P[3];
                for (int k = 0; k < 3; ++k) {
                    const float e = b * Dh[k];
                    P[k] = Op[k] + e;
                }
                const float pp = (P[0] * P[0] + P[1] * P[1]) + P[2] * P[2];
                const float disc = 1.0f - pp;
                if (!(disc > 1e-7f)) continue;
                const float sd = std::sqrt(disc);
                const float q = (b > 0.0f) ? b + sd : b - sd;
                r1v[g] = (c / q) * l;
                r2v[g] = q * l;
                cand.push_back((int32_t)g);
            }
            const float te = tenter[r], tx = texit[r];
            if (te < tx) {
                float tb = te;
                for (int32_t s = 0; s < max_slabs && tb < tx; ++s) {
                    const float tmin = std::fmax(te, tb);
                    const float tmax = std::fmin(tx, tb + slab);
                    if (tmax > te)
                        for (int32_t g : cand)
                            if (std::fmax(tmin, r1v[g]) <= std::fmin(tmax, r2v[g])) out.emplace_back(s, g);
                    tb = tb + slab;
                }
            }
            counts[r] = (int64_t)out.size();
            std::sort(out.begin(), out.end());
            const int64_t m = std::min<int64_t>((int64_t)out.size(), max_pairs);
            for (int64_t i = 0; i < m; ++i) {
                pairs[(r * (int64_t)max_pairs + i) * 2 + 0] = out[i].first;
                pairs[(r * (int64_t)max_pairs + i) * 2 + 1] = out[i].second;
            }
        }
    }
}

#else  // ORC_BASELINE

// One training pass of the ray caster over `nrays` rays of one view, in `real` arithmetic:
// per-Gaussian geometry and the CPU BVH, the forward, the L2 loss gradient 2 (C - target), the
// backward.  times[0..2] = seconds (omp_get_wtime) of the three phases.  rgb (n,3) per-Gaussian
// colour of the view; out_rgb (nrays,3); gradients accumulate (+=) as in the oracle's backward.
int orc_baseline_step(const orc_params* p, const orc_scene* S, const real* rgb, const float* bmin, const float* bmax,
                      const orc_camera* cam, int64_t nrays, const int32_t* pixels, const real* target,
                      real* out_rgb, real* dpos, real* dscale, real* dquat, real* ddens, real* drgb, double* times,
                      int nthreads) {
    const double t0 = omp_get_wtime();
    const std::vector<Geo> geo = make_geos(*S, p->sigma_thr, nthreads);
    const CpuBVH B = bvh_build(geo);
    const double t1 = omp_get_wtime();
    std::vector<real> dLdC(3 * nrays);
#pragma omp parallel num_threads(nthreads)
    {
        std::vector<Hit> hits;
        std::vector<int64_t> scratch;
#pragma omp for schedule(dynamic, 16)
        for (int64_t r = 0; r < nrays; ++r) {
            const RayCtx c = make_ctx(*p, *cam, pixels ? pixels[r] : r, bmin, bmax);
            const RayOut o = march_forward(*p, geo, B, rgb, c, hits, scratch);
            for (int k = 0; k < 3; ++k) {
                out_rgb[3 * r + k] = o.C[k];
                dLdC[3 * r + k] = 2.0 * (o.C[k] - target[3 * r + k]);
            }
        }
    }
    const double t2 = omp_get_wtime();
#pragma omp parallel num_threads(nthreads)
    {
        std::vector<Hit> hits;
        std::vector<int64_t> scratch;
#pragma omp for schedule(dynamic, 16)
        for (int64_t r = 0; r < nrays; ++r) {
            const RayCtx c = make_ctx(*p, *cam, pixels ? pixels[r] : r, bmin, bmax);
            march_backward(*p, geo, B, rgb, c, out_rgb + 3 * r, dLdC.data() + 3 * r, dpos, dscale, dquat, ddens,
                           drgb, *S, hits, scratch);
        }
    }
    const double t3 = omp_get_wtime();
    times[0] = t1 - t0;
    times[1] = t2 - t1;
    times[2] = t3 - t2;
    return 0;
}

int orc_baseline_real_bytes() { return (int)sizeof(real); }

#endif  // ORC_BASELINE

}  // extern "C"
