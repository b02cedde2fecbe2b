// Ray kernels of the gfx950 ray caster: forward render, recompute-based backward (optionally
// forming a per-pixel loss gradient itself), and the hit-set export used for verification.
//
// Per ray (one lane), the reference's per-slab loop (cuda_train/forward/gaussians_aabb.cu:521-612,
// backward/gaussians_aabb.cu:505-775) is kept exactly -- slab grid tb accumulated in FP32 from
// tenter, 8-sample buffers, thresholded density sums, per-slab transmittance exit -- but the hit
// sets are found differently: instead of one BVH traversal per slab, one traversal gathers every
// bounding-ellipsoid interval [r1, r2] overlapping a SEGMENT of m slabs into a per-lane list in
// LDS (capacity hit_depth); each slab's hit set is then {entries: max(tmin,r1) <= min(tmax,r2)},
// the reference's own predicate.  m adapts to the list fill (doubling across empty space,
// halving on overflow); a slab whose own hit set exceeds the list is processed by a per-slab
// traversal that accumulates on the fly (no storage).  Forward and backward step through slabs
// with the same SegState machine, so both make identical segment decisions and sum each slab in
// the same order (the traversal's DFS order, which is increasing slot: slots are numbered in the
// tree's depth-first leaf order by either builder).
//
// Sampling uses the closed form of the Gaussian along the ray: with the predicate's quantities
// (t_c = b l, P = O' + b D^, D' = M' d, h = ln|sigma/thr|),
//     0.5 |M (o + t d - mu)|^2 = h |P|^2 + h |D'|^2 (t - t_c)^2,
// so each (sample, hit) costs one FMA pair and one exp.  The backward accumulates, per
// (slab, hit), four scalars (sum w phi, sum w phi tau, sum w phi tau^2, sum w B) and expands
// them into the 14 parameter gradients of backward/gaussians_aabb.cu:700-763.
//
// Gradient scatter is wave-cooperative: every lane's list is sorted by slot, so after each
// lane has prepared one slab, the 64 lanes (an 8x8 pixel tile) walk the union of their hit sets in
// increasing slot order (a wave-wide min per step); the lanes holding the current Gaussian compute
// their contributions together, and the wave sums them and issues the 14 atomic adds as one
// 14-lane instruction.

#include <hip/hip_runtime.h>

#include <climits>

#include "rg_device.h"
#include "rg_internal.h"

namespace rg {
namespace {

constexpr int kWave = 64;
constexpr int kMaxSeg = 4096;
constexpr int kGradVals = 14;

enum CounterIdx { C_K = 0, C_P = 1, C_OVF = 2, C_FALLBACK = 3, C_GATHER = 4, C_LOSS = 6, C_REC_WORDS = 8,
                  C_REC_FULL = 9, C_DROP = 10, C_BAD = 11 };

struct RayOut {
    float C[3];
    float depth;
    float T;
};

// Per-lane interval list in LDS, [entry][lane] so lanes access consecutive words: the Gaussian's
// slot, its predicate interval and its sampling quadratic (t_c, A, Cq) from the gather.
template <int CAP>
struct IList {
    int g[CAP][kWave];
    float r1[CAP][kWave];
    float r2[CAP][kWave];
};

// Bit e set for the list entries (e < cnt) whose interval meets the slab [tmin, tmax]: the
// lanes then loop over their own hits only (in increasing entry = slot order), so a wave
// iterates max-over-lanes(hits) times instead of max-over-lanes(cnt).
//
// Every listed interval met its gathered segment, so r1 <= r2 (or one of them is NaN), and every
// slab has tmin <= tmax; for such operands overlaps(tmin, tmax, r1, r2) is !(r1 > tmax) and
// !(r2 < tmin), NaN ends included (overlaps() lets fmaxf / fminf drop a NaN).  That form needs
// no NaN-quieting max/min and no branch per entry: the loop ends at the longest list among the
// lanes executing the call (a ballot, which counts only those lanes), and each lane masks off the
// entries past its own count.
template <int CAP>
__device__ __forceinline__ unsigned int slab_mask(const IList<CAP>& L, int cnt, float tmin, float tmax, int lane) {
    unsigned int m = 0u;
#pragma unroll
    for (int e = 0; e < CAP; ++e) {
        if (__builtin_amdgcn_ballot_w64(e < cnt) == 0ULL) break;
        const bool in = (e < cnt) & !(L.r1[e][lane] > tmax) & !(L.r2[e][lane] < tmin);
        m |= in ? 1u << e : 0u;
    }
    return m;
}

// Sampling quadratic of list entry e (as computed by intersect + refine_sampling at gather time).

// Rays whose view or pixel index is out of range (skipped), and gradient records or replayed hits
// skipped by a bounds check (an internal error: every parity test asserts none), go straight to
// their device counters with an atomic on those rare paths, so they cost no registers.
struct Counters {
    unsigned int K = 0, P = 0, ovf = 0, fallback = 0, gathers = 0, rec_words = 0, rec_full = 0;
};

__device__ __forceinline__ void count_drop(unsigned long long* drops) { atomicAdd(drops, 1ull); }

// FP32 gradient add (device scope, relaxed: the result is read by a later kernel).
__device__ __forceinline__ void gadd(float* p, float v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Address of gradient value k (0..13) of a slot: 0-2 d mu, 3 d sigma, 4-6 d s, 7-10 d q in the
// slot's accumulator record (offsets 0-6, 8-11), 11-13 the view's colour gradient.
__device__ __forceinline__ float* grad_addr(const DevGrad& gr, int n, int slot, int view, int k) {
    if (k < 11) return gr.acc + (size_t)kAccStride * slot + (k < 7 ? k : k + 1);
    return reinterpret_cast<float*>(gr.drgb + (size_t)view * n + slot) + (k - 11);
}

__device__ __forceinline__ void load_rec(const float4* __restrict__ geo, int s, float4& g0, float4& g1, float4& g2,
                                         float4& g3) {
    g0 = geo[4 * s];
    g1 = geo[4 * s + 1];
    g2 = geo[4 * s + 2];
    g3 = geo[4 * s + 3];
}

// Sampling terms of a listed hit, recomputed from the Gaussian's record (a list entry keeps only
// the slot and the interval, 12 bytes, so the list holds twice the entries in the same LDS).
// The same operations as the predicate's give the same values; the roots are skipped.
__device__ __forceinline__ Isect entry_hit(const DevScene& sc, const Ray& ray, int slot, float r1, float r2,
                                           float& sigma) {
    float4 g0, g1, g2, g3;
    load_rec(sc.geo, slot, g0, g1, g2, g3);
    Isect I;
    float b, c;
    isect_terms(g0, g1, g2, g3, ray, I, b, c);  // the interval is the entry's: no roots needed
    I.r1 = r1;
    I.r2 = r2;
    refine_sampling(g0, g1, g2, g3, ray, I);
    sigma = g0.w;
    return I;
}

// The Gaussian's weight along the ray, exp(-(A + Cq tau^2)), as one native exp2 with -log2(e)
// folded into the per-hit terms; every pass (forward, backward buffers, gradients) evaluates it
// the same way, so they agree bit for bit.  Weights that small enough to flush to zero are far
// below the sample threshold.
struct WExp {
    float a, c;  // -A log2(e), -Cq log2(e)
};

__device__ __forceinline__ WExp wexp(const Isect& I) {
    constexpr float kLog2e = 1.4426950408889634f;
    return WExp{-I.A * kLog2e, -I.Cq * kLog2e};
}

__device__ __forceinline__ float weight(const WExp& k, float tau) {
    return __builtin_amdgcn_exp2f(fmaf(k.c, tau * tau, k.a));  // explicit fma: same rounding in every pass
}

// Accumulate one hit's contribution to the slab's sample buffers (computeBufferForward,
// forward/gaussians_aabb.cu:437-498).
template <int SPB, bool RNG>
__device__ __forceinline__ void accum_hit(const Isect& I, float sigma, float4 col, float tb, unsigned int seed,
                                          const DevParams& p, float* sig, float* cr, float* cg, float* cb) {
    unsigned int st = seed;
    const WExp k = wexp(I);
#pragma unroll
    for (int i = 0; i < SPB; ++i) {
        if (SPB > 8 && i >= p.spb) break;
        const float t = sample_t(tb, i, p.dt, RNG ? p.rnd_sample : 0, st);
        const float sw = sigma * weight(k, t - I.tc);
        // a sample at or below the threshold adds exact zeros (no branch per sample)
        const float m = sw > p.sigma_thr ? sw : 0.0f;
        sig[i] += m;
        cr[i] += m * col.x;
        cg[i] += m * col.y;
        cb[i] += m * col.z;
    }
}

// Opacity of one sample, 1 - exp(-sigma dt) (forward/gaussians_aabb.cu:509), with the native
// exp2 (1 ulp) and the native reciprocal for the colour below: the forward's compositing and the
// backward's coefficients use the same two, so the backward retraces the forward's transmittance.
__device__ __forceinline__ float sample_alpha(float s, float dt) {
    constexpr float kNegLog2e = -1.4426950408889634f;
    return 1.0f - __builtin_amdgcn_exp2f(s * (dt * kNegLog2e));
}

// colorBlending (forward/gaussians_aabb.cu:500-519) + depth (cuda_gui/gaussians_aabb.cu:415).
template <int SPB, bool RNG>
__device__ __forceinline__ void composite(const float* sig, const float* cr, const float* cg, const float* cb, float tb,
                                          unsigned int seed, const DevParams& p, RayOut& o) {
    unsigned int st = seed;
#pragma unroll
    for (int i = 0; i < SPB; ++i) {
        if (SPB > 8 && i >= p.spb) break;
        const float s = sig[i];
        const float alpha = sample_alpha(s, p.dt);
        // colour = sums / sigma: one reciprocal per sample (a sample without hits has zero sums)
        const float inv = s > 0.0f ? __builtin_amdgcn_rcpf(s) : 0.0f;
        const float r = cr[i] * inv, g = cg[i] * inv, b = cb[i] * inv;
        const float ta = o.T * alpha;
        o.C[0] += ta * r;
        o.C[1] += ta * g;
        o.C[2] += ta * b;
        o.depth += ta * sample_t(tb, i, p.dt, RNG ? p.rnd_sample : 0, st);
        o.T *= 1.0f - alpha;
    }
}

// Per-sample backward coefficients of one slab, from its buffers: overwrites (sig, cr) with
// (G_i, Bq_i) and advances T and the projected colour remainder Dr = dLdC . Crem past the
// slab (Crem_i: the ray's colour not yet composited before sample i; only its projection on
// dLdC enters the coefficients, so one scalar is carried instead of three).
//   A_i  = dt * dLdC . (T_i c^_i - Crem_i) = dt * (T_i E_i - Dr_i)   (dcolor_dsigma_aux, backward :706)
//   Bq_i = T_i alpha_i / sigma_i                (0 when sigma_i == 0)
//   E_i  = dLdC . c^_i = (dLdC . sums_i) / sigma_i
//   G_i  = A_i - Bq_i E_i, so that a hit of colour c weighs sample i by phi_i = G_i + Bq_i (dLdC . c)
template <int SPB>
__device__ __forceinline__ void slab_coeffs(float* sig, float* cr, float* cg, float* cb, const DevParams& p,
                                            const float* dL, float& T, float& Dr) {
#pragma unroll
    for (int i = 0; i < SPB; ++i) {
        if (SPB > 8 && i >= p.spb) break;
        const float s = sig[i];
        const float alpha = sample_alpha(s, p.dt);
        // the forward's colour is sums / sigma (composite); only its projection E on dLdC enters
        const float inv = s > 0.0f ? __builtin_amdgcn_rcpf(s) : 0.0f;
        const float E = (dL[0] * cr[i] + dL[1] * cg[i] + dL[2] * cb[i]) * inv;
        const float A = p.dt * (T * E - Dr);
        const float ta = T * alpha;
        const float Bq = ta * inv;
        Dr -= ta * E;
        T *= 1.0f - alpha;
        sig[i] = A - Bq * E;
        cr[i] = Bq;
    }
}

// Gradient moments of one (slab, hit) for one ray.  Over the slab's samples the hit contributes
// four sums (Ss = sum w phi, S1 = sum w phi tau, S2 = sum w phi tau^2, Sc = sum w Bq; tau = t - t_c,
// phi = G + Bq F with F = dLdC . colour; G, Bq from slab_coeffs).
// Every gradient of the Gaussian is linear in the ray-dependent moments
//   Gy'   = P S0 + D T1                                     (3)
//   Y'    = P P^T S0 + (P D^T + D P^T) T1 + D D^T T2         (6, symmetric)
//   Ss, and dL_c sigma Sc                                   (1 + 3)
// (S0, T1, T2 = sigma Ss, sigma S1, sigma S2; P, D the ray's ellipsoid-space offset and direction)
// with coefficients that depend on the Gaussian alone, so the wave sums these 13 moments over its
// lanes and expand the sum once per Gaussian (expand_grad) instead of once per lane.
constexpr int kMoments = 13;

template <int SPB, bool RNG>
__device__ __forceinline__ void grad_moments(const Isect& I, float sigma, float4 col, float tb, unsigned int seed,
                                             const DevParams& p, const float* Gq, const float* Bq, const float* dL,
                                             float* v) {
    const float F = dL[0] * col.x + dL[1] * col.y + dL[2] * col.z;
    float Ss = 0.0f, S1 = 0.0f, S2 = 0.0f, Sc = 0.0f;
    unsigned int st = seed;
    const WExp k = wexp(I);
#pragma unroll
    for (int i = 0; i < SPB; ++i) {
        if (SPB > 8 && i >= p.spb) break;
        const float t = sample_t(tb, i, p.dt, RNG ? p.rnd_sample : 0, st);
        const float tau = t - I.tc;
        const float w0 = weight(k, tau);
        const float w = sigma * w0 > p.sigma_thr ? w0 : 0.0f;  // below the threshold: exact zeros
        const float phi = Gq[i] + Bq[i] * F;
        const float wp = w * phi;
        Ss += wp;
        S1 += wp * tau;
        S2 += wp * (tau * tau);
        Sc += w * Bq[i];
    }
    const float S0 = sigma * Ss, T1 = sigma * S1, T2 = sigma * S2;
    const float P[3] = {I.Px, I.Py, I.Pz};
    const float D[3] = {I.Dx, I.Dy, I.Dz};
    // Gy' = U = P S0 + D T1; with W = P T1 + D T2, Y'_ab = P_a U_b + D_a W_b (symmetric)
    float W[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        v[a] = P[a] * S0 + D[a] * T1;
        W[a] = P[a] * T1 + D[a] * T2;
    }
    // Y' entries (0,0) (1,1) (2,2) (0,1) (0,2) (1,2)
    const int ya[6] = {0, 1, 2, 0, 0, 1}, yb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        const int a = ya[e], b = yb[e];
        v[3 + e] = P[a] * v[b] + D[a] * W[b];
    }
    v[9] = Ss;
    const float sc_ = sigma * Sc;
    v[10] = dL[0] * sc_;
    v[11] = dL[1] * sc_;
    v[12] = dL[2] * sc_;
}

// The 14 gradient values of one Gaussian from its summed moments m (grad_moments):
// v = (d mu[3], d sigma, d s[3], d q[4] (w,x,y,z), d rgb[3]), with the Gaussian's record rows of
// M' (g1..g3), scales s and delta (brec[0]), quaternion (brec[1]); y = d (P + tau D) in the
// unenlarged ellipsoid frame, so Gy = d Gy', Y = d^2 Y'.
__device__ __forceinline__ void expand_grad(const float* mo, float4 g1, float4 g2, float4 g3, float4 b0, float4 q,
                                            float* v) {
    const float d = b0.w, d2 = d * d;
    const float Gy[3] = {d * mo[0], d * mo[1], d * mo[2]};
    const float Y[3][3] = {{d2 * mo[3], d2 * mo[6], d2 * mo[7]},
                           {d2 * mo[6], d2 * mo[4], d2 * mo[8]},
                           {d2 * mo[7], d2 * mo[8], d2 * mo[5]}};
    const float m[3][3] = {{g1.x, g1.y, g1.z}, {g2.x, g2.y, g2.z}, {g3.x, g3.y, g3.z}};  // rows of M'
    const float s[3] = {b0.x, b0.y, b0.z};
    float dcol[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = d * (m[0][c] * Gy[0] + m[1][c] * Gy[1] + m[2][c] * Gy[2]);
    // reciprocal scales: native reciprocal (1 ulp), well inside the gradients' tolerance
    const float is[3] = {__builtin_amdgcn_rcpf(s[0]), __builtin_amdgcn_rcpf(s[1]), __builtin_amdgcn_rcpf(s[2])};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float f = -d * is[k];
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
    v[3] = mo[9];
    v[4] = Y[0][0] * is[0];
    v[5] = Y[1][1] * is[1];
    v[6] = Y[2][2] * is[2];
    v[7] = 2.0f * (qk * c0[1] - qj * c0[2] - qk * c1[0] + qi * c1[2] + qj * c2[0] - qi * c2[1]);
    v[8] = 2.0f * (qj * c0[1] + qk * c0[2] + qj * c1[0] + r * c1[2] + qk * c2[0] - r * c2[1]) -
           4.0f * qi * (c1[1] + c2[2]);
    v[9] = 2.0f * (qi * c0[1] - r * c0[2] + qi * c1[0] + qk * c1[2] + r * c2[0] + qk * c2[1]) -
           4.0f * qj * (c0[0] + c2[2]);
    v[10] = 2.0f * (r * c0[1] + qi * c0[2] - r * c1[0] + qj * c1[2] + qi * c2[0] + qj * c2[1]) -
            4.0f * qk * (c0[0] + c1[1]);
    v[11] = mo[10];
    v[12] = mo[11];
    v[13] = mo[12];
}

// Staging area of the wave-merged gradient sums in LDS.  Global atomics are issued in bursts of
// kStage records: every vector memory operation of a wave completes in order, so an atomic
// issued between two loads makes the wave wait for its round trip at the next load; a burst
// pays that wait once per kStage records instead of once per record.
constexpr int kStage = 16;
struct GradStage {
    float val[kStage][16];
    int key[kStage];  // view * n + slot, decoded at the flush
    int slot[kStage];
    int view[kStage];
};

// Issues the first cnt staged records: four records per instruction (lanes 16e..16e+13).  The
// records were written by other lanes: the workgroup is one wavefront, and the barriers order
// the LDS writes before these reads and these reads before the next records' writes.
__device__ __forceinline__ void stage_flush(GradStage& G, int cnt, const DevScene& sc, const DevGrad& gr,
                                            int lane, unsigned long long* drops) {
    __syncthreads();
    // lane e decodes record e's key and expands its summed moments into its 14 gradient values,
    // in place
    if (lane < cnt) {
        const int key = G.key[lane];
        const int view = key / sc.n;
        const int slot = key - view * sc.n;
        G.slot[lane] = slot;
        G.view[lane] = view;
        if ((unsigned)slot < (unsigned)sc.n) {
            float mo[kMoments], v[kGradVals];
#pragma unroll
            for (int k = 0; k < kMoments; ++k) mo[k] = G.val[lane][k];
            const float4 g1 = sc.geo[4 * slot + 1], g2 = sc.geo[4 * slot + 2], g3 = sc.geo[4 * slot + 3];
            expand_grad(mo, g1, g2, g3, sc.brec[2 * slot], sc.brec[2 * slot + 1], v);
#pragma unroll
            for (int k = 0; k < kGradVals; ++k) G.val[lane][k] = v[k];
        }
    }
    __syncthreads();
    const int j = lane & 15;
    for (int e0 = 0; e0 < cnt; e0 += 4) {
        const int e = e0 + (lane >> 4);
        if (e < cnt && j < kGradVals) {
            const int slot = G.slot[e], view = G.view[e];
            if ((unsigned)slot < (unsigned)sc.n && (unsigned)view < (unsigned)gr.num_views)
                gadd(grad_addr(gr, sc.n, slot, view, j), G.val[e][j]);
            else if (j == 0)
                count_drop(drops);  // never expected: counted so that tests see it
        }
    }
    __syncthreads();
}

// Wave-wide sum of the 13 gradient moments of one Gaussian (all lanes hold a contribution, zero
// for lanes without one), appended to the staging area as record nst under its key view * n + slot
// (decoded, expanded and flushed when the area is full).  The sums are formed by a
// reduce-scatter: the lanes below 32 keep values 0..7 and add the upper lanes' copies of them,
// the upper lanes values 8..15 (v_permlane32_swap); then the same over row pairs
// (v_permlane16_swap), over lane bit 3 (DPP row rotation by 8) and bit 2 (DPP half-row mirror:
// the partner differs in bits 2..0, and the quad sum below covers bits 1..0); after that lane l
// holds value bitrev4(l >> 2) summed over its quad's partners, and two DPP steps complete the sum
// over the quad.  No LDS traffic.  Called by all 64 lanes together.
__device__ __forceinline__ void grad_scatter_wave(const float* v, int key, const DevScene& sc,
                                                  const DevGrad& gr, int lane, GradStage& G, int& nst,
                                                  unsigned long long* drops) {
    float a[16], b[8], c[4], d[2];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = k < kMoments ? v[k] : 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) b[j] = pair_add32(a[j], a[j + 8]);
#pragma unroll
    for (int j = 0; j < 4; ++j) c[j] = pair_add16(b[j], b[j + 4]);
    const bool up8 = (lane & 8) != 0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const float keep = up8 ? c[j + 2] : c[j];
        const float send = up8 ? c[j] : c[j + 2];
        d[j] = keep + dpp_f<kDppRor8>(send);
    }
    const bool up4 = (lane & 4) != 0;
    float x = (up4 ? d[1] : d[0]) + dpp_f<kDppHalfMirror>(up4 ? d[0] : d[1]);
    x += dpp_f<kDppXor2>(x);
    x += dpp_f<kDppXor1>(x);
    const int idx = ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 + ((lane >> 2) & 1);
    if ((lane & 3) == 0) G.val[nst][idx] = x;
    if (lane == 0) G.key[nst] = key;
    if (++nst == kStage) {
        stage_flush(G, kStage, sc, gr, lane, drops);
        nst = 0;
    }
}

__device__ __forceinline__ int wave_lane() { return threadIdx.x & (kWave - 1); }

// LBVH traversal of the lanes executing the call: wave-uniform packet walk or per-lane walk.
template <typename F>
__device__ __forceinline__ void trav(const DevParams& p, const DevScene& sc, const Ray& ray, float ta, float tb,
                                     F&& on_leaf) {
    if (p.packet) traverse_packet(sc.nodes, sc.total_nodes, true, ray, ta, tb, on_leaf);
    else traverse(sc.nodes, sc.total_nodes, ray, ta, tb, on_leaf);
}

// Writes list entry e of the lane from Gaussian `slot` (false when the ray misses its ellipsoid).
template <int CAP>
__device__ __forceinline__ bool load_entry(const DevScene& sc, const Ray& ray, int slot, IList<CAP>& L, int e,
                                           int lane) {
    float4 g0, g1, g2, g3;
    load_rec(sc.geo, slot, g0, g1, g2, g3);
    Isect I;
    if (!intersect(g0, g1, g2, g3, ray, I)) return false;
    L.g[e][lane] = slot;
    L.r1[e][lane] = I.r1;
    L.r2[e][lane] = I.r2;
    return true;
}

// Gathers the intervals overlapping [ta, tb] of the Gaussians with slot > after into the lane's
// list; returns the count, or CAP + 1 on overflow (the list then holds the first CAP).  Entries
// come out in DFS order of the tree = increasing slot.
template <int CAP>
__device__ __forceinline__ int gather(const DevScene& sc, const DevParams& p, const Ray& ray, float ta, float tb,
                                      IList<CAP>& L, int lane, int after = -1) {
    int cnt = 0;
    trav(p, sc, ray, ta, tb, [&](int slot) -> bool {
        if (slot <= after) return true;
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

// ------------------------------------------------------------------ slab stepper -------------
//
// The reference's slab loop, with hit sets served from gathered segments.  next_slab() returns
// the next slab to process (tb, tmin, tmax) in either list mode (s.cnt entries in the lane's
// list) or direct mode (the slab's hit set overflows the list: per-slab traversal); empty
// segments are skipped without returning.  advance_slab() moves past the slab just processed.

struct SegState {
    float tb;       // start of the current slab (reference tbuffer)
    int k;          // index of that slab along the ray
    float seg_hi;   // end of the gathered segment
    float seg_lo;   // start of the gathered segment
    int m;          // segment length in slabs for the next gather
    int cnt;        // entries in the lane's list
    bool valid;     // list holds the current segment
    bool direct;    // current slab is processed by per-slab traversal
    int dpos;       // replay of a direct slab: first word of its recorded hit list
    int dcnt;       //   and its length (-1: no recorded list, traverse)
    bool dense;     // the last direct slab held more hits than the list: walk the next one too
    int knext;      // replay: slab index of the next record not yet taken (INT_MAX: none)
};

__device__ __forceinline__ void seg_init(SegState& s, const Ray& ray, const DevParams& p) {
    s.tb = ray.tenter;
    s.k = 0;
    s.seg_hi = ray.tenter;
    s.seg_lo = ray.tenter;
    s.m = p.seg_slabs;
    s.cnt = 0;
    s.valid = false;
    s.direct = false;
    s.dpos = 0;
    s.dcnt = -1;
    s.dense = false;
    s.knext = INT_MAX;
}

// ------------------------------------------------------------------ hit-list records ---------
//
// The forward records, per ray, every list it gathered and every slab it served by per-slab
// traversal; the backward over the same rays replays them instead of walking the LBVH.  Each
// record carries the index of the slab where the forward gathered it, and the backward takes
// every record that starts at or before its current slab, so it holds the forward's list for
// every slab: the same Gaussians in the same order.  A ray whose records end before its backward
// does (space ran out, or the backward goes further) continues with its own gathers from there.
// A record is a whole number of 32-byte sectors written by full 16-byte stores (a partly written
// sector costs a read-modify-write in the memory):
//   word 0: bit 31 set, bit 30 direct slab, bits 29..24 list count, bits 21..0 slab index
//   word 1: list record: end of the gathered segment (float bits);
//           direct slab: number of hits recorded after it (kRecNoList: none, the slab is walked)
//   words 2..: the list's slots, or the direct slab's hits in traversal order, zero padded
constexpr unsigned int kRecOn = 0x80000000u, kRecDirect = 0x40000000u, kRecNoList = 0xffffffffu;
constexpr int kRecMaxSlab = (1 << 22) - 1;
constexpr int kRecMaxDirect = 256;  // longest recorded hit list of a direct slab

// Layout: the rays of a group of 64 (ray index / 64) share one region; sector c of ray r sits at
// ((group * C + c) * 64 + r % 64) * 8 words, C = words per ray / 8.  The 64 rays of a wavefront
// thus write neighbouring sectors (no stride that would send every lane to one memory channel).
struct Rec {
    unsigned int* w = nullptr;  // sector 0 of this ray; nullptr: off
    int pos = 0;                // next word (a multiple of 8)
    int end = 0;                // capacity (recording) or words recorded (replay)
    bool full = false;          // recording ran out of space
};

__device__ __forceinline__ unsigned int* rec_base(unsigned int* rec, int64_t ray, int words) {
    return rec + (((ray >> 6) * (int64_t)(words >> 3)) * 64 + (ray & 63)) * 8;
}

__device__ __forceinline__ unsigned int* rec_at(const Rec& rc, int w) { return rc.w + ((w >> 3) << 9) + (w & 7); }

__device__ __forceinline__ int rec_len(int cnt) { return (2 + cnt + 7) & ~7; }

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int CAP>
__device__ __forceinline__ void record(Rec& rc, int k, bool direct, float hi, const IList<CAP>& L, int cnt,
                                       int lane) {
    if (!rc.w || rc.full) return;
    const int len = rec_len(cnt);
    if (k > kRecMaxSlab || rc.pos + len > rc.end) {
        rc.full = true;
        return;
    }
    const unsigned int h0 = kRecOn | (direct ? kRecDirect : 0u) | ((unsigned int)cnt << 24) | (unsigned int)k;
    for (int c = 0; c < len; c += 4) {
        unsigned int v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int e = c + j - 2;
            v[j] = c + j == 0 ? h0 : c + j == 1 ? __float_as_uint(hi) : (e < cnt ? (unsigned int)L.g[e][lane] : 0u);
        }
        // streaming stores: the records are read once, by the next kernel, and must not evict
        // the tree and Gaussian records the traversal keeps in L2
        const u32x4 q = {v[0], v[1], v[2], v[3]};
        __builtin_nontemporal_store(q, reinterpret_cast<u32x4*>(rec_at(rc, rc.pos + c)));
    }
    rc.pos += len;
}

// Direct slab (forward): the record is opened before the slab's traversal, its hits appended as
// the traversal reaches them, and closed with their count (a list that does not fit is dropped:
// the backward then walks that slab).
__device__ __forceinline__ int rec_direct_open(Rec& rc) {
    if (!rc.w || rc.full) return -1;
    if (rc.pos + 8 > rc.end) {
        rc.full = true;
        return -1;
    }
    return rc.pos;
}

__device__ __forceinline__ void rec_direct_push(const Rec& rc, int start, int n, int slot) {
    const int w = start + 2 + n;
    if (start >= 0 && n < kRecMaxDirect && w < rc.end) *rec_at(rc, w) = (unsigned int)slot;
}

__device__ __forceinline__ void rec_direct_close(Rec& rc, int start, int n, int k) {
    if (start < 0) return;
    const bool listed = n <= kRecMaxDirect && start + rec_len(n) <= rc.end && k <= kRecMaxSlab;
    *rec_at(rc, start) = kRecOn | kRecDirect | (unsigned int)min(k, kRecMaxSlab);
    *rec_at(rc, start + 1) = listed ? (unsigned int)n : kRecNoList;
    rc.pos = start + (listed ? rec_len(n) : 8);
}

// Replay: takes every record that starts at or before the current slab.  Anything that does not
// fit the lane's list or does not reproduce (never expected) ends the replay.
template <int CAP>
__device__ __forceinline__ void take_records(SegState& s, const DevScene& sc, const Ray& ray, IList<CAP>& L,
                                             int lane, Rec& rc) {
    s.knext = INT_MAX;
    while (rc.pos + 8 <= rc.end) {
        const u32x4 h = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(rec_at(rc, rc.pos)));
        if (!(h.x & kRecOn)) {
            rc.w = nullptr;
            s.valid = false;
            return;
        }
        if ((int)(h.x & (unsigned int)kRecMaxSlab) > s.k) {
            s.knext = (int)(h.x & (unsigned int)kRecMaxSlab);
            return;
        }
        const int cnt = (int)((h.x >> 24) & 63u);
        const int len = rec_len(cnt);
        if (h.x & kRecDirect) {
            s.direct = true;
            s.cnt = 0;
            s.valid = true;
            const unsigned int n = h.y;
            if (n != kRecNoList && n <= (unsigned int)kRecMaxDirect && rc.pos + rec_len((int)n) <= rc.end) {
                s.dpos = rc.pos + 2;
                s.dcnt = (int)n;
                rc.pos += rec_len((int)n);
            } else {
                s.dcnt = -1;
                rc.pos += 8;
            }
            continue;
        }
        bool ok = cnt <= CAP && rc.pos + len <= rc.end;
        for (int e = 0; ok && e < cnt; ++e) {
            const int slot = (int)__builtin_nontemporal_load(rec_at(rc, rc.pos + 2 + e));
            ok = slot >= 0 && slot < sc.n && load_entry<CAP>(sc, ray, slot, L, e, lane);
        }
        if (!ok) {
            rc.w = nullptr;
            s.valid = false;
            return;
        }
        s.direct = false;
        s.cnt = cnt;
        s.seg_hi = __uint_as_float(h.y);
        s.valid = true;
        rc.pos += len;
    }
}

template <int CAP, int RM>
__device__ __forceinline__ bool next_slab(SegState& s, const DevScene& sc, const DevParams& p, const Ray& ray,
                                          IList<CAP>& L, int lane, Counters& ctr, float& tmin, float& tmax, Rec& rc) {
    const float slab = p.dt * (float)p.spb;
    while (true) {
        if (!(s.tb < ray.texit)) return false;
        if constexpr (RM == 2) {
            if (rc.w) {
                take_records<CAP>(s, sc, ray, L, lane, rc);
                if (!s.valid) rc.w = nullptr;  // no record holds this slab: own gathers from here on
            }
        }
        if (!s.valid && s.dense) {
            // inside a dense stretch (the last slab needed its own walk): a gather would most
            // likely overflow too, so the slab is walked without trying it
            ctr.fallback++;
            s.direct = true;
            s.cnt = 0;
            s.dcnt = -1;
            s.valid = true;
        }
        if (!s.valid) {
            s.seg_hi = fminf(ray.texit, s.tb + (float)s.m * slab);
            s.seg_lo = s.tb;
            const int cnt = gather<CAP>(sc, p, ray, s.tb, s.seg_hi, L, lane);
            ctr.gathers++;
            if (cnt > CAP) {
                // shorter segments down to two slabs; a single slab is walked rather than
                // gathered (one traversal either way, and the walk needs no list)
                if (s.m > 2) {
                    s.m >>= 1;
                    continue;
                }
                ctr.fallback++;
                s.direct = true;
                s.cnt = 0;
                s.dcnt = -1;  // recorded (with its hits) by the forward's traversal of the slab
            } else {
                s.direct = false;
                s.cnt = cnt;
                if constexpr (RM == 1) record<CAP>(rc, s.k, false, s.seg_hi, L, cnt, lane);
            }
            s.valid = true;
        }
        tmax = fminf(ray.texit, add_exact(s.tb, slab));
        if (!s.direct) {
            if (tmax > s.seg_hi) {  // segment done: adapt its length to the list fill, gather the next
                if (s.cnt == 0 || 4 * s.cnt < CAP) s.m = min(s.m * 2, kMaxSeg);
                else if (4 * s.cnt > 3 * CAP) s.m = max(s.m >> 1, 1);
                s.valid = false;
                continue;
            }
            if (s.cnt == 0) {  // empty segment: its slabs have no hits
                const float nt = add_exact(s.tb, slab);
                if (!(nt > s.tb)) return false;
                s.tb = nt;
                ++s.k;
                continue;
            }
        }
        tmin = fmaxf(ray.tenter, s.tb);
        return true;
    }
}

// Whether the lane's next slab needs a new gather because its list is spent (a lane crossing an
// empty segment does not count: it only steps to the segment end).
__device__ __forceinline__ bool needs_gather(const SegState& s, const Ray& ray, const DevParams& p) {
    const float slab = p.dt * (float)p.spb;
    return !s.valid || (!s.direct && s.cnt > 0 && fminf(ray.texit, add_exact(s.tb, slab)) > s.seg_hi);
}

// Wave-synchronous gathers: when one lane executing the call must gather, all of them gather
// anew from their current slab, so that one packet traversal serves every lane.  The gathered
// lists hold the same hit sets, so the result does not change.  Lanes replaying recorded lists
// (own = false) neither gather nor ask others to.
__device__ __forceinline__ void sync_gathers(SegState& s, const Ray& ray, const DevParams& p, bool own = true) {
    const bool need = own && needs_gather(s, ray, p);
    // a lane that has used up less than half of its segment keeps its list for a later round
    const bool fresh = s.valid && !s.direct && s.cnt > 0 && 2.0f * (s.tb - s.seg_lo) < s.seg_hi - s.seg_lo;
    if (__ballot(need) != 0ULL && own && (need || !fresh)) s.valid = false;
}

__device__ __forceinline__ bool advance_slab(SegState& s, const DevParams& p) {
    const float nt = add_exact(s.tb, p.dt * (float)p.spb);
    if (s.direct) {
        s.direct = false;
        s.valid = false;
    }
    if (!(nt > s.tb)) return false;  // slab below half an ulp of t: no progress possible
    s.tb = nt;
    ++s.k;
    return true;
}

struct RayRef {
    int view;
    int pixel;
    int64_t out;  // output row
    bool active;
};

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
    unsigned int v[7] = {c.K, c.P, c.ovf, c.fallback, c.gathers, c.rec_words, c.rec_full};
    const int idx[7] = {C_K, C_P, C_OVF, C_FALLBACK, C_GATHER, C_REC_WORDS, C_REC_FULL};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        unsigned int x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        v[k] = x;
    }
    if (wave_lane() == 0) {
#pragma unroll
        for (int k = 0; k < 7; ++k)
            if (v[k]) atomicAdd(&st.c[idx[k]], (unsigned long long)v[k]);
    }
}

// ------------------------------------------------------------------ per-ray bodies -----------

// Starts a ray: clears its output and sets up its slab stepper; false when the ray has no slab
// to process (it misses the scene box, or its first slab is below half an ulp of t).
// Wavefront-wide step over slabs without hits, after a round in which no lane had one: the lanes
// stepping through a list (step) advance together over the slabs empty for them, up to the
// wavefront's earliest possible next hit -- the smallest start (or the segment end) among the list
// entries that have not ended before each lane's slab; lanes that cannot step bound it by their
// slab (bound_tb).  Every slab passed ends before that point, so none overlaps an entry: the slabs
// with hits and their order are unchanged, and the lanes stay in step in depth.  A replaying lane
// stops at the slab of its next record (knext), which it takes there.  Called by all 64 lanes.
template <int CAP>
__device__ __forceinline__ void jump_empty(SegState& s, const Ray& ray, const DevParams& p, const IList<CAP>& L,
                                           int lane, bool step, bool bound_tb, int knext) {
    float bound = bound_tb ? s.tb : INFINITY;
    if (step) {
        bound = s.seg_hi;
        for (int e = 0; e < s.cnt; ++e)
            if (!(L.r2[e][lane] < s.tb)) bound = fminf(bound, fmaxf(L.r1[e][lane], s.tb));
    }
    const float tstar = wave_min64f(bound);
    if (step && tstar < INFINITY) {
        const float slab = p.dt * (float)p.spb;
        while (s.k < knext) {
            const float nt = add_exact(s.tb, slab);
            if (!(fminf(ray.texit, nt) < tstar) || !(nt > s.tb)) break;
            s.tb = nt;
            ++s.k;
        }
    }
}

__device__ __forceinline__ bool forward_begin(const DevParams& p, const Ray& ray, RayOut& o, SegState& s) {
    o.C[0] = o.C[1] = o.C[2] = 0.0f;
    o.depth = 0.0f;
    o.T = 1.0f;
    if (!(ray.tenter < ray.texit)) return false;
    if (!(add_exact(ray.tenter, p.dt * (float)p.spb) > ray.tenter)) return false;
    seg_init(s, ray, p);
    return true;
}

// One slab of the forward (the reference's raygen loop body, forward/gaussians_aabb.cu:546-606):
// the next slab with hits, its sample buffers, compositing.  Returns false when the ray is done.
template <int SPB, int CAP, bool RNG, int RM>
__device__ __forceinline__ bool forward_step(const DevScene& sc, const DevParams& p, const Ray& ray,
                                             unsigned int& seed, const float4* __restrict__ rgbv, IList<CAP>& L,
                                             int lane, Counters& ctr, RayOut& o, Rec& rc, SegState& s,
                                             bool* hit = nullptr) {
    if (!(o.T > p.t_eps)) return false;
    const float first_tb = ray.tenter;
    sync_gathers(s, ray, p);
    float tmin, tmax;
    if (!next_slab<CAP, RM>(s, sc, p, ray, L, lane, ctr, tmin, tmax, rc)) return false;
    const float tb = s.tb;
    float sig[SPB], cr[SPB], cg[SPB], cb[SPB];
#pragma unroll
    for (int i = 0; i < SPB; ++i) sig[i] = cr[i] = cg[i] = cb[i] = 0.0f;
    int nh = 0;
    auto eval_hit = [&](int slot, const Isect& I, float sigma) {
        accum_hit<SPB, RNG>(I, sigma, rgbv[slot], tb, seed, p, sig, cr, cg, cb);
        if (tb == first_tb || I.r1 > tmin) ctr.P++;
        ++nh;
    };
    if (s.direct) {
        int rstart = -1;
        if constexpr (RM == 1) rstart = rec_direct_open(rc);
        trav(p, sc, ray, tmin, tmax, [&](int slot) -> bool {
            float4 g0, g1, g2, g3;
            load_rec(sc.geo, slot, g0, g1, g2, g3);
            Isect I;
            if (!intersect(g0, g1, g2, g3, ray, I)) return true;
            if (!overlaps(tmin, tmax, I.r1, I.r2)) return true;
            refine_sampling(g0, g1, g2, g3, ray, I);
            if constexpr (RM == 1) rec_direct_push(rc, rstart, nh, slot);
            eval_hit(slot, I, g0.w);
            return true;
        });
        if constexpr (RM == 1) rec_direct_close(rc, rstart, nh, s.k);
        s.dense = nh > CAP;
    } else {
        unsigned int m = slab_mask<CAP>(L, s.cnt, tmin, tmax, lane);
        while (m != 0u) {
            const int e = __builtin_ctz(m);
            m &= m - 1u;
            const int slot = L.g[e][lane];
            float sigma;
            const Isect I = entry_hit(sc, ray, slot, L.r1[e][lane], L.r2[e][lane], sigma);
            eval_hit(slot, I, sigma);
        }
    }
    ctr.K += nh;
    if (hit) *hit = nh > 0;
    if (p.max_hits > 0 && nh > p.max_hits) ctr.ovf++;
    if (nh > 0) {
        composite<SPB, RNG>(sig, cr, cg, cb, tb, seed, p, o);
        if (RNG && p.rnd_sample == 1) seed = skip_draws(seed, p.spb);  // the slab's draws are spent
    }
    return advance_slab(s, p);
}

template <int SPB, int CAP, bool RNG, int RM>
__device__ void forward_ray(const DevScene& sc, const DevParams& p, const Ray& ray, unsigned int seed,
                            const float4* __restrict__ rgbv, IList<CAP>& L, int lane, Counters& ctr, RayOut& o,
                            Rec& rc) {
    SegState s;
    bool alive = forward_begin(p, ray, o, s);
    bool jump = false;
    while (__ballot(alive) != 0ULL) {
        if (jump) jump_empty<CAP>(s, ray, p, L, lane, alive && s.valid && !s.direct, alive, INT_MAX);
        bool hit = false;
        if (alive) alive = forward_step<SPB, CAP, RNG, RM>(sc, p, ray, seed, rgbv, L, lane, ctr, o, rc, s, &hit);
        jump = p.fwd_jump && __ballot(hit) == 0ULL;
    }
}

// Scatter of the slabs whose hit set overflowed the list (lanes with dsc), called by all 64
// lanes together: a packet traversal over the lanes' slab intervals; at each leaf the lanes whose
// slab holds the Gaussian compute its gradient contribution and the wave sums them into one
// staged record per view (as in the list scatter), instead of 14 atomics per lane and hit.
template <int SPB, bool RNG>
__device__ void direct_scatter_wave(const DevScene& sc, const DevParams& p, bool dsc, const Ray& ray,
                                    const float4* __restrict__ rgbv, float cur_tb, unsigned int cur_seed,
                                    float cur_tmin, float cur_tmax,
                                    const float* Gq, const float* Bq, const float* dL, GradStage& G,
                                    int& nst, int lane, const DevGrad& gr, int view, unsigned long long* drops) {
    const float ix = 1.0f / ray.dx, iy = 1.0f / ray.dy, iz = 1.0f / ray.dz;
    const int total = sc.total_nodes;
    bool live = dsc;
    int idx = 0;
    while (idx < total && __ballot(live) != 0ULL) {
        idx = __builtin_amdgcn_readfirstlane(idx);
        const float4 lo = ld_node(sc.nodes, 2 * idx);
        const float4 hi = ld_node(sc.nodes, 2 * idx + 1);
        const int miss = __float_as_int(lo.w);
        const int leaf = __float_as_int(hi.w);
        if (miss <= idx) break;  // malformed link (uniform): never loop
        const bool h = live && box_hit(lo, hi, ray, ix, iy, iz, cur_tmin, cur_tmax);
        if (__ballot(h) == 0ULL) {
            idx = miss;
            continue;
        }
        if (leaf < 0) {
            idx = idx + 1;
            continue;
        }
        if (leaf >= sc.n) break;  // malformed leaf index (uniform)
        float v[kMoments];
#pragma unroll
        for (int k = 0; k < kMoments; ++k) v[k] = 0.0f;
        bool c = false;
        if (h) {
            float4 g0, g1, g2, g3;
            load_rec(sc.geo, leaf, g0, g1, g2, g3);
            Isect I;
            if (intersect(g0, g1, g2, g3, ray, I) && overlaps(cur_tmin, cur_tmax, I.r1, I.r2)) {
                refine_sampling(g0, g1, g2, g3, ray, I);
                grad_moments<SPB, RNG>(I, g0.w, rgbv[leaf], cur_tb, cur_seed, p, Gq, Bq, dL, v);
                c = true;
            }
        }
        unsigned long long rem = __ballot(c);
        while (rem != 0ULL) {  // one record per view among the contributing lanes (normally one)
            const int vw = __builtin_amdgcn_readlane(view, __ffsll((long long)rem) - 1);
            const bool mine = c && view == vw;
            float w[kMoments];
#pragma unroll
            for (int k = 0; k < kMoments; ++k) w[k] = mine ? v[k] : 0.0f;
            grad_scatter_wave(w, vw * sc.n + leaf, sc, gr, lane, G, nst, drops);
            rem &= ~__ballot(mine);
        }
        idx = miss;
    }
}

// Backward of one ray per lane, called by all 64 lanes of the wave together (inactive lanes pass
// active = false).  Each outer iteration every live lane prepares its next slab that has hits
// (recomputed buffers, per-sample coefficients, T and the colour remainder advanced); then the
// wave walks the union of the prepared slabs' hits in increasing (view, slot) order and scatters
// one summed contribution per Gaussian.  Slabs served by per-slab traversal scatter per lane.
template <int SPB, int CAP, bool RNG, int RM>
__device__ void backward_wave(const DevScene& sc, const DevParams& p, bool active, const Ray& ray, unsigned int seed,
                              const float4* __restrict__ rgbv, const float* Cfwd, const float* dL, IList<CAP>& L,
                              GradStage& G, int lane, const DevGrad& gr, int view, Counters& ctr, Rec& rc,
                              unsigned long long* drops) {
    int nst = 0;  // records in the staging area
    bool alive = active && (ray.tenter < ray.texit) && (add_exact(ray.tenter, p.dt * (float)p.spb) > ray.tenter);
    float T = 1.0f;
    float Dr = dL[0] * Cfwd[0] + dL[1] * Cfwd[1] + dL[2] * Cfwd[2];  // dLdC . (colour not yet composited)
    SegState s;
    seg_init(s, ray, p);
    const int key_base = view * sc.n;
    // lanes within p.bwd_window slabs of the laggard step together (half a slab at 1M Gaussians:
    // 4 % faster than one, the same as a quarter; one slab 12 % faster on a 300k-Gaussian batch)
    const float window = p.bwd_window * p.dt * (float)p.spb;
    bool pend = false;  // next_slab() has produced a slab not yet processed
    bool jump = false;  // the last round found no hit in any lane's slab (jump_empty)
    float tmin = 0.0f, tmax = 0.0f;
    while (__ballot(alive) != 0ULL) {
        if (jump)
            jump_empty<CAP>(s, ray, p, L, lane, alive && !pend && s.valid && !s.direct, alive, rc.w ? s.knext : INT_MAX);
        // prepared slab: list mode (has: cur_cnt list entries) or per-slab traversal (dsc)
        bool has = false, dsc = false, prepped = false;
        float cur_tb = 0.0f, cur_tmin = 0.0f, cur_tmax = 0.0f;
        unsigned int cur_seed = 0u;
        int cur_cnt = 0;
        bool cur_dlist = false;  // the entries are a direct slab's recorded hits (all in the slab)
        int cur_dpos = 0;
        unsigned int cur_m = 0u;  // list mode: the entries in the slab (slab_mask), not yet scattered
        // the prepared slab's sample buffers; after slab_coeffs sig and cr hold the per-sample
        // backward coefficients G_i, Bq_i read by the scatter below
        float sig[SPB], cr[SPB], cg[SPB], cb[SPB];
        if (alive && !pend) {
            if (!(T > p.t_eps)) {
                alive = false;
            } else {
                sync_gathers(s, ray, p, rc.w == nullptr);
                if (!next_slab<CAP, RM>(s, sc, p, ray, L, lane, ctr, tmin, tmax, rc)) alive = false;
                else pend = true;
            }
        }
        // depth synchronisation: the wave's merged scatter only pays off when the lanes' current
        // slabs are at similar depth (similar Gaussians), so lanes ahead of the laggard wait
        const float tsync = wave_min64f(alive && pend ? s.tb : INFINITY);
        if (alive && pend && s.tb <= tsync + window) {
            pend = false;
            prepped = true;
            {
                const float tb = s.tb;
#pragma unroll
                for (int i = 0; i < SPB; ++i) sig[i] = cr[i] = cg[i] = cb[i] = 0.0f;
                int nh = 0;
                auto direct_hit = [&](int slot) -> bool {
                    float4 g0, g1, g2, g3;
                    load_rec(sc.geo, slot, g0, g1, g2, g3);
                    Isect I;
                    if (!intersect(g0, g1, g2, g3, ray, I)) return true;
                    if (!overlaps(tmin, tmax, I.r1, I.r2)) return true;
                    refine_sampling(g0, g1, g2, g3, ray, I);
                    accum_hit<SPB, RNG>(I, g0.w, rgbv[slot], tb, seed, p, sig, cr, cg, cb);
                    ++nh;
                    return true;
                };
                if (s.direct && RM == 2 && s.dcnt >= 0) {
                    // the forward's hits of this slab: they passed the exact test in the forward over
                    // this same ray (the token, the handle's generation and the row's (view, pixel)
                    // are checked), so only their sampling terms are computed, as for list entries
                    for (int i = 0; i < s.dcnt; ++i) {
                        const int slot = (int)*rec_at(rc, s.dpos + i);
                        if ((unsigned)slot < (unsigned)sc.n) {
                            float sigma;
                            const Isect I = entry_hit(sc, ray, slot, 0.0f, 0.0f, sigma);
                            accum_hit<SPB, RNG>(I, sigma, rgbv[slot], tb, seed, p, sig, cr, cg, cb);
                            ++nh;
                        } else {
                            count_drop(drops);
                        }
                    }
                    s.dense = nh > CAP;
                } else if (s.direct) {
                    trav(p, sc, ray, tmin, tmax, direct_hit);
                    s.dense = nh > CAP;
                } else {
                    unsigned int m = slab_mask<CAP>(L, s.cnt, tmin, tmax, lane);
                    cur_m = m;
                    while (m != 0u) {
                        const int e = __builtin_ctz(m);
                        m &= m - 1u;
                        const int slot = L.g[e][lane];
                        float sigma;
                        const Isect I = entry_hit(sc, ray, slot, L.r1[e][lane], L.r2[e][lane], sigma);
                        accum_hit<SPB, RNG>(I, sigma, rgbv[slot], tb, seed, p, sig, cr, cg, cb);
                        ++nh;
                    }
                }
                if (nh > 0) {
                    slab_coeffs<SPB>(sig, cr, cg, cb, p, dL, T, Dr);
                    cur_tb = tb;
                    cur_seed = seed;
                    cur_tmin = tmin;
                    cur_tmax = tmax;
                    if (RNG && p.rnd_sample == 1) seed = skip_draws(seed, p.spb);  // the slab's draws are spent
                    if (s.direct && RM == 2 && s.dcnt >= 0) {
                        has = true;
                        cur_dlist = true;
                        cur_dpos = s.dpos;
                        cur_cnt = s.dcnt;
                    } else if (s.direct) {
                        dsc = true;
                    } else {
                        has = true;
                        cur_cnt = s.cnt;
                    }
                }
                if (!advance_slab(s, p)) alive = false;
            }
        }
        jump = __ballot(prepped) != 0ULL && __ballot(has || dsc) == 0ULL;
        // converged: slabs whose hit set overflowed the list scatter from a packet traversal
        if (__ballot(dsc) != 0ULL)
            direct_scatter_wave<SPB, RNG>(sc, p, dsc, ray, rgbv, cur_tb, cur_seed, cur_tmin, cur_tmax, sig, cr, dL, G,
                                     nst, lane, gr, view, drops);
        // wave-merged scatter over the prepared slabs' list hits.  A lane's key (view * n + slot of
        // its next hit; INT_MAX when it has none) changes only when its hit is scattered, and the
        // following hit's slot is loaded ahead of that hit's arithmetic.
        int e = 0;
        auto dlist_key = [&]() -> int {
            while (e < cur_cnt) {
                const int sl = (int)*rec_at(rc, cur_dpos + e);
                if ((unsigned)sl < (unsigned)sc.n) return key_base + sl;
                ++e;  // an out-of-range slot was counted when the slab was prepared
            }
            return INT_MAX;
        };
        int my = INT_MAX;
        if (has && cur_dlist) my = dlist_key();
        else if (has && cur_m != 0u) my = key_base + L.g[__builtin_ctz(cur_m)][lane];  // slot order
        while (true) {
            const int mn = wave_min64(my);
            if (mn == INT_MAX) break;
            const bool act = (my == mn);
            float v[kMoments];
#pragma unroll
            for (int k = 0; k < kMoments; ++k) v[k] = 0.0f;
            if (act) {
                const int slot = mn - key_base;
                int nxt = -1;
                if (cur_dlist) {
                    if (e + 1 < cur_cnt) nxt = (int)*rec_at(rc, cur_dpos + e + 1);
                } else {
                    cur_m &= cur_m - 1u;
                    if (cur_m != 0u) nxt = L.g[__builtin_ctz(cur_m)][lane];
                }
                float4 g0, g1, g2, g3;
                load_rec(sc.geo, slot, g0, g1, g2, g3);
                Isect I;
                // a list entry or a recorded direct hit passed the exact test when it was gathered
                // or recorded, so only its sampling terms are computed (the same values, without the
                // roots)
                {
                    float b, c;
                    isect_terms(g0, g1, g2, g3, ray, I, b, c);
                    refine_sampling(g0, g1, g2, g3, ray, I);
                    grad_moments<SPB, RNG>(I, g0.w, rgbv[slot], cur_tb, cur_seed, p, sig, cr, dL, v);
                }
                if (cur_dlist) {
                    ++e;
                    my = (unsigned)nxt < (unsigned)sc.n ? key_base + nxt : dlist_key();
                } else {
                    my = cur_m != 0u ? key_base + nxt : INT_MAX;
                }
            }
            grad_scatter_wave(v, mn, sc, gr, lane, G, nst, drops);
        }
    }
    stage_flush(G, nst, sc, gr, lane, drops);
}

// ------------------------------------------------------------------ kernels ------------------

enum RayStatus { RAY_NONE = 0, RAY_OK = 1, RAY_BAD = 2 };

// Work item -> (view, pixel, output row).  RAY_NONE: past the end of the job, or a padding item
// of the implicit single-view mode (whose items are enumerated in 8x8 tiles, num_items being the
// padded tile count * 64).  RAY_BAD: a listed view or pixel index outside the views set by
// rg_set_views; such a ray is skipped (its output row reads NaN) and counted.
__device__ __forceinline__ int resolve(const RayJob& job, int64_t w, RayRef& q) {
    q.view = job.view;
    q.pixel = 0;
    q.out = w;
    q.active = false;
    if (job.ray_view || job.pixels) {
        if (w >= job.num_rays) return RAY_NONE;
        if (job.ray_view) {
            q.view = job.ray_view[w];
            q.pixel = job.ray_pixel[w];
        } else {
            q.pixel = job.pixels[w];
        }
        if ((unsigned)q.view >= (unsigned)job.num_views) return RAY_BAD;
        const rg_camera& cam = job.cams[q.view];
        if (q.pixel < 0 || (int64_t)q.pixel >= (int64_t)cam.width * cam.height) return RAY_BAD;
        q.active = true;
        return RAY_OK;
    }
    if (w >= job.num_rays) return RAY_NONE;
    const rg_camera& cam = job.cams[job.view];
    int pixel;
    if (!tile_pixel(w, cam.width, cam.height, pixel)) return RAY_NONE;
    q.pixel = pixel;
    q.out = pixel;
    q.active = true;
    return RAY_OK;
}

// Writes a finished ray's colour (with the background behind its remaining transmittance),
// depth and opacity, and closes its hit-list records.
template <int RM>
__device__ __forceinline__ void forward_finish(const DevParams& p, const RayRef& q, const RayOut& o, const Rec& rc,
                                               Counters& ctr, float* __restrict__ out_rgb, float* __restrict__ out_depth,
                                               float* __restrict__ out_opac) {
    if constexpr (RM == 1) {
        p.rec_len[q.out] = rc.pos | (rc.full ? (int)0x80000000u : 0);
        p.rec_id[q.out] = make_int2(q.view, q.pixel);  // the backward replays only the same ray
        ctr.rec_words += rc.pos;
        ctr.rec_full += rc.full ? 1u : 0u;
    }
    out_rgb[3 * q.out + 0] = o.C[0] + o.T * p.bg;
    out_rgb[3 * q.out + 1] = o.C[1] + o.T * p.bg;
    out_rgb[3 * q.out + 2] = o.C[2] + o.T * p.bg;
    if (out_depth) out_depth[q.out] = o.depth;
    if (out_opac) out_opac[q.out] = 1.0f - o.T;
}

// A ray whose view or pixel index is out of range: NaN outputs, no records, counted.
template <int RM>
__device__ __forceinline__ void forward_bad(const DevParams& p, const RayRef& q, DevStats st,
                                            float* __restrict__ out_rgb, float* __restrict__ out_depth,
                                            float* __restrict__ out_opac) {
    atomicAdd(&st.c[C_BAD], 1ull);
    const float nan = __builtin_nanf("");
    out_rgb[3 * q.out + 0] = nan;
    out_rgb[3 * q.out + 1] = nan;
    out_rgb[3 * q.out + 2] = nan;
    if (out_depth) out_depth[q.out] = nan;
    if (out_opac) out_opac[q.out] = nan;
    if constexpr (RM == 1) {
        p.rec_len[q.out] = 0;
        p.rec_id[q.out] = make_int2(-1, -1);
    }
}

// XCD-aware block order: consecutive work blocks land on the same XCD's L2.
__device__ __forceinline__ int64_t block_item(int64_t b, int64_t nblocks) {
    const int64_t per = (nblocks + 7) >> 3;
    const int64_t xcd = b & 7, k = b >> 3;
    return xcd * per + k;
}

template <int SPB, int CAP, bool RNG, int RM>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(SPB == 8 ? 4 : 3))) k_forward(DevScene sc, DevParams p, RayJob job, int64_t num_items,
                                                   float* __restrict__ out_rgb, float* __restrict__ out_depth,
                                                   float* __restrict__ out_opac, DevStats st) {
    __shared__ IList<CAP> L;
    const int lane = threadIdx.x;
    const int64_t nb = gridDim.x;
    const int64_t item = block_item(blockIdx.x, nb) * kWave + lane;
    Counters ctr;
    RayRef q;
    const int rs = item < num_items ? resolve(job, item, q) : (int)RAY_NONE;
    if (rs == RAY_BAD) {
        forward_bad<RM>(p, q, st, out_rgb, out_depth, out_opac);
    } else if (rs == RAY_OK) {
        const rg_camera cam = job.cams[q.view];
        unsigned int seed = RNG ? tea4(p.iteration, (unsigned int)q.pixel) : 0u;
        const Ray ray = make_ray(cam, q.pixel, sc.box, p.t_near, p.t_far, RNG ? p.jitter : 0, &seed);
        RayOut o;
        Rec rc;
        if constexpr (RM == 1) {
            rc.w = rec_base(p.rec, q.out, p.rec_words);
            rc.end = p.rec_words;
        }
        forward_ray<SPB, CAP, RNG, RM>(sc, p, ray, seed, sc.rgb + (size_t)q.view * sc.n, L, lane, ctr, o, rc);
        forward_finish<RM>(p, q, o, rc, ctr, out_rgb, out_depth, out_opac);
    }
    flush_counters(ctr, st);
}

// One lane's ray of a backward pass.
struct BwdRay {
    bool active = false;
    Ray ray = {};
    unsigned int seed = 0u;
    float dL[3] = {0.0f, 0.0f, 0.0f};
    float Cf[3] = {0.0f, 0.0f, 0.0f};
    int view = 0;
    Rec rc;
};

// Sets up work item `item` of a backward: its rendered colour, dL/dC (given, or formed from the
// per-pixel loss against target, whose value is added to loss), its hit-list records when they
// were written for this same ray, and the ray itself.  A ray with zero dL/dC stays inactive.
template <bool RNG, int RM>
__device__ __forceinline__ void backward_setup(const DevScene& sc, const DevParams& p, const RayJob& job, int64_t item,
                                               const float* __restrict__ ray_colors, const float* __restrict__ dLdC,
                                               const float* __restrict__ target, int loss_kind, DevStats st,
                                               BwdRay& b, float& loss) {
    RayRef q;
    const int rs = resolve(job, item, q);
    if (rs == RAY_BAD) atomicAdd(&st.c[C_BAD], 1ull);
    if (rs != RAY_OK) return;
    b.Cf[0] = ray_colors[3 * q.out];
    b.Cf[1] = ray_colors[3 * q.out + 1];
    b.Cf[2] = ray_colors[3 * q.out + 2];
    if (target) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float e = b.Cf[c] - target[3 * q.out + c];
            if (loss_kind == 1) {
                loss += fabsf(e);
                b.dL[c] = e > 0.0f ? 1.0f : (e < 0.0f ? -1.0f : 0.0f);
            } else {
                loss += e * e;
                b.dL[c] = 2.0f * e;
            }
        }
    } else {
        b.dL[0] = dLdC[3 * q.out];
        b.dL[1] = dLdC[3 * q.out + 1];
        b.dL[2] = dLdC[3 * q.out + 2];
    }
    if (b.dL[0] == 0.0f && b.dL[1] == 0.0f && b.dL[2] == 0.0f) return;
    b.active = true;
    b.view = q.view;
    if constexpr (RM == 2) {
        // records of this output row are replayed only when the forward wrote them for the same
        // (view, pixel); otherwise the ray walks the tree
        const int2 id = p.rec_id[q.out];
        if (id.x == q.view && id.y == q.pixel) {
            b.rc.w = rec_base(p.rec, q.out, p.rec_words);
            b.rc.end = min(p.rec_len[q.out] & 0x7fffffff, p.rec_words);
        }
    }
    const rg_camera cam = job.cams[q.view];
    b.seed = RNG ? tea4(p.iteration, (unsigned int)q.pixel) : 0u;
    b.ray = make_ray(cam, q.pixel, sc.box, p.t_near, p.t_far, RNG ? p.jitter : 0, &b.seed);
}

// Backward over a batch.  dL/dC comes from dLdC, or -- with target != nullptr -- from the
// per-pixel loss of the rendered colours against target (loss_kind 0: L2 sum (C-t)^2, 1: L1
// sum |C-t|), whose value is summed into the loss counter.  count_stats = 0 leaves the
// traversal counters to the forward of the same step.
template <int SPB, int CAP, bool RNG, int RM>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(3))) k_backward(DevScene sc, DevParams p, RayJob job, int64_t num_items,
                                                    const float* __restrict__ ray_colors, const float* __restrict__ dLdC,
                                                    const float* __restrict__ target, int loss_kind, int count_stats,
                                                    DevGrad gr, DevStats st) {
    __shared__ IList<CAP> L;
    __shared__ GradStage G;
    const int lane = threadIdx.x;
    const int64_t nb = gridDim.x;
    const int64_t item = block_item(blockIdx.x, nb) * kWave + lane;
    Counters ctr;
    BwdRay b;
    float loss = 0.0f;
    if (item < num_items) backward_setup<RNG, RM>(sc, p, job, item, ray_colors, dLdC, target, loss_kind, st, b, loss);
    backward_wave<SPB, CAP, RNG, RM>(sc, p, b.active, b.ray, b.seed, sc.rgb + (size_t)b.view * sc.n, b.Cf, b.dL, L, G,
                                     lane, gr, b.view, ctr, b.rc, st.c + C_DROP);
    if (count_stats) flush_counters(ctr, st);
    if (target) {
        const float ls = wave_sum64(loss);
        if (lane == 0 && ls != 0.0f) atomicAdd(reinterpret_cast<float*>(&st.c[C_LOSS]), ls);
    }
}

// ------------------------------------------------------------------ persistent kernels --------
//
// The same per-ray work as k_forward / k_backward, but launched as as many wavefronts as the GPU
// holds at once, each taking p.ray_chunk work items (whole 8x8 tiles) at a time from a device work
// queue until every queue is empty.  The items are cut into 8 contiguous bands, one per XCD (a
// wavefront's XCD is its block index mod 8: blocks are dispatched round-robin over the XCDs), so
// each XCD's L2 sees a compact part of the scene as with the grid launch; a wavefront whose band is
// exhausted takes chunks from the other bands (work stealing), so no XCD idles while another still
// has rays -- the grid launch's static bands (one view per XCD on the 8-view workload) leave the
// XCDs with the cheaper views idle at the end.

// Next chunk [lo, hi) of work items for a wavefront on XCD x: its own band's queue first, then the
// other bands' in turn.  spent: bands this wavefront found exhausted.  Called by all active lanes
// together; the results are wave-uniform.
__device__ __forceinline__ bool take_chunk(const DevParams& p, int64_t num_items, int x, int64_t& lo, int64_t& hi,
                                           unsigned int& spent) {
    const int64_t tiles = (num_items + kWave - 1) / kWave;
    const int64_t band = ((tiles + 7) / 8) * kWave;  // items per band, whole tiles
    const bool leader = (int)threadIdx.x == __ffsll((long long)__ballot(1)) - 1;
    for (int j = 0; j < 8; ++j) {
        const int b = (x + j) & 7;
        if (spent & (1u << b)) continue;
        const int64_t start = (int64_t)b * band;
        const int64_t len = min(band, num_items - start);
        if (len <= 0) {
            spent |= 1u << b;
            continue;
        }
        unsigned long long got = 0ull;
        if (leader) got = atomicAdd(&p.queue[b], (unsigned long long)p.ray_chunk);
        const unsigned int glo = __builtin_amdgcn_readfirstlane((unsigned int)(got & 0xffffffffull));
        const unsigned int ghi = __builtin_amdgcn_readfirstlane((unsigned int)(got >> 32));
        const int64_t g = (int64_t)(((unsigned long long)ghi << 32) | glo);
        if (g < len) {
            lo = start + g;
            hi = start + min(g + (int64_t)p.ray_chunk, len);
            return true;
        }
        spent |= 1u << b;
    }
    return false;
}

// Lanes of mask below this lane.
__device__ __forceinline__ int lanes_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((unsigned int)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned int)mask, 0u));
}

// Persistent forward over whole tiles: each wavefront takes chunks from the queues and renders
// them one tile (64 rays, one per lane) at a time, exactly as k_forward renders its tile.
template <int SPB, int CAP, bool RNG, int RM>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(SPB == 8 ? 4 : 3)))
k_forward_pq(DevScene sc, DevParams p, RayJob job, int64_t num_items, float* __restrict__ out_rgb,
             float* __restrict__ out_depth, float* __restrict__ out_opac, DevStats st) {
    __shared__ IList<CAP> L;
    const int lane = threadIdx.x;
    int64_t lo = 0, hi = 0;
    unsigned int spent = 0u;
    while (take_chunk(p, num_items, blockIdx.x & 7, lo, hi, spent)) {
        for (int64_t base = lo; base < hi; base += kWave) {
            const int64_t item = base + lane;
            Counters ctr;  // flushed per tile: no per-lane state lives across tiles
            RayRef q;
            const int rs = item < hi ? resolve(job, item, q) : (int)RAY_NONE;
            if (rs == RAY_BAD) {
                forward_bad<RM>(p, q, st, out_rgb, out_depth, out_opac);
            } else if (rs == RAY_OK) {
                const rg_camera cam = job.cams[q.view];
                unsigned int seed = RNG ? tea4(p.iteration, (unsigned int)q.pixel) : 0u;
                const Ray ray = make_ray(cam, q.pixel, sc.box, p.t_near, p.t_far, RNG ? p.jitter : 0, &seed);
                RayOut o;
                Rec rc;
                if constexpr (RM == 1) {
                    rc.w = rec_base(p.rec, q.out, p.rec_words);
                    rc.end = p.rec_words;
                }
                forward_ray<SPB, CAP, RNG, RM>(sc, p, ray, seed, sc.rgb + (size_t)q.view * sc.n, L, lane, ctr, o, rc);
                forward_finish<RM>(p, q, o, rc, ctr, out_rgb, out_depth, out_opac);
            }
            flush_counters(ctr, st);
        }
    }
}

// Persistent forward that refills idle lanes (refill_lanes = k, 0 < k < 64): as soon as k lanes
// are done they take the next rays of the chunk (and further chunks).  A lane's ray advances one
// slab per wave step (forward_step), but a refilled lane starts at the scene box while the others
// are deeper in, so the wavefront's lanes no longer query similar ray segments and its packet
// traversals visit the union of two depth ranges: measured slower than whole tiles on the
// 1M-Gaussian workload, kept as a tuning option.
template <int SPB, int CAP, bool RNG, int RM>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(SPB == 8 ? 4 : 3)))
k_forward_refill(DevScene sc, DevParams p, RayJob job, int64_t num_items, float* __restrict__ out_rgb,
                 float* __restrict__ out_depth, float* __restrict__ out_opac, DevStats st) {
    __shared__ IList<CAP> L;
    const int lane = threadIdx.x;
    const int xcd = blockIdx.x & 7;
    const int refill = p.refill_lanes;
    Counters ctr;
    int64_t clo = 0, chi = 0;  // items of the current chunk not yet handed out (wave-uniform)
    unsigned int spent = 0u;
    bool more = true;          // the queues may hold more items (wave-uniform)
    bool alive = false;        // this lane holds a ray that is not done
    RayRef q;
    q.view = 0;
    q.pixel = 0;
    q.out = 0;
    q.active = false;
    Ray ray = {};
    unsigned int seed = 0u;
    SegState s;
    seg_init(s, ray, p);
    RayOut o;
    o.C[0] = o.C[1] = o.C[2] = 0.0f;
    o.depth = 0.0f;
    o.T = 1.0f;
    Rec rc;
    while (true) {
        if (more && __popcll(__ballot(!alive)) >= refill) {
            // hand the next items to the idle lanes, taking chunks from the queues as needed
            bool need = !alive;
            while (true) {
                const unsigned long long nm = __ballot(need);
                if (nm == 0ull) break;
                if (clo >= chi && !take_chunk(p, num_items, xcd, clo, chi, spent)) {
                    more = false;
                    break;
                }
                const int64_t avail = chi - clo;
                const int rank = lanes_below(nm);
                if (need && rank < avail) {
                    need = false;
                    const int rs = resolve(job, clo + rank, q);
                    if (rs == RAY_OK) {
                        const rg_camera cam = job.cams[q.view];
                        seed = RNG ? tea4(p.iteration, (unsigned int)q.pixel) : 0u;
                        ray = make_ray(cam, q.pixel, sc.box, p.t_near, p.t_far, RNG ? p.jitter : 0, &seed);
                        if constexpr (RM == 1) {
                            rc.w = rec_base(p.rec, q.out, p.rec_words);
                            rc.pos = 0;
                            rc.end = p.rec_words;
                            rc.full = false;
                        }
                        alive = forward_begin(p, ray, o, s);
                        if (!alive) forward_finish<RM>(p, q, o, rc, ctr, out_rgb, out_depth, out_opac);
                    } else if (rs == RAY_BAD) {
                        forward_bad<RM>(p, q, st, out_rgb, out_depth, out_opac);
                    }
                }
                clo += min((int64_t)__popcll(nm), avail);
            }
        }
        if (__ballot(alive) == 0ull) {
            if (!more) break;
            continue;
        }
        if (alive && !forward_step<SPB, CAP, RNG, RM>(sc, p, ray, seed, sc.rgb + (size_t)q.view * sc.n, L, lane, ctr, o,
                                                      rc, s)) {
            forward_finish<RM>(p, q, o, rc, ctr, out_rgb, out_depth, out_opac);
            alive = false;
        }
    }
    flush_counters(ctr, st);
}

// Persistent backward: each wavefront takes chunks of whole tiles from the queues and runs the
// depth-synchronised, wave-merged backward on one tile (64 rays) at a time.
template <int SPB, int CAP, bool RNG, int RM>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(3)))
k_backward_pq(DevScene sc, DevParams p, RayJob job, int64_t num_items, const float* __restrict__ ray_colors,
              const float* __restrict__ dLdC, const float* __restrict__ target, int loss_kind, int count_stats,
              DevGrad gr, DevStats st) {
    __shared__ IList<CAP> L;
    __shared__ GradStage G;
    const int lane = threadIdx.x;
    const int xcd = blockIdx.x & 7;
    int64_t lo = 0, hi = 0;
    unsigned int spent = 0u;
    while (take_chunk(p, num_items, xcd, lo, hi, spent)) {
        for (int64_t base = lo; base < hi; base += kWave) {
            const int64_t item = base + lane;
            Counters ctr;  // flushed per tile: no per-lane state lives across tiles
            float loss = 0.0f;
            BwdRay b;
            if (item < hi) backward_setup<RNG, RM>(sc, p, job, item, ray_colors, dLdC, target, loss_kind, st, b, loss);
            backward_wave<SPB, CAP, RNG, RM>(sc, p, b.active, b.ray, b.seed, sc.rgb + (size_t)b.view * sc.n, b.Cf, b.dL,
                                             L, G, lane, gr, b.view, ctr, b.rc, st.c + C_DROP);
            if (count_stats) flush_counters(ctr, st);
            if (target) {
                const float ls = wave_sum64(loss);
                if (lane == 0 && ls != 0.0f) atomicAdd(reinterpret_cast<float*>(&st.c[C_LOSS]), ls);
            }
        }
    }
}

__global__ void k_reset_queue(unsigned long long* q) {
    if (threadIdx.x < 8) q[threadIdx.x] = 0ull;
}

// Verification: per-slab hit sets from the same traversal + predicate (every slab to texit).
__global__ void __launch_bounds__(kWave) k_debug_hits(DevScene sc, DevParams p, RayJob job, int max_slabs,
                                                      int max_pairs, int32_t* pairs, int64_t* counts, float* ro,
                                                      float* rd, float* te, float* tx) {
    const int64_t r = (int64_t)blockIdx.x * kWave + threadIdx.x;
    if (r >= job.num_rays) return;
    RayRef q;
    if (job.pixels || job.ray_view) {
        if (resolve(job, r, q) != RAY_OK) {
            counts[r] = 0;
            return;
        }
    } else {  // no pixel list: ray r is pixel r (row-major)
        q.view = job.view;
        q.pixel = (int)r;
    }
    const rg_camera cam = job.cams[q.view];
    unsigned int seed = tea4(p.iteration, (unsigned int)q.pixel);
    const Ray ray = make_ray(cam, q.pixel, sc.box, p.t_near, p.t_far, p.jitter, &seed);
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
                trav(p, sc, ray, tmin, tmax, [&](int slot) -> bool {
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
            const float nt = add_exact(tb, slab);
            if (!(nt > tb)) break;
            tb = nt;
        }
    }
    counts[r] = n;
}

__global__ void k_reset_counters(unsigned long long* c) {
    if (threadIdx.x < 16) c[threadIdx.x] = 0;
}

}  // namespace

// In the implicit single-view mode RayJob.num_rays carries the padded 8x8-tile item count.
#define RG_DISPATCH_CAP(SPBC, RNGC, RMC, CAPV, KERNEL, ...)                                                       \
    do {                                                                                                             \
        if (CAPV <= 8) hipLaunchKernelGGL((KERNEL<SPBC, 8, RNGC, RMC>), __VA_ARGS__);                                \
        else if (CAPV <= 16) hipLaunchKernelGGL((KERNEL<SPBC, 16, RNGC, RMC>), __VA_ARGS__);                         \
        else hipLaunchKernelGGL((KERNEL<SPBC, 32, RNGC, RMC>), __VA_ARGS__);                                         \
    } while (0)
#define RG_DISPATCH_CAP6(SPBC, RNGC, RMC, CAPV, KERNEL, ...)                                                      \
    do {                                                                                                             \
        if (CAPV <= 6) hipLaunchKernelGGL((KERNEL<SPBC, 6, RNGC, RMC>), __VA_ARGS__);                                \
        else if (CAPV <= 12) hipLaunchKernelGGL((KERNEL<SPBC, 12, RNGC, RMC>), __VA_ARGS__);                         \
        else if (CAPV <= 16) hipLaunchKernelGGL((KERNEL<SPBC, 16, RNGC, RMC>), __VA_ARGS__);                         \
        else hipLaunchKernelGGL((KERNEL<SPBC, 32, RNGC, RMC>), __VA_ARGS__);                                         \
    } while (0)
// Kernels are specialised on SPB (8 or 16 samples per slab), the list capacity, on whether the
// random generator is used at all (jitter or rnd_sample), which costs registers, and on the
// hit-list record mode (RMV: 0 none, or the kernel's recording / replaying mode), so that kernels
// without records carry none of their code.
#define RG_DISPATCH_ANY(CAPMACRO, SPBV, CAPV, RNGV, RMV, RMON, KERNEL, ...)                                         \
    do {                                                                                                             \
        if (RMV) {                                                                                                   \
            if (SPBV == 8) {                                                                                         \
                if (RNGV) CAPMACRO(8, true, RMON, CAPV, KERNEL, __VA_ARGS__);                                        \
                else CAPMACRO(8, false, RMON, CAPV, KERNEL, __VA_ARGS__);                                            \
            } else {                                                                                                 \
                if (RNGV) CAPMACRO(16, true, RMON, CAPV, KERNEL, __VA_ARGS__);                                       \
                else CAPMACRO(16, false, RMON, CAPV, KERNEL, __VA_ARGS__);                                           \
            }                                                                                                        \
        } else {                                                                                                     \
            if (SPBV == 8) {                                                                                         \
                if (RNGV) CAPMACRO(8, true, 0, CAPV, KERNEL, __VA_ARGS__);                                           \
                else CAPMACRO(8, false, 0, CAPV, KERNEL, __VA_ARGS__);                                               \
            } else {                                                                                                 \
                if (RNGV) CAPMACRO(16, true, 0, CAPV, KERNEL, __VA_ARGS__);                                          \
                else CAPMACRO(16, false, 0, CAPV, KERNEL, __VA_ARGS__);                                              \
            }                                                                                                        \
        }                                                                                                            \
    } while (0)
#define RG_DISPATCH(SPBV, CAPV, RNGV, RMV, ...) RG_DISPATCH_ANY(RG_DISPATCH_CAP, SPBV, CAPV, RNGV, RMV, 2, __VA_ARGS__)
#define RG_DISPATCH_FWD(SPBV, CAPV, RNGV, RMV, ...) \
    RG_DISPATCH_ANY(RG_DISPATCH_CAP6, SPBV, CAPV, RNGV, RMV, 1, __VA_ARGS__)

hipError_t launch_reset_counters(DevStats st, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, st.c);
    return hipGetLastError();
}

// Number of wavefronts a persistent kernel is launched with: as many as the GPU holds at once
// (occupancy x CUs, a multiple of 8 so that every XCD gets the same number), and no more than
// there are chunks of work.
template <typename K>
static int64_t persistent_grid(K kernel, int64_t chunks) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kernel), kWave, 0) !=
            hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    const int64_t resident = (int64_t)cus * per_cu;
    int64_t g = chunks < resident ? chunks : resident;
    g = ((g + 7) / 8) * 8;
    return g > 0 ? g : 8;
}

template <typename K, typename... A>
static void launch_persistent(K kernel, int64_t items, const DevParams& p, hipStream_t s, A... args) {
    const int64_t chunks = (items + p.ray_chunk - 1) / p.ray_chunk;
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, s, p.queue);
    hipLaunchKernelGGL(kernel, dim3((unsigned)persistent_grid(kernel, chunks)), dim3(kWave), 0, s, args...);
}

#define RG_PQ_CAP(SPBC, RNGC, RMC, CAPV, KERNEL, ...)                                                           \
    do {                                                                                                             \
        if (CAPV <= 8) launch_persistent(KERNEL<SPBC, 8, RNGC, RMC>, __VA_ARGS__);                                  \
        else if (CAPV <= 16) launch_persistent(KERNEL<SPBC, 16, RNGC, RMC>, __VA_ARGS__);                           \
        else launch_persistent(KERNEL<SPBC, 32, RNGC, RMC>, __VA_ARGS__);                                           \
    } while (0)
#define RG_PQ_CAP6(SPBC, RNGC, RMC, CAPV, KERNEL, ...)                                                          \
    do {                                                                                                             \
        if (CAPV <= 6) launch_persistent(KERNEL<SPBC, 6, RNGC, RMC>, __VA_ARGS__);                                  \
        else if (CAPV <= 12) launch_persistent(KERNEL<SPBC, 12, RNGC, RMC>, __VA_ARGS__);                           \
        else if (CAPV <= 16) launch_persistent(KERNEL<SPBC, 16, RNGC, RMC>, __VA_ARGS__);                           \
        else launch_persistent(KERNEL<SPBC, 32, RNGC, RMC>, __VA_ARGS__);                                           \
    } while (0)

hipError_t launch_forward(const DevScene& sc, const DevParams& p, const RayJob& job, int hit_depth, float* out_rgb,
                          float* out_depth, float* out_opacity, DevStats st, hipStream_t s) {
    const int64_t items = job.num_rays;  // already the padded item count in implicit mode
    if (items <= 0) return hipSuccess;
    const bool rng = p.jitter || p.rnd_sample;
    if (p.ray_chunk > 0 && p.refill_lanes > 0 && p.refill_lanes < kWave) {
        RG_DISPATCH_ANY(RG_PQ_CAP6, p.spb <= 8 ? 8 : 16, hit_depth, rng, p.rec_mode == 1, 1, k_forward_refill, items, p,
                        s, sc, p, job, items, out_rgb, out_depth, out_opacity, st);
        return hipGetLastError();
    }
    if (p.ray_chunk > 0) {
        RG_DISPATCH_ANY(RG_PQ_CAP6, p.spb <= 8 ? 8 : 16, hit_depth, rng, p.rec_mode == 1, 1, k_forward_pq, items, p, s,
                        sc, p, job, items, out_rgb, out_depth, out_opacity, st);
        return hipGetLastError();
    }
    const int64_t nb = (items + kWave - 1) / kWave;
    const int64_t nbp = ((nb + 7) / 8) * 8;
    RG_DISPATCH_FWD(p.spb <= 8 ? 8 : 16, hit_depth, rng, p.rec_mode == 1, k_forward, dim3((unsigned)nbp), dim3(kWave), 0, s, sc, p, job,
                items, out_rgb, out_depth, out_opacity, st);
    return hipGetLastError();
}

hipError_t launch_backward(const DevScene& sc, const DevParams& p, const RayJob& job, int hit_depth,
                           const float* ray_colors, const float* dLdC, const float* target, int loss_kind,
                           int count_stats, DevGrad gr, DevStats st, hipStream_t s) {
    const int64_t items = job.num_rays;
    if (items <= 0) return hipSuccess;
    const bool rng = p.jitter || p.rnd_sample;
    if (p.ray_chunk > 0) {
        RG_DISPATCH_ANY(RG_PQ_CAP, p.spb <= 8 ? 8 : 16, hit_depth, rng, p.rec_mode == 2, 2, k_backward_pq, items, p, s,
                        sc, p, job, items, ray_colors, dLdC, target, loss_kind, count_stats, gr, st);
        return hipGetLastError();
    }
    const int64_t nb = (items + kWave - 1) / kWave;
    const int64_t nbp = ((nb + 7) / 8) * 8;
    RG_DISPATCH(p.spb <= 8 ? 8 : 16, hit_depth, rng, p.rec_mode == 2, k_backward, dim3((unsigned)nbp), dim3(kWave), 0, s, sc, p, job,
                items, ray_colors, dLdC, target, loss_kind, count_stats, gr, st);
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
