// Densification and initialisation utilities of the training loop:
//
//   knn      distCUDA2 of simple-knn (submodules/simple-knn/simple_knn.cu:186-222): per point, the
//            mean squared distance to its 3 nearest other points.  The reference sorts by 30-bit
//            Morton code, cuts the order into boxes of 1024 points and lets every point scan each
//            box that may hold a closer point; that search is exact.  Here one wavefront owns a
//            bucket of 64 Morton-consecutive points: it seeds each lane's three best from its own
//            bucket, then walks 64-bucket groups and buckets whose boxes may still hold a closer
//            point for ANY lane (wave ballots), staging each such bucket in LDS once for all 64
//            lanes -- uniform control flow and one coalesced load per visited bucket instead of a
//            divergent per-thread scan of 1024 points.
//   filter3d PointCloud.compute_3D_filter (classes/point_cloud.py:1152-1215): one thread per
//            point walks the cameras (uniform loads), a wave max feeds one atomic per wave.
//   densify  densify_and_clone + densify_and_split + prune (classes/point_cloud.py:1067-1143) as
//            a plan (decisions + a three-phase scan giving every source row its destinations) and
//            an apply (one streaming copy over all per-Gaussian arrays and their Adam moments,
//            then the transformed rows), instead of the reference's chain of boolean masks and
//            concatenations per tensor.
//   gradient statistics (classes/point_cloud.py:632-640).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <mutex>

#include "rg_internal.h"

#pragma clang fp contract(off)

namespace rg {
namespace {

constexpr int kB = 64;          // points per kNN bucket = one wavefront
constexpr int kBoxBlocks = 512;  // partial boxes of the point set

inline int nblocks(int64_t n, int b) { return (int)((n + b - 1) / b); }
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// ------------------------------------------------------------------ kNN ----------------------

__device__ __forceinline__ float wmin(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wmax(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Bounding box of the points: grid-stride partials (lo, hi) per block.
__global__ void __launch_bounds__(256) k_knn_box_partial(int64_t n, const float* __restrict__ p, float4* part) {
    float lx = FLT_MAX, ly = FLT_MAX, lz = FLT_MAX, hx = -FLT_MAX, hy = -FLT_MAX, hz = -FLT_MAX;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
        lx = fminf(lx, x);
        ly = fminf(ly, y);
        lz = fminf(lz, z);
        hx = fmaxf(hx, x);
        hy = fmaxf(hy, y);
        hz = fmaxf(hz, z);
    }
    __shared__ float red[6][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    lx = wmin(lx), ly = wmin(ly), lz = wmin(lz), hx = wmax(hx), hy = wmax(hy), hz = wmax(hz);
    if (lane == 0) {
        red[0][w] = lx, red[1][w] = ly, red[2][w] = lz, red[3][w] = hx, red[4][w] = hy, red[5][w] = hz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; ++k) {
            lx = fminf(lx, red[0][k]), ly = fminf(ly, red[1][k]), lz = fminf(lz, red[2][k]);
            hx = fmaxf(hx, red[3][k]), hy = fmaxf(hy, red[4][k]), hz = fmaxf(hz, red[5][k]);
        }
        part[2 * blockIdx.x] = make_float4(lx, ly, lz, 0.0f);
        part[2 * blockIdx.x + 1] = make_float4(hx, hy, hz, 0.0f);
    }
}

__global__ void __launch_bounds__(64) k_knn_box_final(int m, const float4* __restrict__ part, float4* box) {
    float lx = FLT_MAX, ly = FLT_MAX, lz = FLT_MAX, hx = -FLT_MAX, hy = -FLT_MAX, hz = -FLT_MAX;
    for (int i = threadIdx.x; i < m; i += 64) {
        const float4 lo = part[2 * i], hi = part[2 * i + 1];
        lx = fminf(lx, lo.x), ly = fminf(ly, lo.y), lz = fminf(lz, lo.z);
        hx = fmaxf(hx, hi.x), hy = fmaxf(hy, hi.y), hz = fmaxf(hz, hi.z);
    }
    lx = wmin(lx), ly = wmin(ly), lz = wmin(lz), hx = wmax(hx), hy = wmax(hy), hz = wmax(hz);
    if (threadIdx.x == 0) {
        box[0] = make_float4(lx, ly, lz, 0.0f);
        box[1] = make_float4(hx, hy, hz, 0.0f);
    }
}

__device__ __forceinline__ unsigned int spread10(unsigned int v) {
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_knn_morton(int n, const float* __restrict__ p, const float4* __restrict__ box,
                             unsigned long long* keys, int* vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 lo = box[0], hi = box[1];
    const float ex = fmaxf(hi.x - lo.x, 1e-30f), ey = fmaxf(hi.y - lo.y, 1e-30f), ez = fmaxf(hi.z - lo.z, 1e-30f);
    const unsigned int qx = (unsigned int)fminf(fmaxf((p[3 * i] - lo.x) / ex * 1023.0f, 0.0f), 1023.0f);
    const unsigned int qy = (unsigned int)fminf(fmaxf((p[3 * i + 1] - lo.y) / ey * 1023.0f, 0.0f), 1023.0f);
    const unsigned int qz = (unsigned int)fminf(fmaxf((p[3 * i + 2] - lo.z) / ez * 1023.0f, 0.0f), 1023.0f);
    keys[i] = (unsigned long long)(spread10(qx) | (spread10(qy) << 1) | (spread10(qz) << 2));
    vals[i] = i;
}

__global__ void k_knn_gather(int n, const float* __restrict__ p, const int* __restrict__ order, float4* sp) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int g = order[s];
    sp[s] = make_float4(p[3 * g], p[3 * g + 1], p[3 * g + 2], 0.0f);
}

// Box of each group of 64 consecutive items; items are points (lo = hi = src[i]) or boxes
// (lo = src[2i], hi = src[2i+1]).  One wavefront per group.
__global__ void __launch_bounds__(kB) k_knn_group_box(int count, const float4* __restrict__ src, int is_box,
                                                      float4* out) {
    const int g = blockIdx.x, lane = threadIdx.x;
    const int i = g * kB + lane;
    float4 lo = make_float4(FLT_MAX, FLT_MAX, FLT_MAX, 0.0f), hi = make_float4(-FLT_MAX, -FLT_MAX, -FLT_MAX, 0.0f);
    if (i < count) {
        lo = is_box ? src[2 * i] : src[i];
        hi = is_box ? src[2 * i + 1] : lo;
    }
    lo.x = wmin(lo.x), lo.y = wmin(lo.y), lo.z = wmin(lo.z);
    hi.x = wmax(hi.x), hi.y = wmax(hi.y), hi.z = wmax(hi.z);
    if (lane == 0) {
        out[2 * g] = make_float4(lo.x, lo.y, lo.z, 0.0f);
        out[2 * g + 1] = make_float4(hi.x, hi.y, hi.z, 0.0f);
    }
}

// Squared FP32 distance.  Every rounding step is monotone in |dx|, |dy|, |dz|, so the box
// distance below is a lower bound of the distance to every point inside the box.
__device__ __forceinline__ float sqdist(float dx, float dy, float dz) { return fmaf(dz, dz, fmaf(dy, dy, dx * dx)); }

__device__ __forceinline__ float axis_gap(float p, float lo, float hi) {
    return (p < lo || p > hi) ? fminf(fabsf(p - lo), fabsf(p - hi)) : 0.0f;
}

__device__ __forceinline__ float box_dist(const float4& lo, const float4& hi, const float4& p) {
    return sqdist(axis_gap(p.x, lo.x, hi.x), axis_gap(p.y, lo.y, hi.y), axis_gap(p.z, lo.z, hi.z));
}

__device__ __forceinline__ void keep3(float d, float* b) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        if (b[j] > d) {
            const float t = b[j];
            b[j] = d;
            d = t;
        }
    }
}

// One wavefront per bucket of 64 sorted points (lane = query).
__global__ void __launch_bounds__(kB) k_knn_search(int n, const float4* __restrict__ sp, const int* __restrict__ order,
                                                   int nb, const float4* __restrict__ bbox, int ns,
                                                   const float4* __restrict__ sbox, float* __restrict__ out) {
    __shared__ float4 tile[kB];
    const int b = blockIdx.x, lane = threadIdx.x;
    const int s = b * kB + lane;
    const bool valid = s < n;
    const float4 p = valid ? sp[s] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float best[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    // own bucket
    tile[lane] = p;
    __syncthreads();
    {
        const int cnt = min(kB, n - b * kB);
        for (int j = 0; j < cnt; ++j) {
            if (j == lane) continue;
            const float4 q = tile[j];
            keep3(sqdist(q.x - p.x, q.y - p.y, q.z - p.z), best);
        }
    }
    // other buckets, by groups of 64
    for (int S = 0; S < ns; ++S) {
        const float ds = valid ? box_dist(sbox[2 * S], sbox[2 * S + 1], p) : FLT_MAX;
        if (__ballot(valid && ds <= best[2]) == 0ULL) continue;
        const int b1 = min(nb, (S + 1) * kB);
        for (int bb = S * kB; bb < b1; ++bb) {
            if (bb == b) continue;
            const float db = valid ? box_dist(bbox[2 * bb], bbox[2 * bb + 1], p) : FLT_MAX;
            const bool need = valid && db <= best[2];
            if (__ballot(need) == 0ULL) continue;
            const int cnt = min(kB, n - bb * kB);
            __syncthreads();
            if (lane < cnt) tile[lane] = sp[bb * kB + lane];
            __syncthreads();
            if (need) {
                for (int j = 0; j < cnt; ++j) {
                    const float4 q = tile[j];
                    keep3(sqdist(q.x - p.x, q.y - p.y, q.z - p.z), best);
                }
            }
        }
    }
    if (valid) out[order[s]] = (best[0] + best[1] + best[2]) / 3.0f;
}

struct KnnWs {
    unsigned long long* keys[2];
    int* vals[2];
    unsigned int* hist;
    float4* sp;
    float4* bbox;
    float4* sbox;
    float4* part;
    float4* box;
    size_t bytes;
};

KnnWs knn_layout(int64_t n, void* base) {
    KnnWs w{};
    const int nb = nblocks(n, kB), ns = nblocks(nb, kB);
    char* p = static_cast<char*>(base);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* q = p ? p + off : nullptr;
        off += align256(bytes);
        return q;
    };
    w.keys[0] = reinterpret_cast<unsigned long long*>(take(8 * (size_t)n));
    w.keys[1] = reinterpret_cast<unsigned long long*>(take(8 * (size_t)n));
    w.vals[0] = reinterpret_cast<int*>(take(4 * (size_t)n));
    w.vals[1] = reinterpret_cast<int*>(take(4 * (size_t)n));
    w.hist = reinterpret_cast<unsigned int*>(take((size_t)radix_hist_bytes((int)n)));
    w.sp = reinterpret_cast<float4*>(take(16 * (size_t)n));
    w.bbox = reinterpret_cast<float4*>(take(32 * (size_t)nb));
    w.sbox = reinterpret_cast<float4*>(take(32 * (size_t)ns));
    w.part = reinterpret_cast<float4*>(take(32 * (size_t)kBoxBlocks));
    w.box = reinterpret_cast<float4*>(take(32));
    w.bytes = off;
    return w;
}

// ------------------------------------------------------------------ 3D filter ----------------

__global__ void k_put_u32(unsigned int* p, unsigned int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void __launch_bounds__(256) k_filter_depth(int64_t n, const float* __restrict__ p,
                                                      const rg_filter_camera* __restrict__ cams, int V,
                                                      float* __restrict__ out, unsigned int* dmax_bits) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float d = 100000.0f;
    bool seen = false;
    if (i < n) {
        const float x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
        for (int v = 0; v < V; ++v) {
            const rg_filter_camera& c = cams[v];
            // row vector times R: x_cam[j] = x R[0][j] + y R[1][j] + z R[2][j] + T[j]
            const float xc = fmaf(z, c.R[6], fmaf(y, c.R[3], x * c.R[0])) + c.T[0];
            const float yc = fmaf(z, c.R[7], fmaf(y, c.R[4], x * c.R[1])) + c.T[1];
            const float zc = fmaf(z, c.R[8], fmaf(y, c.R[5], x * c.R[2])) + c.T[2];
            const bool front = zc > 0.2f;
            const float zz = fmaxf(zc, 0.001f);
            const float sx = xc / zz * c.fx + c.cx;
            const float sy = yc / zz * c.fy + c.cy;
            const bool in = sx >= c.x_lo && sx <= c.x_hi && sy >= c.y_lo && sy <= c.y_hi;
            if (front && in) {
                d = fminf(d, zz);
                seen = true;
            }
        }
        out[i] = seen ? d : -1.0f;
    }
    // seen depths are >= 0.2 > 0, so their bit patterns order as the values do
    const float m = wmax(seen ? d : 0.0f);
    if ((threadIdx.x & 63) == 0 && m > 0.0f) atomicMax(dmax_bits, __float_as_uint(m));
}

__global__ void __launch_bounds__(256) k_filter_finish(int64_t n, const unsigned int* __restrict__ dmax_bits,
                                                       float inv_focal, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float d = out[i];
    if (d < 0.0f) d = __uint_as_float(*dmax_bits);
    out[i] = logf((0.5f * d) * inv_focal);
}

// ------------------------------------------------------------------ gradient statistics -------

__global__ void __launch_bounds__(256) k_accum_grad(int64_t n, const float* __restrict__ g, float* accum, float* num,
                                                    float* gnv) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float x = g[3 * i], y = g[3 * i + 1], z = g[3 * i + 2];
    const float nrm = sqrtf(x * x + y * y + z * z);
    accum[i] += nrm;
    const float one = nrm > 0.0f ? 1.0f : 0.0f;
    num[i] += one;
    if (gnv) gnv[i] += one;
}

// ------------------------------------------------------------------ densify -------------------

constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

struct DensWs {
    int* dst_orig;    // destination row of the source row itself, -1 when it goes
    int* dst_clone;   // destination row of its clone, -1 when none
    int* split_rank;  // rank among all split rows, -1 when not split
    int* dst_child;   // destination row of its first child, -1 when none; child c at + c * K_children
    uint4* part;      // per tile: counts of (kept originals, kept clones, split, kept children)
    size_t bytes;
};

DensWs dens_layout(int64_t n, void* base) {
    DensWs w{};
    char* p = static_cast<char*>(base);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* q = p ? p + off : nullptr;
        off += align256(bytes);
        return q;
    };
    w.dst_orig = reinterpret_cast<int*>(take(4 * (size_t)n));
    w.dst_clone = reinterpret_cast<int*>(take(4 * (size_t)n));
    w.split_rank = reinterpret_cast<int*>(take(4 * (size_t)n));
    w.dst_child = reinterpret_cast<int*>(take(4 * (size_t)n));
    w.part = reinterpret_cast<uint4*>(take(16 * (size_t)(nblocks(n, kScanTile) + 1)));
    w.bytes = off;
    return w;
}

struct DensDecide {
    int64_t n;
    const float* accum;
    const float* num;
    const float* log_scales;
    const float* densities;
    float thr, scale_limit, min_density;
};

// (kept original, kept clone, split, kept children) of one source row.
__device__ __forceinline__ uint4 dens_flags(const DensDecide& d, int64_t i) {
    float g = d.accum[i] / d.num[i];
    if (g != g) g = 0.0f;  // 0/0 of a Gaussian that never received a gradient
    const float smax = fmaxf(fmaxf(expf(d.log_scales[3 * i]), expf(d.log_scales[3 * i + 1])),
                             expf(d.log_scales[3 * i + 2]));
    const bool big_grad = g >= d.thr;
    const bool clone = big_grad && smax <= d.scale_limit;
    const bool split = big_grad && smax > d.scale_limit;
    const float sig = d.densities[i];
    const float half = sig * 0.5f;
    const float own = clone ? half : sig;
    uint4 f;
    f.x = (!split && !(own <= d.min_density)) ? 1u : 0u;
    f.y = (clone && !(half <= d.min_density)) ? 1u : 0u;
    f.z = split ? 1u : 0u;
    f.w = (split && !(half <= d.min_density)) ? 1u : 0u;
    return f;
}

__device__ __forceinline__ uint4 add4(uint4 a, uint4 b) { return make_uint4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }

__global__ void __launch_bounds__(kScanBlock) k_dens_count(DensDecide d, uint4* part) {
    __shared__ uint4 red[kScanBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    uint4 c = make_uint4(0, 0, 0, 0);
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + (int64_t)k * kScanBlock + threadIdx.x;
        if (i < d.n) c = add4(c, dens_flags(d, i));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        c.x += __shfl_xor(c.x, o);
        c.y += __shfl_xor(c.y, o);
        c.z += __shfl_xor(c.z, o);
        c.w += __shfl_xor(c.w, o);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint4 t = red[0];
        for (int k = 1; k < kScanBlock / 64; ++k) t = add4(t, red[k]);
        part[blockIdx.x] = t;
    }
}

// Exclusive scan of the tile counts (one block); the totals go to part[m] and counts[0..3].
__global__ void __launch_bounds__(1024) k_dens_scan(int m, uint4* part, int64_t* counts) {
    __shared__ uint4 sh[1024];
    const int t = threadIdx.x;
    const int chunk = (m + 1023) / 1024;
    const int b = t * chunk, e = min(m, b + chunk);
    uint4 s = make_uint4(0, 0, 0, 0);
    for (int i = b; i < e; ++i) s = add4(s, part[i]);
    sh[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint4 v = t >= o ? sh[t - o] : make_uint4(0, 0, 0, 0);
        __syncthreads();
        sh[t] = add4(sh[t], v);
        __syncthreads();
    }
    uint4 run = make_uint4(sh[t].x - s.x, sh[t].y - s.y, sh[t].z - s.z, sh[t].w - s.w);
    for (int i = b; i < e; ++i) {
        const uint4 v = part[i];
        part[i] = run;
        run = add4(run, v);
    }
    if (t == 1023) {
        part[m] = sh[t];
        counts[0] = sh[t].x;
        counts[1] = sh[t].y;
        counts[2] = sh[t].z;
        counts[3] = sh[t].w;
    }
}

__global__ void __launch_bounds__(kScanBlock) k_dens_index(DensDecide d, const uint4* __restrict__ part, int m,
                                                           DensWs w) {
    __shared__ uint4 sh[kScanBlock];
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)t * kScanItems;  // 4 consecutive rows
    uint4 f[kScanItems];
    uint4 s = make_uint4(0, 0, 0, 0);
    for (int k = 0; k < kScanItems; ++k) {
        f[k] = (base + k < d.n) ? dens_flags(d, base + k) : make_uint4(0, 0, 0, 0);
        s = add4(s, f[k]);
    }
    sh[t] = s;
    __syncthreads();
    for (int o = 1; o < kScanBlock; o <<= 1) {
        const uint4 v = t >= o ? sh[t - o] : make_uint4(0, 0, 0, 0);
        __syncthreads();
        sh[t] = add4(sh[t], v);
        __syncthreads();
    }
    const uint4 tot = part[m];
    const uint4 blk = part[blockIdx.x];
    uint4 r = add4(blk, make_uint4(sh[t].x - s.x, sh[t].y - s.y, sh[t].z - s.z, sh[t].w - s.w));
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k;
        if (i >= d.n) break;
        w.dst_orig[i] = f[k].x ? (int)r.x : -1;
        w.dst_clone[i] = f[k].y ? (int)(tot.x + r.y) : -1;
        w.split_rank[i] = f[k].z ? (int)r.z : -1;
        w.dst_child[i] = f[k].w ? (int)(tot.x + tot.y + r.w) : -1;
        r = add4(r, f[k]);
    }
}

constexpr int kMaxDensArrays = 32;
struct DensArray {
    const float* src;
    float* dst;
    int64_t width;
    int zero_new;
};
struct DensArrays {
    DensArray a[kMaxDensArrays];
};

// blockIdx.y = array; grid-stride over its n*width source floats.
__global__ void __launch_bounds__(256) k_dens_copy(int64_t n, DensArrays arrs, DensWs w, int split_n, int64_t kch) {
    const DensArray A = arrs.a[blockIdx.y];
    const int64_t total = n * A.width;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t i = e / A.width, k = e - i * A.width;
        const float v = A.src[e];
        const float nv = A.zero_new ? 0.0f : v;
        const int o = w.dst_orig[i], c = w.dst_clone[i], ch = w.dst_child[i];
        if (o >= 0) A.dst[(int64_t)o * A.width + k] = v;
        if (c >= 0) A.dst[(int64_t)c * A.width + k] = nv;
        if (ch >= 0)
            for (int j = 0; j < split_n; ++j) A.dst[((int64_t)ch + (int64_t)j * kch) * A.width + k] = nv;
    }
}

struct DensFix {
    int64_t n;
    const float* pos;
    const float* ls;
    const float* q;
    const float* sig;
    const float* normals;
    float* out_pos;
    float* out_ls;
    float* out_sig;
    int split_n;
    int64_t n_split;
    int64_t kch;
    float inv_split;  // float(1 / float(0.8 N))
};

__global__ void __launch_bounds__(256) k_dens_fix(DensFix f, DensWs w) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= f.n) return;
    const int o = w.dst_orig[i], c = w.dst_clone[i], ch = w.dst_child[i];
    const float half = f.sig[i] * 0.5f;
    if (c >= 0) {  // a cloned Gaussian: its own density and its clone's are halved
        if (o >= 0) f.out_sig[o] = half;
        f.out_sig[c] = half;
    }
    if (ch < 0) return;
    // children of a split Gaussian: positions R(q) (s * z) + mu with q normalised as build_rotation
    const float qw = f.q[4 * i], qx = f.q[4 * i + 1], qy = f.q[4 * i + 2], qz = f.q[4 * i + 3];
    const float nrm = sqrtf(qw * qw + qx * qx + qy * qy + qz * qz);
    const float r = qw / nrm, x = qx / nrm, y = qy / nrm, z = qz / nrm;
    const float R00 = 1.0f - 2.0f * (y * y + z * z), R01 = 2.0f * (x * y - r * z), R02 = 2.0f * (x * z + r * y);
    const float R10 = 2.0f * (x * y + r * z), R11 = 1.0f - 2.0f * (x * x + z * z), R12 = 2.0f * (y * z - r * x);
    const float R20 = 2.0f * (x * z - r * y), R21 = 2.0f * (y * z + r * x), R22 = 1.0f - 2.0f * (x * x + y * y);
    const float s0 = expf(f.ls[3 * i]), s1 = expf(f.ls[3 * i + 1]), s2 = expf(f.ls[3 * i + 2]);
    const float l0 = logf(s0 * f.inv_split), l1 = logf(s1 * f.inv_split), l2 = logf(s2 * f.inv_split);
    const int64_t rank = w.split_rank[i];
    for (int j = 0; j < f.split_n; ++j) {
        const int64_t row = (int64_t)ch + (int64_t)j * f.kch;
        const float* zr = f.normals + 3 * ((int64_t)j * f.n_split + rank);
        const float a = zr[0] * s0, b = zr[1] * s1, cc = zr[2] * s2;
        f.out_pos[3 * row] = fmaf(R02, cc, fmaf(R01, b, R00 * a)) + f.pos[3 * i];
        f.out_pos[3 * row + 1] = fmaf(R12, cc, fmaf(R11, b, R10 * a)) + f.pos[3 * i + 1];
        f.out_pos[3 * row + 2] = fmaf(R22, cc, fmaf(R21, b, R20 * a)) + f.pos[3 * i + 2];
        f.out_ls[3 * row] = l0;
        f.out_ls[3 * row + 1] = l1;
        f.out_ls[3 * row + 2] = l2;
        f.out_sig[row] = half;
    }
}

}  // namespace

// ------------------------------------------------------------------ launchers -----------------

int64_t knn_workspace_bytes(int64_t n) { return n <= 0 ? 0 : (int64_t)knn_layout(n, nullptr).bytes; }

hipError_t launch_knn_mean_dist(const float* points, int64_t n, float* out, void* ws, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    KnnWs w = knn_layout(n, ws);
    const int N = (int)n;
    const int nb = nblocks(n, kB), ns = nblocks(nb, kB);
    const int pb = min(kBoxBlocks, nblocks(n, 256));
    hipLaunchKernelGGL(k_knn_box_partial, dim3(pb), dim3(256), 0, s, n, points, w.part);
    hipLaunchKernelGGL(k_knn_box_final, dim3(1), dim3(64), 0, s, pb, w.part, w.box);
    hipLaunchKernelGGL(k_knn_morton, dim3(nblocks(n, 256)), dim3(256), 0, s, N, points, w.box, w.keys[0], w.vals[0]);
    const int cur = launch_radix_sort(N, 4, w.keys, w.vals, w.hist, s);
    hipLaunchKernelGGL(k_knn_gather, dim3(nblocks(n, 256)), dim3(256), 0, s, N, points, w.vals[cur], w.sp);
    hipLaunchKernelGGL(k_knn_group_box, dim3(nb), dim3(kB), 0, s, N, w.sp, 0, w.bbox);
    hipLaunchKernelGGL(k_knn_group_box, dim3(ns), dim3(kB), 0, s, nb, w.bbox, 1, w.sbox);
    hipLaunchKernelGGL(k_knn_search, dim3(nb), dim3(kB), 0, s, N, w.sp, w.vals[cur], nb, w.bbox, ns, w.sbox, out);
    return hipGetLastError();
}

// ---- kNN without a workspace argument ---------------------------------------------------------
//
// For callers that cannot allocate device memory (code under stream capture, serving loops):
// the workspace of launch_knn_mean_dist lives in this file-scope array instead.  Every GPU the
// library runs on holds its own copy until the process ends.  Host code never passes the array's
// name to a kernel (on the host it is only a stand-in for the symbol): it asks the runtime for the
// array's device address once per device, with that device current, and uses that pointer.  The
// kernels of one call fully write every scratch word they later read, so nothing needs clearing
// between calls; two calls that overlap on one GPU share the array and race.
constexpr int64_t kKnnScratchBytes = (int64_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kKnnScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxScratchDevices] = {};
}  // namespace

int64_t knn_scratch_bytes() { return kKnnScratchBytes; }

// Device address of this device's scratch array, looked up on the first call per device.
static hipError_t knn_scratch_ptr(void** out) {
    *out = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return hipSuccess;
}

hipError_t launch_knn_mean_dist_noalloc(const float* points, int64_t n, float* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (knn_workspace_bytes(n) > kKnnScratchBytes) return hipErrorInvalidValue;  // checked by the caller too
    void* ws = nullptr;
    const hipError_t e = knn_scratch_ptr(&ws);
    if (e != hipSuccess) return e;
    return launch_knn_mean_dist(points, n, out, ws, s);
}

hipError_t launch_filter_3d(const float* points, int64_t n, const rg_filter_camera* cams_dev, int V, float inv_focal,
                            float* out, unsigned int* dmax_bits, hipStream_t s) {
    hipLaunchKernelGGL(k_put_u32, dim3(1), dim3(64), 0, s, dmax_bits, 0u);
    if (n <= 0) return hipGetLastError();
    hipLaunchKernelGGL(k_filter_depth, dim3(nblocks(n, 256)), dim3(256), 0, s, n, points, cams_dev, V, out,
                       dmax_bits);
    hipLaunchKernelGGL(k_filter_finish, dim3(nblocks(n, 256)), dim3(256), 0, s, n, dmax_bits, inv_focal, out);
    return hipGetLastError();
}

hipError_t launch_accum_grad(const float* g, int64_t n, float* accum, float* num, float* gnv, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_accum_grad, dim3(nblocks(n, 256)), dim3(256), 0, s, n, g, accum, num, gnv);
    return hipGetLastError();
}

int64_t densify_workspace_bytes(int64_t n) { return n <= 0 ? 256 : (int64_t)dens_layout(n, nullptr).bytes; }

static DensDecide decide_of(const rg_densify_in& in) {
    DensDecide d;
    d.n = in.n;
    d.accum = in.grad_accum;
    d.num = in.num_accum;
    d.log_scales = in.log_scales;
    d.densities = in.densities;
    d.thr = in.grad_threshold;
    d.scale_limit = in.scale_limit;
    d.min_density = in.min_density;
    return d;
}

hipError_t launch_densify_plan(const rg_densify_in& in, void* ws, int64_t* counts, hipStream_t s) {
    DensWs w = dens_layout(in.n, ws);
    const int m = nblocks(in.n, kScanTile);
    const DensDecide d = decide_of(in);
    if (m > 0) hipLaunchKernelGGL(k_dens_count, dim3(m), dim3(kScanBlock), 0, s, d, w.part);
    hipLaunchKernelGGL(k_dens_scan, dim3(1), dim3(1024), 0, s, m, w.part, counts);
    if (m > 0) hipLaunchKernelGGL(k_dens_index, dim3(m), dim3(kScanBlock), 0, s, d, w.part, m, w);
    return hipGetLastError();
}

hipError_t launch_densify_apply(const rg_densify_in& in, const void* ws, const int64_t* counts,
                                const rg_densify_array* arrays, int num_arrays, const float* normals, float* out_pos,
                                float* out_ls, float* out_sig, hipStream_t s) {
    if (in.n <= 0) return hipSuccess;
    DensWs w = dens_layout(in.n, const_cast<void*>(ws));
    DensArrays A{};
    int64_t maxw = 1;
    for (int k = 0; k < num_arrays; ++k) {
        A.a[k].src = arrays[k].src;
        A.a[k].dst = arrays[k].dst;
        A.a[k].width = arrays[k].width;
        A.a[k].zero_new = arrays[k].zero_new;
        if (arrays[k].width > maxw) maxw = arrays[k].width;
    }
    const int64_t kch = counts[3];
    if (num_arrays > 0) {
        const int gx = (int)std::min<int64_t>(4096, nblocks(in.n * maxw, 256));
        hipLaunchKernelGGL(k_dens_copy, dim3(gx, num_arrays), dim3(256), 0, s, in.n, A, w, in.split_n, kch);
    }
    DensFix f;
    f.n = in.n;
    f.pos = in.positions;
    f.ls = in.log_scales;
    f.q = in.quats;
    f.sig = in.densities;
    f.normals = normals;
    f.out_pos = out_pos;
    f.out_ls = out_ls;
    f.out_sig = out_sig;
    f.split_n = in.split_n;
    f.n_split = counts[2];
    f.kch = kch;
    const float div = (float)(0.8 * (double)in.split_n);
    f.inv_split = 1.0f / div;
    hipLaunchKernelGGL(k_dens_fix, dim3(nblocks(in.n, 256)), dim3(256), 0, s, f, w);
    return hipGetLastError();
}

int densify_max_arrays() { return kMaxDensArrays; }

}  // namespace rg
