// LBVH over the per-Gaussian bounding-ellipsoid AABBs: replaces the OptiX GAS build/update of
// the reference (optix_utils.py:68-75, rebuilt every forward at render_optix.py:107-110).
//
//   build : 63-bit Morton codes of the Gaussian means (21 bits per axis over the scene box)
//           -> LSD radix sort (8 passes of 8 bits; per-tile LDS histograms, one scan, stable
//              LDS-ranked scatter using wavefront ballots)
//           -> Karras (2012) binary radix tree, one thread per internal node
//           -> DFS preorder renumbering with "miss" (escape) links, so traversal needs no stack:
//              a node's preorder index is 2*first_leaf + (left turns on its root path), and its
//              subtree occupies the next 2*leaves-1 indices.
//   refit : bottom-up, one thread per leaf; the second thread to reach an internal node (atomic
//           flag) writes the union of its children's boxes.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rg_internal.h"

namespace rg {
namespace {

constexpr int kSortBlock = 256;
constexpr int kSortItems = 8;
constexpr int kTile = kSortBlock * kSortItems;

__global__ void k_iota(int n, int* perm, int* iperm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        perm[i] = i;
        iperm[i] = i;
    }
}

__device__ __forceinline__ unsigned long long spread21(unsigned int v) {
    unsigned long long x = v & 0x1fffff;
    x = (x | (x << 32)) & 0x1f00000000ffffULL;
    x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}

__global__ void k_morton(int n, const float4* __restrict__ leafbox_orig, const float* __restrict__ box,
                         unsigned long long* keys, int* vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // centroid of the (unpadded-equivalent) box; empty boxes (never-hit Gaussians) sort to the end
    const float4 lo = leafbox_orig[2 * i], hi = leafbox_orig[2 * i + 1];
    unsigned long long key;
    if (!(lo.x <= hi.x)) {
        key = 0x7fffffffffffffffULL;
    } else {
        const float cx = 0.5f * (lo.x + hi.x), cy = 0.5f * (lo.y + hi.y), cz = 0.5f * (lo.z + hi.z);
        const float ex = fmaxf(box[3] - box[0], 1e-30f), ey = fmaxf(box[4] - box[1], 1e-30f),
                    ez = fmaxf(box[5] - box[2], 1e-30f);
        const float s = 2097151.0f;
        const unsigned int qx = (unsigned int)fminf(fmaxf((cx - box[0]) / ex * s, 0.0f), s);
        const unsigned int qy = (unsigned int)fminf(fmaxf((cy - box[1]) / ey * s, 0.0f), s);
        const unsigned int qz = (unsigned int)fminf(fmaxf((cz - box[2]) / ez * s, 0.0f), s);
        key = (spread21(qx) << 2) | (spread21(qy) << 1) | spread21(qz);
    }
    keys[i] = key;
    vals[i] = i;
}

// ---------------------------------------------------------------- radix sort ----------------

__global__ void __launch_bounds__(kSortBlock) k_hist(int n, int shift, const unsigned long long* __restrict__ keys,
                                                     unsigned int* __restrict__ hist, int tiles) {
    __shared__ unsigned int cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kTile;
    for (int r = 0; r < kSortItems; ++r) {
        const int i = base + r * kSortBlock + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 0xff], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of hist[0..m) in place, one block of 1024 threads.
__global__ void __launch_bounds__(1024) k_scan(unsigned int* hist, int m) {
    __shared__ unsigned int part[1024];
    const int t = threadIdx.x;
    const int chunk = (m + 1023) / 1024;
    const int b = t * chunk, e = min(m, b + chunk);
    unsigned int s = 0;
    for (int i = b; i < e; ++i) s += hist[i];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const unsigned int v = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    unsigned int run = part[t] - s;
    for (int i = b; i < e; ++i) {
        const unsigned int v = hist[i];
        hist[i] = run;
        run += v;
    }
}

// Stable scatter: keys of a tile are ranked round by round (256 keys per round, in order);
// within a wavefront equal digits are matched with 8 ballots.
__global__ void __launch_bounds__(kSortBlock) k_scatter(int n, int shift, const unsigned long long* __restrict__ kin,
                                                        const int* __restrict__ vin, unsigned long long* __restrict__ kout,
                                                        int* __restrict__ vout, const unsigned int* __restrict__ hist,
                                                        int tiles) {
    __shared__ unsigned int base[256];
    __shared__ unsigned int wcnt[kSortBlock / 64][256];
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    base[t] = hist[t * tiles + blockIdx.x];
    const unsigned long long lt = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    const int tile0 = blockIdx.x * kTile;
    for (int r = 0; r < kSortItems; ++r) {
        for (int d = lane; d < 256; d += 64) wcnt[w][d] = 0;
        const int i = tile0 + r * kSortBlock + t;
        const bool valid = i < n;
        unsigned long long key = 0;
        int val = 0;
        unsigned int dg = 0;
        if (valid) {
            key = kin[i];
            val = vin[i];
            dg = (unsigned int)((key >> shift) & 0xff);
        }
        unsigned long long mask = __ballot(valid);
        for (int bit = 0; bit < 8; ++bit) {
            const unsigned long long bb = __ballot(valid && ((dg >> bit) & 1));
            mask &= ((dg >> bit) & 1) ? bb : ~bb;
        }
        const unsigned int rank = (unsigned int)__popcll(mask & lt);
        const unsigned int cnt = (unsigned int)__popcll(mask);
        __syncthreads();
        if (valid && rank == 0) wcnt[w][dg] = cnt;
        __syncthreads();
        if (valid) {
            unsigned int pre = 0;
            for (int ww = 0; ww < w; ++ww) pre += wcnt[ww][dg];
            const unsigned int pos = base[dg] + pre + rank;
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        {
            unsigned int add = 0;
            for (int ww = 0; ww < kSortBlock / 64; ++ww) add += wcnt[ww][t];
            base[t] += add;
        }
        __syncthreads();
    }
}

__global__ void k_set_perm(int n, const int* __restrict__ sorted_vals, int* perm, int* iperm) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int g = sorted_vals[s];
    perm[s] = g;
    iperm[g] = s;
}

// ---------------------------------------------------------------- Karras tree ---------------

__device__ __forceinline__ int kdelta(const unsigned long long* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const unsigned long long a = k[i], b = k[j];
    if (a == b) return 64 + __clz((unsigned int)(i ^ j));
    return __clzll(a ^ b);
}

__global__ void k_fill(int* p, int count, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

__global__ void k_karras(int n, const unsigned long long* __restrict__ keys, int* kl, int* kr, int* kf, int* kla,
                         int* kpar) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (kdelta(keys, n, i, i + 1) - kdelta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = kdelta(keys, n, i, i - d);
    int lmax = 2;
    while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = kdelta(keys, n, i, j);
    int s = 0;
    for (int div = 2;; div <<= 1) {
        const int t = (l + div - 1) / div;
        if (kdelta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int gamma = i + s * d + min(d, 0);
    const int first = min(i, j), last = max(i, j);
    const int left = (first == gamma) ? ~gamma : gamma;
    const int right = (last == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    kl[i] = left;
    kr[i] = right;
    kf[i] = first;
    kla[i] = last;
    // parent links in the combined numbering: internal c -> c, leaf k -> n-1+k
    kpar[left >= 0 ? left : (n - 1 + ~left)] = i;
    kpar[right >= 0 ? right : (n - 1 + ~right)] = i;
}

// DFS preorder renumbering + threaded links.
__global__ void k_preorder(int n, const int* __restrict__ kl, const int* __restrict__ kf, const int* __restrict__ kla,
                           const int* __restrict__ kpar, float4* nodes, int* parent, int* right, int* leaf_node,
                           int* kpre) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = 2 * n - 1;
    if (x >= total) return;
    const bool internal = x < n - 1;
    const int leafk = internal ? -1 : x - (n - 1);
    const int first = internal ? kf[x] : leafk;
    const int last = internal ? kla[x] : leafk;
    // left turns on the root path
    int lt = 0;
    bool x_is_left = false;
    {
        int c = x;
        int p = kpar[c];
        bool firststep = true;
        for (int guard = 0; p >= 0 && guard < 192; ++guard) {
            const int code = (c < n - 1) ? c : ~(c - (n - 1));
            const bool isleft = kl[p] == code;
            if (firststep) {
                x_is_left = isleft;
                firststep = false;
            }
            lt += isleft ? 1 : 0;
            c = p;
            p = kpar[c];
        }
    }
    const int pre = 2 * first + lt;
    const int size = 2 * (last - first + 1) - 1;
    const int miss = pre + size;  // == total when nothing follows
    nodes[2 * pre].w = __int_as_float(miss);
    nodes[2 * pre + 1].w = __int_as_float(leafk);
    const int p = kpar[x];
    if (p < 0) {
        parent[pre] = -1;
    } else if (x_is_left) {
        parent[pre] = pre - 1;
    } else {
        const int ppre = 2 * kf[p] + lt;  // right edge adds no left turn
        parent[pre] = ppre;
        right[ppre] = pre;
    }
    if (!internal) leaf_node[leafk] = pre;
    else kpre[x] = pre;
}

// ---------------------------------------------------------------- refit ----------------------

// Leaf k's node (preorder index p): its box, miss link p + 1, slot k.
__device__ __forceinline__ void refit_leaf(int k, int p, const float4* __restrict__ leafbox, float4* nodes) {
    const float4 lo = leafbox[2 * k], hi = leafbox[2 * k + 1];
    nodes[2 * p] = make_float4(lo.x, lo.y, lo.z, __int_as_float(p + 1));
    nodes[2 * p + 1] = make_float4(hi.x, hi.y, hi.z, __int_as_float(k));
}

// Internal node q: the union of its children's boxes (left child q + 1, right child right[q]),
// keeping its miss link.
__device__ __forceinline__ void refit_internal(int q, float4* nodes, const int* __restrict__ right) {
    const int a = q + 1, b = right[q];
    const float4 alo = nodes[2 * a], ahi = nodes[2 * a + 1];
    const float4 blo = nodes[2 * b], bhi = nodes[2 * b + 1];
    const float qmiss = nodes[2 * q].w;
    nodes[2 * q] = make_float4(fminf(alo.x, blo.x), fminf(alo.y, blo.y), fminf(alo.z, blo.z), qmiss);
    nodes[2 * q + 1] = make_float4(fmaxf(ahi.x, bhi.x), fmaxf(ahi.y, bhi.y), fmaxf(ahi.z, bhi.z), __int_as_float(-1));
}

__device__ __forceinline__ bool same_bits(float a, float b) { return __float_as_int(a) == __float_as_int(b); }

// Refit of any tree in one launch: each leaf climbs towards the root; the second of two siblings
// to arrive at a parent (a flag per node) computes it and continues.
__global__ void k_refit(int n, const float4* __restrict__ leafbox, float4* nodes, const int* __restrict__ parent,
                        const int* __restrict__ right, int* counter, const int* __restrict__ leaf_node) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int p = leaf_node[k];
    refit_leaf(k, p, leafbox, nodes);
    for (int guard = 0; guard < 192; ++guard) {  // tree depth <= 96 (augmented 96-bit keys)
        const int q = parent[p];
        if (q < 0) break;
        __threadfence();
        if (atomicAdd(&counter[q], 1) == 0) return;  // sibling not finished: it continues
        __threadfence();
        refit_internal(q, nodes, right);
        p = q;
    }
}

// Level-synchronous refit of a binned-SAH tree: the leaves, then one launch per level of internal
// nodes (Karras ids first .. first + count - 1), deepest level first; no flags, no fences.
__global__ void k_refit_leaves(int n, const float4* __restrict__ leafbox, float4* nodes,
                               const int* __restrict__ leaf_node) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) refit_leaf(k, leaf_node[k], leafbox, nodes);
}

__global__ void k_refit_level(int first, int count, const int* __restrict__ kpre, float4* nodes,
                              const int* __restrict__ right) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) refit_internal(kpre[first + i], nodes, right);
}

// Structure and boxes of the tree: links consistent, each leaf's box its Gaussian's, each internal
// box exactly the union of its children's (what every refit computes).
__global__ void k_check(int n, const float4* __restrict__ nodes, const int* __restrict__ right,
                        const int* __restrict__ leaf_node, const float4* __restrict__ leafbox,
                        unsigned long long* errors) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = 2 * n - 1;
    if (p >= total) return;
    const float4 lo = nodes[2 * p], hi = nodes[2 * p + 1];
    const int miss = __float_as_int(lo.w), leaf = __float_as_int(hi.w);
    unsigned long long e = 0;
    if (miss <= p || miss > total) e++;
    if (leaf >= 0) {
        if (leaf >= n || leaf_node[leaf] != p || miss != p + 1) {
            e++;
        } else {
            const float4 blo = leafbox[2 * leaf], bhi = leafbox[2 * leaf + 1];
            const bool same = same_bits(lo.x, blo.x) && same_bits(lo.y, blo.y) && same_bits(lo.z, blo.z) &&
                              same_bits(hi.x, bhi.x) && same_bits(hi.y, bhi.y) && same_bits(hi.z, bhi.z);
            if (!same) e++;
        }
    } else {
        const int a = p + 1, b = right[p];
        if (a >= total || b <= a || b >= total) {
            e++;
        } else {
            const float4 alo = nodes[2 * a], ahi = nodes[2 * a + 1], blo = nodes[2 * b], bhi = nodes[2 * b + 1];
            const bool ok = same_bits(lo.x, fminf(alo.x, blo.x)) && same_bits(lo.y, fminf(alo.y, blo.y)) &&
                            same_bits(lo.z, fminf(alo.z, blo.z)) && same_bits(hi.x, fmaxf(ahi.x, bhi.x)) &&
                            same_bits(hi.y, fmaxf(ahi.y, bhi.y)) && same_bits(hi.z, fmaxf(ahi.z, bhi.z));
            if (!ok) e++;
            if (__float_as_int(ahi.w) < 0 && __float_as_int(alo.w) != b) e++;  // left subtree ends at right
            if (__float_as_int(ahi.w) >= 0 && b != a + 1) e++;
        }
    }
    if (e) atomicAdd(errors, e);
}

// Sum of the node boxes' surface areas (FP64): with the root's area it gives the tree's
// surface-area cost, the expected number of node visits of a random ray that crosses the root.
// Boxes with a negative or non-finite extent count zero.  One FP64 atomic per wavefront.
__global__ void k_tree_area(int total, const float4* __restrict__ nodes, double* out) {
    double a = 0.0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
        const float4 lo = nodes[2 * p], hi = nodes[2 * p + 1];
        const double dx = (double)hi.x - (double)lo.x, dy = (double)hi.y - (double)lo.y,
                     dz = (double)hi.z - (double)lo.z;
        const double area = 2.0 * (dx * dy + dy * dz + dz * dx);
        if (dx >= 0.0 && dy >= 0.0 && dz >= 0.0 && area < INFINITY) a += area;
    }
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, a);
}

inline int blocks(int64_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

// Build: expects leafbox in ORIGINAL order (preprocess run with identity iperm) and the scene box.
// Produces perm/iperm, topology, links.  Boxes are written by a following refit.
int64_t radix_hist_bytes(int n) { return (int64_t)256 * blocks(n, kTile) * sizeof(unsigned int); }

int launch_radix_sort(int n, int passes, unsigned long long* keys[2], int* vals[2], unsigned int* hist,
                      hipStream_t s) {
    const int tiles = blocks(n, kTile);
    int cur = 0;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = pass * 8;
        hipLaunchKernelGGL(k_hist, dim3(tiles), dim3(kSortBlock), 0, s, n, shift, keys[cur], hist, tiles);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, hist, 256 * tiles);
        hipLaunchKernelGGL(k_scatter, dim3(tiles), dim3(kSortBlock), 0, s, n, shift, keys[cur], vals[cur],
                           keys[cur ^ 1], vals[cur ^ 1], hist, tiles);
        cur ^= 1;
    }
    return cur;
}

hipError_t launch_lbvh_build(int n, const float4* leafbox, const float* box, BvhBuffers& b, hipStream_t s) {
    b.levels = 0;  // Karras ids are not grouped by depth: the refit climbs with flags
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_morton, dim3(blocks(n, 256)), dim3(256), 0, s, n, leafbox, box, b.keys[0], b.vals[0]);
    const int cur = launch_radix_sort(n, 8, b.keys, b.vals, b.hist, s);
    const int total = 2 * n - 1;
    hipLaunchKernelGGL(k_fill, dim3(blocks(total, 256)), dim3(256), 0, s, b.k_parent, total, -1);
    if (n > 1)
        hipLaunchKernelGGL(k_karras, dim3(blocks(n - 1, 256)), dim3(256), 0, s, n, b.keys[cur], b.k_left, b.k_right,
                           b.k_first, b.k_last, b.k_parent);
    return launch_lbvh_finish(n, b.vals[cur], b, s);
}

// Slot order and preorder renumbering of a Karras-form topology whose leaves are numbered in
// depth-first order (slot s = leaf s; sorted_vals[s] = original index).
hipError_t launch_lbvh_finish(int n, const int* sorted_vals, BvhBuffers& b, hipStream_t s) {
    const int total = 2 * n - 1;
    hipLaunchKernelGGL(k_set_perm, dim3(blocks(n, 256)), dim3(256), 0, s, n, sorted_vals, b.perm, b.iperm);
    hipLaunchKernelGGL(k_preorder, dim3(blocks(total, 256)), dim3(256), 0, s, n, b.k_left, b.k_first, b.k_last,
                       b.k_parent, b.nodes, b.parent, b.right, b.leaf_node, b.k_pre);
    return hipGetLastError();
}

hipError_t launch_iota(int n, int* perm, int* iperm, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_iota, dim3(blocks(n, 256)), dim3(256), 0, s, n, perm, iperm);
    return hipGetLastError();
}

hipError_t launch_lbvh_refit(int n, const float4* leafbox_sorted, BvhBuffers& b, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (b.levels > 0) {
        hipLaunchKernelGGL(k_refit_leaves, dim3(blocks(n, 256)), dim3(256), 0, s, n, leafbox_sorted, b.nodes,
                           b.leaf_node);
        for (int d = b.levels - 1; d >= 0; --d) {
            const int first = b.level_start[d], count = b.level_start[d + 1] - first;
            if (count > 0)
                hipLaunchKernelGGL(k_refit_level, dim3(blocks(count, 256)), dim3(256), 0, s, first, count, b.k_pre,
                                   b.nodes, b.right);
        }
        return hipGetLastError();
    }
    const int total = 2 * n - 1;
    hipLaunchKernelGGL(k_fill, dim3(blocks(total, 256)), dim3(256), 0, s, b.counter, total, 0);
    hipLaunchKernelGGL(k_refit, dim3(blocks(n, 256)), dim3(256), 0, s, n, leafbox_sorted, b.nodes, b.parent, b.right,
                       b.counter, b.leaf_node);
    return hipGetLastError();
}

hipError_t launch_tree_area(int n, const BvhBuffers& b, double* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int total = 2 * n - 1;
    hipLaunchKernelGGL(k_tree_area, dim3(std::min(blocks(total, 256), 2048)), dim3(256), 0, s, total, b.nodes, out);
    return hipGetLastError();
}

hipError_t launch_lbvh_check(int n, const BvhBuffers& b, const float4* leafbox_sorted, unsigned long long* errors,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int total = 2 * n - 1;
    hipLaunchKernelGGL(k_check, dim3(blocks(total, 256)), dim3(256), 0, s, n, b.nodes, b.right, b.leaf_node,
                       leafbox_sorted, errors);
    return hipGetLastError();
}

}  // namespace rg
