// Binned-SAH build of the BVH over the per-Gaussian bounding-ellipsoid AABBs: an alternative to
// the Morton-order LBVH (lbvh.hip) that replaces the same OptiX GAS build (optix_utils.py:68-75).
// It costs more to build and its tree needs fewer node visits per ray segment (the LBVH orders
// leaves by centre only, so a large box deep in the tree widens every box above it).
//
//   top down, one tree level per launch (one 256-thread workgroup per node with two or more
//   Gaussians): the Gaussians of a node occupy a contiguous range of a ping-pong array of
//   (box, original index) records; the workgroup bins their box centres into 16 bins per axis
//   (LDS, four private copies per wavefront), evaluates the 45 bin-boundary splits by the surface
//   area heuristic, and partitions the range stably (left part first) into the other array.
//   Ranges nest, so the final array order is the depth-first leaf order: leaf slot = position.
//   Nodes deeper than kSahMaxDepth, and sets whose centres do not spread, split at the median,
//   which bounds the depth by kSahMaxDepth + log2(n).  Gaussians without a bounding ellipsoid
//   (empty boxes) are placed as points at their means.
//
// Output is the Karras-form topology that lbvh.hip's preorder renumbering consumes (left/right
// child codes, leaf range, parents) plus the slot order; refit and traversal are shared.
#include <hip/hip_runtime.h>

#include "rg_internal.h"

namespace rg {
namespace {

constexpr int kBins = 16;
constexpr int kSahBlock = 256;
constexpr int kSahWaves = kSahBlock / 64;
constexpr int kCopies = 4;  // private bin copies per wavefront (lanes 16c..16c+15)
constexpr int kSahMaxDepth = 48;

struct SahTask {
    int first, count, node, depth;
    int median;  // 1: split at the median
    int pad[3];
    float cmin[4], cmax[4];
};

inline int blocks(int64_t n, int b) { return (int)((n + b - 1) / b); }

__global__ void k_sah_fill(int* p, int count, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

__device__ __forceinline__ int bin_of(float c, float cmin, float scale) {
    const float f = (c - cmin) * scale;
    return (int)fminf(fmaxf(f, 0.0f), (float)(kBins - 1));  // NaN -> 0
}

__device__ __forceinline__ float half_area(float lx, float ly, float lz, float hx, float hy, float hz) {
    const float x = hx - lx, y = hy - ly, z = hz - lz;
    return x * y + y * z + z * x;
}

// (box, original index) records in original order.  A Gaussian without a bounding ellipsoid (an
// empty box, never hit) is placed by its mean as a point: it then sits in the tree next to its
// neighbours, where its leaf is reached only when they are, instead of in a subtree of its own
// whose empty boxes every traversal would have to enter.
__global__ void k_sah_records(int n, const float4* __restrict__ leafbox, const float4* __restrict__ geo,
                              float4* prims) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 lo = leafbox[2 * i], hi = leafbox[2 * i + 1];
    if (!(lo.x <= hi.x)) {
        const float4 mu = geo[4 * i];
        lo = make_float4(mu.x, mu.y, mu.z, 0.0f);
        hi = lo;
    }
    lo.w = __int_as_float(i);
    hi.w = 0.0f;
    prims[2 * i] = lo;
    prims[2 * i + 1] = hi;
}

__device__ void push_child(int first, int count, int parent_depth, int median, const float* cmin, const float* cmax,
                           int n, SahTask* tasks, int max_tasks, int* ctr, int* kpar, int parent, int* code,
                           int* vals_out, const float4* prims_out) {
    // ctr: [1] next internal node id, [2] tasks of the next level, [3] error flags
    if (count == 1) {
        *code = ~first;
        kpar[n - 1 + first] = parent;
        vals_out[first] = __float_as_int(prims_out[2 * first].w);
        return;
    }
    const int id = atomicAdd(&ctr[1], 1);
    const int t = atomicAdd(&ctr[2], 1);
    if (id >= n - 1 || t >= max_tasks) {
        atomicOr(&ctr[3], 1);
        *code = ~first;
        return;
    }
    *code = id;
    kpar[id] = parent;
    SahTask k;
    k.first = first;
    k.count = count;
    k.node = id;
    k.depth = parent_depth + 1;
    k.median = (median || k.depth >= kSahMaxDepth) ? 1 : 0;
    k.pad[0] = k.pad[1] = k.pad[2] = 0;
    for (int a = 0; a < 3; ++a) {
        k.cmin[a] = cmin[a];
        k.cmax[a] = cmax[a];
    }
    k.cmin[3] = k.cmax[3] = 0.0f;
    tasks[t] = k;
}

// Level 0: the root over all n Gaussians; its centre bounds are the scene box (bins clamp
// centres outside it).
__global__ void k_sah_root(int n, const float* __restrict__ box, SahTask* tasks, int* ctr) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    ctr[0] = 0;
    ctr[1] = 1;
    ctr[2] = 1;
    ctr[3] = 0;
    SahTask k;
    k.first = 0;
    k.count = n;
    k.node = 0;
    k.depth = 0;
    k.median = 0;
    k.pad[0] = k.pad[1] = k.pad[2] = 0;
    for (int a = 0; a < 3; ++a) {
        k.cmin[a] = box[a];
        k.cmax[a] = box[3 + a];
    }
    k.cmin[3] = k.cmax[3] = 0.0f;
    tasks[0] = k;
}

__device__ __forceinline__ void lds_min(float* p, float v) {
    __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_max(float* p, float v) {
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__global__ void __launch_bounds__(kSahBlock) k_sah_level(int n, const SahTask* __restrict__ tasks_in,
                                                         const float4* __restrict__ prims_in, float4* prims_out,
                                                         SahTask* tasks_out, int max_tasks, int* ctr, int* kl,
                                                         int* kr, int* kf, int* kla, int* kpar, int* vals_out) {
    // bins: [copy][axis][bin] with count and box; copy = wave * kCopies + lane / 16
    __shared__ int s_cnt[kSahWaves * kCopies][3][kBins];
    __shared__ float s_lo[kSahWaves * kCopies][3][kBins][3];
    __shared__ float s_hi[kSahWaves * kCopies][3][kBins][3];
    __shared__ float s_cost[kSahBlock];
    __shared__ int s_nl[kSahBlock];
    __shared__ float s_child[2][6];  // centroid bounds of the children: min xyz, max xyz
    __shared__ int s_wave_l[kSahWaves], s_wave_r[kSahWaves];
    __shared__ int s_split[3];       // axis, bin, left count (axis -1: median)

    const SahTask t = tasks_in[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int first = t.first, count = t.count;
    float scale[3];
    for (int a = 0; a < 3; ++a) {
        const float e = t.cmax[a] - t.cmin[a];
        scale[a] = (e > 0.0f && !t.median) ? (float)kBins / e : 0.0f;
    }
    const bool any_axis = scale[0] > 0.0f || scale[1] > 0.0f || scale[2] > 0.0f;
    // private bin copies: as many as the set is large enough to use
    const int ncp = count >= 1024 ? kSahWaves * kCopies : (count >= 256 ? kSahWaves : 1);
    for (int i = tid; any_axis && i < ncp * 3 * kBins; i += kSahBlock) {
        const int c = i / (3 * kBins), r = i % (3 * kBins), a = r / kBins, b = r % kBins;
        s_cnt[c][a][b] = 0;
        for (int d = 0; d < 3; ++d) {
            s_lo[c][a][b][d] = INFINITY;
            s_hi[c][a][b][d] = -INFINITY;
        }
    }
    if (tid < 12) (&s_child[0][0])[tid] = (tid % 6) < 3 ? INFINITY : -INFINITY;
    __syncthreads();

    // ---- binning ----
    if (any_axis) {
        const int cp = ncp == 1 ? 0 : (ncp == kSahWaves ? wave : wave * kCopies + (lane >> 4));
        for (int i = tid; i < count; i += kSahBlock) {
            const float4 lo = prims_in[2 * (first + i)], hi = prims_in[2 * (first + i) + 1];
            const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
            for (int a = 0; a < 3; ++a) {
                if (scale[a] == 0.0f) continue;
                const int b = bin_of(c[a], t.cmin[a], scale[a]);
                atomicAdd(&s_cnt[cp][a][b], 1);
                lds_min(&s_lo[cp][a][b][0], lo.x);
                lds_min(&s_lo[cp][a][b][1], lo.y);
                lds_min(&s_lo[cp][a][b][2], lo.z);
                lds_max(&s_hi[cp][a][b][0], hi.x);
                lds_max(&s_hi[cp][a][b][1], hi.y);
                lds_max(&s_hi[cp][a][b][2], hi.z);
            }
        }
    }
    __syncthreads();
    // merge the copies into copy 0 (thread per (axis, bin))
    if (any_axis && tid < 3 * kBins) {
        const int a = tid / kBins, b = tid % kBins;
        int cnt = 0;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int c = 0; c < ncp; ++c) {
            cnt += s_cnt[c][a][b];
            for (int d = 0; d < 3; ++d) {
                lo[d] = fminf(lo[d], s_lo[c][a][b][d]);
                hi[d] = fmaxf(hi[d], s_hi[c][a][b][d]);
            }
        }
        s_cnt[0][a][b] = cnt;
        for (int d = 0; d < 3; ++d) {
            s_lo[0][a][b][d] = lo[d];
            s_hi[0][a][b][d] = hi[d];
        }
    }
    __syncthreads();

    // ---- split evaluation: thread j < 45 tests axis j/15, boundary j%15+1 ----
    {
        float cost = INFINITY;
        int nl = 0;
        if (any_axis && tid < 3 * (kBins - 1)) {
            const int a = tid / (kBins - 1), s = tid % (kBins - 1) + 1;
            if (scale[a] > 0.0f) {
                float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
                float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
                int nr = 0;
                for (int b = 0; b < kBins; ++b) {
                    const int c = s_cnt[0][a][b];
                    if (c == 0) continue;
                    float* lo = b < s ? llo : rlo;
                    float* hi = b < s ? lhi : rhi;
                    for (int d = 0; d < 3; ++d) {
                        lo[d] = fminf(lo[d], s_lo[0][a][b][d]);
                        hi[d] = fmaxf(hi[d], s_hi[0][a][b][d]);
                    }
                    if (b < s) nl += c;
                    else nr += c;
                }
                if (nl > 0 && nr > 0)
                    cost = (float)nl * half_area(llo[0], llo[1], llo[2], lhi[0], lhi[1], lhi[2]) +
                           (float)nr * half_area(rlo[0], rlo[1], rlo[2], rhi[0], rhi[1], rhi[2]);
                if (!(cost < INFINITY)) cost = INFINITY;  // NaN areas never win
            }
        }
        s_cost[tid] = cost;
        s_nl[tid] = nl;
    }
    __syncthreads();
    if (tid == 0) {
        int best = -1;
        float bc = INFINITY;
        for (int j = 0; j < 3 * (kBins - 1); ++j)
            if (s_cost[j] < bc) {
                bc = s_cost[j];
                best = j;
            }
        if (best < 0) {
            s_split[0] = -1;
            s_split[1] = 0;
            s_split[2] = count >> 1;
        } else {
            s_split[0] = best / (kBins - 1);
            s_split[1] = best % (kBins - 1) + 1;
            s_split[2] = s_nl[best];
        }
    }
    __syncthreads();
    const int axis = s_split[0], sbin = s_split[1], nl = s_split[2];
    const int nr = count - nl;

    // ---- stable partition into prims_out (left part first) ----
    int lbase = 0, rbase = 0;
    for (int base = 0; base < count; base += kSahBlock) {
        const int i = base + tid;
        const bool in = i < count;
        float4 lo = make_float4(0.f, 0.f, 0.f, 0.f), hi = lo;
        bool left = false;
        if (in) {
            lo = prims_in[2 * (first + i)];
            hi = prims_in[2 * (first + i) + 1];
            if (axis < 0) {
                left = i < nl;
            } else {
                const float c = axis == 0 ? 0.5f * (lo.x + hi.x) : axis == 1 ? 0.5f * (lo.y + hi.y) : 0.5f * (lo.z + hi.z);
                left = bin_of(c, t.cmin[axis], scale[axis]) < sbin;
            }
        }
        const unsigned long long bl = __builtin_amdgcn_ballot_w64(in && left);
        const unsigned long long br = __builtin_amdgcn_ballot_w64(in && !left);
        if (lane == 0) {
            s_wave_l[wave] = __popcll(bl);
            s_wave_r[wave] = __popcll(br);
        }
        __syncthreads();
        int lo_off = lbase, ro_off = rbase;
        for (int w = 0; w < wave; ++w) {
            lo_off += s_wave_l[w];
            ro_off += s_wave_r[w];
        }
        int add_l = 0, add_r = 0;
        for (int w = 0; w < kSahWaves; ++w) {
            add_l += s_wave_l[w];
            add_r += s_wave_r[w];
        }
        const unsigned long long lt = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
        if (in) {
            const int dest = left ? first + lo_off + __popcll(bl & lt) : first + nl + ro_off + __popcll(br & lt);
            const bool ok = left ? (lo_off + __popcll(bl & lt) < nl) : (ro_off + __popcll(br & lt) < nr);
            if (ok) {
                prims_out[2 * dest] = lo;
                prims_out[2 * dest + 1] = hi;
                const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
                float* cb = s_child[left ? 0 : 1];
                if (!t.median) {
                    for (int d = 0; d < 3; ++d) {
                        lds_min(&cb[d], c[d]);
                        lds_max(&cb[3 + d], c[d]);
                    }
                }
            } else {
                atomicOr(&ctr[3], 2);  // bin counts and partition disagree (never expected)
            }
        }
        lbase += add_l;
        rbase += add_r;
        __syncthreads();
    }
    if (lbase != nl && tid == 0) atomicOr(&ctr[3], 2);
    __syncthreads();

    if (tid == 0) {
        kf[t.node] = first;
        kla[t.node] = first + count - 1;
        float lmn[3], lmx[3], rmn[3], rmx[3];
        for (int d = 0; d < 3; ++d) {
            lmn[d] = t.median ? t.cmin[d] : s_child[0][d];
            lmx[d] = t.median ? t.cmax[d] : s_child[0][3 + d];
            rmn[d] = t.median ? t.cmin[d] : s_child[1][d];
            rmx[d] = t.median ? t.cmax[d] : s_child[1][3 + d];
        }
        int lc, rc;
        push_child(first, nl, t.depth, t.median, lmn, lmx, n, tasks_out, max_tasks, ctr, kpar, t.node, &lc, vals_out,
                   prims_out);
        push_child(first + nl, nr, t.depth, t.median, rmn, rmx, n, tasks_out, max_tasks, ctr, kpar, t.node, &rc,
                   vals_out, prims_out);
        kl[t.node] = lc;
        kr[t.node] = rc;
    }
}

}  // namespace

int64_t sah_task_capacity(int n) { return (int64_t)n / 2 + 2; }
int64_t sah_task_bytes(int n) { return sah_task_capacity(n) * (int64_t)sizeof(SahTask); }

// Binned-SAH build into the Karras-form arrays of b (then the shared preorder renumbering), with
// the slot order in b.perm / b.iperm.  Reads one counter back per tree level (synchronises the
// stream): not for graph capture.  Internal node ids are handed out level by level, so the
// counter read after each level also gives the id range of the next level (b.levels,
// b.level_start: the refit's launches).  Returns hipErrorUnknown when a consistency check failed.
hipError_t launch_sah_build(int n, const float4* leafbox, const float4* geo, const float* box, BvhBuffers& b,
                            hipStream_t s) {
    if (n <= 1) return launch_lbvh_build(n, leafbox, box, b, s);
    b.levels = 0;
    const int total = 2 * n - 1;
    const int max_tasks = (int)sah_task_capacity(n);
    hipLaunchKernelGGL(k_sah_fill, dim3(blocks(total, 256)), dim3(256), 0, s, b.k_parent, total, -1);
    float4* prims[2] = {b.sah_prims[0], b.sah_prims[1]};
    SahTask* tasks[2] = {reinterpret_cast<SahTask*>(b.sah_tasks[0]), reinterpret_cast<SahTask*>(b.sah_tasks[1])};
    int* ctr = b.sah_ctr;
    int* vals_out = b.vals[0];  // slot -> original index
    hipLaunchKernelGGL(k_sah_records, dim3(blocks(n, 256)), dim3(256), 0, s, n, leafbox, geo, prims[0]);
    hipLaunchKernelGGL(k_sah_root, dim3(1), dim3(64), 0, s, n, box, tasks[0], ctr);
    int host[4] = {0, 0, 0, 0};
    hipError_t e;
    if ((e = hipMemcpyAsync(host, ctr, sizeof host, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    int pp = 0, levels = 0;
    b.level_start[0] = 0;
    b.level_start[1] = host[1];  // the root
    for (int level = 0; host[2] > 0; ++level) {
        if (level >= kTreeMaxLevels || host[3] != 0) return hipErrorUnknown;
        const int ntasks = host[2];
        hipLaunchKernelGGL(k_sah_fill, dim3(1), dim3(64), 0, s, ctr + 2, 1, 0);
        hipLaunchKernelGGL(k_sah_level, dim3(ntasks), dim3(kSahBlock), 0, s, n, tasks[pp], prims[pp], prims[pp ^ 1],
                           tasks[pp ^ 1], max_tasks, ctr, b.k_left, b.k_right, b.k_first, b.k_last, b.k_parent,
                           vals_out);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // the untouched (finished) ranges of this level's input stay valid in the input array
        // only; leaves are written to vals_out when they are made, so nothing else is needed
        if ((e = hipMemcpyAsync(host, ctr, sizeof host, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
        b.level_start[level + 2] = host[1];  // ids made by this level: the next level's nodes
        levels = level + 1;
        pp ^= 1;
    }
    if (host[3] != 0 || host[1] != n - 1) return hipErrorUnknown;
    if ((e = launch_lbvh_finish(n, vals_out, b, s)) != hipSuccess) return e;
    b.levels = levels;
    return hipSuccess;
}

}  // namespace rg
