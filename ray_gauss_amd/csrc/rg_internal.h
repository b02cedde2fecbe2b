// Internal layouts shared by the kernels and the C-ABI layer of the gfx950 ray caster.
//
// HBM layout (n Gaussians, V views; "slot" = position in the LBVH's Morton order):
//   perm[slot] -> original index, iperm[orig] -> slot                     int32
//   geo[4*slot+0] = (mu.x, mu.y, mu.z, sigma)                             float4   } 64-B record
//   geo[4*slot+1] = (M'00, M'01, M'02, h)    h = ln|sigma/thr|            float4   } gathered by
//   geo[4*slot+2] = (M'10, M'11, M'12, 0)    M' rows = col_k(R)/(s_k*d)   float4   } traversal and
//   geo[4*slot+3] = (M'20, M'21, M'22, 0)    d = sqrt(2h)                 float4   } sampling
//   brec[2*slot+0] = (s.x, s.y, s.z, d), brec[2*slot+1] = (q.w, q.x, q.y, q.z)   backward only
//   leafbox[2*slot+0..1] = padded AABB (lo.xyz, -), (hi.xyz, -)
//   rgb[v*n + slot] = (r, g, b, 0) per-view colour                           float4
//   nodes[2*p+0] = (lo.xyz, miss), nodes[2*p+1] = (hi.xyz, leaf)  p = DFS preorder index
//     internal node: left child p+1, right child right[p]; leaf: leaf = slot, -1 otherwise;
//     miss = index of the first node after p's subtree (= total when none): a threaded tree,
//     traversed with no stack.
//   acc[16*slot + ...] gradient accumulators: 0-2 d mu, 3 d sigma, 4-6 d s, 8-11 d q(w,x,y,z)
//   drgb[v*n + slot] per-view colour gradient (float4)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/raygauss_hip.h"

namespace rg {

constexpr int kAccStride = 16;

struct DevScene {
    int n;
    int total_nodes;  // 2n-1 (0 when n == 0)
    const float4* geo;
    const float4* brec;
    const float4* nodes;
    const float4* rgb;  // base of the per-view tables
    const float* box;   // scene box, 6 floats (device)
    const int* perm;
};

struct DevParams {
    float dt;
    int spb;
    float t_eps;
    float sigma_thr;
    float t_near;
    float t_far;
    float bg;
    int max_hits;    // >0: count slabs with more hits (reference max_prim_slice)
    int seg_slabs;   // initial segment length
    int packet;      // 1: wave-uniform packet traversal, 0: per-lane traversal
    int jitter;      // sub-pixel jitter of the rays
    int rnd_sample;  // sample offsets: 0 none, 1 random, 2 half a sample
    unsigned int iteration;  // seeds each ray's generator: tea4(iteration, pixel)
    // hit-list records (rg_train_batch): ray r owns words rec[r * rec_words ...]; rec_len[r] = words
    // written (bit 31: the records stop early).  rec_mode 0 off, 1 forward writes, 2 backward reads.
    unsigned int* rec;
    int* rec_len;
    int2* rec_id;  // (view, pixel) of the ray whose records row r holds
    int rec_words;
    int rec_mode;
    // launch structure of the ray kernels: ray_chunk 0 = one 8x8 tile per wavefront, grid-launched,
    // tiles mapped to XCDs in contiguous bands; > 0 = persistent wavefronts taking ray_chunk rays
    // at a time from per-XCD work queues (queue: 8 device counters, stealing from the other XCDs'
    // queues when their own is empty); refill_lanes (forward, persistent): 0 = a wavefront takes new
    // rays when all its lanes are done, k = as soon as k lanes are idle
    int ray_chunk;
    int refill_lanes;
    unsigned long long* queue;
    // backward depth synchronisation, in slabs (lanes within it of the wave's laggard step together)
    float bwd_window;
    // forward: step over slabs empty for every lane of a wavefront together (jump_empty)
    int fwd_jump;
};

// Gradient accumulators (device-scope atomic adds; the finish kernel reads and clears them).
struct DevGrad {
    float* acc;     // 16 floats per slot
    float4* drgb;   // views x slots
    int num_views;  // rows of drgb
};

// Counters (device, 16): 0 K, 1 P, 2 overflow slabs, 3 fallback slabs, 4 gathers, 6 loss (float),
// 7 debug scratch, 8 record words, 9 truncated records, 10 dropped by bounds checks, 11 rays with
// an out-of-range view or pixel, 12..15 unused
struct DevStats {
    unsigned long long* c;
};

// ---- launchers (preprocess.hip) ---------------------------------------------------------------
hipError_t launch_init_box(float* box, hipStream_t s);
hipError_t launch_preprocess(const rg_gaussians& g, const int* iperm, float sigma_thr, float4* geo, float4* brec,
                             float4* leafbox, float* box, hipStream_t s);
hipError_t launch_sh_forward(const rg_gaussians& g, const int* iperm, const rg_camera* cams_dev, int num_views,
                             float4* rgb, hipStream_t s);
hipError_t launch_finish_grads(const rg_gaussians& g, const int* iperm, const rg_camera* cams_dev, int num_views,
                               const DevGrad& gr, int exact, const rg_grads& out, hipStream_t s);
hipError_t launch_zero(float* p, int64_t count, hipStream_t s);
hipError_t launch_set_colors(int n, const float* rgb, const int* iperm, float4* table, hipStream_t s);
hipError_t launch_rgba8(int64_t n, const float* rgb, uchar4* out, hipStream_t s);
hipError_t launch_export_predicate(int n, const int* iperm, const float4* geo, float* mu, float* Mp, int* valid,
                                   float* hlog, hipStream_t s);

// ---- launchers (lbvh.hip) ---------------------------------------------------------------------
// Depth bound of the binned-SAH build (its level loop fails beyond it).
constexpr int kTreeMaxLevels = 201;
struct BvhBuffers {
    unsigned long long* keys[2];
    int* vals[2];
    unsigned int* hist;     // radix histogram, 256 * tiles
    int* perm;
    int* iperm;
    float4* nodes;
    int* parent;            // preorder
    int* right;             // preorder
    int* counter;           // refit flags
    int* leaf_node;         // slot -> preorder index of its leaf
    int* k_parent;          // Karras numbering scratch: parent of internal (n-1) and leaves (n)
    int* k_left;            // Karras: left child code, right child code, range first/last
    int* k_right;
    int* k_first;
    int* k_last;
    float4* sah_prims[2];   // binned-SAH build: (box, original index) records, ping-pong
    float4* sah_tasks[2];   // binned-SAH build: node tasks of one level, ping-pong
    int* sah_ctr;           // binned-SAH build: counters
    int* k_pre;             // Karras internal node id -> preorder index
    // Binned-SAH trees number their internal nodes level by level: the nodes of depth d have the
    // Karras ids [level_start[d], level_start[d + 1]), so a refit runs one launch per level,
    // deepest first.  levels == 0 for an LBVH, whose refit climbs from the leaves with a flag per
    // node.
    int levels = 0;
    int level_start[kTreeMaxLevels + 2] = {};
};
hipError_t launch_lbvh_build(int n, const float4* leafbox, const float* box, BvhBuffers& b, hipStream_t s);
// Binned-SAH build (sah.hip): same outputs as launch_lbvh_build (geo: the preprocessed records in
// original order, for the means of Gaussians without a box); synchronises the stream once per
// tree level.  sah_task_bytes(n): size of each sah_tasks array.
hipError_t launch_sah_build(int n, const float4* leafbox, const float4* geo, const float* box, BvhBuffers& b,
                            hipStream_t s);
int64_t sah_task_bytes(int n);
hipError_t launch_lbvh_finish(int n, const int* sorted_vals, BvhBuffers& b, hipStream_t s);
hipError_t launch_lbvh_refit(int n, const float4* leafbox_sorted, BvhBuffers& b, hipStream_t s);
hipError_t launch_lbvh_check(int n, const BvhBuffers& b, const float4* leafbox_sorted, unsigned long long* errors,
                             hipStream_t s);
// Sum of the node boxes' surface areas, FP64, added (atomically) to *out.
hipError_t launch_tree_area(int n, const BvhBuffers& b, double* out, hipStream_t s);
hipError_t launch_iota(int n, int* perm, int* iperm, hipStream_t s);
// Stable LSD radix sort of (keys, vals) over the low 8*passes bits; returns the index of the
// buffer pair holding the result.  hist needs radix_hist_bytes(n).
int launch_radix_sort(int n, int passes, unsigned long long* keys[2], int* vals[2], unsigned int* hist,
                      hipStream_t s);
int64_t radix_hist_bytes(int n);

// ---- launchers (optim.hip) ---------------------------------------------------------------------
constexpr int kMaxAdamTensors = 16;
// Per-tensor scalars of one Adam step, rounded to float once on the host.
struct AdamScalars {
    float w1;        // 1 - beta1
    float b2;        // beta2
    float w2;        // 1 - beta2
    float bc2_sqrt;  // sqrt(1 - beta2^step)
    float eps;
    float neg_step;  // -lr / (1 - beta1^step)
};
struct AdamTensor {
    float* param;
    const float* grad;
    float* exp_avg;
    float* exp_avg_sq;
    int64_t n;
    int aligned;  // all four arrays 16-B aligned
};
// Work items are float4 groups: tensor k owns items [item_start[k], item_start[k+1]).
struct AdamBatch {
    AdamTensor t[kMaxAdamTensors];
    AdamScalars s[kMaxAdamTensors];
    int64_t item_start[kMaxAdamTensors + 1];
    int count;
};
hipError_t launch_adam(const AdamBatch& b, int64_t total_items, hipStream_t s);
hipError_t launch_project(const rg_projection& pr, hipStream_t s);

// ---- launchers (densify.hip) -------------------------------------------------------------------
int64_t knn_workspace_bytes(int64_t n);
hipError_t launch_knn_mean_dist(const float* points, int64_t n, float* out, void* ws, hipStream_t s);
// The same with the workspace in a static per-device scratch array of knn_scratch_bytes() bytes.
int64_t knn_scratch_bytes();
hipError_t launch_knn_mean_dist_noalloc(const float* points, int64_t n, float* out, hipStream_t s);
hipError_t launch_filter_3d(const float* points, int64_t n, const rg_filter_camera* cams_dev, int V, float inv_focal,
                            float* out, unsigned int* dmax_bits, hipStream_t s);
hipError_t launch_accum_grad(const float* g, int64_t n, float* accum, float* num, float* gnv, hipStream_t s);
int64_t densify_workspace_bytes(int64_t n);
hipError_t launch_densify_plan(const rg_densify_in& in, void* ws, int64_t* counts, hipStream_t s);
hipError_t launch_densify_apply(const rg_densify_in& in, const void* ws, const int64_t* counts,
                                const rg_densify_array* arrays, int num_arrays, const float* normals, float* out_pos,
                                float* out_ls, float* out_sig, hipStream_t s);
int densify_max_arrays();

// ---- launchers (loss.hip) ----------------------------------------------------------------------
int64_t loss_workspace_bytes(int W, int H);
hipError_t launch_loss_l1_ssim(const float* rays, const float* target, int W, int H, int fx, int fy,
                               const float* window11, float lambda, float* loss, float* parts, float* dl_rays,
                               void* ws, hipStream_t s);

// ---- launchers (trace.hip) --------------------------------------------------------------------
struct RayJob {
    int64_t num_rays;
    const int32_t* pixels;     // single-view mode (nullable)
    const int32_t* ray_view;   // batch mode (nullable)
    const int32_t* ray_pixel;  // batch mode
    int view;                  // single-view mode
    const rg_camera* cams;     // device array
    int num_views;
};
hipError_t launch_forward(const DevScene& sc, const DevParams& p, const RayJob& job, int hit_depth, float* out_rgb,
                          float* out_depth, float* out_opacity, DevStats st, hipStream_t s);
hipError_t launch_backward(const DevScene& sc, const DevParams& p, const RayJob& job, int hit_depth,
                           const float* ray_colors, const float* dLdC, const float* target, int loss_kind,
                           int count_stats, DevGrad gr, DevStats st, hipStream_t s);
hipError_t launch_reset_counters(DevStats st, hipStream_t s);
hipError_t launch_debug_hits(const DevScene& sc, const DevParams& p, const RayJob& job, int max_slabs, int max_pairs,
                             int32_t* pairs, int64_t* counts, float* ro, float* rd, float* te, float* tx,
                             hipStream_t s);

}  // namespace rg
