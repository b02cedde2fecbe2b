/*
 * raygauss_hip.h -- C ABI of the MI355X (gfx950) differentiable Gaussian ray caster.
 *
 * This library replaces, for the RayGauss training path, the OptiX/CuPy machinery behind
 * RenderOptixFunction (reference optix_raycasting/render_optix.py:90-190).  Every entry point
 * below names the reference interface it stands in for.  The ABI uses plain pointers and sizes
 * only: all array pointers are DEVICE pointers (hipMalloc'd or torch-allocated), cameras are
 * host structs, `stream` is a hipStream_t passed as void* (NULL = default stream).
 *
 * Error handling: every function returns 0 on success and a negative rg_status otherwise; the
 * message is available from rg_last_error(handle).  Nothing throws across the ABI.
 * Threading: a handle is used from one host thread at a time (one handle per stream).
 * Ownership: the caller owns inputs and outputs; the handle owns its workspace (per-Gaussian
 * records, LBVH, per-view colour tables, gradient accumulators), grown on demand.
 */
#ifndef RAYGAUSS_HIP_H
#define RAYGAUSS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rg_status {
    RG_OK = 0,
    RG_ERR_ARG = -1,       /* bad argument (null pointer, size, unsupported degree ...) */
    RG_ERR_HIP = -2,       /* HIP runtime error */
    RG_ERR_OOM = -3,       /* device allocation failed */
    RG_ERR_STATE = -4,     /* call order (e.g. forward before prepare) */
    RG_ERR_OVERFLOW = -5   /* max_hits_per_slab > 0 and a slab exceeded it */
} rg_status;

typedef struct rg_handle rg_handle;

/* Runtime form of the reference's compile-time constants
 * (optix_raycasting/cuda_train/forward/gaussians_aabb.h:2-5) plus launch tuning. */
typedef struct rg_config {
    float dt;                  /* sample spacing                 (reference DT = 0.0025)       */
    int32_t samples_per_slab;  /* samples per slab, 1..16         (reference BUFFER_SIZE = 8)   */
    float t_eps;               /* transmittance cut-off          (TRANSMITTANCE_EPSILON 0.003) */
    float sigma_thr;           /* density threshold              (SIGMA_THRESHOLD = 0.1)       */
    float t_near;              /* ray interval clip; 0 = reference                               */
    float t_far;               /* ray interval clip; +inf = reference                            */
    int32_t white_background;  /* RenderOptixSettings.white_background (render_optix.py:70)     */
    int32_t exact_sh_grad;     /* 0: backward_sh as the reference (utils/sh_utils.cu:254)        */
    int32_t hit_depth;         /* per-ray interval list capacity: 8, 16 or 32                    */
    int32_t seg_slabs;         /* initial slabs per interval gather: 1..4096                     */
    int32_t max_hits_per_slab; /* >0: count slabs whose hit set exceeds it (max_prim_slice)      */
    int32_t packet_traversal;  /* 1: the lanes of a wavefront walk the LBVH together (one node
                                  sequence, per-lane box tests); 0: per-lane walks                */
    int32_t jitter;            /* RenderOptixSettings.jitter: sub-pixel jitter (gaussians_aabb.cu:409) */
    int32_t rnd_sample;        /* RenderOptixSettings.rnd_sample: 0 grid, 1 random, 2 mid-sample     */
    uint32_t iteration;        /* RenderOptixSettings.iteration: seeds tea<4>(iteration, pixel)       */
    int32_t fwd_hit_depth;     /* list capacity of the forward kernel: 6, 8 (as 12), 12, 16, 32
                                  (0: hit_depth); a smaller list lets more wavefronts share a CU's
                                  LDS (12 bytes per entry and lane)                             */
    int32_t replay_words;      /* words per ray, a multiple of 8 (0: off), in which a forward records the hit lists
                                  it gathered and returns a token for them; a backward given
                                  that token (no rg_prepare / rg_set_views / rg_set_config /
                                  other recording forward in between) reads them instead of
                                  walking the LBVH again, for each ray whose (view, pixel) is
                                  the one recorded in its output row; rays whose lists did not
                                  fit walk the tree for the rest                               */
    int32_t tree_builder;      /* BVH build at rg_prepare(rebuild): 0 LBVH (Morton order, a few
                                  ms at 1M), 1 binned SAH (tens of ms; fewer node visits per
                                  ray segment; reads back one counter per tree level, so not
                                  under stream capture)                                       */
    int32_t ray_chunk;         /* launch structure of the forward kernel: 0 = one 8x8 pixel tile
                                  per wavefront, grid launch; a multiple of 64 in 64..65536 =
                                  persistent wavefronts (as many as the GPU holds) that take
                                  ray_chunk rays at a time from per-XCD device work queues,
                                  stealing from the other XCDs' queues when their own is empty */
    int32_t refill_lanes;      /* persistent forward: 0 = a wavefront takes new rays when all its
                                  lanes are done; 1..63 = as soon as that many lanes are idle   */
    int32_t bwd_ray_chunk;     /* launch structure of the backward kernel, as ray_chunk         */
    float bwd_window;          /* backward: lanes within this many slabs of the wave's laggard
                                  prepare their next slab together and merge their gradients
                                  (<= 0: the default, 0.5; the best value depends on the scene) */
    int32_t fwd_round_per_slab; /* forward: 0 (default) = after a round in which no lane of a
                                  wavefront had a hit, the lanes stepping through a list step
                                  together over the slabs empty for them (the backward always
                                  does); 1 = one round per slab                                */
} rg_config;

/* One pinhole view: the launch record of optix_utils.py:166-181. u, v, w are columns 0..2 of
 * world_view_transform[:3,:3]; tanx/tany = tan(FoV/2). */
typedef struct rg_camera {
    float eye[3];
    float u[3];
    float v[3];
    float w[3];
    float tanx;
    float tany;
    int32_t width;
    int32_t height;
} rg_camera;

/* The boundary inputs of RenderOptixFunction.forward (render_optix.py:92-93), already activated
 * (classes/point_cloud.py:729-739).  Layouts: positions (n,3), scales (n,3), quats (n,4) w,x,y,z,
 * densities (n), sh (n,3,K) K=(deg+1)^2, sg (n,3,L), sharpness (n,L), lobe_axis (n,L,3). */
typedef struct rg_gaussians {
    int64_t n;
    int32_t sh_degree;  /* 0..3 */
    int32_t num_sg;     /* L >= 0; sg/sharpness/lobe_axis may be NULL when 0 */
    const float* positions;
    const float* scales;
    const float* quats;
    const float* densities;
    const float* sh;
    const float* sg;
    const float* sharpness;
    const float* lobe_axis;
} rg_gaussians;

/* Gradient outputs, original Gaussian order, same layouts as rg_gaussians (overwritten).
 * The sg/sharpness/lobe_axis pointers may be NULL when num_sg == 0. */
typedef struct rg_grads {
    float* positions;
    float* scales;
    float* quats;
    float* densities;
    float* sh;
    float* sg;
    float* sharpness;
    float* lobe_axis;
} rg_grads;

/* A batch of rays over the views given to rg_set_views: ray r is pixel ray_pixel[r] of view
 * ray_view[r] (device int32 arrays).  This is the multi-view generalisation of the reference's
 * one-camera launch (optix_utils.py:132,201). */
typedef struct rg_ray_batch {
    int64_t num_rays;
    const int32_t* ray_view;
    const int32_t* ray_pixel;
} rg_ray_batch;

/* Per-call counters (device-side, read back by rg_get_stats). */
typedef struct rg_stats {
    int64_t slab_hits;        /* K: sum over processed slabs of |hit set|                     */
    int64_t pairs;            /* P: distinct (ray, Gaussian) pairs                            */
    int64_t overflow_slabs;   /* slabs with more hits than max_hits_per_slab (optix_utils:205) */
    int64_t fallback_slabs;   /* slabs whose hit set exceeded hit_depth (per-slab traversal)   */
    int64_t gathers;          /* interval-list gathers (BVH traversals)                       */
    int64_t replay_words;     /* words of hit-list records written by the forward (rg_train_batch) */
    int64_t replay_truncated; /* rays whose records did not fit replay_words                    */
    int64_t dropped;          /* gradient records or replayed hits skipped by a bounds check:
                                 an internal error, never expected (tests assert zero)         */
    int64_t bad_rays;         /* batch / pixel-list rays whose view or pixel index is out of
                                 range: skipped (forward output NaN, no gradient)              */
} rg_stats;

/* Creates a handle.  Replaces create_context/module/pipeline (optix_utils.py:77-110): the
 * kernels are compiled ahead of time for gfx950. */
int rg_create(rg_handle** out, const rg_config* cfg);
int rg_destroy(rg_handle* h);
int rg_set_config(rg_handle* h, const rg_config* cfg);
const char* rg_last_error(const rg_handle* h);

/* Per-Gaussian preprocessing and acceleration structure.  Replaces quaternion_to_rotation +
 * compute_ellipsoids_bbox + the scene-box reductions + create/update_acceleration_structure
 * (optix_utils.py:34-75, render_optix.py:103-110).  rebuild != 0 (or a change of n) rebuilds
 * the LBVH (Morton codes, radix sort, radix tree); otherwise the existing tree is refit. */
int rg_prepare(rg_handle* h, const rg_gaussians* g, int32_t rebuild, void* stream);

/* Uploads the views and computes each view's per-Gaussian colour table (num_views * n must stay
 * below 2^31 - 1: RG_ERR_ARG otherwise).  Replaces
 * compute_cupy_rgb / forward_sh (render_optix.py:19-32, utils/sh_utils.cu:25-119) and the
 * camera part of the launch record (optix_utils.py:166-181). */
int rg_set_views(rg_handle* h, int32_t num_views, const rg_camera* cams, void* stream);

/* Replaces the colour table of one view (after rg_set_views) with caller-computed per-Gaussian
 * colours rgb (n,3), original order: the test path renders colours computed outside the caster
 * (compute_cupy_rgb -> launch_pipeline_test, scripts/test.py:107-116, optix_utils.py:302-373). */
int rg_set_view_colors(rg_handle* h, int32_t view, const float* rgb, void* stream);

/* Forward render of num_rays rays of one view (pixels == NULL: ray r is pixel r).  out_rgb
 * (num_rays,3) is required; out_depth / out_opacity (num_rays) are optional.  out_token
 * (optional, host) receives the token of the hit-list records this forward wrote (0: none; see
 * rg_config.replay_words).  Replaces launch_pipeline_forward (optix_utils.py:128-209) and the
 * forward raygen (cuda_train/forward/gaussians_aabb.cu:521-612). */
int rg_forward(rg_handle* h, int32_t view, int64_t num_rays, const int32_t* pixels, float* out_rgb,
               float* out_depth, float* out_opacity, uint64_t* out_token, void* stream);

/* Backward of rg_forward for the same rays: accumulates parameter gradients into the handle.
 * ray_colors are rg_forward's outputs, dL_dC the incoming gradient (num_rays,3); replay_token
 * is the forward's out_token (0: walk the LBVH again).  Replaces launch_pipeline_backward
 * (optix_utils.py:213-299) and the backward raygen (cuda_train/backward/gaussians_aabb.cu:505-775). */
int rg_backward(rg_handle* h, int32_t view, int64_t num_rays, const int32_t* pixels, const float* ray_colors,
                const float* dL_dC, uint64_t replay_token, void* stream);

/* Multi-view versions: ray r of the batch is pixel ray_pixel[r] of view ray_view[r]. */
int rg_forward_batch(rg_handle* h, const rg_ray_batch* b, float* out_rgb, float* out_depth, float* out_opacity,
                     uint64_t* out_token, void* stream);
int rg_backward_batch(rg_handle* h, const rg_ray_batch* b, const float* ray_colors, const float* dL_dC,
                      uint64_t replay_token, void* stream);

/* Train step over a batch: forward, per-ray loss against target (loss_kind 0: L2 =
 * sum (C-target)^2, 1: L1 = sum |C-target|), backward -- two launches, the loss gradient formed
 * inside the backward kernel (no dL/dC buffer).  out_rgb (optional) receives the rendered
 * colours, out_loss (optional, device float[1]) the summed loss.  No reference counterpart: it
 * fuses RenderOptixFunction.forward + the caller's per-pixel loss + backward
 * (scripts/train.py:104-114) for per-pixel losses. */
int rg_train_batch(rg_handle* h, const rg_ray_batch* b, const float* target, int32_t loss_kind, float* out_rgb,
                   float* out_loss, void* stream);

/* Zeroes the gradient accumulators (done automatically by rg_finish_grads). */
int rg_zero_grads(rg_handle* h, void* stream);

/* Chains the accumulated per-view colour gradients through the SH/SG model (backward_sh,
 * utils/sh_utils.cu:121-315, summed over views), adds the view-direction position term
 * (render_optix.py:175), writes all gradients in the original Gaussian order and clears the
 * accumulators.  Replaces compute_cupy_sh_grad + the gradient assembly (render_optix.py:34-52,
 * :170-190). */
int rg_finish_grads(rg_handle* h, const rg_grads* out, void* stream);

/* Device scene box (6 floats: min xyz, max xyz) of the last rg_prepare: the reference's
 * bb_min/bb_max (render_optix.py:105-106). */
int rg_scene_box(rg_handle* h, float* out6_device, void* stream);

/* Counters of the last forward/backward/train call (synchronises the stream). */
int rg_get_stats(rg_handle* h, rg_stats* out, void* stream);

/* Allocation epoch of the handle's device workspace: it changes whenever a call (re)allocates a
 * buffer (a larger n, more views or rays, a first recording forward) and at every binned-SAH
 * rebuild (the refit launches one kernel per level of that tree).  A HIP graph captured over
 * this handle's calls bakes in those addresses and launches and stays valid only while the
 * epoch is unchanged.  No reference counterpart (the reference captures no graphs). */
int64_t rg_alloc_epoch(const rg_handle* h);

/* Device memory the handle holds: the sum of its live allocations (the tree, per-Gaussian
 * records, colour tables, gradient accumulators, hit-list records), bytes; record_bytes (may be
 * NULL) receives the share of the hit-list records that rg_backward replays (rays x replay_words
 * x 4 B plus 12 B per ray).  -1 for a null handle.  No reference counterpart: the reference
 * keeps no state between its forward and backward launches. */
int64_t rg_device_bytes(const rg_handle* h, int64_t* record_bytes);

/* Debug/verification: the FP32 per-Gaussian records in original order, exactly as the kernels use
 * them -- mu (n,3), Mp (n,9) rows of S'^-1 R^T, valid (n), and (optional, may be NULL) hlog (n) =
 * h = ln|sigma / thr|, the scale of the sampling exponent -- so that a checker can take the same
 * FP32 decisions (hit sets, sample thresholds) as the kernels. */
int rg_export_predicate(rg_handle* h, float* mu, float* Mp, int32_t* valid, float* hlog, void* stream);

/* Debug/verification: per-slab hit sets found by the LBVH traversal for num_rays rays of a view,
 * enumerating slabs from tenter to texit without transmittance exit.  pairs (num_rays,
 * max_pairs, 2) receives (slab index, original Gaussian index) sorted, counts (num_rays) the
 * number of pairs (may exceed max_pairs).  Also returns the FP32 rays and their interval. */
int rg_debug_hits(rg_handle* h, int32_t view, int64_t num_rays, const int32_t* pixels, int32_t max_slabs,
                  int32_t max_pairs, int32_t* pairs, int64_t* counts, float* rays_o, float* rays_d,
                  float* tenter, float* texit, void* stream);

/* Surface-area cost of the current tree: the sum of its node boxes' surface areas over the
 * root's, i.e. the expected number of node boxes a random ray crossing the root box visits.  A
 * refit keeps the topology and lets the boxes grow as Gaussians move; the ratio of this cost to
 * its value just after a rebuild tells when a rebuild pays (BatchedTrainer's rebuild policy).
 * Synchronises the stream.  No reference counterpart: the reference rebuilds its acceleration
 * structure on every forward (render_optix.py:107-110). */
int rg_tree_cost(rg_handle* h, double* cost, void* stream);

/* Tree check (debug, either builder): *errors = 0 when the links are consistent, every leaf
 * appears exactly once with its Gaussian's box, and every internal box is exactly (bit for bit)
 * the union of its children's boxes, as the refit computes it; otherwise the number of faults. */
int rg_debug_check_bvh(rg_handle* h, int64_t* errors, void* stream);

/* ---- optimiser step (handle-free; errors via rg_last_error(NULL)) ------------------------- */

/* One parameter tensor of an Adam step with its torch.optim.Adam state (exp_avg, exp_avg_sq),
 * all device pointers of n floats.  step = the state's step count AFTER this update. */
typedef struct rg_adam_tensor {
    float* param;
    const float* grad;
    float* exp_avg;
    float* exp_avg_sq;
    int64_t n;
    double lr;
    double beta1;
    double beta2;
    double eps;
    int64_t step;
} rg_adam_tensor;

/* Adam update of up to 16 tensors in one launch.  Replaces the per-tensor torch.optim.Adam(eps=
 * 1e-15) steps of the training loop (classes/optimizer.py:13-20 OptimizerManager.step,
 * utils/optim_utils.py:4-83 define_optimizer_manager, scripts/train.py:121-124). */
int rg_adam_step(const rg_adam_tensor* tensors, int32_t count, void* stream);

/* Post-step parameter projections (scripts/train.py:168-183): densities clamped at 0,
 * quaternions (n,4) normalised (a zero one becomes (1,0,0,0), classes/point_cloud.py:703-706),
 * lobe axes (n,num_sg,3) normalised (a zero one becomes (0,0,1), :721-724), log-scales (n,3)
 * clamped below by filter_3d (n, may be NULL) and above by max_scale.  NULL arrays are skipped. */
typedef struct rg_projection {
    int64_t n;
    int32_t num_sg;
    float* densities;
    float* log_scales;
    const float* filter_3d;
    float max_scale;
    float* quats;
    int32_t normalize_quaternion;
    float* lobe_axis;
    int32_t normalize_lobe_axis;
} rg_projection;
int rg_project_params(const rg_projection* pr, void* stream);

/* Viewer output (handle-free): n colours (n,3) -> RGBA8 (n,4) as the viewer's make_color
 * (cuda_gui/helpers.h:35-57): clamp to [0,1], min(floor(256 x), 255), alpha 255. */
int rg_colors_to_rgba8(const float* rgb, int64_t n, uint8_t* rgba, void* stream);

/* ---- photometric loss (handle-free) -------------------------------------------------------- */

/* Device workspace bytes rg_loss_l1_ssim needs for a width x height image. */
int64_t rg_loss_workspace_bytes(int32_t width, int32_t height);

/* L = (1 - lambda) * mean|x - y| + lambda * (1 - SSIM(x, y)), SSIM with the 11x11 Gaussian window
 * (sigma 1.5), zero padding, C1 = 0.01^2, C2 = 0.03^2 (utils/loss_utils.py:17-64; train.py:110-112
 * uses lambda = 0.2), where x is the rendered image after the supersampling mean
 * (utils/general_utils.py:489-502).  ray_colors: (height*ss_y, width*ss_x, 3) row-major rays;
 * target: (3, height, width).  out_loss: device float[1]; out_parts (optional): device float[2] =
 * (L1, SSIM); dL_dC (optional): gradient w.r.t. ray_colors, same layout.  No reference counterpart
 * as one routine: it fuses the caller's reduce_supersampling + l1_loss + ssim + their autograd. */
int rg_loss_l1_ssim(const float* ray_colors, const float* target, int32_t width, int32_t height, int32_t ss_x,
                    int32_t ss_y, float lambda_ssim, float* out_loss, float* out_parts, float* dL_dC,
                    void* workspace, int64_t workspace_bytes, void* stream);

/* ---- densification and initialisation utilities (handle-free) ------------------------------ */

/* Device workspace bytes rg_knn_mean_dist needs for n points. */
int64_t rg_knn_workspace_bytes(int64_t n);

/* distCUDA2 of simple-knn (submodules/simple-knn/simple_knn.cu:186-222, called at
 * classes/point_cloud.py:206,376,807): out[i] = mean of the squared distances from point i to its
 * 3 nearest other points (FP32 distances; the reference's bucket search is exact, so is this
 * one).  points: (n, 3) float32; out: (n) float32.  With fewer than 4 points the missing
 * neighbours count as FLT_MAX, as in the reference. */
int rg_knn_mean_dist(const float* points, int64_t n, float* out, void* workspace, int64_t workspace_bytes,
                     void* stream);

/* rg_knn_mean_dist without a workspace argument and without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The workspace is a static 64 MiB
 * device array that every GPU the library runs on holds until the process ends (one copy per
 * device; the calling thread's current device must be the stream's).  n is limited to what fits
 * it (about 1.6M points: the message of RG_ERR_ARG names the exact limit); larger n returns
 * RG_ERR_ARG.  Calls that overlap on one GPU (two streams, two threads) share the array and race:
 * serialise them. */
int rg_knn_mean_dist_noalloc(const float* points, int64_t n, float* out, void* stream);

/* One camera of the 3D filter, prepared on the host from the reference's camera:
 * R (row-major, as camera.R: x_cam = x R + T), T, fx = fov2focal(FoVx, W), fy, cx = W/2, cy = H/2,
 * and the screen window [-0.15 W, 1.15 W] x [-0.15 H, 1.15 H], all rounded to float. */
typedef struct rg_filter_camera {
    float R[9];
    float T[3];
    float fx, fy, cx, cy;
    float x_lo, x_hi, y_lo, y_hi;
} rg_filter_camera;

/* PointCloud.compute_3D_filter (classes/point_cloud.py:1152-1215): out[i] =
 * log(0.5 * d_i / focal) with d_i the smallest camera-space depth z > 0.2 over the cameras whose
 * widened screen window holds the projection (100000 if none is smaller), points seen by no
 * camera taking the largest d_i of the seen ones.  cams: DEVICE array of num_cams;
 * inv_focal = float(1 / max over cameras of fx, fy); workspace: device, >= 256 bytes; its first
 * uint32 holds, after the call, the bits of the largest seen d_i (0 when no point is seen, where
 * the reference raises: the outputs are then -inf). */
int rg_filter_3d(const float* points, int64_t n, const rg_filter_camera* cams, int32_t num_cams, float inv_focal,
                 float* out, void* workspace, int64_t workspace_bytes, void* stream);

/* Per-iteration gradient statistics of densification (classes/point_cloud.py:632-640):
 * accum += |g_i|, num += (|g_i| > 0), and num_gnv += (|g_i| > 0) when num_gnv is not NULL;
 * grad_positions: (n, 3). */
int rg_accumulate_grad_norm(const float* grad_positions, int64_t n, float* accum, float* num, float* num_gnv,
                            void* stream);

/* Inputs of densify_and_prune (classes/point_cloud.py:1067-1143): gradient statistics, the
 * parameters that decide and transform, and the thresholds.  scale_limit = float(0.01 *
 * scene_extent); split_n = N children per split Gaussian. */
typedef struct rg_densify_in {
    int64_t n;
    const float* grad_accum;   /* xyz_gradient_accum_norm (n) */
    const float* num_accum;    /* num_accum (n) */
    const float* positions;    /* (n, 3) */
    const float* log_scales;   /* (n, 3) */
    const float* quats;        /* (n, 4) raw (w, x, y, z) */
    const float* densities;    /* (n) */
    float grad_threshold;
    float scale_limit;
    float min_density;
    int32_t split_n;
} rg_densify_in;

/* One per-Gaussian array to carry through densification: rows of `width` floats.  zero_new = 0
 * copies a source row to each of its new rows (parameters, filter_3D), 1 gives new rows zeros
 * (Adam moments, visibility counters). */
typedef struct rg_densify_array {
    const float* src;
    float* dst;
    int64_t width;
    int32_t zero_new;
} rg_densify_array;

int64_t rg_densify_workspace_bytes(int64_t n);

/* First half of densify_and_prune: clone / split / prune decisions and the destination of every
 * row.  counts (device int64[4]): kept originals, kept clones, split Gaussians, split Gaussians
 * whose children are kept.  The output has counts[0] + counts[1] + split_n * counts[3] rows:
 * surviving originals, then clones, then children (child-major), each in source order, as the
 * reference's cat / boolean-mask sequence leaves them. */
int rg_densify_plan(const rg_densify_in* in, void* workspace, int64_t workspace_bytes, int64_t* counts,
                    void* stream);

/* Second half: writes every array's rows to their destinations, then the transformed rows:
 * halved densities of cloned Gaussians and their clones, and for the children
 * position = R(q) (s * z) + mu, log-scale = log(exp(log s) / (0.8 N)), density / 2, where
 * z = normals[c * counts[2] + rank] (the split Gaussian's rank among all split ones, child c)
 * are standard normal draws.  counts_host: the four counts read back from the plan; the
 * positions / log_scales / densities arrays must be among `arrays` and their destinations be
 * out_positions / out_log_scales / out_densities. */
int rg_densify_apply(const rg_densify_in* in, const void* workspace, int64_t workspace_bytes,
                     const int64_t* counts_host, const rg_densify_array* arrays, int32_t num_arrays,
                     const float* normals, float* out_positions, float* out_log_scales, float* out_densities,
                     void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RAYGAUSS_HIP_H */
