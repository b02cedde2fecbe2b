// C ABI of the gfx950 ray caster (declared in include/raygauss_hip.h).  Owns the per-handle
// workspace and sequences the kernels; never throws across the ABI.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "rg_internal.h"

struct rg_handle {
    rg_config cfg;
    std::string err;
    // Gaussians
    int64_t n = 0;
    int64_t cap = 0;
    bool built = false;
    rg_gaussians g{};
    int generation = 0;
    // per-Gaussian buffers (slot order)
    float4* geo = nullptr;
    float4* brec = nullptr;
    float4* leafbox = nullptr;
    float* acc = nullptr;
    float* box = nullptr;
    unsigned long long* counters = nullptr;
    unsigned long long* queue = nullptr;  // work queues of the persistent ray kernels (8 counters)
    rg::BvhBuffers bvh{};
    int64_t hist_cap = 0;
    // views
    int num_views = 0;
    int views_cap = 0;
    rg_camera* cams = nullptr;
    std::vector<rg_camera> host_cams;
    float4* rgb = nullptr;
    float4* drgb = nullptr;
    int64_t rgb_cap = 0;  // floats4 allocated for rgb and drgb
    // colour buffer of rg_train_batch when the caller passes no out_rgb
    float* train_rgb = nullptr;
    int64_t train_rgb_cap = 0;
    // hit-list records of the last forward (rays x replay_words), the (view, pixel) each row
    // belongs to, and the token the forward returned for them
    unsigned int* rec = nullptr;
    int* rec_len = nullptr;
    int2* rec_id = nullptr;
    int64_t rec_cap_words = 0;
    int64_t rec_cap_rays = 0;
    uint64_t tokens_issued = 0;
    struct {
        bool valid = false;
        uint64_t token = 0;
        int gen = 0;
        int64_t num_rays = 0;
        int words = 0;
    } rec_set;
    // bumped by every device (re)allocation of the handle's workspace
    int64_t alloc_epoch = 0;
    // bytes of every live device allocation of the handle, and their sum
    std::unordered_map<const void*, int64_t> sizes;
    int64_t device_bytes = 0;
};

namespace {
// Message of the last failed handle-free call (rg_adam_step, rg_project_params) on this thread.
thread_local std::string g_free_error = "no error";

int fail_free(int code, const std::string& msg) {
    g_free_error = msg;
    return code;
}

// Row limit of the handle-free per-Gaussian utilities (int32 row indices inside).
constexpr int64_t kMaxUtilRows = (int64_t)1 << 30;
}  // namespace

namespace {


int fail(rg_handle* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

#define HIPCHK(h, expr)                                                                              \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return fail(h, RG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

template <typename T>
int dalloc(rg_handle* h, T** p, int64_t count) {
    if (*p) {
        if (h) {
            auto it = h->sizes.find(*p);
            if (it != h->sizes.end()) {
                h->device_bytes -= it->second;
                h->sizes.erase(it);
            }
        }
        hipFree(*p);
        *p = nullptr;
    }
    if (count <= 0) count = 1;
    if (h) h->alloc_epoch++;
    const int64_t bytes = (int64_t)sizeof(T) * count;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), (size_t)bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(h, RG_ERR_OOM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    if (h) {
        h->sizes[*p] = bytes;
        h->device_bytes += bytes;
    }
    return RG_OK;
}

int check_config(rg_handle* h, const rg_config* c) {
    if (!c) return fail(h, RG_ERR_ARG, "null config");
    if (!(c->dt > 0.0f) || !std::isfinite(c->dt)) return fail(h, RG_ERR_ARG, "dt must be positive");
    if (c->samples_per_slab < 1 || c->samples_per_slab > 16) return fail(h, RG_ERR_ARG, "samples_per_slab in 1..16");
    if (!(c->sigma_thr > 0.0f)) return fail(h, RG_ERR_ARG, "sigma_thr must be positive");
    if (!(c->t_near <= c->t_far)) return fail(h, RG_ERR_ARG, "t_near must be <= t_far");
    if (c->hit_depth != 8 && c->hit_depth != 16 && c->hit_depth != 32)
        return fail(h, RG_ERR_ARG, "hit_depth must be 8, 16 or 32");
    if (c->seg_slabs < 1 || c->seg_slabs > 4096) return fail(h, RG_ERR_ARG, "seg_slabs in 1..4096");
    if (c->fwd_hit_depth != 0 && c->fwd_hit_depth != 6 && c->fwd_hit_depth != 8 && c->fwd_hit_depth != 12 &&
        c->fwd_hit_depth != 16 && c->fwd_hit_depth != 32)
        return fail(h, RG_ERR_ARG, "fwd_hit_depth must be 0, 6, 8, 12, 16 or 32");
    if (c->jitter != 0 && c->jitter != 1) return fail(h, RG_ERR_ARG, "jitter must be 0 or 1");
    if (c->rnd_sample < 0 || c->rnd_sample > 2) return fail(h, RG_ERR_ARG, "rnd_sample must be 0, 1 or 2");
    if (c->replay_words < 0 || c->replay_words > (1 << 20) || (c->replay_words & 7))
        return fail(h, RG_ERR_ARG, "replay_words must be a multiple of 8 in 0..2^20");
    if (c->tree_builder != 0 && c->tree_builder != 1) return fail(h, RG_ERR_ARG, "tree_builder must be 0 or 1");
    if (c->ray_chunk != 0 && (c->ray_chunk < 64 || c->ray_chunk > 65536 || (c->ray_chunk & 63)))
        return fail(h, RG_ERR_ARG, "ray_chunk must be 0 or a multiple of 64 in 64..65536");
    if (c->refill_lanes < 0 || c->refill_lanes > 64) return fail(h, RG_ERR_ARG, "refill_lanes must be in 0..64");
    if (c->bwd_ray_chunk != 0 && (c->bwd_ray_chunk < 64 || c->bwd_ray_chunk > 65536 || (c->bwd_ray_chunk & 63)))
        return fail(h, RG_ERR_ARG, "bwd_ray_chunk must be 0 or a multiple of 64 in 64..65536");
    if (!(c->bwd_window <= 1024.0f)) return fail(h, RG_ERR_ARG, "bwd_window must be at most 1024 slabs");
    if (c->fwd_round_per_slab != 0 && c->fwd_round_per_slab != 1)
        return fail(h, RG_ERR_ARG, "fwd_round_per_slab must be 0 or 1");
    return RG_OK;
}

int ensure_gaussian_capacity(rg_handle* h, int64_t n) {
    if (n <= h->cap && h->geo) return RG_OK;
    const int64_t cap = n;
    const int64_t total = 2 * cap;  // >= 2n-1
    int rc;
    if ((rc = dalloc(h, &h->geo, 4 * cap))) return rc;
    if ((rc = dalloc(h, &h->brec, 2 * cap))) return rc;
    if ((rc = dalloc(h, &h->leafbox, 2 * cap))) return rc;
    if ((rc = dalloc(h, &h->acc, (int64_t)rg::kAccStride * cap))) return rc;
    rg::BvhBuffers& b = h->bvh;
    if ((rc = dalloc(h, &b.keys[0], cap))) return rc;
    if ((rc = dalloc(h, &b.keys[1], cap))) return rc;
    if ((rc = dalloc(h, &b.vals[0], cap))) return rc;
    if ((rc = dalloc(h, &b.vals[1], cap))) return rc;
    const int64_t tiles = (cap + 2047) / 2048;
    if ((rc = dalloc(h, &b.hist, 256 * tiles))) return rc;
    if ((rc = dalloc(h, &b.perm, cap))) return rc;
    if ((rc = dalloc(h, &b.iperm, cap))) return rc;
    if ((rc = dalloc(h, &b.nodes, 2 * total))) return rc;
    if ((rc = dalloc(h, &b.parent, total))) return rc;
    if ((rc = dalloc(h, &b.right, total))) return rc;
    if ((rc = dalloc(h, &b.counter, total))) return rc;
    if ((rc = dalloc(h, &b.leaf_node, cap))) return rc;
    if ((rc = dalloc(h, &b.k_parent, total))) return rc;
    if ((rc = dalloc(h, &b.k_left, cap))) return rc;
    if ((rc = dalloc(h, &b.k_right, cap))) return rc;
    if ((rc = dalloc(h, &b.k_first, cap))) return rc;
    if ((rc = dalloc(h, &b.k_last, cap))) return rc;
    if ((rc = dalloc(h, &b.sah_prims[0], 2 * cap))) return rc;
    if ((rc = dalloc(h, &b.sah_prims[1], 2 * cap))) return rc;
    const int64_t task_f4 = (rg::sah_task_bytes((int)cap) + 15) / 16;
    if ((rc = dalloc(h, &b.sah_tasks[0], task_f4))) return rc;
    if ((rc = dalloc(h, &b.sah_tasks[1], task_f4))) return rc;
    if ((rc = dalloc(h, &b.sah_ctr, 4))) return rc;
    if ((rc = dalloc(h, &b.k_pre, cap))) return rc;
    h->cap = cap;
    h->built = false;
    return RG_OK;
}

int ensure_view_capacity(rg_handle* h, int views) {
    const int64_t need = (int64_t)views * std::max<int64_t>(h->n, 1);
    int rc;
    if (views > h->views_cap || !h->cams) {
        if ((rc = dalloc(h, &h->cams, views))) return rc;
        h->views_cap = views;
    }
    if (need > h->rgb_cap || !h->rgb) {
        if ((rc = dalloc(h, &h->rgb, need))) return rc;
        if ((rc = dalloc(h, &h->drgb, need))) return rc;
        h->rgb_cap = need;
        HIPCHK(h, hipMemset(h->drgb, 0, sizeof(float4) * need));
    }
    return RG_OK;
}

rg::DevScene dev_scene(rg_handle* h) {
    rg::DevScene s;
    s.n = (int)h->n;
    s.total_nodes = h->n > 0 ? (int)(2 * h->n - 1) : 0;
    s.geo = h->geo;
    s.brec = h->brec;
    s.nodes = h->bvh.nodes;
    s.rgb = h->rgb;
    s.box = h->box;
    s.perm = h->bvh.perm;
    return s;
}

rg::DevParams dev_params(const rg_handle* h) {
    rg::DevParams p;
    p.dt = h->cfg.dt;
    p.spb = h->cfg.samples_per_slab;
    p.t_eps = h->cfg.t_eps;
    p.sigma_thr = h->cfg.sigma_thr;
    p.t_near = h->cfg.t_near;
    p.t_far = h->cfg.t_far;
    p.bg = h->cfg.white_background ? 1.0f : 0.0f;
    p.max_hits = h->cfg.max_hits_per_slab;
    p.seg_slabs = h->cfg.seg_slabs;
    p.packet = h->cfg.packet_traversal;
    p.jitter = h->cfg.jitter;
    p.rnd_sample = h->cfg.rnd_sample;
    p.iteration = h->cfg.iteration;
    p.rec = nullptr;
    p.rec_len = nullptr;
    p.rec_id = nullptr;
    p.rec_words = 0;
    p.rec_mode = 0;
    p.ray_chunk = h->cfg.ray_chunk;
    p.refill_lanes = h->cfg.refill_lanes;
    p.queue = h->queue;
    p.bwd_window = h->cfg.bwd_window > 0.0f ? h->cfg.bwd_window : 0.5f;
    p.fwd_jump = h->cfg.fwd_round_per_slab ? 0 : 1;
    return p;
}

int ready_for_rays(rg_handle* h, int view) {
    if (!h->built) return fail(h, RG_ERR_STATE, "rg_prepare has not been called");
    if (h->num_views <= 0) return fail(h, RG_ERR_STATE, "rg_set_views has not been called");
    if (view < 0 || view >= h->num_views) return fail(h, RG_ERR_ARG, "view index out of range");
    return RG_OK;
}

rg::RayJob single_job(rg_handle* h, int view, int64_t num_rays, const int32_t* pixels) {
    rg::RayJob j;
    j.pixels = pixels;
    j.ray_view = nullptr;
    j.ray_pixel = nullptr;
    j.view = view;
    j.cams = h->cams;
    j.num_views = h->num_views;
    if (pixels) {
        j.num_rays = num_rays;
    } else {
        // implicit full view: 8x8 tiles
        const rg_camera& c = h->host_cams[view];
        j.num_rays = (int64_t)((c.width + 7) / 8) * ((c.height + 7) / 8) * 64;
    }
    return j;
}

int fwd_depth(const rg_handle* h) { return h->cfg.fwd_hit_depth ? h->cfg.fwd_hit_depth : h->cfg.hit_depth; }

rg::DevGrad dev_grad(rg_handle* h) {
    rg::DevGrad g;
    g.acc = h->acc;
    g.drgb = h->drgb;
    g.num_views = h->num_views;
    return g;
}

rg::DevStats dev_stats(rg_handle* h) {
    rg::DevStats s;
    s.c = h->counters;
    return s;
}

}  // namespace

extern "C" {

int rg_create(rg_handle** out, const rg_config* cfg) {
    if (!out) return RG_ERR_ARG;
    *out = nullptr;
    rg_handle* h = new (std::nothrow) rg_handle();
    if (!h) return RG_ERR_OOM;
    int rc = check_config(h, cfg);
    if (rc) {
        delete h;
        return rc;
    }
    h->cfg = *cfg;
    if ((rc = dalloc(h, &h->box, 8)) || (rc = dalloc(h, &h->counters, 16)) || (rc = dalloc(h, &h->queue, 16))) {
        if (h->box) hipFree(h->box);
        if (h->counters) hipFree(h->counters);
        delete h;
        return rc;
    }
    if (hipMemset(h->counters, 0, 16 * sizeof(unsigned long long)) != hipSuccess) {
        delete h;
        return RG_ERR_HIP;
    }
    *out = h;
    return RG_OK;
}

int rg_destroy(rg_handle* h) {
    if (!h) return RG_OK;
    void* ptrs[] = {h->geo, h->brec, h->leafbox, h->acc, h->box, h->counters, h->queue, h->bvh.keys[0], h->bvh.keys[1],
                    h->bvh.vals[0], h->bvh.vals[1], h->bvh.hist, h->bvh.perm, h->bvh.iperm, h->bvh.nodes,
                    h->bvh.parent, h->bvh.right, h->bvh.counter, h->bvh.leaf_node, h->bvh.k_parent,
                    h->bvh.k_left, h->bvh.k_right, h->bvh.k_first, h->bvh.k_last, h->cams, h->rgb, h->drgb,
                    h->train_rgb, h->rec, h->rec_len, h->rec_id, h->bvh.sah_prims[0], h->bvh.sah_prims[1],
                    h->bvh.sah_tasks[0], h->bvh.sah_tasks[1], h->bvh.sah_ctr, h->bvh.k_pre};
    for (void* p : ptrs)
        if (p) hipFree(p);
    delete h;
    return RG_OK;
}

int rg_set_config(rg_handle* h, const rg_config* cfg) {
    if (!h) return RG_ERR_ARG;
    int rc = check_config(h, cfg);
    if (rc) return rc;
    const bool thr_changed = cfg->sigma_thr != h->cfg.sigma_thr;
    h->cfg = *cfg;
    h->rec_set.valid = false;  // recorded hit lists belong to the previous settings
    if (thr_changed) h->built = false;  // records depend on the threshold: prepare again
    return RG_OK;
}

const char* rg_last_error(const rg_handle* h) { return h ? h->err.c_str() : g_free_error.c_str(); }

int rg_prepare(rg_handle* h, const rg_gaussians* g, int32_t rebuild, void* stream) {
    if (!h || !g) return fail(h, RG_ERR_ARG, "null argument");
    if (g->n < 0 || g->n > (int64_t)0x3fffffff) return fail(h, RG_ERR_ARG, "n out of range");
    if (g->sh_degree < 0 || g->sh_degree > 3) return fail(h, RG_ERR_ARG, "sh_degree must be 0..3");
    if (g->num_sg < 0 || g->num_sg > 8) return fail(h, RG_ERR_ARG, "num_sg must be 0..8");
    if (g->n > 0 && (!g->positions || !g->scales || !g->quats || !g->densities || !g->sh))
        return fail(h, RG_ERR_ARG, "null Gaussian array");
    if (g->n > 0 && g->num_sg > 0 && (!g->sg || !g->sharpness || !g->lobe_axis))
        return fail(h, RG_ERR_ARG, "null SG array with num_sg > 0");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc;
    const bool topo = rebuild || !h->built || g->n != h->n;
    if (topo && h->cfg.tree_builder == 1) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
            return fail(h, RG_ERR_ARG, "tree_builder 1 (SAH) cannot build under stream capture");
    }
    // from here on the handle's state changes: it is 'built' again only when everything succeeded
    h->built = false;
    h->rec_set.valid = false;
    if ((rc = ensure_gaussian_capacity(h, g->n))) return rc;
    h->n = g->n;
    h->g = *g;
    h->generation++;
    const int n = (int)g->n;
    if (topo) {
        h->num_views = 0;  // colour tables are per slot: views must be set again
        // records in original order -> Morton order -> tree
        HIPCHK(h, rg::launch_iota(n, h->bvh.perm, h->bvh.iperm, s));
        HIPCHK(h, rg::launch_preprocess(*g, h->bvh.iperm, h->cfg.sigma_thr, h->geo, h->brec, h->leafbox, h->box, s));
        if (h->cfg.tree_builder == 1) {
            const hipError_t e = rg::launch_sah_build(n, h->leafbox, h->geo, h->box, h->bvh, s);
            if (e == hipErrorUnknown) return fail(h, RG_ERR_HIP, "SAH build: inconsistent partition (internal error)");
            HIPCHK(h, e);
            // the refit's launches follow this tree's levels: a graph captured before is stale
            h->alloc_epoch++;
        } else {
            HIPCHK(h, rg::launch_lbvh_build(n, h->leafbox, h->box, h->bvh, s));
        }
        // gradient accumulators are per slot: clear what an unfinished backward left in them
        HIPCHK(h, rg::launch_zero(h->acc, (int64_t)rg::kAccStride * std::max<int64_t>(h->cap, 1), s));
        if (h->drgb) HIPCHK(h, rg::launch_zero(reinterpret_cast<float*>(h->drgb), 4 * h->rgb_cap, s));
    }
    HIPCHK(h, rg::launch_preprocess(*g, h->bvh.iperm, h->cfg.sigma_thr, h->geo, h->brec, h->leafbox, h->box, s));
    HIPCHK(h, rg::launch_lbvh_refit(n, h->leafbox, h->bvh, s));
    if (!topo && h->num_views > 0)
        HIPCHK(h, rg::launch_sh_forward(h->g, h->bvh.iperm, h->cams, h->num_views, h->rgb, s));
    h->built = true;
    return RG_OK;
}

int rg_set_views(rg_handle* h, int32_t num_views, const rg_camera* cams, void* stream) {
    if (!h || !cams) return fail(h, RG_ERR_ARG, "null argument");
    if (num_views <= 0) return fail(h, RG_ERR_ARG, "num_views must be positive");
    if (!h->built) return fail(h, RG_ERR_STATE, "rg_prepare has not been called");
    // the backward's merge key is view * n + slot in 32 bits
    if ((int64_t)num_views * std::max<int64_t>(h->n, 1) >= (int64_t)INT_MAX)
        return fail(h, RG_ERR_ARG, "num_views * n must stay below 2^31 - 1 (set fewer views per batch)");
    for (int v = 0; v < num_views; ++v)
        if (cams[v].width <= 0 || cams[v].height <= 0) return fail(h, RG_ERR_ARG, "camera size must be positive");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc;
    h->rec_set.valid = false;  // recorded hit lists belong to the previous cameras
    if ((rc = ensure_view_capacity(h, num_views))) return rc;
    h->host_cams.assign(cams, cams + num_views);
    HIPCHK(h, hipMemcpyAsync(h->cams, cams, sizeof(rg_camera) * num_views, hipMemcpyHostToDevice, s));
    h->num_views = num_views;
    HIPCHK(h, rg::launch_sh_forward(h->g, h->bvh.iperm, h->cams, num_views, h->rgb, s));
    return RG_OK;
}

int rg_set_view_colors(rg_handle* h, int32_t view, const float* rgb, void* stream) {
    if (!h || !rgb) return fail(h, RG_ERR_ARG, "null argument");
    if (!h->built || h->num_views <= 0) return fail(h, RG_ERR_STATE, "rg_set_views has not been called");
    if (view < 0 || view >= h->num_views) return fail(h, RG_ERR_ARG, "view out of range");
    HIPCHK(h, rg::launch_set_colors((int)h->n, rgb, h->bvh.iperm, h->rgb + (size_t)view * h->n,
                                    static_cast<hipStream_t>(stream)));
    return RG_OK;
}

// Hit-list records: with cfg.replay_words > 0 a forward records, per ray, the lists it gathered
// and returns a token for them; a backward given that token (and nothing in between: no
// rg_prepare, rg_set_views, rg_set_config or other recording forward) replays them instead of
// walking the LBVH again.  Each output row's records carry the (view, pixel) they were made
// for, and the backward replays a row only for that same ray, so a caller that rewrites its ray
// arrays between the two calls still gets correct gradients (those rays walk the tree).
static uint64_t rec_forward(rg_handle* h, const rg::RayJob& j, rg::DevParams& p) {
    h->rec_set.valid = false;
    const int words = h->cfg.replay_words;
    if (words <= 0 || j.num_rays <= 0) return 0;
    if (h->rec_cap_rays < j.num_rays || h->rec_cap_words < words) {
        h->rec_cap_rays = 0;
        h->rec_cap_words = 0;
        // without the memory the pass runs without records (the backward walks the tree)
        const int64_t rays = (j.num_rays + 63) & ~(int64_t)63;  // whole groups of 64 rays
        if (dalloc(h, &h->rec, rays * (int64_t)words) != RG_OK || dalloc(h, &h->rec_len, j.num_rays) != RG_OK ||
            dalloc(h, &h->rec_id, j.num_rays) != RG_OK)
            return 0;
        h->rec_cap_rays = j.num_rays;
        h->rec_cap_words = words;
    }
    p.rec = h->rec;
    p.rec_len = h->rec_len;
    p.rec_id = h->rec_id;
    p.rec_words = words;
    p.rec_mode = 1;
    h->rec_set.valid = true;
    h->rec_set.token = ++h->tokens_issued;
    h->rec_set.gen = h->generation;
    h->rec_set.num_rays = j.num_rays;
    h->rec_set.words = words;
    return h->rec_set.token;
}

static void rec_backward(const rg_handle* h, const rg::RayJob& j, uint64_t token, rg::DevParams& p) {
    const auto& r = h->rec_set;
    if (token == 0 || !r.valid || r.token != token || r.gen != h->generation || r.words != h->cfg.replay_words ||
        r.num_rays < j.num_rays)
        return;
    p.rec = h->rec;
    p.rec_len = h->rec_len;
    p.rec_id = h->rec_id;
    p.rec_words = r.words;
    p.rec_mode = 2;
}

int rg_forward(rg_handle* h, int32_t view, int64_t num_rays, const int32_t* pixels, float* out_rgb, float* out_depth,
               float* out_opacity, uint64_t* out_token, void* stream) {
    if (out_token) *out_token = 0;
    if (!h || !out_rgb) return fail(h, RG_ERR_ARG, "null argument");
    int rc;
    if ((rc = ready_for_rays(h, view))) return rc;
    if (num_rays < 0) return fail(h, RG_ERR_ARG, "num_rays < 0");
    const rg_camera& c = h->host_cams[view];
    if (!pixels && num_rays != (int64_t)c.width * c.height)
        return fail(h, RG_ERR_ARG, "without a pixel list num_rays must equal width*height");
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIPCHK(h, rg::launch_reset_counters(dev_stats(h), s));
    const rg::RayJob job = single_job(h, view, num_rays, pixels);
    rg::DevParams p = dev_params(h);
    const uint64_t token = rec_forward(h, job, p);
    HIPCHK(h, rg::launch_forward(dev_scene(h), p, job, fwd_depth(h), out_rgb, out_depth, out_opacity, dev_stats(h), s));
    if (out_token) *out_token = token;
    return RG_OK;
}

int rg_backward(rg_handle* h, int32_t view, int64_t num_rays, const int32_t* pixels, const float* ray_colors,
                const float* dL_dC, uint64_t replay_token, void* stream) {
    if (!h || !ray_colors || !dL_dC) return fail(h, RG_ERR_ARG, "null argument");
    int rc;
    if ((rc = ready_for_rays(h, view))) return rc;
    const rg_camera& c = h->host_cams[view];
    if (!pixels && num_rays != (int64_t)c.width * c.height)
        return fail(h, RG_ERR_ARG, "without a pixel list num_rays must equal width*height");
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIPCHK(h, rg::launch_reset_counters(dev_stats(h), s));
    const rg::RayJob job = single_job(h, view, num_rays, pixels);
    const rg::DevGrad gr = dev_grad(h);
    rg::DevParams p = dev_params(h);
    rec_backward(h, job, replay_token, p);
    p.ray_chunk = h->cfg.bwd_ray_chunk;
    HIPCHK(h, rg::launch_backward(dev_scene(h), p, job, h->cfg.hit_depth, ray_colors, dL_dC, nullptr, 0, 1, gr,
                                  dev_stats(h), s));
    return RG_OK;
}

static int batch_job(rg_handle* h, const rg_ray_batch* b, rg::RayJob& j) {
    if (!b) return fail(h, RG_ERR_ARG, "null batch");
    if (!h->built) return fail(h, RG_ERR_STATE, "rg_prepare has not been called");
    if (h->num_views <= 0) return fail(h, RG_ERR_STATE, "rg_set_views has not been called");
    if (b->num_rays < 0 || (b->num_rays > 0 && (!b->ray_view || !b->ray_pixel)))
        return fail(h, RG_ERR_ARG, "bad ray batch");
    j.num_rays = b->num_rays;
    j.pixels = nullptr;
    j.ray_view = b->ray_view;
    j.ray_pixel = b->ray_pixel;
    j.view = 0;
    j.cams = h->cams;
    j.num_views = h->num_views;
    return RG_OK;
}

int rg_forward_batch(rg_handle* h, const rg_ray_batch* b, float* out_rgb, float* out_depth, float* out_opacity,
                     uint64_t* out_token, void* stream) {
    if (out_token) *out_token = 0;
    if (!h || !out_rgb) return fail(h, RG_ERR_ARG, "null argument");
    rg::RayJob j;
    int rc;
    if ((rc = batch_job(h, b, j))) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIPCHK(h, rg::launch_reset_counters(dev_stats(h), s));
    rg::DevParams p = dev_params(h);
    const uint64_t token = rec_forward(h, j, p);
    HIPCHK(h, rg::launch_forward(dev_scene(h), p, j, fwd_depth(h), out_rgb, out_depth, out_opacity, dev_stats(h), s));
    if (out_token) *out_token = token;
    return RG_OK;
}

int rg_backward_batch(rg_handle* h, const rg_ray_batch* b, const float* ray_colors, const float* dL_dC,
                      uint64_t replay_token, void* stream) {
    if (!h || !ray_colors || !dL_dC) return fail(h, RG_ERR_ARG, "null argument");
    rg::RayJob j;
    int rc;
    if ((rc = batch_job(h, b, j))) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIPCHK(h, rg::launch_reset_counters(dev_stats(h), s));
    const rg::DevGrad gr = dev_grad(h);
    rg::DevParams p = dev_params(h);
    rec_backward(h, j, replay_token, p);
    p.ray_chunk = h->cfg.bwd_ray_chunk;
    HIPCHK(h, rg::launch_backward(dev_scene(h), p, j, h->cfg.hit_depth, ray_colors, dL_dC, nullptr, 0, 1, gr,
                                  dev_stats(h), s));
    return RG_OK;
}

int rg_train_batch(rg_handle* h, const rg_ray_batch* b, const float* target, int32_t loss_kind, float* out_rgb,
                   float* out_loss, void* stream) {
    if (!h || !target) return fail(h, RG_ERR_ARG, "null argument");
    if (loss_kind != 0 && loss_kind != 1) return fail(h, RG_ERR_ARG, "loss_kind must be 0 (L2) or 1 (L1)");
    rg::RayJob j;
    int rc;
    if ((rc = batch_job(h, b, j))) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIPCHK(h, rg::launch_reset_counters(dev_stats(h), s));
    const rg::DevGrad gr = dev_grad(h);
    float* rgb = out_rgb;
    if (!rgb) {
        if (h->train_rgb_cap < b->num_rays) {
            if ((rc = dalloc(h, &h->train_rgb, 3 * b->num_rays))) return rc;
            h->train_rgb_cap = b->num_rays;
        }
        rgb = h->train_rgb;
    }
    const rg::DevScene sc = dev_scene(h);
    rg::DevParams p = dev_params(h);
    const uint64_t token = rec_forward(h, j, p);
    HIPCHK(h, rg::launch_forward(sc, p, j, fwd_depth(h), rgb, nullptr, nullptr, dev_stats(h), s));
    p = dev_params(h);
    rec_backward(h, j, token, p);
    p.ray_chunk = h->cfg.bwd_ray_chunk;
    HIPCHK(h, rg::launch_backward(sc, p, j, h->cfg.hit_depth, rgb, nullptr, target, loss_kind, 0, dev_grad(h),
                                  dev_stats(h), s));
    if (out_loss)
        HIPCHK(h, hipMemcpyAsync(out_loss, h->counters + 6, sizeof(float), hipMemcpyDeviceToDevice, s));
    return RG_OK;
}

int rg_zero_grads(rg_handle* h, void* stream) {
    if (!h) return RG_ERR_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->acc) HIPCHK(h, rg::launch_zero(h->acc, (int64_t)rg::kAccStride * h->cap, s));
    if (h->drgb) HIPCHK(h, rg::launch_zero(reinterpret_cast<float*>(h->drgb), 4 * h->rgb_cap, s));
    return RG_OK;
}

int rg_finish_grads(rg_handle* h, const rg_grads* out, void* stream) {
    if (!h || !out) return fail(h, RG_ERR_ARG, "null argument");
    if (!h->built || h->num_views <= 0) return fail(h, RG_ERR_STATE, "nothing to finish");
    if (h->n > 0 && (!out->positions || !out->scales || !out->quats || !out->densities || !out->sh))
        return fail(h, RG_ERR_ARG, "null gradient array");
    if (h->n > 0 && h->g.num_sg > 0 && (!out->sg || !out->sharpness || !out->lobe_axis))
        return fail(h, RG_ERR_ARG, "null SG gradient array with num_sg > 0");
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIPCHK(h, rg::launch_finish_grads(h->g, h->bvh.iperm, h->cams, h->num_views, dev_grad(h),
                                      h->cfg.exact_sh_grad, *out, s));
    return RG_OK;
}

int rg_scene_box(rg_handle* h, float* out6, void* stream) {
    if (!h || !out6) return fail(h, RG_ERR_ARG, "null argument");
    if (!h->built) return fail(h, RG_ERR_STATE, "rg_prepare has not been called");
    HIPCHK(h, hipMemcpyAsync(out6, h->box, 6 * sizeof(float), hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
    return RG_OK;
}

int rg_get_stats(rg_handle* h, rg_stats* out, void* stream) {
    if (!h || !out) return fail(h, RG_ERR_ARG, "null argument");
    unsigned long long c[12];
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIPCHK(h, hipMemcpyAsync(c, h->counters, sizeof(c), hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipStreamSynchronize(s));
    out->slab_hits = (int64_t)c[0];
    out->pairs = (int64_t)c[1];
    out->overflow_slabs = (int64_t)c[2];
    out->fallback_slabs = (int64_t)c[3];
    out->gathers = (int64_t)c[4];
    out->replay_words = (int64_t)c[8];
    out->replay_truncated = (int64_t)c[9];
    out->dropped = (int64_t)c[10];
    out->bad_rays = (int64_t)c[11];
    return RG_OK;
}

int64_t rg_alloc_epoch(const rg_handle* h) { return h ? h->alloc_epoch : -1; }

int64_t rg_device_bytes(const rg_handle* h, int64_t* record_bytes) {
    if (!h) return -1;
    if (record_bytes) {
        int64_t r = 0;
        for (const void* q : {(const void*)h->rec, (const void*)h->rec_len, (const void*)h->rec_id}) {
            auto it = h->sizes.find(q);
            if (q && it != h->sizes.end()) r += it->second;
        }
        *record_bytes = r;
    }
    return h->device_bytes;
}

int rg_export_predicate(rg_handle* h, float* mu, float* Mp, int32_t* valid, float* hlog, void* stream) {
    if (!h || !mu || !Mp || !valid) return fail(h, RG_ERR_ARG, "null argument");
    if (!h->built) return fail(h, RG_ERR_STATE, "rg_prepare has not been called");
    HIPCHK(h, rg::launch_export_predicate((int)h->n, h->bvh.iperm, h->geo, mu, Mp, valid, hlog,
                                          static_cast<hipStream_t>(stream)));
    return RG_OK;
}

int rg_debug_hits(rg_handle* h, int32_t view, int64_t num_rays, const int32_t* pixels, int32_t max_slabs,
                  int32_t max_pairs, int32_t* pairs, int64_t* counts, float* rays_o, float* rays_d, float* tenter,
                  float* texit, void* stream) {
    if (!h || !pairs || !counts || !rays_o || !rays_d || !tenter || !texit) return fail(h, RG_ERR_ARG, "null argument");
    int rc;
    if ((rc = ready_for_rays(h, view))) return rc;
    if (max_pairs <= 0 || max_slabs <= 0) return fail(h, RG_ERR_ARG, "max_pairs and max_slabs must be positive");
    const rg_camera& c = h->host_cams[view];
    if (!pixels && num_rays > (int64_t)c.width * c.height) return fail(h, RG_ERR_ARG, "num_rays > width*height");
    rg::RayJob j;
    j.num_rays = num_rays;
    j.pixels = pixels;
    j.ray_view = nullptr;
    j.ray_pixel = nullptr;
    j.view = view;
    j.cams = h->cams;
    j.num_views = h->num_views;
    HIPCHK(h, rg::launch_debug_hits(dev_scene(h), dev_params(h), j, max_slabs, max_pairs, pairs, counts, rays_o, rays_d,
                                    tenter, texit, static_cast<hipStream_t>(stream)));
    return RG_OK;
}

int rg_debug_check_bvh(rg_handle* h, int64_t* errors, void* stream) {
    if (!h || !errors) return fail(h, RG_ERR_ARG, "null argument");
    if (!h->built) return fail(h, RG_ERR_STATE, "rg_prepare has not been called");
    hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned long long* d = h->counters + 7;
    HIPCHK(h, hipMemsetAsync(d, 0, sizeof(unsigned long long), s));
    HIPCHK(h, rg::launch_lbvh_check((int)h->n, h->bvh, h->leafbox, d, s));
    unsigned long long e = 0;
    HIPCHK(h, hipMemcpyAsync(&e, d, sizeof(e), hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipStreamSynchronize(s));
    *errors = (int64_t)e;
    return RG_OK;
}

int rg_tree_cost(rg_handle* h, double* cost, void* stream) {
    if (!h || !cost) return fail(h, RG_ERR_ARG, "null argument");
    if (!h->built) return fail(h, RG_ERR_STATE, "rg_prepare has not been called");
    hipStream_t s = static_cast<hipStream_t>(stream);
    double* d = reinterpret_cast<double*>(h->counters + 7);  // scratch slot (as rg_debug_check_bvh)
    HIPCHK(h, hipMemsetAsync(d, 0, sizeof(double), s));
    HIPCHK(h, rg::launch_tree_area((int)h->n, h->bvh, d, s));
    float4 root[2];
    double sum = 0.0;
    HIPCHK(h, hipMemcpyAsync(&sum, d, sizeof(sum), hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipMemcpyAsync(root, h->bvh.nodes, sizeof(root), hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipStreamSynchronize(s));
    const double dx = (double)root[1].x - root[0].x, dy = (double)root[1].y - root[0].y,
                 dz = (double)root[1].z - root[0].z;
    const double a = 2.0 * (dx * dy + dy * dz + dz * dx);
    *cost = a > 0.0 && std::isfinite(a) ? sum / a : 0.0;
    return RG_OK;
}

int rg_adam_step(const rg_adam_tensor* tensors, int32_t count, void* stream) {
    if (count < 0 || count > rg::kMaxAdamTensors) return fail_free(RG_ERR_ARG, "rg_adam_step: 0..16 tensors");
    if (count > 0 && !tensors) return fail_free(RG_ERR_ARG, "rg_adam_step: null tensor table");
    rg::AdamBatch b{};
    b.count = count;
    int64_t items = 0;
    for (int k = 0; k < count; ++k) {
        const rg_adam_tensor& t = tensors[k];
        if (t.n < 0) return fail_free(RG_ERR_ARG, "rg_adam_step: negative size");
        if (t.n > 0 && (!t.param || !t.grad || !t.exp_avg || !t.exp_avg_sq))
            return fail_free(RG_ERR_ARG, "rg_adam_step: null array");
        if (t.step < 1) return fail_free(RG_ERR_ARG, "rg_adam_step: step must be >= 1 (count after the update)");
        if (!(t.beta1 >= 0.0 && t.beta1 < 1.0 && t.beta2 >= 0.0 && t.beta2 < 1.0 && t.eps >= 0.0 && t.lr >= 0.0))
            return fail_free(RG_ERR_ARG, "rg_adam_step: lr, eps >= 0 and betas in [0, 1) required");
        b.t[k].param = t.param;
        b.t[k].grad = t.grad;
        b.t[k].exp_avg = t.exp_avg;
        b.t[k].exp_avg_sq = t.exp_avg_sq;
        b.t[k].n = t.n;
        auto al = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15u) == 0; };
        b.t[k].aligned = al(t.param) && al(t.grad) && al(t.exp_avg) && al(t.exp_avg_sq);
        const double bc1 = 1.0 - std::pow(t.beta1, (double)t.step);
        const double bc2 = 1.0 - std::pow(t.beta2, (double)t.step);
        b.s[k].w1 = (float)(1.0 - t.beta1);
        b.s[k].b2 = (float)t.beta2;
        b.s[k].w2 = (float)(1.0 - t.beta2);
        b.s[k].bc2_sqrt = (float)std::sqrt(bc2);
        b.s[k].eps = (float)t.eps;
        b.s[k].neg_step = (float)(-(t.lr / bc1));
        b.item_start[k] = items;
        items += (t.n + 3) / 4;
    }
    b.item_start[count] = items;
    const hipError_t e = rg::launch_adam(b, items, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail_free(RG_ERR_HIP, std::string("rg_adam_step: ") + hipGetErrorString(e));
    return RG_OK;
}

int rg_project_params(const rg_projection* pr, void* stream) {
    if (!pr) return fail_free(RG_ERR_ARG, "rg_project_params: null argument");
    if (pr->n < 0 || pr->num_sg < 0) return fail_free(RG_ERR_ARG, "rg_project_params: negative size");
    if (pr->normalize_lobe_axis && pr->num_sg > 0 && pr->n > 0 && !pr->lobe_axis)
        return fail_free(RG_ERR_ARG, "rg_project_params: null lobe_axis with num_sg > 0");
    const hipError_t e = rg::launch_project(*pr, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail_free(RG_ERR_HIP, std::string("rg_project_params: ") + hipGetErrorString(e));
    return RG_OK;
}

int rg_colors_to_rgba8(const float* rgb, int64_t n, uint8_t* rgba, void* stream) {
    if (n < 0 || (n > 0 && (!rgb || !rgba))) return fail_free(RG_ERR_ARG, "rg_colors_to_rgba8: bad argument");
    const hipError_t e = rg::launch_rgba8(n, rgb, reinterpret_cast<uchar4*>(rgba), static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail_free(RG_ERR_HIP, std::string("rg_colors_to_rgba8: ") + hipGetErrorString(e));
    return RG_OK;
}

int64_t rg_knn_workspace_bytes(int64_t n) {
    if (n <= 0 || n > kMaxUtilRows) return 0;
    return rg::knn_workspace_bytes(n);
}

int rg_knn_mean_dist(const float* points, int64_t n, float* out, void* workspace, int64_t workspace_bytes,
                     void* stream) {
    if (n < 0 || n > kMaxUtilRows) return fail_free(RG_ERR_ARG, "rg_knn_mean_dist: n out of range");
    if (n == 0) return RG_OK;
    if (!points || !out || !workspace) return fail_free(RG_ERR_ARG, "rg_knn_mean_dist: null argument");
    if (workspace_bytes < rg::knn_workspace_bytes(n))
        return fail_free(RG_ERR_ARG, "rg_knn_mean_dist: workspace smaller than rg_knn_workspace_bytes()");
    const hipError_t e = rg::launch_knn_mean_dist(points, n, out, workspace, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail_free(RG_ERR_HIP, std::string("rg_knn_mean_dist: ") + hipGetErrorString(e));
    return RG_OK;
}

int rg_knn_mean_dist_noalloc(const float* points, int64_t n, float* out, void* stream) {
    if (n < 0 || n > kMaxUtilRows) return fail_free(RG_ERR_ARG, "rg_knn_mean_dist_noalloc: n out of range");
    if (n == 0) return RG_OK;
    if (!points || !out) return fail_free(RG_ERR_ARG, "rg_knn_mean_dist_noalloc: null argument");
    const int64_t cap = rg::knn_scratch_bytes();
    if (rg::knn_workspace_bytes(n) > cap) {
        // largest n whose workspace fits (the workspace grows monotonically with n)
        int64_t lo = 0, hi = n;
        while (hi - lo > 1) {
            const int64_t mid = lo + (hi - lo) / 2;
            (rg::knn_workspace_bytes(mid) <= cap ? lo : hi) = mid;
        }
        return fail_free(RG_ERR_ARG, "rg_knn_mean_dist_noalloc: n = " + std::to_string(n) + " needs " +
                                         std::to_string(rg::knn_workspace_bytes(n)) + " bytes of scratch; the static " +
                                         "scratch holds " + std::to_string(cap) + " bytes (n <= " + std::to_string(lo) +
                                         "): use rg_knn_mean_dist with a workspace");
    }
    const hipError_t e = rg::launch_knn_mean_dist_noalloc(points, n, out, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail_free(RG_ERR_HIP, std::string("rg_knn_mean_dist_noalloc: ") + hipGetErrorString(e));
    return RG_OK;
}

int rg_filter_3d(const float* points, int64_t n, const rg_filter_camera* cams, int32_t num_cams, float inv_focal,
                 float* out, void* workspace, int64_t workspace_bytes, void* stream) {
    if (n < 0 || n > kMaxUtilRows || num_cams < 0) return fail_free(RG_ERR_ARG, "rg_filter_3d: size out of range");
    if (!workspace || workspace_bytes < 256) return fail_free(RG_ERR_ARG, "rg_filter_3d: workspace < 256 bytes");
    if (n > 0 && (!points || !out || (num_cams > 0 && !cams)))
        return fail_free(RG_ERR_ARG, "rg_filter_3d: null argument");
    const hipError_t e = rg::launch_filter_3d(points, n, cams, num_cams, inv_focal, out,
                                              static_cast<unsigned int*>(workspace), static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail_free(RG_ERR_HIP, std::string("rg_filter_3d: ") + hipGetErrorString(e));
    return RG_OK;
}

int rg_accumulate_grad_norm(const float* grad_positions, int64_t n, float* accum, float* num, float* num_gnv,
                            void* stream) {
    if (n < 0) return fail_free(RG_ERR_ARG, "rg_accumulate_grad_norm: n < 0");
    if (n > 0 && (!grad_positions || !accum || !num))
        return fail_free(RG_ERR_ARG, "rg_accumulate_grad_norm: null argument");
    const hipError_t e = rg::launch_accum_grad(grad_positions, n, accum, num, num_gnv, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail_free(RG_ERR_HIP, std::string("rg_accumulate_grad_norm: ") + hipGetErrorString(e));
    return RG_OK;
}

int64_t rg_densify_workspace_bytes(int64_t n) {
    if (n < 0 || n > kMaxUtilRows) return 0;
    return rg::densify_workspace_bytes(n);
}

static int densify_check(const rg_densify_in* in, int64_t workspace_bytes, const char* who) {
    if (!in) return fail_free(RG_ERR_ARG, std::string(who) + ": null argument");
    if (in->n < 0 || in->n > kMaxUtilRows) return fail_free(RG_ERR_ARG, std::string(who) + ": n out of range");
    if (in->split_n < 1 || in->split_n > 64) return fail_free(RG_ERR_ARG, std::string(who) + ": split_n in [1, 64]");
    if (in->n > 0 && (!in->grad_accum || !in->num_accum || !in->positions || !in->log_scales || !in->quats ||
                      !in->densities))
        return fail_free(RG_ERR_ARG, std::string(who) + ": null input array");
    if (workspace_bytes < rg::densify_workspace_bytes(in->n))
        return fail_free(RG_ERR_ARG, std::string(who) + ": workspace smaller than rg_densify_workspace_bytes()");
    return RG_OK;
}

int rg_densify_plan(const rg_densify_in* in, void* workspace, int64_t workspace_bytes, int64_t* counts,
                    void* stream) {
    int rc;
    if ((rc = densify_check(in, workspace_bytes, "rg_densify_plan"))) return rc;
    if (!workspace || !counts) return fail_free(RG_ERR_ARG, "rg_densify_plan: null argument");
    const hipError_t e = rg::launch_densify_plan(*in, workspace, counts, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail_free(RG_ERR_HIP, std::string("rg_densify_plan: ") + hipGetErrorString(e));
    return RG_OK;
}

int rg_densify_apply(const rg_densify_in* in, const void* workspace, int64_t workspace_bytes,
                     const int64_t* counts_host, const rg_densify_array* arrays, int32_t num_arrays,
                     const float* normals, float* out_positions, float* out_log_scales, float* out_densities,
                     void* stream) {
    int rc;
    if ((rc = densify_check(in, workspace_bytes, "rg_densify_apply"))) return rc;
    if (!workspace || !counts_host) return fail_free(RG_ERR_ARG, "rg_densify_apply: null argument");
    if (num_arrays < 0 || num_arrays > rg::densify_max_arrays() || (num_arrays > 0 && !arrays))
        return fail_free(RG_ERR_ARG, "rg_densify_apply: num_arrays out of range");
    const int64_t ko = counts_host[0], kc = counts_host[1], ns = counts_host[2], kch = counts_host[3];
    if (ko < 0 || kc < 0 || ns < 0 || kch < 0 || kch > ns || ko + ns > in->n || kc > in->n)
        return fail_free(RG_ERR_ARG, "rg_densify_apply: counts inconsistent with n");
    if (in->n > 0 && (!out_positions || !out_log_scales || !out_densities))
        return fail_free(RG_ERR_ARG, "rg_densify_apply: null output");
    if (ns > 0 && !normals) return fail_free(RG_ERR_ARG, "rg_densify_apply: normals needed for split Gaussians");
    for (int k = 0; k < num_arrays; ++k)
        if (arrays[k].width <= 0 || (in->n > 0 && (!arrays[k].src || !arrays[k].dst)))
            return fail_free(RG_ERR_ARG, "rg_densify_apply: bad array entry");
    const hipError_t e = rg::launch_densify_apply(*in, workspace, counts_host, arrays, num_arrays, normals,
                                                  out_positions, out_log_scales, out_densities,
                                                  static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail_free(RG_ERR_HIP, std::string("rg_densify_apply: ") + hipGetErrorString(e));
    return RG_OK;
}

int64_t rg_loss_workspace_bytes(int32_t width, int32_t height) {
    if (width <= 0 || height <= 0) return 0;
    return rg::loss_workspace_bytes(width, height);
}

int rg_loss_l1_ssim(const float* ray_colors, const float* target, int32_t width, int32_t height, int32_t ss_x,
                    int32_t ss_y, float lambda_ssim, float* out_loss, float* out_parts, float* dL_dC,
                    void* workspace, int64_t workspace_bytes, void* stream) {
    if (!ray_colors || !target || !out_loss || !workspace)
        return fail_free(RG_ERR_ARG, "rg_loss_l1_ssim: null argument");
    if (width <= 0 || height <= 0 || ss_x <= 0 || ss_y <= 0)
        return fail_free(RG_ERR_ARG, "rg_loss_l1_ssim: sizes must be positive");
    if (workspace_bytes < rg::loss_workspace_bytes(width, height))
        return fail_free(RG_ERR_ARG, "rg_loss_l1_ssim: workspace smaller than rg_loss_workspace_bytes()");
    // the window as the reference builds it: float Gaussian taps normalised by their float sum
    float w[11];
    float sum = 0.0f;
    for (int i = 0; i < 11; ++i) {
        const double d = (double)(i - 5);
        w[i] = (float)std::exp(-(d * d) / (2.0 * 1.5 * 1.5));
        sum += w[i];
    }
    for (int i = 0; i < 11; ++i) w[i] /= sum;
    const hipError_t e = rg::launch_loss_l1_ssim(ray_colors, target, width, height, ss_x, ss_y, w, lambda_ssim,
                                                 out_loss, out_parts, dL_dC, workspace,
                                                 static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail_free(RG_ERR_HIP, std::string("rg_loss_l1_ssim: ") + hipGetErrorString(e));
    return RG_OK;
}

}  // extern "C"
