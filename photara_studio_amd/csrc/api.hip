// C ABI of the gfx950 3DGS training step: workspace carving and launch orchestration.
//
// Replaces splat_drender::Rasterizer::{forward, backward} (third_party/splat_drender/src/
// pipeline.cu:178-312, 457-712) and the CudaRasterizer bridge (photara/src/splat/
// rasterizer.cu:340-743). Differences by design: every view of a batched step goes through one
// launch per stage; all counts stay on the device (fixed-capacity buffers + an overflow flag
// instead of the reference's host readback of the instance count); workspace memory is carved
// from one caller-owned allocation instead of grow-callbacks; every launch goes to the caller's
// stream, so the step can be captured as one hipGraph.
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "launch.h"

using namespace pht;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) return fail(PHT_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
    return PHT_OK;
}

int next_pow2_log2(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

int bits_for(uint64_t max_value) {
    int b = 0;
    while (b < 32 && (max_value >> b) != 0) ++b;
    return b;
}

struct Carve {
    char* base;
    size_t off = 0;
    template <class T> T* take(size_t count) {
        off = (off + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base ? base + off : nullptr);
        off += count * sizeof(T);
        return p;
    }
};

} // namespace

struct pht_workspace {
    pht_capacity cap{};
    int vp_log2 = 0;
    uint32_t slot_cap = 0, inst_cap = 0;
    DevView* views = nullptr;
    Counters* counters = nullptr;
    uint32_t* key_parts = nullptr; // partial ranges of the visible depth keys (kKeyParts pairs)
    SplatRec* rec = nullptr;
    uint2* meta = nullptr;
    uint32_t* vis_flag = nullptr;
    float* sgrad = nullptr;
    uint32_t* view_mask = nullptr;
    uint32_t* dkeys[3] = {nullptr, nullptr, nullptr}; // depth sort: two ping-pong buffers and the result
    uint32_t* dvals[3] = {nullptr, nullptr, nullptr};
    uint32_t* dtiles[3] = {nullptr, nullptr, nullptr};
    uint32_t* offsets = nullptr;
    uint32_t* tkeys[2] = {nullptr, nullptr};
    uint32_t* tvals[2] = {nullptr, nullptr};
    uint2* ranges = nullptr;
    uint32_t* tile_order = nullptr;
    uint32_t* order_hist = nullptr;
    float* final_T = nullptr;
    uint32_t* n_contrib = nullptr;
    uint32_t* scan_scratch = nullptr;
    uint32_t* hist = nullptr;
    uint32_t* digit_total = nullptr;
    Counters* host_counters = nullptr; // pinned
    hipEvent_t join_ev = nullptr;      // end of this workspace's blend backward in a grouped backward
    // state of the last forward
    bool fwd_valid = false;
    // a backward has accumulated into sgrad since the forward cleared it: the next backward of
    // the same forward clears the records again first (pipeline.cu:680-688 zeroes the gradient
    // state on every backward)
    bool bwd_done = false;
    int n_views = 0;
    int tile_sel = 0;
    int64_t gaussians = 0;
    int64_t tiles_total = 0;
    int64_t pixels_total = 0;
    pht_settings settings{};
    std::vector<DevView> host_views;
    // per-stage timing
    bool profiling = false;
    hipEvent_t ev[PHT_STAGES + 1][2] = {};
    bool ev_recorded[PHT_STAGES] = {};

    void mark(int stage, int edge, hipStream_t st) {
        if (!profiling) return;
        if (!ev[stage][edge]) (void)hipEventCreate(&ev[stage][edge]);
        (void)hipEventRecord(ev[stage][edge], st);
        if (edge == 1) ev_recorded[stage] = true;
    }
};

static void carve_workspace(pht_workspace& w, char* base, const pht_capacity& cap, size_t* total) {
    Carve c{base};
    const int vp = 1 << w.vp_log2;
    const size_t slots = size_t(cap.gaussians) * size_t(vp);
    const size_t inst = size_t(cap.instances);
    w.views = c.take<DevView>(PHT_MAX_VIEWS);
    w.counters = c.take<Counters>(1);
    w.key_parts = c.take<uint32_t>(2 * kKeyParts);
    w.rec = c.take<SplatRec>(slots);
    w.meta = c.take<uint2>(slots);
    w.vis_flag = c.take<uint32_t>(slots);
    w.sgrad = c.take<float>(slots * kGradStride);
    w.view_mask = c.take<uint32_t>(size_t(cap.gaussians));
    for (int k = 0; k < 3; ++k) {
        w.dkeys[k] = c.take<uint32_t>(slots);
        w.dvals[k] = c.take<uint32_t>(slots);
        w.dtiles[k] = c.take<uint32_t>(slots);
    }
    w.offsets = c.take<uint32_t>(slots);
    for (int k = 0; k < 2; ++k) {
        w.tkeys[k] = c.take<uint32_t>(inst);
        w.tvals[k] = c.take<uint32_t>(inst);
    }
    w.ranges = c.take<uint2>(size_t(cap.tiles));
    w.tile_order = c.take<uint32_t>(size_t(cap.tiles));
    w.order_hist = c.take<uint32_t>(512);
    w.final_T = c.take<float>(size_t(cap.pixels));
    w.n_contrib = c.take<uint32_t>(size_t(cap.pixels));
    w.scan_scratch = c.take<uint32_t>(kScanBlocks);
    w.hist = c.take<uint32_t>(256 * size_t(kScanBlocks));
    w.digit_total = c.take<uint32_t>(256);
    if (total) *total = (c.off + 255) & ~size_t(255);
}

static int validate_capacity(const pht_capacity* cap) {
    if (!cap) return fail(PHT_EINVAL, "capacity is null");
    if (cap->gaussians <= 0 || cap->views <= 0 || cap->views > PHT_MAX_VIEWS || cap->tiles <= 0 ||
        cap->pixels <= 0 || cap->instances <= 0)
        return fail(PHT_EINVAL, "capacity fields must be positive (views <= 32)");
    if (cap->tile_size != 8 && cap->tile_size != 16) return fail(PHT_EINVAL, "tile_size must be 8 or 16");
    const int vp = 1 << next_pow2_log2(cap->views);
    if (double(cap->gaussians) * vp >= 2147483647.0) return fail(PHT_EINVAL, "gaussians * views exceeds 2^31");
    if (cap->instances >= 4294967295LL) return fail(PHT_EINVAL, "instance capacity exceeds 2^32-1");
    if (cap->tiles >= 2147483647LL || cap->pixels >= 4294967295LL) return fail(PHT_EINVAL, "tiles/pixels too large");
    return PHT_OK;
}

extern "C" {

int pht_abi_version(void) { return PHT_ABI_VERSION; }

const char* pht_last_error(void) { return g_last_error.c_str(); }

int pht_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

size_t pht_workspace_bytes(const pht_capacity* cap) {
    if (validate_capacity(cap) != PHT_OK) return 0;
    pht_workspace w;
    w.vp_log2 = next_pow2_log2(cap->views);
    size_t total = 0;
    carve_workspace(w, nullptr, *cap, &total);
    return total;
}

int pht_workspace_create(void* device_bytes, size_t nbytes, const pht_capacity* cap, pht_workspace** out) {
    if (!out) return fail(PHT_EINVAL, "out is null");
    *out = nullptr;
    if (int r = validate_capacity(cap)) return r;
    if (!device_bytes) return fail(PHT_EINVAL, "device_bytes is null");
    if ((reinterpret_cast<uintptr_t>(device_bytes) & 255u) != 0) return fail(PHT_EINVAL, "workspace must be 256-byte aligned");
    pht_workspace* w = new pht_workspace();
    w->cap = *cap;
    w->vp_log2 = next_pow2_log2(cap->views);
    w->slot_cap = uint32_t(cap->gaussians << w->vp_log2);
    w->inst_cap = uint32_t(cap->instances);
    size_t need = 0;
    carve_workspace(*w, static_cast<char*>(device_bytes), *cap, &need);
    if (nbytes < need) {
        delete w;
        return fail(PHT_EINVAL, "workspace too small: need " + std::to_string(need) + " bytes");
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&w->host_counters), sizeof(Counters), hipHostMallocDefault) !=
        hipSuccess) {
        delete w;
        return fail(PHT_EDEVICE, "hipHostMalloc failed for the counter readback buffer");
    }
    std::memset(w->host_counters, 0, sizeof(Counters));
    *out = w;
    return PHT_OK;
}

void pht_workspace_destroy(pht_workspace* ws) {
    if (!ws) return;
    if (ws->host_counters) (void)hipHostFree(ws->host_counters);
    if (ws->join_ev) (void)hipEventDestroy(ws->join_ev);
    for (auto& pair : ws->ev)
        for (hipEvent_t& e : pair)
            if (e) (void)hipEventDestroy(e);
    delete ws;
}

// The fused L1 loss of pht_forward_l1: per-view targets and gradient outputs, the loss sum (or
// null) and the photometric weight.
struct FusedL1 {
    const float* const* targets;
    float* const* grads;
    float* loss_sum;
    float weight;
};

static int forward_impl(pht_workspace* ws, const pht_gaussians* g, const pht_camera* cams, int n_views,
                        const pht_settings* s, pht_view_outputs* out, int check_overflow, void* stream,
                        const FusedL1* l1) {
    if (!ws || !g || !cams || !s || !out) return fail(PHT_EINVAL, "null argument");
    ws->fwd_valid = false;
    if (s->require_depth)
        return fail(PHT_ENOTSUP, "median-depth / normal channels are not built (use require_depth = 0)");
    if (g->count <= 0) return fail(PHT_EINVAL, "empty model");
    if (g->count > ws->cap.gaussians) return fail(PHT_EINVAL, "more gaussians than the workspace capacity");
    if (n_views <= 0 || n_views > ws->cap.views) return fail(PHT_EINVAL, "view count outside the workspace capacity");
    if (s->tile_size != ws->cap.tile_size) return fail(PHT_EINVAL, "tile_size differs from the workspace's");
    if (s->active_sh_degree < 0 || s->active_sh_degree > 3) return fail(PHT_EINVAL, "active_sh_degree must be 0..3");
    if (!g->means || !g->quaternions) return fail(PHT_EINVAL, "means and quaternions are required");
    const bool lp = g->lp_format != PHT_LP_NONE;
    if (lp && g->lp_format != PHT_LP_BF16 && g->lp_format != PHT_LP_FP16) return fail(PHT_EINVAL, "bad lp_format");
    if (lp ? (!g->log_scales_lp || !g->opacity_logits_lp) : (!g->log_scales || !g->opacity_logits))
        return fail(PHT_EINVAL, "log_scales / opacity_logits are required");
    if (!g->colors_precomp) {
        if (lp ? !g->sh_lp : !g->sh) return fail(PHT_EINVAL, "sh (or colors_precomp) is required");
        if ((s->active_sh_degree + 1) * (s->active_sh_degree + 1) > s->sh_bases)
            return fail(PHT_EINVAL, "active_sh_degree needs more SH bases than the model stores");
    }
    if (s->batch != 32 && s->batch != 64 && s->batch != 128 && s->batch != 256)
        return fail(PHT_EINVAL, "batch must be 32, 64, 128 or 256");
    if (s->tile_size == 8 && s->batch > 128) return fail(PHT_EINVAL, "8x8 tiles support batch <= 128");
    if (s->blend_waves != 0 && s->blend_waves != 1 && s->blend_waves != 2 && s->blend_waves != 4)
        return fail(PHT_EINVAL, "blend_waves must be 1, 2 or 4");
    hipStream_t st = static_cast<hipStream_t>(stream);
    // per-view tables
    std::vector<DevView> hv(static_cast<size_t>(n_views));
    int64_t tiles = 0, pixels = 0;
    const int ts = s->tile_size;
    for (int v = 0; v < n_views; ++v) {
        const pht_camera& c = cams[v];
        if (c.width <= 0 || c.height <= 0) return fail(PHT_EINVAL, "camera dimensions must be positive");
        if (!out[v].color || !out[v].alpha) return fail(PHT_EINVAL, "color and alpha outputs are required");
        DevView& d = hv[size_t(v)];
        std::memset(&d, 0, sizeof(d));
        d.vi.cam = c;
        d.vi.grid_x = (c.width + ts - 1) / ts;
        d.vi.grid_y = (c.height + ts - 1) / ts;
        if (d.vi.grid_x >= 8192 || d.vi.grid_y >= 8192)
            return fail(PHT_EINVAL, "camera too large: at most 8191 tiles per image axis");
        d.vi.tile_base = int32_t(tiles);
        d.vi.tile_count = d.vi.grid_x * d.vi.grid_y;
        d.vi.pixel_base = pixels;
        d.color = out[v].color;
        d.alpha = out[v].alpha;
        d.radii = out[v].radii;
        d.visibility = out[v].visibility;
        if (l1) {
            if (!l1->targets[v] || !l1->grads[v]) return fail(PHT_EINVAL, "the fused L1 loss needs a target and a gradient per view");
            d.l1_target = l1->targets[v];
            d.l1_grad = l1->grads[v];
            // the separate loss kernel's scale, in the same float operations
            const float inv_p = 1.0f / float(int64_t(c.width) * c.height);
            d.l1_scale = (l1->weight * inv_p) / 3.0f;
        }
        tiles += d.vi.tile_count;
        pixels += int64_t(c.width) * c.height;
    }
    if (tiles > ws->cap.tiles || pixels > ws->cap.pixels)
        return fail(PHT_EINVAL, "views exceed the workspace's tile/pixel capacity");

    launch_upload_views(ws->views, hv.data(), n_views, st);
    ws->mark(0, 0, st);
    launch_clear(ws->counters, ws->ranges, uint32_t(tiles), ws->key_parts, st);
    PreArgs pa;
    pa.g = *g;
    pa.s = *s;
    pa.views = ws->views;
    pa.n_views = n_views;
    pa.vp_log2 = ws->vp_log2;
    pa.rec = ws->rec;
    pa.meta = ws->meta;
    pa.vis_flag = ws->vis_flag;
    pa.view_mask = ws->view_mask;
    pa.counters = ws->counters;
    pa.key_parts = ws->key_parts;
    const int64_t n_slots = g->count << ws->vp_log2;
    launch_preprocess(pa, n_slots, st);
    launch_key_range(ws->key_parts, ws->counters, st);
    ws->mark(0, 1, st);
    ws->mark(1, 0, st);
    launch_compact_visible(ws->vis_flag, ws->meta, uint32_t(n_slots), ws->dkeys[0], ws->dvals[0], ws->dtiles[0],
                           ws->scan_scratch, &ws->counters->visible, st);
    launch_depth_sort(ws->dkeys, ws->dvals, ws->dtiles, &ws->counters->visible, uint32_t(n_slots),
                      &ws->counters->key_nmin, ws->hist, ws->digit_total, st);
    const int dsel = 2; // the depth sort's result buffer
    ws->mark(1, 1, st);
    ws->mark(2, 0, st);
    launch_tile_offsets(ws->dtiles[dsel], &ws->counters->visible, uint32_t(n_slots), ws->offsets, ws->scan_scratch,
                        &ws->counters->instances, &ws->counters->overflow, ws->inst_cap, st);
    // 16-bit tile keys whenever every global tile id fits (the bench's 8 x 8160 tiles do)
    const bool narrow = tiles <= 65536;
    launch_emit_instances(ws->dvals[dsel], ws->dtiles[dsel], ws->offsets, ws->counters, uint32_t(n_slots), ws->rec,
                          ws->views, ws->vp_log2, ts, ws->inst_cap, ws->tkeys[0], ws->tvals[0], narrow, st);
    ws->mark(2, 1, st);
    ws->mark(3, 0, st);
    const int tbits = bits_for(uint64_t(tiles - 1));
    const int tsel = launch_radix_sort(ws->tkeys, ws->tvals, nullptr, &ws->counters->instances, ws->inst_cap, tbits,
                                       ws->hist, ws->digit_total, st, narrow);
    launch_ranges(ws->tkeys[tsel], ws->counters, ws->inst_cap, uint32_t(tiles), ws->ranges, narrow, st);
    launch_tile_order(ws->ranges, uint32_t(tiles), ws->tile_order, ws->order_hist, st);
    ws->mark(3, 1, st);
    BlendArgs ba;
    ba.ranges = ws->ranges;
    ba.tile_order = ws->tile_order;
    ba.inst_vals = ws->tvals[tsel];
    ba.rec = ws->rec;
    ba.views = ws->views;
    ba.n_views = n_views;
    ba.vp_log2 = ws->vp_log2;
    ba.alpha_clip = s->alpha_clip;
    ba.final_T = ws->final_T;
    ba.n_contrib = ws->n_contrib;
    ba.sgrad = ws->sgrad;
    ba.counters = ws->counters;
    ba.slot_cap = ws->slot_cap;
    ba.n_gauss = g->count;
    ba.want_refine = 0;
    ba.queue = s->blend_queue ? ws->counters->queue : nullptr;
    ba.loss_sum = l1 ? l1->loss_sum : nullptr;
    ws->mark(4, 0, st);
    launch_blend_forward(ba, int(tiles), ts, s->batch, s->blend_waves, st);
    ws->mark(4, 1, st);
    if (int r = check_hip(hipGetLastError(), "forward launch")) return r;

    ws->fwd_valid = true;
    ws->bwd_done = false;
    ws->n_views = n_views;
    ws->tile_sel = tsel;
    ws->gaussians = g->count;
    ws->tiles_total = tiles;
    ws->pixels_total = pixels;
    ws->settings = *s;
    ws->host_views = hv;
    if (check_overflow) {
        if (int r = check_hip(hipMemcpyAsync(ws->host_counters, ws->counters, sizeof(Counters), hipMemcpyDeviceToHost, st),
                              "counter readback"))
            return r;
        if (int r = check_hip(hipStreamSynchronize(st), "forward synchronise")) return r;
        if (ws->host_counters->overflow) {
            ws->fwd_valid = false;
            return fail(PHT_ECAPACITY, "instance list (" + std::to_string(ws->host_counters->instances) +
                                           ") exceeds the workspace capacity (" + std::to_string(ws->inst_cap) + ")");
        }
    }
    return PHT_OK;
}

int pht_forward(pht_workspace* ws, const pht_gaussians* g, const pht_camera* cams, int n_views,
                const pht_settings* s, pht_view_outputs* out, int check_overflow, void* stream) {
    return forward_impl(ws, g, cams, n_views, s, out, check_overflow, stream, nullptr);
}

int pht_forward_l1(pht_workspace* ws, const pht_gaussians* g, const pht_camera* cams, int n_views,
                   const pht_settings* s, pht_view_outputs* out, const float* const* targets,
                   float* const* grad_color, float* loss_sum, float weight, int check_overflow, void* stream) {
    if (!targets || !grad_color) return fail(PHT_EINVAL, "targets and grad_color are required");
    const FusedL1 l1{targets, grad_color, loss_sum, weight};
    return forward_impl(ws, g, cams, n_views, s, out, check_overflow, stream, &l1);
}

int pht_backward(pht_workspace* ws, const pht_gaussians* g, const pht_view_grads* dl, pht_model_grads* grads,
                 const pht_adam* adam, void* stream) {
    return pht_backward_densify(ws, g, dl, grads, adam, nullptr, stream);
}

// Argument checks shared by the backward entry points: a completed forward of this model on the
// workspace, and either gradient outputs or a complete fused-Adam description.
static int check_backward(const pht_workspace* ws, const pht_gaussians* g, const pht_model_grads* grads,
                          const pht_adam* adam) {
    if (!ws->fwd_valid) return fail(PHT_ESTATE, "backward requires a completed forward on this workspace");
    if (g->count != ws->gaussians) return fail(PHT_EINVAL, "model size differs from the forward's");
    if (!grads && !adam) return fail(PHT_EINVAL, "either gradients or an Adam update are required");
    if (adam) {
        if (!adam->means || !adam->log_scales || !adam->quaternions || !adam->opacity_logits || !adam->m_means ||
            !adam->v_means || !adam->m_log_scales || !adam->v_log_scales || !adam->m_quaternions ||
            !adam->v_quaternions || !adam->m_opacity_logits || !adam->v_opacity_logits)
            return fail(PHT_EINVAL, "fused Adam requires every structure parameter and moment");
        if (adam->sh && !adam->sh_first && (!adam->m_sh || !adam->v_sh))
            return fail(PHT_EINVAL, "SH Adam requires its moments");
        if (adam->sh && adam->sh_first && (!adam->sh_packed || !adam->sh_bounds))
            return fail(PHT_EINVAL, "quantised SH Adam requires sh_first, sh_packed and sh_bounds");
        if (adam->lp_format != g->lp_format) return fail(PHT_EINVAL, "Adam lp_format differs from the model's");
    }
    return PHT_OK;
}

// The blend backward of one workspace's views into its splat-gradient records (dl: one entry per
// view of the workspace's forward).
static int blend_backward_part(pht_workspace* ws, const pht_view_grads* dl, bool want_refine, hipStream_t st) {
    std::vector<DevView> hv = ws->host_views;
    for (int v = 0; v < ws->n_views; ++v) {
        if (!dl[v].grad_color) return fail(PHT_EINVAL, "grad_color is required for every view");
        hv[size_t(v)].grad_color = dl[v].grad_color;
        hv[size_t(v)].grad_alpha = dl[v].grad_alpha;
    }
    launch_upload_views(ws->views, hv.data(), ws->n_views, st);
    BlendArgs ba;
    ba.ranges = ws->ranges;
    ba.tile_order = ws->tile_order;
    ba.inst_vals = ws->tvals[ws->tile_sel];
    ba.rec = ws->rec;
    ba.views = ws->views;
    ba.n_views = ws->n_views;
    ba.vp_log2 = ws->vp_log2;
    ba.alpha_clip = ws->settings.alpha_clip;
    ba.final_T = ws->final_T;
    ba.n_contrib = ws->n_contrib;
    ba.sgrad = ws->sgrad;
    ba.counters = ws->counters;
    ba.slot_cap = ws->slot_cap;
    ba.n_gauss = ws->gaussians;
    ba.want_refine = want_refine ? 1 : 0;
    ba.queue = ws->settings.blend_queue ? ws->counters->queue : nullptr;
    ba.loss_sum = nullptr;
    ws->mark(5, 0, st);
    if (ws->bwd_done)
        launch_zero_f32(ws->sgrad, int64_t(ws->n_views) * ws->gaussians * kGradStride, st);
    ws->bwd_done = true;
    launch_blend_backward(ba, int(ws->tiles_total), ws->settings.tile_size, ws->settings.batch,
                          ws->settings.blend_waves, st);
    ws->mark(5, 1, st);
    return PHT_OK;
}

static GaussGroup gauss_group(const pht_workspace* ws) {
    GaussGroup gr;
    gr.views = ws->views;
    gr.n_views = ws->n_views;
    gr.vp_log2 = ws->vp_log2;
    gr.meta = ws->meta;
    gr.sgrad = ws->sgrad;
    gr.view_mask = ws->view_mask;
    gr.counters = ws->counters;
    return gr;
}

int pht_backward_densify(pht_workspace* ws, const pht_gaussians* g, const pht_view_grads* dl, pht_model_grads* grads,
                         const pht_adam* adam, const pht_densify_step* densify, void* stream) {
    if (!ws || !g || !dl) return fail(PHT_EINVAL, "null argument");
    if (int r = check_backward(ws, g, grads, adam)) return r;
    DensifyStepArgs da{};
    if (densify) {
        const pht_densify_stats& d = densify->stats;
        if (!d.gradient || !d.count || !d.max_screen_radius || !d.priority || !d.max_geometry_gradient ||
            !d.first_view || !d.view_support || !d.accumulated_error)
            return fail(PHT_EINVAL, "densify: null statistics array");
        for (int v = 0; v < ws->n_views; ++v) {
            const DevView& hv = ws->host_views[size_t(v)];
            if (!hv.radii || !hv.visibility)
                return fail(PHT_EINVAL, "densify: the forward must write radii and visibility for every view");
            const int32_t side = std::max<int32_t>(1, std::min(hv.vi.cam.width, hv.vi.cam.height));
            da.inverse_resolution[v] = 1.0f / float(side);
            da.view_index[v] = densify->view_index[v];
        }
        da.stats = d;
        da.rules.use_maximum = densify->use_maximum;
        da.rules.require_contribution_visibility = densify->require_contribution_visibility;
        da.rules.step_score_power = densify->step_score_power;
        da.rules.oversize_screen_threshold = densify->oversize_screen_threshold;
        da.geometry_from_refine = densify->geometry_from_refine;
        da.views = ws->views;
        da.n_views = ws->n_views;
        da.sgrad = ws->sgrad;
        da.n = ws->gaussians;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (int r = blend_backward_part(ws, dl, (grads && grads->refine_weight) || densify, st)) return r;
    GaussArgs ga;
    ga.g = *g;
    ga.s = ws->settings;
    ga.n_groups = 1;
    ga.grp[0] = gauss_group(ws);
    std::memset(&ga.out, 0, sizeof(ga.out));
    if (grads) ga.out = *grads;
    std::memset(&ga.adam, 0, sizeof(ga.adam));
    if (adam) ga.adam = *adam;
    ga.dens = da;
    ga.has_dens = densify && adam ? 1 : 0; // folded into the fused-Adam kernel
    ws->mark(6, 0, st);
    launch_gaussian_backward(ga, adam != nullptr, st);
    if (densify && !adam) launch_densify_step(da, st);
    ws->mark(6, 1, st);
    return check_hip(hipGetLastError(), "backward launch");
}

int pht_backward_groups(pht_workspace* const* ws, int n_groups, const pht_gaussians* g, const pht_view_grads* dl,
                        pht_model_grads* grads, const pht_adam* adam, void* const* group_streams, void* stream) {
    if (!ws || !g || !dl || !group_streams) return fail(PHT_EINVAL, "null argument");
    if (n_groups < 1 || n_groups > kMaxViewGroups) return fail(PHT_EINVAL, "n_groups must be 1..4");
    for (int k = 0; k < n_groups; ++k) {
        if (!ws[k]) return fail(PHT_EINVAL, "null workspace");
        if (int r = check_backward(ws[k], g, grads, adam)) return r;
        if (ws[k]->settings.active_sh_degree != ws[0]->settings.active_sh_degree ||
            ws[k]->settings.sh_bases != ws[0]->settings.sh_bases)
            return fail(PHT_EINVAL, "view groups must share the SH settings");
        for (int j = 0; j < k; ++j)
            if (ws[j] == ws[k]) return fail(PHT_EINVAL, "each view group needs its own workspace");
    }
    const bool want_refine = grads && grads->refine_weight;
    hipStream_t st = static_cast<hipStream_t>(stream);
    GaussArgs ga;
    ga.g = *g;
    ga.s = ws[0]->settings;
    ga.n_groups = n_groups;
    int first_view = 0;
    for (int k = 0; k < n_groups; ++k) {
        hipStream_t gs = static_cast<hipStream_t>(group_streams[k]);
        if (int r = blend_backward_part(ws[k], dl + first_view, want_refine, gs)) return r;
        first_view += ws[k]->n_views;
        // the per-Gaussian backward on `stream` waits for this group's blend backward
        if (gs != st) {
            if (!ws[k]->join_ev)
                if (int r = check_hip(hipEventCreateWithFlags(&ws[k]->join_ev, hipEventDisableTiming), "event"))
                    return r;
            if (int r = check_hip(hipEventRecord(ws[k]->join_ev, gs), "join record")) return r;
            if (int r = check_hip(hipStreamWaitEvent(st, ws[k]->join_ev, 0), "join wait")) return r;
        }
        ga.grp[k] = gauss_group(ws[k]);
    }
    std::memset(&ga.out, 0, sizeof(ga.out));
    if (grads) ga.out = *grads;
    std::memset(&ga.adam, 0, sizeof(ga.adam));
    if (adam) ga.adam = *adam;
    std::memset(&ga.dens, 0, sizeof(ga.dens));
    ga.has_dens = 0;
    ws[0]->mark(6, 0, st);
    launch_gaussian_backward(ga, adam != nullptr, st);
    ws[0]->mark(6, 1, st);
    return check_hip(hipGetLastError(), "backward launch");
}

int pht_set_profiling(pht_workspace* ws, int enable) {
    if (!ws) return fail(PHT_EINVAL, "null workspace");
    ws->profiling = enable != 0;
    for (bool& r : ws->ev_recorded) r = false;
    return PHT_OK;
}

int pht_stage_times(pht_workspace* ws, float* ms_out, int max_stages) {
    if (!ws || !ms_out) return fail(PHT_EINVAL, "null argument");
    const int n = max_stages < PHT_STAGES ? max_stages : PHT_STAGES;
    for (int s = 0; s < n; ++s) {
        ms_out[s] = -1.0f;
        if (!ws->ev_recorded[s]) continue;
        if (int r = check_hip(hipEventSynchronize(ws->ev[s][1]), "event synchronise")) return r;
        float ms = 0.0f;
        if (int r = check_hip(hipEventElapsedTime(&ms, ws->ev[s][0], ws->ev[s][1]), "event elapsed")) return r;
        ms_out[s] = ms;
    }
    return n;
}

int pht_l1_loss(const float* color, const float* target, float* grad_color, float* loss_sum, int64_t pixels,
                float weight, void* stream) {
    if (!color || !target || !grad_color || pixels <= 0) return fail(PHT_EINVAL, "bad L1 arguments");
    launch_l1_loss(color, target, grad_color, loss_sum, pixels, weight, static_cast<hipStream_t>(stream));
    return check_hip(hipGetLastError(), "l1 launch");
}

size_t pht_l1_ssim_scratch_bytes(int width, int height) {
    if (width <= 10 || height <= 10) return 0;
    return l1_ssim_scratch_floats(width, height) * sizeof(float);
}

int pht_l1_ssim_loss(const float* color, const float* target, const float* mask, float* grad_color, int accumulate,
                     float* loss_sum, void* scratch, int width, int height, float ssim_weight,
                     float photometric_weight, void* stream) {
    if (!color || !target || !grad_color) return fail(PHT_EINVAL, "color, target and grad_color are required");
    if (width <= 10 || height <= 10) return fail(PHT_EINVAL, "fused SSIM requires width and height > 10");
    if (loss_sum && !scratch) return fail(PHT_EINVAL, "loss_sum needs scratch (pht_l1_ssim_scratch_bytes)");
    launch_l1_ssim(color, target, mask, grad_color, accumulate, loss_sum, static_cast<float*>(scratch), width, height,
                   ssim_weight, photometric_weight, static_cast<hipStream_t>(stream));
    return check_hip(hipGetLastError(), "l1+ssim launch");
}

int pht_l1_ssim_loss_noalloc(const float* color, const float* target, const float* mask, float* grad_color,
                             int accumulate, float* loss_sum, int width, int height, float ssim_weight,
                             float photometric_weight, void* stream) {
    if (!color || !target || !grad_color) return fail(PHT_EINVAL, "color, target and grad_color are required");
    if (width <= 10 || height <= 10) return fail(PHT_EINVAL, "fused SSIM requires width and height > 10");
    if (l1_ssim_scratch_floats(width, height) > size_t(PHT_NOALLOC_SSIM_TILES))
        return fail(PHT_EINVAL, "view too large for the no-allocation L1+SSIM loss: ceil(W/16) * ceil(H/16) must "
                                "be at most " + std::to_string(PHT_NOALLOC_SSIM_TILES));
    float* partials = nullptr;
    if (loss_sum) {
        hipError_t err = hipSuccess;
        partials = l1_ssim_static_scratch(&err);
        if (!partials) return check_hip(err == hipSuccess ? hipErrorInvalidSymbol : err, "static scratch address");
    }
    launch_l1_ssim(color, target, mask, grad_color, accumulate, loss_sum, partials, width, height, ssim_weight,
                   photometric_weight, static_cast<hipStream_t>(stream));
    return check_hip(hipGetLastError(), "l1+ssim launch");
}

int pht_adam_step(float* param, const float* grad, float* m, float* v, int64_t rows, int64_t row_stride,
                  int64_t active_cols, float lr, float lr_secondary, int64_t secondary_from_col, float beta1,
                  float beta2, float correction1, float correction2, float epsilon, float clamp_min, float clamp_max,
                  int32_t prefix_semantics, void* stream) {
    if (!param || !grad || !m || !v || rows < 0 || row_stride <= 0 || active_cols < 0 || active_cols > row_stride)
        return fail(PHT_EINVAL, "bad Adam arguments");
    launch_adam_standalone(param, grad, m, v, rows, row_stride, active_cols, lr, lr_secondary, secondary_from_col,
                           beta1, beta2, correction1, correction2, epsilon, clamp_min, clamp_max, prefix_semantics,
                           static_cast<hipStream_t>(stream));
    return check_hip(hipGetLastError(), "adam launch");
}

int pht_sh_adam_quant_step(float* parameter, const float* gradient, uint16_t* first, uint8_t* packed, float* bounds,
                           int64_t rows, int32_t stride, int32_t active, float lr, float rest_lr, float beta1,
                           float beta2, float correction1, float correction2, float epsilon, float regularization,
                           void* stream) {
    if (rows < 0 || stride < 1 || stride > kShqMaxStride || active < 0 || active > stride)
        return fail(PHT_EINVAL, "quantised SH Adam needs rows >= 0, stride in [1, 48] and active in [0, stride]");
    if (rows > 0 && (!parameter || !first || !packed || !bounds || (active > 0 && !gradient)))
        return fail(PHT_EINVAL, "null quantised SH Adam array");
    launch_sh_adam_quant(parameter, gradient, first, packed, bounds, rows, stride,
                         ShqStep{lr, rest_lr, beta1, beta2, correction1, correction2, epsilon, regularization, active},
                         static_cast<hipStream_t>(stream));
    return check_hip(hipGetLastError(), "quantised SH Adam launch");
}

// ---------------------------------------------------------------- densification
int pht_densify_accumulate(const pht_densify_stats* stats, const pht_densify_observation* obs, int n_obs, int64_t n,
                           void* stream) {
    if (!stats || (n_obs > 0 && !obs) || n_obs < 0 || n_obs > PHT_MAX_VIEWS || n < 0)
        return fail(PHT_EINVAL, "densify_accumulate: bad arguments (at most 32 observations)");
    const pht_densify_stats& s = *stats;
    if (n > 0 && (!s.gradient || !s.count || !s.max_screen_radius || !s.priority || !s.max_geometry_gradient ||
                  !s.first_view || !s.view_support || !s.accumulated_error))
        return fail(PHT_EINVAL, "densify_accumulate: null statistics array");
    DensifyObsSet set{};
    set.stats = s;
    set.n = n_obs;
    for (int v = 0; v < n_obs; ++v) {
        const pht_densify_observation& o = obs[v];
        if (n > 0 && (!o.refine_weight || !o.visibility || !o.radii))
            return fail(PHT_EINVAL, "densify_accumulate: refine_weight, visibility and radii are required");
        set.obs[v] = o;
        const int32_t side = std::max<int32_t>(1, std::min(o.width, o.height));
        set.inverse_resolution[v] = 1.0f / float(side);
    }
    launch_densify_accumulate(set, n, static_cast<hipStream_t>(stream));
    return check_hip(hipGetLastError(), "densify_accumulate launch");
}

int pht_densify_prune_masks(const pht_gaussians* g, int32_t sh_stride, float minimum_opacity, float maximum_bounds,
                            const float center[3], uint8_t* keep, uint8_t* hard, float* opacities, void* stream) {
    if (!g || !center || g->count < 0 || sh_stride < 0) return fail(PHT_EINVAL, "prune_masks: bad arguments");
    if (g->count > 0 && (!g->means || !g->log_scales || !g->quaternions || !g->opacity_logits ||
                         (sh_stride > 0 && !g->sh) || !keep || !hard || !opacities))
        return fail(PHT_EINVAL, "prune_masks: null array");
    launch_prune_masks(*g, sh_stride, minimum_opacity, maximum_bounds, center, keep, hard, opacities,
                       static_cast<hipStream_t>(stream));
    return check_hip(hipGetLastError(), "prune_masks launch");
}

size_t pht_select_workspace_bytes(int64_t n) { return select_workspace_bytes(n); }

float pht_select_key_value(uint64_t key, int mode) {
    const uint32_t hi = uint32_t(key >> 32);
    return float_from_order_bits(mode == PHT_SELECT_SMALLEST ? ~hi : hi);
}

int pht_select_rows(const float* values, const uint8_t* mask, const uint8_t* exclude, int64_t n, int64_t k, int mode,
                    uint64_t seed, uint8_t* out, uint8_t out_value, uint32_t* counts, uint64_t* threshold,
                    void* workspace, size_t workspace_bytes, void* stream) {
    if (n < 0 || n > int64_t(0xfffffffe) || mode < PHT_SELECT_LARGEST || mode > PHT_SELECT_LARGEST_POSITIVE)
        return fail(PHT_EINVAL, "select_rows: n must be in [0, 2^32 - 2] and mode one of PHT_SELECT_*");
    if (!workspace || workspace_bytes < select_workspace_bytes(n))
        return fail(PHT_EINVAL, "select_rows: workspace smaller than pht_select_workspace_bytes(n)");
    if (n > 0 && (!values || !out)) return fail(PHT_EINVAL, "select_rows: null values or out");
    launch_select_rows(values, mask, exclude, n, k, mode, seed, out, out_value, counts, threshold, workspace,
                       static_cast<hipStream_t>(stream));
    return check_hip(hipGetLastError(), "select_rows launch");
}

int pht_igs_prepare(const pht_densify_stats* stats, const int32_t* keep_rows, int64_t retained,
                    float gradient_threshold, float screen_threshold, float* priority, uint8_t* positive,
                    uint32_t* counts, void* stream) {
    if (!stats || retained < 0 || !counts) return fail(PHT_EINVAL, "igs_prepare: bad arguments");
    if (retained > 0 && (!stats->count || !stats->priority || !stats->gradient || !stats->max_screen_radius ||
                         !priority || !positive))
        return fail(PHT_EINVAL, "igs_prepare: null array");
    launch_igs_prepare(*stats, keep_rows, retained, gradient_threshold, screen_threshold, priority, positive, counts,
                       static_cast<hipStream_t>(stream));
    return check_hip(hipGetLastError(), "igs_prepare launch");
}

int pht_igs_weights(const pht_densify_stats* stats, const int32_t* keep_rows, int64_t retained,
                    const float* opacities, float median, int32_t allow_growth,
                    float gradient_threshold, float screen_threshold, float* replacement, float* growth,
                    float* oversize, void* stream) {
    if (!stats || retained < 0) return fail(PHT_EINVAL, "igs_weights: bad arguments");
    if (retained > 0 && (!opacities || !replacement || !stats->count || !stats->priority || !stats->gradient ||
                         !stats->max_screen_radius || (allow_growth && (!growth || !oversize))))
        return fail(PHT_EINVAL, "igs_weights: null array");
    launch_igs_weights(*stats, keep_rows, retained, opacities, median, allow_growth, gradient_threshold,
                       screen_threshold, replacement, growth, oversize, static_cast<hipStream_t>(stream));
    return check_hip(hipGetLastError(), "igs_weights launch");
}

int pht_gather_rows(const pht_row_buffer* buffers, int n_buffers, const int32_t* rows, int64_t n_rows,
                    void* stream) {
    if (n_buffers < 0 || n_buffers > PHT_MAX_ROW_BUFFERS || n_rows < 0 || (n_buffers > 0 && !buffers) ||
        (n_rows > 0 && !rows))
        return fail(PHT_EINVAL, "gather_rows: bad arguments (at most 24 buffers)");
    RowBuffers set{};
    set.n = n_buffers;
    for (int i = 0; i < n_buffers; ++i) {
        const pht_row_buffer& b = buffers[i];
        if (b.row_bytes <= 0 || (b.row_bytes & 1) || !b.dst || (n_rows > 0 && !b.src))
            return fail(PHT_EINVAL, "gather_rows: each buffer needs src, dst and a positive even row size");
        const uintptr_t a = reinterpret_cast<uintptr_t>(b.src) | reinterpret_cast<uintptr_t>(b.dst);
        int64_t unit = 2;
        if (b.row_bytes % 16 == 0 && a % 16 == 0) unit = 16;
        else if (b.row_bytes % 4 == 0 && a % 4 == 0) unit = 4;
        else if (a % 2 != 0) return fail(PHT_EINVAL, "gather_rows: buffers must be 2-byte aligned");
        set.b[i] = RowBuffer{b.src, b.dst, b.row_bytes, unit};
    }
    launch_gather_rows(set, rows, n_rows, static_cast<hipStream_t>(stream));
    return check_hip(hipGetLastError(), "gather_rows launch");
}

int pht_densify_split(const pht_split_params* p, const int32_t* parents, int64_t count, int64_t child_base,
                      const float* samples, const float* screen_sizes, void* stream) {
    if (!p || count < 0 || child_base < 0 || p->mode < PHT_SPLIT_DENSE_TANGENT || p->mode > PHT_SPLIT_LONG_AXIS)
        return fail(PHT_EINVAL, "densify_split: bad arguments");
    if (count > 0 && (!parents || !samples || !p->means || !p->log_scales || !p->opacity_logits || !p->quaternions))
        return fail(PHT_EINVAL, "densify_split: null array");
    if (p->lp_format != PHT_LP_NONE && (!p->log_scales_lp || !p->opacity_logits_lp))
        return fail(PHT_EINVAL, "densify_split: lp_format set without reduced-precision copies");
    launch_densify_split(*p, parents, count, child_base, samples, screen_sizes, static_cast<hipStream_t>(stream));
    return check_hip(hipGetLastError(), "densify_split launch");
}

int pht_adc_decay(float* log_scales, float* opacity_logits, uint16_t* log_scales_lp, uint16_t* opacity_logits_lp,
                  int32_t lp_format, int64_t n, float opacity_decay, float log_scale_decay, void* stream) {
    if (n < 0 || (n > 0 && (!log_scales || !opacity_logits)) ||
        (lp_format != PHT_LP_NONE && n > 0 && (!log_scales_lp || !opacity_logits_lp)))
        return fail(PHT_EINVAL, "adc_decay: bad arguments");
    launch_adc_decay(log_scales, opacity_logits, log_scales_lp, opacity_logits_lp, lp_format, n, opacity_decay,
                     log_scale_decay, static_cast<hipStream_t>(stream));
    return check_hip(hipGetLastError(), "adc_decay launch");
}

namespace {
// MT19937 (Matsumoto & Nishimura 1998) over a caller-held state of 624 words and a position, the
// state layout of std::mt19937 and of numpy's MT19937; a UniformRandomBitGenerator, so the
// standard library's distributions consume it exactly as they consume std::mt19937.
struct Mt19937State {
    using result_type = uint32_t;
    uint32_t* key;
    int32_t pos;
    static constexpr result_type min() { return 0u; }
    static constexpr result_type max() { return 0xffffffffu; }
    void regenerate() {
        for (int i = 0; i < 624; ++i) {
            const uint32_t y = (key[i] & 0x80000000u) | (key[(i + 1) % 624] & 0x7fffffffu);
            key[i] = key[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        pos = 0;
    }
    result_type operator()() {
        if (pos >= 624) regenerate();
        uint32_t y = key[pos++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        return y ^ (y >> 18);
    }
};
} // namespace

int pht_host_normal_floats(uint32_t* mt_key, int32_t* mt_pos, float* out, int64_t count) {
    if (!mt_key || !mt_pos || count < 0 || (count > 0 && !out) || *mt_pos < 0 || *mt_pos > 624)
        return fail(PHT_EINVAL, "host_normal_floats: bad arguments");
    Mt19937State eng{mt_key, *mt_pos};
    std::normal_distribution<float> normal(0.0f, 1.0f);
    for (int64_t i = 0; i < count; ++i) out[i] = normal(eng);
    *mt_pos = eng.pos;
    return PHT_OK;
}

int pht_get_counts(pht_workspace* ws, pht_counts* out, void* stream) {
    if (!ws || !out) return fail(PHT_EINVAL, "null argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (int r = check_hip(hipMemcpyAsync(ws->host_counters, ws->counters, sizeof(Counters), hipMemcpyDeviceToHost, st),
                          "counter readback"))
        return r;
    if (int r = check_hip(hipStreamSynchronize(st), "counter synchronise")) return r;
    out->visible = ws->host_counters->visible;
    out->instances = ws->host_counters->instances;
    out->fp64_conics = ws->host_counters->fp64;
    out->overflow = int32_t(ws->host_counters->overflow);
    out->reserved = 0;
    return PHT_OK;
}

int pht_get_counts_async(pht_workspace* ws, uint32_t* dst, void* stream) {
    if (!ws || !dst) return fail(PHT_EINVAL, "null argument");
    static_assert(offsetof(Counters, visible) == 0 && offsetof(Counters, instances) == 4 &&
                      offsetof(Counters, overflow) == 8 && offsetof(Counters, fp64) == 12,
                  "the first four counters are copied as one block");
    return check_hip(hipMemcpyAsync(dst, ws->counters, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                    static_cast<hipStream_t>(stream)),
                     "counter copy");
}

int pht_debug_lists(pht_workspace* ws, uint64_t* keys, uint32_t* values, uint32_t* ranges, uint32_t* n_contrib,
                    float* final_T, void* stream) {
    if (!ws) return fail(PHT_EINVAL, "null workspace");
    if (!ws->fwd_valid) return fail(PHT_ESTATE, "no forward recorded");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (keys || values)
        launch_debug_keys(ws->tkeys[ws->tile_sel], ws->tvals[ws->tile_sel], ws->meta, ws->counters, ws->inst_cap,
                          ws->vp_log2, keys, values, ws->tiles_total <= 65536, st);
    if (ranges)
        if (int r = check_hip(hipMemcpyAsync(ranges, ws->ranges, size_t(ws->tiles_total) * sizeof(uint2),
                                             hipMemcpyDeviceToDevice, st), "ranges copy"))
            return r;
    if (n_contrib)
        if (int r = check_hip(hipMemcpyAsync(n_contrib, ws->n_contrib, size_t(ws->pixels_total) * 4,
                                             hipMemcpyDeviceToDevice, st), "n_contrib copy"))
            return r;
    if (final_T)
        if (int r = check_hip(hipMemcpyAsync(final_T, ws->final_T, size_t(ws->pixels_total) * 4,
                                             hipMemcpyDeviceToDevice, st), "final_T copy"))
            return r;
    return check_hip(hipGetLastError(), "debug lists");
}

int pht_debug_splat_grads(pht_workspace* ws, float* out, void* stream) {
    if (!ws || !out) return fail(PHT_EINVAL, "null argument");
    if (!ws->fwd_valid) return fail(PHT_ESTATE, "no forward recorded");
    launch_debug_splat_grads(ws->sgrad, ws->view_mask, ws->gaussians, ws->n_views, ws->vp_log2, out,
                             static_cast<hipStream_t>(stream));
    return check_hip(hipGetLastError(), "debug splat grads");
}

// ---------------------------------------------------------------- whole-step graph
struct pht_step_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int pht_step_graph_create(pht_workspace* ws, const pht_gaussians* g, const pht_camera* cams, int n_views,
                          const pht_settings* s, pht_view_outputs* out, const float* const* targets,
                          float* const* grad_color_bufs, float* loss_sum, const pht_loss_config* loss,
                          const pht_adam* adam, void* stream, pht_step_graph** out_graph) {
    return pht_step_graph_create_densify(ws, g, cams, n_views, s, out, targets, grad_color_bufs, loss_sum, loss, adam,
                                         nullptr, stream, out_graph);
}

int pht_step_graph_create_densify(pht_workspace* ws, const pht_gaussians* g, const pht_camera* cams, int n_views,
                                  const pht_settings* s, pht_view_outputs* out, const float* const* targets,
                                  float* const* grad_color_bufs, float* loss_sum, const pht_loss_config* loss,
                                  const pht_adam* adam, const pht_densify_step* densify, void* stream,
                                  pht_step_graph** out_graph) {
    if (!out_graph || !targets || !grad_color_bufs || !adam) return fail(PHT_EINVAL, "null argument");
    pht_loss_config lc{};
    lc.kind = PHT_LOSS_L1;
    lc.photometric_weight = 1.0f;
    if (loss) lc = *loss;
    if (lc.kind != PHT_LOSS_L1 && lc.kind != PHT_LOSS_L1_SSIM) return fail(PHT_EINVAL, "unknown loss kind");
    if (lc.kind == PHT_LOSS_L1_SSIM && loss_sum && !lc.ssim_scratch)
        return fail(PHT_EINVAL, "the L1+SSIM loss sum needs per-view scratch");
    *out_graph = nullptr;
    hipStream_t cap_stream = nullptr;
    if (int r = check_hip(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking), "capture stream")) return r;
    (void)stream;
    pht_step_graph* sg = new pht_step_graph();
    const bool was_profiling = ws && ws->profiling;
    if (ws) ws->profiling = false;
    int rc = check_hip(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc == PHT_OK) {
        if (loss_sum) launch_zero_f32(loss_sum, 1, cap_stream);
        // the L1 loss is fused into the forward's blend; L1+SSIM runs as its own kernel per view
        if (lc.kind == PHT_LOSS_L1) {
            const FusedL1 l1{targets, grad_color_bufs, loss_sum, lc.photometric_weight};
            rc = forward_impl(ws, g, cams, n_views, s, out, 0, cap_stream, &l1);
        } else {
            rc = pht_forward(ws, g, cams, n_views, s, out, 0, cap_stream);
        }
        std::vector<pht_view_grads> dl(static_cast<size_t>(n_views));
        for (int v = 0; rc == PHT_OK && v < n_views; ++v) {
            if (lc.kind == PHT_LOSS_L1_SSIM)
                rc = pht_l1_ssim_loss(out[v].color, targets[v], nullptr, grad_color_bufs[v], 0, loss_sum,
                                      loss_sum ? lc.ssim_scratch[v] : nullptr, cams[v].width, cams[v].height,
                                      lc.ssim_weight, lc.photometric_weight, cap_stream);
            dl[size_t(v)].grad_color = grad_color_bufs[v];
            dl[size_t(v)].grad_alpha = nullptr;
        }
        if (rc == PHT_OK) rc = pht_backward_densify(ws, g, dl.data(), nullptr, adam, densify, cap_stream);
        hipGraph_t graph = nullptr;
        const int rc_end = check_hip(hipStreamEndCapture(cap_stream, &graph), "end capture");
        if (rc == PHT_OK) rc = rc_end;
        sg->graph = graph;
        if (rc == PHT_OK) rc = check_hip(hipGraphInstantiate(&sg->exec, graph, nullptr, nullptr, 0), "instantiate");
    }
    (void)hipStreamDestroy(cap_stream);
    if (ws) ws->profiling = was_profiling;
    if (rc != PHT_OK) {
        pht_step_graph_destroy(sg);
        return rc;
    }
    *out_graph = sg;
    return PHT_OK;
}

int pht_step_graph_create_groups(pht_workspace* const* ws, int n_groups, const int* group_views,
                                 const pht_gaussians* g, const pht_camera* cams, int n_views, const pht_settings* s,
                                 pht_view_outputs* out, const float* const* targets, float* const* grad_color_bufs,
                                 float* loss_sum, const pht_loss_config* loss, const pht_adam* adam, void* stream,
                                 pht_step_graph** out_graph) {
    if (!out_graph || !ws || !group_views || !cams || !out || !targets || !grad_color_bufs || !adam)
        return fail(PHT_EINVAL, "null argument");
    if (n_groups < 1 || n_groups > kMaxViewGroups) return fail(PHT_EINVAL, "n_groups must be 1..4");
    int total = 0;
    for (int k = 0; k < n_groups; ++k) {
        if (!ws[k] || group_views[k] <= 0) return fail(PHT_EINVAL, "every view group needs a workspace and views");
        total += group_views[k];
    }
    if (total != n_views) return fail(PHT_EINVAL, "the view groups must cover the step's views");
    pht_loss_config lc{};
    lc.kind = PHT_LOSS_L1;
    lc.photometric_weight = 1.0f;
    if (loss) lc = *loss;
    if (lc.kind != PHT_LOSS_L1 && lc.kind != PHT_LOSS_L1_SSIM) return fail(PHT_EINVAL, "unknown loss kind");
    if (lc.kind == PHT_LOSS_L1_SSIM && loss_sum && !lc.ssim_scratch)
        return fail(PHT_EINVAL, "the L1+SSIM loss sum needs per-view scratch");
    *out_graph = nullptr;
    (void)stream;
    // the capture stream forks into one stream per group (default priority) and joins back
    hipStream_t cap_stream = nullptr;
    hipStream_t gs[kMaxViewGroups] = {};
    hipEvent_t fork = nullptr;
    int rc = check_hip(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking), "capture stream");
    for (int k = 0; rc == PHT_OK && k < n_groups; ++k)
        rc = check_hip(hipStreamCreateWithFlags(&gs[k], hipStreamNonBlocking), "group stream");
    if (rc == PHT_OK) rc = check_hip(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "fork event");
    pht_step_graph* sg = new pht_step_graph();
    bool was_profiling[kMaxViewGroups] = {};
    for (int k = 0; k < n_groups; ++k) {
        was_profiling[k] = ws[k]->profiling;
        ws[k]->profiling = false;
    }
    if (rc == PHT_OK) rc = check_hip(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc == PHT_OK) {
        if (loss_sum) launch_zero_f32(loss_sum, 1, cap_stream);
        rc = check_hip(hipEventRecord(fork, cap_stream), "fork record");
        std::vector<pht_view_grads> dl(static_cast<size_t>(n_views));
        int v0 = 0;
        for (int k = 0; rc == PHT_OK && k < n_groups; ++k) {
            rc = check_hip(hipStreamWaitEvent(gs[k], fork, 0), "fork wait");
            const int nk = group_views[k];
            if (rc == PHT_OK && lc.kind == PHT_LOSS_L1) {
                const FusedL1 l1{targets + v0, grad_color_bufs + v0, loss_sum, lc.photometric_weight};
                rc = forward_impl(ws[k], g, cams + v0, nk, s, out + v0, 0, gs[k], &l1);
            } else if (rc == PHT_OK) {
                rc = pht_forward(ws[k], g, cams + v0, nk, s, out + v0, 0, gs[k]);
                for (int v = v0; rc == PHT_OK && v < v0 + nk; ++v)
                    rc = pht_l1_ssim_loss(out[v].color, targets[v], nullptr, grad_color_bufs[v], 0, loss_sum,
                                          loss_sum ? lc.ssim_scratch[v] : nullptr, cams[v].width, cams[v].height,
                                          lc.ssim_weight, lc.photometric_weight, gs[k]);
            }
            for (int v = v0; v < v0 + nk; ++v) {
                dl[size_t(v)].grad_color = grad_color_bufs[v];
                dl[size_t(v)].grad_alpha = nullptr;
            }
            v0 += nk;
        }
        void* gsv[kMaxViewGroups];
        for (int k = 0; k < n_groups; ++k) gsv[k] = gs[k];
        if (rc == PHT_OK) rc = pht_backward_groups(ws, n_groups, g, dl.data(), nullptr, adam, gsv, cap_stream);
        hipGraph_t graph = nullptr;
        const int rc_end = check_hip(hipStreamEndCapture(cap_stream, &graph), "end capture");
        if (rc == PHT_OK) rc = rc_end;
        sg->graph = graph;
        if (rc == PHT_OK) rc = check_hip(hipGraphInstantiate(&sg->exec, graph, nullptr, nullptr, 0), "instantiate");
    }
    for (int k = 0; k < n_groups; ++k) {
        ws[k]->profiling = was_profiling[k];
        if (gs[k]) (void)hipStreamDestroy(gs[k]);
    }
    if (fork) (void)hipEventDestroy(fork);
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    if (rc != PHT_OK) {
        pht_step_graph_destroy(sg);
        return rc;
    }
    *out_graph = sg;
    return PHT_OK;
}

int pht_step_graph_launch(pht_step_graph* graph, void* stream) {
    if (!graph || !graph->exec) return fail(PHT_EINVAL, "null graph");
    return check_hip(hipGraphLaunch(graph->exec, static_cast<hipStream_t>(stream)), "graph launch");
}

void pht_step_graph_destroy(pht_step_graph* graph) {
    if (!graph) return;
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    delete graph;
}

} // extern "C"
