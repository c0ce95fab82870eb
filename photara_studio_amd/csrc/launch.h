// Host-side launchers shared between the kernel translation units and the C-ABI layer.
#pragma once

#include "device_common.h"

namespace pht {

// Arguments of the per-(gaussian, view) preprocess.
struct PreArgs {
    pht_gaussians g;
    pht_settings s;
    const DevView* views;
    int n_views;
    int vp_log2;
    SplatRec* rec;
    uint2* meta;
    uint32_t* vis_flag;
    uint32_t* view_mask;
    Counters* counters;
    uint32_t* key_parts; // [kKeyParts][2] partial (max of ~key, max of key) of the visible depth keys
};
// The visible depth keys' range is gathered in kKeyParts partial pairs (wave w of block b adds
// to pair (4b + w) mod kKeyParts), so that no single address takes every wave's atomics, and
// folded into the counters by launch_key_range.
constexpr int kKeyParts = 256;

struct BlendArgs {
    const uint2* ranges;
    const uint32_t* inst_vals;
    const SplatRec* rec;
    const DevView* views;
    int n_views;
    int vp_log2;
    float alpha_clip;
    float* final_T;
    uint32_t* n_contrib;
    float* sgrad;
    const Counters* counters;
    uint32_t slot_cap;
    int64_t n_gauss;
    int want_refine; // backward: accumulate the refine statistic (only when a caller reads it)
    const uint32_t* tile_order; // tile of each work item (heaviest first), or null for the XCD-contiguous map
    uint32_t* queue; // persistent work-queue counters [forward, backward] (zeroed per launch), or null: flat grid
    float* loss_sum; // forward with the fused L1 loss: += sum over views of scale * sum |C - target|, or null
};

// The statistics of a batched backward's views, from its per-(gaussian, view) refine statistic.
struct DensifyStepArgs {
    pht_densify_stats stats;
    pht_densify_observation rules; // use_maximum, visibility rule, power, oversize threshold
    int32_t view_index[PHT_MAX_VIEWS];
    float inverse_resolution[PHT_MAX_VIEWS];
    int geometry_from_refine;
    const DevView* views;
    int n_views;
    const float* sgrad;
    int64_t n;
};
// One workspace's views as the per-Gaussian backward reads them: a step's views may be split into
// view groups, each forwarded and blended on its own workspace (and stream); the per-Gaussian
// backward then sums over every group's views, in group order.
constexpr int kMaxViewGroups = 4;
struct GaussGroup {
    const DevView* views;
    int n_views;
    int vp_log2;
    const uint2* meta;
    const float* sgrad;
    const uint32_t* view_mask;
    const Counters* counters;
};
struct GaussArgs {
    pht_gaussians g;
    pht_settings s;
    int n_groups;
    GaussGroup grp[kMaxViewGroups];
    pht_model_grads out;
    pht_adam adam;
    DensifyStepArgs dens; // with the fused Adam: the step's densification statistics (has_dens)
    int has_dens;
};

// k_gaussian.hip: standalone quantised SH Adam over rows x stride (stride <= kShqMaxStride).
constexpr int kShqMaxStride = 48;
void launch_sh_adam_quant(float* param, const float* grad, uint16_t* first, uint8_t* packed, float* bounds,
                          int64_t rows, int stride, const ShqStep& step, hipStream_t st);

// k_sort.hip
int scan_blocks_for(uint32_t capacity);
// Compacts the visible slots (slot order) into (depth key, slot, tile count) lists.
void launch_compact_visible(const uint32_t* flags, const uint2* meta, uint32_t n_slots, uint32_t* keys,
                            uint32_t* vals, uint32_t* tiles, uint32_t* block_scratch, uint32_t* visible_out,
                            hipStream_t st);
// Exclusive scan of the depth-ordered tile counts into instance offsets; stores the instance
// total and the overflow flag (total > inst_cap).
void launch_tile_offsets(const uint32_t* sorted_tiles, const uint32_t* visible_ptr, uint32_t slot_cap,
                         uint32_t* offsets, uint32_t* block_scratch, uint32_t* instances_out,
                         uint32_t* overflow_out, uint32_t inst_cap, hipStream_t st);
// Stable sort of (keys, vals[, aux]) by the low `bits` bits of the keys (aux may be null);
// returns the buffer index (0/1) holding the result. hist holds 256 * kScanBlocks counts,
// digit_total 256.
int launch_radix_sort(uint32_t* keys[2], uint32_t* vals[2], uint32_t* aux[2], const uint32_t* n_ptr, uint32_t cap,
                      int bits, uint32_t* hist, uint32_t* digit_total, hipStream_t st, bool narrow = false);
// Stable sort of the visible (depth key, slot, tile count) lists by depth key, ranking key - the
// least visible key (krange: the preprocess's {max of ~key, max of key} counters); the top pass
// is skipped when the keys' range fits 24 bits. The result is in buffer 2 either way.
void launch_depth_sort(uint32_t* keys[3], uint32_t* vals[3], uint32_t* aux[3], const uint32_t* n_ptr, uint32_t cap,
                       const uint32_t* krange, uint32_t* hist, uint32_t* digit_total, hipStream_t st);

// k_preprocess.hip
void launch_upload_views(DevView* dst, const DevView* host_views, int n_views, hipStream_t st);
void launch_clear(Counters* counters, uint2* ranges, uint32_t n_tiles, uint32_t* key_parts, hipStream_t st);
void launch_key_range(const uint32_t* key_parts, Counters* counters, hipStream_t st);
void launch_preprocess(const PreArgs& a, int64_t n_slots, hipStream_t st);
void launch_emit_instances(const uint32_t* sorted_slots, const uint32_t* sorted_tiles, const uint32_t* offsets,
                           const Counters* counters, uint32_t slot_cap, const SplatRec* rec, const DevView* views,
                           int vp_log2, int tile_size, uint32_t inst_cap, uint32_t* keys_out, uint32_t* vals_out,
                           bool narrow, hipStream_t st);
void launch_ranges(const uint32_t* keys, const Counters* counters, uint32_t inst_cap, uint32_t n_tiles,
                   uint2* ranges, bool narrow, hipStream_t st);
void launch_debug_keys(const uint32_t* tile_keys, const uint32_t* slots, const uint2* meta,
                       const Counters* counters, uint32_t inst_cap, int vp_log2, uint64_t* keys_out,
                       uint32_t* vals_out, bool narrow, hipStream_t st);

void launch_debug_splat_grads(const float* sgrad, const uint32_t* view_mask, int64_t n, int n_views, int vp_log2,
                              float* out, hipStream_t st);

// k_blend.hip
// order: n_tiles entries; hist: 512 words of scratch.
// The tiles heaviest first, by instance count.
void launch_tile_order(const uint2* ranges, uint32_t n_tiles, uint32_t* order, uint32_t* hist, hipStream_t st);
void launch_blend_forward(const BlendArgs& a, int n_tiles, int tile_size, int batch, int waves, hipStream_t st);
void launch_blend_backward(const BlendArgs& a, int n_tiles, int tile_size, int batch, int waves, hipStream_t st);

// k_gaussian.hip
void launch_gaussian_backward(const GaussArgs& a, bool adam, hipStream_t st);
void launch_adam_standalone(float* p, const float* g, float* m, float* v, int64_t rows, int64_t stride,
                            int64_t active, float lr, float lr2, int64_t secondary_from, float b1, float b2,
                            float c1, float c2, float eps, float lo, float hi, int prefix, hipStream_t st);
// k_ssim.hip
size_t l1_ssim_scratch_floats(int width, int height);
void launch_l1_ssim(const float* color, const float* target, const float* mask, float* grad, int accumulate,
                    float* loss_sum, float* scratch, int width, int height, float ssim_weight,
                    float photometric_weight, hipStream_t st);

void launch_l1_loss(const float* color, const float* target, float* grad, float* loss_sum, int64_t pixels,
                    float weight, hipStream_t st);
void launch_zero_f32(float* p, int64_t n, hipStream_t st);
// Device address of the static per-tile partials array of the no-allocation L1+SSIM entry point on
// the current device (looked up once per device), or null with *err set.
float* l1_ssim_static_scratch(hipError_t* err);

// k_densify.hip
struct DensifyObsSet {
    pht_densify_stats stats;
    pht_densify_observation obs[PHT_MAX_VIEWS];
    float inverse_resolution[PHT_MAX_VIEWS]; // 1 / max(1, min(width, height)) per observation
    int n;
};
void launch_densify_accumulate(const DensifyObsSet& set, int64_t n, hipStream_t st);
void launch_densify_step(const DensifyStepArgs& a, hipStream_t st);
void launch_prune_masks(const pht_gaussians& g, int sh_stride, float min_opacity, float max_bounds,
                        const float center[3], uint8_t* keep, uint8_t* hard, float* opacities, hipStream_t st);
// Device state of one radix select (in the caller's workspace, after the keys).
struct SelectState {
    uint32_t hist[256];
    uint32_t eligible;  // rows with a key
    uint32_t remaining; // rank still to find below the current prefix
    uint32_t chosen;    // min(k, eligible)
    uint32_t pad;
    uint64_t prefix;    // digits chosen so far (the full threshold key after eight passes)
};
size_t select_workspace_bytes(int64_t n);
void launch_select_rows(const float* values, const uint8_t* mask, const uint8_t* exclude, int64_t n, int64_t k,
                        int mode, uint64_t seed, uint8_t* out, uint8_t out_value, uint32_t* counts,
                        uint64_t* threshold, void* workspace, hipStream_t st);
void launch_igs_prepare(const pht_densify_stats& s, const int32_t* keep_rows, int64_t retained, float gthr,
                        float sthr, float* priority, uint8_t* positive, uint32_t* counts, hipStream_t st);
void launch_igs_weights(const pht_densify_stats& s, const int32_t* keep_rows, int64_t retained,
                        const float* opacities, float median, int allow_growth, float gthr, float sthr,
                        float* replacement, float* growth, float* oversize, hipStream_t st);
struct RowBuffer {
    const void* src;
    void* dst;
    int64_t row_bytes;
    int64_t unit; // bytes moved per access: 16, 4 or 2
};
struct RowBuffers {
    RowBuffer b[PHT_MAX_ROW_BUFFERS];
    int n;
};
void launch_gather_rows(const RowBuffers& set, const int32_t* rows, int64_t n_rows, hipStream_t st);
void launch_densify_split(const pht_split_params& p, const int32_t* parents, int64_t count, int64_t child_base,
                          const float* samples, const float* screens, hipStream_t st);
void launch_adc_decay(float* log_scales, float* logits, uint16_t* ls_lp, uint16_t* logit_lp, int fmt, int64_t n,
                      float opacity_decay, float log_scale_decay, hipStream_t st);

} // namespace pht
