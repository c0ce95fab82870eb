/*
 * pht_splat.h — C ABI of the MI355X (gfx950) 3D Gaussian Splatting training step.
 *
 * This is the drop-in boundary for the hot path of the reference's splat trainer:
 *   photara::splat::Rasterizer::forward / ::backward
 *     (photara/include/splat/rasterizer.hpp:128-165, impl photara/src/splat/rasterizer.cu:340-743)
 *   splat_drender::Rasterizer::forward / ::backward over POD views + WorkspacePools
 *     (third_party/splat_drender/include/splat_drender/api.h:19-306, src/pipeline.cu:457-712)
 *   detail::adam_step / adam_step_structure / adam_step_active_prefix
 *     (photara/src/splat/optimizer.cpp:461-534, photara/src/splat/cuda_ops.cu:1291-1546, 2863-2957)
 *   loss_kernel, L1 part (photara/src/splat/cuda_ops.cu:846-898)
 *
 * Everything crossing this boundary is plain C: device pointers, sizes, POD structs and a
 * hipStream_t passed as void*. No C++ or torch types. All device pointers are caller-owned;
 * the library never allocates device memory on the compute path (the caller supplies the
 * workspace bytes reported by pht_workspace_bytes()).
 *
 * Layouts follow the reference: means [N,3], log_scales [N,3], quaternions [N,4] scalar-first,
 * opacity_logits [N] (reference shape [N,1]), sh [N,K,3] (K = bases per channel, 16 at degree 3),
 * images planar CHW: color [3,H,W], alpha [H,W]. world_to_camera is the reference's column-major
 * 4x4 (photara/include/splat/types.hpp:17-34; transform t = v[0..2]x + v[4..6]y + v[8..10]z + v[12..14]).
 */
#ifndef PHT_SPLAT_H
#define PHT_SPLAT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHT_ABI_VERSION 4
#define PHT_MAX_VIEWS 32

/* Status codes. The reference throws std::invalid_argument for bad inputs
 * (third_party/splat_drender/src/pipeline.cu:63-66,460-468), std::runtime_error for device
 * errors (:68-72) and std::logic_error for a backward without recorded forward state
 * (:655-656). The host mirrors map PHT_EINVAL -> invalid_argument / ValueError,
 * PHT_EDEVICE -> runtime_error / RuntimeError, PHT_ESTATE -> logic_error / RuntimeError. */
enum pht_status {
    PHT_OK = 0,
    PHT_EINVAL = 1,    /* invalid argument, shape or pointer */
    PHT_EDEVICE = 2,   /* HIP runtime error */
    PHT_ECAPACITY = 3, /* instance list exceeded the workspace capacity (see pht_counts) */
    PHT_ESTATE = 4,    /* backward without a matching forward in this workspace */
    PHT_ENOTSUP = 5    /* a reference feature outside this build (depth/normal channels, non-pinhole) */
};

/* Reduced-precision storage formats for SH / log-scales / opacity logits. */
enum pht_lp_format { PHT_LP_NONE = 0, PHT_LP_BF16 = 1, PHT_LP_FP16 = 2 };

/* One pinhole camera. Mirrors photara::splat::Camera (types.hpp:17-34) after
 * camera_view_of() (rasterizer.cu:200-224): cx/cy already include -pixel_center_offset. */
typedef struct pht_camera {
    float world_to_camera[16]; /* column-major 4x4 */
    float position[3];         /* camera centre in world space (SH view direction) */
    float fx, fy, cx, cy;
    int32_t width, height;
    float background[3];
    int32_t reserved;
} pht_camera;

/* Raster settings. Mirrors RasterizeOptions / RasterizationPolicy
 * (rasterizer.hpp:9-54, options.hpp:32-37) for the pinhole RGB training path. */
typedef struct pht_settings {
    int32_t active_sh_degree; /* min(requested, model degree), rasterizer.cu:361-362 */
    int32_t sh_bases;         /* K: coefficients per channel stored per Gaussian */
    float kernel_size;        /* 2D dilation added to the projected covariance */
    float scale_modifier;
    float alpha_clip;         /* 0.99 */
    float near_plane;         /* 0.2 */
    int32_t compensate_2d;    /* opacity *= sqrt(det0/det1) */
    int32_t tile_size;        /* 16 (reference) or 8 */
    int32_t batch;            /* LDS batch depth of the blend kernels: 32, 64, 128 or 256 */
    int32_t fp64_conic;       /* route near-singular 2x2 inverses (and the backward of
                                 ill-conditioned ones) through FP64 */
    int32_t require_depth;    /* must be 0: median-depth/normal channels are not built */
    int16_t blend_waves;      /* blend launch shape: waves (one (view, tile) item each) per
                                 workgroup, 1, 2 or 4 (0 = 1) */
    int16_t blend_queue;      /* 1: blends run as one persistent work-queue launch (resident
                                 workgroups pull (view, tile) items from a device counter);
                                 0: one flat grid of work items */
} pht_settings;

/* Gaussian parameters (raw, pre-activation), reference GaussianModel (types.hpp:36-56)
 * passed through gaussians_of() (rasterizer.cu:226-251). When lp_format != PHT_LP_NONE the
 * *_lp arrays are read instead of sh / log_scales / opacity_logits. */
typedef struct pht_gaussians {
    int64_t count;
    const float* means;          /* [N,3] */
    const float* log_scales;     /* [N,3] */
    const float* quaternions;    /* [N,4] scalar-first, unnormalised */
    const float* opacity_logits; /* [N] */
    const float* sh;             /* [N,K,3] */
    const float* filter_3d;      /* [N] optional Mip 3D filter, or NULL */
    const float* colors_precomp; /* [N,3] optional: bypasses SH (RasterizeOptions::colors_precomp) */
    const uint16_t* sh_lp;             /* [N,K,3] bf16/fp16 bits */
    const uint16_t* log_scales_lp;     /* [N,3] */
    const uint16_t* opacity_logits_lp; /* [N] */
    int32_t lp_format;
    int32_t reserved;
} pht_gaussians;

/* Per-view forward outputs (RenderResult, types.hpp:99-111). Optional pointers may be NULL. */
typedef struct pht_view_outputs {
    float* color;        /* [3,H,W] required */
    float* alpha;        /* [H,W]   required: 1 - final transmittance */
    int32_t* radii;      /* [N]     optional */
    float* visibility;   /* [N]     optional: 1 where the Gaussian blended into >= 1 pixel */
} pht_view_outputs;

/* Per-view loss gradients (LossGradients, api.h). grad_alpha may be NULL (treated as 0). */
typedef struct pht_view_grads {
    const float* grad_color; /* [3,H,W] */
    const float* grad_alpha; /* [H,W] or NULL */
} pht_view_grads;

/* Model gradients wrt raw parameters (ModelGradients, types.hpp:71-87), summed over views.
 * Arrays are overwritten (not accumulated). Any pointer may be NULL to skip that output. */
typedef struct pht_model_grads {
    float* means;          /* [N,3] */
    float* log_scales;     /* [N,3] */
    float* quaternions;    /* [N,4] */
    float* opacity_logits; /* [N]   */
    float* sh;             /* [N,K,3] */
    float* colors_precomp; /* [N,3] only when forward used colors_precomp */
    float* refine_weight;  /* [N]; when NULL (and on the fused-Adam path) the backward skips the
                            * refine statistic, whose per-splat value then reads 0 */
} pht_model_grads;

/* Fused Adam over every Gaussian parameter (StructureAdamUpdate rasterizer.hpp:84-104 plus the
 * FP32 active-prefix SH Adam, cuda_ops.cu:1506-1546). Masters are FP32 and updated in place;
 * when lp_format != PHT_LP_NONE the *_lp mirrors are rewritten from the updated masters. */
typedef struct pht_adam {
    float* means; float* log_scales; float* quaternions; float* opacity_logits; float* sh;
    float* m_means; float* v_means;
    float* m_log_scales; float* v_log_scales;
    float* m_quaternions; float* v_quaternions;
    float* m_opacity_logits; float* v_opacity_logits;
    float* m_sh; float* v_sh;
    float lr_means, lr_scales, lr_quaternions, lr_opacity, lr_sh0, lr_sh_rest;
    float beta1, beta2, correction1, correction2, epsilon;
    float min_log_scale, max_log_scale;
    /* Optional device float[3] = {correction1, correction2, lr_means}: when non-NULL it
     * overrides the three fields above, so a captured step graph can be replayed with the
     * per-step bias corrections and learning-rate schedule written before each launch. */
    const float* device_hparams;
    uint16_t* sh_lp; uint16_t* log_scales_lp; uint16_t* opacity_logits_lp;
    int32_t lp_format;
    int32_t reserved;
    /* Quantised SH Adam state, the reference trainer's SH optimizer (ShAdamQuant,
     * photara/src/splat/sh_adam_quant.hpp:12-24; trainer.cpp:1569-1577). When sh_first is
     * non-NULL it replaces m_sh / v_sh: sh_first fp16 [N, stride] normalised update, sh_packed
     * uint8 [N, stride] log-second codes, sh_bounds float [N, 4] (DC min, DC max, rest min, rest
     * max), stride = sh_bases * 3, all zero for a new model. sh_regularization is the factor of
     * the L2 term added to the active non-DC gradients (2 * weight / (N * (stride - 3)) in
     * rasterizer.cu:598-602; 0 for none). */
    uint16_t* sh_first; uint8_t* sh_packed; float* sh_bounds;
    float sh_regularization;
    int32_t reserved1;
} pht_adam;

/* Workspace capacities. instances = capacity of the (tile, splat) instance list summed over views. */
typedef struct pht_capacity {
    int64_t gaussians;
    int32_t views;
    int32_t tile_size;
    int64_t tiles;     /* sum over views of ceil(W/ts)*ceil(H/ts) */
    int64_t pixels;    /* sum over views of W*H */
    int64_t instances;
} pht_capacity;

/* Device-side counters of the last forward. */
typedef struct pht_counts {
    int64_t visible;      /* visible (gaussian, view) splats */
    int64_t instances;    /* emitted (tile, splat) instances (may exceed capacity -> overflow) */
    int64_t fp64_conics;  /* splats whose 2x2 inverse was routed through FP64 */
    int32_t overflow;     /* 1 when instances > capacity: outputs of that forward are invalid */
    int32_t reserved;
} pht_counts;

typedef struct pht_workspace pht_workspace; /* opaque host handle over caller-owned device bytes */

/* ---- library ---- */
int pht_abi_version(void);
const char* pht_last_error(void);
/* Number of visible HIP devices (0 without a GPU); never initialises a context on failure paths. */
int pht_device_count(void);

/* ---- workspace (replaces WorkspacePools grow-callbacks, api.h:227-240, rasterizer.cu:177-185) ---- */
size_t pht_workspace_bytes(const pht_capacity* cap);
int pht_workspace_create(void* device_bytes, size_t nbytes, const pht_capacity* cap,
                         pht_workspace** out);
void pht_workspace_destroy(pht_workspace* ws);

/* ---- forward: replaces splat_drender::Rasterizer::forward (pipeline.cu:457-615) and the
 * CudaRasterizer::forward bridge (rasterizer.cu:340-402) for n_views pinhole views at once.
 * Enqueues on `stream` only. check_overflow=1 synchronises once at the end to report
 * PHT_ECAPACITY (the reference reads its counters back too, pipeline.cu:206-212);
 * check_overflow=0 never synchronises (graph capture) — query pht_counts afterwards. ---- */
int pht_forward(pht_workspace* ws, const pht_gaussians* g, const pht_camera* cams, int n_views,
                const pht_settings* s, pht_view_outputs* out, int check_overflow, void* stream);

/* ---- pht_forward followed by the L1 loss of every view (the trainer's compute_step_loss L1 leg,
 * trainer.cpp:2201 / training_dispatch.cpp:9-23 into loss_kernel, cuda_ops.cu:886-898, full image,
 * no mask), fused: the blend's epilogue writes
 * grad_color[v] = weight * sign(color - target[v]) / (3 W H) from the colour it composites and
 * adds sum |color - target[v]| * weight / (3 W H) to *loss_sum (device float, may be NULL).
 * The same colours, gradients and arguments as pht_forward then pht_l1_loss per view; the loss
 * sum is added in another order. No single reference function: it fuses the two calls. ---- */
int pht_forward_l1(pht_workspace* ws, const pht_gaussians* g, const pht_camera* cams, int n_views,
                   const pht_settings* s, pht_view_outputs* out, const float* const* targets,
                   float* const* grad_color, float* loss_sum, float weight, int check_overflow, void* stream);

/* ---- backward: replaces splat_drender::Rasterizer::backward (pipeline.cu:651-712) and
 * CudaRasterizer::backward (rasterizer.cu:495-743). Uses the state recorded by the last
 * pht_forward on this workspace (PHT_ESTATE if none). With adam != NULL the per-Gaussian
 * kernel applies the fused Adam update instead of (or in addition to) writing grads. ---- */
int pht_backward(pht_workspace* ws, const pht_gaussians* g, const pht_view_grads* dl,
                 pht_model_grads* grads, const pht_adam* adam, void* stream);

/* ---- L1 photometric loss gradient (cuda_ops.cu:886-898, full image, no mask):
 * grad = weight * sign(color - target) / (3 * W * H); loss_sum (device float, may be NULL)
 * receives += sum |color - target| * weight / (3 W H). ---- */
int pht_l1_loss(const float* color, const float* target, float* grad_color, float* loss_sum,
                int64_t pixels, float weight, void* stream);

/* ---- fused L1 + SSIM photometric loss of one view: replaces cuda_backend::fused_l1_ssim_loss
 * (photara/src/splat/fused_ssim.cu:473-560; semantics :120-470). 11x11 Gaussian window
 * (sigma 1.5, zero padding), C1 = 0.01^2, C2 = 0.03^2; over the valid window
 * [5, W-5) x [5, H-5): loss = norm * sum (ssim_weight (1 - SSIM) + (1 - ssim_weight) |x - y|),
 * norm = photometric_weight / (3 (W-10) (H-10)). color/target/grad_color are [3,H,W]; mask
 * ([H,W], may be NULL) multiplies both images and the gradient. accumulate != 0 adds the
 * gradient to grad_color (the reference's +=), else overwrites it. loss_sum (device float, may
 * be NULL) receives += loss and then needs scratch (pht_l1_ssim_scratch_bytes bytes of device
 * memory). W and H must exceed 10 (PHT_EINVAL otherwise, like the reference's
 * std::invalid_argument). ---- */
size_t pht_l1_ssim_scratch_bytes(int width, int height);
int pht_l1_ssim_loss(const float* color, const float* target, const float* mask, float* grad_color,
                     int accumulate, float* loss_sum, void* scratch, int width, int height,
                     float ssim_weight, float photometric_weight, void* stream);

/* ---- pht_l1_ssim_loss without caller scratch and without any allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The per-tile loss partials live in a static
 * device array of PHT_NOALLOC_SSIM_TILES floats that the library holds on each GPU it runs on
 * until the process ends; a view needs ceil(W/16) * ceil(H/16) <= PHT_NOALLOC_SSIM_TILES
 * (PHT_EINVAL otherwise). Same arguments and results as pht_l1_ssim_loss with loss_sum; loss_sum
 * may be NULL (then the array is not used). Calls that overlap on one GPU (concurrent streams)
 * race on that array. ---- */
#define PHT_NOALLOC_SSIM_TILES 65536
int pht_l1_ssim_loss_noalloc(const float* color, const float* target, const float* mask, float* grad_color,
                             int accumulate, float* loss_sum, int width, int height, float ssim_weight,
                             float photometric_weight, void* stream);

/* ---- standalone Adam (adam_kernel / adam_active_prefix_kernel, cuda_ops.cu:1336-1358,1506-1546):
 * group_stride/active/secondary lr follow the reference signature. ---- */
int pht_adam_step(float* param, const float* grad, float* m, float* v, int64_t rows,
                  int64_t row_stride, int64_t active_cols, float lr, float lr_secondary,
                  int64_t secondary_from_col, float beta1, float beta2, float correction1,
                  float correction2, float epsilon, float clamp_min, float clamp_max,
                  int32_t prefix_semantics, void* stream);

/* ---- standalone quantised SH Adam (sh_adam_quant_step, photara/src/splat/sh_adam_quant.hpp:34-49,
 * kernel sh_adam_quant_warp_row, sh_adam_quant.cuh:146-234): rows x stride coefficients, columns
 * [0, active) updated (lr for columns 0-2, rest_lr after), columns [active, stride) re-encoded
 * under the new bounds with their parameters untouched; regularization is the factor added times
 * the parameter to active non-DC gradients (0 when the gradient already holds it). first is fp16
 * bits, packed uint8, bounds float [rows, 4]; stride in [1, 48]. ---- */
int pht_sh_adam_quant_step(float* parameter, const float* gradient, uint16_t* first, uint8_t* packed,
                           float* bounds, int64_t rows, int32_t stride, int32_t active, float lr, float rest_lr,
                           float beta1, float beta2, float correction1, float correction2, float epsilon,
                           float regularization, void* stream);

/* ---- counters and debug views of the last forward ---- */
int pht_get_counts(pht_workspace* ws, pht_counts* out, void* stream); /* synchronises */
/* Enqueues on `stream`, without synchronising, a copy of the last forward's counters into dst[4]
 * (visible, instances, overflow, fp64_conics as uint32; dst should be pinned host memory): the
 * step driver reads it after an event to find a replayed step whose instance list overflowed
 * (that step left the model unchanged) without stalling the replay loop. The reference reads its
 * counts back synchronously instead (pipeline.cu:206-212). */
int pht_get_counts_async(pht_workspace* ws, uint32_t* dst, void* stream);
/* Copies the sorted instance list as 64-bit keys ((global_tile << 32) | depth_bits) and
 * values (gaussian index) plus per-tile [start,end) ranges and per-pixel contributor counts
 * into caller device buffers (any may be NULL). Sizes: keys/values >= instances, ranges
 * >= 2*tiles, n_contrib >= pixels. */
int pht_debug_lists(pht_workspace* ws, uint64_t* keys, uint32_t* values, uint32_t* ranges,
                    uint32_t* n_contrib, float* final_T, void* stream);
/* Copies the per-(gaussian, view) blend gradients of the last backward (the reference's GradState,
 * buffers.h): out[(g * n_views + v) * 10 + k], k = d_mean2d.xy, d_conic.xyz, d_opacity, d_color.rgb,
 * refine; zero for (gaussian, view) pairs that were not visible. out holds N * n_views * 10 floats. */
int pht_debug_splat_grads(pht_workspace* ws, float* out, void* stream);

/* ---- per-stage timing (the reference's CudaTrainingProfiler, trainer.cpp:216-258, as hipEvents):
 * with profiling on, forward/backward record an event at each stage boundary on their stream.
 * Stages: 0 preprocess, 1 visibility compaction + depth sort, 2 tile offsets + instance emission,
 * 3 tile sort + ranges, 4 blend forward, 5 blend backward, 6 per-Gaussian backward (+ Adam).
 * pht_stage_times synchronises and writes the last recorded milliseconds of each stage. ---- */
#define PHT_STAGES 7
int pht_set_profiling(pht_workspace* ws, int enable);
int pht_stage_times(pht_workspace* ws, float* ms_out, int max_stages);

/* Photometric loss of a captured step (the trainer's compute_step_loss, trainer.cpp:2024-2379):
 * kind PHT_LOSS_L1 = pht_l1_loss with weight photometric_weight; PHT_LOSS_L1_SSIM =
 * pht_l1_ssim_loss with ssim_weight and photometric_weight, no mask. ssim_scratch: per view,
 * pht_l1_ssim_scratch_bytes of device memory (needed when the step sums the loss). */
#define PHT_LOSS_L1 0
#define PHT_LOSS_L1_SSIM 1
typedef struct pht_loss_config {
    int32_t kind;
    float ssim_weight;
    float photometric_weight;
    int32_t reserved0;
    void* const* ssim_scratch;
    int64_t reserved[2];
} pht_loss_config;

/* ---- whole training step as one hipGraph (forward + loss + backward + fused Adam); loss may
 * be NULL (L1 with weight 1) ---- */
typedef struct pht_step_graph pht_step_graph;
int pht_step_graph_create(pht_workspace* ws, const pht_gaussians* g, const pht_camera* cams,
                          int n_views, const pht_settings* s, pht_view_outputs* out,
                          const float* const* targets, float* const* grad_color_bufs,
                          float* loss_sum, const pht_loss_config* loss, const pht_adam* adam,
                          void* stream, pht_step_graph** out_graph);
int pht_step_graph_launch(pht_step_graph* graph, void* stream);
void pht_step_graph_destroy(pht_step_graph* graph);

/* ======================================================================================
 * Densification (ADC-IGS refinement, photara/src/splat/densification_igs.cpp:291-515): the
 * per-row device work of a refinement. The host sequences these calls (and reads the few
 * counts the refinement branches on), as the reference's IgsStrategy::refine does.
 * ====================================================================================== */

/* detail::DensificationStats (photara/src/splat/adc_ops.cpp:8-19), [N] each. */
typedef struct pht_densify_stats {
    float* gradient;
    float* count;
    float* max_screen_radius;
    float* priority;
    float* max_geometry_gradient;
    int32_t* first_view;
    float* view_support;
    float* accumulated_error;
} pht_densify_stats;

/* One view's observation for the statistics (accumulate_densification_stats,
 * photara/src/splat/adc_ops.cpp:23-62; kernel photara/src/splat/cuda_ops.cu:1548-1606).
 * Optional arrays may be NULL. */
typedef struct pht_densify_observation {
    const float* refine_weight;     /* [N] */
    const float* visibility;        /* [N] */
    const int32_t* radii;           /* [N] */
    const float* geometry_gradient; /* [N] or NULL */
    const float* image_error;       /* [N] or NULL */
    const float* normalized_size;   /* [N] or NULL: replaces radius / min(width, height) */
    int32_t width, height;
    int32_t use_maximum;            /* gradient = max(gradient, w) instead of += w */
    int32_t require_contribution_visibility;
    float step_score_power;         /* w = max(refine_weight, 0) ^ power */
    float oversize_screen_threshold; /* > 0: priority += max(log2(screen / threshold), 0) */
    int32_t view_index;             /* >= 0: distinct-view support */
    int32_t reserved;
} pht_densify_observation;

/* Accumulates n_obs observations (in order: the same as n_obs calls of the reference's
 * accumulate_densification_stats) into stats in one pass over the N rows. */
int pht_densify_accumulate(const pht_densify_stats* stats, const pht_densify_observation* obs, int n_obs,
                           int64_t n, void* stream);

/* prune_masks (photara/src/splat/ops.hpp:131-141; kernel adc_plus_prune_kernel,
 * photara/src/splat/cuda_ops.cu:263-302) over the FP32 master parameters of g: keep = !hard &&
 * sigmoid(logit) >= minimum_opacity; hard = any non-finite parameter (SH included, sh_stride
 * floats per row), a mean outside the box |mean - center| <= maximum_bounds, or max
 * exp(log_scale) > maximum_bounds; opacities = sigmoid(logit). */
int pht_densify_prune_masks(const pht_gaussians* g, int32_t sh_stride, float minimum_opacity,
                            float maximum_bounds, const float center[3], uint8_t* keep, uint8_t* hard,
                            float* opacities, void* stream);

/* Row selection by radix select over 64-bit keys (no sort of the n rows).
 *   PHT_SELECT_LARGEST / PHT_SELECT_SMALLEST: the k rows with the largest / smallest value,
 *     ties to the lower index: select_front's sort-and-take-front (densification_igs.cpp:91-121).
 *   PHT_SELECT_GUMBEL: weighted sampling without replacement by Gumbel top-k, keys of
 *     igs_weighted_sample_cuda (photara/src/splat/igs_selection_cuda.cu:31-70): rows with a
 *     positive finite weight, score log(w) - log(-log(u)), u from splitmix64(seed + golden * (i+1)),
 *     ties to the higher index.
 *   PHT_SELECT_LARGEST_POSITIVE: LARGEST over the rows with a positive finite value.
 * A row is eligible when mask is NULL or mask[i] != 0, exclude is NULL or exclude[i] == 0, its
 * value is not NaN and (GUMBEL, LARGEST_POSITIVE) positive and finite. k is clamped to the eligible count. Each selected
 * row gets out[i] = out_value (out may alias mask or exclude). counts (device, optional) receives
 * {eligible, selected}; threshold (device, optional) the key of the last selected row, whose high
 * 32 bits give its value for LARGEST / SMALLEST (pht_select_key_value). */
#define PHT_SELECT_LARGEST 0
#define PHT_SELECT_SMALLEST 1
#define PHT_SELECT_GUMBEL 2
#define PHT_SELECT_LARGEST_POSITIVE 3
size_t pht_select_workspace_bytes(int64_t n);
int pht_select_rows(const float* values, const uint8_t* mask, const uint8_t* exclude, int64_t n, int64_t k,
                    int mode, uint64_t seed, uint8_t* out, uint8_t out_value, uint32_t* counts,
                    uint64_t* threshold, void* workspace, size_t workspace_bytes, void* stream);
/* The value encoded in a LARGEST / SMALLEST key (host helper). */
float pht_select_key_value(uint64_t key, int mode);

/* IGS parent scores over the retained rows r < retained (row src = keep_rows[r], or r when
 * keep_rows is NULL), densification_igs.cpp:439-497:
 *   pht_igs_prepare: priority[r] = stats.priority / max(count, 1); positive[r] = count > 0 &&
 *     priority > 0; counts (device) += {positive, growth candidates (count > 0 && gradient >
 *     gradient_threshold), oversized (count > 0 && max_screen_radius > screen_threshold)};
 *   pht_igs_weights: edge = 1 + 0.25 min(max(priority, 0) / median, 10); replacement = opacity *
 *     edge (candidates); growth = gradient * edge * (oversized ? 2 : 1) (growth candidates) and
 *     oversize = max_screen_radius * edge (oversized), both only when allow_growth (else 0). */
int pht_igs_prepare(const pht_densify_stats* stats, const int32_t* keep_rows, int64_t retained,
                    float gradient_threshold, float screen_threshold, float* priority, uint8_t* positive,
                    uint32_t* counts, void* stream);
int pht_igs_weights(const pht_densify_stats* stats, const int32_t* keep_rows, int64_t retained,
                    const float* opacities, float median, int32_t allow_growth,
                    float gradient_threshold, float screen_threshold, float* replacement, float* growth,
                    float* oversize, void* stream);

/* Row remap of model and optimizer arrays: dst[r] = src[rows[r]] (rows[r] < 0: zero row) for
 * r < n_rows, for each buffer (row_bytes per row, any multiple of 2). Replaces the reference's
 * index_select / cat / zero_adam_rows sequence (densification_adc_plus.cpp:18-95, 148-196;
 * zero_adam_rows_kernel, cuda_ops.cu:1407-1421) with one launch. dst must not alias src. */
#define PHT_MAX_ROW_BUFFERS 24
typedef struct pht_row_buffer {
    const void* src;
    void* dst;
    int64_t row_bytes;
} pht_row_buffer;
int pht_gather_rows(const pht_row_buffer* buffers, int n_buffers, const int32_t* rows, int64_t n_rows,
                    void* stream);

/* split_gaussians_kernel (photara/src/splat/cuda_ops.cu:1720-1845) in place on FP32 masters:
 * for j < count, parent row parents[j] and child row child_base + j, both holding the parent's
 * parameters on entry, receive the split parameters. samples [count,3] normal variates;
 * screen_sizes [count] or NULL (ADC covariance mode). When lp_format != PHT_LP_NONE the rows'
 * reduced-precision copies of log-scales and logits are refreshed too. */
#define PHT_SPLIT_DENSE_TANGENT 0
#define PHT_SPLIT_ADC_COVARIANCE 1
#define PHT_SPLIT_IGS_RANDOM 2
#define PHT_SPLIT_LONG_AXIS 3
typedef struct pht_split_params {
    float* means;
    float* log_scales;
    float* opacity_logits;
    const float* quaternions;
    uint16_t* log_scales_lp;
    uint16_t* opacity_logits_lp;
    int32_t lp_format;
    int32_t mode;
    float minimum_opacity;
    float split_at_screen_size;
    float split_opacity_k;
    int32_t reserved;
} pht_split_params;
int pht_densify_split(const pht_split_params* p, const int32_t* parents, int64_t count, int64_t child_base,
                      const float* samples, const float* screen_sizes, void* stream);

/* adc_decay_kernel (photara/src/splat/cuda_ops.cu:1847-1861): opacity -= opacity_decay
 * (clamped to [1e-12, 1 - 1e-12]), log_scales += log_scale_decay, over n rows; refreshes the
 * reduced-precision copies when given (lp_format != PHT_LP_NONE). */
int pht_adc_decay(float* log_scales, float* opacity_logits, uint16_t* log_scales_lp, uint16_t* opacity_logits_lp,
                  int32_t lp_format, int64_t n, float opacity_decay, float log_scale_decay, void* stream);

/* Statistics of a batched training step, accumulated inside its backward: view v of the batch is
 * observed as accumulate_densification_stats would observe its render (its radii, visibility and
 * refine weight, the trainer's call at photara/src/splat/trainer.cpp:2441-2449), views in order.
 * view_index[v] is the dataset index of the batch's view v (or -1). geometry_from_refine: the
 * refine weight is also the geometry gradient (the trainer's call). The forward must have
 * written radii and visibility for every view. */
typedef struct pht_densify_step {
    pht_densify_stats stats;
    int32_t view_index[PHT_MAX_VIEWS];
    int32_t use_maximum;
    int32_t require_contribution_visibility;
    float step_score_power;
    float oversize_screen_threshold;
    int32_t geometry_from_refine;
    int32_t reserved;
} pht_densify_step;
/* pht_backward + the step's densification statistics (computes the refine statistic). */
int pht_backward_densify(pht_workspace* ws, const pht_gaussians* g, const pht_view_grads* dl,
                         pht_model_grads* grads, const pht_adam* adam, const pht_densify_step* densify,
                         void* stream);
/* pht_step_graph_create whose captured backward is pht_backward_densify. */
int pht_step_graph_create_densify(pht_workspace* ws, const pht_gaussians* g, const pht_camera* cams,
                                  int n_views, const pht_settings* s, pht_view_outputs* out,
                                  const float* const* targets, float* const* grad_color_bufs,
                                  float* loss_sum, const pht_loss_config* loss, const pht_adam* adam,
                                  const pht_densify_step* densify, void* stream,
                                  pht_step_graph** out_graph);

/* ---- view groups: a step's views split into n_groups (1..4) groups, each forwarded (pht_forward
 * or pht_forward_l1) on its own workspace and stream, so that one group's sorts can run beside
 * another group's blends. pht_backward_groups runs each group's blend backward on that group's
 * stream (group_streams[k], which must already hold the group's forward), makes `stream` wait for
 * all of them, then runs ONE per-Gaussian backward (+ fused Adam) over every group's views in
 * group order on `stream`. dl lists the views of group 0, then group 1, ... No densification
 * statistics (use one workspace for those). No reference counterpart: the reference renders one
 * view per call (trainer.cpp:2024-2379); this is the batched step's scheduling. ---- */
int pht_backward_groups(pht_workspace* const* ws, int n_groups, const pht_gaussians* g,
                        const pht_view_grads* dl, pht_model_grads* grads, const pht_adam* adam,
                        void* const* group_streams, void* stream);
/* The whole grouped step as one hipGraph: the capture forks into one stream per group (group k
 * takes the next group_views[k] cameras, outputs, targets and gradient buffers), each group runs
 * its forward and loss, and the graph joins into pht_backward_groups. */
int pht_step_graph_create_groups(pht_workspace* const* ws, int n_groups, const int* group_views,
                                 const pht_gaussians* g, const pht_camera* cams, int n_views,
                                 const pht_settings* s, pht_view_outputs* out, const float* const* targets,
                                 float* const* grad_color_bufs, float* loss_sum, const pht_loss_config* loss,
                                 const pht_adam* adam, void* stream, pht_step_graph** out_graph);

/* Host only: count draws of a fresh std::normal_distribution<float>(0, 1) over the std::mt19937
 * whose state (624 words, position in [0, 624]) the caller holds and which is advanced in place:
 * the split offsets of grow_igs_random_gpu (densification_adc_plus.cpp:113-120). */
int pht_host_normal_floats(uint32_t* mt_key, int32_t* mt_pos, float* out, int64_t count);

#ifdef __cplusplus
}
#endif
#endif /* PHT_SPLAT_H */
