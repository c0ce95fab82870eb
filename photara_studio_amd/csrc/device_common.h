// Device-side shared definitions for the gfx950 kernels: workspace tables, wave64 helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "splat_math.h"

namespace pht {

// Per-view table the kernels read: geometry + this call's caller-owned output/input pointers.
struct DevView {
    ViewInfo vi;
    float* color;          // [3,H,W]
    float* alpha;          // [H,W]
    int32_t* radii;        // [N] or null
    float* visibility;     // [N] or null
    const float* grad_color; // [3,H,W] (backward)
    const float* grad_alpha; // [H,W] or null (backward)
    // forward with the L1 loss fused into the blend's epilogue (pht_forward_l1), else null
    const float* l1_target; // [3,H,W]
    float* l1_grad;         // [3,H,W] dL/dC out
    float l1_scale;         // weight / (3 P), the per-pixel gradient magnitude
    float l1_pad;
};

// One L1 term (loss_kernel, photara/src/splat/cuda_ops.cu:886-898): returns |c - t| and sets
// g = scale * sign(c - t).
__device__ __forceinline__ float l1_term(float c, float t, float scale, float& g) {
    const float d = c - t;
    g = scale * float((d > 0.0f) - (d < 0.0f));
    return fabsf(d);
}

// Device counters, cleared at the start of every forward.
struct Counters {
    uint32_t visible;     // compacted visible splats
    uint32_t instances;   // emitted instances (exact, may exceed capacity)
    uint32_t overflow;    // instances > capacity
    uint32_t fp64;        // FP64-routed conics
    uint32_t scan_total;  // scratch total of the last chunked scan
    uint32_t queue[2];    // persistent blend work-queue heads (forward, backward)
    uint32_t key_nmin;    // max over visible slots of ~(depth key): the least key is ~key_nmin (k_key_range)
    uint32_t key_max;     // max over visible slots of the depth key (k_key_range)
    uint32_t pad[2];
};

// Blend record, one per (gaussian, view) slot (slot = g * vp + v), 64 bytes so that every
// gather of it by the tile kernels is one aligned 64-byte HBM sector:
//   a = (mean.x, mean.y, a2, b2)          d = (conic.x, conic.y, conic.z, opacity)
//   b = (c2, opacity, r, g)               c = (b, ext.x, ext.y, qmax)
// (a2, b2, c2) = the conic scaled for base-2 exponents (conic_log2), read by the per-pair test;
// ext = half extents of the box around the alpha >= 1/255 ellipse (grown by 1% + 1 px) and
// qmax = the ellipse's grown bound on the quadratic form, used to cull 8x8 pixel patches
// (+inf = no culling, -inf = reaches no pixel); the conic itself feeds the tile walk and the
// gradient terms. The members are laid out a, d, b, c: the emission's random gathers (a and d
// only, in depth order) then touch the record's first 32 bytes alone.
struct SplatRec {
    float4 a, d, b, c;
};
// Splat gradient (10 floats, read as five 8-byte pairs): dmean.xy, dconic.xyz, dopacity,
// dcolor.rgb, refine. The records are stored view-major: slot g * vp + v owns record v * n + g
// (n Gaussians), so a wave of consecutive Gaussians reads one view's records contiguously.
constexpr int kGradStride = 10;

// Per-slot meta record (uint2) written by the preprocess for visible slots: x = depth bits (the
// depth sort key), y = touched tile count | SH clamp flags << kMetaClampShift. A view has fewer
// than 2^26 tiles (at most 8191 per axis), so the count never reaches the flag bits.
constexpr uint32_t kMetaClampShift = 29;
constexpr uint32_t kMetaCountMask = (1u << kMetaClampShift) - 1u;

// Per-Gaussian view-mask word: bit v = visible in view v. With at most kMaskClampViews view slots
// bits 8 + v, 16 + v and 24 + v also hold the SH clamp flags of view v (also in the slot's meta record), so
// the per-Gaussian backward reads one word per Gaussian instead of one 16-byte record per view.
constexpr int kMaskClampViewsLog2 = 3;
constexpr int kMaskClampViews = 1 << kMaskClampViewsLog2;
__device__ __forceinline__ uint32_t mask_clamp_flags(uint32_t mask, int v) {
    return ((mask >> (8 + v)) & 1u) | (((mask >> (16 + v)) & 1u) << 1) | (((mask >> (24 + v)) & 1u) << 2);
}
__device__ __forceinline__ int64_t grad_offset(int64_t slot, int vp_log2, int64_t n) {
    return ((slot & ((int64_t(1) << vp_log2) - 1)) * n + (slot >> vp_log2)) * kGradStride;
}

// Fixed block count of the chunked scans / radix-sort passes (each block owns one contiguous
// chunk of the list, so the passes are stable and need no device-wide synchronisation).
constexpr int kScanBlocks = 1024;
constexpr int kScanThreads = 256;

// ---------------------------------------------------------------- wave64 helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_move(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL,
                                                                  ROW_MASK, 0xf, false));
}

// Sum over the 64 lanes of a wave; every lane must be active. The total lands in lane 63.
// quad_perm xor1, xor2, row_half_mirror, row_mirror, row_bcast15 (rows 1,3), row_bcast31 (rows 2,3).
__device__ __forceinline__ float wave_sum_to_last(float v) {
    v += dpp_move<0xb1, 0xf>(v);
    v += dpp_move<0x4e, 0xf>(v);
    v += dpp_move<0x141, 0xf>(v);
    v += dpp_move<0x140, 0xf>(v);
    v += dpp_move<0x142, 0xa>(v);
    v += dpp_move<0x143, 0xc>(v);
    return v;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// Inclusive prefix sum over the wave (Hillis-Steele on shuffles).
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t w = __shfl_up(v, o, 64);
        if (lane >= o) v += w;
    }
    return v;
}

__device__ __forceinline__ void atomic_add_f32(float* p, float v) { unsafeAtomicAdd(p, v); }

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2): map block b of
// n to a work item so that every XCD gets one contiguous run of items (neighbouring tiles share
// splats, so their records stay in one L2).
__device__ __forceinline__ int xcd_contiguous(int b, int n) {
    const int q = n >> 3, r = n & 7;
    const int x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ const DevView& view_of_tile(const DevView* views, int n_views, int tile,
                                                       int& v_out) {
    int v = 0;
    while (v + 1 < n_views && views[v + 1].vi.tile_base <= tile) ++v;
    v_out = v;
    return views[v];
}

} // namespace pht
