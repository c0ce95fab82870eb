// Per-(gaussian, view) preprocessing, instance emission and tile ranges for gfx950.
//
// k_preprocess fuses what preprocess_gaussians does per Gaussian for one camera
// (third_party/splat_drender/src/render_forward.cu:30-139: activation, EWA projection, conic,
// radius, AccuTile tile count, radial depth key, SH colour) for every view of a batched step.
// Threads are laid out slot = gaussian * vp + view (vp = next power of two >= views), so the
// lanes of one Gaussian sit in one wave: its parameters are fetched from HBM once and served to
// the other views' lanes from cache, and the per-Gaussian view mask is one ballot.
#include "launch.h"

namespace pht {

struct ViewBatch {
    DevView v[8];
};

__global__ void k_upload_views(DevView* dst, ViewBatch b, int n) {
    const int i = threadIdx.x;
    if (i < n) dst[i] = b.v[i];
}

__global__ __launch_bounds__(256) void k_clear(Counters* counters, uint2* ranges, uint32_t n_tiles) {
    const uint32_t i0 = blockIdx.x * 256u + threadIdx.x;
    if (i0 < sizeof(Counters) / 4) reinterpret_cast<uint32_t*>(counters)[i0] = 0u;
    for (uint32_t i = i0; i < n_tiles; i += gridDim.x * 256u) ranges[i] = make_uint2(0u, 0u);
}

template <int LP> __device__ __forceinline__ float load_lp(const float* f, const uint16_t* lp, int64_t i) {
    if constexpr (LP == PHT_LP_NONE) return f[i];
    else return lp_to_f32(lp[i], LP);
}

template <int DEG, int LP>
__global__ __launch_bounds__(256) void k_preprocess(PreArgs a, int64_t n_slots) {
    const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
    const int vp = 1 << a.vp_log2;
    const int64_t gi = t >> a.vp_log2;
    const int v = int(t & int64_t(vp - 1));
    const bool in_range = t < n_slots;
    const bool valid = in_range && v < a.n_views;
    bool visible = false, fp64 = false;
    if (valid) {
        const DevView& dv = a.views[v];
        const float mean[3] = {a.g.means[3 * gi], a.g.means[3 * gi + 1], a.g.means[3 * gi + 2]};
        float ls[3], q[4];
        for (int k = 0; k < 3; ++k) ls[k] = load_lp<LP>(a.g.log_scales, a.g.log_scales_lp, 3 * gi + k);
        for (int k = 0; k < 4; ++k) q[k] = a.g.quaternions[4 * gi + k];
        const float logit = load_lp<LP>(a.g.opacity_logits, a.g.opacity_logits_lp, gi);
        const bool hf = a.g.filter_3d != nullptr;
        const Activated<float> act = activate<float>(ls, q, logit, hf, hf ? a.g.filter_3d[gi] : 0.0f);
        Projected<float> p;
        int radius = 0;
        if (project_pinhole(mean, act, dv.vi.cam, a.s, p)) {
            const uint32_t touched = enumerate_tiles(p.mean2d[0], p.mean2d[1], p.conic, p.opacity, dv.vi.grid_x,
                                                     dv.vi.grid_y, a.s.tile_size, [](int, int) {});
            if (touched) {
                visible = true;
                fp64 = p.fp64;
                radius = p.radius;
                float rgb[3];
                uint32_t clamp = 0;
                if (a.g.colors_precomp) {
                    for (int ch = 0; ch < 3; ++ch) rgb[ch] = a.g.colors_precomp[3 * gi + ch];
                } else {
                    constexpr int NB = (DEG + 1) * (DEG + 1);
                    float coef[NB * 3];
                    const int64_t row = gi * int64_t(a.s.sh_bases) * 3;
#pragma unroll
                    for (int k = 0; k < NB * 3; ++k) coef[k] = load_lp<LP>(a.g.sh, a.g.sh_lp, row + k);
                    float d[3], raw[3];
                    view_dir(mean, dv.vi.cam.position, d, raw);
                    sh_eval(DEG, d, coef, rgb, clamp);
                }
                a.rec_a[t] = make_float4(p.mean2d[0], p.mean2d[1], p.conic[0], p.conic[1]);
                a.rec_b[t] = make_float4(p.conic[2], p.opacity, rgb[0], rgb[1]);
                a.rec_c[t] = make_uint4(f2u(rgb[2]), f2u(p.depth), touched, clamp | (fp64 ? 8u : 0u));
                float4* sg = reinterpret_cast<float4*>(a.sgrad + t * kGradStride);
                sg[0] = make_float4(0.f, 0.f, 0.f, 0.f);
                sg[1] = make_float4(0.f, 0.f, 0.f, 0.f);
                sg[2] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
        if (dv.radii) dv.radii[gi] = visible ? radius : 0;
        if (dv.visibility) dv.visibility[gi] = 0.0f;
    }
    if (in_range) a.vis_flag[t] = visible ? 1u : 0u;
    const uint64_t bal = __ballot(visible);
    const int lane = lane_id();
    if (in_range && v == 0) {
        const int base = lane & ~(vp - 1);
        const uint64_t grp = vp >= 64 ? bal : ((bal >> base) & ((uint64_t(1) << vp) - 1u));
        a.view_mask[gi] = uint32_t(grp);
    }
    const uint64_t fb = __ballot(fp64);
    if (lane == 0 && fb) atomicAdd(&a.counters->fp64, uint32_t(__popcll(fb)));
}

// Emits (global tile, slot) pairs for the depth-sorted visible splats at their scanned offsets
// (emit_instances, render_forward.cu:167-181; tile keys carry the view's tile_base).
__global__ __launch_bounds__(256) void k_emit(const uint32_t* sorted_slots, const uint32_t* offsets,
                                              const Counters* counters, uint32_t slot_cap, const float4* rec_a,
                                              const float4* rec_b, const uint4* rec_c, const DevView* views,
                                              int vp_log2, int ts, uint32_t inst_cap, uint32_t* keys_out,
                                              uint32_t* vals_out) {
    const uint32_t nv = counters->visible;
    const uint32_t n = nv < slot_cap ? nv : slot_cap;
    const uint32_t vmask = (1u << vp_log2) - 1u;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const uint32_t slot = sorted_slots[i];
        const uint32_t off = offsets[i];
        const uint4 c = rec_c[slot];
        if (uint64_t(off) + c.z > inst_cap) continue;
        const float4 A = rec_a[slot], B = rec_b[slot];
        const ViewInfo& vi = views[slot & vmask].vi;
        const float conic[3] = {A.z, A.w, B.x};
        uint32_t k = 0;
        const int gx = vi.grid_x;
        const uint32_t tb = uint32_t(vi.tile_base);
        enumerate_tiles(A.x, A.y, conic, B.y, vi.grid_x, vi.grid_y, ts, [&](int tx, int ty) {
            if (k < c.z) {
                keys_out[off + k] = tb + uint32_t(ty * gx + tx);
                vals_out[off + k] = slot;
            }
            ++k;
        });
    }
}

// Per-tile [start, end) of the tile-sorted list (extract_tile_ranges, render_forward.cu:215-232).
__global__ __launch_bounds__(256) void k_ranges(const uint32_t* keys, const Counters* counters, uint32_t inst_cap,
                                                uint32_t n_tiles, uint2* ranges) {
    if (counters->overflow) return; // the list is incomplete: leave every range empty
    const uint32_t ni = counters->instances;
    const uint32_t n = ni < inst_cap ? ni : inst_cap;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const uint32_t tile = keys[i];
        if (tile >= n_tiles) continue;
        if (i == 0 || keys[i - 1] != tile) ranges[tile].x = i;
        if (i + 1 == n || keys[i + 1] != tile) ranges[tile].y = i + 1;
    }
}

__global__ __launch_bounds__(256) void k_debug_keys(const uint32_t* tile_keys, const uint32_t* slots,
                                                    const uint4* rec_c, const Counters* counters, uint32_t inst_cap,
                                                    int vp_log2, uint64_t* keys_out, uint32_t* vals_out) {
    const uint32_t ni = counters->instances;
    const uint32_t n = ni < inst_cap ? ni : inst_cap;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const uint32_t slot = slots[i];
        if (keys_out) keys_out[i] = (uint64_t(tile_keys[i]) << 32) | rec_c[slot].y;
        if (vals_out) vals_out[i] = slot >> vp_log2;
    }
}

__global__ __launch_bounds__(256) void k_debug_splat_grads(const float* sgrad, const uint32_t* view_mask, int64_t n,
                                                           int n_views, int vp_log2, float* out) {
    const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (t >= n * n_views) return;
    const int64_t g = t / n_views;
    const int v = int(t % n_views);
    const bool vis = (view_mask[g] >> v) & 1u;
    const float* src = sgrad + ((g << vp_log2) + v) * kGradStride;
    for (int k = 0; k < 10; ++k) out[t * 10 + k] = vis ? src[k] : 0.0f;
}

void launch_debug_splat_grads(const float* sgrad, const uint32_t* view_mask, int64_t n, int n_views, int vp_log2,
                              float* out, hipStream_t st) {
    const int64_t total = n * n_views;
    hipLaunchKernelGGL(k_debug_splat_grads, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, sgrad, view_mask, n,
                       n_views, vp_log2, out);
}

static int grid_stride_blocks(uint64_t n) {
    const uint64_t b = (n + 255u) / 256u;
    return int(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

void launch_upload_views(DevView* dst, const DevView* host_views, int n_views, hipStream_t st) {
    for (int base = 0; base < n_views; base += 8) {
        ViewBatch b;
        const int n = n_views - base < 8 ? n_views - base : 8;
        for (int i = 0; i < n; ++i) b.v[i] = host_views[base + i];
        hipLaunchKernelGGL(k_upload_views, dim3(1), dim3(64), 0, st, dst + base, b, n);
    }
}

void launch_clear(Counters* counters, uint2* ranges, uint32_t n_tiles, hipStream_t st) {
    hipLaunchKernelGGL(k_clear, dim3(grid_stride_blocks(n_tiles)), dim3(256), 0, st, counters, ranges, n_tiles);
}

template <int DEG, int LP> static void pre_dispatch(const PreArgs& a, int64_t n_slots, hipStream_t st) {
    const int64_t blocks = (n_slots + 255) / 256;
    hipLaunchKernelGGL((k_preprocess<DEG, LP>), dim3(unsigned(blocks)), dim3(256), 0, st, a, n_slots);
}
template <int LP> static void pre_deg(const PreArgs& a, int64_t n_slots, hipStream_t st) {
    switch (a.s.active_sh_degree) {
    case 0: pre_dispatch<0, LP>(a, n_slots, st); break;
    case 1: pre_dispatch<1, LP>(a, n_slots, st); break;
    case 2: pre_dispatch<2, LP>(a, n_slots, st); break;
    default: pre_dispatch<3, LP>(a, n_slots, st); break;
    }
}
void launch_preprocess(const PreArgs& a, int64_t n_slots, hipStream_t st) {
    if (a.g.lp_format == PHT_LP_BF16) pre_deg<PHT_LP_BF16>(a, n_slots, st);
    else if (a.g.lp_format == PHT_LP_FP16) pre_deg<PHT_LP_FP16>(a, n_slots, st);
    else pre_deg<PHT_LP_NONE>(a, n_slots, st);
}

void launch_emit_instances(const uint32_t* sorted_slots, const uint32_t* offsets, const Counters* counters,
                           uint32_t slot_cap, const float4* rec_a, const float4* rec_b, const uint4* rec_c,
                           const DevView* views, int vp_log2, int tile_size, uint32_t inst_cap,
                           uint32_t* keys_out, uint32_t* vals_out, hipStream_t st) {
    hipLaunchKernelGGL(k_emit, dim3(grid_stride_blocks(slot_cap)), dim3(256), 0, st, sorted_slots, offsets, counters,
                       slot_cap, rec_a, rec_b, rec_c, views, vp_log2, tile_size, inst_cap, keys_out, vals_out);
}

void launch_ranges(const uint32_t* keys, const Counters* counters, uint32_t inst_cap, uint32_t n_tiles,
                   uint2* ranges, hipStream_t st) {
    hipLaunchKernelGGL(k_ranges, dim3(grid_stride_blocks(inst_cap)), dim3(256), 0, st, keys, counters, inst_cap,
                       n_tiles, ranges);
}

void launch_debug_keys(const uint32_t* tile_keys, const uint32_t* slots, const uint4* rec_c,
                       const Counters* counters, uint32_t inst_cap, int vp_log2, uint64_t* keys_out,
                       uint32_t* vals_out, hipStream_t st) {
    hipLaunchKernelGGL(k_debug_keys, dim3(grid_stride_blocks(inst_cap)), dim3(256), 0, st, tile_keys, slots, rec_c,
                       counters, inst_cap, vp_log2, keys_out, vals_out);
}

} // namespace pht
