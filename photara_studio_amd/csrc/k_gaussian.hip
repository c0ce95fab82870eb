// Per-Gaussian backward (+ fused Adam), standalone Adam and the L1 loss gradient for gfx950.
//
// k_gaussian_bwd runs one thread per Gaussian and loops over the views of the batched step in
// lockstep (view constants stay wave-uniform): for each view the Gaussian was visible in, it
// chains the splat's blend gradients through the EWA projection and SH colour
// (gaussian_backward, third_party/splat_drender/src/render_backward.cu:708-832), sums over views,
// applies the activation chain (update_structure_adam_row, :650-678) and then either writes the
// raw-parameter gradients or updates every parameter with Adam in the same pass
// (adam_scalar :525-550 for structure, the active-prefix SH Adam of cuda_ops.cu:1506-1546),
// including rows culled in every view, whose moments still decay. With reduced-precision storage
// the bf16/fp16 mirrors read by the next forward are rewritten from the FP32 masters here.
#include <algorithm>

#include "launch.h"

namespace pht {

// SH clamp flags of Gaussian i in view v: from its view-mask word when there are at most
// kMaskClampViews view slots, else from the slot's meta record.
__device__ __forceinline__ uint32_t slot_clamp(const GaussGroup& a, uint32_t mask, int64_t i, int v) {
    if (a.vp_log2 <= kMaskClampViewsLog2) return mask_clamp_flags(mask, v);
    return a.meta[(i << a.vp_log2) + v].y >> kMetaClampShift;
}

template <int LP> __device__ __forceinline__ float ld_lp(const float* f, const uint16_t* lp, int64_t i) {
    if constexpr (LP == PHT_LP_NONE) return f[i];
    else return lp_to_f32(lp[i], LP);
}

// SH rows of the block's 256 Gaussians live in LDS with an odd row pitch, so the per-thread row
// reads of the backward are bank-conflict free; loads and stores of the SH arrays (values,
// moments, reduced-precision mirror) are done cooperatively over the block's contiguous range,
// one coalesced pass each, instead of one 192-byte row per lane.
constexpr int kShPitch = 49;
// With bf16 SH storage the backward stages the coefficient rows as loaded, two per 32-bit word
// (odd pitch of 25 words), and the LDS this frees holds, per thread, the colour gradients of its
// first kStashViews visible views, written by the geometry pass for the SH-coefficient pass. That
// pass then reads no splat-gradient record a second time: by the time it runs, the records the
// geometry pass read are long gone from L2, and re-reading them cost about one extra pass over
// the records in HBM traffic.
constexpr int kPackPitch = 25;
constexpr int kStashViews = (kShPitch - kPackPitch) / 3;
static_assert(kPackPitch * 2 >= 48 && kStashViews * 3 + kPackPitch <= kShPitch, "packed rows + stash fit the float rows");
template <int LP> struct PackedCoef {
    const uint32_t* w; // w[e / 2]: coefficient e in the low half when e is even
    __device__ __forceinline__ float operator[](int e) const {
        const uint32_t x = w[e >> 1];
        return lp_to_f32(uint16_t((e & 1) ? (x >> 16) : (x & 0xffffu)), LP);
    }
};

// Walks e = t, t + 256, ... over [0, rows * stride) keeping (row, k) = divmod(e, stride)
// incrementally; calls f(e, row, k) for the active coefficients k < active.
template <class F>
__device__ __forceinline__ void for_block_rows(int rows, int stride, int active, F&& f) {
    const int total = rows * stride;
    const int step_row = kScanThreads / stride, step_k = kScanThreads % stride;
    int row = int(threadIdx.x) / stride, k = int(threadIdx.x) % stride;
    for (int e = threadIdx.x; e < total; e += kScanThreads) {
        if (k < active) f(e, row, k);
        row += step_row;
        k += step_k;
        if (k >= stride) {
            k -= stride;
            ++row;
        }
    }
}

// Walks the block's contiguous range [0, rows * stride) in chunks of four consecutive
// coefficients (stride % 4 == 0, so a chunk never straddles two rows): chunk c = t + 256 u. For
// each group of U chunks it calls load(u, c, row, k0) for all of them before step(u, c, row, k0)
// for any, so a thread keeps U chunks of every array in flight: the SH passes are bound by how
// many bytes are outstanding, and one scalar per array per iteration leaves HBM mostly idle.
template <int U, class Load, class Step>
__device__ __forceinline__ void for_block_chunks(int rows, int stride, Load&& load, Step&& step) {
    const int n = (rows * stride) >> 2;
    for (int c0 = int(threadIdx.x); c0 < n; c0 += U * kScanThreads) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int c = c0 + u * kScanThreads;
            if (c < n) load(u, c, (4 * c) / stride, (4 * c) % stride);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int c = c0 + u * kScanThreads;
            if (c < n) step(u, c, (4 * c) / stride, (4 * c) % stride);
        }
    }
}

// Four stored SH values as floats (fp32, or four bf16/fp16 in one 8-byte load).
template <int LP> __device__ __forceinline__ float4 ld4_lp(const float* f, const uint16_t* lp, int64_t e) {
    if constexpr (LP == PHT_LP_NONE) {
        return *reinterpret_cast<const float4*>(f + e);
    } else {
        const uint2 h = *reinterpret_cast<const uint2*>(lp + e);
        return make_float4(lp_to_f32(uint16_t(h.x & 0xffffu), LP), lp_to_f32(uint16_t(h.x >> 16), LP),
                           lp_to_f32(uint16_t(h.y & 0xffffu), LP), lp_to_f32(uint16_t(h.y >> 16), LP));
    }
}
__device__ __forceinline__ uint2 pack4_lp(float a, float b, float c, float d, int fmt) {
    return make_uint2(uint32_t(f32_to_lp(a, fmt)) | (uint32_t(f32_to_lp(b, fmt)) << 16),
                      uint32_t(f32_to_lp(c, fmt)) | (uint32_t(f32_to_lp(d, fmt)) << 16));
}
__device__ __forceinline__ bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Quantised SH Adam over the block's `rows` consecutive rows starting at row_begin
// (sh_adam_quant_warp_row, third_party/splat_drender/include/splat_drender/sh_adam_quant.cuh:146-234,
// with the per-coefficient step of shq_column). On entry s_row holds each row's gradient in
// columns [0, q.step.active) (pitch kShPitch); its contents are consumed. Two passes:
// (1) cooperative and coalesced over the block's contiguous coefficient range: decode under the
//     row's old bounds, update, write parameter, fp16 update (and the reduced-precision mirror),
//     leave log_s in LDS;
// (2) one thread per row: the DC and non-DC (min, max) of its log_s, the new bounds, and the
//     row's codes, packed four to a word when the row length allows.
struct ShqArgs {
    float* param;
    uint16_t* first;
    uint8_t* packed;
    float* bounds;
    uint16_t* lp;
    ShqStep step;
};

template <int LP>
__device__ void shq_block(const ShqArgs& q, int64_t row_begin, int rows, int stride, float* s_row) {
    const int64_t e0 = row_begin * stride;
    const ShqConst kc = shq_const(q.step);
    const bool chunked = (stride & 3) == 0 && aligned16(q.param + e0) && aligned16(q.bounds) &&
                         (reinterpret_cast<uintptr_t>(q.first + e0) & 7u) == 0 &&
                         (reinterpret_cast<uintptr_t>(q.packed + e0) & 3u) == 0 &&
                         (!q.lp || (reinterpret_cast<uintptr_t>(q.lp + e0) & 7u) == 0);
    if (chunked) {
        constexpr int U = 2;
        float4 pv[U], bv[U];
        uint2 fv[U];
        uint32_t cv[U];
        float* P = q.param + e0;
        uint16_t* F = q.first + e0;
        uint8_t* Cd = q.packed + e0;
        for_block_chunks<U>(
            rows, stride,
            [&](int u, int c, int row, int) {
                pv[u] = reinterpret_cast<const float4*>(P)[c];
                fv[u] = reinterpret_cast<const uint2*>(F)[c];
                cv[u] = reinterpret_cast<const uint32_t*>(Cd)[c];
                bv[u] = reinterpret_cast<const float4*>(q.bounds)[row_begin + row];
            },
            [&](int u, int c, int row, int k0) {
                // the row's (min, max) pairs, a non-finite pair read as (0, 0)
                const bool dc_ok = finite_(bv[u].x) && finite_(bv[u].y), rest_ok = finite_(bv[u].z) && finite_(bv[u].w);
                float p[4] = {pv[u].x, pv[u].y, pv[u].z, pv[u].w};
                const uint32_t h[4] = {fv[u].x & 0xffffu, fv[u].x >> 16, fv[u].y & 0xffffu, fv[u].y >> 16};
                uint32_t nh[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int k = k0 + j;
                    const bool dc = k < 3;
                    const float lo = dc ? (dc_ok ? bv[u].x : 0.0f) : (rest_ok ? bv[u].z : 0.0f);
                    const float hi = dc ? (dc_ok ? bv[u].y : 0.0f) : (rest_ok ? bv[u].w : 0.0f);
                    float uu = fp16_to_f32(uint16_t(h[j]));
                    const bool act = k < q.step.active;
                    const float ls = shq_column(kc, k, p[j], act ? s_row[row * kShPitch + k] : 0.0f, uu,
                                                shq_decode((cv[u] >> (8 * j)) & 255u, lo, hi));
                    nh[j] = f32_to_fp16(uu);
                    s_row[row * kShPitch + k] = ls;
                }
                reinterpret_cast<uint2*>(F)[c] = make_uint2(nh[0] | (nh[1] << 16), nh[2] | (nh[3] << 16));
                if (k0 < q.step.active) {
                    reinterpret_cast<float4*>(P)[c] = make_float4(p[0], p[1], p[2], p[3]);
                    if constexpr (LP != PHT_LP_NONE) {
                        if (q.lp) reinterpret_cast<uint2*>(q.lp + e0)[c] = pack4_lp(p[0], p[1], p[2], p[3], LP);
                    }
                }
            });
    } else for_block_rows(rows, stride, stride, [&](int e, int row, int k) {
        float lo, hi;
        shq_bounds(q.bounds + (row_begin + row) * 4, k, lo, hi);
        float p = q.param[e0 + e];
        float u = fp16_to_f32(q.first[e0 + e]);
        const bool act = k < q.step.active;
        const float ls = shq_column(kc, k, p, act ? s_row[row * kShPitch + k] : 0.0f, u,
                                    shq_decode(q.packed[e0 + e], lo, hi));
        if (act) {
            q.param[e0 + e] = p;
            if constexpr (LP != PHT_LP_NONE) {
                if (q.lp) q.lp[e0 + e] = f32_to_lp(p, LP);
            }
        }
        q.first[e0 + e] = f32_to_fp16(u);
        s_row[row * kShPitch + k] = ls;
    });
    __syncthreads();
    if (int(threadIdx.x) >= rows) return;
    const float* r = s_row + threadIdx.x * kShPitch;
    float b[4] = {1e30f, -1e30f, 1e30f, -1e30f};
    for (int k = 0; k < stride; ++k) {
        const int o = k < 3 ? 0 : 2;
        b[o] = fminf(b[o], r[k]);
        b[o + 1] = fmaxf(b[o + 1], r[k]);
    }
    if (stride <= 3) b[2] = b[3] = 0.0f;
    const int64_t row = row_begin + threadIdx.x;
    for (int o = 0; o < 4; ++o) q.bounds[row * 4 + o] = b[o];
    const float sc[2] = {shq_scale(b[0], b[1]), shq_scale(b[2], b[3])};
    uint8_t* dst = q.packed + row * stride;
    if ((stride & 3) == 0 && (reinterpret_cast<uintptr_t>(q.packed) & 3u) == 0) {
        for (int k = 0; k < stride; k += 4) {
            uint32_t w = 0;
            for (int j = 0; j < 4; ++j) {
                const int g = k + j < 3 ? 0 : 1;
                w |= shq_encode(r[k + j], b[2 * g], sc[g]) << (8 * j);
            }
            *reinterpret_cast<uint32_t*>(dst + k) = w;
        }
    } else {
        for (int k = 0; k < stride; ++k) {
            const int g = k < 3 ? 0 : 1;
            dst[k] = uint8_t(shq_encode(r[k], b[2 * g], sc[g]));
        }
    }
}

// Adam over one structure array of the block's rows (W floats per row), flat and coalesced: element
// e of the block's range is row e / W, component e % W, with its gradient at
// s_g[row * kStructGrads + first + component]. The reduced-precision mirror is rewritten when given.
constexpr int kStructGrads = 11;
template <int W, int LP>
__device__ __forceinline__ void struct_adam_pass(float* P, float* M, float* V, uint16_t* L, int rows, const float* s_g,
                                                 int first, float lr, const pht_adam& A, float c1, float c2, float lo,
                                                 float hi) {
    // rows <= 256, so a thread has at most W elements: all of their loads are issued before any
    // update (a load per array per iteration, behind the previous iteration's stores, left one
    // round trip per element)
    float p[W], m[W], v[W];
#pragma unroll
    for (int it = 0; it < W; ++it) {
        const int e = int(threadIdx.x) + it * kScanThreads;
        if (e < rows * W) {
            p[it] = P[e];
            m[it] = M[e];
            v[it] = V[e];
        }
    }
#pragma unroll
    for (int it = 0; it < W; ++it) {
        const int e = int(threadIdx.x) + it * kScanThreads;
        if (e >= rows * W) break;
        const int row = e / W, k = e - row * W;
        adam_update(p[it], m[it], v[it], s_g[row * kStructGrads + first + k], lr, A.beta1, A.beta2, c1, c2, A.epsilon,
                    lo, hi, false);
        P[e] = p[it];
        M[e] = m[it];
        V[e] = v[it];
        if constexpr (LP != PHT_LP_NONE) {
            if (L) L[e] = f32_to_lp(p[it], LP);
        }
    }
}

// DENS: also fold each view's densification observation of the Gaussian into its statistics
// (the refine statistic is the record's slot 9, already loaded for the geometry pass).
template <int DEG, int LP, bool ADAM, bool DENS = false>
__global__ __launch_bounds__(256) void k_gaussian_bwd(GaussArgs a) {
    constexpr int NB = (DEG + 1) * (DEG + 1);
    __shared__ float s_sh[kScanThreads * kShPitch];
    if constexpr (ADAM) {
        // uniform over the grid: a group whose instance list overflowed skipped its blends
        for (int gr = 0; gr < a.n_groups; ++gr)
            if (a.grp[gr].counters->overflow) return;
    }
    const int64_t i0 = int64_t(blockIdx.x) * kScanThreads;
    const int64_t i = i0 + threadIdx.x;
    const bool live = i < a.g.count;
    const int64_t n = a.g.count;
    const int rows = int(a.g.count - i0 < kScanThreads ? a.g.count - i0 : kScanThreads);
    const int stride = a.s.sh_bases * 3;
    const int64_t row0 = i0 * int64_t(stride);
    const bool precomp = a.g.colors_precomp != nullptr;
    const bool use_sh = !precomp;
    // (bf16 only: the fp16 variant already spills, and the stash's registers would cost it a wave)
    constexpr bool kPacked = LP == PHT_LP_BF16;
    uint32_t* const s_pack = reinterpret_cast<uint32_t*>(s_sh);
    float* const s_stash = s_sh + kScanThreads * kPackPitch; // [(u * 3 + ch) * 256 + thread]
    if (use_sh && kPacked) {
        if ((stride & 3) == 0 && (reinterpret_cast<uintptr_t>(a.g.sh_lp + row0) & 7u) == 0) {
            uint2 v[4];
            for_block_chunks<4>(
                rows, stride,
                [&](int u, int c, int, int k0) {
                    if (k0 < NB * 3) v[u] = reinterpret_cast<const uint2*>(a.g.sh_lp + row0)[c];
                },
                [&](int u, int, int row, int k0) {
                    if (k0 >= NB * 3) return;
                    uint32_t* d = s_pack + row * kPackPitch + (k0 >> 1);
                    d[0] = v[u].x;
                    d[1] = v[u].y;
                });
        } else {
            uint16_t* h = reinterpret_cast<uint16_t*>(s_pack);
            for_block_rows(rows, stride, NB * 3,
                           [&](int e, int row, int k) { h[row * 2 * kPackPitch + k] = a.g.sh_lp[row0 + e]; });
        }
    } else if (use_sh) {
        const bool chunked = (stride & 3) == 0 && aligned16(a.g.sh + row0);
        if (chunked) {
            float4 v[4];
            for_block_chunks<4>(
                rows, stride,
                [&](int u, int c, int, int k0) {
                    if (k0 < NB * 3) v[u] = ld4_lp<LP>(a.g.sh, a.g.sh_lp, row0 + 4 * int64_t(c));
                },
                [&](int u, int, int row, int k0) {
                    if (k0 >= NB * 3) return;
                    float* d = s_sh + row * kShPitch + k0;
                    d[0] = v[u].x;
                    d[1] = v[u].y;
                    d[2] = v[u].z;
                    d[3] = v[u].w;
                });
        } else {
            for_block_rows(rows, stride, NB * 3, [&](int e, int row, int k) {
                s_sh[row * kShPitch + k] = ld_lp<LP>(a.g.sh, a.g.sh_lp, row0 + e);
            });
        }
    }
    __syncthreads();
    float* my_sh = s_sh + threadIdx.x * kShPitch;
    const PackedCoef<LP> my_pack{s_pack + threadIdx.x * kPackPitch};

    SplatGrad<float> sg;
    for (int k = 0; k < 3; ++k) { sg.mean[k] = 0.0f; sg.scale[k] = 0.0f; }
    for (int k = 0; k < 4; ++k) sg.rot[k] = 0.0f;
    sg.opacity = 0.0f;
    float gsh[NB * 3];
#pragma unroll
    for (int k = 0; k < NB * 3; ++k) gsh[k] = 0.0f;
    float gcol[3] = {0.0f, 0.0f, 0.0f};
    float refine = 0.0f;
    float gls[3] = {0.0f, 0.0f, 0.0f}, gq[4] = {0.0f, 0.0f, 0.0f, 0.0f}, glogit = 0.0f;
    if (live) {
        const float mean[3] = {a.g.means[3 * i], a.g.means[3 * i + 1], a.g.means[3 * i + 2]};
        // Pass 1: geometry (EWA chain) and the SH view-direction term of the mean gradient.
        // Pass 2 (below) accumulates the SH coefficient gradients; splitting the view loop keeps
        // the 48 coefficient accumulators out of the geometry pass's registers.
        {
            float ls[3], q[4];
            for (int k = 0; k < 3; ++k) ls[k] = ld_lp<LP>(a.g.log_scales, a.g.log_scales_lp, 3 * i + k);
            for (int k = 0; k < 4; ++k) q[k] = a.g.quaternions[4 * i + k];
            const float logit = ld_lp<LP>(a.g.opacity_logits, a.g.opacity_logits_lp, i);
            const bool hf = a.g.filter_3d != nullptr;
            const float filt = hf ? a.g.filter_3d[i] : 0.0f;
            const Activated<float> act = activate<float>(ls, q, logit, hf, filt);
            int u = 0; // visible views so far: the stash slot of the next one
            // the views' d(3D covariance) is summed and chained to scale and rotation once
            float dcov[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            for (int gr = 0; gr < a.n_groups; ++gr) {
                const GaussGroup& G = a.grp[gr];
                const uint32_t mask = G.view_mask[i];
                for (int v = 0; v < G.n_views; ++v) {
                    if (!((mask >> v) & 1u)) continue;
                    const float2* gp = reinterpret_cast<const float2*>(G.sgrad + (v * n + i) * kGradStride);
                    const float2 g0 = gp[0], g1 = gp[1], g2 = gp[2], g3 = gp[3], g4 = gp[4];
                    const float dconic[4] = {g1.x, g1.y, g2.x, g2.y};
                    const float dm2[2] = {g0.x, g0.y};
                    const pht_camera& cam = G.views[v].vi.cam;
                    splat_backward_pinhole_view(mean, act, cam, a.s, dconic, dm2, sg, dcov);
                    const float dcol[3] = {g3.x, g3.y, g4.x};
                    if (precomp) {
                        gcol[0] += dcol[0]; gcol[1] += dcol[1]; gcol[2] += dcol[2];
                    } else {
                        float d[3], raw[3];
                        view_dir(mean, cam.position, d, raw);
                        if constexpr (kPacked) {
                            if (u < kStashViews) {
#pragma unroll
                                for (int ch = 0; ch < 3; ++ch)
                                    s_stash[(u * 3 + ch) * kScanThreads + threadIdx.x] = dcol[ch];
                            }
                            sh_backward_dir(DEG, raw, d, my_pack, slot_clamp(G, mask, i, v), dcol, sg.mean);
                        } else {
                            sh_backward_dir(DEG, raw, d, my_sh, slot_clamp(G, mask, i, v), dcol, sg.mean);
                        }
                    }
                    refine += g4.y;
                    ++u;
                }
            }
            splat_backward_cov_chain(act, a.s, dcov, sg);
            chain_raw(act, hf, filt, sg, gls, gq, glogit);
        }
        if (a.out.refine_weight) a.out.refine_weight[i] = refine;
        if constexpr (DENS) {
            // its own pass over the views (the records are in cache), so the statistics' registers
            // do not add to the geometry pass's
            // (one view group: the host refuses densification statistics with several)
            const GaussGroup& G = a.grp[0];
            const uint32_t mask = G.view_mask[i];
            DensifyRow dr = densify_load_row(a.dens.stats, i);
            for (int v = 0; v < G.n_views; ++v) {
                if (!((mask >> v) & 1u)) continue;
                const float rw = G.sgrad[(v * n + i) * kGradStride + 9];
                const DevView& dv = G.views[v];
                densify_observe(dr, a.dens.rules, a.dens.view_index[v], a.dens.inverse_resolution[v], dv.radii[i],
                                dv.visibility[i], rw, a.dens.geometry_from_refine != 0, rw, false, 0.0f, false, 0.0f);
            }
            densify_store_row(a.dens.stats, i, dr);
        }
        if (!precomp) {
            int u = 0;
            for (int gr = 0; gr < a.n_groups; ++gr) {
                const GaussGroup& G = a.grp[gr];
                const uint32_t mask = G.view_mask[i];
                for (int v = 0; v < G.n_views; ++v) {
                    if (!((mask >> v) & 1u)) continue;
                    float dcol[3];
                    if (kPacked && u < kStashViews) {
#pragma unroll
                        for (int ch = 0; ch < 3; ++ch) dcol[ch] = s_stash[(u * 3 + ch) * kScanThreads + threadIdx.x];
                    } else {
                        const float2* gp = reinterpret_cast<const float2*>(G.sgrad + (v * n + i) * kGradStride);
                        const float2 g3 = gp[3], g4 = gp[4];
                        dcol[0] = g3.x;
                        dcol[1] = g3.y;
                        dcol[2] = g4.x;
                    }
                    ++u;
                    float d[3], raw[3];
                    view_dir(mean, G.views[v].vi.cam.position, d, raw);
                    sh_backward_coef(DEG, d, slot_clamp(G, mask, i, v), dcol, gsh);
                }
            }
        }
    }
    // own row only: the coefficients just read are replaced by their gradients (packed rows and
    // stashes overlap other threads' float rows: all of them are read first)
    if (use_sh) {
        if constexpr (kPacked) __syncthreads();
#pragma unroll
        for (int k = 0; k < NB * 3; ++k) my_sh[k] = gsh[k];
    }
    __syncthreads();
    if constexpr (!ADAM) {
        const pht_model_grads& o = a.out;
        if (live) {
            for (int k = 0; k < 3; ++k) {
                if (o.means) o.means[3 * i + k] = sg.mean[k];
                if (o.log_scales) o.log_scales[3 * i + k] = gls[k];
                if (o.colors_precomp) o.colors_precomp[3 * i + k] = gcol[k];
            }
            for (int k = 0; k < 4; ++k)
                if (o.quaternions) o.quaternions[4 * i + k] = gq[k];
            if (o.opacity_logits) o.opacity_logits[i] = glogit;
        }
        if (o.sh) {
            float* dst = o.sh + row0;
            for_block_rows(rows, stride, stride, [&](int e, int row, int k) {
                dst[e] = (use_sh && k < NB * 3) ? s_sh[row * kShPitch + k] : 0.0f;
            });
        }
    } else {
        const pht_adam& A = a.adam;
        const float inf = __builtin_huge_valf();
        float c1 = A.correction1, c2 = A.correction2, lr_means = A.lr_means;
        if (A.device_hparams) {
            c1 = A.device_hparams[0];
            c2 = A.device_hparams[1];
            lr_means = A.device_hparams[2];
        }
        if (A.sh && use_sh && A.sh_first) {
            ShqArgs qa;
            qa.param = A.sh;
            qa.first = A.sh_first;
            qa.packed = A.sh_packed;
            qa.bounds = A.sh_bounds;
            qa.lp = LP != PHT_LP_NONE ? A.sh_lp : nullptr;
            qa.step = ShqStep{A.lr_sh0, A.lr_sh_rest, A.beta1, A.beta2, c1, c2, A.epsilon, A.sh_regularization, NB * 3};
            shq_block<LP>(qa, i0, rows, stride, s_sh);
        } else if (A.sh && use_sh) {
            float* P = A.sh + row0;
            float* M = A.m_sh + row0;
            float* V = A.v_sh + row0;
            uint16_t* L = (LP != PHT_LP_NONE && A.sh_lp) ? A.sh_lp + row0 : nullptr;
            const bool chunked = (stride & 3) == 0 && aligned16(P) && aligned16(M) && aligned16(V) &&
                                 (!L || (reinterpret_cast<uintptr_t>(L) & 7u) == 0);
            if (chunked) {
                constexpr int U = 4;
                float4 pv[U], mv[U], vv[U];
                for_block_chunks<U>(
                    rows, stride,
                    [&](int u, int c, int, int k0) {
                        if (k0 >= NB * 3) return;
                        pv[u] = reinterpret_cast<const float4*>(P)[c];
                        mv[u] = reinterpret_cast<const float4*>(M)[c];
                        vv[u] = reinterpret_cast<const float4*>(V)[c];
                    },
                    [&](int u, int c, int row, int k0) {
                        if (k0 >= NB * 3) return;
                        float p[4] = {pv[u].x, pv[u].y, pv[u].z, pv[u].w};
                        float m[4] = {mv[u].x, mv[u].y, mv[u].z, mv[u].w};
                        float v[4] = {vv[u].x, vv[u].y, vv[u].z, vv[u].w};
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int k = k0 + j;
                            if (k < NB * 3)
                                adam_update(p[j], m[j], v[j], s_sh[row * kShPitch + k], k < 3 ? A.lr_sh0 : A.lr_sh_rest,
                                            A.beta1, A.beta2, c1, c2, A.epsilon, -inf, inf, true);
                        }
                        reinterpret_cast<float4*>(P)[c] = make_float4(p[0], p[1], p[2], p[3]);
                        reinterpret_cast<float4*>(M)[c] = make_float4(m[0], m[1], m[2], m[3]);
                        reinterpret_cast<float4*>(V)[c] = make_float4(v[0], v[1], v[2], v[3]);
                        if constexpr (LP != PHT_LP_NONE) {
                            if (L) reinterpret_cast<uint2*>(L)[c] = pack4_lp(p[0], p[1], p[2], p[3], LP);
                        }
                    });
            } else for_block_rows(rows, stride, NB * 3, [&](int e, int row, int k) {
                float pv = P[e], mv = M[e], vv = V[e];
                adam_update(pv, mv, vv, s_sh[row * kShPitch + k], k < 3 ? A.lr_sh0 : A.lr_sh_rest, A.beta1,
                            A.beta2, c1, c2, A.epsilon, -inf, inf, true);
                P[e] = pv;
                M[e] = mv;
                V[e] = vv;
                if constexpr (LP != PHT_LP_NONE) {
                    if (L) L[e] = f32_to_lp(pv, LP);
                }
            });
        }
        // Structure parameters: the rows' 11 gradients go through LDS (the SH rows are done with
        // it), then each parameter array of the block's contiguous rows is updated in one
        // coalesced pass, as the SH values are; per-lane rows of 3, 4 and 1 floats made every
        // load and store a strided access.
        __syncthreads();
        float* s_g = s_sh; // [row][kStructGrads]: mean 0-2, log-scale 3-5, quaternion 6-9, logit 10
        if (live) {
            float* d = s_g + threadIdx.x * kStructGrads;
            for (int k = 0; k < 3; ++k) {
                d[k] = sg.mean[k];
                d[3 + k] = gls[k];
            }
            for (int k = 0; k < 4; ++k) d[6 + k] = gq[k];
            d[10] = glogit;
        }
        __syncthreads();
        struct_adam_pass<3, LP>(A.means + 3 * i0, A.m_means + 3 * i0, A.v_means + 3 * i0, nullptr, rows, s_g, 0,
                                lr_means, A, c1, c2, -inf, inf);
        struct_adam_pass<3, LP>(A.log_scales + 3 * i0, A.m_log_scales + 3 * i0, A.v_log_scales + 3 * i0,
                                A.log_scales_lp ? A.log_scales_lp + 3 * i0 : nullptr, rows, s_g, 3, A.lr_scales, A,
                                c1, c2, A.min_log_scale, A.max_log_scale);
        struct_adam_pass<4, LP>(A.quaternions + 4 * i0, A.m_quaternions + 4 * i0, A.v_quaternions + 4 * i0, nullptr,
                                rows, s_g, 6, A.lr_quaternions, A, c1, c2, -inf, inf);
        struct_adam_pass<1, LP>(A.opacity_logits + i0, A.m_opacity_logits + i0, A.v_opacity_logits + i0,
                                A.opacity_logits_lp ? A.opacity_logits_lp + i0 : nullptr, rows, s_g, 10, A.lr_opacity,
                                A, c1, c2, -12.0f, 12.0f);
    }
}

template <int DEG, int LP> static void gb_launch(const GaussArgs& a, bool adam, hipStream_t st) {
    const unsigned blocks = unsigned((a.g.count + 255) / 256);
    if (adam && a.has_dens) hipLaunchKernelGGL((k_gaussian_bwd<DEG, LP, true, true>), dim3(blocks), dim3(256), 0, st, a);
    else if (adam) hipLaunchKernelGGL((k_gaussian_bwd<DEG, LP, true>), dim3(blocks), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_gaussian_bwd<DEG, LP, false>), dim3(blocks), dim3(256), 0, st, a);
}
template <int LP> static void gb_deg(const GaussArgs& a, bool adam, hipStream_t st) {
    switch (a.s.active_sh_degree) {
    case 0: gb_launch<0, LP>(a, adam, st); break;
    case 1: gb_launch<1, LP>(a, adam, st); break;
    case 2: gb_launch<2, LP>(a, adam, st); break;
    default: gb_launch<3, LP>(a, adam, st); break;
    }
}
void launch_gaussian_backward(const GaussArgs& a, bool adam, hipStream_t st) {
    if (a.g.count <= 0) return;
    if (a.g.lp_format == PHT_LP_BF16) gb_deg<PHT_LP_BF16>(a, adam, st);
    else if (a.g.lp_format == PHT_LP_FP16) gb_deg<PHT_LP_FP16>(a, adam, st);
    else gb_deg<PHT_LP_NONE>(a, adam, st);
}

// ---------------------------------------------------------------- standalone quantised SH Adam
// sh_adam_quant_step's kernel (photara/src/splat/sh_adam_quant.cu): 256 rows per block, the
// active gradients staged into LDS with one coalesced pass, then shq_block.
__global__ __launch_bounds__(256) void k_sh_adam_quant(float* param, const float* grad, uint16_t* first,
                                                       uint8_t* packed, float* bounds, int64_t n_rows, int stride,
                                                       ShqStep step) {
    __shared__ float s_row[kScanThreads * kShPitch];
    const int64_t r0 = int64_t(blockIdx.x) * kScanThreads;
    const int rows = int(n_rows - r0 < kScanThreads ? n_rows - r0 : kScanThreads);
    const float* g = grad ? grad + r0 * stride : nullptr;
    if (g) for_block_rows(rows, stride, step.active, [&](int e, int row, int k) { s_row[row * kShPitch + k] = g[e]; });
    __syncthreads();
    ShqArgs q;
    q.param = param;
    q.first = first;
    q.packed = packed;
    q.bounds = bounds;
    q.lp = nullptr;
    q.step = step;
    shq_block<PHT_LP_NONE>(q, r0, rows, stride, s_row);
}

void launch_sh_adam_quant(float* param, const float* grad, uint16_t* first, uint8_t* packed, float* bounds,
                          int64_t rows, int stride, const ShqStep& step, hipStream_t st) {
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_sh_adam_quant, dim3(unsigned((rows + kScanThreads - 1) / kScanThreads)), dim3(kScanThreads),
                       0, st, param, grad, first, packed, bounds, rows, stride, step);
}

// ---------------------------------------------------------------- standalone Adam
__global__ __launch_bounds__(256) void k_adam(float* p, const float* g, float* m, float* v, int64_t rows,
                                              int64_t stride, int64_t active, float lr, float lr2,
                                              int64_t secondary_from, float b1, float b2, float c1, float c2,
                                              float eps, float lo, float hi, int prefix) {
    const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (t >= rows * active) return;
    const int64_t r = t / active, c = t % active;
    const int64_t j = r * stride + c;
    const float lrx = (secondary_from > 0 && c >= secondary_from) ? lr2 : lr;
    adam_update(p[j], m[j], v[j], g[j], lrx, b1, b2, c1, c2, eps, lo, hi, prefix != 0);
}

void launch_adam_standalone(float* p, const float* g, float* m, float* v, int64_t rows, int64_t stride,
                            int64_t active, float lr, float lr2, int64_t secondary_from, float b1, float b2,
                            float c1, float c2, float eps, float lo, float hi, int prefix, hipStream_t st) {
    const int64_t n = rows * active;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_adam, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, p, g, m, v, rows, stride, active,
                       lr, lr2, secondary_from, b1, b2, c1, c2, eps, lo, hi, prefix);
}

// ---------------------------------------------------------------- L1 loss
// loss_kernel L1 part (photara/src/splat/cuda_ops.cu:886-898), no mask, full image. A bounded
// grid strides over the pixels (four at a time when the planes are 16-byte aligned), so the
// optional loss sum takes one atomic per block.

// One view's L1 term over a bounded grid of `nblocks` blocks (this block is `block`).
template <bool VEC>
__device__ __forceinline__ void l1_view(const float* color, const float* target, float* grad, float* loss_sum,
                                        int64_t pixels, float weight, int64_t block, int64_t nblocks) {
    __shared__ float s_w[4];
    const float inv_p = 1.0f / float(pixels);
    const float scale = (weight * inv_p) / 3.0f;
    const int64_t stride = nblocks * 256;
    float part = 0.0f;
    if constexpr (VEC) {
        const int64_t n4 = pixels >> 2;
        for (int64_t q = block * 256 + threadIdx.x; q < n4; q += stride) {
            for (int ch = 0; ch < 3; ++ch) {
                const int64_t o = (ch * pixels >> 2) + q;
                const float4 c = reinterpret_cast<const float4*>(color)[o];
                const float4 t = reinterpret_cast<const float4*>(target)[o];
                float4 g;
                part += l1_term(c.x, t.x, scale, g.x);
                part += l1_term(c.y, t.y, scale, g.y);
                part += l1_term(c.z, t.z, scale, g.z);
                part += l1_term(c.w, t.w, scale, g.w);
                reinterpret_cast<float4*>(grad)[o] = g;
            }
        }
    } else {
        for (int64_t p = block * 256 + threadIdx.x; p < pixels; p += stride) {
            for (int ch = 0; ch < 3; ++ch) {
                const int64_t o = ch * pixels + p;
                float g;
                part += l1_term(color[o], target[o], scale, g);
                grad[o] = g;
            }
        }
    }
    if (loss_sum) {
        part = wave_sum_to_last(part);
        if (lane_id() == 63) s_w[threadIdx.x >> 6] = part;
        __syncthreads();
        if (threadIdx.x == 0) atomic_add_f32(loss_sum, ((s_w[0] + s_w[1]) + (s_w[2] + s_w[3])) * scale);
    }
}

template <bool VEC>
__global__ __launch_bounds__(256) void k_l1(const float* color, const float* target, float* grad, float* loss_sum,
                                            int64_t pixels, float weight) {
    l1_view<VEC>(color, target, grad, loss_sum, pixels, weight, blockIdx.x, gridDim.x);
}

void launch_l1_loss(const float* color, const float* target, float* grad, float* loss_sum, int64_t pixels,
                    float weight, hipStream_t st) {
    if (pixels <= 0) return;
    const bool vec = (pixels & 3) == 0 && ((reinterpret_cast<uintptr_t>(color) | reinterpret_cast<uintptr_t>(target) |
                                            reinterpret_cast<uintptr_t>(grad)) & 15u) == 0;
    const int64_t items = vec ? pixels >> 2 : pixels;
    int64_t blocks = (items + 255) / 256;
    if (blocks > 512) blocks = 512;
    if (vec) hipLaunchKernelGGL(k_l1<true>, dim3(unsigned(blocks)), dim3(256), 0, st, color, target, grad, loss_sum, pixels, weight);
    else hipLaunchKernelGGL(k_l1<false>, dim3(unsigned(blocks)), dim3(256), 0, st, color, target, grad, loss_sum, pixels, weight);
}

__global__ void k_zero_f32(float* p, int64_t n) {
    const int64_t stride = int64_t(gridDim.x) * 256;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) p[i] = 0.0f;
}

// Zero fill by a kernel (no memset node, so it may be recorded into a graph). Grid-stride over at
// most 4096 blocks.
void launch_zero_f32(float* p, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_zero_f32, dim3(unsigned(blocks)), dim3(256), 0, st, p, n);
}

} // namespace pht
