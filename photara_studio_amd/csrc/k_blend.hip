// Per-tile alpha blending, forward and backward, for gfx950 (wave64, 160 KiB LDS per CU).
//
// One wave per (view, tile) work item of the batched launch, blocks mapped XCD-contiguously.
// A 16x16 tile is four 8x8 quadrants; lane l owns pixel (l & 7, l >> 3) of every quadrant, so
// each lane carries four pixels and each quadrant is a wave-uniform unit of culling (an 8x8 tile
// is one quadrant). Instances of the tile's sorted list are staged through LDS in batches of
// BATCH by the wave itself (no workgroup barrier beyond the wave's own); a per-instance quadrant
// mask, computed by the loading lane from the record's precomputed ellipse extents, lets the wave
// skip quadrants an instance cannot reach, without changing any blend decision.
//
// Forward  : blend_tile (third_party/splat_drender/src/render_forward.cu:274-538), colour path.
// Backward : the closed-form per-pair terms of blend_bucket_backward (render_backward.cu:383-489),
//            walked back to front per pixel from the stored final transmittance and contributor
//            count (T_before = T_after / (1 - alpha)), with no per-bucket snapshots. Each lane
//            sums its pixels' terms, the wave sums the ten values through LDS (wave_reduce10_lds),
//            and ten lanes commit them with one global atomic instruction per (instance, tile).
#include <algorithm>

#include "launch.h"

namespace pht {

// Minimum over the rectangle [u0, u1] x [v0, v1] of the positive-definite form
// q(u, v) = a u^2 + 2 b u v + c v^2: 0 when the rectangle holds the origin, else the least of
// its four edges' minima (each edge's unconstrained minimiser clamped to the edge).
__device__ __forceinline__ float form_min_box(float a, float b, float c, float ia, float ic, float u0, float u1,
                                              float v0, float v1) {
    if (u0 <= 0.0f && u1 >= 0.0f && v0 <= 0.0f && v1 >= 0.0f) return 0.0f;
    auto q = [&](float u, float v) { return (a * u + 2.0f * b * v) * u + c * v * v; };
    auto on_u = [&](float u) { return q(u, fminf(fmaxf(-b * u * ic, v0), v1)); };
    auto on_v = [&](float v) { return q(fminf(fmaxf(-b * v * ia, u0), u1), v); };
    return fminf(fminf(on_u(u0), on_u(u1)), fminf(on_v(v0), on_v(v1)));
}

// Quadrant k of a TS x TS tile: bit k set when the instance can reach that 8x8 patch: its culling
// box overlaps the patch and its alpha >= 1/255 ellipse (with the record's grown bound on q)
// meets the patch's pixel-centre rectangle. Conservative; the exact per-pixel test still decides.
template <int Q>
__device__ __forceinline__ uint32_t quadrant_mask(const SplatRec& r, int x0, int y0) {
    const float mx = r.a.x, my = r.a.y, ex = r.c.y, ey = r.c.z, qmax = r.c.w;
    const float a = r.d.x, b = r.d.y, c = r.d.z;
    const bool exact = qmax < __builtin_huge_valf();
    const float ia = exact ? 1.0f / a : 0.0f, ic = exact ? 1.0f / c : 0.0f;
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const float px0 = float(x0 + (k & 1) * 8), py0 = float(y0 + (k >> 1) * 8);
        if (!(mx + ex >= px0 && mx - ex <= px0 + 7.0f && my + ey >= py0 && my - ey <= py0 + 7.0f)) continue;
        if (exact && form_min_box(a, b, c, ia, ic, mx - (px0 + 7.0f), mx - px0, my - (py0 + 7.0f), my - py0) > qmax)
            continue;
        m |= 1u << k;
    }
    return m;
}

// Round-to-integer float to int32 by v_cvt_i32_f32, which saturates out-of-range values and maps
// NaN to 0. Written as asm so that an out-of-range argument is a defined value (a C++ conversion
// of it would be undefined behaviour).
__device__ __forceinline__ int cvt_i32_sat(float x) {
    int k;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(k) : "v"(x));
    return k;
}

// pexp2f_blend (splat_math.h) for an argument in [kSkipPower2, 0]: the same rounding and
// polynomial, and the scaling by 2^k as one v_ldexp. p * 2^k is exact for k >= -8 (p is near 1),
// so the result equals pexp2f_blend's product bit for bit. Any other argument (including inf and
// NaN) gives some float, with no undefined behaviour.
__device__ __forceinline__ float exp2_pair(float x) {
    const float kf = rintf(x);
    return __builtin_amdgcn_ldexpf(pexp2f_poly(x - kf), cvt_i32_sat(kf));
}

// pair_alpha2 (splat_math.h) without early returns, so a wave evaluates it with no exec-mask
// branches: identical G, alpha and decision for every pair pair_alpha2 accepts (the exponential's
// argument is the same exponent), including a NaN exponent. With FINITE_G a rejected pair gets
// G = 1 (so alpha is finite and below 1); without it, a rejected pair's G and alpha are
// unspecified (possibly inf or NaN) and must not be used.
template <bool FINITE_G>
__device__ __forceinline__ bool pair_alpha_flat(float a2, float b2, float c2, float op, float dx, float dy,
                                                float clip, float& alpha, float& G) {
    const float p2 = pair_power2(a2, b2, c2, dx, dy);
    const bool ok = !(p2 > 0.0f) && !(p2 < kSkipPower2);
    G = exp2_pair(FINITE_G ? (ok ? p2 : 0.0f) : p2);
    alpha = min_(clip, op * G);
    return ok && !(alpha < K<float>::alpha_floor());
}

// The backward's alpha of a pair: the same decision as pair_alpha_flat (and so as the forward),
// with G from the hardware exponential (v_exp_f32, within a few ulp of the forward's polynomial)
// instead of the nine-instruction polynomial. The decision compares alpha with the 1/255 floor, so
// a lane whose alpha lies within kFloorMargin of it (about 64 ulp; the two exponentials differ by
// far less) takes the forward's exponential and its exact decision; every other lane's decision
// cannot differ from it. G and alpha feed only the gradients, which are compared to a tolerance.
// A rejected pair's G and alpha are unspecified (possibly inf or NaN) and must not be used.
constexpr float kFloorMargin = 3.0e-8f;
__device__ __forceinline__ bool pair_alpha_bwd(float a2, float b2, float c2, float op, float dx, float dy,
                                               float clip, float& alpha, float& G) {
    // (non-short-circuit & on the predicates: lane masks, no branches)
    const float p2 = pair_power2(a2, b2, c2, dx, dy);
    G = __builtin_amdgcn_exp2f(p2);
    alpha = min_(clip, op * G);
    if (fabsf(alpha - K<float>::alpha_floor()) <= kFloorMargin) { // rare: skipped by the wave almost always
        G = exp2_pair(p2);
        alpha = min_(clip, op * G);
    }
    return !(p2 > 0.0f) & !(p2 < kSkipPower2) & !(alpha < K<float>::alpha_floor());
}

struct TileCoord {
    int tile, v, tx, ty;
};

// Work item -> (view, tile): the heaviest-first order when one is given, else the flat grid's
// XCD-contiguous map (consecutive items of one XCD take neighbouring tiles).
__device__ __forceinline__ TileCoord tile_of_item(const BlendArgs& a, int item, int n_items, const DevView*& dv) {
    TileCoord c;
    c.tile = a.tile_order ? int(a.tile_order[item]) : (a.queue ? item : xcd_contiguous(item, n_items));
    dv = &view_of_tile(a.views, a.n_views, c.tile, c.v);
    const int local = c.tile - dv->vi.tile_base;
    c.tx = local % dv->vi.grid_x;
    c.ty = local / dv->vi.grid_x;
    return c;
}

// LDS hand-over between the lanes of one wave. Each wave stages into its own LDS region and no
// other wave reads it, so no workgroup barrier is needed (the waves of a workgroup run different
// tiles, with different trip counts): a wave's LDS operations complete in issue order, and the
// fences keep the compiler from moving LDS accesses across this point.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The work loop of a blend launch. Flat grid: wave w of block b takes item b * WPB + w.
// Persistent queue (a.queue set): the grid holds only as many workgroups as are resident at once,
// and each wave takes the next item from a device counter until the items run out (every wave
// reaches the exit: the counter only grows, and an item >= n_items ends the loop).
template <int WPB, class F>
__device__ __forceinline__ void for_each_item(const BlendArgs& a, uint32_t* queue, int n_items, F&& body) {
    const int wave = int(__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)), lane = int(threadIdx.x) & 63;
    if (!queue) {
        const int item = int(blockIdx.x) * WPB + wave;
        if (item < n_items) body(item);
        return;
    }
    for (;;) {
        uint32_t got = 0;
        if (lane == 0) got = atomicAdd(queue, 1u);
        const int item = int(__builtin_amdgcn_readfirstlane(got));
        if (item >= n_items) return;
        body(item);
    }
}

// One (view, tile) work item of the forward: front-to-back blend of the tile's sorted list.
template <int TS, int BATCH>
__device__ __forceinline__ void blend_fwd_item(const BlendArgs& a, int item, int n_items, float4* sA, float4* sB,
                                               float* sC, uint32_t* sM, uint32_t* sV) {
    constexpr int Q = (TS / 8) * (TS / 8);
    const DevView* dvp;
    const TileCoord tc = tile_of_item(a, item, n_items, dvp);
    const DevView& dv = *dvp;
    const int lane = int(threadIdx.x) & 63;
    const int W = dv.vi.cam.width, H = dv.vi.cam.height;
    // A pixel that has stopped (or lies outside the image) carries its transmittance negated: the
    // saturation test T (1 - alpha) < floor then holds for it at every later pair, so it needs no
    // flag of its own, and |T| is its final transmittance.
    float pxf[Q], pyf[Q], T[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const int px = tc.tx * TS + (k & 1) * 8 + (lane & 7);
        const int py = tc.ty * TS + (k >> 1) * 8 + (lane >> 3);
        pxf[k] = float(px);
        pyf[k] = float(py);
        T[k] = px < W && py < H ? 1.0f : -1.0f;
    }
    const uint2 range = a.ranges[tc.tile];

    float C0[Q], C1[Q], C2[Q];
    uint32_t last[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        C0[k] = C1[k] = C2[k] = 0.0f;
        last[k] = 0;
    }
    uint32_t live = 0; // quadrants with a pixel still blending (wave-uniform)
#pragma unroll
    for (int k = 0; k < Q; ++k)
        if (__ballot(T[k] > 0.0f) != 0) live |= 1u << k;
    // canonical once, so the per-pair min(clip, .) needs no canonicalisation of its own
    const float clip = __builtin_canonicalizef(a.alpha_clip);
    const float t_floor = K<float>::transmittance_floor();
    uint32_t my_slot[BATCH / 64 > 0 ? BATCH / 64 : 1];
    for (uint32_t base = range.x; base < range.y && live != 0; base += BATCH) {
        const uint32_t cnt = min(uint32_t(BATCH), range.y - base);
#pragma unroll
        for (int q = 0; q < (BATCH + 63) / 64; ++q) {
            const uint32_t j = uint32_t(q * 64 + lane);
            my_slot[q] = 0xffffffffu;
            if (j < cnt) {
                const uint32_t slot = a.inst_vals[base + j];
                uint32_t m = 0;
                if (slot < a.slot_cap) {
                    const SplatRec& r = a.rec[slot];
                    sA[j] = r.a;
                    sB[j] = r.b;
                    sC[j] = r.c.x;
                    m = quadrant_mask<Q>(r, tc.tx * TS, tc.ty * TS);
                    my_slot[q] = slot;
                }
                sM[j] = m;
                sV[j] = 0u;
            }
        }
        wave_lds_sync();
        for (uint32_t j = 0; j < cnt; ++j) {
            const uint32_t m = __builtin_amdgcn_readfirstlane(sM[j]) & live;
            if (m == 0) continue;
            const float4 A = sA[j], B = sB[j];
            const float cb = sC[j];
            bool blended = false;
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                if (!((m >> k) & 1u)) continue;
                float alpha, G;
                if (pair_alpha_flat<false>(A.z, A.w, B.x, B.y, A.x - pxf[k], A.y - pyf[k], clip, alpha, G)) {
                    const float test_t = T[k] * (1.0f - alpha);
                    if (test_t < t_floor) {
                        T[k] = -fabsf(T[k]);
                    } else {
                        const float a_t = alpha * T[k];
                        C0[k] = fmaf(B.z, a_t, C0[k]);
                        C1[k] = fmaf(B.w, a_t, C1[k]);
                        C2[k] = fmaf(cb, a_t, C2[k]);
                        T[k] = test_t;
                        last[k] = base - range.x + j + 1;
                        blended = true;
                    }
                }
            }
            if (__ballot(blended) != 0 && lane == 0) sV[j] = 1u;
        }
#pragma unroll
        for (int k = 0; k < Q; ++k)
            if (__ballot(T[k] > 0.0f) == 0) live &= ~(1u << k);
        wave_lds_sync();
        if (dv.visibility) {
#pragma unroll
            for (int q = 0; q < (BATCH + 63) / 64; ++q) {
                const uint32_t j = uint32_t(q * 64 + lane);
                if (j < cnt && sV[j] && my_slot[q] != 0xffffffffu) dv.visibility[my_slot[q] >> a.vp_log2] = 1.0f;
            }
        }
    }
    const int64_t P = int64_t(W) * H;
    // With the fused L1 loss the epilogue also writes dL/dC = scale * sign(C - target) from the
    // colour it has just composited (the same float value the separate loss kernel would read
    // back) and adds the tile's sum of |C - target| to the loss.
    const bool l1 = dv.l1_target != nullptr; // uniform over the item
    float l1_part = 0.0f;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const int px = int(pxf[k]), py = int(pyf[k]);
        if (px < W && py < H) {
            const int64_t pid = int64_t(py) * W + px;
            const float Tk = fabsf(T[k]);
            const float c[3] = {C0[k] + Tk * dv.vi.cam.background[0], C1[k] + Tk * dv.vi.cam.background[1],
                                C2[k] + Tk * dv.vi.cam.background[2]};
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) dv.color[ch * P + pid] = c[ch];
            dv.alpha[pid] = 1.0f - Tk;
            a.final_T[dv.vi.pixel_base + pid] = Tk;
            a.n_contrib[dv.vi.pixel_base + pid] = last[k];
            if (l1) {
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) {
                    float g;
                    l1_part += l1_term(c[ch], dv.l1_target[ch * P + pid], dv.l1_scale, g);
                    dv.l1_grad[ch * P + pid] = g;
                }
            }
        }
    }
    if (l1 && a.loss_sum) {
        l1_part = wave_sum_to_last(l1_part);
        if (lane == 63) atomic_add_f32(a.loss_sum, l1_part * dv.l1_scale);
    }
}

template <int TS, int BATCH, int WPB>
// 8 waves per SIMD: the kernel fits 64 VGPRs without spilling (69 unconstrained, 7 waves), and the
// extra wave hides more of the per-instance LDS reads (measured: 1.92 -> 1.83 ms on the batched
// step). The backward does not fit its bound without spilling and keeps the default.
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_blend_fwd(BlendArgs a,
                                                                                                   int n_items) {
    {
        // Clear the splat-gradient records the backward accumulates into (n_views * N records of
        // kGradStride floats, view-major): each block streams zeros over its share. The stores
        // retire in the background of this VALU-bound kernel instead of taking HBM time of their own.
        static_assert(kGradStride % 2 == 0, "records are cleared as float pairs");
        const int64_t n2 = int64_t(a.n_views) * a.n_gauss * (kGradStride / 2);
        const int64_t nb = gridDim.x, b = blockIdx.x;
        const int64_t lo = n2 * b / nb, hi = n2 * (b + 1) / nb;
        float2* g2 = reinterpret_cast<float2*>(a.sgrad);
        for (int64_t j = lo + threadIdx.x; j < hi; j += 64 * WPB) g2[j] = make_float2(0.f, 0.f);
    }
    if (a.counters->overflow) return; // incomplete instance list: the host re-runs with more capacity
    __shared__ float4 sA[WPB][BATCH];
    __shared__ float4 sB[WPB][BATCH];
    __shared__ float sC[WPB][BATCH];
    __shared__ uint32_t sM[WPB][BATCH];
    __shared__ uint32_t sV[WPB][BATCH];
    const int w = int(__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)); // wave-uniform
    for_each_item<WPB>(a, a.queue ? a.queue : nullptr, n_items, [&](int item) {
        blend_fwd_item<TS, BATCH>(a, item, n_items, sA[w], sB[w], sC[w], sM[w], sV[w]);
    });
}

// Sum over the wave of ten per-lane values, through the wave's own LDS region red[10 * kRedPitch]:
// each lane stores its values value-major (red[g * kRedPitch + lane]); lane 4g + q (g < 10,
// q < 4) adds value g of source lanes 16q .. 16q + 15 (four 16-byte reads), and two quad-permute
// adds complete the sum. Lane 4g ends with the total of value g (slot_of_lane). The stores and
// reads go to the LDS pipe, which the backward leaves nearly idle: 17 VALU instructions against
// the butterfly's ~45. Every lane must be active.
constexpr int kRedPitch = 68; // 64 lanes + 4: value rows start on different banks
__device__ __forceinline__ int slot_of_lane(int lane) { return (lane & 3) == 0 && lane < 40 ? lane >> 2 : -1; }
__device__ __forceinline__ float wave_reduce10_lds(const float (&r)[10], int lane, float* red) {
#pragma unroll
    for (int g = 0; g < 10; ++g) red[g * kRedPitch + lane] = r[g];
    wave_lds_sync();
    float s = 0.0f;
    if (lane < 40) {
        const float4* p = reinterpret_cast<const float4*>(red + (lane >> 2) * kRedPitch + (lane & 3) * 16);
        const float4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
        s = (((q0.x + q0.y) + (q0.z + q0.w)) + ((q1.x + q1.y) + (q1.z + q1.w))) +
            (((q2.x + q2.y) + (q2.z + q2.w)) + ((q3.x + q3.y) + (q3.z + q3.w)));
    }
    wave_lds_sync(); // the region is free for the next instance's values
    s += dpp_move<0xb1, 0xf>(s); // quad_perm [1, 0, 3, 2]
    s += dpp_move<0x4e, 0xf>(s); // quad_perm [2, 3, 0, 1]
    return s;
}

// One (view, tile) work item of the backward: back-to-front from the stored final transmittance.
template <int TS, int BATCH, bool REFINE>
__device__ __forceinline__ void blend_bwd_item(const BlendArgs& a, int item, int n_items, float4* sA, float4* sB,
                                               float4* sC, uint32_t* sM, uint32_t* sS, float* sR) {
    constexpr int Q = (TS / 8) * (TS / 8);
    constexpr int NG = 10;
    const DevView* dvp;
    const TileCoord tc = tile_of_item(a, item, n_items, dvp);
    const DevView& dv = *dvp;
    const int lane = int(threadIdx.x) & 63;
    const int slot_g = slot_of_lane(lane);
    const int W = dv.vi.cam.width, H = dv.vi.cam.height;
    const uint2 range = a.ranges[tc.tile];

    // E = (colour composited behind) . dL/dC + T_final * dL/dT_final, the pixel's whole
    // "behind" term of dL/dalpha (times 1/(1 - alpha)); it is carried as one scalar instead of
    // the three colour channels.
    float pxf[Q], pyf[Q], gp0[Q], gp1[Q], gp2[Q], refine_den[Q];
    float T[Q], E[Q];
    uint32_t last[Q], qmax[Q];
    const int64_t P = int64_t(W) * H;
    uint32_t tile_max = 0;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const int px = tc.tx * TS + (k & 1) * 8 + (lane & 7);
        const int py = tc.ty * TS + (k >> 1) * 8 + (lane >> 3);
        pxf[k] = float(px);
        pyf[k] = float(py);
        last[k] = 0;
        float Tfin = 1.0f, dL_dTfin = 0.0f;
        gp0[k] = gp1[k] = gp2[k] = refine_den[k] = 0.0f;
        if (px < W && py < H) {
            const int64_t pid = int64_t(py) * W + px;
            last[k] = a.n_contrib[dv.vi.pixel_base + pid];
            Tfin = a.final_T[dv.vi.pixel_base + pid];
            gp0[k] = dv.grad_color[pid];
            gp1[k] = dv.grad_color[P + pid];
            gp2[k] = dv.grad_color[2 * P + pid];
            const float ga = dv.grad_alpha ? dv.grad_alpha[pid] : 0.0f;
            const float* bg = dv.vi.cam.background;
            dL_dTfin = ((bg[0] * gp0[k] + bg[1] * gp1[k]) + bg[2] * gp2[k]) - ga;
            if constexpr (REFINE) refine_den[k] = 1.0f / fmaxf(1.0f - Tfin, 1.0e-5f);
        }
        T[k] = Tfin;
        E[k] = Tfin * dL_dTfin;
        qmax[k] = __builtin_amdgcn_readfirstlane(wave_max_u32(last[k]));
        tile_max = qmax[k] > tile_max ? qmax[k] : tile_max;
    }

    // canonical once, so the per-pair min(clip, .) needs no canonicalisation of its own
    const float clip = __builtin_canonicalizef(a.alpha_clip);
    const float fW = float(W), fH = float(H);
    for (int hi = int(tile_max); hi > 0; hi -= BATCH) {
        const int lo = hi - BATCH > 0 ? hi - BATCH : 0;
        const int cnt = hi - lo;
        wave_lds_sync(); // the previous batch's LDS reads are complete
#pragma unroll
        for (int q = 0; q < (BATCH + 63) / 64; ++q) {
            const int j = q * 64 + lane;
            if (j < cnt) {
                const uint32_t slot = a.inst_vals[range.x + lo + j];
                uint32_t m = 0;
                if (slot < a.slot_cap) {
                    const SplatRec& r = a.rec[slot];
                    sA[j] = r.a;
                    sB[j] = r.b;
                    sC[j] = make_float4(r.c.x, r.d.x, r.d.y, r.d.z);
                    m = quadrant_mask<Q>(r, tc.tx * TS, tc.ty * TS);
                }
                sM[j] = m;
                sS[j] = slot;
            }
        }
        wave_lds_sync();
        for (int j = cnt - 1; j >= 0; --j) {
            const uint32_t idx = uint32_t(lo + j);
            uint32_t m = __builtin_amdgcn_readfirstlane(sM[j]);
#pragma unroll
            for (int k = 0; k < Q; ++k)
                if (idx >= qmax[k]) m &= ~(1u << k);
            if (m == 0) continue;
            const float4 A = sA[j], B = sB[j], Cc = sC[j];
            const float cb = Cc.x;
            float r[NG];
#pragma unroll
            for (int g = 0; g < NG; ++g) r[g] = 0.0f;
            bool any = false;
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                if (!((m >> k) & 1u)) continue;
                const float dx = A.x - pxf[k], dy = A.y - pyf[k];
                float alpha, G;
                const bool contrib = pair_alpha_bwd(A.z, A.w, B.x, B.y, dx, dy, clip, alpha, G) & (idx < last[k]);
                if (__ballot(contrib) == 0) continue;
                any = true;
                // Branch-free over the wave: a lane without a contributing pair takes alpha = 0 and a
                // zero s (below), so its blend weight and every gradient term are exact zeros and E
                // keeps its value; its G (unspecified, possibly inf or NaN) is never used.
                // (The recovered T and the gradients are not compared bit for bit: the hardware
                // reciprocal / square root and fused multiply-adds are used throughout.)
                const float ac = contrib ? alpha : 0.0f;
                const float inv_1ma = __builtin_amdgcn_rcpf(1.0f - ac);
                const float Tb = T[k] * inv_1ma;
                const float wgt = ac * Tb;
                const float c_dot = fmaf(cb, gp2[k], fmaf(B.w, gp1[k], B.z * gp0[k]));
                const float dL_dalpha = fmaf(Tb, c_dot, -E[k] * inv_1ma);
                // s = G dL/dalpha, the pair's dL/dopacity; dL/dG = opacity * dL/dalpha, so
                // dL/dG * G = opacity * s
                const float s = contrib ? G * dL_dalpha : 0.0f;
                if constexpr (REFINE) {
                    const float wG = B.y * s;
                    // dL/dmean2d = dL/dG * dG/d(dx, dy), dG/ddx = -G (a dx + b dy), dG/ddy = -G (b dx + c dy)
                    const float ddx = -wG * fmaf(dx, Cc.y, dy * Cc.z);
                    const float ddy = -wG * fmaf(dy, Cc.w, dx * Cc.z);
                    r[0] += ddx;
                    r[1] += ddy;
                    // dL/dconic = -dL/dG * G * (dx^2, dx dy, dy^2) / 2
                    const float h = -0.5f * wG;
                    const float hx = h * dx;
                    r[2] = fmaf(hx, dx, r[2]);
                    r[3] = fmaf(hx, dy, r[3]);
                    r[4] = fmaf(h * dy, dy, r[4]);
                    // the refine statistic is a norm per pixel: it needs the per-pair dL/dmean2d
                    const float rx = ddx * fW, ry = ddy * fH;
                    r[9] = fmaf(__builtin_amdgcn_sqrtf(fmaf(rx, rx, ry * ry)), refine_den[k], r[9]);
                } else {
                    // Without it, both gradients are linear in the moments of the offsets weighted
                    // by s: sum s (dx, dy) and sum s (dx^2, dx dy, dy^2). The opacity, the conic and
                    // the -1/2 are applied once per (instance, tile) after the reduction.
                    const float t = s * dx, u = s * dy;
                    r[0] += t;
                    r[1] += u;
                    r[2] = fmaf(t, dx, r[2]);
                    r[3] = fmaf(t, dy, r[3]);
                    r[4] = fmaf(u, dy, r[4]);
                }
                r[5] += s;
                r[6] = fmaf(wgt, gp0[k], r[6]);
                r[7] = fmaf(wgt, gp1[k], r[7]);
                r[8] = fmaf(wgt, gp2[k], r[8]);
                E[k] = fmaf(c_dot, wgt, E[k]);
                T[k] = contrib ? Tb : T[k];
            }
            if (!any) continue;
            float tot = wave_reduce10_lds(r, lane, sR);
            if constexpr (!REFINE) {
                // moments -> gradients: dL/dmean2d = -opacity (conic . (sum s dx, sum s dy)),
                // dL/dconic = -opacity / 2 (sum s dx^2, sum s dx dy, sum s dy^2); the first-order
                // moments sit in lanes 0 and 4 (slot_of_lane)
                const float sx = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, tot), 0));
                const float sy = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, tot), 4));
                const float4 Cc = sC[j];
                const float op = sB[j].y;
                if (slot_g == 0) tot = -op * fmaf(Cc.y, sx, Cc.z * sy);
                else if (slot_g == 1) tot = -op * fmaf(Cc.w, sy, Cc.z * sx);
                else if (slot_g >= 2 && slot_g <= 4) tot = (-0.5f * op) * tot;
            }
            if (slot_g >= 0 && tot != 0.0f)
                atomic_add_f32(a.sgrad + grad_offset(sS[j], a.vp_log2, a.n_gauss) + slot_g, tot);
        }
    }
}

// Waves per SIMD the backward is compiled for. Without the refine statistic it fits 80 VGPRs
// (6 waves) without spilling. With a 32-deep LDS batch a wave's LDS (staging plus the 2.7 KB wave
// sum region, 4.5 KB) leaves room for 7 waves: there it is held to 72 VGPRs at the cost of a few
// spilled dwords, which measured faster (blend backward 2.83 -> 2.74 ms, bench step). With deeper
// batches the LDS allows no more than 6 anyway. With the refine statistic, 5.
constexpr int bwd_waves(int batch, bool refine) { return refine ? 5 : (batch <= 32 ? 7 : 6); }
template <int TS, int BATCH, int WPB, bool REFINE>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(bwd_waves(BATCH, REFINE)))) void k_blend_bwd(BlendArgs a,
                                                                                                      int n_items) {
    if (a.counters->overflow) return;
    __shared__ float4 sA[WPB][BATCH];
    __shared__ float4 sB[WPB][BATCH];
    __shared__ float4 sC[WPB][BATCH]; // (colour b, conic a, b, c): the gradient terms take the conic itself
    __shared__ uint32_t sM[WPB][BATCH];
    __shared__ uint32_t sS[WPB][BATCH];
    __shared__ __align__(16) float sR[WPB][10 * kRedPitch];
    const int w = int(__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)); // wave-uniform
    for_each_item<WPB>(a, a.queue ? a.queue + 1 : nullptr, n_items, [&](int item) {
        blend_bwd_item<TS, BATCH, REFINE>(a, item, n_items, sA[w], sB[w], sC[w], sM[w], sS[w], sR[w]);
    });
}

// Workgroups of `kernel` resident on the whole device at once (occupancy per CU x CUs): the grid of
// a persistent work-queue launch.
static int resident_blocks(const void* kernel, int threads) {
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
    return per_cu * cus;
}

template <class K> static unsigned grid_for(K kernel, const BlendArgs& a, int n_items, int wpb) {
    int blocks = (n_items + wpb - 1) / wpb;
    if (a.queue) blocks = std::min(blocks, resident_blocks(reinterpret_cast<const void*>(kernel), 64 * wpb));
    return unsigned(blocks);
}

template <int TS, int BATCH, int WPB> static void fwd_launch(const BlendArgs& a, int n_items, hipStream_t st) {
    const auto k = k_blend_fwd<TS, BATCH, WPB>;
    hipLaunchKernelGGL(k, dim3(grid_for(k, a, n_items, WPB)), dim3(64 * WPB), 0, st, a, n_items);
}
template <int TS, int BATCH> static void fwd_shape(const BlendArgs& a, int n_items, int waves, hipStream_t st) {
    if (waves >= 4) fwd_launch<TS, BATCH, 4>(a, n_items, st);
    else if (waves == 2) fwd_launch<TS, BATCH, 2>(a, n_items, st);
    else fwd_launch<TS, BATCH, 1>(a, n_items, st);
}
// ---------------------------------------------------------------- dispatch order of the tiles
// Blocks start in blockIdx order. A heavy tile that starts last runs as a lone wave after the rest
// of the grid has drained, so the blend launches take tiles longest first: a counting sort of the
// tiles by instance count (buckets of 128 instances, the last bucket open-ended), heaviest bucket
// first. Any order gives the same pixels; only the schedule changes.
constexpr int kOrderBuckets = 256;
__device__ __forceinline__ uint32_t order_bucket(const uint2* ranges, uint32_t t) {
    const uint32_t b = (ranges[t].y - ranges[t].x) >> 7;
    return uint32_t(kOrderBuckets - 1) - (b < uint32_t(kOrderBuckets - 1) ? b : uint32_t(kOrderBuckets - 1));
}
__global__ void k_order_clear(uint32_t* hist) {
    for (int i = threadIdx.x; i < 2 * kOrderBuckets; i += blockDim.x) hist[i] = 0;
}
__global__ __launch_bounds__(256) void k_order_hist(const uint2* ranges, uint32_t n_tiles, uint32_t* hist) {
    __shared__ uint32_t h[kOrderBuckets];
    h[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t t = blockIdx.x * 256 + threadIdx.x; t < n_tiles; t += gridDim.x * 256)
        atomicAdd(&h[order_bucket(ranges, t)], 1u);
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}
__global__ __launch_bounds__(256) void k_order_scan(uint32_t* hist) {
    __shared__ uint32_t wave_tot[4];
    const uint32_t c = hist[threadIdx.x];
    const uint32_t incl_w = wave_inclusive_scan(c);
    if (lane_id() == 63) wave_tot[threadIdx.x >> 6] = incl_w;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < int(threadIdx.x >> 6); ++w) before += wave_tot[w];
    hist[kOrderBuckets + threadIdx.x] = before + incl_w - c; // exclusive cursor of the bucket
}
// Ranks within a block come from LDS atomics; each (block, bucket) takes its global range with
// one atomic, so the few buckets that hold most tiles are not a serial hot spot.
__global__ __launch_bounds__(256) void k_order_scatter(const uint2* ranges, uint32_t n_tiles, uint32_t* hist,
                                                       uint32_t* order) {
    __shared__ uint32_t h[kOrderBuckets], base[kOrderBuckets];
    for (uint32_t t0 = blockIdx.x * 256; t0 < n_tiles; t0 += gridDim.x * 256) {
        h[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t t = t0 + threadIdx.x;
        const bool own = t < n_tiles;
        const uint32_t b = own ? order_bucket(ranges, t) : 0u;
        const uint32_t local = own ? atomicAdd(&h[b], 1u) : 0u;
        __syncthreads();
        const uint32_t c = h[threadIdx.x];
        if (c) base[threadIdx.x] = atomicAdd(&hist[kOrderBuckets + threadIdx.x], c);
        __syncthreads();
        if (own) order[base[b] + local] = t;
        __syncthreads();
    }
}

void launch_tile_order(const uint2* ranges, uint32_t n_tiles, uint32_t* order, uint32_t* hist, hipStream_t st) {
    const unsigned blocks = unsigned(std::min<uint32_t>((n_tiles + 255) / 256, 256u));
    k_order_clear<<<1, 256, 0, st>>>(hist);
    k_order_hist<<<blocks, 256, 0, st>>>(ranges, n_tiles, hist);
    k_order_scan<<<1, 256, 0, st>>>(hist);
    k_order_scatter<<<blocks, 256, 0, st>>>(ranges, n_tiles, hist, order);
}

template <int TS, int BATCH, int WPB, bool R> static void bwd_launch(const BlendArgs& a, int n_items, hipStream_t st) {
    const auto k = k_blend_bwd<TS, BATCH, WPB, R>;
    hipLaunchKernelGGL(k, dim3(grid_for(k, a, n_items, WPB)), dim3(64 * WPB), 0, st, a, n_items);
}
template <int TS, int BATCH> static void bwd_shape(const BlendArgs& a, int n_items, int waves, hipStream_t st) {
    if (a.want_refine) {
        if (waves >= 4) bwd_launch<TS, BATCH, 4, true>(a, n_items, st);
        else if (waves == 2) bwd_launch<TS, BATCH, 2, true>(a, n_items, st);
        else bwd_launch<TS, BATCH, 1, true>(a, n_items, st);
    } else {
        if (waves >= 4) bwd_launch<TS, BATCH, 4, false>(a, n_items, st);
        else if (waves == 2) bwd_launch<TS, BATCH, 2, false>(a, n_items, st);
        else bwd_launch<TS, BATCH, 1, false>(a, n_items, st);
    }
}

__global__ void k_queue_reset(uint32_t* q) { *q = 0u; }

void launch_blend_forward(const BlendArgs& a, int n_tiles, int tile_size, int batch, int waves, hipStream_t st) {
    if (n_tiles <= 0) return;
    if (tile_size == 8) {
        if (batch <= 32) fwd_shape<8, 32>(a, n_tiles, waves, st);
        else if (batch <= 64) fwd_shape<8, 64>(a, n_tiles, waves, st);
        else fwd_shape<8, 128>(a, n_tiles, waves, st);
    } else {
        if (batch <= 32) fwd_shape<16, 32>(a, n_tiles, waves, st);
        else if (batch <= 64) fwd_shape<16, 64>(a, n_tiles, waves, st);
        else if (batch <= 128) fwd_shape<16, 128>(a, n_tiles, waves, st);
        else fwd_shape<16, 256>(a, n_tiles, waves, st);
    }
}

void launch_blend_backward(const BlendArgs& a, int n_tiles, int tile_size, int batch, int waves, hipStream_t st) {
    if (n_tiles <= 0) return;
    // the backward's queue counter starts from zero at every backward (a forward may be
    // differentiated more than once)
    if (a.queue) hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(1), 0, st, a.queue + 1);
    if (tile_size == 8) {
        if (batch <= 32) bwd_shape<8, 32>(a, n_tiles, waves, st);
        else if (batch <= 64) bwd_shape<8, 64>(a, n_tiles, waves, st);
        else bwd_shape<8, 128>(a, n_tiles, waves, st);
    } else {
        if (batch <= 32) bwd_shape<16, 32>(a, n_tiles, waves, st);
        else if (batch <= 64) bwd_shape<16, 64>(a, n_tiles, waves, st);
        else if (batch <= 128) bwd_shape<16, 128>(a, n_tiles, waves, st);
        else bwd_shape<16, 256>(a, n_tiles, waves, st);
    }
}

} // namespace pht
