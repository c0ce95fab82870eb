// Per-tile alpha blending, forward and backward, for gfx950 (wave64, 160 KiB LDS per CU).
//
// One workgroup per (view, tile) work item of the batched launch. A 16x16 tile is 4 waves, each
// wave owning an 8x8 pixel patch (an 8x8 tile is one wave). Instances of the tile's sorted list
// are staged through LDS in batches of BATCH; each wave tests a per-instance patch bit computed
// once by the loader (the alpha >= 1/255 ellipse's box, grown by a safety margin) and skips
// instances that cannot reach its pixels, without changing any blend decision.
//
// Forward  : blend_tile (third_party/splat_drender/src/render_forward.cu:274-538), colour path.
// Backward : the closed-form per-pair terms of blend_bucket_backward (render_backward.cu:383-489),
//            walked back to front per pixel from the stored final transmittance and contributor
//            count (T_before = T_after / (1 - alpha)), with no per-bucket snapshots. The per-pair
//            gradients are summed over the wave with DPP, combined over the tile's waves with
//            LDS atomics, and committed to the splat's gradient with one global atomic per
//            value per (instance, tile).
#include "launch.h"

namespace pht {

// Bit w set when the instance can reach the 8x8 patch of wave w of a 16x16 tile whose
// top-left pixel is (x0, y0). Conservative: the exact per-pixel test still decides.
__device__ __forceinline__ uint32_t patch_mask16(float4 A, float4 B, int x0, int y0) {
    const float a = A.z, b = A.w, c = B.x, op = B.y;
    if (!(op >= K<float>::alpha_floor())) return 0u;
    const float det = a * c - b * b;
    const float thr = 2.0f * plogf(op * 255.0f);
    const float tr = a + c;
    if (!(det > 0.0f) || !(a > 0.0f) || !(c > 0.0f) || !(thr >= 0.0f) || !(tr * tr < 1.6e5f * det))
        return 0xfu;
    const float ex = sqrtf(thr * c / det) * 1.01f + 1.0f;
    const float ey = sqrtf(thr * a / det) * 1.01f + 1.0f;
    if (!(ex < 1.0e6f) || !(ey < 1.0e6f)) return 0xfu;
    uint32_t m = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const float px0 = float(x0 + (w & 1) * 8), py0 = float(y0 + (w >> 1) * 8);
        if (A.x + ex >= px0 && A.x - ex <= px0 + 7.0f && A.y + ey >= py0 && A.y - ey <= py0 + 7.0f)
            m |= 1u << w;
    }
    return m;
}

template <int TS>
__device__ __forceinline__ void pixel_of(int tx, int ty, int w, int lane, int& px, int& py) {
    if constexpr (TS == 16) {
        px = tx * 16 + (w & 1) * 8 + (lane & 7);
        py = ty * 16 + (w >> 1) * 8 + (lane >> 3);
    } else {
        px = tx * 8 + (lane & 7);
        py = ty * 8 + (lane >> 3);
    }
}

template <int TS, int BATCH>
__global__ __launch_bounds__(TS* TS) void k_blend_fwd(BlendArgs a) {
    constexpr int NT = TS * TS;
    const int tile = blockIdx.x;
    int v;
    const DevView& dv = view_of_tile(a.views, a.n_views, tile, v);
    const int local = tile - dv.vi.tile_base;
    const int gx = dv.vi.grid_x;
    const int tx = local % gx, ty = local / gx;
    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    int px, py;
    pixel_of<TS>(tx, ty, w, lane, px, py);
    const int W = dv.vi.cam.width, H = dv.vi.cam.height;
    const bool inside = px < W && py < H;
    const float pxf = float(px), pyf = float(py);
    if (a.counters->overflow) return; // incomplete instance list: the host re-runs with more capacity
    const uint2 range = a.ranges[tile];

    __shared__ float4 sA[BATCH];
    __shared__ float4 sB[BATCH];
    __shared__ float sC[BATCH];
    __shared__ uint32_t sM[BATCH];
    __shared__ uint32_t sS[BATCH];
    __shared__ uint32_t sV[BATCH];

    float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
    uint32_t last = 0;
    bool done = !inside;
    const float clip = a.alpha_clip;
    const float t_floor = K<float>::transmittance_floor();
    for (uint32_t base = range.x; base < range.y; base += BATCH) {
        const uint32_t cnt = min(uint32_t(BATCH), range.y - base);
        for (uint32_t j = tid; j < cnt; j += NT) {
            const uint32_t slot = a.inst_vals[base + j];
            sV[j] = 0u;
            sS[j] = 0u;
            sM[j] = 0u;
            if (slot < a.slot_cap) {
                const float4 A = a.rec_a[slot], B = a.rec_b[slot];
                sA[j] = A;
                sB[j] = B;
                sC[j] = __uint_as_float(a.rec_c[slot].x);
                sS[j] = slot;
                sM[j] = TS == 16 ? patch_mask16(A, B, tx * 16, ty * 16) : 1u;
            }
        }
        __syncthreads();
        if (__ballot(!done) != 0) {
            for (uint32_t j = 0; j < cnt; ++j) {
                const uint32_t m = __builtin_amdgcn_readfirstlane(sM[j]);
                if (!((m >> w) & 1u)) continue;
                bool blended = false;
                if (!done) {
                    const float4 A = sA[j], B = sB[j];
                    float alpha, G;
                    if (pair_alpha(A.z, A.w, B.x, B.y, A.x - pxf, A.y - pyf, clip, alpha, G)) {
                        const float test_t = T * (1.0f - alpha);
                        if (test_t < t_floor) {
                            done = true;
                        } else {
                            const float a_t = alpha * T;
                            C0 = fmaf(B.z, a_t, C0);
                            C1 = fmaf(B.w, a_t, C1);
                            C2 = fmaf(sC[j], a_t, C2);
                            T = test_t;
                            last = base - range.x + j + 1;
                            blended = true;
                        }
                    }
                }
                if (__ballot(blended) != 0 && lane == 0) sV[j] = 1u;
            }
        }
        const int all_done = __syncthreads_and(done ? 1 : 0);
        if (dv.visibility)
            for (uint32_t j = tid; j < cnt; j += NT)
                if (sV[j]) dv.visibility[sS[j] >> a.vp_log2] = 1.0f;
        if (all_done) break;
    }
    if (inside) {
        const int64_t P = int64_t(W) * H;
        const int64_t pid = int64_t(py) * W + px;
        dv.color[pid] = C0 + T * dv.vi.cam.background[0];
        dv.color[P + pid] = C1 + T * dv.vi.cam.background[1];
        dv.color[2 * P + pid] = C2 + T * dv.vi.cam.background[2];
        dv.alpha[pid] = 1.0f - T;
        a.final_T[dv.vi.pixel_base + pid] = T;
        a.n_contrib[dv.vi.pixel_base + pid] = last;
    }
}

template <int TS, int BATCH>
__global__ __launch_bounds__(TS* TS) void k_blend_bwd(BlendArgs a) {
    constexpr int NT = TS * TS;
    constexpr int NW = NT / 64;
    constexpr int NG = 10;
    const int tile = blockIdx.x;
    int v;
    const DevView& dv = view_of_tile(a.views, a.n_views, tile, v);
    const int local = tile - dv.vi.tile_base;
    const int gx = dv.vi.grid_x;
    const int tx = local % gx, ty = local / gx;
    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    int px, py;
    pixel_of<TS>(tx, ty, w, lane, px, py);
    const int W = dv.vi.cam.width, H = dv.vi.cam.height;
    const bool inside = px < W && py < H;
    const float pxf = float(px), pyf = float(py);
    if (a.counters->overflow) return;
    const uint2 range = a.ranges[tile];

    __shared__ float4 sA[BATCH];
    __shared__ float4 sB[BATCH];
    __shared__ float sC[BATCH];
    __shared__ uint32_t sM[BATCH];
    __shared__ uint32_t sS[BATCH];
    __shared__ float sAcc[BATCH * NG];
    __shared__ uint32_t sMax[NW];

    uint32_t last = 0;
    float Tfin = 1.0f, gp0 = 0.0f, gp1 = 0.0f, gp2 = 0.0f, dL_dTfin = 0.0f, refine_den = 0.0f;
    if (inside) {
        const int64_t P = int64_t(W) * H;
        const int64_t pid = int64_t(py) * W + px;
        last = a.n_contrib[dv.vi.pixel_base + pid];
        Tfin = a.final_T[dv.vi.pixel_base + pid];
        gp0 = dv.grad_color[pid];
        gp1 = dv.grad_color[P + pid];
        gp2 = dv.grad_color[2 * P + pid];
        const float ga = dv.grad_alpha ? dv.grad_alpha[pid] : 0.0f;
        const float* bg = dv.vi.cam.background;
        dL_dTfin = ((bg[0] * gp0 + bg[1] * gp1) + bg[2] * gp2) - ga;
        refine_den = 1.0f / fmaxf(1.0f - Tfin, 1.0e-5f);
    }
    const uint32_t wmax = wave_max_u32(last);
    if (lane == 0) sMax[w] = wmax;
    __syncthreads();
    uint32_t tile_max = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) tile_max = sMax[k] > tile_max ? sMax[k] : tile_max;

    const float clip = a.alpha_clip;
    const float fW = float(W), fH = float(H);
    float T = Tfin, Ca0 = 0.0f, Ca1 = 0.0f, Ca2 = 0.0f;
    for (int hi = int(tile_max); hi > 0; hi -= BATCH) {
        const int lo = hi - BATCH > 0 ? hi - BATCH : 0;
        const int cnt = hi - lo;
        for (int j = tid; j < cnt; j += NT) {
            const uint32_t slot = a.inst_vals[range.x + lo + j];
            sS[j] = 0u;
            sM[j] = 0u;
            if (slot < a.slot_cap) {
                const float4 A = a.rec_a[slot], B = a.rec_b[slot];
                sA[j] = A;
                sB[j] = B;
                sC[j] = __uint_as_float(a.rec_c[slot].x);
                sS[j] = slot;
                sM[j] = TS == 16 ? patch_mask16(A, B, tx * 16, ty * 16) : 1u;
            }
#pragma unroll
            for (int k = 0; k < NG; ++k) sAcc[j * NG + k] = 0.0f;
        }
        __syncthreads();
        for (int j = cnt - 1; j >= 0; --j) {
            const uint32_t m = __builtin_amdgcn_readfirstlane(sM[j]);
            if (!((m >> w) & 1u)) continue;
            const float4 A = sA[j], B = sB[j];
            const float dx = A.x - pxf, dy = A.y - pyf;
            float alpha = 0.0f, G = 0.0f;
            const bool contrib = uint32_t(lo + j) < last && pair_alpha(A.z, A.w, B.x, B.y, dx, dy, clip, alpha, G);
            if (__ballot(contrib) == 0) continue;
            float r[NG];
#pragma unroll
            for (int k = 0; k < NG; ++k) r[k] = 0.0f;
            if (contrib) {
                const float cb = sC[j];
                const float one_m = 1.0f - alpha;
                const float Tb = T / one_m;
                const float inv_1ma = 1.0f / one_m;
                const float wgt = alpha * Tb;
                const float c_dot = (B.z * gp0 + B.w * gp1) + cb * gp2;
                const float a_dot = ((Ca0 * gp0 + Ca1 * gp1) + Ca2 * gp2) * inv_1ma;
                const float dL_dalpha = (Tb * c_dot - a_dot) - Tfin * inv_1ma * dL_dTfin;
                const float dL_dG = B.y * dL_dalpha;
                const float gdx = G * dx, gdy = G * dy;
                const float dG_dx = -gdx * A.z - gdy * A.w;
                const float dG_dy = -gdy * B.x - gdx * A.w;
                const float ddx = dL_dG * dG_dx, ddy = dL_dG * dG_dy;
                r[0] = ddx;
                r[1] = ddy;
                r[2] = -0.5f * gdx * dx * dL_dG;
                r[3] = -0.5f * gdx * dy * dL_dG;
                r[4] = -0.5f * gdy * dy * dL_dG;
                r[5] = G * dL_dalpha;
                r[6] = wgt * gp0;
                r[7] = wgt * gp1;
                r[8] = wgt * gp2;
                const float rx = ddx * fW, ry = ddy * fH;
                r[9] = sqrtf(rx * rx + ry * ry) * refine_den;
                Ca0 = fmaf(B.z, wgt, Ca0);
                Ca1 = fmaf(B.w, wgt, Ca1);
                Ca2 = fmaf(cb, wgt, Ca2);
                T = Tb;
            }
#pragma unroll
            for (int k = 0; k < NG; ++k) r[k] = wave_sum_to_last(r[k]);
            if (lane == 63) {
#pragma unroll
                for (int k = 0; k < NG; ++k) atomicAdd(&sAcc[j * NG + k], r[k]);
            }
        }
        __syncthreads();
        for (int j = tid; j < cnt; j += NT) {
            if (sM[j] == 0u) continue; // no wave reached this instance: nothing accumulated
            float* dst = a.sgrad + size_t(sS[j]) * kGradStride;
#pragma unroll
            for (int k = 0; k < NG; ++k) {
                const float x = sAcc[j * NG + k];
                if (x != 0.0f) atomic_add_f32(dst + k, x);
            }
        }
    }
}

template <int TS, int BATCH> static void fwd_launch(const BlendArgs& a, int n_tiles, hipStream_t st) {
    hipLaunchKernelGGL((k_blend_fwd<TS, BATCH>), dim3(n_tiles), dim3(TS * TS), 0, st, a);
}
template <int TS, int BATCH> static void bwd_launch(const BlendArgs& a, int n_tiles, hipStream_t st) {
    hipLaunchKernelGGL((k_blend_bwd<TS, BATCH>), dim3(n_tiles), dim3(TS * TS), 0, st, a);
}

void launch_blend_forward(const BlendArgs& a, int n_tiles, int tile_size, int batch, hipStream_t st) {
    if (n_tiles <= 0) return;
    if (tile_size == 8) {
        if (batch <= 32) fwd_launch<8, 32>(a, n_tiles, st);
        else if (batch <= 64) fwd_launch<8, 64>(a, n_tiles, st);
        else fwd_launch<8, 128>(a, n_tiles, st);
    } else {
        if (batch <= 32) fwd_launch<16, 32>(a, n_tiles, st);
        else if (batch <= 64) fwd_launch<16, 64>(a, n_tiles, st);
        else if (batch <= 128) fwd_launch<16, 128>(a, n_tiles, st);
        else fwd_launch<16, 256>(a, n_tiles, st);
    }
}

void launch_blend_backward(const BlendArgs& a, int n_tiles, int tile_size, int batch, hipStream_t st) {
    if (n_tiles <= 0) return;
    if (tile_size == 8) {
        if (batch <= 32) bwd_launch<8, 32>(a, n_tiles, st);
        else if (batch <= 64) bwd_launch<8, 64>(a, n_tiles, st);
        else bwd_launch<8, 128>(a, n_tiles, st);
    } else {
        if (batch <= 32) bwd_launch<16, 32>(a, n_tiles, st);
        else if (batch <= 64) bwd_launch<16, 64>(a, n_tiles, st);
        else if (batch <= 128) bwd_launch<16, 128>(a, n_tiles, st);
        else bwd_launch<16, 256>(a, n_tiles, st);
    }
}

} // namespace pht
