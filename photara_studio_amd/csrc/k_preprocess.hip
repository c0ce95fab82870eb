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

__global__ __launch_bounds__(256) void k_clear(Counters* counters, uint2* ranges, uint32_t n_tiles,
                                               uint32_t* key_parts) {
    const uint32_t i0 = blockIdx.x * 256u + threadIdx.x;
    if (i0 < sizeof(Counters) / 4) reinterpret_cast<uint32_t*>(counters)[i0] = 0u;
    for (uint32_t i = i0; i < n_tiles; i += gridDim.x * 256u) ranges[i] = make_uint2(0u, 0u);
    for (uint32_t i = i0; i < 2u * kKeyParts; i += gridDim.x * 256u) key_parts[i] = 0u;
}

// Folds the partial key ranges into the counters (one block of kKeyParts threads).
__global__ __launch_bounds__(kKeyParts) void k_key_range(const uint32_t* key_parts, Counters* counters) {
    __shared__ uint32_t s[2][kKeyParts / 64];
    const int t = threadIdx.x, lane = lane_id(), w = t >> 6;
    const uint32_t a = wave_max_u32(key_parts[2 * t]), b = wave_max_u32(key_parts[2 * t + 1]);
    if (lane == 0) {
        s[0][w] = a;
        s[1][w] = b;
    }
    __syncthreads();
    if (t == 0) {
        uint32_t x = 0, y = 0;
        for (int k = 0; k < kKeyParts / 64; ++k) {
            x = s[0][k] > x ? s[0][k] : x;
            y = s[1][k] > y ? s[1][k] : y;
        }
        counters->key_nmin = x;
        counters->key_max = y;
    }
}

// Half extents of the box around the alpha >= 1/255 ellipse {q(d) <= 2 log(255 op)} of the
// conic (a, b, c), grown by 1% + 1 px, and the ellipse's bound on q grown by 2% + 0.05 (qmax):
// +inf when the conic is degenerate or too anisotropic for a safe box (no culling), -inf when the
// opacity is below 1/255 (reaches no pixel).
__device__ __forceinline__ void patch_extent(float a, float b, float c, float op, float& ex, float& ey, float& qmax) {
    const float inf = __builtin_huge_valf();
    if (!(op >= K<float>::alpha_floor())) {
        ex = ey = qmax = -inf;
        return;
    }
    const float det = a * c - b * b;
    const float thr = 2.0f * plogf(op * 255.0f);
    const float tr = a + c;
    ex = ey = qmax = inf;
    if (!(det > 0.0f) || !(a > 0.0f) || !(c > 0.0f) || !(thr >= 0.0f) || !(tr * tr < 1.6e5f * det)) return;
    const float x = sqrtf(thr * c / det) * 1.01f + 1.0f;
    const float y = sqrtf(thr * a / det) * 1.01f + 1.0f;
    if (!(x < 1.0e6f) || !(y < 1.0e6f)) return;
    ex = x;
    ey = y;
    qmax = thr * 1.02f + 0.05f;
}

template <int LP> __device__ __forceinline__ float load_lp(const float* f, const uint16_t* lp, int64_t i) {
    if constexpr (LP == PHT_LP_NONE) return f[i];
    else return lp_to_f32(lp[i], LP);
}

// The first NC coefficients of one SH row (row_elems stored elements). When the array is 16-byte
// aligned and a row is a whole number of 16-byte chunks, the row is read as uint4 chunks (NC
// rounded up to a chunk never leaves the row); otherwise element by element.
template <int LP, int NC>
__device__ __forceinline__ void load_sh_row(const float* f, const uint16_t* lp, int64_t row, int row_elems,
                                            float (&coef)[NC]) {
    constexpr int ES = LP == PHT_LP_NONE ? 4 : 2;   // bytes per stored element
    constexpr int PER = 16 / ES;                    // elements per 16-byte chunk
    constexpr int CHUNKS = (NC + PER - 1) / PER;
    const uintptr_t base = LP == PHT_LP_NONE ? reinterpret_cast<uintptr_t>(f) : reinterpret_cast<uintptr_t>(lp);
    if ((row_elems * ES) % 16 == 0 && (base & 15u) == 0) {
        const uint4* src = LP == PHT_LP_NONE ? reinterpret_cast<const uint4*>(f + row)
                                             : reinterpret_cast<const uint4*>(lp + row);
#pragma unroll
        for (int c = 0; c < CHUNKS; ++c) {
            const uint4 q = src[c];
            const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int e = 0; e < PER; ++e) {
                const int k = c * PER + e;
                if (k < NC) {
                    if constexpr (LP == PHT_LP_NONE) coef[k] = u2f(w[e]);
                    else coef[k] = lp_to_f32(uint16_t(w[e >> 1] >> ((e & 1) * 16)), LP);
                }
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < NC; ++k) coef[k] = load_lp<LP>(f, lp, row + k);
    }
}

// sh_eval on 16-bit stored coefficients kept packed two per 32-bit word (w[e / 2], element e in
// the low half when e is even): each is widened where sh_eval reads it, with the same conversion
// and the same per-channel fused multiply-add order, so the colours are sh_eval's bit for bit;
// the row takes half the registers of the widened coefficients.
template <int LP, int NC>
__device__ __forceinline__ void load_sh_words(const uint16_t* lp, int64_t row, int row_elems, uint32_t (&w)[(NC + 1) / 2]) {
    constexpr int NW = (NC + 1) / 2;
    if ((row_elems * 2) % 16 == 0 && (reinterpret_cast<uintptr_t>(lp) & 15u) == 0) {
        const uint4* src = reinterpret_cast<const uint4*>(lp + row);
#pragma unroll
        for (int c = 0; c < (NW + 3) / 4; ++c) {
            const uint4 q = src[c];
            const uint32_t v[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (4 * c + e < NW) w[4 * c + e] = v[e];
        }
    } else {
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            const uint32_t lo = lp[row + 2 * k];
            const uint32_t hi = 2 * k + 1 < NC ? uint32_t(lp[row + 2 * k + 1]) : 0u;
            w[k] = lo | (hi << 16);
        }
    }
}
template <int LP, int NC>
__device__ __forceinline__ float sh_word(const uint32_t (&w)[(NC + 1) / 2], int e) {
    return lp_to_f32(uint16_t((e & 1) ? (w[e >> 1] >> 16) : (w[e >> 1] & 0xffffu)), LP);
}
template <int LP, int DEG>
__device__ __forceinline__ void sh_eval_words(const float d[3], const uint32_t (&w)[((DEG + 1) * (DEG + 1) * 3 + 1) / 2],
                                              float rgb[3], uint32_t& clamp_bits) {
    constexpr int NB = (DEG + 1) * (DEG + 1), NC = NB * 3;
    float Y[16];
    sh_basis(DEG, d, Y);
    clamp_bits = 0;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        float acc = Y[0] * sh_word<LP, NC>(w, ch);
#pragma unroll
        for (int k = 1; k < NB; ++k) acc = fma_(Y[k], sh_word<LP, NC>(w, k * 3 + ch), acc);
        acc = acc + 0.5f;
        if (acc < 0.0f) clamp_bits |= 1u << ch;
        rgb[ch] = max_(acc, 0.0f);
    }
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
    uint32_t clamp_flags = 0, dkey = 0;
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
                    const int64_t row = gi * int64_t(a.s.sh_bases) * 3;
                    float d[3], raw[3];
                    view_dir(mean, dv.vi.cam.position, d, raw);
                    if constexpr (LP == PHT_LP_NONE) {
                        float coef[NB * 3];
                        load_sh_row<LP, NB * 3>(a.g.sh, a.g.sh_lp, row, a.s.sh_bases * 3, coef);
                        sh_eval(DEG, d, coef, rgb, clamp);
                    } else {
                        uint32_t w[(NB * 3 + 1) / 2];
                        load_sh_words<LP, NB * 3>(a.g.sh_lp, row, a.s.sh_bases * 3, w);
                        sh_eval_words<LP, DEG>(d, w, rgb, clamp);
                    }
                }
                float ex, ey, qmax;
                patch_extent(p.conic[0], p.conic[1], p.conic[2], p.opacity, ex, ey, qmax);
                float a2, b2, c2;
                conic_log2(p.conic[0], p.conic[1], p.conic[2], a2, b2, c2);
                SplatRec& r = a.rec[t];
                r.a = make_float4(p.mean2d[0], p.mean2d[1], a2, b2);
                r.b = make_float4(c2, p.opacity, rgb[0], rgb[1]);
                r.c = make_float4(rgb[2], ex, ey, qmax);
                r.d = make_float4(p.conic[0], p.conic[1], p.conic[2], p.opacity);
                dkey = f2u(p.depth);
                a.meta[t] = make_uint2(dkey, touched | (clamp << kMetaClampShift));
                clamp_flags = clamp;
            }
        }
        if (dv.radii) dv.radii[gi] = visible ? radius : 0;
        if (dv.visibility) dv.visibility[gi] = 0.0f;
    }
    if (in_range) a.vis_flag[t] = visible ? 1u : 0u;
    const uint64_t bal = __ballot(visible);
    const int lane = lane_id();
    // with at most 8 view slots the SH clamp flags of every view ride in the same word
    uint64_t cb[3] = {0, 0, 0};
    if (vp <= kMaskClampViews) {
#pragma unroll
        for (int b = 0; b < 3; ++b) cb[b] = __ballot(visible && ((clamp_flags >> b) & 1u));
    }
    if (in_range && v == 0) {
        const int base = lane & ~(vp - 1);
        const uint64_t sel = vp >= 64 ? ~uint64_t(0) : ((uint64_t(1) << vp) - 1u);
        uint32_t word = uint32_t((bal >> base) & sel);
        if (vp <= kMaskClampViews) {
#pragma unroll
            for (int b = 0; b < 3; ++b) word |= uint32_t((cb[b] >> base) & sel) << (8 * (b + 1));
        }
        a.view_mask[gi] = word;
    }
    const uint64_t fb = __ballot(fp64);
    if (lane == 0 && fb) atomicAdd(&a.counters->fp64, uint32_t(__popcll(fb)));
    // the range of the visible depth keys: the depth sort ranks keys relative to the least, and
    // skips its top pass when the range fits 24 bits
    if (bal != 0) {
        const uint32_t kmax = wave_max_u32(visible ? dkey : 0u), knmin = wave_max_u32(visible ? ~dkey : 0u);
        if (lane == 0) {
            uint32_t* part = a.key_parts + 2 * ((blockIdx.x * 4u + (threadIdx.x >> 6)) % uint32_t(kKeyParts));
            atomicMax(part, knmin);
            atomicMax(part + 1, kmax);
        }
    }
}

// Emits (global tile, slot) pairs for the depth-sorted visible splats at their scanned offsets
// (emit_instances, render_forward.cu:167-181; tile keys carry the view's tile_base).
//
// Each wave takes 64 consecutive sorted splats, whose outputs form one contiguous run. Every lane
// walks its splat's columns (AccuTile) and records each column's tile interval in LDS with its
// position in the run; the wave then writes the run in order, 64 consecutive outputs at a time,
// each lane locating its output's column by binary search. Writes are coalesced and the work per
// lane no longer follows one splat's tile count. A group whose columns do not fit the LDS table,
// or that would overflow the instance capacity, falls back to per-lane emission.
constexpr int kEmitCols = 512; // column records per wave

template <class K>
__global__ __launch_bounds__(256) void k_emit(const uint32_t* sorted_slots, const uint32_t* sorted_tiles,
                                              const uint32_t* offsets, const Counters* counters, uint32_t slot_cap,
                                              const SplatRec* rec, const DevView* views, int vp_log2, int ts,
                                              uint32_t inst_cap, K* keys_out, uint32_t* vals_out) {
    __shared__ uint2 s_col[4][kEmitCols]; // (run position of the column, lane << 26 | u << 13 | v_min)
    __shared__ uint32_t s_slot[4][64];
    __shared__ uint32_t s_tb[4][64];
    __shared__ uint32_t s_gx[4][64];      // grid_x | iterate_x << 31
    const uint32_t nv = counters->visible;
    const uint32_t n = nv < slot_cap ? nv : slot_cap;
    const uint32_t vmask = (1u << vp_log2) - 1u;
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const uint32_t groups = (n + 63u) >> 6;
    for (uint32_t grp = blockIdx.x * 4u + uint32_t(w); grp < groups; grp += gridDim.x * 4u) {
        const uint32_t i = grp * 64u + uint32_t(lane);
        const bool valid = i < n;
        uint32_t slot = 0, off = 0, count = 0;
        if (valid) {
            slot = sorted_slots[i];
            off = offsets[i];
            count = sorted_tiles[i];
        }
        const bool over = valid && uint64_t(off) + count > inst_cap;
        TileWalk<float> tw;
        tw.ok = false;
        uint32_t tb = 0, gx = 0;
        if (valid && !over) {
            const float4 A = rec[slot].a, D = rec[slot].d;
            const ViewInfo& vi = views[slot & vmask].vi;
            const float conic[3] = {D.x, D.y, D.z};
            tw = tile_walk_setup(A.x, A.y, conic, D.w, vi.grid_x, vi.grid_y, ts);
            tb = uint32_t(vi.tile_base);
            gx = uint32_t(vi.grid_x);
        }
        const uint32_t ncols = tw.ok ? uint32_t(tw.u_max - tw.u_min) : 0u;
        const uint32_t col_incl = wave_inclusive_scan(ncols);
        const uint32_t total_cols = uint32_t(__builtin_amdgcn_readlane(int(col_incl), 63));
        const uint32_t cnt_incl = wave_inclusive_scan(valid && !over ? count : 0u);
        const uint32_t total = uint32_t(__builtin_amdgcn_readlane(int(cnt_incl), 63));
        const uint32_t run_base = uint32_t(__builtin_amdgcn_readfirstlane(int(off)));
        if (__ballot(over) != 0 || total_cols > uint32_t(kEmitCols)) {
            if (tw.ok) {
                uint32_t k = 0;
#pragma unroll 1
                for (int tu = tw.u_min; tu < tw.u_max; ++tu) {
                    int v0, v1;
                    tile_walk_column(tw, tu, v0, v1);
#pragma unroll 1
                    for (int tv = v0; tv < v1; ++tv) {
                        if (k < count) {
                            const int tx = tw.iterate_x ? tu : tv, ty = tw.iterate_x ? tv : tu;
                            keys_out[off + k] = K(tb + uint32_t(ty) * gx + uint32_t(tx));
                            vals_out[off + k] = slot;
                        }
                        ++k;
                    }
                }
            }
            continue;
        }
        // column table: this lane's columns start at col_incl - ncols, its outputs at cnt_incl - count
        {
            uint32_t r = col_incl - ncols;
            const uint32_t pos0 = cnt_incl - count;
            uint32_t done = 0;
#pragma unroll 1
            for (int tu = tw.u_min; tw.ok && tu < tw.u_max; ++tu) {
                int v0, v1;
                tile_walk_column(tw, tu, v0, v1);
                uint32_t len = v1 > v0 ? uint32_t(v1 - v0) : 0u;
                if (done + len > count) len = count - done;
                s_col[w][r++] = make_uint2(pos0 + done,
                                           (uint32_t(lane) << 26) | (uint32_t(tu) << 13) | uint32_t(v0 > 0 ? v0 : 0));
                done += len;
            }
            s_slot[w][lane] = slot;
            s_tb[w][lane] = tb;
            s_gx[w][lane] = gx | (tw.iterate_x ? 0x80000000u : 0u);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 1
        for (uint32_t o = uint32_t(lane); o < total; o += 64u) {
            int lo = 0, hi = int(total_cols) - 1;
#pragma unroll 1
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_col[w][mid].x <= o) lo = mid;
                else hi = mid - 1;
            }
            const uint2 c = s_col[w][lo];
            const uint32_t L = c.y >> 26;
            const uint32_t u = (c.y >> 13) & 0x1fffu;
            const uint32_t v = (c.y & 0x1fffu) + (o - c.x);
            const uint32_t g = s_gx[w][L];
            const bool it = (g >> 31) != 0;
            const uint32_t tx = it ? u : v, ty = it ? v : u;
            keys_out[run_base + o] = K(s_tb[w][L] + ty * (g & 0x7fffffffu) + tx);
            vals_out[run_base + o] = s_slot[w][L];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Four consecutive keys as one vector load (16 bytes of 32-bit keys, 8 bytes of 16-bit keys).
// kRangeKeys consecutive keys per thread, loaded as 16-byte vectors (two for 16-bit keys, four for
// 32-bit): with one key vector per thread the walk had too few bytes in flight per CU to stream at
// HBM rate.
constexpr int kRangeKeys = 16;
__device__ __forceinline__ void load_keys(const uint32_t* keys, uint32_t g, uint32_t out[kRangeKeys]) {
    const uint4* p = reinterpret_cast<const uint4*>(keys) + uint64_t(g) * (kRangeKeys / 4);
#pragma unroll
    for (int v = 0; v < kRangeKeys / 4; ++v) {
        const uint4 x = p[v];
        out[4 * v] = x.x; out[4 * v + 1] = x.y; out[4 * v + 2] = x.z; out[4 * v + 3] = x.w;
    }
}
__device__ __forceinline__ void load_keys(const uint16_t* keys, uint32_t g, uint32_t out[kRangeKeys]) {
    const uint4* p = reinterpret_cast<const uint4*>(keys) + uint64_t(g) * (kRangeKeys / 8);
#pragma unroll
    for (int v = 0; v < kRangeKeys / 8; ++v) {
        const uint4 x = p[v];
        const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            out[8 * v + 2 * e] = w[e] & 0xffffu;
            out[8 * v + 2 * e + 1] = w[e] >> 16;
        }
    }
}

// Per-tile [start, end) of the tile-sorted list (extract_tile_ranges, render_forward.cu:215-232).
// Each thread takes kRangeKeys consecutive keys and their two outer neighbours.
template <class K>
__global__ __launch_bounds__(256) void k_ranges(const K* keys, const Counters* counters, uint32_t inst_cap,
                                                uint32_t n_tiles, uint2* ranges) {
    if (counters->overflow) return; // the list is incomplete: leave every range empty
    constexpr uint32_t E = kRangeKeys;
    const uint32_t ni = counters->instances;
    const uint32_t n = ni < inst_cap ? ni : inst_cap;
    const uint32_t groups = (n + E - 1u) / E;
    for (uint32_t g = blockIdx.x * 256u + threadIdx.x; g < groups; g += gridDim.x * 256u) {
        const uint32_t i0 = g * E;
        uint32_t k[E + 2]; // k[0] = keys[i0 - 1], k[1..E] = keys[i0 .. i0 + E - 1], k[E + 1] = keys[i0 + E]
        if (i0 + E - 1u < n) {
            load_keys(keys, g, k + 1);
        } else {
#pragma unroll
            for (uint32_t e = 0; e < E; ++e) k[1 + e] = i0 + e < n ? uint32_t(keys[i0 + e]) : 0xffffffffu;
        }
        // The neighbours come from the adjacent lanes, which hold the adjacent groups (lane l + 1
        // holds group g + 1 whenever key i0 + E exists); only the wave's end lanes load them.
        const int lane = int(threadIdx.x & 63u);
        const uint32_t up = __shfl_up(k[E], 1, 64), down = __shfl_down(k[1], 1, 64);
        k[0] = i0 == 0u ? 0xffffffffu : (lane > 0 ? up : uint32_t(keys[i0 - 1u]));
        k[E + 1] = i0 + E < n ? (lane < 63 ? down : uint32_t(keys[i0 + E])) : 0xffffffffu;
#pragma unroll
        for (uint32_t e = 0; e < E; ++e) {
            const uint32_t i = i0 + e;
            const uint32_t tile = k[1 + e];
            if (i >= n || tile >= n_tiles) continue;
            if (i == 0 || k[e] != tile) ranges[tile].x = i;
            if (i + 1 == n || k[2 + e] != tile) ranges[tile].y = i + 1;
        }
    }
}

template <class K>
__global__ __launch_bounds__(256) void k_debug_keys(const K* tile_keys, const uint32_t* slots,
                                                    const uint2* meta, const Counters* counters, uint32_t inst_cap,
                                                    int vp_log2, uint64_t* keys_out, uint32_t* vals_out) {
    const uint32_t ni = counters->instances;
    const uint32_t n = ni < inst_cap ? ni : inst_cap;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const uint32_t slot = slots[i];
        if (keys_out) keys_out[i] = (uint64_t(tile_keys[i]) << 32) | meta[slot].x;
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
    const float* src = sgrad + (int64_t(v) * n + g) * kGradStride;
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

void launch_clear(Counters* counters, uint2* ranges, uint32_t n_tiles, uint32_t* key_parts, hipStream_t st) {
    hipLaunchKernelGGL(k_clear, dim3(grid_stride_blocks(n_tiles)), dim3(256), 0, st, counters, ranges, n_tiles,
                       key_parts);
}
void launch_key_range(const uint32_t* key_parts, Counters* counters, hipStream_t st) {
    hipLaunchKernelGGL(k_key_range, dim3(1), dim3(kKeyParts), 0, st, key_parts, counters);
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

// Tile keys are 16-bit when every global tile id fits (narrow), else 32-bit; the key buffers are
// allocated as 32-bit arrays either way.
void launch_emit_instances(const uint32_t* sorted_slots, const uint32_t* sorted_tiles, const uint32_t* offsets,
                           const Counters* counters, uint32_t slot_cap, const SplatRec* rec, const DevView* views,
                           int vp_log2, int tile_size, uint32_t inst_cap, uint32_t* keys_out, uint32_t* vals_out,
                           bool narrow, hipStream_t st) {
    const dim3 grid(grid_stride_blocks(slot_cap));
    if (narrow)
        hipLaunchKernelGGL(k_emit<uint16_t>, grid, dim3(256), 0, st, sorted_slots, sorted_tiles, offsets, counters,
                           slot_cap, rec, views, vp_log2, tile_size, inst_cap, reinterpret_cast<uint16_t*>(keys_out),
                           vals_out);
    else
        hipLaunchKernelGGL(k_emit<uint32_t>, grid, dim3(256), 0, st, sorted_slots, sorted_tiles, offsets, counters,
                           slot_cap, rec, views, vp_log2, tile_size, inst_cap, keys_out, vals_out);
}

void launch_ranges(const uint32_t* keys, const Counters* counters, uint32_t inst_cap, uint32_t n_tiles,
                   uint2* ranges, bool narrow, hipStream_t st) {
    // one group of kRangeKeys keys per thread up to 65535 blocks, grid-stride beyond
    const uint64_t blocks = (uint64_t(inst_cap) + 256u * kRangeKeys - 1u) / (256u * kRangeKeys);
    const dim3 grid(unsigned(blocks < 1 ? 1 : (blocks > 65535 ? 65535 : blocks)));
    if (narrow)
        hipLaunchKernelGGL(k_ranges<uint16_t>, grid, dim3(256), 0, st, reinterpret_cast<const uint16_t*>(keys),
                           counters, inst_cap, n_tiles, ranges);
    else
        hipLaunchKernelGGL(k_ranges<uint32_t>, grid, dim3(256), 0, st, keys, counters, inst_cap, n_tiles, ranges);
}

void launch_debug_keys(const uint32_t* tile_keys, const uint32_t* slots, const uint2* meta,
                       const Counters* counters, uint32_t inst_cap, int vp_log2, uint64_t* keys_out,
                       uint32_t* vals_out, bool narrow, hipStream_t st) {
    const dim3 grid(grid_stride_blocks(inst_cap));
    if (narrow)
        hipLaunchKernelGGL(k_debug_keys<uint16_t>, grid, dim3(256), 0, st,
                           reinterpret_cast<const uint16_t*>(tile_keys), slots, meta, counters, inst_cap, vp_log2,
                           keys_out, vals_out);
    else
        hipLaunchKernelGGL(k_debug_keys<uint32_t>, grid, dim3(256), 0, st, tile_keys, slots, meta, counters,
                           inst_cap, vp_log2, keys_out, vals_out);
}

} // namespace pht
