// Densification (ADC-IGS refinement) row kernels: statistics, prune masks, radix-select row
// selection, IGS parent scores, the model/optimizer row remap, the split and the decay.
//
// The reference runs these as separate tensor ops with a host readback after most of them
// (photara/src/splat/densification_igs.cpp:291-515) and selects rows by sorting every row
// (a full 64-bit radix sort for the Gumbel sample, igs_selection_cuda.cu:71-136). Here:
// - the statistics of all views of a batched step are folded into one pass over the rows;
// - selection is a radix select (eight 8-bit digit passes over the keys, MSB first) that finds
//   the k-th largest key without ordering the rest; keys are unique (the row is in the low bits),
//   so the selected set is exactly {key >= threshold};
// - the remap of every model and optimizer array (keep rows, append children, zero the moments
//   of split rows) is one launch over a row map.
// Row math lives in splat_math.h, shared with the CPU oracle.
#include <algorithm>

#include "launch.h"

namespace pht {

namespace {

constexpr int kThreads = 256;

int blocks_for(int64_t n, int per_thread = 1, int cap = 4096) {
    const int64_t b = (n + int64_t(kThreads) * per_thread - 1) / (int64_t(kThreads) * per_thread);
    return int(b < 1 ? 1 : (b > cap ? cap : b));
}

// Block-wide sum of a per-thread count (kThreads threads), result valid in thread 0.
__device__ uint32_t block_count(uint32_t v, uint32_t* lds) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int wave = threadIdx.x >> 6;
    if (lane_id() == 0) lds[wave] = v;
    __syncthreads();
    uint32_t t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kThreads / 64; ++w) t += lds[w];
    return t;
}

} // namespace

// ---------------------------------------------------------------- statistics
__global__ __launch_bounds__(kThreads) void k_densify_accumulate(DensifyObsSet set, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        DensifyRow r = densify_load_row(set.stats, i);
        for (int v = 0; v < set.n; ++v) densify_accumulate_row(r, set.obs[v], set.inverse_resolution[v], i);
        densify_store_row(set.stats, i, r);
    }
}

void launch_densify_accumulate(const DensifyObsSet& set, int64_t n, hipStream_t st) {
    if (n <= 0 || set.n <= 0) return;
    k_densify_accumulate<<<blocks_for(n), kThreads, 0, st>>>(set, n);
}

// Fused into a batched backward: view v's observation of Gaussian i is its radius and visibility
// in that view's render and the refine statistic the blend backward left in record (i, v).
__global__ __launch_bounds__(kThreads) void k_densify_step(DensifyStepArgs a) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < a.n; i += int64_t(gridDim.x) * kThreads) {
        DensifyRow r;
        bool loaded = false;
        for (int v = 0; v < a.n_views; ++v) {
            const DevView& dv = a.views[v];
            const int32_t radius = dv.radii[i];
            if (radius <= 0) continue;
            if (!loaded) {
                r = densify_load_row(a.stats, i);
                loaded = true;
            }
            const float rw = a.sgrad[(int64_t(v) * a.n + i) * kGradStride + 9];
            densify_observe(r, a.rules, a.view_index[v], a.inverse_resolution[v], radius, dv.visibility[i], rw,
                            a.geometry_from_refine != 0, rw, false, 0.0f, false, 0.0f);
        }
        if (loaded) densify_store_row(a.stats, i, r);
    }
}

void launch_densify_step(const DensifyStepArgs& a, hipStream_t st) {
    if (a.n <= 0 || a.n_views <= 0) return;
    k_densify_step<<<blocks_for(a.n), kThreads, 0, st>>>(a);
}

// ---------------------------------------------------------------- prune masks
// One row per thread, 256 contiguous rows per block. The rows' SH values (the bulk of the bytes)
// are checked for finiteness by the whole block in one coalesced pass over its contiguous range,
// the result kept per row in LDS; a lane walking its own row would touch 64 lines per load.
__global__ __launch_bounds__(kThreads) void k_prune_masks(pht_gaussians g, int sh_stride, float min_opacity,
                                                          float max_bounds, float cx, float cy, float cz,
                                                          uint8_t* keep, uint8_t* hard, float* opacities) {
    __shared__ uint32_t s_bad[kThreads];
    const float center[3] = {cx, cy, cz};
    const int64_t i0 = int64_t(blockIdx.x) * kThreads;
    const int rows = int(g.count - i0 < kThreads ? g.count - i0 : kThreads);
    s_bad[threadIdx.x] = 0;
    __syncthreads();
    if (sh_stride > 0) {
        const float* sh = g.sh + i0 * sh_stride;
        const int total = rows * sh_stride;
        if ((sh_stride & 3) == 0 && (reinterpret_cast<uintptr_t>(sh) & 15u) == 0) {
            const float4* sh4 = reinterpret_cast<const float4*>(sh);
            for (int c = int(threadIdx.x); c < (total >> 2); c += kThreads) {
                const float4 v = sh4[c];
                if (!(finite_(v.x) && finite_(v.y) && finite_(v.z) && finite_(v.w))) s_bad[(4 * c) / sh_stride] = 1u;
            }
        } else {
            for (int e = int(threadIdx.x); e < total; e += kThreads)
                if (!finite_(sh[e])) s_bad[e / sh_stride] = 1u;
        }
    }
    __syncthreads();
    const int64_t i = i0 + threadIdx.x;
    if (i >= g.count) return;
    bool k, h;
    opacities[i] = densify_prune_row(g.means, g.log_scales, g.quaternions, g.opacity_logits, nullptr, 0, i,
                                     min_opacity, max_bounds, center, k, h);
    if (s_bad[threadIdx.x]) {
        h = true;
        k = false;
    }
    keep[i] = k ? 1 : 0;
    hard[i] = h ? 1 : 0;
}

void launch_prune_masks(const pht_gaussians& g, int sh_stride, float min_opacity, float max_bounds,
                        const float center[3], uint8_t* keep, uint8_t* hard, float* opacities, hipStream_t st) {
    if (g.count <= 0) return;
    const unsigned blocks = unsigned((g.count + kThreads - 1) / kThreads); // one row per thread
    k_prune_masks<<<blocks, kThreads, 0, st>>>(g, sh_stride, min_opacity, max_bounds, center[0], center[1], center[2],
                                               keep, hard, opacities);
}

// ---------------------------------------------------------------- radix select
__global__ void k_select_init(SelectState* s) {
    for (int t = threadIdx.x; t < 256; t += blockDim.x) s->hist[t] = 0;
    if (threadIdx.x == 0) {
        s->eligible = 0;
        s->remaining = 0;
        s->chosen = 0;
        s->prefix = 0;
    }
}

__global__ __launch_bounds__(kThreads) void k_select_keys(const float* values, const uint8_t* mask,
                                                          const uint8_t* exclude, int64_t n, int mode,
                                                          uint64_t seed, uint64_t* keys, SelectState* s) {
    __shared__ uint32_t lds[kThreads / 64];
    uint32_t eligible = 0;
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        uint64_t key = 0;
        if ((mask == nullptr || mask[i] != 0) && (exclude == nullptr || exclude[i] == 0)) {
            const float v = values[i];
            if (mode == PHT_SELECT_GUMBEL) {
                key = gumbel_key(v, seed, uint32_t(i));
            } else if (v == v && (mode != PHT_SELECT_LARGEST_POSITIVE || (v > 0.0f && finite_(v)))) {
                key = select_value_key(v, mode != PHT_SELECT_SMALLEST, uint32_t(i));
            }
        }
        keys[i] = key;
        eligible += key != 0 ? 1u : 0u;
    }
    const uint32_t t = block_count(eligible, lds);
    if (threadIdx.x == 0 && t) atomicAdd(&s->eligible, t);
}

__global__ void k_select_begin(SelectState* s, int64_t k) {
    if (threadIdx.x == 0) {
        const uint64_t e = s->eligible;
        const uint64_t kk = k < 0 ? 0 : uint64_t(k);
        s->remaining = uint32_t(kk < e ? kk : e);
        s->chosen = s->remaining;
    }
}

// Histogram of digit `pass` (bits 63-8p .. 56-8p) over the keys whose higher digits equal the
// prefix chosen so far.
__global__ __launch_bounds__(kThreads) void k_select_hist(const uint64_t* keys, int64_t n, SelectState* s, int pass) {
    __shared__ uint32_t h[256];
    if (s->remaining == 0) return; // uniform over the grid
    h[threadIdx.x] = 0;
    __syncthreads();
    const int shift = 56 - 8 * pass;
    const uint64_t prefix = s->prefix;
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        const uint64_t key = keys[i];
        if (key == 0) continue;
        if (pass > 0 && (key >> (shift + 8)) != prefix) continue;
        atomicAdd(&h[uint32_t(key >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    const uint32_t c = h[threadIdx.x];
    if (c) atomicAdd(&s->hist[threadIdx.x], c);
}

// One block of 256: the digit d (from the top) at which the running count of larger digits
// first reaches `remaining`; appends d to the prefix and clears the histogram.
__global__ __launch_bounds__(256) void k_select_pick(SelectState* s) {
    __shared__ uint32_t wave_tot[4];
    if (s->remaining == 0) return;
    const int t = threadIdx.x;          // t = 0 is digit 255
    const uint32_t c = s->hist[255 - t];
    const uint32_t incl_wave = wave_inclusive_scan(c);
    if (lane_id() == 63) wave_tot[t >> 6] = incl_wave;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < (t >> 6); ++w) before += wave_tot[w];
    const uint32_t incl = before + incl_wave, excl = incl - c;
    const uint32_t rem = s->remaining;
    __syncthreads(); // every thread has read remaining
    if (excl < rem && rem <= incl) {
        s->remaining = rem - excl;
        s->prefix = (s->prefix << 8) | uint64_t(255 - t);
    }
    s->hist[255 - t] = 0;
}

__global__ __launch_bounds__(kThreads) void k_select_flag(const uint64_t* keys, int64_t n, const SelectState* s,
                                                          uint8_t* out, uint8_t out_value, uint32_t* counts,
                                                          uint64_t* threshold) {
    const bool any = s->chosen != 0;
    const uint64_t thr = s->prefix; // after eight passes: the full key of the k-th largest
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (counts) {
            counts[0] = s->eligible;
            counts[1] = s->chosen;
        }
        if (threshold) *threshold = any ? thr : 0;
    }
    if (!any) return;
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        const uint64_t key = keys[i];
        if (key != 0 && key >= thr) out[i] = out_value;
    }
}

size_t select_workspace_bytes(int64_t n) {
    return ((size_t(n < 1 ? 1 : n) * sizeof(uint64_t) + 255) & ~size_t(255)) + sizeof(SelectState);
}

void launch_select_rows(const float* values, const uint8_t* mask, const uint8_t* exclude, int64_t n, int64_t k,
                        int mode, uint64_t seed, uint8_t* out, uint8_t out_value, uint32_t* counts,
                        uint64_t* threshold, void* workspace, hipStream_t st) {
    uint64_t* keys = static_cast<uint64_t*>(workspace);
    SelectState* s = reinterpret_cast<SelectState*>(static_cast<char*>(workspace) +
                                                    ((size_t(n < 1 ? 1 : n) * sizeof(uint64_t) + 255) & ~size_t(255)));
    k_select_init<<<1, 256, 0, st>>>(s);
    const int blocks = blocks_for(n, 4, 2048);
    if (n > 0) k_select_keys<<<blocks, kThreads, 0, st>>>(values, mask, exclude, n, mode, seed, keys, s);
    k_select_begin<<<1, 64, 0, st>>>(s, k);
    for (int pass = 0; pass < 8 && n > 0; ++pass) {
        k_select_hist<<<blocks, kThreads, 0, st>>>(keys, n, s, pass);
        k_select_pick<<<1, 256, 0, st>>>(s);
    }
    k_select_flag<<<blocks, kThreads, 0, st>>>(keys, n, s, out, out_value, counts, threshold);
}

// ---------------------------------------------------------------- IGS parent scores
__global__ __launch_bounds__(kThreads) void k_igs_prepare(pht_densify_stats s, const int32_t* keep_rows,
                                                          int64_t retained, float gthr, float sthr, float* priority,
                                                          uint8_t* positive, uint32_t* counts) {
    __shared__ uint32_t lds[kThreads / 64];
    uint32_t c_pos = 0, c_grow = 0, c_over = 0;
    for (int64_t r = int64_t(blockIdx.x) * kThreads + threadIdx.x; r < retained; r += int64_t(gridDim.x) * kThreads) {
        const int64_t src = keep_rows ? keep_rows[r] : r;
        const IgsRow row = igs_row(s, src, gthr, sthr);
        priority[r] = row.priority;
        positive[r] = row.positive ? 1 : 0;
        c_pos += row.positive;
        c_grow += row.growth;
        c_over += row.oversized;
    }
    uint32_t t = block_count(c_pos, lds);
    if (threadIdx.x == 0 && t) atomicAdd(&counts[0], t);
    __syncthreads();
    t = block_count(c_grow, lds);
    if (threadIdx.x == 0 && t) atomicAdd(&counts[1], t);
    __syncthreads();
    t = block_count(c_over, lds);
    if (threadIdx.x == 0 && t) atomicAdd(&counts[2], t);
}

__global__ void k_zero_u32(uint32_t* p, int n) {
    if (int(threadIdx.x) < n) p[threadIdx.x] = 0;
}

void launch_igs_prepare(const pht_densify_stats& s, const int32_t* keep_rows, int64_t retained, float gthr,
                        float sthr, float* priority, uint8_t* positive, uint32_t* counts, hipStream_t st) {
    k_zero_u32<<<1, 64, 0, st>>>(counts, 3);
    if (retained <= 0) return;
    k_igs_prepare<<<blocks_for(retained, 4, 2048), kThreads, 0, st>>>(s, keep_rows, retained, gthr, sthr, priority,
                                                                     positive, counts);
}

__global__ __launch_bounds__(kThreads) void k_igs_weights(pht_densify_stats s, const int32_t* keep_rows,
                                                          int64_t retained, const float* opacities, float median,
                                                          int allow_growth, float gthr, float sthr,
                                                          float* replacement, float* growth, float* oversize) {
    for (int64_t r = int64_t(blockIdx.x) * kThreads + threadIdx.x; r < retained; r += int64_t(gridDim.x) * kThreads) {
        const int64_t src = keep_rows ? keep_rows[r] : r;
        float rw, gw, ow;
        igs_weights_row(s, src, opacities[src], median, allow_growth != 0, gthr, sthr, rw, gw, ow);
        replacement[r] = rw;
        if (growth) growth[r] = gw;
        if (oversize) oversize[r] = ow;
    }
}

void launch_igs_weights(const pht_densify_stats& s, const int32_t* keep_rows, int64_t retained,
                        const float* opacities, float median, int allow_growth, float gthr, float sthr,
                        float* replacement, float* growth, float* oversize, hipStream_t st) {
    if (retained <= 0) return;
    k_igs_weights<<<blocks_for(retained), kThreads, 0, st>>>(s, keep_rows, retained, opacities, median, allow_growth,
                                                            gthr, sthr, replacement, growth, oversize);
}

// ---------------------------------------------------------------- row remap
// blockIdx.y selects the buffer; each buffer moves in 16-, 4- or 2-byte units, whichever its row
// size and alignment allow.
template <class U>
__device__ void gather_units(const RowBuffer& b, const int32_t* rows, int64_t n_rows) {
    const int64_t per_row = b.row_bytes / int64_t(sizeof(U));
    const int64_t total = n_rows * per_row;
    const U* src = static_cast<const U*>(b.src);
    U* dst = static_cast<U*>(b.dst);
    for (int64_t w = int64_t(blockIdx.x) * kThreads + threadIdx.x; w < total; w += int64_t(gridDim.x) * kThreads) {
        const int64_t r = w / per_row, c = w - r * per_row;
        const int32_t s = rows[r];
        U v;
        if (s >= 0) {
            v = src[int64_t(s) * per_row + c];
        } else {
            v = U{};
        }
        dst[w] = v;
    }
}

__global__ __launch_bounds__(kThreads) void k_gather_rows(RowBuffers set, const int32_t* rows, int64_t n_rows) {
    const RowBuffer& b = set.b[blockIdx.y];
    switch (b.unit) {
    case 16: gather_units<uint4>(b, rows, n_rows); break;
    case 4: gather_units<uint32_t>(b, rows, n_rows); break;
    default: gather_units<uint16_t>(b, rows, n_rows); break;
    }
}

void launch_gather_rows(const RowBuffers& set, const int32_t* rows, int64_t n_rows, hipStream_t st) {
    if (n_rows <= 0 || set.n <= 0) return;
    int64_t most = 0;
    for (int i = 0; i < set.n; ++i) most = std::max(most, n_rows * set.b[i].row_bytes / set.b[i].unit);
    dim3 grid(blocks_for(most, 4, 2048), set.n);
    k_gather_rows<<<grid, kThreads, 0, st>>>(set, rows, n_rows);
}

// ---------------------------------------------------------------- split and decay
__global__ __launch_bounds__(kThreads) void k_densify_split(pht_split_params p, const int32_t* parents, int64_t count,
                                                            int64_t child_base, const float* samples,
                                                            const float* screens) {
    for (int64_t j = int64_t(blockIdx.x) * kThreads + threadIdx.x; j < count; j += int64_t(gridDim.x) * kThreads) {
        const int64_t parent = parents[j], child = child_base + j;
        densify_split_row(p.means, p.log_scales, p.opacity_logits, p.quaternions, parent, child, samples + 3 * j,
                          screens ? screens[j] : 0.0f, screens != nullptr, p.mode, p.minimum_opacity,
                          p.split_at_screen_size, p.split_opacity_k);
        if (p.lp_format != PHT_LP_NONE) {
            for (int k = 0; k < 3; ++k) {
                p.log_scales_lp[3 * parent + k] = f32_to_lp(p.log_scales[3 * parent + k], p.lp_format);
                p.log_scales_lp[3 * child + k] = f32_to_lp(p.log_scales[3 * child + k], p.lp_format);
            }
            p.opacity_logits_lp[parent] = f32_to_lp(p.opacity_logits[parent], p.lp_format);
            p.opacity_logits_lp[child] = f32_to_lp(p.opacity_logits[child], p.lp_format);
        }
    }
}

void launch_densify_split(const pht_split_params& p, const int32_t* parents, int64_t count, int64_t child_base,
                          const float* samples, const float* screens, hipStream_t st) {
    if (count <= 0) return;
    k_densify_split<<<blocks_for(count), kThreads, 0, st>>>(p, parents, count, child_base, samples, screens);
}

__global__ __launch_bounds__(kThreads) void k_adc_decay(float* log_scales, float* logits, uint16_t* ls_lp,
                                                        uint16_t* logit_lp, int fmt, int64_t n, float od, float lsd) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        densify_decay_row(log_scales, logits, i, od, lsd);
        if (fmt != PHT_LP_NONE) {
            for (int k = 0; k < 3; ++k) ls_lp[3 * i + k] = f32_to_lp(log_scales[3 * i + k], fmt);
            logit_lp[i] = f32_to_lp(logits[i], fmt);
        }
    }
}

void launch_adc_decay(float* log_scales, float* logits, uint16_t* ls_lp, uint16_t* logit_lp, int fmt, int64_t n,
                      float opacity_decay, float log_scale_decay, hipStream_t st) {
    if (n <= 0) return;
    k_adc_decay<<<blocks_for(n), kThreads, 0, st>>>(log_scales, logits, ls_lp, logit_lp, fmt, n, opacity_decay,
                                                   log_scale_decay);
}

} // namespace pht
