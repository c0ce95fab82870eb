// Device-count-driven chunked scans and a stable LSD radix sort for gfx950.
//
// Replaces the CUB calls of the reference pipeline (cub::DeviceScan::InclusiveSum and
// cub::DeviceRadixSort::SortPairs, third_party/splat_drender/src/pipeline.cu:234-305).
// Every pass reads its element count from device memory, so the whole forward needs no host
// readback and can be captured in a hipGraph. A fixed number of blocks (<= kScanBlocks) each
// own one contiguous chunk of the list; within a chunk elements are processed in order, which
// makes every scan exclusive-prefix exact and every sort pass stable:
//   reduce  : per-block partial (sum, or 256-bin digit histogram)
//   scan    : one block turns the partials into exclusive block offsets
//   write   : per-block in-order rescan / wave-level multi-split ranking + scatter
#include "launch.h"

namespace pht {

__device__ __forceinline__ uint32_t load_count(const uint32_t* n_ptr, uint32_t n_const, uint32_t cap) {
    const uint32_t n = n_ptr ? *n_ptr : n_const;
    return n < cap ? n : cap;
}

__device__ __forceinline__ void chunk_of(uint32_t n, int nb, int b, uint32_t& lo, uint32_t& hi) {
    uint32_t chunk = (n + uint32_t(nb) - 1u) / uint32_t(nb);
    chunk = (chunk + 255u) & ~255u;
    const uint64_t l = uint64_t(b) * chunk;
    lo = l < n ? uint32_t(l) : n;
    const uint64_t h = uint64_t(lo) + chunk;
    hi = h < n ? uint32_t(h) : n;
}

// Block-wide inclusive scan of one value per thread (256 threads = 4 waves).
__device__ __forceinline__ uint32_t block_inclusive_scan(uint32_t v, uint32_t* s_w, uint32_t& total) {
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (int k = 0; k < 4; ++k) {
        const uint32_t t = s_w[k];
        if (k < w) before += t;
        total += t;
    }
    __syncthreads();
    return inc + before;
}

// ---------------------------------------------------------------- value sources / sinks
struct SrcFlags { // visibility flags of the dense slot array
    const uint32_t* flags;
    __device__ uint32_t operator()(uint32_t i) const { return flags[i]; }
};
struct SrcTiles { // tile counts gathered in depth order
    const uint32_t* sorted_slots;
    const uint4* rec_c;
    __device__ uint32_t operator()(uint32_t i) const { return rec_c[sorted_slots[i]].z; }
};
struct SinkCompact { // visible slot -> (depth key, slot) pair at its compact position
    const uint4* rec_c;
    uint32_t* keys;
    uint32_t* vals;
    __device__ void operator()(uint32_t i, uint32_t excl, uint32_t v) const {
        if (v) {
            keys[excl] = rec_c[i].y;
            vals[excl] = i;
        }
    }
};
struct SinkOffsets {
    uint32_t* out;
    __device__ void operator()(uint32_t i, uint32_t excl, uint32_t) const { out[i] = excl; }
};

template <class Src>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(Src src, const uint32_t* n_ptr, uint32_t n_const,
                                                              uint32_t cap, int nb, uint32_t* block_sums) {
    __shared__ uint32_t s_w[4];
    const uint32_t n = load_count(n_ptr, n_const, cap);
    uint32_t lo, hi;
    chunk_of(n, nb, blockIdx.x, lo, hi);
    uint32_t acc = 0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kScanThreads) acc += src(i);
    uint32_t total;
    block_inclusive_scan(acc, s_w, total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// One block of 1024 threads: exclusive scan of nb <= 1024 block sums in place; the grand
// total is stored to *total_out and, when overflow_out is set, *overflow_out = total > cap.
__global__ __launch_bounds__(1024) void k_scan_blocks(uint32_t* block_sums, int nb, uint32_t* total_out,
                                                      uint32_t* overflow_out, uint32_t cap) {
    __shared__ uint32_t s_w[16];
    const int t = threadIdx.x, lane = lane_id(), w = t >> 6;
    const uint32_t v = t < nb ? block_sums[t] : 0u;
    const uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int k = 0; k < 16; ++k) {
        const uint32_t x = s_w[k];
        if (k < w) before += x;
        total += x;
    }
    if (t < nb) block_sums[t] = inc + before - v;
    if (t == 0) {
        if (total_out) *total_out = total;
        if (overflow_out) *overflow_out = total > cap ? 1u : 0u;
    }
}

template <class Src, class Sink>
__global__ __launch_bounds__(kScanThreads) void k_scan_write(Src src, Sink sink, const uint32_t* n_ptr,
                                                             uint32_t n_const, uint32_t cap, int nb,
                                                             const uint32_t* block_offsets) {
    __shared__ uint32_t s_w[4];
    const uint32_t n = load_count(n_ptr, n_const, cap);
    uint32_t lo, hi;
    chunk_of(n, nb, blockIdx.x, lo, hi);
    uint32_t running = block_offsets[blockIdx.x];
    for (uint32_t base = lo; base < hi; base += kScanThreads) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < hi ? src(i) : 0u;
        uint32_t total;
        const uint32_t inc = block_inclusive_scan(v, s_w, total);
        if (i < hi) sink(i, running + inc - v, v);
        running += total;
    }
}

// ---------------------------------------------------------------- radix sort (8-bit digits)
__global__ __launch_bounds__(kScanThreads) void k_radix_upsweep(const uint32_t* keys, const uint32_t* n_ptr,
                                                                uint32_t n_const, uint32_t cap, int shift, int nb,
                                                                uint32_t* hist) {
    __shared__ uint32_t s_h[256];
    s_h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t n = load_count(n_ptr, n_const, cap);
    uint32_t lo, hi;
    chunk_of(n, nb, blockIdx.x, lo, hi);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kScanThreads)
        atomicAdd(&s_h[(keys[i] >> shift) & 255u], 1u);
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = s_h[threadIdx.x];
}

// Exclusive scan of the digit-major histogram hist[256 * nb] by one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_radix_scan(uint32_t* hist, int nb) {
    __shared__ uint32_t s_w[16];
    const int total = 256 * nb;
    const int per = (total + 1023) / 1024;
    const int t = threadIdx.x, lane = lane_id(), w = t >> 6;
    const int b = t * per, e = min(total, b + per);
    uint32_t sum = 0;
    for (int i = b; i < e; ++i) sum += hist[i];
    const uint32_t inc = wave_inclusive_scan(sum);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (int k = 0; k < w; ++k) before += s_w[k];
    uint32_t run = inc - sum + before;
    for (int i = b; i < e; ++i) {
        const uint32_t x = hist[i];
        hist[i] = run;
        run += x;
    }
}

// Stable scatter of one chunk. Tiles of 256 keys are ranked with a wave-level multi-split
// (8 ballots give each lane the mask of same-digit lanes), per-wave digit counts are combined
// across the 4 waves in LDS, and a running per-digit base carries the order across tiles.
__global__ __launch_bounds__(kScanThreads) void k_radix_downsweep(const uint32_t* keys_in, const uint32_t* vals_in,
                                                                  uint32_t* keys_out, uint32_t* vals_out,
                                                                  const uint32_t* n_ptr, uint32_t n_const,
                                                                  uint32_t cap, int shift, int nb,
                                                                  const uint32_t* hist) {
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_wh[4][256];
    const int t = threadIdx.x, lane = lane_id(), w = t >> 6;
    s_base[t] = hist[t * nb + blockIdx.x];
    const uint32_t n = load_count(n_ptr, n_const, cap);
    uint32_t lo, hi;
    chunk_of(n, nb, blockIdx.x, lo, hi);
    const uint64_t lt = __lanemask_lt();
    for (uint32_t base = lo; base < hi; base += kScanThreads) {
        for (int k = 0; k < 4; ++k) s_wh[k][t] = 0;
        __syncthreads();
        const uint32_t i = base + t;
        const bool valid = i < hi;
        uint32_t key = 0, val = 0, d = 0;
        if (valid) {
            key = keys_in[i];
            val = vals_in[i];
            d = (key >> shift) & 255u;
        }
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool set = (d >> bit) & 1u;
            const uint64_t bal = __ballot(valid && set);
            peers &= set ? bal : ~bal;
        }
        const uint32_t rank = uint32_t(__popcll(peers & lt));
        if (valid && (peers & lt) == 0) s_wh[w][d] = uint32_t(__popcll(peers));
        __syncthreads();
        {
            uint32_t s = s_base[t];
            for (int k = 0; k < 4; ++k) {
                const uint32_t c = s_wh[k][t];
                s_wh[k][t] = s;
                s += c;
            }
            s_base[t] = s;
        }
        __syncthreads();
        if (valid) {
            const uint32_t dst = s_wh[w][d] + rank;
            keys_out[dst] = key;
            vals_out[dst] = val;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- host launchers
static int blocks_for(uint32_t capacity) {
    const uint64_t b = (uint64_t(capacity) + 4095u) / 4096u;
    return int(b < 1 ? 1 : (b > uint64_t(kScanBlocks) ? kScanBlocks : b));
}

int scan_blocks_for(uint32_t capacity) { return blocks_for(capacity); }

void launch_compact_visible(const uint32_t* flags, const uint4* rec_c, uint32_t n_slots, uint32_t* keys,
                            uint32_t* vals, uint32_t* block_scratch, uint32_t* visible_out,
                            hipStream_t st) {
    const int nb = blocks_for(n_slots);
    SrcFlags src{flags};
    hipLaunchKernelGGL(k_scan_reduce<SrcFlags>, dim3(nb), dim3(kScanThreads), 0, st, src, nullptr, n_slots,
                       n_slots, nb, block_scratch);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, block_scratch, nb, visible_out, nullptr,
                       0xffffffffu);
    SinkCompact sink{rec_c, keys, vals};
    hipLaunchKernelGGL((k_scan_write<SrcFlags, SinkCompact>), dim3(nb), dim3(kScanThreads), 0, st, src, sink,
                       nullptr, n_slots, n_slots, nb, block_scratch);
}

void launch_tile_offsets(const uint32_t* sorted_slots, const uint4* rec_c, const uint32_t* visible_ptr,
                         uint32_t slot_cap, uint32_t* offsets, uint32_t* block_scratch,
                         uint32_t* instances_out, uint32_t* overflow_out, uint32_t inst_cap, hipStream_t st) {
    const int nb = blocks_for(slot_cap);
    SrcTiles src{sorted_slots, rec_c};
    hipLaunchKernelGGL(k_scan_reduce<SrcTiles>, dim3(nb), dim3(kScanThreads), 0, st, src, visible_ptr, 0u,
                       slot_cap, nb, block_scratch);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, block_scratch, nb, instances_out,
                       overflow_out, inst_cap);
    SinkOffsets sink{offsets};
    hipLaunchKernelGGL((k_scan_write<SrcTiles, SinkOffsets>), dim3(nb), dim3(kScanThreads), 0, st, src, sink,
                       visible_ptr, 0u, slot_cap, nb, block_scratch);
}

int launch_radix_sort(uint32_t* keys[2], uint32_t* vals[2], const uint32_t* n_ptr, uint32_t cap, int bits,
                      uint32_t* hist, hipStream_t st) {
    const int nb = blocks_for(cap);
    int cur = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_upsweep, dim3(nb), dim3(kScanThreads), 0, st, keys[cur], n_ptr, 0u, cap, shift,
                           nb, hist);
        hipLaunchKernelGGL(k_radix_scan, dim3(1), dim3(1024), 0, st, hist, nb);
        hipLaunchKernelGGL(k_radix_downsweep, dim3(nb), dim3(kScanThreads), 0, st, keys[cur], vals[cur],
                           keys[cur ^ 1], vals[cur ^ 1], n_ptr, 0u, cap, shift, nb, hist);
        cur ^= 1;
    }
    return cur;
}

} // namespace pht
