// Device-count-driven chunked scans and a stable LSD radix sort for gfx950.
//
// Replaces the CUB calls of the reference pipeline (cub::DeviceScan::InclusiveSum and
// cub::DeviceRadixSort::SortPairs, third_party/splat_drender/src/pipeline.cu:234-305).
// Every pass reads its element count from device memory, so the whole forward needs no host
// readback and can be captured in a hipGraph. A fixed number of blocks (<= kScanBlocks) each
// own one contiguous chunk of the list; within a chunk elements are processed in order, which
// makes every scan exclusive-prefix exact and every sort pass stable:
//   reduce  : per-block partial (sum, or 256-bin digit histogram)
//   scan    : partials -> exclusive block offsets (one block; one block per digit for the sort)
//   write   : per-block in-order rescan / LDS tile ranking + per-digit contiguous writes
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

// Block-wide inclusive scan of one value per thread (NW waves; 256 threads = 4 waves).
template <int NW = 4>
__device__ __forceinline__ uint32_t block_inclusive_scan(uint32_t v, uint32_t* s_w, uint32_t& total) {
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (int k = 0; k < NW; ++k) {
        const uint32_t t = s_w[k];
        if (k < w) before += t;
        total += t;
    }
    __syncthreads();
    return inc + before;
}

// ---------------------------------------------------------------- value sources / sinks
// kBinary: every value is 0 or 1, so a wave's inclusive scan is a ballot and a popcount.
struct SrcFlags { // visibility flags of the dense slot array
    static constexpr bool kBinary = true;
    const uint32_t* flags;
    __device__ uint32_t operator()(uint32_t i) const { return flags[i]; }
};
struct SrcTiles { // tile counts, already permuted into depth order by the depth sort
    static constexpr bool kBinary = false;
    const uint32_t* counts;
    __device__ uint32_t operator()(uint32_t i) const { return counts[i]; }
};

template <class Src> __device__ __forceinline__ uint32_t src_wave_inclusive_scan(uint32_t v) {
    if constexpr (Src::kBinary) return uint32_t(__popcll(__ballot(v != 0u) & __lanemask_lt())) + v;
    else return wave_inclusive_scan(v);
}
// A sink's load(i, v) fetches what it stores for element i before the scan needs it, so that a
// block's loads of every row are in flight together; store(i, excl, v, payload) writes it.
struct SinkCompact { // visible slot -> (depth key, slot, tile count) at its compact position
    const uint2* meta;
    uint32_t* keys;
    uint32_t* vals;
    uint32_t* tiles;
    using Payload = uint2;
    __device__ Payload load(uint32_t i, uint32_t v) const {
        if (!v) return make_uint2(0u, 0u);
        const uint2 c = meta[i];
        return make_uint2(c.x, c.y & kMetaCountMask);
    }
    __device__ void store(uint32_t i, uint32_t excl, uint32_t v, Payload c) const {
        if (v) {
            keys[excl] = c.x;
            vals[excl] = i;
            tiles[excl] = c.y;
        }
    }
};
struct SinkOffsets {
    uint32_t* out;
    struct Payload {};
    __device__ Payload load(uint32_t, uint32_t) const { return {}; }
    __device__ void store(uint32_t i, uint32_t excl, uint32_t, Payload) const { out[i] = excl; }
};

// A block walks its chunk kScanRows rows of kScanThreads consecutive elements at a time (row k of
// an iteration = base + 256 k .. base + 256 k + 255): every row's loads and stores are coalesced,
// all rows' loads are issued before any value is used, and the block synchronises once per
// kScanRows * 256 elements instead of once per row.
constexpr int kScanRows = 8;

template <class Src>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(Src src, const uint32_t* n_ptr, uint32_t n_const,
                                                              uint32_t cap, int nb, uint32_t* block_sums) {
    __shared__ uint32_t s_w[4];
    const uint32_t n = load_count(n_ptr, n_const, cap);
    uint32_t lo, hi;
    chunk_of(n, nb, blockIdx.x, lo, hi);
    uint32_t acc = 0;
    for (uint32_t base = lo; base < hi; base += kScanRows * kScanThreads) {
        uint32_t v[kScanRows];
#pragma unroll
        for (int k = 0; k < kScanRows; ++k) {
            const uint32_t i = base + uint32_t(k) * kScanThreads + threadIdx.x;
            v[k] = i < hi ? src(i) : 0u;
        }
#pragma unroll
        for (int k = 0; k < kScanRows; ++k) acc += v[k];
    }
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
    constexpr int NW = kScanThreads / 64;
    // per-row wave totals, double-buffered: iteration j + 2 rewrites a buffer only after every
    // thread has passed iteration j + 1's barrier, so after its reads of iteration j
    __shared__ uint32_t s_w[2][kScanRows][NW];
    const uint32_t n = load_count(n_ptr, n_const, cap);
    uint32_t lo, hi;
    chunk_of(n, nb, blockIdx.x, lo, hi);
    const int lane = lane_id(), w = int(threadIdx.x) >> 6;
    uint32_t running = block_offsets[blockIdx.x];
    int buf = 0;
    for (uint32_t base = lo; base < hi; base += kScanRows * kScanThreads, buf ^= 1) {
        uint32_t v[kScanRows], inc[kScanRows];
        typename Sink::Payload pay[kScanRows];
#pragma unroll
        for (int k = 0; k < kScanRows; ++k) {
            const uint32_t i = base + uint32_t(k) * kScanThreads + threadIdx.x;
            v[k] = i < hi ? src(i) : 0u;
        }
#pragma unroll
        for (int k = 0; k < kScanRows; ++k) {
            const uint32_t i = base + uint32_t(k) * kScanThreads + threadIdx.x;
            pay[k] = sink.load(i < hi ? i : 0u, i < hi ? v[k] : 0u);
            inc[k] = src_wave_inclusive_scan<Src>(v[k]);
            if (lane == 63) s_w[buf][k][w] = inc[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kScanRows; ++k) {
            uint32_t before_w = 0, row_total = 0;
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                const uint32_t x = s_w[buf][k][q];
                before_w += q < w ? x : 0u;
                row_total += x;
            }
            const uint32_t i = base + uint32_t(k) * kScanThreads + threadIdx.x;
            if (i < hi) sink.store(i, running + before_w + inc[k] - v[k], v[k], pay[k]);
            running += row_total;
        }
    }
}

// ---------------------------------------------------------------- radix sort (8-bit digits)
// One pass = upsweep (per-block digit histogram of its chunk), digit scan (one block per digit
// scans that digit's counts over the blocks and stores the digit total) and downsweep. The
// downsweep ranks tiles of kSortTile keys in LDS: each wave ranks its keys row by row (the lanes
// sharing a digit found through an LDS lane-mask table) against per-wave running digit counters, the tile is
// reordered by digit in LDS, and each digit's run is written out contiguously. A block of BT
// threads ranks BT * R keys per iteration (R keys per lane): the tile sort uses 1024 x 16, whose
// 16384-key tiles give each digit runs of 64 keys on average (whole 128-byte lines per digit and
// array in the write-out, and half the per-tile bookkeeping of 8192-key tiles); the depth sort,
// with a third array in LDS, 1024 x 8.

// Depth sort: its passes rank key - base, base = the least visible key, from the preprocess's
// range counters (krange = {max of ~key, max of key}), so a pass whose digit is then the same for
// every key can be skipped. Its pass roles: 0 = an ordinary pass; 1 = the pass before the top
// one, which writes to the final (alternate) buffers when the top pass is not needed; 2 = the
// top pass, which does nothing when the range fits 24 bits. The tile sort passes krange = null.
__device__ __forceinline__ uint32_t key_base(const uint32_t* krange) { return krange ? ~krange[0] : 0u; }
__device__ __forceinline__ bool top_pass_needed(const uint32_t* krange) {
    return !krange || ((krange[1] - ~krange[0]) >> 24) != 0u; // no visible key: 0 - ~0 = 1, not needed
}

// 1024 threads per chunk (16 waves sharing the block's LDS sub-histograms): the chunks are the
// downsweep's, one per CU, so the block's width sets how many loads a CU has in
// flight (four times as many as 256-thread blocks; the six passes of the bench workload went
// from 456 to 409 us, still short of HBM rate, with the LDS atomics the other limit).
constexpr int kUpThreads = 1024;
// K = uint32_t, or uint16_t for tile keys when every global tile id fits 16 bits (half the key
// bytes of every pass). A 16-byte vector holds KV keys.
template <class K>
__global__ __launch_bounds__(kUpThreads) void k_radix_upsweep(const K* keys, const uint32_t* n_ptr,
                                                              uint32_t n_const, uint32_t cap, int shift, int nb,
                                                              uint32_t* hist, const uint32_t* krange, int role) {
    constexpr uint32_t KV = 16u / sizeof(K);
    if (role == 2 && !top_pass_needed(krange)) return; // uniform over the grid
    const uint32_t kb = key_base(krange);
    // kUpSubs sub-histograms, lane l counting into sub-histogram l % kUpSubs. Their rows are 257
    // words apart, so the lanes of one instruction that hold the same digit hit kUpSubs different
    // banks: a pass whose keys share one or two digit values (the top byte of depth keys) no
    // longer serialises 32 lanes on one LDS word.
    constexpr int kUpSubs = 8, kUpPitch = 257;
    __shared__ uint32_t s_h[kUpSubs * kUpPitch];
    const int t = threadIdx.x;
    for (int k = t; k < kUpSubs * kUpPitch; k += kUpThreads) s_h[k] = 0;
    __syncthreads();
    uint32_t* h = s_h + (t & (kUpSubs - 1)) * kUpPitch;
    const uint32_t n = load_count(n_ptr, n_const, cap);
    uint32_t lo, hi;
    chunk_of(n, nb, blockIdx.x, lo, hi);
    // lo is a multiple of 256 keys, so the 16-byte view of [lo, hiv) is aligned; four vector
    // loads per thread are issued before any of them is counted (bytes in flight)
    const uint32_t hiv = lo + ((hi - lo) & ~(KV - 1u));
    const uint4* k4 = reinterpret_cast<const uint4*>(keys);
    constexpr uint32_t kStep = KV * kUpThreads;
    auto count_word = [&](uint32_t x) {
        if constexpr (sizeof(K) == 4) {
            atomicAdd(&h[((x - kb) >> shift) & 255u], 1u);
        } else {
            atomicAdd(&h[((x & 0xffffu) >> shift) & 255u], 1u);
            atomicAdd(&h[((x >> 16) >> shift) & 255u], 1u);
        }
    };
    auto count_vec = [&](const uint4& q) {
        count_word(q.x);
        count_word(q.y);
        count_word(q.z);
        count_word(q.w);
    };
    uint32_t i = lo + KV * t;
    for (; i + 3u * kStep < hiv; i += 4u * kStep) {
        uint4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = k4[(i + uint32_t(u) * kStep) / KV];
#pragma unroll
        for (int u = 0; u < 4; ++u) count_vec(q[u]);
    }
    for (; i < hiv; i += kStep) count_vec(k4[i / KV]);
    for (uint32_t i = hiv + t; i < hi; i += kUpThreads) atomicAdd(&h[((uint32_t(keys[i]) - kb) >> shift) & 255u], 1u);
    __syncthreads();
    if (t < 256) {
        uint32_t c = 0;
#pragma unroll
        for (int q = 0; q < kUpSubs; ++q) c += s_h[q * kUpPitch + t];
        hist[size_t(t) * nb + blockIdx.x] = c;
    }
}

// Block d (256 blocks of 1024 threads): exclusive scan of digit d's counts over the nb <= 1024
// blocks, in place; the digit's total goes to digit_total[d].
__global__ __launch_bounds__(1024) void k_radix_scan(uint32_t* hist, int nb, uint32_t* digit_total) {
    __shared__ uint32_t s_w[16];
    const int t = threadIdx.x, lane = lane_id(), w = t >> 6;
    uint32_t* row = hist + size_t(blockIdx.x) * nb;
    const uint32_t v = t < nb ? row[t] : 0u;
    const uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int k = 0; k < 16; ++k) {
        const uint32_t x = s_w[k];
        if (k < w) before += x;
        total += x;
    }
    if (t < nb) row[t] = inc + before - v;
    if (t == 0) digit_total[blockIdx.x] = total;
}

// The final buffers of a depth sort whose top pass may be skipped (role 1 writes there then).
template <class K> struct SortOut {
    K* keys;
    uint32_t* vals;
    uint32_t* aux;
};
template <class K, bool AUX, int BT, int R>
__global__ __launch_bounds__(BT) void k_radix_downsweep(const K* keys_in, const uint32_t* vals_in,
                                                                  const uint32_t* aux_in, K* keys_out,
                                                                  uint32_t* vals_out, uint32_t* aux_out,
                                                                  const uint32_t* n_ptr, uint32_t n_const,
                                                                  uint32_t cap, int shift, int nb,
                                                                  const uint32_t* hist, const uint32_t* digit_total,
                                                                  const uint32_t* krange, int role, SortOut<K> alt) {
    constexpr int NW = BT / 64;
    constexpr int kSortTile = BT * R;
    static_assert(kSortTile * 4 >= BT / 64 * 256 * 8, "the digit-match tables live in s_key");
    // keys are staged as 32-bit words in LDS whatever their width in memory: scattering 16-bit keys
    // to random positions would put two lanes on one 32-bit bank word
    __shared__ __align__(8) uint32_t s_key[kSortTile];
    __shared__ uint32_t s_val[kSortTile];
    __shared__ uint32_t s_aux[AUX ? kSortTile : 1];
    __shared__ uint32_t s_cnt[NW][256]; // per-wave running digit counts, then per-wave offsets
    __shared__ uint32_t s_start[256];   // digit start inside the reordered tile
    __shared__ uint32_t s_shift[256];   // global position of the digit's run minus its start in the tile
    __shared__ uint32_t s_w[NW];
    const int t = threadIdx.x, lane = lane_id(), w = t >> 6;
    const bool digit_thread = t < 256; // thread t < 256 owns digit t in the per-tile bookkeeping
    if (role != 0) { // uniform over the grid
        const bool top = top_pass_needed(krange);
        if (role == 2 && !top) return;
        if (role == 1 && !top) {
            keys_out = alt.keys;
            vals_out = alt.vals;
            aux_out = alt.aux;
        }
    }
    const uint32_t kb = key_base(krange);
    // global digit base of this block = exclusive digit prefix + this block's offset in the digit
    uint32_t run = 0;
    {
        const uint32_t tot = digit_thread ? digit_total[t] : 0u;
        uint32_t all;
        const uint32_t inc = block_inclusive_scan<NW>(tot, s_w, all);
        if (digit_thread) run = inc - tot + hist[size_t(t) * nb + blockIdx.x];
    }
    const uint32_t n = load_count(n_ptr, n_const, cap);
    uint32_t lo, hi;
    chunk_of(n, nb, blockIdx.x, lo, hi);
    const uint64_t lt = __lanemask_lt();
    // The next tile's keys are loaded into the same registers as soon as the current tile has
    // been reordered into LDS, so its global loads overlap the current tile's write-out. (Two
    // register sets, with the next tile's loads issued before the current tile is ranked, made
    // both sorts slower on the bench step: the block is bound by its LDS work, not by HBM idling.)
    uint32_t key[R], val[R], aux[R], rank[R];
    auto load_tile = [&](uint32_t b) {
        const uint32_t wb = b + uint32_t(w) * (kSortTile / NW);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t i = wb + uint32_t(r) * 64u + uint32_t(lane);
            const bool valid = i < hi;
            key[r] = valid ? uint32_t(keys_in[i]) : 0u;
            val[r] = valid ? vals_in[i] : 0u;
            if constexpr (AUX) aux[r] = valid ? aux_in[i] : 0u;
        }
    };
    if (lo < hi) load_tile(lo);
    // per-wave digit-match tables (256 lane masks), in s_key's storage while the tile is ranked
    uint64_t* s_tab = reinterpret_cast<uint64_t*>(s_key) + w * 256;
    for (uint32_t base = lo; base < hi; base += kSortTile) {
        for (int k = t; k < NW * 256; k += BT) {
            (&s_cnt[0][0])[k] = 0;
            reinterpret_cast<uint64_t*>(s_key)[k] = 0ull;
        }
        __syncthreads();
        const uint32_t wbase = base + uint32_t(w) * (kSortTile / NW);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t i = wbase + uint32_t(r) * 64u + uint32_t(lane);
            const bool valid = i < hi;
            const uint32_t d = ((key[r] - kb) >> shift) & 255u;
            // lanes of this row holding digit d: each lane ORs its bit into the wave's table entry,
            // reads the entry back and clears it (the wave's LDS operations complete in order;
            // OR is order-free, so the mask and the ranks below do not depend on atomic order)
            uint64_t peers = 0;
            if (valid) {
                atomicOr(reinterpret_cast<unsigned long long*>(&s_tab[d]), 1ull << lane);
                peers = s_tab[d];
                s_tab[d] = 0ull;
            }
            const uint32_t before = valid ? s_cnt[w][d] : 0u;
            rank[r] = before + uint32_t(__popcll(peers & lt));
            // the highest lane of each digit group advances the wave's running count
            if (valid && (peers >> lane) == 1ull) s_cnt[w][d] = before + uint32_t(__popcll(peers));
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        // thread t = digit t: per-wave offsets inside the digit, digit starts inside the tile
        {
            uint32_t s = 0;
            if (digit_thread) {
                for (int k = 0; k < NW; ++k) {
                    const uint32_t c = s_cnt[k][t];
                    s_cnt[k][t] = s;
                    s += c;
                }
            }
            uint32_t all;
            const uint32_t inc = block_inclusive_scan<NW>(s, s_w, all);
            if (digit_thread) {
                s_start[t] = inc - s;
                s_shift[t] = run - (inc - s); // modulo 2^32: + a position inside the run = its destination
                run += s;
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t i = wbase + uint32_t(r) * 64u + uint32_t(lane);
            if (i < hi) {
                const uint32_t d = ((key[r] - kb) >> shift) & 255u;
                const uint32_t p = s_start[d] + s_cnt[w][d] + rank[r];
                s_key[p] = key[r];
                s_val[p] = val[r];
                if constexpr (AUX) s_aux[p] = aux[r];
            }
        }
        if (base + kSortTile < hi) load_tile(base + kSortTile);
        __syncthreads();
        const uint32_t tile_n = hi - base < uint32_t(kSortTile) ? hi - base : uint32_t(kSortTile);
        for (uint32_t j = t; j < tile_n; j += BT) {
            const uint32_t k = s_key[j];
            const uint32_t d = ((k - kb) >> shift) & 255u;
            const uint32_t dst = s_shift[d] + j;
            if (dst >= n) continue; // unreachable for consistent ranks; keeps every store in bounds
            keys_out[dst] = K(k);
            vals_out[dst] = s_val[j];
            if constexpr (AUX) aux_out[dst] = s_aux[j];
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

void launch_compact_visible(const uint32_t* flags, const uint2* meta, uint32_t n_slots, uint32_t* keys,
                            uint32_t* vals, uint32_t* tiles, uint32_t* block_scratch, uint32_t* visible_out,
                            hipStream_t st) {
    const int nb = blocks_for(n_slots);
    SrcFlags src{flags};
    hipLaunchKernelGGL(k_scan_reduce<SrcFlags>, dim3(nb), dim3(kScanThreads), 0, st, src, nullptr, n_slots,
                       n_slots, nb, block_scratch);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, block_scratch, nb, visible_out, nullptr,
                       0xffffffffu);
    SinkCompact sink{meta, keys, vals, tiles};
    hipLaunchKernelGGL((k_scan_write<SrcFlags, SinkCompact>), dim3(nb), dim3(kScanThreads), 0, st, src, sink,
                       nullptr, n_slots, n_slots, nb, block_scratch);
}

void launch_tile_offsets(const uint32_t* sorted_tiles, const uint32_t* visible_ptr, uint32_t slot_cap,
                         uint32_t* offsets, uint32_t* block_scratch, uint32_t* instances_out,
                         uint32_t* overflow_out, uint32_t inst_cap, hipStream_t st) {
    const int nb = blocks_for(slot_cap);
    SrcTiles src{sorted_tiles};
    hipLaunchKernelGGL(k_scan_reduce<SrcTiles>, dim3(nb), dim3(kScanThreads), 0, st, src, visible_ptr, 0u,
                       slot_cap, nb, block_scratch);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, block_scratch, nb, instances_out,
                       overflow_out, inst_cap);
    SinkOffsets sink{offsets};
    hipLaunchKernelGGL((k_scan_write<SrcTiles, SinkOffsets>), dim3(nb), dim3(kScanThreads), 0, st, src, sink,
                       visible_ptr, 0u, slot_cap, nb, block_scratch);
}

int launch_radix_sort(uint32_t* keys[2], uint32_t* vals[2], uint32_t* aux[2], const uint32_t* n_ptr, uint32_t cap,
                      int bits, uint32_t* hist, uint32_t* digit_total, hipStream_t st, bool narrow) {
    // 1024-thread blocks, one resident per CU: at most 256 chunks. (Measured on the bench step:
    // two 512-thread blocks per CU with 8192- or 4096-key tiles, or 256-thread blocks, made both
    // sorts slower; the per-tile bookkeeping favours the largest tile that fits LDS.)
    const int nb = blocks_for(cap) < 256 ? blocks_for(cap) : 256;
    int cur = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        if (narrow) {
            uint16_t* kin = reinterpret_cast<uint16_t*>(keys[cur]);
            uint16_t* kout = reinterpret_cast<uint16_t*>(keys[cur ^ 1]);
            hipLaunchKernelGGL(k_radix_upsweep<uint16_t>, dim3(nb), dim3(kUpThreads), 0, st, kin, n_ptr, 0u, cap,
                               shift, nb, hist, nullptr, 0);
            hipLaunchKernelGGL(k_radix_scan, dim3(256), dim3(1024), 0, st, hist, nb, digit_total);
            hipLaunchKernelGGL((k_radix_downsweep<uint16_t, false, 1024, 16>), dim3(nb), dim3(1024), 0, st, kin,
                               vals[cur], nullptr, kout, vals[cur ^ 1], nullptr, n_ptr, 0u, cap, shift, nb, hist,
                               digit_total, nullptr, 0, SortOut<uint16_t>{nullptr, nullptr, nullptr});
        } else {
            hipLaunchKernelGGL(k_radix_upsweep<uint32_t>, dim3(nb), dim3(kUpThreads), 0, st, keys[cur], n_ptr, 0u,
                               cap, shift, nb, hist, nullptr, 0);
            hipLaunchKernelGGL(k_radix_scan, dim3(256), dim3(1024), 0, st, hist, nb, digit_total);
            if (aux)
                hipLaunchKernelGGL((k_radix_downsweep<uint32_t, true, 1024, 8>), dim3(nb), dim3(1024), 0, st,
                                   keys[cur], vals[cur], aux[cur], keys[cur ^ 1], vals[cur ^ 1], aux[cur ^ 1], n_ptr,
                                   0u, cap, shift, nb, hist, digit_total, nullptr, 0,
                                   SortOut<uint32_t>{nullptr, nullptr, nullptr});
            else
                hipLaunchKernelGGL((k_radix_downsweep<uint32_t, false, 1024, 16>), dim3(nb), dim3(1024), 0, st,
                                   keys[cur], vals[cur], nullptr, keys[cur ^ 1], vals[cur ^ 1], nullptr, n_ptr, 0u,
                                   cap, shift, nb, hist, digit_total, nullptr, 0,
                                   SortOut<uint32_t>{nullptr, nullptr, nullptr});
        }
        cur ^= 1;
    }
    return cur;
}

void launch_depth_sort(uint32_t* keys[3], uint32_t* vals[3], uint32_t* aux[3], const uint32_t* n_ptr, uint32_t cap,
                       const uint32_t* krange, uint32_t* hist, uint32_t* digit_total, hipStream_t st) {
    const int nb = blocks_for(cap) < 256 ? blocks_for(cap) : 256;
    // buffers 0 -> 1 -> 0 -> 1 -> 2, or 0 -> 1 -> 0 -> 2 when the top pass is not needed
    const int src[4] = {0, 1, 0, 1}, dst[4] = {1, 0, 1, 2}, role[4] = {0, 0, 1, 2};
    for (int p = 0; p < 4; ++p) {
        const int a = src[p], b = dst[p];
        hipLaunchKernelGGL(k_radix_upsweep<uint32_t>, dim3(nb), dim3(kUpThreads), 0, st, keys[a], n_ptr, 0u, cap,
                           8 * p, nb, hist, krange, role[p]);
        hipLaunchKernelGGL(k_radix_scan, dim3(256), dim3(1024), 0, st, hist, nb, digit_total);
        hipLaunchKernelGGL((k_radix_downsweep<uint32_t, true, 1024, 8>), dim3(nb), dim3(1024), 0, st, keys[a],
                           vals[a], aux[a], keys[b], vals[b], aux[b], n_ptr, 0u, cap, 8 * p, nb, hist, digit_total,
                           krange, role[p], SortOut<uint32_t>{keys[2], vals[2], aux[2]});
    }
}

} // namespace pht
