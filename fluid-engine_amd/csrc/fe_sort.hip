// fe_sort.hip — the sort stage of the neighbour search on CDNA4.
//
// Radix: a stable LSD sort of (hash, index) pairs with 8-bit digits, three kernels a pass:
//   hist    per-tile digit counts, stored digit-major [digit][tile]
//   scan    one block per digit: exclusive scan of its row, plus the digit total
//   scatter wave64 ballot ranking (a 64-lane match on the 8 digit bits), tile-local
//           reorder through LDS, then digit runs written out contiguously
// Stability makes the permutation equal to the reference RadixSorter's
// (radix_count.comp / radix_scan.comp / radix_reorder.comp, sorter.rs:325-455), whose
// result is the stable sort by hash.
//
// Bitonic: the compare-exchange network of bitonic_sort.comp:19-43 as driven by
// GpuSorter::execute (sorter.rs:73-115); the stages whose pairs stay inside a 2048-entry
// block run together in LDS. The network is the same, so the (unstable) result is the same.
#include "fe_internal.h"

namespace fe {
namespace {

constexpr int kSortThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kSortThreads * kItems;  // 2048 pairs per tile
constexpr int kWaves = kSortThreads / 64;
constexpr int kWaveElems = kTile / kWaves;     // 512

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread over a 256-thread block; returns the prefix and
// writes the block total to *total. `tmp` is LDS of kWaves words.
__device__ __forceinline__ uint32_t block_exclusive_scan256(uint32_t v, uint32_t* tmp, uint32_t* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) tmp[w] = inc;
    __syncthreads();
    uint32_t wbase = 0, all = 0;
#pragma unroll
    for (int k = 0; k < kWaves; k++) {
        uint32_t t = tmp[k];
        if (k < w) wbase += t;
        all += t;
    }
    __syncthreads();
    *total = all;
    return wbase + inc - v;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                             uint32_t shift, uint32_t num_tiles,
                                                             uint32_t* __restrict__ counts) {
    __shared__ uint32_t hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kTile;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        uint32_t i = base + k * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&hist[(keys[i] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    counts[threadIdx.x * num_tiles + blockIdx.x] = hist[threadIdx.x];
}

// One block per digit: exclusive scan of counts[d][0..num_tiles), total into totals[d].
__global__ __launch_bounds__(kSortThreads) void k_radix_scan_rows(uint32_t* __restrict__ counts,
                                                                  uint32_t num_tiles,
                                                                  uint32_t* __restrict__ totals) {
    __shared__ uint32_t tmp[kWaves];
    uint32_t* row = counts + (size_t)blockIdx.x * num_tiles;
    const uint32_t per = (num_tiles + kSortThreads - 1) / kSortThreads;
    const uint32_t beg = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < per; k++) {
        uint32_t t = beg + k;
        if (t < num_tiles) sum += row[t];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan256(sum, tmp, &total);
    for (uint32_t k = 0; k < per; k++) {
        uint32_t t = beg + k;
        if (t < num_tiles) {
            uint32_t v = row[t];
            row[t] = run;
            run += v;
        }
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, uint32_t n, uint32_t shift, uint32_t num_tiles,
    const uint32_t* __restrict__ counts, const uint32_t* __restrict__ totals) {
    __shared__ uint32_t s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];
    __shared__ uint32_t wave_hist[kWaves][256];
    __shared__ uint32_t tile_start[256];
    __shared__ uint32_t out_base[256];
    __shared__ uint32_t tmp[kWaves];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t tile_base = blockIdx.x * kTile;
#pragma unroll
    for (int k = 0; k < 4; k++) wave_hist[w][lane + 64 * k] = 0;

    uint32_t key[kItems], val[kItems], rank[kItems];
#pragma unroll
    for (int r = 0; r < kItems; r++) {
        uint32_t i = tile_base + w * kWaveElems + r * 64 + lane;
        bool ok = i < n;
        key[r] = ok ? keys_in[i] : 0u;
        val[r] = ok ? vals_in[i] : 0u;
    }
    __builtin_amdgcn_wave_barrier();
    __syncthreads();

    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < kItems; r++) {
        uint32_t i = tile_base + w * kWaveElems + r * 64 + lane;
        bool ok = i < n;
        uint32_t d = (key[r] >> shift) & 0xFFu;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t before = (uint32_t)__popcll(peers & lt_mask);
        uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t base = ok ? wave_hist[w][d] : 0u;
        __builtin_amdgcn_wave_barrier();
        rank[r] = base + before;
        if (ok && before == 0) wave_hist[w][d] = base + cnt;
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // thread = digit: per-wave prefixes, tile-local digit starts, global run starts
    {
        const uint32_t d = tid;
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < kWaves; k++) {
            uint32_t c = wave_hist[k][d];
            wave_hist[k][d] = sum;
            sum += c;
        }
        uint32_t tot_unused;
        uint32_t ts = block_exclusive_scan256(sum, tmp, &tot_unused);
        uint32_t db = block_exclusive_scan256(totals[d], tmp, &tot_unused);
        tile_start[d] = ts;
        out_base[d] = db + counts[d * num_tiles + blockIdx.x];
    }
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kItems; r++) {
        uint32_t i = tile_base + w * kWaveElems + r * 64 + lane;
        if (i < n) {
            uint32_t d = (key[r] >> shift) & 0xFFu;
            uint32_t p = tile_start[d] + wave_hist[w][d] + rank[r];
            s_keys[p] = key[r];
            s_vals[p] = val[r];
        }
    }
    __syncthreads();

    const uint32_t valid = (n - tile_base) < (uint32_t)kTile ? (n - tile_base) : (uint32_t)kTile;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        uint32_t idx = k * kSortThreads + tid;
        if (idx < valid) {
            uint32_t kk = s_keys[idx];
            uint32_t d = (kk >> shift) & 0xFFu;
            uint32_t o = out_base[d] + (idx - tile_start[d]);
            keys_out[o] = kk;
            vals_out[o] = s_vals[idx];
        }
    }
}

// ---- bitonic --------------------------------------------------------------------------

constexpr int kBitonicLocal = 2048;   // entries per block in the LDS stages
constexpr int kBitonicThreads = 256;

__device__ __forceinline__ void cmp_swap(uint32_t& kl, uint32_t& vl, uint32_t& kr, uint32_t& vr, bool asc) {
    if ((asc && kl > kr) || (!asc && kl < kr)) {
        uint32_t tk = kl, tv = vl;
        kl = kr; vl = vr;
        kr = tk; vr = tv;
    }
}

// One stage (h, step) over all entries: bitonic_sort.comp:19-43.
__global__ __launch_bounds__(kBitonicThreads) void k_bitonic_global(uint32_t* __restrict__ keys,
                                                                    uint32_t* __restrict__ vals,
                                                                    uint32_t num, uint32_t h, uint32_t step) {
    uint32_t i = blockIdx.x * kBitonicThreads + threadIdx.x;
    if (i >= num / 2) return;
    uint32_t l = (i / step) * step * 2 + (i % step);
    uint32_t r = l + step;
    uint32_t kl = keys[l], kr = keys[r], vl = vals[l], vr = vals[r];
    bool asc = (l & h) == 0u;
    if ((asc && kl > kr) || (!asc && kl < kr)) {
        keys[l] = kr; vals[l] = vr;
        keys[r] = kl; vals[r] = vl;
    }
}

// Stages (h, step) for h in [h_lo, h_hi] and step from `step_first` (for h_lo, else h/2)
// down to 1, all with 2*step <= `span` entries, on one aligned block of `span` entries.
__global__ __launch_bounds__(kBitonicThreads) void k_bitonic_local(uint32_t* __restrict__ keys,
                                                                   uint32_t* __restrict__ vals, uint32_t span,
                                                                   uint32_t h_lo, uint32_t h_hi,
                                                                   uint32_t step_first) {
    __shared__ uint32_t sk[kBitonicLocal];
    __shared__ uint32_t sv[kBitonicLocal];
    const uint32_t base = blockIdx.x * span;
    for (uint32_t e = threadIdx.x; e < span; e += kBitonicThreads) {
        sk[e] = keys[base + e];
        sv[e] = vals[base + e];
    }
    const uint32_t pairs = span / 2;
    for (uint32_t h = h_lo; h <= h_hi; h *= 2) {
        for (uint32_t step = (h == h_lo ? step_first : h / 2); step > 0; step /= 2) {
            __syncthreads();
            for (uint32_t p = threadIdx.x; p < pairs; p += kBitonicThreads) {
                uint32_t gi = blockIdx.x * pairs + p;          // global invocation index
                uint32_t l = (gi / step) * step * 2 + (gi % step);
                uint32_t ll = l - base, rl = ll + step;
                bool asc = (l & h) == 0u;
                cmp_swap(sk[ll], sv[ll], sk[rl], sv[rl], asc);
            }
        }
        if (h == 0x80000000u) break;
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < span; e += kBitonicThreads) {
        keys[base + e] = sk[e];
        vals[base + e] = sv[e];
    }
}

}  // namespace

hipError_t sort_scratch_alloc(SortScratch* s, uint32_t max_n) {
    s->max_tiles = (max_n + kTile - 1) / kTile;
    if (s->max_tiles == 0) s->max_tiles = 1;
    hipError_t e = hipMalloc(&s->counts, sizeof(uint32_t) * 256 * (size_t)s->max_tiles);
    if (e != hipSuccess) return e;
    return hipMalloc(&s->totals, sizeof(uint32_t) * 256);
}

void sort_scratch_free(SortScratch* s) {
    if (s->counts) (void)hipFree(s->counts);
    if (s->totals) (void)hipFree(s->totals);
    s->counts = s->totals = nullptr;
}

namespace {
__global__ __launch_bounds__(256) void k_copy_pairs(const uint32_t* __restrict__ ks, const uint32_t* __restrict__ vs,
                                                    uint32_t* __restrict__ kd, uint32_t* __restrict__ vd, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        kd[i] = ks[i];
        vd[i] = vs[i];
    }
}
}  // namespace

hipError_t radix_sort_pairs(fe_ctx* c, uint32_t* keys[2], uint32_t* vals[2], uint32_t n, uint32_t key_bits,
                            const char* tag) {
    if (n <= 1) return hipSuccess;
    const uint32_t tiles = (n + kTile - 1) / kTile;
    if (tiles > c->sort.max_tiles) return hipErrorInvalidValue;
    const uint32_t passes = (key_bits + 7) / 8;
    (void)tag;
    for (uint32_t p = 0; p < passes; p++) {
        const uint32_t shift = 8 * p;
        {
            Launch L(c, "radix_hist");
            hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kSortThreads), 0, c->stream, keys[0], n, shift, tiles,
                               c->sort.counts);
        }
        {
            Launch L(c, "radix_scan");
            hipLaunchKernelGGL(k_radix_scan_rows, dim3(256), dim3(kSortThreads), 0, c->stream, c->sort.counts, tiles,
                               c->sort.totals);
        }
        {
            Launch L(c, "radix_scatter");
            hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kSortThreads), 0, c->stream, keys[0], vals[0],
                               keys[1], vals[1], n, shift, tiles, c->sort.counts, c->sort.totals);
        }
        std::swap(keys[0], keys[1]);
        std::swap(vals[0], vals[1]);
    }
    if ((passes & 1u) && c->batch_graph) {
        // A frame replayed from a graph uses the buffers of its capture: bring an odd pass
        // count's result back to the buffer the sort started from.
        Launch L(c, "radix_home");
        hipLaunchKernelGGL(k_copy_pairs, dim3((n + 255) / 256), dim3(256), 0, c->stream, keys[0], vals[0], keys[1],
                           vals[1], n);
        std::swap(keys[0], keys[1]);
        std::swap(vals[0], vals[1]);
    }
    return hipGetLastError();
}

hipError_t radix_sort(fe_ctx* c, uint32_t n, uint32_t key_bits) {
    return radix_sort_pairs(c, c->keys, c->vals, n, key_bits, "entries");
}

hipError_t bitonic_sort(fe_ctx* c) {
    const uint32_t num = c->T;
    if (num < 2) return hipSuccess;
    const uint32_t span = num < (uint32_t)kBitonicLocal ? num : (uint32_t)kBitonicLocal;
    const uint32_t blocks_local = num / span;
    const uint32_t blocks_global = (num / 2 + kBitonicThreads - 1) / kBitonicThreads;
    {
        // every stage with h <= span
        Launch L(c, "bitonic_local");
        hipLaunchKernelGGL(k_bitonic_local, dim3(blocks_local), dim3(kBitonicThreads), 0, c->stream, c->keys[0],
                           c->vals[0], span, 2u, span, 1u);
    }
    for (uint64_t h = 2ull * span; h <= num; h *= 2) {
        for (uint64_t step = h / 2; step > span / 2; step /= 2) {
            Launch L(c, "bitonic_global");
            hipLaunchKernelGGL(k_bitonic_global, dim3(blocks_global), dim3(kBitonicThreads), 0, c->stream,
                               c->keys[0], c->vals[0], num, (uint32_t)h, (uint32_t)step);
        }
        Launch L(c, "bitonic_local");
        hipLaunchKernelGGL(k_bitonic_local, dim3(blocks_local), dim3(kBitonicThreads), 0, c->stream, c->keys[0],
                           c->vals[0], span, (uint32_t)h, (uint32_t)h, span / 2);
    }
    return hipGetLastError();
}

}  // namespace fe
