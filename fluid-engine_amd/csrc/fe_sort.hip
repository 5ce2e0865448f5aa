// fe_sort.hip — the sort stage of the neighbour search on CDNA4.
//
// Radix: a stable LSD sort of (hash, index) pairs with 8-bit digits in the one-sweep form:
//   hist     one pass over the keys counts the digits of every pass at once (LDS counts per
//            block, then one atomic per digit and block into the pass's global counts)
//   scatter  one launch per pass. Each block takes the next tile ticket, counts the digits of
//            its 4096 keys and publishes the counts, ranks the keys by wave64 ballots (a
//            64-lane match on the 8 digit bits), and finds the counts of all earlier tiles by a
//            two-level decoupled look-back: the earlier tiles of its group of 32 tiles are read
//            at once, and the
//            earlier groups through the group words that each group's last tile publishes
//            (its group's count, then the inclusive count of all groups up to it). The tile
//            then reorders through LDS and writes its digit runs contiguously.
// Tickets are handed out in launch order, so every tile a block waits for belongs to a block
// that is already running. At these sizes every tile of a pass is resident at once; the two
// levels keep each wait to a few device round trips instead of a walk over every earlier
// tile. Stability makes the permutation equal to the reference RadixSorter's
// (radix_count.comp / radix_scan.comp / radix_reorder.comp, sorter.rs:325-455), whose result
// is the stable sort by hash.
//
// Bitonic: the compare-exchange network of bitonic_sort.comp:19-43 as driven by
// GpuSorter::execute (sorter.rs:73-115); the stages whose pairs stay inside a 2048-entry
// block run together in LDS. The network is the same, so the (unstable) result is the same.
#include <array>
#include <mutex>
#include <optional>

#include "fe_abi.h"
#include "fe_internal.h"

namespace fe {
namespace {

constexpr int kSortThreads = 256;
constexpr int kWaves = kSortThreads / 64;
constexpr int kHistItems = 16;
constexpr int kMinTile = 4096;                 // the smallest scatter tile above kSmallSort keys
constexpr uint32_t kSmallSort = 1u << 16;      // sorts up to this size scatter in 1024-key tiles
constexpr uint32_t kSmallTile = 1024;
constexpr uint32_t kGroup = 32;                // tiles per look-back group
constexpr uint32_t kGroupBatch = 8;            // group words read per poll
static_assert(kGroup == kSortGroupTiles, "group size");
// Look-back word: bits 31:30 the state (0 nothing yet, 1 a count of its own tile or group,
// 2 an inclusive count from tile or group 0), bits 29:0 the count.
constexpr uint32_t kFlagAgg = 1u << 30, kFlagPre = 2u << 30, kFlagMask = 3u << 30, kCountMask = (1u << 30) - 1u;
// A look-back wait gives up (and raises the error word) after this many polls; a wait only
// ever covers tiles of running blocks, so it ends long before.
constexpr uint32_t kSpinLimit = 1u << 22;

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread over a block of W waves; returns the prefix and
// writes the block total to *total. `tmp` is LDS of W words.
template <int W = kWaves>
__device__ __forceinline__ uint32_t block_exclusive_scan256(uint32_t v, uint32_t* tmp, uint32_t* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) tmp[w] = inc;
    __syncthreads();
    uint32_t wbase = 0, all = 0;
#pragma unroll
    for (int k = 0; k < W; k++) {
        uint32_t t = tmp[k];
        if (k < w) wbase += t;
        all += t;
    }
    __syncthreads();
    *total = all;
    return wbase + inc - v;
}

// Device-coherent (agent scope) word access for the look-back: the tiles of one pass run on
// all XCDs, whose L2 caches are separate.
__device__ __forceinline__ uint32_t lb_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Digit counts of every pass over keys[0, n), each block over chunks of kHt x kHistItems keys
// (all loads of a thread in flight before its LDS counts). kRuns: each thread counts
// kHistItems consecutive keys and adds each run of one digit with a single LDS atomic, which
// suits keys that arrive almost sorted (a hash-ordered state's new cell hashes, where whole
// cells keep their hash); otherwise the threads stride over the chunk, one atomic per key.
// Also clears the look-back words of every pass and resets the tile tickets, ahead of the
// scatter launches that follow on the stream. hist[] must be zero on entry (the previous
// sort's last tile leaves it so).
template <int kHt, bool kRuns>
__global__ __launch_bounds__(kHt) void k_onesweep_hist(const uint32_t* __restrict__ keys, uint32_t n, uint32_t passes,
                                                       uint32_t status_words, uint32_t* __restrict__ hist,
                                                       uint32_t* __restrict__ status, uint32_t* __restrict__ ctrl) {
    constexpr uint32_t kChunk = kHt * kHistItems;
    __shared__ uint32_t h[kSortMaxPasses][256];
    const uint32_t tid = threadIdx.x;
    for (uint32_t q = tid; q < kSortMaxPasses * 256u; q += kHt) (&h[0][0])[q] = 0;
    if (blockIdx.x == 0 && tid < kSortMaxPasses) ctrl[tid] = 0;
    for (uint32_t q = (blockIdx.x * kHt + tid) * 4u; q < status_words; q += gridDim.x * kHt * 4u)
        *reinterpret_cast<uint4*>(status + q) = make_uint4(0u, 0u, 0u, 0u);   // status_words is a multiple of 256
    __syncthreads();
    const uint32_t chunks = (n + kChunk - 1) / kChunk;
    for (uint32_t t = blockIdx.x; t < chunks; t += gridDim.x) {
        uint32_t k[kHistItems];
        if constexpr (kRuns) {
            const uint32_t b = t * kChunk + tid * kHistItems;
            if (b + kHistItems <= n) {
#pragma unroll
                for (int r = 0; r < kHistItems; r += 4) {
                    const uint4 v = *reinterpret_cast<const uint4*>(keys + b + r);   // b is a multiple of 16
                    k[r] = v.x;
                    k[r + 1] = v.y;
                    k[r + 2] = v.z;
                    k[r + 3] = v.w;
                }
            } else {
#pragma unroll
                for (int r = 0; r < kHistItems; r++) k[r] = b + r < n ? keys[b + r] : 0u;
            }
            const uint32_t m = b < n ? (n - b < (uint32_t)kHistItems ? n - b : (uint32_t)kHistItems) : 0u;
            if (m) {
                for (uint32_t p = 0; p < passes; p++) {
                    const uint32_t sh = 8u * p;
                    uint32_t cur = (k[0] >> sh) & 0xFFu, c = 0;
#pragma unroll
                    for (int r = 0; r < kHistItems; r++) {
                        if ((uint32_t)r < m) {
                            const uint32_t d = (k[r] >> sh) & 0xFFu;
                            if (d != cur) {
                                atomicAdd(&h[p][cur], c);
                                cur = d;
                                c = 0;
                            }
                            c++;
                        }
                    }
                    atomicAdd(&h[p][cur], c);
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < kHistItems; r++) {
                const uint32_t i = t * kChunk + r * kHt + tid;
                k[r] = i < n ? keys[i] : 0u;
            }
#pragma unroll
            for (int r = 0; r < kHistItems; r++) {
                if (t * kChunk + r * kHt + tid < n)
                    for (uint32_t p = 0; p < passes; p++) atomicAdd(&h[p][(k[r] >> (8u * p)) & 0xFFu], 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t q = tid; q < passes * 256u; q += kHt) {
        const uint32_t v = (&h[0][0])[q];
        if (v) atomicAdd(hist + q, v);
    }
}

// Count of digit d in the tiles before `tile` (thread d of the tile's block), by the two-level
// look-back, once the tile has published its own count. tw: this pass's tile words of digit d
// (stride 256), gw: its group words.
__device__ __forceinline__ uint32_t lookback(uint32_t tile, uint32_t num_tiles, uint32_t count, uint32_t* tw,
                                             uint32_t* gw, uint32_t* err) {
    const uint32_t g = tile / kGroup, j = tile % kGroup, g0 = g * kGroup;
    // earlier tiles of the group: all words read at once, missing ones polled again
    uint32_t in_group = 0, missing = j ? (0xFFFFFFFFu >> (32u - j)) : 0u, spins = 0;
    while (missing) {
        uint32_t v[kGroup];
#pragma unroll
        for (uint32_t i = 0; i < kGroup; i++) v[i] = (missing >> i) & 1u ? lb_load(tw + (size_t)(g0 + i) * 256u) : 0u;
#pragma unroll
        for (uint32_t i = 0; i < kGroup; i++) {
            if (v[i] & kFlagMask) {
                in_group += v[i] & kCountMask;
                missing &= ~(1u << i);
            }
        }
        if (missing) {
            if (++spins > kSpinLimit) {
                atomicOr(err, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    const bool last_of_group = j == kGroup - 1;
    if (last_of_group) lb_store(gw + (size_t)g * 256u, kFlagAgg | (in_group + count));
    // earlier groups: walk back over their words, kGroupBatch at a time, until an inclusive one
    uint32_t before = 0;
    if (g > 0) {
        int32_t k = (int32_t)g - 1;
        spins = 0;
        for (;;) {
            uint32_t v[kGroupBatch];
#pragma unroll
            for (uint32_t i = 0; i < kGroupBatch; i++)
                v[i] = k - (int32_t)i >= 0 ? lb_load(gw + (size_t)(k - (int32_t)i) * 256u) : kFlagPre;
            bool done = false;
            int32_t next = k - (int32_t)kGroupBatch;
#pragma unroll
            for (uint32_t i = 0; i < kGroupBatch; i++) {
                if (done || next != k - (int32_t)kGroupBatch) continue;
                const uint32_t f = v[i] & kFlagMask;
                if (f == 0u) {
                    next = k - (int32_t)i;   // poll again from this group
                } else {
                    before += v[i] & kCountMask;
                    if (f == kFlagPre) done = true;
                }
            }
            if (done) break;
            if (next != k - (int32_t)kGroupBatch) {
                if (++spins > kSpinLimit) {
                    atomicOr(err, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            k = next;
            if (k < 0) break;   // unreachable: group 0's word is inclusive
        }
    }
    if (last_of_group) lb_store(gw + (size_t)g * 256u, kFlagPre | (before + in_group + count));
    (void)num_tiles;
    return before + in_group;
}

// One pass of the sort: digit `pass` (bits 8 pass .. 8 pass + 7) of keys_in -> keys_out, in
// tiles of kThr x kIt keys; threads 0..255 do the per-digit work.
template <int kThr, int kIt>
__global__ __launch_bounds__(kThr) void k_onesweep_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, uint32_t n, uint32_t pass, uint32_t passes, uint32_t num_tiles,
    uint32_t* __restrict__ hist, uint32_t* __restrict__ tile_words, uint32_t* __restrict__ group_words,
    uint32_t* __restrict__ ctrl) {
    constexpr int kTile = kThr * kIt, kW = kThr / 64, kWaveElems = 64 * kIt;
    __shared__ uint32_t s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];
    __shared__ uint32_t wave_hist[kW][256];
    __shared__ uint32_t tile_start[256];
    __shared__ uint32_t out_base[256];
    __shared__ uint32_t tmp[kW];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t tile_count[256];
    __shared__ uint32_t digit_total[256];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t shift = 8u * pass;
    if (tid == 0) s_tile = atomicAdd(ctrl + pass, 1u);
#pragma unroll
    for (int k = 0; k < 4; k++) wave_hist[w][lane + 64 * k] = 0;
    if (tid < 256) {
        tile_count[tid] = 0;
        // through LDS, so the read has completed before this tile publishes anything
        digit_total[tid] = hist[pass * 256u + tid];
    }
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t tile_base = tile * kTile;
    const uint32_t groups = (num_tiles + kGroup - 1) / kGroup;
    uint32_t* tw = tile_words + (size_t)pass * num_tiles * 256u;

    uint32_t key[kIt], val[kIt], rank[kIt];
#pragma unroll
    for (int r = 0; r < kIt; r++) {
        uint32_t i = tile_base + w * kWaveElems + r * 64 + lane;
        bool ok = i < n;
        key[r] = ok ? keys_in[i] : 0u;
        val[r] = ok ? vals_in[i] : 0u;
    }
    // the tile's digit counts, published before the ranking so that later tiles' look-backs
    // find them early
#pragma unroll
    for (int r = 0; r < kIt; r++)
        if (tile_base + w * kWaveElems + r * 64 + lane < n) atomicAdd(&tile_count[(key[r] >> shift) & 0xFFu], 1u);
    __syncthreads();
    if (tid < 256) lb_store(tw + (size_t)tile * 256u + tid, kFlagAgg | tile_count[tid]);

    // rank every key among the equal digits of its wave, in input order
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < kIt; r++) {
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

    // thread = digit: per-wave prefixes, the tile's digit starts, the digit's global start,
    // then the count of this digit in all earlier tiles
    {
        const uint32_t d = tid;
        uint32_t count = 0;
        if (d < 256) {
#pragma unroll
            for (int k = 0; k < kW; k++) {
                uint32_t c = wave_hist[k][d];
                wave_hist[k][d] = count;
                count += c;
            }
        }
        uint32_t tot_unused;
        const uint32_t ts = block_exclusive_scan256<kW>(count, tmp, &tot_unused);
        const uint32_t gb = block_exclusive_scan256<kW>(d < 256 ? digit_total[d] : 0u, tmp, &tot_unused);
        if (d < 256) {
            const uint32_t before = lookback(tile, num_tiles, count, tw + d, group_words + (size_t)pass * groups * 256u + d,
                                             ctrl + kSortMaxPasses);
            tile_start[d] = ts;
            out_base[d] = gb + before;
        }
    }
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kIt; r++) {
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
    for (int k = 0; k < kIt; k++) {
        uint32_t idx = k * kThr + tid;
        if (idx < valid) {
            uint32_t kk = s_keys[idx];
            uint32_t d = (kk >> shift) & 0xFFu;
            uint32_t o = out_base[d] + (idx - tile_start[d]);
            keys_out[o] = kk;
            vals_out[o] = s_vals[idx];
        }
    }

    // The last tile of the last pass clears the digit counts for the next sort. Its look-back
    // saw every tile of its group and the word of every earlier group, and a group's word
    // appears only once all its tiles have published; every tile reads its digit total before
    // publishing.
    if (pass + 1 == passes && tile == num_tiles - 1 && tid < 256) {
        for (uint32_t p = 0; p < passes; p++) hist[p * 256u + tid] = 0u;
    }
}

__global__ __launch_bounds__(kSortThreads) void k_sort_scratch_clear(uint32_t* __restrict__ hist,
                                                                     uint32_t* __restrict__ ctrl) {
    for (uint32_t q = threadIdx.x; q < kSortMaxPasses * 256u; q += kSortThreads) hist[q] = 0u;
    if (threadIdx.x <= kSortMaxPasses) ctrl[threadIdx.x] = 0u;
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

uint32_t sort_max_tiles(uint32_t max_n) {
    const uint32_t small = max_n < kSmallSort ? max_n : kSmallSort;
    const uint32_t ts = (small + kSmallTile - 1) / kSmallTile;
    uint32_t t = (uint32_t)(((uint64_t)max_n + kMinTile - 1) / kMinTile);
    t = t > ts ? t : ts;
    return t ? t : 1u;
}

size_t sort_status_words(uint32_t max_tiles) {
    return (size_t)kSortMaxPasses * (max_tiles + (max_tiles + kGroup - 1) / kGroup) * 256u;
}

hipError_t sort_scratch_alloc(SortScratch* s, uint32_t max_n, hipStream_t stream) {
    s->max_tiles = sort_max_tiles(max_n);
    hipError_t e = hipMalloc(&s->hist, sizeof(uint32_t) * 256 * kSortMaxPasses);
    if (e == hipSuccess) e = hipMalloc(&s->status, sizeof(uint32_t) * sort_status_words(s->max_tiles));
    if (e == hipSuccess) e = hipMalloc(&s->ctrl, sizeof(uint32_t) * (kSortMaxPasses + 1));
    if (e == hipSuccess) e = sort_scratch_clear(*s, stream);
    return e;
}

void sort_scratch_free(SortScratch* s) {
    for (uint32_t** p : {&s->hist, &s->status, &s->ctrl}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
}

hipError_t sort_scratch_clear(const SortScratch& s, hipStream_t stream) {
    hipLaunchKernelGGL(k_sort_scratch_clear, dim3(1), dim3(kSortThreads), 0, stream, s.hist, s.ctrl);
    return hipGetLastError();
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

// A profiling record when the sort runs for a context.
struct SortLaunch {
    std::optional<Launch> L;
    SortLaunch(fe_ctx* c, const char* name) {
        if (c) L.emplace(c, name);
    }
};
}  // namespace

namespace {
// Tile shape of the scatter: measured on MI355X (1M dam-break neighbour search, 2^20 and 2^22
// random-key sorts), 512 x 8 keys is fastest up to 2^21 keys and 512 x 16 above (half the
// tiles of 256 x 16, so half the look-back words; 0.184 against 0.196 ms at 2^22).
struct ScatterCfg {
    uint32_t threads, items;
};
ScatterCfg scatter_cfg(uint32_t n) {
    if (n <= kSmallSort) return ScatterCfg{256, 4};   // more, shorter tiles for a small sort
    return n <= (1u << 21) ? ScatterCfg{512, 8} : ScatterCfg{512, 16};
}
// 256 threads x 16 keys per chunk: a 2^20-key sort counts in 256 blocks, one per CU
constexpr uint32_t kHistThreads = 256;
constexpr uint32_t kHistBlocks = 512;
}  // namespace

hipError_t radix_sort_core(const SortScratch& s, hipStream_t stream, uint32_t* keys[2], uint32_t* vals[2],
                           uint32_t n, uint32_t key_bits, fe_ctx* prof, bool runs) {
    if (n <= 1 || key_bits == 0) return hipSuccess;
    if (n > kSortMaxKeys) return hipErrorInvalidValue;
    const ScatterCfg cfg = scatter_cfg(n);
    const uint32_t tile = cfg.threads * cfg.items;
    const uint32_t tiles = (uint32_t)(((uint64_t)n + tile - 1) / tile);
    if (tiles > s.max_tiles) return hipErrorInvalidValue;
    const uint32_t passes = key_bits >= 8 * kSortMaxPasses ? kSortMaxPasses : (key_bits + 7) / 8;
    const uint32_t hchunk = kHistThreads * kHistItems;
    const uint32_t chunks = (n + hchunk - 1) / hchunk;
    const uint32_t hblocks = chunks < kHistBlocks ? chunks : kHistBlocks;
    const uint32_t groups = (tiles + kGroup - 1) / kGroup;
    // this sort's look-back words: [pass][tile][digit], then [pass][group][digit]
    uint32_t* tile_words = s.status;
    uint32_t* group_words = s.status + (size_t)passes * tiles * 256u;
    const uint32_t words = passes * (tiles + groups) * 256u;
    {
        SortLaunch L(prof, "radix_hist");
        if (runs)
            hipLaunchKernelGGL((k_onesweep_hist<kHistThreads, true>), dim3(hblocks), dim3(kHistThreads), 0, stream,
                               keys[0], n, passes, words, s.hist, s.status, s.ctrl);
        else
            hipLaunchKernelGGL((k_onesweep_hist<kHistThreads, false>), dim3(hblocks), dim3(kHistThreads), 0, stream,
                               keys[0], n, passes, words, s.hist, s.status, s.ctrl);
    }
    for (uint32_t p = 0; p < passes; p++) {
        {
            SortLaunch L(prof, "radix_scatter");
#define FE_SCATTER(T, I)                                                                                      \
    hipLaunchKernelGGL((k_onesweep_scatter<T, I>), dim3(tiles), dim3(T), 0, stream, keys[0], vals[0], keys[1], \
                       vals[1], n, p, passes, tiles, s.hist, tile_words, group_words, s.ctrl)
            if (cfg.items == 4) FE_SCATTER(256, 4);
            else if (cfg.items == 8) FE_SCATTER(512, 8);
            else FE_SCATTER(512, 16);
#undef FE_SCATTER
        }
        std::swap(keys[0], keys[1]);
        std::swap(vals[0], vals[1]);
    }
    return hipGetLastError();
}

hipError_t radix_sort_pairs(fe_ctx* c, uint32_t* keys[2], uint32_t* vals[2], uint32_t n, uint32_t key_bits,
                            bool runs) {
    if (n <= 1 || key_bits == 0) return hipSuccess;
    hipError_t e = radix_sort_core(c->sort, c->stream, keys, vals, n, key_bits, c, runs);
    if (e != hipSuccess) return e;
    const uint32_t passes = key_bits >= 8 * kSortMaxPasses ? kSortMaxPasses : (key_bits + 7) / 8;
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

hipError_t radix_sort(fe_ctx* c, uint32_t n, uint32_t key_bits, bool runs) {
    return radix_sort_pairs(c, c->keys, c->vals, n, key_bits, runs);
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

// ---- the sort without allocation (fe_sort_pairs_noalloc) --------------------------------
// The same radix sort on caller-owned entries, with every buffer beyond them in one fixed
// device array per GPU (the split keys and values with their ping-pong partners, the digit
// counts, tickets and look-back words of a sort of FE_SORT_NOALLOC_MAX_PAIRS pairs).
namespace {
constexpr uint32_t kNaPairs = FE_SORT_NOALLOC_MAX_PAIRS;
constexpr uint32_t kNaTiles = kNaPairs / kMinTile;
constexpr size_t kNaHist = 4 * (size_t)kNaPairs;                 // word offsets in the array
constexpr size_t kNaCtrl = kNaHist + kSortMaxPasses * 256u;
constexpr size_t kNaStatus = kNaCtrl + 64;
constexpr size_t kNaWords = kNaStatus + (size_t)kSortMaxPasses * (kNaTiles + (kNaTiles + kGroup - 1) / kGroup) * 256u;
static_assert(kNaPairs % kMinTile == 0, "whole tiles");
static_assert(kNaWords * 4 == FE_SORT_NOALLOC_SCRATCH_BYTES, "documented scratch size");
}  // namespace

static __device__ __align__(256) uint32_t scratch[kNaWords];

namespace {
constexpr int kNaMaxDevices = 64;
std::mutex g_na_mutex;
std::array<uint32_t*, kNaMaxDevices> g_na_ptr{};   // device address of `scratch` per device

__global__ __launch_bounds__(256) void k_split_pairs(const uint2* __restrict__ e, uint32_t* __restrict__ keys,
                                                     uint32_t* __restrict__ vals, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const uint2 v = e[i];
        keys[i] = v.x;
        vals[i] = v.y;
    }
}

__global__ __launch_bounds__(256) void k_join_pairs(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                    uint2* __restrict__ e, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) e[i] = make_uint2(keys[i], vals[i]);
}
}  // namespace

}  // namespace fe

extern "C" fe_status fe_sort_pairs_noalloc(int device, uint32_t* entries, uint32_t n, void* stream) {
    using namespace fe;
    if (n > kNaPairs)
        return fail(FE_ERR_INVALID_ARGUMENT, "fe_sort_pairs_noalloc: n = " + std::to_string(n) +
                                                 " exceeds FE_SORT_NOALLOC_MAX_PAIRS (" + std::to_string(kNaPairs) + ")");
    if (!entries && n > 0) return fail(FE_ERR_INVALID_ARGUMENT, "fe_sort_pairs_noalloc: null entries");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(FE_ERR_DEVICE, "no HIP device");
    if (device < 0 && hipGetDevice(&device) != hipSuccess) return fail(FE_ERR_DEVICE, "hipGetDevice failed");
    if (device >= count || device >= kNaMaxDevices)
        return fail(FE_ERR_INVALID_ARGUMENT, "fe_sort_pairs_noalloc: bad device " + std::to_string(device));
    DeviceGuard g(device);
    uint32_t* base;
    {
        std::lock_guard<std::mutex> lock(g_na_mutex);
        if (!g_na_ptr[device]) {
            // once per device, with that device current: its copy of the array
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
            g_na_ptr[device] = static_cast<uint32_t*>(p);
        }
        base = g_na_ptr[device];
    }
    if (n <= 1) return FE_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    SortScratch s;
    s.hist = base + kNaHist;
    s.ctrl = base + kNaCtrl;
    s.status = base + kNaStatus;
    s.max_tiles = kNaTiles;
    uint32_t* keys[2] = {base, base + 2 * (size_t)kNaPairs};
    uint32_t* vals[2] = {base + kNaPairs, base + 3 * (size_t)kNaPairs};
    // the array keeps the previous call's contents: the counts and tickets start from zero
    FE_HIP(sort_scratch_clear(s, st), "fe_sort_pairs_noalloc clear");
    const uint32_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_split_pairs, dim3(blocks), dim3(256), 0, st, reinterpret_cast<const uint2*>(entries), keys[0],
                       vals[0], n);
    FE_HIP(hipGetLastError(), "fe_sort_pairs_noalloc split");
    FE_HIP(radix_sort_core(s, st, keys, vals, n, 32, nullptr, false), "fe_sort_pairs_noalloc sort");
    hipLaunchKernelGGL(k_join_pairs, dim3(blocks), dim3(256), 0, st, keys[0], vals[0], reinterpret_cast<uint2*>(entries),
                       n);
    FE_HIP(hipGetLastError(), "fe_sort_pairs_noalloc join");
    return FE_OK;
}

namespace fe {
}  // namespace fe
