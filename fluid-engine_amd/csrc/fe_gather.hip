// fe_gather.hip — the neighbour gathers of the DFSPH solver on CDNA4, in the reference's
// own arithmetic.
//
// Every gather of the reference walks the 27 buckets around its particle (z outer, x
// inner, each bucket in sorted order, a bucket visited twice when two neighbour cells
// share a hash slot; density_and_alpha.comp:31-45 and the same loop elsewhere) and uses the
// candidates with r^2 <= h^2. After each neighbour search that walk is done once
// (build_list) and every accepted pair is recorded in visit order; density_alpha,
// viscosity, both source terms and each pressure_force / pressure_update sweep then
// visit exactly those pairs in exactly that order — a rejected candidate contributes
// nothing in the reference either. With IEEE division and square root and the GLSL
// operation order, each particle's sums are the reference's sums, bit for bit.
//
// Memory layout for the gathers. The reference's arrays stay in hash-sorted order (they
// are the observable state), but hash order scatters neighbouring cells across memory,
// so a 128-byte line of a scalar field holds particles of four unrelated cells. The work
// is therefore scheduled in Morton order of the particles' cells (work-item t handles
// particle sched[t]) and the data a gather reads from its neighbours is kept in
// schedule-ordered mirrors that the producing stage writes next to its hash-ordered output:
// pos_m (position_b, with p / rho^2 in w), vb_m (velocity_b, with m / rho in w), va_m,
// dens_m, and pa_m / pa2_m (position and pressure acceleration packed for pressure_update). The list holds schedule
// positions in tiles of 16 work-items x 4 slots, so 16 lanes reading four entries each
// read 256 contiguous bytes, a cell group's entries fill whole lines while the group is
// built, and each lane keeps four independent neighbour loads in flight.
// Blocks are remapped so that each XCD (one L2 each) walks a contiguous range of the
// Morton order. A particle with more than kMaxNbr pairs, or a stale list, falls back to
// the bucket walk, which produces the same pairs in the same order.
#include <cmath>
#include <cstring>
#include <type_traits>

#include "fe_device.h"

namespace fe {
namespace {

// Workgroups are dealt round-robin over the 8 XCDs; give every XCD a contiguous range of
// the Morton-ordered work.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t g = gridDim.x, b = blockIdx.x;
    const uint32_t per = g / 8u, rem = g % 8u, x = b % 8u, k = b / 8u;
    return x * per + (x < rem ? x : rem) + k;
}

template <int B>
__device__ __forceinline__ uint32_t work_item() { return xcd_block() * B + threadIdx.x; }

// The 27-bucket walk in reference order over the hash-ordered positions. The nine ranges
// of one z slab are loaded together and candidates are fetched four at a time.
template <class F>
__device__ __forceinline__ void visit27(const uint2* __restrict__ range, const float4* __restrict__ pos,
                                        const Sph& s, float px, float py, float pz, F&& f) {
    const int cx = (int)floorf(px * s.inv_h), cy = (int)floorf(py * s.inv_h), cz = (int)floorf(pz * s.inv_h);
    const uint32_t hx[3] = {(uint32_t)(cx - 1) * kP1, (uint32_t)cx * kP1, (uint32_t)(cx + 1) * kP1};
    const uint32_t hy[3] = {(uint32_t)(cy - 1) * kP2, (uint32_t)cy * kP2, (uint32_t)(cy + 1) * kP2};
    const uint32_t hz[3] = {(uint32_t)(cz - 1) * kP3, (uint32_t)cz * kP3, (uint32_t)(cz + 1) * kP3};
#pragma unroll
    for (int z = 0; z < 3; z++) {
        uint2 rr[9];
#pragma unroll
        for (int y = 0; y < 3; y++)
#pragma unroll
            for (int x = 0; x < 3; x++) rr[y * 3 + x] = range[s.tab0 + ((hx[x] ^ hy[y] ^ hz[z]) & s.mask)];
#pragma unroll
        for (int b = 0; b < 9; b++) {
            if (rr[b].x == kSentinel) continue;
            const uint32_t e = rr[b].y;
            for (uint32_t j = rr[b].x; j < e; j += 4) {
                float4 p[4];
#pragma unroll
                for (int u = 0; u < 4; u++) p[u] = pos[(j + u < e) ? j + u : j];
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (j + u < e) f(j + u, p[u]);
            }
        }
    }
}

// ---- mirrors -----------------------------------------------------------------------------

// After the schedule sort: inverse schedule and the position, velocity_b and velocity_a
// mirrors (velocity_a is not permuted by a sort, so its mirror holds velocity_a at the new
// order's indices, which is what divergence_source reads).
__global__ __launch_bounds__(kBlock) void k_schedule_mirrors(const uint32_t* __restrict__ sched, uint32_t n,
                                                             const float4* __restrict__ pos_b,
                                                             const float4* __restrict__ vel_b,
                                                             const float4* __restrict__ vel_a,
                                                             uint32_t* __restrict__ inv, float4* __restrict__ pos_m,
                                                             float4* __restrict__ vb_m, float4* __restrict__ va_m) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = sched[t];
    inv[i] = t;
    pos_m[t] = pos_b[i];
    vb_m[t] = vel_b[i];
    va_m[t] = vel_a[i];
}

// reorder.comp:44 for the velocities alone, written to velocity_b and the xyz of its mirror
// (whose w keeps the m / rho of the unchanged densities).
__global__ __launch_bounds__(kBlock) void k_reorder_velocity(const uint32_t* __restrict__ sched,
                                                             const uint32_t* __restrict__ vals, uint32_t n,
                                                             const float4* __restrict__ vel_a,
                                                             float4* __restrict__ vel_b, float4* __restrict__ vb_m) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = sched[t];
    const float4 v = vel_a[vals[i]];
    vel_b[i] = v;
    float* m = reinterpret_cast<float*>(vb_m + t);
    m[0] = v.x;
    m[1] = v.y;
    m[2] = v.z;
}

// Every mirror from the hash-ordered fields (after uploads or out-of-sequence calls).
__global__ __launch_bounds__(kBlock) void k_refresh_mirrors(Sph s0, Scenes sc, const uint32_t* __restrict__ sched,
                                                            uint32_t n,
                                                            const float4* __restrict__ pos_b,
                                                            const float4* __restrict__ vel_b,
                                                            const float4* __restrict__ vel_a,
                                                            const float* __restrict__ dens,
                                                            const float* __restrict__ pres,
                                                            const float4* __restrict__ pacc,
                                                            const float* __restrict__ fac, const float* __restrict__ src,
                                                            uint32_t* __restrict__ inv, float4* __restrict__ pos_m,
                                                            float4* __restrict__ vb_m, float4* __restrict__ va_m,
                                                            float* __restrict__ dens_m, float4* __restrict__ pa_m,
                                                            float2* __restrict__ pa2_m, float* __restrict__ fac_m,
                                                            float* __restrict__ src_m, float* __restrict__ pres_m) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = sched[t];
    inv[i] = t;
    const float4 x = pos_b[i];
    const float rho = dens[i], p = pres[i];
    pos_m[t] = make_float4(x.x, x.y, x.z, prho_of(p, rho));
    const float4 vb = vel_b[i];
    vb_m[t] = make_float4(vb.x, vb.y, vb.z, scene_at(sc, s0, i).mass / rho);
    va_m[t] = vel_a[i];
    dens_m[t] = rho;
    const float4 a = pacc[i];
    pa_m[t] = make_float4(x.x, x.y, x.z, a.x);
    pa2_m[t] = make_float2(a.y, a.z);
    fac_m[t] = fac[i];
    src_m[t] = src[i];
    pres_m[t] = p;
}

// ---- neighbour list build -----------------------------------------------------------------

// The tiled list: entry `slot` of work-item t sits at
// ((t / 16) (kMaxNbr / 4) + slot / 4) 64 + (t % 16) 4 + slot % 4, in tiles of 16 work-items x 4
// slots (256 B), so 16 lanes reading 4 slots each touch 256 contiguous bytes.

constexpr uint32_t kCand = 256;   // candidates staged per pass per wave

struct BuildStage {
    float4 pos[kCand];   // candidate positions (hash order position_b)
    uint32_t idx[kCand]; // hash index, then schedule position of each candidate
};
// A test round reads pos[q] for q < kCand + 63, which stays inside the stage.
static_assert(sizeof(BuildStage) >= (kCand + 63) * sizeof(float4), "stage read bound");

__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x - v;
}

// One cell group (a run of consecutive work-items whose particles share a cell): lanes
// 0..26 fetch the 27 bucket ranges in visit order, the wave stages every candidate's
// position and schedule position in LDS, then S = 64 / (group size) lanes per particle
// test S consecutive candidates per round; a ballot numbers the accepted ones so each
// particle's list comes out in candidate (= reference visit) order.
__device__ void build_group(const Sph& s, const uint2* __restrict__ range, const float4* __restrict__ pos_b,
                            const float4* __restrict__ pos_m, const uint32_t* __restrict__ inv,
                            uint32_t* __restrict__ nbr, uint32_t* __restrict__ cnt, BuildStage& st, uint32_t gs,
                            uint32_t ge, int3 gc) {
    const int lane = threadIdx.x & 63;
    uint32_t bst = 0, bcnt = 0;
    if (lane < 27) {
        const int dz = lane / 9 - 1, dy = (lane / 3) % 3 - 1, dx = lane % 3 - 1;
        const uint2 r = range[s.tab0 + cell_hash(gc.x + dx, gc.y + dy, gc.z + dz, s.mask)];
        if (r.x != kSentinel) {
            bst = r.x;
            bcnt = r.y - r.x;
        }
    }
    const uint32_t boff = wave_excl_scan(bcnt);
    // wave-uniform loop bounds (scalar registers)
    const uint32_t C = __builtin_amdgcn_readfirstlane(__shfl(boff + bcnt, 26, 64));
    const uint32_t np = __builtin_amdgcn_readfirstlane(ge - gs);
    uint32_t S = 64;
    while (S > 1 && S * np > 64u) S >>= 1;
    const uint32_t per_chunk = 64u / S;
    const uint64_t segmask = S == 64 ? ~0ull : ((1ull << S) - 1ull);
    const uint32_t sub = (uint32_t)lane & (S - 1u);
    const uint32_t seg_shift = (uint32_t)lane & ~(S - 1u);
    const uint64_t segabs = segmask << seg_shift;   // this lane's particle's lanes
    uint32_t staged = 0xFFFFFFFFu;

    for (uint32_t pb = gs; pb < ge; pb += per_chunk) {
        const uint32_t tp = pb + (uint32_t)lane / S;
        const bool act = tp < ge;
        const float4 pi = pos_m[act ? tp : gs];
        // tp's row, as a byte offset from the tile of pb (tp - pb < 64); slot sl sits
        // (sl / 4) 256 + (sl % 4) 4 bytes from the row
        char* const tiles = reinterpret_cast<char*>(nbr + (size_t)(pb >> 4) * (kMaxNbr / 4) * 64u);
        const uint32_t row = ((((tp >> 4) - (pb >> 4)) * (kMaxNbr / 4)) << 8) + ((tp & 15u) << 4);
        uint32_t m = 0;
        for (uint32_t cb = 0; cb < C; cb += kCand) {
            const uint32_t cn = (C - cb) < kCand ? (C - cb) : kCand;
            if (staged != cb) {
                wave_sync();
                if (bcnt) {
                    const uint32_t lo = boff > cb ? boff : cb;
                    const uint32_t hi = (boff + bcnt) < (cb + cn) ? (boff + bcnt) : (cb + cn);
                    for (uint32_t q = lo; q < hi; q++) st.idx[q - cb] = bst + (q - boff);
                }
                wave_sync();
                for (uint32_t q = lane; q < cn; q += 64) {
                    const uint32_t j = st.idx[q];
                    st.pos[q] = pos_b[j];
                    st.idx[q] = inv[j];
                }
                wave_sync();
                staged = cb;
            }
            for (uint32_t r = 0; r < cn; r += S) {
                // q < kCand + 63: past the chunk's end (q >= cn) a lane reads other words of
                // its wave's stage, and rejects them
                const uint32_t q = r + sub;
                const float4 pj = st.pos[q];
                const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
                const float r2 = rx * rx + ry * ry + rz * rz;
                const bool ok = !((act & (sub < cn - r) ? r2 : __builtin_inff()) > s.hh);
                // the particle's accepted candidates, numbered in lane order
                const uint64_t seg = __builtin_amdgcn_ballot_w64(ok) & segabs;
                if (ok) {
                    uint32_t slot = m + __builtin_amdgcn_mbcnt_hi((uint32_t)(seg >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)seg, 0u));
                    // a particle with more than kMaxNbr pairs is swept by the bucket walk,
                    // which never reads its row: its extra pairs may land on its last slot
                    slot = slot < kMaxNbr ? slot : kMaxNbr - 1u;
                    // (slot / 4) 256 + (slot % 4) 4 = 4 slot + 240 (slot / 4)
                    const uint32_t off = row + (slot << 2) + (slot >> 2) * 240u;
                    *reinterpret_cast<uint32_t*>(tiles + off) = st.idx[q];
                }
                m += (uint32_t)__popcll(seg);
            }
        }
        if (act && sub == 0) cnt[tp] = m;
    }
}

// A wave takes the cell groups that start in its window of kWin consecutive schedule
// positions (each group runs on past the window's end if its cell does). kWin 64 (tune items
// 1): about eight groups one after another per wave; kWin 8 (items 8): about one, so a small
// scene's groups are built side by side instead of in a chain.
template <bool kBatch, int kWaves, int kWin = 64>
__global__ __launch_bounds__(64 * kWaves) void k_build_list(Sph s0, Scenes sc, uint32_t n, const uint2* __restrict__ range,
                                                       const float4* __restrict__ pos_b,
                                                       const float4* __restrict__ pos_m,
                                                       const uint32_t* __restrict__ inv, uint32_t* __restrict__ nbr,
                                                       uint32_t* __restrict__ cnt) {
    static_assert(kWin == 64 || kWin == 8, "window");
    __shared__ BuildStage stage[kWaves];
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t base = (xcd_block() * kWaves + w) * (uint32_t)kWin;
    if (base >= n) return;
    const uint32_t t = base + lane;
    const bool valid = t < n;
    // a group is a run of work-items of one scene whose particles share a cell
    const uint32_t sid = kBatch ? sc.sid[valid ? t : base] : 0u;
    const float inv_h = kBatch ? sc.tab[sid].inv_h : s0.inv_h;
    const float4 p = valid ? pos_m[t] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int3 c = cell_of(p.x, p.y, p.z, inv_h);
    int3 cp = make_int3(__shfl_up(c.x, 1, 64), __shfl_up(c.y, 1, 64), __shfl_up(c.z, 1, 64));
    uint32_t sp = __shfl_up(sid, 1, 64);
    if (lane == 0 && base > 0) {
        const float4 q = pos_m[base - 1];
        sp = kBatch ? sc.sid[base - 1] : 0u;
        cp = cell_of(q.x, q.y, q.z, kBatch ? sc.tab[sp].inv_h : s0.inv_h);
    }
    const bool start = valid && lane < kWin && (t == 0 || c.x != cp.x || c.y != cp.y || c.z != cp.z || sid != sp);
    uint64_t starts = __ballot(start);
    while (starts) {
        const int sl = __builtin_ctzll(starts);
        starts &= starts - 1;
        const uint32_t gs = base + (uint32_t)sl;
        const int3 gc = make_int3(__shfl(c.x, sl, 64), __shfl(c.y, sl, 64), __shfl(c.z, sl, 64));
        const uint32_t gsid = __shfl(sid, sl, 64);
        const Sph s = kBatch ? sc.tab[__builtin_amdgcn_readfirstlane(gsid)] : s0;
        uint32_t ge;
        if (starts) {
            ge = base + (uint32_t)__builtin_ctzll(starts);
        } else {
            // the group runs to the end of the window, and possibly beyond it
            ge = base + (uint32_t)kWin < n ? base + (uint32_t)kWin : n;
            while (ge < n) {
                const uint32_t j = ge + lane;
                const bool vj = j < n;
                const float4 q = vj ? pos_m[j] : make_float4(0.f, 0.f, 0.f, 0.f);
                const int3 cj = cell_of(q.x, q.y, q.z, s.inv_h);
                const uint64_t diff = __ballot(!vj || cj.x != gc.x || cj.y != gc.y || cj.z != gc.z ||
                                               (kBatch && sc.sid[vj ? j : 0] != gsid));
                if (diff) {
                    ge += (uint32_t)__builtin_ctzll(diff);
                    break;
                }
                ge += 64u;
            }
        }
        build_group(s, range, pos_b, pos_m, inv, nbr, cnt, stage[w], gs, ge, gc);
    }
}

// ---- sweep windows ----------------------------------------------------------------------
// A Jacobi sweep reads, for every listed pair, the neighbour's position and one solver
// field. The 64 particles of a window (64 consecutive schedule positions) list ~2100 pairs
// but only ~300 distinct neighbours, which lie in a few dozen runs of consecutive schedule
// positions (a cell's particles are consecutive in the Morton schedule). k_window_tables
// finds that union once per list build: it inserts the window's listed schedule positions
// into an LDS hash set, takes the runs of present positions (a gap of at most G absent
// positions inside a run is staged too), lays the runs out one after another as the
// window's staging slots, and rewrites each list entry as a 16-bit staging slot. A sweep
// then stages the window's slots with coalesced loads into LDS and walks the 16-bit lists
// over LDS. The staged values are the mirrors' own values and each particle still visits
// its pairs in list order, so the sums are unchanged bit for bit. A window whose union
// needs more than kWinMaxK slots or kWinMaxRuns runs, or that holds a particle with more
// than kMaxNbr pairs, is marked kWinFallback and swept from the 32-bit list.

constexpr uint32_t kHashSlots = 1024;   // > 2 kWinMaxK, so probes stay short
constexpr uint32_t kHashEmpty = 0xFFFFFFFFu;

struct WindowStage {
    uint32_t key[kHashSlots];    // schedule position, or kHashEmpty
    uint32_t val[kHashSlots];    // run << 16 | position - run start
    uint32_t rstart[kWinMaxRuns];
    uint32_t rlen[kWinMaxRuns];
    uint32_t roff[kWinMaxRuns];
    unsigned long long mask[kWinRounds];
    uint32_t rbase[kWinRounds];
    uint32_t count, over;
};

__device__ __forceinline__ uint32_t hash_slot(uint32_t x) { return (x * 2654435761u) >> 22; }

__device__ __forceinline__ uint32_t hash_find(const WindowStage& st, uint32_t x) {
    uint32_t h = hash_slot(x);
    for (uint32_t p = 0; p < kHashSlots; p++) {
        const uint32_t k = st.key[h];
        if (k == x) return h;
        if (k == kHashEmpty) return kHashEmpty;
        h = (h + 1u) & (kHashSlots - 1u);
    }
    return kHashEmpty;
}

__device__ __forceinline__ void hash_insert(WindowStage& st, uint32_t x) {
    uint32_t h = hash_slot(x);
    for (uint32_t p = 0; p < kHashSlots; p++) {
        const uint32_t old = atomicCAS(&st.key[h], kHashEmpty, x);
        if (old == kHashEmpty) {
            atomicAdd(&st.count, 1u);
            return;
        }
        if (old == x) return;
        h = (h + 1u) & (kHashSlots - 1u);
    }
    st.over = 1u;
}

// Runs of the hash set with gaps of at most G: run starts in slot order, their lengths and
// staging offsets. Returns the number of staging slots K, or kWinFallback when the runs do
// not fit.
__device__ uint32_t window_runs(WindowStage& st, uint32_t G, uint32_t& nruns_out) {
    const int lane = threadIdx.x & 63;
    uint32_t nruns = 0;
    for (uint32_t s0 = 0; s0 < kHashSlots; s0 += 64) {
        const uint32_t x = st.key[s0 + lane];
        bool start = x != kHashEmpty;
        for (uint32_t d = 1; start && d <= G + 1u; d++)
            if (x >= d && hash_find(st, x - d) != kHashEmpty) start = false;
        const uint64_t b = __ballot(start);
        if (start) {
            const uint32_t rid = nruns + (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
            if (rid < kWinMaxRuns) st.rstart[rid] = x;
        }
        nruns += (uint32_t)__popcll(b);
    }
    nruns_out = nruns;
    if (nruns > kWinMaxRuns) return kWinFallback;
    wave_sync();
    for (uint32_t r = lane; r < nruns; r += 64) {
        const uint32_t x0 = st.rstart[r];
        st.val[hash_find(st, x0)] = r << 16;
        uint32_t last = x0;
        for (;;) {
            uint32_t y = last + 1u, h = kHashEmpty;
            for (; y <= last + G + 1u; y++)
                if ((h = hash_find(st, y)) != kHashEmpty) break;
            if (h == kHashEmpty) break;
            if (y - x0 >= kWinMaxK) {   // longer than the staging area: K exceeds kWinMaxK
                last = x0 + kWinMaxK;
                break;
            }
            st.val[h] = (r << 16) | (y - x0);
            last = y;
        }
        st.rlen[r] = last - x0 + 1u;
    }
    wave_sync();
    uint32_t carry = 0;
    for (uint32_t r0 = 0; r0 < nruns; r0 += 64) {
        const uint32_t r = r0 + lane;
        const uint32_t len = r < nruns ? st.rlen[r] : 0u;
        const uint32_t ex = wave_excl_scan(len);
        if (r < nruns) st.roff[r] = carry + ex;
        carry += __shfl(ex + len, 63, 64);
    }
    wave_sync();
    return carry <= kWinMaxK ? carry : kWinFallback;
}

// The 16-bit list of work-item t: tiles of 16 work-items x 8 slots (uint4 chunk k / 8).
__device__ __forceinline__ uint4* list16_tile(uint16_t* nbr16, uint32_t t) {
    return reinterpret_cast<uint4*>(nbr16) + (size_t)(t >> 4) * (kMaxNbr / 8) * 16 + (t & 15u);
}
__device__ __forceinline__ const uint4* list16_tile(const uint16_t* nbr16, uint32_t t) {
    return reinterpret_cast<const uint4*>(nbr16) + (size_t)(t >> 4) * (kMaxNbr / 8) * 16 + (t & 15u);
}

template <int kWaves>
__global__ __launch_bounds__(64 * kWaves) void k_window_tables(uint32_t n, const uint32_t* __restrict__ nbr,
                                                               const uint32_t* __restrict__ cnt,
                                                               uint32_t* __restrict__ win_meta,
                                                               uint4* __restrict__ win_round,
                                                               uint32_t* __restrict__ win_delta,
                                                               uint16_t* __restrict__ nbr16) {
    __shared__ WindowStage stage[kWaves];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t win = xcd_block() * kWaves + wv;
    if (win * kWin >= n) return;
    WindowStage& st = stage[wv];
    const uint32_t t = win * kWin + lane;
    const bool valid = t < n;
    const uint32_t m = valid ? cnt[t] : 0u;
    if (__any(m > kMaxNbr)) {
        if (lane == 0) win_meta[win] = kWinFallback;
        return;
    }
    for (uint32_t q = lane; q < kHashSlots; q += 64) st.key[q] = kHashEmpty;
    if (lane == 0) {
        st.count = 0;
        st.over = 0;
    }
    wave_sync();
    const uint4* tile = list_tile(nbr, t);
    for (uint32_t k = 0; k < m; k += 4) {
        if (st.count > kWinMaxK) break;   // the window falls back in any case
        const uint4 q4 = tile[(k >> 2) * 16];
#pragma unroll
        for (uint32_t u = 0; u < 4; u++)
            if (k + u < m) hash_insert(st, list_pick(q4, u));
    }
    wave_sync();
    uint32_t K = kWinFallback, nruns = 0;
    if (st.count <= kWinMaxK && !st.over) {
        // gap 4 stages ~30% more slots in half as many runs; narrower gaps when that overflows
        for (uint32_t G : {4u, 1u, 0u}) {
            K = window_runs(st, G, nruns);
            if (K != kWinFallback) break;
        }
    }
    if (K == kWinFallback) {
        if (lane == 0) win_meta[win] = kWinFallback;
        return;
    }
    // staging rounds: the runs that start inside each round, and the run holding its first slot
    if (lane < (int)kWinRounds) st.mask[lane] = 0ull;
    wave_sync();
    for (uint32_t r = lane; r < nruns; r += 64) {
        const uint32_t off = st.roff[r], len = st.rlen[r];
        if (off & 63u) atomicOr(&st.mask[off >> 6], 1ull << (off & 63u));
        for (uint32_t k = (off + 63u) >> 6; k * 64u < off + len; k++) st.rbase[k] = r;
    }
    wave_sync();
    const uint32_t rounds = (K + 63u) >> 6;
    if (lane < (int)rounds) {
        const unsigned long long mk = st.mask[lane];
        win_round[(size_t)win * kWinRounds + lane] =
            make_uint4((uint32_t)mk, (uint32_t)(mk >> 32), st.rbase[lane], 0u);
    }
    for (uint32_t r = lane; r < nruns; r += 64) win_delta[(size_t)win * kWinMaxRuns + r] = st.rstart[r] - st.roff[r];
    if (lane == 0) win_meta[win] = K | (nruns << 16);
    // the list as staging slots (slots past the pair count hold 0)
    if (valid) {
        uint4* out = list16_tile(nbr16, t);
        for (uint32_t k = 0; k < m; k += 8) {
            uint32_t loc[8];
            const uint4 qa = tile[(k >> 2) * 16];
            const uint4 qb = k + 4 < m ? tile[((k >> 2) + 1) * 16] : qa;
#pragma unroll
            for (uint32_t u = 0; u < 8; u++) {
                loc[u] = 0u;
                if (k + u < m) {
                    const uint32_t x = list_pick(u < 4 ? qa : qb, u & 3u);
                    const uint32_t v = st.val[hash_find(st, x)];
                    loc[u] = st.roff[v >> 16] + (v & 0xFFFFu);
                }
            }
            out[(k >> 3) * 16] = make_uint4(loc[0] | (loc[1] << 16), loc[2] | (loc[3] << 16),
                                            loc[4] | (loc[5] << 16), loc[6] | (loc[7] << 16));
        }
    }
}

// ---- list consumers: one policy per reference shader ------------------------------------
// pair() repeats the reference's own tests (r^2 > h^2 rejects, gradient terms need
// r > 1e-6), so the same code serves list entries (which pass them) and the bucket walk.
// The tests select instead of branching: a skipped pair adds +0 or -0. Every sum starts at
// +0 and in round-to-nearest only -0 + -0 gives -0, so a sum is never -0 and adding a zero
// leaves it unchanged bit for bit.
// geo_of() gives the record that holds a neighbour's position (pos_m, or pa_m for
// pressure_update); load(jt, g) reads the rest of the neighbour's data by schedule position
// jt, given that record g, and self(i, t, g) the particle's own.

// The wall samples (FE_BOUNDARY_PARTICLES) enter every gather through two per-particle sums
// over the samples, fixed while the positions are: G = sum psi_b grad W_ib and
// rho_w = sum psi_b W_ib (wall_g = (G, rho_w), k_wall_terms). Because the samples are at
// rest, their velocity-divergence terms are v . G, their A p term a . G and, mirroring the
// pressure, their pressure acceleration (p / rho^2) G; each policy's wall_add adds that once
// after the fluid pairs.
__device__ __forceinline__ float dot_g(float x, float y, float z, const float4& g) {
    return x * g.x + y * g.y + z * g.z;
}

// density_and_alpha.comp:14-72 over every pair with r^2 <= h^2 (the list includes the
// particle itself and any r <= 1e-6 pair, which add to the density only).
struct DensityAlpha {
    float* dens;
    float* fac;
    float* dens_m;
    float* fac_m;
    float4* vb_w;   // the velocity_b mirror, whose w takes m / rho
    struct Self {};
    struct Cand {};
    struct Acc { float rho, sgs, gx, gy, gz; };
    static constexpr bool kPaGeo = false;
    static constexpr bool kWalls = false;
    static constexpr bool kAdditive = true;   // pair() adds one term per Acc component
    __device__ Self self(uint32_t, uint32_t, const float4&) const { return {}; }
    __device__ Cand load(uint32_t, const float4&) const { return {}; }
    __device__ Acc zero() const { return {0.f, 0.f, 0.f, 0.f, 0.f}; }
    // The pair's terms; returns its gradient scalar g (grad W = g r), whether the pair has a
    // gradient term (r^2 <= h^2 and r > 1e-6) and its distance r.
    __device__ __forceinline__ static float terms(Acc& a, const Sph& s, float rx, float ry, float rz, bool& gok,
                                                  float& r) {
        const float r2 = rx * rx + ry * ry + rz * rz;
        const bool in = !(r2 > s.hh);
        r = sqrt_exact(r2);
        const float q = q_of(r, s);
        const float w = w_of_q(q, s);
        a.rho += in ? s.mass * w : 0.0f;
        gok = in && r > 1e-6f;
        const float g = grad_of_q(r, q, s);
        const float mx = s.mass * (g * rx), my = s.mass * (g * ry), mz = s.mass * (g * rz);
        a.gx += gok ? mx : 0.0f;
        a.gy += gok ? my : 0.0f;
        a.gz += gok ? mz : 0.0f;
        a.sgs += gok ? mx * mx + my * my + mz * mz : 0.0f;
        return g;
    }
    __device__ __forceinline__ void pair(Acc& a, const Self&, const Sph& s, float rx, float ry, float rz,
                                         const Cand&) const {
        bool gok;
        float r;
        (void)terms(a, s, rx, ry, rz, gok, r);
    }
    // the samples' density and gradient sum (they do not move: no term in the sum of squares)
    __device__ __forceinline__ void wall_add(Acc& a, const Self&, const Sph&, const float4& g) const {
        a.rho += g.w;
        a.gx += g.x;
        a.gy += g.y;
        a.gz += g.z;
    }
    __device__ unsigned long long finish(uint32_t i, uint32_t t, const Acc& a, const Self&, const Sph& s) const {
        dens[i] = a.rho;
        dens_m[t] = a.rho;
        vb_w[t].w = s.mass / a.rho;
        const float alpha = a.sgs + (a.gx * a.gx + a.gy * a.gy + a.gz * a.gz);
        const float f = alpha > 1e-6f ? alpha : 0.0f;
        fac[i] = f;
        fac_m[t] = f;
        return 0;
    }
};

// viscosity.comp:17-66 (velocity_b in, velocity_a out with w = 0, gravity fused). The
// neighbour's m / rho_j comes in the w of its velocity_b mirror (written with its density,
// the same IEEE quotient the shader forms per pair); W is needed only where r > 1e-6.
struct Viscosity {
    const float4* vb_m;
    float4* vel_a;
    float4* va_m;
    struct Self { float4 v; };
    struct Cand { float4 v; float mrho; };
    struct Acc { float x, y, z; };
    static constexpr bool kPaGeo = false;
    static constexpr bool kWalls = false;
    static constexpr bool kAdditive = true;   // pair() adds one term per Acc component
    __device__ Self self(uint32_t, uint32_t t, const float4&) const { return {vb_m[t]}; }
    __device__ Cand load(uint32_t jt, const float4&) const {
        const float4 v = vb_m[jt];
        return {v, v.w};
    }
    __device__ Acc zero() const { return {0.f, 0.f, 0.f}; }
    __device__ __forceinline__ void pair(Acc& a, const Self& si, const Sph& s, float rx, float ry, float rz,
                                         const Cand& c) const {
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float r = sqrt_grad(r2);
        const bool ok = !(r2 > s.hh) && r > 1e-6f;
        const float w = w_of_q(q_big(r, s), s);
        const float m_rho = c.mrho;
        a.x += ok ? (m_rho * (c.v.x - si.v.x)) * w : 0.0f;
        a.y += ok ? (m_rho * (c.v.y - si.v.y)) * w : 0.0f;
        a.z += ok ? (m_rho * (c.v.z - si.v.z)) * w : 0.0f;
    }
    __device__ unsigned long long finish(uint32_t i, uint32_t t, const Acc& a, const Self& si, const Sph& s) const {
        const float vx = si.v.x + s.nu * a.x, vy = si.v.y + s.nu * a.y, vz = si.v.z + s.nu * a.z;
        const float4 v = make_float4(vx + s.gx * s.dt, vy + s.gy * s.dt, vz + s.gz * s.dt, 0.0f);
        vel_a[i] = v;
        va_m[t] = v;
        return 0;
    }
};

// density_source_term.comp:16-74: velocity_a of the current order; pressures zeroed.
struct DensitySource {
    const float4* va_m;
    const float* dens;
    float* src;
    float* pres;
    float4* pos_w;   // the position mirror, whose w takes p / rho^2
    float* src_m;
    float* pres_m;
    struct Self { float4 v; };
    struct Cand { float4 v; };
    struct Acc { float div; };
    static constexpr bool kPaGeo = false;
    static constexpr bool kWalls = false;
    static constexpr bool kAdditive = true;   // pair() adds one term per Acc component
    __device__ Self self(uint32_t, uint32_t t, const float4&) const { return {va_m[t]}; }
    __device__ Cand load(uint32_t jt, const float4&) const { return {va_m[jt]}; }
    __device__ Acc zero() const { return {0.f}; }
    __device__ __forceinline__ void pair(Acc& a, const Self& si, const Sph& s, float rx, float ry, float rz,
                                         const Cand& c) const {
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float r = sqrt_grad(r2);
        const bool ok = !(r2 > s.hh) && r > 1e-6f;
        const float g = exact_grad(r, s);
        const float dx = si.v.x - c.v.x, dy = si.v.y - c.v.y, dz = si.v.z - c.v.z;
        a.div += ok ? s.mass * (dx * (g * rx) + dy * (g * ry) + dz * (g * rz)) : 0.0f;
    }
    __device__ __forceinline__ void wall_add(Acc& a, const Self& si, const Sph&, const float4& g) const {
        a.div += dot_g(si.v.x, si.v.y, si.v.z, g);
    }
    __device__ unsigned long long finish(uint32_t i, uint32_t t, const Acc& a, const Self&, const Sph& s) const {
        const float rho = dens[i];
        float source = 0.0f;
        if (s.dt > 1e-6f) source = ((s.rho0 - rho) / s.dt) - a.div;
        src[i] = source;
        pres[i] = 0.0f;
        src_m[t] = source;
        pres_m[t] = 0.0f;
        pos_w[t].w = prho_of(0.0f, rho);
        return 0;
    }
};

// divergence_source_term.comp:14-62. velocity_a still holds the previous sort order and is
// read at the current order's indices (as the reference does): its mirror va_m, written
// after the sort as velocity_a[sched[t]], holds exactly those values. Also seeds p / rho^2
// for the divergence solve from the pressures the density solve left (never reset, never
// permuted).
struct DivergenceSource {
    const float4* va_m;
    const float* dens;
    const float* pres;
    float* src;
    float4* pos_w;
    float* src_m;
    float* pres_m;
    struct Self { float4 v; };
    struct Cand { float4 v; };
    struct Acc { float div; };
    static constexpr bool kPaGeo = false;
    static constexpr bool kWalls = false;
    static constexpr bool kAdditive = true;   // pair() adds one term per Acc component
    __device__ Self self(uint32_t, uint32_t t, const float4&) const { return {va_m[t]}; }
    __device__ Cand load(uint32_t jt, const float4&) const { return {va_m[jt]}; }
    __device__ Acc zero() const { return {0.f}; }
    __device__ __forceinline__ void pair(Acc& a, const Self& si, const Sph& s, float rx, float ry, float rz,
                                         const Cand& c) const {
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float r = sqrt_grad(r2);
        const bool ok = !(r2 > s.hh) && r > 1e-6f;
        const float g = exact_grad(r, s);
        const float dx = si.v.x - c.v.x, dy = si.v.y - c.v.y, dz = si.v.z - c.v.z;
        a.div += ok ? s.mass * (dx * (g * rx) + dy * (g * ry) + dz * (g * rz)) : 0.0f;
    }
    __device__ __forceinline__ void wall_add(Acc& a, const Self& si, const Sph&, const float4& g) const {
        a.div += dot_g(si.v.x, si.v.y, si.v.z, g);
    }
    __device__ unsigned long long finish(uint32_t i, uint32_t t, const Acc& a, const Self&, const Sph& s) const {
        src[i] = -a.div;
        src_m[t] = -a.div;
        const float p = pres[i];
        pres_m[t] = p;
        pos_w[t].w = prho_of(p, dens[i]);
        return 0;
    }
};

// density_and_alpha.comp:14-72 and divergence_source_term.comp:14-62 in one pass over the
// pairs of the post-integrate search, which the reference runs back to back: the divergence
// source reads no neighbour density, and its pair term uses the same gradient scalar. Where
// DensityAlpha has a gradient term (gok) its g equals DivergenceSource's exact_grad of the
// same r^2 (sqrt_exact and sqrt_grad agree from 2^-96 on, q_of and q_big from r = 1e-6 on),
// and elsewhere both add nothing, so every sum is the two kernels' own, in list order. The
// particle's own density comes from the accumulator instead of the densities field.
struct DensityAlphaDivergence {
    DensityAlpha da;
    DivergenceSource dv;
    struct Self { float4 v; };
    struct Cand { float4 v; };
    struct Acc { DensityAlpha::Acc a; float div; };
    static constexpr bool kPaGeo = false;
    static constexpr bool kWalls = false;
    static constexpr bool kAdditive = true;   // pair() adds one term per Acc component
    __device__ Self self(uint32_t, uint32_t t, const float4&) const { return {dv.va_m[t]}; }
    __device__ Cand load(uint32_t jt, const float4&) const { return {dv.va_m[jt]}; }
    __device__ Acc zero() const { return {da.zero(), 0.f}; }
    __device__ __forceinline__ void pair(Acc& a, const Self& si, const Sph& s, float rx, float ry, float rz,
                                         const Cand& c) const {
        bool gok;
        float r;
        const float g = DensityAlpha::terms(a.a, s, rx, ry, rz, gok, r);
        const float dx = si.v.x - c.v.x, dy = si.v.y - c.v.y, dz = si.v.z - c.v.z;
        a.div += gok ? s.mass * (dx * (g * rx) + dy * (g * ry) + dz * (g * rz)) : 0.0f;
    }
    // each kernel's wall term, after its fluid pairs
    __device__ __forceinline__ void wall_add(Acc& a, const Self& si, const Sph& s, const float4& g) const {
        da.wall_add(a.a, DensityAlpha::Self{}, s, g);
        a.div += dot_g(si.v.x, si.v.y, si.v.z, g);
    }
    __device__ unsigned long long finish(uint32_t i, uint32_t t, const Acc& a, const Self&, const Sph& s) const {
        (void)da.finish(i, t, a.a, DensityAlpha::Self{}, s);
        dv.src[i] = -a.div;
        dv.src_m[t] = -a.div;
        const float p = dv.pres[i];
        dv.pres_m[t] = p;
        dv.pos_w[t].w = prho_of(p, a.a.rho);
        return 0;
    }
};

// The warm-start features of a solve, gathered in the same pass as its source term (over
// the same listed pairs, no extra neighbour walk): per particle, over its pairs with
// r^2 <= h^2 in visit order (the first kMaxNbr of them; the count itself is not capped),
//   0 (rho - rho0) / rho0        1 s dt / rho0              2 alpha
//   3 pair count                 4 previous pressure        5 sum W
//   6 sum W p_prev_j             7 |sum r_ij| / h           8 sum |r_ij|^2 / h^2
//   9 sum (v_i - v_j).r_ij / h   10 sum |v_i - v_j|^2       11 -s rho^2 / (dt alpha)
// with v the velocity the source term uses, written feature-major: feat[k n + t].
struct FeatAcc {
    uint32_t cnt;
    float sw, swp, cx, cy, cz, sr2, svr, sv2;
};

__device__ __forceinline__ FeatAcc feat_zero() { return {0u, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}; }

// One accepted pair (r2 <= h^2 already checked; r = sqrt(r2), dv = v_i - v_j).
__device__ __forceinline__ void feat_pair(FeatAcc& f, const Sph& s, float rx, float ry, float rz, float r2, float r,
                                          float dx, float dy, float dz, float pprev_j) {
    if (++f.cnt > kMaxNbr) return;
    const float w = spline_w(r, s);
    f.sw += w;
    f.swp += w * pprev_j;
    f.cx += rx;
    f.cy += ry;
    f.cz += rz;
    f.sr2 += r2;
    f.svr += dx * rx + dy * ry + dz * rz;
    f.sv2 += dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ void feat_write(float* __restrict__ feat, uint32_t n, uint32_t t, const Sph& s,
                                           const FeatAcc& f, float rho, float alpha, float src, float pprev_i,
                                           float rho0, float dt) {
    float v[kMlpIn];
    v[0] = (rho - rho0) / rho0;
    v[1] = src * dt / rho0;
    v[2] = alpha;
    v[3] = (float)f.cnt;
    v[4] = pprev_i;
    v[5] = f.sw;
    v[6] = f.swp;
    v[7] = sqrtf(f.cx * f.cx + f.cy * f.cy + f.cz * f.cz) / s.h;
    v[8] = f.sr2 / s.hh;
    v[9] = f.svr / s.h;
    v[10] = f.sv2;
    v[11] = (alpha > 1e-6f && rho > 1e-6f && dt > 1e-6f) ? -src * (rho * rho) / (dt * alpha) : 0.0f;
#pragma unroll
    for (int k = 0; k < kMlpIn; k++) feat[(size_t)k * n + t] = v[k];
}

// DensitySource plus the features of the density solve.
struct DensitySourceFeat {
    DensitySource base;
    const float* fac;
    const float* pprev_m;
    float* feat;
    struct Self { float4 v; float pp; };
    struct Cand { float4 v; float pp; };
    struct Acc { float div; FeatAcc f; };
    static constexpr bool kPaGeo = false;
    static constexpr bool kWalls = false;
    __device__ Self self(uint32_t, uint32_t t, const float4&) const { return {base.va_m[t], pprev_m[t]}; }
    __device__ Cand load(uint32_t jt, const float4&) const { return {base.va_m[jt], pprev_m[jt]}; }
    __device__ Acc zero() const { return {0.f, feat_zero()}; }
    // wall samples enter the source term only (the features describe fluid neighbours)
    __device__ __forceinline__ void wall_add(Acc& a, const Self& si, const Sph&, const float4& g) const {
        a.div += dot_g(si.v.x, si.v.y, si.v.z, g);
    }
    __device__ void pair(Acc& a, const Self& si, const Sph& s, float rx, float ry, float rz, const Cand& c) const {
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        const float r = sqrt_exact(r2);
        const float dx = si.v.x - c.v.x, dy = si.v.y - c.v.y, dz = si.v.z - c.v.z;
        if (r > 1e-6f) {
            const float g = exact_grad(r, s);
            a.div += s.mass * (dx * (g * rx) + dy * (g * ry) + dz * (g * rz));
        }
        feat_pair(a.f, s, rx, ry, rz, r2, r, dx, dy, dz, c.pp);
    }
    __device__ unsigned long long finish(uint32_t i, uint32_t t, const Acc& a, const Self& si, const Sph& s) const {
        const float rho = base.dens[i];
        float source = 0.0f;
        if (s.dt > 1e-6f) source = ((s.rho0 - rho) / s.dt) - a.div;
        base.src[i] = source;
        base.pres[i] = 0.0f;
        base.src_m[t] = source;
        base.pres_m[t] = 0.0f;
        base.pos_w[t].w = prho_of(0.0f, rho);
        feat_write(feat, n_, t, s, a.f, rho, fac[i], source, si.pp, s.rho0, s.dt);
        return 0;
    }
    uint32_t n_;
};

// DivergenceSource plus the features of the divergence solve (velocity_a read at the
// current order's indices, as the source term reads it).
struct DivergenceSourceFeat {
    DivergenceSource base;
    const float* fac;
    const float* pprev_m;
    float* feat;
    struct Self { float4 v; float pp; };
    struct Cand { float4 v; float pp; };
    struct Acc { float div; FeatAcc f; };
    static constexpr bool kPaGeo = false;
    static constexpr bool kWalls = false;
    __device__ Self self(uint32_t, uint32_t t, const float4&) const { return {base.va_m[t], pprev_m[t]}; }
    __device__ Cand load(uint32_t jt, const float4&) const { return {base.va_m[jt], pprev_m[jt]}; }
    __device__ Acc zero() const { return {0.f, feat_zero()}; }
    // wall samples enter the source term only (the features describe fluid neighbours)
    __device__ __forceinline__ void wall_add(Acc& a, const Self& si, const Sph&, const float4& g) const {
        a.div += dot_g(si.v.x, si.v.y, si.v.z, g);
    }
    __device__ void pair(Acc& a, const Self& si, const Sph& s, float rx, float ry, float rz, const Cand& c) const {
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        const float r = sqrt_exact(r2);
        const float dx = si.v.x - c.v.x, dy = si.v.y - c.v.y, dz = si.v.z - c.v.z;
        if (r > 1e-6f) {
            const float g = exact_grad(r, s);
            a.div += s.mass * (dx * (g * rx) + dy * (g * ry) + dz * (g * rz));
        }
        feat_pair(a.f, s, rx, ry, rz, r2, r, dx, dy, dz, c.pp);
    }
    __device__ unsigned long long finish(uint32_t i, uint32_t t, const Acc& a, const Self& si, const Sph& s) const {
        const float source = -a.div;
        base.src[i] = source;
        base.src_m[t] = source;
        const float rho = base.dens[i], p = base.pres[i];
        base.pres_m[t] = p;
        base.pos_w[t].w = prho_of(p, rho);
        feat_write(feat, n_, t, s, a.f, rho, fac[i], source, si.pp, s.rho0, s.dt);
        return 0;
    }
    uint32_t n_;
};

// DensityAlphaDivergence plus the warm-start features of the divergence solve, the fused form
// of DensityAlpha followed by DivergenceSourceFeat: the features take the particle's own
// density and factor, which this pass writes before it forms them.
struct DensityAlphaDivergenceFeat {
    DensityAlpha da;
    DivergenceSource dv;
    const float* pprev_m;
    float* feat;
    uint32_t n_;
    struct Self { float4 v; float pp; };
    struct Cand { float4 v; float pp; };
    struct Acc { DensityAlpha::Acc a; float div; FeatAcc f; };
    static constexpr bool kPaGeo = false;
    static constexpr bool kWalls = false;
    __device__ Self self(uint32_t, uint32_t t, const float4&) const { return {dv.va_m[t], pprev_m[t]}; }
    __device__ Cand load(uint32_t jt, const float4&) const { return {dv.va_m[jt], pprev_m[jt]}; }
    __device__ Acc zero() const { return {da.zero(), 0.f, feat_zero()}; }
    // the samples enter density, factor and source term; the features describe fluid neighbours
    __device__ __forceinline__ void wall_add(Acc& a, const Self& si, const Sph& s, const float4& g) const {
        da.wall_add(a.a, DensityAlpha::Self{}, s, g);
        a.div += dot_g(si.v.x, si.v.y, si.v.z, g);
    }
    __device__ void pair(Acc& a, const Self& si, const Sph& s, float rx, float ry, float rz, const Cand& c) const {
        bool gok;
        float r;
        const float g = DensityAlpha::terms(a.a, s, rx, ry, rz, gok, r);
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        const float dx = si.v.x - c.v.x, dy = si.v.y - c.v.y, dz = si.v.z - c.v.z;
        if (gok) a.div += s.mass * (dx * (g * rx) + dy * (g * ry) + dz * (g * rz));
        feat_pair(a.f, s, rx, ry, rz, r2, r, dx, dy, dz, c.pp);
    }
    __device__ unsigned long long finish(uint32_t i, uint32_t t, const Acc& a, const Self& si, const Sph& s) const {
        (void)da.finish(i, t, a.a, DensityAlpha::Self{}, s);
        const float source = -a.div;
        dv.src[i] = source;
        dv.src_m[t] = source;
        const float rho = a.a.rho, p = dv.pres[i];
        dv.pres_m[t] = p;
        dv.pos_w[t].w = prho_of(p, rho);
        feat_write(feat, n_, t, s, a.f, rho, da.fac[i], source, si.pp, s.rho0, s.dt);
        return 0;
    }
};

// pressure_force.comp:15-78. Neighbours supply p_j / rho_j^2 in the w of their position
// record (kNoRho when rho_j <= 1e-6, which the reference skips); the particle's own p_i /
// rho_i^2 is the same field, with kNoRho read as the reference's 0. Writes the pressure
// acceleration to the hash-ordered field and packed with the position for pressure_update.
// A staged window entry is the position record itself.
struct PressureForce {
    float4* pacc;
    float4* pa_m;
    float2* pa2_m;
    static constexpr int kParts = 1;
    static constexpr bool kPaGeo = false;
    static constexpr bool kWalls = false;
    static constexpr bool kAdditive = true;   // pair() adds one term per Acc component
    struct Self { float prho, px, py, pz; };
    struct Cand { float prho; };
    struct Acc { float x, y, z; };
    __device__ __forceinline__ Self self(uint32_t, uint32_t, const float4& g) const {
        return {__float_as_uint(g.w) == kNoRho ? 0.0f : g.w, g.x, g.y, g.z};
    }
    __device__ __forceinline__ Cand load(uint32_t, const float4& g) const { return {g.w}; }
    __device__ __forceinline__ void stage(float4* st, float2*, uint32_t q, uint32_t jt,
                                          const float4* pos_m) const {
        st[q] = pos_m[jt];
    }
    __device__ __forceinline__ Cand staged(const float2*, uint32_t, const float4& e0) const { return {e0.w}; }
    __device__ __forceinline__ Acc zero() const { return {0.f, 0.f, 0.f}; }
    __device__ __forceinline__ void pair(Acc& a, const Self& si, const Sph& s, float rx, float ry, float rz, const Cand& c) const {
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float r = sqrt_grad(r2);
        const bool ok = !(r2 > s.hh) && r > 1e-6f && __float_as_uint(c.prho) != kNoRho;
        const float g = exact_grad(r, s);   // finite, so a skipped pair adds (+0) g r = +/-0
        const float mt = ok ? s.mass * (si.prho + c.prho) : 0.0f;
        a.x += mt * (g * rx);
        a.y += mt * (g * ry);
        a.z += mt * (g * rz);
    }
    // the samples mirror the particle's pressure: (p_i / rho_i^2) G
    __device__ __forceinline__ void wall_add(Acc& a, const Self& si, const Sph&, const float4& g) const {
        a.x += si.prho * g.x;
        a.y += si.prho * g.y;
        a.z += si.prho * g.z;
    }
    __device__ __forceinline__ unsigned long long finish(uint32_t i, uint32_t t, const Acc& a, const Self& si, const Sph& s) const {
        const float4 v = make_float4(-a.x, -a.y, -a.z, 0.0f);
        if (pacc) pacc[i] = v;   // null inside a solve: written back afterwards
        pa_m[t] = make_float4(si.px, si.py, si.pz, v.x);
        pa2_m[t] = make_float2(v.y, v.z);
        return 0;
    }
};

// pressure_update.comp:18-84 (in place; neighbours' pressures are not read), plus the
// sweep's quantised compression residual. Positions and pressure accelerations come packed
// in pa_m / pa2_m; the particle's own factor, source term, density and pressure from their
// schedule-ordered mirrors. The new pressure goes to the hash-ordered field and its mirror,
// p / rho^2 to the w of the position mirror. A staged window entry is pa_m and pa2_m.
struct PressureUpdate {
    const float4* pa_m;
    const float2* pa2_m;
    const float* fac_m;
    const float* src_m;
    const float* dens_m;
    float* pres;
    float* pres_m;
    float4* pos_w;
    uint32_t two_sided;   // FE_RESIDUAL_TWO_SIDED: |s - Ap| where the entering pressure is > 0
    static constexpr int kParts = 2;
    static constexpr bool kPaGeo = true;
    static constexpr bool kWalls = false;
    static constexpr bool kAdditive = true;   // pair() adds one term per Acc component
    struct Self { float ax, ay, az; };
    struct Cand { float ax, ay, az; };
    struct Acc { float ap; };
    __device__ __forceinline__ Self self(uint32_t, uint32_t t, const float4& g) const {
        const float2 b = pa2_m[t];
        return {g.w, b.x, b.y};
    }
    __device__ __forceinline__ Cand load(uint32_t jt, const float4& g) const {
        const float2 b = pa2_m[jt];
        return {g.w, b.x, b.y};
    }
    __device__ __forceinline__ void stage(float4* st, float2* st2, uint32_t q, uint32_t jt, const float4*) const {
        st[q] = pa_m[jt];
        st2[q] = pa2_m[jt];
    }
    __device__ __forceinline__ Cand staged(const float2* st2, uint32_t loc, const float4& e0) const {
        const float2 b = st2[loc];
        return {e0.w, b.x, b.y};
    }
    __device__ __forceinline__ Acc zero() const { return {0.f}; }
    __device__ __forceinline__ void pair(Acc& a, const Self& si, const Sph& s, float rx, float ry, float rz, const Cand& c) const {
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float r = sqrt_grad(r2);
        const bool ok = !(r2 > s.hh) && r > 1e-6f;
        const float g = exact_grad(r, s);
        const float dx = si.ax - c.ax, dy = si.ay - c.ay, dz = si.az - c.az;
        a.ap += ok ? s.mass * (dx * (g * rx) + dy * (g * ry) + dz * (g * rz)) : 0.0f;
    }
    // the samples have no pressure acceleration: a_i . G
    __device__ __forceinline__ void wall_add(Acc& a, const Self& si, const Sph&, const float4& g) const {
        a.ap += dot_g(si.ax, si.ay, si.az, g);
    }
    __device__ __forceinline__ unsigned long long finish(uint32_t i, uint32_t t, const Acc& a, const Self&, const Sph& s) const {
        const float alpha_i = fac_m[t], rho_i = dens_m[t], src_i = src_m[t];
        float p = pres_m[t];
        const float ap = s.dt * a.ap;
        const float err = src_i - ap;
        const unsigned long long q = residual_q(two_sided && p > 0.0f ? -fabsf(err) : err, s.dt, s.rho0);
        if (alpha_i > 1e-6f && rho_i > 1e-6f) {
            const float a_ii = -(s.dt / (rho_i * rho_i)) * alpha_i;
            if (fabsf(a_ii) > 1e-20f) {
                const float corr = (s.omega * err) / a_ii;
                const float np = p + corr;
                p = np > 0.0f ? np : 0.0f;
                if (pres) pres[i] = p;   // null inside a solve: written back afterwards
                pres_m[t] = p;
            }
        }
        pos_w[t].w = prho_of(p, rho_i);
        return q;
    }
};

// The exit report of a solve (no reference counterpart): pressure_update's sum over the
// listed pairs with the accelerations of the solve's last sweep, keeping r = s - Ap of every
// particle (by schedule position) instead of updating its pressure.
struct ResidualReport : PressureUpdate {
    float* err_out;
    __device__ __forceinline__ unsigned long long finish(uint32_t, uint32_t t, const Acc& a, const Self&, const Sph& s) const {
        err_out[t] = src_m[t] - s.dt * a.ap;
        return 0;
    }
};

// Sums of a report: |r|, max(0, r) and max(0, r) where the final pressure is > 0, each as
// floor(x dt / rho0 2^24), plus the largest |r| term (slots 0..3). Each thread folds
// kReportItems particles (strided by the block), and the block adds once: integer sums and a
// maximum, so the order does not change them, and four atomics per block instead of per wave
// keep the four words from serialising the launch.
constexpr uint32_t kReportItems = 16;

__global__ __launch_bounds__(kBlock) void k_residual_reduce(const float* __restrict__ err, const float* __restrict__ pres_m,
                                                            uint32_t n, float dt, float rho0,
                                                            unsigned long long* __restrict__ out) {
    __shared__ unsigned long long part[4][kBlock / 64];
    const uint32_t base = blockIdx.x * (kBlock * kReportItems) + threadIdx.x;
    const float scale = dt / rho0 * 16777216.0f;
    unsigned long long a = 0, e = 0, ea = 0, mx = 0;
    for (uint32_t k = 0; k < kReportItems; k++) {
        const uint32_t t = base + k * kBlock;
        if (t >= n) break;
        const float r = err[t];
        const float ab = fminf(fabsf(r) * scale, 1.0e15f), ex = fminf(fmaxf(r, 0.0f) * scale, 1.0e15f);
        const unsigned long long at = (unsigned long long)ab, et = (unsigned long long)ex;
        a += at;
        e += et;
        ea += pres_m[t] > 0.0f ? et : 0ull;
        mx = at > mx ? at : mx;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        e += __shfl_xor(e, off, 64);
        ea += __shfl_xor(ea, off, 64);
        const unsigned long long o = __shfl_xor(mx, off, 64);
        mx = o > mx ? o : mx;
    }
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[0][w] = a;
        part[1][w] = e;
        part[2][w] = ea;
        part[3][w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t u = 1; u < kBlock / 64; u++) {
            a += part[0][u];
            e += part[1][u];
            ea += part[2][u];
            mx = part[3][u] > mx ? part[3][u] : mx;
        }
        if (a) atomicAdd(out, a);
        if (e) atomicAdd(out + 1, e);
        if (ea) atomicAdd(out + 2, ea);
        if (mx) atomicMax(out + 3, mx);
    }
}

// A policy with the wall samples of FE_BOUNDARY_PARTICLES: after its fluid pairs it adds its
// wall_add term of the particle's (G, rho_w).
template <class P>
struct Walled : P {
    const float4* wall_g;
    static constexpr bool kWalls = true;
};

// (G, rho_w) of every particle by schedule position: G = sum psi grad W over the samples with
// r^2 <= h^2 and r > 1e-6, rho_w = sum psi W over those with r^2 <= h^2, each summed in the
// enumeration order of for_each_wall_sample.
__global__ __launch_bounds__(kBlock) void k_wall_terms(Walls w, Sph s, const float4* __restrict__ pos_m, uint32_t n,
                                                       float4* __restrict__ wall_g) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const float4 p = pos_m[t];
    float gx = 0.0f, gy = 0.0f, gz = 0.0f, rho = 0.0f;
    for_each_wall_sample(w, s.h, p.x, p.y, p.z, [&](float rx, float ry, float rz, uint32_t b) {
        const float psi = w.psi[b];
        const float r2 = rx * rx + ry * ry + rz * rz;
        const bool in = !(r2 > s.hh);
        const float r = sqrt_exact(r2);
        const float q = q_of(r, s);
        rho += in ? psi * w_of_q(q, s) : 0.0f;
        const bool gok = in && r > 1e-6f;
        const float g = grad_of_q(r, q, s);
        gx += gok ? psi * (g * rx) : 0.0f;
        gy += gok ? psi * (g * ry) : 0.0f;
        gz += gok ? psi * (g * rz) : 0.0f;
    });
    wall_g[t] = make_float4(gx, gy, gz, rho);
}

// Boundary mass of every wall sample: psi_b = rho0 / sum_k W(x_b - x_k) over the samples k
// within h (the sample itself included), in the enumeration order of for_each_wall_sample.
__global__ __launch_bounds__(kBlock) void k_wall_psi(Walls w, Sph s, uint32_t total, float* __restrict__ psi) {
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= total) return;
    int f = 0;
    while (f < 5 && b >= w.off[f + 1]) f++;
    const int a = f >> 1;
    const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
    const uint32_t local = b - w.off[f];
    const int k = (int)(local % (uint32_t)w.n[f][0]), m = (int)(local / (uint32_t)w.n[f][0]);
    float x[3];
    x[a] = w.pl[f];
    x[u] = w.mn[u] + ((float)k + 0.5f) * w.sb;
    x[v] = w.mn[v] + ((float)m + 0.5f) * w.sb;
    float sum = 0.0f;
    for_each_wall_sample(w, s.h, x[0], x[1], x[2], [&](float rx, float ry, float rz, uint32_t) {
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float r = sqrt_exact(r2);
        sum += !(r2 > s.hh) ? exact_w(r, s) : 0.0f;
    });
    psi[b] = s.rho0 / sum;
}

// The record that holds a neighbour's position: pa_m for pressure_update, else pos_m.
template <class P>
__device__ __forceinline__ float4 geo_of(const P& pol, const float4* pos_m, uint32_t jt) {
    if constexpr (P::kPaGeo)
        return pol.pa_m[jt];
    else
        return pos_m[jt];
}

template <int B>
__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v) {
    __shared__ unsigned long long part[B / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < B / 64; k++) t += part[k];
    return t;
}

// A sweep block's residual output, summed into the sweep's slot by integer atomics (so the
// order does not matter): one per block for one scene; for a batch one per block into its
// scene's slot when all its waves lie in that scene, else one per wave, or per lane where a
// wave straddles two scenes.
template <bool kBatch, int B>
__device__ __forceinline__ void sweep_epilogue(unsigned long long q, uint32_t scene,
                                               unsigned long long* __restrict__ slot_out) {
    if (!slot_out) return;
    if constexpr (!kBatch) {
        const unsigned long long b = block_sum_u64<B>(q);
        if (threadIdx.x == 0 && b) atomicAdd(slot_out, b);
    } else {
        const uint32_t s_first = __shfl(scene, 0, 64);
        const bool uni = __all(scene == s_first);
        unsigned long long v = q;
        if (uni) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        }
        if constexpr (B > 64) {
            __shared__ uint32_t wsc[B / 64];
            __shared__ unsigned long long wv[B / 64];
            const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
            if (lane == 0) {
                wsc[w] = uni ? s_first : kSentinel;
                wv[w] = v;
            }
            __syncthreads();
            bool block_uni = wsc[0] != kSentinel;
#pragma unroll
            for (int u = 1; u < B / 64; u++) block_uni = block_uni && wsc[u] == wsc[0];
            if (block_uni) {
                if (threadIdx.x == 0) {
                    unsigned long long t = 0;
                    for (int u = 0; u < B / 64; u++) t += wv[u];
                    if (t) atomicAdd(slot_out + (size_t)wsc[0] * kMaxSolveIters, t);
                }
                return;
            }
        }
        if (uni) {
            if ((threadIdx.x & 63) == 0 && v) atomicAdd(slot_out + (size_t)s_first * kMaxSolveIters, v);
        } else if (q) {
            atomicAdd(slot_out + (size_t)scene * kMaxSolveIters, q);
        }
    }
}

// One work-item's gather: the listed pairs of schedule position t, kDepth neighbour loads in
// flight (4, or 16 for small scenes, where a gather is a chain of dependent list and
// neighbour loads rather than a stream), or the bucket walk when the list is stale or the
// particle overflowed it. Either depth visits the pairs in list order.
template <class P, int kDepth = 4>
__device__ __forceinline__ unsigned long long gather_one(const P& pol, const Sph& s, uint32_t t,
                                                         const uint2* __restrict__ range,
                                                         const float4* __restrict__ pos_b,
                                                         const float4* pos_m,
                                                         const uint32_t* __restrict__ sched,
                                                         const uint32_t* __restrict__ inv,
                                                         const uint32_t* __restrict__ nbr,
                                                         const uint32_t* __restrict__ cnt, uint32_t list_ok) {
    const uint32_t i = sched[t];
    const float4 pi = geo_of(pol, pos_m, t);
    const typename P::Self si = pol.self(i, t, pi);
    typename P::Acc acc = pol.zero();
    const uint32_t m = list_ok ? cnt[t] : 0xFFFFFFFFu;
    if (m <= kMaxNbr) {
        static_assert(kDepth % 4 == 0 && kMaxNbr % kDepth == 0, "whole list chunks");
        const uint4* tile = list_tile(nbr, t);
        for (uint32_t k = 0; k < m; k += kDepth) {
            // the row's chunks k / 4 .. k / 4 + kDepth / 4 - 1 (inside its kMaxNbr slots); slots
            // at m and beyond load the batch's first neighbour again and add nothing
            uint4 q4[kDepth / 4];
#pragma unroll
            for (int b = 0; b < kDepth / 4; b++) q4[b] = tile[((k >> 2) + b) * 16];
            uint32_t jt[kDepth];
#pragma unroll
            for (int b = 0; b < kDepth / 4; b++) {
                const uint32_t k0 = k + 4 * b;
                jt[4 * b] = b == 0 || k0 < m ? q4[b].x : q4[0].x;
                jt[4 * b + 1] = k0 + 1 < m ? q4[b].y : q4[0].x;
                jt[4 * b + 2] = k0 + 2 < m ? q4[b].z : q4[0].x;
                jt[4 * b + 3] = k0 + 3 < m ? q4[b].w : q4[0].x;
            }
            float4 pj[kDepth];
            typename P::Cand cj[kDepth];
#pragma unroll
            for (int u = 0; u < kDepth; u++) {
                pj[u] = geo_of(pol, pos_m, jt[u]);
                cj[u] = pol.load(jt[u], pj[u]);
            }
#pragma unroll
            for (int u = 0; u < kDepth; u++)
                if (k + u < m) pol.pair(acc, si, s, pi.x - pj[u].x, pi.y - pj[u].y, pi.z - pj[u].z, cj[u]);
        }
    } else {
        visit27(range, pos_b, s, pi.x, pi.y, pi.z, [&](uint32_t j, const float4& pj) {
            const uint32_t jt = inv[j];
            pol.pair(acc, si, s, pi.x - pj.x, pi.y - pj.y, pi.z - pj.z, pol.load(jt, geo_of(pol, pos_m, jt)));
        });
    }
    if constexpr (P::kWalls) pol.wall_add(acc, si, s, pol.wall_g[t]);
    return pol.finish(i, t, acc, si, s);
}

// One launch of a list consumer over all work-items. One scene (kBatch false): the scene's
// constants are kernel arguments, and the sweep returns at once when the previous sweep
// converged. A batch: each work-item takes its scene's constants from the table and skips
// when its own scene converged (a batch's slots are kMaxSolveIters apart per scene).
template <class P, bool kBatch, int B, int kDepth = 4>
__global__ __launch_bounds__(B) void k_list_gather(P pol, Sph s0, Scenes sc, uint32_t n,
                                                        const uint2* __restrict__ range,
                                                        const float4* __restrict__ pos_b,
                                                        const float4* pos_m,
                                                        const uint32_t* __restrict__ sched,
                                                        const uint32_t* __restrict__ inv,
                                                        const uint32_t* __restrict__ nbr,
                                                        const uint32_t* __restrict__ cnt, uint32_t list_ok,
                                                        const unsigned long long* __restrict__ prev,
                                                        unsigned long long thr, uint32_t kind,
                                                        unsigned long long* __restrict__ slot_out) {
    if (!kBatch && prev && *prev <= thr) return;
    const uint32_t t = work_item<B>();
    unsigned long long q = 0;
    uint32_t scene = 0;
    if (t < n) {
        if constexpr (kBatch) {
            scene = sc.sid[t];
            const Sph s = sc.tab[scene];
            if (!(prev && prev[(size_t)scene * kMaxSolveIters] <= (kind ? s.thr[1] : s.thr[0])))
                q = gather_one(pol, s, t, range, pos_b, pos_m, sched, inv, nbr, cnt, list_ok);
        } else {
            q = gather_one<P, kDepth>(pol, s0, t, range, pos_b, pos_m, sched, inv, nbr, cnt, list_ok);
        }
    }
    sweep_epilogue<kBatch, B>(q, scene, slot_out);
}

// ---- four lanes per particle (tune items 8, one scene) ------------------------------------
// Lane r of a particle's quad (four consecutive lanes) takes the listed slots r, r + 4, r + 8,
// ..., loads kQuadDepth of them at a time and forms each slot's term from a zero
// accumulator (0 + term = term, and a sum that starts at +0 never holds -0, so adding the
// term alone is adding the pair). The quad then adds the four terms of every group of four
// slots in slot order through DPP broadcasts, so each lane holds the particle's sums exactly
// as the one-lane gather forms them; lane 0 finishes the particle. A small scene's sweep then
// runs four times as many waves, each with a quarter of the pair arithmetic in its chain.
// Policies whose pair() only adds a term to each accumulator component (kAdditive) can take
// it; the warm-start feature gathers, which count their pairs, cannot.
constexpr int kQuadDepth = 8;   // slots per lane loaded at once: 32 per particle

template <class P, class = void>
struct Additive : std::false_type {};
template <class P>
struct Additive<P, std::void_t<decltype(P::kAdditive)>> : std::bool_constant<P::kAdditive && !P::kWalls> {};

// x of quad lane u, in every lane of the quad
__device__ __forceinline__ float quad_bcast(float x, int u) {
    const int ctrl = u | (u << 2) | (u << 4) | (u << 6);   // quad_perm [u, u, u, u]
    int v;
    switch (ctrl) {
        case 0x00: v = __builtin_amdgcn_mov_dpp(__float_as_int(x), 0x00, 0xF, 0xF, false); break;
        case 0x55: v = __builtin_amdgcn_mov_dpp(__float_as_int(x), 0x55, 0xF, 0xF, false); break;
        case 0xAA: v = __builtin_amdgcn_mov_dpp(__float_as_int(x), 0xAA, 0xF, 0xF, false); break;
        default: v = __builtin_amdgcn_mov_dpp(__float_as_int(x), 0xFF, 0xF, 0xF, false); break;
    }
    return __int_as_float(v);
}

// acc += term of quad lane 0, then 1, 2, 3 (component by component).
template <class Acc>
__device__ __forceinline__ void quad_add(Acc& acc, const Acc& term) {
    constexpr int N = sizeof(Acc) / sizeof(float);
    static_assert(sizeof(Acc) == N * sizeof(float), "float accumulators");
    float a[N], t[N];
    __builtin_memcpy(a, &acc, sizeof(Acc));
    __builtin_memcpy(t, &term, sizeof(Acc));
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int k = 0; k < N; k++) a[k] += quad_bcast(t[k], u);
    __builtin_memcpy(&acc, a, sizeof(Acc));
}

template <class P>
__device__ __forceinline__ unsigned long long gather_quad(const P& pol, const Sph& s, uint32_t t, uint32_t r,
                                                          const uint2* __restrict__ range,
                                                          const float4* __restrict__ pos_b, const float4* pos_m,
                                                          const uint32_t* __restrict__ sched,
                                                          const uint32_t* __restrict__ inv,
                                                          const uint32_t* __restrict__ nbr,
                                                          const uint32_t* __restrict__ cnt, uint32_t list_ok) {
    const uint32_t i = sched[t];
    const float4 pi = geo_of(pol, pos_m, t);
    const typename P::Self si = pol.self(i, t, pi);
    typename P::Acc acc = pol.zero();
    const uint32_t m = list_ok ? cnt[t] : 0xFFFFFFFFu;   // the same in the four lanes
    if (m <= kMaxNbr) {
        // slot k of t's row: uint32 ((t / 16) (kMaxNbr / 4) + k / 4) 64 + (t % 16) 4 + k % 4
        const uint32_t* row = nbr + (((size_t)(t >> 4) * (kMaxNbr / 4)) << 6) + ((t & 15u) << 2) + r;
        for (uint32_t k0 = 0; k0 < m; k0 += 4 * kQuadDepth) {
            // slots k0 + 4 b + r < kMaxNbr: inside the row; past m the particle's own record
            // is loaded and nothing is added
            uint32_t jt[kQuadDepth];
#pragma unroll
            for (int b = 0; b < kQuadDepth; b++) {
                const uint32_t k = k0 + 4 * b;
                const uint32_t e = row[(k >> 2) << 6];
                jt[b] = k + r < m ? e : t;
            }
            float4 pj[kQuadDepth];
            typename P::Cand cj[kQuadDepth];
#pragma unroll
            for (int b = 0; b < kQuadDepth; b++) {
                pj[b] = geo_of(pol, pos_m, jt[b]);
                cj[b] = pol.load(jt[b], pj[b]);
            }
#pragma unroll
            for (int b = 0; b < kQuadDepth; b++) {
                typename P::Acc term = pol.zero();
                if (k0 + 4 * b + r < m) pol.pair(term, si, s, pi.x - pj[b].x, pi.y - pj[b].y, pi.z - pj[b].z, cj[b]);
                quad_add(acc, term);
            }
        }
    } else if (r == 0) {
        visit27(range, pos_b, s, pi.x, pi.y, pi.z, [&](uint32_t j, const float4& pj) {
            const uint32_t jt = inv[j];
            pol.pair(acc, si, s, pi.x - pj.x, pi.y - pj.y, pi.z - pj.z, pol.load(jt, geo_of(pol, pos_m, jt)));
        });
    }
    return r == 0 ? pol.finish(i, t, acc, si, s) : 0ull;
}

template <class P, int B>
__global__ __launch_bounds__(B) void k_quad_gather(P pol, Sph s0, uint32_t n, const uint2* __restrict__ range,
                                                   const float4* __restrict__ pos_b, const float4* pos_m,
                                                   const uint32_t* __restrict__ sched,
                                                   const uint32_t* __restrict__ inv,
                                                   const uint32_t* __restrict__ nbr,
                                                   const uint32_t* __restrict__ cnt, uint32_t list_ok,
                                                   const unsigned long long* __restrict__ prev,
                                                   unsigned long long thr,
                                                   unsigned long long* __restrict__ slot_out) {
    static_assert(Additive<P>::value, "only additive policies without wall samples");
    if (prev && *prev <= thr) return;
    const uint32_t wi = work_item<B>();
    const uint32_t t = wi >> 2, r = wi & 3u;
    unsigned long long q = 0;
    if (t < n) q = gather_quad(pol, s0, t, r, range, pos_b, pos_m, sched, inv, nbr, cnt, list_ok);
    sweep_epilogue<false, B>(q, 0u, slot_out);
}

// The staged sweep of one window (one wave): the window's staging slots are loaded from the
// schedule-ordered mirrors into LDS, then each lane walks its 16-bit list over LDS.
// run: this lane's particle is swept (in range, its scene not converged).
template <class P>
__device__ __forceinline__ unsigned long long sweep_window(
    const P& pol, const Sph& s, bool run, uint32_t t, uint32_t win, uint32_t last, float4* st, float2* st2, uint32_t* dl,
    const uint2* __restrict__ range, const float4* __restrict__ pos_b, const float4* pos_m,
    const uint32_t* __restrict__ sched, const uint32_t* __restrict__ inv, const uint32_t* __restrict__ nbr,
    const uint32_t* __restrict__ cnt, const uint16_t* __restrict__ nbr16, const uint32_t* __restrict__ win_meta,
    const uint4* __restrict__ win_round, const uint32_t* __restrict__ win_delta) {
    if (!__any(run)) return 0;
    const int lane = threadIdx.x & 63;
    const uint32_t meta = win_meta[win];
    if (meta == kWinFallback) return run ? gather_one(pol, s, t, range, pos_b, pos_m, sched, inv, nbr, cnt, 1u) : 0ull;
    // the clamps keep every access in bounds whatever the tables hold
    const uint32_t K = min(meta & 0xFFFFu, kWinMaxK), R = min(meta >> 16, kWinMaxRuns);
    for (uint32_t r = lane; r < R; r += 64) dl[r] = win_delta[(size_t)win * kWinMaxRuns + r];
    wave_sync();
    const uint64_t below = (2ull << lane) - 1ull;   // bits 0..lane
    for (uint32_t k = 0; k * 64u < K; k++) {
        const uint4 rd = win_round[(size_t)win * kWinRounds + k];
        const uint64_t mk = (uint64_t)rd.x | ((uint64_t)rd.y << 32);
        const uint32_t slot = k * 64u + (uint32_t)lane;
        if (slot < K) {
            const uint32_t r = min(rd.z + (uint32_t)__popcll(mk & below), kWinMaxRuns - 1u);
            pol.stage(st, st2, slot, min(slot + dl[r], last), pos_m);
        }
    }
    wave_sync();
    if (!run) return 0;
    const uint32_t i = sched[t];
    const float4 pi = geo_of(pol, pos_m, t);
    const typename P::Self si = pol.self(i, t, pi);
    typename P::Acc acc = pol.zero();
    const uint32_t m = min(cnt[t], kMaxNbr);   // a window holds no overflowing list
    const uint4* tile = list16_tile(nbr16, t);
    for (uint32_t k = 0; k < m; k += 8) {
        const uint4 qv = tile[(k >> 3) * 16];
        const uint32_t w4[4] = {qv.x, qv.y, qv.z, qv.w};
        float4 e[8];
        typename P::Cand cj[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint32_t loc = ((w4[u >> 1] >> (16 * (u & 1))) & 0xFFFFu) & (kWinMaxK - 1u);
            e[u] = st[loc];
            cj[u] = pol.staged(st2, loc, e[u]);
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (k + u < m) pol.pair(acc, si, s, pi.x - e[u].x, pi.y - e[u].y, pi.z - e[u].z, cj[u]);
    }
    return pol.finish(i, t, acc, si, s);
}

template <class P, bool kBatch, int kWaves>
__global__ __launch_bounds__(64 * kWaves) void k_sweep_win(P pol, Sph s0, Scenes sc, uint32_t n,
                                                         const uint2* __restrict__ range,
                                                         const float4* __restrict__ pos_b,
                                                         const float4* pos_m,
                                                         const uint32_t* __restrict__ sched,
                                                         const uint32_t* __restrict__ inv,
                                                         const uint32_t* __restrict__ nbr,
                                                         const uint32_t* __restrict__ cnt,
                                                         const uint16_t* __restrict__ nbr16,
                                                         const uint32_t* __restrict__ win_meta,
                                                         const uint4* __restrict__ win_round,
                                                         const uint32_t* __restrict__ win_delta,
                                                         const unsigned long long* __restrict__ prev,
                                                         unsigned long long thr, uint32_t kind,
                                                         unsigned long long* __restrict__ slot_out) {
    constexpr int B = 64 * kWaves;
    __shared__ float4 stage[kWaves][kWinMaxK];
    __shared__ float2 stage2[kWaves][P::kParts > 1 ? kWinMaxK : 1];
    __shared__ uint32_t sdelta[kWaves][kWinMaxRuns];
    if (!kBatch && prev && *prev <= thr) return;
    const int wv = threadIdx.x >> 6;
    const uint32_t win = __builtin_amdgcn_readfirstlane(xcd_block() * kWaves + wv);
    const uint32_t t = win * kWin + (threadIdx.x & 63);
    unsigned long long q = 0;
    uint32_t scene = 0;
    if constexpr (kBatch) {
        scene = t < n ? sc.sid[t] : 0u;
        const Sph s = sc.tab[scene];
        const bool run = t < n && !(prev && prev[(size_t)scene * kMaxSolveIters] <= (kind ? s.thr[1] : s.thr[0]));
        q = sweep_window(pol, s, run, t, win, n - 1u, stage[wv], stage2[wv], sdelta[wv], range, pos_b, pos_m, sched, inv,
                         nbr, cnt,
                         nbr16, win_meta, win_round, win_delta);
    } else {
        q = sweep_window(pol, s0, t < n, t, win, n - 1u, stage[wv], stage2[wv], sdelta[wv], range, pos_b, pos_m, sched,
                         inv, nbr,
                         cnt, nbr16, win_meta, win_round, win_delta);
    }
    sweep_epilogue<kBatch, B>(q, scene, slot_out);
}

inline dim3 grid_for(uint32_t count) { return dim3((count + kBlock - 1) / kBlock); }

// The hash-ordered pressures and pressure accelerations of the sweeps that deferred them.
__global__ __launch_bounds__(kBlock) void k_sweep_writeback(const uint32_t* __restrict__ sched, uint32_t n,
                                                            const float* __restrict__ pres_m,
                                                            const float4* __restrict__ pa_m,
                                                            const float2* __restrict__ pa2_m, float* __restrict__ pres,
                                                            float4* __restrict__ pacc) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = sched[t];
    if (pres) pres[i] = pres_m[t];
    if (pacc) {
        const float ax = pa_m[t].w;
        const float2 b = pa2_m[t];
        pacc[i] = make_float4(ax, b.x, b.y, 0.0f);
    }
}

// Mirror bits each consumer needs; a missing one is rebuilt from the hash-ordered fields
// (after any deferred sweep writes have reached them).
hipError_t need(fe_ctx* c, uint32_t bits) {
    if ((c->mirrors & bits) == bits) return hipSuccess;
    hipError_t e = flush_sweep_writes(c);
    if (e != hipSuccess) return e;
    Launch L(c, "refresh_mirrors");
    hipLaunchKernelGGL(k_refresh_mirrors, grid_for(c->n), dim3(kBlock), 0, c->stream, sph_of(c), scenes_of(c),
                       c->sched[0], c->n, c->pos_b, c->vel_b, c->vel_a, c->dens, c->pres, c->pacc, c->fac, c->src,
                       c->inv, c->pos_m, c->vb_m, c->va_m, c->dens_m, c->pa_m, c->pa2_m, c->fac_m,
                       c->src_m, c->pres_m);
    c->mirrors = kMirrorAll;
    return hipGetLastError();
}

template <class P, bool kBatch, int B, int kDepth = 4>
void launch_gather(fe_ctx* c, const P& pol, const unsigned long long* prev,
                   unsigned long long thr, uint32_t kind, unsigned long long* slot_out) {
    hipLaunchKernelGGL((k_list_gather<P, kBatch, B, kDepth>), dim3((c->n + B - 1) / B), dim3(B), 0, c->stream, pol,
                       sph_of(c), scenes_of(c), c->n, c->range, c->pos_b, c->pos_m, c->sched[0], c->inv, c->nbr,
                       c->nbr_count, c->list_valid ? 1u : 0u, prev, thr, kind, slot_out);
}

// items 4 (one scene, blocks of 64 to 256, no wall samples): 16 neighbour loads in flight
// per work-item
template <class P, bool kBatch>
void launch_gather_b(fe_ctx* c, uint32_t block, const P& pol, const unsigned long long* prev, unsigned long long thr, uint32_t kind,
                     unsigned long long* slot_out, uint32_t items = 1) {
    if constexpr (!kBatch && !P::kWalls) {
        if (items == 4) {
            switch (block) {
                case 64: launch_gather<P, false, 64, 16>(c, pol, prev, thr, kind, slot_out); return;
                case 128: launch_gather<P, false, 128, 16>(c, pol, prev, thr, kind, slot_out); return;
                default: launch_gather<P, false, 256, 16>(c, pol, prev, thr, kind, slot_out); return;
            }
        }
    }
    switch (block) {
        case 64: launch_gather<P, kBatch, 64>(c, pol, prev, thr, kind, slot_out); break;
        case 128: launch_gather<P, kBatch, 128>(c, pol, prev, thr, kind, slot_out); break;
        case 512: launch_gather<P, kBatch, 512>(c, pol, prev, thr, kind, slot_out); break;
        case 1024: launch_gather<P, kBatch, 1024>(c, pol, prev, thr, kind, slot_out); break;
        default: launch_gather<P, kBatch, 256>(c, pol, prev, thr, kind, slot_out); break;
    }
}

template <class P, int B>
void launch_quad(fe_ctx* c, const P& pol, const unsigned long long* prev, unsigned long long thr,
                 unsigned long long* slot_out) {
    const uint64_t items = 4ull * c->n;
    hipLaunchKernelGGL((k_quad_gather<P, B>), dim3((unsigned)((items + B - 1) / B)), dim3(B), 0, c->stream, pol,
                       sph_of(c), c->n, c->range, c->pos_b, c->pos_m, c->sched[0], c->inv, c->nbr, c->nbr_count,
                       c->list_valid ? 1u : 0u, prev, thr, slot_out);
}

template <class P>
hipError_t run_list(fe_ctx* c, const char* name, int tid, const P& pol, const unsigned long long* prev = nullptr, unsigned long long thr = 0, uint32_t kind = 0,
                    unsigned long long* slot_out = nullptr) {
    Launch L(c, name);
    const uint32_t block = tid >= 0 ? c->tune[tid].block : 256u, items = tid >= 0 ? c->tune[tid].items : 1u;
    if constexpr (Additive<P>::value) {
        if (items == 8 && c->nscenes == 1) {   // four lanes per particle
            switch (block) {
                case 64: launch_quad<P, 64>(c, pol, prev, thr, slot_out); break;
                case 128: launch_quad<P, 128>(c, pol, prev, thr, slot_out); break;
                default: launch_quad<P, 256>(c, pol, prev, thr, slot_out); break;
            }
            return hipGetLastError();
        }
    }
    if (c->nscenes > 1)
        launch_gather_b<P, true>(c, block, pol, prev, thr, kind, slot_out);
    else
        launch_gather_b<P, false>(c, block, pol, prev, thr, kind, slot_out, items);
    return hipGetLastError();
}

template <class P, bool kBatch, int W>
void launch_win(fe_ctx* c, const P& pol, const unsigned long long* prev,
                unsigned long long thr, uint32_t kind, unsigned long long* slot_out) {
    const uint32_t nwin = (c->n + kWin - 1) / kWin;
    hipLaunchKernelGGL((k_sweep_win<P, kBatch, W>), dim3((nwin + W - 1) / W), dim3(64 * W), 0, c->stream, pol,
                       sph_of(c), scenes_of(c), c->n, c->range, c->pos_b, c->pos_m, c->sched[0], c->inv, c->nbr,
                       c->nbr_count, c->nbr16, c->win_meta, c->win_round, c->win_delta, prev, thr, kind,
                       slot_out);
}

template <class P, bool kBatch>
void launch_win_b(fe_ctx* c, uint32_t block, const P& pol, const unsigned long long* prev, unsigned long long thr, uint32_t kind,
                  unsigned long long* slot_out) {
    switch (block) {
        case 64: launch_win<P, kBatch, 1>(c, pol, prev, thr, kind, slot_out); break;
        case 128: launch_win<P, kBatch, 2>(c, pol, prev, thr, kind, slot_out); break;
        default: launch_win<P, kBatch, 4>(c, pol, prev, thr, kind, slot_out); break;
    }
}

// A Jacobi sweep: staged windows over the current list (tuned items 2), or per-particle
// gathers of the listed pairs (items 1; the bucket walk when there is no current list).
template <class P>
hipError_t run_sweep(fe_ctx* c, const char* name, int tid, const P& pol, const unsigned long long* prev, unsigned long long thr, uint32_t kind,
                     unsigned long long* slot_out) {
    if (!c->list_valid || c->tune[tid].items != 2) return run_list(c, name, tid, pol, prev, thr, kind, slot_out);
    if (c->win_gen != c->list_gen) {   // the current list has no window tables yet
        Launch W(c, "window_tables");
        constexpr int kTabWaves = 4;
        const uint32_t nwin = (c->n + kWin - 1) / kWin;
        hipLaunchKernelGGL(k_window_tables<kTabWaves>, dim3((nwin + kTabWaves - 1) / kTabWaves), dim3(64 * kTabWaves),
                           0, c->stream, c->n, c->nbr, c->nbr_count, c->win_meta, c->win_round, c->win_delta,
                           c->nbr16);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        c->win_gen = c->list_gen;
    }
    Launch L(c, name);
    const uint32_t block = c->tune[tid].block;
    if (c->nscenes > 1)
        launch_win_b<P, true>(c, block, pol, prev, thr, kind, slot_out);
    else
        launch_win_b<P, false>(c, block, pol, prev, thr, kind, slot_out);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_schedule_mirrors(fe_ctx* c) {
    Launch L(c, "schedule_mirrors");
    hipLaunchKernelGGL(k_schedule_mirrors, grid_for(c->n), dim3(kBlock), 0, c->stream, c->sched[0], c->n, c->pos_b,
                       c->vel_b, c->vel_a, c->inv, c->pos_m, c->vb_m, c->va_m);
    c->mirrors = kMirrorPos | kMirrorVelB | kMirrorVelA;
    c->list_valid = false;
    return hipGetLastError();
}

hipError_t launch_reorder_velocity(fe_ctx* c) {
    Launch L(c, "reorder_velocity");
    hipLaunchKernelGGL(k_reorder_velocity, grid_for(c->n), dim3(kBlock), 0, c->stream, c->sched[0], c->vals[0], c->n,
                       c->vel_a, c->vel_b, c->vb_m);
    c->mirrors |= kMirrorVelB;
    return hipGetLastError();
}

hipError_t launch_build_list(fe_ctx* c) {
    hipError_t e = need(c, kMirrorPos);   // positions by schedule position + inverse schedule
    if (e != hipSuccess) return e;
    Launch L(c, "build_list");
    const uint32_t waves = c->tune[kTuneBuildList].block / 64;
    const uint32_t win = c->tune[kTuneBuildList].items == 8 ? 8u : 64u;
    const dim3 grid((c->n + win * waves - 1) / (win * waves)), blk(64 * waves);
    const Sph s0 = sph_of(c);
    const Scenes sc = scenes_of(c);
#define FE_BUILD(BATCH, W, K)                                                                                   \
    hipLaunchKernelGGL((k_build_list<BATCH, W, K>), grid, blk, 0, c->stream, s0, sc, c->n, c->range, c->pos_b, \
                       c->pos_m, c->inv, c->nbr, c->nbr_count)
#define FE_BUILD_W(BATCH, K)                                                                \
    do {                                                                                     \
        if (waves == 1) FE_BUILD(BATCH, 1, K); else if (waves == 2) FE_BUILD(BATCH, 2, K); else FE_BUILD(BATCH, 4, K); \
    } while (0)
    if (c->nscenes > 1) {
        if (win == 8) FE_BUILD_W(true, 8); else FE_BUILD_W(true, 64);
    } else {
        if (win == 8) FE_BUILD_W(false, 8); else FE_BUILD_W(false, 64);
    }
#undef FE_BUILD_W
#undef FE_BUILD
    e = hipGetLastError();
    if (e == hipSuccess) {
        c->list_valid = true;
        c->list_gen++;   // window tables are built from it on first use
    }
    return e;
}

bool walls_on(const fe_ctx* c) { return c->options.boundary == FE_BOUNDARY_PARTICLES; }

// The wall lattice of the current box and particle radius (psi: the context's table).
Walls walls_of(const fe_ctx* c, uint32_t* total) {
    Walls w{};
    w.psi = c->walls_psi;
    const fe_params& p = c->params;
    for (int a = 0; a < 3; a++) {
        w.mn[a] = p.box_min[a];
        w.mx[a] = p.box_max[a];
    }
    w.sb = 2.0f * p.particle_radius;
    w.R = (int)std::ceil(p.smoothing_radius / w.sb) + 1;
    uint32_t off = 0;
    for (int f = 0; f < 6; f++) {
        const int a = f >> 1;
        const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
        w.pl[f] = (f & 1) ? w.mx[a] + p.particle_radius : w.mn[a] - p.particle_radius;
        const float eu = std::floor((w.mx[u] - w.mn[u]) / w.sb), ev = std::floor((w.mx[v] - w.mn[v]) / w.sb);
        w.n[f][0] = eu > 0.0f ? (int)eu : 0;
        w.n[f][1] = ev > 0.0f ? (int)ev : 0;
        w.off[f] = off;
        off += (uint32_t)w.n[f][0] * (uint32_t)w.n[f][1];
    }
    if (total) *total = off;
    return w;
}

void walls_key_of(const fe_params& p, float key[9]) {
    const float k[9] = {p.box_min[0], p.box_min[1], p.box_min[2], p.box_max[0], p.box_max[1], p.box_max[2],
                        p.particle_radius, p.smoothing_radius, p.target_density};
    std::memcpy(key, k, sizeof(k));
}

bool walls_current(const fe_ctx* c) {
    if (!walls_on(c)) return true;
    float key[9];
    walls_key_of(c->params, key);
    return c->walls_ready && std::memcmp(key, c->walls_key, sizeof(key)) == 0;
}

// (Re)compute the samples' boundary masses when the box, r, h or rho0 changed.
hipError_t ensure_walls(fe_ctx* c) {
    float key[9];
    walls_key_of(c->params, key);
    if (c->walls_ready && std::memcmp(key, c->walls_key, sizeof(key)) == 0) return hipSuccess;
    uint32_t total = 0;
    walls_of(c, &total);
    if (total > (1u << 26)) return hipErrorInvalidValue;   // the box grew past the checked size
    if (total > c->walls_cap) {
        if (c->walls_psi) (void)hipFree(c->walls_psi);
        c->walls_psi = nullptr;
        c->walls_cap = 0;
        hipError_t e = hipMalloc(&c->walls_psi, sizeof(float) * (total ? total : 1));
        if (e != hipSuccess) return e;
        c->walls_cap = total;
    }
    const Walls w = walls_of(c, nullptr);
    if (total) {
        Launch L(c, "wall_psi");
        hipLaunchKernelGGL(k_wall_psi, grid_for(total), dim3(kBlock), 0, c->stream, w, sph_of(c), total, c->walls_psi);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    std::memcpy(c->walls_key, key, sizeof(key));
    c->walls_ready = true;
    c->wall_g_fresh = false;
    return hipSuccess;
}

// (G, rho_w) of the current positions (schedule order), computed once per neighbour search
// or sample change and read by every gather until then.
hipError_t ensure_wall_g(fe_ctx* c) {
    if (!c->walls_ready) return hipErrorInvalidValue;
    if (c->wall_g_fresh) return hipSuccess;
    hipError_t e = need(c, kMirrorPos);
    if (e != hipSuccess) return e;
    if (!c->wall_g && (e = hipMalloc(&c->wall_g, sizeof(float4) * (size_t)c->n)) != hipSuccess) return e;
    {
        Launch L(c, "wall_terms");
        hipLaunchKernelGGL(k_wall_terms, grid_for(c->n), dim3(kBlock), 0, c->stream, walls_of(c, nullptr), sph_of(c),
                           c->pos_m, c->n, c->wall_g);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    c->wall_g_fresh = true;
    return hipSuccess;
}

// A gather with the wall samples (one scene, the per-particle lists or the bucket walk), in
// the block size tuned for kernel `tid` (the tuner times these instances when the walls are
// on); tid < 0 (the feature gathers) runs 256 threads.
template <class P>
hipError_t run_walled(fe_ctx* c, const char* name, int tid, const P& pol, const unsigned long long* prev = nullptr,
                      unsigned long long thr = 0, uint32_t kind = 0, unsigned long long* slot_out = nullptr) {
    hipError_t e = ensure_walls(c);
    if (e == hipSuccess) e = ensure_wall_g(c);
    if (e != hipSuccess) return e;
    Walled<P> wp;
    static_cast<P&>(wp) = pol;
    wp.wall_g = c->wall_g;
    Launch L(c, name);
    if (tid < 0)
        launch_gather<Walled<P>, false, 256>(c, wp, prev, thr, kind, slot_out);
    else
        launch_gather_b<Walled<P>, false>(c, c->tune[tid].block, wp, prev, thr, kind, slot_out);
    return hipGetLastError();
}

hipError_t launch_density_alpha_only(fe_ctx* c) {
    DensityAlpha pol{c->dens, c->fac, c->dens_m, c->fac_m, c->vb_m};
    hipError_t e = walls_on(c) ? run_walled(c, "density_alpha", kTuneDensityAlpha, pol) : run_list(c, "density_alpha", kTuneDensityAlpha, pol);
    c->mirrors = (c->mirrors | kMirrorDens | kMirrorFac | kMirrorMrho) & ~kMirrorPrho;   // p / rho^2 used the old densities
    c->da_fresh = c->ns_fresh;
    return e;
}

hipError_t launch_density_alpha(fe_ctx* c) {
    hipError_t e = launch_build_list(c);
    return e == hipSuccess ? launch_density_alpha_only(c) : e;
}

namespace {

DensityAlphaDivergence density_alpha_divergence_of(const fe_ctx* c) {
    // The consistent order's velocities are the velocity_b mirror, whose w this pass writes
    // (m / rho); the pass reads only its xyz.
    const bool consistent = c->options.divergence_order == FE_DIVERGENCE_CONSISTENT;
    return DensityAlphaDivergence{{c->dens, c->fac, c->dens_m, c->fac_m, c->vb_m},
                                  {consistent ? c->vb_m : c->va_m, c->dens, c->pres, c->src, c->pos_m, c->src_m,
                                   c->pres_m}};
}

hipError_t density_alpha_divergence_mirrors(fe_ctx* c) {
    const bool consistent = c->options.divergence_order == FE_DIVERGENCE_CONSISTENT;
    return need(c, kMirrorPos | (consistent ? kMirrorVelB : kMirrorVelA));
}

void density_alpha_divergence_done(fe_ctx* c) {
    c->mirrors |= kMirrorDens | kMirrorFac | kMirrorMrho | kMirrorPrho | kMirrorSrc | kMirrorPres;
    c->da_fresh = c->ns_fresh;
}

}  // namespace

hipError_t launch_density_alpha_divergence_source(fe_ctx* c) {
    hipError_t e = launch_build_list(c);   // also brings the position mirror
    if (e == hipSuccess) e = density_alpha_divergence_mirrors(c);
    if (e != hipSuccess) return e;
    const DensityAlphaDivergence pol = density_alpha_divergence_of(c);
    if (learn_features_wanted(c, 1)) {
        if ((e = learn_prev(c, 1)) != hipSuccess) return e;
        DensityAlphaDivergenceFeat fp{pol.da, pol.dv, c->ws.pprev, c->ws.feat, c->n};
        e = walls_on(c) ? run_walled(c, "density_alpha_divergence_source_feat", -1, fp)
                        : run_list(c, "density_alpha_divergence_source_feat", kTuneDensityAlphaDivergence, fp);
        learn_features_done(c, 1);
    } else {
        e = walls_on(c) ? run_walled(c, "density_alpha_divergence_source", kTuneDensityAlphaDivergence, pol)
                        : run_list(c, "density_alpha_divergence_source", kTuneDensityAlphaDivergence, pol);
    }
    density_alpha_divergence_done(c);
    return e;
}

hipError_t launch_density_alpha_divergence_only(fe_ctx* c) {
    hipError_t e = density_alpha_divergence_mirrors(c);
    if (e != hipSuccess) return e;
    const DensityAlphaDivergence pol = density_alpha_divergence_of(c);
    e = walls_on(c) ? run_walled(c, "density_alpha_divergence_source", kTuneDensityAlphaDivergence, pol)
                    : run_list(c, "density_alpha_divergence_source", kTuneDensityAlphaDivergence, pol);
    density_alpha_divergence_done(c);
    return e;
}

hipError_t launch_viscosity(fe_ctx* c) {
    hipError_t e = need(c, kMirrorPos | kMirrorVelB | kMirrorMrho);
    if (e != hipSuccess) return e;
    Viscosity pol{c->vb_m, c->vel_a, c->va_m};
    e = run_list(c, "viscosity", kTuneViscosity, pol);
    c->mirrors |= kMirrorVelA;
    return e;
}

hipError_t launch_density_source(fe_ctx* c) {
    hipError_t e = need(c, kMirrorPos | kMirrorVelA);
    if (e != hipSuccess) return e;
    DensitySource pol{c->va_m, c->dens, c->src, c->pres, c->pos_m, c->src_m, c->pres_m};
    if (learn_features_wanted(c, 0)) {
        if ((e = learn_prev(c, 0)) != hipSuccess) return e;
        DensitySourceFeat fp{pol, c->fac, c->ws.pprev, c->ws.feat, c->n};
        e = walls_on(c) ? run_walled(c, "density_source_feat", -1, fp)
                        : run_list(c, "density_source_feat", kTuneDensitySource, fp);
        learn_features_done(c, 0);
    } else {
        e = walls_on(c) ? run_walled(c, "density_source", kTuneDensitySource, pol) : run_list(c, "density_source", kTuneDensitySource, pol);
    }
    c->mirrors |= kMirrorPrho | kMirrorSrc | kMirrorPres;
    return e;
}

hipError_t launch_divergence_source(fe_ctx* c) {
    // The consistent order reads velocity_b, which the reorder filled in the current order,
    // through its mirror (xyz; w holds m / rho and is not read).
    const bool consistent = c->options.divergence_order == FE_DIVERGENCE_CONSISTENT;
    hipError_t e = need(c, kMirrorPos | (consistent ? kMirrorVelB : kMirrorVelA));
    if (e != hipSuccess) return e;
    DivergenceSource pol{consistent ? c->vb_m : c->va_m, c->dens, c->pres, c->src, c->pos_m, c->src_m, c->pres_m};
    if (learn_features_wanted(c, 1)) {
        if ((e = learn_prev(c, 1)) != hipSuccess) return e;
        DivergenceSourceFeat fp{pol, c->fac, c->ws.pprev, c->ws.feat, c->n};
        e = walls_on(c) ? run_walled(c, "divergence_source_feat", -1, fp)
                        : run_list(c, "divergence_source_feat", kTuneDivergenceSource, fp);
        learn_features_done(c, 1);
    } else {
        e = walls_on(c) ? run_walled(c, "divergence_source", kTuneDivergenceSource, pol)
                        : run_list(c, "divergence_source", kTuneDivergenceSource, pol);
    }
    c->mirrors |= kMirrorPrho | kMirrorSrc | kMirrorPres;
    return e;
}

hipError_t launch_pressure_force(fe_ctx* c, const unsigned long long* prev, unsigned long long thr, uint32_t kind) {
    hipError_t e = need(c, kMirrorPos | kMirrorPrho);
    if (e != hipSuccess) return e;
    PressureForce pol{c->defer_sweep_writes ? nullptr : c->pacc, c->pa_m, c->pa2_m};
    e = walls_on(c) ? run_walled(c, "pressure_force", kTunePressureForce, pol, prev, thr, kind, nullptr)
                    : run_sweep(c, "pressure_force", kTunePressureForce, pol, prev, thr, kind, nullptr);
    c->mirrors |= kMirrorPacc;
    c->pending_pacc |= c->defer_sweep_writes;
    return e;
}

hipError_t launch_pressure_update(fe_ctx* c, unsigned long long* out, const unsigned long long* prev,
                                  unsigned long long thr, uint32_t kind) {
    hipError_t e = need(c, kMirrorPacc | kMirrorDens | kMirrorFac | kMirrorSrc | kMirrorPres);
    if (e != hipSuccess) return e;
    PressureUpdate pol{c->pa_m,   c->pa2_m,  c->fac_m, c->src_m, c->dens_m, c->defer_sweep_writes ? nullptr : c->pres,
                       c->pres_m, c->pos_m, c->options.residual == FE_RESIDUAL_TWO_SIDED ? 1u : 0u};
    e = walls_on(c) ? run_walled(c, "pressure_update", kTunePressureUpdate, pol, prev, thr, kind, out)
                    : run_sweep(c, "pressure_update", kTunePressureUpdate, pol, prev, thr, kind, out);
    c->pending_pres |= c->defer_sweep_writes;
    return e;
}

hipError_t launch_residual_report(fe_ctx* c, int kind) {
    hipError_t e = need(c, kMirrorPacc | kMirrorDens | kMirrorFac | kMirrorSrc | kMirrorPres);
    if (e != hipSuccess) return e;
    ResidualReport pol;
    static_cast<PressureUpdate&>(pol) = PressureUpdate{c->pa_m, c->pa2_m, c->fac_m, c->src_m, c->dens_m, nullptr,
                                                       c->pres_m, c->pos_m, 0u};
    pol.err_out = c->report_err;
    e = walls_on(c) ? run_walled(c, "residual_report", -1, pol) : run_list(c, "residual_report", kTunePressureUpdate, pol);
    if (e != hipSuccess) return e;
    Launch L(c, "residual_reduce");
    const uint32_t per = kBlock * kReportItems;
    hipLaunchKernelGGL(k_residual_reduce, dim3((c->n + per - 1) / per), dim3(kBlock), 0, c->stream, c->report_err,
                       c->pres_m, c->n, c->params.dt, c->params.target_density, c->report + 4 * kind);
    return hipGetLastError();
}

hipError_t flush_sweep_writes(fe_ctx* c) {
    if (!c->pending_pres && !c->pending_pacc) return hipSuccess;
    Launch L(c, "sweep_writeback");
    hipLaunchKernelGGL(k_sweep_writeback, grid_for(c->n), dim3(kBlock), 0, c->stream, c->sched[0], c->n, c->pres_m,
                       c->pa_m, c->pa2_m, c->pending_pres ? c->pres : nullptr, c->pending_pacc ? c->pacc : nullptr);
    c->pending_pres = c->pending_pacc = false;
    return hipGetLastError();
}

}  // namespace fe
