// fe_sph.hip — neighbour-search plumbing and the streaming (per-particle) DFSPH stages
// on CDNA4: hash, offsets + reorder, both integrations, stats, and transfer helpers.
// Each expression follows the reference shader's FP32 operation order with contraction
// off and IEEE division / square root.
//
// Memory layout (HBM, SOA, 16-byte vec4 rows as in GpuPhysicsData): positions, velocities,
// pressure accelerations and colours are float4[N]; densities, alpha factors, source terms
// and pressures are float[N]; the sorted entries are two uint32[T] arrays (hash, index) and
// each hash bucket's [start, end) is one uint2 in `range` (start = grid_start).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include "fe_device.h"

namespace fe {
namespace {

__device__ __forceinline__ uint32_t gid() { return blockIdx.x * kBlock + threadIdx.x; }

// ---- neighbour search ---------------------------------------------------------------

// spatial_hash.comp:23-36, plus the grid_start fill of neighbor_search.rs:207-208 (the
// table is not read between the hash and the offsets). The cell coordinate
// ivec3(floor(pos / h)) is evaluated as floor(pos * (1 / h)), the form the reference GPU
// executed (it reproduces the bucket order of the reference's rest-density capture).
// In a batch the key is the scene's first table entry plus the scene's hash, so one stable
// sort of all keys orders every scene exactly as its own sort would, within its own range.
__global__ __launch_bounds__(kBlock) void k_hash(const float4* __restrict__ pos, uint32_t n, uint32_t T, Sph s0,
                                                 Scenes sc, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                 uint32_t* __restrict__ keys_alt, uint32_t* __restrict__ vals_alt,
                                                 uint2* __restrict__ range) {
    uint32_t i = gid();
    if (i >= T) return;
    range[i] = make_uint2(kSentinel, 0u);
    if (i < n) {
        const Sph s = scene_at(sc, s0, i);
        float4 p = pos[i];
        keys[i] = s.tab0 + cell_hash((int)floorf(p.x * s.inv_h), (int)floorf(p.y * s.inv_h),
                                     (int)floorf(p.z * s.inv_h), s.mask);
        vals[i] = i;
    } else {
        keys[i] = kSentinel;
        vals[i] = kSentinel;
        keys_alt[i] = kSentinel;
        vals_alt[i] = kSentinel;
    }
}

// Morton key of a particle's cell (the low `bits` bits of each coordinate relative to the
// box's minimum cell, interleaved): sorted with the particle's hash-order index, it orders
// the gathers' work so that neighbouring cells are processed close together in time. In a
// batch the scene number sits above the 3 x bits Morton bits, which keeps scenes apart.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t morton_key(const Scenes& sc, const Sph& s, uint32_t i, const float4& p,
                                               uint32_t bits) {
    const uint32_t mask = (1u << bits) - 1u;
    const int3 c = cell_of(p.x, p.y, p.z, s.inv_h);
    const uint32_t x = (uint32_t)(c.x - s.c0x) & mask, y = (uint32_t)(c.y - s.c0y) & mask,
                   z = (uint32_t)(c.z - s.c0z) & mask;
    return (scene_id(sc, i) << (3u * bits)) | spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

// grid_offsets.comp:20-32 (bucket starts) fused with reorder.comp:37-45 (pos/vel gather),
// the bucket ends the gathers use, and the Morton key of each reordered particle's cell with
// its hash-order index (the input of the gather schedule's sort). Entries at i >= n are
// sentinels.
__global__ __launch_bounds__(kBlock) void k_offsets_reorder(const uint32_t* __restrict__ keys,
                                                            const uint32_t* __restrict__ vals, uint32_t n,
                                                            uint2* __restrict__ range,
                                                            const float4* __restrict__ pos_a,
                                                            const float4* __restrict__ vel_a,
                                                            float4* __restrict__ pos_b, float4* __restrict__ vel_b,
                                                            Sph s0, Scenes sc, uint32_t bits,
                                                            uint32_t* __restrict__ mkeys,
                                                            uint32_t* __restrict__ midx) {
    uint32_t i = gid();
    if (i >= n) return;
    uint32_t key = keys[i];
    if (key != kSentinel) {
        if (i == 0 || keys[i - 1] != key) range[key].x = i;
        if (i + 1 == n || keys[i + 1] != key) range[key].y = i + 1;
    }
    uint32_t src = vals[i];
    const float4 p = pos_a[src];
    pos_b[i] = p;
    vel_b[i] = vel_a[src];
    mkeys[i] = morton_key(sc, scene_at(sc, s0, i), i, p, bits);
    midx[i] = i;
}

__global__ __launch_bounds__(kBlock) void k_iota(uint32_t* __restrict__ p, uint32_t count) {
    const uint32_t i = gid();
    if (i < count) p[i] = i;
}

// grid_start fill + grid_offsets.comp over all T entries (used after a stand-alone sort).
__global__ __launch_bounds__(kBlock) void k_fill_ranges(uint2* __restrict__ range, uint32_t T) {
    uint32_t i = gid();
    if (i < T) range[i] = make_uint2(kSentinel, 0u);
}

__global__ __launch_bounds__(kBlock) void k_offsets_only(const uint32_t* __restrict__ keys, uint32_t T,
                                                         uint2* __restrict__ range) {
    uint32_t i = gid();
    if (i >= T) return;
    uint32_t key = keys[i];
    if (key == kSentinel || key >= T) return;
    uint32_t prev = i == 0 ? kSentinel : keys[i - 1];
    if (key != prev) range[key].x = i;
}

// Bucket ends from starts: the scan `for j = start; j < N && hash[j] == key` of the gathers.
__global__ __launch_bounds__(kBlock) void k_rebuild_ends(const uint32_t* __restrict__ keys, uint32_t n,
                                                         uint32_t T, uint2* __restrict__ range) {
    uint32_t k = gid();
    if (k >= T) return;
    uint32_t s = range[k].x;
    if (s == kSentinel) return;
    uint32_t e = s;
    while (e < n && keys[e] == k) e++;
    range[k].y = e;
}

// ---- streaming stages ----------------------------------------------------------------

// pressure_integration.comp:16-77 (position_b -> position_a; collision order y, x, z).
__device__ __forceinline__ void pressure_integration_one(uint32_t i, const Sph& s0, const Scenes& sc,
                                                         const float4* __restrict__ pos_b,
                                                         const float4* __restrict__ pacc, float4* __restrict__ vel,
                                                         float4* __restrict__ pos_a) {
    const Sph s = scene_at(sc, s0, i);
    const float dt = s.dt, rad = s.rad;
    const float3 mn = sc.box ? make_float3(sc.box[0], sc.box[1], sc.box[2]) : make_float3(s.mnx, s.mny, s.mnz);
    const float3 mx = sc.box ? make_float3(sc.box[3], sc.box[4], sc.box[5]) : make_float3(s.mxx, s.mxy, s.mxz);
    const float4 p = pos_b[i], v = vel[i], a = pacc[i];
    float vx = v.x + a.x * dt, vy = v.y + a.y * dt, vz = v.z + a.z * dt;
    float px = p.x + vx * dt, py = p.y + vy * dt, pz = p.z + vz * dt;
    const float rest = 0.5f, fric = 0.95f, eps = 0.001f;
    if (py < mn.y + rad) {
        py = mn.y + rad + eps; vy *= -rest; vx *= fric; vz *= fric;
    } else if (py > mx.y - rad) {
        py = mx.y - rad - eps; vy *= -rest; vx *= fric; vz *= fric;
    }
    if (px < mn.x + rad) {
        px = mn.x + rad + eps; vx *= -rest; vy *= fric; vz *= fric;
    } else if (px > mx.x - rad) {
        px = mx.x - rad - eps; vx *= -rest; vy *= fric; vz *= fric;
    }
    if (pz < mn.z + rad) {
        pz = mn.z + rad + eps; vz *= -rest; vx *= fric; vy *= fric;
    } else if (pz > mx.z - rad) {
        pz = mx.z - rad - eps; vz *= -rest; vx *= fric; vy *= fric;
    }
    pos_a[i] = make_float4(px, py, pz, 1.0f);
    vel[i] = make_float4(vx, vy, vz, 0.0f);
}

// B threads per block, IPT particles per thread (block-contiguous, so each round coalesces).
template <int B, int IPT>
__global__ __launch_bounds__(B) void k_pressure_integration(uint32_t n, Sph s0, Scenes sc,
                                                            const float4* __restrict__ pos_b,
                                                            const float4* __restrict__ pacc, float4* __restrict__ vel,
                                                            float4* __restrict__ pos_a) {
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const uint32_t i = (blockIdx.x * IPT + k) * B + threadIdx.x;
        if (i < n) pressure_integration_one(i, s0, sc, pos_b, pacc, vel, pos_a);
    }
}

__device__ __forceinline__ float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }

// divergence_integration.comp:14-43.
// divergence_integration.comp: v + a dt and the speed colour ramp of one particle.
__device__ __forceinline__ void divergence_step(float dt, const float4& v, const float4& a, float4& v_out,
                                                float4& colour) {
    float vx = v.x + a.x * dt, vy = v.y + a.y * dt, vz = v.z + a.z * dt;
    float speed = sqrtf(vx * vx + vy * vy + vz * vz);
    float t = speed / 4.0f;
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    float c0, c1, c2;
    if (t < 0.5f) {
        float u = t * 2.0f;
        c0 = mixf(0.05f, 0.2f, u); c1 = mixf(0.1f, 0.5f, u); c2 = mixf(0.5f, 1.0f, u);
    } else {
        float u = (t - 0.5f) * 2.0f;
        c0 = mixf(0.2f, 1.0f, u); c1 = mixf(0.5f, 1.0f, u); c2 = mixf(1.0f, 1.0f, u);
    }
    v_out = make_float4(vx, vy, vz, 0.0f);
    colour = make_float4(c0, c1, c2, 1.0f);
}

__device__ __forceinline__ void divergence_integration_one(uint32_t i, const Sph& s0, const Scenes& sc,
                                                           float4* __restrict__ vel, const float4* __restrict__ pacc,
                                                           float4* __restrict__ colors) {
    float4 v, col;
    divergence_step(scene_at(sc, s0, i).dt, vel[i], pacc[i], v, col);
    vel[i] = v;
    colors[i] = col;
}

template <int B, int IPT>
__global__ __launch_bounds__(B) void k_divergence_integration(uint32_t n, Sph s0, Scenes sc, float4* __restrict__ vel,
                                                              const float4* __restrict__ pacc,
                                                              float4* __restrict__ colors) {
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const uint32_t i = (blockIdx.x * IPT + k) * B + threadIdx.x;
        if (i < n) divergence_integration_one(i, s0, sc, vel, pacc, colors);
    }
}

// The consistent divergence order: the velocity of the particle at sorted index i is
// velocity_b[i]; its corrected value and colour go back to the same particle, index vals[i]
// of velocity_a / colors, and to velocity_b[i] and the xyz of its schedule mirror (whose w
// keeps m / rho).
template <int B, int IPT>
__global__ __launch_bounds__(B) void k_divergence_integration_consistent(
    uint32_t n, Sph s0, Scenes sc, const uint32_t* __restrict__ vals, const uint32_t* __restrict__ inv,
    float4* __restrict__ vel_a, float4* __restrict__ vel_b, float4* __restrict__ vb_m, const float4* __restrict__ pacc,
    float4* __restrict__ colors) {
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const uint32_t i = (blockIdx.x * IPT + k) * B + threadIdx.x;
        if (i >= n) continue;
        float4 v, col;
        divergence_step(scene_at(sc, s0, i).dt, vel_b[i], pacc[i], v, col);
        const uint32_t j = vals[i];
        vel_a[j] = v;
        colors[j] = col;
        vel_b[i] = v;
        if (vb_m) {
            float* m = reinterpret_cast<float*>(vb_m + inv[i]);
            m[0] = v.x;
            m[1] = v.y;
            m[2] = v.z;
        }
    }
}

// Dispatch of a streaming kernel template over the tuned (block, items) configurations.
#define FE_STREAM_DISPATCH(KERNEL, cfg, n, ...)                                                          \
    do {                                                                                                 \
        const uint32_t b_ = (cfg).block, it_ = (cfg).items;                                              \
        const dim3 g_((unsigned)(((n) + b_ * it_ - 1) / (b_ * it_)));                                     \
        switch (b_ * 8 + it_) {                                                                          \
            case 64 * 8 + 1: hipLaunchKernelGGL((KERNEL<64, 1>), g_, dim3(64), 0, __VA_ARGS__); break;     \
            case 64 * 8 + 2: hipLaunchKernelGGL((KERNEL<64, 2>), g_, dim3(64), 0, __VA_ARGS__); break;     \
            case 64 * 8 + 4: hipLaunchKernelGGL((KERNEL<64, 4>), g_, dim3(64), 0, __VA_ARGS__); break;     \
            case 128 * 8 + 1: hipLaunchKernelGGL((KERNEL<128, 1>), g_, dim3(128), 0, __VA_ARGS__); break;  \
            case 128 * 8 + 2: hipLaunchKernelGGL((KERNEL<128, 2>), g_, dim3(128), 0, __VA_ARGS__); break;  \
            case 128 * 8 + 4: hipLaunchKernelGGL((KERNEL<128, 4>), g_, dim3(128), 0, __VA_ARGS__); break;  \
            case 256 * 8 + 2: hipLaunchKernelGGL((KERNEL<256, 2>), g_, dim3(256), 0, __VA_ARGS__); break;  \
            case 256 * 8 + 4: hipLaunchKernelGGL((KERNEL<256, 4>), g_, dim3(256), 0, __VA_ARGS__); break;  \
            case 512 * 8 + 1: hipLaunchKernelGGL((KERNEL<512, 1>), g_, dim3(512), 0, __VA_ARGS__); break;  \
            case 512 * 8 + 2: hipLaunchKernelGGL((KERNEL<512, 2>), g_, dim3(512), 0, __VA_ARGS__); break;  \
            case 512 * 8 + 4: hipLaunchKernelGGL((KERNEL<512, 4>), g_, dim3(512), 0, __VA_ARGS__); break;  \
            case 1024 * 8 + 1: hipLaunchKernelGGL((KERNEL<1024, 1>), g_, dim3(1024), 0, __VA_ARGS__); break; \
            case 1024 * 8 + 2: hipLaunchKernelGGL((KERNEL<1024, 2>), g_, dim3(1024), 0, __VA_ARGS__); break; \
            case 1024 * 8 + 4: hipLaunchKernelGGL((KERNEL<1024, 4>), g_, dim3(1024), 0, __VA_ARGS__); break; \
            default: hipLaunchKernelGGL((KERNEL<256, 1>), g_, dim3(256), 0, __VA_ARGS__); break;           \
        }                                                                                                \
    } while (0)

__device__ __forceinline__ uint32_t to_u32_sat(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

// stats.comp:18-30 as a two-level reduction: a grid-stride pass writes per-block partials
// [max speed bits, density error sum, divergence error sum], then one block combines them
// (which replaces the fill-0 of stats_pipeline.rs:56-66). Max and wrapping u32 sums are
// order-independent, so the result equals the reference's atomics bit for bit.
// With kHealth (the watchdog, no reference counterpart) the pass also counts the particles
// whose position_a, velocity_a, density, source term or pressure is not finite and keeps the
// lowest such index: partials [3 blocks + b] (count) and [4 blocks + b] (index).
template <bool kHealth>
__global__ __launch_bounds__(kBlock) void k_stats(uint32_t n, const float4* __restrict__ vel,
                                                  const float* __restrict__ dens, const float* __restrict__ src,
                                                  float rho0, uint32_t* __restrict__ partial,
                                                  const float4* __restrict__ pos, const float* __restrict__ pres) {
    __shared__ uint32_t part[5][kBlock / 64];
    uint32_t mx = 0, s1 = 0, s2 = 0, bad = 0, first = 0xFFFFFFFFu;
    for (uint32_t i = gid(); i < n; i += gridDim.x * kBlock) {
        const float4 v = vel[i];
        const float speed = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
        mx = max(mx, __float_as_uint(speed));
        const float d = dens[i], sr = src[i];
        s1 += min(to_u32_sat(fabsf(d - rho0) * 1.0f), 10000u);
        s2 += min(to_u32_sat(fabsf(sr) * 10.0f), 10000u);
        if constexpr (kHealth) {
            const float4 p = pos[i];
            const bool fin = isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(v.x) && isfinite(v.y) &&
                             isfinite(v.z) && isfinite(d) && isfinite(sr) && isfinite(pres[i]);
            if (!fin) {
                bad++;
                first = min(first, i);
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
        s1 += (uint32_t)__shfl_xor((int)s1, off, 64);
        s2 += (uint32_t)__shfl_xor((int)s2, off, 64);
        if constexpr (kHealth) {
            bad += (uint32_t)__shfl_xor((int)bad, off, 64);
            first = min(first, (uint32_t)__shfl_xor((int)first, off, 64));
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        part[0][w] = mx;
        part[1][w] = s1;
        part[2][w] = s2;
        part[3][w] = bad;
        part[4][w] = first;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(kBlock / 64); k++) {
            mx = max(mx, part[0][k]);
            s1 += part[1][k];
            s2 += part[2][k];
            bad += part[3][k];
            first = min(first, part[4][k]);
        }
        partial[blockIdx.x] = mx;
        partial[gridDim.x + blockIdx.x] = s1;
        partial[2 * gridDim.x + blockIdx.x] = s2;
        if constexpr (kHealth) {
            partial[3 * gridDim.x + blockIdx.x] = bad;
            partial[4 * gridDim.x + blockIdx.x] = first;
        }
    }
}

// Word 6 of the stats buffer carries the radix sort's error word (a look-back that stopped
// waiting), so the host sees it with the frame's stats.
template <bool kHealth>
__global__ __launch_bounds__(kBlock) void k_stats_final(const uint32_t* __restrict__ partial, uint32_t blocks,
                                                        uint32_t* __restrict__ stats,
                                                        const uint32_t* __restrict__ sort_err) {
    __shared__ uint32_t part[5][kBlock / 64];
    uint32_t mx = 0, s1 = 0, s2 = 0, bad = 0, first = 0xFFFFFFFFu;
    for (uint32_t k = threadIdx.x; k < blocks; k += kBlock) {
        mx = max(mx, partial[k]);
        s1 += partial[blocks + k];
        s2 += partial[2 * blocks + k];
        if constexpr (kHealth) {
            bad += partial[3 * blocks + k];
            first = min(first, partial[4 * blocks + k]);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
        s1 += (uint32_t)__shfl_xor((int)s1, off, 64);
        s2 += (uint32_t)__shfl_xor((int)s2, off, 64);
        if constexpr (kHealth) {
            bad += (uint32_t)__shfl_xor((int)bad, off, 64);
            first = min(first, (uint32_t)__shfl_xor((int)first, off, 64));
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        part[0][w] = mx;
        part[1][w] = s1;
        part[2][w] = s2;
        part[3][w] = bad;
        part[4][w] = first;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(kBlock / 64); k++) {
            mx = max(mx, part[0][k]);
            s1 += part[1][k];
            s2 += part[2][k];
            bad += part[3][k];
            first = min(first, part[4][k]);
        }
        if constexpr (kHealth) {
            stats[4] = bad;
            stats[5] = first;
        }
        stats[6] = *sort_err;
        stats[0] = mx;
        stats[1] = s1;
        stats[2] = s2;
    }
}

// stats.comp:18-30 for a batch: each scene's three values by atomics (max and wrapping u32
// sums are order-independent, as in the reference). A scene's particles are contiguous in
// index, so a thread folds its kStatsItems particles (strided by the block, for coalesced
// loads) into running values, adding them to the scene's words only where the scene changes;
// a wave whose lanes end in one scene adds once, and a block whose waves all do so adds
// once. That keeps the atomics on the few words of each scene to about three per block
// (one set per wave put 175,000 atomics on 192 words for 3.7M particles and took 0.83 ms).
// The per-scene triples are zeroed first (k_fill_u32).
constexpr uint32_t kStatsItems = 16;

__device__ __forceinline__ void add_scene_stats(uint32_t* __restrict__ stats, uint32_t scene, uint32_t mx,
                                                uint32_t s1, uint32_t s2) {
    atomicMax(stats + 3 * scene, mx);
    atomicAdd(stats + 3 * scene + 1, s1);
    atomicAdd(stats + 3 * scene + 2, s2);
}

__global__ __launch_bounds__(kBlock) void k_stats_scenes(uint32_t n, Scenes sc, const float4* __restrict__ vel,
                                                         const float* __restrict__ dens,
                                                         const float* __restrict__ src,
                                                         uint32_t* __restrict__ stats) {
    constexpr uint32_t kWaves = kBlock / 64, kMixed = kSentinel - 1u;
    __shared__ uint32_t wsc[kWaves], wmx[kWaves], ws1[kWaves], ws2[kWaves];
    const uint32_t base = blockIdx.x * (kBlock * kStatsItems) + threadIdx.x;
    uint32_t mx = 0, s1 = 0, s2 = 0, cur = kSentinel;   // kSentinel: no particle yet
    float rho0 = 0.0f;
    for (uint32_t k = 0; k < kStatsItems; k++) {
        const uint32_t i = base + k * kBlock;
        if (i >= n) break;
        const uint32_t scene = sc.sid[i];
        if (scene != cur) {
            if (cur != kSentinel) add_scene_stats(stats, cur, mx, s1, s2);
            cur = scene;
            rho0 = sc.tab[scene].rho0;
            mx = s1 = s2 = 0;
        }
        const float4 v = vel[i];
        const float speed = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
        mx = max(mx, __float_as_uint(speed));
        s1 += min(to_u32_sat(fabsf(dens[i] - rho0) * 1.0f), 10000u);
        s2 += min(to_u32_sat(fabsf(src[i]) * 10.0f), 10000u);
    }
    const uint32_t first = __shfl(cur, 0, 64);
    const bool uni = __all(cur == first);   // the wave's lanes end in one scene (or none)
    if (uni) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
            s1 += (uint32_t)__shfl_xor((int)s1, off, 64);
            s2 += (uint32_t)__shfl_xor((int)s2, off, 64);
        }
    }
    const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63u;
    if (l == 0) {
        wsc[w] = uni ? first : kMixed;
        wmx[w] = mx;
        ws1[w] = s1;
        ws2[w] = s2;
    }
    __syncthreads();
    bool block_uni = wsc[0] < kMixed;
#pragma unroll
    for (uint32_t u = 1; u < kWaves; u++) block_uni = block_uni && wsc[u] == wsc[0];
    if (block_uni) {
        if (threadIdx.x == 0) {
            uint32_t bm = wmx[0], b1 = ws1[0], b2 = ws2[0];
            for (uint32_t u = 1; u < kWaves; u++) {
                bm = max(bm, wmx[u]);
                b1 += ws1[u];
                b2 += ws2[u];
            }
            add_scene_stats(stats, wsc[0], bm, b1, b2);
        }
    } else if (uni) {
        if (l == 0 && first != kSentinel) add_scene_stats(stats, first, mx, s1, s2);
    } else if (cur != kSentinel) {
        add_scene_stats(stats, cur, mx, s1, s2);
    }
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* p, uint32_t count) {
    uint32_t i = gid();
    if (i < count) p[i] = 0u;
}

__global__ __launch_bounds__(kBlock) void k_fill_u64(unsigned long long* p, uint32_t count) {
    uint32_t i = gid();
    if (i < count) p[i] = 0ull;
}

__global__ __launch_bounds__(kBlock) void k_pack_entries(const uint32_t* __restrict__ keys,
                                                         const uint32_t* __restrict__ vals, uint32_t T,
                                                         uint2* __restrict__ out) {
    uint32_t i = gid();
    if (i < T) out[i] = make_uint2(keys[i], vals[i]);
}

__global__ __launch_bounds__(kBlock) void k_unpack_entries(const uint2* __restrict__ in, uint32_t T,
                                                           uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t i = gid();
    if (i < T) {
        uint2 e = in[i];
        keys[i] = e.x;
        vals[i] = e.y;
    }
}

__global__ __launch_bounds__(kBlock) void k_pack_start(const uint2* __restrict__ range, uint32_t T,
                                                       uint32_t* __restrict__ out) {
    uint32_t i = gid();
    if (i < T) out[i] = range[i].x;
}

__global__ __launch_bounds__(kBlock) void k_unpack_start(const uint32_t* __restrict__ in, uint32_t T,
                                                         uint2* __restrict__ range) {
    uint32_t i = gid();
    if (i < T) range[i] = make_uint2(in[i], 0u);
}

inline dim3 grid_for(uint32_t count) { return dim3((count + kBlock - 1) / kBlock); }


}  // namespace

hipError_t launch_hash(fe_ctx* c) {
    Launch L(c, "spatial_hash");
    hipLaunchKernelGGL(k_hash, grid_for(c->T), dim3(kBlock), 0, c->stream, c->pos_a, c->n, c->T, sph_of(c),
                       scenes_of(c), c->keys[0], c->vals[0], c->keys[1], c->vals[1], c->range);
    return hipGetLastError();
}

hipError_t launch_offsets_reorder(fe_ctx* c) {
    Launch L(c, "offsets_reorder");
    hipLaunchKernelGGL(k_offsets_reorder, grid_for(c->n), dim3(kBlock), 0, c->stream, c->keys[0], c->vals[0], c->n,
                       c->range, c->pos_a, c->vel_a, c->pos_b, c->vel_b, sph_of(c), scenes_of(c), c->morton_bits / 3,
                       c->mkeys[0], c->sched[0]);
    return hipGetLastError();
}

namespace {

// Whether the residual-corrected product equals the IEEE quotient r / h for every float r in
// [lo, lo + count): any mismatch sets *bad.
__global__ __launch_bounds__(kBlock) void k_check_div_h(float h, float inv_h, uint32_t lo, uint32_t count,
                                                        uint32_t* __restrict__ bad) {
    const uint32_t stride = gridDim.x * kBlock;
    uint32_t miss = 0;
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < count; k += stride) {
        const float r = __uint_as_float(lo + k);
        const float q = r / h;
        const float q0 = r * inv_h;
        const float e = fmaf(-q0, h, r);
        const float f = fmaf(e, inv_h, q0);
        miss |= __float_as_uint(q) != __float_as_uint(f);
    }
    if (__any(miss) && (threadIdx.x & 63) == 0) atomicOr(bad, 1u);
}

// The gradient division gf / (h r) over the pairs a scene can produce, 1e-6 <= r <= 1.001 h:
// h r within [1e-6 h, 1.001 h^2], |gf| within [min(3 k 1e-6 / h, 6 k 2^-48), 2 k] with
// k = 8 / (pi h^3), all inside [2^-40, 2^40) with a factor-4 margin.
bool div_box_ok(float h) {
    const double hd = h, k = 8.0 / (3.14159265358979323846 * hd * hd * hd);
    const double lo = std::ldexp(1.0, -38), hi = std::ldexp(1.0, 38);
    const double vals[] = {1e-6 * hd, 1.001 * hd * hd, std::min(3.0 * k * 1e-6 / hd, 6.0 * k * std::ldexp(1.0, -48)),
                           2.0 * k};
    for (double v : vals)
        if (!(v >= lo && v <= hi)) return false;
    return true;
}

std::mutex g_div_h_mutex;
std::unordered_map<uint32_t, bool> g_div_h_ok;   // by the bits of h

}  // namespace

namespace {

// Mismatch counts of the exact shortcuts against the IEEE operations (fe_check_exact_arith).
__global__ __launch_bounds__(kBlock) void k_check_sqrt(uint32_t lo, uint32_t count,
                                                       unsigned long long* __restrict__ bad) {
    const uint32_t stride = gridDim.x * kBlock;
    unsigned long long miss = 0;
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < count; k += stride) {
        const float x = __uint_as_float(lo + k);
        miss += __float_as_uint(sqrt_corrected(x)) != __float_as_uint(sqrtf(x));
    }
    if (miss) atomicAdd(bad, miss);
}

__global__ __launch_bounds__(kBlock) void k_check_div(unsigned long long count, unsigned long long* __restrict__ bad) {
    const unsigned long long stride = (unsigned long long)gridDim.x * kBlock;
    unsigned long long miss = 0;
    for (unsigned long long k = blockIdx.x * kBlock + threadIdx.x; k < count; k += stride) {
        // splitmix64 of the sample index: signs, exponents in [-40, 39], mantissas
        unsigned long long z = (k + 1ull) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const uint32_t a = (uint32_t)z, b = (uint32_t)(z >> 32);
        const uint32_t ea = 87u + (a >> 23) % 80u, eb = 87u + (b >> 23) % 80u;   // biased 87..166
        const float n = __uint_as_float(((a & 1u) << 31) | (ea << 23) | (a & 0x7FFFFFu));
        const float d = __uint_as_float(((b & 1u) << 31) | (eb << 23) | (b & 0x7FFFFFu));
        miss += __float_as_uint(div_boxed(n, d)) != __float_as_uint(n / d);
    }
    if (miss) atomicAdd(bad, miss);
}

__global__ __launch_bounds__(kBlock) void k_check_div_h_count(float h, float inv_h, uint32_t lo, uint32_t count,
                                                              unsigned long long* __restrict__ bad) {
    const uint32_t stride = gridDim.x * kBlock;
    unsigned long long miss = 0;
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < count; k += stride) {
        const float r = __uint_as_float(lo + k);
        const float q0 = r * inv_h;
        const float e = fmaf(-q0, h, r);
        miss += __float_as_uint(fmaf(e, inv_h, q0)) != __float_as_uint(r / h);
    }
    if (miss) atomicAdd(bad, miss);
}

}  // namespace

hipError_t check_exact_arith(hipStream_t st, float h, unsigned long long samples, unsigned long long* dev_bad,
                             unsigned long long counts[6]) {
    const float sq_lo_f = 0x1p-96f, sq_hi_f = 0x1p16f, r_lo_f = 1e-6f, r_hi_f = h * 1.001f;
    uint32_t sq_lo, sq_hi, r_lo, r_hi;
    std::memcpy(&sq_lo, &sq_lo_f, 4);
    std::memcpy(&sq_hi, &sq_hi_f, 4);
    std::memcpy(&r_lo, &r_lo_f, 4);
    std::memcpy(&r_hi, &r_hi_f, 4);
    counts[0] = sq_hi - sq_lo;
    counts[2] = samples;
    counts[4] = (h > 1e-6f && std::isfinite(r_hi_f)) ? (unsigned long long)(r_hi - r_lo + 1u) : 0ull;
    hipError_t e = hipMemsetAsync(dev_bad, 0, 3 * sizeof(unsigned long long), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_check_sqrt, dim3(4096), dim3(kBlock), 0, st, sq_lo, (uint32_t)counts[0], dev_bad);
    if (samples) hipLaunchKernelGGL(k_check_div, dim3(4096), dim3(kBlock), 0, st, samples, dev_bad + 1);
    if (counts[4])
        hipLaunchKernelGGL(k_check_div_h_count, dim3(4096), dim3(kBlock), 0, st, h, 1.0f / h, r_lo,
                           (uint32_t)counts[4], dev_bad + 2);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    unsigned long long bad[3];
    e = hipMemcpyAsync(bad, dev_bad, sizeof(bad), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    counts[1] = bad[0];
    counts[3] = bad[1];
    counts[5] = bad[2];
    return hipSuccess;
}

hipError_t exact_math_flags(fe_ctx* c, const fe_params& p, uint32_t* flags) {
    const float h = p.smoothing_radius;
    uint32_t hb;
    std::memcpy(&hb, &h, 4);
    bool ok_h = false, known = false;
    {
        std::lock_guard<std::mutex> lk(g_div_h_mutex);
        auto it = g_div_h_ok.find(hb);
        if (it != g_div_h_ok.end()) {
            ok_h = it->second;
            known = true;
        }
    }
    if (!known) {
        const float lo_f = 1e-6f, hi_f = h * 1.001f;
        uint32_t lo, hi;
        std::memcpy(&lo, &lo_f, 4);
        std::memcpy(&hi, &hi_f, 4);
        if (!(h > 1e-6f) || !std::isfinite(hi_f)) {
            ok_h = false;
        } else {
            uint32_t* bad = nullptr;
            hipError_t e = hipMalloc((void**)&bad, sizeof(uint32_t));
            if (e != hipSuccess) return e;
            uint32_t host_bad = 1u;
            e = hipMemsetAsync(bad, 0, sizeof(uint32_t), c->stream);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_check_div_h, dim3(2048), dim3(kBlock), 0, c->stream, h, 1.0f / h, lo, hi - lo + 1u,
                                   bad);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipMemcpyAsync(&host_bad, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            (void)hipFree(bad);
            if (e != hipSuccess) return e;
            ok_h = host_bad == 0u;
        }
        std::lock_guard<std::mutex> lk(g_div_h_mutex);
        g_div_h_ok[hb] = ok_h;
    }
    *flags = (ok_h ? kFastDivH : 0u) | (div_box_ok(h) ? kFastDivBox : 0u);
    return hipSuccess;
}

hipError_t launch_iota(fe_ctx* c, uint32_t* p, uint32_t count) {
    hipLaunchKernelGGL(k_iota, grid_for(count), dim3(kBlock), 0, c->stream, p, count);
    return hipGetLastError();
}

hipError_t launch_fill_ranges(fe_ctx* c) {
    Launch L(c, "fill_ranges");
    hipLaunchKernelGGL(k_fill_ranges, grid_for(c->T), dim3(kBlock), 0, c->stream, c->range, c->T);
    return hipGetLastError();
}

hipError_t launch_offsets_only(fe_ctx* c) {
    Launch L(c, "grid_offsets");
    hipLaunchKernelGGL(k_offsets_only, grid_for(c->T), dim3(kBlock), 0, c->stream, c->keys[0], c->T, c->range);
    return hipGetLastError();
}

hipError_t launch_rebuild_ends(fe_ctx* c) {
    Launch L(c, "rebuild_ends");
    hipLaunchKernelGGL(k_rebuild_ends, grid_for(c->T), dim3(kBlock), 0, c->stream, c->keys[0], c->n, c->T, c->range);
    return hipGetLastError();
}






hipError_t launch_pressure_integration(fe_ctx* c) {
    c->mirrors &= ~kMirrorVelA;
    c->ns_fresh = false;   // position_a moves
    c->da_fresh = false;
    Launch L(c, "pressure_integration");
    FE_STREAM_DISPATCH(k_pressure_integration, c->tune[kTunePressureIntegration], c->n, c->stream, c->n, sph_of(c),
                       scenes_of(c), c->pos_b, c->pacc, c->vel_a, c->pos_a);
    return hipGetLastError();
}


hipError_t launch_divergence_integration(fe_ctx* c) {
    c->mirrors &= ~kMirrorVelA;
    Launch L(c, "divergence_integration");
    if (c->options.divergence_order == FE_DIVERGENCE_CONSISTENT) {
        // vals and inv come from the last neighbour search (the callers check it); the
        // velocity_b mirror is updated in place when it is current.
        float4* vbm = (c->mirrors & kMirrorVelB) ? c->vb_m : nullptr;
        FE_STREAM_DISPATCH(k_divergence_integration_consistent, c->tune[kTuneDivergenceIntegration], c->n, c->stream,
                           c->n, sph_of(c), scenes_of(c), c->vals[0], c->inv, c->vel_a, c->vel_b, vbm, c->pacc,
                           c->colors);
    } else {
        FE_STREAM_DISPATCH(k_divergence_integration, c->tune[kTuneDivergenceIntegration], c->n, c->stream, c->n,
                           sph_of(c), scenes_of(c), c->vel_a, c->pacc, c->colors);
    }
    return hipGetLastError();
}

hipError_t launch_stats(fe_ctx* c) {
    if (c->nscenes > 1) {
        hipLaunchKernelGGL(k_fill_u32, grid_for(3 * c->nscenes), dim3(kBlock), 0, c->stream, c->scene_stats,
                           3 * c->nscenes);
        Launch L(c, "stats");
        const uint32_t per = kBlock * kStatsItems;
        hipLaunchKernelGGL(k_stats_scenes, dim3((c->n + per - 1) / per), dim3(kBlock), 0, c->stream, c->n,
                           scenes_of(c), c->vel_a, c->dens, c->src, c->scene_stats);
        return hipGetLastError();
    }
    uint32_t blocks = (c->n + kBlock - 1) / kBlock;
    if (blocks > kStatsBlocks) blocks = kStatsBlocks;
    {
        Launch L(c, "stats");
        if (c->watchdog)
            hipLaunchKernelGGL(k_stats<true>, dim3(blocks), dim3(kBlock), 0, c->stream, c->n, c->vel_a, c->dens, c->src,
                               c->params.target_density, c->stats_partial, c->pos_a, c->pres);
        else
            hipLaunchKernelGGL(k_stats<false>, dim3(blocks), dim3(kBlock), 0, c->stream, c->n, c->vel_a, c->dens,
                               c->src, c->params.target_density, c->stats_partial, c->pos_a, c->pres);
    }
    Launch L(c, "stats_final");
    const uint32_t* sort_err = c->sort.ctrl + kSortMaxPasses;
    if (c->watchdog)
        hipLaunchKernelGGL(k_stats_final<true>, dim3(1), dim3(kBlock), 0, c->stream, c->stats_partial, blocks, c->stats,
                           sort_err);
    else
        hipLaunchKernelGGL(k_stats_final<false>, dim3(1), dim3(kBlock), 0, c->stream, c->stats_partial, blocks,
                           c->stats, sort_err);
    return hipGetLastError();
}

hipError_t launch_fill_u64(fe_ctx* c, unsigned long long* p, uint32_t count) {
    Launch L(c, "fill_u64");
    hipLaunchKernelGGL(k_fill_u64, grid_for(count), dim3(kBlock), 0, c->stream, p, count);
    return hipGetLastError();
}

hipError_t launch_pack_entries(fe_ctx* c, uint32_t* out) {
    hipLaunchKernelGGL(k_pack_entries, grid_for(c->T), dim3(kBlock), 0, c->stream, c->keys[0], c->vals[0], c->T,
                       (uint2*)out);
    return hipGetLastError();
}

hipError_t launch_unpack_entries(fe_ctx* c, const uint32_t* in) {
    hipLaunchKernelGGL(k_unpack_entries, grid_for(c->T), dim3(kBlock), 0, c->stream, (const uint2*)in, c->T,
                       c->keys[0], c->vals[0]);
    return hipGetLastError();
}

hipError_t launch_pack_start(fe_ctx* c, uint32_t* out) {
    hipLaunchKernelGGL(k_pack_start, grid_for(c->T), dim3(kBlock), 0, c->stream, c->range, c->T, out);
    return hipGetLastError();
}

const char* tune_name(int id) {
    static const char* names[kTuneCount] = {"density_alpha",     "viscosity",      "density_source",
                                            "divergence_source", "pressure_force", "pressure_update",
                                            "build_list",        "pressure_integration", "divergence_integration",
                                            "density_alpha_divergence_source"};
    return id >= 0 && id < kTuneCount ? names[id] : "";
}

bool tune_valid(int id, uint32_t block, uint32_t items) {
    const bool pow2 = block == 64 || block == 128 || block == 256 || block == 512 || block == 1024;
    // build_list: items 1 = 64 schedule positions per wave, 8 = 8 per wave
    if (id == kTuneBuildList) return (block == 64 || block == 128 || block == 256) && (items == 1 || items == 8);
    if (id == kTunePressureIntegration || id == kTuneDivergenceIntegration)
        return pow2 && (items == 1 || items == 2 || items == 4);
    // the list gathers: items 1 gathers each listed pair from the mirrors with 4 loads in
    // flight, items 4 with 16, items 8 with four lanes per particle (one scene, blocks of 64 to
    // 256); for the sweeps items 2 stages one window per wave in LDS (1, 2 or 4 windows per
    // block)
    const bool small = block == 64 || block == 128 || block == 256;
    if (id == kTunePressureForce || id == kTunePressureUpdate)
        return (pow2 && items == 1) || (small && (items == 2 || items == 4 || items == 8));
    return id >= 0 && id < kTuneCount && ((pow2 && items == 1) || (small && (items == 4 || items == 8)));
}

hipError_t launch_unpack_start(fe_ctx* c, const uint32_t* in) {
    hipLaunchKernelGGL(k_unpack_start, grid_for(c->T), dim3(kBlock), 0, c->stream, in, c->T, c->range);
    return hipGetLastError();
}

}  // namespace fe
