// fe_splat.hip — the render hand-off: the raymarching density texture of splat_density.comp
// (shaders/compute/splat_density.comp:11-51), cleared and rebuilt from position_a
// (src/renderer/mod.rs:386-391).
//
// The shader scatters: each particle inside the box adds uint(200 (1 - (d/8)^2)^3) to every
// texel at distance d < 8 of its centre texel, with one image atomic per texel (about 2145
// per particle). The contribution depends only on the integer offset, so the texture is
// the centre counts convolved with a fixed integer stencil:
//     tex[v] = sum over offsets o with |o|^2 <= 53 of w[|o|^2] * count[v - o]  (mod 2^32),
// which equals the sum of the atomics exactly (unsigned addition modulo 2^32 is associative
// and commutative; offsets with |o|^2 in [54, 63] contribute uint(...) = 0). Two kernels:
//   k_splat_count  the centre-count grid (grid_res + 1 per axis, since uvw = 1 puts the centre
//                  one past the last texel): the lanes of a wave with the same centre texel
//                  add their number with one atomic, and mark the 8^3 occupancy block;
//   k_splat_gather one workgroup per 32 x 8 x 4 texel tile: the tile's centre counts with a
//                  7-texel halo (no offset with |o|^2 <= 53 has a component of 8), 46 x 22 x 18
//                  words (71 KB, two workgroups per CU), are staged in LDS with x fastest, so
//                  the 32 lanes of a half-wave read 32 consecutive words. Each lane forms 4
//                  z-consecutive texels. The stencil is unrolled at compile time: its (dx, dy)
//                  columns are taken in groups of equal dx^2 + dy^2, whose weights
//                  w[dx^2 + dy^2 + dz^2] agree; a group's 2m + 4 counts per column are summed
//                  per z, then multiplied by the group's weights (per lane and tile: 2178 LDS
//                  words, 2178 adds and 1084 multiply-adds instead of 6588 multiply-adds). Column
//                  offsets are immediate LDS offsets and the weights are constants. Tiles whose
//                  halo holds no particle write zeros.
// The constant weights are checked, before the first launch, against the host evaluation of
// the shader's FP32 operation sequence.
#include <cmath>
#include <utility>

#include "fe_device.h"

namespace fe {
namespace {

constexpr int kR = 8;                 // splat radius in texels (splat_density.comp:30)
constexpr int kMaxD2 = 53;            // largest |offset|^2 with a non-zero contribution
constexpr int kH = 7;                 // largest offset component among those: the halo width
constexpr int kTX = 32, kTY = 8, kTZ = 4;
constexpr int kHX = kTX + 2 * kH, kHY = kTY + 2 * kH, kHZ = kTZ + 2 * kH;   // 46, 22, 18
static_assert(kH * kH <= kMaxD2 && (kH + 1) * (kH + 1) > kMaxD2, "halo width");
constexpr int kPlane = kHX * kHY;
constexpr int kHalo = kPlane * kHZ;
constexpr int kOcc = 8;               // occupancy block edge, in centre texels
constexpr uint32_t kThreads = kTX * kTY;

// uint(200 (1 - (sqrt(d2) / 8)^2)^3) for d2 = 0 .. 53 in the shader's FP32 sequence (zero for
// d2 = 54 .. 63); splat_weights() recomputes it and the launcher refuses to run on a mismatch.
constexpr uint32_t kWeights[kMaxD2 + 1] = {200, 190, 181, 173, 164, 156, 148, 141, 133, 126, 120, 113, 107, 101,
                                           95,  89,  84,  79,  74,  69,  64,  60,  56,  52,  48,  45,  41,  38,
                                           35,  32,  29,  27,  25,  22,  20,  18,  16,  15,  13,  11,  10,  9,
                                           8,   7,   6,   5,   4,   3,   3,   2,   2,   1,   1,   1};

// The columns (dx, dy) with dx^2 + dy^2 = s2 and their half-height m (largest |dz| with
// s2 + dz^2 <= 53).
struct Columns {
    int n = 0, m = 0;
    int dx[12] = {}, dy[12] = {};
};
constexpr Columns columns_of(int s2) {
    Columns c;
    while ((c.m + 1) * (c.m + 1) + s2 <= kMaxD2) c.m++;
    for (int dy = -kH; dy <= kH; dy++)
        for (int dx = -kH; dx <= kH; dx++)
            if (dx * dx + dy * dy == s2) {
                c.dx[c.n] = dx;
                c.dy[c.n] = dy;
                c.n++;
            }
    return c;
}
constexpr int count_taps() {
    int t = 0;
    for (int s2 = 0; s2 <= kMaxD2; s2++) t += columns_of(s2).n * (2 * columns_of(s2).m + 1);
    return t;
}
static_assert(count_taps() == 1647, "the stencil has 1647 non-zero offsets");

struct SplatGeom {
    int32_t rx, ry, rz;      // grid_res
    float mnx, mny, mnz;     // box_min
    float ix, iy, iz;        // 1 / (box_max - box_min)
    uint32_t ox, oy, oz;     // occupancy blocks per axis
};

__global__ __launch_bounds__(256) void k_splat_clear(uint4* __restrict__ p, size_t quads) {
    for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < quads; i += (size_t)gridDim.x * 256u)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// splat_density.comp:12-28: the particle's uvw and centre texel. A NaN uvw (no defined
// result in the shader) is skipped. The particles arrive in hash order, so the lanes of a
// wave share few centre texels: each distinct texel takes one atomic of its lane count.
__global__ __launch_bounds__(256) void k_splat_count(const float4* __restrict__ pos, uint32_t n, SplatGeom g,
                                                     uint32_t* __restrict__ count, uint32_t* __restrict__ occ) {
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    uint32_t idx = kNone, blk = 0u;
    if (i < n) {
        const float4 p = pos[i];
        const float u = (p.x - g.mnx) * g.ix, v = (p.y - g.mny) * g.iy, w = (p.z - g.mnz) * g.iz;
        if (u >= 0.0f && u <= 1.0f && v >= 0.0f && v <= 1.0f && w >= 0.0f && w <= 1.0f) {
            const uint32_t cx = (uint32_t)(int32_t)(u * (float)g.rx), cy = (uint32_t)(int32_t)(v * (float)g.ry),
                           cz = (uint32_t)(int32_t)(w * (float)g.rz);
            idx = (cz * (uint32_t)(g.ry + 1) + cy) * (uint32_t)(g.rx + 1) + cx;
            blk = ((cz / kOcc) * g.oy + cy / kOcc) * g.ox + cx / kOcc;
        }
    }
    unsigned long long todo = __ballot(idx != kNone);
    while (todo) {
        const int leader = __builtin_ctzll(todo);
        const uint32_t lidx = (uint32_t)__builtin_amdgcn_readlane((int)idx, leader);
        const unsigned long long same = __ballot(idx == lidx);
        if ((int)__lane_id() == leader) {
            atomicAdd(&count[lidx], (uint32_t)__popcll(same));
            occ[blk] = 1u;
        }
        todo &= ~same;
    }
}

template <bool kWide>
__device__ __forceinline__ uint32_t mac(uint32_t w, uint32_t c, uint32_t acc) {
    if constexpr (kWide) return w * c + acc;
    else return __umul24(w, c) + acc;
}

// One group (all columns with dx^2 + dy^2 = S2) for the lane's 4 texels z0 .. z0 + 3. `base`
// is the lane's word for (dx, dy) = (7, 7) in halo plane 0, so every column's count at halo
// z is base + z * plane + (7 - dy) * 46 + (7 - dx): a constant offset.
template <int S2, bool kWide>
__device__ __forceinline__ void group(const uint32_t* __restrict__ base, uint32_t acc[kTZ]) {
    constexpr Columns C = columns_of(S2);
    if constexpr (C.n > 0) {
        constexpr int M = C.m;
        uint32_t sum[2 * M + kTZ];
#pragma unroll
        for (int j = 0; j < 2 * M + kTZ; j++) sum[j] = 0u;
#pragma unroll
        for (int c = 0; c < C.n; c++) {
            const uint32_t* col = base + (kH - C.dy[c]) * kHX + (kH - C.dx[c]);
#pragma unroll
            for (int j = 0; j < 2 * M + kTZ; j++) sum[j] += col[(kH - M + j) * kPlane];
        }
#pragma unroll
        for (int dz = -M; dz <= M; dz++) {
            const uint32_t w = kWeights[S2 + dz * dz];
#pragma unroll
            for (int k = 0; k < kTZ; k++) acc[k] = mac<kWide>(w, sum[k - dz + M], acc[k]);
        }
        // Close the group: the empty asm makes the accumulators opaque, so the compiler does
        // not re-associate the whole 1647-term sum across groups (which keeps hundreds of
        // partial products live), and the scheduler does not hoist later groups' loads.
#pragma unroll
        for (int k = 0; k < kTZ; k++) asm volatile("" : "+v"(acc[k]));
        __builtin_amdgcn_sched_barrier(0);
    }
}

template <bool kWide, int... S>
__device__ __forceinline__ void stencil(const uint32_t* __restrict__ base, uint32_t acc[kTZ],
                                        std::integer_sequence<int, S...>) {
    (group<S, kWide>(base, acc), ...);
}

template <bool kWide>
__global__ __launch_bounds__(kThreads) void k_splat_gather(const uint32_t* __restrict__ count,
                                                           const uint32_t* __restrict__ occ, SplatGeom g,
                                                           uint32_t tiles_x, uint32_t tiles_y,
                                                           uint32_t* __restrict__ tex) {
    __shared__ uint32_t halo[kHalo];
    // Tiles in launch order: the dispatcher deals consecutive blocks to the 8 XCDs in turn,
    // which spreads the busy tiles (those near the fluid) over every XCD.
    const uint32_t t = blockIdx.x;
    const int32_t X0 = (int32_t)(t % tiles_x) * kTX, Y0 = (int32_t)((t / tiles_x) % tiles_y) * kTY,
                  Z0 = (int32_t)(t / (tiles_x * tiles_y)) * kTZ;
    const int tid = (int)threadIdx.x;
    const int x = tid % kTX, y = tid / kTX;

    // Occupancy of the centre region the tile reads: [X0 - 7, X0 + 38] x [Y0 - 7, Y0 + 14] x
    // [Z0 - 7, Z0 + 10], clipped to the count grid [0, r] per axis.
    const int bx0 = max(X0 - kH, 0) / kOcc, bx1 = min(X0 + kTX - 1 + kH, g.rx) / kOcc;
    const int by0 = max(Y0 - kH, 0) / kOcc, by1 = min(Y0 + kTY - 1 + kH, g.ry) / kOcc;
    const int bz0 = max(Z0 - kH, 0) / kOcc, bz1 = min(Z0 + kTZ - 1 + kH, g.rz) / kOcc;
    const int nbx = bx1 - bx0 + 1, nby = by1 - by0 + 1, nbz = bz1 - bz0 + 1;
    int mine = 0;
    for (int k = tid; k < nbx * nby * nbz; k += (int)kThreads) {
        const int bx = bx0 + k % nbx, by = by0 + (k / nbx) % nby, bz = bz0 + k / (nbx * nby);
        mine |= occ[((uint32_t)bz * g.oy + (uint32_t)by) * g.ox + (uint32_t)bx] != 0u;
    }
    const bool busy = __syncthreads_or(mine) != 0;

    uint32_t acc[kTZ] = {0u, 0u, 0u, 0u};
    if (busy) {
        // Stage the counts plane by plane: thread tid covers plane positions tid + 256 q,
        // q = 0 .. 3 (1012 per plane); all 72 loads are issued before the LDS stores.
        constexpr int kPasses = (kPlane + (int)kThreads - 1) / (int)kThreads;   // 4
        const uint32_t sx = (uint32_t)(g.rx + 1), sxy = sx * (uint32_t)(g.ry + 1);
        int64_t off[kPasses];
        bool ok[kPasses];
#pragma unroll
        for (int q = 0; q < kPasses; q++) {
            const int p = tid + (int)kThreads * q;
            const int hx = p % kHX, hy = p / kHX;
            const int cx = X0 - kH + hx, cy = Y0 - kH + hy;
            ok[q] = p < kPlane && cx >= 0 && cy >= 0 && cx <= g.rx && cy <= g.ry;
            off[q] = (int64_t)cy * sx + cx;
        }
        uint32_t v[kHZ][kPasses];
#pragma unroll
        for (int hz = 0; hz < kHZ; hz++) {
            const int cz = Z0 - kH + hz;
            const bool okz = cz >= 0 && cz <= g.rz;
#pragma unroll
            for (int q = 0; q < kPasses; q++)
                v[hz][q] = (okz && ok[q]) ? count[(size_t)cz * sxy + (size_t)off[q]] : 0u;
        }
#pragma unroll
        for (int hz = 0; hz < kHZ; hz++)
#pragma unroll
            for (int q = 0; q < kPasses; q++) {
                const int p = tid + (int)kThreads * q;
                if (p < kPlane) halo[hz * kPlane + p] = v[hz][q];
            }
        __syncthreads();
        stencil<kWide>(halo + y * kHX + x, acc, std::make_integer_sequence<int, kMaxD2 + 1>{});
    }
    const int X = X0 + x, Y = Y0 + y;
    if (X >= g.rx || Y >= g.ry) return;
#pragma unroll
    for (int k = 0; k < kTZ; k++) {
        const int Z = Z0 + k;
        if (Z < g.rz) tex[((size_t)Z * (uint32_t)g.ry + (uint32_t)Y) * (uint32_t)g.rx + (uint32_t)X] = acc[k];
    }
}

}  // namespace

// splat_density.comp:36-44 per squared offset d2 = x^2 + y^2 + z^2 (the float sum of the
// integer squares is exact): dist = sqrt(d2); if dist < 8: weight = (1 - (dist / 8)^2)^3,
// contribution uint(weight * 200), each step rounded to float in the shader's order.
void splat_weights(uint32_t wt[64]) {
    const float r = (float)kR;
    for (int d2 = 0; d2 < 64; d2++) {
        const float dist = std::sqrt((float)d2);
        uint32_t add = 0u;
        if (dist < r) {
            const float xn = dist / r;
            float w = 1.0f - xn * xn;
            w = w * w * w;
            add = (uint32_t)(w * 200.0f);
        }
        wt[d2] = add;
    }
}

hipError_t launch_splat_density(fe_ctx* c) {
    const int32_t rx = c->params.grid_res[0], ry = c->params.grid_res[1], rz = c->params.grid_res[2];
    const size_t cells = (size_t)rx * (size_t)ry * (size_t)rz;
    const size_t centres = (size_t)(rx + 1) * (size_t)(ry + 1) * (size_t)(rz + 1);
    const uint32_t ox = (uint32_t)(rx + kOcc) / kOcc, oy = (uint32_t)(ry + kOcc) / kOcc, oz = (uint32_t)(rz + kOcc) / kOcc;
    const size_t occs = (size_t)ox * oy * oz;
    const size_t quads = (centres + occs + 3) / 4;   // the count buffer is cleared as uint4
    hipError_t e = hipSuccess;
    if (c->splat.cells != cells || c->splat.centres != centres) {
        for (void* p : {(void*)c->splat.tex, (void*)c->splat.count})
            if (p) (void)hipFree(p);
        c->splat.tex = c->splat.count = nullptr;
        c->splat.cells = c->splat.centres = 0;
        if ((e = hipMalloc(&c->splat.tex, 4 * cells)) != hipSuccess) return e;
        if ((e = hipMalloc(&c->splat.count, 16 * quads)) != hipSuccess) return e;
        c->splat.cells = cells;
        c->splat.centres = centres;
    }
    if (!c->splat.checked) {
        uint32_t wt[64];
        splat_weights(wt);
        for (int d2 = 0; d2 < 64; d2++)
            if (wt[d2] != (d2 <= kMaxD2 ? kWeights[d2] : 0u)) return hipErrorInvalidValue;
        c->splat.checked = true;
    }
    SplatGeom g;
    g.rx = rx;
    g.ry = ry;
    g.rz = rz;
    g.mnx = c->params.box_min[0];
    g.mny = c->params.box_min[1];
    g.mnz = c->params.box_min[2];
    g.ix = 1.0f / (c->params.box_max[0] - c->params.box_min[0]);
    g.iy = 1.0f / (c->params.box_max[1] - c->params.box_min[1]);
    g.iz = 1.0f / (c->params.box_max[2] - c->params.box_min[2]);
    g.ox = ox;
    g.oy = oy;
    g.oz = oz;
    uint32_t* occ = c->splat.count + centres;
    {
        Launch L(c, "splat_clear");
        const uint32_t blocks = (uint32_t)std::min<size_t>((quads + 255) / 256, 8192);
        hipLaunchKernelGGL(k_splat_clear, dim3(blocks), dim3(256), 0, c->stream, (uint4*)c->splat.count, quads);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (c->n) {
        Launch L(c, "splat_count");
        hipLaunchKernelGGL(k_splat_count, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->pos_a, c->n, g,
                           c->splat.count, occ);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint32_t tx = (uint32_t)(rx + kTX - 1) / kTX, ty = (uint32_t)(ry + kTY - 1) / kTY,
                   tz = (uint32_t)(rz + kTZ - 1) / kTZ;
    {
        Launch L(c, "splat_gather");
        // summed counts fit the 24-bit multiply when fewer than 2^24 particles exist
        if (c->n < (1u << 24))
            hipLaunchKernelGGL(k_splat_gather<false>, dim3(tx * ty * tz), dim3(kThreads), 0, c->stream, c->splat.count,
                               occ, g, tx, ty, c->splat.tex);
        else
            hipLaunchKernelGGL(k_splat_gather<true>, dim3(tx * ty * tz), dim3(kThreads), 0, c->stream, c->splat.count,
                               occ, g, tx, ty, c->splat.tex);
    }
    return hipGetLastError();
}

void splat_free(fe_ctx* c) {
    for (void* p : {(void*)c->splat.tex, (void*)c->splat.count})
        if (p) (void)hipFree(p);
    c->splat = SplatState{};
}

}  // namespace fe
