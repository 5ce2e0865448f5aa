// fe_sph.hip — neighbour-search plumbing and the DFSPH solver kernels on CDNA4.
//
// One work-item per particle (sorted order) for every stage, as the reference dispatches
// its compute shaders. Each per-pair expression is evaluated in the GLSL order of the
// reference shader it restates, with FP contraction off and IEEE division / square root,
// so a particle's sum is the same sequence of FP32 operations as the reference's.
//
// Memory layout (HBM, SOA, 16-byte vec4 rows as in GpuPhysicsData): positions, velocities,
// pressure accelerations and colours are float4[N]; densities, alpha factors, source terms
// and pressures are float[N]; the sorted entries are two uint32[T] arrays (hash, index) and
// each hash bucket's [start, end) is one uint2 in `range` (start = grid_start).
#include "fe_internal.h"

namespace fe {
namespace {

constexpr float kPi = 3.1415926535897932384626433832795f;
constexpr uint32_t P1 = 73856093u, P2 = 19349663u, P3 = 83492791u;

struct Sph {
    float h, hh, mass, k;   // k = 8 / (pi h^3) of common.glsl:42-43 / :61-62
    float inv_h;            // 1 / h: the cell coordinate is floor(pos * (1 / h))
    uint32_t n, mask;       // particles, T - 1 (T is a power of two, so % T == & (T - 1))
};

// common.glsl:38-54 kernel_w.
__device__ __forceinline__ float kernel_w(float r, const Sph& s) {
    float q = r / s.h;
    if (q >= 1.0f) return 0.0f;
    if (q <= 0.5f) {
        float q2 = q * q;
        float q3 = q2 * q;
        return s.k * (6.0f * q3 - 6.0f * q2 + 1.0f);
    }
    float omq = 1.0f - q;
    float omq3 = omq * omq * omq;
    return s.k * 2.0f * omq3;
}

// common.glsl:56-73 kernel_grad; returns the scalar that multiplies r_vec.
__device__ __forceinline__ float kernel_grad_scale(float r, const Sph& s) {
    float q = r / s.h;
    if (q >= 1.0f || r < 1e-6f) return 0.0f;
    float gf;
    if (q <= 0.5f) {
        gf = s.k * (18.0f * q * q - 12.0f * q);
    } else {
        float omq = 1.0f - q;
        gf = s.k * -6.0f * omq * omq;
    }
    return gf / (s.h * r);
}

__device__ __forceinline__ uint32_t cell_hash(int cx, int cy, int cz, uint32_t mask) {
    return ((uint32_t)cx * P1 ^ (uint32_t)cy * P2 ^ (uint32_t)cz * P3) & mask;
}

// The 27-cell bucket walk of density_and_alpha.comp:31-45 (z outer, x inner). f(j) is
// called for every sorted entry j of every non-empty neighbour bucket, in reference order,
// visiting a bucket twice when two neighbour cells share a hash slot.
template <class F>
__device__ __forceinline__ void visit27(const uint2* __restrict__ range, const Sph& s, float px, float py,
                                        float pz, F&& f) {
    const int cx = (int)floorf(px * s.inv_h), cy = (int)floorf(py * s.inv_h), cz = (int)floorf(pz * s.inv_h);
    const uint32_t hx[3] = {(uint32_t)(cx - 1) * P1, (uint32_t)cx * P1, (uint32_t)(cx + 1) * P1};
    const uint32_t hy[3] = {(uint32_t)(cy - 1) * P2, (uint32_t)cy * P2, (uint32_t)(cy + 1) * P2};
    const uint32_t hz[3] = {(uint32_t)(cz - 1) * P3, (uint32_t)cz * P3, (uint32_t)(cz + 1) * P3};
#pragma unroll
    for (int z = 0; z < 3; z++) {
#pragma unroll
        for (int y = 0; y < 3; y++) {
            const uint32_t hyz = hy[y] ^ hz[z];
#pragma unroll
            for (int x = 0; x < 3; x++) {
                const uint2 r = range[(hx[x] ^ hyz) & s.mask];
                if (r.x == kSentinel) continue;
                for (uint32_t j = r.x; j < r.y; j++) f(j);
            }
        }
    }
}

__device__ __forceinline__ uint32_t gid() { return blockIdx.x * kBlock + threadIdx.x; }

// ---- neighbour search ---------------------------------------------------------------

// spatial_hash.comp:23-36, plus the grid_start fill of neighbor_search.rs:207-208 (the
// table is not read between the hash and the offsets). The cell coordinate
// ivec3(floor(pos / h)) is evaluated as floor(pos * (1 / h)), the form the reference GPU
// executed (it reproduces the bucket order of the reference's rest-density capture).
__global__ __launch_bounds__(kBlock) void k_hash(const float4* __restrict__ pos, uint32_t n, uint32_t T,
                                                 float inv_h, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                 uint32_t* __restrict__ keys_alt, uint32_t* __restrict__ vals_alt,
                                                 uint2* __restrict__ range) {
    uint32_t i = gid();
    if (i >= T) return;
    range[i] = make_uint2(kSentinel, 0u);
    if (i < n) {
        float4 p = pos[i];
        keys[i] = cell_hash((int)floorf(p.x * inv_h), (int)floorf(p.y * inv_h), (int)floorf(p.z * inv_h), T - 1u);
        vals[i] = i;
    } else {
        keys[i] = kSentinel;
        vals[i] = kSentinel;
        keys_alt[i] = kSentinel;
        vals_alt[i] = kSentinel;
    }
}

// grid_offsets.comp:20-32 (bucket starts) fused with reorder.comp:37-45 (pos/vel gather)
// and the bucket ends the gathers use. Entries at i >= n are sentinels.
__global__ __launch_bounds__(kBlock) void k_offsets_reorder(const uint32_t* __restrict__ keys,
                                                            const uint32_t* __restrict__ vals, uint32_t n,
                                                            uint2* __restrict__ range,
                                                            const float4* __restrict__ pos_a,
                                                            const float4* __restrict__ vel_a,
                                                            float4* __restrict__ pos_b, float4* __restrict__ vel_b) {
    uint32_t i = gid();
    if (i >= n) return;
    uint32_t key = keys[i];
    if (key != kSentinel) {
        if (i == 0 || keys[i - 1] != key) range[key].x = i;
        if (i + 1 == n || keys[i + 1] != key) range[key].y = i + 1;
    }
    uint32_t src = vals[i];
    pos_b[i] = pos_a[src];
    vel_b[i] = vel_a[src];
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

// ---- solver gathers -------------------------------------------------------------------

// density_and_alpha.comp:14-72.
__global__ __launch_bounds__(kBlock) void k_density_alpha(Sph s, const uint2* __restrict__ range,
                                                          const float4* __restrict__ pos,
                                                          float* __restrict__ dens, float* __restrict__ fac) {
    uint32_t i = gid();
    if (i >= s.n) return;
    const float4 pi = pos[i];
    float density = 0.0f, sum_grad_sq = 0.0f;
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    visit27(range, s, pi.x, pi.y, pi.z, [&](uint32_t j) {
        const float4 pj = pos[j];
        float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        float r = sqrtf(r2);
        float w = kernel_w(r, s);
        density += s.mass * w;
        if (r > 1e-6f) {
            float g = kernel_grad_scale(r, s);
            float mx = s.mass * (g * rx), my = s.mass * (g * ry), mz = s.mass * (g * rz);
            gx += mx; gy += my; gz += mz;
            sum_grad_sq += mx * mx + my * my + mz * mz;
        }
    });
    dens[i] = density;
    float alpha = sum_grad_sq + (gx * gx + gy * gy + gz * gz);
    fac[i] = alpha > 1e-6f ? alpha : 0.0f;
}

// viscosity.comp:17-66: XSPH over velocity_b, gravity fused, written to velocity_a (w = 0).
__global__ __launch_bounds__(kBlock) void k_viscosity(Sph s, const uint2* __restrict__ range,
                                                      const float4* __restrict__ pos,
                                                      const float4* __restrict__ vel, const float* __restrict__ dens,
                                                      float4* __restrict__ out, float nu, float gdx, float gdy,
                                                      float gdz, float dt) {
    uint32_t i = gid();
    if (i >= s.n) return;
    const float4 pi = pos[i];
    const float4 vi = vel[i];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    visit27(range, s, pi.x, pi.y, pi.z, [&](uint32_t j) {
        const float4 pj = pos[j];
        float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        float r = sqrtf(r2);
        float w = kernel_w(r, s);
        if (r > 1e-6f) {
            const float4 vj = vel[j];
            float m_rho = s.mass / dens[j];
            sx += (m_rho * (vj.x - vi.x)) * w;
            sy += (m_rho * (vj.y - vi.y)) * w;
            sz += (m_rho * (vj.z - vi.z)) * w;
        }
    });
    float vx = vi.x + nu * sx, vy = vi.y + nu * sy, vz = vi.z + nu * sz;
    out[i] = make_float4(vx + gdx * dt, vy + gdy * dt, vz + gdz * dt, 0.0f);
}

// density_source_term.comp:16-74 (velocity_a; pressures zeroed).
__global__ __launch_bounds__(kBlock) void k_density_source(Sph s, const uint2* __restrict__ range,
                                                           const float4* __restrict__ pos,
                                                           const float4* __restrict__ vel,
                                                           const float* __restrict__ dens, float* __restrict__ src,
                                                           float* __restrict__ pres, float dt, float rho0) {
    uint32_t i = gid();
    if (i >= s.n) return;
    const float4 pi = pos[i];
    const float4 vi = vel[i];
    float div = 0.0f;
    visit27(range, s, pi.x, pi.y, pi.z, [&](uint32_t j) {
        const float4 pj = pos[j];
        float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        float r = sqrtf(r2);
        if (r > 1e-6f) {
            float g = kernel_grad_scale(r, s);
            const float4 vj = vel[j];
            float dx = vi.x - vj.x, dy = vi.y - vj.y, dz = vi.z - vj.z;
            div += s.mass * (dx * (g * rx) + dy * (g * ry) + dz * (g * rz));
        }
    });
    float source = 0.0f;
    if (dt > 1e-6f) source = ((rho0 - dens[i]) / dt) - div;
    src[i] = source;
    pres[i] = 0.0f;
}

// divergence_source_term.comp:14-62 (velocity_a indexed by the current order).
__global__ __launch_bounds__(kBlock) void k_divergence_source(Sph s, const uint2* __restrict__ range,
                                                              const float4* __restrict__ pos,
                                                              const float4* __restrict__ vel,
                                                              float* __restrict__ src) {
    uint32_t i = gid();
    if (i >= s.n) return;
    const float4 pi = pos[i];
    const float4 vi = vel[i];
    float div = 0.0f;
    visit27(range, s, pi.x, pi.y, pi.z, [&](uint32_t j) {
        const float4 pj = pos[j];
        float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        float r = sqrtf(r2);
        if (r > 1e-6f) {
            float g = kernel_grad_scale(r, s);
            const float4 vj = vel[j];
            float dx = vi.x - vj.x, dy = vi.y - vj.y, dz = vi.z - vj.z;
            div += s.mass * (dx * (g * rx) + dy * (g * ry) + dz * (g * rz));
        }
    });
    src[i] = -div;
}

// pressure_force.comp:15-78. With `prev` set, a converged previous sweep skips the launch.
__global__ __launch_bounds__(kBlock) void k_pressure_force(Sph s, const uint2* __restrict__ range,
                                                           const float4* __restrict__ pos,
                                                           const float* __restrict__ pres,
                                                           const float* __restrict__ dens,
                                                           float4* __restrict__ pacc,
                                                           const unsigned long long* __restrict__ prev,
                                                           unsigned long long thr) {
    if (prev && *prev <= thr) return;
    uint32_t i = gid();
    if (i >= s.n) return;
    const float4 pi = pos[i];
    const float rho_i = dens[i], p_i = pres[i];
    float p_rho_i = 0.0f;
    if (rho_i > 1e-6f) p_rho_i = p_i / (rho_i * rho_i);
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    visit27(range, s, pi.x, pi.y, pi.z, [&](uint32_t j) {
        const float4 pj = pos[j];
        float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        float r = sqrtf(r2);
        if (r > 1e-6f) {
            float rho_j = dens[j];
            float p_j = pres[j];
            if (rho_j > 1e-6f) {
                float g = kernel_grad_scale(r, s);
                float p_rho_j = p_j / (rho_j * rho_j);
                float mt = s.mass * (p_rho_i + p_rho_j);
                ax += mt * (g * rx);
                ay += mt * (g * ry);
                az += mt * (g * rz);
            }
        }
    });
    pacc[i] = make_float4(-ax, -ay, -az, 0.0f);
}

// Quantised compression residual of one particle (FE_SOLVER_CONVERGE): the decision sums
// integers, so it does not depend on the order of the atomics.
__device__ __forceinline__ unsigned long long residual_q(float s_minus_ap, float dt, float rho0) {
    float neg = -s_minus_ap;
    if (!(neg > 0.0f)) return 0ull;
    float e = (neg * dt) / rho0;
    if (!(e < 256.0f)) e = 256.0f;
    return (unsigned long long)(e * 4294967296.0f);
}

__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v) {
    __shared__ unsigned long long part[kBlock / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(kBlock / 64); k++) t += part[k];
    return t;
}

// pressure_update.comp:18-84 (in place: neighbours' pressures are not read), plus the
// residual of the sweep into *out when `out` is set.
__global__ __launch_bounds__(kBlock) void k_pressure_update(
    Sph s, const uint2* __restrict__ range, const float4* __restrict__ pos, const float4* __restrict__ pacc,
    const float* __restrict__ fac, const float* __restrict__ src, const float* __restrict__ dens,
    float* __restrict__ pres, float dt, float omega, float rho0, unsigned long long* __restrict__ out,
    const unsigned long long* __restrict__ prev, unsigned long long thr) {
    if (prev && *prev <= thr) return;
    uint32_t i = gid();
    unsigned long long q = 0;
    if (i < s.n) {
        const float4 pi = pos[i];
        const float4 ai = pacc[i];
        const float alpha_i = fac[i], rho_i = dens[i], src_i = src[i], p_i = pres[i];
        float sum_ap = 0.0f;
        visit27(range, s, pi.x, pi.y, pi.z, [&](uint32_t j) {
            const float4 pj = pos[j];
            float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
            float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 > s.hh) return;
            float r = sqrtf(r2);
            if (r > 1e-6f) {
                float g = kernel_grad_scale(r, s);
                const float4 aj = pacc[j];
                float dx = ai.x - aj.x, dy = ai.y - aj.y, dz = ai.z - aj.z;
                sum_ap += s.mass * (dx * (g * rx) + dy * (g * ry) + dz * (g * rz));
            }
        });
        float ap = dt * sum_ap;
        float err = src_i - ap;
        q = residual_q(err, dt, rho0);
        if (alpha_i > 1e-6f && rho_i > 1e-6f) {
            float a_ii = -(dt / (rho_i * rho_i)) * alpha_i;
            if (fabsf(a_ii) > 1e-20f) {
                float corr = (omega * err) / a_ii;
                float np = p_i + corr;
                pres[i] = np > 0.0f ? np : 0.0f;
            }
        }
    }
    if (out) {
        unsigned long long t = block_sum_u64(q);
        if (threadIdx.x == 0 && t) atomicAdd(out, t);
    }
}

// ---- streaming stages ----------------------------------------------------------------

// pressure_integration.comp:16-77 (position_b -> position_a; collision order y, x, z).
__global__ __launch_bounds__(kBlock) void k_pressure_integration(
    uint32_t n, const float4* __restrict__ pos_b, const float4* __restrict__ pacc, float4* __restrict__ vel,
    float4* __restrict__ pos_a, float dt, float rad, float3 mn, float3 mx) {
    uint32_t i = gid();
    if (i >= n) return;
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

__device__ __forceinline__ float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }

// divergence_integration.comp:14-43.
__global__ __launch_bounds__(kBlock) void k_divergence_integration(uint32_t n, float4* __restrict__ vel,
                                                                   const float4* __restrict__ pacc,
                                                                   float4* __restrict__ colors, float dt) {
    uint32_t i = gid();
    if (i >= n) return;
    const float4 v = vel[i], a = pacc[i];
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
    vel[i] = make_float4(vx, vy, vz, 0.0f);
    colors[i] = make_float4(c0, c1, c2, 1.0f);
}

__device__ __forceinline__ uint32_t to_u32_sat(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

// stats_pipeline.rs:56-66 (zeroing done by the launcher) + stats.comp:18-30, reduced per
// block before the atomics (max and wrapping sums are order-independent).
__global__ __launch_bounds__(kBlock) void k_stats(uint32_t n, const float4* __restrict__ vel,
                                                  const float* __restrict__ dens, const float* __restrict__ src,
                                                  float rho0, uint32_t* __restrict__ stats) {
    __shared__ uint32_t part[3][kBlock / 64];
    uint32_t i = gid();
    uint32_t mx = 0, s1 = 0, s2 = 0;
    if (i < n) {
        const float4 v = vel[i];
        float speed = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
        mx = __float_as_uint(speed);
        s1 = min(to_u32_sat(fabsf(dens[i] - rho0) * 1.0f), 10000u);
        s2 = min(to_u32_sat(fabsf(src[i]) * 10.0f), 10000u);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
        s1 += (uint32_t)__shfl_xor((int)s1, off, 64);
        s2 += (uint32_t)__shfl_xor((int)s2, off, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        part[0][w] = mx; part[1][w] = s1; part[2][w] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(kBlock / 64); k++) {
            mx = max(mx, part[0][k]); s1 += part[1][k]; s2 += part[2][k];
        }
        atomicMax(&stats[0], mx);
        atomicAdd(&stats[1], s1);
        atomicAdd(&stats[2], s2);
    }
}

__global__ __launch_bounds__(kBlock) void k_zero_u32(uint32_t* p, uint32_t count) {
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

inline Sph sph_of(const fe_ctx* c) {
    Sph s;
    s.h = c->params.smoothing_radius;
    s.hh = s.h * s.h;
    s.mass = c->params.particle_mass;
    float h3 = s.h * s.h * s.h;
    s.k = 8.0f / (kPi * h3);
    s.inv_h = 1.0f / s.h;
    s.n = c->n;
    s.mask = c->T - 1u;
    return s;
}

}  // namespace

hipError_t launch_hash(fe_ctx* c) {
    Launch L(c, "spatial_hash");
    hipLaunchKernelGGL(k_hash, grid_for(c->T), dim3(kBlock), 0, c->stream, c->pos_a, c->n, c->T,
                       1.0f / c->params.smoothing_radius, c->keys[0], c->vals[0], c->keys[1], c->vals[1], c->range);
    return hipGetLastError();
}

hipError_t launch_offsets_reorder(fe_ctx* c) {
    Launch L(c, "offsets_reorder");
    hipLaunchKernelGGL(k_offsets_reorder, grid_for(c->n), dim3(kBlock), 0, c->stream, c->keys[0], c->vals[0], c->n,
                       c->range, c->pos_a, c->vel_a, c->pos_b, c->vel_b);
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

hipError_t launch_density_alpha(fe_ctx* c) {
    Launch L(c, "density_alpha");
    hipLaunchKernelGGL(k_density_alpha, grid_for(c->n), dim3(kBlock), 0, c->stream, sph_of(c), c->range, c->pos_b,
                       c->dens, c->fac);
    return hipGetLastError();
}

hipError_t launch_viscosity(fe_ctx* c) {
    Launch L(c, "viscosity");
    const fe_params& p = c->params;
    hipLaunchKernelGGL(k_viscosity, grid_for(c->n), dim3(kBlock), 0, c->stream, sph_of(c), c->range, c->pos_b,
                       c->vel_b, c->dens, c->vel_a, p.viscosity, p.gravity[0], p.gravity[1], p.gravity[2], p.dt);
    return hipGetLastError();
}

hipError_t launch_density_source(fe_ctx* c) {
    Launch L(c, "density_source");
    hipLaunchKernelGGL(k_density_source, grid_for(c->n), dim3(kBlock), 0, c->stream, sph_of(c), c->range, c->pos_b,
                       c->vel_a, c->dens, c->src, c->pres, c->params.dt, c->params.target_density);
    return hipGetLastError();
}

hipError_t launch_pressure_force(fe_ctx* c, const unsigned long long* prev, unsigned long long thr) {
    Launch L(c, "pressure_force");
    hipLaunchKernelGGL(k_pressure_force, grid_for(c->n), dim3(kBlock), 0, c->stream, sph_of(c), c->range, c->pos_b,
                       c->pres, c->dens, c->pacc, prev, thr);
    return hipGetLastError();
}

hipError_t launch_pressure_update(fe_ctx* c, unsigned long long* out, const unsigned long long* prev,
                                  unsigned long long thr) {
    Launch L(c, "pressure_update");
    hipLaunchKernelGGL(k_pressure_update, grid_for(c->n), dim3(kBlock), 0, c->stream, sph_of(c), c->range, c->pos_b,
                       c->pacc, c->fac, c->src, c->dens, c->pres, c->params.dt, c->params.relax_factor,
                       c->params.target_density, out, prev, thr);
    return hipGetLastError();
}

hipError_t launch_pressure_integration(fe_ctx* c) {
    Launch L(c, "pressure_integration");
    const fe_params& p = c->params;
    hipLaunchKernelGGL(k_pressure_integration, grid_for(c->n), dim3(kBlock), 0, c->stream, c->n, c->pos_b, c->pacc,
                       c->vel_a, c->pos_a, p.dt, p.particle_radius,
                       make_float3(p.box_min[0], p.box_min[1], p.box_min[2]),
                       make_float3(p.box_max[0], p.box_max[1], p.box_max[2]));
    return hipGetLastError();
}

hipError_t launch_divergence_source(fe_ctx* c) {
    Launch L(c, "divergence_source");
    hipLaunchKernelGGL(k_divergence_source, grid_for(c->n), dim3(kBlock), 0, c->stream, sph_of(c), c->range,
                       c->pos_b, c->vel_a, c->src);
    return hipGetLastError();
}

hipError_t launch_divergence_integration(fe_ctx* c) {
    Launch L(c, "divergence_integration");
    hipLaunchKernelGGL(k_divergence_integration, grid_for(c->n), dim3(kBlock), 0, c->stream, c->n, c->vel_a, c->pacc,
                       c->colors, c->params.dt);
    return hipGetLastError();
}

hipError_t launch_stats(fe_ctx* c) {
    {
        Launch L(c, "stats_zero");
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(kBlock), 0, c->stream, c->stats, 3u);
    }
    Launch L(c, "stats");
    hipLaunchKernelGGL(k_stats, grid_for(c->n), dim3(kBlock), 0, c->stream, c->n, c->vel_a, c->dens, c->src,
                       c->params.target_density, c->stats);
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

hipError_t launch_unpack_start(fe_ctx* c, const uint32_t* in) {
    hipLaunchKernelGGL(k_unpack_start, grid_for(c->T), dim3(kBlock), 0, c->stream, in, c->T, c->range);
    return hipGetLastError();
}

}  // namespace fe
