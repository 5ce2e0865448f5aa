// fe_gather.hip — the neighbour gathers of the DFSPH solver on CDNA4, in the reference's
// own arithmetic.
//
// density_alpha (density_and_alpha.comp:14-72) is the one stage that walks the 27 buckets
// around each particle (z outer, x inner, each bucket in sorted order, a bucket visited
// twice when two neighbour cells share a hash slot). While it does, it records every pair
// that contributes a gradient term (r^2 <= h^2 and r > 1e-6) in that visit order. Every
// later gather of the same neighbour structure — viscosity, both source terms and each
// pressure_force / pressure_update sweep — visits exactly those pairs in exactly that
// order: a pair the reference skips (r^2 > h^2, or r <= 1e-6 where only the density uses
// it) contributes nothing there either. With IEEE division and square root and the GLSL
// operation order, each particle's sums are the reference's sums, bit for bit.
//
// Memory layout for the gathers. The reference's arrays stay in hash-sorted order (they
// are the observable state), but hash order scatters neighbouring cells across memory,
// so a 128-byte line of a scalar field holds particles of four unrelated cells. The work
// is therefore scheduled in Morton order of the particles' cells (work-item t handles
// particle sched[t]) and the data a gather reads from its neighbours is kept in
// schedule-ordered mirrors (pos_m, vb_m, va_m, dens_m, prho_m = p / rho^2, pacc_m) that the
// producing stage writes next to its hash-ordered output. The list holds schedule
// positions, slot-major (nbr[k * N + t]), so a wave reads slot k of 64 work-items as 256
// contiguous bytes and each lane keeps four independent neighbour loads in flight.
// Blocks are remapped so that each XCD (one L2 each) walks a contiguous range of the
// Morton order. A particle with more than kMaxNbr pairs, or a stale list, falls back to
// the bucket walk, which produces the same pairs in the same order.
#include "fe_device.h"

namespace fe {
namespace {

// p / rho^2 of a particle whose density is <= 1e-6: pressure_force skips such neighbours.
// A signalling-NaN payload that no arithmetic result carries.
constexpr uint32_t kNoRho = 0x7FA00001u;

__device__ __forceinline__ float prho_of(float p, float rho) {
    return rho > 1e-6f ? p / (rho * rho) : __uint_as_float(kNoRho);
}

// Workgroups are dealt round-robin over the 8 XCDs; give every XCD a contiguous range of
// the Morton-ordered work.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t g = gridDim.x, b = blockIdx.x;
    const uint32_t per = g / 8u, rem = g % 8u, x = b % 8u, k = b / 8u;
    return x * per + (x < rem ? x : rem) + k;
}

__device__ __forceinline__ uint32_t work_item() { return xcd_block() * kBlock + threadIdx.x; }

// common.glsl:38-54 kernel_w with IEEE q = r / h.
__device__ __forceinline__ float exact_w(float r, const Sph& s) {
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

// common.glsl:56-73 kernel_grad; the scalar that multiplies r_vec.
__device__ __forceinline__ float exact_grad(float r, const Sph& s) {
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
            for (int x = 0; x < 3; x++) rr[y * 3 + x] = range[(hx[x] ^ hy[y] ^ hz[z]) & s.mask];
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

// After the schedule sort: inverse schedule and the position / velocity_b mirrors.
__global__ __launch_bounds__(kBlock) void k_schedule_mirrors(const uint32_t* __restrict__ sched, uint32_t n,
                                                             const float4* __restrict__ pos_b,
                                                             const float4* __restrict__ vel_b,
                                                             uint32_t* __restrict__ inv, float4* __restrict__ pos_m,
                                                             float4* __restrict__ vb_m) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = sched[t];
    inv[i] = t;
    pos_m[t] = pos_b[i];
    vb_m[t] = vel_b[i];
}

// Every mirror from the hash-ordered fields (after uploads or out-of-sequence calls).
__global__ __launch_bounds__(kBlock) void k_refresh_mirrors(const uint32_t* __restrict__ sched, uint32_t n,
                                                            const float4* __restrict__ pos_b,
                                                            const float4* __restrict__ vel_b,
                                                            const float4* __restrict__ vel_a,
                                                            const float* __restrict__ dens,
                                                            const float* __restrict__ pres,
                                                            const float4* __restrict__ pacc, uint32_t* __restrict__ inv,
                                                            float4* __restrict__ pos_m, float4* __restrict__ vb_m,
                                                            float4* __restrict__ va_m, float* __restrict__ dens_m,
                                                            float* __restrict__ prho_m, float4* __restrict__ pacc_m) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = sched[t];
    inv[i] = t;
    pos_m[t] = pos_b[i];
    vb_m[t] = vel_b[i];
    va_m[t] = vel_a[i];
    const float rho = dens[i];
    dens_m[t] = rho;
    prho_m[t] = prho_of(pres[i], rho);
    pacc_m[t] = pacc[i];
}

// ---- density_alpha: bucket walk + list build ---------------------------------------------

__global__ __launch_bounds__(kBlock) void k_density_alpha(Sph s, const uint2* __restrict__ range,
                                                          const float4* __restrict__ pos,
                                                          const uint32_t* __restrict__ sched,
                                                          const uint32_t* __restrict__ inv, float* __restrict__ dens,
                                                          float* __restrict__ fac, float* __restrict__ dens_m,
                                                          uint32_t* __restrict__ nbr, uint32_t* __restrict__ cnt) {
    const uint32_t t = work_item();
    if (t >= s.n) return;
    const uint32_t i = sched[t];
    const float4 pi = pos[i];
    float density = 0.0f, sum_grad_sq = 0.0f;
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    uint32_t m = 0;
    visit27(range, pos, s, pi.x, pi.y, pi.z, [&](uint32_t j, const float4& pj) {
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        const float r = sqrtf(r2);
        const float w = exact_w(r, s);
        density += s.mass * w;
        if (r > 1e-6f) {
            if (m < kMaxNbr) nbr[(size_t)m * s.n + t] = inv[j];
            m++;
            const float g = exact_grad(r, s);
            const float mx = s.mass * (g * rx), my = s.mass * (g * ry), mz = s.mass * (g * rz);
            gx += mx;
            gy += my;
            gz += mz;
            sum_grad_sq += mx * mx + my * my + mz * mz;
        }
    });
    cnt[t] = m;
    dens[i] = density;
    dens_m[t] = density;
    const float alpha = sum_grad_sq + (gx * gx + gy * gy + gz * gz);
    fac[i] = alpha > 1e-6f ? alpha : 0.0f;
}

// ---- list consumers: one policy per reference shader ------------------------------------
// pair() repeats the reference's own tests (r^2 > h^2 rejects, gradient terms need
// r > 1e-6), so the same code serves list entries (which pass them) and the bucket walk.
// load(jt) reads neighbour data by schedule position jt.

// viscosity.comp:17-66 (velocity_b in, velocity_a out with w = 0, gravity fused)
struct Viscosity {
    const float4* vb_m;
    const float* dens_m;
    float4* vel_a;
    float4* va_m;
    float nu, gdx, gdy, gdz, dt;
    struct Self { float4 v; };
    struct Cand { float4 v; float rho; };
    struct Acc { float x, y, z; };
    __device__ Self self(uint32_t, uint32_t t) const { return {vb_m[t]}; }
    __device__ Cand load(uint32_t jt) const { return {vb_m[jt], dens_m[jt]}; }
    __device__ Acc zero() const { return {0.f, 0.f, 0.f}; }
    __device__ void pair(Acc& a, const Self& si, const Sph& s, float rx, float ry, float rz, const Cand& c) const {
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        const float r = sqrtf(r2);
        const float w = exact_w(r, s);
        if (r > 1e-6f) {
            const float m_rho = s.mass / c.rho;
            a.x += (m_rho * (c.v.x - si.v.x)) * w;
            a.y += (m_rho * (c.v.y - si.v.y)) * w;
            a.z += (m_rho * (c.v.z - si.v.z)) * w;
        }
    }
    __device__ unsigned long long finish(uint32_t i, uint32_t t, const Acc& a, const Self& si) const {
        const float vx = si.v.x + nu * a.x, vy = si.v.y + nu * a.y, vz = si.v.z + nu * a.z;
        const float4 v = make_float4(vx + gdx * dt, vy + gdy * dt, vz + gdz * dt, 0.0f);
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
    float* prho_m;
    float dt, rho0;
    struct Self { float4 v; };
    struct Cand { float4 v; };
    struct Acc { float div; };
    __device__ Self self(uint32_t, uint32_t t) const { return {va_m[t]}; }
    __device__ Cand load(uint32_t jt) const { return {va_m[jt]}; }
    __device__ Acc zero() const { return {0.f}; }
    __device__ void pair(Acc& a, const Self& si, const Sph& s, float rx, float ry, float rz, const Cand& c) const {
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        const float r = sqrtf(r2);
        if (r > 1e-6f) {
            const float g = exact_grad(r, s);
            const float dx = si.v.x - c.v.x, dy = si.v.y - c.v.y, dz = si.v.z - c.v.z;
            a.div += s.mass * (dx * (g * rx) + dy * (g * ry) + dz * (g * rz));
        }
    }
    __device__ unsigned long long finish(uint32_t i, uint32_t t, const Acc& a, const Self&) const {
        const float rho = dens[i];
        float source = 0.0f;
        if (dt > 1e-6f) source = ((rho0 - rho) / dt) - a.div;
        src[i] = source;
        pres[i] = 0.0f;
        prho_m[t] = prho_of(0.0f, rho);
        return 0;
    }
};

// divergence_source_term.comp:14-62. velocity_a still holds the previous sort order and is
// read at the current order's indices (as the reference does), so it is gathered through
// the hash index sched[jt]. Also seeds p / rho^2 for the divergence solve from the
// pressures the density solve left (never reset, never permuted).
struct DivergenceSource {
    const float4* vel_a;
    const uint32_t* sched;
    const float* dens;
    const float* pres;
    float* src;
    float* prho_m;
    struct Self { float4 v; };
    struct Cand { float4 v; };
    struct Acc { float div; };
    __device__ Self self(uint32_t i, uint32_t) const { return {vel_a[i]}; }
    __device__ Cand load(uint32_t jt) const { return {vel_a[sched[jt]]}; }
    __device__ Acc zero() const { return {0.f}; }
    __device__ void pair(Acc& a, const Self& si, const Sph& s, float rx, float ry, float rz, const Cand& c) const {
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        const float r = sqrtf(r2);
        if (r > 1e-6f) {
            const float g = exact_grad(r, s);
            const float dx = si.v.x - c.v.x, dy = si.v.y - c.v.y, dz = si.v.z - c.v.z;
            a.div += s.mass * (dx * (g * rx) + dy * (g * ry) + dz * (g * rz));
        }
    }
    __device__ unsigned long long finish(uint32_t i, uint32_t t, const Acc& a, const Self&) const {
        src[i] = -a.div;
        prho_m[t] = prho_of(pres[i], dens[i]);
        return 0;
    }
};

// pressure_force.comp:15-78. Neighbours supply p_j / rho_j^2 (kNoRho when rho_j <= 1e-6,
// which the reference skips).
struct PressureForce {
    const float* pres;
    const float* dens;
    const float* prho_m;
    float4* pacc;
    float4* pacc_m;
    struct Self { float prho; };
    struct Cand { float prho; };
    struct Acc { float x, y, z; };
    __device__ Self self(uint32_t i, uint32_t) const {
        const float rho = dens[i];
        float pr = 0.0f;
        if (rho > 1e-6f) pr = pres[i] / (rho * rho);
        return {pr};
    }
    __device__ Cand load(uint32_t jt) const { return {prho_m[jt]}; }
    __device__ Acc zero() const { return {0.f, 0.f, 0.f}; }
    __device__ void pair(Acc& a, const Self& si, const Sph& s, float rx, float ry, float rz, const Cand& c) const {
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        const float r = sqrtf(r2);
        if (r > 1e-6f && __float_as_uint(c.prho) != kNoRho) {
            const float g = exact_grad(r, s);
            const float mt = s.mass * (si.prho + c.prho);
            a.x += mt * (g * rx);
            a.y += mt * (g * ry);
            a.z += mt * (g * rz);
        }
    }
    __device__ unsigned long long finish(uint32_t i, uint32_t t, const Acc& a, const Self&) const {
        const float4 v = make_float4(-a.x, -a.y, -a.z, 0.0f);
        pacc[i] = v;
        pacc_m[t] = v;
        return 0;
    }
};

// pressure_update.comp:18-84 (in place; neighbours' pressures are not read), plus the
// sweep's quantised compression residual.
struct PressureUpdate {
    const float4* pacc_m;
    const float* fac;
    const float* src;
    const float* dens;
    float* pres;
    float* prho_m;
    float dt, omega, rho0;
    struct Self { float4 a; };
    struct Cand { float4 a; };
    struct Acc { float ap; };
    __device__ Self self(uint32_t, uint32_t t) const { return {pacc_m[t]}; }
    __device__ Cand load(uint32_t jt) const { return {pacc_m[jt]}; }
    __device__ Acc zero() const { return {0.f}; }
    __device__ void pair(Acc& a, const Self& si, const Sph& s, float rx, float ry, float rz, const Cand& c) const {
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        const float r = sqrtf(r2);
        if (r > 1e-6f) {
            const float g = exact_grad(r, s);
            const float dx = si.a.x - c.a.x, dy = si.a.y - c.a.y, dz = si.a.z - c.a.z;
            a.ap += s.mass * (dx * (g * rx) + dy * (g * ry) + dz * (g * rz));
        }
    }
    __device__ unsigned long long finish(uint32_t i, uint32_t t, const Acc& a, const Self&) const {
        const float alpha_i = fac[i], rho_i = dens[i], src_i = src[i];
        float p = pres[i];
        const float ap = dt * a.ap;
        const float err = src_i - ap;
        const unsigned long long q = residual_q(err, dt, rho0);
        if (alpha_i > 1e-6f && rho_i > 1e-6f) {
            const float a_ii = -(dt / (rho_i * rho_i)) * alpha_i;
            if (fabsf(a_ii) > 1e-20f) {
                const float corr = (omega * err) / a_ii;
                const float np = p + corr;
                p = np > 0.0f ? np : 0.0f;
                pres[i] = p;
            }
        }
        prho_m[t] = prho_of(p, rho_i);
        return q;
    }
};

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

// One work-item per particle: the listed pairs, four neighbour loads in flight, or the
// bucket walk when the list is stale or the particle overflowed it. With `prev` set, a
// converged previous sweep skips the launch; with `partial` set, each block stores the sum
// of its particles' finish() values.
template <class P>
__global__ __launch_bounds__(kBlock) void k_list_gather(P pol, Sph s, const uint2* __restrict__ range,
                                                        const float4* __restrict__ pos_b,
                                                        const float4* __restrict__ pos_m,
                                                        const uint32_t* __restrict__ sched,
                                                        const uint32_t* __restrict__ inv,
                                                        const uint32_t* __restrict__ nbr,
                                                        const uint32_t* __restrict__ cnt, uint32_t list_ok,
                                                        unsigned long long* __restrict__ partial,
                                                        const unsigned long long* __restrict__ prev,
                                                        unsigned long long thr) {
    if (prev && *prev <= thr) return;
    const uint32_t t = work_item();
    unsigned long long q = 0;
    if (t < s.n) {
        const uint32_t i = sched[t];
        const float4 pi = pos_m[t];
        const typename P::Self si = pol.self(i, t);
        typename P::Acc acc = pol.zero();
        const uint32_t m = list_ok ? cnt[t] : 0xFFFFFFFFu;
        if (m <= kMaxNbr) {
            for (uint32_t k = 0; k < m; k += 4) {
                uint32_t jt[4];
#pragma unroll
                for (int u = 0; u < 4; u++) jt[u] = nbr[(size_t)((k + u < m) ? k + u : k) * s.n + t];
                float4 pj[4];
                typename P::Cand cj[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    pj[u] = pos_m[jt[u]];
                    cj[u] = pol.load(jt[u]);
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (k + u < m) pol.pair(acc, si, s, pi.x - pj[u].x, pi.y - pj[u].y, pi.z - pj[u].z, cj[u]);
            }
        } else {
            visit27(range, pos_b, s, pi.x, pi.y, pi.z, [&](uint32_t j, const float4& pj) {
                pol.pair(acc, si, s, pi.x - pj.x, pi.y - pj.y, pi.z - pj.z, pol.load(inv[j]));
            });
        }
        q = pol.finish(i, t, acc, si);
    }
    if (partial) {
        const unsigned long long b = block_sum_u64(q);
        if (threadIdx.x == 0) partial[blockIdx.x] = b;
    }
}

// Sum of per-block residual partials into one sweep slot (single block, deterministic).
__global__ __launch_bounds__(kBlock) void k_resid_finalize(const unsigned long long* __restrict__ partial,
                                                           uint32_t count, unsigned long long* __restrict__ out,
                                                           const unsigned long long* __restrict__ prev,
                                                           unsigned long long thr) {
    if (prev && *prev <= thr) return;
    __shared__ unsigned long long part[kBlock / 64];
    unsigned long long v = 0;
    for (uint32_t k = threadIdx.x; k < count; k += kBlock) v += partial[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int k = 0; k < (int)(kBlock / 64); k++) tot += part[k];
        *out = tot;
    }
}

inline dim3 grid_for(uint32_t count) { return dim3((count + kBlock - 1) / kBlock); }

// Mirror bits each consumer needs; a missing one is rebuilt from the hash-ordered fields.
hipError_t need(fe_ctx* c, uint32_t bits) {
    if ((c->mirrors & bits) == bits) return hipSuccess;
    Launch L(c, "refresh_mirrors");
    hipLaunchKernelGGL(k_refresh_mirrors, grid_for(c->n), dim3(kBlock), 0, c->stream, c->sched[0], c->n, c->pos_b,
                       c->vel_b, c->vel_a, c->dens, c->pres, c->pacc, c->inv, c->pos_m, c->vb_m, c->va_m, c->dens_m,
                       c->prho_m, c->pacc_m);
    c->mirrors = kMirrorAll;
    return hipGetLastError();
}

template <class P>
hipError_t run_list(fe_ctx* c, const char* name, const P& pol, unsigned long long* partial = nullptr,
                    const unsigned long long* prev = nullptr, unsigned long long thr = 0) {
    Launch L(c, name);
    hipLaunchKernelGGL(k_list_gather<P>, grid_for(c->n), dim3(kBlock), 0, c->stream, pol, sph_of(c), c->range,
                       c->pos_b, c->pos_m, c->sched[0], c->inv, c->nbr, c->nbr_count, c->list_valid ? 1u : 0u,
                       partial, prev, thr);
    return hipGetLastError();
}

}  // namespace

uint32_t gather_partials(uint32_t n) { return (n + kBlock - 1) / kBlock; }

hipError_t launch_schedule_mirrors(fe_ctx* c) {
    Launch L(c, "schedule_mirrors");
    hipLaunchKernelGGL(k_schedule_mirrors, grid_for(c->n), dim3(kBlock), 0, c->stream, c->sched[0], c->n, c->pos_b,
                       c->vel_b, c->inv, c->pos_m, c->vb_m);
    c->mirrors = kMirrorPos | kMirrorVelB;
    c->list_valid = false;
    return hipGetLastError();
}

hipError_t launch_density_alpha(fe_ctx* c) {
    hipError_t e = need(c, kMirrorPos);   // the inverse schedule
    if (e != hipSuccess) return e;
    Launch L(c, "density_alpha");
    hipLaunchKernelGGL(k_density_alpha, grid_for(c->n), dim3(kBlock), 0, c->stream, sph_of(c), c->range, c->pos_b,
                       c->sched[0], c->inv, c->dens, c->fac, c->dens_m, c->nbr, c->nbr_count);
    c->list_valid = true;
    c->mirrors = (c->mirrors | kMirrorDens) & ~kMirrorPrho;   // p / rho^2 used the old densities
    return hipGetLastError();
}

hipError_t launch_viscosity(fe_ctx* c) {
    hipError_t e = need(c, kMirrorPos | kMirrorVelB | kMirrorDens);
    if (e != hipSuccess) return e;
    const fe_params& p = c->params;
    Viscosity pol{c->vb_m, c->dens_m, c->vel_a, c->va_m, p.viscosity, p.gravity[0], p.gravity[1], p.gravity[2], p.dt};
    e = run_list(c, "viscosity", pol);
    c->mirrors |= kMirrorVelA;
    return e;
}

hipError_t launch_density_source(fe_ctx* c) {
    hipError_t e = need(c, kMirrorPos | kMirrorVelA);
    if (e != hipSuccess) return e;
    DensitySource pol{c->va_m, c->dens, c->src, c->pres, c->prho_m, c->params.dt, c->params.target_density};
    e = run_list(c, "density_source", pol);
    c->mirrors |= kMirrorPrho;
    return e;
}

hipError_t launch_divergence_source(fe_ctx* c) {
    hipError_t e = need(c, kMirrorPos);
    if (e != hipSuccess) return e;
    DivergenceSource pol{c->vel_a, c->sched[0], c->dens, c->pres, c->src, c->prho_m};
    e = run_list(c, "divergence_source", pol);
    c->mirrors |= kMirrorPrho;
    return e;
}

hipError_t launch_pressure_force(fe_ctx* c, const unsigned long long* prev, unsigned long long thr) {
    hipError_t e = need(c, kMirrorPos | kMirrorPrho);
    if (e != hipSuccess) return e;
    PressureForce pol{c->pres, c->dens, c->prho_m, c->pacc, c->pacc_m};
    e = run_list(c, "pressure_force", pol, nullptr, prev, thr);
    c->mirrors |= kMirrorPacc;
    return e;
}

hipError_t launch_pressure_update(fe_ctx* c, unsigned long long* out, const unsigned long long* prev,
                                  unsigned long long thr) {
    hipError_t e = need(c, kMirrorPos | kMirrorPacc | kMirrorPrho);
    if (e != hipSuccess) return e;
    PressureUpdate pol{c->pacc_m, c->fac, c->src, c->dens, c->pres, c->prho_m, c->params.dt, c->params.relax_factor,
                       c->params.target_density};
    e = run_list(c, "pressure_update", pol, out ? c->partial : nullptr, prev, thr);
    if (e != hipSuccess || !out) return e;
    Launch L(c, "resid_finalize");
    hipLaunchKernelGGL(k_resid_finalize, dim3(1), dim3(kBlock), 0, c->stream, c->partial, gather_partials(c->n), out,
                       prev, thr);
    return hipGetLastError();
}

}  // namespace fe
