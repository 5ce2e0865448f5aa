// fe_exact.hip — the solver gathers in the reference's own arithmetic, driven by a
// per-particle neighbour list.
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
// Work-item t handles particle sched[t]: the neighbour search orders particles by the
// Morton code of their cell, so a block's particles and their neighbours are close in
// space and their cache lines are reused from L2 instead of being fetched per particle
// (the arrays themselves stay in the reference's hash order). The list is slot-major by
// work-item, nbr[k * N + t] for slot k < kMaxNbr, so a wave reading slot k reads 256
// contiguous bytes; each lane then has four independent candidate loads in flight instead
// of one dependent bucket walk. A particle with more than kMaxNbr pairs, or a list made
// stale by an upload, falls back to the walk.
#include "fe_device.h"

namespace fe {
namespace {


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

// The 27-bucket walk in reference order. Per z slab the nine bucket ranges are loaded
// together and the slab's candidates are treated as one sequence fetched eight at a time
// (few dependent round trips per lane); f(j, pos_j) still sees them in visit order.
template <class F>
__device__ __forceinline__ void visit27(const uint2* __restrict__ range, const float4* __restrict__ pos,
                                        const Sph& s, float px, float py, float pz, F&& f) {
    const int cx = (int)floorf(px * s.inv_h), cy = (int)floorf(py * s.inv_h), cz = (int)floorf(pz * s.inv_h);
    const uint32_t hx[3] = {(uint32_t)(cx - 1) * kP1, (uint32_t)cx * kP1, (uint32_t)(cx + 1) * kP1};
    const uint32_t hy[3] = {(uint32_t)(cy - 1) * kP2, (uint32_t)cy * kP2, (uint32_t)(cy + 1) * kP2};
    const uint32_t hz[3] = {(uint32_t)(cz - 1) * kP3, (uint32_t)cz * kP3, (uint32_t)(cz + 1) * kP3};
#pragma unroll
    for (int z = 0; z < 3; z++) {
        uint32_t st[9], pre[10];
        pre[0] = 0;
#pragma unroll
        for (int y = 0; y < 3; y++)
#pragma unroll
            for (int x = 0; x < 3; x++) {
                const uint2 r = range[(hx[x] ^ hy[y] ^ hz[z]) & s.mask];
                const bool ok = r.x != kSentinel;
                st[y * 3 + x] = ok ? r.x : 0u;
                pre[y * 3 + x + 1] = pre[y * 3 + x] + (ok ? r.y - r.x : 0u);
            }
        const uint32_t total = pre[9];
        for (uint32_t k = 0; k < total; k += 8) {
            uint32_t jj[8];
            float4 p[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const uint32_t q = (k + u < total) ? k + u : k;
                uint32_t b = 0;
#pragma unroll
                for (int bb = 1; bb < 9; bb++)
                    if (q >= pre[bb]) b = bb;
                // select st[b] / pre[b] without dynamic register indexing
                uint32_t sb = st[0], pb = pre[0];
#pragma unroll
                for (int bb = 1; bb < 9; bb++)
                    if (b == (uint32_t)bb) {
                        sb = st[bb];
                        pb = pre[bb];
                    }
                jj[u] = sb + (q - pb);
                p[u] = pos[jj[u]];
            }
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (k + u < total) f(jj[u], p[u]);
        }
    }
}

// Workgroups are dealt round-robin over the 8 XCDs (each with its own L2); give every XCD a
// contiguous range of the Morton-ordered work so neighbouring cells share one L2.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t g = gridDim.x, b = blockIdx.x;
    const uint32_t per = g / 8u, rem = g % 8u, x = b % 8u, k = b / 8u;
    return x * per + (x < rem ? x : rem) + k;
}

// density_and_alpha.comp:14-72, recording the gradient pairs into the neighbour list.
__global__ __launch_bounds__(kBlock) void kx_density_alpha(Sph s, const uint2* __restrict__ range,
                                                           const float4* __restrict__ pos, float* __restrict__ dens,
                                                           float* __restrict__ fac, const uint32_t* __restrict__ sched,
                                                           uint32_t* __restrict__ nbr, uint32_t* __restrict__ cnt) {
    const uint32_t t = xcd_block() * kBlock + threadIdx.x;
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
            if (m < kMaxNbr) nbr[(size_t)m * s.n + t] = j;
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
    const float alpha = sum_grad_sq + (gx * gx + gy * gy + gz * gz);
    fac[i] = alpha > 1e-6f ? alpha : 0.0f;
}

// ---- list consumers: policy per reference shader ---------------------------------------
// pair() repeats the reference's own tests (r^2 > h^2 rejects, gradient terms need
// r > 1e-6), so the same code serves list entries (which pass them) and the bucket walk.

// viscosity.comp:17-66 (velocity_b in, velocity_a out, w = 0, gravity fused)
struct XViscosity {
    const float4* vel;
    const float* dens;
    float4* out;
    float nu, gdx, gdy, gdz, dt;
    struct Self { float4 v; };
    struct Cand { float4 v; float rho; };
    struct Acc { float x, y, z; };
    __device__ Self self(uint32_t i) const { return {vel[i]}; }
    __device__ Cand load(uint32_t j) const { return {vel[j], dens[j]}; }
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
    __device__ unsigned long long finish(uint32_t i, const Acc& a, const Self& si, const Sph&) const {
        const float vx = si.v.x + nu * a.x, vy = si.v.y + nu * a.y, vz = si.v.z + nu * a.z;
        out[i] = make_float4(vx + gdx * dt, vy + gdy * dt, vz + gdz * dt, 0.0f);
        return 0;
    }
};

// density_source_term.comp:16-74 (kDensity) / divergence_source_term.comp:14-62
template <bool kDensity>
struct XSource {
    const float4* vel;
    const float* dens;
    float* src;
    float* pres;
    float dt, rho0;
    struct Self { float4 v; };
    struct Cand { float4 v; };
    struct Acc { float div; };
    __device__ Self self(uint32_t i) const { return {vel[i]}; }
    __device__ Cand load(uint32_t j) const { return {vel[j]}; }
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
    __device__ unsigned long long finish(uint32_t i, const Acc& a, const Self&, const Sph&) const {
        if (kDensity) {
            float source = 0.0f;
            if (dt > 1e-6f) source = ((rho0 - dens[i]) / dt) - a.div;
            src[i] = source;
            pres[i] = 0.0f;
        } else {
            src[i] = -a.div;
        }
        return 0;
    }
};

// pressure_force.comp:15-78
struct XPressureForce {
    const float* pres;
    const float* dens;
    float4* pacc;
    struct Self { float prho; };
    struct Cand { float p, rho; };
    struct Acc { float x, y, z; };
    __device__ Self self(uint32_t i) const {
        const float rho = dens[i];
        float pr = 0.0f;
        if (rho > 1e-6f) pr = pres[i] / (rho * rho);
        return {pr};
    }
    __device__ Cand load(uint32_t j) const { return {pres[j], dens[j]}; }
    __device__ Acc zero() const { return {0.f, 0.f, 0.f}; }
    __device__ void pair(Acc& a, const Self& si, const Sph& s, float rx, float ry, float rz, const Cand& c) const {
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        const float r = sqrtf(r2);
        if (r > 1e-6f && c.rho > 1e-6f) {
            const float g = exact_grad(r, s);
            const float p_rho_j = c.p / (c.rho * c.rho);
            const float mt = s.mass * (si.prho + p_rho_j);
            a.x += mt * (g * rx);
            a.y += mt * (g * ry);
            a.z += mt * (g * rz);
        }
    }
    __device__ unsigned long long finish(uint32_t i, const Acc& a, const Self&, const Sph&) const {
        pacc[i] = make_float4(-a.x, -a.y, -a.z, 0.0f);
        return 0;
    }
};

// pressure_update.comp:18-84 (in place: neighbours' pressures are not read), plus the
// sweep's quantised compression residual.
struct XPressureUpdate {
    const float4* pacc;
    const float* fac;
    const float* src;
    const float* dens;
    float* pres;
    float dt, omega, rho0;
    struct Self { float4 a; };
    struct Cand { float4 a; };
    struct Acc { float ap; };
    __device__ Self self(uint32_t i) const { return {pacc[i]}; }
    __device__ Cand load(uint32_t j) const { return {pacc[j]}; }
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
    __device__ unsigned long long finish(uint32_t i, const Acc& a, const Self&, const Sph&) const {
        const float alpha_i = fac[i], rho_i = dens[i], src_i = src[i], p_i = pres[i];
        const float ap = dt * a.ap;
        const float err = src_i - ap;
        const unsigned long long q = residual_q(err, dt, rho0);
        if (alpha_i > 1e-6f && rho_i > 1e-6f) {
            const float a_ii = -(dt / (rho_i * rho_i)) * alpha_i;
            if (fabsf(a_ii) > 1e-20f) {
                const float corr = (omega * err) / a_ii;
                const float np = p_i + corr;
                pres[i] = np > 0.0f ? np : 0.0f;
            }
        }
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

// One work-item per particle: the listed pairs (four candidate loads in flight), or the
// bucket walk when the list is stale or the particle overflowed it. With `prev` set, a
// converged previous sweep skips the launch; with `partial` set, each block stores the sum
// of its particles' finish() values.
template <class P>
__global__ __launch_bounds__(kBlock) void kx_list_gather(P pol, Sph s, const uint2* __restrict__ range,
                                                         const float4* __restrict__ pos,
                                                         const uint32_t* __restrict__ sched,
                                                         const uint32_t* __restrict__ nbr,
                                                         const uint32_t* __restrict__ cnt, uint32_t list_ok,
                                                         unsigned long long* __restrict__ partial,
                                                         const unsigned long long* __restrict__ prev,
                                                         unsigned long long thr) {
    if (prev && *prev <= thr) return;
    const uint32_t t = xcd_block() * kBlock + threadIdx.x;
    unsigned long long q = 0;
    if (t < s.n) {
        const uint32_t i = sched[t];
        const float4 pi = pos[i];
        const typename P::Self si = pol.self(i);
        typename P::Acc acc = pol.zero();
        const uint32_t m = list_ok ? cnt[t] : 0xFFFFFFFFu;
        if (m <= kMaxNbr) {
            for (uint32_t k = 0; k < m; k += 4) {
                uint32_t jj[4];
#pragma unroll
                for (int u = 0; u < 4; u++) jj[u] = nbr[(size_t)((k + u < m) ? k + u : k) * s.n + t];
                float4 pj[4];
                typename P::Cand cj[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    pj[u] = pos[jj[u]];
                    cj[u] = pol.load(jj[u]);
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (k + u < m) pol.pair(acc, si, s, pi.x - pj[u].x, pi.y - pj[u].y, pi.z - pj[u].z, cj[u]);
            }
        } else {
            visit27(range, pos, s, pi.x, pi.y, pi.z, [&](uint32_t j, const float4& pj) {
                pol.pair(acc, si, s, pi.x - pj.x, pi.y - pj.y, pi.z - pj.z, pol.load(j));
            });
        }
        q = pol.finish(i, acc, si, s);
    }
    if (partial) {
        const unsigned long long t = block_sum_u64(q);
        if (threadIdx.x == 0) partial[blockIdx.x] = t;
    }
}

inline dim3 grid_for(uint32_t count) { return dim3((count + kBlock - 1) / kBlock); }

template <class P>
hipError_t run_list(fe_ctx* c, const char* name, const P& pol, unsigned long long* partial = nullptr,
                    const unsigned long long* prev = nullptr, unsigned long long thr = 0) {
    Launch L(c, name);
    hipLaunchKernelGGL(kx_list_gather<P>, grid_for(c->n), dim3(kBlock), 0, c->stream, pol, sph_of(c), c->range,
                       c->pos_b, c->sched[0], c->nbr, c->nbr_count, c->list_valid ? 1u : 0u, partial, prev, thr);
    return hipGetLastError();
}

}  // namespace

hipError_t exact_density_alpha(fe_ctx* c) {
    Launch L(c, "density_alpha_ref");
    hipLaunchKernelGGL(kx_density_alpha, grid_for(c->n), dim3(kBlock), 0, c->stream, sph_of(c), c->range, c->pos_b,
                       c->dens, c->fac, c->sched[0], c->nbr, c->nbr_count);
    c->list_valid = true;
    return hipGetLastError();
}

hipError_t exact_viscosity(fe_ctx* c) {
    const fe_params& p = c->params;
    XViscosity pol{c->vel_b, c->dens, c->vel_a, p.viscosity, p.gravity[0], p.gravity[1], p.gravity[2], p.dt};
    return run_list(c, "viscosity_ref", pol);
}

hipError_t exact_density_source(fe_ctx* c) {
    XSource<true> pol{c->vel_a, c->dens, c->src, c->pres, c->params.dt, c->params.target_density};
    return run_list(c, "density_source_ref", pol);
}

hipError_t exact_divergence_source(fe_ctx* c) {
    XSource<false> pol{c->vel_a, c->dens, c->src, c->pres, c->params.dt, c->params.target_density};
    return run_list(c, "divergence_source_ref", pol);
}

hipError_t exact_pressure_force(fe_ctx* c, const unsigned long long* prev, unsigned long long thr) {
    XPressureForce pol{c->pres, c->dens, c->pacc};
    return run_list(c, "pressure_force_ref", pol, nullptr, prev, thr);
}

hipError_t exact_pressure_update(fe_ctx* c, unsigned long long* partial, const unsigned long long* prev,
                                 unsigned long long thr, uint32_t* partial_count) {
    XPressureUpdate pol{c->pacc, c->fac, c->src, c->dens, c->pres, c->params.dt, c->params.relax_factor,
                        c->params.target_density};
    *partial_count = grid_for(c->n).x;
    return run_list(c, "pressure_update_ref", pol, partial, prev, thr);
}

}  // namespace fe
