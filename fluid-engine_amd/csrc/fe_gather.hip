// fe_gather.hip — the 27-cell neighbour gathers of the DFSPH solver on CDNA4.
//
// The reference runs one invocation per particle, each walking the 27 buckets around its
// own cell (density_and_alpha.comp:31-45 and the same loop in viscosity, both source terms,
// pressure_force and pressure_update). Particles that share a cell share those 27 buckets,
// and the sort makes them contiguous. So here one wave64 takes a *cell group* (a maximal
// run of sorted particles with the same cell coordinate):
//   1. lanes 0..26 each load one neighbour bucket's [start, end) (z outer, x inner: the
//      reference visit order, repeated slots included) and a wave prefix sum numbers the
//      candidates;
//   2. the wave stages every candidate's fields into LDS with independent, coalesced-by-
//      bucket loads (up to 256 candidates per pass);
//   3. S = 64 / (group size) lanes per particle each test a strided share of the
//      candidates against r^2 <= h^2 from LDS (same-address reads broadcast), and the S
//      partial sums are combined with cross-lane shuffles.
// Every candidate the reference visits is visited here, with the same acceptance tests
// (r^2 > h^2 rejects, r > 1e-6 for gradient terms). Each particle's sum is accumulated in
// a different order than the reference's serial loop, and pair distances use the hardware
// square root and reciprocal; the results agree with the oracle to ~1e-7 relative.
#include "fe_device.h"

namespace fe {
namespace {

constexpr int kWaves = kBlock / 64;   // waves per block
constexpr uint32_t kCand = 256;       // candidates staged per pass per wave

template <bool B>
struct Stage {
    float4 a[kCand];
    float4 b[kCand];
    uint32_t idx[kCand];
};
template <>
struct Stage<false> {
    float4 a[kCand];
    float4 b[1];
    uint32_t idx[kCand];
};

__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(x, off, 64);
        if (lane >= off) x += t;
    }
    return x - v;
}

__device__ __forceinline__ float sum_lanes(float v, int S) {
    for (int o = S >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---- per-kernel policies -----------------------------------------------------------

// density_and_alpha.comp:14-72
struct DensityAlpha {
    static constexpr bool kB = false;
    const float4* pos;
    float* dens;
    float* fac;
    struct Self { float x, y, z; };
    struct Acc { float rho, sgs, gx, gy, gz; };
    __device__ void load(uint32_t j, float4& a, float4&) const { a = pos[j]; }
    __device__ Self self(uint32_t i) const {
        const float4 p = pos[i];
        return {p.x, p.y, p.z};
    }
    __device__ Acc zero() const { return {0.f, 0.f, 0.f, 0.f, 0.f}; }
    __device__ void pair(Acc& c, const Self& si, const float4& a, const float4*, const Sph& s) const {
        const float rx = si.x - a.x, ry = si.y - a.y, rz = si.z - a.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        const float r = pair_sqrt(r2);
        c.rho += s.mass * spline_w(r, s);
        if (r > 1e-6f) {
            const float g = spline_grad(r, s);
            const float mx = s.mass * (g * rx), my = s.mass * (g * ry), mz = s.mass * (g * rz);
            c.gx += mx;
            c.gy += my;
            c.gz += mz;
            c.sgs += mx * mx + my * my + mz * mz;
        }
    }
    __device__ void reduce(Acc& c, int S) const {
        c.rho = sum_lanes(c.rho, S);
        c.sgs = sum_lanes(c.sgs, S);
        c.gx = sum_lanes(c.gx, S);
        c.gy = sum_lanes(c.gy, S);
        c.gz = sum_lanes(c.gz, S);
    }
    __device__ unsigned long long finish(uint32_t i, const Acc& c, const Self&, const Sph&) const {
        dens[i] = c.rho;
        const float alpha = c.sgs + (c.gx * c.gx + c.gy * c.gy + c.gz * c.gz);
        fac[i] = alpha > 1e-6f ? alpha : 0.0f;
        return 0;
    }
};

// viscosity.comp:17-66 (velocity_b in, velocity_a out, gravity fused)
struct Viscosity {
    static constexpr bool kB = true;
    const float4* pos;
    const float4* vel;
    const float* dens;
    float4* out;
    float nu, gdx, gdy, gdz, dt, mass;
    struct Self { float x, y, z, vx, vy, vz; };
    struct Acc { float sx, sy, sz; };
    __device__ void load(uint32_t j, float4& a, float4& b) const {
        const float4 p = pos[j];
        a = make_float4(p.x, p.y, p.z, mass / dens[j]);
        b = vel[j];
    }
    __device__ Self self(uint32_t i) const {
        const float4 p = pos[i], v = vel[i];
        return {p.x, p.y, p.z, v.x, v.y, v.z};
    }
    __device__ Acc zero() const { return {0.f, 0.f, 0.f}; }
    __device__ void pair(Acc& c, const Self& si, const float4& a, const float4* bp, const Sph& s) const {
        const float rx = si.x - a.x, ry = si.y - a.y, rz = si.z - a.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        const float r = pair_sqrt(r2);
        if (r > 1e-6f) {
            const float w = spline_w(r, s);
            const float4 vj = *bp;
            c.sx += (a.w * (vj.x - si.vx)) * w;
            c.sy += (a.w * (vj.y - si.vy)) * w;
            c.sz += (a.w * (vj.z - si.vz)) * w;
        }
    }
    __device__ void reduce(Acc& c, int S) const {
        c.sx = sum_lanes(c.sx, S);
        c.sy = sum_lanes(c.sy, S);
        c.sz = sum_lanes(c.sz, S);
    }
    __device__ unsigned long long finish(uint32_t i, const Acc& c, const Self& si, const Sph&) const {
        const float vx = si.vx + nu * c.sx, vy = si.vy + nu * c.sy, vz = si.vz + nu * c.sz;
        out[i] = make_float4(vx + gdx * dt, vy + gdy * dt, vz + gdz * dt, 0.0f);
        return 0;
    }
};

// density_source_term.comp:16-74 (kDensity) and divergence_source_term.comp:14-62
template <bool kDensity>
struct Source {
    static constexpr bool kB = true;
    const float4* pos;
    const float4* vel;
    const float* dens;
    float* src;
    float* pres;
    float dt, rho0;
    struct Self { float x, y, z, vx, vy, vz; };
    struct Acc { float div; };
    __device__ void load(uint32_t j, float4& a, float4& b) const {
        a = pos[j];
        b = vel[j];
    }
    __device__ Self self(uint32_t i) const {
        const float4 p = pos[i], v = vel[i];
        return {p.x, p.y, p.z, v.x, v.y, v.z};
    }
    __device__ Acc zero() const { return {0.f}; }
    __device__ void pair(Acc& c, const Self& si, const float4& a, const float4* bp, const Sph& s) const {
        const float rx = si.x - a.x, ry = si.y - a.y, rz = si.z - a.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        const float r = pair_sqrt(r2);
        if (r > 1e-6f) {
            const float g = spline_grad(r, s);
            const float4 vj = *bp;
            const float dx = si.vx - vj.x, dy = si.vy - vj.y, dz = si.vz - vj.z;
            c.div += s.mass * (dx * (g * rx) + dy * (g * ry) + dz * (g * rz));
        }
    }
    __device__ void reduce(Acc& c, int S) const { c.div = sum_lanes(c.div, S); }
    __device__ unsigned long long finish(uint32_t i, const Acc& c, const Self&, const Sph&) const {
        if (kDensity) {
            float source = 0.0f;
            if (dt > 1e-6f) source = ((rho0 - dens[i]) / dt) - c.div;
            src[i] = source;
            pres[i] = 0.0f;
        } else {
            src[i] = -c.div;
        }
        return 0;
    }
};

// pressure_force.comp:15-78. Candidates carry p_j / rho_j^2 and whether rho_j > 1e-6.
struct PressureForce {
    static constexpr bool kB = true;
    const float4* pos;
    const float* pres;
    const float* dens;
    float4* pacc;
    struct Self { float x, y, z, prho; };
    struct Acc { float ax, ay, az; };
    __device__ void load(uint32_t j, float4& a, float4& b) const {
        const float4 p = pos[j];
        const float rho = dens[j];
        const bool ok = rho > 1e-6f;
        a = make_float4(p.x, p.y, p.z, ok ? pres[j] / (rho * rho) : 0.0f);
        b = make_float4(ok ? 1.0f : 0.0f, 0.f, 0.f, 0.f);
    }
    __device__ Self self(uint32_t i) const {
        const float4 p = pos[i];
        const float rho = dens[i];
        return {p.x, p.y, p.z, rho > 1e-6f ? pres[i] / (rho * rho) : 0.0f};
    }
    __device__ Acc zero() const { return {0.f, 0.f, 0.f}; }
    __device__ void pair(Acc& c, const Self& si, const float4& a, const float4* bp, const Sph& s) const {
        const float rx = si.x - a.x, ry = si.y - a.y, rz = si.z - a.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        const float r = pair_sqrt(r2);
        if (r > 1e-6f && bp->x != 0.0f) {
            const float g = spline_grad(r, s);
            const float mt = s.mass * (si.prho + a.w);
            c.ax += mt * (g * rx);
            c.ay += mt * (g * ry);
            c.az += mt * (g * rz);
        }
    }
    __device__ void reduce(Acc& c, int S) const {
        c.ax = sum_lanes(c.ax, S);
        c.ay = sum_lanes(c.ay, S);
        c.az = sum_lanes(c.az, S);
    }
    __device__ unsigned long long finish(uint32_t i, const Acc& c, const Self&, const Sph&) const {
        pacc[i] = make_float4(-c.ax, -c.ay, -c.az, 0.0f);
        return 0;
    }
};

// pressure_update.comp:18-84 (in place; neighbours' pressures are not read)
struct PressureUpdate {
    static constexpr bool kB = true;
    const float4* pos;
    const float4* pacc;
    const float* fac;
    const float* src;
    const float* dens;
    float* pres;
    float dt, omega, rho0;
    struct Self { float x, y, z, ax, ay, az; };
    struct Acc { float ap; };
    __device__ void load(uint32_t j, float4& a, float4& b) const {
        a = pos[j];
        b = pacc[j];
    }
    __device__ Self self(uint32_t i) const {
        const float4 p = pos[i], a = pacc[i];
        return {p.x, p.y, p.z, a.x, a.y, a.z};
    }
    __device__ Acc zero() const { return {0.f}; }
    __device__ void pair(Acc& c, const Self& si, const float4& a, const float4* bp, const Sph& s) const {
        const float rx = si.x - a.x, ry = si.y - a.y, rz = si.z - a.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > s.hh) return;
        const float r = pair_sqrt(r2);
        if (r > 1e-6f) {
            const float g = spline_grad(r, s);
            const float4 aj = *bp;
            const float dx = si.ax - aj.x, dy = si.ay - aj.y, dz = si.az - aj.z;
            c.ap += s.mass * (dx * (g * rx) + dy * (g * ry) + dz * (g * rz));
        }
    }
    __device__ void reduce(Acc& c, int S) const { c.ap = sum_lanes(c.ap, S); }
    __device__ unsigned long long finish(uint32_t i, const Acc& c, const Self&, const Sph&) const {
        const float alpha = fac[i], rho = dens[i], p = pres[i];
        const float ap = dt * c.ap;
        const float err = src[i] - ap;
        const unsigned long long q = residual_q(err, dt, rho0);
        if (alpha > 1e-6f && rho > 1e-6f) {
            const float a_ii = -(dt / (rho * rho)) * alpha;
            if (fabsf(a_ii) > 1e-20f) {
                const float np = p + (omega * err) / a_ii;
                pres[i] = np > 0.0f ? np : 0.0f;
            }
        }
        return q;
    }
};

// ---- the cell-group engine ---------------------------------------------------------

template <class P>
__device__ void process_group(const P& pol, const Sph& s, const uint2* __restrict__ range, Stage<P::kB>& st,
                              uint32_t gs, uint32_t ge, int3 gc, unsigned long long& qsum) {
    const int lane = threadIdx.x & 63;
    uint32_t bst = 0, bcnt = 0;
    if (lane < 27) {
        const int dz = lane / 9 - 1, dy = (lane / 3) % 3 - 1, dx = lane % 3 - 1;
        const uint2 r = range[cell_hash(gc.x + dx, gc.y + dy, gc.z + dz, s.mask)];
        if (r.x != kSentinel) {
            bst = r.x;
            bcnt = r.y - r.x;
        }
    }
    const uint32_t boff = wave_excl_scan(bcnt);
    const uint32_t C = __shfl(boff + bcnt, 26, 64);
    const uint32_t np = ge - gs;
    int S = 64;
    while (S > 1 && (uint32_t)S * np > 64u) S >>= 1;
    const uint32_t per_chunk = 64u / (uint32_t)S;
    uint32_t staged = 0xFFFFFFFFu;

    for (uint32_t pb = gs; pb < ge; pb += per_chunk) {
        const uint32_t pi = pb + (uint32_t)lane / (uint32_t)S;
        const int sub = lane & (S - 1);
        const bool act = pi < ge;
        const typename P::Self self = pol.self(act ? pi : gs);
        typename P::Acc acc = pol.zero();
        for (uint32_t cb = 0; cb < C; cb += kCand) {
            const uint32_t cn = (C - cb) < kCand ? (C - cb) : kCand;
            if (staged != cb) {
                wave_sync();
                if (bcnt) {
                    const uint32_t lo = boff > cb ? boff : cb;
                    const uint32_t hi = (boff + bcnt) < (cb + cn) ? (boff + bcnt) : (cb + cn);
                    for (uint32_t t = lo; t < hi; t++) st.idx[t - cb] = bst + (t - boff);
                }
                wave_sync();
                for (uint32_t t = lane; t < cn; t += 64) pol.load(st.idx[t], st.a[t], st.b[P::kB ? t : 0]);
                wave_sync();
                staged = cb;
            }
            if (act)
                for (uint32_t t = sub; t < cn; t += (uint32_t)S) pol.pair(acc, self, st.a[t], &st.b[P::kB ? t : 0], s);
        }
        pol.reduce(acc, S);
        if (act && sub == 0) qsum += pol.finish(pi, acc, self, s);
    }
}

// One wave per 64-particle window; the wave processes every cell group that starts in its
// window (a group running past the window end is finished by the same wave).
template <class P>
__global__ __launch_bounds__(kBlock) void k_gather(P pol, Sph s, const uint2* __restrict__ range,
                                                   const float4* __restrict__ pos,
                                                   unsigned long long* __restrict__ partial,
                                                   const unsigned long long* __restrict__ prev,
                                                   unsigned long long thr) {
    if (prev && *prev <= thr) return;
    __shared__ Stage<P::kB> stage[kWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t wid = blockIdx.x * kWaves + w;
    const uint32_t base = wid * 64u;
    unsigned long long qsum = 0;
    if (base < s.n) {
        const uint32_t i = base + lane;
        const bool valid = i < s.n;
        const float4 p = valid ? pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        const int3 c = cell_of(p.x, p.y, p.z, s.inv_h);
        int3 cp = make_int3(__shfl_up(c.x, 1, 64), __shfl_up(c.y, 1, 64), __shfl_up(c.z, 1, 64));
        bool first = i == 0;
        if (lane == 0 && base > 0) {
            const float4 q = pos[base - 1];
            cp = cell_of(q.x, q.y, q.z, s.inv_h);
        }
        const bool start = valid && (first || c.x != cp.x || c.y != cp.y || c.z != cp.z);
        uint64_t starts = __ballot(start);
        while (starts) {
            const int sl = __builtin_ctzll(starts);
            starts &= starts - 1;
            const uint32_t gs = base + (uint32_t)sl;
            const int3 gc = make_int3(__shfl(c.x, sl, 64), __shfl(c.y, sl, 64), __shfl(c.z, sl, 64));
            uint32_t ge;
            if (starts) {
                ge = base + (uint32_t)__builtin_ctzll(starts);
            } else {
                ge = base + 64u < s.n ? base + 64u : s.n;
                while (ge < s.n) {
                    const uint32_t j = ge + lane;
                    const bool vj = j < s.n;
                    const float4 q = vj ? pos[j] : make_float4(0.f, 0.f, 0.f, 0.f);
                    const int3 cj = cell_of(q.x, q.y, q.z, s.inv_h);
                    const uint64_t m = __ballot(!vj || cj.x != gc.x || cj.y != gc.y || cj.z != gc.z);
                    if (m) {
                        ge += (uint32_t)__builtin_ctzll(m);
                        break;
                    }
                    ge += 64u;
                }
            }
            process_group(pol, s, range, stage[w], gs, ge, gc, qsum);
        }
    }
    if (partial) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) qsum += __shfl_xor(qsum, off, 64);
        if (lane == 0) partial[wid] = qsum;
    }
}

// Sum of the per-wave residual partials into one sweep slot (single block, deterministic).
__global__ __launch_bounds__(kBlock) void k_resid_finalize(const unsigned long long* __restrict__ partial,
                                                           uint32_t count, unsigned long long* __restrict__ out,
                                                           const unsigned long long* __restrict__ prev,
                                                           unsigned long long thr) {
    if (prev && *prev <= thr) return;
    __shared__ unsigned long long part[kWaves];
    unsigned long long v = 0;
    for (uint32_t k = threadIdx.x; k < count; k += kBlock) v += partial[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int k = 0; k < kWaves; k++) t += part[k];
        *out = t;
    }
}

inline uint32_t gather_blocks(uint32_t n) { return (n + kBlock - 1) / kBlock; }

template <class P>
hipError_t run_gather(fe_ctx* c, const char* name, const P& pol, const float4* pos, unsigned long long* partial,
                      const unsigned long long* prev, unsigned long long thr) {
    Launch L(c, name);
    hipLaunchKernelGGL(k_gather<P>, dim3(gather_blocks(c->n)), dim3(kBlock), 0, c->stream, pol, sph_of(c),
                       c->range, pos, partial, prev, thr);
    return hipGetLastError();
}

}  // namespace

uint32_t gather_partials(uint32_t n) { return gather_blocks(n) * kWaves; }

hipError_t launch_density_alpha(fe_ctx* c) {
    if (c->options.arithmetic == FE_ARITH_REFERENCE) return exact_density_alpha(c);
    DensityAlpha p{c->pos_b, c->dens, c->fac};
    c->list_valid = false;   // this path does not record the neighbour list
    return run_gather(c, "density_alpha", p, c->pos_b, nullptr, nullptr, 0);
}

hipError_t launch_viscosity(fe_ctx* c) {
    if (c->options.arithmetic == FE_ARITH_REFERENCE) return exact_viscosity(c);
    const fe_params& q = c->params;
    Viscosity p{c->pos_b, c->vel_b, c->dens, c->vel_a, q.viscosity, q.gravity[0], q.gravity[1], q.gravity[2],
                q.dt, q.particle_mass};
    return run_gather(c, "viscosity", p, c->pos_b, nullptr, nullptr, 0);
}

hipError_t launch_density_source(fe_ctx* c) {
    if (c->options.arithmetic == FE_ARITH_REFERENCE) return exact_density_source(c);
    Source<true> p{c->pos_b, c->vel_a, c->dens, c->src, c->pres, c->params.dt, c->params.target_density};
    return run_gather(c, "density_source", p, c->pos_b, nullptr, nullptr, 0);
}

hipError_t launch_divergence_source(fe_ctx* c) {
    if (c->options.arithmetic == FE_ARITH_REFERENCE) return exact_divergence_source(c);
    Source<false> p{c->pos_b, c->vel_a, c->dens, c->src, c->pres, c->params.dt, c->params.target_density};
    return run_gather(c, "divergence_source", p, c->pos_b, nullptr, nullptr, 0);
}

hipError_t launch_pressure_force(fe_ctx* c, const unsigned long long* prev, unsigned long long thr) {
    if (c->options.arithmetic == FE_ARITH_REFERENCE) return exact_pressure_force(c, prev, thr);
    PressureForce p{c->pos_b, c->pres, c->dens, c->pacc};
    return run_gather(c, "pressure_force", p, c->pos_b, nullptr, prev, thr);
}

hipError_t launch_pressure_update(fe_ctx* c, unsigned long long* out, const unsigned long long* prev,
                                  unsigned long long thr) {
    hipError_t e;
    uint32_t parts = gather_partials(c->n);
    if (c->options.arithmetic == FE_ARITH_REFERENCE) {
        e = exact_pressure_update(c, out ? c->partial : nullptr, prev, thr, &parts);
    } else {
        PressureUpdate p{c->pos_b, c->pacc, c->fac, c->src, c->dens, c->pres, c->params.dt, c->params.relax_factor,
                         c->params.target_density};
        e = run_gather(c, "pressure_update", p, c->pos_b, out ? c->partial : nullptr, prev, thr);
    }
    if (e != hipSuccess || !out) return e;
    Launch L(c, "resid_finalize");
    hipLaunchKernelGGL(k_resid_finalize, dim3(1), dim3(kBlock), 0, c->stream, c->partial, parts, out, prev, thr);
    return hipGetLastError();
}

}  // namespace fe
