// fe_device.h — device helpers shared by the HIP kernels: the SPH constants, the cell hash
// (common.glsl:30-35) and the cubic-spline kernel (common.glsl:38-73).
#pragma once

#include "fe_internal.h"

namespace fe {

constexpr float kPi = 3.1415926535897932384626433832795f;
constexpr uint32_t kP1 = 73856093u, kP2 = 19349663u, kP3 = 83492791u;

// One scene's constants, as the kernels see them. A context holds one scene, or a batch of
// scenes laid out one after another: scene s owns particles [part0, part0 + n) in every
// particle order (position_a, hash order and schedule order keep scenes apart) and hash
// table entries [tab0, tab0 + T).
struct Sph {
    float h, hh, mass, k;   // k = 8 / (pi h^3) of common.glsl:42-43 / :61-62
    float inv_h;            // 1 / h: the cell coordinate is floor(pos * (1 / h))
    uint32_t n, mask;       // particles, T - 1 (T is a power of two, so % T == & (T - 1))
    uint32_t tab0, part0;   // first hash-table entry and first particle of the scene
    float dt, rho0, nu, omega, rad;
    float gx, gy, gz;
    float mnx, mny, mnz, mxx, mxy, mxz;   // box_min, box_max
    int c0x, c0y, c0z;                     // origin cell of the Morton schedule key
    uint32_t fast;                         // kFast* bits: exact shortcuts valid for this scene
    unsigned long long thr[2];             // converge thresholds (density, divergence)
};

// Exact shortcuts of the pair arithmetic, enabled per scene once shown to give the IEEE
// results for every input the kernels can pass them:
//  kFastDivH:   r / h as q0 = r (1/h), q0 + (r - q0 h)(1/h) with fused residuals, checked on
//               the device against the IEEE quotient for every float r in [1e-6, 1.001 h];
//  kFastDivBox: the gradient division gf / (h r) as the IEEE expansion's reciprocal and
//               residual steps alone: its operand scaling and final fix-up are identities
//               when both operands lie in [2^-40, 2^40), which the scene's h and mass bound.
constexpr uint32_t kFastDivH = 1u, kFastDivBox = 2u;

// The scene table of a batch (tab == nullptr: one scene, whose constants are passed by value).
struct Scenes {
    const Sph* tab;
    const uint8_t* sid;   // scene of each particle index (the same in every particle order)
    uint32_t count;
    // one scene replayed from a captured frame: the box (min xyz, max xyz) of the current
    // frame, copied to the device before each replay (a moving wall changes it every frame)
    const float* box = nullptr;
};

__device__ __forceinline__ Sph scene_at(const Scenes& sc, const Sph& s0, uint32_t idx) {
    return sc.tab ? sc.tab[sc.sid[idx]] : s0;
}
__device__ __forceinline__ uint32_t scene_id(const Scenes& sc, uint32_t idx) { return sc.tab ? sc.sid[idx] : 0u; }

// Scene constants from a parameter block (host side); fast: the scene's kFast* bits.
inline Sph sph_from(const fe_params& p, uint32_t n, uint32_t T, uint32_t tab0, uint32_t part0, uint32_t fast) {
    Sph s;
    s.h = p.smoothing_radius;
    s.hh = s.h * s.h;
    s.mass = p.particle_mass;
    const float h3 = s.h * s.h * s.h;
    s.k = 8.0f / (kPi * h3);
    s.inv_h = 1.0f / s.h;
    s.n = n;
    s.mask = T - 1u;
    s.tab0 = tab0;
    s.part0 = part0;
    s.dt = p.dt;
    s.rho0 = p.target_density;
    s.nu = p.viscosity;
    s.omega = p.relax_factor;
    s.rad = p.particle_radius;
    s.gx = p.gravity[0];
    s.gy = p.gravity[1];
    s.gz = p.gravity[2];
    s.mnx = p.box_min[0];
    s.mny = p.box_min[1];
    s.mnz = p.box_min[2];
    s.mxx = p.box_max[0];
    s.mxy = p.box_max[1];
    s.mxz = p.box_max[2];
    s.c0x = (int)floorf(p.box_min[0] * s.inv_h) - 1;
    s.c0y = (int)floorf(p.box_min[1] * s.inv_h) - 1;
    s.c0z = (int)floorf(p.box_min[2] * s.inv_h) - 1;
    s.fast = fast;
    s.thr[0] = s.thr[1] = 0ull;
    return s;
}

inline Sph sph_of(const fe_ctx* c) { return sph_from(c->params, c->n, c->T, 0u, 0u, c->exact_fast); }

inline Scenes scenes_of(const fe_ctx* c) { return Scenes{c->scene_tab, c->sid, c->nscenes, c->live_box ? c->box_dev : nullptr}; }

__device__ __forceinline__ uint32_t cell_hash(int cx, int cy, int cz, uint32_t mask) {
    return ((uint32_t)cx * kP1 ^ (uint32_t)cy * kP2 ^ (uint32_t)cz * kP3) & mask;
}

// ivec3(floor(pos / h)) evaluated as floor(pos * (1 / h)), the form the reference GPU ran.
__device__ __forceinline__ int3 cell_of(float x, float y, float z, float inv_h) {
    return make_int3((int)floorf(x * inv_h), (int)floorf(y * inv_h), (int)floorf(z * inv_h));
}

// Cubic spline of common.glsl:38-54 for an accepted pair (r <= h). The quotient r / h is
// taken as r * (1 / h) (the form the reference GPU used for pos / h).
__device__ __forceinline__ float spline_w(float r, const Sph& s) {
    const float q = r * s.inv_h;
    if (q >= 1.0f) return 0.0f;
    if (q <= 0.5f) {
        const float q2 = q * q;
        return s.k * (6.0f * (q2 * q) - 6.0f * q2 + 1.0f);
    }
    const float omq = 1.0f - q;
    return s.k * 2.0f * (omq * omq * omq);
}

// common.glsl:56-73: the scalar g with grad W = g * r_vec (zero for q >= 1 or r < 1e-6).
__device__ __forceinline__ float spline_grad(float r, const Sph& s) {
    const float q = r * s.inv_h;
    if (q >= 1.0f || r < 1e-6f) return 0.0f;
    float gf;
    if (q <= 0.5f) {
        gf = s.k * (18.0f * q * q - 12.0f * q);
    } else {
        const float omq = 1.0f - q;
        gf = s.k * -6.0f * omq * omq;
    }
    return gf * __builtin_amdgcn_rcpf(s.h * r);
}

// Quantised compression residual of one particle (FE_SOLVER_CONVERGE):
// q = floor(min(max(0, -(s - Ap)) * dt / rho0, 256) * 2^32); integer sums make the decision
// independent of summation order.
__device__ __forceinline__ unsigned long long residual_q(float s_minus_ap, float dt, float rho0) {
    const float neg = -s_minus_ap;
    if (!(neg > 0.0f)) return 0ull;
    float e = (neg * dt) / rho0;
    if (!(e < 256.0f)) e = 256.0f;
    return (unsigned long long)(e * 4294967296.0f);
}

// Hardware square root (1 ulp) for pair distances of the approximate kernels.
__device__ __forceinline__ float pair_sqrt(float r2) { return __builtin_amdgcn_sqrtf(r2); }

// The correctly rounded square root: the hardware result s (within 1 ulp) moved to s - ulp or
// s + ulp by the signs of the fused residuals x - (s -/+ ulp) s. This is the compiler's IEEE
// expansion without its rescaling of 0 < x < 2^-96 (such x take sqrtf) and without its
// pass-through of infinities and NaN, which no pair distance is.
__device__ __forceinline__ float sqrt_corrected(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u);
    const float sp = __uint_as_float(__float_as_uint(s) + 1u);
    const float rm = fmaf(-sm, s, x);
    const float rp = fmaf(-sp, s, x);
    const float t = rm <= 0.0f ? sm : s;
    return rp > 0.0f ? sp : t;
}
__device__ __forceinline__ float sqrt_exact(float x) {
    if (__builtin_expect(x > 0.0f && x < 0x1p-96f, 0)) return sqrtf(x);
    return sqrt_corrected(x);
}
// The pair distance of a gradient term: sqrt_corrected alone. For x < 2^-96 it may differ
// from sqrtf(x), but both are below 1e-6, where every gradient term is zero or skipped.
__device__ __forceinline__ float sqrt_grad(float x) { return sqrt_corrected(x); }

// n / d correctly rounded for 2^-40 <= |n|, |d| < 2^40: the reciprocal refinement and the
// two residual corrections of the compiler's IEEE expansion, whose operand scaling and
// fix-up steps change nothing inside that range.
__device__ __forceinline__ float div_boxed(float n, float d) {
    float y = __builtin_amdgcn_rcpf(d);
    const float e = fmaf(-d, y, 1.0f);
    y = fmaf(e, y, y);
    float q = n * y;
    float r = fmaf(-d, q, n);
    q = fmaf(r, y, q);
    r = fmaf(-d, q, n);
    return fmaf(r, y, q);
}

// r / h by the residual-corrected product with 1/h (r = 0 or 1e-6 <= r <= 1.001 h).
__device__ __forceinline__ float div_by_h(float r, const Sph& s) {
    const float q0 = r * s.inv_h;
    const float e = fmaf(-q0, s.h, r);
    return fmaf(e, s.inv_h, q0);
}

// q = r / h (IEEE) for a pair distance r <= 1.001 h.
__device__ __forceinline__ float q_of(float r, const Sph& s) {
    if ((s.fast & kFastDivH) && !(r > 0.0f && r < 1e-6f)) return div_by_h(r, s);
    return r / s.h;
}

// q = r / h (IEEE) for 1e-6 <= r <= 1.001 h, without the range test; for smaller r the
// result may differ, so only terms that skip r <= 1e-6 may use it.
__device__ __forceinline__ float q_big(float r, const Sph& s) {
    return (s.fast & kFastDivH) ? div_by_h(r, s) : r / s.h;
}

// common.glsl:38-54 kernel_w of q = r / h; both polynomial pieces are evaluated and one is
// selected.
__device__ __forceinline__ float w_of_q(float q, const Sph& s) {
    const float q2 = q * q;
    const float q3 = q2 * q;
    const float inner = s.k * (6.0f * q3 - 6.0f * q2 + 1.0f);
    const float omq = 1.0f - q;
    const float omq3 = omq * omq * omq;
    const float outer = s.k * 2.0f * omq3;
    const float w = q <= 0.5f ? inner : outer;
    return q >= 1.0f ? 0.0f : w;
}
__device__ __forceinline__ float exact_w(float r, const Sph& s) { return w_of_q(q_of(r, s), s); }

// common.glsl:56-73 kernel_grad of r and q = r / h; the scalar that multiplies r_vec
// (r <= 1.001 h). It is 0 for r < 1e-6, so q needs to be exact only from 1e-6 on.
__device__ __forceinline__ float grad_of_q(float r, float q, const Sph& s) {
    const float inner = s.k * (18.0f * q * q - 12.0f * q);
    const float omq = 1.0f - q;
    const float outer = s.k * -6.0f * omq * omq;
    const float gf = q <= 0.5f ? inner : outer;
    const float d = s.h * r;
    float g;
    if (s.fast & kFastDivBox)
        g = div_boxed(gf, d);
    else
        g = gf / d;
    return (q >= 1.0f || r < 1e-6f) ? 0.0f : g;
}
__device__ __forceinline__ float exact_grad(float r, const Sph& s) { return grad_of_q(r, q_big(r, s), s); }

// Wall samples of FE_BOUNDARY_PARTICLES (no reference counterpart): each face f of the box
// (f = 2 a + side: axis a, side 0 at box_min, 1 at box_max) carries a lattice of static samples
// at spacing sb = 2 r in the plane pl[f], r outside the face (box_min - r, box_max + r: one
// spacing from the fluid layer the collision keeps r inside), sample (k, m) at tangential
// coordinates min_u + (k + 1/2) sb and min_v + (m + 1/2) sb (u, v the other two axes in
// increasing order), k < n[f][0], m < n[f][1]. psi holds each sample's
// rho0 / sum_k W(x_b - x_k) (its boundary mass), faces concatenated from off[f], m-major.
struct Walls {
    const float* psi;
    float mn[3], mx[3];
    float pl[6];
    float sb;
    int R;               // window half-width in samples: ceil(h / sb) + 1
    int n[6][2];
    uint32_t off[6];
};

// Every wall sample within the window of point p, in a fixed order (faces 0..5, then m, then
// k), as fn(rx, ry, rz, index) with r = p - x_b. A face is visited when |p_a - plane| < h.
// The window is a superset of the samples within h; callers test r^2 <= h^2 themselves.
template <class F>
__host__ __device__ __forceinline__ void for_each_wall_sample(const Walls& w, float h, float px, float py, float pz,
                                                              F&& fn) {
    const float pc[3] = {px, py, pz};
#pragma unroll
    for (int f = 0; f < 6; f++) {
        const int a = f >> 1;
        const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
        const float plane = w.pl[f];
        if (!(fabsf(pc[a] - plane) < h)) continue;
        const int nu = w.n[f][0], nv = w.n[f][1];
        // nearest lower sample index, clamped before the conversion (a NaN coordinate gives -1)
        const float fu = fminf(fmaxf(floorf((pc[u] - w.mn[u]) / w.sb), -1.0f), (float)nu);
        const float fv = fminf(fmaxf(floorf((pc[v] - w.mn[v]) / w.sb), -1.0f), (float)nv);
        const int ku = (int)fu, kv = (int)fv;
        const int m0 = kv - w.R < 0 ? 0 : kv - w.R, m1 = kv + w.R > nv - 1 ? nv - 1 : kv + w.R;
        const int k0 = ku - w.R < 0 ? 0 : ku - w.R, k1 = ku + w.R > nu - 1 ? nu - 1 : ku + w.R;
        for (int m = m0; m <= m1; m++) {
            const float cv = w.mn[v] + ((float)m + 0.5f) * w.sb;
            for (int k = k0; k <= k1; k++) {
                const float cu = w.mn[u] + ((float)k + 0.5f) * w.sb;
                float r[3];
                r[a] = pc[a] - plane;
                r[u] = pc[u] - cu;
                r[v] = pc[v] - cv;
                fn(r[0], r[1], r[2], w.off[f] + (uint32_t)m * (uint32_t)nu + (uint32_t)k);
            }
        }
    }
}

// p / rho^2 of a particle whose density is <= 1e-6: pressure_force skips such neighbours.
// A signalling-NaN payload that no arithmetic result carries.
constexpr uint32_t kNoRho = 0x7FA00001u;

__device__ __forceinline__ float prho_of(float p, float rho) {
    return rho > 1e-6f ? p / (rho * rho) : __uint_as_float(kNoRho);
}

// Schedule position of the k-th listed pair of work-item t: the list is stored in tiles of
// 16 work-items x 4 slots (256 B), uint4 chunk (k / 4) of lane (t % 16).
__device__ __forceinline__ const uint4* list_tile(const uint32_t* nbr, uint32_t t) {
    return reinterpret_cast<const uint4*>(nbr) + (size_t)(t >> 4) * (kMaxNbr / 4) * 16 + (t & 15u);
}
__device__ __forceinline__ uint32_t list_pick(const uint4& q, uint32_t k) {
    const uint32_t u = k & 3u;
    return u == 0 ? q.x : u == 1 ? q.y : u == 2 ? q.z : q.w;
}

}  // namespace fe
