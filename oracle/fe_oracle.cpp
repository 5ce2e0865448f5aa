// fe_oracle.cpp — CPU restatement of the reference DFSPH frame (the oracle).
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP path; never the product path.
//
// Each function cites the reference file:line it restates. FP32 arithmetic follows the
// GLSL expression order exactly (built with -ffp-contract=off, IEEE division and sqrt);
// FP64 contexts run the same expressions in double. OpenMP parallelises the per-particle
// loops; every particle's sum is accumulated serially in the reference visit order, so
// results do not depend on the thread count.
#include "fe_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t kSentinel = 0xFFFFFFFFu;

template <class R>
struct V3 {
    R x, y, z;
};
template <class R>
inline V3<R> sub(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R>
inline V3<R> add(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R>
inline V3<R> scale(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R>
inline V3<R> scale_r(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
// GLSL dot(a, b), evaluated as (x*x' + y*y') + z*z'.
template <class R>
inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Literals of common.glsl / the shaders: float literals in FP32, their decimal value in FP64.
template <class R> struct Lit;
template <> struct Lit<float> {
    static constexpr float pi = 3.1415926535897932384626433832795f;
    static constexpr float eps6 = 1e-6f;
    static constexpr float eps20 = 1e-20f;
};
template <> struct Lit<double> {
    static constexpr double pi = 3.1415926535897932384626433832795;
    static constexpr double eps6 = 1e-6;
    static constexpr double eps20 = 1e-20;
};

// common.glsl:30-35 get_cell_hash. uint(ivec3) is a two's-complement reinterpretation.
inline uint32_t cell_hash(int32_t cx, int32_t cy, int32_t cz, uint32_t T) {
    return ((uint32_t)cx * 73856093u ^ (uint32_t)cy * 19349663u ^ (uint32_t)cz * 83492791u) % T;
}

// spatial_hash.comp:29 / density_and_alpha.comp:27: ivec3(floor(pos / h)), always in FP32.
// The vec3-by-scalar division is evaluated as pos * (1 / h) with the correctly rounded
// reciprocal: that is the form the reference GPU executed, fixed by the bucket order of
// scripts/rest_density.csv (all 4800 rows reproduce in place only with it; the IEEE
// quotient moves 240 lattice particles that sit exactly on a cell face).
inline float inv_h_of(float h) { return 1.0f / h; }
inline void cell_of(float px, float py, float pz, float h, int32_t c[3]) {
    const float ih = inv_h_of(h);
    c[0] = (int32_t)std::floor(px * ih);
    c[1] = (int32_t)std::floor(py * ih);
    c[2] = (int32_t)std::floor(pz * ih);
}

// common.glsl:38-54 kernel_w (cubic spline, support h).
template <class R>
R kernel_w(R r, R h) {
    R q = r / h;
    if (q >= R(1)) return R(0);
    R h3 = h * h * h;
    R k = R(8) / (Lit<R>::pi * h3);
    if (q <= R(0.5)) {
        R q2 = q * q;
        R q3 = q2 * q;
        return k * (R(6) * q3 - R(6) * q2 + R(1));
    }
    R omq = R(1) - q;
    R omq3 = omq * omq * omq;
    return k * R(2) * omq3;
}

// common.glsl:56-73 kernel_grad.
template <class R>
V3<R> kernel_grad(V3<R> rv, R r, R h) {
    R q = r / h;
    if (q >= R(1) || r < Lit<R>::eps6) return {R(0), R(0), R(0)};
    R h3 = h * h * h;
    R k = R(8) / (Lit<R>::pi * h3);
    R gf;
    if (q <= R(0.5)) {
        gf = k * (R(18) * q * q - R(12) * q);
    } else {
        R omq = R(1) - q;
        gf = k * R(-6) * omq * omq;
    }
    return scale(gf / (h * r), rv);
}

template <class R>
struct State {
    fe_params p{};
    fe_options o{};
    uint32_t n = 0, T = 0;
    std::vector<R> pos_a, pos_b, vel_a, vel_b, colors, pacc;  // 4*n
    std::vector<R> dens, fac, src, pres;                      // n
    std::vector<uint32_t> entries;                            // 2*T, {hash, index}
    std::vector<uint32_t> grid_start;                         // T
    uint32_t stats[3] = {0, 0, 0};
    uint32_t last_iters[2] = {0, 0};
    bool have_stats = false;
    fe_wall wall{};
    bool wall_on = false;
    std::vector<R> wall_psi;      // FE_BOUNDARY_PARTICLES: boundary mass of each wall sample
    float wall_key[9] = {};
    bool wall_ready = false;

    void alloc(uint32_t n_) {
        n = n_;
        T = 1;
        while (T < n) T <<= 1;
        for (auto* v : {&pos_a, &pos_b, &vel_a, &vel_b, &colors, &pacc}) v->assign(4 * (size_t)n, R(0));
        for (auto* v : {&dens, &fac, &src, &pres}) v->assign(n, R(0));
        entries.assign(2 * (size_t)T, 0);
        grid_start.assign(T, 0);
    }
    V3<R> v3(const std::vector<R>& a, uint32_t i) const { return {a[4 * i], a[4 * i + 1], a[4 * i + 2]}; }
    void put4(std::vector<R>& a, uint32_t i, V3<R> v, R w) {
        a[4 * i] = v.x; a[4 * i + 1] = v.y; a[4 * i + 2] = v.z; a[4 * i + 3] = w;
    }
    R P(float x) const { return R(x); }
};

// Visit the 27 cells around particle i (z outer, y, x inner) and every sorted entry j of
// each bucket (density_and_alpha.comp:31-45 and the same skeleton in the other gathers).
// f(j) is called in the reference visit order, including repeated buckets when two
// neighbour cells hash to the same slot.
template <class R, class F>
inline void for_each_candidate(const State<R>& s, uint32_t i, F&& f) {
    const float h = s.p.smoothing_radius;
    int32_t c[3];
    cell_of((float)s.pos_b[4 * i], (float)s.pos_b[4 * i + 1], (float)s.pos_b[4 * i + 2], h, c);
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                uint32_t key = cell_hash(c[0] + x, c[1] + y, c[2] + z, s.T);
                uint32_t start = s.grid_start[key];
                if (start == kSentinel) continue;
                for (uint32_t j = start; j < s.n; j++) {
                    if (s.entries[2 * (size_t)j] != key) break;
                    f(j);
                }
            }
}

// ---------------------------------------------------------------------------------------
// Neighbour search: spatial_hash.comp:23-36, sort, grid_offsets.comp:20-32, reorder.comp:37-45
// driven as in NeighborSearch::execute (neighbor_search.rs:176-231).

template <class R>
void spatial_hash(State<R>& s) {
    const float h = s.p.smoothing_radius;
    for (uint32_t i = 0; i < s.T; i++) {
        if (i < s.n) {
            int32_t c[3];
            cell_of((float)s.pos_a[4 * i], (float)s.pos_a[4 * i + 1], (float)s.pos_a[4 * i + 2], h, c);
            s.entries[2 * (size_t)i] = cell_hash(c[0], c[1], c[2], s.T);
            s.entries[2 * (size_t)i + 1] = i;
        } else {
            s.entries[2 * (size_t)i] = kSentinel;
            s.entries[2 * (size_t)i + 1] = kSentinel;
        }
    }
}

// radix_count.comp:26-47, radix_scan.comp:18-51, radix_reorder.comp:35-97 with the
// RadixSorter::execute driver (sorter.rs:325-455): 4 passes of 8 bits, WG = 256,
// EPI = 8 (chunk 2048), counts stored column-major [digit * num_wg + wg], ping-pong
// between the entry buffer and a temporary; the result lands back in the entry buffer.
void radix_sort_literal(uint32_t* pairs, uint32_t n) {
    const uint32_t WG = 256, EPI = 8, CHUNK = WG * EPI;
    const uint32_t num_wg = (n + CHUNK - 1) / CHUNK;
    std::vector<uint32_t> tmp(2 * (size_t)n);
    std::vector<uint32_t> counts(256 * (size_t)num_wg);
    uint32_t* bufs[2] = {pairs, tmp.data()};
    for (uint32_t pass = 0; pass < 4; pass++) {
        const uint32_t shift = 8 * pass;
        const uint32_t* in = bufs[pass & 1];
        uint32_t* out = bufs[(pass & 1) ^ 1];
        // count: one WG per chunk, every invocation counts EPI consecutive elements.
        for (uint32_t wg = 0; wg < num_wg; wg++) {
            uint32_t local[256] = {0};
            for (uint32_t lid = 0; lid < WG; lid++) {
                uint32_t base = wg * WG * EPI + lid * EPI;
                for (uint32_t e = 0; e < EPI; e++) {
                    uint32_t gi = base + e;
                    if (gi >= n) break;
                    local[(in[2 * (size_t)gi] >> shift) & 0xFFu]++;
                }
            }
            for (uint32_t d = 0; d < 256; d++) counts[(size_t)d * num_wg + wg] = local[d];
        }
        // scan: each digit's stripe serially (exclusive), then an inclusive prefix over the
        // 256 stripe totals added back as block offsets.
        uint32_t sums[256];
        for (uint32_t d = 0; d < 256; d++) {
            uint32_t running = 0;
            for (uint32_t k = d * num_wg; k < (d + 1) * num_wg; k++) {
                uint32_t v = counts[k];
                counts[k] = running;
                running += v;
            }
            sums[d] = running;
        }
        for (uint32_t stride = 1; stride < 256; stride *= 2) {
            uint32_t prev[256];
            std::memcpy(prev, sums, sizeof(sums));
            for (uint32_t d = stride; d < 256; d++) sums[d] += prev[d - stride];
        }
        for (uint32_t d = 0; d < 256; d++) {
            uint32_t block_offset = d == 0 ? 0 : sums[d - 1];
            for (uint32_t k = d * num_wg; k < (d + 1) * num_wg; k++) counts[k] += block_offset;
        }
        // reorder: invocation `lid` of WG `wg` walks the whole chunk in order and writes the
        // elements of digit `lid` to consecutive slots from its global offset (stable).
        for (uint32_t wg = 0; wg < num_wg; wg++) {
            uint32_t base = wg * CHUNK;
            uint32_t pos[256];
            for (uint32_t d = 0; d < 256; d++) pos[d] = counts[(size_t)d * num_wg + wg];
            for (uint32_t e = 0; e < CHUNK; e++) {
                uint32_t gi = base + e;
                if (gi >= n) continue;
                uint32_t d = (in[2 * (size_t)gi] >> shift) & 0xFFu;
                uint32_t o = pos[d]++;
                out[2 * (size_t)o] = in[2 * (size_t)gi];
                out[2 * (size_t)o + 1] = in[2 * (size_t)gi + 1];
            }
        }
    }
}

// bitonic_sort.comp:19-43 driven by GpuSorter::execute (sorter.rs:73-115): for every
// block height h = 2..n and step = h/2..1, invocation i < n/2 compares
// l = (i/step)*step*2 + i%step with l+step, ascending iff (l & h) == 0, and swaps the whole
// entry when out of order. Unstable but deterministic.
void bitonic_sort_literal(uint32_t* pairs, uint32_t n) {
    for (uint32_t h = 2; h <= n && h != 0; h *= 2) {
        for (uint32_t step = h / 2; step > 0; step /= 2) {
            for (uint32_t i = 0; i < n / 2; i++) {
                uint32_t l = (i / step) * step * 2 + (i % step);
                uint32_t r = l + step;
                uint32_t kl = pairs[2 * (size_t)l], kr = pairs[2 * (size_t)r];
                bool asc = (l & h) == 0;
                if ((asc && kl > kr) || (!asc && kl < kr)) {
                    std::swap(pairs[2 * (size_t)l], pairs[2 * (size_t)r]);
                    std::swap(pairs[2 * (size_t)l + 1], pairs[2 * (size_t)r + 1]);
                }
            }
        }
        if (h == 0x80000000u) break;
    }
}

template <class R>
void sort_entries(State<R>& s) {
    if (s.o.sort_algorithm == FE_SORT_BITONIC)
        bitonic_sort_literal(s.entries.data(), s.T);
    else
        radix_sort_literal(s.entries.data(), s.T);
}

// neighbor_search.rs:207-208 fill_buffer(grid_start, 0xFFFFFFFF) + grid_offsets.comp:20-32.
template <class R>
void grid_offsets(State<R>& s) {
    std::fill(s.grid_start.begin(), s.grid_start.end(), kSentinel);
    for (uint32_t i = 0; i < s.T; i++) {
        uint32_t key = s.entries[2 * (size_t)i];
        if (key == kSentinel) continue;
        uint32_t prev = i == 0 ? kSentinel : s.entries[2 * (size_t)(i - 1)];
        if (key != prev) s.grid_start[key] = i;
    }
}

// reorder.comp:37-45: pos_b[i] = pos_a[e[i].index], vel_b[i] = vel_a[e[i].index], i < N.
template <class R>
void reorder(State<R>& s) {
    for (uint32_t i = 0; i < s.n; i++) {
        uint32_t src = s.entries[2 * (size_t)i + 1];
        for (int k = 0; k < 4; k++) {
            s.pos_b[4 * (size_t)i + k] = s.pos_a[4 * (size_t)src + k];
            s.vel_b[4 * (size_t)i + k] = s.vel_a[4 * (size_t)src + k];
        }
    }
}

template <class R>
void neighbor_search(State<R>& s) {
    spatial_hash(s);
    sort_entries(s);
    grid_offsets(s);
    reorder(s);
}

// ---------------------------------------------------------------------------------------
// Solver kernels.

// density_and_alpha.comp:14-72.
// FE_BOUNDARY_PARTICLES (no reference counterpart): the wall samples of the six box faces.
// Face f = 2 a + side (axis a, side 0 at box_min, 1 at box_max) holds a lattice of spacing
// sb = 2 r in the plane box_min - r (side 0) or box_max + r (side 1), sample (k, m) at
// tangential coordinates min_u + (k + 1/2) sb, min_v + (m + 1/2) sb (u, v the other two axes in
// increasing order), k < n[f][0], m < n[f][1]; sample indices run face by face, m-major.
struct WallGeom {
    float mn[3], mx[3], pl[6], sb;
    int R;
    int n[6][2];
    uint32_t off[6], total;
};

inline WallGeom wall_geom(const fe_params& p) {
    WallGeom w{};
    for (int a = 0; a < 3; a++) {
        w.mn[a] = p.box_min[a];
        w.mx[a] = p.box_max[a];
    }
    w.sb = 2.0f * p.particle_radius;
    w.R = (int)std::ceil(p.smoothing_radius / w.sb) + 1;
    uint32_t off = 0;
    for (int f = 0; f < 6; f++) {
        const int a = f >> 1, u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
        w.pl[f] = (f & 1) ? w.mx[a] + p.particle_radius : w.mn[a] - p.particle_radius;
        const float eu = std::floor((w.mx[u] - w.mn[u]) / w.sb), ev = std::floor((w.mx[v] - w.mn[v]) / w.sb);
        w.n[f][0] = eu > 0.0f ? (int)eu : 0;
        w.n[f][1] = ev > 0.0f ? (int)ev : 0;
        w.off[f] = off;
        off += (uint32_t)w.n[f][0] * (uint32_t)w.n[f][1];
    }
    w.total = off;
    return w;
}

// Every sample in the window around p: faces whose plane is closer than h, then the samples
// within R lattice steps of p's own lattice cell, m outer, k inner; fn(r = p - x_b, index).
template <class R, class F>
void for_each_wall_sample(const WallGeom& w, R h, const V3<R>& p, F&& fn) {
    const R pc[3] = {p.x, p.y, p.z};
    for (int f = 0; f < 6; f++) {
        const int a = f >> 1, u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
        const R plane = R(w.pl[f]);
        if (!(std::fabs(pc[a] - plane) < h)) continue;
        const int nu = w.n[f][0], nv = w.n[f][1];
        const R fu = std::fmin(std::fmax(std::floor((pc[u] - R(w.mn[u])) / R(w.sb)), R(-1)), R(nu));
        const R fv = std::fmin(std::fmax(std::floor((pc[v] - R(w.mn[v])) / R(w.sb)), R(-1)), R(nv));
        const int ku = (int)fu, kv = (int)fv;
        const int m0 = std::max(kv - w.R, 0), m1 = std::min(kv + w.R, nv - 1);
        const int k0 = std::max(ku - w.R, 0), k1 = std::min(ku + w.R, nu - 1);
        for (int m = m0; m <= m1; m++) {
            const R cv = R(w.mn[v]) + (R(m) + R(0.5)) * R(w.sb);
            for (int k = k0; k <= k1; k++) {
                const R cu = R(w.mn[u]) + (R(k) + R(0.5)) * R(w.sb);
                R r[3];
                r[a] = pc[a] - plane;
                r[u] = pc[u] - cu;
                r[v] = pc[v] - cv;
                fn(V3<R>{r[0], r[1], r[2]}, w.off[f] + (uint32_t)m * (uint32_t)nu + (uint32_t)k);
            }
        }
    }
}

// Each sample's boundary mass rho0 / sum_k W(x_b - x_k) (the sample itself included), for
// the current box, r, h and rho0; recomputed when one of them changes.
template <class R>
const WallGeom* walls(State<R>& s) {
    if (s.o.boundary != FE_BOUNDARY_PARTICLES) return nullptr;
    static thread_local WallGeom geom;
    geom = wall_geom(s.p);
    const float key[9] = {s.p.box_min[0], s.p.box_min[1], s.p.box_min[2], s.p.box_max[0], s.p.box_max[1],
                          s.p.box_max[2], s.p.particle_radius, s.p.smoothing_radius, s.p.target_density};
    if (s.wall_ready && std::memcmp(key, s.wall_key, sizeof(key)) == 0) return &geom;
    const R h = s.P(s.p.smoothing_radius), hh = h * h, rho0 = s.P(s.p.target_density);
    s.wall_psi.assign(geom.total, R(0));
    for (int f = 0; f < 6; f++) {
        const int a = f >> 1, u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
        for (int m = 0; m < geom.n[f][1]; m++)
            for (int k = 0; k < geom.n[f][0]; k++) {
                R x[3];
                x[a] = R(geom.pl[f]);
                x[u] = R(geom.mn[u]) + (R(k) + R(0.5)) * R(geom.sb);
                x[v] = R(geom.mn[v]) + (R(m) + R(0.5)) * R(geom.sb);
                R sum = R(0);
                for_each_wall_sample(geom, h, V3<R>{x[0], x[1], x[2]}, [&](const V3<R>& rv, uint32_t) {
                    const R r2 = dot(rv, rv);
                    if (r2 > hh) return;
                    sum += kernel_w(std::sqrt(r2), h);
                });
                s.wall_psi[geom.off[f] + (uint32_t)m * (uint32_t)geom.n[f][0] + (uint32_t)k] = rho0 / sum;
            }
    }
    std::memcpy(s.wall_key, key, sizeof(key));
    s.wall_ready = true;
    return &geom;
}

// A particle's sums over the wall samples, fixed while the positions are: G = sum psi grad W
// over the samples with r^2 <= h^2 and r > 1e-6, rho_w = sum psi W over those with r^2 <= h^2,
// in enumeration order. The samples are at rest, so they add v . G to a velocity divergence,
// a . G to A p, (p / rho^2) G to the pressure acceleration (mirrored pressure), and G and
// rho_w to the density pass's gradient sum and density.
template <class R>
struct WallTerms {
    V3<R> g{R(0), R(0), R(0)};
    R rho = R(0);
};

template <class R>
WallTerms<R> wall_terms(const State<R>& s, const WallGeom& w, R h, const V3<R>& p) {
    const R hh = h * h;
    WallTerms<R> t;
    for_each_wall_sample(w, h, p, [&](const V3<R>& rv, uint32_t b) {
        R r2 = dot(rv, rv);
        if (r2 > hh) return;
        R r = std::sqrt(r2);
        t.rho += s.wall_psi[b] * kernel_w(r, h);
        if (r > Lit<R>::eps6) t.g = add(t.g, scale(s.wall_psi[b], kernel_grad(rv, r, h)));
    });
    return t;
}

template <class R>
void density_alpha(State<R>& s) {
    const R h = s.P(s.p.smoothing_radius), mass = s.P(s.p.particle_mass);
    const R hh = h * h;
    const WallGeom* wg = walls(s);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s.n; ii++) {
        uint32_t i = (uint32_t)ii;
        V3<R> pi = s.v3(s.pos_b, i);
        R density = R(0), sum_grad_sq = R(0);
        V3<R> grad_sum{R(0), R(0), R(0)};
        for_each_candidate(s, i, [&](uint32_t j) {
            V3<R> rv = sub(pi, s.v3(s.pos_b, j));
            R r2 = dot(rv, rv);
            if (r2 > hh) return;
            R r = std::sqrt(r2);
            R w = kernel_w(r, h);
            density += mass * w;
            if (r > Lit<R>::eps6) {
                V3<R> g = kernel_grad(rv, r, h);
                V3<R> mg = scale(mass, g);
                grad_sum = add(grad_sum, mg);
                sum_grad_sq += dot(mg, mg);
            }
        });
        if (wg) {
            const WallTerms<R> wt = wall_terms(s, *wg, h, pi);
            density += wt.rho;
            grad_sum = add(grad_sum, wt.g);
        }
        s.dens[i] = density;
        R alpha = sum_grad_sq + dot(grad_sum, grad_sum);
        s.fac[i] = alpha > Lit<R>::eps6 ? alpha : R(0);
    }
}

// viscosity.comp:17-66 (reads velocity_b, writes velocity_a with w = 0; gravity fused).
template <class R>
void viscosity(State<R>& s) {
    const R h = s.P(s.p.smoothing_radius), mass = s.P(s.p.particle_mass);
    const R hh = h * h, nu = s.P(s.p.viscosity), dt = s.P(s.p.dt);
    const V3<R> g{s.P(s.p.gravity[0]), s.P(s.p.gravity[1]), s.P(s.p.gravity[2])};
    std::vector<R> out(4 * (size_t)s.n);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s.n; ii++) {
        uint32_t i = (uint32_t)ii;
        V3<R> pi = s.v3(s.pos_b, i);
        V3<R> sum{R(0), R(0), R(0)};
        for_each_candidate(s, i, [&](uint32_t j) {
            V3<R> rv = sub(pi, s.v3(s.pos_b, j));
            R r2 = dot(rv, rv);
            if (r2 > hh) return;
            R r = std::sqrt(r2);
            R w = kernel_w(r, h);
            if (r > Lit<R>::eps6) {
                V3<R> vd = sub(s.v3(s.vel_b, j), s.v3(s.vel_b, i));
                R m_rho = mass / s.dens[j];
                sum = add(sum, scale_r(scale(m_rho, vd), w));
            }
        });
        V3<R> v = s.v3(s.vel_b, i);
        V3<R> visco = add(v, scale(nu, sum));
        V3<R> fin = add(visco, scale_r(g, dt));
        out[4 * (size_t)i] = fin.x; out[4 * (size_t)i + 1] = fin.y;
        out[4 * (size_t)i + 2] = fin.z; out[4 * (size_t)i + 3] = R(0);
    }
    s.vel_a.swap(out);
}

// density_source_term.comp:16-74 (velocity_a; pressures zero-initialised).
template <class R>
void density_source(State<R>& s) {
    const R h = s.P(s.p.smoothing_radius), mass = s.P(s.p.particle_mass);
    const R hh = h * h, dt = s.P(s.p.dt), rho0 = s.P(s.p.target_density);
    const WallGeom* wg = walls(s);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s.n; ii++) {
        uint32_t i = (uint32_t)ii;
        V3<R> pi = s.v3(s.pos_b, i), vi = s.v3(s.vel_a, i);
        R rho_i = s.dens[i];
        R div = R(0);
        for_each_candidate(s, i, [&](uint32_t j) {
            V3<R> rv = sub(pi, s.v3(s.pos_b, j));
            R r2 = dot(rv, rv);
            if (r2 > hh) return;
            R r = std::sqrt(r2);
            if (r > Lit<R>::eps6) {
                V3<R> g = kernel_grad(rv, r, h);
                V3<R> vd = sub(vi, s.v3(s.vel_a, j));
                div += mass * dot(vd, g);
            }
        });
        if (wg) div += dot(vi, wall_terms(s, *wg, h, pi).g);
        R source = R(0);
        if (dt > Lit<R>::eps6) source = ((rho0 - rho_i) / dt) - div;
        s.src[i] = source;
        s.pres[i] = R(0);
    }
}

// pressure_force.comp:15-78.
template <class R>
void pressure_force(State<R>& s) {
    const R h = s.P(s.p.smoothing_radius), mass = s.P(s.p.particle_mass);
    const R hh = h * h;
    const WallGeom* wg = walls(s);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s.n; ii++) {
        uint32_t i = (uint32_t)ii;
        V3<R> pi = s.v3(s.pos_b, i);
        R rho_i = s.dens[i], p_i = s.pres[i];
        R p_rho_i = R(0);
        if (rho_i > Lit<R>::eps6) p_rho_i = p_i / (rho_i * rho_i);
        V3<R> acc{R(0), R(0), R(0)};
        for_each_candidate(s, i, [&](uint32_t j) {
            V3<R> rv = sub(pi, s.v3(s.pos_b, j));
            R r2 = dot(rv, rv);
            if (r2 > hh) return;
            R r = std::sqrt(r2);
            if (r > Lit<R>::eps6) {
                R rho_j = s.dens[j], p_j = s.pres[j];
                if (rho_j > Lit<R>::eps6) {
                    V3<R> g = kernel_grad(rv, r, h);
                    R p_rho_j = p_j / (rho_j * rho_j);
                    R term = p_rho_i + p_rho_j;
                    acc = add(acc, scale(mass * term, g));
                }
            }
        });
        if (wg) acc = add(acc, scale(p_rho_i, wall_terms(s, *wg, h, pi).g));
        s.put4(s.pacc, i, V3<R>{-acc.x, -acc.y, -acc.z}, R(0));
    }
}

// Convergence residual quantisation (no reference counterpart; documented in the header).
inline uint64_t residual_q(float s_minus_ap, float dt, float rho0) {
    float neg = -s_minus_ap;
    if (!(neg > 0.0f)) return 0;
    float e = (neg * dt) / rho0;
    if (!(e < 256.0f)) e = 256.0f;
    return (uint64_t)(e * 4294967296.0f);
}

// pressure_update.comp:18-84 (in place; neighbours' pressures are not read).
template <class R>
uint64_t pressure_update(State<R>& s) {
    const R h = s.P(s.p.smoothing_radius), mass = s.P(s.p.particle_mass);
    const R hh = h * h, dt = s.P(s.p.dt), omega = s.P(s.p.relax_factor);
    const WallGeom* wg = walls(s);
    uint64_t resid = 0;
#pragma omp parallel for schedule(static) reduction(+ : resid)
    for (int64_t ii = 0; ii < (int64_t)s.n; ii++) {
        uint32_t i = (uint32_t)ii;
        V3<R> pi = s.v3(s.pos_b, i), ai = s.v3(s.pacc, i);
        R alpha_i = s.fac[i], rho_i = s.dens[i], src_i = s.src[i], p_i = s.pres[i];
        R sum_ap = R(0);
        for_each_candidate(s, i, [&](uint32_t j) {
            V3<R> rv = sub(pi, s.v3(s.pos_b, j));
            R r2 = dot(rv, rv);
            if (r2 > hh) return;
            R r = std::sqrt(r2);
            if (r > Lit<R>::eps6) {
                V3<R> g = kernel_grad(rv, r, h);
                V3<R> ad = sub(ai, s.v3(s.pacc, j));
                sum_ap += mass * dot(ad, g);
            }
        });
        if (wg) sum_ap += dot(ai, wall_terms(s, *wg, h, pi).g);
        R ap = dt * sum_ap;
        R err = src_i - ap;
        // FE_RESIDUAL_TWO_SIDED: a particle entering the sweep with pressure counts |s - Ap|
        const R counted = s.o.residual == FE_RESIDUAL_TWO_SIDED && p_i > R(0) ? -std::fabs(err) : err;
        resid += residual_q((float)counted, (float)dt, s.p.target_density);
        if (alpha_i > Lit<R>::eps6 && rho_i > Lit<R>::eps6) {
            R a_ii = -(dt / (rho_i * rho_i)) * alpha_i;
            if (std::fabs(a_ii) > Lit<R>::eps20) {
                R corr = (omega * err) / a_ii;
                R np = p_i + corr;
                s.pres[i] = np > R(0) ? np : R(0);
            }
        }
    }
    return resid;
}

// pressure_integration.comp:16-77 (reads position_b, writes position_a; collision order
// y, x, z with restitution 0.5, friction 0.95, eps 0.001).
template <class R>
void pressure_integration(State<R>& s) {
    const R dt = s.P(s.p.dt), rad = s.P(s.p.particle_radius);
    const R mn[3] = {s.P(s.p.box_min[0]), s.P(s.p.box_min[1]), s.P(s.p.box_min[2])};
    const R mx[3] = {s.P(s.p.box_max[0]), s.P(s.p.box_max[1]), s.P(s.p.box_max[2])};
    const R rest = R(0.5), fric = R(0.95), eps = R(0.001);
    const int order[3] = {1, 0, 2};
    for (uint32_t i = 0; i < s.n; i++) {
        R pos[3], vel[3], ap[3];
        for (int k = 0; k < 3; k++) {
            pos[k] = s.pos_b[4 * (size_t)i + k];
            vel[k] = s.vel_a[4 * (size_t)i + k];
            ap[k] = s.pacc[4 * (size_t)i + k];
        }
        R nv[3], np[3];
        for (int k = 0; k < 3; k++) nv[k] = vel[k] + ap[k] * dt;
        for (int k = 0; k < 3; k++) np[k] = pos[k] + nv[k] * dt;
        for (int a : order) {
            int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
            if (np[a] < mn[a] + rad) {
                np[a] = mn[a] + rad + eps;
                nv[a] *= -rest;
                nv[t1] *= fric;
                nv[t2] *= fric;
            } else if (np[a] > mx[a] - rad) {
                np[a] = mx[a] - rad - eps;
                nv[a] *= -rest;
                nv[t1] *= fric;
                nv[t2] *= fric;
            }
        }
        s.put4(s.pos_a, i, V3<R>{np[0], np[1], np[2]}, R(1));
        s.put4(s.vel_a, i, V3<R>{nv[0], nv[1], nv[2]}, R(0));
    }
}

// divergence_source_term.comp:14-62 (velocity_a indexed by the current sort order).
template <class R>
void divergence_source(State<R>& s) {
    const R h = s.P(s.p.smoothing_radius), mass = s.P(s.p.particle_mass);
    const R hh = h * h;
    const WallGeom* wg = walls(s);
    // FE_DIVERGENCE_CONSISTENT (no reference counterpart): each particle's own velocity, as
    // reordered with the positions into velocity_b; the reference reads velocity_a, still in
    // the previous order, at the current indices.
    const std::vector<R>& vel = s.o.divergence_order == FE_DIVERGENCE_CONSISTENT ? s.vel_b : s.vel_a;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s.n; ii++) {
        uint32_t i = (uint32_t)ii;
        V3<R> pi = s.v3(s.pos_b, i), vi = s.v3(vel, i);
        R div = R(0);
        for_each_candidate(s, i, [&](uint32_t j) {
            V3<R> rv = sub(pi, s.v3(s.pos_b, j));
            R r2 = dot(rv, rv);
            if (r2 > hh) return;
            R r = std::sqrt(r2);
            if (r > Lit<R>::eps6) {
                V3<R> g = kernel_grad(rv, r, h);
                V3<R> vd = sub(vi, s.v3(vel, j));
                div += mass * dot(vd, g);
            }
        });
        if (wg) div += dot(vi, wall_terms(s, *wg, h, pi).g);
        s.src[i] = -div;
    }
}

// GLSL mix(x, y, a) = x * (1 - a) + y * a.
template <class R>
inline R mixf(R x, R y, R a) { return x * (R(1) - a) + y * a; }

// divergence_integration.comp:14-43.
template <class R>
void divergence_integration(State<R>& s) {
    const R dt = s.P(s.p.dt);
    const R deep[3] = {R(0.05), R(0.1), R(0.5)};
    const R water[3] = {R(0.2), R(0.5), R(1.0)};
    const R foam[3] = {R(1), R(1), R(1)};
    // FE_DIVERGENCE_CONSISTENT: the particle at sorted index i is entries[i].index of
    // velocity_a / colors; its corrected velocity also goes to velocity_b[i].
    const bool consistent = s.o.divergence_order == FE_DIVERGENCE_CONSISTENT;
    for (uint32_t i = 0; i < s.n; i++) {
        V3<R> v = s.v3(consistent ? s.vel_b : s.vel_a, i), a = s.v3(s.pacc, i);
        V3<R> nv = add(v, scale_r(a, dt));
        R speed = std::sqrt(dot(nv, nv));
        R t = speed / R(4);
        t = std::min(std::max(t, R(0)), R(1));
        R col[3];
        for (int k = 0; k < 3; k++)
            col[k] = t < R(0.5) ? mixf(deep[k], water[k], t * R(2))
                                : mixf(water[k], foam[k], (t - R(0.5)) * R(2));
        const uint32_t dst = consistent ? s.entries[2 * (size_t)i + 1] : i;
        s.put4(s.vel_a, dst, nv, R(0));
        s.put4(s.colors, dst, V3<R>{col[0], col[1], col[2]}, R(1));
        if (consistent) s.put4(s.vel_b, i, nv, R(0));
    }
}

// uint(x) for x >= 0 with the conversion saturating at the uint range (NaN -> 0).
inline uint32_t to_u32_sat(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

// stats_pipeline.rs:56-66 fill 0 + stats.comp:18-30.
template <class R>
void stats(State<R>& s) {
    uint32_t st[3] = {0, 0, 0};
    const float rho0 = s.p.target_density;
    for (uint32_t i = 0; i < s.n; i++) {
        float vx = (float)s.vel_a[4 * (size_t)i], vy = (float)s.vel_a[4 * (size_t)i + 1],
              vz = (float)s.vel_a[4 * (size_t)i + 2];
        float speed = std::sqrt(vx * vx + vy * vy + vz * vz);
        uint32_t bits;
        std::memcpy(&bits, &speed, 4);
        st[0] = std::max(st[0], bits);
        float de = std::fabs((float)s.dens[i] - rho0);
        st[1] += std::min(to_u32_sat(de * 1.0f), 10000u);
        float ve = std::fabs((float)s.src[i]);
        st[2] += std::min(to_u32_sat(ve * 10.0f), 10000u);
    }
    std::memcpy(s.stats, st, sizeof(st));
    s.have_stats = true;
}

// Jacobi loop: src/renderer/mod.rs:301-306 (density) and :327-332 (divergence); CONVERGE
// mode stops on the quantised compression residual of the sweep.
template <class R>
uint32_t solve(State<R>& s, uint32_t fixed_iters, float tol) {
    if (s.o.solver_mode != FE_SOLVER_CONVERGE) {
        for (uint32_t k = 0; k < fixed_iters; k++) {
            pressure_force(s);
            pressure_update(s);
        }
        return fixed_iters;
    }
    const uint64_t thr = (uint64_t)((double)tol * (double)s.n * 4294967296.0);
    uint32_t maxit = s.o.max_iterations ? s.o.max_iterations : 1;
    for (uint32_t k = 0; k < maxit; k++) {
        pressure_force(s);
        uint64_t q = pressure_update(s);
        if (q <= thr) return k + 1;
    }
    return maxit;
}

// Renderer::step host control (src/renderer/mod.rs:231-267) — CFL dt and early exit.
template <class R>
void frame_control(State<R>& s, uint32_t* dit, uint32_t* vit) {
    *dit = s.p.density_solver_iterations;
    *vit = s.p.divergence_solver_iterations;
    // stats_buffer starts as [0, 0, 0] (particle.rs:270-281) and is read every frame.
    float n = (float)s.n;
    float max_speed;
    std::memcpy(&max_speed, &s.stats[0], 4);
    float avg_de = (float)s.stats[1] / (1.0f * n);
    float avg_ve = (float)s.stats[2] / (10.0f * n);
    if (s.o.use_cfl && max_speed > 0.01f) {
        float h = s.p.smoothing_radius;
        float cfl = std::min(std::max(0.4f * h / max_speed, 0.001f), 0.05f);
        s.p.dt = std::min(cfl, s.p.dt * 1.1f);
    }
    float rho0 = s.p.target_density;
    float dthr = s.o.density_error_pct / 100.0f * rho0;
    float vthr = s.o.divergence_error_pct / 100.0f * rho0;
    if (s.o.use_solver_error_threshold && avg_de < dthr) *dit = 2;
    if (s.o.use_solver_error_threshold && avg_ve < vthr) *vit = 1;
}

// CollisionBox::update (src/entities/collision.rs:31-35): `self.time += dt; let offset =
// self.wave_amplitude * (self.time * self.wave_frequency * TAU).sin(); self.min.x =
// self.base_min_x + offset;` in f32, left to right, TAU = std::f32::consts::TAU.
void wall_update(fe_wall* w, float dt) {
    const float tau = (float)(2.0 * 3.14159265358979323846);
    w->time += dt;
    float arg = w->time * w->wave_frequency;
    arg = arg * tau;
    float offset = w->wave_amplitude * sinf(arg);
    w->min[0] = w->base_min_x + offset;
}

// Renderer::step physics frame (src/renderer/mod.rs:279-342). With a moving wall, the UI
// pass that precedes the frame copies the box into the parameters (src/renderer/ui.rs:
// 125-126) and the wall advances by max_dt once the parameters are uploaded (mod.rs:269-270).
template <class R>
void step(State<R>& s, float max_dt) {
    if (s.wall_on) {
        for (int a = 0; a < 3; a++) {
            s.p.box_min[a] = s.wall.min[a];
            s.p.box_max[a] = s.wall.max[a];
        }
        s.p.box_min[3] = 0.0f;
        s.p.box_max[3] = 0.0f;
    }
    uint32_t dit, vit;
    frame_control(s, &dit, &vit);
    if (s.wall_on) wall_update(&s.wall, max_dt);
    neighbor_search(s);
    density_alpha(s);
    float t = 0.0f;
    while (t < max_dt) {
        viscosity(s);
        density_source(s);
        s.last_iters[0] = solve(s, dit, s.o.density_tol);
        pressure_integration(s);
        t += s.p.dt;
        neighbor_search(s);
        density_alpha(s);
        divergence_source(s);
        s.last_iters[1] = solve(s, vit, s.o.divergence_tol);
        divergence_integration(s);
    }
    stats(s);
}

}  // namespace

// splat_density.comp:11-51 after the texture clear of src/renderer/mod.rs:386-388, over
// position_a with the parameter block's box and grid_res (density_texture.rs:30-55 binds
// position_a, the R32Uint image and sim_params). The texture is grid_res.x * grid_res.y *
// grid_res.z uint32, x fastest. Every particle whose uvw lies in [0, 1]^3 adds
// uint(weight * 200) for each texel of the 17^3 cube around its centre texel that is inside
// the image and at distance < 8. uvw = (pos - box_min) / box_size is evaluated as
// (pos - box_min) * (1 / box_size), the form the reference GPU used for the hash's division by
// a uniform (see cell_of). A particle with a NaN uvw (undefined in the shader: every
// comparison is false and ivec3(NaN) has no value) is skipped. Each thread accumulates a
// private texture that is summed at the end: unsigned addition modulo 2^32 is the atomic's.
template <class R>
void splat_density(const State<R>& s, uint32_t* out) {
    const int32_t rx = s.p.grid_res[0], ry = s.p.grid_res[1], rz = s.p.grid_res[2];
    const size_t cells = (size_t)rx * (size_t)ry * (size_t)rz;
    std::fill(out, out + cells, 0u);
    float inv[3], mn[3];
    for (int a = 0; a < 3; a++) {
        mn[a] = s.p.box_min[a];
        inv[a] = 1.0f / (s.p.box_max[a] - s.p.box_min[a]);
    }
    const int radius = 8;
    const float r = (float)radius;
    const float res[3] = {(float)rx, (float)ry, (float)rz};
#pragma omp parallel
    {
        std::vector<uint32_t> mine(cells, 0u);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < (int64_t)s.n; i++) {
            float uvw[3];
            bool nan = false, outside = false;
            for (int a = 0; a < 3; a++) {
                const float pa = (float)s.pos_a[4 * (size_t)i + a];
                uvw[a] = (pa - mn[a]) * inv[a];
                nan |= std::isnan(uvw[a]);
                outside |= uvw[a] < 0.0f || uvw[a] > 1.0f;
            }
            if (outside || nan) continue;
            const int32_t cx = (int32_t)(uvw[0] * res[0]), cy = (int32_t)(uvw[1] * res[1]),
                          cz = (int32_t)(uvw[2] * res[2]);
            for (int z = -radius; z <= radius; z++)
                for (int y = -radius; y <= radius; y++)
                    for (int x = -radius; x <= radius; x++) {
                        const int32_t vx = cx + x, vy = cy + y, vz = cz + z;
                        if (vx < 0 || vy < 0 || vz < 0 || vx >= rx || vy >= ry || vz >= rz) continue;
                        const float fx = (float)x, fy = (float)y, fz = (float)z;
                        const float dist = std::sqrt(fx * fx + fy * fy + fz * fz);
                        if (dist < r) {
                            const float xn = dist / r;
                            float w = 1.0f - xn * xn;
                            w = w * w * w;
                            const uint32_t add = (uint32_t)(w * 200.0f);
                            mine[((size_t)vz * ry + vy) * rx + vx] += add;
                        }
                    }
        }
#pragma omp critical
        for (size_t k = 0; k < cells; k++) out[k] += mine[k];
    }
}

// ---------------------------------------------------------------------------------------
// C interface.

// Warm-start features (no reference counterpart; the definition the HIP path's learned warm
// start uses): for particle i, over its accepted pairs (r^2 <= h^2) in the reference visit
// order, at most kFeatPairs of them (the pair count itself is not capped):
//   0 (rho - rho0) / rho0        1 s dt / rho0              2 alpha
//   3 pair count                 4 previous pressure        5 sum W
//   6 sum W p_prev_j             7 |sum r_ij| / h           8 sum |r_ij|^2 / h^2
//   9 sum (v_i - v_j).r_ij / h   10 sum |v_i - v_j|^2       11 -s rho^2 / (dt alpha)
// with v = velocity_a and p_prev the given previous pressures (hash order).
constexpr uint32_t kFeatPairs = 64;

template <class R>
void features(const State<R>& s, const float* pprev, float* out) {
    const R h = s.P(s.p.smoothing_radius);
    const R hh = h * h, rho0 = s.P(s.p.target_density), dt = s.P(s.p.dt);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s.n; ii++) {
        const uint32_t i = (uint32_t)ii;
        const V3<R> pi = s.v3(s.pos_b, i), vi = s.v3(s.vel_a, i);
        uint32_t count = 0;
        R sw = R(0), swp = R(0), cx = R(0), cy = R(0), cz = R(0), sr2 = R(0), svr = R(0), sv2 = R(0);
        for_each_candidate(s, i, [&](uint32_t j) {
            const V3<R> rv = sub(pi, s.v3(s.pos_b, j));
            const R r2 = dot(rv, rv);
            if (r2 > hh) return;
            if (++count > kFeatPairs) return;
            const R w = kernel_w(R(std::sqrt(r2)), h);
            const V3<R> d = sub(vi, s.v3(s.vel_a, j));
            sw += w;
            swp += w * R(pprev[j]);
            cx += rv.x;
            cy += rv.y;
            cz += rv.z;
            sr2 += r2;
            svr += dot(d, rv);
            sv2 += dot(d, d);
        });
        const R rho = s.dens[i], alpha = s.fac[i], si = s.src[i];
        float* f = out + (size_t)i * 12;
        f[0] = (float)((rho - rho0) / rho0);
        f[1] = (float)(si * dt / rho0);
        f[2] = (float)alpha;
        f[3] = (float)count;
        f[4] = pprev[i];
        f[5] = (float)sw;
        f[6] = (float)swp;
        f[7] = (float)(R(std::sqrt(cx * cx + cy * cy + cz * cz)) / h);
        f[8] = (float)(sr2 / hh);
        f[9] = (float)(svr / h);
        f[10] = (float)sv2;
        f[11] = (float)((alpha > Lit<R>::eps6 && rho > Lit<R>::eps6 && dt > Lit<R>::eps6)
                            ? -si * (rho * rho) / (dt * alpha)
                            : R(0));
    }
}

struct oracle_ctx {
    int fp64;
    State<float> f;
    State<double> d;
};

#define DISPATCH(c, call) ((c)->fp64 ? (call((c)->d)) : (call((c)->f)))

template <class R>
static std::vector<R>* fp_field(State<R>& s, int field, size_t* per) {
    *per = 4;
    switch (field) {
        case FE_FIELD_POSITION_A: return &s.pos_a;
        case FE_FIELD_POSITION_B: return &s.pos_b;
        case FE_FIELD_VELOCITY_A: return &s.vel_a;
        case FE_FIELD_VELOCITY_B: return &s.vel_b;
        case FE_FIELD_COLORS: return &s.colors;
        case FE_FIELD_PRESSURE_ACCELERATIONS: return &s.pacc;
        default: break;
    }
    *per = 1;
    switch (field) {
        case FE_FIELD_DENSITIES: return &s.dens;
        case FE_FIELD_FACTORS: return &s.fac;
        case FE_FIELD_SOURCE_TERMS: return &s.src;
        case FE_FIELD_PRESSURES: return &s.pres;
        default: return nullptr;
    }
}

template <class R>
static int upload_impl(State<R>& s, int field, const void* src, size_t bytes) {
    if (field == FE_FIELD_GRID_ENTRIES) {
        if (bytes != 8 * (size_t)s.T) return 1;
        std::memcpy(s.entries.data(), src, bytes);
        return 0;
    }
    if (field == FE_FIELD_GRID_START) {
        if (bytes != 4 * (size_t)s.T) return 1;
        std::memcpy(s.grid_start.data(), src, bytes);
        return 0;
    }
    size_t per;
    std::vector<R>* v = fp_field(s, field, &per);
    if (!v || bytes != 4 * per * (size_t)s.n) return 1;
    const float* f = (const float*)src;
    for (size_t k = 0; k < per * (size_t)s.n; k++) (*v)[k] = (R)f[k];
    return 0;
}

template <class R>
static int download_impl(State<R>& s, int field, void* dst, size_t bytes, int as_double) {
    if (field == FE_FIELD_GRID_ENTRIES) {
        if (bytes != 8 * (size_t)s.T) return 1;
        std::memcpy(dst, s.entries.data(), bytes);
        return 0;
    }
    if (field == FE_FIELD_GRID_START) {
        if (bytes != 4 * (size_t)s.T) return 1;
        std::memcpy(dst, s.grid_start.data(), bytes);
        return 0;
    }
    size_t per;
    std::vector<R>* v = fp_field(s, field, &per);
    if (!v) return 1;
    size_t cnt = per * (size_t)s.n;
    if (as_double) {
        if (bytes != 8 * cnt) return 1;
        double* o = (double*)dst;
        for (size_t k = 0; k < cnt; k++) o[k] = (double)(*v)[k];
    } else {
        if (bytes != 4 * cnt) return 1;
        float* o = (float*)dst;
        for (size_t k = 0; k < cnt; k++) o[k] = (float)(*v)[k];
    }
    return 0;
}

extern "C" {

oracle_ctx* oracle_create(const fe_params* params, const float* xyz, uint32_t n, int fp64) {
    oracle_ctx* c = new oracle_ctx();
    c->fp64 = fp64 ? 1 : 0;
    fe_options o;
    std::memset(&o, 0, sizeof(o));
    o.sort_algorithm = FE_SORT_RADIX;
    o.density_error_pct = 0.1f;
    o.divergence_error_pct = 0.5f;
    o.density_tol = 1e-3f;
    o.divergence_tol = 1e-2f;
    o.max_iterations = 100;
    auto init = [&](auto& s) {
        s.p = *params;
        s.o = o;
        s.alloc(n);
        for (uint32_t i = 0; i < n; i++) {
            for (int k = 0; k < 3; k++) s.pos_a[4 * (size_t)i + k] = xyz[3 * (size_t)i + k];
            s.pos_a[4 * (size_t)i + 3] = 1;
        }
    };
    if (c->fp64) init(c->d); else init(c->f);
    return c;
}

void oracle_destroy(oracle_ctx* c) { delete c; }
void oracle_set_params(oracle_ctx* c, const fe_params* p) { c->f.p = *p; c->d.p = *p; }
void oracle_get_params(oracle_ctx* c, fe_params* out) { *out = c->fp64 ? c->d.p : c->f.p; }
void oracle_set_options(oracle_ctx* c, const fe_options* o) { c->f.o = *o; c->d.o = *o; }
void oracle_set_threads(int t) {
#ifdef _OPENMP
    omp_set_num_threads(t > 0 ? t : 1);
#else
    (void)t;
#endif
}

int oracle_upload(oracle_ctx* c, int field, const void* src, size_t bytes) {
    return c->fp64 ? upload_impl(c->d, field, src, bytes) : upload_impl(c->f, field, src, bytes);
}
int oracle_download(oracle_ctx* c, int field, void* dst, size_t bytes, int as_double) {
    return c->fp64 ? download_impl(c->d, field, dst, bytes, as_double)
                   : download_impl(c->f, field, dst, bytes, as_double);
}

void oracle_neighbor_search(oracle_ctx* c) { DISPATCH(c, neighbor_search); }
void oracle_sort(oracle_ctx* c) { DISPATCH(c, sort_entries); }
void oracle_density_alpha(oracle_ctx* c) { DISPATCH(c, density_alpha); }
void oracle_viscosity(oracle_ctx* c) { DISPATCH(c, viscosity); }
void oracle_density_source(oracle_ctx* c) { DISPATCH(c, density_source); }
void oracle_pressure_force(oracle_ctx* c) { DISPATCH(c, pressure_force); }
uint64_t oracle_pressure_update(oracle_ctx* c) {
    return c->fp64 ? pressure_update(c->d) : pressure_update(c->f);
}
void oracle_pressure_integration(oracle_ctx* c) { DISPATCH(c, pressure_integration); }
void oracle_divergence_source(oracle_ctx* c) { DISPATCH(c, divergence_source); }
void oracle_divergence_integration(oracle_ctx* c) { DISPATCH(c, divergence_integration); }
void oracle_stats(oracle_ctx* c) { DISPATCH(c, stats); }
void oracle_features(oracle_ctx* c, const float* pprev, float* out) {
    if (c->fp64) features(c->d, pprev, out); else features(c->f, pprev, out);
}
void oracle_read_stats(oracle_ctx* c, uint32_t out[3]) {
    std::memcpy(out, c->fp64 ? c->d.stats : c->f.stats, 12);
}

uint32_t oracle_solve_density(oracle_ctx* c) {
    if (c->fp64) return c->d.last_iters[0] = solve(c->d, c->d.p.density_solver_iterations, c->d.o.density_tol);
    return c->f.last_iters[0] = solve(c->f, c->f.p.density_solver_iterations, c->f.o.density_tol);
}
uint32_t oracle_solve_divergence(oracle_ctx* c) {
    if (c->fp64) return c->d.last_iters[1] = solve(c->d, c->d.p.divergence_solver_iterations, c->d.o.divergence_tol);
    return c->f.last_iters[1] = solve(c->f, c->f.p.divergence_solver_iterations, c->f.o.divergence_tol);
}
void oracle_step(oracle_ctx* c, float max_dt) {
    if (c->fp64) step(c->d, max_dt); else step(c->f, max_dt);
}
size_t oracle_wall_psi(oracle_ctx* c, float* out, size_t cap) {
    if (c->fp64) {
        const WallGeom* w = walls(c->d);
        if (!w) return 0;
        for (size_t k = 0; k < c->d.wall_psi.size() && k < cap; k++) out[k] = (float)c->d.wall_psi[k];
        return c->d.wall_psi.size();
    }
    const WallGeom* w = walls(c->f);
    if (!w) return 0;
    for (size_t k = 0; k < c->f.wall_psi.size() && k < cap; k++) out[k] = c->f.wall_psi[k];
    return c->f.wall_psi.size();
}
int oracle_splat_density(oracle_ctx* c, uint32_t* out, size_t cells) {
    const fe_params& p = c->fp64 ? c->d.p : c->f.p;
    for (int a = 0; a < 3; a++)
        if (p.grid_res[a] < 1) return 1;
    if (cells != (size_t)p.grid_res[0] * (size_t)p.grid_res[1] * (size_t)p.grid_res[2]) return 1;
    if (c->fp64) splat_density(c->d, out); else splat_density(c->f, out);
    return 0;
}
void oracle_wall_update(fe_wall* w, float dt) { wall_update(w, dt); }
void oracle_set_wall(oracle_ctx* c, const fe_wall* w) {
    c->f.wall_on = c->d.wall_on = w != nullptr;
    c->f.wall = c->d.wall = w ? *w : fe_wall{};
}
int oracle_get_wall(oracle_ctx* c, fe_wall* out) {
    const auto& s = c->fp64 ? c->d.wall : c->f.wall;
    *out = s;
    return (c->fp64 ? c->d.wall_on : c->f.wall_on) ? 1 : 0;
}
void oracle_last_iterations(oracle_ctx* c, uint32_t* density, uint32_t* divergence) {
    const uint32_t* it = c->fp64 ? c->d.last_iters : c->f.last_iters;
    *density = it[0];
    *divergence = it[1];
}

uint32_t oracle_cell_hash(int32_t cx, int32_t cy, int32_t cz, uint32_t T) { return cell_hash(cx, cy, cz, T); }
float oracle_kernel_w(float r, float h) { return kernel_w<float>(r, h); }
double oracle_kernel_w64(double r, double h) { return kernel_w<double>(r, h); }
void oracle_kernel_grad(const float rv[3], float r, float h, float out[3]) {
    V3<float> g = kernel_grad<float>(V3<float>{rv[0], rv[1], rv[2]}, r, h);
    out[0] = g.x; out[1] = g.y; out[2] = g.z;
}
void oracle_radix_sort_pairs(uint32_t* pairs, uint32_t n) { radix_sort_literal(pairs, n); }
void oracle_bitonic_sort_pairs(uint32_t* pairs, uint32_t n) { bitonic_sort_literal(pairs, n); }
uint64_t oracle_residual_q(float s_minus_ap, float dt, float rho0) { return residual_q(s_minus_ap, dt, rho0); }

}  // extern "C"
