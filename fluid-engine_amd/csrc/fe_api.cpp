// fe_api.cpp — the C ABI: context lifetime, field transfer, one enqueue per reference
// ComputeStep, the Jacobi solves and the Renderer::step frame sequencer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fe_abi.h"

static_assert(sizeof(fe_params) == 112, "fe_params must match SimulationParams (112 bytes)");
static_assert(offsetof(fe_params, dt) == 24, "dt offset");
static_assert(offsetof(fe_params, density_solver_iterations) == 28, "iterations offset");
static_assert(offsetof(fe_params, gravity) == 48, "gravity offset");
static_assert(offsetof(fe_params, box_min) == 64, "box_min offset");
static_assert(offsetof(fe_params, box_max) == 80, "box_max offset");
static_assert(offsetof(fe_params, grid_res) == 96, "grid_res offset");

namespace {
thread_local std::string g_last_error;
}  // namespace

fe_status fe::fail(fe_status code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

namespace {

using fe::DeviceGuard;
using fe::fail;
using fe::hip_fail;

bool field_info(const fe_ctx* c, int field, size_t* bytes, void** dev) {
    const size_t n = c->n;
    switch (field) {
        case FE_FIELD_POSITION_A: *bytes = 16 * n; *dev = c->pos_a; return true;
        case FE_FIELD_POSITION_B: *bytes = 16 * n; *dev = c->pos_b; return true;
        case FE_FIELD_VELOCITY_A: *bytes = 16 * n; *dev = c->vel_a; return true;
        case FE_FIELD_VELOCITY_B: *bytes = 16 * n; *dev = c->vel_b; return true;
        case FE_FIELD_COLORS: *bytes = 16 * n; *dev = c->colors; return true;
        case FE_FIELD_PRESSURE_ACCELERATIONS: *bytes = 16 * n; *dev = c->pacc; return true;
        case FE_FIELD_DENSITIES: *bytes = 4 * n; *dev = c->dens; return true;
        case FE_FIELD_FACTORS: *bytes = 4 * n; *dev = c->fac; return true;
        case FE_FIELD_SOURCE_TERMS: *bytes = 4 * n; *dev = c->src; return true;
        case FE_FIELD_PRESSURES: *bytes = 4 * n; *dev = c->pres; return true;
        case FE_FIELD_GRID_ENTRIES: *bytes = 8 * (size_t)c->T; *dev = nullptr; return true;
        case FE_FIELD_GRID_START: *bytes = 4 * (size_t)c->T; *dev = nullptr; return true;
        default: return false;
    }
}

void free_ctx(fe_ctx* c) {
    if (!c) return;
    for (void* p : {(void*)c->pos_a, (void*)c->pos_b, (void*)c->vel_a, (void*)c->vel_b, (void*)c->colors,
                    (void*)c->pacc, (void*)c->dens, (void*)c->fac, (void*)c->src, (void*)c->pres,
                    (void*)c->keys[0], (void*)c->keys[1], (void*)c->vals[0], (void*)c->vals[1],
                    (void*)c->range, (void*)c->stats, (void*)c->resid, (void*)c->staging,
                    (void*)c->partial, (void*)c->stats_partial, (void*)c->nbr, (void*)c->nbr_count,
                    (void*)c->sched[0], (void*)c->sched[1], (void*)c->mkeys[0], (void*)c->mkeys[1],
                    (void*)c->inv, (void*)c->pos_m, (void*)c->vb_m, (void*)c->va_m, (void*)c->dens_m,
                    (void*)c->prho_m, (void*)c->pacc_m})
        if (p) (void)hipFree(p);
    fe::sort_scratch_free(&c->sort);
    fe::learn_free(c);
    if (c->host_stats) (void)hipHostFree(c->host_stats);
    if (c->host_resid) (void)hipHostFree(c->host_resid);
    for (auto& r : c->pending) {
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    for (hipEvent_t e : {c->stats_event, c->timer_start, c->timer_stop})
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

void default_options(fe_options* o) {
    std::memset(o, 0, sizeof(*o));
    o->sort_algorithm = FE_SORT_RADIX;
    o->use_cfl = 0;
    o->use_solver_error_threshold = 0;
    o->density_error_pct = 0.1f;
    o->divergence_error_pct = 0.5f;
    o->solver_mode = FE_SOLVER_FIXED;
    o->density_tol = 1e-3f;
    o->divergence_tol = 1e-2f;
    o->max_iterations = 100;
    o->warm_start = FE_WARM_START_REFERENCE;
}

// Bits per axis of the Morton schedule key: enough for the box's cell extent (+ margin).
uint32_t morton_bits_for(const fe_params& p) {
    uint32_t bits = 1;
    for (int a = 0; a < 3; a++) {
        const float cells = (p.box_max[a] - p.box_min[a]) / p.smoothing_radius + 3.0f;
        uint32_t b = 1;
        while (b < 10 && (float)(1u << b) < cells) b++;
        bits = b > bits ? b : bits;
    }
    return 3 * bits;
}

fe_status ensure_ends(fe_ctx* c) {
    if (!c->ends_dirty) return FE_OK;
    FE_HIP(fe::launch_rebuild_ends(c), "rebuild_ends");
    c->ends_dirty = false;
    return FE_OK;
}

fe_status ensure_staging(fe_ctx* c) {
    if (c->staging) return FE_OK;
    FE_HIP(hipMalloc(&c->staging, 8 * (size_t)c->T), "hipMalloc staging");
    return FE_OK;
}

// Frame control of Renderer::step (src/renderer/mod.rs:231-267).
void frame_control(const uint32_t stats[3], uint32_t n, fe_params* p, const fe_options* o, uint32_t* dit,
                   uint32_t* vit) {
    const float nf = (float)n;
    float max_speed;
    std::memcpy(&max_speed, &stats[0], 4);
    const float avg_de = (float)stats[1] / (1.0f * nf);
    const float avg_ve = (float)stats[2] / (10.0f * nf);
    if (o->use_cfl && max_speed > 0.01f) {
        const float h = p->smoothing_radius;
        float cfl = 0.4f * h / max_speed;
        if (cfl < 0.001f) cfl = 0.001f;
        if (cfl > 0.05f) cfl = 0.05f;
        const float grown = p->dt * 1.1f;
        p->dt = std::isnan(grown) ? cfl : std::min(cfl, grown);
    }
    const float rho0 = p->target_density;
    const float dthr = o->density_error_pct / 100.0f * rho0;
    const float vthr = o->divergence_error_pct / 100.0f * rho0;
    *dit = (o->use_solver_error_threshold && avg_de < dthr) ? 2u : p->density_solver_iterations;
    *vit = (o->use_solver_error_threshold && avg_ve < vthr) ? 1u : p->divergence_solver_iterations;
}

fe_status jacobi(fe_ctx* c, int kind, uint32_t fixed_iters, float tol, uint32_t* iterations);

// Jacobi loop (src/renderer/mod.rs:301-306 / :327-332) between the warm-start hooks.
// kind 0 = density, 1 = divergence.
fe_status solve(fe_ctx* c, int kind, uint32_t fixed_iters, float tol, uint32_t* iterations) {
    FE_CHECK_CTX(c);
    fe_status s = ensure_ends(c);
    if (s == FE_OK) s = fe::learn_before_solve(c, kind);
    if (s == FE_OK) s = jacobi(c, kind, fixed_iters, tol, iterations);
    if (s == FE_OK) s = fe::learn_after_solve(c, kind);
    return s;
}

fe_status jacobi(fe_ctx* c, int kind, uint32_t fixed_iters, float tol, uint32_t* iterations) {
    if (c->options.solver_mode != FE_SOLVER_CONVERGE) {
        for (uint32_t k = 0; k < fixed_iters; k++) {
            FE_HIP(fe::launch_pressure_force(c, nullptr, 0), "pressure_force");
            FE_HIP(fe::launch_pressure_update(c, nullptr, nullptr, 0), "pressure_update");
        }
        c->last_iters[kind] = fixed_iters;
        c->total_iters[kind] += fixed_iters;
        if (iterations) *iterations = fixed_iters;
        return FE_OK;
    }
    // FE_SOLVER_CONVERGE: sweep k writes its residual to slot[k]; sweep k+1 skips itself on
    // the device once slot[k] <= thr. The host launches chunks of sweeps ahead and stops
    // launching once a finished chunk shows convergence.
    unsigned long long* slot = c->resid + (size_t)kind * fe::kMaxSolveIters;
    unsigned long long* hslot = c->host_resid + (size_t)kind * fe::kMaxSolveIters;
    uint32_t maxit = c->options.max_iterations;
    if (maxit < 1) maxit = 1;
    if (maxit > fe::kMaxSolveIters) maxit = fe::kMaxSolveIters;
    const unsigned long long thr =
        (unsigned long long)((double)tol * (double)c->n * 4294967296.0);
    FE_HIP(fe::launch_fill_u64(c, slot, maxit), "fill residuals");
    // Chunk sizes: the previous frame's sweep count first (the usual answer), then 2, 4, 8,
    // 16, ... Chunk j's residuals are read back while chunk j+1 runs.
    struct Chunk {
        hipEvent_t ev;
        uint32_t beg, end;
    };
    std::vector<Chunk> chunks;
    uint32_t first = c->last_iters[kind] ? c->last_iters[kind] : 4;
    if (first > maxit) first = maxit;
    uint32_t chunk_beg = 0, chunk_end = first, grow = 2;
    uint32_t launched = 0, found = 0;
    size_t checked_chunks = 0;
    bool done = false;
    fe_status st = FE_OK;
    for (uint32_t k = 0; k < maxit && !done; k++) {
        const unsigned long long* prev = k == 0 ? nullptr : slot + (k - 1);
        hipError_t e = fe::launch_pressure_force(c, prev, thr);
        if (e == hipSuccess) e = fe::launch_pressure_update(c, slot + k, prev, thr);
        if (e != hipSuccess) { st = hip_fail(e, "jacobi sweep"); break; }
        launched = k + 1;
        if (launched < chunk_end && launched < maxit) continue;
        hipEvent_t ev;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess)
            e = hipMemcpyAsync(hslot + chunk_beg, slot + chunk_beg,
                               sizeof(unsigned long long) * (launched - chunk_beg), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipEventRecord(ev, c->stream);
        if (e != hipSuccess) { st = hip_fail(e, "residual readback"); break; }
        chunks.push_back(Chunk{ev, chunk_beg, launched});
        chunk_beg = launched;
        chunk_end = launched + grow;
        grow = grow < 16 ? grow * 2 : 16;
        // look at the chunk before this one (finished or nearly so)
        if (chunks.size() >= 2) {
            const Chunk& ch = chunks[chunks.size() - 2];
            e = hipEventSynchronize(ch.ev);
            if (e != hipSuccess) { st = hip_fail(e, "residual wait"); break; }
            for (uint32_t q = ch.beg; q < ch.end; q++)
                if (hslot[q] <= thr) { found = q + 1; done = true; break; }
            checked_chunks = chunks.size() - 1;
        }
    }
    if (st == FE_OK && !chunks.empty()) {
        hipError_t e = hipEventSynchronize(chunks.back().ev);
        if (e != hipSuccess) st = hip_fail(e, "residual wait");
        for (size_t j = checked_chunks; st == FE_OK && !found && j < chunks.size(); j++)
            for (uint32_t q = chunks[j].beg; q < chunks[j].end; q++)
                if (hslot[q] <= thr) { found = q + 1; break; }
    }
    for (const Chunk& ch : chunks) (void)hipEventDestroy(ch.ev);
    if (st != FE_OK) return st;
    const uint32_t its = found ? found : launched;
    c->last_iters[kind] = its;
    c->total_iters[kind] += its;
    if (iterations) *iterations = its;
    return FE_OK;
}

fe_status neighbor_search(fe_ctx* c) {
    FE_HIP(fe::launch_hash(c), "spatial_hash");
    if (c->options.sort_algorithm == FE_SORT_BITONIC) {
        FE_HIP(fe::bitonic_sort(c), "bitonic sort");
    } else {
        // Real hashes are < T and the T - n sentinels stay in place under a stable sort,
        // so sorting the first n entries on log2(T) bits gives the reference permutation.
        FE_HIP(fe::radix_sort(c, c->n, c->log2T), "radix sort");
    }
    FE_HIP(fe::launch_offsets_reorder(c), "offsets/reorder");
    c->ends_dirty = false;
    c->list_valid = false;
    c->ws.feat_fresh = false;
    // gather schedule: particles in Morton order of their cells
    FE_HIP(fe::launch_morton_keys(c), "morton keys");
    FE_HIP(fe::radix_sort_pairs(c, c->mkeys, c->sched, c->n, c->morton_bits, "schedule"), "schedule sort");
    FE_HIP(fe::launch_schedule_mirrors(c), "schedule mirrors");
    c->ns_fresh = true;
    c->da_fresh = false;
    c->ns_sort = c->options.sort_algorithm;
    c->ns_h = c->params.smoothing_radius;
    return FE_OK;
}

fe_status wait_prev_stats(fe_ctx* c) {
    if (c->stats_pending) {
        FE_HIP(hipEventSynchronize(c->stats_event), "stats wait");
        std::memcpy(c->prev_stats, c->host_stats, sizeof(c->prev_stats));
        c->stats_pending = false;
    }
    return FE_OK;
}

fe_status step(fe_ctx* c, float max_dt) {
    uint32_t dit = c->params.density_solver_iterations, vit = c->params.divergence_solver_iterations;
    if (c->options.use_cfl || c->options.use_solver_error_threshold) {
        fe_status s = wait_prev_stats(c);
        if (s != FE_OK) return s;
        frame_control(c->prev_stats, c->n, &c->params, &c->options, &dit, &vit);
    }
    fe_status s;
    // Frame-start neighbour search + density_alpha (src/renderer/mod.rs:279-286). When the
    // particles have not moved since the last search (one substep per frame: the divergence
    // phase leaves position_a alone), hash, sort, offsets, position_b, the densities and the
    // factors would come out bit-identical, so only velocity_b is gathered again.
    if (c->ns_fresh && c->da_fresh && c->ns_sort == c->options.sort_algorithm &&
        c->ns_h == c->params.smoothing_radius) {
        FE_HIP(fe::launch_reorder_velocity(c), "reorder velocity");
    } else {
        if ((s = neighbor_search(c)) != FE_OK) return s;
        FE_HIP(fe::launch_density_alpha(c), "density_alpha");
    }
    float t = 0.0f;
    while (t < max_dt) {
        FE_HIP(fe::launch_viscosity(c), "viscosity");
        FE_HIP(fe::launch_density_source(c), "density_source");
        if ((s = solve(c, 0, dit, c->options.density_tol, nullptr)) != FE_OK) return s;
        FE_HIP(fe::launch_pressure_integration(c), "pressure_integration");
        t += c->params.dt;
        if ((s = neighbor_search(c)) != FE_OK) return s;
        FE_HIP(fe::launch_density_alpha(c), "density_alpha");
        FE_HIP(fe::launch_divergence_source(c), "divergence_source");
        if ((s = solve(c, 1, vit, c->options.divergence_tol, nullptr)) != FE_OK) return s;
        FE_HIP(fe::launch_divergence_integration(c), "divergence_integration");
        if (!(c->params.dt > 0.0f)) break;   // a non-positive dt would never end the frame
    }
    FE_HIP(fe::launch_stats(c), "stats");
    FE_HIP(hipMemcpyAsync(c->host_stats, c->stats, 12, hipMemcpyDeviceToHost, c->stream), "stats copy");
    FE_HIP(hipEventRecord(c->stats_event, c->stream), "stats event");
    c->stats_pending = true;
    c->frames++;
    return FE_OK;
}

}  // namespace

namespace fe {

Launch::Launch(fe_ctx* ctx, const char* name) : c(ctx), on(ctx->profiling) {
    if (!on) return;
    rec.name = name;
    if (hipEventCreate(&rec.start) != hipSuccess || hipEventCreate(&rec.stop) != hipSuccess) {
        on = false;
        return;
    }
    (void)hipEventRecord(rec.start, c->stream);
}

Launch::~Launch() {
    if (!on) return;
    (void)hipEventRecord(rec.stop, c->stream);
    c->pending.push_back(rec);
}

}  // namespace fe

static fe_status resolve_profile(fe_ctx* c) {
    for (auto& r : c->pending) {
        FE_HIP(hipEventSynchronize(r.stop), "profile wait");
        float ms = 0.0f;
        FE_HIP(hipEventElapsedTime(&ms, r.start, r.stop), "profile elapsed");
        auto it = std::find_if(c->totals.begin(), c->totals.end(),
                               [&](const fe::ProfileTotal& t) { return t.name == r.name; });
        if (it == c->totals.end()) {
            c->totals.push_back(fe::ProfileTotal{r.name, 0, 0.0});
            it = c->totals.end() - 1;
        }
        it->launches++;
        it->total_ms += ms;
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    c->pending.clear();
    return FE_OK;
}

extern "C" {

const char* fe_version(void) { return "fluid_engine-mi355x 0.1 (gfx950)"; }
const char* fe_last_error(void) { return g_last_error.c_str(); }

fe_status fe_default_options(fe_options* out) {
    if (!out) return fail(FE_ERR_INVALID_ARGUMENT, "null options");
    default_options(out);
    return FE_OK;
}

fe_status fe_create(const fe_params* params, const float* xyz, uint32_t n, int device, fe_ctx** out) {
    if (!params || !out) return fail(FE_ERR_INVALID_ARGUMENT, "null params or out pointer");
    *out = nullptr;
    if (n == 0 || n > (1u << 30)) return fail(FE_ERR_INVALID_ARGUMENT, "particle count must be in [1, 2^30]");
    if (!xyz) return fail(FE_ERR_INVALID_ARGUMENT, "null positions");
    if (!(params->smoothing_radius > 0.0f)) return fail(FE_ERR_INVALID_ARGUMENT, "smoothing_radius must be > 0");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail(FE_ERR_DEVICE, "no HIP device available");
    if (device < 0) {
        if (hipGetDevice(&device) != hipSuccess) device = 0;
    }
    if (device >= ndev) return fail(FE_ERR_INVALID_ARGUMENT, "device index out of range");
    DeviceGuard guard(device);

    fe_ctx* c = new fe_ctx();
    c->device = device;
    c->params = *params;
    c->morton_bits = morton_bits_for(*params);
    default_options(&c->options);
    c->n = n;
    c->T = 1;
    c->log2T = 0;
    while (c->T < n) {
        c->T <<= 1;
        c->log2T++;
    }
#define ALLOC(ptr, bytes)                                        \
    do {                                                         \
        hipError_t ee = hipMalloc((void**)&(ptr), (bytes));      \
        if (ee != hipSuccess) {                                  \
            free_ctx(c);                                         \
            return hip_fail(ee, "hipMalloc " #ptr);              \
        }                                                        \
    } while (0)
    const size_t v4 = 16 * (size_t)n, s1 = 4 * (size_t)n, tb = 4 * (size_t)c->T;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { free_ctx(c); return hip_fail(e, "hipStreamCreate"); }
    ALLOC(c->pos_a, v4); ALLOC(c->pos_b, v4); ALLOC(c->vel_a, v4); ALLOC(c->vel_b, v4);
    ALLOC(c->colors, v4); ALLOC(c->pacc, v4);
    ALLOC(c->dens, s1); ALLOC(c->fac, s1); ALLOC(c->src, s1); ALLOC(c->pres, s1);
    ALLOC(c->keys[0], tb); ALLOC(c->keys[1], tb); ALLOC(c->vals[0], tb); ALLOC(c->vals[1], tb);
    ALLOC(c->range, 2 * tb);
    ALLOC(c->stats, 16);
    ALLOC(c->resid, sizeof(unsigned long long) * 2 * fe::kMaxSolveIters);
    ALLOC(c->partial, sizeof(unsigned long long) * fe::gather_partials(n));
    ALLOC(c->stats_partial, sizeof(uint32_t) * 3 * fe::kStatsBlocks);
    ALLOC(c->nbr, sizeof(uint32_t) * fe::kMaxNbr * (((size_t)n + 15) / 16 * 16));   // whole 16-item tiles
    ALLOC(c->nbr_count, sizeof(uint32_t) * (size_t)n);
    for (int k = 0; k < 2; k++) {
        ALLOC(c->sched[k], sizeof(uint32_t) * (size_t)n);
        ALLOC(c->mkeys[k], sizeof(uint32_t) * (size_t)n);
    }
    ALLOC(c->inv, s1); ALLOC(c->pos_m, v4); ALLOC(c->vb_m, v4); ALLOC(c->va_m, v4);
    ALLOC(c->dens_m, s1); ALLOC(c->prho_m, s1); ALLOC(c->pacc_m, v4);
#undef ALLOC
    e = fe::sort_scratch_alloc(&c->sort, c->T);
    if (e != hipSuccess) { free_ctx(c); return hip_fail(e, "sort scratch"); }
    e = hipHostMalloc((void**)&c->host_stats, 16, hipHostMallocDefault);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&c->host_resid, sizeof(unsigned long long) * 2 * fe::kMaxSolveIters,
                          hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->stats_event, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&c->timer_start);
    if (e == hipSuccess) e = hipEventCreate(&c->timer_stop);
    if (e != hipSuccess) { free_ctx(c); return hip_fail(e, "host buffers / events"); }

    // GpuPhysicsData::new: position_a = (xyz, 1); velocities, pressures and stats zero.
    std::vector<float> p4(4 * (size_t)n);
    for (size_t i = 0; i < n; i++) {
        p4[4 * i] = xyz[3 * i];
        p4[4 * i + 1] = xyz[3 * i + 1];
        p4[4 * i + 2] = xyz[3 * i + 2];
        p4[4 * i + 3] = 1.0f;
    }
    e = hipMemcpy(c->pos_a, p4.data(), v4, hipMemcpyHostToDevice);
    for (void* z : {(void*)c->pos_b, (void*)c->vel_a, (void*)c->vel_b, (void*)c->colors, (void*)c->pacc})
        if (e == hipSuccess) e = hipMemset(z, 0, v4);
    for (void* z : {(void*)c->dens, (void*)c->fac, (void*)c->src, (void*)c->pres})
        if (e == hipSuccess) e = hipMemset(z, 0, s1);
    if (e == hipSuccess) e = hipMemset(c->stats, 0, 16);
    if (e == hipSuccess) e = hipMemset(c->range, 0xFF, 2 * tb);
    if (e == hipSuccess) e = hipMemset(c->keys[0], 0xFF, tb);
    if (e == hipSuccess) e = hipMemset(c->vals[0], 0xFF, tb);
    if (e == hipSuccess) e = fe::launch_iota(c, c->sched[0], n);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) { free_ctx(c); return hip_fail(e, "initial upload"); }
    std::memset(c->host_stats, 0, 16);
    *out = c;
    return FE_OK;
}

fe_status fe_destroy(fe_ctx* c) {
    if (!c) return FE_OK;
    DeviceGuard g(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_ctx(c);
    return FE_OK;
}

fe_status fe_set_params(fe_ctx* c, const fe_params* p) {
    FE_CHECK_CTX(c);
    if (!p) return fail(FE_ERR_INVALID_ARGUMENT, "null params");
    if (!(p->smoothing_radius > 0.0f)) return fail(FE_ERR_INVALID_ARGUMENT, "smoothing_radius must be > 0");
    c->params = *p;
    c->morton_bits = morton_bits_for(*p);
    return FE_OK;
}

fe_status fe_get_params(const fe_ctx* c, fe_params* out) {
    FE_CHECK_CTX(c);
    if (!out) return fail(FE_ERR_INVALID_ARGUMENT, "null out");
    *out = c->params;
    return FE_OK;
}

fe_status fe_set_options(fe_ctx* c, const fe_options* o) {
    FE_CHECK_CTX(c);
    if (!o) return fail(FE_ERR_INVALID_ARGUMENT, "null options");
    if (o->sort_algorithm != FE_SORT_BITONIC && o->sort_algorithm != FE_SORT_RADIX)
        return fail(FE_ERR_INVALID_ARGUMENT, "unknown sort_algorithm");
    if (o->solver_mode != FE_SOLVER_FIXED && o->solver_mode != FE_SOLVER_CONVERGE)
        return fail(FE_ERR_INVALID_ARGUMENT, "unknown solver_mode");
    if (o->warm_start != FE_WARM_START_REFERENCE && o->warm_start != FE_WARM_START_LEARNED)
        return fail(FE_ERR_INVALID_ARGUMENT, "unknown warm_start");
    if (o->warm_start == FE_WARM_START_LEARNED && !fe::learn_ready(c, 0) && !fe::learn_ready(c, 1))
        return fail(FE_ERR_UNSUPPORTED,
                    "warm_start LEARNED needs a network: fe_learn_train or fe_learn_set_weights first");
    if (o->solver_mode == FE_SOLVER_CONVERGE && (o->max_iterations < 1 || o->max_iterations > fe::kMaxSolveIters))
        return fail(FE_ERR_INVALID_ARGUMENT, "max_iterations must be in [1, 1024]");
    c->options = *o;
    return FE_OK;
}

fe_status fe_get_options(const fe_ctx* c, fe_options* out) {
    FE_CHECK_CTX(c);
    if (!out) return fail(FE_ERR_INVALID_ARGUMENT, "null out");
    *out = c->options;
    return FE_OK;
}

fe_status fe_num_particles(const fe_ctx* c, uint32_t* n, uint32_t* table_size) {
    FE_CHECK_CTX(c);
    if (n) *n = c->n;
    if (table_size) *table_size = c->T;
    return FE_OK;
}

fe_status fe_upload(fe_ctx* c, int field, const void* src, size_t bytes) {
    FE_CHECK_CTX(c);
    size_t want;
    void* dev;
    if (!field_info(c, field, &want, &dev)) return fail(FE_ERR_INVALID_ARGUMENT, "unknown field");
    if (bytes != want)
        return fail(FE_ERR_SIZE_MISMATCH, "field size mismatch: expected " + std::to_string(want) + " bytes");
    if (!src) return fail(FE_ERR_INVALID_ARGUMENT, "null source");
    DeviceGuard g(c->device);
    if (field == FE_FIELD_POSITION_B || field == FE_FIELD_GRID_ENTRIES || field == FE_FIELD_GRID_START)
        c->list_valid = false;
    c->ns_fresh = false;
    c->da_fresh = false;
    fe::learn_invalidate(c);
    // schedule-ordered mirrors of the uploaded field are stale now
    switch (field) {
        case FE_FIELD_POSITION_B: c->mirrors &= ~fe::kMirrorPos; break;
        case FE_FIELD_VELOCITY_B: c->mirrors &= ~fe::kMirrorVelB; break;
        case FE_FIELD_VELOCITY_A: c->mirrors &= ~fe::kMirrorVelA; break;
        case FE_FIELD_DENSITIES: c->mirrors &= ~(fe::kMirrorDens | fe::kMirrorPrho); break;
        case FE_FIELD_PRESSURES: c->mirrors &= ~fe::kMirrorPrho; break;
        case FE_FIELD_PRESSURE_ACCELERATIONS: c->mirrors &= ~fe::kMirrorPacc; break;
        default: break;
    }
    if (dev) {
        FE_HIP(hipMemcpyAsync(dev, src, bytes, hipMemcpyHostToDevice, c->stream), "upload");
        FE_HIP(hipStreamSynchronize(c->stream), "upload wait");
        return FE_OK;
    }
    fe_status s = ensure_staging(c);
    if (s != FE_OK) return s;
    FE_HIP(hipMemcpyAsync(c->staging, src, bytes, hipMemcpyHostToDevice, c->stream), "upload");
    if (field == FE_FIELD_GRID_ENTRIES)
        FE_HIP(fe::launch_unpack_entries(c, c->staging), "unpack entries");
    else
        FE_HIP(fe::launch_unpack_start(c, c->staging), "unpack grid_start");
    FE_HIP(hipStreamSynchronize(c->stream), "upload wait");
    c->ends_dirty = true;
    return FE_OK;
}

fe_status fe_download(fe_ctx* c, int field, void* dst, size_t bytes) {
    FE_CHECK_CTX(c);
    size_t want;
    void* dev;
    if (!field_info(c, field, &want, &dev)) return fail(FE_ERR_INVALID_ARGUMENT, "unknown field");
    if (bytes != want)
        return fail(FE_ERR_SIZE_MISMATCH, "field size mismatch: expected " + std::to_string(want) + " bytes");
    if (!dst) return fail(FE_ERR_INVALID_ARGUMENT, "null destination");
    DeviceGuard g(c->device);
    if (dev) {
        FE_HIP(hipMemcpyAsync(dst, dev, bytes, hipMemcpyDeviceToHost, c->stream), "download");
        FE_HIP(hipStreamSynchronize(c->stream), "download wait");
        return FE_OK;
    }
    fe_status s = ensure_staging(c);
    if (s != FE_OK) return s;
    if (field == FE_FIELD_GRID_ENTRIES)
        FE_HIP(fe::launch_pack_entries(c, c->staging), "pack entries");
    else
        FE_HIP(fe::launch_pack_start(c, c->staging), "pack grid_start");
    FE_HIP(hipMemcpyAsync(dst, c->staging, bytes, hipMemcpyDeviceToHost, c->stream), "download");
    FE_HIP(hipStreamSynchronize(c->stream), "download wait");
    return FE_OK;
}

fe_status fe_device_ptr(fe_ctx* c, int field, void** out) {
    FE_CHECK_CTX(c);
    if (!out) return fail(FE_ERR_INVALID_ARGUMENT, "null out");
    size_t want;
    void* dev;
    if (!field_info(c, field, &want, &dev)) return fail(FE_ERR_INVALID_ARGUMENT, "unknown field");
    if (!dev) return fail(FE_ERR_UNSUPPORTED, "grid fields use an internal layout");
    *out = dev;
    return FE_OK;
}

#define FE_STAGE(fn_name, body)                  \
    fe_status fn_name(fe_ctx* c) {               \
        FE_CHECK_CTX(c);                         \
        DeviceGuard g(c->device);                \
        body;                                    \
        return FE_OK;                            \
    }

FE_STAGE(fe_neighbor_search, {
    fe_status s = neighbor_search(c);
    if (s != FE_OK) return s;
})
FE_STAGE(fe_sort, {
    if (c->options.sort_algorithm == FE_SORT_BITONIC)
        FE_HIP(fe::bitonic_sort(c), "bitonic sort");
    else
        FE_HIP(fe::radix_sort(c, c->T, 32), "radix sort");
    c->ends_dirty = true;
    c->ns_fresh = false;
    c->da_fresh = false;
    c->list_valid = false;
    fe::learn_invalidate(c);
})
FE_STAGE(fe_density_alpha, {
    fe_status s = ensure_ends(c);
    if (s != FE_OK) return s;
    FE_HIP(fe::launch_density_alpha(c), "density_alpha");
})
FE_STAGE(fe_viscosity, {
    fe_status s = ensure_ends(c);
    if (s != FE_OK) return s;
    FE_HIP(fe::launch_viscosity(c), "viscosity");
})
FE_STAGE(fe_density_source, {
    fe_status s = ensure_ends(c);
    if (s != FE_OK) return s;
    FE_HIP(fe::launch_density_source(c), "density_source");
})
FE_STAGE(fe_pressure_force, {
    fe_status s = ensure_ends(c);
    if (s != FE_OK) return s;
    FE_HIP(fe::launch_pressure_force(c, nullptr, 0), "pressure_force");
})
FE_STAGE(fe_pressure_update, {
    fe_status s = ensure_ends(c);
    if (s != FE_OK) return s;
    FE_HIP(fe::launch_pressure_update(c, nullptr, nullptr, 0), "pressure_update");
})
FE_STAGE(fe_pressure_integration, { FE_HIP(fe::launch_pressure_integration(c), "pressure_integration"); })
FE_STAGE(fe_divergence_source, {
    fe_status s = ensure_ends(c);
    if (s != FE_OK) return s;
    FE_HIP(fe::launch_divergence_source(c), "divergence_source");
})
FE_STAGE(fe_divergence_integration, { FE_HIP(fe::launch_divergence_integration(c), "divergence_integration"); })
FE_STAGE(fe_stats, { FE_HIP(fe::launch_stats(c), "stats"); })

fe_status fe_solve_density(fe_ctx* c, uint32_t* iterations) {
    FE_CHECK_CTX(c);
    DeviceGuard g(c->device);
    return solve(c, 0, c->params.density_solver_iterations, c->options.density_tol, iterations);
}

fe_status fe_solve_divergence(fe_ctx* c, uint32_t* iterations) {
    FE_CHECK_CTX(c);
    DeviceGuard g(c->device);
    return solve(c, 1, c->params.divergence_solver_iterations, c->options.divergence_tol, iterations);
}

fe_status fe_step(fe_ctx* c, float max_dt) {
    FE_CHECK_CTX(c);
    DeviceGuard g(c->device);
    return step(c, max_dt);
}

fe_status fe_run(fe_ctx* c, float max_dt, uint32_t count) {
    FE_CHECK_CTX(c);
    DeviceGuard g(c->device);
    for (uint32_t k = 0; k < count; k++) {
        fe_status s = step(c, max_dt);
        if (s != FE_OK) return s;
    }
    return FE_OK;
}

fe_status fe_read_stats(fe_ctx* c, uint32_t out[3]) {
    FE_CHECK_CTX(c);
    if (!out) return fail(FE_ERR_INVALID_ARGUMENT, "null out");
    DeviceGuard g(c->device);
    FE_HIP(hipMemcpyAsync(c->host_stats, c->stats, 12, hipMemcpyDeviceToHost, c->stream), "stats copy");
    FE_HIP(hipStreamSynchronize(c->stream), "stats wait");
    std::memcpy(out, c->host_stats, 12);
    return FE_OK;
}

fe_status fe_last_iterations(fe_ctx* c, uint32_t* density, uint32_t* divergence) {
    FE_CHECK_CTX(c);
    if (density) *density = c->last_iters[0];
    if (divergence) *divergence = c->last_iters[1];
    return FE_OK;
}

fe_status fe_iteration_totals(fe_ctx* c, uint64_t* density, uint64_t* divergence, uint64_t* frames) {
    FE_CHECK_CTX(c);
    if (density) *density = c->total_iters[0];
    if (divergence) *divergence = c->total_iters[1];
    if (frames) *frames = c->frames;
    return FE_OK;
}

fe_status fe_reset_iteration_totals(fe_ctx* c) {
    FE_CHECK_CTX(c);
    c->total_iters[0] = c->total_iters[1] = 0;
    c->frames = 0;
    return FE_OK;
}

fe_status fe_sync(fe_ctx* c) {
    FE_CHECK_CTX(c);
    DeviceGuard g(c->device);
    FE_HIP(hipStreamSynchronize(c->stream), "sync");
    return FE_OK;
}

fe_status fe_timer_start(fe_ctx* c) {
    FE_CHECK_CTX(c);
    DeviceGuard g(c->device);
    FE_HIP(hipEventRecord(c->timer_start, c->stream), "timer start");
    return FE_OK;
}

fe_status fe_timer_stop(fe_ctx* c, float* ms) {
    FE_CHECK_CTX(c);
    if (!ms) return fail(FE_ERR_INVALID_ARGUMENT, "null out");
    DeviceGuard g(c->device);
    FE_HIP(hipEventRecord(c->timer_stop, c->stream), "timer stop");
    FE_HIP(hipEventSynchronize(c->timer_stop), "timer wait");
    FE_HIP(hipEventElapsedTime(ms, c->timer_start, c->timer_stop), "timer elapsed");
    return FE_OK;
}

fe_status fe_profile_enable(fe_ctx* c, int enable) {
    FE_CHECK_CTX(c);
    c->profiling = enable != 0;
    return FE_OK;
}

fe_status fe_profile_count(fe_ctx* c, uint32_t* count) {
    FE_CHECK_CTX(c);
    if (!count) return fail(FE_ERR_INVALID_ARGUMENT, "null out");
    DeviceGuard g(c->device);
    fe_status s = resolve_profile(c);
    if (s != FE_OK) return s;
    *count = (uint32_t)c->totals.size();
    return FE_OK;
}

fe_status fe_profile_entry(fe_ctx* c, uint32_t i, char* name, size_t name_cap, uint64_t* launches,
                           double* total_ms) {
    FE_CHECK_CTX(c);
    DeviceGuard g(c->device);
    fe_status s = resolve_profile(c);
    if (s != FE_OK) return s;
    if (i >= c->totals.size()) return fail(FE_ERR_INVALID_ARGUMENT, "profile index out of range");
    const fe::ProfileTotal& t = c->totals[i];
    if (name && name_cap) {
        std::strncpy(name, t.name.c_str(), name_cap - 1);
        name[name_cap - 1] = '\0';
    }
    if (launches) *launches = t.launches;
    if (total_ms) *total_ms = t.total_ms;
    return FE_OK;
}

fe_status fe_profile_reset(fe_ctx* c) {
    FE_CHECK_CTX(c);
    DeviceGuard g(c->device);
    fe_status s = resolve_profile(c);
    if (s != FE_OK) return s;
    c->totals.clear();
    return FE_OK;
}

fe_status fe_frame_control(const uint32_t stats[3], uint32_t n, const fe_params* params,
                           const fe_options* options, float* dt_out, uint32_t* density_iters,
                           uint32_t* divergence_iters) {
    if (!stats || !params || !options || n == 0) return fail(FE_ERR_INVALID_ARGUMENT, "null argument or n == 0");
    fe_params p = *params;
    uint32_t dit, vit;
    frame_control(stats, n, &p, options, &dit, &vit);
    if (dt_out) *dt_out = p.dt;
    if (density_iters) *density_iters = dit;
    if (divergence_iters) *divergence_iters = vit;
    return FE_OK;
}

}  // extern "C"
