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
#include "fe_device.h"

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
                    (void*)c->stats_partial, (void*)c->nbr, (void*)c->nbr_count,
                    (void*)c->sched[0], (void*)c->sched[1], (void*)c->mkeys[0], (void*)c->mkeys[1],
                    (void*)c->inv, (void*)c->pos_m, (void*)c->vb_m, (void*)c->va_m, (void*)c->dens_m,
                    (void*)c->pa_m, (void*)c->pa2_m, (void*)c->fac_m, (void*)c->src_m, (void*)c->pres_m,
                    (void*)c->win_meta, (void*)c->win_round, (void*)c->win_delta, (void*)c->nbr16,
                    (void*)c->report_err, (void*)c->report})
        if (p) (void)hipFree(p);
    fe::sort_scratch_free(&c->sort);
    fe::learn_free(c);
    fe::splat_free(c);
    if (c->walls_psi) (void)hipFree(c->walls_psi);
    if (c->wall_g) (void)hipFree(c->wall_g);
    for (void* p : {(void*)c->scene_tab, (void*)c->sid, (void*)c->scene_stats})
        if (p) (void)hipFree(p);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->host_stats) (void)hipHostFree(c->host_stats);
    if (c->host_resid) (void)hipHostFree(c->host_resid);
    for (auto& r : c->pending) {
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    for (hipEvent_t e : {c->stats_event, c->timer_start, c->timer_stop})
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->tab_event)
        if (e) (void)hipEventDestroy(e);
    if (c->tab_ring) (void)hipHostFree(c->tab_ring);
    for (hipEvent_t e : c->box_event)
        if (e) (void)hipEventDestroy(e);
    if (c->box_ring) (void)hipHostFree(c->box_ring);
    if (c->box_dev) (void)hipFree(c->box_dev);
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

// The scene table of a batch: per-scene constants and converge thresholds (tol x N_s x 2^32).
fe_status fill_scene_tab(fe_ctx* c, fe::Sph* tab) {
    for (uint32_t k = 0; k < c->nscenes; k++) {
        uint32_t fast = 0;
        FE_HIP(fe::exact_math_flags(c, c->scene_params[k], &fast), "exact arithmetic check");
        tab[k] = fe::sph_from(c->scene_params[k], c->scene_n[k], c->scene_T[k], c->scene_tab0[k], c->scene_part0[k],
                              fast);
        tab[k].thr[0] = (unsigned long long)((double)c->options.density_tol * (double)c->scene_n[k] * 4294967296.0);
        tab[k].thr[1] = (unsigned long long)((double)c->options.divergence_tol * (double)c->scene_n[k] * 4294967296.0);
    }
    return FE_OK;
}

fe_status upload_scene_tab(fe_ctx* c) {
    if (c->nscenes <= 1) return FE_OK;
    std::vector<fe::Sph> tab(c->nscenes);
    fe_status s = fill_scene_tab(c, tab.data());
    if (s != FE_OK) return s;
    FE_HIP(hipStreamSynchronize(c->stream), "scene table wait");
    FE_HIP(hipMemcpy(c->scene_tab, tab.data(), sizeof(fe::Sph) * tab.size(), hipMemcpyHostToDevice), "scene table");
    return FE_OK;
}

// The same copy ordered on the stream without waiting for it: the table is written into one
// of kTabRing pinned host slots, whose previous copy (kTabRing frames ago) has long finished.
fe_status enqueue_scene_tab(fe_ctx* c) {
    const size_t per = sizeof(fe::Sph) * c->nscenes;
    if (!c->tab_ring) {
        FE_HIP(hipHostMalloc((void**)&c->tab_ring, per * fe::kTabRing, hipHostMallocDefault), "hipHostMalloc tables");
        for (hipEvent_t& e : c->tab_event) FE_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming), "table event");
    }
    const uint32_t slot = c->tab_slot;
    c->tab_slot = (slot + 1) % fe::kTabRing;
    FE_HIP(hipEventSynchronize(c->tab_event[slot]), "table slot wait");
    fe::Sph* tab = c->tab_ring + (size_t)slot * c->nscenes;
    fe_status s = fill_scene_tab(c, tab);
    if (s != FE_OK) return s;
    FE_HIP(hipMemcpyAsync(c->scene_tab, tab, per, hipMemcpyHostToDevice, c->stream), "scene table");
    FE_HIP(hipEventRecord(c->tab_event[slot], c->stream), "table event");
    return FE_OK;
}

// One scene in graph mode: the frame's box copied to the device on the stream, from one of
// kTabRing pinned host slots (the slot's previous copy, kTabRing frames ago, has finished).
fe_status enqueue_live_box(fe_ctx* c) {
    if (!c->box_ring) {
        FE_HIP(hipMalloc((void**)&c->box_dev, 8 * sizeof(float)), "hipMalloc box");
        FE_HIP(hipHostMalloc((void**)&c->box_ring, 8 * sizeof(float) * fe::kTabRing, hipHostMallocDefault),
               "hipHostMalloc box");
        for (hipEvent_t& e : c->box_event) FE_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming), "box event");
    }
    const uint32_t slot = c->box_slot;
    c->box_slot = (slot + 1) % fe::kTabRing;
    FE_HIP(hipEventSynchronize(c->box_event[slot]), "box slot wait");
    float* b = c->box_ring + 8 * (size_t)slot;
    for (int a = 0; a < 3; a++) {
        b[a] = c->params.box_min[a];
        b[3 + a] = c->params.box_max[a];
    }
    b[6] = b[7] = 0.0f;
    FE_HIP(hipMemcpyAsync(c->box_dev, b, 8 * sizeof(float), hipMemcpyHostToDevice, c->stream), "box copy");
    FE_HIP(hipEventRecord(c->box_event[slot], c->stream), "box event");
    return FE_OK;
}

// CollisionBox::update (src/entities/collision.rs:31-35). Each operation rounds to float in
// the reference's order (no contraction: the library is built with -ffp-contract=off), and
// the sine is the C library's sinf, which Rust's f32::sin calls. The constant is
// std::f32::consts::TAU rounded to float.
void wall_update(fe_wall* w, float dt) {
    constexpr float kTau = 6.28318530717958647692f;
    w->time = w->time + dt;
    const float phase = w->time * w->wave_frequency * kTau;
    const float offset = w->wave_amplitude * std::sin(phase);
    w->min[0] = w->base_min_x + offset;
}

// The UI pass of the reference frame (src/renderer/ui.rs:125-126): the box enters the
// parameter block with w = 0.
void wall_to_params(const fe_wall& w, fe_params* p) {
    for (int a = 0; a < 3; a++) {
        p->box_min[a] = w.min[a];
        p->box_max[a] = w.max[a];
    }
    p->box_min[3] = 0.0f;
    p->box_max[3] = 0.0f;
}

// Before a frame: every attached wall's box goes into its scene's parameters.
fe_status apply_walls(fe_ctx* c) {
    if (!c->walls_on) return FE_OK;
    if (c->nscenes == 1) {
        wall_to_params(c->walls[0], &c->params);
        c->morton_bits = morton_bits_for(c->params);
        return FE_OK;
    }
    for (uint32_t k = 0; k < c->nscenes; k++)
        if (c->wall_on[k]) wall_to_params(c->walls[k], &c->scene_params[k]);
    return enqueue_scene_tab(c);
}

// After a frame (src/renderer/mod.rs:270): every attached wall advances by max_dt.
void advance_walls(fe_ctx* c, float max_dt) {
    for (uint32_t k = 0; k < c->walls.size(); k++)
        if (c->wall_on[k]) wall_update(&c->walls[k], max_dt);
}

bool wall_finite(const fe_wall& w) {
    const float* f = &w.min[0];
    for (size_t k = 0; k < sizeof(fe_wall) / sizeof(float); k++)
        if (!std::isfinite(f[k])) return false;
    return true;
}

fe_status set_wall(fe_ctx* c, uint32_t scene, const fe_wall* w) {
    if (w && !wall_finite(*w)) return fail(FE_ERR_INVALID_ARGUMENT, "wall fields must be finite");
    if (c->walls.size() < c->nscenes) {
        c->walls.resize(c->nscenes, fe_wall{});
        c->wall_on.resize(c->nscenes, 0);
    }
    const bool was = c->wall_on[scene] != 0;
    c->wall_on[scene] = w ? 1 : 0;
    c->walls[scene] = w ? *w : fe_wall{};
    c->walls_on = c->walls_on - (was ? 1 : 0) + (w ? 1 : 0);
    return FE_OK;
}

fe_status get_wall(const fe_ctx* c, uint32_t scene, fe_wall* out, int* attached) {
    if (!out) return fail(FE_ERR_INVALID_ARGUMENT, "null out");
    const bool on = scene < c->wall_on.size() && c->wall_on[scene];
    *out = on ? c->walls[scene] : fe_wall{};
    if (attached) *attached = on ? 1 : 0;
    return FE_OK;
}

}  // namespace

// A captured frame holds the kernel arguments of its capture; anything that changes them
// drops it.
void fe::drop_graph(fe_ctx* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
}

namespace {
using fe::drop_graph;

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
// kind 0 = density, 1 = divergence. The sweeps keep the new pressures and pressure
// accelerations in their schedule-ordered mirrors; the hash-ordered fields are written once
// when the loop ends.
fe_status solve(fe_ctx* c, int kind, uint32_t fixed_iters, float tol, uint32_t* iterations) {
    FE_CHECK_CTX(c);
    fe_status s = ensure_ends(c);
    if (s == FE_OK) s = fe::learn_before_solve(c, kind);
    if (s == FE_OK) {
        c->defer_sweep_writes = true;
        s = jacobi(c, kind, fixed_iters, tol, iterations);
        if (s == FE_OK && c->report_on) {
            hipError_t e = fe::launch_residual_report(c, kind);
            if (e != hipSuccess) s = hip_fail(e, "residual report");
            c->report_solves[kind]++;
        }
        c->defer_sweep_writes = false;
        hipError_t e = fe::flush_sweep_writes(c);
        if (s == FE_OK && e != hipSuccess) s = hip_fail(e, "sweep write-back");
    }
    if (s == FE_OK) s = fe::learn_after_solve(c, kind);
    return s;
}

fe_status jacobi(fe_ctx* c, int kind, uint32_t fixed_iters, float tol, uint32_t* iterations) {
    if (c->options.solver_mode != FE_SOLVER_CONVERGE) {
        for (uint32_t k = 0; k < fixed_iters; k++) {
            FE_HIP(fe::launch_pressure_force(c, nullptr, 0, 0), "pressure_force");
            FE_HIP(fe::launch_pressure_update(c, nullptr, nullptr, 0, 0), "pressure_update");
        }
        c->last_iters[kind] = fixed_iters;
        c->total_iters[kind] += fixed_iters;
        for (uint32_t sc = 0; c->nscenes > 1 && sc < c->nscenes; sc++) {
            c->scene_iters[kind][sc] = fixed_iters;
            c->scene_total_iters[kind][sc] += fixed_iters;
        }
        if (iterations) *iterations = fixed_iters;
        return FE_OK;
    }
    // FE_SOLVER_CONVERGE: sweep k writes its residual to slot[k]; sweep k+1 skips itself on
    // the device once slot[k] <= thr. The host launches chunks of sweeps ahead and stops
    // launching once a finished chunk shows convergence. A batch keeps one row of slots per
    // scene (kMaxSolveIters apart) and a threshold per scene; each scene's particles skip on
    // their own, and the host stops once every scene has converged.
    const uint32_t S = c->nscenes;
    unsigned long long* slot = c->resid + (size_t)kind * S * fe::kMaxSolveIters;
    unsigned long long* hslot = c->host_resid + (size_t)kind * S * fe::kMaxSolveIters;
    uint32_t maxit = c->options.max_iterations;
    if (maxit < 1) maxit = 1;
    if (maxit > fe::kMaxSolveIters) maxit = fe::kMaxSolveIters;
    const unsigned long long thr =
        (unsigned long long)((double)tol * (double)c->n * 4294967296.0);
    std::vector<unsigned long long> thr_s(S, thr);
    for (uint32_t k = 0; S > 1 && k < S; k++)
        thr_s[k] = (unsigned long long)((double)tol * (double)c->scene_n[k] * 4294967296.0);
    FE_HIP(fe::launch_fill_u64(c, slot, S > 1 ? S * fe::kMaxSolveIters : maxit), "fill residuals");
    // Chunk sizes: the previous frame's sweep count first (the usual answer), then 1, 2, 4,
    // 8, 16, ... Chunk j's residuals are read back while chunk j+1 runs, so a solve that ends
    // with chunk j also launches chunk j+1, whose sweeps return at once (about 4.5 us a
    // kernel at 1M): a one-sweep second chunk keeps that to two kernels.
    struct Chunk {
        hipEvent_t ev;
        uint32_t beg, end;
    };
    std::vector<Chunk> chunks;
    std::vector<uint32_t> found(S, 0);
    uint32_t pending = S;
    auto scan = [&](const Chunk& ch) {
        for (uint32_t sc = 0; sc < S; sc++) {
            if (found[sc]) continue;
            for (uint32_t q = ch.beg; q < ch.end; q++)
                if (hslot[(size_t)sc * fe::kMaxSolveIters + q] <= thr_s[sc]) {
                    found[sc] = q + 1;
                    pending--;
                    break;
                }
        }
    };
    uint32_t first = c->last_iters[kind] ? c->last_iters[kind] : 4;
    if (first > maxit) first = maxit;
    uint32_t chunk_beg = 0, chunk_end = first, grow = 1;
    uint32_t launched = 0;
    size_t checked_chunks = 0;
    fe_status st = FE_OK;
    for (uint32_t k = 0; k < maxit && pending; k++) {
        const unsigned long long* prev = k == 0 ? nullptr : slot + (k - 1);
        hipError_t e = fe::launch_pressure_force(c, prev, thr, (uint32_t)kind);
        if (e == hipSuccess) e = fe::launch_pressure_update(c, slot + k, prev, thr, (uint32_t)kind);
        if (e != hipSuccess) { st = hip_fail(e, "jacobi sweep"); break; }
        launched = k + 1;
        if (launched < chunk_end && launched < maxit) continue;
        hipEvent_t ev;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) {
            const size_t w = sizeof(unsigned long long) * (launched - chunk_beg);
            const size_t pitch = sizeof(unsigned long long) * fe::kMaxSolveIters;
            e = S == 1 ? hipMemcpyAsync(hslot + chunk_beg, slot + chunk_beg, w, hipMemcpyDeviceToHost, c->stream)
                       : hipMemcpy2DAsync(hslot + chunk_beg, pitch, slot + chunk_beg, pitch, w, S,
                                          hipMemcpyDeviceToHost, c->stream);
        }
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
            scan(ch);
            checked_chunks = chunks.size() - 1;
        }
    }
    // Once a scanned chunk showed convergence the sweeps still queued skip themselves on the
    // device, so the host returns without waiting for them and the next stages queue behind.
    if (st == FE_OK && !chunks.empty() && pending) {
        hipError_t e = hipEventSynchronize(chunks.back().ev);
        if (e != hipSuccess) st = hip_fail(e, "residual wait");
        for (size_t j = checked_chunks; st == FE_OK && pending && j < chunks.size(); j++) scan(chunks[j]);
    }
    for (const Chunk& ch : chunks) (void)hipEventDestroy(ch.ev);
    if (st != FE_OK) return st;
    uint32_t its = 0;
    for (uint32_t sc = 0; sc < S; sc++) {
        const uint32_t it = found[sc] ? found[sc] : launched;
        its = it > its ? it : its;
        if (S > 1) {
            c->scene_iters[kind][sc] = it;
            c->scene_total_iters[kind][sc] += it;
        }
    }
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
        // the state is in the previous search's hash order: the new hashes arrive almost sorted
        FE_HIP(fe::radix_sort(c, c->n, c->log2T, true), "radix sort");
    }
    FE_HIP(fe::launch_offsets_reorder(c), "offsets/reorder");
    c->ends_dirty = false;
    c->list_valid = false;
    c->ws.feat_fresh = false;
    // gather schedule: particles in Morton order of their cells (keys from offsets/reorder).
    // A small single scene keeps hash order (offsets/reorder wrote the identity): its fields
    // sit in the L2 caches whatever the order, and a cell's particles are consecutive in hash
    // order as well. (With the learned warm start the Morton order stays, so that its samples
    // are drawn as before.)
    if (!(c->nscenes == 1 && c->n <= fe::kHashScheduleMax && !c->ws.alloc))
        FE_HIP(fe::radix_sort_pairs(c, c->mkeys, c->sched, c->n, c->morton_bits + c->scene_bits, false),
               "schedule sort");
    FE_HIP(fe::launch_schedule_mirrors(c), "schedule mirrors");
    c->ns_fresh = true;
    c->perm_valid = true;
    c->wall_g_fresh = false;
    c->da_fresh = false;
    c->ns_sort = c->options.sort_algorithm;
    c->ns_h = c->params.smoothing_radius;
    return FE_OK;
}

// The frame's stats (and, with the watchdog, its health words) read back behind the frame.
fe_status enqueue_stats_copy(fe_ctx* c) {
    // words 0-2 the stats, 4-5 the watchdog's, 6 the radix sort's error word
    FE_HIP(hipMemcpyAsync(c->host_stats, c->stats, 28, hipMemcpyDeviceToHost, c->stream), "stats copy");
    FE_HIP(hipEventRecord(c->stats_event, c->stream), "stats event");
    c->stats_pending = true;
    c->stats_frame = c->frames;
    c->stats_health = c->watchdog;
    return FE_OK;
}

fe_status wait_prev_stats(fe_ctx* c) {
    if (c->stats_pending) {
        FE_HIP(hipEventSynchronize(c->stats_event), "stats wait");
        std::memcpy(c->prev_stats, c->host_stats, sizeof(c->prev_stats));
        c->stats_pending = false;
        if (c->host_stats[6]) c->sort_failed = true;
        if (c->stats_health) {
            fe_health& h = c->health;
            h.frame = c->stats_frame;
            h.nonfinite = c->host_stats[4];
            h.first_nonfinite = c->host_stats[5];
            std::memcpy(&h.max_speed, &c->host_stats[0], sizeof(float));
            h.speed_limit = c->speed_limit;
            h.flags = (h.nonfinite ? FE_HEALTH_NONFINITE : 0u) |
                      (c->speed_limit > 0.0f && !(h.max_speed <= c->speed_limit) ? FE_HEALTH_SPEED : 0u);
            h.sort_error = c->host_stats[6];
        }
    }
    return FE_OK;
}

// stats_copy: read the frame's stats back to the host (a captured frame leaves that copy to
// its replays).
fe_status step(fe_ctx* c, float max_dt, bool stats_copy = true) {
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
    // (With wall boundary particles the densities also depend on the box, which a moving wall
    // changes between frames.)
    if (c->ns_fresh && c->da_fresh && c->ns_sort == c->options.sort_algorithm &&
        c->ns_h == c->params.smoothing_radius && fe::walls_current(c)) {
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
        FE_HIP(fe::launch_density_alpha_divergence_source(c), "density_alpha + divergence_source");
        if ((s = solve(c, 1, vit, c->options.divergence_tol, nullptr)) != FE_OK) return s;
        FE_HIP(fe::launch_divergence_integration(c), "divergence_integration");
        if (!(c->params.dt > 0.0f)) break;   // a non-positive dt would never end the frame
    }
    FE_HIP(fe::launch_stats(c), "stats");
    c->frames++;
    if (c->nscenes == 1 && stats_copy) {
        fe_status s2 = enqueue_stats_copy(c);
        if (s2 != FE_OK) return s2;
    }
    return FE_OK;
}

// A frame in FIXED mode from a captured graph: the steady-state frame (frame-start search
// reused) is captured once and replayed; the first frame after creation, an upload or a
// parameter change runs eagerly. The capture holds kernels only; a single scene's stats
// read-back is enqueued after each replay.
fe_status step_graph(fe_ctx* c, float max_dt) {
    if (!(c->ns_fresh && c->da_fresh)) return step(c, max_dt);
    if (c->graph_exec && c->graph_morton_bits != c->morton_bits) drop_graph(c);   // the box grew
    if (!c->graph_exec) {
        c->graph_morton_bits = c->morton_bits;
        FE_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "graph capture");
        fe_status st = step(c, max_dt, false);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (st != FE_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return hip_fail(e, "graph capture");
        c->graph = g;
        e = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            drop_graph(c);
            return hip_fail(e, "graph instantiate");
        }
        // the capture recorded the frame; it has not run yet
        c->frames--;
        for (int k = 0; k < 2; k++) c->total_iters[k] -= c->last_iters[k];
        for (int k = 0; k < 2; k++)
            for (uint32_t sc = 0; c->nscenes > 1 && sc < c->nscenes; sc++)
                c->scene_total_iters[k][sc] -= c->scene_iters[k][sc];
    }
    FE_HIP(hipGraphLaunch(c->graph_exec, c->stream), "graph launch");
    c->frames++;
    if (c->nscenes == 1) {
        fe_status s2 = enqueue_stats_copy(c);
        if (s2 != FE_OK) return s2;
    }
    for (int k = 0; k < 2; k++) c->total_iters[k] += c->last_iters[k];
    for (int k = 0; k < 2; k++)
        for (uint32_t sc = 0; c->nscenes > 1 && sc < c->nscenes; sc++) c->scene_total_iters[k][sc] += c->scene_iters[k][sc];
    return FE_OK;
}

// Before a frame (and outside any capture): a finished frame's stats are taken without
// waiting; a radix sort whose look-back stopped waiting left a wrong permutation, so the
// context fails from then on; with the watchdog's halt, an unhealthy previous frame stops this
// one.
fe_status frame_checks(fe_ctx* c) {
    if (c->stats_pending) {
        const hipError_t q = hipEventQuery(c->stats_event);
        if (q == hipSuccess) {
            fe_status s0 = wait_prev_stats(c);
            if (s0 != FE_OK) return s0;
        } else {
            (void)hipGetLastError();   // "not ready" must not reach the launch checks that follow
            if (q != hipErrorNotReady) return hip_fail(q, "stats query");
        }
    }
    if (c->sort_failed)
        return fail(FE_ERR_DEVICE, "a radix sort's look-back stopped waiting for an earlier tile; the neighbour "
                                   "search is not trustworthy (recreate the context)");
    if (c->watchdog && c->watchdog_halt) {
        fe_status s0 = wait_prev_stats(c);
        if (s0 != FE_OK) return s0;
        const fe_health& h = c->health;
        if (h.flags & FE_HEALTH_NONFINITE)
            return fail(FE_ERR_UNSTABLE, "watchdog: frame " + std::to_string(h.frame) + " left " +
                                             std::to_string(h.nonfinite) + " particles with a non-finite field (first: " +
                                             std::to_string(h.first_nonfinite) + ")");
        if (h.flags & FE_HEALTH_SPEED)
            return fail(FE_ERR_UNSTABLE, "watchdog: frame " + std::to_string(h.frame) + " max speed " +
                                             std::to_string(h.max_speed) + " exceeds the limit " +
                                             std::to_string(h.speed_limit));
    }
    return FE_OK;
}

// Whether a single scene's frame can be replayed from a graph: fixed sweeps and nothing that
// makes the host decide inside the frame (frame control, learning, exit reports) or changes
// kernel arguments from frame to frame other than the box (wall samples follow the box).
bool single_graph_ok(const fe_ctx* c) {
    const fe_options& o = c->options;
    return !o.use_cfl && !o.use_solver_error_threshold && o.warm_start == FE_WARM_START_REFERENCE &&
           o.boundary == FE_BOUNDARY_CLAMP && !c->report_on && !fe::learn_features_wanted(c, 0) &&
           !fe::learn_features_wanted(c, 1);
}

fe_status run_frame(fe_ctx* c, float max_dt) {
    fe_status s = frame_checks(c);
    if (s != FE_OK) return s;
    if ((s = apply_walls(c)) != FE_OK) return s;
    const bool graph = c->batch_graph && c->options.solver_mode == FE_SOLVER_FIXED && !c->profiling &&
                       (c->nscenes > 1 || single_graph_ok(c));
    if (graph && c->nscenes == 1) {
        // every frame of a graph-mode scene takes its box from the device copy, eager or replayed
        if ((s = enqueue_live_box(c)) != FE_OK) return s;
        c->live_box = true;
    }
    s = graph ? step_graph(c, max_dt) : step(c, max_dt);
    c->live_box = false;
    if (s == FE_OK) advance_walls(c, max_dt);
    return s;
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

// Device buffers of a context whose n, T, log2T and nscenes are set; position_a from xyz
// (n x 3), everything else zero (GpuPhysicsData::new, particle.rs:164-299).
static fe_status alloc_ctx(fe_ctx* c, const float* xyz) {
    const uint32_t n = c->n;
#define ALLOC(ptr, bytes)                                        \
    do {                                                         \
        hipError_t ee = hipMalloc((void**)&(ptr), (bytes));      \
        if (ee != hipSuccess) return hip_fail(ee, "hipMalloc " #ptr); \
    } while (0)
    const size_t v4 = 16 * (size_t)n, s1 = 4 * (size_t)n, tb = 4 * (size_t)c->T;
    const size_t slots = (size_t)2 * c->nscenes * fe::kMaxSolveIters;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    ALLOC(c->pos_a, v4); ALLOC(c->pos_b, v4); ALLOC(c->vel_a, v4); ALLOC(c->vel_b, v4);
    ALLOC(c->colors, v4); ALLOC(c->pacc, v4);
    ALLOC(c->dens, s1); ALLOC(c->fac, s1); ALLOC(c->src, s1); ALLOC(c->pres, s1);
    ALLOC(c->keys[0], tb); ALLOC(c->keys[1], tb); ALLOC(c->vals[0], tb); ALLOC(c->vals[1], tb);
    ALLOC(c->range, 2 * tb);
    ALLOC(c->stats, 32);
    ALLOC(c->resid, sizeof(unsigned long long) * slots);
    ALLOC(c->stats_partial, sizeof(uint32_t) * 5 * fe::kStatsBlocks);
    ALLOC(c->nbr, sizeof(uint32_t) * fe::kMaxNbr * (((size_t)n + 15) / 16 * 16));   // whole 16-item tiles
    ALLOC(c->nbr_count, sizeof(uint32_t) * (size_t)n);
    for (int k = 0; k < 2; k++) {
        ALLOC(c->sched[k], sizeof(uint32_t) * (size_t)n);
        ALLOC(c->mkeys[k], sizeof(uint32_t) * (size_t)n);
    }
    ALLOC(c->inv, s1); ALLOC(c->pos_m, v4); ALLOC(c->vb_m, v4); ALLOC(c->va_m, v4);
    ALLOC(c->dens_m, s1); ALLOC(c->pa_m, v4); ALLOC(c->pa2_m, 8 * (size_t)n);
    ALLOC(c->fac_m, s1); ALLOC(c->src_m, s1); ALLOC(c->pres_m, s1);
    const size_t nwin = ((size_t)n + fe::kWin - 1) / fe::kWin;
    ALLOC(c->win_meta, sizeof(uint32_t) * nwin);
    ALLOC(c->win_round, sizeof(uint4) * fe::kWinRounds * nwin);
    ALLOC(c->win_delta, sizeof(uint32_t) * fe::kWinMaxRuns * nwin);
    ALLOC(c->nbr16, sizeof(uint16_t) * fe::kMaxNbr * fe::kWin * nwin);   // whole windows
#undef ALLOC
    e = fe::sort_scratch_alloc(&c->sort, c->T, c->stream);
    if (e != hipSuccess) return hip_fail(e, "sort scratch");
    e = hipHostMalloc((void**)&c->host_stats, 32, hipHostMallocDefault);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&c->host_resid, sizeof(unsigned long long) * slots, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->stats_event, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&c->timer_start);
    if (e == hipSuccess) e = hipEventCreate(&c->timer_stop);
    if (e != hipSuccess) return hip_fail(e, "host buffers / events");

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
    if (e == hipSuccess) e = hipMemset(c->stats, 0, 32);
    if (e == hipSuccess) e = hipMemset(c->range, 0xFF, 2 * tb);
    if (e == hipSuccess) e = hipMemset(c->keys[0], 0xFF, tb);
    if (e == hipSuccess) e = hipMemset(c->vals[0], 0xFF, tb);
    if (e == hipSuccess) e = fe::launch_iota(c, c->sched[0], n);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(e, "initial upload");
    std::memset(c->host_stats, 0, 32);
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
    fe_status st = alloc_ctx(c, xyz);
    if (st == FE_OK) {
        hipError_t ee = fe::exact_math_flags(c, c->params, &c->exact_fast);
        if (ee != hipSuccess) st = hip_fail(ee, "exact arithmetic check");
    }
    if (st != FE_OK) {
        free_ctx(c);
        return st;
    }
    std::memset(c->host_stats, 0, 32);
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
    if (c->nscenes > 1) return fail(FE_ERR_UNSUPPORTED, "a batch takes per-scene parameters (fe_batch_set_params)");
    DeviceGuard g(c->device);
    drop_graph(c);
    uint32_t fast = 0;
    FE_HIP(fe::exact_math_flags(c, *p, &fast), "exact arithmetic check");
    // the densities of the last density_alpha used the old mass
    if (std::memcmp(&p->particle_mass, &c->params.particle_mass, sizeof(float)) != 0) c->da_fresh = false;
    // the list holds the pairs within the old h
    if (std::memcmp(&p->smoothing_radius, &c->params.smoothing_radius, sizeof(float)) != 0) c->list_valid = false;
    c->params = *p;
    c->exact_fast = fast;
    c->morton_bits = morton_bits_for(*p);
    c->mirrors &= ~fe::kMirrorMrho;   // holds the old mass
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
    if (o->divergence_order != FE_DIVERGENCE_REFERENCE && o->divergence_order != FE_DIVERGENCE_CONSISTENT)
        return fail(FE_ERR_INVALID_ARGUMENT, "unknown divergence_order");
    if (o->boundary != FE_BOUNDARY_CLAMP && o->boundary != FE_BOUNDARY_PARTICLES)
        return fail(FE_ERR_INVALID_ARGUMENT, "unknown boundary");
    if (o->boundary == FE_BOUNDARY_PARTICLES) {
        if (c->nscenes > 1) return fail(FE_ERR_UNSUPPORTED, "boundary particles take a single-scene context");
        const fe_params& p = c->params;
        const float sb = 2.0f * p.particle_radius;
        if (!(sb > 0.0f) || !std::isfinite(sb) || !(p.smoothing_radius / sb <= 16.0f))
            return fail(FE_ERR_INVALID_ARGUMENT, "boundary particles need 0 < smoothing_radius <= 32 particle_radius");
        double samples = 0.0;
        for (int a = 0; a < 3; a++) {
            const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
            const double eu = (p.box_max[u] - p.box_min[u]) / sb, ev = (p.box_max[v] - p.box_min[v]) / sb;
            if (!(eu >= 1.0 && ev >= 1.0) || !std::isfinite(eu * ev))
                return fail(FE_ERR_INVALID_ARGUMENT, "boundary particles need a box at least 2 r wide on each axis");
            samples += 2.0 * eu * ev;
        }
        if (samples > (double)(1u << 26)) return fail(FE_ERR_INVALID_ARGUMENT, "boundary particles: too many wall samples");
    }
    if (o->warm_start == FE_WARM_START_LEARNED && !fe::learn_ready(c, 0) && !fe::learn_ready(c, 1))
        return fail(FE_ERR_UNSUPPORTED,
                    "warm_start LEARNED needs a network: fe_learn_train or fe_learn_set_weights first");
    if (o->solver_mode == FE_SOLVER_CONVERGE && (o->max_iterations < 1 || o->max_iterations > fe::kMaxSolveIters))
        return fail(FE_ERR_INVALID_ARGUMENT, "max_iterations must be in [1, 1024]");
    if (o->residual > FE_RESIDUAL_TWO_SIDED) return fail(FE_ERR_INVALID_ARGUMENT, "unknown residual");
    if (c->nscenes > 1 && (o->sort_algorithm != FE_SORT_RADIX || o->use_cfl || o->use_solver_error_threshold ||
                           o->warm_start != FE_WARM_START_REFERENCE))
        return fail(FE_ERR_UNSUPPORTED,
                    "a batch runs the radix sort, fixed dt, no early-exit flags and the reference warm start");
    drop_graph(c);
    // the last densities were formed with (or without) the wall samples
    if (o->boundary != c->options.boundary) c->da_fresh = false;
    c->options = *o;
    return upload_scene_tab(c);
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
    if (field == FE_FIELD_GRID_ENTRIES) c->perm_valid = false;
    if (field == FE_FIELD_POSITION_B) c->wall_g_fresh = false;
    c->ns_fresh = false;
    c->da_fresh = false;
    fe::learn_invalidate(c);
    drop_graph(c);
    if (c->nscenes > 1 && (field == FE_FIELD_GRID_ENTRIES || field == FE_FIELD_GRID_START))
        return fail(FE_ERR_UNSUPPORTED, "a batch's grid fields are per scene (fe_batch_download)");
    // schedule-ordered mirrors of the uploaded field are stale now
    switch (field) {
        case FE_FIELD_POSITION_B: c->mirrors &= ~(fe::kMirrorPos | fe::kMirrorPacc); break;
        case FE_FIELD_VELOCITY_B: c->mirrors &= ~fe::kMirrorVelB; break;
        case FE_FIELD_VELOCITY_A: c->mirrors &= ~fe::kMirrorVelA; break;
        case FE_FIELD_DENSITIES: c->mirrors &= ~(fe::kMirrorDens | fe::kMirrorPrho | fe::kMirrorMrho); break;
        case FE_FIELD_FACTORS: c->mirrors &= ~fe::kMirrorFac; break;
        case FE_FIELD_SOURCE_TERMS: c->mirrors &= ~fe::kMirrorSrc; break;
        case FE_FIELD_PRESSURES: c->mirrors &= ~(fe::kMirrorPrho | fe::kMirrorPres); break;
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
    if (c->nscenes > 1 && (field == FE_FIELD_GRID_ENTRIES || field == FE_FIELD_GRID_START))
        return fail(FE_ERR_UNSUPPORTED, "a batch's grid fields are per scene (fe_batch_download)");
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
    if (c->nscenes > 1) return fail(FE_ERR_UNSUPPORTED, "the stand-alone sort stage takes one scene");
    drop_graph(c);
    c->perm_valid = false;
    if (c->options.sort_algorithm == FE_SORT_BITONIC)
        FE_HIP(fe::bitonic_sort(c), "bitonic sort");
    else
        FE_HIP(fe::radix_sort(c, c->T, 32, false), "radix sort");
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
    FE_HIP(fe::launch_pressure_force(c, nullptr, 0, 0), "pressure_force");
})
FE_STAGE(fe_pressure_update, {
    fe_status s = ensure_ends(c);
    if (s != FE_OK) return s;
    FE_HIP(fe::launch_pressure_update(c, nullptr, nullptr, 0, 0), "pressure_update");
})
FE_STAGE(fe_pressure_integration, { FE_HIP(fe::launch_pressure_integration(c), "pressure_integration"); })
FE_STAGE(fe_divergence_source, {
    fe_status s = ensure_ends(c);
    if (s != FE_OK) return s;
    FE_HIP(fe::launch_divergence_source(c), "divergence_source");
})
FE_STAGE(fe_divergence_integration, {
    if (c->options.divergence_order == FE_DIVERGENCE_CONSISTENT && !c->perm_valid)
        return fail(FE_ERR_UNSUPPORTED, "the consistent divergence order needs a neighbour search first");
    FE_HIP(fe::launch_divergence_integration(c), "divergence_integration");
})
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
    return run_frame(c, max_dt);
}

fe_status fe_run(fe_ctx* c, float max_dt, uint32_t count) {
    FE_CHECK_CTX(c);
    DeviceGuard g(c->device);
    for (uint32_t k = 0; k < count; k++) {
        fe_status s = run_frame(c, max_dt);
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

fe_status fe_solve_report_enable(fe_ctx* c, int enable) {
    FE_CHECK_CTX(c);
    if (c->nscenes > 1) return fail(FE_ERR_UNSUPPORTED, "solve reports take one scene");
    DeviceGuard g(c->device);
    if (enable) {
        if (!c->report_err) FE_HIP(hipMalloc((void**)&c->report_err, sizeof(float) * (size_t)c->n), "report buffer");
        if (!c->report) FE_HIP(hipMalloc((void**)&c->report, sizeof(unsigned long long) * 8), "report sums");
        FE_HIP(hipMemsetAsync(c->report, 0, sizeof(unsigned long long) * 8, c->stream), "report reset");
        c->report_solves[0] = c->report_solves[1] = 0;
    }
    c->report_on = enable != 0;
    return FE_OK;
}

fe_status fe_solve_report(fe_ctx* c, int kind, uint32_t* solves, double out[4]) {
    FE_CHECK_CTX(c);
    if (kind != 0 && kind != 1) return fail(FE_ERR_INVALID_ARGUMENT, "kind must be 0 (density) or 1 (divergence)");
    if (!out) return fail(FE_ERR_INVALID_ARGUMENT, "null out");
    const uint32_t k = c->report ? c->report_solves[kind] : 0;
    if (solves) *solves = k;
    for (int j = 0; j < 4; j++) out[j] = 0.0;
    if (!k) return FE_OK;
    DeviceGuard g(c->device);
    unsigned long long h[4];
    FE_HIP(hipMemcpyAsync(h, c->report + 4 * kind, sizeof(h), hipMemcpyDeviceToHost, c->stream), "report read");
    FE_HIP(hipStreamSynchronize(c->stream), "report wait");
    const double unit = 1.0 / 16777216.0;   // the sums hold r dt / rho0 in units of 2^-24
    out[0] = (double)h[0] * unit / ((double)k * c->n);
    out[1] = (double)h[3] * unit;
    out[2] = (double)h[1] * unit / ((double)k * c->n);
    out[3] = (double)h[2] * unit / ((double)k * c->n);
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

/* ---- render hand-off ------------------------------------------------------------------------- */

fe_status fe_splat_density(fe_ctx* c) {
    FE_CHECK_CTX(c);
    if (c->nscenes > 1) return fail(FE_ERR_UNSUPPORTED, "a batch has no density texture");
    uint64_t centres = 1;
    for (int a = 0; a < 3; a++) {
        if (c->params.grid_res[a] < 1 || c->params.grid_res[a] > 2048)
            return fail(FE_ERR_INVALID_ARGUMENT, "grid_res must be in [1, 2048] per axis");
        centres *= (uint64_t)c->params.grid_res[a] + 1;
    }
    if (centres > (1ull << 31)) return fail(FE_ERR_INVALID_ARGUMENT, "grid_res too large: (x+1)(y+1)(z+1) > 2^31");
    DeviceGuard g(c->device);
    c->splat.valid = false;
    FE_HIP(fe::launch_splat_density(c), "splat_density");
    c->splat.valid = true;
    return FE_OK;
}

fe_status fe_density_texture(fe_ctx* c, void** device_ptr, size_t* bytes) {
    FE_CHECK_CTX(c);
    if (!c->splat.valid) return fail(FE_ERR_UNSUPPORTED, "no density texture: call fe_splat_density first");
    if (device_ptr) *device_ptr = c->splat.tex;
    if (bytes) *bytes = 4 * c->splat.cells;
    return FE_OK;
}

fe_status fe_download_density_texture(fe_ctx* c, uint32_t* dst, size_t bytes) {
    FE_CHECK_CTX(c);
    if (!c->splat.valid) return fail(FE_ERR_UNSUPPORTED, "no density texture: call fe_splat_density first");
    if (bytes != 4 * c->splat.cells)
        return fail(FE_ERR_SIZE_MISMATCH, "density texture is " + std::to_string(4 * c->splat.cells) + " bytes");
    if (!dst) return fail(FE_ERR_INVALID_ARGUMENT, "null destination");
    DeviceGuard g(c->device);
    FE_HIP(hipMemcpyAsync(dst, c->splat.tex, bytes, hipMemcpyDeviceToHost, c->stream), "texture download");
    FE_HIP(hipStreamSynchronize(c->stream), "texture download wait");
    return FE_OK;
}

fe_status fe_copy_field_device(fe_ctx* c, int field, void* dst, size_t bytes) {
    FE_CHECK_CTX(c);
    size_t want;
    void* dev;
    if (!field_info(c, field, &want, &dev)) return fail(FE_ERR_INVALID_ARGUMENT, "unknown field");
    if (!dev) return fail(FE_ERR_UNSUPPORTED, "grid fields have no device layout of the reference's");
    if (bytes != want)
        return fail(FE_ERR_SIZE_MISMATCH, "field size mismatch: expected " + std::to_string(want) + " bytes");
    if (!dst) return fail(FE_ERR_INVALID_ARGUMENT, "null destination");
    DeviceGuard g(c->device);
    FE_HIP(hipMemcpyAsync(dst, dev, bytes, hipMemcpyDeviceToDevice, c->stream), "device copy");
    return FE_OK;
}

/* ---- moving wall ---------------------------------------------------------------------------- */

fe_status fe_wall_init(fe_wall* w, const float min[3], const float max[3]) {
    if (!w || !min || !max) return fail(FE_ERR_INVALID_ARGUMENT, "null argument");
    // CollisionBox::new (src/entities/collision.rs:20-29)
    for (int a = 0; a < 3; a++) {
        w->min[a] = min[a];
        w->max[a] = max[a];
    }
    w->wave_amplitude = 0.0f;
    w->wave_frequency = 1.0f;
    w->time = 0.0f;
    w->base_min_x = min[0];
    return FE_OK;
}

fe_status fe_wall_update(fe_wall* w, float dt) {
    if (!w) return fail(FE_ERR_INVALID_ARGUMENT, "null wall");
    wall_update(w, dt);
    return FE_OK;
}

fe_status fe_set_wall(fe_ctx* c, const fe_wall* wall) {
    FE_CHECK_CTX(c);
    if (c->nscenes > 1) return fail(FE_ERR_UNSUPPORTED, "a batch takes per-scene walls (fe_batch_set_wall)");
    return set_wall(c, 0, wall);
}

fe_status fe_get_wall(const fe_ctx* c, fe_wall* out, int* attached) {
    FE_CHECK_CTX(c);
    if (c->nscenes > 1) return fail(FE_ERR_UNSUPPORTED, "a batch has per-scene walls (fe_batch_get_wall)");
    return get_wall(c, 0, out, attached);
}

/* ---- batched scenes ----------------------------------------------------------------------- */

fe_status fe_batch_create(const fe_params* params, const float* const* xyz, const uint32_t* counts, uint32_t scenes,
                          int device, fe_ctx** out) {
    if (!params || !xyz || !counts || !out) return fail(FE_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (scenes == 0 || scenes > 256) return fail(FE_ERR_INVALID_ARGUMENT, "scenes must be in [1, 256]");
    uint64_t n_all = 0, t_all = 0;
    for (uint32_t k = 0; k < scenes; k++) {
        if (counts[k] == 0 || counts[k] > (1u << 28) || !xyz[k])
            return fail(FE_ERR_INVALID_ARGUMENT, "scene " + std::to_string(k) + ": bad particle count or positions");
        if (!(params[k].smoothing_radius > 0.0f))
            return fail(FE_ERR_INVALID_ARGUMENT, "scene " + std::to_string(k) + ": smoothing_radius must be > 0");
        if (std::memcmp(&params[k].dt, &params[0].dt, sizeof(float)) != 0 ||
            params[k].density_solver_iterations != params[0].density_solver_iterations ||
            params[k].divergence_solver_iterations != params[0].divergence_solver_iterations)
            return fail(FE_ERR_INVALID_ARGUMENT, "scenes advance in lockstep: dt and sweep counts must be equal");
        uint64_t T = 1;
        while (T < counts[k]) T <<= 1;
        n_all += counts[k];
        t_all += T;
    }
    if (n_all > (1u << 30) || t_all > (1u << 31)) return fail(FE_ERR_INVALID_ARGUMENT, "batch too large");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(FE_ERR_DEVICE, "no HIP device available");
    if (device < 0 && hipGetDevice(&device) != hipSuccess) device = 0;
    if (device >= ndev) return fail(FE_ERR_INVALID_ARGUMENT, "device index out of range");
    DeviceGuard guard(device);

    fe_ctx* c = new fe_ctx();
    c->device = device;
    c->params = params[0];
    default_options(&c->options);
    c->nscenes = scenes;
    c->n = (uint32_t)n_all;
    c->T = (uint32_t)t_all;
    c->log2T = 0;
    while ((1ull << c->log2T) < t_all) c->log2T++;
    uint32_t sbits = 0;
    while ((1u << sbits) < scenes) sbits++;
    uint32_t axis = 1;
    for (uint32_t k = 0; k < scenes; k++) axis = std::max(axis, morton_bits_for(params[k]) / 3);
    axis = std::min(axis, (32u - sbits) / 3u);
    c->morton_bits = 3 * axis;
    c->scene_bits = sbits;
    c->scene_params.assign(params, params + scenes);
    std::vector<float> xyz_all(3 * n_all);
    std::vector<uint8_t> sid(n_all);
    uint32_t part = 0, tab = 0;
    for (uint32_t k = 0; k < scenes; k++) {
        uint32_t T = 1;
        while (T < counts[k]) T <<= 1;
        c->scene_n.push_back(counts[k]);
        c->scene_T.push_back(T);
        c->scene_part0.push_back(part);
        c->scene_tab0.push_back(tab);
        std::memcpy(&xyz_all[3 * (size_t)part], xyz[k], sizeof(float) * 3 * (size_t)counts[k]);
        std::memset(&sid[part], (int)k, counts[k]);
        part += counts[k];
        tab += T;
    }
    for (int k = 0; k < 2; k++) {
        c->scene_iters[k].assign(scenes, 0);
        c->scene_total_iters[k].assign(scenes, 0);
    }
    fe_status st = alloc_ctx(c, xyz_all.data());
    hipError_t e = hipSuccess;
    if (st == FE_OK) {
        e = hipMalloc((void**)&c->scene_tab, sizeof(fe::Sph) * scenes);
        if (e == hipSuccess) e = hipMalloc((void**)&c->sid, n_all);
        if (e == hipSuccess) e = hipMalloc((void**)&c->scene_stats, sizeof(uint32_t) * 3 * scenes);
        if (e == hipSuccess) e = hipMemcpy(c->sid, sid.data(), n_all, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(c->scene_stats, 0, sizeof(uint32_t) * 3 * scenes);
        if (e == hipSuccess) e = fe::exact_math_flags(c, c->params, &c->exact_fast);
        if (e != hipSuccess) st = hip_fail(e, "batch tables");
    }
    if (st == FE_OK) st = upload_scene_tab(c);
    if (st != FE_OK) {
        free_ctx(c);
        return st;
    }
    *out = c;
    return FE_OK;
}

fe_status fe_batch_scenes(const fe_ctx* c, uint32_t* scenes) {
    FE_CHECK_CTX(c);
    if (!scenes) return fail(FE_ERR_INVALID_ARGUMENT, "null out");
    *scenes = c->nscenes;
    return FE_OK;
}

static fe_status check_scene(const fe_ctx* c, uint32_t scene) {
    if (c->nscenes <= 1) return fail(FE_ERR_INVALID_ARGUMENT, "not a batch context (fe_batch_create)");
    if (scene >= c->nscenes) return fail(FE_ERR_INVALID_ARGUMENT, "scene index out of range");
    return FE_OK;
}

fe_status fe_batch_scene_info(const fe_ctx* c, uint32_t scene, uint32_t* n, uint32_t* table_size) {
    FE_CHECK_CTX(c);
    fe_status s = check_scene(c, scene);
    if (s != FE_OK) return s;
    if (n) *n = c->scene_n[scene];
    if (table_size) *table_size = c->scene_T[scene];
    return FE_OK;
}

fe_status fe_batch_set_params(fe_ctx* c, uint32_t scene, const fe_params* p) {
    FE_CHECK_CTX(c);
    fe_status s = check_scene(c, scene);
    if (s != FE_OK) return s;
    if (!p || !(p->smoothing_radius > 0.0f)) return fail(FE_ERR_INVALID_ARGUMENT, "bad params");
    if (std::memcmp(&p->dt, &c->params.dt, sizeof(float)) != 0 ||
        p->density_solver_iterations != c->params.density_solver_iterations ||
        p->divergence_solver_iterations != c->params.divergence_solver_iterations)
        return fail(FE_ERR_INVALID_ARGUMENT, "scenes advance in lockstep: dt and sweep counts must stay equal");
    DeviceGuard g(c->device);
    drop_graph(c);
    if (std::memcmp(&p->smoothing_radius, &c->scene_params[scene].smoothing_radius, sizeof(float)) != 0)
        c->list_valid = false;   // the list holds the pairs within the old h
    c->scene_params[scene] = *p;
    c->ns_fresh = c->da_fresh = false;   // h may have changed
    c->mirrors &= ~fe::kMirrorMrho;      // holds the old mass
    return upload_scene_tab(c);
}

// One scene's field in its own layout: particle fields are the scene's range; the grid
// fields are rebuilt as the scene's own T_s-entry table (indices relative to the scene).
fe_status fe_batch_download(fe_ctx* c, uint32_t scene, int field, void* dst, size_t bytes) {
    FE_CHECK_CTX(c);
    fe_status s = check_scene(c, scene);
    if (s != FE_OK) return s;
    if (!dst) return fail(FE_ERR_INVALID_ARGUMENT, "null destination");
    DeviceGuard g(c->device);
    const uint32_t n = c->scene_n[scene], T = c->scene_T[scene], p0 = c->scene_part0[scene], t0 = c->scene_tab0[scene];
    if (field == FE_FIELD_GRID_ENTRIES || field == FE_FIELD_GRID_START) {
        s = ensure_ends(c);
        if (s != FE_OK) return s;
        const size_t want = field == FE_FIELD_GRID_ENTRIES ? 8 * (size_t)T : 4 * (size_t)T;
        if (bytes != want) return fail(FE_ERR_SIZE_MISMATCH, "field size mismatch: expected " + std::to_string(want));
        uint32_t* o = static_cast<uint32_t*>(dst);
        if (field == FE_FIELD_GRID_ENTRIES) {
            std::vector<uint32_t> k(n), v(n);
            FE_HIP(hipMemcpyAsync(k.data(), c->keys[0] + p0, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream), "download");
            FE_HIP(hipMemcpyAsync(v.data(), c->vals[0] + p0, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream), "download");
            FE_HIP(hipStreamSynchronize(c->stream), "download wait");
            for (uint32_t i = 0; i < T; i++) {
                o[2 * (size_t)i] = i < n ? k[i] - t0 : fe::kSentinel;
                o[2 * (size_t)i + 1] = i < n ? v[i] - p0 : fe::kSentinel;
            }
        } else {
            std::vector<uint2> r(T);
            FE_HIP(hipMemcpyAsync(r.data(), c->range + t0, 8 * (size_t)T, hipMemcpyDeviceToHost, c->stream), "download");
            FE_HIP(hipStreamSynchronize(c->stream), "download wait");
            for (uint32_t i = 0; i < T; i++) o[i] = r[i].x == fe::kSentinel ? fe::kSentinel : r[i].x - p0;
        }
        return FE_OK;
    }
    size_t whole;
    void* dev;
    if (!field_info(c, field, &whole, &dev) || !dev) return fail(FE_ERR_INVALID_ARGUMENT, "unknown field");
    const size_t per = whole / c->n;
    if (bytes != per * n) return fail(FE_ERR_SIZE_MISMATCH, "field size mismatch: expected " + std::to_string(per * n));
    FE_HIP(hipMemcpyAsync(dst, static_cast<char*>(dev) + per * p0, bytes, hipMemcpyDeviceToHost, c->stream), "download");
    FE_HIP(hipStreamSynchronize(c->stream), "download wait");
    return FE_OK;
}

fe_status fe_batch_upload(fe_ctx* c, uint32_t scene, int field, const void* src, size_t bytes) {
    FE_CHECK_CTX(c);
    fe_status s = check_scene(c, scene);
    if (s != FE_OK) return s;
    if (!src) return fail(FE_ERR_INVALID_ARGUMENT, "null source");
    if (field == FE_FIELD_GRID_ENTRIES || field == FE_FIELD_GRID_START)
        return fail(FE_ERR_UNSUPPORTED, "a batch's grid fields are rebuilt by the neighbour search");
    size_t whole;
    void* dev;
    if (!field_info(c, field, &whole, &dev) || !dev) return fail(FE_ERR_INVALID_ARGUMENT, "unknown field");
    const size_t per = whole / c->n;
    const uint32_t n = c->scene_n[scene], p0 = c->scene_part0[scene];
    if (bytes != per * n) return fail(FE_ERR_SIZE_MISMATCH, "field size mismatch: expected " + std::to_string(per * n));
    DeviceGuard g(c->device);
    drop_graph(c);
    c->ns_fresh = c->da_fresh = false;
    c->list_valid = false;
    c->mirrors = 0;
    FE_HIP(hipMemcpyAsync(static_cast<char*>(dev) + per * p0, src, bytes, hipMemcpyHostToDevice, c->stream), "upload");
    FE_HIP(hipStreamSynchronize(c->stream), "upload wait");
    return FE_OK;
}

fe_status fe_batch_read_stats(fe_ctx* c, uint32_t scene, uint32_t out[3]) {
    FE_CHECK_CTX(c);
    fe_status s = check_scene(c, scene);
    if (s != FE_OK) return s;
    if (!out) return fail(FE_ERR_INVALID_ARGUMENT, "null out");
    DeviceGuard g(c->device);
    FE_HIP(hipMemcpyAsync(c->host_stats, c->scene_stats + 3 * (size_t)scene, 12, hipMemcpyDeviceToHost, c->stream),
           "stats copy");
    FE_HIP(hipStreamSynchronize(c->stream), "stats wait");
    std::memcpy(out, c->host_stats, 12);
    return FE_OK;
}

fe_status fe_batch_iterations(fe_ctx* c, uint32_t scene, uint32_t* density, uint32_t* divergence,
                              uint64_t* density_total, uint64_t* divergence_total) {
    FE_CHECK_CTX(c);
    fe_status s = check_scene(c, scene);
    if (s != FE_OK) return s;
    if (density) *density = c->scene_iters[0][scene];
    if (divergence) *divergence = c->scene_iters[1][scene];
    if (density_total) *density_total = c->scene_total_iters[0][scene];
    if (divergence_total) *divergence_total = c->scene_total_iters[1][scene];
    return FE_OK;
}

fe_status fe_batch_set_wall(fe_ctx* c, uint32_t scene, const fe_wall* wall) {
    FE_CHECK_CTX(c);
    fe_status s = check_scene(c, scene);
    if (s != FE_OK) return s;
    return set_wall(c, scene, wall);
}

fe_status fe_batch_get_wall(const fe_ctx* c, uint32_t scene, fe_wall* out, int* attached) {
    FE_CHECK_CTX(c);
    fe_status s = check_scene(c, scene);
    if (s != FE_OK) return s;
    return get_wall(c, scene, out, attached);
}

fe_status fe_batch_use_graph(fe_ctx* c, int enable) {
    FE_CHECK_CTX(c);
    if (c->nscenes <= 1) return fail(FE_ERR_INVALID_ARGUMENT, "not a batch context (fe_batch_create)");
    DeviceGuard g(c->device);
    drop_graph(c);
    c->batch_graph = enable != 0;
    return FE_OK;
}

fe_status fe_use_graph(fe_ctx* c, int enable) {
    FE_CHECK_CTX(c);
    DeviceGuard g(c->device);
    drop_graph(c);
    c->batch_graph = enable != 0;
    return FE_OK;
}

fe_status fe_set_watchdog(fe_ctx* c, int enable, float speed_limit, int halt) {
    FE_CHECK_CTX(c);
    if (enable && c->nscenes > 1) return fail(FE_ERR_UNSUPPORTED, "the watchdog takes a single-scene context");
    if (!(speed_limit >= 0.0f)) return fail(FE_ERR_INVALID_ARGUMENT, "speed_limit must be >= 0 (0: no limit)");
    DeviceGuard g(c->device);
    FE_HIP(hipStreamSynchronize(c->stream), "watchdog wait");
    fe_status s = wait_prev_stats(c);
    if (s != FE_OK) return s;
    drop_graph(c);   // the stats kernels differ
    c->watchdog = enable != 0;
    c->watchdog_halt = enable && halt;
    c->speed_limit = speed_limit;
    c->health = fe_health{};
    c->health.first_nonfinite = 0xFFFFFFFFu;
    c->health.speed_limit = speed_limit;
    return FE_OK;
}

fe_status fe_get_health(fe_ctx* c, fe_health* out) {
    FE_CHECK_CTX(c);
    if (!out) return fail(FE_ERR_INVALID_ARGUMENT, "null out");
    if (!c->watchdog) return fail(FE_ERR_INVALID_ARGUMENT, "the watchdog is off (fe_set_watchdog)");
    DeviceGuard g(c->device);
    fe_status s = wait_prev_stats(c);
    if (s != FE_OK) return s;
    *out = c->health;
    return FE_OK;
}

}  // extern "C"
