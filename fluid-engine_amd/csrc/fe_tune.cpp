// fe_tune.cpp — the launch-configuration autotuner (no reference counterpart): for each
// tunable kernel, every configuration it is built for is timed on the context's current
// state (the median of `reps` launches, each bracketed by HIP events on the context's
// stream) and the fastest one is kept. Block sizes and items per thread never change
// results, so tuning changes speed only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "fe_abi.h"

namespace {

using fe::DeviceGuard;
using fe::fail;
using fe::hip_fail;

hipError_t launch_tuned(fe_ctx* c, int id) {
    switch (id) {
        case fe::kTuneDensityAlpha: return fe::launch_density_alpha_only(c);
        case fe::kTuneViscosity: return fe::launch_viscosity(c);
        case fe::kTuneDensitySource: return fe::launch_density_source(c);
        case fe::kTuneDivergenceSource: return fe::launch_divergence_source(c);
        case fe::kTunePressureForce: return fe::launch_pressure_force(c, nullptr, 0, 0);
        // in converge mode each block of a sweep adds its residual into the sweep's slot, as
        // in the frame (the slots are cleared when a solve starts): a configuration with many
        // blocks pays for those same-word atomics here too
        case fe::kTunePressureUpdate:
            return fe::launch_pressure_update(c, c->options.solver_mode == FE_SOLVER_CONVERGE ? c->resid : nullptr,
                                              nullptr, 0, 0);
        case fe::kTuneBuildList: return fe::launch_build_list(c);
        case fe::kTunePressureIntegration: return fe::launch_pressure_integration(c);
        case fe::kTuneDivergenceIntegration: return fe::launch_divergence_integration(c);
        case fe::kTuneDensityAlphaDivergence: return fe::launch_density_alpha_divergence_only(c);
        default: return hipErrorInvalidValue;
    }
}

// The staged-window sweeps (items 2) are not candidates: their window tables are built once
// per list, a cost that timing sweep launches alone would leave out. fe_tune_set selects them.
std::vector<fe::TuneCfg> candidates(int id) {
    std::vector<fe::TuneCfg> out;
    const bool sweep = id == fe::kTunePressureForce || id == fe::kTunePressureUpdate;
    for (uint32_t b : {64u, 128u, 256u, 512u, 1024u})
        for (uint32_t it : {1u, 2u, 4u, 8u})
            if (fe::tune_valid(id, b, it) && !(sweep && it == 2u)) {
                fe::TuneCfg t;
                t.block = b;
                t.items = it;
                out.push_back(t);
            }
    return out;
}

// Device copies of the fields the timed launches overwrite, restored afterwards.
struct Snapshot {
    struct Item {
        void* dev;
        void* copy;
        size_t bytes;
    };
    std::vector<Item> items;
    ~Snapshot() {
        for (auto& it : items)
            if (it.copy) (void)hipFree(it.copy);
    }
};

}  // namespace

extern "C" {

fe_status fe_autotune(fe_ctx* c, uint32_t reps) {
    FE_CHECK_CTX(c);
    if (reps == 0) reps = 20;
    if (reps > 1000) return fail(FE_ERR_INVALID_ARGUMENT, "reps must be <= 1000");
    if (!c->list_valid)
        return fail(FE_ERR_INVALID_ARGUMENT, "tuning times the frame's kernels on a stepped state: run a frame first");
    DeviceGuard g(c->device);
    fe::drop_graph(c);
    FE_HIP(hipStreamSynchronize(c->stream), "tune wait");
    const size_t n = c->n;
    Snapshot snap;
    // velocity_b: the consistent divergence order writes it too
    for (auto f : std::vector<std::pair<void*, size_t>>{{c->pos_a, 16 * n}, {c->vel_a, 16 * n}, {c->vel_b, 16 * n},
                                                        {c->colors, 16 * n}, {c->pacc, 16 * n}, {c->pres, 4 * n},
                                                        {c->src, 4 * n}}) {
        Snapshot::Item it{f.first, nullptr, f.second};
        FE_HIP(hipMalloc(&it.copy, it.bytes), "hipMalloc tune snapshot");
        snap.items.push_back(it);
        FE_HIP(hipMemcpyAsync(it.copy, it.dev, it.bytes, hipMemcpyDeviceToDevice, c->stream), "tune snapshot");
    }
    const bool ns_fresh = c->ns_fresh, da_fresh = c->da_fresh, profiling = c->profiling;
    c->profiling = false;
    std::vector<hipEvent_t> ev(2 * reps);
    fe_status st = FE_OK;
    for (auto& e : ev)
        if (st == FE_OK && hipEventCreate(&e) != hipSuccess) st = fail(FE_ERR_DEVICE, "hipEventCreate");
    for (int id = 0; id < fe::kTuneCount && st == FE_OK; id++) {
        const fe::TuneCfg keep = c->tune[id];
        fe::TuneCfg best = keep;
        float best_us = 0.0f, default_us = 0.0f;
        for (const fe::TuneCfg& cand : candidates(id)) {
            c->tune[id].block = cand.block;
            c->tune[id].items = cand.items;
            hipError_t e = hipSuccess;
            for (int w = 0; w < 2 && e == hipSuccess; w++) e = launch_tuned(c, id);
            for (uint32_t r = 0; r < reps && e == hipSuccess; r++) {
                e = hipEventRecord(ev[2 * r], c->stream);
                if (e == hipSuccess) e = launch_tuned(c, id);
                if (e == hipSuccess) e = hipEventRecord(ev[2 * r + 1], c->stream);
            }
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            std::vector<float> us(reps);
            for (uint32_t r = 0; r < reps && e == hipSuccess; r++) {
                float ms = 0.0f;
                e = hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]);
                us[r] = ms * 1e3f;
            }
            if (e != hipSuccess) {
                st = hip_fail(e, "tuning launch");
                break;
            }
            std::nth_element(us.begin(), us.begin() + reps / 2, us.end());
            const float med = us[reps / 2];
            const fe::TuneCfg def;
            const bool is_default = cand.block == def.block && cand.items == def.items;
            if (is_default) default_us = med;
            if (best_us == 0.0f || med < best_us) {
                best_us = med;
                best = cand;
            }
        }
        c->tune[id] = best;
        c->tune[id].median_us = best_us;
        c->tune[id].default_us = default_us;
        (void)keep;
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& it : snap.items)
        if (st == FE_OK) {
            hipError_t e = hipMemcpyAsync(it.dev, it.copy, it.bytes, hipMemcpyDeviceToDevice, c->stream);
            if (e != hipSuccess) st = hip_fail(e, "tune restore");
        }
    if (st == FE_OK) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) st = hip_fail(e, "tune restore wait");
    }
    // the mirrors are rebuilt from the restored fields before their next use
    c->mirrors = 0;
    c->ns_fresh = ns_fresh;
    c->da_fresh = da_fresh;
    c->profiling = profiling;
    fe::learn_invalidate(c);
    return st;
}

fe_status fe_tune_count(const fe_ctx* c, uint32_t* count) {
    FE_CHECK_CTX(c);
    if (!count) return fail(FE_ERR_INVALID_ARGUMENT, "null out");
    *count = fe::kTuneCount;
    return FE_OK;
}

fe_status fe_tune_get(const fe_ctx* c, uint32_t index, fe_tune_entry* out) {
    FE_CHECK_CTX(c);
    if (!out) return fail(FE_ERR_INVALID_ARGUMENT, "null out");
    if (index >= (uint32_t)fe::kTuneCount) return fail(FE_ERR_INVALID_ARGUMENT, "tune index out of range");
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->kernel, fe::tune_name((int)index), sizeof(out->kernel) - 1);
    out->block = c->tune[index].block;
    out->items = c->tune[index].items;
    out->median_us = c->tune[index].median_us;
    out->default_us = c->tune[index].default_us;
    return FE_OK;
}

fe_status fe_tune_set(fe_ctx* c, const char* kernel, uint32_t block, uint32_t items) {
    FE_CHECK_CTX(c);
    if (!kernel) return fail(FE_ERR_INVALID_ARGUMENT, "null kernel name");
    for (int id = 0; id < fe::kTuneCount; id++) {
        if (std::strcmp(kernel, fe::tune_name(id)) != 0) continue;
        if (!fe::tune_valid(id, block, items))
            return fail(FE_ERR_INVALID_ARGUMENT, std::string(kernel) + ": no build for block " + std::to_string(block) +
                                                     ", items " + std::to_string(items));
        fe::drop_graph(c);
        c->tune[id].block = block;
        c->tune[id].items = items;
        return FE_OK;
    }
    return fail(FE_ERR_INVALID_ARGUMENT, std::string("no tunable kernel named ") + kernel);
}

fe_status fe_check_exact_arith(int device, float h, uint64_t samples, uint64_t counts[6]) {
    if (!counts) return fail(FE_ERR_INVALID_ARGUMENT, "null counts");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(FE_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(FE_ERR_INVALID_ARGUMENT, "device index out of range");
    if (!(h > 0.0f)) return fail(FE_ERR_INVALID_ARGUMENT, "h must be > 0");
    DeviceGuard g(device);
    hipStream_t st = nullptr;
    unsigned long long* bad = nullptr;
    FE_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    hipError_t e = hipMalloc((void**)&bad, 3 * sizeof(unsigned long long));
    unsigned long long c[6] = {0, 0, 0, 0, 0, 0};
    if (e == hipSuccess) e = fe::check_exact_arith(st, h, samples, bad, c);
    if (bad) (void)hipFree(bad);
    (void)hipStreamDestroy(st);
    if (e != hipSuccess) return hip_fail(e, "exact arithmetic check");
    for (int k = 0; k < 6; k++) counts[k] = c[k];
    return FE_OK;
}

fe_status fe_device_name(const fe_ctx* c, char* name, size_t cap) {
    FE_CHECK_CTX(c);
    if (!name || cap == 0) return fail(FE_ERR_INVALID_ARGUMENT, "null name buffer");
    hipDeviceProp_t prop;
    FE_HIP(hipGetDeviceProperties(&prop, c->device), "hipGetDeviceProperties");
    const std::string s = std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount) + "cu";
    std::strncpy(name, s.c_str(), cap - 1);
    name[cap - 1] = '\0';
    return FE_OK;
}

}  // extern "C"
