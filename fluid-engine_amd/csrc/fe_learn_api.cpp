// fe_learn_api.cpp — host side of the learned warm start: buffers, the solve hooks, dataset
// collection, training orchestration and the fe_learn_* entry points. No reference
// counterpart; the kernels are in fe_learn.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "fe_abi.h"

static_assert(FE_MLP_PARAMS == fe::kMlpParams, "FE_MLP_PARAMS");
static_assert(FE_MLP_NORM == fe::kMlpNorm, "FE_MLP_NORM");
static_assert(FE_MLP_INPUTS == fe::kMlpIn && FE_MLP_HIDDEN == fe::kMlpHidden, "MLP shape");

namespace {

using fe::DeviceGuard;
using fe::fail;
using fe::hip_fail;
using fe::kMlpIn;
using fe::kMlpNorm;
using fe::kMlpParams;

fe_status check_net(int net) {
    if (net != FE_NET_DENSITY && net != FE_NET_DIVERGENCE) return fail(FE_ERR_INVALID_ARGUMENT, "net must be 0 or 1");
    return FE_OK;
}

// Per-particle buffers and the parameter sets (the datasets are sized by fe_learn_collect).
fe_status ensure_alloc(fe_ctx* c) {
    fe::LearnState& w = c->ws;
    if (w.alloc) return FE_OK;
    if (c->nscenes > 1) return fail(FE_ERR_UNSUPPORTED, "the learned warm start takes one scene per context");
    const size_t n = c->n;
    FE_HIP(hipMalloc(&w.feat, sizeof(float) * kMlpIn * n), "hipMalloc features");
    FE_HIP(hipMalloc(&w.pprev, sizeof(float) * n), "hipMalloc pprev");
    for (int k = 0; k < 2; k++) {
        FE_HIP(hipMalloc(&w.plast[k], sizeof(float) * n), "hipMalloc plast");
        FE_HIP(hipMalloc(&w.params[k], sizeof(float) * kMlpParams), "hipMalloc params");
        FE_HIP(hipMalloc(&w.m1[k], sizeof(float) * kMlpParams), "hipMalloc adam");
        FE_HIP(hipMalloc(&w.m2[k], sizeof(float) * kMlpParams), "hipMalloc adam");
        FE_HIP(hipMalloc(&w.norm[k], sizeof(float) * kMlpNorm), "hipMalloc norm");
        FE_HIP(hipMalloc(&w.packed[k], sizeof(float) * fe::kMlpPackedFloats), "hipMalloc packed weights");
    }
    const uint32_t parts = fe::kMlpMaxBatch / fe::kMlpSamplesPerBlock;
    FE_HIP(hipMalloc(&w.grads, sizeof(float) * kMlpParams * (size_t)parts), "hipMalloc grads");
    FE_HIP(hipMalloc(&w.loss, sizeof(float) * parts), "hipMalloc loss");
    w.alloc = true;
    return FE_OK;
}

fe_status ensure_io(fe_ctx* c, size_t floats) {
    fe::LearnState& w = c->ws;
    if (w.io_floats >= floats) return FE_OK;
    if (w.io) (void)hipFree(w.io);
    w.io = nullptr;
    w.io_floats = 0;
    FE_HIP(hipMalloc(&w.io, sizeof(float) * floats), "hipMalloc learn staging");
    w.io_floats = floats;
    return FE_OK;
}

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Glorot-uniform weights, zero biases; the output layer starts small.
void init_params(std::vector<float>& p, uint32_t seed) {
    p.assign(kMlpParams, 0.0f);
    uint64_t s = 0x5EEDull * 1000003ull + seed;
    auto uni = [&](float a) { return (float)(((double)(splitmix64(s) >> 11) * 0x1.0p-53 * 2.0 - 1.0) * a); };
    const int H = fe::kMlpHidden;
    int o = 0;
    const float a1 = std::sqrt(6.0f / (kMlpIn + H)), a2 = std::sqrt(6.0f / (2 * H)), a4 = std::sqrt(6.0f / (H + 1));
    for (int e = 0; e < kMlpIn * H; e++) p[o++] = uni(a1);
    o += H;
    for (int e = 0; e < H * H; e++) p[o++] = uni(a2);
    o += H;
    for (int e = 0; e < H * H; e++) p[o++] = uni(a2);
    o += H;
    for (int e = 0; e < H; e++) p[o++] = 0.1f * uni(a4);
}

// Dataset statistics in double: feature means / stds, target mean / std (std floored).
void dataset_norm(const std::vector<float>& x, const std::vector<float>& t, uint32_t n, float* norm) {
    for (int k = 0; k < kMlpIn; k++) {
        double m = 0.0, v = 0.0;
        for (uint32_t i = 0; i < n; i++) m += x[(size_t)i * kMlpIn + k];
        m /= n;
        for (uint32_t i = 0; i < n; i++) {
            const double d = x[(size_t)i * kMlpIn + k] - m;
            v += d * d;
        }
        const double sd = std::sqrt(v / n);
        norm[k] = (float)m;
        norm[kMlpIn + k] = (float)(sd > 1e-12 * (1.0 + std::fabs(m)) && sd > 0.0 ? sd : 1.0);
    }
    double m = 0.0, v = 0.0;
    for (uint32_t i = 0; i < n; i++) m += t[i];
    m /= n;
    for (uint32_t i = 0; i < n; i++) v += (t[i] - m) * (t[i] - m);
    const double sd = std::sqrt(v / n);
    norm[2 * kMlpIn] = (float)m;
    norm[2 * kMlpIn + 1] = (float)(sd > 0.0 ? sd : 1.0);
}

}  // namespace

namespace fe {

bool learn_ready(const fe_ctx* c, int kind) { return c->ws.alloc && c->ws.ready[kind]; }

void learn_invalidate(fe_ctx* c) {
    c->ws.plast_valid[0] = c->ws.plast_valid[1] = false;
    c->ws.feat_fresh = false;
}

bool learn_features_wanted(const fe_ctx* c, int kind) {
    const LearnState& w = c->ws;
    if (!w.alloc) return false;
    return w.force_features || (c->options.warm_start == FE_WARM_START_LEARNED && w.ready[kind]) ||
           (w.per_solve > 0 && w.capacity > 0);
}

void learn_features_done(fe_ctx* c, int kind) {
    c->ws.feat_fresh = true;
    c->ws.feat_kind = kind;
}

void learn_free(fe_ctx* c) {
    LearnState& w = c->ws;
    for (void* p : {(void*)w.feat, (void*)w.pprev, (void*)w.plast[0], (void*)w.plast[1], (void*)w.params[0],
                    (void*)w.params[1], (void*)w.m1[0], (void*)w.m1[1], (void*)w.m2[0], (void*)w.m2[1],
                    (void*)w.norm[0], (void*)w.norm[1], (void*)w.data_x[0], (void*)w.data_x[1], (void*)w.data_t[0],
                    (void*)w.data_t[1], (void*)w.train_x, (void*)w.train_t, (void*)w.grads, (void*)w.loss,
                    (void*)w.io, (void*)w.packed[0], (void*)w.packed[1]})
        if (p) (void)hipFree(p);
    w = LearnState{};
}

// The features come from the solve's source-term stage (fe_density_source /
// fe_divergence_source), which gathers them in the same pass whenever they are wanted.
fe_status learn_before_solve(fe_ctx* c, int kind) {
    LearnState& w = c->ws;
    if (!w.alloc) return FE_OK;
    if (c->options.warm_start != FE_WARM_START_LEARNED || !w.ready[kind]) return FE_OK;
    if (!w.feat_fresh || w.feat_kind != kind)
        return fail(FE_ERR_INVALID_ARGUMENT,
                    "the learned warm start needs the features of this solve's source-term stage: run "
                    "fe_density_source / fe_divergence_source (with warm_start LEARNED set) right before the solve");
    FE_HIP(learn_infer(c, kind), "warm-start inference");
    return FE_OK;
}

fe_status learn_after_solve(fe_ctx* c, int kind) {
    LearnState& w = c->ws;
    if (!w.alloc) return FE_OK;
    if (w.per_solve > 0 && w.capacity > 0 && w.feat_fresh && w.feat_kind == kind)
        FE_HIP(learn_collect(c, kind), "warm-start samples");
    FE_HIP(hipMemcpyAsync(w.plast[kind], c->pres, sizeof(float) * c->n, hipMemcpyDeviceToDevice, c->stream),
           "pressure history");
    w.plast_valid[kind] = true;
    w.feat_fresh = false;
    return FE_OK;
}

}  // namespace fe

extern "C" {

fe_status fe_learn_collect(fe_ctx* c, uint32_t samples_per_solve, uint32_t capacity) {
    FE_CHECK_CTX(c);
    DeviceGuard g(c->device);
    fe_status s = ensure_alloc(c);
    if (s != FE_OK) return s;
    fe::LearnState& w = c->ws;
    if (samples_per_solve > 0 && capacity == 0) return fail(FE_ERR_INVALID_ARGUMENT, "capacity must be > 0");
    if (capacity != 0 && capacity != w.capacity) {
        FE_HIP(hipStreamSynchronize(c->stream), "collect wait");
        for (int k = 0; k < 2; k++) {
            if (w.data_x[k]) (void)hipFree(w.data_x[k]);
            if (w.data_t[k]) (void)hipFree(w.data_t[k]);
            w.data_x[k] = w.data_t[k] = nullptr;
            w.head[k] = w.size[k] = 0;
        }
        if (w.train_x) (void)hipFree(w.train_x);
        if (w.train_t) (void)hipFree(w.train_t);
        w.train_x = w.train_t = nullptr;
        w.capacity = 0;
        for (int k = 0; k < 2; k++) {
            FE_HIP(hipMalloc(&w.data_x[k], sizeof(float) * kMlpIn * (size_t)capacity), "hipMalloc dataset");
            FE_HIP(hipMalloc(&w.data_t[k], sizeof(float) * (size_t)capacity), "hipMalloc dataset");
        }
        FE_HIP(hipMalloc(&w.train_x, sizeof(float) * kMlpIn * (size_t)capacity), "hipMalloc training set");
        FE_HIP(hipMalloc(&w.train_t, sizeof(float) * (size_t)capacity), "hipMalloc training set");
        w.capacity = capacity;
    }
    w.per_solve = samples_per_solve;
    return FE_OK;
}

fe_status fe_learn_dataset_size(fe_ctx* c, int net, uint32_t* size) {
    FE_CHECK_CTX(c);
    if (check_net(net) != FE_OK) return FE_ERR_INVALID_ARGUMENT;
    if (!size) return fail(FE_ERR_INVALID_ARGUMENT, "null out");
    *size = c->ws.size[net];
    return FE_OK;
}

fe_status fe_learn_dataset(fe_ctx* c, int net, float* x, float* t, uint32_t count) {
    FE_CHECK_CTX(c);
    if (check_net(net) != FE_OK) return FE_ERR_INVALID_ARGUMENT;
    if (count > c->ws.size[net]) return fail(FE_ERR_SIZE_MISMATCH, "count exceeds the dataset size");
    if (count == 0) return FE_OK;
    DeviceGuard g(c->device);
    if (x)
        FE_HIP(hipMemcpyAsync(x, c->ws.data_x[net], sizeof(float) * kMlpIn * count, hipMemcpyDeviceToHost, c->stream),
               "dataset download");
    if (t)
        FE_HIP(hipMemcpyAsync(t, c->ws.data_t[net], sizeof(float) * count, hipMemcpyDeviceToHost, c->stream),
               "dataset download");
    FE_HIP(hipStreamSynchronize(c->stream), "dataset wait");
    return FE_OK;
}

fe_status fe_learn_train(fe_ctx* c, int net, uint32_t steps, uint32_t batch, float lr, uint32_t seed,
                         float* losses) {
    FE_CHECK_CTX(c);
    if (check_net(net) != FE_OK) return FE_ERR_INVALID_ARGUMENT;
    if (batch == 0 || batch > fe::kMlpMaxBatch) return fail(FE_ERR_INVALID_ARGUMENT, "batch must be in [1, 65536]");
    if (!(lr > 0.0f)) return fail(FE_ERR_INVALID_ARGUMENT, "lr must be > 0");
    DeviceGuard g(c->device);
    fe_status s = ensure_alloc(c);
    if (s != FE_OK) return s;
    fe::LearnState& w = c->ws;
    const uint32_t n = w.size[net];
    if (n == 0) return fail(FE_ERR_INVALID_ARGUMENT, "empty dataset: collect samples first (fe_learn_collect)");
    // normalisation from the dataset, computed on the host in double
    std::vector<float> x((size_t)n * kMlpIn), t(n);
    FE_HIP(hipMemcpyAsync(x.data(), w.data_x[net], sizeof(float) * x.size(), hipMemcpyDeviceToHost, c->stream),
           "dataset download");
    FE_HIP(hipMemcpyAsync(t.data(), w.data_t[net], sizeof(float) * n, hipMemcpyDeviceToHost, c->stream),
           "dataset download");
    FE_HIP(hipStreamSynchronize(c->stream), "dataset wait");
    float norm[kMlpNorm];
    dataset_norm(x, t, n, norm);
    for (uint32_t i = 0; i < n; i++) {
        for (int k = 0; k < kMlpIn; k++) {
            float& v = x[(size_t)i * kMlpIn + k];
            v = (v - norm[k]) / norm[kMlpIn + k];
        }
        t[i] = (t[i] - norm[2 * kMlpIn]) / norm[2 * kMlpIn + 1];
    }
    std::vector<float> p0;
    init_params(p0, seed);
    FE_HIP(hipMemcpyAsync(w.train_x, x.data(), sizeof(float) * x.size(), hipMemcpyHostToDevice, c->stream),
           "training set upload");
    FE_HIP(hipMemcpyAsync(w.train_t, t.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->stream),
           "training set upload");
    FE_HIP(hipMemcpyAsync(w.params[net], p0.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice, c->stream),
           "parameter upload");
    FE_HIP(hipMemcpyAsync(w.norm[net], norm, sizeof(norm), hipMemcpyHostToDevice, c->stream), "norm upload");
    FE_HIP(hipMemsetAsync(w.m1[net], 0, sizeof(float) * kMlpParams, c->stream), "adam reset");
    FE_HIP(hipMemsetAsync(w.m2[net], 0, sizeof(float) * kMlpParams, c->stream), "adam reset");
    w.adam_t[net] = 0;
    w.ready[net] = false;
    float* dloss = nullptr;
    if (losses && steps) FE_HIP(hipMalloc(&dloss, sizeof(float) * steps), "hipMalloc losses");
    // a batch of at least the dataset's size takes every sample once per step, in order
    const uint32_t rows = batch >= n ? n : batch, mod = batch >= n ? 0u : n;
    fe_status st = FE_OK;
    for (uint32_t k = 0; k < steps && st == FE_OK; k++) {
        // cosine decay from lr to lr / 20
        const double frac = steps > 1 ? (double)k / (double)(steps - 1) : 0.0;
        const float lrk = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.14159265358979323846 * frac))));
        hipError_t e = fe::learn_step(c, net, w.train_x, w.train_t, rows, mod, seed * 0x9E3779B9u + k, lrk, 1u,
                                      nullptr, dloss ? dloss + k : nullptr);
        if (e != hipSuccess) st = hip_fail(e, "training step");
    }
    if (st == FE_OK && dloss) {
        hipError_t e = hipMemcpyAsync(losses, dloss, sizeof(float) * steps, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) st = hip_fail(e, "loss download");
        for (uint32_t k = 0; st == FE_OK && k < steps; k++) losses[k] /= (float)rows;
    }
    if (st == FE_OK) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) st = hip_fail(e, "training wait");
    }
    if (dloss) (void)hipFree(dloss);
    if (st == FE_OK) {
        hipError_t e = fe::learn_pack(c, net);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) st = hip_fail(e, "weight packing");
    }
    if (st == FE_OK) w.ready[net] = true;   // steps = 0 leaves the initialisation
    return st;
}

fe_status fe_learn_get_weights(fe_ctx* c, int net, float* params, float* norm) {
    FE_CHECK_CTX(c);
    if (check_net(net) != FE_OK) return FE_ERR_INVALID_ARGUMENT;
    if (!fe::learn_ready(c, net)) return fail(FE_ERR_INVALID_ARGUMENT, "network not trained or set");
    DeviceGuard g(c->device);
    if (params)
        FE_HIP(hipMemcpyAsync(params, c->ws.params[net], sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, c->stream),
               "weights download");
    if (norm)
        FE_HIP(hipMemcpyAsync(norm, c->ws.norm[net], sizeof(float) * kMlpNorm, hipMemcpyDeviceToHost, c->stream),
               "norm download");
    FE_HIP(hipStreamSynchronize(c->stream), "weights wait");
    return FE_OK;
}

fe_status fe_learn_set_weights(fe_ctx* c, int net, const float* params, const float* norm) {
    FE_CHECK_CTX(c);
    if (check_net(net) != FE_OK) return FE_ERR_INVALID_ARGUMENT;
    if (!params || !norm) return fail(FE_ERR_INVALID_ARGUMENT, "null params or norm");
    for (int k = kMlpIn; k < 2 * kMlpIn; k++)
        if (!(norm[k] > 0.0f)) return fail(FE_ERR_INVALID_ARGUMENT, "feature standard deviations must be > 0");
    if (!(norm[2 * kMlpIn + 1] > 0.0f)) return fail(FE_ERR_INVALID_ARGUMENT, "target standard deviation must be > 0");
    DeviceGuard g(c->device);
    fe_status s = ensure_alloc(c);
    if (s != FE_OK) return s;
    FE_HIP(hipMemcpyAsync(c->ws.params[net], params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, c->stream),
           "weights upload");
    FE_HIP(hipMemcpyAsync(c->ws.norm[net], norm, sizeof(float) * kMlpNorm, hipMemcpyHostToDevice, c->stream),
           "norm upload");
    FE_HIP(hipMemsetAsync(c->ws.m1[net], 0, sizeof(float) * kMlpParams, c->stream), "adam reset");
    FE_HIP(hipMemsetAsync(c->ws.m2[net], 0, sizeof(float) * kMlpParams, c->stream), "adam reset");
    FE_HIP(fe::learn_pack(c, net), "weight packing");
    FE_HIP(hipStreamSynchronize(c->stream), "weights wait");
    c->ws.adam_t[net] = 0;
    c->ws.ready[net] = true;
    return FE_OK;
}

fe_status fe_learn_forward(fe_ctx* c, int net, const float* x, uint32_t n, float* y) {
    FE_CHECK_CTX(c);
    if (check_net(net) != FE_OK) return FE_ERR_INVALID_ARGUMENT;
    if (!fe::learn_ready(c, net)) return fail(FE_ERR_INVALID_ARGUMENT, "network not trained or set");
    if (n == 0) return FE_OK;
    if (!x || !y) return fail(FE_ERR_INVALID_ARGUMENT, "null x or y");
    DeviceGuard g(c->device);
    fe_status s = ensure_io(c, (size_t)n * (kMlpIn + 1));
    if (s != FE_OK) return s;
    float* dx = c->ws.io;
    float* dy = dx + (size_t)n * kMlpIn;
    std::vector<float> xt((size_t)n * kMlpIn);   // feature-major, as the inference kernel reads
    for (uint32_t b = 0; b < n; b++)
        for (int k = 0; k < kMlpIn; k++) xt[(size_t)k * n + b] = x[(size_t)b * kMlpIn + k];
    FE_HIP(hipMemcpyAsync(dx, xt.data(), sizeof(float) * xt.size(), hipMemcpyHostToDevice, c->stream),
           "input upload");
    FE_HIP(fe::learn_forward_raw(c, net, dx, n, dy), "mlp forward");
    FE_HIP(hipMemcpyAsync(y, dy, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream), "output download");
    FE_HIP(hipStreamSynchronize(c->stream), "forward wait");
    return FE_OK;
}

fe_status fe_learn_gradient(fe_ctx* c, int net, const float* x, const float* t, uint32_t n, float* grad,
                            float* sum_sq) {
    FE_CHECK_CTX(c);
    if (check_net(net) != FE_OK) return FE_ERR_INVALID_ARGUMENT;
    if (!fe::learn_ready(c, net)) return fail(FE_ERR_INVALID_ARGUMENT, "network not trained or set");
    if (n == 0 || n > fe::kMlpMaxBatch) return fail(FE_ERR_INVALID_ARGUMENT, "n must be in [1, 65536]");
    if (!x || !t || !grad) return fail(FE_ERR_INVALID_ARGUMENT, "null x, t or grad");
    DeviceGuard g(c->device);
    fe_status s = ensure_io(c, (size_t)n * (kMlpIn + 1) + kMlpParams + 1);
    if (s != FE_OK) return s;
    float* dx = c->ws.io;
    float* dt = dx + (size_t)n * kMlpIn;
    float* dg = dt + n;
    float* dl = dg + kMlpParams;
    FE_HIP(hipMemcpyAsync(dx, x, sizeof(float) * kMlpIn * n, hipMemcpyHostToDevice, c->stream), "input upload");
    FE_HIP(hipMemcpyAsync(dt, t, sizeof(float) * n, hipMemcpyHostToDevice, c->stream), "target upload");
    FE_HIP(fe::learn_step(c, net, dx, dt, n, 0, 0, 0.0f, 0u, dg, dl), "mlp gradient");
    FE_HIP(hipMemcpyAsync(grad, dg, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, c->stream), "grad download");
    float l = 0.0f;
    FE_HIP(hipMemcpyAsync(&l, dl, sizeof(float), hipMemcpyDeviceToHost, c->stream), "loss download");
    FE_HIP(hipStreamSynchronize(c->stream), "gradient wait");
    if (sum_sq) *sum_sq = l;
    return FE_OK;
}

fe_status fe_learn_features(fe_ctx* c, int net, float* features, float* prev_pressure) {
    FE_CHECK_CTX(c);
    if (check_net(net) != FE_OK) return FE_ERR_INVALID_ARGUMENT;
    if (!features) return fail(FE_ERR_INVALID_ARGUMENT, "null features");
    {
        DeviceGuard g(c->device);
        fe_status s = ensure_alloc(c);
        if (s != FE_OK) return s;
    }
    c->ws.force_features = true;
    fe_status s = net == FE_NET_DENSITY ? fe_density_source(c) : fe_divergence_source(c);
    c->ws.force_features = false;
    if (s != FE_OK) return s;
    DeviceGuard g(c->device);
    const size_t n = c->n;
    std::vector<float> f(n * kMlpIn), pp(n);
    std::vector<uint32_t> sched(n);
    FE_HIP(hipMemcpyAsync(f.data(), c->ws.feat, sizeof(float) * f.size(), hipMemcpyDeviceToHost, c->stream),
           "features download");
    FE_HIP(hipMemcpyAsync(pp.data(), c->ws.pprev, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream),
           "previous pressure download");
    FE_HIP(hipMemcpyAsync(sched.data(), c->sched[0], sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream),
           "schedule download");
    FE_HIP(hipStreamSynchronize(c->stream), "features wait");
    for (size_t t = 0; t < n; t++) {
        const size_t i = sched[t];
        for (int k = 0; k < kMlpIn; k++) features[i * kMlpIn + k] = f[(size_t)k * n + t];
        if (prev_pressure) prev_pressure[i] = pp[t];
    }
    return FE_OK;
}

fe_status fe_learn_predict(fe_ctx* c, int net) {
    FE_CHECK_CTX(c);
    if (check_net(net) != FE_OK) return FE_ERR_INVALID_ARGUMENT;
    if (!fe::learn_ready(c, net)) return fail(FE_ERR_INVALID_ARGUMENT, "network not trained or set");
    c->ws.force_features = true;
    fe_status s = net == FE_NET_DENSITY ? fe_density_source(c) : fe_divergence_source(c);
    c->ws.force_features = false;
    if (s != FE_OK) return s;
    DeviceGuard g(c->device);
    FE_HIP(fe::learn_infer(c, net), "warm-start inference");
    return FE_OK;
}

}  // extern "C"
