// fe_internal.h — device buffers, launch plumbing and kernel launchers shared by the
// C ABI (fe_api.cpp) and the kernels (fe_sort.hip, fe_sph.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "fluid_engine.h"

namespace fe {

constexpr uint32_t kSentinel = 0xFFFFFFFFu;
constexpr uint32_t kBlock = 256;          // threads per block of the per-particle kernels
constexpr uint32_t kMaxSolveIters = 1024; // residual slots per solve
constexpr uint32_t kStatsBlocks = 1024;   // blocks of the stats reduction
constexpr uint32_t kMaxNbr = 64;          // neighbour-list slots per particle
constexpr uint32_t kHashScheduleMax = 1u << 16;   // single scenes up to this size gather in hash order
constexpr uint32_t kTabRing = 4;          // pinned per-frame scene tables of a batch with moving walls
// schedule-ordered mirrors of the fields the gathers read from neighbours, and of the
// per-particle solver fields the Jacobi sweeps read for their own particle
constexpr uint32_t kMirrorPos = 1u, kMirrorVelB = 2u, kMirrorVelA = 4u, kMirrorDens = 8u, kMirrorPrho = 16u,
                   kMirrorPacc = 32u, kMirrorFac = 64u, kMirrorSrc = 128u, kMirrorPres = 256u,
                   kMirrorMrho = 512u, kMirrorAll = 1023u;
// Windows of the Jacobi sweeps: 64 consecutive schedule positions (one wave). The union of
// the window's listed neighbours is staged in LDS once per sweep, as at most kWinMaxRuns
// runs of consecutive schedule positions holding at most kWinMaxK entries in total.
constexpr uint32_t kWin = 64;
constexpr uint32_t kWinMaxK = 512;
constexpr uint32_t kWinRounds = kWinMaxK / kWin;   // 64-entry staging rounds
constexpr uint32_t kWinMaxRuns = 192;
constexpr uint32_t kWinFallback = 0xFFFFFFFFu;     // window meta: use the per-particle list

// Per-kernel event profiling (enabled by fe_profile_enable).
struct ProfileRecord {
    std::string name;
    hipEvent_t start, stop;
};
struct ProfileTotal {
    std::string name;
    uint64_t launches = 0;
    double total_ms = 0.0;
};

// Scratch of the one-sweep radix sort (fe_sort.hip): the digit counts of every pass over all
// keys (zero between sorts: a sort's last tile clears them), the look-back word of every
// (pass, tile of 2048 keys, digit), and the tile tickets of every pass followed by an error
// word that a look-back sets if it ever stops waiting.
constexpr uint32_t kSortMaxPasses = 4;
constexpr uint32_t kSortMaxKeys = (1u << 30) - 1u;   // counts share a word with a 2-bit state
constexpr uint32_t kSortTileKeys = 4096;             // keys per tile
constexpr uint32_t kSortGroupTiles = 32;             // tiles per look-back group
struct SortScratch {
    uint32_t* hist = nullptr;     // kSortMaxPasses x 256
    uint32_t* status = nullptr;   // sort_status_words(max_tiles): tile words, then group words
    uint32_t* ctrl = nullptr;     // kSortMaxPasses tickets, then the error word
    uint32_t max_tiles = 0;
};

// The learned warm start: two per-particle MLPs (density solve, divergence solve), each
// kMlpIn -> kMlpHidden -> kMlpHidden -> kMlpHidden -> 1 with GELU.
constexpr int kMlpIn = 12;
constexpr int kMlpHidden = 64;
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + 2 * (kMlpHidden * kMlpHidden + kMlpHidden) +
                           kMlpHidden + 1;
constexpr int kMlpNorm = 2 * kMlpIn + 2;        // feature means, feature stds, target mean, target std
constexpr uint32_t kMlpMaxBatch = 1u << 16;     // samples per training step
constexpr uint32_t kMlpSamplesPerBlock = 32;    // one wave of the training kernel
constexpr int kMlpPackedOps = 20;               // 16x32 weight tiles: 4 (layer 1) + 8 + 8
constexpr int kMlpPackedFloats = kMlpPackedOps * 64 * 4 + 288 + 2 * 2048;   // + the GELU table

struct LearnState {
    bool alloc = false;
    // per network (0 density, 1 divergence): FP32 parameters, Adam moments, normalisation
    float* params[2] = {nullptr, nullptr};
    float* m1[2] = {nullptr, nullptr};
    float* m2[2] = {nullptr, nullptr};
    float* norm[2] = {nullptr, nullptr};
    bool ready[2] = {false, false};
    uint32_t adam_t[2] = {0, 0};
    // inference copy of each network: BF16 weight operands in the matrix-core lane layout,
    // then biases, the output layer and the normalisation (kMlpPackedFloats floats)
    float* packed[2] = {nullptr, nullptr};
    // per particle: features of the coming solve (feature-major, feat[k n + t] for schedule
    // position t), the previous pressure of each particle by schedule position, and the last
    // solve's final pressures (hash order of that solve)
    float* feat = nullptr;
    float* pprev = nullptr;
    float* plast[2] = {nullptr, nullptr};
    bool plast_valid[2] = {false, false};
    // the source-term stage of solve feat_kind gathered features that are still current
    bool feat_fresh = false;
    int feat_kind = -1;
    bool force_features = false;
    // dataset (ring buffer per network) of (features, final pressure)
    float* data_x[2] = {nullptr, nullptr};
    float* data_t[2] = {nullptr, nullptr};
    uint32_t capacity = 0;
    uint32_t head[2] = {0, 0};
    uint32_t size[2] = {0, 0};
    uint32_t per_solve = 0;
    uint32_t salt = 0;
    // training scratch
    float* train_x = nullptr;   // standardised copy of a dataset
    float* train_t = nullptr;
    float* grads = nullptr;     // per-block partial gradients
    float* loss = nullptr;      // per-block partial losses
    float* io = nullptr;        // staging of fe_learn_forward / fe_learn_gradient
    size_t io_floats = 0;
};

// The density texture of the render hand-off and its scratch (allocated on first use).
struct SplatState {
    uint32_t* tex = nullptr;     // grid_res.x * .y * .z texels, x fastest
    uint32_t* count = nullptr;   // (grid_res + 1)^3 centre counts, then the 8^3 occupancy blocks
    size_t cells = 0, centres = 0;
    bool checked = false;        // the kernel's constant weights equal the host evaluation
    bool valid = false;          // a splat has run since the last reallocation
};

struct Launch;
struct Sph;

// Launch configurations the autotuner chooses (block = threads per block, items = particles
// per thread for the streaming kernels). Results never depend on them.
enum TuneId {
    kTuneDensityAlpha,
    kTuneViscosity,
    kTuneDensitySource,
    kTuneDivergenceSource,
    kTunePressureForce,
    kTunePressureUpdate,
    kTuneBuildList,
    kTunePressureIntegration,
    kTuneDivergenceIntegration,
    kTuneDensityAlphaDivergence,   // density_alpha + divergence source in one pass
    kTuneCount
};
struct TuneCfg {
    uint32_t block = 256, items = 1;
    float median_us = 0.0f, default_us = 0.0f;   // of the last fe_autotune
};
const char* tune_name(int id);
// Whether (block, items) is a configuration kernel `id` is built for.
bool tune_valid(int id, uint32_t block, uint32_t items);

}  // namespace fe

// The context: SOA particle state in HBM, laid out as GpuPhysicsData
// (src/entities/particle.rs:137-161), plus the internal neighbour-search arrays.
struct fe_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    fe_params params{};
    fe_options options{};
    uint32_t n = 0;      // particles (all scenes)
    uint32_t T = 0;      // hash-table entries = next_power_of_two(n) (the sum over scenes)
    uint32_t log2T = 0;  // bits of a table index: log2(T) rounded up

    // A batch of independent scenes (nscenes > 1): per-scene constants on the device, the
    // scene of each particle index, and the host copies of the layout.
    uint32_t nscenes = 1;
    fe::Sph* scene_tab = nullptr;
    uint8_t* sid = nullptr;
    std::vector<fe_params> scene_params;
    std::vector<uint32_t> scene_n, scene_T, scene_part0, scene_tab0;
    std::vector<uint32_t> scene_iters[2];       // sweeps of the last solve, per scene
    std::vector<uint64_t> scene_total_iters[2];
    uint32_t* scene_stats = nullptr;            // 3 per scene
    bool batch_graph = false;                   // replay the frame from a captured graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // one scene in graph mode: the frame's box is read from box_dev (copied before every
    // frame from a ring of pinned host slots) while live_box is set; the capture holds the
    // Morton key width it was made with
    float* box_dev = nullptr;
    float* box_ring = nullptr;                  // pinned, kTabRing x 8
    hipEvent_t box_event[fe::kTabRing] = {};
    uint32_t box_slot = 0;
    bool live_box = false;
    uint32_t graph_morton_bits = 0;

    // Moving walls (one per scene) applied at the start of every frame. A batch whose walls
    // move re-copies its scene constants each frame from a ring of pinned host tables.
    std::vector<fe_wall> walls;
    std::vector<uint8_t> wall_on;
    uint32_t walls_on = 0;
    fe::Sph* tab_ring = nullptr;                // pinned, kTabRing x nscenes
    hipEvent_t tab_event[fe::kTabRing] = {};
    uint32_t tab_slot = 0;

    fe::SplatState splat;

    // FE_BOUNDARY_PARTICLES: the wall samples' boundary masses and the geometry they were
    // computed for (recomputed when the box, r, h or rho0 change).
    float* walls_psi = nullptr;
    size_t walls_cap = 0;
    float walls_key[9] = {};
    bool walls_ready = false;
    // per schedule position: (sum psi grad W, sum psi W) over the samples, for the current
    // positions and samples (cleared by every neighbour search and position upload)
    float4* wall_g = nullptr;
    bool wall_g_fresh = false;

    // GpuPhysicsData fields
    float4* pos_a = nullptr;
    float4* pos_b = nullptr;
    float4* vel_a = nullptr;
    float4* vel_b = nullptr;
    float4* colors = nullptr;
    float4* pacc = nullptr;
    float* dens = nullptr;
    float* fac = nullptr;
    float* src = nullptr;
    float* pres = nullptr;

    // grid_entries as two arrays (hash, index) with a ping-pong partner for the sort;
    // keys[0]/vals[0] always hold the current entries after a sort.
    uint32_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    // grid_start (x) and one-past-the-end of each bucket (y); x = 0xFFFFFFFF when empty.
    uint2* range = nullptr;
    bool ends_dirty = false;   // y must be rebuilt from x and the keys (after uploads)
    // neighbour list of the current (pos_b, range): one row of kMaxNbr entries per work-item
    // of the gather schedule, and the pair count of each (which may exceed kMaxNbr)
    uint32_t* nbr = nullptr;
    uint32_t* nbr_count = nullptr;
    bool list_valid = false;
    // work schedule of the solver gathers: particle indices ordered by the Morton code of
    // their cell (sched[0] holds it; [1] is the sort's partner), and the Morton keys
    uint32_t* sched[2] = {nullptr, nullptr};
    uint32_t* mkeys[2] = {nullptr, nullptr};
    uint32_t morton_bits = 30;   // bits of the Morton key (3 x bits per axis)
    uint32_t exact_fast = 0;     // exact arithmetic shortcuts valid for params (kFast* bits)
    uint32_t scene_bits = 0;     // bits of the scene number above them (a batch)
    uint32_t* inv = nullptr;     // inverse schedule: hash index -> schedule position
    float4* pos_m = nullptr;     // position_b by schedule position
    float4* vb_m = nullptr;      // velocity_b, with mass / density in w (kMirrorMrho)
    float4* va_m = nullptr;      // velocity_a
    float* dens_m = nullptr;     // densities
    // pos_m.w holds pressure / density^2 (kMirrorPrho); pa_m = (position_b xyz, pressure
    // acceleration x) and pa2_m = (pressure acceleration y, z) (kMirrorPacc), so a sweep
    // gathers one 16-byte record (pressure_force) or 16 + 8 bytes (pressure_update) per pair
    float4* pa_m = nullptr;
    float2* pa2_m = nullptr;
    float* fac_m = nullptr;      // factors
    float* src_m = nullptr;      // source_terms
    float* pres_m = nullptr;     // pressures
    uint32_t mirrors = 0;        // kMirror* bits that are current
    // Inside a solve the sweeps leave the hash-ordered pressures / pressure accelerations to
    // their mirrors; one write-back after the solve (or before any mirror refresh) stores them.
    bool defer_sweep_writes = false;
    bool pending_pres = false, pending_pacc = false;
    // sweep windows of the current list (valid with it): per window the staged entry count K
    // and run count R (K | R << 16, or kWinFallback), per staging round the 64-bit mask of
    // runs that start inside the round and the run holding its first entry, per run the
    // offset from staging slot to schedule position, and the list as 16-bit staging slots
    // (tiles of 16 work-items x 8 slots)
    uint64_t list_gen = 0;           // lists built so far
    uint64_t win_gen = ~0ull;        // the list the window tables were built from
    uint32_t* win_meta = nullptr;
    uint4* win_round = nullptr;
    uint32_t* win_delta = nullptr;
    uint16_t* nbr16 = nullptr;
    // The neighbour structure and densities still describe position_a (nothing moved the
    // particles since the last neighbour search + density_alpha): the next frame-start
    // search would reproduce them bit for bit, so only velocity_b is re-gathered.
    bool ns_fresh = false;
    bool perm_valid = false;   // vals / inv hold the permutation and schedule of the last search
    bool da_fresh = false;
    uint32_t ns_sort = 0;
    float ns_h = 0.0f;

    fe::SortScratch sort;
    fe::LearnState ws;
    fe::TuneCfg tune[fe::kTuneCount];

    // stats_buffer and the convergence residuals (one slot per sweep, two solves)
    uint32_t* stats = nullptr;
    unsigned long long* resid = nullptr;   // 2 * kMaxSolveIters
    uint32_t* stats_partial = nullptr;     // 3 x kStatsBlocks per-block stats partials
    uint32_t* host_stats = nullptr;         // pinned [3]
    unsigned long long* host_resid = nullptr;  // pinned 2 * kMaxSolveIters
    hipEvent_t stats_event = nullptr;
    bool stats_pending = false;
    bool have_prev_stats = false;
    uint32_t prev_stats[3] = {0, 0, 0};
    uint64_t stats_frame = 0;      // frames run when the pending stats were enqueued
    bool stats_health = false;     // the pending copy holds the health words
    bool sort_failed = false;      // a radix sort's look-back gave up (stats word 6)
    // the watchdog (fe_set_watchdog): health words in stats[4..5], the last report
    bool watchdog = false, watchdog_halt = false;
    float speed_limit = 0.0f;
    fe_health health{};

    // exit reports of the Jacobi solves (fe_solve_report_enable): r = s - Ap by schedule
    // position, and per solve kind the sums of k_residual_reduce
    bool report_on = false;
    float* report_err = nullptr;
    unsigned long long* report = nullptr;   // 2 x 4
    uint32_t report_solves[2] = {0, 0};

    uint32_t last_iters[2] = {0, 0};
    bool last_iters_pending[2] = {false, false};
    uint64_t total_iters[2] = {0, 0};
    uint64_t frames = 0;

    // transfer staging for the interleaved entry layout
    uint32_t* staging = nullptr;   // 2 * T

    // timing
    hipEvent_t timer_start = nullptr, timer_stop = nullptr;
    bool profiling = false;
    std::vector<fe::ProfileRecord> pending;
    std::vector<fe::ProfileTotal> totals;
};

namespace fe {

// RAII helper: brackets one kernel launch with profiling events when enabled.
struct Launch {
    fe_ctx* c;
    ProfileRecord rec;
    bool on;
    Launch(fe_ctx* ctx, const char* name);
    ~Launch();
};

// ---- launchers (fe_sort.hip) ----
// Stable LSD radix sort of n (key, value) pairs on bits [0, key_bits) of the key.
// The result is in keys[0]/vals[0] on return (buffers are swapped when needed). runs: the
// keys arrive almost sorted, so the digit counts add runs of equal digits (any order sorts
// correctly either way; this only picks the faster count).
hipError_t radix_sort(fe_ctx* c, uint32_t n, uint32_t key_bits, bool runs);
// The same sort on an explicit ping-pong pair of buffers (result in keys[0]/vals[0]).
hipError_t radix_sort_pairs(fe_ctx* c, uint32_t* keys[2], uint32_t* vals[2], uint32_t n, uint32_t key_bits,
                            bool runs);
// The bitonic network of bitonic_sort.comp over all T entries, in place in keys[0]/vals[0].
hipError_t bitonic_sort(fe_ctx* c);
// Stable LSD radix sort of n pairs on bits [0, key_bits) on `stream` with explicit scratch;
// result in keys[0]/vals[0] (buffers swapped as needed). prof (may be null) records launches;
// runs as for radix_sort.
hipError_t radix_sort_core(const SortScratch& s, hipStream_t stream, uint32_t* keys[2], uint32_t* vals[2],
                           uint32_t n, uint32_t key_bits, fe_ctx* prof, bool runs);
// Tiles of a sort of max_n keys, and the look-back words a scratch for that many tiles holds.
uint32_t sort_max_tiles(uint32_t max_n);
size_t sort_status_words(uint32_t max_tiles);
hipError_t sort_scratch_alloc(SortScratch* s, uint32_t max_n, hipStream_t stream);
// Zero the digit counts, tickets and error word (one small kernel, no memset).
hipError_t sort_scratch_clear(const SortScratch& s, hipStream_t stream);
void sort_scratch_free(SortScratch* s);

// ---- launchers (fe_gather.hip) ----
// Inverse schedule and the position / velocity_b / velocity_a mirrors after the schedule sort.
hipError_t launch_schedule_mirrors(fe_ctx* c);
// velocity_b = velocity_a gathered through the current sort (reorder.comp's velocity half).
hipError_t launch_reorder_velocity(fe_ctx* c);

// ---- launchers (fe_sph.hip) ----
// The exact arithmetic shortcuts (kFast* bits) valid for a scene with parameters p: the
// constant-divisor quotient is checked on the device once per smoothing radius (cached per
// process), the operand box of the gradient division from h on the host.
hipError_t exact_math_flags(fe_ctx* c, const fe_params& p, uint32_t* flags);
// fe_check_exact_arith on stream st with a 3-counter device scratch.
hipError_t check_exact_arith(hipStream_t st, float h, unsigned long long samples, unsigned long long* dev_bad,
                             unsigned long long counts[6]);
hipError_t launch_hash(fe_ctx* c);
// Bucket ranges, position_b / velocity_b, and the Morton keys of position_b's cells with the
// identity schedule payload (mkeys[0], sched[0]).
hipError_t launch_offsets_reorder(fe_ctx* c);
hipError_t launch_iota(fe_ctx* c, uint32_t* p, uint32_t count);
hipError_t launch_fill_ranges(fe_ctx* c);
hipError_t launch_offsets_only(fe_ctx* c);
hipError_t launch_rebuild_ends(fe_ctx* c);
hipError_t launch_density_alpha(fe_ctx* c);   // build_list, then the density_alpha gather
hipError_t launch_build_list(fe_ctx* c);
hipError_t launch_density_alpha_only(fe_ctx* c);
// build_list, then density_alpha and divergence_source (with the warm-start features when
// wanted) in one gather: the post-integrate pair of a frame
hipError_t launch_density_alpha_divergence_source(fe_ctx* c);
// the same pass over the current list, without the build and the features (the autotuner's)
hipError_t launch_density_alpha_divergence_only(fe_ctx* c);
hipError_t launch_viscosity(fe_ctx* c);
hipError_t launch_density_source(fe_ctx* c);
// Jacobi sweep kernels. prev: the previous sweep's residual slot (nullptr: no early exit);
// thr: the threshold of one scene; kind selects a batch's per-scene threshold (0 density,
// 1 divergence). A batch's slots are kMaxSolveIters apart per scene.
hipError_t launch_pressure_force(fe_ctx* c, const unsigned long long* prev, unsigned long long thr, uint32_t kind);
hipError_t launch_pressure_update(fe_ctx* c, unsigned long long* out, const unsigned long long* prev,
                                  unsigned long long thr, uint32_t kind);
// The hash-ordered pressures / pressure accelerations from their mirrors after sweeps that
// deferred them (no-op when nothing is pending).
hipError_t flush_sweep_writes(fe_ctx* c);
// The exit report of the solve that just ended (kind 0 density, 1 divergence): r = s - Ap of
// its last sweep, reduced into report[4 kind .. 4 kind + 3].
hipError_t launch_residual_report(fe_ctx* c, int kind);
hipError_t launch_pressure_integration(fe_ctx* c);
hipError_t launch_divergence_source(fe_ctx* c);
hipError_t launch_divergence_integration(fe_ctx* c);
hipError_t launch_stats(fe_ctx* c);
hipError_t launch_fill_u64(fe_ctx* c, unsigned long long* p, uint32_t count);
hipError_t launch_pack_entries(fe_ctx* c, uint32_t* out);
hipError_t launch_unpack_entries(fe_ctx* c, const uint32_t* in);
hipError_t launch_pack_start(fe_ctx* c, uint32_t* out);
hipError_t launch_unpack_start(fe_ctx* c, const uint32_t* in);

// ---- launchers (fe_learn.hip) ----
// Previous pressure of every particle by schedule position (ws.pprev) from ws.plast[which]
// through the sort.
hipError_t learn_prev(fe_ctx* c, int which);
// Inference copy of network `which` (ws.packed) from its parameters and normalisation.
hipError_t learn_pack(fe_ctx* c, int which);
// Predicted initial pressures of solve `which` into pressures and the p / rho^2 mirror.
hipError_t learn_infer(fe_ctx* c, int which);
// Append ws.per_solve samples (features, current pressures) to dataset `which`.
hipError_t learn_collect(fe_ctx* c, int which);
// Raw network output for standardised inputs x (feature-major, x[k n + b], device) into y.
hipError_t learn_forward_raw(fe_ctx* c, int which, const float* x, uint32_t n, float* y);
// (fe_learn_api.cpp) Whether the source-term stage of solve `kind` also gathers the
// features, and the record that it did.
bool learn_features_wanted(const fe_ctx* c, int kind);
void learn_features_done(fe_ctx* c, int kind);
// One training step on `batch` samples of (x, t): sample b is row b when sample_mod == 0,
// else row hash(seed, b) % sample_mod. Sums the gradient into grad_out (if not null) and
// the summed squared error into loss_out (if not null); applies Adam when apply != 0.
hipError_t learn_step(fe_ctx* c, int which, const float* x, const float* t, uint32_t batch, uint32_t sample_mod,
                      uint32_t seed, float lr, uint32_t apply, float* grad_out, float* loss_out);

// ---- wall boundary particles (fe_gather.hip) ----
// Whether the wall samples (when FE_BOUNDARY_PARTICLES is on) were computed for the current
// box, r, h and rho0: the densities of the last density_alpha depend on them.
bool walls_current(const fe_ctx* c);

// ---- launchers (fe_splat.hip) ----
// Clear and rebuild the density texture from position_a with the current box and grid_res.
hipError_t launch_splat_density(fe_ctx* c);
// Contribution of a texel at squared offset d2 (0..63) from a particle's centre texel.
void splat_weights(uint32_t wt[64]);
void splat_free(fe_ctx* c);

}  // namespace fe
