/*
 * fluid_engine.h — C ABI of the MI355X-native DFSPH physics frame.
 *
 * Drop-in boundary for the physics half of Nikita-Lysiuk/Fluid-Engine:
 *   - the `ComputeStep` operator trait (src/renderer/pipelines/mod.rs:52-76) and the
 *     `ComputePipelines` aggregate (:78-124) become one `fe_ctx` with one enqueue call
 *     per reference ComputeStep;
 *   - `GpuPhysicsData` (src/entities/particle.rs:137-308) becomes the device buffers the
 *     context owns, reachable through fe_upload / fe_download;
 *   - `SimulationParams` (src/entities/particle.rs:310-331) is `fe_params`, byte-identical
 *     (112 bytes, std140, 16-byte aligned), so a Rust caller can pass `&SimulationParams`;
 *   - `Renderer::step` (src/renderer/mod.rs:227-364) physics part is `fe_step`.
 *
 * Conventions: every call returns an fe_status (0 = FE_OK); no exception crosses the ABI;
 * the message of the last failure on the calling thread is in fe_last_error().
 * All work is enqueued on the context's own HIP stream (default priority) in call order,
 * which replaces vulkano's automatic barriers; calls on one context are not thread-safe,
 * separate contexts may be driven from separate threads. fe_sync() waits for the stream.
 * Host arrays belong to the caller; uploads and downloads copy.
 */
#ifndef FLUID_ENGINE_H
#define FLUID_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------------------- */
typedef int32_t fe_status;
#define FE_OK 0
#define FE_ERR_INVALID_ARGUMENT 1
#define FE_ERR_OUT_OF_MEMORY 2
#define FE_ERR_DEVICE 3
#define FE_ERR_UNSUPPORTED 4
#define FE_ERR_SIZE_MISMATCH 5
#define FE_ERR_UNSTABLE 6   /* the watchdog halted a frame (fe_set_watchdog) */

/* ---- parameter block ---------------------------------------------------------------------
 * Byte layout of `SimulationParams` (src/entities/particle.rs:310-331) and of the std140
 * uniform block in shaders/include/common.glsl:11-28. Offsets: 0 particle_radius,
 * 4 particle_mass, 8 smoothing_radius, 12 target_density, 16 viscosity, 20 relax_factor,
 * 24 dt, 28 density_solver_iterations, 32 divergence_solver_iterations, 36 pad[3],
 * 48 gravity, 64 box_min, 80 box_max, 96 grid_res. */
typedef struct fe_params {
    float particle_radius;
    float particle_mass;
    float smoothing_radius;
    float target_density;
    float viscosity;
    float relax_factor;
    float dt;
    uint32_t density_solver_iterations;
    uint32_t divergence_solver_iterations;
    float _pad[3];
    float gravity[4];
    float box_min[4];
    float box_max[4];
    int32_t grid_res[4];
} fe_params;

/* Sort switch: `SortAlgorithm` (src/renderer/pipelines/sorter.rs:19-23). */
#define FE_SORT_BITONIC 0u
#define FE_SORT_RADIX 1u

/* Pressure-solver termination. FIXED runs density_solver_iterations /
 * divergence_solver_iterations Jacobi sweeps as the reference does
 * (src/renderer/mod.rs:301-306, :327-332). CONVERGE runs sweeps until the
 * compression-only mean density error is <= density_tol (resp. divergence_tol),
 * at least 1 and at most max_iterations sweeps; this mode has no reference counterpart. */
#define FE_SOLVER_FIXED 0u
#define FE_SOLVER_CONVERGE 1u

/* Initial pressure of the two Jacobi solves. REFERENCE: zero before the density solve
 * (shaders/compute/density_source_term.comp:73) and the density solve's last pressure
 * before the divergence solve (pressures is never reset, src/renderer/mod.rs:323-332).
 * LEARNED: the per-particle MLP warm start (fe_learn_* below). */
#define FE_WARM_START_REFERENCE 0u
#define FE_WARM_START_LEARNED 1u

/* Velocity order of the divergence phase. REFERENCE reproduces the reference: after the
 * post-integrate reorder, divergence_source_term.comp:47 and divergence_integration.comp:49
 * read and correct velocity_a, still in the previous sort order, at the new order's indices,
 * and the next substep's viscosity reads velocity_b from before the correction
 * (src/renderer/mod.rs:289-336). CONSISTENT (no reference counterpart, not a parity mode):
 * the divergence source uses each particle's own velocity (velocity_b, reordered with the
 * positions), and the divergence integration writes the corrected velocity and the colour back
 * to the same particle in velocity_a / colors and to velocity_b. */
#define FE_DIVERGENCE_REFERENCE 0u
#define FE_DIVERGENCE_CONSISTENT 1u

/* Boundary handling. CLAMP is the reference's: the box acts only through the collision clamp of
 * pressure_integration.comp:36-73. PARTICLES (no reference counterpart, not a parity mode) adds
 * static boundary samples on the six box faces (Akinci et al. 2012, as in DFSPH): a square
 * lattice of spacing 2 particle_radius per face, sample b with boundary mass
 * psi_b = rho0 / sum_k W(x_b - x_k) over the samples. A fluid particle within h of a face adds
 * psi_b W to its density, psi_b grad W to the gradient sum of alpha, psi_b v_i . grad W to both
 * source terms' velocity divergence, psi_b (p_i / rho_i^2) grad W to its pressure acceleration
 * and psi_b a_i . grad W to A p (samples at rest, pressure mirrored). The clamp stays as a
 * guard. Single-scene contexts; the Jacobi sweeps use the per-particle lists. */
#define FE_BOUNDARY_CLAMP 0u
#define FE_BOUNDARY_PARTICLES 1u

/* Residual of a Jacobi sweep that FE_SOLVER_CONVERGE compares with the tolerance (no
 * reference counterpart). COMPRESSION: mean over particles of max(0, -(s - Ap)) dt / rho0, the
 * DFSPH compression-only error (under-dense particles at the free surface are not errors).
 * TWO_SIDED: the same, except that a particle whose pressure entering the sweep is > 0 counts
 * |s - Ap| dt / rho0, so a start that over-predicts pressure is not taken as converged; a particle
 * with zero pressure still counts only compression (the clamp p >= 0 cannot push it further). */
#define FE_RESIDUAL_COMPRESSION 0u
#define FE_RESIDUAL_TWO_SIDED 1u

/* Run options. The first five mirror the runtime flags of `AppUI`
 * (src/renderer/ui.rs:12-49): sort_algorithm, use_cfl, use_solver_error_threshold,
 * density_error_pct, divergence_error_pct, whose host logic is in Renderer::step
 * (src/renderer/mod.rs:231-267). */
typedef struct fe_options {
    uint32_t sort_algorithm;             /* FE_SORT_RADIX (app default, ui.rs:36) */
    uint32_t use_cfl;                    /* 0 */
    uint32_t use_solver_error_threshold; /* 0: early-exit to 2/1 sweeps, mod.rs:251-265 */
    float density_error_pct;             /* 0.1 */
    float divergence_error_pct;          /* 0.5 */
    uint32_t solver_mode;                /* FE_SOLVER_FIXED */
    float density_tol;                   /* 1e-3 (relative) for FE_SOLVER_CONVERGE */
    float divergence_tol;                /* 1e-2 (relative) for FE_SOLVER_CONVERGE */
    uint32_t max_iterations;             /* 100 */
    uint32_t warm_start;                 /* FE_WARM_START_REFERENCE */
    uint32_t divergence_order;           /* FE_DIVERGENCE_REFERENCE */
    uint32_t boundary;                   /* FE_BOUNDARY_CLAMP */
    uint32_t residual;                   /* FE_RESIDUAL_COMPRESSION */
    uint32_t reserved[3];
} fe_options;

/* Device buffers of `GpuPhysicsData` (src/entities/particle.rs:137-161).
 * Element size per particle: vec4 fields 16 B (w as the reference writes it),
 * scalar fields 4 B; GRID_ENTRIES is T x {uint32 hash, uint32 index} (8 B) and
 * GRID_START is T x uint32, with T = next_power_of_two(n) (particle.rs:255). */
#define FE_FIELD_POSITION_A 0
#define FE_FIELD_POSITION_B 1
#define FE_FIELD_VELOCITY_A 2
#define FE_FIELD_VELOCITY_B 3
#define FE_FIELD_COLORS 4
#define FE_FIELD_DENSITIES 5
#define FE_FIELD_FACTORS 6
#define FE_FIELD_SOURCE_TERMS 7
#define FE_FIELD_PRESSURES 8
#define FE_FIELD_PRESSURE_ACCELERATIONS 9
#define FE_FIELD_GRID_ENTRIES 10
#define FE_FIELD_GRID_START 11
#define FE_FIELD_COUNT 12

typedef struct fe_ctx fe_ctx;

/* Library version string and the last error message of the calling thread. */
const char* fe_version(void);
const char* fe_last_error(void);

/* Replaces GpuPhysicsData::new (particle.rs:164-299) + ComputePipelines::new
 * (pipelines/mod.rs:93-123) + the prepare() calls (src/renderer/mod.rs:137-192):
 * xyz is n x 3 floats; position_a = (xyz, 1), velocities and pressures zero.
 * Uses HIP device `device` (the current device if negative). */
fe_status fe_create(const fe_params* params, const float* xyz, uint32_t n, int device,
                    fe_ctx** out);
fe_status fe_destroy(fe_ctx* ctx);

/* GpuSceneResources::sync_with_scene (src/renderer/resources.rs:85-92). */
fe_status fe_set_params(fe_ctx* ctx, const fe_params* params);
fe_status fe_get_params(const fe_ctx* ctx, fe_params* out);
fe_status fe_set_options(fe_ctx* ctx, const fe_options* options);
fe_status fe_get_options(const fe_ctx* ctx, fe_options* out);
/* Fills `out` with the defaults listed in fe_options. */
fe_status fe_default_options(fe_options* out);

fe_status fe_num_particles(const fe_ctx* ctx, uint32_t* n, uint32_t* table_size);

/* Copy a whole field between host and device (bytes must equal the field size). */
fe_status fe_upload(fe_ctx* ctx, int field, const void* src, size_t bytes);
fe_status fe_download(fe_ctx* ctx, int field, void* dst, size_t bytes);
/* Device address of a field for zero-copy interop (valid for the context's lifetime).
 * GRID_ENTRIES and GRID_START are stored in a different internal layout and are not
 * exposed this way (FE_ERR_UNSUPPORTED). */
fe_status fe_device_ptr(fe_ctx* ctx, int field, void** out);

/* ---- one enqueue per reference ComputeStep::execute --------------------------------------- */
/* NeighborSearch::execute (neighbor_search.rs:176-231): hash, sort, offsets, reorder. */
fe_status fe_neighbor_search(fe_ctx* ctx);
/* The sort stage alone on GRID_ENTRIES (RadixSorter::execute sorter.rs:325-455 or
 * GpuSorter::execute sorter.rs:73-115, by options.sort_algorithm). */
fe_status fe_sort(fe_ctx* ctx);
/* fe_sort's radix sort on caller-owned device memory, allocating nothing. No counterpart in
 * the reference: for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). `entries` is a device
 * pointer to n interleaved {uint32 hash, uint32 index} pairs (the GRID_ENTRIES layout); they
 * are sorted in place, stably by all 32 bits of the hash, on HIP stream `stream` (a
 * hipStream_t; NULL is the null stream) of HIP device `device` (the current one if negative),
 * asynchronously. The result equals fe_sort's on the same pairs. n <= FE_SORT_NOALLOC_MAX_PAIRS,
 * else FE_ERR_INVALID_ARGUMENT and nothing runs. Every buffer beyond the entries lives in one
 * fixed device array of FE_SORT_NOALLOC_SCRATCH_BYTES per GPU, which that GPU holds until the
 * process ends; every call on a GPU shares it, so calls that overlap on one GPU (other streams,
 * other threads) race. */
#define FE_SORT_NOALLOC_MAX_PAIRS 4194304u
#define FE_SORT_NOALLOC_SCRATCH_BYTES 71438592u
fe_status fe_sort_pairs_noalloc(int device, uint32_t* entries, uint32_t n, void* stream);
fe_status fe_density_alpha(fe_ctx* ctx);            /* density_and_alpha.comp */
fe_status fe_viscosity(fe_ctx* ctx);                /* viscosity.comp */
fe_status fe_density_source(fe_ctx* ctx);           /* density_source_term.comp */
fe_status fe_pressure_force(fe_ctx* ctx);           /* pressure_force.comp */
fe_status fe_pressure_update(fe_ctx* ctx);          /* pressure_update.comp */
fe_status fe_pressure_integration(fe_ctx* ctx);     /* pressure_integration.comp */
fe_status fe_divergence_source(fe_ctx* ctx);        /* divergence_source_term.comp */
fe_status fe_divergence_integration(fe_ctx* ctx);   /* divergence_integration.comp */
fe_status fe_stats(fe_ctx* ctx);                    /* fill 0 + stats.comp */

/* Jacobi solves: density (after fe_density_source) and divergence (after
 * fe_divergence_source), FIXED or CONVERGE per the options. `iterations` (may be NULL)
 * receives the number of sweeps that did work; in CONVERGE mode this waits for the
 * stream. */
fe_status fe_solve_density(fe_ctx* ctx, uint32_t* iterations);
fe_status fe_solve_divergence(fe_ctx* ctx, uint32_t* iterations);
/* Exit residuals of the Jacobi solves (no reference counterpart; one scene). While enabled,
 * each solve ends with a report of its last sweep's residual r = s - Ap (the residual the exit
 * test saw), as the relative rate r dt / rho0 that the tolerances are stated in.
 * fe_solve_report gives, for solve kind 0 (density) or 1 (divergence), the solves reported
 * since the last enable and, averaged over them: out[0] mean |r|, out[1] the largest |r| of any
 * particle, out[2] mean max(0, r) (the expansion part), out[3] the expansion part summed over
 * the particles whose final pressure is > 0, over N. Enabling resets the counts; each report
 * adds one gather and one reduction to the solve. */
fe_status fe_solve_report_enable(fe_ctx* ctx, int enable);
fe_status fe_solve_report(fe_ctx* ctx, int kind, uint32_t* solves, double out[4]);

/* The physics frame of Renderer::step (src/renderer/mod.rs:227-364): frame-start neighbour
 * search and density+alpha, substeps while t < max_dt, then stats. With max_dt == dt it is
 * exactly one substep. Host flags (CFL, early exit) read the previous frame's stats. */
fe_status fe_step(fe_ctx* ctx, float max_dt);
/* Run `count` frames back to back without host synchronisation (FE_SOLVER_FIXED and no
 * host flags), or with the per-frame control when those are on. */
fe_status fe_run(fe_ctx* ctx, float max_dt, uint32_t count);

/* stats_buffer (particle.rs:270-281): [max_speed_bits, density_error_fp, divergence_error_fp]. */
fe_status fe_read_stats(fe_ctx* ctx, uint32_t out[3]);
/* Iteration counts of the last frame's density / divergence solves. */
fe_status fe_last_iterations(fe_ctx* ctx, uint32_t* density, uint32_t* divergence);
/* Cumulative sweep counts since creation or the last reset. */
fe_status fe_iteration_totals(fe_ctx* ctx, uint64_t* density, uint64_t* divergence,
                              uint64_t* frames);
fe_status fe_reset_iteration_totals(fe_ctx* ctx);

fe_status fe_sync(fe_ctx* ctx);

/* ---- timing on the context's stream --------------------------------------------------------
 * fe_timer_start/stop bracket enqueued work with HIP events on the context's stream;
 * fe_timer_stop waits and returns the elapsed milliseconds. */
fe_status fe_timer_start(fe_ctx* ctx);
fe_status fe_timer_stop(fe_ctx* ctx, float* ms);
/* Per-kernel profiling: when enabled, every kernel launch of the context is bracketed by
 * events and its duration accumulated under its kernel name. */
fe_status fe_profile_enable(fe_ctx* ctx, int enable);
/* Number of profiled kernel names. */
fe_status fe_profile_count(fe_ctx* ctx, uint32_t* count);
/* Entry i: name (NUL-terminated, truncated to name_cap), launch count, total ms. Waits. */
fe_status fe_profile_entry(fe_ctx* ctx, uint32_t i, char* name, size_t name_cap,
                           uint64_t* launches, double* total_ms);
fe_status fe_profile_reset(fe_ctx* ctx);

/* ---- host frame control (src/renderer/mod.rs:231-267) -------------------------------------
 * Pure host function: from the previous frame's stats decide the next dt (CFL) and the two
 * sweep counts (early exit). Exposed for tests. */
fe_status fe_frame_control(const uint32_t stats[3], uint32_t n, const fe_params* params,
                           const fe_options* options, float* dt_out,
                           uint32_t* density_iters, uint32_t* divergence_iters);

/* ---- render hand-off (src/renderer/mod.rs:344-355, :385-391) ----------------------------------
 * The density texture of the raymarching view, `GpuSceneResources::density_texture`
 * (src/renderer/resources.rs:55-69): R32Uint, grid_res.x x grid_res.y x grid_res.z, held by
 * the context as uint32 with x fastest, texel (x, y, z) at (z * grid_res.y + y) * grid_res.x
 * + x. fe_splat_density enqueues the clear to 0 (mod.rs:386-388) and splat_density.comp
 * (shaders/compute/splat_density.comp:11-51) over position_a with the current parameter
 * block's box and grid_res (1..2048 per axis, (x+1)(y+1)(z+1) <= 2^31): every particle with
 * uvw = (pos - box_min) / (box_max - box_min) in [0, 1]^3 adds uint(200 (1 - (d / 8)^2)^3) to
 * each texel at distance d < 8 texels from its centre texel ivec3(uvw * grid_res).
 * Single-scene contexts only. */
fe_status fe_splat_density(fe_ctx* ctx);
/* Device address and size of the texture of the last fe_splat_density (zero-copy interop;
 * valid until the next splat with another grid_res). */
fe_status fe_density_texture(fe_ctx* ctx, void** device_ptr, size_t* bytes);
/* Copy the texture to host memory (bytes must equal its size). Waits for the stream. */
fe_status fe_download_density_texture(fe_ctx* ctx, uint32_t* dst, size_t bytes);
/* copy_buffer of a particle field into caller device memory, enqueued on the context's
 * stream (mod.rs:347-355 copies position_a and colors into the render buffers of the next
 * frame in flight). bytes must equal the field's size; grid fields are refused. */
fe_status fe_copy_field_device(fe_ctx* ctx, int field, void* dst_device, size_t bytes);

/* ---- moving wall: `CollisionBox` (src/entities/collision.rs:11-35) ---------------------------
 * The simulation box; its min.x wall oscillates about base_min_x. Each reference frame first
 * copies the box into the parameter block in the UI pass (src/renderer/ui.rs:125-126:
 * box_min = (min, 0), box_max = (max, 0)), runs the physics with it, and advances the wall
 * by max_dt after uploading the parameters (src/renderer/mod.rs:269-270). Frame k therefore
 * runs with the wall at time t_k, the float sum of the k earlier max_dt (t_0 = 0 gives
 * min.x = base_min_x as set by fe_wall_init). The scene of the app sets amplitude 0.3 and
 * frequency 0.5 (src/core/scene.rs:19-23). */
typedef struct fe_wall {
    float min[3];
    float max[3];
    float wave_amplitude;   /* 0 after fe_wall_init (collision.rs:24) */
    float wave_frequency;   /* 1 after fe_wall_init (collision.rs:25) */
    float time;             /* seconds advanced so far (private in the reference) */
    float base_min_x;       /* min[0] at fe_wall_init (private in the reference) */
} fe_wall;
/* CollisionBox::new (collision.rs:20-29). Pure host function. */
fe_status fe_wall_init(fe_wall* wall, const float min[3], const float max[3]);
/* CollisionBox::update (collision.rs:31-35): time += dt; min[0] = base_min_x +
 * wave_amplitude * sin(time * wave_frequency * 2 pi), each operation in float, sin being the
 * C library's sinf (which Rust's f32::sin calls). Pure host function. */
fe_status fe_wall_update(fe_wall* wall, float dt);
/* Attach a copy of `wall` to a single-scene context (NULL detaches; every field must be
 * finite). While one is attached, each fe_step / fe_run frame writes the wall's box into the
 * parameters before the frame (replacing the box fe_set_params gave) and advances the wall by
 * the frame's max_dt after it. */
fe_status fe_set_wall(fe_ctx* ctx, const fe_wall* wall);
/* The attached wall as advanced by the frames run since; *attached (may be NULL) is 0 and
 * *out zeroed when none is attached. */
fe_status fe_get_wall(const fe_ctx* ctx, fe_wall* out, int* attached);

/* ---- launch-configuration autotuner (no reference counterpart) --------------------------------
 * Tunable kernels: the seven list gathers (block 64..1024), build_list (block 64, 128, 256;
 * items 1 = 64 schedule positions per wave, 8 = 8 per wave, for small scenes) and the two
 * integrations (block 64..1024 x 1, 2 or 4 particles per thread). In fe_step:
 *   density_alpha                    the frame-start density pass (when the last search
 *                                    cannot be reused) and fe_density_alpha
 *   density_alpha_divergence_source  the post-integrate density + divergence-source pass
 *   divergence_source                fe_divergence_source only
 *   viscosity, density_source, pressure_force, pressure_update, build_list and the two
 *   integrations                     every launch of that kernel. 
 * For the seven list gathers, items selects the variant: 1 gathers each listed pair from the
 * schedule-ordered mirrors with 4 neighbour records in flight per work-item; 4 does the same
 * with 16 in flight (block 64, 128 or 256; single-scene contexts without wall samples, a
 * batch or the wall mode runs variant 1), which shortens a small scene's chain of dependent
 * loads; for the two Jacobi sweeps, 2 stages each 64-particle window's neighbour union in LDS
 * and walks 16-bit lists (block 64, 128 or 256 = 1, 2 or 4 windows per block). For every list
 * gather, 8 runs four lanes per particle (block 64, 128 or 256; single-scene contexts without
 * wall samples, otherwise variant 1; the warm-start feature passes run variant 1), each lane
 * forming the terms of every fourth listed pair, added in list order across the four lanes.
 * fe_autotune tunes variants 1, 4 and 8 (variant 2 builds window tables once per list, which
 * timing sweep launches alone would not count; fe_tune_set selects it). In converge mode it
 * times pressure_update with its residual atomics, as the frame runs it.
 * fe_autotune times every configuration on the context's current state (the median of
 * `reps` launches, 0 = 20) and keeps the fastest; the fields the timed launches overwrite
 * are restored, so the next frame is the one the context would have run. Configurations
 * never change results. A JSON cache keyed by device and workload is kept by the Python
 * host (fluid_engine.tuning). */
typedef struct fe_tune_entry {
    char kernel[32];
    uint32_t block;
    uint32_t items;
    float median_us;    /* of the chosen configuration in the last fe_autotune (0 if none) */
    float default_us;   /* of the default configuration (block 256, 1 item) in that run */
} fe_tune_entry;
fe_status fe_autotune(fe_ctx* ctx, uint32_t reps);
fe_status fe_tune_count(const fe_ctx* ctx, uint32_t* count);
fe_status fe_tune_get(const fe_ctx* ctx, uint32_t index, fe_tune_entry* out);
fe_status fe_tune_set(fe_ctx* ctx, const char* kernel, uint32_t block, uint32_t items);
/* "<gcn arch>/<compute units>cu" of the context's device (the tuning cache key). */
fe_status fe_device_name(const fe_ctx* ctx, char* name, size_t cap);

/* ---- exact arithmetic self-check (no reference counterpart) ----------------------------------
 * The pair kernels compute square roots and quotients with shortcuts that equal the IEEE
 * operations on the inputs they receive. This runs on `device` and compares each shortcut
 * with the IEEE operation: the corrected square root for every float in [2^-96, 2^16); the
 * boxed division for `samples` pseudo-random pairs with 2^-40 <= |n|, |d| < 2^40; r / h by
 * the residual-corrected product for every float r in [1e-6, 1.001 h]. counts receives
 * {sqrt checked, sqrt mismatches, division checked, division mismatches, r/h checked,
 * r/h mismatches}. */
fe_status fe_check_exact_arith(int device, float h, uint64_t samples, uint64_t counts[6]);

/* ---- batched scenes (no reference counterpart) -----------------------------------------------
 * Many independent scenes of uneven sizes in one context, advanced together: every kernel of
 * a frame is one launch over all scenes. Each scene keeps its own parameters (mass, rest
 * density, viscosity, box, h), its own T_s = next_power_of_two(N_s)-entry hash table, its
 * own convergence (a converged scene's particles skip the remaining sweeps) and its own stats,
 * and evolves exactly as the same scene would in its own context (bit for bit). The scenes
 * share dt and the fixed sweep counts (they advance in lockstep) and run the radix sort
 * without the early-exit / CFL flags and with the reference warm start. fe_step, fe_run,
 * fe_set_options, the per-ComputeStep calls (except fe_sort) and whole-field transfers of
 * particle fields work on a batch context; per-scene transfers use fe_batch_*. */
fe_status fe_batch_create(const fe_params* params, const float* const* xyz, const uint32_t* counts,
                          uint32_t scenes, int device, fe_ctx** out);
fe_status fe_batch_scenes(const fe_ctx* ctx, uint32_t* scenes);
fe_status fe_batch_scene_info(const fe_ctx* ctx, uint32_t scene, uint32_t* n, uint32_t* table_size);
/* Per-scene parameters (dt and sweep counts must stay those of the batch). */
fe_status fe_batch_set_params(fe_ctx* ctx, uint32_t scene, const fe_params* params);
/* One scene's field as its own context would hold it: particle fields are the scene's N_s
 * rows; GRID_ENTRIES / GRID_START are the scene's own T_s-entry table with indices relative
 * to the scene. Uploads take particle fields only. */
fe_status fe_batch_download(fe_ctx* ctx, uint32_t scene, int field, void* dst, size_t bytes);
fe_status fe_batch_upload(fe_ctx* ctx, uint32_t scene, int field, const void* src, size_t bytes);
/* The scene's stats_buffer after the last frame. */
fe_status fe_batch_read_stats(fe_ctx* ctx, uint32_t scene, uint32_t out[3]);
/* The scene's sweeps in the last frame and since creation. */
fe_status fe_batch_iterations(fe_ctx* ctx, uint32_t scene, uint32_t* density, uint32_t* divergence,
                              uint64_t* density_total, uint64_t* divergence_total);
/* With FE_SOLVER_FIXED, capture the steady-state frame of the batch once in a HIP graph and
 * replay it for every later frame (parameter, option and field changes drop the graph). */
fe_status fe_batch_use_graph(fe_ctx* ctx, int enable);
/* The same for any context (no reference counterpart; the reference records each physics
 * frame into one command buffer, src/renderer/mod.rs:279-363). A single scene replays its
 * steady-state frame when it runs FE_SOLVER_FIXED with no CFL or early-exit flags, the
 * reference warm start, the clamp walls and no learning or exit reports; other frames run
 * eagerly. An attached moving wall keeps moving: the frame's box is copied to the device
 * before each frame. Results equal the eager frames bit for bit. */
fe_status fe_use_graph(fe_ctx* ctx, int enable);

/* ---- watchdog (no reference counterpart) ------------------------------------------------------
 * The reference's only health signal is the max-speed atomic of stats.comp:22-23. With the
 * watchdog on, the stats pass of every frame also counts the particles whose position_a,
 * velocity_a, density, source term or pressure is not finite, and keeps the lowest such index
 * (hash order). fe_get_health reports the last frame (it waits for that frame). With halt set,
 * fe_step / fe_run check the previous frame before running one and return FE_ERR_UNSTABLE
 * when it left a non-finite particle, or a max speed above speed_limit (speed_limit > 0; 0
 * sets no limit). Single-scene contexts; results are unchanged. */
#define FE_HEALTH_NONFINITE 1u
#define FE_HEALTH_SPEED 2u
typedef struct fe_health {
    uint64_t frame;            /* frames run when the reported frame ended (0: none yet) */
    uint32_t nonfinite;        /* particles with a non-finite field */
    uint32_t first_nonfinite;  /* the lowest such index (hash order), 0xFFFFFFFF if none */
    float max_speed;           /* the frame's max |velocity_a| (the stats buffer's word 0) */
    float speed_limit;
    uint32_t flags;            /* FE_HEALTH_NONFINITE | FE_HEALTH_SPEED */
    uint32_t sort_error;       /* nonzero: a radix sort's look-back stopped waiting (any context
                                  with it fails its next fe_step with FE_ERR_DEVICE) */
} fe_health;
fe_status fe_set_watchdog(fe_ctx* ctx, int enable, float speed_limit, int halt);
fe_status fe_get_health(fe_ctx* ctx, fe_health* out);
/* A moving wall per scene, as fe_set_wall / fe_get_wall: the scene's box comes from its
 * wall at the start of every frame (the per-scene constants are copied to the device on the
 * context's stream, also before graph replays). */
fe_status fe_batch_set_wall(fe_ctx* ctx, uint32_t scene, const fe_wall* wall);
fe_status fe_batch_get_wall(const fe_ctx* ctx, uint32_t scene, fe_wall* out, int* attached);

/* ---- learned warm start (no reference counterpart) -----------------------------------------
 * Two networks, FE_NET_DENSITY and FE_NET_DIVERGENCE, one per Jacobi solve. Each maps the
 * FE_MLP_INPUTS features of a particle through three GELU layers of FE_MLP_HIDDEN units to
 * one output, in BF16 with FP32 accumulation on the matrix cores. Features, at the start of
 * a solve, over the particle's pairs with r^2 <= h^2 (at most 64, reference visit order):
 *   0 (rho - rho0) / rho0          1 source_term dt / rho0        2 factor (alpha)
 *   3 pair count                   4 previous pressure            5 sum W
 *   6 sum W p_prev_j               7 |sum r_ij| / h               8 sum |r_ij|^2 / h^2
 *   9 sum (v_i - v_j).r_ij / h     10 sum |v_i - v_j|^2           11 -s rho^2 / (dt alpha)
 * with v = velocity_a and "previous pressure" the final pressure of the same particle in the
 * previous solve of the same kind (0 on the first solve and after uploads).
 * Parameters, FE_MLP_PARAMS floats: W1[12][64], b1[64], W2[64][64], b2[64], W3[64][64],
 * b3[64], w4[64], b4 (W[in][out], row-major). Normalisation, FE_MLP_NORM floats: feature
 * means[12], feature standard deviations[12], target mean, target standard deviation; the
 * network sees (feature - mean) / std and predicts (p - target mean) / target std.
 * With warm_start = FE_WARM_START_LEARNED each solve whose network is loaded starts from
 * max(0, prediction); the other solve keeps the reference start. */
#define FE_NET_DENSITY 0
#define FE_NET_DIVERGENCE 1
#define FE_MLP_INPUTS 12
#define FE_MLP_HIDDEN 64
#define FE_MLP_PARAMS 9217
#define FE_MLP_NORM 26

/* Record training samples: after every solve, samples_per_solve particles (chosen by a
 * hash) contribute (features at the start of the solve, final pressure) to that solve's
 * dataset, a ring buffer of `capacity` samples. samples_per_solve = 0 stops collecting but
 * keeps the pressure history the "previous pressure" features need. */
fe_status fe_learn_collect(fe_ctx* ctx, uint32_t samples_per_solve, uint32_t capacity);
/* Number of samples held for network `net`. */
fe_status fe_learn_dataset_size(fe_ctx* ctx, int net, uint32_t* size);
/* Copy the first `count` samples of the dataset: x (count x 12, raw features), t (count). */
fe_status fe_learn_dataset(fe_ctx* ctx, int net, float* x, float* t, uint32_t count);
/* Train network `net` from a fresh initialisation (seeded) on its dataset: normalisation
 * from the dataset, then `steps` Adam steps (beta 0.9 / 0.999, eps 1e-8) on mini-batches of
 * `batch` <= 65536 samples drawn by a hash of (seed, step, row); a batch of at least the
 * dataset's size uses every sample once, in order. The learning rate follows a cosine from
 * lr (first step) to lr / 20 (last step). losses (optional, `steps` floats) receives each
 * step's mean squared error in normalised units. steps = 0 only initialises. Deterministic. */
fe_status fe_learn_train(fe_ctx* ctx, int net, uint32_t steps, uint32_t batch, float lr, uint32_t seed,
                         float* losses);
fe_status fe_learn_get_weights(fe_ctx* ctx, int net, float* params, float* norm);
fe_status fe_learn_set_weights(fe_ctx* ctx, int net, const float* params, const float* norm);
/* Network output for n normalised inputs (n x 12 row-major) into y (n). */
fe_status fe_learn_forward(fe_ctx* ctx, int net, const float* x, uint32_t n, float* y);
/* Gradient of L = sum_b (y_b - t_b)^2 / n over n <= 65536 normalised samples, as the
 * training step computes it (FE_MLP_PARAMS floats), and L * n (optional). */
fe_status fe_learn_gradient(fe_ctx* ctx, int net, const float* x, const float* t, uint32_t n, float* grad,
                            float* sum_sq);
/* Runs the source-term stage of solve `net` (exactly fe_density_source or
 * fe_divergence_source, with all their effects), which gathers the features in the same
 * pass, and returns them (n x 12, hash order) with the previous pressures (n, optional).
 * In fe_step the source-term stages gather the features whenever a learned solve or sample
 * collection needs them. */
fe_status fe_learn_features(fe_ctx* ctx, int net, float* features, float* prev_pressure);
/* The source-term stage of solve `net` followed by the warm start alone: the prediction
 * is written to the pressures as a learned solve begins. */
fe_status fe_learn_predict(fe_ctx* ctx, int net);

#ifdef __cplusplus
}
#endif

#endif /* FLUID_ENGINE_H */
