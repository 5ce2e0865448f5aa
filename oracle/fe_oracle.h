/*
 * fe_oracle.h — CPU restatement of the reference DFSPH physics frame.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP path is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it. It is never part of the product path.
 *
 * It restates, line for line in semantics, the GLSL of Nikita-Lysiuk/Fluid-Engine
 * (shaders/include/common.glsl and the shaders/compute/<stage>.comp files) and the Rust sequencers that
 * drive them (src/renderer/mod.rs:227-364, src/renderer/pipelines/neighbor_search.rs:176-231,
 * src/renderer/pipelines/sorter.rs:73-115 and :325-455), on host memory laid out like
 * `GpuPhysicsData` (src/entities/particle.rs:137-161). Field ids and the parameter block
 * are those of include/fluid_engine.h.
 *
 * Every context runs in FP32 (the reference arithmetic, no FMA contraction, IEEE
 * division and square root) or, with fp64 != 0, in FP64 for the physics arithmetic while
 * the integer structure (cell coordinates, hashes, sort, offsets) is still derived from
 * FP32 positions exactly as in the FP32 path.
 */
#ifndef FE_ORACLE_H
#define FE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/fluid_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

oracle_ctx* oracle_create(const fe_params* params, const float* xyz, uint32_t n, int fp64);
void oracle_destroy(oracle_ctx* c);
void oracle_set_params(oracle_ctx* c, const fe_params* params);
/* Current parameter block (dt changes under the CFL flag). */
void oracle_get_params(oracle_ctx* c, fe_params* out);
void oracle_set_options(oracle_ctx* c, const fe_options* options);
void oracle_set_threads(int threads);

/* Field transfer. FP fields are float arrays on the wire (converted for fp64 contexts),
 * unless as_double != 0 on download. Returns 0 on success, nonzero on a size mismatch. */
int oracle_upload(oracle_ctx* c, int field, const void* src, size_t bytes);
int oracle_download(oracle_ctx* c, int field, void* dst, size_t bytes, int as_double);

/* One call per reference ComputeStep. */
void oracle_neighbor_search(oracle_ctx* c);
void oracle_sort(oracle_ctx* c);
void oracle_density_alpha(oracle_ctx* c);
void oracle_viscosity(oracle_ctx* c);
void oracle_density_source(oracle_ctx* c);
void oracle_pressure_force(oracle_ctx* c);
/* Returns the quantised compression residual of this sweep (see oracle_residual_q). */
uint64_t oracle_pressure_update(oracle_ctx* c);
void oracle_pressure_integration(oracle_ctx* c);
void oracle_divergence_source(oracle_ctx* c);
void oracle_divergence_integration(oracle_ctx* c);
void oracle_stats(oracle_ctx* c);
void oracle_read_stats(oracle_ctx* c, uint32_t out[3]);

/* The 12 warm-start features of every particle (n x 12, hash order) for the current state,
 * given the previous pressures (n, hash order). No reference counterpart: the definition is
 * the one of include/fluid_engine.h (fe_learn_*). */
void oracle_features(oracle_ctx* c, const float* prev_pressure, float* out);

/* Jacobi solves (FIXED or CONVERGE per the options); return the sweep count. */
uint32_t oracle_solve_density(oracle_ctx* c);
uint32_t oracle_solve_divergence(oracle_ctx* c);
/* Renderer::step physics frame with the current params and options (host flags from the
 * previous frame's stats when enabled). */
void oracle_step(oracle_ctx* c, float max_dt);
/* CollisionBox::update (src/entities/collision.rs:31-35) on a wall; attach a wall (NULL
 * detaches) that oracle_step copies into the box before each frame and then advances by
 * max_dt (src/renderer/ui.rs:125-126, src/renderer/mod.rs:269-270); read it back (returns
 * whether one is attached). */
void oracle_wall_update(fe_wall* w, float dt);
/* splat_density.comp (shaders/compute/splat_density.comp:11-51) after the texture clear
 * (src/renderer/mod.rs:386-388): the density texture of position_a, grid_res.x * grid_res.y
 * * grid_res.z uint32 with x fastest. Returns nonzero if `cells` is not that product. */
int oracle_splat_density(oracle_ctx* c, uint32_t* out, size_t cells);
/* FE_BOUNDARY_PARTICLES: the wall samples' boundary masses for the current box (faces 0..5 =
 * -x, +x, -y, +y, -z, +z, each m-major), up to cap into out; returns the sample count (0 when
 * the option is off). */
size_t oracle_wall_psi(oracle_ctx* c, float* out, size_t cap);
void oracle_set_wall(oracle_ctx* c, const fe_wall* w);
int oracle_get_wall(oracle_ctx* c, fe_wall* out);
void oracle_last_iterations(oracle_ctx* c, uint32_t* density, uint32_t* divergence);

/* Stand-alone pieces for unit tests. */
uint32_t oracle_cell_hash(int32_t cx, int32_t cy, int32_t cz, uint32_t table_size);
float oracle_kernel_w(float r, float h);
void oracle_kernel_grad(const float r_vec[3], float r, float h, float out[3]);
double oracle_kernel_w64(double r, double h);
/* Literal 4-pass radix sort (count / scan / reorder with 256-thread, 8-element chunks) and
 * the literal bitonic network, on interleaved {hash, index} pairs. n must be a power of two
 * for the bitonic network (GpuSorter::prepare asserts it, sorter.rs:56-59). */
void oracle_radix_sort_pairs(uint32_t* pairs, uint32_t n);
void oracle_bitonic_sort_pairs(uint32_t* pairs, uint32_t n);

/* Quantisation of the per-particle compression residual used by FE_SOLVER_CONVERGE:
 * q_i = floor(min(max(0, -(s_i - Ap_i)) * dt / rho0, 2^8) * 2^32) as uint64; the sweep's
 * error is sum(q_i) / (N * 2^32). Integer sums make the decision order-independent. */
uint64_t oracle_residual_q(float s_minus_ap, float dt, float rho0);

#ifdef __cplusplus
}
#endif

#endif
