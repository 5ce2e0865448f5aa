"""fluid_engine — MI355X-native DFSPH physics frame (HIP/CDNA4 kernels behind a C ABI).

`Simulation` mirrors the reference's operator surface: one method per `ComputeStep`
of `ComputePipelines` (src/renderer/pipelines/mod.rs:78-90), `step` for the physics part
of `Renderer::step` (src/renderer/mod.rs:227-364), and field access to the buffers of
`GpuPhysicsData` (src/entities/particle.rs:137-161). All work runs in libfluid_engine.so
on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_double, c_float, c_uint32, c_uint64, c_void_p

import numpy as np

from . import scene, tuning
from ._ffi import (EXPORTED_SYMBOLS, FE_MLP_HIDDEN, FE_MLP_INPUTS, FE_MLP_NORM, FE_MLP_PARAMS, FE_NET_DENSITY,
                   FE_NET_DIVERGENCE, FE_SOLVER_CONVERGE, FE_SOLVER_FIXED, FE_SORT_BITONIC, FE_SORT_RADIX,
                   FE_WARM_START_LEARNED, FE_WARM_START_REFERENCE, FE_DIVERGENCE_REFERENCE, FE_DIVERGENCE_CONSISTENT,
                   FE_BOUNDARY_CLAMP, FE_BOUNDARY_PARTICLES, FE_RESIDUAL_COMPRESSION, FE_RESIDUAL_TWO_SIDED,
                   FE_ERR_UNSTABLE, FE_HEALTH_NONFINITE, FE_HEALTH_SPEED,
                   FIELDS, INT_FIELDS, LIB_PATH, FeHealth, FeOptions,
                   FeParams, FeTuneEntry, FeWall, FluidEngineError, check, load_library)
from .collision import CollisionBox

__all__ = [
    "Simulation", "FeParams", "FeOptions", "FluidEngineError", "load_library", "scene", "frame_control",
    "default_options", "FE_SORT_BITONIC", "FE_SORT_RADIX", "FE_SOLVER_FIXED", "FE_SOLVER_CONVERGE",
    "FE_WARM_START_REFERENCE", "FE_WARM_START_LEARNED", "FIELDS", "EXPORTED_SYMBOLS", "LIB_PATH",
    "FE_NET_DENSITY", "FE_NET_DIVERGENCE", "FE_MLP_INPUTS", "FE_MLP_HIDDEN", "FE_MLP_PARAMS", "FE_MLP_NORM",
    "Batch", "CollisionBox", "FeWall", "FE_DIVERGENCE_REFERENCE", "FE_DIVERGENCE_CONSISTENT",
    "FE_BOUNDARY_CLAMP", "FE_BOUNDARY_PARTICLES", "FE_RESIDUAL_COMPRESSION", "FE_RESIDUAL_TWO_SIDED",
]


def default_options() -> FeOptions:
    lib = load_library()
    o = FeOptions()
    check(lib.fe_default_options(byref(o)), "fe_default_options")
    return o


def frame_control(stats, n, params: FeParams, options: FeOptions):
    """Host control of Renderer::step (src/renderer/mod.rs:231-267): returns
    (dt, density_iterations, divergence_iterations) for the next frame."""
    lib = load_library()
    st = (c_uint32 * 3)(*[int(x) for x in stats])
    dt = c_float()
    di = c_uint32()
    vi = c_uint32()
    check(lib.fe_frame_control(st, n, byref(params), byref(options), byref(dt), byref(di), byref(vi)),
          "fe_frame_control")
    return dt.value, di.value, vi.value


class Simulation:
    """One fe_ctx: the particle state in HBM plus every physics stage."""

    def __init__(self, params: FeParams, positions: np.ndarray, device: int = -1,
                 options: FeOptions | None = None):
        self._lib = load_library()
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        ctx = c_void_p()
        check(self._lib.fe_create(byref(params), pos.ctypes.data, pos.shape[0], device, byref(ctx)), "fe_create")
        self._ctx = ctx
        n = c_uint32()
        t = c_uint32()
        check(self._lib.fe_num_particles(ctx, byref(n), byref(t)), "fe_num_particles")
        self.n = n.value
        self.table_size = t.value
        if options is not None:
            self.set_options(options)

    # -- lifetime -------------------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            check(self._lib.fe_destroy(self._ctx), "fe_destroy")
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- parameters / options -------------------------------------------------------
    @property
    def params(self) -> FeParams:
        p = FeParams()
        check(self._lib.fe_get_params(self._ctx, byref(p)), "fe_get_params")
        return p

    def set_params(self, params: FeParams):
        check(self._lib.fe_set_params(self._ctx, byref(params)), "fe_set_params")

    @property
    def options(self) -> FeOptions:
        o = FeOptions()
        check(self._lib.fe_get_options(self._ctx, byref(o)), "fe_get_options")
        return o

    def set_options(self, options: FeOptions):
        check(self._lib.fe_set_options(self._ctx, byref(options)), "fe_set_options")

    def configure(self, **kw):
        o = self.options
        for k, v in kw.items():
            setattr(o, k, v)
        self.set_options(o)

    # -- fields -------------------------------------------------------------------------
    def _shape(self, name):
        fid, per, per_table = FIELDS[name]
        count = self.table_size if per_table else self.n
        return fid, (count, per) if per > 1 else (count,)

    def download(self, name: str) -> np.ndarray:
        fid, shape = self._shape(name)
        out = np.empty(shape, dtype=np.uint32 if name in INT_FIELDS else np.float32)
        check(self._lib.fe_download(self._ctx, fid, out.ctypes.data, out.nbytes), f"fe_download({name})")
        return out

    def upload(self, name: str, data) -> None:
        fid, shape = self._shape(name)
        arr = np.ascontiguousarray(data, dtype=np.uint32 if name in INT_FIELDS else np.float32).reshape(shape)
        check(self._lib.fe_upload(self._ctx, fid, arr.ctypes.data, arr.nbytes), f"fe_upload({name})")

    def device_ptr(self, name: str) -> int:
        p = c_void_p()
        check(self._lib.fe_device_ptr(self._ctx, FIELDS[name][0], byref(p)), "fe_device_ptr")
        return p.value

    # -- ComputeSteps -------------------------------------------------------------------
    def neighbor_search(self):
        check(self._lib.fe_neighbor_search(self._ctx), "fe_neighbor_search")

    def sort(self):
        check(self._lib.fe_sort(self._ctx), "fe_sort")

    def density_alpha(self):
        check(self._lib.fe_density_alpha(self._ctx), "fe_density_alpha")

    def viscosity(self):
        check(self._lib.fe_viscosity(self._ctx), "fe_viscosity")

    def density_source(self):
        check(self._lib.fe_density_source(self._ctx), "fe_density_source")

    def pressure_force(self):
        check(self._lib.fe_pressure_force(self._ctx), "fe_pressure_force")

    def pressure_update(self):
        check(self._lib.fe_pressure_update(self._ctx), "fe_pressure_update")

    def pressure_integration(self):
        check(self._lib.fe_pressure_integration(self._ctx), "fe_pressure_integration")

    def divergence_source(self):
        check(self._lib.fe_divergence_source(self._ctx), "fe_divergence_source")

    def divergence_integration(self):
        check(self._lib.fe_divergence_integration(self._ctx), "fe_divergence_integration")

    def stats(self):
        check(self._lib.fe_stats(self._ctx), "fe_stats")

    def solve_density(self) -> int:
        it = c_uint32()
        check(self._lib.fe_solve_density(self._ctx, byref(it)), "fe_solve_density")
        return it.value

    def solve_divergence(self) -> int:
        it = c_uint32()
        check(self._lib.fe_solve_divergence(self._ctx, byref(it)), "fe_solve_divergence")
        return it.value

    def step(self, max_dt: float | None = None):
        dt = self.params.dt if max_dt is None else max_dt
        check(self._lib.fe_step(self._ctx, c_float(dt)), "fe_step")

    def run(self, frames: int, max_dt: float | None = None):
        dt = self.params.dt if max_dt is None else max_dt
        check(self._lib.fe_run(self._ctx, c_float(dt), frames), "fe_run")

    # -- render hand-off ---------------------------------------------------------------------
    def splat_density(self):
        """Clear and rebuild the density texture (splat_density.comp) from position_a."""
        check(self._lib.fe_splat_density(self._ctx), "fe_splat_density")

    def density_texture(self) -> np.ndarray:
        """The last splat's texture as uint32 [grid_res.z, grid_res.y, grid_res.x]."""
        ptr = c_void_p()
        size = ctypes.c_size_t()
        check(self._lib.fe_density_texture(self._ctx, byref(ptr), byref(size)), "fe_density_texture")
        p = self.params
        shape = (int(p.grid_res[2]), int(p.grid_res[1]), int(p.grid_res[0]))
        out = np.empty(size.value // 4, dtype=np.uint32)
        check(self._lib.fe_download_density_texture(self._ctx, out.ctypes.data, out.nbytes),
              "fe_download_density_texture")
        return out.reshape(shape) if out.size == shape[0] * shape[1] * shape[2] else out

    def density_texture_ptr(self):
        """(device address, bytes) of the last splat's texture."""
        ptr = c_void_p()
        size = ctypes.c_size_t()
        check(self._lib.fe_density_texture(self._ctx, byref(ptr), byref(size)), "fe_density_texture")
        return ptr.value, size.value

    def copy_field_device(self, name: str, dst_ptr: int, nbytes: int):
        """Enqueue a device-to-device copy of a particle field into caller memory."""
        fid = FIELDS[name][0]
        check(self._lib.fe_copy_field_device(self._ctx, fid, c_void_p(dst_ptr), nbytes), "fe_copy_field_device")

    # -- moving wall (CollisionBox) ---------------------------------------------------------
    def set_wall(self, box: CollisionBox | None):
        """Attach a copy of `box` (None detaches): every later frame takes the parameter box
        from the wall, then advances the wall by the frame's max_dt."""
        check(self._lib.fe_set_wall(self._ctx, None if box is None else byref(box.wall())), "fe_set_wall")

    def set_watchdog(self, enable: bool = True, speed_limit: float = 0.0, halt: bool = False):
        """Count non-finite particles (and the max speed against speed_limit) in every frame's
        stats pass; with halt, the next frame refuses to run (FE_ERR_UNSTABLE) after an
        unhealthy one."""
        check(self._lib.fe_set_watchdog(self._ctx, 1 if enable else 0, speed_limit, 1 if halt else 0),
              "fe_set_watchdog")

    def health(self) -> dict:
        """The watchdog's report of the last frame."""
        h = FeHealth()
        check(self._lib.fe_get_health(self._ctx, byref(h)), "fe_get_health")
        return {"frame": h.frame, "nonfinite": h.nonfinite, "first_nonfinite": h.first_nonfinite,
                "max_speed": h.max_speed, "speed_limit": h.speed_limit, "flags": h.flags, "sort_error": h.sort_error}

    def use_graph(self, enable: bool = True):
        """Replay the steady-state FIXED frame from a captured HIP graph (fe_use_graph)."""
        check(self._lib.fe_use_graph(self._ctx, 1 if enable else 0), "fe_use_graph")

    def wall(self) -> CollisionBox | None:
        """The attached wall as advanced so far, or None."""
        w = FeWall()
        on = ctypes.c_int()
        check(self._lib.fe_get_wall(self._ctx, byref(w), byref(on)), "fe_get_wall")
        return CollisionBox.from_wall(w) if on.value else None

    # -- readback / timing ----------------------------------------------------------------
    def read_stats(self) -> np.ndarray:
        out = (c_uint32 * 3)()
        check(self._lib.fe_read_stats(self._ctx, out), "fe_read_stats")
        return np.array(out[:], dtype=np.uint32)

    def last_iterations(self):
        d = c_uint32()
        v = c_uint32()
        check(self._lib.fe_last_iterations(self._ctx, byref(d), byref(v)), "fe_last_iterations")
        return d.value, v.value

    def iteration_totals(self):
        d = c_uint64()
        v = c_uint64()
        f = c_uint64()
        check(self._lib.fe_iteration_totals(self._ctx, byref(d), byref(v), byref(f)), "fe_iteration_totals")
        return d.value, v.value, f.value

    def reset_iteration_totals(self):
        check(self._lib.fe_reset_iteration_totals(self._ctx), "fe_reset_iteration_totals")

    def solve_report_enable(self, enable: bool = True):
        """Report each Jacobi solve's exit residual from now on (resets the averages)."""
        check(self._lib.fe_solve_report_enable(self._ctx, 1 if enable else 0), "fe_solve_report_enable")

    def solve_report(self, kind: int) -> dict:
        """Exit residuals of the solves of `kind` (0 density, 1 divergence) since the last
        enable, in units of rho0 / dt, each averaged over the solves."""
        n = c_uint32()
        out = (c_double * 4)()
        check(self._lib.fe_solve_report(self._ctx, kind, byref(n), out), "fe_solve_report")
        return {"solves": n.value, "mean_abs": out[0], "max_abs": out[1], "mean_expansion": out[2],
                "mean_expansion_pressured": out[3]}

    def sync(self):
        check(self._lib.fe_sync(self._ctx), "fe_sync")

    def timer_start(self):
        check(self._lib.fe_timer_start(self._ctx), "fe_timer_start")

    def timer_stop(self) -> float:
        ms = c_float()
        check(self._lib.fe_timer_stop(self._ctx, byref(ms)), "fe_timer_stop")
        return ms.value

    def profile(self, enable: bool = True):
        check(self._lib.fe_profile_enable(self._ctx, 1 if enable else 0), "fe_profile_enable")

    def profile_reset(self):
        check(self._lib.fe_profile_reset(self._ctx), "fe_profile_reset")

    def profile_results(self) -> dict:
        cnt = c_uint32()
        check(self._lib.fe_profile_count(self._ctx, byref(cnt)), "fe_profile_count")
        res = {}
        buf = ctypes.create_string_buffer(128)
        for i in range(cnt.value):
            launches = c_uint64()
            total = c_double()
            check(self._lib.fe_profile_entry(self._ctx, i, buf, 128, byref(launches), byref(total)),
                  "fe_profile_entry")
            res[buf.value.decode()] = (launches.value, total.value)
        return res

    # -- launch-configuration autotuner ---------------------------------------------------
    def device_name(self) -> str:
        buf = ctypes.create_string_buffer(128)
        check(self._lib.fe_device_name(self._ctx, buf, 128), "fe_device_name")
        return buf.value.decode()

    def autotune_now(self, reps: int = 20) -> dict:
        """Time every launch configuration on the current state and keep the fastest."""
        check(self._lib.fe_autotune(self._ctx, reps), "fe_autotune")
        return self.tune_entries()

    def tune_entries(self) -> dict:
        cnt = c_uint32()
        check(self._lib.fe_tune_count(self._ctx, byref(cnt)), "fe_tune_count")
        out = {}
        e = FeTuneEntry()
        for i in range(cnt.value):
            check(self._lib.fe_tune_get(self._ctx, i, byref(e)), "fe_tune_get")
            out[e.kernel.decode()] = {"block": e.block, "items": e.items, "median_us": round(e.median_us, 3),
                                      "default_us": round(e.default_us, 3)}
        return out

    def set_tune(self, kernel: str, block: int, items: int = 1):
        check(self._lib.fe_tune_set(self._ctx, kernel.encode(), block, items), "fe_tune_set")

    def autotune(self, cache_path: str | None = None, reps: int = 20, retune: bool = False) -> dict:
        """The tuned configuration for this device and workload: from the JSON cache when it
        holds one, else measured now (and stored there)."""
        return tuning.autotune(self, cache_path, reps, retune)

    # -- learned warm start (no reference counterpart) -------------------------------------
    def learn_collect(self, samples_per_solve: int, capacity: int = 0):
        """Record (features, final pressure) samples after every solve; 0 stops collecting."""
        check(self._lib.fe_learn_collect(self._ctx, samples_per_solve, capacity), "fe_learn_collect")

    def learn_dataset_size(self, net: int) -> int:
        n = c_uint32()
        check(self._lib.fe_learn_dataset_size(self._ctx, net, byref(n)), "fe_learn_dataset_size")
        return n.value

    def learn_dataset(self, net: int, count: int | None = None):
        count = self.learn_dataset_size(net) if count is None else count
        x = np.empty((count, FE_MLP_INPUTS), np.float32)
        t = np.empty(count, np.float32)
        check(self._lib.fe_learn_dataset(self._ctx, net, x.ctypes.data, t.ctypes.data, count), "fe_learn_dataset")
        return x, t

    def learn_train(self, net: int, steps: int, batch: int = 4096, lr: float = 2e-3, seed: int = 0) -> np.ndarray:
        """Train network `net` on its dataset; returns the per-step normalised MSE."""
        losses = np.empty(max(steps, 1), np.float32)
        check(self._lib.fe_learn_train(self._ctx, net, steps, batch, lr, seed, losses.ctypes.data), "fe_learn_train")
        return losses[:steps]

    def learn_weights(self, net: int):
        p = np.empty(FE_MLP_PARAMS, np.float32)
        nm = np.empty(FE_MLP_NORM, np.float32)
        check(self._lib.fe_learn_get_weights(self._ctx, net, p.ctypes.data, nm.ctypes.data), "fe_learn_get_weights")
        return p, nm

    def learn_set_weights(self, net: int, params, norm):
        p = np.ascontiguousarray(params, np.float32).reshape(FE_MLP_PARAMS)
        nm = np.ascontiguousarray(norm, np.float32).reshape(FE_MLP_NORM)
        check(self._lib.fe_learn_set_weights(self._ctx, net, p.ctypes.data, nm.ctypes.data), "fe_learn_set_weights")

    def learn_forward(self, net: int, x) -> np.ndarray:
        xx = np.ascontiguousarray(x, np.float32).reshape(-1, FE_MLP_INPUTS)
        y = np.empty(xx.shape[0], np.float32)
        check(self._lib.fe_learn_forward(self._ctx, net, xx.ctypes.data, xx.shape[0], y.ctypes.data),
              "fe_learn_forward")
        return y

    def learn_gradient(self, net: int, x, t):
        xx = np.ascontiguousarray(x, np.float32).reshape(-1, FE_MLP_INPUTS)
        tt = np.ascontiguousarray(t, np.float32).reshape(xx.shape[0])
        g = np.empty(FE_MLP_PARAMS, np.float32)
        ss = c_float()
        check(self._lib.fe_learn_gradient(self._ctx, net, xx.ctypes.data, tt.ctypes.data, xx.shape[0],
                                          g.ctypes.data, byref(ss)), "fe_learn_gradient")
        return g, ss.value

    def learn_features(self, net: int):
        """(features n x 12, previous pressures n), both in hash order."""
        f = np.empty((self.n, FE_MLP_INPUTS), np.float32)
        pp = np.empty(self.n, np.float32)
        check(self._lib.fe_learn_features(self._ctx, net, f.ctypes.data, pp.ctypes.data), "fe_learn_features")
        return f, pp

    def learn_predict(self, net: int):
        check(self._lib.fe_learn_predict(self._ctx, net), "fe_learn_predict")


class Batch(Simulation):
    """Many independent scenes in one context (fe_batch_*): every kernel of a frame is one
    launch over all scenes; each scene evolves as it would alone. Whole-context methods
    (step, run, set_options, timers, profiling) are those of Simulation; per-scene fields,
    stats and sweep counts take a scene index."""

    def __init__(self, workloads, device: int = -1, options: FeOptions | None = None):
        self._lib = load_library()
        count = len(workloads)
        params = (FeParams * count)(*[w.params for w in workloads])
        pos = [np.ascontiguousarray(w.positions, dtype=np.float32).reshape(-1, 3) for w in workloads]
        ptrs = (c_void_p * count)(*[p.ctypes.data for p in pos])
        counts = (c_uint32 * count)(*[p.shape[0] for p in pos])
        ctx = c_void_p()
        check(self._lib.fe_batch_create(params, ptrs, counts, count, device, byref(ctx)), "fe_batch_create")
        self._ctx = ctx
        n = c_uint32()
        t = c_uint32()
        check(self._lib.fe_num_particles(ctx, byref(n), byref(t)), "fe_num_particles")
        self.n = n.value
        self.table_size = t.value
        self.scenes = count
        self.scene_n = []
        self.scene_table_size = []
        for k in range(count):
            check(self._lib.fe_batch_scene_info(ctx, k, byref(n), byref(t)), "fe_batch_scene_info")
            self.scene_n.append(n.value)
            self.scene_table_size.append(t.value)
        if options is not None:
            self.set_options(options)

    def _scene_shape(self, scene, name):
        fid, per, table = FIELDS[name]
        rows = self.scene_table_size[scene] if table else self.scene_n[scene]
        return fid, (rows, per) if per > 1 else (rows,)

    def download_scene(self, scene: int, name: str) -> np.ndarray:
        fid, shape = self._scene_shape(scene, name)
        out = np.empty(shape, dtype=np.uint32 if name in INT_FIELDS else np.float32)
        check(self._lib.fe_batch_download(self._ctx, scene, fid, out.ctypes.data, out.nbytes), "fe_batch_download")
        return out

    def upload_scene(self, scene: int, name: str, data) -> None:
        fid, shape = self._scene_shape(scene, name)
        arr = np.ascontiguousarray(data, dtype=np.float32).reshape(shape)
        check(self._lib.fe_batch_upload(self._ctx, scene, fid, arr.ctypes.data, arr.nbytes), "fe_batch_upload")

    def set_scene_params(self, scene: int, params: FeParams):
        check(self._lib.fe_batch_set_params(self._ctx, scene, byref(params)), "fe_batch_set_params")

    def read_scene_stats(self, scene: int) -> np.ndarray:
        out = (c_uint32 * 3)()
        check(self._lib.fe_batch_read_stats(self._ctx, scene, out), "fe_batch_read_stats")
        return np.array(out[:], dtype=np.uint32)

    def scene_iterations(self, scene: int):
        """(density, divergence) sweeps of the last frame and (density, divergence) totals."""
        d, v = c_uint32(), c_uint32()
        dt, vt = c_uint64(), c_uint64()
        check(self._lib.fe_batch_iterations(self._ctx, scene, byref(d), byref(v), byref(dt), byref(vt)),
              "fe_batch_iterations")
        return (d.value, v.value), (dt.value, vt.value)

    def use_graph(self, enable: bool = True):
        check(self._lib.fe_batch_use_graph(self._ctx, 1 if enable else 0), "fe_batch_use_graph")

    def set_scene_wall(self, scene: int, box: CollisionBox | None):
        check(self._lib.fe_batch_set_wall(self._ctx, scene, None if box is None else byref(box.wall())),
              "fe_batch_set_wall")

    def scene_wall(self, scene: int) -> CollisionBox | None:
        w = FeWall()
        on = ctypes.c_int()
        check(self._lib.fe_batch_get_wall(self._ctx, scene, byref(w), byref(on)), "fe_batch_get_wall")
        return CollisionBox.from_wall(w) if on.value else None
