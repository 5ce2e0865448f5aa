"""ctypes binding of the CPU oracle (oracle/libfe_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker. It shares field names and the parameter
block with the product binding so that comparisons read stage by stage.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_float, c_int, c_int32, c_size_t, c_uint32, c_uint64, c_void_p, byref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FE_ORACLE_LIB names another build of fe_oracle.cpp (tests/test_oracle_mutants.py loads
# variants with a planted mistake to show that the pins catch it).
LIB = os.environ.get("FE_ORACLE_LIB") or os.path.join(HERE, "libfe_oracle.so")

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    if force and os.path.exists(LIB):
        os.remove(LIB)
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = ctypes.CDLL(LIB)
    P = c_void_p
    sig = {
        "oracle_create": ([c_void_p, c_void_p, c_uint32, c_int], P),
        "oracle_destroy": ([P], None),
        "oracle_set_params": ([P, c_void_p], None),
        "oracle_set_options": ([P, c_void_p], None),
        "oracle_get_params": ([P, c_void_p], None),
        "oracle_set_threads": ([c_int], None),
        "oracle_upload": ([P, c_int, c_void_p, c_size_t], c_int),
        "oracle_download": ([P, c_int, c_void_p, c_size_t, c_int], c_int),
        "oracle_pressure_update": ([P], c_uint64),
        "oracle_read_stats": ([P, POINTER(c_uint32)], None),
        "oracle_solve_density": ([P], c_uint32),
        "oracle_solve_divergence": ([P], c_uint32),
        "oracle_step": ([P, c_float], None),
        "oracle_last_iterations": ([P, POINTER(c_uint32), POINTER(c_uint32)], None),
        "oracle_cell_hash": ([c_int32, c_int32, c_int32, c_uint32], c_uint32),
        "oracle_kernel_w": ([c_float, c_float], c_float),
        "oracle_kernel_w64": ([ctypes.c_double, ctypes.c_double], ctypes.c_double),
        "oracle_kernel_grad": ([POINTER(c_float), c_float, c_float, POINTER(c_float)], None),
        "oracle_radix_sort_pairs": ([c_void_p, c_uint32], None),
        "oracle_bitonic_sort_pairs": ([c_void_p, c_uint32], None),
        "oracle_residual_q": ([c_float, c_float, c_float], c_uint64),
        "oracle_features": ([P, c_void_p, c_void_p], None),
        "oracle_wall_update": ([c_void_p, c_float], None),
        "oracle_splat_density": ([P, c_void_p, c_size_t], c_int),
        "oracle_wall_psi": ([P, c_void_p, c_size_t], c_size_t),
        "oracle_set_wall": ([P, c_void_p], None),
        "oracle_get_wall": ([P, c_void_p], c_int),
    }
    for stage in ("neighbor_search", "sort", "density_alpha", "viscosity", "density_source", "pressure_force",
                  "pressure_integration", "divergence_source", "divergence_integration", "stats"):
        sig["oracle_" + stage] = ([P], None)
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


FIELDS = {
    "position_a": (0, 4, False), "position_b": (1, 4, False), "velocity_a": (2, 4, False),
    "velocity_b": (3, 4, False), "colors": (4, 4, False), "densities": (5, 1, False),
    "factors": (6, 1, False), "source_terms": (7, 1, False), "pressures": (8, 1, False),
    "pressure_accelerations": (9, 4, False), "grid_entries": (10, 2, True), "grid_start": (11, 1, True),
}
INT_FIELDS = ("grid_entries", "grid_start")


class Wall(ctypes.Structure):
    """The fe_wall layout of include/fluid_engine.h (CollisionBox, src/entities/collision.rs:11-18)."""

    _fields_ = [("min", c_float * 3), ("max", c_float * 3), ("wave_amplitude", c_float),
                ("wave_frequency", c_float), ("time", c_float), ("base_min_x", c_float)]


def make_wall(box_min, box_max, amplitude=0.0, frequency=1.0):
    """CollisionBox::new (collision.rs:20-29) followed by setting the two wave fields."""
    w = Wall()
    for a in range(3):
        w.min[a] = box_min[a]
        w.max[a] = box_max[a]
    w.wave_amplitude = amplitude
    w.wave_frequency = frequency
    w.time = 0.0
    w.base_min_x = box_min[0]
    return w


def wall_update(wall, dt):
    """CollisionBox::update (collision.rs:31-35), in place."""
    lib().oracle_wall_update(ctypes.addressof(wall), c_float(dt))
    return wall


def set_threads(t: int):
    lib().oracle_set_threads(int(t))


class Oracle:
    """CPU twin of fluid_engine.Simulation (FP32, or FP64 arithmetic with fp64=True)."""

    def __init__(self, params, positions, fp64=False, options=None):
        self._l = lib()
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        self.n = pos.shape[0]
        t = 1
        while t < self.n:
            t <<= 1
        self.table_size = t
        self.fp64 = bool(fp64)
        self._params = params
        self._c = self._l.oracle_create(ctypes.addressof(params), pos.ctypes.data, self.n, 1 if fp64 else 0)
        if options is not None:
            self.set_options(options)

    def close(self):
        if getattr(self, "_c", None):
            self._l.oracle_destroy(self._c)
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        self._params = params
        self._l.oracle_set_params(self._c, ctypes.addressof(params))

    @property
    def params(self):
        p = type(self._params)()
        self._l.oracle_get_params(self._c, ctypes.addressof(p))
        return p

    def set_options(self, options):
        self._l.oracle_set_options(self._c, ctypes.addressof(options))

    def _shape(self, name):
        fid, per, per_table = FIELDS[name]
        count = self.table_size if per_table else self.n
        return fid, (count, per) if per > 1 else (count,)

    def download(self, name, as_double=False):
        fid, shape = self._shape(name)
        if name in INT_FIELDS:
            dt = np.uint32
        else:
            dt = np.float64 if as_double else np.float32
        out = np.empty(shape, dtype=dt)
        rc = self._l.oracle_download(self._c, fid, out.ctypes.data, out.nbytes, 1 if as_double else 0)
        if rc:
            raise ValueError(f"oracle download {name} size mismatch")
        return out

    def upload(self, name, data):
        fid, shape = self._shape(name)
        arr = np.ascontiguousarray(data, dtype=np.uint32 if name in INT_FIELDS else np.float32).reshape(shape)
        if self._l.oracle_upload(self._c, fid, arr.ctypes.data, arr.nbytes):
            raise ValueError(f"oracle upload {name} size mismatch")

    def __getattr__(self, stage):
        if stage.startswith("_"):
            raise AttributeError(stage)
        fn = getattr(lib(), "oracle_" + stage, None)
        if fn is None:
            raise AttributeError(stage)
        return lambda: fn(self._c)

    def solve_density(self):
        return self._l.oracle_solve_density(self._c)

    def solve_divergence(self):
        return self._l.oracle_solve_divergence(self._c)

    def step(self, max_dt=None):
        self._l.oracle_step(self._c, c_float(self._params.dt if max_dt is None else max_dt))

    def set_wall(self, wall):
        """Attach a moving wall (None detaches): each step copies its box into the
        parameters, then advances it by max_dt."""
        self._l.oracle_set_wall(self._c, None if wall is None else ctypes.addressof(wall))

    def wall(self):
        w = Wall()
        on = self._l.oracle_get_wall(self._c, ctypes.addressof(w))
        return w if on else None

    def wall_psi(self):
        """Boundary masses of the wall samples (empty unless options.boundary is PARTICLES)."""
        n = self._l.oracle_wall_psi(self._c, None, 0)
        out = np.empty(n, np.float32)
        if n:
            self._l.oracle_wall_psi(self._c, out.ctypes.data, n)
        return out

    def splat_density(self):
        """The density texture (grid_res z, y, x uint32) of the current position_a."""
        p = self.params
        rx, ry, rz = int(p.grid_res[0]), int(p.grid_res[1]), int(p.grid_res[2])
        out = np.empty((rz, ry, rx), np.uint32)
        if self._l.oracle_splat_density(self._c, out.ctypes.data, out.size):
            raise ValueError("oracle splat: bad grid_res")
        return out

    def read_stats(self):
        out = (c_uint32 * 3)()
        self._l.oracle_read_stats(self._c, out)
        return np.array(out[:], dtype=np.uint32)

    def features(self, prev_pressure):
        """The 12 warm-start features of every particle (n x 12, hash order)."""
        pp = np.ascontiguousarray(prev_pressure, dtype=np.float32).reshape(self.n)
        out = np.empty((self.n, 12), np.float32)
        self._l.oracle_features(self._c, pp.ctypes.data, out.ctypes.data)
        return out

    def last_iterations(self):
        d = c_uint32()
        v = c_uint32()
        self._l.oracle_last_iterations(self._c, byref(d), byref(v))
        return d.value, v.value


def cell_hash(cx, cy, cz, table_size):
    return lib().oracle_cell_hash(cx, cy, cz, table_size)


def kernel_w(r, h):
    return lib().oracle_kernel_w(r, h)


def kernel_w64(r, h):
    return lib().oracle_kernel_w64(r, h)


def kernel_grad(rv, r, h):
    a = (c_float * 3)(*rv)
    o = (c_float * 3)()
    lib().oracle_kernel_grad(a, r, h, o)
    return np.array(o[:], dtype=np.float32)


def radix_sort_pairs(pairs: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(pairs, dtype=np.uint32).copy()
    lib().oracle_radix_sort_pairs(p.ctypes.data, p.shape[0])
    return p


def bitonic_sort_pairs(pairs: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(pairs, dtype=np.uint32).copy()
    lib().oracle_bitonic_sort_pairs(p.ctypes.data, p.shape[0])
    return p


def residual_q(s_minus_ap, dt, rho0):
    return lib().oracle_residual_q(s_minus_ap, dt, rho0)
