"""`CollisionBox` (src/entities/collision.rs:11-42): the simulation box with an oscillating
min.x wall, held as the library's fe_wall. `update` is fe_wall_update; attaching a box to a
Simulation (`Simulation.set_wall`) makes every frame take its box from the wall and advance
the wall by the frame's max_dt, as the reference's UI pass and Renderer::step do
(src/renderer/ui.rs:125-126, src/renderer/mod.rs:269-270).
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_float

import numpy as np

from ._ffi import FeWall, check, load_library


class CollisionBox:
    def __init__(self, min, max):
        """CollisionBox::new (collision.rs:20-29): wave amplitude 0, frequency 1, time 0."""
        lib = load_library()
        self._w = FeWall()
        lo = (c_float * 3)(*[float(x) for x in min])
        hi = (c_float * 3)(*[float(x) for x in max])
        check(lib.fe_wall_init(byref(self._w), lo, hi), "fe_wall_init")

    @classmethod
    def from_wall(cls, wall: FeWall) -> "CollisionBox":
        box = cls.__new__(cls)
        box._w = FeWall()
        ctypes.memmove(byref(box._w), byref(wall), ctypes.sizeof(FeWall))
        return box

    @classmethod
    def from_params(cls, params, amplitude: float = 0.0, frequency: float = 1.0) -> "CollisionBox":
        """The box of a parameter block, with the given wave (Scene::new, src/core/scene.rs:19-23)."""
        box = cls(params.box_min[:3], params.box_max[:3])
        box.wave_amplitude = amplitude
        box.wave_frequency = frequency
        return box

    def wall(self) -> FeWall:
        """A copy of the fe_wall this box holds."""
        w = FeWall()
        ctypes.memmove(byref(w), byref(self._w), ctypes.sizeof(FeWall))
        return w

    @property
    def min(self) -> np.ndarray:
        return np.array(self._w.min[:], dtype=np.float32)

    @min.setter
    def min(self, v):
        for a in range(3):
            self._w.min[a] = float(v[a])

    @property
    def max(self) -> np.ndarray:
        return np.array(self._w.max[:], dtype=np.float32)

    @max.setter
    def max(self, v):
        for a in range(3):
            self._w.max[a] = float(v[a])

    @property
    def wave_amplitude(self) -> float:
        return self._w.wave_amplitude

    @wave_amplitude.setter
    def wave_amplitude(self, v: float):
        self._w.wave_amplitude = v

    @property
    def wave_frequency(self) -> float:
        return self._w.wave_frequency

    @wave_frequency.setter
    def wave_frequency(self, v: float):
        self._w.wave_frequency = v

    @property
    def time(self) -> float:
        return self._w.time

    @property
    def base_min_x(self) -> float:
        return self._w.base_min_x

    def update(self, dt: float) -> None:
        """CollisionBox::update (collision.rs:31-35)."""
        check(load_library().fe_wall_update(byref(self._w), c_float(dt)), "fe_wall_update")

    def contains(self, point) -> bool:
        """CollisionBox::contains (collision.rs:37-41): the closed box test, in float."""
        p = np.asarray(point, dtype=np.float32)
        lo = self.min
        hi = self.max
        return bool(np.all(p >= lo) and np.all(p <= hi))

    def __repr__(self) -> str:
        return (f"CollisionBox(min={self.min.tolist()}, max={self.max.tolist()}, "
                f"wave_amplitude={self.wave_amplitude}, wave_frequency={self.wave_frequency}, "
                f"time={self.time})")
