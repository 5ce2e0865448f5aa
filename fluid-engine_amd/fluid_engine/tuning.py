"""JSON cache of autotuned launch configurations (fe_autotune / fe_tune_set).

Entries are keyed by device ("<gcn arch>/<compute units>cu"), particle count and scene count,
so a configuration measured once on a device is reused by later runs of the same workload.
File format: {"version": 2, "entries": {key: {kernel: {"block", "items", "median_us",
"default_us"}}}}.
"""
from __future__ import annotations

import json
import os

from ._ffi import FluidEngineError

VERSION = 2


def cache_key(device: str, particles: int, scenes: int = 1) -> str:
    return f"{device}|n={particles}|scenes={scenes}"


def load_cache(path: str | None) -> dict:
    if not path or not os.path.exists(path):
        return {"version": VERSION, "entries": {}}
    with open(path) as f:
        data = json.load(f)
    if not isinstance(data, dict) or data.get("version") != VERSION or not isinstance(data.get("entries"), dict):
        return {"version": VERSION, "entries": {}}
    return data


def save_cache(path: str, data: dict) -> None:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def autotune(sim, cache_path: str | None = None, reps: int = 20, retune: bool = False) -> dict:
    """Apply the cached configuration of `sim`'s device and workload, or measure it with
    fe_autotune and store it. Returns {kernel: {"block", "items", ...}}."""
    key = cache_key(sim.device_name(), sim.n, getattr(sim, "scenes", 1))
    cache = load_cache(cache_path)
    entry = cache["entries"].get(key)
    # an entry written by a library with other tunable kernels would leave the missing ones
    # at their defaults: measure again
    if entry is not None and set(entry) != set(sim.tune_entries()):
        entry = None
    if entry and not retune:
        try:
            for kernel, cfg in entry.items():
                sim.set_tune(kernel, int(cfg["block"]), int(cfg["items"]))
            return entry
        except FluidEngineError:   # a configuration this library no longer builds: measure again
            pass
    entry = sim.autotune_now(reps)
    if cache_path:
        cache["entries"][key] = entry
        save_cache(cache_path, cache)
    return entry
