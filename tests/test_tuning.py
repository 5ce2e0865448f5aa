"""The autotuner's JSON cache (fluid_engine.tuning) with a stand-in context: a cached
workload is applied without measuring, a new one is measured once and stored, and a
damaged or foreign file is ignored."""
import json

from fluid_engine import tuning
from fluid_engine._ffi import FluidEngineError


class FakeSim:
    def __init__(self, n=1000, scenes=1, kernels=("pressure_force",)):
        self.n = n
        self.scenes = scenes
        self.kernels = kernels
        self.measured = 0
        self.applied = {}

    def tune_entries(self):
        return {k: {"block": 256, "items": 1, "median_us": 0.0, "default_us": 0.0} for k in self.kernels}

    def device_name(self):
        return "gfx950/256cu"

    def autotune_now(self, reps):
        self.measured += 1
        return {k: {"block": 128, "items": 1, "median_us": 10.0, "default_us": 12.0} for k in self.kernels}

    def set_tune(self, kernel, block, items):
        if block not in (64, 128, 256, 512, 1024):
            raise FluidEngineError(1, "fe_tune_set", "no build for block %d" % block)
        self.applied[kernel] = (block, items)


def test_cache_roundtrip(tmp_path):
    path = str(tmp_path / "tune.json")
    a = FakeSim()
    r = tuning.autotune(a, path, reps=3)
    assert a.measured == 1 and r["pressure_force"]["block"] == 128
    data = json.load(open(path))
    assert data["version"] == tuning.VERSION
    assert tuning.cache_key("gfx950/256cu", 1000, 1) in data["entries"]
    b = FakeSim()
    tuning.autotune(b, path)
    assert b.measured == 0 and b.applied == {"pressure_force": (128, 1)}
    c = FakeSim(n=2000)                       # another workload: measured, added
    tuning.autotune(c, path)
    assert c.measured == 1
    assert len(json.load(open(path))["entries"]) == 2
    d = FakeSim()
    tuning.autotune(d, path, retune=True)     # forced re-measure
    assert d.measured == 1


def test_bad_cache_is_ignored(tmp_path):
    path = tmp_path / "tune.json"
    path.write_text(json.dumps({"version": 99, "entries": {"x": 1}}))
    s = FakeSim()
    tuning.autotune(s, str(path))
    assert s.measured == 1
    assert json.load(open(path))["version"] == tuning.VERSION


def test_stale_entry_is_measured_again(tmp_path):
    """An entry naming a configuration the library no longer builds is re-measured."""
    path = str(tmp_path / "tune.json")
    key = tuning.cache_key("gfx950/256cu", 1000, 1)
    stale = {"pressure_force": {"block": 96, "items": 1, "median_us": 1.0, "default_us": 1.0}}
    tuning.save_cache(path, {"version": tuning.VERSION, "entries": {key: stale}})
    s = FakeSim()
    r = tuning.autotune(s, path)
    assert s.measured == 1 and r["pressure_force"]["block"] == 128
    assert json.load(open(path))["entries"][key]["pressure_force"]["block"] == 128


def test_entry_without_a_tunable_kernel_is_measured_again(tmp_path):
    """A cached entry from a library with fewer tunable kernels would leave the new kernel at
    its default: it is measured again, and the new entry names every kernel."""
    path = str(tmp_path / "tune.json")
    tuning.autotune(FakeSim(), path)
    s = FakeSim(kernels=("pressure_force", "density_alpha_divergence_source"))
    r = tuning.autotune(s, path)
    assert s.measured == 1 and set(r) == {"pressure_force", "density_alpha_divergence_source"}
    t = FakeSim(kernels=("pressure_force", "density_alpha_divergence_source"))
    tuning.autotune(t, path)
    assert t.measured == 0 and t.applied["density_alpha_divergence_source"] == (128, 1)
