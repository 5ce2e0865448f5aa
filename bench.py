#!/usr/bin/env python3
"""DFSPH throughput on one MI355X: the 1,048,576-particle dam break.

One step = one physics frame of Renderer::step with max_dt == dt = 1 ms (frame-start
neighbour search + density/alpha, viscosity, density Jacobi solve to 0.1 % compression
error, integration, neighbour rebuild, density/alpha, divergence Jacobi solve to 1 %,
divergence integration, stats; at most 100 sweeps per solve). By default 500 steps are timed
after 5 warm-up steps from rest. Inputs are resident in HBM before the timed region; the region
is bracketed by HIP events on the context's stream.

Prints one JSON line: steps/s (value), particle-updates/s, Jacobi sweeps per step, the
roofline of the dominant kernel (algorithmic bytes / event-timed average launch, and its
VALU issue rate: the gathers are bound by vector-ALU issue, not by HBM) and a CPU baseline
(the oracle on the host cores, bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fluid-engine_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per particle per launch: logical FP32 fields, vec3 = 12 B, each array
# touched once per launch, neighbour re-reads not counted; T = N and ~N/8 occupied cells.
ALGO_BYTES = {
    "spatial_hash": 20.0 + 4.0,        # pos 12 -> key 4 + index 4; grid_start fill 4
    "radix_hist": 4.0,                  # key 4 per pass
    "radix_scan": 0.0,
    "radix_scatter": 16.0,              # key+index 8 in, 8 out per pass
    "offsets_reorder": 56.5,            # key 4 + starts 0.5 + index 4 + pos/vel 24 in, 24 out
    "density_alpha": 28.0,
    "viscosity": 48.0,
    "density_source": 44.0,
    "pressure_force": 40.0,
    "pressure_update": 52.0,
    "pressure_integration": 60.0,
    "divergence_source": 36.0,
    "divergence_integration": 48.0,
    "stats": 20.0,
}
# The source terms with the warm-start features gathered in the same pass also write the 12
# features (48 B) and read the previous pressure (4 B).
ALGO_BYTES["density_source_feat"] = ALGO_BYTES["density_source"] + 52.0
ALGO_BYTES["divergence_source_feat"] = ALGO_BYTES["divergence_source"] + 52.0
# density_alpha and divergence_source fused into one pass after the post-integrate search: the
# position (12 B) both read is read once
ALGO_BYTES["density_alpha_divergence_source"] = ALGO_BYTES["density_alpha"] + ALGO_BYTES["divergence_source"] - 12.0
ALGO_BYTES["density_alpha_divergence_source_feat"] = ALGO_BYTES["density_alpha_divergence_source"] + 52.0
# Warm-start inference: 2 (12 x 64 + 2 x 64 x 64 + 64) multiply-adds per particle on the matrix
# cores (BF16 dense peak 2.5 PFLOP/s).
MLP_FLOP = 2.0 * (12 * 64 + 2 * 64 * 64 + 64)
MFMA_BF16_PEAK_TFLOPS = 2500.0
# The splat gather's stencil: 1647 offsets with |o|^2 <= 53 (the non-zero contributions),
# formed for every texel of a 32 x 8 x 4 tile whose halo holds a particle. Per lane (4 texels)
# and busy tile it reads 2178 LDS words and stores its share of the 46 x 22 x 18-word halo.
SPLAT_TAPS = 1647
SPLAT_TILE = (32, 8, 4)
SPLAT_LDS_WORDS_PER_LANE = 2178 + 46 * 22 * 18 / 256
# LDS read peak: 128 B/clk per CU (the CDNA figure) x 256 CUs x 2.4 GHz.
LDS_PEAK_TBS = 128 * 256 * 2.4e9 / 1e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="dam_break_1m", choices=["dam_break_1m", "dam_break_32k", "block_4m"])
    ap.add_argument("--solver", default="converge", choices=["converge", "fixed"])
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scenes measurement")
    ap.add_argument("--no-tune", action="store_true", help="run the default launch configurations")
    ap.add_argument("--tune-cache", default=None, help="JSON cache of tuned configurations (read and written)")
    ap.add_argument("--batch-frames", type=int, default=10)
    ap.add_argument("--no-sort-bench", action="store_true", help="skip the isolated 2^22 sort measurement")
    ap.add_argument("--no-splat", action="store_true", help="skip the density-splat (render hand-off) measurement")
    ap.add_argument("--no-consistent", action="store_true", help="skip the consistent divergence-order measurement")
    ap.add_argument("--no-boundary", action="store_true", help="skip the wall boundary-particles measurement")
    ap.add_argument("--no-app", action="store_true", help="skip the app-scene frame measurement")
    ap.add_argument("--no-search-4m", action="store_true",
                    help="skip the 2^22-particle neighbour search + density_alpha measurement")
    ap.add_argument("--learned-collect", type=int, default=100, help="frames sampled before training")
    ap.add_argument("--learned-eval", type=int, default=400, help="frames compared after training")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--app-eager-only", action="store_true",
                    help="only time the reference app's 25,000-particle frame, eagerly launched, physics only "
                         "(prints that section's JSON; for profiling its kernels alone)")
    return ap.parse_args()


def copy_probe():
    """A plain device-to-device copy of 1 GiB whose source was pushed out of the 256 MB
    MALL by a later 1 GiB write: the HBM bytes per dispatch are known exactly (1 GiB read,
    1 GiB written), which calibrates the traffic counters."""
    import torch
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    z = torch.full((n,), 2.0, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "check": float(y[-1]) + float(z[0])}), flush=True)


def traffic_for(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", "traffic_dam_break_1m.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f).get("kernels", {}).get(kernel)
    return None if t is None else t.get("hbm_bytes_per_launch")


def valu_for(kernel, avg_launch_us):
    """VALU issue roofline of `kernel`: its wave64 VALU instructions per launch (from the
    committed rocprofv3 SQ counter summary) over the event-timed average launch, against the
    chip's issue rate (one wave64 VALU instruction per SIMD per 4 cycles)."""
    path = os.path.join(ROOT, "profiles", "valu_dam_break_1m.json")
    if not os.path.exists(path) or not avg_launch_us:
        return None
    with open(path) as f:
        doc = json.load(f)
    k = doc.get("kernels", {}).get(kernel)
    if k is None:
        return None
    achieved = k["valu_per_launch"] / (avg_launch_us * 1e3)   # G wave-instructions/s
    peak = doc["issue_peak_ginstr_per_s"]
    return {"achieved": round(achieved, 1), "peak": peak, "unit": "G wave64 VALU instr/s",
            "frac": round(achieved / peak, 4), "valu_per_launch": k["valu_per_launch"]}


def main():
    args = parse()
    if args.copy_probe:
        copy_probe()
        return
    if args.app_eager_only:
        import fluid_engine as fe
        from fluid_engine import scene
        print(json.dumps(app_frame_measurement(fe, scene, False, contexts=((False, False),))), flush=True)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build runs one scene on one GPU; --gpus {args.gpus} / WORLD_SIZE {world} "
              "is not supported (multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import fluid_engine as fe
    from fluid_engine import scene

    w = getattr(scene, args.config)(seed=0)
    opts = fe.default_options()
    opts.sort_algorithm = fe.FE_SORT_RADIX
    opts.solver_mode = fe.FE_SOLVER_CONVERGE if args.solver == "converge" else fe.FE_SOLVER_FIXED
    opts.density_tol = 1e-3
    opts.divergence_tol = 1e-2
    opts.max_iterations = 100
    print(f"[bench] {w.name} n={w.n} solver={args.solver} warmup={args.warmup} steps={args.steps}",
          file=sys.stderr, flush=True)
    sim = fe.Simulation(w.params, w.positions, device=0, options=opts)
    dt = w.params.dt

    sim.run(args.warmup, dt)
    sim.sync()
    tuned = None
    if not args.no_tune:
        t0 = time.perf_counter()
        tuned = sim.autotune(args.tune_cache, reps=20)
        print(f"[bench] autotune {time.perf_counter() - t0:.2f} s: " +
              json.dumps({k: (v["block"], v["items"]) for k, v in tuned.items()}), file=sys.stderr, flush=True)
    sim.reset_iteration_totals()
    t0 = time.perf_counter()
    sim.timer_start()
    sim.run(args.steps, dt)
    ms = sim.timer_stop()
    wall = time.perf_counter() - t0
    d_it, v_it, frames = sim.iteration_totals()
    print(f"[bench] timed {args.steps} steps: {ms:.2f} ms (wall {wall * 1e3:.2f} ms)", file=sys.stderr, flush=True)
    untuned = None
    if tuned is not None:
        # the same frames with the default configurations, from the same initial state
        ref = fe.Simulation(w.params, w.positions, device=0, options=opts)
        ref.run(args.warmup, dt)
        ref.sync()
        ref.timer_start()
        ref.run(args.steps, dt)
        ums = ref.timer_stop()
        ref.close()
        untuned = round(args.steps / (ums / 1e3), 3)

    # Per-kernel durations: every launch of a few more steps bracketed by HIP events on the
    # context's stream. These steps run the timed run's mean sweep counts as fixed sweeps, so
    # every pressure_force / pressure_update launch does a full sweep (in converge mode the
    # launches queued past convergence return at once and would dilute the average).
    d_avg = max(1, int(round(d_it / max(frames, 1))))
    v_avg = max(1, int(round(v_it / max(frames, 1))))
    prof_opts = fe.FeOptions.from_buffer_copy(opts)
    prof_opts.solver_mode = fe.FE_SOLVER_FIXED
    prof_params = fe.FeParams.from_buffer_copy(sim.params)
    prof_params.density_solver_iterations = d_avg
    prof_params.divergence_solver_iterations = v_avg
    sim.set_options(prof_opts)
    sim.set_params(prof_params)
    sim.profile(True)
    sim.profile_reset()
    sim.run(args.profile_steps, dt)
    prof = sim.profile_results()
    sim.profile(False)
    kernels = {}
    for name, (launches, total_ms) in prof.items():
        per = ALGO_BYTES.get(name)
        avg_us = total_ms * 1e3 / max(launches, 1)
        gbs = per * w.n / (avg_us * 1e-6) / 1e9 if per else None
        kernels[name] = {"launches_per_step": launches / args.profile_steps, "avg_us": round(avg_us, 3),
                         "share": None, "algo_gbs": round(gbs, 1) if gbs is not None else None}
    tot = sum(v[1] for v in prof.values())
    for name in kernels:
        kernels[name]["share"] = round(prof[name][1] / tot, 4) if tot else None
    dom = max((k for k in kernels if ALGO_BYTES.get(k)), key=lambda k: prof[k][1])
    dom_avg_s = kernels[dom]["avg_us"] * 1e-6
    dom_bytes = ALGO_BYTES[dom]
    achieved = dom_bytes * w.n / dom_avg_s / 1e9
    print("[bench] kernels: " + json.dumps(kernels), file=sys.stderr, flush=True)

    steps_per_s = args.steps / (ms / 1e3)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(w, opts, d_it / max(frames, 1), v_it / max(frames, 1), args.cpu_seconds)

    out = {
        "metric": "DFSPH steps/s at 1M particles (dam break, dt 1 ms, Jacobi to 0.1% density / 1% divergence error)",
        "value": round(steps_per_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered block)",
        "config": {"workload": w.name, "particles": w.n, "table_size": sim.table_size, "dt": dt,
                   "solver": args.solver, "density_tol": 1e-3, "divergence_tol": 1e-2, "max_iterations": 100,
                   "sort": "radix"},
        "particle_updates_per_s": round(steps_per_s * w.n, 1),
        "jacobi_sweeps_per_step": {"density": round(d_it / max(frames, 1), 3),
                                   "divergence": round(v_it / max(frames, 1), 3)},
        "profile_sweeps": {"density": d_avg, "divergence": v_avg},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_for(dom),
                     "algo_bytes_per_particle": dom_bytes, "avg_launch_us": kernels[dom]["avg_us"],
                     "valu_issue": valu_for(dom, kernels[dom]["avg_us"])},
        "cpu_baseline": cpu,
        "autotune": None if tuned is None else {
            "untuned_steps_per_s": untuned,
            "configs": {k: [v["block"], v["items"]] for k, v in tuned.items()},
            "median_us_tuned_vs_default": {k: [v["median_us"], v["default_us"]] for k, v in tuned.items()}},
    }
    if not args.no_splat:
        out["render_handoff"] = splat_measurement(fe, sim, w, args.no_cpu_baseline)
    sim.close()
    if not args.no_consistent:
        out["consistent_divergence"] = divergence_order_measurement(fe, w, opts, args.steps, args.warmup, tuned)
    if not args.no_app:
        out["app_frame"] = app_frame_measurement(fe, scene, not args.no_cpu_baseline)
    if not args.no_boundary:
        out["boundary_particles"] = boundary_measurement(fe, w, opts, args.steps, args.warmup)
    if not args.no_learned and args.solver == "converge":
        out["learned_warm_start"] = learned_comparison(fe, w, opts, args.learned_collect, args.learned_eval, tuned)
    if not args.no_batch:
        out["batched_scenes"] = batched_measurement(fe, scene, args.batch_frames)
    if not args.no_sort_bench:
        out["isolated_sort_2p22"] = sort_measurement(fe, scene)
    if not args.no_search_4m:
        out["search_4m"] = search_measurement(fe, scene, args.no_cpu_baseline)
    print(json.dumps(out), flush=True)


# Algorithmic bytes per particle of one neighbour search (hash 20, four-pass sort 68, fill +
# offsets 8.5, reorder 52) and of density + alpha (28), T = N.
SEARCH_BYTES = 148.5
DENSITY_ALPHA_BYTES = 28.0


def search_measurement(fe, scene, no_cpu, warmup=3, reps=20):
    """The neighbour search and density + alpha of the 2^22-particle block (256 x 128 x 128,
    spacing 0.02 m, in a 6 x 3 x 3 m box, T = 2^22): hash -> sort -> grid_start -> reorder
    (plus the gather schedule) and the neighbour list + density_alpha, each stage timed alone
    with HIP events on the context's stream, median of `reps` after `warmup`. The roofline
    prices the reference's algorithmic bytes of those stages (the schedule and the list are
    this build's own extra work inside the same time). The context is autotuned after one
    frame (the tuner times the frame's kernels on a stepped state), and the stages are timed
    on that state."""
    w = scene.block_4m(seed=0)
    o = fe.default_options()
    o.sort_algorithm = fe.FE_SORT_RADIX
    sim = fe.Simulation(w.params, w.positions, device=0, options=o)
    sim.step()
    tuned = sim.autotune_now(10)
    ns, da = [], []
    for r in range(warmup + reps):
        sim.sync()
        sim.timer_start()
        sim.neighbor_search()
        a = sim.timer_stop()
        sim.timer_start()
        sim.density_alpha()
        b = sim.timer_stop()
        if r >= warmup:
            ns.append(a)
            da.append(b)
    sim.close()
    med = lambda v: sorted(v)[len(v) // 2]
    t_ns, t_da = med(ns), med(da)
    t = med([a + b for a, b in zip(ns, da)])
    byts = (SEARCH_BYTES + DENSITY_ALPHA_BYTES) * w.n
    out = {"workload": f"{w.name} (256 x 128 x 128 block, 6 x 3 x 3 m box)", "particles": w.n, "sort": "radix",
           "protocol": f"one frame, autotune, then the median of {reps} after {warmup} warm-up, HIP events per stage",
           "tuned": {k: [tuned[k]["block"], tuned[k]["items"]] for k in ("build_list", "density_alpha")},
           "neighbor_search_ms": round(t_ns, 4), "density_alpha_ms": round(t_da, 4), "total_ms": round(t, 4),
           "particles_per_s": round(w.n / (t * 1e-3), 1),
           "roofline": {"bound": "hbm", "achieved": round(byts / (t * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(byts / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "algo_bytes_per_particle": SEARCH_BYTES + DENSITY_ALPHA_BYTES,
                        "hbm_floor_ms": round(byts / (HBM_PEAK_GBS * 1e9) * 1e3, 4)}}
    if not no_cpu:
        from oracle import oracle
        cores = min(os.cpu_count() or 1, 16)
        oracle.set_threads(cores)
        orc = oracle.Oracle(w.params, w.positions, options=o)
        t0 = time.perf_counter()
        orc.neighbor_search()
        orc.density_alpha()
        el = time.perf_counter() - t0
        orc.close()
        out["cpu_baseline"] = {"value": round(1e3 * el, 1), "unit": "ms per search + density_alpha", "cores": cores,
                               "kind": "port", "sample": "one search + density_alpha of the same block by the oracle"}
    return out


def sort_measurement(fe, scene, log2n=22, warmup=3, reps=15):
    """The reference's isolated sort benchmark (sorter.rs:869-924, scripts/sort_benchmark.csv:20-21):
    2^22 {hash, index} pairs with random u32 keys, median of `reps` sorts after `warmup`, each
    sort timed alone with HIP events on the context's stream (fresh random keys uploaded
    before each one, outside the timed region). fe_sort sorts all entries on 32 bits (radix:
    four 8-bit passes)."""
    import numpy as np
    n = 1 << log2n
    rng = np.random.default_rng(22)
    params = scene.make_params(0.01, 0.008, 0.04, 1000, 0.01, 0.5, 1e-3, 1, 1, (0, -9.81, 0), (0, 0, 0), (1, 1, 1))
    out = {"pairs": n, "keys": "random u32", "protocol": f"median of {reps} after {warmup} warm-up, HIP events"}
    for name, algo in (("radix", fe.FE_SORT_RADIX), ("bitonic", fe.FE_SORT_BITONIC)):
        o = fe.default_options()
        o.sort_algorithm = algo
        sim = fe.Simulation(params, np.zeros((n, 3), np.float32), device=0, options=o)
        times = []
        for r in range(warmup + reps):
            keys = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
            sim.upload("grid_entries", np.stack([keys, np.arange(n, dtype=np.uint32)], axis=1))
            sim.sync()
            sim.timer_start()
            sim.sort()
            ms = sim.timer_stop()
            if r >= warmup:
                times.append(ms)
        sim.close()
        times.sort()
        out[f"{name}_ms"] = round(times[len(times) // 2], 4)
    out["reference"] = {"radix_ms": 27.5775, "bitonic_ms": 82.766, "hardware": "unstated GPU (Windows, Vulkan)",
                        "measures": "submit-to-fence wall clock", "source": "scripts/sort_benchmark.csv:20-21"}
    # 68 algorithmic bytes per pair for a four-pass LSD sort of 8-byte pairs, at 8 TB/s
    out["radix_hbm_floor_ms"] = round(68 * n / 8e12 * 1e3, 4)
    return out


def divergence_order_measurement(fe, w, opts, steps, warmup, tuned=None):
    """The same frames (warm-up, then `steps` timed with HIP events) with the reference and the
    consistent divergence order (FE_DIVERGENCE_CONSISTENT, a non-parity mode): steps/s and the
    Jacobi sweeps per step."""
    out = {"workload": w.name, "frames": f"{warmup} warm-up + {steps} timed, from rest"}
    for name, order in (("reference", fe.FE_DIVERGENCE_REFERENCE), ("consistent", fe.FE_DIVERGENCE_CONSISTENT)):
        o = fe.FeOptions.from_buffer_copy(opts)
        o.divergence_order = order
        sim = fe.Simulation(w.params, w.positions, device=0, options=o)
        if tuned:
            for k, v in tuned.items():
                sim.set_tune(k, v["block"], v["items"])
        sim.run(warmup, w.params.dt)
        sim.reset_iteration_totals()
        sim.timer_start()
        sim.run(steps, w.params.dt)
        ms = sim.timer_stop()
        d, v, frames = sim.iteration_totals()
        sim.close()
        out[name] = {"steps_per_s": round(steps / (ms / 1e3), 3),
                     "tuned_blocks": {k: v["block"] for k, v in tuned.items()},
                     "sweeps_per_step": {"density": round(d / max(frames, 1), 3),
                                         "divergence": round(v / max(frames, 1), 3)}}
    return out


def app_frame_measurement(fe, scene, with_cpu, warmup=20, frames=200,
                          contexts=((False, True), (True, True), (False, False))):
    """The reference app's own frame (src/core/scene.rs:14-86, src/renderer/mod.rs:227-391): its
    25,000-particle scene, the min.x wall oscillating 0.3 m at 0.5 Hz, fixed 4/4 sweeps, radix
    sort, dt = max_dt = 5 ms, and the raymarching density splat (128^3) after each frame.
    Frames/s over `frames` frames (HIP events), with and without the splat, each context
    autotuned after its warm-up; `contexts` lists the (splat, graph) pairs to time."""
    w = scene.app_scene()
    o = fe.default_options()
    o.sort_algorithm = fe.FE_SORT_RADIX
    out = {"workload": f"{w.name}, moving wall {w.wall_wave[0]} m / {w.wall_wave[1]} Hz, fixed 4/4 sweeps",
           "frames": f"{warmup} warm-up, autotune, 5 frames + {frames} timed"}
    # the reference records each frame into one command buffer; here the frame is replayed
    # from a captured HIP graph (fe_use_graph), and also timed eagerly
    for splat, graph in contexts:
        sim = fe.Simulation(w.params, w.positions, device=0, options=o)
        sim.set_wall(fe.CollisionBox.from_params(w.params, *w.wall_wave))
        for _ in range(warmup):
            sim.step()
            if splat:
                sim.splat_density()
        # tuned on this scene after the warm-up (at 25,000 particles the deep gathers win)
        tuned = sim.autotune_now(20)
        sim.use_graph(graph)
        for _ in range(5):
            sim.step()
            if splat:
                sim.splat_density()
        sim.sync()
        sim.timer_start()
        for _ in range(frames):
            sim.step()
            if splat:
                sim.splat_density()
        ms = sim.timer_stop()
        wall = sim.wall()
        sim.close()
        key = ("with_splat" if splat else "physics_only") + ("" if graph else "_eager")
        out[key] = {"frames_per_s": round(frames / (ms / 1e3), 2), "ms_per_frame": round(ms / frames, 4),
                    "graph": graph}
        out["tuned"] = {k: [v["block"], v["items"]] for k, v in tuned.items()}
        out["wall_min_x_after"] = round(float(wall.min[0]), 5)
    if with_cpu:
        # the oracle on the host cores: the same frames with the moving wall, physics only
        from oracle import oracle
        cores = min(os.cpu_count() or 1, 16)
        oracle.set_threads(cores)
        orc = oracle.Oracle(w.params, w.positions, options=o)
        orc.set_wall(oracle.make_wall(w.params.box_min[:3], w.params.box_max[:3], *w.wall_wave))
        n = 0
        t0 = time.perf_counter()
        while n < frames and time.perf_counter() - t0 < 5.0:
            orc.step()
            n += 1
        el = time.perf_counter() - t0
        orc.close()
        out["cpu_baseline"] = {"value": round(n / el, 2), "unit": "frames/s", "cores": cores, "kind": "port",
                               "sample": f"{n} physics frames of the same scene from rest by the oracle, {el:.1f} s"}
    return out


def boundary_measurement(fe, w, opts, steps, warmup):
    """The same frames with the reference's clamp-only walls and with wall boundary particles
    (FE_BOUNDARY_PARTICLES, a non-parity mode), each autotuned after the warm-up: steps/s,
    Jacobi sweeps per step, and after the
    frames the mean density of the fluid layer on the floor (particles within 2 r of it) and the
    number of particles the floor's collision clamp holds."""
    import numpy as np
    out = {"workload": w.name, "frames": f"{warmup} warm-up + {steps} timed, from rest"}
    r = float(w.params.particle_radius)
    rho0 = float(w.params.target_density)
    for name, b in (("clamp", fe.FE_BOUNDARY_CLAMP), ("particles", fe.FE_BOUNDARY_PARTICLES)):
        o = fe.FeOptions.from_buffer_copy(opts)
        o.boundary = b
        sim = fe.Simulation(w.params, w.positions, device=0, options=o)
        sim.run(warmup, w.params.dt)
        # each mode tuned on its own state (with the samples on, the tuner times the walled
        # gather instances)
        tuned = sim.autotune(None, reps=10)
        sim.reset_iteration_totals()
        sim.timer_start()
        sim.run(steps, w.params.dt)
        ms = sim.timer_stop()
        d, v, frames = sim.iteration_totals()
        pos = sim.download("position_b")
        rho = sim.download("densities")
        pos_a = sim.download("position_a")
        sim.close()
        floor = pos[:, 1] < float(w.params.box_min[1]) + 2 * r
        # resting on the collision clamp: the clamp puts a particle at box_min.y + r + 0.001
        on_clamp = pos_a[:, 1] <= float(w.params.box_min[1]) + r + 0.0015
        out[name] = {"steps_per_s": round(steps / (ms / 1e3), 3),
                     "tuned_blocks": {k: v["block"] for k, v in tuned.items()},
                     "sweeps_per_step": {"density": round(d / max(frames, 1), 3),
                                         "divergence": round(v / max(frames, 1), 3)},
                     "floor_layer": {"particles": int(floor.sum()),
                                     "mean_rho_over_rho0": round(float(rho[floor].mean()) / rho0, 4)},
                     "held_by_the_floor_clamp": int(on_clamp.sum()),
                     "all_mean_rho_over_rho0": round(float(rho.mean()) / rho0, 4)}
    return out


def splat_work(pos, params):
    """What one splat of these positions costs the reference and this build: the reference's
    image atomics (texels at d^2 < 64 inside the image around each in-box particle's centre
    texel, splat_density.comp:29-47) and the gather's busy 32 x 8 x 4 tiles (a tile works
    when an 8^3 block of centre texels within 7 texels of it holds a particle)."""
    import numpy as np
    from scipy.signal import fftconvolve
    f = np.float32
    res = np.array([int(params.grid_res[a]) for a in range(3)])
    mn = np.array(params.box_min[:3], f)
    inv = (f(1.0) / (np.array(params.box_max[:3], f) - mn)).astype(f)
    uvw = ((pos[:, :3].astype(f) - mn) * inv).astype(f)
    ok = np.all((uvw >= 0) & (uvw <= 1), axis=1)
    c = (uvw[ok] * res.astype(f)).astype(np.int64)
    cnt = np.zeros(tuple(res + 1), np.float64)
    np.add.at(cnt, (c[:, 0], c[:, 1], c[:, 2]), 1.0)
    o = np.arange(-8, 9)
    ball = ((o[:, None, None] ** 2 + o[None, :, None] ** 2 + o[None, None, :] ** 2) < 64).astype(np.float64)
    atomics = float(np.rint(fftconvolve(cnt, ball, mode="full")[8:8 + res[0], 8:8 + res[1], 8:8 + res[2]]).sum())
    occ = np.zeros(tuple(res // 8 + 1), bool)
    occ[c[:, 0] // 8, c[:, 1] // 8, c[:, 2] // 8] = True
    tiles = [(res[a] + SPLAT_TILE[a] - 1) // SPLAT_TILE[a] for a in range(3)]
    busy = 0
    for tz in range(tiles[2]):
        z0 = tz * SPLAT_TILE[2]
        bz = slice(max(z0 - 7, 0) // 8, min(z0 + SPLAT_TILE[2] - 1 + 7, res[2]) // 8 + 1)
        for ty in range(tiles[1]):
            y0 = ty * SPLAT_TILE[1]
            by = slice(max(y0 - 7, 0) // 8, min(y0 + SPLAT_TILE[1] - 1 + 7, res[1]) // 8 + 1)
            for tx in range(tiles[0]):
                x0 = tx * SPLAT_TILE[0]
                bx = slice(max(x0 - 7, 0) // 8, min(x0 + SPLAT_TILE[0] - 1 + 7, res[0]) // 8 + 1)
                busy += bool(occ[bx, by, bz].any())
    return {"in_box": int(ok.sum()), "reference_atomics": atomics, "busy_tiles": busy,
            "tiles": int(np.prod(tiles))}


def splat_measurement(fe, sim, w, no_cpu, warmup=3, reps=20):
    """The render hand-off on the benchmark's state (after the timed frames): clear + splat of
    the raymarching density texture (grid_res^3 uint32) from position_a, timed with HIP events
    (the mean over `reps`), with per-kernel durations, the gather's integer multiply-add issue
    rate against the VALU peak, and the oracle's splat on the host cores."""
    for _ in range(warmup):
        sim.splat_density()
    sim.sync()
    sim.timer_start()
    for _ in range(reps):
        sim.splat_density()
    ms = sim.timer_stop() / reps
    sim.profile(True)
    sim.profile_reset()
    for _ in range(reps):
        sim.splat_density()
    prof = sim.profile_results()
    sim.profile(False)
    kernels = {k: round(v[1] * 1e3 / max(v[0], 1), 3) for k, v in prof.items() if k.startswith("splat")}
    pos = sim.download("position_a")
    work = splat_work(pos, sim.params)
    gather_s = kernels["splat_gather"] * 1e-6
    lanes = work["busy_tiles"] * SPLAT_TILE[0] * SPLAT_TILE[1]
    lds_bytes = lanes * SPLAT_LDS_WORDS_PER_LANE * 4
    texel_taps = lanes * SPLAT_TILE[2] * SPLAT_TAPS
    res = [int(sim.params.grid_res[a]) for a in range(3)]
    out = {"texture": f"{res[0]}x{res[1]}x{res[2]} uint32", "particles": w.n, "ms_per_splat": round(ms, 4),
           "kernels_avg_us": kernels, "reference_atomics_per_splat": work["reference_atomics"],
           "reference_atomics_per_s": round(work["reference_atomics"] / (ms * 1e-3), 1),
           "busy_tiles": work["busy_tiles"], "tiles": work["tiles"],
           "roofline": {"bound": "lds", "kernel": "splat_gather", "achieved": round(lds_bytes / gather_s / 1e12, 2),
                        "peak": round(LDS_PEAK_TBS, 2), "unit": "TB/s",
                        "frac": round(lds_bytes / gather_s / 1e12 / LDS_PEAK_TBS, 4),
                        "lds_bytes_per_launch": round(lds_bytes), "texel_taps_per_s": round(texel_taps / gather_s, 1),
                        "avg_launch_us": kernels["splat_gather"]}}
    if not no_cpu:
        from oracle import oracle
        cores = min(os.cpu_count() or 1, 16)
        oracle.set_threads(cores)
        orc = oracle.Oracle(sim.params, w.positions)
        orc.upload("position_a", pos)
        t0 = time.perf_counter()
        orc.splat_density()
        el = time.perf_counter() - t0
        orc.close()
        out["cpu_baseline"] = {"value": round(1e3 * el, 1), "unit": "ms per splat", "cores": cores, "kind": "port",
                               "sample": "one splat of the same position_a by the oracle (OpenMP, private textures)"}
    return out


def learned_comparison(fe, w, opts, collect, evaluate, tuned=None, report_frames=100):
    """Learned warm start on the same workload. A context samples 2048 (features, final
    pressure) pairs per solve over the first `collect` frames and both networks are trained on
    the GPU (2000 Adam steps, batch 4096). Then, for each exit residual (FE_RESIDUAL_TWO_SIDED,
    the fair comparison: an over-predicted pressure is not taken as converged; and the
    compression-only X1/X2 residual), a context from the reference start and one from the
    learned start run `evaluate` frames from the same state at frame `collect` to the same
    0.1 % / 1 % thresholds (sweeps per frame and frames per second, HIP events), then
    `report_frames` more frames with the exit-residual report on (fe_solve_report, outside the
    timed frames)."""
    def make(collecting):
        sim = fe.Simulation(w.params, w.positions, device=0, options=opts)
        for k, v in (tuned or {}).items():   # every context runs the tuned configurations
            sim.set_tune(k, v["block"], v["items"])
        if collecting:
            sim.learn_collect(2048, 1 << 18)   # also keeps each particle's previous pressures
        sim.run(collect)
        sim.sync()
        if collecting:
            sim.learn_collect(0)
        return sim

    trainer = make(True)
    t0 = time.perf_counter()
    l0 = trainer.learn_train(fe.FE_NET_DENSITY, steps=2000, batch=4096, lr=2e-3, seed=1)
    l1 = trainer.learn_train(fe.FE_NET_DIVERGENCE, steps=2000, batch=4096, lr=2e-3, seed=2)
    train_s = time.perf_counter() - t0
    nets = {net: trainer.learn_weights(net) for net in (fe.FE_NET_DENSITY, fe.FE_NET_DIVERGENCE)}
    prof = {}
    modes = {}
    for mode, resid in (("two_sided", fe.FE_RESIDUAL_TWO_SIDED), ("compression", fe.FE_RESIDUAL_COMPRESSION)):
        res = {}
        for name in ("reference", "learned"):
            if name == "learned" and mode == "two_sided":
                sim, trainer = trainer, None          # continues from its own sampled frames
            else:
                sim = make(name == "learned")
            o = fe.FeOptions.from_buffer_copy(opts)
            o.residual = resid
            if name == "learned":
                for net, (prm, norm) in nets.items():
                    sim.learn_set_weights(net, prm, norm)
                o.warm_start = fe.FE_WARM_START_LEARNED
            sim.set_options(o)
            if mode == "compression":
                # the one-sided exit takes an over-predicted pressure as converged, so a learned
                # start can blow the fluid apart; the watchdog stops such a run (both starts
                # run with it, and its per-frame check is in their timings)
                sim.set_watchdog(True, speed_limit=30.0, halt=True)
            sim.reset_iteration_totals()
            sim.timer_start()
            try:
                sim.run(evaluate)
            except fe.FluidEngineError as err:
                if err.status != fe.FE_ERR_UNSTABLE:
                    raise
                sim.sync()
                d, v, f = sim.iteration_totals()
                res[name] = {"unstable": str(err), "frames_before_halt": f}
                sim.close()
                continue
            ms = sim.timer_stop()
            d, v, f = sim.iteration_totals()
            r = {"steps_per_s": round(evaluate / (ms / 1e3), 3),
                 "sweeps_per_step": {"density": round(d / f, 3), "divergence": round(v / f, 3)},
                 "mean_abs_density_error": round(float(sim.read_stats()[1]) / w.n, 3)}
            sim.solve_report_enable(True)
            try:
                sim.run(report_frames)
                r["exit_residual"] = {kind: {k: (round(x, 7) if isinstance(x, float) else x)
                                             for k, x in sim.solve_report(kid).items()}
                                      for kind, kid in (("density", 0), ("divergence", 1))}
            except fe.FluidEngineError as err:
                if err.status != fe.FE_ERR_UNSTABLE:
                    raise
                r["exit_residual"] = {"unstable": str(err)}
            sim.solve_report_enable(False)
            if name == "learned" and mode == "two_sided":
                # inference kernel, event-timed over a few learned frames
                sim.profile(True)
                sim.profile_reset()
                sim.run(5)
                prof = sim.profile_results()
                sim.profile(False)
            sim.close()
            res[name] = r
        if all("sweeps_per_step" in res[k] for k in ("reference", "learned")):
            rs, ls = res["reference"]["sweeps_per_step"], res["learned"]["sweeps_per_step"]
            res["sweep_reduction"] = round((rs["density"] + rs["divergence"]) / (ls["density"] + ls["divergence"]), 3)
            res["speedup"] = round(res["learned"]["steps_per_s"] / res["reference"]["steps_per_s"], 3)
        modes[mode] = res
    # the density network alone: the divergence solve keeps the reference start, which
    # already needs few sweeps, and skips that network's inference and feature gathers
    sim = make(False)
    o = fe.FeOptions.from_buffer_copy(opts)
    o.residual = fe.FE_RESIDUAL_TWO_SIDED
    sim.learn_set_weights(fe.FE_NET_DENSITY, *nets[fe.FE_NET_DENSITY])
    o.warm_start = fe.FE_WARM_START_LEARNED
    sim.set_options(o)
    sim.reset_iteration_totals()
    sim.timer_start()
    sim.run(evaluate)
    ms = sim.timer_stop()
    d, v, f = sim.iteration_totals()
    sim.close()
    density_only = {"steps_per_s": round(evaluate / (ms / 1e3), 3),
                    "sweeps_per_step": {"density": round(d / f, 3), "divergence": round(v / f, 3)}}
    density_only["speedup"] = round(density_only["steps_per_s"] / modes["two_sided"]["reference"]["steps_per_s"], 3)
    out = {
        "workload": f"{w.name}: frames {collect}-{collect + evaluate - 1} after sampling frames 0-{collect - 1}",
        "samples_per_network": collect * 2048,
        "training": {"steps": 2000, "batch": 4096, "seconds_both_networks": round(train_s, 3),
                     "density_loss_first_last": [round(float(l0[:20].mean()), 4), round(float(l0[-50:].mean()), 4)],
                     "divergence_loss_first_last": [round(float(l1[:20].mean()), 4),
                                                    round(float(l1[-50:].mean()), 4)]},
        "exit_residual_frames": f"{collect + evaluate}-{collect + evaluate + report_frames - 1}, "
                                "units of the relative rate r dt / rho0",
        # the fair comparison: both starts stop on the two-sided residual
        "reference": modes["two_sided"]["reference"],
        "learned": modes["two_sided"]["learned"],
        "sweep_reduction": modes["two_sided"]["sweep_reduction"],
        "speedup": modes["two_sided"]["speedup"],
        "compression_residual": modes["compression"],
        "learned_density_only": density_only,
    }
    if "ws_infer" in prof:
        launches, total_ms = prof["ws_infer"]
        avg_s = total_ms / launches / 1e3
        tf = MLP_FLOP * w.n / avg_s / 1e12
        out["inference_roofline"] = {"bound": "mfma", "kernel": "ws_infer", "achieved": round(tf, 2),
                                     "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                                     "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 4),
                                     "avg_launch_us": round(avg_s * 1e6, 2)}
    return out


def batched_measurement(fe, scene, frames):
    """64 independent dam breaks (particle counts log-uniform in [4096, 262144], per-scene rest
    density and viscosity) in one batch context: every kernel of a frame is one launch over all
    scenes. Frames per second of the whole batch with fixed 4/4 sweeps (eager, and replayed from
    a captured HIP graph) and converging to 0.1 % / 1 % (each scene stops on its own), next to
    the same 64 scenes stepped as 64 separate contexts."""
    ws = scene.batch_scenes(count=64, seed=0)
    n_all = sum(w.n for w in ws)
    res = {"workload": "64 dam breaks, N_s log-uniform in [4096, 262144] (seed 0)", "particles": n_all,
           "scenes": len(ws)}

    def timed(sim, k):
        sim.run(2)
        sim.sync()
        sim.timer_start()
        sim.run(k)
        ms = sim.timer_stop()
        return k / (ms / 1e3)

    def opt(mode):
        o = fe.default_options()
        o.solver_mode = mode
        o.density_tol = 1e-3
        o.divergence_tol = 1e-2
        o.max_iterations = 100
        return o

    b = fe.Batch(ws, device=0, options=opt(fe.FE_SOLVER_FIXED))
    b.run(2)
    b.sync()
    tuned = b.autotune_now(reps=10)
    res["autotuned"] = {k: [v["block"], v["items"]] for k, v in tuned.items()}
    fps = timed(b, frames)
    res["fixed_4_4_eager"] = {"frames_per_s": round(fps, 3), "particle_updates_per_s": round(fps * n_all, 1)}
    b.use_graph(True)
    fps = timed(b, frames)
    res["fixed_4_4_graph"] = {"frames_per_s": round(fps, 3), "particle_updates_per_s": round(fps * n_all, 1)}
    b.close()
    b = fe.Batch(ws, device=0, options=opt(fe.FE_SOLVER_CONVERGE))
    for k, (blk, it) in res["autotuned"].items():
        b.set_tune(k, blk, it)
    fps = timed(b, frames)
    its = [b.scene_iterations(k)[1][0] for k in range(len(ws))]
    res["converge_eager"] = {"frames_per_s": round(fps, 3), "particle_updates_per_s": round(fps * n_all, 1),
                             "density_sweeps_per_frame_min_max": [min(its) / (frames + 2), max(its) / (frames + 2)]}
    b.close()
    # the same scenes as separate contexts, one frame each in turn
    sims = [fe.Simulation(w.params, w.positions, device=0, options=opt(fe.FE_SOLVER_FIXED)) for w in ws]
    for s in sims:
        s.run(2)
    for s in sims:
        s.sync()
    t0 = time.perf_counter()
    for _ in range(frames):
        for s in sims:
            s.step()
    for s in sims:
        s.sync()
    fps = frames / (time.perf_counter() - t0)
    for s in sims:
        s.close()
    res["fixed_4_4_separate_contexts"] = {"frames_per_s": round(fps, 3),
                                          "particle_updates_per_s": round(fps * n_all, 1)}
    res["batch_gain_fixed"] = round(res["fixed_4_4_graph"]["frames_per_s"] / fps, 3)
    return res


def cpu_baseline(w, opts, d_avg, v_avg, budget_s):
    """The oracle (FP32 C++ restatement, OpenMP) on the host cores: frames of the same
    workload with the GPU run's mean sweep counts, until the time budget is spent."""
    from oracle import oracle

    cores = min(os.cpu_count() or 1, 16)
    oracle.set_threads(cores)
    p = type(w.params).from_buffer_copy(w.params)
    p.density_solver_iterations = max(1, int(round(d_avg)))
    p.divergence_solver_iterations = max(1, int(round(v_avg)))
    orc = oracle.Oracle(p, w.positions)
    o = type(opts).from_buffer_copy(opts)
    o.solver_mode = 0
    orc.set_options(o)
    frames = 0
    t0 = time.perf_counter()
    while True:
        orc.step(p.dt)
        frames += 1
        el = time.perf_counter() - t0
        if el >= budget_s or el * (frames + 1) / frames > 2 * budget_s:
            break
    orc.close()
    return {"value": round(frames / el, 4), "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": f"{frames} frame(s) of {w.name} from rest, FP32 oracle, fixed "
                      f"{p.density_solver_iterations}/{p.divergence_solver_iterations} sweeps "
                      f"(the GPU run's mean per step), {el:.1f} s"}


if __name__ == "__main__":
    main()
