"""Throughput of the MI355X SPH step (one JSON line on stdout).

Default workload: 1,048,576 particles (1024x1024 block at 0.5h, seeded
jitter) in a 600h x 600h box with 4 walls and 64 circles; h=1, rho0=10,
k=0.004, k_near=0.01, sigma=0, beta=0.3, g=(0,-9.8), dt=0.005.  A step is one
full SPH step of that scene.  W warm-up steps run untimed; then K steps are
timed with the inputs resident in HBM: the warm-up runs on a throwaway
instance and the timed steps are steps 0..K-1 of a fresh instance.

roofline: the dominant kernel's algorithmic bytes per launch divided by its
average duration, measured with HIP events on the library's stream over the
timed steps.  cpu_baseline: the reference's multi-threaded step restated in
C++ (oracle/, Gauss-Seidel order, pooled phases), timed on this host on a
bounded sample that starts from the GPU state at the end of the timed window.

surrogate (1M dam break): the learned relaxation is trained on the GPU on
plain frames of 200 seeded 16,384-particle dam breaks (one batched
simulation, --train-steps Adam steps after --train-warm plain steps), then
the same K steps are timed with the network relaxing every step and the
plain density check every 8th step (fallback above 1% density error).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s @1M particles: plain vs surrogate at ≤1% density error"
HBM_PEAK_GBS = 8000.0
DT = 0.005

# Algorithmic bytes per particle of each kernel group: every particle record
# crosses HBM once per phase, neighbour re-reads are on-chip.
#   viscosity_predict: R x,v (16) + W xpred (8) + W key, fill (8) + R key (4)  = 36
#   grid:              R key, fill (8) + W slot, id (8) + R slot, id (8)
#                      + R xpred, x (16) + W x, prev, id, key (24)              = 64
#   density:           R x, key (12) + W density, pressure (16)                = 28
#   delta:             R x, key, pressure, prev (28) + W x, v (16)             = 44
KERNEL_BYTES = {"viscosity_predict": 36, "grid": 64, "density": 28, "delta": 44}
STEP_ALG_BYTES = 172  # reference phases, one SoA FP32 pass each
MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16
HID, FEAT = 64, 16


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name):
    from nbodysimulation_experiment_amd import scenes
    if name == "dam-break-1m":
        return [scenes.dam_break_1m()], "dam break, 1,048,576 particles (1024x1024 at 0.5h), 600h x 600h box, 4 walls + 64 circles"
    if name == "dam-break-4k":
        return [scenes.dam_break_4k()], "dam break, 4,096 particles (64x64 at 0.5h), 80h x 60h box, 4 walls"
    if name == "batched-256":
        sp = scenes.batched_scenes()
        return sp, f"{len(sp)} independent scenes, {sum(s.n for s in sp):,} particles, counts log-uniform in [2048, 65536]"
    if name == "batched-surrogate":
        sp = scenes.batched_scenes() + [scenes.dam_break_4m()]
        return sp, (f"{len(sp) - 1} independent scenes (counts log-uniform in [2048, 65536]) plus one 4,194,304-particle "
                    f"dam break (2048x2048 at 0.5h, 1536h x 1536h box), {sum(s.n for s in sp):,} particles")
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(state, specs, threads, sample_steps, warm_steps=1):
    """Times the oracle's multi-threaded reference-order step on a sample that
    starts from the GPU state after warm-up."""
    from nbodysimulation_experiment_amd.scenes import SPHParams
    from oracle.oracle import GS_MT, OracleSim
    o = OracleSim(SPHParams(), 32, GS_MT, threads)
    for s in specs:
        o.add_scene_spec(s)
    o.set_gravity(0.0, -9.8)
    o.set_state(state["x"], state["v"])
    o.step(DT, warm_steps)  # rebuilds the neighbour lists
    t0 = time.perf_counter()
    o.step(DT, sample_steps)
    t = time.perf_counter() - t0
    n = o.n
    o.close()
    return n * sample_steps / t, t


def train_surrogate(args):
    """Trains the learned relaxation on plain frames of the seeded training
    scenes; returns the parameter vector and a summary."""
    from nbodysimulation_experiment_amd import scenes
    from nbodysimulation_experiment_amd.scenes import SPHParams
    from nbodysimulation_experiment_amd.sim import ParticleSimulation
    specs = scenes.training_scenes(count=args.train_scenes)
    sim = ParticleSimulation.from_scenes(specs, SPHParams())
    sim.Update(DT, args.train_warm)
    sim.surrogate_init(0)
    sim.surrogate_calibrate(DT)
    t0 = time.perf_counter()
    losses, rel = sim.surrogate_train(DT, args.train_steps, args.lr, relative=True)
    t = time.perf_counter() - t0
    params = sim.surrogate_params()
    n = sim.GetParticleCount()
    sim.close()
    k = max(1, len(losses) // 10)
    return params, {"scenes": len(specs), "particles": n, "warm_steps": args.train_warm, "steps": args.train_steps,
                    "learning_rate": args.lr, "seconds": round(t, 3),
                    "particle_steps_per_s": round(n * args.train_steps / t, 1),
                    "relative_error_first": float(rel[:k].mean()), "relative_error_last": float(rel[-k:].mean()),
                    "relative_error_definition": "sum|o - t/os|^2 / sum|t/os|^2 per training frame (first / last "
                                                 "10% of the steps)"}


def scenario_benchmark(args):
    """The reference application's built-in benchmark (64 frames x 16
    iterations per scenario, per-phase min / max / avg) on the GPU, beside the
    reference-order multi-threaded CPU step on the same frames (fewer
    iterations, since its cost is the same in every iteration)."""
    import numpy as np

    from nbodysimulation_experiment_amd import scenarios as sc
    from nbodysimulation_experiment_amd.scenes import volume_positions
    from oracle.oracle import GS_MT, OracleSim
    threads = args.cpu_threads or min(16, int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1)
    dt = float(sc.CONSTANTS["delta_time"])
    rows = []
    for idx, s in enumerate(sc.SCENARIOS):
        with sc.make_simulation(s) as sim:
            g = sc.run_benchmark(sim, idx, frames=sc.BENCHMARK_FRAMES, iterations=sc.BENCHMARK_ITERATIONS)
        cpu_ms = []
        for _ in range(args.scenario_cpu_iterations):
            o = OracleSim(sc.scenario_params(s), 32, GS_MT, threads)
            c = sc.CONSTANTS
            scene = o.add_scene(float(c["boundary_width"]), float(c["boundary_height"]), float(c["grid_cell_size"]))
            o.set_spacing(s.params["particle_spacing"])
            o.set_gravity(*map(float, s.gravity))
            for b in s.bodies:
                r = sc.resolve_body(b)
                if r["kind"] == "plane":
                    o.add_plane(scene, r["normal"][0], r["normal"][1], r["distance"])
                elif r["kind"] == "circle":
                    o.add_circle(scene, r["center"][0], r["center"][1], r["radius"])
                else:
                    o.add_polygon(scene, np.array(r["verts"], np.float32))
            for v in s.volumes:
                nx, ny = sc.volume_counts(v, s.params["particle_spacing"])
                pos = volume_positions(v.position, nx, ny, float(s.params["particle_spacing"]),
                                       float(s.params["kernel_height"]), 0)
                o.add_particles(scene, pos, np.tile(np.array(v.force, np.float32), (pos.shape[0], 1)))
            for e in s.emitters:
                o.add_emitter(scene, e.position, e.direction, e.radius, e.speed, e.rate, e.duration, seed=0)
            for _ in range(sc.BENCHMARK_FRAMES):
                t0 = time.perf_counter()
                o.step(dt, 1)
                cpu_ms.append((time.perf_counter() - t0) * 1e3)
            o.close()
        row = {"scenario": s.name, "particles_max": int(g["particles"][1]),
               "gpu_ms": [round(x, 4) for x in g["simulation_ms"]],
               "gpu_phases_avg_ms": {k: round(v[2], 4) for k, v in g["phases_ms"].items() if v[2] > 0},
               "cpu_ms": [round(min(cpu_ms), 4), round(max(cpu_ms), 4), round(sum(cpu_ms) / len(cpu_ms), 4)]}
        row["cpu_over_gpu_avg"] = round(row["cpu_ms"][2] / row["gpu_ms"][2], 2)
        rows.append(row)
        log(f"{s.name}: GPU avg {row['gpu_ms'][2]:.3f} ms/frame, CPU ({threads} threads) avg {row['cpu_ms'][2]:.3f}")
    out = {"metric": "built-in benchmark: ms per frame (min, max, avg) per scenario", "unit": "ms",
           "higher_is_better": False, "frames": sc.BENCHMARK_FRAMES, "iterations": sc.BENCHMARK_ITERATIONS,
           "cpu": {"kind": "port", "cores": threads, "iterations": args.scenario_cpu_iterations,
                   "order": "reference (Gauss-Seidel, pooled phases)"},
           "dt": dt, "scenarios": rows}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="dam-break-1m",
                    choices=["dam-break-4k", "dam-break-1m", "batched-256", "batched-surrogate", "scenarios"])
    ap.add_argument("--scenario-cpu-iterations", type=int, default=2)
    ap.add_argument("--graph", action="store_true", help="replay steps as HIP graphs (no per-kernel timing)")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="also run the library's copy kernel on MiB of data (3 times) for counter calibration")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-relaxation leg")
    ap.add_argument("--no-graph-leg", action="store_true", help="skip the HIP-graph replay leg")
    ap.add_argument("--train-scenes", type=int, default=200)
    ap.add_argument("--train-warm", type=int, default=100)
    ap.add_argument("--train-steps", type=int, default=500)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--check-interval", type=int, default=8)
    ap.add_argument("--max-density-error", type=float, default=0.01)
    ap.add_argument("--tile-cap", default="auto",
                    help="LDS tile capacities of the viscosity, density and relaxation kernels, e.g. 2048,2048,1024; "
                         "'auto' lets the library's autotuner choose")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build computes on one GPU; runs on more than one GPU are not supported", file=sys.stderr)
        sys.exit(2)

    if args.workload == "scenarios":
        scenario_benchmark(args)
        return

    from nbodysimulation_experiment_amd.sim import ParticleSimulation
    from nbodysimulation_experiment_amd.scenes import SPHParams

    from nbodysimulation_experiment_amd.sim import bench_copy
    if args.calibrate_copy:
        log(f"copy kernel on {args.calibrate_copy} MiB: {bench_copy(args.calibrate_copy << 20, 3):.1f} GB/s")
    copy_gbs = bench_copy(1 << 30, 20)
    log(f"device copy ceiling {copy_gbs:.1f} GB/s")

    specs, desc = build_workload(args.workload)
    n = sum(s.n for s in specs)
    log(f"workload {args.workload}: {desc}")

    def make_sim():
        sim = ParticleSimulation.from_scene(specs[0], SPHParams()) if len(specs) == 1 else \
            ParticleSimulation.from_scenes(specs, SPHParams())
        sim.set_options(hip_graph=args.graph)
        if args.tile_cap != "auto":
            sim.set_tuning(autotune=False, tile_cap=[int(c) for c in args.tile_cap.split(",")])
        return sim

    # Warm-up on a throwaway instance, then K timed steps of a fresh instance
    # from the scene's initial state (the inputs are already in HBM).
    log(f"warm-up {args.warmup} steps")
    warm = make_sim()
    warm.Update(DT, args.warmup)
    warm.close()
    # The headline: the library as a caller runs it by default (per-kernel
    # event timing off).
    sim = make_sim()
    sim.ResetStats()
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    sim.Update(DT, args.steps)  # returns after the stream has drained
    elapsed = time.perf_counter() - t0
    value = n * args.steps / elapsed
    tuning = sim.get_tuning()
    stats = sim.GetStats()
    diag = sim.diagnostics()
    end_state = sim.get_state() if not args.no_cpu_baseline else None
    sim.close()
    log(f"{value:.4g} particle-steps/s, {1e3 * elapsed / args.steps:.4f} ms/step")

    # Per-kernel averages for the roofline: the same K steps of a fresh
    # instance with a HIP event pair around every kernel group on the library
    # stream (the events add a little time at each boundary).
    kt, evented = {}, None
    if not args.graph:
        esim = make_sim()
        esim.set_timing(True)
        log(f"timing {args.steps} steps with per-kernel events")
        t0 = time.perf_counter()
        esim.Update(DT, args.steps)
        e_elapsed = time.perf_counter() - t0
        kt = esim.kernel_times()
        esim.close()
        evented = {"value": round(n * args.steps / e_elapsed, 1), "unit": "particle-steps/s",
                   "ms_per_step": round(1e3 * e_elapsed / args.steps, 5)}
        log(f"with events {n * args.steps / e_elapsed:.4g} particle-steps/s, "
            f"{1e3 * e_elapsed / args.steps:.4f} ms/step; kernels {kt}")

    graph_leg = None
    if not args.graph and not args.no_graph_leg:
        # The same K steps of a fresh instance with steady-state steps replayed
        # as HIP graphs (no per-kernel events; tuning steps still run eagerly).
        gsim = make_sim()
        gsim.set_options(hip_graph=True)
        log(f"timing {args.steps} steps with HIP-graph replay")
        t0 = time.perf_counter()
        gsim.Update(DT, args.steps)
        g_elapsed = time.perf_counter() - t0
        gsim.close()
        graph_leg = {"value": round(n * args.steps / g_elapsed, 1), "unit": "particle-steps/s",
                     "ms_per_step": round(1e3 * g_elapsed / args.steps, 5)}
        log(f"graph replay {n * args.steps / g_elapsed:.4g} particle-steps/s")

    roofline = None
    if not args.graph:
        dom = max(kt, key=lambda k: kt[k][0])
        avg_ms = kt[dom][0]
        alg = KERNEL_BYTES[dom] * n
        achieved = alg / (avg_ms * 1e-3) / 1e9
        traffic = None
        tpath = os.path.join(ROOT, "profiles", "traffic.json")
        if os.path.exists(tpath):
            t = json.load(open(tpath))
            if t.get("workload") == args.workload and dom in t.get("kernels", {}):
                traffic = t["kernels"][dom]
        # the kernel's actual limiter: VALU issue, from the instruction-counter profile
        valu = None
        vpath = os.path.join(ROOT, "profiles", "valu_issue.json")
        if os.path.exists(vpath):
            v = json.load(open(vpath))
            if v.get("workload") == args.workload and dom in v.get("kernels", {}):
                valu = v["kernels"][dom]
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "kernel": dom,
                    "kernel_avg_ms": round(avg_ms, 5), "alg_bytes_per_launch": alg,
                    "copy_ceiling_gbs": round(copy_gbs, 1),
                    "step_alg_bytes_frac": round(value * STEP_ALG_BYTES / 1e9 / HBM_PEAK_GBS, 5),
                    "valu_issue_frac": valu,
                    "valu_issue_source": "profiles/valu_issue.json (rocprofv3 SQ_INSTS_VALU / GRBM_GUI_ACTIVE)"
                    if valu is not None else None}

    surrogate = None
    online = args.workload == "batched-surrogate"  # in-run training on each check step's plain frame
    if args.workload in ("dam-break-1m", "batched-surrogate") and not args.no_surrogate and not args.graph:
        log(f"training the learned relaxation: {args.train_scenes} scenes, {args.train_steps} steps")
        params, train = train_surrogate(args)
        log(f"training: {train}")
        warm = make_sim()
        warm.surrogate_set_params(params)
        warm.surrogate_set_mode(True, args.check_interval, args.max_density_error)
        warm.Update(DT, args.warmup)
        warm.close()
        ssim = make_sim()
        ssim.surrogate_set_params(params)
        ssim.surrogate_set_mode(True, args.check_interval, args.max_density_error)
        if online:
            ssim.surrogate_set_online_training(args.lr)
        ssim.set_timing(True)
        ssim.ResetStats()
        log(f"timing {args.steps} surrogate steps")
        t0 = time.perf_counter()
        ssim.Update(DT, args.steps)
        s_elapsed = time.perf_counter() - t0
        s_kt = ssim.kernel_times()
        rep = ssim.surrogate_report()
        sdiag = ssim.diagnostics()
        scene_err = ssim.surrogate_scene_errors() if online else None
        ssim.close()
        s_value = n * args.steps / s_elapsed
        plain_evented = evented["value"] if evented else value
        relax_ms = s_kt["delta"][0]
        pairs = rep["last_pair_count"]
        flops = 2.0 * HID * FEAT * pairs + 2.0 * 2 * HID * HID * n
        surrogate = {
            "value": round(s_value, 1), "unit": "particle-steps/s", "ms_per_step": round(1e3 * s_elapsed / args.steps, 5),
            # both runs with per-kernel events
            "speedup_vs_plain": round(s_value / plain_evented, 4),
            "density_error": {"definition": "sum|rho(surrogate) - rho(plain)| / sum rho(plain) on check steps",
                              "max": rep["max_density_error"], "mean": rep["mean_density_error"],
                              "limit": args.max_density_error, "check_interval": args.check_interval,
                              "checks": rep["check_steps"], "fallbacks": rep["fallbacks"]},
            "steps_network": rep["surrogate_steps"], "steps_plain": rep["plain_steps"],
            "relaxation_ms": round(relax_ms, 5),
            "mfma": {"flops_per_step": flops, "pairs_per_step": pairs,
                     "achieved_tflops": round(flops / (relax_ms * 1e-3) / 1e12, 3),
                     "peak_tflops": MFMA_BF16_PEAK_TFLOPS,
                     "frac": round(flops / (relax_ms * 1e-3) / 1e12 / MFMA_BF16_PEAK_TFLOPS, 6),
                     "note": "pair layer 2*64*16 flops per pair + head 2*2*64*64 per particle, over the relaxation "
                             "phase's average time (check steps add a plain relaxation and two density passes)"},
            "training": train,
            "in_run_training": None if not online else {
                "learning_rate": args.lr, "adam_steps": rep["train_steps"],
                "scene_density_error": {"scenes": int(scene_err.size), "min": float(scene_err.min()),
                                        "max": float(scene_err.max()), "mean": float(scene_err.mean()),
                                        "largest_scene": float(scene_err[-1]),
                                        "note": "last check step, per scene; the last scene is the 4M dam break"}},
            "kernels_ms": {k: round(v[0], 5) for k, v in s_kt.items()},
            "diagnostics_fp64": {k: sdiag[k] for k in ("mean_density", "max_density", "non_finite_count")},
        }
        log(f"surrogate {s_value:.4g} particle-steps/s ({s_value / plain_evented:.3f}x plain); report {rep}")

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1)
        log(f"cpu baseline: {threads} threads, {args.cpu_steps} steps from step {args.steps}")
        cv, ct = cpu_baseline(end_state, specs, threads, args.cpu_steps)
        cpu = {"value": round(cv, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} steps of the same scene from the GPU state at the end of the timed "
                         f"window (step {args.steps}; +1 untimed step to rebuild the neighbour lists), {ct:.1f} s",
               "gpu_over_cpu": round(value / cv, 1)}
        log(f"cpu baseline {cv:.4g} particle-steps/s")

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": desc, "particles": n, "dt": DT, "method": "plain", "hip_graph": args.graph,
                   "tuning": tuning},
        "roofline": roofline, "cpu_baseline": cpu, "per_kernel_events": evented, "hip_graph_replay": graph_leg,
        "surrogate": surrogate,
        "stats": {k: v for k, v in stats.items() if k != "time"},
        "diagnostics_fp64": {k: diag[k] for k in ("mean_density", "max_density", "mean_abs_density_error",
                                                  "kinetic_energy", "non_finite_count")},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
