"""Throughput of the MI355X SPH step (one JSON line on stdout).

Default workload: 1,048,576 particles (1024x1024 block at 0.5h, seeded
jitter) in a 600h x 600h box with 4 walls and 64 circles; h=1, rho0=10,
k=0.004, k_near=0.01, sigma=0, beta=0.3, g=(0,-9.8), dt=0.005.  A step is one
full SPH step of that scene.

Protocol (inputs resident in HBM in every timed region):
  warm-up   W untimed steps on a throwaway instance, continued until the
            library's autotuner has finished its first tuning; its choice of
            kernel variants is carried into the timed instances;
  value     the median over --repeats fresh instances of K timed steps from
            the scene's initial state (steps 0..K-1), library defaults
            otherwise (eager launches, no per-kernel events, periodic
            re-tuning every 256 steps);
  tuning    the same K steps of a fresh instance that tunes from scratch
            inside the timed window, reported separately with its overhead.

roofline: the dominant kernel group's limiter.  Its wave64 VALU instructions
per launch (the rocprofv3 SQ_INSTS_VALU summary of this command in
profiles/valu_issue.json) over its average launch duration, measured here with
HIP events on the library's stream, against the VALU issue ceiling of 256 CUs
x 2.4 GHz x one instruction per CU per cycle; the HBM figure (algorithmic
bytes per launch over the same duration, against 8 TB/s) follows as the
secondary bound.

cpu_baseline: the reference's multi-threaded step restated in C++ (oracle/,
Gauss-Seidel order, pooled phases) on every CPU this process may use, and on
one thread, timed from GPU states spread over the timed window (steps 0..K);
the median over the samples.

surrogate (1M dam break): the learned relaxation is trained on the GPU on
plain frames of 200 seeded 16,384-particle dam breaks (one batched
simulation), then the same K steps are timed with the network relaxing every
step and the plain density check every 8th step (fallback above 1% density
error in any scene); beside it the ceiling of any relaxation-only network
and the trajectory of a 500-step surrogate run against the plain run.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s @1M particles: plain vs surrogate at ≤1% density error"
HBM_PEAK_GBS = 8000.0
# MI355X: 256 CUs at 2.4 GHz; a CU issues at most one wave64 VALU instruction
# per cycle (4 SIMD16 units, 4 cycles each).
VALU_PEAK_GINST = 256 * 2.4
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
TRAJECTORY_MARKS = (100, 250, 496, 500)  # 496: the last check step before 500 at the default interval


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


def host_cpu() -> dict:
    """The host CPUs this process may use: the affinity mask, capped by the
    cgroup's CPU quota when there is one (the reference sizes its pool to the
    logical core count, threading.h:135-138), and the model name (the
    reference shows it, app.cpp:28-34)."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, period = fh.read().split()[:2]
        if q != "max":
            quota = max(1, math.floor(int(q) / int(period)))
    except (OSError, ValueError):
        pass
    threads = min(affinity, quota) if quota else affinity
    return {"model": model, "os_cpu_count": os.cpu_count(), "affinity_cpus": affinity, "cgroup_cpu_quota": quota,
            "threads": threads}


def cpu_rates(states, specs, threads, steps):
    """Particle-steps/s of the oracle's reference-order step (GS_MT: the
    reference's pooled phases and partitioning) from each state: one untimed
    step rebuilds the neighbour lists, then `steps` timed steps."""
    from nbodysimulation_experiment_amd.scenes import SPHParams
    from oracle.oracle import GS_MT, OracleSim
    rates = []
    for st in states:
        o = OracleSim(SPHParams(), 32, GS_MT, threads)
        for s in specs:
            o.add_scene_spec(s)
        o.set_gravity(0.0, -9.8)
        o.set_state(st["x"], st["v"])
        o.step(DT, 1)
        t0 = time.perf_counter()
        o.step(DT, steps)
        t = time.perf_counter() - t0
        rates.append(o.n * steps / t)
        o.close()
    return rates


def train_surrogate(args):
    """Trains the learned relaxation on plain frames of the seeded training
    scenes; returns the parameter vector and a summary."""
    from nbodysimulation_experiment_amd import scenes
    from nbodysimulation_experiment_amd.scenes import SPHParams
    from nbodysimulation_experiment_amd.sim import ParticleSimulation
    specs = scenes.training_scenes(count=args.train_scenes, tall=args.train_tall)
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
    return params, {"scenes": len(specs), "tall_columns": args.train_tall, "particles": n,
                    "warm_steps": args.train_warm, "steps": args.train_steps,
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
    from nbodysimulation_experiment_amd import scenarios as sc
    from nbodysimulation_experiment_amd.scenes import volume_positions
    from oracle.oracle import GS_MT, OracleSim
    cpu = host_cpu()
    threads = args.cpu_threads or cpu["threads"]
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
           "cpu": {"kind": "port", "cores": threads, "model": cpu["model"], "iterations": args.scenario_cpu_iterations,
                   "order": "reference (Gauss-Seidel, pooled phases)"},
           "dt": dt, "scenarios": rows}
    print(json.dumps(out), flush=True)


def first_tuning_steps(n):
    """Steps until the autotuner's first tuning is done: one step that builds
    the first binning, then two rounds over the kernel variants (five below
    131,072 particles, where the wide kernels are a candidate, else four)."""
    return 1 + 2 * (5 if n <= 131072 else 4)


def timed_update(sim, steps):
    t0 = time.perf_counter()
    sim.Update(DT, steps)  # returns after the library's stream has drained
    return time.perf_counter() - t0


def compare_states(a, b):
    """Trajectory divergence of state a from state b (particle-index order)."""
    xa, xb = a["x"].astype(np.float64), b["x"].astype(np.float64)
    ra, rb = a["density"][:, 0].astype(np.float64), b["density"][:, 0].astype(np.float64)
    d = np.linalg.norm(xa - xb, axis=1)
    ka = 0.5 * float((a["v"].astype(np.float64) ** 2).sum())
    kb = 0.5 * float((b["v"].astype(np.float64) ** 2).sum())
    return {"density_error": float(np.abs(ra - rb).sum() / np.abs(rb).sum()),
            "position_max_h": float(d.max()), "position_mean_h": float(d.mean()),
            "mean_density_rel": float(abs(ra.mean() - rb.mean()) / rb.mean()),
            "kinetic_energy_rel": float(abs(ka - kb) / kb) if kb else None}


def trajectory(make_sim, spec, params, args):
    """A surrogate run (checks every --check-interval steps) and a plain run
    from the same initial state, compared at TRAJECTORY_MARKS; beside them a
    plain run whose initial positions differ by one ulp, which measures how
    far the flow itself amplifies the smallest perturbation."""
    from nbodysimulation_experiment_amd.scenes import SceneSpec
    bumped = SceneSpec(spec.width, spec.height, spec.cell_size, np.nextafter(spec.positions, np.float32(np.inf)),
                       planes=spec.planes, circles=spec.circles)
    plain, ulp, sur, car = make_sim(), make_sim([bumped]), make_sim(), make_sim()
    for s, carried in ((sur, False), (car, True)):
        s.surrogate_set_params(params)
        s.surrogate_set_mode(True, args.check_interval, args.max_density_error, carried_density=carried)
    out = {"surrogate_vs_plain": {}, "carried_density_vs_plain": {}, "plain_one_ulp_vs_plain": {}}
    done = 0
    for m in TRAJECTORY_MARKS:
        for s in (plain, ulp, sur, car):
            s.Update(DT, m - done)
        done = m
        p = plain.get_state()
        out["surrogate_vs_plain"][m] = compare_states(sur.get_state(), p)
        out["carried_density_vs_plain"][m] = compare_states(car.get_state(), p)
        out["plain_one_ulp_vs_plain"][m] = compare_states(ulp.get_state(), p)
    for key, s in (("surrogate_checks", sur), ("carried_density_checks", car)):
        rep = s.surrogate_report()
        out[key] = {k: rep[k] for k in ("check_steps", "fallbacks", "max_density_error", "mean_density_error")}
    out["note"] = ("between checks the carried-density run's state density is the last check's, so at marks that are "
                   "not check steps (100, 250, 500 with checks every 8 steps) its density is compared as of that "
                   "check; 496 is a check step")
    for s in (plain, ulp, sur, car):
        s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--repeats", type=int, default=5, help="timed repeats; value is their median")
    ap.add_argument("--workload", default="dam-break-1m",
                    choices=["dam-break-4k", "dam-break-1m", "batched-256", "batched-surrogate", "scenarios"])
    ap.add_argument("--scenario-cpu-iterations", type=int, default=2)
    ap.add_argument("--graph", action="store_true", help="replay steps as HIP graphs (no per-kernel timing)")
    ap.add_argument("--cpu-steps", type=int, default=3, help="timed CPU steps per sample (multi-threaded)")
    ap.add_argument("--cpu-samples", type=int, default=5, help="CPU samples spread over steps 0..K")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: every CPU this process may use")
    ap.add_argument("--no-cpu-single-thread", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="also run the library's copy kernel on MiB of data (3 times) for counter calibration")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-relaxation leg")
    ap.add_argument("--no-trajectory", action="store_true", help="skip the surrogate trajectory comparison")
    ap.add_argument("--no-graph-leg", action="store_true", help="skip the HIP-graph replay leg")
    ap.add_argument("--train-scenes", type=int, default=168, help="16,384-particle training dam breaks")
    ap.add_argument("--train-tall", type=int, default=8,
                    help="64 x 1,024-particle training columns, as tall as the 1M dam break's")
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
        print("bench.py: the library computes on one GPU; runs on more than one GPU are not supported", file=sys.stderr)
        sys.exit(2)

    if args.workload == "scenarios":
        scenario_benchmark(args)
        return

    from nbodysimulation_experiment_amd.scenes import SPHParams
    from nbodysimulation_experiment_amd.sim import ParticleSimulation, bench_copy

    if args.calibrate_copy:
        log(f"copy kernel on {args.calibrate_copy} MiB: {bench_copy(args.calibrate_copy << 20, 3):.1f} GB/s")
    copy_gbs = bench_copy(1 << 30, 20)
    log(f"device copy ceiling {copy_gbs:.1f} GB/s")

    specs, desc = build_workload(args.workload)
    n = sum(s.n for s in specs)
    K = args.steps
    log(f"workload {args.workload}: {desc}")
    fixed_caps = None if args.tile_cap == "auto" else [int(c) for c in args.tile_cap.split(",")]

    def make_sim(scene_specs=None, caps=None, hip_graph=False):
        """A fresh instance; caps: an earlier tuning's choice (the periodic
        re-tuning still runs), fixed_caps: no tuning at all."""
        sp = scene_specs or specs
        sim = ParticleSimulation.from_scene(sp[0], SPHParams()) if len(sp) == 1 else \
            ParticleSimulation.from_scenes(sp, SPHParams())
        sim.set_options(hip_graph=hip_graph or args.graph)
        if fixed_caps is not None:
            sim.set_tuning(autotune=False, tile_cap=fixed_caps)
        elif caps is not None:
            sim.set_tuning(autotune=True, tile_cap=caps, tuned=True)
        return sim

    # Warm-up on a throwaway instance, continued until the first tuning is done.
    warm_steps = args.warmup if fixed_caps is not None else max(args.warmup, first_tuning_steps(n))
    log(f"warm-up {warm_steps} steps")
    warm = make_sim()
    warm.Update(DT, warm_steps)
    caps = list(warm.get_tuning()["tile_cap"]) if fixed_caps is None else fixed_caps
    warm.close()
    log(f"kernel variants carried into the timed instances: {caps}")

    # The headline: K steps of fresh instances, median over the repeats.
    times = []
    for r in range(args.repeats):
        sim = make_sim(caps=caps)
        sim.ResetStats()
        times.append(timed_update(sim, K))
        if r == args.repeats - 1:
            tuning_state = sim.get_tuning()
            stats = sim.GetStats()
            diag = sim.diagnostics()
        sim.close()
    med = statistics.median(times)
    value = n * K / med
    log(f"{value:.4g} particle-steps/s, {1e3 * med / K:.4f} ms/step (median of {len(times)}: "
        f"{', '.join(f'{1e3 * t / K:.4f}' for t in times)} ms/step)")

    tuning = {"carried_tile_cap": caps, "warm_up_steps": warm_steps, "autotune": fixed_caps is None,
              "retune_interval": tuning_state["retune_interval"], "tile_cap_at_end": list(tuning_state["tile_cap"])}
    if fixed_caps is None:
        # What a fresh caller pays: the first tuning runs inside the window.
        tsim = make_sim()
        t_tune = timed_update(tsim, K)
        tsim.close()
        tuning["from_scratch_in_window"] = {
            "value": round(n * K / t_tune, 1), "ms_per_step": round(1e3 * t_tune / K, 5),
            "tuning_steps": min(K, first_tuning_steps(n)), "overhead_ms": round(1e3 * (t_tune - med), 3)}
        log(f"tuning from scratch inside the window: {n * K / t_tune:.4g} particle-steps/s")

    # Per-kernel averages for the roofline: the same K steps with a HIP event
    # pair around every kernel group on the library stream (the events add a
    # little time at each boundary).
    kt, evented = {}, None
    if not args.graph:
        esim = make_sim(caps=caps)
        esim.set_timing(True)
        e_el = timed_update(esim, K)
        kt = esim.kernel_times()
        esim.close()
        evented = {"value": round(n * K / e_el, 1), "unit": "particle-steps/s", "ms_per_step": round(1e3 * e_el / K, 5),
                   "kernels_ms": {k: round(v[0], 5) for k, v in kt.items()}}
        log(f"with events {n * K / e_el:.4g} particle-steps/s; kernels {evented['kernels_ms']}")

    graph_leg = None
    if not args.graph and not args.no_graph_leg:
        gsim = make_sim(caps=caps, hip_graph=True)
        g_el = timed_update(gsim, K)
        gsim.close()
        graph_leg = {"value": round(n * K / g_el, 1), "unit": "particle-steps/s", "ms_per_step": round(1e3 * g_el / K, 5)}
        log(f"graph replay {n * K / g_el:.4g} particle-steps/s")

    roofline = None
    if kt:
        dom = max(kt, key=lambda k: kt[k][0])
        avg_ms = kt[dom][0]
        alg = KERNEL_BYTES[dom] * n
        hbm_gbs = alg / (avg_ms * 1e-3) / 1e9
        traffic = None
        tpath = os.path.join(ROOT, "profiles", "traffic.json")
        if os.path.exists(tpath):
            t = json.load(open(tpath))
            if t.get("workload") == args.workload and dom in t.get("kernels", {}):
                traffic = t["kernels"][dom]
        hbm = {"bound": "hbm", "achieved": round(hbm_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(hbm_gbs / HBM_PEAK_GBS, 5), "traffic": traffic, "alg_bytes_per_launch": alg,
               "copy_ceiling_gbs": round(copy_gbs, 1),
               "step_alg_bytes_frac": round(value * STEP_ALG_BYTES / 1e9 / HBM_PEAK_GBS, 5)}
        valu = None
        vpath = os.path.join(ROOT, "profiles", "valu_issue.json")
        if os.path.exists(vpath):
            v = json.load(open(vpath))
            # the counters of this command's window: the profiled step count,
            # or a window of another step count from the 500-step counter run
            win = v if v.get("steps") == K else v.get("windows", {}).get(str(K))
            ins = (win or {}).get("valu_insts_per_launch", {}).get(dom)
            if v.get("workload") == args.workload and ins:
                g = ins / (avg_ms * 1e-3) / 1e9
                valu = {"bound": "valu", "achieved": round(g, 2), "peak": VALU_PEAK_GINST,
                        "unit": "G wave64 VALU instructions/s", "frac": round(g / VALU_PEAK_GINST, 5),
                        "valu_insts_per_launch": ins, "issue_frac_under_counters": win["kernels"].get(dom),
                        "source": "profiles/valu_issue.json (rocprofv3 SQ_INSTS_VALU of this command's window)"}
        primary = valu if valu is not None and valu["frac"] > hbm["frac"] else hbm
        roofline = dict(primary)
        roofline.update({"traffic": traffic, "kernel": dom, "kernel_avg_ms": round(avg_ms, 5),
                         "hbm": hbm, "valu": valu})

    surrogate = None
    online = args.workload == "batched-surrogate"  # in-run training on each check step's plain frame
    if args.workload in ("dam-break-1m", "batched-surrogate") and not args.no_surrogate and not args.graph:
        log(f"training the learned relaxation: {args.train_scenes} scenes + {args.train_tall} tall columns, "
            f"{args.train_steps} steps")
        params, train = train_surrogate(args)
        log(f"training: {train}")

        def sur_sim(timing=False, carried=False):
            s = make_sim(caps=caps)
            s.surrogate_set_params(params)
            s.surrogate_set_mode(True, args.check_interval, args.max_density_error, carried_density=carried)
            if online:
                s.surrogate_set_online_training(args.lr)
            s.set_timing(timing)
            return s

        warm = sur_sim()
        warm.Update(DT, args.warmup)
        warm.close()
        s_times = []
        for _ in range(max(1, min(3, args.repeats))):
            ssim = sur_sim()
            s_times.append(timed_update(ssim, K))
            rep = ssim.surrogate_report()
            ssim.close()
        s_med = statistics.median(s_times)
        s_value = n * K / s_med
        carried = None
        if not online:
            # the density pass only on check steps (features from the last check's density)
            c_times = []
            for _ in range(max(1, min(3, args.repeats))):
                csim = sur_sim(carried=True)
                c_times.append(timed_update(csim, K))
                crep = csim.surrogate_report()
                csim.close()
            c_med = statistics.median(c_times)
            carried = {"value": round(n * K / c_med, 1), "ms_per_step": round(1e3 * c_med / K, 5),
                       "speedup_vs_plain": round((n * K / c_med) / value, 4),
                       "density_error": {"max": crep["max_density_error"], "mean": crep["mean_density_error"],
                                         "checks": crep["check_steps"], "fallbacks": crep["fallbacks"]},
                       "definition": "density pass only on check steps; the network steps in between take their "
                                     "density features from the last accepted check, carried through the binning"}
            log(f"surrogate with carried density {carried['value']:.4g} particle-steps/s "
                f"({carried['speedup_vs_plain']:.3f}x plain)")
        esim = sur_sim(timing=True)
        s_el = timed_update(esim, K)
        s_kt = esim.kernel_times()
        erep = esim.surrogate_report()
        sdiag = esim.diagnostics()
        scene_err = esim.surrogate_scene_errors() if online else None
        esim.close()
        relax_ms = s_kt["delta"][0]
        pairs = erep["last_pair_count"]
        flops = 2.0 * HID * FEAT * pairs + 2.0 * 2 * HID * HID * n
        plain_ms = 1e3 * med / K
        surrogate = {
            "value": round(s_value, 1), "unit": "particle-steps/s", "ms_per_step": round(1e3 * s_med / K, 5),
            "speedup_vs_plain": round(s_value / value, 4),
            "speedup_vs_plain_with_events": round((n * K / s_el) / evented["value"], 4) if evented else None,
            "density_error": {"definition": "sum|rho(surrogate) - rho(plain)| / sum rho(plain) on check steps, "
                                            "from the same input (one step)",
                              "max": rep["max_density_error"], "mean": rep["mean_density_error"],
                              "limit": args.max_density_error, "check_interval": args.check_interval,
                              "checks": rep["check_steps"], "fallbacks": rep["fallbacks"]},
            "carried_density": carried,
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
                "learning_rate": args.lr, "adam_steps": erep["train_steps"],
                "scene_density_error": {"scenes": int(scene_err.size), "min": float(scene_err.min()),
                                        "max": float(scene_err.max()), "mean": float(scene_err.mean()),
                                        "largest_scene": float(scene_err[-1]),
                                        "note": "last check step, per scene; the last scene is the 4M dam break"}},
            "kernels_ms": {k: round(v[0], 5) for k, v in s_kt.items()},
            "diagnostics_fp64": {k: sdiag[k] for k in ("mean_density", "max_density", "non_finite_count")},
        }
        if kt:
            step_ms = sum(v[0] for v in kt.values())
            surrogate["ceiling_vs_plain"] = {
                "value": round(step_ms / (step_ms - kt["delta"][0]), 4),
                "definition": "plain step / (plain step - plain relaxation kernel), per-kernel event averages: the "
                              "speed-up of a relaxation network that cost nothing (the relaxation kernel also "
                              "holds collisions and the velocity update, which the network step still runs)"}
        log(f"surrogate {s_value:.4g} particle-steps/s ({s_value / value:.3f}x plain, plain {plain_ms:.4f} ms/step); "
            f"report {rep}")
        if args.workload == "dam-break-1m" and not args.no_trajectory:
            log(f"trajectory: surrogate and plain runs to step {TRAJECTORY_MARKS[-1]}")
            surrogate["trajectory"] = trajectory(lambda sp=None: make_sim(sp, caps=caps), specs[0], params, args)
            log(f"trajectory: {surrogate['trajectory']}")

    cpu = None
    if not args.no_cpu_baseline:
        host = host_cpu()
        threads = args.cpu_threads or host["threads"]
        marks = sorted({round(k * K / max(1, args.cpu_samples - 1)) for k in range(args.cpu_samples)})
        ssim = make_sim(caps=caps)
        states, done = [], 0
        for m in marks:
            ssim.Update(DT, m - done)
            done = m
            s = ssim.get_state()
            states.append({"x": s["x"], "v": s["v"]})
        ssim.close()
        log(f"cpu baseline: {threads} threads, {args.cpu_steps} steps from each of steps {marks}")
        t0 = time.perf_counter()
        mt = cpu_rates(states, specs, threads, args.cpu_steps)
        t_mt = time.perf_counter() - t0
        cv = statistics.median(mt)
        cpu = {"value": round(cv, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "model": host["model"], "host": host,
               "sample": f"median of {len(mt)} samples: {args.cpu_steps} timed steps of the same scene from the GPU "
                         f"states at steps {marks} (the timed window; +1 untimed step each to rebuild the neighbour "
                         f"lists), {t_mt:.1f} s in all",
               "samples": [round(r, 1) for r in mt], "gpu_over_cpu": round(value / cv, 1)}
        log(f"cpu baseline {cv:.4g} particle-steps/s ({threads} threads)")
        if not args.no_cpu_single_thread:
            st = cpu_rates(states, specs, 1, 1)
            cpu["single_thread"] = {"value": round(statistics.median(st), 1), "cores": 1,
                                    "samples": [round(r, 1) for r in st],
                                    "sample": f"median of {len(st)} samples, 1 timed step from each state",
                                    "gpu_over_cpu": round(value / statistics.median(st), 1)}
            log(f"cpu single thread {statistics.median(st):.4g} particle-steps/s")

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": round(1e3 * med / K, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": desc, "particles": n, "dt": DT, "method": "plain", "hip_graph": args.graph,
                   "repeats": len(times), "tuning": tuning},
        "repeats_ms_per_step": [round(1e3 * t / K, 5) for t in times],
        "roofline": roofline, "cpu_baseline": cpu, "per_kernel_events": evented, "hip_graph_replay": graph_leg,
        "surrogate": surrogate,
        "stats": {k: v for k, v in stats.items() if k != "time"},
        "diagnostics_fp64": {k: diag[k] for k in ("mean_density", "max_density", "mean_abs_density_error",
                                                  "kinetic_energy", "non_finite_count")},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
