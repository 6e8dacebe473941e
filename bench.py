#!/usr/bin/env python3
"""AVBD step throughput on one MI355X.

Workload (default): 100,000 random boxes in a 50 x 50 column lattice, 40
layers, spacing 2 m, jitter +-0.1 m, half extents U[0.25, 0.75] m, density
1000, friction U[0.3, 0.8], uniform random orientations, dropped into a
110 x 110 m walled container; dt 1/60, 10 iterations, FP32.  A step is one
full AVBD substep (TOI, broadphase, narrowphase and merge, colouring, 10 x
(primal sweep + dual update), BDF1, sleep) over the whole scene.

`--warmup W` untimed steps are followed by exactly `--steps K` timed steps
(defaults 10 + 290: the workload's 300-step trajectory).  The JSON line
reports body-steps/s, the primal-solve kernel against MI355X's 8 TB/s HBM
roofline (algorithmic bytes / summed launch time, from HIP events around
every launch), and the CPU oracle timed on this host on a bounded sample of
the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "body-steps/s @100k boxes (10 AVBD iters) + % HBM roofline; batched scenes/s"


TAIL = (-1, 0)  # fused primal tail setting for every engine (-1: library default)


def new_engine(**kw):
    from ornis_amd.avbd import AvbdEngine
    eng = AvbdEngine(**kw)
    if TAIL[0] >= 0:
        eng.set_primal_tail(*TAIL)
    return eng


def host_cores():
    """Host threads this process may use (the affinity mask; at most 16,
    the CPU share of one GPU on the box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(scene, state, start_step, seconds, max_steps, threads):
    """The CPU oracle in the reference's own order (reverse-handle sweep,
    history-ordered pairs, libm) with pair discovery on `threads` threads
    and a single-threaded sweep and dual update, as the reference runs them
    (rows.rs:427, :769-782; mod.rs:1052-1059), on a bounded sample of the
    same workload: the steps that follow step `start_step` of the GPU
    trajectory (its state there becomes the oracle's initial state; contact
    history is rebuilt), until `seconds` of CPU time or `max_steps` steps."""
    from oracle import pyoracle as po
    s = dict(scene)
    s["pos"], s["quat"], s["vel"], s["angvel"] = (np.ascontiguousarray(state[:, 0:3]),
                                                  np.ascontiguousarray(state[:, 3:7]),
                                                  np.ascontiguousarray(state[:, 7:10]),
                                                  np.ascontiguousarray(state[:, 10:13]))
    n_dyn = int((scene["body_type"] == 1).sum())
    e = po.OracleEngine(s, order=po.ORDER_REFERENCE, math=po.MATH_LIBM)
    e.set_threads(threads)
    done, t0 = 0, time.perf_counter()
    while done < max_steps:
        e.step_inner(1)
        done += 1
        print(f"bench.py: cpu baseline step {done}, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": n_dyn * done / dt, "unit": "body-steps/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (reference sweep order, grid-accelerated discovery on {threads} threads, "
                      f"single-threaded sweep and dual, FP32): steps {start_step + 1}-{start_step + done} of the "
                      f"same 100k-box workload, starting from the GPU trajectory's state at step {start_step} "
                      f"(contact history rebuilt), {dt:.1f} s"}


def cpu_batched_sample(parts, n_scenes, steps, threads):
    """The C4 scenes on the host: independent scenes, one CPU oracle each, run
    `threads` at a time (ctypes releases the interpreter lock during a step),
    on a sample of n_scenes scenes x steps steps; scene-steps/s."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import pyoracle as po
    engines = [po.OracleEngine(parts[i], order=po.ORDER_REFERENCE, math=po.MATH_LIBM) for i in range(n_scenes)]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda e: e.step_inner(steps), engines))
    dt = time.perf_counter() - t0
    bodies = sum(int((parts[i]["body_type"] == 1).sum()) for i in range(n_scenes))
    return {"value": n_scenes * steps / dt, "unit": "scene-steps/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle, scenes 0-{n_scenes - 1} ({bodies} dynamic bodies) x {steps} steps from rest, one "
                      f"oracle per scene on {threads} threads, {dt:.1f} s"}


def batched_leg(steps, loop_scenes, loop_steps):
    """4,096 independent box-drop scenes (8-512 bodies each, log-uniform,
    each with its own 10 x 10 m floor) stepped together in one context,
    against the same scenes stepped one context per scene on the same GPU
    (timed on a sample of scenes and steps, reported per scene-step)."""
    from ornis_amd import scenes
    parts, allsc = scenes.batched_drops()
    n_scenes = len(parts)
    eng = new_engine(iters=10)
    eng.load(allsc, scene_ids=allsc["scene_id"])
    eng.step_n(5)
    eng.sync()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(50, steps - done)
        eng.step_n(k)
        done += k
        eng.sync()
        print(f"bench.py: batched {done}/{steps} steps, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    eng.close()
    batched = n_scenes * steps / wall
    # one context per scene, the loop a caller without batching would run
    engines = [new_engine(scene=parts[i], iters=10) for i in range(loop_scenes)]
    for e in engines:
        e.step_n(5)
        e.sync()
    t0 = time.perf_counter()
    for _ in range(loop_steps):
        for e in engines:
            e.step_n(1)
    for e in engines:
        e.sync()
    lwall = time.perf_counter() - t0
    for e in engines:
        e.close()
    loop = loop_scenes * loop_steps / lwall
    return {"metric": "batched scenes/s", "value": batched, "unit": "scene-steps/s", "scenes": n_scenes,
            "bodies": int(len(allsc["shape"])), "steps": steps, "ms_per_step": 1e3 * wall / steps,
            "per_scene_loop": {"value": loop, "unit": "scene-steps/s",
                               "sample": f"scenes 0-{loop_scenes - 1}, {loop_steps} steps each after 5 warmup, "
                                         f"one context per scene, {lwall:.2f} s"},
            "speedup_vs_loop": batched / loop}


def differentiable_leg(steps, reps, cpu_scenes, threads):
    """1,024 throw scenes x 64 boxes (65,536 bodies): `steps` steps forward
    with a checkpoint every 10 steps, then the adjoint back to the initial
    velocities, masses and frictions (each checkpoint interval recomputed on
    the device, then differentiated).  One untimed pass, then `reps` timed
    passes; scene-steps/s of the forward alone and of forward + backward.
    The CPU line is the oracle's forward only (the reference has no adjoint),
    scenes across `threads` threads."""
    import torch
    from ornis_amd import diff, scenes
    parts, allsc = scenes.diff_throws(n_scenes=1024)
    n_scenes = len(parts)
    targets = torch.tensor(np.stack([p["target"] for p in parts]), dtype=torch.float32)
    dyn = allsc["body_type"] == 1
    eng = new_engine(iters=10)
    fwd, bwd = [], []
    for r in range(reps + 1):
        v0 = torch.tensor(allsc["vel"], requires_grad=True)
        mass = torch.tensor(allsc["mass"], requires_grad=True)
        fric = torch.tensor(allsc["friction"], requires_grad=True)
        t0 = time.perf_counter()
        final = diff.rollout(eng, allsc, v0, mass, fric, steps, every=10)
        t1 = time.perf_counter()
        diff.centroid_loss(final, allsc["scene_id"], dyn, targets).backward()
        t2 = time.perf_counter()
        if r:
            fwd.append(t1 - t0)
            bwd.append(t2 - t1)
        print(f"bench.py: differentiable pass {r}/{reps}: forward {t1 - t0:.3f} s, backward {t2 - t1:.3f} s",
              file=sys.stderr, flush=True)
    eng.close()
    f, b = float(np.mean(fwd)), float(np.mean(bwd))
    out = {"metric": "differentiable scene-steps/s (forward + backward)", "value": n_scenes * steps / (f + b),
           "unit": "scene-steps/s", "scenes": n_scenes, "bodies": int(len(allsc["shape"])), "steps": steps,
           "checkpoint_every": 10, "forward_s": f, "backward_s": b,
           "forward_only": {"value": n_scenes * steps / f, "unit": "scene-steps/s"},
           "timing": f"host clock around each call (inputs cross to the device and the gradients back), mean of "
                     f"{reps} passes after one untimed"}
    if cpu_scenes:
        from concurrent.futures import ThreadPoolExecutor
        from oracle import pyoracle as po
        engines = [po.OracleEngine({k: v for k, v in parts[i].items() if k != "target"}, order=po.ORDER_REFERENCE,
                                   math=po.MATH_LIBM) for i in range(cpu_scenes)]
        t0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda e: e.step_inner(steps), engines))
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cpu_scenes * steps / dt, "unit": "scene-steps/s (forward only)",
                               "cores": threads, "kind": "port",
                               "sample": f"CPU oracle forward, scenes 0-{cpu_scenes - 1} x {steps} steps, one oracle "
                                         f"per scene on {threads} threads, {dt:.1f} s"}
    return out


def measured_traffic(window, kernel="k_primal8"):
    """HBM bytes per launch of `kernel` (k_primal8 or k_dual) counted by
    rocprofv3 (FETCH_SIZE + WRITE_SIZE, each scaled by the counter's reading
    for a plain copy of known size) over the same steps of this workload,
    from the committed summary profiles/k_primal_traffic_100k.json; None
    when no summary covers `window` (first-last step)."""
    path = os.path.join(ROOT, "profiles", "k_primal_traffic_100k.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        j = json.load(f)
    w = j.get("windows", {}).get(window)
    return None if w is None else w.get(kernel, {}).get("hbm_bytes_per_launch")


def trajectory_leg(scene, n_dyn, steps=300, windows=100, late=50, sample_every=5, capture_step=280):
    """The workload's whole trajectory in a fresh context, end to end:
    ms/step per window of `windows` steps, and over the last `late` steps
    the per-phase times and the primal / dual launch rates (HIP events on
    one step in `sample_every`).  The state after `capture_step` is
    returned for the CPU baseline (read while the clock is stopped)."""
    eng = new_engine(scene=scene, iters=10)
    eng.sync()
    first_late = steps - late + 1
    sampled = {i for i in range(first_late, steps + 1) if (i - first_late) % sample_every == 0}
    phases = {"setup": 0.0, "primal": 0.0, "dual": 0.0, "finish": 0.0, "step": 0.0}
    per_window, wall, captured = [], 0.0, None
    eng.kernel_times(reset=True)
    i = 1  # next step to run (1-based)
    for w0 in range(1, steps + 1, windows):
        w1 = min(w0 + windows - 1, steps)
        t0 = time.perf_counter()
        while i <= w1:
            if i in sampled:
                eng.set_timing(True)
                eng.step_n(1)
                eng.set_timing(False)
                for k, v in eng.timing().items():
                    phases[k] += v
                i += 1
                continue
            j = i  # run up to the next sampled or captured step, inside the window
            while j + 1 <= w1 and (j + 1) not in sampled and j != capture_step:
                j += 1
            eng.step_n(j - i + 1)
            i = j + 1
            if j == capture_step:
                eng.sync()
                t_pause = time.perf_counter()
                captured = eng.state()
                t0 += time.perf_counter() - t_pause
        eng.sync()
        dt = time.perf_counter() - t0
        wall += dt
        per_window.append({"steps": f"{w0}-{w1}", "ms_per_step": 1e3 * dt / (w1 - w0 + 1)})
        print(f"bench.py: trajectory {w1}/{steps} steps, {wall:.1f} s", file=sys.stderr, flush=True)
    kt = eng.kernel_times()
    stats_end = eng.stats()
    eng.close()
    nph = max(len(sampled), 1)
    primal_gbs = kt["primal_bytes"] / (kt["primal_ms"] * 1e-3) / 1e9 if kt["primal_ms"] > 0 else 0.0
    dual_gbs = kt["dual_bytes"] / (kt["dual_ms"] * 1e-3) / 1e9 if kt["dual_ms"] > 0 else 0.0
    late_window = f"{first_late}-{steps}"
    return {
        "steps": steps, "ms_per_step": 1e3 * wall / steps, "body_steps_per_s": n_dyn * steps / wall,
        "windows": per_window,
        "late_window": f"steps {late_window} (HIP events on every {sample_every}th step)",
        "late_phases_ms_per_step": {k: v / nph for k, v in phases.items()},
        "late_primal": {"achieved_GBs": primal_gbs, "frac": primal_gbs / HBM_PEAK_GBS,
                        "bytes_per_launch": kt["primal_bytes"] / max(kt["primal_launches"], 1),
                        "avg_launch_us": 1e3 * kt["primal_ms"] / max(kt["primal_launches"], 1),
                        "launches_per_step": kt["primal_launches"] / max(kt["steps"], 1),
                        "traffic": measured_traffic(late_window), "traffic_window": late_window},
        "late_dual": {"achieved_GBs": dual_gbs, "frac": dual_gbs / HBM_PEAK_GBS,
                      "bytes_per_launch": kt["dual_bytes"] / max(kt["dual_launches"], 1),
                      "avg_launch_us": 1e3 * kt["dual_ms"] / max(kt["dual_launches"], 1)},
        "final_stats": stats_end,
    }, captured


def chains_leg(steps):
    """Contacts + joints: 2,000 chains of 32 boxes (0.1 x 0.1 x 0.4 m) linked
    by ball joints, roots on a 45 x 45 grid at 20 m, falling onto a plane
    scattered with 5,000 static spheres; 12 iterations, `steps` steps."""
    from ornis_amd import scenes
    s = scenes.chains()
    n_dyn = int((s["body_type"] == 1).sum())
    eng = new_engine(scene=s, iters=12)
    eng.sync()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(60, steps - done)
        eng.step_n(k)
        done += k
        eng.sync()
        print(f"bench.py: chains {done}/{steps} steps, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    st = eng.stats()
    eng.close()
    return {"metric": "body-steps/s", "value": n_dyn * steps / wall, "unit": "body-steps/s", "bodies": n_dyn,
            "joints": len(s["joints"]), "steps": steps, "ms_per_step": 1e3 * wall / steps, "iters": 12,
            "final_stats": st}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=290)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-max-steps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--progress", action="store_true")
    ap.add_argument("--lanes", type=int, default=0, help="primal lanes per body (0: library default)")
    ap.add_argument("--waves", type=int, default=2, help="primal wavefronts per workgroup")
    ap.add_argument("--sweep-shapes", type=int, default=0,
                    help="after warmup, time this many steps per primal launch shape and exit")
    ap.add_argument("--timing-every", type=int, default=10,
                    help="record per-launch kernel events on one timed step in this many")
    ap.add_argument("--trajectory-steps", type=int, default=300,
                    help="steps of the end-to-end trajectory leg (0: skip it)")
    ap.add_argument("--chains-steps", type=int, default=240, help="steps of the chains workload (0: skip it)")
    ap.add_argument("--batched-steps", type=int, default=300,
                    help="timed steps of the 4,096-scene batched workload (0: skip it)")
    ap.add_argument("--loop-scenes", type=int, default=32)
    ap.add_argument("--loop-steps", type=int, default=30)
    ap.add_argument("--cpu-batched-scenes", type=int, default=64)
    ap.add_argument("--cpu-batched-steps", type=int, default=30)
    ap.add_argument("--diff-steps", type=int, default=120,
                    help="steps of the 1,024-scene differentiable workload (0: skip it)")
    ap.add_argument("--diff-reps", type=int, default=3)
    ap.add_argument("--cpu-diff-scenes", type=int, default=64)
    ap.add_argument("--autotune-steps", type=int, default=0,
                    help="after the warmup, autotune the primal launch variants over this many steps "
                         "(state restored; 0: library default table)")
    ap.add_argument("--copy-check", type=int, default=0, metavar="MIB",
                    help="before warmup, copy a MIB-sized device buffer 3 times (the plain-copy reference "
                         "for HBM byte counters)")
    ap.add_argument("--only-main", action="store_true", help="run the timed leg only (profiling runs)")
    ap.add_argument("--tail", type=int, default=-1,
                    help="bodies of the last colour classes fused into one primal launch (0: none; -1: default)")
    ap.add_argument("--tail-waves", type=int, default=0, help="wavefront limit of the fused tail launch")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the 100k-box scene fits one MI355X and multi-GPU "
              "sharding is a design note only (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    from ornis_amd import scenes

    global TAIL
    TAIL = (args.tail, args.tail_waves)
    if args.copy_check:
        import torch
        x = torch.ones(args.copy_check * (1 << 18), dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        del x, y

    scene = scenes.box_drop_lattice()
    n_dyn = int((scene["body_type"] == 1).sum())
    eng = new_engine(scene=scene, iters=10)
    for s in range(args.warmup):
        eng.step_n(1)
        if args.progress:
            print(f"warmup {s} {eng.stats()}", file=sys.stderr, flush=True)
    eng.sync()
    table = None
    eng_tune = None
    if args.autotune_steps:
        table = eng.autotune(args.autotune_steps)
        eng_tune = eng.tune_us
    eng.set_timing(True)
    if args.lanes:
        eng.set_primal_shape(args.lanes, args.waves)
    if args.sweep_shapes:
        for shape in ((16, 1), (16, 2), (16, 4), (16, 8)):
            eng.set_primal_shape(*shape)
            eng.kernel_times(reset=True)
            eng.step_n(args.sweep_shapes)
            kt = eng.kernel_times()
            print(json.dumps({"lanes": shape[0], "waves": shape[1],
                              "primal_ms_per_step": kt["primal_ms"] / kt["steps"],
                              "avg_launch_us": 1e3 * kt["primal_ms"] / kt["primal_launches"],
                              "GBps": kt["primal_bytes"] / (kt["primal_ms"] * 1e-3) / 1e9}), flush=True)
        return
    eng.kernel_times(reset=True)
    # Kernel timing: HIP events bracket every primal and dual launch of one
    # step in every `timing_every` (the events themselves add launch gaps,
    # so they are kept off the other steps).
    every = max(args.timing_every, 1)
    eng.set_timing(False)
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        if done % every == 0:
            eng.set_timing(True)
            eng.step_n(1)
            eng.set_timing(False)
            done += 1
        else:
            k = min(every - done % every, args.steps - done)
            eng.step_n(k)
            done += k
        if args.progress:
            print(f"step {args.warmup + done} {eng.stats()} {eng.timing()}", file=sys.stderr, flush=True)
    eng.sync()
    t1 = time.perf_counter()
    kt = eng.kernel_times()
    st = eng.stats()
    wall = t1 - t0
    ms_per_step = 1e3 * wall / args.steps
    value = n_dyn * args.steps / wall
    achieved = kt["primal_bytes"] / (kt["primal_ms"] * 1e-3) / 1e9 if kt["primal_ms"] > 0 else 0.0
    dual_gbs = kt["dual_bytes"] / (kt["dual_ms"] * 1e-3) / 1e9 if kt["dual_ms"] > 0 else 0.0
    eng.close()
    window = f"{args.warmup + 1}-{args.warmup + args.steps}"
    traffic = measured_traffic(window)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": "100k random boxes, 50x50x40 lattice drop into a 110 m container, "
                               "dt 1/60, 10 iterations, FP32", "bodies": n_dyn, "iters": 10,
                   "parallelism": "single GPU", "timed_steps": window},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "k_primal8",
                     "window": f"steps {window} (HIP events on every {every}th step)",
                     "bytes_per_launch": kt["primal_bytes"] / max(kt["primal_launches"], 1),
                     "avg_launch_us": 1e3 * kt["primal_ms"] / max(kt["primal_launches"], 1)},
        "cpu_baseline": None,
        "phases_ms_per_step": {"primal": kt["primal_ms"] / max(kt["steps"], 1),
                               "dual": kt["dual_ms"] / max(kt["steps"], 1)},
        "dual_roofline": {"achieved": dual_gbs, "frac": dual_gbs / HBM_PEAK_GBS, "kernel": "k_dual",
                          "bytes_per_launch": kt["dual_bytes"] / max(kt["dual_launches"], 1),
                          "avg_launch_us": 1e3 * kt["dual_ms"] / max(kt["dual_launches"], 1),
                          "traffic": measured_traffic(window, "k_dual")},
        "final_stats": st,
        "primal_table": None if table is None else [int(x) for x in table],
        "primal_tune_us": None if table is None else
        {f"2^{b}": [round(float(x), 1) for x in eng_tune[:, b]] for b in range(24) if (eng_tune[:, b] > 0).any()},
    }
    if not args.only_main:
        captured = None
        if args.trajectory_steps:
            traj, captured = trajectory_leg(scene, n_dyn, steps=args.trajectory_steps)
            line["trajectory"] = traj
        if args.chains_steps:
            line["chains"] = chains_leg(args.chains_steps)
        parts = None
        if args.batched_steps:
            line["batched"] = batched_leg(args.batched_steps, args.loop_scenes, args.loop_steps)
        if args.diff_steps:
            line["differentiable"] = differentiable_leg(args.diff_steps, args.diff_reps,
                                                        0 if args.no_cpu else args.cpu_diff_scenes, host_cores())
        if not args.no_cpu:
            cores = host_cores()
            if captured is not None:
                line["cpu_baseline"] = cpu_baseline(scene, captured, 280, args.cpu_seconds, args.cpu_max_steps, cores)
            if args.batched_steps and args.cpu_batched_scenes:
                parts, _ = scenes.batched_drops()
                line["batched"]["cpu_baseline"] = cpu_batched_sample(parts, args.cpu_batched_scenes,
                                                                     args.cpu_batched_steps, cores)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
