"""Benchmark of the GPU rigid-body step (plain per-scene path).

Workload (one GPU): 131,072 spheres (r ~ U[0.1,0.3] m) + 131,072 boxes (half
extents ~ U[0.1,0.3] m) poured from a seeded jittered grid into a 40 x 40 m
walled bin (floor + 4 wall planes), TGS with 8 position / 1 velocity
iterations, dt = 1/60 s, float32. A fixed, untimed pour of --settle steps
(default 300, 5 s of simulated time) lets the pile form; then W warm-up steps
and K timed steps (wall clock around simulate + fetchResults, bracketed by
device synchronisation) run on the settled pile.

Prints one JSON line with the body-steps/s, the roofline of the dominant
kernel (the colour-partition solve, HBM-bound; its launches carry HIP events
recorded by the dispatches themselves on the scene's stream during the timed
steps) and the reference CPU pipeline timed on a bounded sample of the same
state on the host cores at 16, 8 and 1 threads.

The same line carries the batched form ("batched"): 8,192 independent worlds
of log-uniform [8, 512] bodies (box stacks and sphere/box piles, about 1.0M
bodies), TGS 4/1, dt = 1/120 s, stepped as one scene (one launch sequence per
step for all worlds), with its body-steps/s and p50/p99 step latency, beside
the plain per-scene path (each world its own scene, stepped one after the
other with the same kernels) measured on an evenly spaced sample of the worlds.

"served" is the request stream: Poisson arrivals at several rates of step
requests for scenes of log-uniform [16, 4096] bodies, each for k ~ U{1..32}
steps at dt = 1/60, admitted against a 200 GB HBM budget into one batched
step per serving iteration, with body-steps/s, serving-step and request
latency percentiles.

"convex" is the convex-heavy scene: 65,536 hulls of 24 points on seeded
ellipsoids dropped in layers onto a plane, TGS 4/1, all 400 steps timed (the
settled last 40 also reported alone), beside the reference CPU SDK stepping
the final state.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_solve_traffic_sphere_box_bin_262k.md")


def solve_traffic(settle, warmup, steps, constraints_avg):
    """HBM bytes per k_solve launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated against a device copy), used only when that
    summary was taken on this very state: the same settle / warm-up / timed
    step counts and the same mean constraint count over the timed steps
    (within 0.5%). Returns (bytes or None, note)."""
    import re
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None, "no counter summary committed"
    got = {k: re.search(rf"^{k}:\s*([0-9.]+)", text, re.M) for k in
           ("traffic_bytes_per_launch", "settle", "warmup", "steps", "constraints_avg")}
    if not all(got.values()):
        return None, "counter summary lacks the state it was taken on"
    v = {k: float(m.group(1)) for k, m in got.items()}
    if (int(v["settle"]), int(v["warmup"]), int(v["steps"])) != (settle, warmup, steps):
        return None, (f"counter summary taken at settle {int(v['settle'])} warmup {int(v['warmup'])} steps "
                      f"{int(v['steps'])}, not this run's state")
    if abs(v["constraints_avg"] - constraints_avg) > 0.005 * max(constraints_avg, 1.0):
        return None, (f"counter summary state has {v['constraints_avg']:.0f} constraints per step, this run "
                      f"{constraints_avg:.0f}")
    return int(v["traffic_bytes_per_launch"]), os.path.relpath(TRAFFIC_PROFILE, ROOT)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def solve_bytes_per_step(st: dict, iters: int) -> float:
    """Minimal algorithmic HBM bytes of all solve launches of one step, the
    roofline's byte count: per iteration 152 B per touching pair (a 64 B row
    header and two friction rows of 44 B), 60 B per contact point (56 B of row
    read, the 4 B impulse written) and 128 B per body (velocity and motion
    state read and written once)."""
    return iters * (152.0 * st["constraints"] + 60.0 * st["contacts"] + 128.0 * st["bodies"])


def solve_bytes_per_step_endpoints(st: dict, iters: int) -> float:
    """The same launches counted as this solver's row layout moves them. Per
    constraint and iteration: a 32 B header; per contact point 32 B of row
    read and the 4 B impulse read and written; per friction row 32 B of row
    (+16 B of body-B row when B is dynamic) and the 4 B impulse read and
    written; per dynamic body endpoint 64 B of velocity/motion state read and
    32 B written (a body is gathered once for every constraint it is in)."""
    T = st["constraints"]
    K = st["contacts"]
    F = st["friction_rows"]
    Fd = st["body_body_friction_rows"]
    endpoints = T + st["body_body_constraints"]
    return iters * (32.0 * T + 40.0 * K + 40.0 * F + 16.0 * Fd + 96.0 * endpoints)


def _window(rows, window):
    """The dispatches of steps SKIP+1 .. SKIP+STEPS (one k_finalize or
    k_pgs_finalize dispatch ends every step), rows sorted by start time."""
    fin = [i for i, r in enumerate(rows) if "finalize" in r["Kernel_Name"]]
    if window is None:
        return rows, len(fin)
    skip, steps = window
    lo = fin[skip - 1] + 1 if skip > 0 else 0
    hi = fin[skip + steps - 1] + 1
    return rows[lo:hi], steps


def trace_summary(path: str, window) -> None:
    """Per-step kernel table from a rocprofv3 kernel-trace CSV
    (`--kernel-trace --output-format csv`) over a window of steps."""
    import collections
    import csv
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Start_Timestamp"]))
    sel, steps = _window(rows, window)
    steps = max(1, steps)
    agg = collections.defaultdict(lambda: [0, 0])
    for r in sel:
        name = r["Kernel_Name"].split("(")[0].replace("void ", "")
        agg[name][0] += 1
        agg[name][1] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    busy = sum(v[1] for v in agg.values())
    span = int(sel[-1]["End_Timestamp"]) - int(sel[0]["Start_Timestamp"])
    print(f"steps {steps}; dispatches {len(sel)}; per step: kernel busy {busy / 1e6 / steps:.3f} ms, "
          f"first start to last end {span / 1e6 / steps:.3f} ms\n")
    print("| kernel | calls/step | ms/step | avg us | % busy |")
    print("|---|---:|---:|---:|---:|")
    for k, (c, d) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"| {k} | {c / steps:.1f} | {d / 1e6 / steps:.3f} | {d / c / 1e3:.2f} | {100 * d / busy:.1f} |")


def pmc_summary(path: str, kernel: str, window) -> None:
    """Mean counter value per dispatch of `kernel` over a window of steps in a
    rocprofv3 counter-collection CSV (`--pmc ... --output-format csv`), with
    the mean dispatch duration."""
    import csv
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Start_Timestamp"]))
    # one row per counter and dispatch: the step window is found on one counter's rows
    names = sorted({r["Counter_Name"] for r in rows})
    by = {}
    for cn in names:
        sub, _ = _window([r for r in rows if r["Counter_Name"] == cn], window)
        sub = [r for r in sub if kernel in r["Kernel_Name"] and "pgs" not in r["Kernel_Name"]
               and "worlds" not in r["Kernel_Name"]]
        by[cn] = ([float(r["Counter_Value"]) for r in sub],
                  [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in sub])
    for k, (v, d) in by.items():
        print(f"{kernel}: {k} mean {np.mean(v):.1f} KiB per dispatch over {len(v)} dispatches "
              f"(mean duration under counters {np.mean(d) / 1e3:.2f} us)")


def calibrate_copy(mb: int) -> None:
    """Device-to-device copies of a known size (for scaling rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters to bytes): 5 copies of `mb` MiB."""
    import torch
    n = mb * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copied {mb} MiB 5 times")


def run_batched(args, P, scenes):
    """Batched step of many small worlds (every step of the run timed, from
    the initial drop) against the plain per-scene path."""
    import torch
    dt = 1.0 / 120.0
    sc = scenes.batched_piles(n_scenes=args.scenes)
    log(f"batched: {sc.n_scenes} worlds, {sc.n} bodies")
    g = P.Scene(sc, position_iterations=4, velocity_iterations=1)
    # the solve's launch shape is autotuned once the piles have formed, at
    # step --autotune-at of the timed run (every candidate for
    # --autotune-steps real steps; no bit of the result depends on it)
    for s in range(args.batched_warmup):
        g.step(dt)
    torch.cuda.synchronize()
    times = []
    pairs = 0
    t0 = time.perf_counter()
    for s in range(args.batched_steps):
        if s == args.autotune_at and args.autotune_steps > 0:
            g.set_tuning(autotune_steps=args.autotune_steps)
        a = time.perf_counter()
        g.simulate(dt)
        g.fetch_results(True)
        times.append((time.perf_counter() - a) * 1e3)
        pairs += g.stats()["constraints"]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    tuning, tuning_log = g.tuning(), g.tuning_log()
    g.close()
    batched = sc.n * args.batched_steps / elapsed
    # plain path: every k-th world as its own scene, same warm-up, stepped in turn
    stride = max(1, sc.n_scenes // args.plain_sample)
    worlds = list(range(0, sc.n_scenes, stride))[: args.plain_sample]
    subs = [sc.subset(w) for w in worlds]
    plain = [P.Scene(x, position_iterations=4, velocity_iterations=1) for x in subs]
    for s in range(args.plain_warmup):
        for x in plain:
            x.step(dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.plain_steps):
        for x in plain:
            x.simulate(dt)
            x.fetch_results(True)
    torch.cuda.synchronize()
    pe = time.perf_counter() - t0
    nb = sum(x.n for x in subs)
    for x in plain:
        x.close()
    plain_rate = nb * args.plain_steps / pe
    ts = np.array(times)
    out = {"workload": f"batched_piles: {sc.n_scenes} worlds, {sc.n} bodies, log-uniform [8,512] per world, "
                       f"TGS 4/1, dt 1/120", "worlds": sc.n_scenes, "bodies": sc.n, "steps": args.batched_steps,
           "warmup": args.batched_warmup, "body_steps_per_s": round(batched, 1),
           "ms_per_step": round(1e3 * elapsed / args.batched_steps, 3),
           "p50_step_ms": round(float(np.percentile(ts, 50)), 3), "p99_step_ms": round(float(np.percentile(ts, 99)), 3),
           "contact_pairs_per_s": round(pairs / elapsed, 1),
           "plain_body_steps_per_s": round(plain_rate, 1),
           "plain_sample": f"{len(worlds)} worlds (every {stride}th, {nb} bodies), {args.plain_steps} steps each "
                           f"after {args.plain_warmup} warm-up steps, one scene per world stepped in turn",
           "speedup_vs_plain": round(batched / plain_rate, 2),
           "slowest_steps_ms": [[int(i) + 1, round(float(ts[i]), 3)] for i in np.argsort(ts)[::-1][:12]],
           "solve_shape": {"layout": ["auto", "colour launches", "per-world persistent"][tuning[0]],
                           "workgroups": tuning[1], "lds_bodies": tuning[2], "tuned_at_step": args.autotune_at,
                           "autotune_log_ms_per_step": [[a, b, c, round(m, 4)] for a, b, c, m in tuning_log]}}
    log(f"batched: {out}")
    return out


def run_convex(args, P, scenes):
    """Convex-heavy scene: 65,536 hulls of 24 points on seeded ellipsoids
    (semi-axes U[0.2,0.6] m) dropped in layers onto a plane, TGS 4/1,
    dt = 1/60, every step of the run timed, the last --convex-window steps
    (the settled pile) reported separately, with the reference CPU SDK
    stepping the final state on the host cores."""
    import torch
    dt = 1.0 / 60.0
    t0 = time.perf_counter()
    sc = scenes.convex_layers(n_bodies=args.convex_bodies)
    log(f"convex: {sc.n} hulls generated in {time.perf_counter() - t0:.1f}s")
    g = P.Scene(sc, position_iterations=4, velocity_iterations=1)
    torch.cuda.synchronize()
    times, pairs = [], []
    t0 = time.perf_counter()
    for s in range(args.convex_steps):
        a = time.perf_counter()
        g.simulate(dt)
        g.fetch_results(True)
        times.append((time.perf_counter() - a) * 1e3)
        pairs.append(g.stats()["constraints"])
        if (s + 1) % 100 == 0:
            log(f"convex step {s + 1}/{args.convex_steps} {g.stats()}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st = g.stats()
    # stage breakdown from a profiled pass of a window's length
    w = min(args.convex_window, args.convex_steps)
    g.set_profiling(True)
    for s in range(w):
        g.simulate(dt)
        g.fetch_results(True)
    torch.cuda.synchronize()
    prof = g.profile()
    g.set_profiling(False)
    ts = np.array(times)
    tw = ts[-w:]
    out = {"workload": f"convex_layers: {sc.n} hulls of 24 points on ellipsoids U[0.2,0.6] m, layers onto a plane, "
                       f"TGS 4/1, dt 1/60, all {args.convex_steps} steps timed", "bodies": sc.n,
           "steps": args.convex_steps, "body_steps_per_s": round(sc.n * args.convex_steps / elapsed, 1),
           "ms_per_step": round(1e3 * elapsed / args.convex_steps, 3),
           "p50_step_ms": round(float(np.percentile(ts, 50)), 3), "p99_step_ms": round(float(np.percentile(ts, 99)), 3),
           "contact_pairs_per_s": round(sum(pairs) / elapsed, 1), "last_step_counts": st,
           "settled_window": {"steps": f"{args.convex_steps - w + 1}-{args.convex_steps}",
                              "ms_per_step": round(float(tw.mean()), 3),
                              "body_steps_per_s": round(sc.n * 1e3 / float(tw.mean()), 1),
                              "p99_step_ms": round(float(np.percentile(tw, 99)), 3),
                              "contact_pairs_per_step": round(float(np.mean(pairs[-w:])), 1)},
           "stage_ms_per_step": {k: round(v[0] / w, 3) for k, v in prof.items()}}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_libs as L
        if L.have_ref():
            pose, lv, av = g.bodies()
            stt = scenes.Scene(sc.shape_type, sc.dims, pose, lv, av, sc.mass, sc.inertia, sc.scene_id, sc.planes,
                               sc.plane_scene, 1, sc.name, {}, sc.hull_points, sc.hull_counts)
            threads = args.cpu_threads[0]
            ref = L.RefScene(stt, "tgs", 4, 1, threads=threads)
            ref.step(dt, 1)
            n = 2 * args.cpu_steps[0]
            total, _ = ref.step(dt, n)
            cpu = sc.n * n / total
            out["cpu_baseline"] = {"value": round(cpu, 1), "unit": "body-steps/s", "cores": threads,
                                   "kind": "reference", "ms_per_step": round(1e3 * total / n, 2),
                                   "sample": f"{n} steps of the pile state after step "
                                             f"{args.convex_steps + w} (+1 untimed), reference PhysX 5.6.1 CPU SDK "
                                             f"cooking the same point clouds"}
            out["speedup_vs_cpu_settled"] = round(out["settled_window"]["body_steps_per_s"] / cpu, 2)
    g.close()
    log(f"convex: {out}")
    return out


def run_served(args, P, scenes):
    """Served request stream: Poisson arrivals (rate lam per second, wall
    clock) of step requests for scenes of log-uniform [16, 4096] bodies, each
    asking for k ~ U{1..32} steps at dt = 1/60, admitted against an HBM budget
    and merged into one batched step per serving iteration."""
    import torch
    dt = 1.0 / 60.0
    pool_sc = scenes.batched_piles(n_scenes=args.serve_pool, nmin=16, nmax=4096, seed=11)
    pool = [pool_sc.subset(w) for w in range(pool_sc.n_scenes)]
    log(f"served: pool of {len(pool)} scenes, {pool_sc.n} bodies")
    results = []
    for lam in args.serve_rates:
        rng = np.random.default_rng(5)
        sv = P.Server(hbm_budget_bytes=args.serve_budget_gb << 30, max_worlds=65536, bytes_per_body=8192,
                      position_iterations=4, velocity_iterations=1)
        t0 = time.perf_counter()
        t_warm, t_end = t0 + args.serve_warmup_s, t0 + args.serve_warmup_s + args.serve_seconds
        next_arrival = t0 + rng.exponential(1.0 / lam)
        submitted = {}
        step_ms, lat_ms, wait_ms, bodies_active = [], [], [], []
        bs0 = None
        n_req = 0
        while True:
            now = time.perf_counter()
            if now >= t_end:
                break
            while next_arrival <= now:
                x = pool[int(rng.integers(len(pool)))]
                k = int(rng.integers(1, 33))
                rid = sv.submit(x, k)
                submitted[rid] = next_arrival
                n_req += 1
                next_arrival += rng.exponential(1.0 / lam)
            st = sv.stats()
            if st["active"] == 0 and st["queued"] == 0:
                time.sleep(max(0.0, min(next_arrival, t_end) - time.perf_counter()))
                continue
            if bs0 is None and now >= t_warm:
                bs0, tb0 = st["body_steps"], now
            a = time.perf_counter()
            sv.step(dt)
            b = time.perf_counter()
            if bs0 is not None:
                step_ms.append((b - a) * 1e3)
                bodies_active.append(st["active_bodies"])
            for rid in sv.finished():
                _, _, _, tm = sv.result(rid)
                t_arr = submitted.pop(rid)
                if t_arr >= t_warm:
                    lat_ms.append((tm[2] - t_arr) * 1e3)
                    wait_ms.append((tm[1] - t_arr) * 1e3)
        torch.cuda.synchronize()
        st = sv.stats()
        elapsed = time.perf_counter() - tb0 if bs0 is not None else 1.0
        rate = (st["body_steps"] - (bs0 or 0)) / elapsed
        sv.close()
        sm, lm, wm = np.array(step_ms or [0.0]), np.array(lat_ms or [0.0]), np.array(wait_ms or [0.0])
        r = {"lambda_per_s": lam, "requests": n_req, "timed_s": round(elapsed, 2), "body_steps_per_s": round(rate, 1),
             "serving_steps": len(step_ms), "mean_active_bodies": round(float(np.mean(bodies_active or [0])), 1),
             "p50_step_ms": round(float(np.percentile(sm, 50)), 3), "p99_step_ms": round(float(np.percentile(sm, 99)), 3),
             "p50_request_ms": round(float(np.percentile(lm, 50)), 3),
             "p99_request_ms": round(float(np.percentile(lm, 99)), 3),
             "p99_admission_wait_ms": round(float(np.percentile(wm, 99)), 3), "completed_timed": len(lat_ms)}
        log(f"served: {r}")
        results.append(r)
    return {"workload": f"Poisson step requests, scenes of log-uniform [16,4096] bodies (pool of {len(pool)} seeded "
                        f"box stacks / sphere-box piles), k ~ U{{1..32}} steps at dt 1/60, TGS 4/1, HBM budget "
                        f"{args.serve_budget_gb} GB at 8 KiB per body, one batched step per serving iteration",
            "warmup_s": args.serve_warmup_s, "rates": results}


def host_info() -> dict:
    """The host the CPU baseline ran on: logical CPUs, the CPUs this process
    may use, and the CPU model."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model}


def cpu_baseline(sc, pose, lv, av, dt, pos_iters, vel_iters, args):
    """The reference PhysX CPU SDK stepping the GPU's state at each thread
    count of --cpu-threads (one PxScene; the dispatcher is swapped between
    legs), a bounded number of steps per leg after one untimed step."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_libs as L
    from physx_amd import scenes
    if not L.have_ref():
        return None
    st = scenes.Scene(sc.shape_type, sc.dims, pose, lv, av, sc.mass, sc.inertia, sc.scene_id, sc.planes,
                      sc.plane_scene, 1, sc.name)
    threads = args.cpu_threads
    t0 = time.perf_counter()
    ref = L.RefScene(st, "tgs", pos_iters, vel_iters, threads=threads[0])
    ref.step(dt, 1)  # untimed: the first step creates every broadphase pair
    log(f"cpu baseline: reference scene built and first step in {time.perf_counter() - t0:.1f}s")
    legs = []
    for k, (t, n) in enumerate(zip(threads, args.cpu_steps)):
        if k > 0:
            ref.set_threads(t)
            ref.step(dt, 1)  # untimed: the new dispatcher's workers start
        total, ms = ref.step(dt, n)
        legs.append({"threads": t, "steps": n, "ms_per_step": round(1e3 * total / n, 2),
                     "body_steps_per_s": round(sc.n * n / total, 1)})
        log(f"cpu baseline: {legs[-1]}")
    first = legs[0]
    return {"value": first["body_steps_per_s"], "unit": "body-steps/s", "cores": first["threads"],
            "kind": "reference", "ms_per_step": first["ms_per_step"], "legs": legs, "host": host_info(),
            "sample": f"{first['steps']} steps of the settled {sc.n}-body pile state that the GPU reached at step "
                      f"{args.settle + args.warmup + 2 * args.steps} (+1 untimed step per leg), reference PhysX "
                      f"5.6.1 CPU SDK built from source (ABP, PCM, TGS {pos_iters}/{vel_iters}); value is the "
                      f"{first['threads']}-thread leg (the CPU share of one GPU on the host)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--settle", type=int, default=300,
                    help="untimed pour steps before --warmup: the timed steps run on the settled pile")
    ap.add_argument("--bodies", type=int, default=262144)
    ap.add_argument("--solve-event-stride", type=int, default=0,
                    help="timed steps: launch j of step k carries HIP events when (j + k) %% stride == 0 "
                         "(default: the number of timed steps; 1: every launch)")
    ap.add_argument("--cpu-threads", type=int, nargs="+", default=[16, 8, 1])
    ap.add_argument("--cpu-steps", type=int, nargs="+", default=[3, 2, 1])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--scenes", type=int, default=8192)
    ap.add_argument("--batched-steps", type=int, default=1000)
    ap.add_argument("--batched-warmup", type=int, default=0)
    ap.add_argument("--autotune-at", type=int, default=200,
                    help="batched run: the step at which the solve's launch shape is autotuned")
    ap.add_argument("--autotune-steps", type=int, default=2,
                    help="batched run: steps per candidate solve shape in the autotuning (0: none)")
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--plain-warmup", type=int, default=120)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-convex", action="store_true")
    ap.add_argument("--convex-bodies", type=int, default=65536)
    ap.add_argument("--convex-steps", type=int, default=400)
    ap.add_argument("--convex-window", type=int, default=40,
                    help="the last steps of the convex run reported separately (the settled pile)")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[250.0, 1000.0, 4000.0])
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--serve-warmup-s", type=float, default=1.0)
    ap.add_argument("--serve-pool", type=int, default=256)
    ap.add_argument("--serve-budget-gb", type=int, default=200)
    ap.add_argument("--trace-summary", metavar="CSV",
                    help="print the per-step kernel table of a rocprofv3 kernel trace and exit")
    ap.add_argument("--calibrate-copy", type=int, metavar="MIB",
                    help="run 5 device copies of MIB MiB (counter calibration) and exit")
    ap.add_argument("--pmc-summary", metavar="CSV",
                    help="print the mean counter value per k_solve dispatch of a rocprofv3 counter CSV and exit")
    ap.add_argument("--kernel", default="k_solve",
                    help="--pmc-summary: the kernel (name substring) whose dispatches are summarised")
    ap.add_argument("--window", type=int, nargs=2, metavar=("SKIP", "STEPS"),
                    help="--trace-summary / --pmc-summary: the STEPS steps after the first SKIP")
    args = ap.parse_args()
    if len(args.cpu_steps) != len(args.cpu_threads):
        ap.error("--cpu-steps needs one count per --cpu-threads entry")
    if args.pmc_summary:
        pmc_summary(args.pmc_summary, args.kernel, args.window)
        return
    if args.trace_summary:
        trace_summary(args.trace_summary, args.window)
        return
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the step of one scene runs on a single MI355X in this build "
              "(independent worlds would be placed on separate GPUs; one large scene would need slab halos). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import physx_amd as P
    from physx_amd import scenes

    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    dt = 1.0 / 60.0
    pos_iters, vel_iters = 8, 1
    sc = scenes.sphere_box_bin(n_bodies=args.bodies)
    log(f"scene: {sc.n} bodies, {sc.planes.shape[0]} planes")
    gpu = P.Scene(sc, position_iterations=pos_iters, velocity_iterations=vel_iters)

    # the pour: untimed for the headline, its step times kept for the full-run figure
    pour_ms = []
    t0 = time.perf_counter()
    for s in range(args.settle + args.warmup):
        a = time.perf_counter()
        gpu.simulate(dt)
        gpu.fetch_results(True)
        pour_ms.append((time.perf_counter() - a) * 1e3)
        if (s + 1) % 50 == 0:
            log(f"pour step {s + 1}/{args.settle + args.warmup} ({time.perf_counter() - t0:.1f}s) {gpu.stats()}")
    torch.cuda.synchronize()

    # timed steps: the library as a caller runs it; launch j of timed step k
    # carries its own start/end events (recorded by the dispatch) when
    # (j + k) % stride == 0, stride = the number of timed steps: every solve
    # launch position of a step is timed once over the window
    SOLVE = P.Scene.STAGES.index("solve")
    stride = args.solve_event_stride or args.steps
    gpu.set_profiling(1 << SOLVE)
    gpu.set_profiling_stride(stride)
    step_ms = []
    sums = dict(bodies=0, constraints=0, contacts=0, friction_rows=0, body_body_constraints=0, bp_pairs=0,
                plane_pairs=0, colors=0, color_rounds=0, body_body_friction_rows=0, solve_launches=0)
    solve_bytes = solve_bytes_ep = 0.0
    torch.cuda.synchronize()
    t_begin = time.perf_counter()
    for s in range(args.steps):
        a = time.perf_counter()
        gpu.simulate(dt)
        gpu.fetch_results(True)
        step_ms.append((time.perf_counter() - a) * 1e3)
        st = gpu.stats()
        for k in sums:
            sums[k] += st[k]
        solve_bytes += solve_bytes_per_step(st, pos_iters + vel_iters)
        solve_bytes_ep += solve_bytes_per_step_endpoints(st, pos_iters + vel_iters)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_begin
    solve_ms, solve_launches = gpu.profile().get("solve", (0.0, 0))
    log(f"timed {args.steps} steps in {elapsed:.3f}s; timed solve launches {solve_launches}, {solve_ms:.3f} ms")

    # the same number of steps again with every stage timed (events around
    # each launch add marker packets, so this pass is not the timed one)
    gpu.set_profiling(True)
    gpu.set_profiling_stride(1)
    for s in range(args.steps):
        gpu.simulate(dt)
        gpu.fetch_results(True)
    torch.cuda.synchronize()
    prof = gpu.profile()
    gpu.set_profiling(False)
    log(f"profiled {args.steps} steps; stage GPU ms (total, launches): {prof}")

    value = sc.n * args.steps / elapsed
    avg = {k: v / args.steps for k, v in sums.items()}
    # average duration of the sampled launches; bytes per launch over every
    # solve launch of the window
    all_launches = max(sums["solve_launches"], 1)
    avg_launch_s = solve_ms * 1e-3 / max(solve_launches, 1)
    bytes_launch = solve_bytes / all_launches
    achieved = bytes_launch / avg_launch_s / 1e9 if solve_ms > 0 else 0.0
    achieved_ep = solve_bytes_ep / all_launches / avg_launch_s / 1e9 if solve_ms > 0 else 0.0
    traffic, traffic_src = solve_traffic(args.settle, args.warmup, args.steps, avg["constraints"])
    roofline = {"bound": "hbm", "kernel": "k_solve", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "traffic_source": traffic_src,
                "bytes_per_launch": round(bytes_launch, 1),
                "byte_model": "per iteration 152 B per touching pair + 60 B per contact point + 128 B per body",
                "avg_launch_us": round(1e6 * avg_launch_s, 3), "launches": sums["solve_launches"],
                "timed_launches": solve_launches, "event_stride": stride,
                "solve_ms_per_step": round(avg_launch_s * 1e3 * sums["solve_launches"] / args.steps, 4),
                "achieved_row_layout_model": round(achieved_ep, 2),
                "bytes_per_launch_row_layout_model": round(solve_bytes_ep / all_launches, 1)}

    cpu = None
    if not args.no_cpu:
        pose, lv, av = gpu.bodies()
        cpu = cpu_baseline(sc, pose, lv, av, dt, pos_iters, vel_iters, args)
    gpu.close()
    batched = None if args.no_batched else run_batched(args, P, scenes)
    served = None if args.no_served else run_served(args, P, scenes)
    convex = None if args.no_convex else run_convex(args, P, scenes)

    sms = np.array(step_ms)
    pour = np.array(pour_ms)
    first = args.settle + args.warmup + 1
    out = {
        "metric": "body-steps/sec at 262k bodies",
        "value": round(value, 1),
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"sphere_box_bin settled pile: 131072 spheres + 131072 boxes poured into a 40x40 m "
                               f"walled bin, TGS 8/1, dt 1/60; timed steps {first}-{first + args.steps - 1} after "
                               f"an untimed {args.settle}-step pour and {args.warmup} warm-up steps",
                   "bodies": sc.n, "solver": "TGS", "position_iterations": pos_iters,
                   "velocity_iterations": vel_iters, "dt": dt, "settle_steps": args.settle,
                   "parallelism": "single GPU, single scene"},
        "p50_step_ms": round(float(np.percentile(sms, 50)), 3),
        "p99_step_ms": round(float(np.percentile(sms, 99)), 3),
        "contact_pairs_per_s": round(sums["constraints"] / elapsed, 1),
        "avg_counts": {k: round(v, 1) for k, v in avg.items()},
        "stage_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in prof.items()},
        "pour": {"steps": len(pour_ms), "ms_per_step": round(float(pour.mean()), 4) if len(pour) else None,
                 "body_steps_per_s": round(sc.n * 1e3 / float(pour.mean()), 1) if len(pour) else None,
                 "note": "steps 1..settle+warmup (free fall into the forming pile), not the headline"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(value / cpu["value"], 2) if cpu else None,
        "batched": batched,
        "served": served,
        "convex": convex,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
