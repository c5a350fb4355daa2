#!/usr/bin/env python3
"""Benchmark of the Steel step path on one MI355X.

Workload (the batching configuration): 4096 worlds of uneven size, each a floating-base 12-joint quadruped plus
K ~ Geometric(0.1) <= 64 boxes on a ground plane, FP32, dt = 1/200 s, 4 substeps, 12 iterations, all worlds in
one launch per kernel with the frame's substep loop captured in a hipGraph; 1000 frames from rest.

value       batched world-steps/s: worlds x frames / wall time of frames 0..K-1 after a reset, state resident
            in HBM (warmup frames run before the reset)
plain       world-steps/s of the plain per-world step (one world per launch, per-substep launches, as the
            reference dispatches), frames 0..K-1 of a few of the same worlds
latency     p50/p99 of the batched frame time (HIP events per frame)
roofline    dominant kernel: algorithmic bytes per launch / average launch time, both over a replay of the
            same K frames (each kernel bracketed by HIP events on the library's stream), against 8 TB/s
cpu_baseline  the FP32 CPU oracle (port of the reference step) over the same frames of a sample of the
            worlds, one world at a time per process on the host cores
"""
import argparse
import json
import multiprocessing as mp
import os
import pickle
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "world-steps/s batched vs plain per-world (FP32) + p50/p99 step latency"
HBM_PEAK_GBS = 8000.0


def _pack_world(i):
    from steel_webgpu_amd import pack, workloads
    pw = pack.pack_scene(workloads.quadruped_world(i), capacity_policy="compact")
    pw.scene = None
    return pw


def build_worlds(n, workers, cache=None):
    """Pack worlds 0..n-1 (worker processes); with `cache`, load them from / save them to that file."""
    if cache and os.path.exists(cache):
        with open(cache, "rb") as f:
            worlds = pickle.load(f)  # written by this function
        if len(worlds) >= n:
            return worlds[:n]
    if workers <= 1:
        worlds = [_pack_world(i) for i in range(n)]
    else:
        ctx = mp.get_context("fork")
        with ctx.Pool(workers) as pool:
            worlds = pool.map(_pack_world, range(n), chunksize=max(1, n // (workers * 8)))
    if cache:
        with open(cache, "wb") as f:
            pickle.dump(worlds, f, protocol=pickle.HIGHEST_PROTOCOL)
    return worlds


def _pack_request(args):
    i, nmax, hmax = args
    from steel_webgpu_amd import pack, workloads
    sc, h = workloads.serving_request(i, nmax=nmax, hmax=hmax)
    pw = pack.pack_scene(sc, capacity_policy="compact")
    pw.scene = None
    return pw, h


def serving(args, procs, log):
    """Serving: Poisson requests (rate --rate) of N ~ logU[10, --nmax] bodies (30% hulls, 70% boxes, 0-4
    quadrupeds) and horizons ~ U[50, --hmax] frames through the serving pool with an HBM budget."""
    from steel_webgpu_amd import serve, workloads
    t0 = time.time()
    ctx = mp.get_context("fork")
    with ctx.Pool(procs) as pool:
        packed = pool.map(_pack_request, [(i, args.nmax, args.hmax) for i in range(args.requests)], chunksize=1)
    log(f"packed {len(packed)} requests in {time.time() - t0:.1f}s; bodies mean {np.mean([p.NB for p, _ in packed]):.0f}, "
        f"max {max(p.NB for p, _ in packed)}; frames {sum(h for _, h in packed)}")
    from steel_webgpu_amd import sim
    dev = sim.Device(0)
    results = []
    for rate in args.rates:
        rng = workloads.PCG32(workloads.SEED + 5, int(rate))
        arr = serve.poisson_arrivals(len(packed), rate, rng)
        reqs = [serve.Request(rid=i, arrival=arr[i], world=pw, horizon=h, lane=serve.lane_of(pw)) for i, (pw, h) in enumerate(packed)]
        srv = serve.Server(budget_bytes=int(args.budget_gb * 2**30), device=dev)
        m = serve.run(reqs, srv, readback=())
        srv.destroy()
        m["rate"] = rate
        log(f"  rate {rate}/s: {m['world_steps_per_s']:.0f} world-steps/s, step latency p50 {m['step_latency_ms']['p50']:.2f} ms "
            f"p99 {m['step_latency_ms']['p99']:.2f} ms, queue wait p99 {m['queue_wait_ms']['p99']:.1f} ms, {m['wall_s']:.1f}s")
        results.append(m)
    return results


def copy_calibration(nbytes=1 << 30, reps=4):
    """A plain device copy of known size (k_calib_copy, libsteel_calib.so), for scaling the profiler's
    FETCH_SIZE / WRITE_SIZE counters: each launch reads and writes `nbytes`."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "steel_webgpu_amd", "libsteel_calib.so"))
    lib.steel_calib_copy.argtypes = [ctypes.c_size_t, ctypes.c_int]
    rc = lib.steel_calib_copy(nbytes, reps)
    if rc != 0:
        raise RuntimeError(f"calibration copy failed ({rc})")
    return nbytes


def hull_bytes(pw):
    """16 nv + 24 nf + 4 sum|face| + 16 ne per hull (the per-pair hull traffic of a SAT/manifold query)."""
    out = {}
    NH = pw.h("W_NH")
    SU = pw.statU
    for h in range(NH):
        nv, nf = int(SU[pw.h("W_SU_HULL") + h * 4 + 1]), int(SU[pw.h("W_SU_HULL") + h * 4 + 3])
        fo = int(SU[pw.h("W_SU_HULL") + h * 4 + 2])
        fsum = sum(int(SU[pw.h("W_SU_FACE") + (fo + f) * 2 + 1]) for f in range(nf))
        ne = int(SU[pw.h("W_SU_HEDGE") + h * 2 + 1])
        out[h] = 16 * nv + 24 * nf + 4 * fsum + 16 * ne
    return out


def algorithmic_bytes(batch, worlds):
    """Per-launch compulsory bytes of each kernel, summed over worlds and averaged over the substeps since the
    last reset (stats words 12-15: sums of pairs, kept contacts, rows; substep count).

    collide  288 NS + 296 NP + Hp + 304 NC      (shape pose/AABB, pair record + two shape records, hull data of
                                                 hull pairs, contact write + sort read/key/write)
    prepare  96 NC + 432 NR + 1200 NL           (sorted contact read, row build, articulation factor/response)
    solve    (It+1)(384 NR + 576 NL) + 200 NB + 8 ND   (row read + impulse write per sweep, link propagation,
                                                 predictor/integrator, joint state)
    Hp: hull bytes per pair of the world's last substep pair list times the window's mean pair count.
    """
    tot = np.zeros(3)
    for i, pw in enumerate(worlds):
        st, _ = batch.stats_raw(i)
        nsub = max(int(st[15]), 1)
        NP, NC, NR = st[12] / nsub, st[13] / nsub, st[14] / nsub
        NS, NB, NL, ND = pw.NS, pw.NB, pw.NL, pw.ND
        It = pw.h("W_ITERS")
        hb = hull_bytes(pw)
        lastp = int(st[9])
        pairs = batch.read(i, "pairs", 8 * min(lastp, pw.h("W_MAXP")), dtype=np.uint32).reshape(-1, 2) if lastp else np.zeros((0, 2), np.int64)
        Hp = 0
        SF = pw.statF.reshape(-1, 4)
        for a, b in pairs:
            ka, kb = SF[pw.h("W_SF_SHAPE") + a * 7 + 3][0], SF[pw.h("W_SF_SHAPE") + b * 7 + 3][0]
            if ka == 4 or kb == 4:
                if ka == 0 or kb == 0:  # plane against a convex hull
                    s = b if ka == 4 else a
                    Hp += hb[int(SF[pw.h("W_SF_SHAPE") + s * 7][1])]
                continue
            for s, k in ((a, ka), (b, kb)):
                if k == 0:
                    Hp += hb[int(SF[pw.h("W_SF_SHAPE") + s * 7][1])]
        Hp = Hp / max(len(pairs), 1) * NP
        tot += [288 * NS + 296 * NP + Hp + 304 * NC, 96 * NC + 432 * NR + 1200 * NL,
                (It + 1) * (384 * NR + 576 * NL) + 200 * NB + 8 * ND]
    return tot


def _cpu_worker(args):
    idx, frames = args
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import OracleWorld
    n = 0
    t = 0.0
    for i in idx:
        pw = _pack_world(i)
        o = OracleWorld(pw)
        t0 = time.perf_counter()
        o.frames(frames)
        t += time.perf_counter() - t0
        n += frames
    return n, t


def cpu_baseline(nworlds, frames, procs):
    """FP32 CPU oracle (the port of the reference step) on `procs` processes; returns the aggregate rate
    (all world-steps / the busiest process's stepping time), the step count and the wall time."""
    chunks = [list(range(k, nworlds, procs)) for k in range(procs)]
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(procs) as pool:
        res = pool.map(_cpu_worker, [(c, frames) for c in chunks])
    wall = time.perf_counter() - t0
    steps = sum(r[0] for r in res)
    busy = max(r[1] for r in res)
    return steps / busy, steps, wall


def traffic_from_profile(kernel):
    """HBM bytes per launch of k_<kernel> on this workload, from the committed counter summary (its
    `bytes_per_launch:` line: FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel factor)."""
    path = os.path.join(ROOT, "profiles", f"k_{kernel}_quadruped4096_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("bytes_per_launch:"):
                return int(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed frames, from rest (the configuration's 1000 steps)")
    ap.add_argument("--warmup", type=int, default=20, help="untimed frames before the reset that starts the timed run")
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--plain-worlds", type=int, default=3)
    ap.add_argument("--cpu-worlds", type=int, default=48)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=0, help="prepare/solve workgroup size (0: autotune)")
    ap.add_argument("--worlds-cache", default=None, help="pickle file to load packed worlds from / save them to")
    ap.add_argument("--pack-only", action="store_true", help="pack the worlds into --worlds-cache and exit")
    ap.add_argument("--profile", action="store_true",
                    help="profiler run: batched leg only (no autotune, plain or CPU legs), then a 1 GiB copy for counter scaling")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--workload", default="batch", choices=["batch", "serving"])
    ap.add_argument("--requests", type=int, default=200, help="serving: requests per rate")
    ap.add_argument("--rates", type=float, nargs="+", default=[200.0, 1000.0, 5000.0], help="serving: Poisson rates (req/s)")
    ap.add_argument("--nmax", type=int, default=2000, help="serving: largest request body count")
    ap.add_argument("--hmax", type=int, default=1000, help="serving: longest horizon (frames)")
    ap.add_argument("--budget-gb", type=float, default=64.0, help="serving: HBM budget of the pool")
    args = ap.parse_args()
    if args.profile:
        args.no_cpu = True
        args.plain_worlds = 0
        args.block = args.block or 64
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (worlds shard whole across GPUs with no "
              f"per-step exchange, so N GPUs would run N independent batches).", file=sys.stderr)
        sys.exit(2)
    procs = max(1, min(16, os.cpu_count() or 1))
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    if args.workload == "serving":
        res = serving(args, procs, log)
        tot = sum(r["world_steps"] for r in res)
        wall = sum(r["wall_s"] for r in res)
        print(json.dumps({
            "metric": "served world-steps/s + p50/p99 step latency", "value": round(tot / wall, 1), "unit": "world-steps/s",
            "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic (seeded PCG32 scenes)",
            "config": {"workload": "serving", "requests_per_rate": args.requests, "rates": args.rates, "nmax": args.nmax,
                       "hmax": args.hmax, "budget_gb": args.budget_gb, "fps": 240, "substeps": 4, "iterations": 12,
                       "lanes": "NB<100 / <600 / >=600 bodies, blocks 64/128/256"},
            "per_rate": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()} for r in res]}))
        return
    K = args.steps
    # Host-side work that uses worker processes runs before anything touches the GPU.
    t0 = time.time()
    worlds = build_worlds(args.worlds, procs, args.worlds_cache)
    if args.pack_only:
        return
    log(f"built {len(worlds)} worlds in {time.time() - t0:.1f}s; bodies/world mean {np.mean([w.NB for w in worlds]):.1f} "
        f"max {max(w.NB for w in worlds)}")
    cpu = None
    if not args.no_cpu:
        rate, steps, cwall = cpu_baseline(args.cpu_worlds, K, procs)
        cpu = {"value": round(rate, 1), "unit": "world-steps/s", "cores": procs, "kind": "port",
               "sample": f"worlds 0..{args.cpu_worlds - 1} of the {args.worlds}-world quadruped batch, frames 0..{K - 1} from rest "
                         f"(the timed window), FP32 oracle, one world at a time per process ({steps} world-steps, {cwall:.1f}s wall)"}
        log(f"cpu baseline: {cpu['value']} world-steps/s on {procs} processes")
    from steel_webgpu_amd import sim
    dev = sim.Device(0)
    log(f"device {dev.name}, {dev.num_cus} CUs, {dev.total_mem / 2**30:.0f} GiB")
    graph = not args.no_graph
    batch = sim.Batch(worlds, dev, block=256, graph=graph)
    log(f"batch resident: {batch.device_bytes() / 2**30:.2f} GiB")
    # Autotuner: prepare/solve workgroup size by measured frame time over the first frames.
    block = args.block
    if block == 0:
        best = None
        for bs in (64, 128, 256):
            batch.configure(bs, graph)
            batch.reset()
            ms = float(np.sum(batch.step_timed(10)))
            log(f"  autotune block {bs}: {ms / 10:.2f} ms/frame over frames 0-9")
            if best is None or ms < best[1]:
                best = (bs, ms)
        block = best[0]
    batch.configure(block, graph)
    nw = len(worlds)
    S = worlds[0].h("W_SUBSTEPS")
    CH = 100
    value = ms_per_step = None
    frame_ms = np.zeros(0)
    if not args.profile:
        batch.step(args.warmup)
        batch.reset()
        batch.sync()
        parts = []
        t0 = time.perf_counter()
        for f0 in range(0, K, CH):
            parts.append(batch.step_timed(min(CH, K - f0)))
            log(f"  timed frames {f0}..{f0 + len(parts[-1]) - 1}: {np.mean(parts[-1]):.3f} ms/frame")
        batch.sync()
        wall = time.perf_counter() - t0
        frame_ms = np.concatenate(parts)
        value = nw * K / wall
        ms_per_step = wall / K * 1e3
        log(f"batched: {value:.0f} world-steps/s, {ms_per_step:.3f} ms/frame over frames 0..{K - 1} (block {block})")
    # Per-kernel times on a replay of the same frames (the step is bitwise deterministic), each kernel bracketed
    # by HIP events on the library's stream; the algorithmic bytes are averaged over the same substeps.
    batch.reset()
    acc = np.zeros(3)
    for f0 in range(0, K, CH):
        n = min(CH, K - f0)
        km = batch.profile_kernels(n * S)
        acc += np.asarray(km) * n
        log(f"  kernel replay frames {f0}..{f0 + n - 1}: collide {km[0]:.3f} prepare {km[1]:.3f} solve {km[2]:.3f} ms/launch")
    kms = list(acc / K)
    algo = algorithmic_bytes(batch, worlds)
    names = ["collide", "prepare", "solve"]
    k = int(np.argmax(kms))
    achieved = algo[k] / (kms[k] * 1e-3) / 1e9
    overflow = sum(1 for i in range(nw) if batch.stats_raw(i)[0][1] >= worlds[i].h("W_MAXC"))
    log(f"kernels ms/launch collide {kms[0]:.3f} prepare {kms[1]:.3f} solve {kms[2]:.3f}; algorithmic MB/launch "
        f"{algo[0] / 1e6:.1f} {algo[1] / 1e6:.1f} {algo[2] / 1e6:.1f}; worlds at contact capacity: {overflow}")
    batch.destroy()
    if args.profile:
        copy_calibration()
        log("profile run done")
        return
    # Plain per-world step: one world per batch, one launch per kernel per substep (no graph), same frames.
    plain_rates, plain_ms = [], []
    for i in range(min(args.plain_worlds, nw)):
        b1 = sim.Batch([worlds[i * (nw // max(1, args.plain_worlds))]], dev, block=64, graph=False)
        b1.sync()
        t1 = time.perf_counter()
        fm = b1.step_timed(K)
        b1.sync()
        dtp = time.perf_counter() - t1
        plain_rates.append(K / dtp)
        plain_ms.extend(fm.tolist())
        b1.destroy()
        log(f"  plain world {i * (nw // max(1, args.plain_worlds))}: {K / dtp:.1f} world-steps/s")
    plain = float(np.mean(plain_rates))
    log(f"plain per-world: {plain:.1f} world-steps/s ({np.median(plain_ms):.3f} ms/frame GPU)")
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "world-steps/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded PCG32 scenes)",
        "config": {"workload": "quadruped_batch", "worlds": nw, "fps": 200, "substeps": S, "iterations": worlds[0].h("W_ITERS"),
                   "frames": f"0..{K - 1} from rest", "block": block, "graph": graph},
        "plain_world_steps_per_s": round(plain, 2),
        "speedup_vs_plain": round(value / plain, 1),
        "latency_ms": {"p50": round(float(np.percentile(frame_ms, 50)), 4), "p99": round(float(np.percentile(frame_ms, 99)), 4),
                       "plain_p50": round(float(np.percentile(plain_ms, 50)), 4), "plain_p99": round(float(np.percentile(plain_ms, 99)), 4)},
        "roofline": {"bound": "hbm", "kernel": names[k], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profile(names[k]),
                     "kernel_ms": [round(x, 4) for x in kms], "algorithmic_bytes_per_launch": [int(x) for x in algo]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
