#!/usr/bin/env python3
"""Benchmark of the Steel step path on one MI355X.

Workload (the batching configuration): 4096 worlds of uneven size, each a floating-base 12-joint quadruped plus
K ~ Geometric(0.1) <= 64 boxes on a ground plane, FP32, dt = 1/200 s, 4 substeps, 12 iterations, all worlds in
one launch per kernel with the frame's substep loop captured in a hipGraph; the configuration runs 1000 frames
from rest.

window      the timed frames: all 1000 frames when --steps >= 1000, else the last --steps frames of the 1000
            (frames 1000-K..999, after an untimed advance from rest), where the batch has settled
value       batched world-steps/s: worlds x K / wall time of the window, state resident in HBM
plain       world-steps/s of the plain per-world step (one world per launch, per-substep launches, as the
            reference dispatches) over the same window of a few of the same worlds
latency     p50/p99 of the batched frame time (HIP events per frame)
roofline    dominant kernel: algorithmic bytes per launch / average launch time, both over a replay of the same
            window (each kernel bracketed by HIP events on the library's stream), against 8 TB/s; `traffic` is the
            counter-measured HBM bytes per launch of a committed profile of the same window, or null
cpu_baseline  the FP32 CPU oracle (a port of the reference step) over the same window of a sample of the worlds,
            std::threads over the host cores this process may use (at most 16, the GPU box's share)
parity      the sample worlds' state after the window, GPU against the CPU oracle, bit for bit
fp64_error  one-step error of FP32 against the FP64 oracle from the settled state of 16 sample worlds, per
            quantity (largest difference over the largest FP64 magnitude)
"""
import argparse
import json
import math
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
CONFIG_FRAMES = 1000
# A world's state: everything a step reads that is not static (the step's other tables are rebuilt every substep).
STATE = ("poses", "vel", "art", "anchor_keys", "anchor_vals", "warm", "counters", "stats")


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
    pw = pack.pack_scene(sc, capacity_policy="scaled")
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
        srv = serve.Server(budget_bytes=int(args.budget_gb * 2**30), device=dev, blocks=tuple(args.lane_blocks))
        m = serve.run(reqs, srv, readback=(), log=log)
        srv.destroy()
        m["rate"] = rate
        log(f"  rate {rate}/s: {m['world_steps_per_s']:.0f} world-steps/s, step latency p50 {m['step_latency_ms']['p50']:.2f} ms "
            f"p99 {m['step_latency_ms']['p99']:.2f} ms, queue wait p99 {m['queue_wait_ms']['p99']:.1f} ms, {m['wall_s']:.1f}s")
        results.append(m)
    # Plain per-world reference: sample requests stepped alone, one launch per kernel per substep.
    frames, secs = 0, 0.0
    for i in range(0, len(packed), max(1, len(packed) // args.plain_worlds)) if args.plain_worlds else []:
        pw, _ = packed[i]
        b1 = sim.Batch([pw], dev, block=args.lane_blocks[serve.lane_of(pw)], graph=False)
        b1.sync()
        t1 = time.perf_counter()
        b1.step_timed(20)
        b1.sync()
        secs += time.perf_counter() - t1
        frames += 20
        b1.destroy()
    plain = frames / secs if secs else None
    if plain:
        log(f"  plain per-world over {frames // 20} sample requests x 20 frames: {plain:.1f} world-steps/s")
    dev.close()
    return results, plain


PLAIN_CONFIGS = (
    ("boxes64", "ground plane + 64 boxes, fps 240, 4 substeps", 500),
    ("arm_hulls", "32-link revolute arm among 256 hulls in a walled bin, fps 240, 8 substeps", 1000),
)


def plain_worlds(args, log):
    """The single-world configurations: each world stepped alone on the GPU over its whole run (one launch per
    kernel per substep, HIP events per frame), and by the FP32 oracle on one thread over a bounded prefix of the
    same run (frames in chunks of 25 until --cpu-seconds have passed).  The GPU state after that prefix is
    compared with the oracle's bit for bit, and the oracle then takes one FP32 and one FP64 step from it."""
    from steel_webgpu_amd import pack, workloads
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import OracleWorld, one_step_errors
    cases = []
    for name, desc, frames in PLAIN_CONFIGS:
        pw = pack.pack_scene(getattr(workloads, name)(), capacity_policy="scaled")
        o = OracleWorld(pw)
        done, t0 = 0, time.perf_counter()
        while done < frames and time.perf_counter() - t0 < args.cpu_seconds:
            n = min(25, frames - done)
            o.frames(n)
            done += n
        cpu_s = time.perf_counter() - t0
        log(f"{name}: FP32 oracle, one thread: frames 0..{done - 1} in {cpu_s:.1f}s ({done / cpu_s:.1f} world-steps/s)")
        cases.append((name, desc, frames, pw, o, done, cpu_s))
    from steel_webgpu_amd import sim
    dev = sim.Device(0)
    out = []
    from steel_webgpu_amd import serve
    for name, desc, frames, pw, o, done, cpu_s in cases:
        block = serve.LANE_BLOCKS[serve.lane_of(pw)]
        b = sim.Batch([pw], dev, block=block, graph=False)
        b.sync()
        parts, wall, same = [], 0.0, None
        cuts = sorted(set(range(0, frames, 100)) | {done, frames})
        for f0, f1 in zip(cuts[:-1], cuts[1:]):
            t0 = time.perf_counter()
            parts.append(b.step_timed(f1 - f0))
            b.sync()
            wall += time.perf_counter() - t0
            log(f"  {name} frames {f0}..{f1 - 1}: {np.mean(parts[-1]):.3f} ms/frame")
            if f1 == done:
                same = all(np.array_equal(b.read(0, q).view(np.uint32), getattr(o, q)[:b.read(0, q).size].view(np.uint32))
                           for q in ("poses", "vel", "art"))
        st, _ = b.stats_raw(0)
        b.destroy()
        fm = np.concatenate(parts)
        err, walls = one_step_errors(o)
        log(f"{name}: GPU plain {frames / wall:.1f} world-steps/s ({np.median(fm):.3f} ms/frame p50, {np.percentile(fm, 99):.3f} p99); "
            f"block {block}; bit-exact with the oracle after frame {done - 1}: {same}; last substep {int(st[9])} pairs")
        out.append({"workload": name, "description": desc, "frames": frames, "bodies": pw.NB, "shapes": pw.NS, "block": block,
                    "gpu_world_steps_per_s": round(frames / wall, 2),
                    "gpu_frame_ms": {"p50": round(float(np.median(fm)), 4), "p99": round(float(np.percentile(fm, 99)), 4)},
                    "cpu_fp32_one_thread": {"world_steps_per_s": round(done / cpu_s, 2), "frames": f"0..{done - 1}",
                                            "seconds": round(cpu_s, 2)},
                    "parity": {"bitexact": bool(same), "after_frame": done - 1,
                               "compared": "poses, velocities, articulation state, GPU vs FP32 oracle"},
                    "fp64_one_step_error": {q: {"abs": float(f"{v['abs']:.3g}"), "rel": float(f"{v['rel']:.3g}")} for q, v in err.items()},
                    "cpu_fp64_one_step_s": round(float(walls[1]), 4)})
    dev.close()
    return out


LARGE_FRAMES = 300


def large_world(args, procs, log):
    """The large single world (100k boxes and 12-point hulls on a jittered lattice over a 20 x 20 m floor, 256
    quadrupeds; fps 240, 4 substeps, 300 frames), from rest: frames 0..K-1 (K = --steps, at most 300) stepped one
    workgroup per stage with 1024 threads (the value) and with 256 threads (the same per-world kernels at the batch's
    size), compared bit for bit; device-wide stepping when the library has it on.  Per-kernel times and algorithmic
    bytes come from a replay of the same frames.  The FP32 oracle steps the same frames on one host thread for about
    --cpu-seconds."""
    from steel_webgpu_amd import pack, sim, workloads
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import OracleWorld
    t0 = time.time()
    pw = pack.pack_scene(workloads.large_world(nbodies=args.large_bodies, nquads=args.large_quads,
                                               extent=20.0 * math.sqrt(args.large_bodies / 100000.0), workers=procs),
                         capacity_policy="scaled")
    log(f"large world: {pw.NB} bodies, {pw.NS} shapes, {pw.NA} articulations, MAXC {pw.h('W_MAXC')}; built and packed in "
        f"{time.time() - t0:.1f}s")
    K = min(args.steps, LARGE_FRAMES)
    S = pw.h("W_SUBSTEPS")
    dev = sim.Device(0)
    legs, finals = {}, {}
    for block in (1024, 256):
        b = sim.Batch([pw], dev, block=block, graph=False)
        if block == 1024:
            log(f"resident {b.device_bytes() / 2**30:.2f} GiB")
        b.sync()
        t0 = time.perf_counter()
        fm = b.step_timed(K)
        b.sync()
        wall = time.perf_counter() - t0
        st = b.stats_raw(0)[0]
        counts = {"pairs": int(st[9]), "contacts": int(st[5]), "rows": int(st[10]), "colors": int(st[6]), "units": int(st[11])}
        legs[block] = {"world_steps_per_s": K / wall, "frame_ms": fm, "counts": counts}
        finals[block] = [b.read(0, q, dtype=np.uint32) for q in ("poses", "vel", "art")]
        log(f"one workgroup of {block} threads per stage: {K / wall:.3f} world-steps/s over frames 0..{K - 1} "
            f"({np.median(fm):.1f} ms/frame p50); last substep {counts}")
        if block == 1024:
            b.reset()
            base = [b.stats_raw(0)[0].astype(np.int64)]
            acc = np.zeros(3)
            for f in range(K):
                acc += np.asarray(b.profile_kernels(S))
            kms = list(acc / K)
            algo = algorithmic_bytes(b, [pw], base)
            log(f"kernels ms/launch collide {kms[0]:.3f} prepare {kms[1]:.3f} solve {kms[2]:.3f}; algorithmic MB/launch "
                f"{algo[0] / 1e6:.2f} {algo[1] / 1e6:.2f} {algo[2] / 1e6:.2f}")
        b.destroy()
    same = all(np.array_equal(x, y) for x, y in zip(finals[1024], finals[256]))
    wide = None
    b = sim.Batch([pw], dev, block=256, graph=False)
    try:
        b.set_wide(True)
        t0 = time.perf_counter()
        fmw = b.step_timed(K)
        b.sync()
        wide = K / (time.perf_counter() - t0)
        log(f"device-wide: {wide:.3f} world-steps/s")
    except sim.UnsupportedError as e:
        log(f"device-wide: not available ({e})")
    b.destroy()
    dev.close()
    names = ["collide", "prepare", "solve"]
    k = int(np.argmax(kms))
    achieved = algo[k] / (kms[k] * 1e-3) / 1e9
    o = OracleWorld(pw)
    done, t0 = 0, time.perf_counter()
    while done < K and (done == 0 or time.perf_counter() - t0 < args.cpu_seconds):
        o.frames(1)
        done += 1
    cpu_s = time.perf_counter() - t0
    ok = np.array_equal(o.poses.view(np.uint32), finals[1024][0]) if done == K else None
    log(f"FP32 oracle, one thread: {done} frames in {cpu_s:.1f}s ({done / cpu_s:.4f} world-steps/s)"
        + (f"; bit-exact with the GPU after frame {K - 1}: {ok}" if ok is not None else ""))
    L = legs[1024]
    return {
        "metric": "large single world: world-steps/s (FP32)", "value": round(L["world_steps_per_s"], 4), "unit": "world-steps/s",
        "n_gpus": 1, "steps": K, "warmup": 0, "ms_per_step": round(1e3 / L["world_steps_per_s"], 2), "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded PCG32 scene)",
        "config": {"workload": "large_world", "bodies": pw.NB, "shapes": pw.NS, "articulations": pw.NA, "fps": 240, "substeps": S,
                   "iterations": pw.h("W_ITERS"), "frames": f"0..{K - 1} from rest", "path": "one workgroup of 1024 threads per stage",
                   "last_substep": L["counts"]},
        "plain_world_steps_per_s": round(legs[256]["world_steps_per_s"], 4),
        "plain": "the same per-world kernels at the batch's 256 threads per stage",
        "speedup_vs_plain": round(L["world_steps_per_s"] / legs[256]["world_steps_per_s"], 2),
        "device_wide_world_steps_per_s": wide,
        "parity": {"bitexact_1024_vs_256_threads": bool(same), "bitexact_vs_fp32_oracle": ok, "frames": f"0..{K - 1}"},
        "latency_ms": {"p50": round(float(np.median(L["frame_ms"])), 2), "p99": round(float(np.percentile(L["frame_ms"], 99)), 2)},
        "roofline": {"bound": "hbm", "kernel": names[k], "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel_ms": [round(x, 3) for x in kms], "algorithmic_bytes_per_launch": [int(x) for x in algo]},
        "cpu_baseline": {"value": round(done / cpu_s, 5), "unit": "world-steps/s", "cores": 1, "kind": "port",
                         "sample": f"frames 0..{done - 1} of the large world from rest, FP32 oracle, one thread ({cpu_s:.1f}s)"},
    }


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


def algorithmic_bytes(batch, worlds, base=None):
    """Per-launch compulsory bytes of each kernel, summed over worlds and averaged over the substeps of the
    window: stats words 12-15 (running sums of pairs, kept contacts, rows; substep count) now minus `base`
    (the same words at the window's start; None: since the reset).

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
        st = st.astype(np.int64)
        if base is not None:
            st[12:16] -= base[i][12:16]
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


def cpu_threads():
    """Host threads for the CPU legs: the cores this process may run on, at most 16 (the GPU box's CPU share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(worlds, f0, k, threads, log):
    """FP32 CPU oracle over the window: the sample worlds advance frames 0..f0-1 untimed, then step frames
    f0..f0+k-1 timed, on `threads` std::threads in one process.  Returns the rate, the oracle worlds (their
    state at frame f0+k) and a description."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import OracleWorld, run_batch
    ows = [OracleWorld(pw) for pw in worlds]
    t0 = time.perf_counter()
    wall, busy = run_batch(ows, f0, k, threads)
    total = time.perf_counter() - t0
    rate = len(ows) * k / wall
    log(f"cpu baseline: {rate:.1f} world-steps/s on {threads} threads ({len(ows)} worlds x frames {f0}..{f0 + k - 1}, "
        f"{wall:.2f}s timed, {sum(busy):.1f} CPU-s; {total:.1f}s with the advance)")
    return rate, ows, wall, sum(busy)


def fp64_reference(ows, log):
    """One-step error of FP32 against the FP64 oracle from each world's current state, per quantity: the largest
    absolute difference over the worlds, and the norm-wise relative error (largest difference over the largest
    FP64 magnitude of the world) as its median and maximum over the worlds; plus the single-thread FP32 and FP64
    step rates."""
    from oracle_bind import one_step_errors
    per, t32, t64 = {}, 0.0, 0.0
    for o in ows:
        e, walls = one_step_errors(o)
        t32 += walls[0]
        t64 += walls[1]
        for q, v in e.items():
            per.setdefault(q, []).append((v["abs"], v["rel"]))
    err = {}
    for q, vals in per.items():
        a = np.asarray(vals)
        err[q] = {"max_abs": float(f"{a[:, 0].max():.3g}"), "median_rel": float(f"{np.median(a[:, 1]):.3g}"),
                  "max_rel": float(f"{a[:, 1].max():.3g}")}
    log(f"fp64 one-step error over {len(ows)} worlds (median / max relative): " +
        ", ".join(f"{q} {v['median_rel']:.1e}/{v['max_rel']:.1e}" for q, v in err.items()))
    return err, len(ows) / t32, len(ows) / t64


def traffic_from_profile(kernel, window):
    """HBM bytes per launch of a stage (collide, prepare or solve) from the committed counter summary of this workload
    (its `bytes_per_launch <stage>:` lines, FETCH_SIZE + WRITE_SIZE per dispatch scaled by the copy-kernel factor),
    when its `window:` line names the same frames as this run; None otherwise."""
    path = os.path.join(ROOT, "profiles", "quadruped4096_traffic.md")
    if not os.path.exists(path):
        return None, None
    got, win = None, None
    with open(path) as f:
        for line in f:
            if line.startswith(f"bytes_per_launch {kernel}:"):
                got = int(line.split(":")[1])
            if line.startswith("window:"):
                win = line.split(":", 1)[1].strip()
    return (got, win) if win == window else (None, win)


def save_states(batch, idx):
    return [{n: batch.read(i, n, dtype=np.uint8) for n in STATE} for i in idx]


def restore_states(batch, idx, states):
    for i, st in zip(idx, states):
        for n in STATE:
            batch.write(i, n, st[n])


def write_state_file(path, states):
    arrs = {}
    for n in STATE:
        arrs[n] = np.concatenate([st[n] for st in states])
        arrs[n + "_len"] = np.asarray([st[n].size for st in states], np.int64)
    np.savez(path, **arrs)


def read_state_file(path):
    z = np.load(path, allow_pickle=False)
    out = None
    for n in STATE:
        lens = z[n + "_len"]
        parts = np.split(z[n], np.cumsum(lens)[:-1])
        if out is None:
            out = [{} for _ in parts]
        for st, p in zip(out, parts):
            st[n] = p
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=CONFIG_FRAMES,
                    help="timed frames: the configuration's 1000, or the last K of them after an untimed advance")
    ap.add_argument("--warmup", type=int, default=20, help="untimed frames before the reset that starts the run")
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--plain-worlds", type=int, default=3)
    ap.add_argument("--cpu-worlds", type=int, default=0, help="CPU sample worlds (0: enough for about 4000 timed world-steps, 48..256)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=0, help="prepare/solve workgroup size (0: autotune)")
    ap.add_argument("--worlds-cache", default=None, help="pickle file to load packed worlds from / save them to")
    ap.add_argument("--pack-only", action="store_true", help="pack the worlds into --worlds-cache and exit")
    ap.add_argument("--profile", action="store_true",
                    help="profiler run: batched leg over the window only (no autotune, plain or CPU legs), then a 1 GiB copy")
    ap.add_argument("--state-out", default=None, help="write every world's state at the window's start to this .npz and exit")
    ap.add_argument("--state-in", default=None, help="start the window from this .npz (written by --state-out) instead of advancing")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--workload", default="batch", choices=["batch", "serving", "plain", "large"])
    ap.add_argument("--large-bodies", type=int, default=100000, help="large: lattice bodies")
    ap.add_argument("--large-quads", type=int, default=256, help="large: quadrupeds")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="plain: FP32 oracle time budget per configuration")
    ap.add_argument("--requests", type=int, default=200, help="serving: requests per rate")
    ap.add_argument("--rates", type=float, nargs="+", default=[200.0, 1000.0, 5000.0], help="serving: Poisson rates (req/s)")
    ap.add_argument("--nmax", type=int, default=2000, help="serving: largest request body count")
    ap.add_argument("--hmax", type=int, default=1000, help="serving: longest horizon (frames)")
    ap.add_argument("--budget-gb", type=float, default=64.0, help="serving: HBM budget of the pool")
    ap.add_argument("--lane-blocks", type=int, nargs=3, default=None, help="serving: workgroup size of each lane")
    args = ap.parse_args()
    if args.lane_blocks is None:
        from steel_webgpu_amd import serve
        args.lane_blocks = list(serve.LANE_BLOCKS)
    if args.profile:
        args.no_cpu = True
        args.plain_worlds = 0
        args.warmup = 0
        args.block = args.block or 64
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (worlds shard whole across GPUs with no "
              f"per-step exchange, so N GPUs would run N independent batches).", file=sys.stderr)
        sys.exit(2)
    procs = cpu_threads()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    if args.workload == "large":
        print(json.dumps(large_world(args, procs, log)))
        return
    if args.workload == "plain":
        res = plain_worlds(args, log)
        print(json.dumps({"metric": "plain single-world world-steps/s (FP32) + p50/p99 frame time", "unit": "world-steps/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic (seeded PCG32 scenes)",
                          "per_config": res}))
        return
    if args.workload == "serving":
        res, plain = serving(args, procs, log)
        tot = sum(r["world_steps"] for r in res)
        wall = sum(r["wall_s"] for r in res)
        print(json.dumps({
            "metric": "served world-steps/s + p50/p99 step latency", "value": round(tot / wall, 1), "unit": "world-steps/s",
            "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic (seeded PCG32 scenes)",
            "config": {"workload": "serving", "requests_per_rate": args.requests, "rates": args.rates, "nmax": args.nmax,
                       "hmax": args.hmax, "budget_gb": args.budget_gb, "fps": 240, "substeps": 4, "iterations": 12,
                       "lanes": "NB<100 / <600 / >=600 bodies, blocks " + "/".join(str(x) for x in args.lane_blocks)},
            "plain_world_steps_per_s": round(plain, 2) if plain else None,
            "served_vs_plain": {str(r["rate"]): round(r["world_steps_per_s"] / plain, 1) for r in res} if plain else None,
            "per_rate": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()} for r in res]}))
        return
    K = args.steps
    F0 = CONFIG_FRAMES - K if K < CONFIG_FRAMES else 0
    window = f"{F0}..{F0 + K - 1}"
    # Host-side work that uses worker processes runs before anything touches the GPU.
    t0 = time.time()
    worlds = build_worlds(args.worlds, procs, args.worlds_cache)
    if args.pack_only:
        return
    nw = len(worlds)
    log(f"built {nw} worlds in {time.time() - t0:.1f}s; bodies/world mean {np.mean([w.NB for w in worlds]):.1f} "
        f"max {max(w.NB for w in worlds)}; window frames {window}")
    cpu = ows = None
    if not args.no_cpu:
        ncw = args.cpu_worlds or int(min(256, max(48, -(-4000 // K))))
        rate, ows, cwall, cbusy = cpu_baseline(worlds[:ncw], F0, K, procs, log)
        cpu = {"value": round(rate, 1), "unit": "world-steps/s", "cores": procs, "kind": "port",
               "sample": f"worlds 0..{ncw - 1} of the {nw}-world quadruped batch over the timed window (frames {window}, "
                         f"advanced untimed from rest), FP32 oracle, std::threads in one process ({ncw * K} world-steps, "
                         f"{cwall:.2f}s wall, {cbusy:.1f} CPU-s)"}
    from steel_webgpu_amd import sim
    dev = sim.Device(0)
    log(f"device {dev.name}, {dev.num_cus} CUs, {dev.total_mem / 2**30:.0f} GiB")
    graph = not args.no_graph
    batch = sim.Batch(worlds, dev, block=64, graph=graph)
    log(f"batch resident: {batch.device_bytes() / 2**30:.2f} GiB")
    S = worlds[0].h("W_SUBSTEPS")
    everyone = list(range(nw))
    if args.warmup:
        batch.step(args.warmup)
    batch.reset()
    # The window's starting state: advanced from rest (graph replays) or read from a state file.
    t0 = time.perf_counter()
    if args.state_in:
        start = read_state_file(args.state_in)
        restore_states(batch, everyone, start)
        log(f"window start loaded from {args.state_in}")
    else:
        # Under a profiler, frames are advanced one at a time with a sync each: runs that queued 100 frames of
        # kernels per call ended in a SIGSEGV on a profiler thread.
        chunk = 1 if args.profile else 100
        for f in range(0, F0, chunk):
            batch.step(min(chunk, F0 - f))
            if args.profile:
                batch.sync()
        batch.sync()
        start = save_states(batch, everyone) if F0 else None
        log(f"advanced frames 0..{F0 - 1} untimed in {time.perf_counter() - t0:.1f}s")
    if args.state_out:
        write_state_file(args.state_out, start if start is not None else save_states(batch, everyone))
        log(f"window start written to {args.state_out}")
        batch.destroy()
        dev.close()
        return

    def to_start():
        batch.reset()
        if start is not None:
            restore_states(batch, everyone, start)

    # The library's autotuner: workgroup size x narrowphase grid, timed over the window's first frames from its
    # starting state (restored after each trial), cached per device.
    block, tune = args.block, None
    if block == 0:
        to_start()
        tune = batch.autotune(frames=10)
        block = tune["block"]
        log(f"  autotune: block {block}, narrowphase {tune['narrow_per_cu']} workgroups/CU "
            f"({tune['trials']} schedules x 10 frames from frame {F0}; {tune['frame_ms']:.2f} ms/frame)")
    batch.configure(block, graph)
    CH = 100
    value = ms_per_step = None
    frame_ms = np.zeros(0)
    plain_start = []
    parity = None
    if not args.profile:
        to_start()
        pidx = [i * (nw // max(1, args.plain_worlds)) for i in range(min(args.plain_worlds, nw))]
        plain_start = save_states(batch, pidx) if start is not None else []
        batch.sync()
        parts = []
        t0 = time.perf_counter()
        for f0 in range(0, K, CH):
            parts.append(batch.step_timed(min(CH, K - f0)))
            log(f"  timed frames {F0 + f0}..{F0 + f0 + len(parts[-1]) - 1}: {np.mean(parts[-1]):.3f} ms/frame")
        batch.sync()
        wall = time.perf_counter() - t0
        frame_ms = np.concatenate(parts)
        value = nw * K / wall
        ms_per_step = wall / K * 1e3
        log(f"batched: {value:.0f} world-steps/s, {ms_per_step:.3f} ms/frame over frames {window} (block {block})")
        if ows is not None:
            same = 0
            for i, o in enumerate(ows):
                eq = all(np.array_equal(batch.read(i, n).view(np.uint32), getattr(o, n)[:batch.read(i, n).size].view(np.uint32))
                         for n in ("poses", "vel", "art"))
                same += int(eq)
            parity = {"worlds": len(ows), "bitexact": same, "frames": f"0..{F0 + K - 1}",
                      "compared": "poses, velocities, articulation state after the window, GPU batch vs FP32 oracle"}
            log(f"parity: {same} of {len(ows)} sample worlds bit-exact against the FP32 oracle after frame {F0 + K - 1}")
    # Per-kernel times on a replay of the same window (the step is bitwise deterministic), each kernel bracketed
    # by HIP events on the library's stream; the algorithmic bytes are the same window's.
    to_start()
    base = [batch.stats_raw(i)[0].astype(np.int64) for i in everyone]
    acc = np.zeros(3)
    for f0 in range(0, K, CH):
        n = min(CH, K - f0)
        km = batch.profile_kernels(n * S)
        acc += np.asarray(km) * n
        log(f"  kernel replay frames {F0 + f0}..{F0 + f0 + n - 1}: collide {km[0]:.3f} prepare {km[1]:.3f} solve {km[2]:.3f} ms/launch")
    kms = list(acc / K)
    algo = algorithmic_bytes(batch, worlds, base)
    names = ["collide", "prepare", "solve"]
    k = int(np.argmax(kms))
    achieved = algo[k] / (kms[k] * 1e-3) / 1e9
    overflow = sum(1 for i in everyone if batch.stats_raw(i)[0][1] >= worlds[i].h("W_MAXC"))
    log(f"kernels ms/launch collide {kms[0]:.3f} prepare {kms[1]:.3f} solve {kms[2]:.3f}; algorithmic MB/launch "
        f"{algo[0] / 1e6:.1f} {algo[1] / 1e6:.1f} {algo[2] / 1e6:.1f}; worlds at contact capacity: {overflow}")
    if args.profile:
        import hashlib
        h = hashlib.sha256()
        for i in everyone:
            for n in ("poses", "vel", "art"):
                h.update(batch.read(i, n, dtype=np.uint8).tobytes())
        log(f"state after the window: sha256 {h.hexdigest()[:16]}")
    batch.destroy()
    if args.profile:
        copy_calibration()
        dev.close()
        log("profile run done")
        return
    # Plain per-world step: one world per batch, one launch per kernel per substep (no graph), same window.
    plain_rates, plain_ms = [], []
    for j, i in enumerate(pidx):
        b1 = sim.Batch([worlds[i]], dev, block=64, graph=False)
        if plain_start:
            restore_states(b1, [0], [plain_start[j]])
        b1.sync()
        t1 = time.perf_counter()
        fm = b1.step_timed(K)
        b1.sync()
        dtp = time.perf_counter() - t1
        plain_rates.append(K / dtp)
        plain_ms.extend(fm.tolist())
        b1.destroy()
        log(f"  plain world {i}: {K / dtp:.1f} world-steps/s")
    dev.close()
    plain = float(np.mean(plain_rates))
    log(f"plain per-world: {plain:.1f} world-steps/s ({np.median(plain_ms):.3f} ms/frame GPU)")
    fp64 = None
    if ows is not None:
        err, r32, r64 = fp64_reference(ows[:16], log)
        fp64 = {"one_step_rel_error": err, "state": f"after frame {F0 + K - 1} of worlds 0..{min(16, len(ows)) - 1}",
                "cpu_fp64_reference_world_steps_per_s": round(r64, 1), "cpu_fp32_single_thread_world_steps_per_s": round(r32, 1),
                "note": "FP64 oracle: the error reference, timed on one thread; not a baseline"}
    traffic, twin = traffic_from_profile(names[k], window)
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "world-steps/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded PCG32 scenes)",
        "config": {"workload": "quadruped_batch", "worlds": nw, "fps": 200, "substeps": S, "iterations": worlds[0].h("W_ITERS"),
                   "frames": window + (f" (frames 0..{F0 - 1} of the 1000-frame run advanced untimed)" if F0 else " from rest"),
                   "block": block, "graph": graph, "autotune": tune},
        "plain_world_steps_per_s": round(plain, 2),
        "speedup_vs_plain": round(value / plain, 1),
        "latency_ms": {"p50": round(float(np.percentile(frame_ms, 50)), 4), "p99": round(float(np.percentile(frame_ms, 99)), 4),
                       "plain_p50": round(float(np.percentile(plain_ms, 50)), 4), "plain_p99": round(float(np.percentile(plain_ms, 99)), 4)},
        "roofline": {"bound": "hbm", "kernel": names[k], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_window": twin if traffic is not None else f"none for frames {window} (profile holds {twin})",
                     "kernel_ms": [round(x, 4) for x in kms], "algorithmic_bytes_per_launch": [int(x) for x in algo]},
        "cpu_baseline": cpu,
        "parity": parity,
        "fp64_error": fp64,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
