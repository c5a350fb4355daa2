"""Serving: a request loop over the library's serving pool (steel_server_* in include/steel_amd.h).

A request is a packed world and a horizon in frames.  Requests arrive over time; each is admitted into the
device pool when the HBM budget allows (first come, first served), then stepped frame by frame in the lane
of its size class together with every other admitted world of that lane.  This is continuous batching:
worlds join and leave a lane between its frames, and lanes run concurrently on their own streams.  When a
request reaches its horizon its state is read back and its pool block released.  Every world is stepped by
the same kernels as a world stepped alone, so results do not depend on the mix (tests check it bit for bit).

Step latency is the host-observed time from a lane frame's launch to its completion, counted once per world
stepped in that frame.
"""
from __future__ import annotations

import collections
import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import sim as _sim

# Lane (size class) by body count: small worlds, medium worlds, large worlds; workgroup size per lane.
LANE_EDGES = (100, 600)
LANE_BLOCKS = (64, 128, 1024)


class AdmitRefused(RuntimeError):
    """The library refused to admit a world (steel_server_admit returned a status other than out-of-memory)."""


def lane_of(pw, edges=LANE_EDGES):
    return sum(pw.NB >= e for e in edges)


class Server:
    """The device side: pool of `budget_bytes`, `lanes` lanes (steel_server_create)."""

    def __init__(self, budget_bytes, lanes=len(LANE_BLOCKS), blocks=LANE_BLOCKS, device=None):
        self.lib = _sim.load_library()
        self.device = device or _sim.default_device()
        h = ctypes.c_void_p()
        _sim._check(self.lib.steel_server_create(self.device.handle, int(budget_bytes), lanes, ctypes.byref(h)))
        self.handle = h
        self.nlanes = lanes
        for l in range(lanes):
            _sim._check(self.lib.steel_server_configure_lane(self.handle, l, blocks[min(l, len(blocks) - 1)]))
        self._sizes = {}

    @staticmethod
    def world_bytes(pw):
        lib = _sim.load_library()
        arr, keep = _sim._descs([pw])
        out = ctypes.c_size_t()
        _sim._check(lib.steel_server_world_bytes(arr, ctypes.byref(out)))
        return out.value

    def admit(self, pw, lane):
        """World id, or None when the pool cannot hold the world now (nothing was allocated).  Raises
        AdmitRefused when the library refuses the world itself (a world it cannot step, or a substep count
        other than that of the lane's current worlds)."""
        arr, keep = _sim._descs([pw])
        wid = ctypes.c_uint32()
        rc = self.lib.steel_server_admit(self.handle, arr, lane, ctypes.byref(wid))
        if rc == _sim.STEEL_OUT_OF_MEMORY:
            return None
        if rc in (_sim.STEEL_UNSUPPORTED, _sim.STEEL_BAD_ARGUMENT):
            raise AdmitRefused(_sim.last_error())
        _sim._check(rc)
        self._sizes[wid.value] = pw
        return wid.value

    def release(self, wid):
        _sim._check(self.lib.steel_server_release(self.handle, wid))
        self._sizes.pop(wid, None)

    def step(self, lane):
        _sim._check(self.lib.steel_server_step_lane(self.handle, lane))

    def poll(self, lane):
        done, n, ms = ctypes.c_int(), ctypes.c_uint32(), ctypes.c_float()
        _sim._check(self.lib.steel_server_lane_poll(self.handle, lane, ctypes.byref(done), ctypes.byref(n), ctypes.byref(ms)))
        return bool(done.value), n.value, ms.value

    def read(self, wid, name, nbytes=None, offset=0, dtype=np.float32):
        pw = self._sizes[wid]
        if nbytes is None:
            nbytes = buffer_bytes(pw, name) - offset
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
        _sim._check(self.lib.steel_server_read(self.handle, wid, _sim.BUF[name], offset, out.nbytes, out.ctypes.data))
        return out

    def write(self, wid, name, data, offset=0):
        """Control input for an admitted world between lane frames (Sim's writeBuffer: joint targets, wrenches,
        poses, twists, the control flags word)."""
        data = np.ascontiguousarray(data)
        _sim._check(self.lib.steel_server_write(self.handle, wid, _sim.BUF[name], offset, data.nbytes, data.ctypes.data))

    def info(self):
        used, cap, n = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_uint32()
        _sim._check(self.lib.steel_server_info(self.handle, ctypes.byref(used), ctypes.byref(cap), ctypes.byref(n)))
        return used.value, cap.value, n.value

    def destroy(self):
        if self.handle:
            self.lib.steel_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def buffer_bytes(pw, name):
    """Bytes of one of a world's state buffers (the sizes the library allocates)."""
    if name == "poses":
        return pw.h("W_NPOSES") * 16
    if name == "vel":
        return max(pw.NB, 1) * 32
    if name == "art":
        return pw.h("W_NART") * 4
    if name == "stats":
        return 64
    if name == "counters":
        return 32
    if name == "warm":
        return pw.h("W_WH") * 40
    raise ValueError(f"buffer {name}: use an explicit byte count")


@dataclass
class Request:
    rid: int
    arrival: float            # seconds after the loop starts
    world: object             # PackedWorld
    horizon: int              # frames
    lane: int = 0
    wid: int = -1
    frames: int = 0
    admitted: float = math.nan
    finished: float = math.nan
    result: dict = field(default_factory=dict)


def poisson_arrivals(n, rate, rng):
    """n arrival times of a Poisson process of `rate` per second (exponential gaps from a uniform stream)."""
    t, out = 0.0, []
    for _ in range(n):
        t += -math.log(1.0 - rng.uniform()) / rate
        out.append(t)
    return out


def run(requests, server, readback=("poses",), clock=time.perf_counter, log=None):
    """Serve `requests` (arrival order) through `server` until every one has finished; returns metrics."""
    pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
    queue = collections.deque()
    L = server.nlanes
    members = [[] for _ in range(L)]
    in_flight = [False] * L
    launched = [0.0] * L
    frame_members = [[] for _ in range(L)]
    lat, lat_w, gpu_ms = [], [], []
    deferrals, rejected, steps, peak = 0, 0, 0, 0
    finished = 0
    total = len(requests)
    t0 = clock()
    next_log = 10.0
    while finished < total:
        now = clock() - t0
        if log is not None and now >= next_log:
            log(f"    {now:.0f}s: {finished} of {total} requests finished, {sum(len(m) for m in members)} worlds resident, "
                f"{len(queue)} waiting, {steps} world-steps")
            next_log += 10.0
        while pending and pending[0].arrival <= now:
            queue.append(pending.popleft())
        while queue:
            r = queue[0]
            if members[r.lane] and members[r.lane][0].world.h("W_SUBSTEPS") != r.world.h("W_SUBSTEPS"):
                deferrals += 1             # its lane steps another substep count: wait until the lane drains
                break
            try:
                wid = server.admit(r.world, r.lane)
            except AdmitRefused:
                rejected += 1              # a world the library cannot step
                queue.popleft()
                r.finished = now
                finished += 1
                continue
            if wid is None:
                if not any(members):
                    rejected += 1          # larger than the whole pool: it can never be admitted
                    queue.popleft()
                    r.finished = now
                    finished += 1
                    continue
                deferrals += 1
                break
            r.wid, r.admitted = wid, now
            members[r.lane].append(r)
            queue.popleft()
            peak = max(peak, server.info()[0])
        for l in range(L):
            if in_flight[l]:
                done, n, ms = server.poll(l)
                if not done:
                    continue
                t = clock() - t0
                lat.append(t - launched[l])
                lat_w.append(len(frame_members[l]))
                gpu_ms.append(ms)
                steps += len(frame_members[l])
                for r in frame_members[l]:
                    r.frames += 1
                    if r.frames >= r.horizon:
                        for name in readback:
                            r.result[name] = server.read(r.wid, name)
                        server.release(r.wid)
                        members[l].remove(r)
                        r.finished = t
                        finished += 1
                in_flight[l] = False
            if not in_flight[l] and members[l]:
                launched[l] = clock() - t0
                server.step(l)
                frame_members[l] = list(members[l])
                in_flight[l] = True
        if not any(in_flight) and not queue and pending:
            wait = pending[0].arrival - (clock() - t0)
            if wait > 0:
                time.sleep(min(wait, 0.01))
    wall = clock() - t0
    lat = np.asarray(lat)
    w = np.asarray(lat_w, np.float64)

    def wpct(q):
        if lat.size == 0:
            return math.nan
        o = np.argsort(lat)
        c = np.cumsum(w[o])
        return float(lat[o][min(np.searchsorted(c, q / 100.0 * c[-1]), lat.size - 1)])

    done_req = [r for r in requests if not math.isnan(r.admitted)]
    first = min(r.arrival for r in requests) if requests else 0.0
    return {
        "requests": total, "rejected": rejected, "world_steps": steps, "wall_s": wall,
        "world_steps_per_s": steps / max(wall - first, 1e-9),
        "requests_per_s": len(done_req) / max(wall - first, 1e-9),
        "step_latency_ms": {"p50": 1e3 * wpct(50), "p99": 1e3 * wpct(99)},
        "lane_frame_gpu_ms": {"p50": float(np.percentile(gpu_ms, 50)) if gpu_ms else math.nan,
                              "p99": float(np.percentile(gpu_ms, 99)) if gpu_ms else math.nan},
        "queue_wait_ms": {"p50": 1e3 * float(np.percentile([r.admitted - r.arrival for r in done_req], 50)) if done_req else math.nan,
                          "p99": 1e3 * float(np.percentile([r.admitted - r.arrival for r in done_req], 99)) if done_req else math.nan},
        "admission_deferrals": deferrals, "peak_pool_bytes": peak,
    }
