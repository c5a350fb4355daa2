"""Host side of the step path: the reference's `Sim` API over the C ABI of libsteel_amd.so.

`Sim` mirrors packages/steel-webgpu/src/sim.js:63-835 (typed in src/sim.d.ts:53-92): the constructor packs
the scene (`pack.pack_scene`, the `build()` half of sim.js) and creates a one-world batch on the device;
`step`, `reset`, `state`, `read`, `poses`, `stats` and the setters keep their names, arguments and layouts.
`Batch` steps many packed worlds in one launch per kernel.  Nothing here computes physics: every step runs
the HIP kernels, and a missing library or device raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import pack as _pack

_HERE = os.path.dirname(os.path.abspath(__file__))
# STEEL_AMD_LIB names another build of the same library (e.g. one compiled with other waves-per-SIMD settings).
LIB_PATH = os.environ.get("STEEL_AMD_LIB") or os.path.join(_HERE, "libsteel_amd.so")

STEEL_OK, STEEL_ERROR, STEEL_UNSUPPORTED, STEEL_BAD_ARGUMENT, STEEL_OUT_OF_MEMORY = 0, -1, -2, -3, -4
BUF = dict(poses=0, vel=1, art=2, anchor_keys=3, anchor_vals=4, pairs=5, contacts=6, rows=7, sched=8, counters=9, stats=10, warm=11)
STAGE_COLLIDE, STAGE_NARROW, STAGE_PREPARE, STAGE_SOLVE, STAGE_ALL = 1, 2, 4, 8, 15
# Symbols include/steel_amd.h declares.
EXPORTS = ["steel_last_error", "steel_device_open", "steel_device_close", "steel_device_info", "steel_batch_create",
           "steel_batch_destroy", "steel_batch_reset", "steel_batch_step", "steel_batch_substep_stages", "steel_batch_sync",
           "steel_batch_read", "steel_batch_write", "steel_batch_buffer_size", "steel_batch_stats", "steel_batch_configure",
           "steel_batch_device_bytes", "steel_estimate_bytes", "steel_batch_step_timed", "steel_batch_profile_kernels",
           "steel_server_create", "steel_server_destroy", "steel_server_configure_lane", "steel_server_world_bytes",
           "steel_server_admit", "steel_server_release", "steel_server_step_lane", "steel_server_lane_poll", "steel_server_read", "steel_server_write",
           "steel_server_info", "steel_world_step_noalloc", "steel_batch_autotune", "steel_batch_set_wide"]


class UnsupportedError(RuntimeError):
    """No usable GPU (the reference's UnsupportedError, src/device.js:1-9)."""


class SteelError(RuntimeError):
    pass


class WorldDeviceArrays(ctypes.Structure):
    """steel_world_device_arrays: device pointers of one world's static arrays and state."""
    _fields_ = [(n, ctypes.c_void_p) for n in ("header", "statF", "statU", "poses", "vel", "art", "anchor_keys", "anchor_vals",
                                                "warm", "counters", "stats")]


class TuneResult(ctypes.Structure):
    """steel_tune_result: the schedule the autotuner picked."""
    _fields_ = [("block", ctypes.c_uint32), ("narrow_per_cu", ctypes.c_uint32), ("from_cache", ctypes.c_uint32),
                ("trials", ctypes.c_uint32), ("frame_ms", ctypes.c_float)]


STEEL_TUNE_REMEASURE = 1


class WorldDesc(ctypes.Structure):
    _fields_ = [("header", ctypes.c_void_p), ("statF", ctypes.c_void_p), ("statU", ctypes.c_void_p),
                ("poses0", ctypes.c_void_p), ("vel0", ctypes.c_void_p), ("art0", ctypes.c_void_p)]


_lib = None


def load_library(path=LIB_PATH):
    """Load libsteel_amd.so (built by __graft_entry__.build); raise if it is not there."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SteelError(f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(path)
    V, U32, I32, SZ = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_size_t
    sig = {
        "steel_last_error": ([], ctypes.c_char_p),
        "steel_device_open": ([ctypes.c_int, ctypes.POINTER(V)], ctypes.c_int),
        "steel_device_close": ([V], ctypes.c_int),
        "steel_device_info": ([V, ctypes.c_char_p, SZ, ctypes.POINTER(SZ), ctypes.POINTER(U32)], ctypes.c_int),
        "steel_batch_create": ([V, U32, ctypes.POINTER(WorldDesc), ctypes.POINTER(V)], ctypes.c_int),
        "steel_batch_destroy": ([V], ctypes.c_int),
        "steel_batch_reset": ([V, I32], ctypes.c_int),
        "steel_batch_step": ([V, U32], ctypes.c_int),
        "steel_batch_substep_stages": ([V, U32], ctypes.c_int),
        "steel_batch_sync": ([V], ctypes.c_int),
        "steel_batch_read": ([V, U32, U32, SZ, SZ, V], ctypes.c_int),
        "steel_batch_write": ([V, U32, U32, SZ, SZ, V], ctypes.c_int),
        "steel_batch_buffer_size": ([V, U32, U32, ctypes.POINTER(SZ)], ctypes.c_int),
        "steel_batch_stats": ([V, U32, ctypes.POINTER(U32), ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "steel_batch_configure": ([V, U32, ctypes.c_int], ctypes.c_int),
        "steel_batch_set_wide": ([V, ctypes.c_int], ctypes.c_int),
        "steel_batch_device_bytes": ([V, ctypes.POINTER(SZ)], ctypes.c_int),
        "steel_estimate_bytes": ([U32, ctypes.POINTER(WorldDesc), ctypes.POINTER(SZ)], ctypes.c_int),
        "steel_batch_step_timed": ([V, U32, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
        "steel_batch_profile_kernels": ([V, U32, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "steel_server_create": ([V, SZ, U32, ctypes.POINTER(V)], ctypes.c_int),
        "steel_server_destroy": ([V], ctypes.c_int),
        "steel_server_configure_lane": ([V, U32, U32], ctypes.c_int),
        "steel_server_world_bytes": ([ctypes.POINTER(WorldDesc), ctypes.POINTER(SZ)], ctypes.c_int),
        "steel_server_admit": ([V, ctypes.POINTER(WorldDesc), U32, ctypes.POINTER(U32)], ctypes.c_int),
        "steel_server_release": ([V, U32], ctypes.c_int),
        "steel_server_step_lane": ([V, U32], ctypes.c_int),
        "steel_server_lane_poll": ([V, U32, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(U32), ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
        "steel_server_read": ([V, U32, U32, SZ, SZ, V], ctypes.c_int),
        "steel_server_write": ([V, U32, U32, SZ, SZ, V], ctypes.c_int),
        "steel_server_info": ([V, ctypes.POINTER(SZ), ctypes.POINTER(SZ), ctypes.POINTER(U32)], ctypes.c_int),
        "steel_world_step_noalloc": ([V, ctypes.POINTER(WorldDesc), ctypes.POINTER(WorldDeviceArrays), U32, V], ctypes.c_int),
        "steel_batch_autotune": ([V, U32, U32, ctypes.POINTER(TuneResult)], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def last_error():
    """The library's message for the calling thread's last failed call."""
    return _lib.steel_last_error().decode()


def _check(rc):
    if rc != STEEL_OK:
        msg = _lib.steel_last_error().decode()
        if rc == STEEL_UNSUPPORTED:
            raise UnsupportedError(msg)
        raise SteelError(f"{msg} (status {rc})")


class Device:
    """initDevice() (src/device.js:1-27): the GPU the worlds live on."""

    def __init__(self, ordinal=0):
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.steel_device_open(ordinal, ctypes.byref(h)))
        self.handle = h
        name = ctypes.create_string_buffer(256)
        mem, cus = ctypes.c_size_t(), ctypes.c_uint32()
        _check(lib.steel_device_info(h, name, 256, ctypes.byref(mem), ctypes.byref(cus)))
        self.name, self.total_mem, self.num_cus = name.value.decode(), mem.value, cus.value

    def close(self):
        if self.handle:
            _lib.steel_device_close(self.handle)
            self.handle = None


_DEFAULT_DEVICE = None


def default_device():
    global _DEFAULT_DEVICE
    if _DEFAULT_DEVICE is None:
        _DEFAULT_DEVICE = Device(0)
    return _DEFAULT_DEVICE


def _descs(worlds):
    keep = []
    arr = (WorldDesc * len(worlds))()
    for i, pw in enumerate(worlds):
        parts = [np.ascontiguousarray(pw.header, np.uint32), np.ascontiguousarray(pw.statF, np.float32),
                 np.ascontiguousarray(pw.statU if pw.statU.size else np.zeros(1, np.uint32), np.uint32),
                 np.ascontiguousarray(pw.poses0, np.float32), np.ascontiguousarray(pw.vel0 if pw.vel0.size else np.zeros(8, np.float32), np.float32),
                 np.ascontiguousarray(pw.art0, np.float32)]
        keep.append(parts)
        arr[i] = WorldDesc(*[p.ctypes.data for p in parts])
    return arr, keep


def estimate_bytes(worlds):
    lib = load_library()
    arr, keep = _descs(worlds)
    out = ctypes.c_size_t()
    _check(lib.steel_estimate_bytes(len(worlds), arr, ctypes.byref(out)))
    return out.value


class Batch:
    """Many packed worlds stepped together: one launch per kernel per substep for all of them."""

    def __init__(self, worlds, device=None, block=256, graph=True):
        self.lib = load_library()
        self.device = device or default_device()
        self.worlds = list(worlds)
        arr, keep = _descs(self.worlds)
        h = ctypes.c_void_p()
        _check(self.lib.steel_batch_create(self.device.handle, len(self.worlds), arr, ctypes.byref(h)))
        self.handle = h
        self.configure(block, graph)

    def configure(self, block=256, graph=True):
        _check(self.lib.steel_batch_configure(self.handle, block, 1 if graph else 0))

    def set_wide(self, enable=True):
        """steel_batch_set_wide: step this batch's one large world with device-wide (cooperative) launches."""
        _check(self.lib.steel_batch_set_wide(self.handle, 1 if enable else 0))

    def autotune(self, frames=5, remeasure=False):
        """steel_batch_autotune: the fastest (workgroup size, narrowphase grid) over `frames` frames from the current
        state, which is restored afterwards; cached per device (and in the file STEEL_TUNE_CACHE names)."""
        r = TuneResult()
        _check(self.lib.steel_batch_autotune(self.handle, frames, STEEL_TUNE_REMEASURE if remeasure else 0, ctypes.byref(r)))
        return {"block": r.block, "narrow_per_cu": r.narrow_per_cu, "from_cache": bool(r.from_cache), "trials": r.trials,
                "frame_ms": round(float(r.frame_ms), 4)}

    def step(self, frames=1):
        _check(self.lib.steel_batch_step(self.handle, frames))

    def step_timed(self, frames):
        """Step `frames` frames; per-frame GPU milliseconds from HIP events."""
        out = np.zeros(frames, np.float32)
        _check(self.lib.steel_batch_step_timed(self.handle, frames, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def profile_kernels(self, substeps):
        """Average GPU milliseconds per launch of (collide, prepare, solve), HIP events on the batch stream."""
        out = (ctypes.c_double * 3)()
        _check(self.lib.steel_batch_profile_kernels(self.handle, substeps, out))
        return list(out)

    def substep_stages(self, stages):
        _check(self.lib.steel_batch_substep_stages(self.handle, stages))

    def sync(self):
        _check(self.lib.steel_batch_sync(self.handle))

    def reset(self, world=-1):
        _check(self.lib.steel_batch_reset(self.handle, world))

    def buffer_size(self, world, name):
        out = ctypes.c_size_t()
        _check(self.lib.steel_batch_buffer_size(self.handle, world, BUF[name], ctypes.byref(out)))
        return out.value

    def read(self, world, name, nbytes=None, offset=0, dtype=np.float32):
        if nbytes is None:
            nbytes = self.buffer_size(world, name) - offset
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
        _check(self.lib.steel_batch_read(self.handle, world, BUF[name], offset, out.nbytes, out.ctypes.data))
        return out

    def write(self, world, name, data, offset=0):
        data = np.ascontiguousarray(data)
        _check(self.lib.steel_batch_write(self.handle, world, BUF[name], offset, data.nbytes, data.ctypes.data))

    def stats_raw(self, world=0):
        out = (ctypes.c_uint32 * 16)()
        ms = ctypes.c_double()
        _check(self.lib.steel_batch_stats(self.handle, world, out, ctypes.byref(ms)))
        return np.frombuffer(out, np.uint32).copy(), ms.value

    def device_bytes(self):
        out = ctypes.c_size_t()
        _check(self.lib.steel_batch_device_bytes(self.handle, ctypes.byref(out)))
        return out.value

    def destroy(self):
        if self.handle:
            self.lib.steel_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Sim:
    """The reference's Sim (sim.js:63-835) on one MI355X world."""

    def __init__(self, device, scene, opts=None, capacity_policy="reference"):
        self.device = device if device is not None else default_device()
        self.scene = scene
        self.packed = _pack.pack_scene(scene, opts, capacity_policy)
        pw = self.packed
        self.opts = pw.opts
        self.solver = pw.solver
        self.NB = pw.NB
        self.bodyKind = pw.bodyKind
        self.initial = {"bodyPose": pw.poses0[:pw.NB * 8].copy()}
        self.hasTimestamps = True
        self.pick = None
        self.userTargets = False
        self.draggable = list(pw.draggable)
        self.batch = Batch([pw], self.device)
        self.time = 0.0
        self.frame = 0

    @staticmethod
    def create(device, scene, opts=None):
        return Sim(device, scene, opts)

    def destroy(self):
        self.batch.destroy()

    def reset(self):
        """Sim.reset (sim.js:613-623): the initial state; user-set targets stay in force, as there."""
        self.batch.reset(0)
        self.time = 0.0
        self.frame = 0
        if self.userTargets:
            self._user_targets_flag()

    def _user_targets_flag(self):
        # counters word 3 bit 0: the user set joint targets, so a recorded target path stops driving them
        # (the reference passes driven = 0 once userTargets is set, sim.js:633)
        self.batch.write(0, "counters", np.asarray([1], np.uint32), 12)

    def step(self, frames=1, timestamps=False):
        self.batch.step(frames)
        self.frame += frames
        self.time += frames / self.packed.fps

    def run(self, frames, onFrame=None):
        for _ in range(frames):
            if onFrame:
                onFrame(self, self.frame)
            self.step(1)
        self.batch.sync()

    def read(self, name, nbytes, offset=0):
        return self.batch.read(0, name, nbytes, offset, np.uint8).tobytes()

    def poses(self):
        return self.batch.read(0, "poses", self.NB * 32)

    def state(self):
        """Newton State (sim.js:732-763): bodyQ, bodyQd, jointQ, jointQd."""
        pw, S = self.packed, self.scene.json
        J, NB, A = S["joints"], pw.NB, pw.A
        pose = self.batch.read(0, "poses", NB * 32)
        vel = self.batch.read(0, "vel", max(NB, 1) * 32)
        art = self.batch.read(0, "art")
        for l, lk in enumerate(pw.links):
            o = A["A_LINK"] + l * _pack.LB + _pack.L_V
            b = lk["body"]
            vel[b * 8:b * 8 + 8] = [art[o + 3], art[o + 4], art[o + 5], 0, art[o], art[o + 1], art[o + 2], 0]
        bodyQ = np.zeros(NB * 7, np.float32)
        bodyQd = np.zeros(NB * 6, np.float32)
        for b in range(NB):
            q = pose[b * 8 + 4:b * 8 + 8].astype(np.float64)
            r = _pack.quat_rotate(q, S["bodies"]["com"][b])
            bodyQ[b * 7:b * 7 + 7] = [pose[b * 8] - r[0], pose[b * 8 + 1] - r[1], pose[b * 8 + 2] - r[2], *q]
            bodyQd[b * 6:b * 6 + 6] = [vel[b * 8], vel[b * 8 + 1], vel[b * 8 + 2], vel[b * 8 + 4], vel[b * 8 + 5], vel[b * 8 + 6]]
        jointQ = np.asarray(J["q"], np.float32).copy()
        jointQd = np.zeros(J["qd_start"][-1], np.float32)
        for nq, q in pw.qMap.items():
            jointQ[nq] = art[A["A_Q"] + q]
        for nd, d in pw.dofMap.items():
            jointQd[nd] = art[A["A_QD"] + d]
        for j, t in enumerate(J["type"]):
            if t not in (4, 5) or J["q_start"][j] in pw.qMap:
                continue
            b = J["child"][j]
            jointQ[J["q_start"][j]:J["q_start"][j] + 7] = bodyQ[b * 7:b * 7 + 7]
            jointQd[J["qd_start"][j]:J["qd_start"][j] + 6] = bodyQd[b * 6:b * 6 + 6]
        return dict(bodyQ=bodyQ, bodyQd=bodyQd, jointQ=jointQ, jointQd=jointQd)

    def setJointTarget(self, newtonDof, pos, vel=0.0):
        """Joint position/velocity target of one Newton dof (sim.js:716-723): art[A_TGT + d], art[A_TGTV + d]."""
        d = self.packed.dofMap.get(newtonDof)
        if d is None:
            raise ValueError(f"dof {newtonDof} is not simulated as an articulation dof")
        A = self.packed.A
        if not self.userTargets:
            self.userTargets = True
            self._user_targets_flag()
        self.batch.write(0, "art", np.asarray([pos], np.float32), (A["A_TGT"] + d) * 4)
        self.batch.write(0, "art", np.asarray([vel], np.float32), (A["A_TGTV"] + d) * 4)

    def setJointTargets(self, pos, vel=None, first=0):
        """Targets for consecutive Newton dofs from `first` (sim.js:726-728)."""
        for k in range(len(pos)):
            self.setJointTarget(first + k, pos[k], vel[k] if vel is not None else 0.0)

    def readPosesPooled(self):
        return self.poses()

    def setJointForce(self, newtonDof, value):
        d = self.packed.dofMap.get(newtonDof)
        if d is None:
            raise ValueError(f"dof {newtonDof} is not simulated as an articulation dof")
        self.batch.write(0, "art", np.asarray([value], np.float32), (self.packed.A["A_JF"] + d) * 4)

    def setBodyWrench(self, body, force, torque):
        self.batch.write(0, "art", np.asarray([*force, 0, *torque, 0], np.float32), (self.packed.A["A_BF"] + body * 8) * 4)

    def setBodyPose(self, body, com, quat):
        self.batch.write(0, "poses", np.asarray([*com, 0, *quat], np.float32), body * 2 * 16)

    def setBodyTwist(self, body, twist):
        t = twist
        self.batch.write(0, "vel", np.asarray([t[0], t[1], t[2], 0, t[3], t[4], t[5], 0], np.float32), body * 32)

    def stats(self):
        s, ms = self.batch.stats_raw(0)
        pw = self.packed
        out = dict(maxContacts=int(s[1]), maxColors=int(s[2]), maxUnits=int(s[3]), maxPairs=int(s[4]), contacts=int(s[5]),
                   colors=int(s[6]), maxDense=int(s[7]), asleep=int(s[8]))
        out["overflow"] = [n for n, v, cap in (("pairs", s[4], pw.h("W_MAXP")), ("contacts", s[1], pw.h("W_MAXC")),
                                                ("units", s[3], pw.h("W_MAXU")), ("dense rows", s[7], pw.h("W_MAXDENSE"))) if v >= cap]
        out["gpuFrameMs"] = ms
        return out
