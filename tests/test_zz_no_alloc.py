"""steel_world_step_noalloc against steel_batch_step: the same world stepped from the same state through the
entry point that allocates nothing (caller-held device arrays, the library's fixed scratch array) and through a
batch of that one world must agree bit for bit, on two calls in a row (the second a smaller world, so the
scratch array holds the first call's larger tables when it starts), and when a world is stepped one frame per
call (nothing may carry in the scratch array from one call to the next)."""
import ctypes

import numpy as np
import pytest

from steel_webgpu_amd import pack, workloads

pytestmark = pytest.mark.gpu

STATE = ("poses", "vel", "art", "anchor_keys", "anchor_vals", "warm", "counters", "stats")


def _hip():
    """The HIP runtime the library itself uses (already loaded by it), for the caller's own device buffers."""
    from steel_webgpu_amd import sim
    sim.load_library()
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


H2D, D2H = 1, 2


class DeviceWorld:
    """One world's static arrays and zero-at-reset state in device memory the caller allocated itself."""

    def __init__(self, pw, sizes):
        self.hip = _hip()
        self.pw = pw
        self.ptr, self.nbytes = {}, {}

        def dev(name, a):
            a = np.ascontiguousarray(a).view(np.uint8)
            p = ctypes.c_void_p()
            assert self.hip.hipMalloc(ctypes.byref(p), a.size) == 0
            assert self.hip.hipMemcpy(p, a.ctypes.data, a.size, H2D) == 0
            self.ptr[name], self.nbytes[name] = p.value, a.size

        dev("header", pw.header.astype(np.uint32))
        dev("statF", pw.statF.astype(np.float32))
        dev("statU", pw.statU.astype(np.uint32) if pw.statU.size else np.zeros(1, np.uint32))
        init = dict(poses=pw.poses0.astype(np.float32), vel=(pw.vel0 if pw.vel0.size else np.zeros(8, np.float32)).astype(np.float32),
                    art=pw.art0.astype(np.float32))
        for name in STATE:
            buf = np.zeros(sizes[name], np.uint8)
            if name in init:
                src = init[name].view(np.uint8)
                buf[:src.size] = src
            dev(name, buf)

    def arrays(self):
        from steel_webgpu_amd import sim
        return sim.WorldDeviceArrays(**self.ptr)

    def read(self, name):
        out = np.zeros(self.nbytes[name], np.uint8)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr[name], out.size, D2H) == 0
        return out

    def free(self):
        for p in self.ptr.values():
            self.hip.hipFree(p)
        self.ptr = {}


def _sizes(batch):
    return {n: batch.buffer_size(0, n) for n in STATE}


def _step_noalloc(lib, dev, pw, dw, frames):
    from steel_webgpu_amd import sim
    arr, keep = sim._descs([pw])
    arrays = dw.arrays()
    rc = lib.steel_world_step_noalloc(dev.handle, arr, ctypes.byref(arrays), frames, None)
    assert rc == 0, lib.steel_last_error().decode()
    assert dw.hip.hipDeviceSynchronize() == 0


def _compare(batch, dw):
    for name in STATE:
        want = batch.read(0, name, dtype=np.uint8)
        got = dw.read(name)
        assert got.size == want.size, name
        assert np.array_equal(got, want), f"{name} differs"


def _world(kind):
    if kind == "large":
        return pack.pack_scene(workloads.quadruped_world(9))   # a quadruped and its boxes, reference capacities
    return pack.pack_scene(workloads.boxes64(n=6))             # six boxes


def test_noalloc_step_matches_batch_step_on_two_calls_the_second_smaller():
    from steel_webgpu_amd import sim
    lib = sim.load_library()
    dev = sim.default_device()
    for kind, frames in (("large", 12), ("small", 5)):
        pw = _world(kind)
        print(f"noalloc two calls: {kind} world NB={pw.NB} NS={pw.NS}, {frames} frames", flush=True)
        b = sim.Batch([pw], dev, block=64, graph=True)
        try:
            dw = DeviceWorld(pw, _sizes(b))
            _step_noalloc(lib, dev, pw, dw, frames)
            b.step(frames)
            b.sync()
            _compare(b, dw)
            assert int(b.stats_raw(0)[0][15]) == frames * pw.h("W_SUBSTEPS")
            dw.free()
        finally:
            b.destroy()


def test_noalloc_one_frame_per_call_matches_continuous_batch_stepping():
    from steel_webgpu_amd import sim
    lib = sim.load_library()
    dev = sim.default_device()
    pw = _world("large")
    print(f"noalloc per-frame calls: NB={pw.NB}", flush=True)
    b = sim.Batch([pw], dev, block=64, graph=True)
    try:
        dw = DeviceWorld(pw, _sizes(b))
        for _ in range(10):
            _step_noalloc(lib, dev, pw, dw, 1)
        b.step(10)
        b.sync()
        _compare(b, dw)
        dw.free()
    finally:
        b.destroy()


def test_noalloc_refuses_a_world_beyond_its_scratch():
    """A world whose step scratch exceeds the fixed array is refused before any GPU work."""
    from steel_webgpu_amd import sim
    lib = sim.load_library()
    dev = sim.default_device()
    pw = pack.pack_scene(workloads.boxes64(n=6), capacity_override=dict(MAXP=200000, MAXC=60000))
    arr, keep = sim._descs([pw])
    arrays = sim.WorldDeviceArrays(*([1] * 11))  # never dereferenced: the size check comes first
    rc = lib.steel_world_step_noalloc(dev.handle, arr, ctypes.byref(arrays), 1, None)
    assert rc == sim.STEEL_BAD_ARGUMENT
    assert b"scratch" in lib.steel_last_error()
