"""GPU checks of the C ABI's argument handling: reads and writes that reach outside one world's buffer are
refused, whatever the offset (a negative Python offset arrives as a huge size_t), and leave every world's
memory untouched."""
import ctypes

import numpy as np
import pytest

from steel_webgpu_amd import pack, serve, workloads

pytestmark = pytest.mark.gpu


def _two_worlds():
    return [pack.pack_scene(workloads.boxes64(workloads.SEED + i, n=4 + i)) for i in range(2)]


@pytest.mark.parametrize("offset", [-4, -16, (1 << 64) - 1, 1 << 63])
def test_batch_write_and_read_refuse_offsets_that_wrap(offset):
    from steel_webgpu_amd import sim
    b = sim.Batch(_two_worlds())
    try:
        before = [b.read(w, "poses").copy() for w in range(2)]
        lib = b.lib
        data = np.full(4, 7.0, np.float32)
        off = ctypes.c_size_t(offset % (1 << 64)).value
        rc = lib.steel_batch_write(b.handle, 1, sim.BUF["poses"], off, data.nbytes, data.ctypes.data)
        assert rc == sim.STEEL_BAD_ARGUMENT
        out = np.zeros(4, np.float32)
        rc = lib.steel_batch_read(b.handle, 1, sim.BUF["poses"], off, out.nbytes, out.ctypes.data)
        assert rc == sim.STEEL_BAD_ARGUMENT
        for w in range(2):
            assert np.array_equal(b.read(w, "poses").view(np.uint32), before[w].view(np.uint32))
    finally:
        b.destroy()


@pytest.mark.parametrize("offset", [-4, (1 << 64) - 1])
def test_server_write_and_read_refuse_offsets_that_wrap(offset):
    from steel_webgpu_amd import sim
    worlds = _two_worlds()
    srv = serve.Server(budget_bytes=256 << 20, lanes=1)
    try:
        ids = [srv.admit(pw, 0) for pw in worlds]
        before = [srv.read(i, "poses").copy() for i in ids]
        data = np.full(4, 7.0, np.float32)
        off = ctypes.c_size_t(offset % (1 << 64)).value
        rc = srv.lib.steel_server_write(srv.handle, ids[1], sim.BUF["poses"], off, data.nbytes, data.ctypes.data)
        assert rc == sim.STEEL_BAD_ARGUMENT
        out = np.zeros(4, np.float32)
        rc = srv.lib.steel_server_read(srv.handle, ids[1], sim.BUF["poses"], off, out.nbytes, out.ctypes.data)
        assert rc == sim.STEEL_BAD_ARGUMENT
        for i, wid in enumerate(ids):
            assert np.array_equal(srv.read(wid, "poses").view(np.uint32), before[i].view(np.uint32))
    finally:
        srv.destroy()


def test_lane_takes_a_new_substep_count_once_empty():
    """A lane fixes its substep count while it holds worlds; once they are released it admits another count."""
    from steel_webgpu_amd import sim
    a = pack.pack_scene(workloads.boxes64(n=4, substeps=4))
    bw = pack.pack_scene(workloads.boxes64(n=5, substeps=8))
    assert a.h("W_SUBSTEPS") != bw.h("W_SUBSTEPS")
    srv = serve.Server(budget_bytes=256 << 20, lanes=1)
    try:
        wa = srv.admit(a, 0)
        with pytest.raises(serve.AdmitRefused):
            srv.admit(bw, 0)
        srv.step(0)
        srv.release(wa)
        wb = srv.admit(bw, 0)
        srv.step(0)
        srv.read(wb, "poses")
    finally:
        srv.destroy()


def test_autotune_leaves_the_state_and_the_trajectory_unchanged_and_caches_per_device():
    """steel_batch_autotune measures every schedule from the batch's current state and restores it: the state right
    after tuning equals the state before, and the tuned batch's next frames equal an untuned batch's bit for bit.
    A second call with the same signature is answered from the device's cache without measuring."""
    from steel_webgpu_amd import sim
    worlds = [pack.pack_scene(workloads.quadruped_world(i), capacity_policy="compact") for i in (0, 3, 7)]
    dev = sim.Device(0)
    try:
        a = sim.Batch(worlds, dev, block=64)
        ref = sim.Batch(worlds, dev, block=64)
        a.step(10)
        ref.step(10)
        before = [[a.read(i, n, dtype=np.uint8) for n in ("poses", "vel", "art", "anchor_keys", "anchor_vals", "counters", "stats")]
                  for i in range(len(worlds))]
        t = a.autotune(frames=2, remeasure=True)
        assert t["trials"] == 9 and not t["from_cache"] and t["block"] in (64, 128, 256) and t["narrow_per_cu"] in (2, 4, 8)
        after = [[a.read(i, n, dtype=np.uint8) for n in ("poses", "vel", "art", "anchor_keys", "anchor_vals", "counters", "stats")]
                 for i in range(len(worlds))]
        assert all(np.array_equal(x, y) for bx, by in zip(before, after) for x, y in zip(bx, by))
        a.step(10)
        ref.step(10)
        for i in range(len(worlds)):
            for n in ("poses", "vel", "art"):
                assert np.array_equal(a.read(i, n, dtype=np.uint8), ref.read(i, n, dtype=np.uint8)), (i, n)
        c = sim.Batch(worlds, dev, block=64)
        t2 = c.autotune(frames=2)
        assert t2["from_cache"] and t2["trials"] == 0 and (t2["block"], t2["narrow_per_cu"]) == (t["block"], t["narrow_per_cu"])
        for b in (a, ref, c):
            b.destroy()
    finally:
        dev.close()
