"""Serving pool on the GPU: continuous batching gives every request the same result as stepping its world
alone (bit for bit), admission follows the HBM budget, and a world larger than the pool is refused."""
import numpy as np
import pytest

from steel_webgpu_amd import pack, serve, workloads

pytestmark = pytest.mark.gpu


def _requests(n, seed_offset=0, nmax=260, hmax=40, rate=400.0):
    rng = workloads.PCG32(workloads.SEED + 77, seed_offset)
    arr = serve.poisson_arrivals(n, rate, rng)
    out = []
    for i in range(n):
        sc, h = workloads.serving_request(i + seed_offset, nmin=10, nmax=nmax, hmin=10, hmax=hmax)
        pw = pack.pack_scene(sc, capacity_policy="compact")
        out.append(serve.Request(rid=i, arrival=arr[i], world=pw, horizon=h, lane=serve.lane_of(pw, edges=(40, 150))))
    return out


def _alone(pw, frames):
    from steel_webgpu_amd import sim
    b = sim.Batch([pw], graph=False, block=64)
    b.step(frames)
    out = {n: b.read(0, n) for n in ("poses", "vel", "art")}
    b.destroy()
    return out


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_served_results_equal_alone_results():
    reqs = _requests(10)
    assert len({r.lane for r in reqs}) >= 2
    srv = serve.Server(budget_bytes=2 << 30, lanes=3)
    m = serve.run(reqs, srv, readback=("poses", "vel", "art"))
    srv.destroy()
    assert m["rejected"] == 0 and m["world_steps"] == sum(r.horizon for r in reqs)
    for r in reqs:
        ref = _alone(r.world, r.horizon)
        for name in ("poses", "vel", "art"):
            assert np.array_equal(_bits(r.result[name]), _bits(ref[name])), (r.rid, name)


def test_admission_waits_for_budget_and_results_stay_exact():
    reqs = _requests(6, seed_offset=50, nmax=120, hmax=30, rate=5000.0)
    need = sorted(serve.Server.world_bytes(r.world) for r in reqs)
    budget = need[-1] + need[-2] // 2  # the largest world fits, but never two of the larger ones together
    srv = serve.Server(budget_bytes=budget, lanes=3)
    m = serve.run(reqs, srv, readback=("poses",))
    used, cap, nworlds = srv.info()
    srv.destroy()
    assert m["rejected"] == 0 and m["admission_deferrals"] > 0
    assert m["peak_pool_bytes"] <= cap and used == 0 and nworlds == 0
    for r in reqs:
        assert np.array_equal(_bits(r.result["poses"]), _bits(_alone(r.world, r.horizon)["poses"])), r.rid


def test_world_larger_than_pool_is_refused():
    reqs = _requests(1, seed_offset=90)
    srv = serve.Server(budget_bytes=serve.Server.world_bytes(reqs[0].world) // 2, lanes=1)
    reqs[0].lane = 0
    assert srv.admit(reqs[0].world, 0) is None
    m = serve.run(reqs, srv)
    srv.destroy()
    assert m["rejected"] == 1 and m["world_steps"] == 0


def test_policy_writes_between_lane_frames_match_a_batch_world():
    """Control in the serving loop: a policy writes new joint targets into an admitted world before each lane
    frame (steel_server_write); the served rollout equals the same writes into a batch world between its frames,
    bit for bit."""
    from steel_webgpu_amd import sim
    pw = pack.pack_scene(workloads.quadruped_world(3, drive=dict(targetKe=40.0, targetKd=1.0)))
    off, nd = pw.A["A_TGT"] * 4, pw.ND
    srv = serve.Server(budget_bytes=1 << 30, lanes=1)
    wid = srv.admit(pw, 0)
    b = sim.Batch([pw], graph=False)
    rng = np.random.default_rng(4)
    for f in range(30):
        tgt = rng.uniform(-0.4, 0.4, nd).astype(np.float32)
        srv.write(wid, "art", tgt, offset=off)
        b.write(0, "art", tgt, offset=off)
        srv.step(0)
        while not srv.poll(0)[0]:
            pass
        b.step(1)
    for name in ("poses", "art"):
        got, ref = srv.read(wid, name), b.read(0, name)
        assert np.array_equal(_bits(got), _bits(ref[:got.size])), name
    assert not np.array_equal(srv.read(wid, "art")[pw.A["A_TGT"]:pw.A["A_TGT"] + nd], pw.art0[pw.A["A_TGT"]:pw.A["A_TGT"] + nd])
    srv.release(wid)
    srv.destroy()
    b.destroy()
