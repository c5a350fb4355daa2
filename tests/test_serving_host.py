"""Host side of serving: arrivals, lanes and pool sizing (no GPU needed)."""
import numpy as np

from steel_webgpu_amd import pack, serve, sim, workloads


def test_poisson_arrivals_have_the_requested_rate():
    rng = workloads.PCG32(1, 2)
    t = np.asarray(serve.poisson_arrivals(20000, 1000.0, rng))
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0)
    assert abs(gaps.mean() - 1e-3) < 3e-5            # mean 1/rate
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.03  # exponential: coefficient of variation 1


def test_serving_request_sizes_follow_the_distribution():
    ns = []
    for i in range(40):
        sc, h = workloads.serving_request(i, nmin=10, nmax=60, hmin=50, hmax=1000)
        ns.append(len(sc.json["bodies"]["mass"]))
        assert 50 <= h <= 1000
    assert min(ns) >= 10 and max(ns) <= 60 + 4 * 13


def test_lane_of_uses_body_count_edges():
    sc, _ = workloads.serving_request(3, nmin=10, nmax=30)
    pw = pack.pack_scene(sc, capacity_policy="compact")
    assert serve.lane_of(pw, edges=(pw.NB + 1, pw.NB + 2)) == 0
    assert serve.lane_of(pw, edges=(pw.NB, pw.NB + 2)) == 1
    assert serve.lane_of(pw, edges=(1, 2)) == 2


def test_world_pool_bytes_match_batch_estimate():
    sc, _ = workloads.serving_request(5, nmin=10, nmax=80)
    pw = pack.pack_scene(sc, capacity_policy="compact")
    per_world = serve.Server.world_bytes(pw)
    table = 256  # one world pointer record, rounded to 256 bytes
    assert per_world + table == sim.estimate_bytes([pw])
    assert per_world > pw.statF.nbytes + pw.art0.nbytes
