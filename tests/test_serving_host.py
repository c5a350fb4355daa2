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


class _FakeServer:
    """A host-only stand-in for serve.Server: lane frames finish at the next poll; it enforces what the library
    enforces at admission (one substep count per non-empty lane) and refuses worlds marked unsteppable."""

    def __init__(self, nlanes):
        self.nlanes = nlanes
        self.lanes = [[] for _ in range(nlanes)]
        self.world = {}
        self.next = 1
        self.mixed = False

    def admit(self, pw, lane):
        if getattr(pw, "refuse", False):
            raise serve.AdmitRefused("world refused")
        S = pw.h("W_SUBSTEPS")
        if any(self.world[w][0].h("W_SUBSTEPS") != S for w in self.lanes[lane]):
            self.mixed = True
            raise sim.SteelError("lane steps another substep count")
        wid = self.next
        self.next += 1
        self.world[wid] = (pw, lane)
        self.lanes[lane].append(wid)
        return wid

    def info(self):
        return len(self.world), 1 << 40, len(self.world)

    def poll(self, lane):
        return True, len(self.lanes[lane]), 1.0

    def step(self, lane):
        pass

    def read(self, wid, name):
        return np.zeros(1, np.float32)

    def release(self, wid):
        pw, lane = self.world.pop(wid)
        self.lanes[lane].remove(wid)


def test_run_defers_other_substep_counts_and_rejects_refused_worlds():
    """Requests of two substep counts share one lane: each waits until the lane has drained of the other count,
    none is refused for it; a world the library refuses is counted as rejected, not raised."""
    sc, _ = workloads.serving_request(3, nmin=10, nmax=20)
    reqs = []
    for i in range(12):
        pw = pack.pack_scene(sc, capacity_policy="compact")
        if i % 3 == 1:
            pw.header[pack.FIELDS["W_SUBSTEPS"]] = 8
        pw.refuse = (i == 7)
        reqs.append(serve.Request(rid=i, arrival=0.0, world=pw, horizon=2 + i % 4, lane=0))
    fake = _FakeServer(1)
    m = serve.run(reqs, fake, clock=iter(np.arange(0, 1e6, 1e-3)).__next__)
    assert not fake.mixed
    assert m["rejected"] == 1
    assert m["world_steps"] == sum(r.horizon for r in reqs if r.rid != 7)
    assert all(not np.isnan(r.finished) for r in reqs)
