"""GPU parity on the large single world: a tenth-size lattice of the 100k-body configuration (10k boxes and hulls, 25
quadrupeds; per-shape filter lists, the solve's link mirror in global memory) from rest and from a piled state, and
a world whose contact keys need body ids past 16 bits.  Pairs, sorted contacts, colors and trajectories match the
CPU oracle bit for bit."""
import sys

import numpy as np
import pytest

from oracle_bind import OracleWorld
from steel_webgpu_amd import pack, workloads

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.fixture(scope="module")
def sim():
    from steel_webgpu_amd import sim as s
    s.default_device()
    return s


@pytest.fixture(scope="module")
def lattice():
    pw = pack.pack_scene(workloads.large_world_small(), capacity_policy="scaled")
    assert pw.NB > 10000 and pw.NA == 25
    return pw


def _same_state(b, o, pw):
    return (np.array_equal(_bits(b.read(0, "poses")), _bits(o.poses)) and
            np.array_equal(_bits(b.read(0, "vel")), _bits(o.vel[:2 * max(pw.NB, 1) * 4])) and
            np.array_equal(_bits(b.read(0, "art")), _bits(o.art)))


def _compare_last_substep(b, o, pw):
    """Pairs (as a set), sorted contact records and per-row colors and units of the last substep (its pair and
    contact counts are stats words 9 and 5; the step's buffers keep that substep's tables)."""
    st = b.stats_raw(0)[0]
    gp = b.read(0, "pairs", dtype=np.uint32)[:2 * min(int(st[9]), pw.h("W_MAXP"))].reshape(-1, 2)
    op = o.dbg_bufs["pairs"][:2 * int(o.dbg_bufs["npairs"][0])].reshape(-1, 2)
    assert st[9] == o.dbg_bufs["npairs"][0]
    assert set(map(tuple, gp.tolist())) == set(map(tuple, op.tolist()))
    nc = int(o.dbg_bufs["ncontacts"][0])
    assert st[5] == nc and nc > 0
    assert np.array_equal(_bits(b.read(0, "contacts", nbytes=nc * 96)), _bits(o.dbg_bufs["contacts"][:nc * 24]))
    L = pack.world_sched_layout(pw)
    sched = b.read(0, "sched", dtype=np.uint32)
    nr = int(o.dbg_bufs["nrows"][0])
    info = o.dbg_bufs["rowinfo"][:nr * 4].reshape(nr, 4)
    assert np.array_equal(sched[L["PCOL"]:L["PCOL"] + nr], info[:, 1])
    assert np.array_equal(sched[L["RUNIT"]:L["RUNIT"] + nr], info[:, 2])
    assert np.array_equal(sched[L["META"]:L["META"] + 4], o.dbg_bufs["meta"][:4])
    return len(op), nc, nr


def test_large_lattice_from_rest_bitexact(sim, lattice):
    pw = lattice
    S = pw.h("W_SUBSTEPS")
    b = sim.Batch([pw], block=256, graph=False)
    o = OracleWorld(pw)
    for k in range(4):
        b.step(5)
        if k < 3:
            o.frames(5)
        else:
            o.frames(4)
            o.substeps(S - 1)
            o.substeps(1, debug=True)
        assert _same_state(b, o, pw), f"frames {5 * k}..{5 * k + 4}"
    npairs, nc, nr = _compare_last_substep(b, o, pw)
    print(f"lattice frame 20: {npairs} pairs, {nc} contacts, {nr} rows", file=sys.stderr)
    b.destroy()


def test_large_lattice_piled_state_bitexact(sim, lattice):
    """The GPU advances the lattice 300 frames (the lowest layers piled on the floor and on the quadrupeds); the
    oracle takes that state and both step 3 more frames."""
    pw = lattice
    S = pw.h("W_SUBSTEPS")
    b = sim.Batch([pw], block=256, graph=True)
    for _ in range(3):
        b.step(100)
    b.sync()
    o = OracleWorld(pw)
    o.poses[:] = b.read(0, "poses")
    o.vel[:2 * pw.NB * 4] = b.read(0, "vel")
    o.art[:] = b.read(0, "art")
    o.akey[:] = b.read(0, "anchor_keys", dtype=np.uint32)
    o.aval[:] = b.read(0, "anchor_vals", dtype=np.uint8)  # the float oracle's anchor record has the kernels' layout
    ctr = b.read(0, "counters", dtype=np.uint32)
    o.count[:4] = ctr[2:6]
    b.step(2)
    o.frames(2)
    assert _same_state(b, o, pw)
    b.step(1)
    o.substeps(S - 1)
    o.substeps(1, debug=True)
    assert _same_state(b, o, pw)
    npairs, nc, nr = _compare_last_substep(b, o, pw)
    ncol = int(o.dbg_bufs["meta"][1])
    print(f"lattice frame 303: {npairs} pairs, {nc} contacts, {nr} rows, {ncol} colors", file=sys.stderr)
    assert nc > 5000 and ncol >= 8
    b.destroy()


def test_body_ids_past_16_bits_bitexact(sim):
    """Two stacked boxes as bodies 65600 and 65601: the GPU matches the oracle bit for bit, and the stack moves as
    it does when numbered 0 and 1."""
    wide = pack.pack_scene(workloads.offset_stack(65600), capacity_policy="scaled")
    small = pack.pack_scene(workloads.offset_stack(0), capacity_policy="scaled")
    b = sim.Batch([wide, small], block=256)
    o = OracleWorld(wide)
    for _ in range(3):
        b.step(20)
        o.frames(20)
        assert np.array_equal(_bits(b.read(0, "poses")), _bits(o.poses))
        assert np.array_equal(_bits(b.read(0, "vel")), _bits(o.vel[:2 * wide.NB * 4]))
    P = wide.h("W_P_BODY")
    stack_w = b.read(0, "poses").reshape(-1, 4)[P + 2 * 65600:P + 2 * 65602]
    stack_s = b.read(1, "poses").reshape(-1, 4)[small.h("W_P_BODY"):small.h("W_P_BODY") + 4]
    assert np.array_equal(_bits(stack_w), _bits(stack_s))
    assert b.stats_raw(0)[0][1] > 0
    b.destroy()


@pytest.mark.skip(reason="device-wide stepping is switched off: its cooperative launches faulted the GPU twice and the "
                         "cause is not yet found")
def test_device_wide_step_equals_the_one_workgroup_step(sim, lattice):
    """The lattice stepped by device-wide cooperative launches (grid teams) matches the one-workgroup-per-stage step
    bit for bit, from rest and from a piled state, and its last substep's pairs, contacts and colors match the
    oracle's."""
    pw = lattice
    S = pw.h("W_SUBSTEPS")
    a = sim.Batch([pw], block=256, graph=False)
    b = sim.Batch([pw], block=256, graph=False)
    b.set_wide(True)
    for k in range(2):
        a.step(5)
        b.step(5)
        for q in ("poses", "vel", "art"):
            assert np.array_equal(_bits(a.read(0, q)), _bits(b.read(0, q))), (k, q)
    for _ in range(3):
        a.step(100)
    state = {q: a.read(0, q, dtype=np.uint8) for q in ("poses", "vel", "art", "anchor_keys", "anchor_vals", "counters", "stats")}
    for q, v in state.items():
        b.write(0, q, v)
    o = OracleWorld(pw)
    o.poses[:] = a.read(0, "poses")
    o.vel[:2 * pw.NB * 4] = a.read(0, "vel")
    o.art[:] = a.read(0, "art")
    o.akey[:] = a.read(0, "anchor_keys", dtype=np.uint32)
    o.aval[:] = a.read(0, "anchor_vals", dtype=np.uint8)
    o.count[:4] = a.read(0, "counters", dtype=np.uint32)[2:6]
    a.step(2)
    b.step(2)
    for q in ("poses", "vel", "art"):
        assert np.array_equal(_bits(a.read(0, q)), _bits(b.read(0, q))), q
    o.frames(1)
    o.substeps(S - 1)
    o.substeps(1, debug=True)
    assert _same_state(b, o, pw)
    npairs, nc, nr = _compare_last_substep(b, o, pw)
    print(f"device-wide, lattice frame 302: {npairs} pairs, {nc} contacts, {nr} rows", file=sys.stderr)
    a.destroy()
    b.destroy()


def test_device_wide_refuses_what_it_does_not_step(sim, lattice):
    small = pack.pack_scene(workloads.quadruped_world(2))
    b = sim.Batch([small])
    with pytest.raises(sim.UnsupportedError):
        b.set_wide(True)
    b.destroy()
    b = sim.Batch([lattice, small])
    with pytest.raises(sim.UnsupportedError):
        b.set_wide(True)
    b.destroy()


@pytest.mark.parametrize("block", [256, 1024], ids=["workgroup", "device_wide_broadphase"])
@pytest.mark.parametrize("far", [False, True], ids=["cluster", "far_clamped"])
def test_cell_broadphase_mixed_sizes_bitexact(sim, far, block):
    """The cell broadphase on a world of 4600 small boxes at random overlapping poses, large free and static boxes
    it does not bin, and the ground plane (with `far`, boxes 5e6 m out clamp every cell coordinate): after one frame
    and after ten, the pair set, sorted contacts, colors and state equal the oracle's.  With 1024-thread workgroups
    the world, alone in its batch, takes the device-wide broadphase (a launch per phase, a multi-launch radix sort)."""
    pw = pack.pack_scene(workloads.cell_mix(far=far), capacity_policy="scaled")
    assert pw.NS > pack.CELL_SHAPES and pw.hf("W_CELL") > 0
    S = pw.h("W_SUBSTEPS")
    b = sim.Batch([pw], block=block, graph=False)
    o = OracleWorld(pw)
    for frames in (1, 9):
        b.step(frames)
        if frames > 1:
            o.frames(frames - 1)
        o.substeps(S - 1)
        o.substeps(1, debug=True)
        assert _same_state(b, o, pw)
        npairs, nc, nr = _compare_last_substep(b, o, pw)
        print(f"cell mix far={far}: {npairs} pairs, {nc} contacts, {nr} rows", file=sys.stderr)
        assert npairs > 4600
    b.destroy()


def test_lattice_1024_thread_workgroups_equal_256(sim, lattice):
    """Workgroups of 1024 threads (broadphase bounds in global memory, 1024-thread sorts) step the lattice bit for
    bit as workgroups of 256 do, over 10 frames."""
    pw = lattice
    a = sim.Batch([pw], block=256, graph=False)
    b = sim.Batch([pw], block=1024, graph=False)
    a.step(10)
    b.step(10)
    for name in ("poses", "vel", "art"):
        assert np.array_equal(_bits(a.read(0, name)), _bits(b.read(0, name))), name
    assert np.array_equal(a.stats_raw(0)[0], b.stats_raw(0)[0])
    a.destroy()
    b.destroy()
