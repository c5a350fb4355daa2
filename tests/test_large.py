"""The large single world on the CPU: its generator, the sorted-list collision filter, and body ids past 16 bits
in the oracle (contact keys, units, coloring and anchors), each pinned against an independent statement."""
import math

import numpy as np
import pytest

from oracle_bind import OracleWorld
from steel_webgpu_amd import pack, workloads


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.fixture(scope="module")
def lattice():
    return workloads.large_world_small(nbodies=2000, nquads=4)


def test_large_world_generator_counts_sizes_and_clearance(lattice):
    """Half boxes, half 12-point hulls, sizes in [0.03, 0.1], quadrupeds of 13 bodies, lattice bodies clear of each
    other by bounding spheres and of the ground."""
    J = lattice.json
    shapes = J["shapes"][1:]  # after the ground plane
    quad_shapes = 4 * 13
    lat = shapes[quad_shapes:]
    assert len(lat) == 2000
    nv = [s["hull"]["nv"] for s in lat]
    assert sum(1 for v in nv[0::2] if v == 8) == 1000          # boxes
    assert all(4 <= v <= 12 for v in nv[1::2])                # hulls of 12 sphere points
    pos = np.asarray(J["bodies"]["q"])[4 * 13:, :3]
    assert pos[:, 2].min() >= 0.19
    # bounding radius of each lattice body <= 0.1 sqrt(3); neighbours on the lattice are >= pitch - 1 cm apart
    d = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=2) + np.eye(len(pos)) * 9
    assert d.min() >= workloads.LARGE_PITCH - 0.0101 > 2 * 0.1 * math.sqrt(3)
    assert len(J["bodies"]["mass"]) == 2000 + 4 * 13


def test_filter_lists_equal_the_bitset(monkeypatch):
    """A world packed with per-shape filter lists steps bit for bit like the same world packed with the bitset (the
    quadrupeds' parent/child filters are the listed pairs)."""
    scene = workloads.large_world_small(nbodies=300, nquads=4)
    pw_bits = pack.pack_scene(scene, capacity_policy="scaled")
    monkeypatch.setattr(pack, "FILT_LIST_SHAPES", 10)
    pw_list = pack.pack_scene(scene, capacity_policy="scaled")
    assert pw_list.h("W_FILT_LIST") == 1 and pw_bits.h("W_FILT_LIST") == 0
    F = pw_list.h("W_SU_FILTER")
    NS = pw_list.NS
    offs = pw_list.statU[F:F + NS + 1]
    listed = {(i, int(j)) for i in range(NS) for j in pw_list.statU[F + NS + 1 + offs[i]:F + NS + 1 + offs[i + 1]]}
    Fb = pw_bits.h("W_SU_FILTER")
    bits = np.unpackbits(pw_bits.statU[Fb:Fb + (NS * NS + 31) // 32].view(np.uint8), bitorder="little")[:NS * NS].reshape(NS, NS)
    assert listed == {(i, j) for i, j in zip(*np.nonzero(bits)) if i < j} and len(listed) == 4 * 12
    a, b = OracleWorld(pw_bits), OracleWorld(pw_list)
    a.frames(30)
    b.frames(30)
    assert np.array_equal(_bits(a.poses), _bits(b.poses)) and np.array_equal(_bits(a.art), _bits(b.art))


def test_body_ids_past_16_bits_leave_the_stack_unchanged():
    """The same two-box stack as bodies 0, 1 and as bodies 65600, 65601 (after 65600 shapeless bodies falling far
    away; the stack's shapes keep their numbers): contact keys, units, colors and friction anchors key the pair by
    ids past 16 bits, and the stack's trajectory is the same bit for bit (a 16-bit truncation of the ids would merge
    the ground pair with another body's)."""
    small = pack.pack_scene(workloads.offset_stack(0), capacity_policy="scaled")
    wide = pack.pack_scene(workloads.offset_stack(65600), capacity_policy="scaled")
    assert wide.NB == 65602 and wide.NS == small.NS == 3
    os_, ow = OracleWorld(small), OracleWorld(wide)
    for _ in range(3):
        os_.frames(20)
        ow.frames(20)
        ps = os_.poses.reshape(-1, 4)[small.h("W_P_BODY"):small.h("W_P_BODY") + 4]
        pw_ = ow.poses.reshape(-1, 4)[wide.h("W_P_BODY") + 2 * 65600:wide.h("W_P_BODY") + 2 * 65602]
        assert np.array_equal(_bits(ps), _bits(pw_))
    assert os_.stats[1] > 0 and ow.stats[1] == os_.stats[1]  # the stack's contacts, and no others


def test_cell_bound_bins_the_small_shapes_only():
    """W_CELL is set past CELL_SHAPES shapes only, bounds every small box's AABB extent under any rotation
    (2 (|half extents| + gap)), and stays below the large boxes' bound so that they are tested against every shape."""
    pw = pack.pack_scene(workloads.cell_mix(n=4200), capacity_policy="scaled")
    small = pack.pack_scene(workloads.offset_stack(0))
    assert small.hf("W_CELL") == 0.0
    HB = pw.hf("W_CELL")
    assert HB > 0
    sf = pw.statF
    rec = sf[pw.h("W_SF_SHAPE") * 4:(pw.h("W_SF_SHAPE") + pw.NS * 7) * 4].reshape(pw.NS, 7, 4)
    hh = sf[pw.h("W_SF_HAABB") * 4:].reshape(-1, 4)[2 * rec[:, 0, 1].astype(np.int64) + 1, :3].astype(np.float64)
    bound = 2 * (np.linalg.norm(hh, axis=1) + rec[:, 0, 3])
    dyn = rec[:, 0, 0] >= 0
    # the 4200 small boxes (half extents <= 0.15 and the default gap) are binned, the 6 large free ones are not
    half = hh.max(1)
    assert np.all(bound[dyn & (half <= 0.15 + 1e-6)] <= HB + 1e-6)
    assert np.all(bound[dyn & (half >= 0.5)] > HB)
    # a rotated box's AABB extent never exceeds the bound: corners of a box with the largest small half extents
    rng = np.random.default_rng(1)
    h = np.array([0.15, 0.15, 0.15])
    for _ in range(200):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        x, y, z, w = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        assert np.all(2 * (np.abs(R) @ h) <= 2 * np.linalg.norm(h) + 1e-12)
