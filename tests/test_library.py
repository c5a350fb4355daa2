"""CPU-side checks of the boundary: the C ABI library loads and exports every declared symbol, the host
packing follows the reference's rules, unsupported options raise."""
import re
import os

import numpy as np
import pytest

from steel_webgpu_amd import SceneBuilder, pack, workloads, scene as scene_mod

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    from steel_webgpu_amd import sim
    lib = sim.load_library()
    hdr = open(os.path.join(ROOT, "include", "steel_amd.h")).read()
    declared = set(re.findall(r"^(?:int|const char\*)\s+(steel_\w+)\(", hdr, re.M))
    assert declared and declared == set(sim.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def test_device_open_without_gpu_is_unsupported():
    from steel_webgpu_amd import sim
    import ctypes
    lib = sim.load_library()
    h = ctypes.c_void_p()
    rc = lib.steel_device_open(0, ctypes.byref(h))
    if rc == 0:
        pytest.skip("a GPU is present")
    assert rc == sim.STEEL_UNSUPPORTED
    assert b"device" in lib.steel_last_error()


def test_header_fields_match_c_enum():
    """The packer reads the header field order from include/steel_world.h; check it against an independent parse."""
    src = open(os.path.join(ROOT, "include", "steel_world.h")).read()
    body = src[src.index("enum steel_world_field"):]
    body = body[body.index("{") + 1:body.index("}")]
    names = [t.split("=")[0].strip() for t in re.sub(r"/\*.*?\*/|//[^\n]*", "", body, flags=re.S).split(",") if t.strip()]
    assert names[-1] == "W_FIELD_COUNT" and len(names) <= pack.HEADER_WORDS
    for i, n in enumerate(names[:-1]):
        assert pack.FIELDS[n] == i, n
    assert len(pack.FIELDS) == len(names) - 1


def test_sched_layout_mirror_matches_device_header():
    src = open(os.path.join(ROOT, "steel_webgpu_amd", "csrc", "steel_dev.h")).read()
    steps = re.findall(r"s\.(\w+) = o; o \+= ([^;]+);|(o = \(o \+ 1\) & ~1u;)", src)
    assert any(al for _, _, al in steps)
    for WH, MAXDENSE in ((0, 7), (16384, 7)):
        L = pack.sched_layout(1001, 300, 256, 64, 40, 2, 12, WH, MAXDENSE)
        env = dict(MAXC=1001, MAXU=300, UH=256, NCOL=64, NB=40, NCOLW=2, NL=12, WH=WH, MAXDENSE=MAXDENSE)
        o = 0
        for name, expr, align in steps:
            if align:
                o = (o + 1) & ~1
                continue
            assert L[name] == o, name
            o += eval(expr.replace("(NB > 0 ? NB : 1)", "max(NB,1)").replace("NB > 0 ? NB : 1", "max(NB,1)")
                      .replace("NL > 0 ? NL : 1", "max(NL,1)").replace("WH ? MAXC + MAXDENSE : 0", "(MAXC + MAXDENSE if WH else 0)"),
                      {"max": max}, env)
        assert L["TOTAL"] == o and L["WWIN"] % 2 == 0


def test_filter_bitset_holds_joint_parent_child_and_explicit_filters():
    b = SceneBuilder()
    b.addGroundPlane()
    p = b.addBody(xform=[0, 0, 1, 0, 0, 0, 1])
    sp = b.addShapeBox(p, hx=0.1, hy=0.1, hz=0.1)
    c = b.addBody(xform=[0, 0, 1.3, 0, 0, 0, 1])
    sc = b.addShapeBox(c, hx=0.1, hy=0.1, hz=0.1)
    d = b.addBody(xform=[1, 0, 1, 0, 0, 0, 1])
    sd = b.addShapeSphere(d, radius=0.1)
    b.addJointRevolute(p, c, axis=[0, 1, 0])
    b.addCollisionFilter(sp, sd)
    pw = pack.pack_scene(b.finalize())
    NS = pw.NS
    bits = pw.statU[pw.h("W_SU_FILTER"):]

    def filtered(i, j):
        k = i * NS + j
        return bool(bits[k >> 5] & (1 << (k & 31)))

    assert filtered(sp, sc) and filtered(sc, sp)
    assert filtered(sp, sd) and filtered(sd, sp)
    assert not filtered(sc, sd) and not filtered(0, sp)


def test_capacities_are_exact_bounds():
    maxp, maxc, maxu, uh, ah, md = pack.capacities(65, 64, 0)
    assert maxp == 65 * 64 // 2 and maxc == min(8192, 5 * maxp)
    assert pack.capacities(300, 300, 1)[0] == 16384  # the reference's MAXP
    c = pack.capacities(25, 24, 1, "compact")
    assert c[1] == 400 and c[3] <= c[1]


@pytest.mark.parametrize("ns", [2, 10, 65, 291, 700, 2000, 5600])
def test_scaled_capacities_floor_at_the_reference_and_grow_with_the_world(ns):
    """The scaled sizes never go below the reference's (where the world can reach them), grow with the shape
    count, keep each hash table within the contact capacity and the anchor table a power of two."""
    ref = pack.capacities(ns, ns, 1)
    maxp, maxc, maxu, uh, ah, md = pack.capacities(ns, ns, 1, "scaled")
    assert maxp >= ref[0] and maxc >= ref[1] and maxu >= ref[2] and uh >= ref[3] and ah >= ref[4] and md == ref[5]
    assert maxp <= ns * (ns - 1) // 2 and maxc <= 5 * maxp and maxu <= maxc and uh <= maxc
    assert ah & (ah - 1) == 0 and ah >= 2 * maxu
    if ns >= 700:  # past the reference's fixed sizes
        assert maxu == 12 * ns and maxc == 32 * ns and maxp == 24 * ns
    pw = pack.pack_scene(workloads.arm_hulls(nlinks=4, nhulls=8), capacity_policy="scaled")
    assert pw.h("W_MAXU") <= pw.h("W_MAXC") and pw.h("W_UH") <= pw.h("W_MAXC")


def test_unsupported_options_raise():
    with pytest.raises(pack.UnsupportedError):
        pack.pack_scene(workloads.boxes64(n=4), opts={"solver": {"row_schedule": "contact_then_internal"}})
    with pytest.raises(ValueError):
        pack.pack_scene(workloads.boxes64(n=4), opts={"substeps": 64})
    with pytest.raises(pack.UnsupportedError):
        pack.pack_scene(workloads.boxes64(n=4), opts={"solver": {"enable_torsional_rolling_friction": True,
                                                                  "torsional_rolling_friction_cone": "pyramidal"}})


def test_convex_hull_faces_are_outward_ccw_planes():
    rng = np.random.default_rng(3)
    pts = rng.normal(size=(14, 3))
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    h = scene_mod.convexHull(pts.tolist())
    V = np.array(h["verts"])
    for f in h["faces"]:
        n, d = np.array(f["n"]), f["d"]
        assert abs(np.linalg.norm(n) - 1) < 1e-9
        assert np.all(V @ n - d <= 1e-9)  # all vertices behind the plane
        ring = V[f["v"]]
        np.testing.assert_allclose(ring @ n, d, atol=1e-9)
        area = sum(np.cross(ring[i] - ring[0], ring[i + 1] - ring[0]) for i in range(1, len(ring) - 1))
        assert area @ n > 0  # counter-clockwise about the outward normal


def test_convex_hull_of_a_tetrahedron_traced_by_hand():
    """The reference quickhull (builder.js:61-173) on the corner tetrahedron, traced by hand: i0 = 0 (smallest x,
    first), i1 = 1 (first farthest from it), i2 = 2 (first farthest from that line), i3 = 3; the initial faces are
    (0,2,1), (0,1,3), (0,3,2) (the first and third flipped to face away from the centroid) and (1,2,3); each is its
    own plane, keeping its triangle's (n, d); each loop is the monotone chain over (u, w) coordinates starting at
    the smallest u."""
    h = scene_mod.convexHull([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    assert h["verts"] == [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]
    s3 = 1 / np.sqrt(3)
    want = [((0, 0, -1), 0.0, [2, 1, 0]), ((0, -1, 0), 0.0, [1, 3, 0]), ((-1, 0, 0), 0.0, [3, 2, 0]), ((s3, s3, s3), s3, [3, 1, 2])]
    assert len(h["faces"]) == 4
    for f, (n, d, v) in zip(h["faces"], want):
        np.testing.assert_allclose(f["n"], n, atol=1e-15)
        assert abs(f["d"] - d) < 1e-15 and f["v"] == v


def test_convex_hull_of_a_cube_traced_by_hand():
    """The unit cube in the builder's corner order (x, then y, then z): the tetrahedron (0, 7, 1, 2), then the
    points 4, 3, 5, 6 grow the hull in that order, and the flood fill meets the faces' planes in the order -x, -y,
    -z, +y, +z, +x; the -x face's loop is 3, 2, 0, 1 (traced through the monotone chain)."""
    P = [[x, y, z] for x in (0.0, 1.0) for y in (0.0, 1.0) for z in (0.0, 1.0)]
    h = scene_mod.convexHull(P)
    assert h["verts"] == P
    normals = [tuple(f["n"]) for f in h["faces"]]
    assert normals == [(-1.0, 0.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0)]
    assert [f["d"] for f in h["faces"]] == [0.0, 0.0, 0.0, 1.0, 1.0, 1.0]
    assert h["faces"][0]["v"] == [3, 2, 0, 1]
    V = np.array(P)
    for f in h["faces"]:
        ring = V[f["v"]]
        assert len(ring) == 4 and np.allclose(ring @ np.array(f["n"]), f["d"])
        area = sum(np.cross(ring[i] - ring[0], ring[i + 1] - ring[0]) for i in range(1, len(ring) - 1))
        assert area @ np.array(f["n"]) == pytest.approx(2.0)  # twice the unit face area, counter-clockwise


def test_convex_hull_invariants_on_random_clouds():
    """Every input point is inside or on the hull, each face is a convex CCW polygon on its plane, and the
    polyhedron is closed (V - E + F = 2, every edge shared by exactly two faces in opposite directions)."""
    rng = np.random.default_rng(9)
    for trial in range(20):
        n = int(rng.integers(5, 40))
        pts = rng.normal(size=(n, 3)) * rng.uniform(0.05, 2.0, size=3)
        if trial % 3 == 0:
            pts = np.round(pts * 4) / 4  # lattice points: coplanar faces to merge
            if np.linalg.matrix_rank(pts - pts.mean(0)) < 3:
                continue
        h = scene_mod.convexHull(pts.tolist())
        V = np.array(h["verts"])
        edges = {}
        for f in h["faces"]:
            nrm, d = np.array(f["n"]), f["d"]
            assert np.all(pts @ nrm - d <= 1e-9 * (1 + np.abs(pts).max()))
            ring = f["v"]
            for i in range(len(ring)):
                a, b = ring[i], ring[(i + 1) % len(ring)]
                assert (a, b) not in edges
                edges[(a, b)] = True
        for a, b in edges:
            assert (b, a) in edges
        assert len(V) - len(edges) // 2 + len(h["faces"]) == 2


def test_scene_folder_round_trip(tmp_path):
    sc = workloads.boxes64(n=5)
    sc.save(str(tmp_path))
    back = scene_mod.loadScene(str(tmp_path))
    a, b = pack.pack_scene(sc), pack.pack_scene(back)
    assert np.array_equal(a.statF, b.statF) and np.array_equal(a.statU, b.statU) and np.array_equal(a.header, b.header)


def test_box_mass_properties_closed_form():
    mp = scene_mod.massProps(scene_mod.GT["BOX"], [0.1, 0.2, 0.3], None)
    m = 8 * 0.1 * 0.2 * 0.3
    assert mp["m"] == pytest.approx(m)
    assert mp["I"][0] == pytest.approx(m / 3 * (0.04 + 0.09))
    h = scene_mod.convexHull(scene_mod.shapePoints(scene_mod.GT["BOX"], [0.1, 0.2, 0.3]))
    mh = scene_mod.massProps(scene_mod.GT["CONVEX_MESH"], None, h)
    assert mh["m"] == pytest.approx(m)
    np.testing.assert_allclose(mh["I"], mp["I"], atol=1e-12)
    assert len(h["faces"]) == 6 and len(h["verts"]) == 8


def test_scene_fixtures_match_a_fresh_pack():
    """The packed reference scene fixtures (tests/golden/scenes) equal what the packer produces now from the
    shipped scene folders (skipped where the reference tree is absent)."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import scene_worlds
    from steel_webgpu_amd.scene import loadScene
    root = "/root/reference/assets/scenes"
    names = scene_worlds.names()
    assert len(names) == 41
    if not os.path.isdir(root):
        pytest.skip("reference scene folders not present")
    for n in names:
        fx = scene_worlds.load(n)
        pw = pack.pack_scene(loadScene(os.path.join(root, n)))
        for k in scene_worlds.FIELDS:
            assert np.array_equal(getattr(fx, k), getattr(pw, k)), (n, k)
